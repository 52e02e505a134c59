// Element stress (simulation.py:123-135, 214-248) and its hand-written adjoint.
//
// Forward, per element, from the deformation gradient F:
//   Gram-Schmidt QR with norm(eps)=sqrt(v.v+1e-9), R = inverse(Q) @ F   (:123-135)
//   Ft = R[0:2,0:2]; polar rotation Rr = U V^T of Ft                       (:24-28, :39-42)
//   P_fc = 2 mu (Ft - Rr) + la (J-1) J inverse(Ft^T)                       (:234-237)
//   A from R[:,2] with shearing/contact stiffness (A[0,0]'s E term is
//   overwritten by the shear-only value, :216-217)                        (:215-229)
//   P = Q @ [P_fc 0; 0 0] + Q @ A @ inverse(R)^T                          (:238-240)
// The 2x2 polar runs in PT (double by default).  Its adjoint is the exact
// derivative of the rotation angle; in reference mode it is scaled by
// min(1, |s1^2 - s0^2| / 1e-6), which equals the reference's PyTorch-style
// SVD backward with 1/clamp(s_j^2 - s_i^2) (simulation.py:44-50, 71-94)
// for every F with s0 + s1 != 0 (the clamped off-diagonal entries are the
// exact ones times |delta|/1e-6, and at delta = 0 both terms cancel).
#pragma once
#include "sim_common.hpp"

struct ElemParams { float E, nu, ks, kc; };

template <typename PT>
struct PolarOut { float c, s; };

template <typename PT>
__device__ __forceinline__ void polar2(float a00, float a01, float a10, float a11, float& c, float& s) {
    PT x = (PT)a00 + (PT)a11;
    PT y = (PT)a10 - (PT)a01;
    PT r = sqrt(x * x + y * y);
    c = (float)(x / r);
    s = (float)(y / r);
}

// d/dFt of <gR, Rr(Ft)>; mode 0 = reference clamped, 1 = exact.
template <typename PT>
__device__ __forceinline__ void polar2_adj(float a00, float a01, float a10, float a11, const float gR[4],
                                           int mode, float gF[4]) {
    PT x = (PT)a00 + (PT)a11;
    PT y = (PT)a10 - (PT)a01;
    PT r2 = x * x + y * y;
    PT r = sqrt(r2);
    PT c = x / r, s = y / r;
    PT gth = -s * (PT)gR[0] - c * (PT)gR[1] + c * (PT)gR[2] - s * (PT)gR[3];
    PT gx = gth * (-y / r2);
    PT gy = gth * (x / r2);
    if (mode == 0) {
        PT hd = ((PT)a00 - (PT)a11) * (PT)0.5;
        PT ho = ((PT)a10 + (PT)a01) * (PT)0.5;
        PT delta = (PT)2 * r * sqrt(hd * hd + ho * ho);   // |s1^2 - s0^2|
        PT sc = delta / (PT)1e-6;
        if (sc < (PT)1) { gx *= sc; gy *= sc; }
    }
    gF[0] += (float)gx;   // a00
    gF[3] += (float)gx;   // a11
    gF[2] += (float)gy;   // a10
    gF[1] -= (float)gy;   // a01
}

struct QRState {
    V3<float> c1, c2, c3, q1, q2, q3, h2, h3;
    float n1, n2, n3, a21, a31, a32;
    M3<float> Q, Qi, R;
};

__device__ __forceinline__ void qr_forward(const M3<float>& F, QRState& S) {
    S.c1 = {F(0, 0), F(1, 0), F(2, 0)};
    S.c2 = {F(0, 1), F(1, 1), F(2, 1)};
    S.c3 = {F(0, 2), F(1, 2), F(2, 2)};
    S.n1 = sqrtf(dot(S.c1, S.c1) + 1e-9f);
    S.q1 = {S.c1.x / S.n1, S.c1.y / S.n1, S.c1.z / S.n1};
    S.a21 = dot(S.c2, S.q1);
    S.h2 = S.c2 - S.a21 * S.q1;
    S.n2 = sqrtf(dot(S.h2, S.h2) + 1e-9f);
    S.q2 = {S.h2.x / S.n2, S.h2.y / S.n2, S.h2.z / S.n2};
    S.a31 = dot(S.c3, S.q1);
    S.a32 = dot(S.c3, S.q2);
    S.h3 = (S.c3 - S.a31 * S.q1) - S.a32 * S.q2;
    S.n3 = sqrtf(dot(S.h3, S.h3) + 1e-9f);
    S.q3 = {S.h3.x / S.n3, S.h3.y / S.n3, S.h3.z / S.n3};
    for (int i = 0; i < 3; ++i) { S.Q(i, 0) = S.q1[i]; S.Q(i, 1) = S.q2[i]; S.Q(i, 2) = S.q3[i]; }
    S.Qi = minv(S.Q);
    S.R = mmul(S.Qi, F);
}

// adjoint of q = h / sqrt(h.h + eps)
__device__ __forceinline__ V3<float> normalize_adj(V3<float> h, float n, V3<float> gq) {
    float hg = dot(h, gq);
    float inv = 1.0f / n;
    float inv3 = inv * inv * inv;
    return {gq.x * inv - h.x * hg * inv3, gq.y * inv - h.y * hg * inv3, gq.z * inv - h.z * hg * inv3};
}

// given dL/dQ and dL/dR, accumulate dL/dF
__device__ __forceinline__ void qr_adjoint(const M3<float>& F, const QRState& S, const M3<float>& gQ_in,
                                           const M3<float>& gR, M3<float>& gF) {
    // R = Qi F
    M3<float> QiT = mT(S.Qi);
    M3<float> gFd = mmul(QiT, gR);
    M3<float> gQi = mmul(gR, mT(F));
    // Qi = inverse(Q): gQ = -Qi^T gQi Qi^T
    M3<float> gQ = madd(gQ_in, mscale(-1.0f, mmul(mmul(QiT, gQi), QiT)));
    V3<float> gq1 = {gQ(0, 0), gQ(1, 0), gQ(2, 0)};
    V3<float> gq2 = {gQ(0, 1), gQ(1, 1), gQ(2, 1)};
    V3<float> gq3 = {gQ(0, 2), gQ(1, 2), gQ(2, 2)};
    // q3 = h3/n3, h3 = c3 - a31 q1 - a32 q2
    V3<float> gh3 = normalize_adj(S.h3, S.n3, gq3);
    V3<float> gc3 = gh3;
    float ga31 = -dot(S.q1, gh3);
    float ga32 = -dot(S.q2, gh3);
    gq1 = gq1 - S.a31 * gh3;
    gq2 = gq2 - S.a32 * gh3;
    gc3 = gc3 + ga31 * S.q1 + ga32 * S.q2;
    gq1 = gq1 + ga31 * S.c3;
    gq2 = gq2 + ga32 * S.c3;
    // q2 = h2/n2, h2 = c2 - a21 q1
    V3<float> gh2 = normalize_adj(S.h2, S.n2, gq2);
    V3<float> gc2 = gh2;
    float ga21 = -dot(S.q1, gh2);
    gq1 = gq1 - S.a21 * gh2;
    gc2 = gc2 + ga21 * S.q1;
    gq1 = gq1 + ga21 * S.c2;
    // q1 = c1/n1
    V3<float> gc1 = normalize_adj(S.c1, S.n1, gq1);
    for (int i = 0; i < 3; ++i) {
        gF(i, 0) += gFd(i, 0) + gc1[i];
        gF(i, 1) += gFd(i, 1) + gc2[i];
        gF(i, 2) += gFd(i, 2) + gc3[i];
    }
}

struct StressState {
    QRState qr;
    float ft[4];          // Ft row-major a00 a01 a10 a11
    float pc, ps;         // polar rotation cos/sin
    float J, mu, la;
    float ftiT[4];        // inverse(Ft^T)
    M3<float> A, Ri, Mx;  // Mx = [P_fc 0;0 0] + A Ri^T  (P = Q Mx up to rounding order)
    float R22le1;
};

__device__ __forceinline__ void lame(float E, float nu, float& mu, float& la) {
    mu = E / (2.0f * (1.0f + nu));                                   // simulation.py:230
    la = E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));                // :230-232
}

template <typename PT>
__device__ __forceinline__ M3<float> stress_forward(const M3<float>& F, const ElemParams& prm, StressState& S) {
    qr_forward(F, S.qr);
    const M3<float>& R = S.qr.R;
    S.ft[0] = R(0, 0); S.ft[1] = R(0, 1); S.ft[2] = R(1, 0); S.ft[3] = R(1, 1);
    // A (:215-229)
    float R02 = R(0, 2), R12 = R(1, 2), R22 = R(2, 2);
    S.R22le1 = (R22 <= 1.0f) ? 1.0f : 0.0f;
    M3<float> A;
    A(0, 0) = prm.ks * R02 * R02;
    A(0, 1) = prm.ks * R02 * R12;
    A(0, 2) = prm.ks * R02 * R22;
    A(1, 1) = prm.ks * R12 * R12;
    A(1, 2) = prm.ks * R12 * R22;
    A(2, 2) = (S.R22le1 * (-prm.kc) * ((1.0f - R22) * (1.0f - R22))) * R22;
    A(1, 0) = A(0, 1); A(2, 0) = A(0, 2); A(2, 1) = A(1, 2);
    S.A = A;
    lame(prm.E, prm.nu, S.mu, S.la);
    S.J = S.ft[0] * S.ft[3] - S.ft[1] * S.ft[2];
    polar2<PT>(S.ft[0], S.ft[1], S.ft[2], S.ft[3], S.pc, S.ps);
    // inverse(Ft^T) with Ft^T = [[a00, a10],[a01, a11]]
    float det = S.ft[0] * S.ft[3] - S.ft[2] * S.ft[1];
    S.ftiT[0] = S.ft[3] / det;  S.ftiT[1] = -S.ft[2] / det;
    S.ftiT[2] = -S.ft[1] / det; S.ftiT[3] = S.ft[0] / det;
    float rr[4] = {S.pc, -S.ps, S.ps, S.pc};
    float jj = S.la * (S.J - 1.0f) * S.J;
    M3<float> Pf = mzero<float>();
    Pf(0, 0) = 2.0f * S.mu * (S.ft[0] - rr[0]) + jj * S.ftiT[0];
    Pf(0, 1) = 2.0f * S.mu * (S.ft[1] - rr[1]) + jj * S.ftiT[1];
    Pf(1, 0) = 2.0f * S.mu * (S.ft[2] - rr[2]) + jj * S.ftiT[2];
    Pf(1, 1) = 2.0f * S.mu * (S.ft[3] - rr[3]) + jj * S.ftiT[3];
    S.Ri = minv(R);
    M3<float> RiT = mT(S.Ri);
    M3<float> P = madd(mmul(S.qr.Q, Pf), mmul(mmul(S.qr.Q, A), RiT));
    S.Mx = madd(Pf, mmul(A, RiT));
    return P;
}

// Given dL/dP accumulate dL/dF and the parameter gradients (E, nu, ks, kc).
template <typename PT>
__device__ __forceinline__ void stress_adjoint(const M3<float>& F, const ElemParams& prm, const StressState& S,
                                               const M3<float>& gP, int svd_mode, M3<float>& gF, float gparam[4]) {
    const M3<float>& Q = S.qr.Q;
    const M3<float>& R = S.qr.R;
    // P = Q Mx
    M3<float> gQ = mmul(gP, mT(S.Mx));
    M3<float> gM = mmul(mT(Q), gP);
    // Mx = Pf + A Ri^T
    M3<float> gA = mmul(gM, S.Ri);                 // d/dA of <gM, A Ri^T> = gM Ri
    M3<float> gRi = mmul(mT(gM), S.A);             // d/dRi = (A^T gM)^T = gM^T A
    M3<float> gR = mscale(-1.0f, mmul(mmul(mT(S.Ri), gRi), mT(S.Ri)));
    // A(R02, R12, R22; ks, kc)
    float R02 = R(0, 2), R12 = R(1, 2), R22 = R(2, 2);
    float g00 = gA(0, 0), g11 = gA(1, 1), g22 = gA(2, 2);
    float g01 = gA(0, 1) + gA(1, 0), g02 = gA(0, 2) + gA(2, 0), g12 = gA(1, 2) + gA(2, 1);
    gR(0, 2) += prm.ks * (2.0f * R02 * g00 + R12 * g01 + R22 * g02);
    gR(1, 2) += prm.ks * (R02 * g01 + 2.0f * R12 * g11 + R22 * g12);
    float om = 1.0f - R22;
    gR(2, 2) += prm.ks * (R02 * g02 + R12 * g12) + g22 * S.R22le1 * (-prm.kc) * (om * om - 2.0f * om * R22);
    float gks = g00 * R02 * R02 + g01 * R02 * R12 + g02 * R02 * R22 + g11 * R12 * R12 + g12 * R12 * R22;
    float gkc = g22 * S.R22le1 * (-(om * om) * R22);
    // P_fc
    float gPf[4] = {gM(0, 0), gM(0, 1), gM(1, 0), gM(1, 1)};
    float rr[4] = {S.pc, -S.ps, S.ps, S.pc};
    float gmu = 0.0f, sfi = 0.0f;
    for (int i = 0; i < 4; ++i) { gmu += 2.0f * gPf[i] * (S.ft[i] - rr[i]); sfi += gPf[i] * S.ftiT[i]; }
    float gla = (S.J - 1.0f) * S.J * sfi;
    float gJ = S.la * sfi * (2.0f * S.J - 1.0f);
    float jj = S.la * (S.J - 1.0f) * S.J;
    float gFt[4];
    for (int i = 0; i < 4; ++i) gFt[i] = 2.0f * S.mu * gPf[i];
    // d inverse(B) with B = Ft^T: gB = -X^T gX X^T, X = ftiT; gFt += gB^T
    {
        float X[4] = {S.ftiT[0], S.ftiT[1], S.ftiT[2], S.ftiT[3]};
        float gX[4] = {jj * gPf[0], jj * gPf[1], jj * gPf[2], jj * gPf[3]};
        // T1 = X^T gX
        float T1[4] = {X[0] * gX[0] + X[2] * gX[2], X[0] * gX[1] + X[2] * gX[3],
                       X[1] * gX[0] + X[3] * gX[2], X[1] * gX[1] + X[3] * gX[3]};
        // gB = -(T1 X^T)
        float gB[4] = {-(T1[0] * X[0] + T1[1] * X[1]), -(T1[0] * X[2] + T1[1] * X[3]),
                       -(T1[2] * X[0] + T1[3] * X[1]), -(T1[2] * X[2] + T1[3] * X[3])};
        gFt[0] += gB[0]; gFt[1] += gB[2]; gFt[2] += gB[1]; gFt[3] += gB[3];
    }
    // J = a00 a11 - a01 a10
    gFt[0] += gJ * S.ft[3]; gFt[3] += gJ * S.ft[0];
    gFt[1] -= gJ * S.ft[2]; gFt[2] -= gJ * S.ft[1];
    // polar rotation
    float gRr[4] = {-2.0f * S.mu * gPf[0], -2.0f * S.mu * gPf[1], -2.0f * S.mu * gPf[2], -2.0f * S.mu * gPf[3]};
    polar2_adj<PT>(S.ft[0], S.ft[1], S.ft[2], S.ft[3], gRr, svd_mode, gFt);
    gR(0, 0) += gFt[0]; gR(0, 1) += gFt[1]; gR(1, 0) += gFt[2]; gR(1, 1) += gFt[3];
    qr_adjoint(F, S.qr, gQ, gR, gF);
    // Lame parameters
    float opn = 1.0f + prm.nu, omn = 1.0f - 2.0f * prm.nu;
    float dmu_dE = 1.0f / (2.0f * opn);
    float dmu_dnu = -prm.E / (2.0f * opn * opn);
    float dla_dE = prm.nu / (opn * omn);
    float dla_dnu = prm.E * (1.0f + 2.0f * prm.nu * prm.nu) / ((opn * omn) * (opn * omn));
    gparam[0] += gmu * dmu_dE + gla * dla_dE;
    gparam[1] += gmu * dmu_dnu + gla * dla_dnu;
    gparam[2] += gks;
    gparam[3] += gkc;
}

// ---------------------------------------------------------------------------
// Live return mapping (optional; simulation.py:408-428 computes it with
// cf = 0.1 and never writes it back, so it is off by default).  On the R of
// the Gram-Schmidt QR of the new F, only the third column changes:
//   ks == 0:       R02 = R12 = 1e-7, R22 = min(R22, 1)
//   else R22 > 1:  R02 = R12 = 1e-7, R22 = 1
//   else R22 < 0:  R02 = R12 = 1e-7, R22 = max(R22, -1)
//   else R22 <= 1: with shear = R02^2 + R12^2, k = ks / kc, when
//                  k^2 shear - (cf (R22-1)^2)^2 > 0 the pair (R02, R12) is
//                  scaled by cf (R22-1)^2 / (k sqrt(shear))
// and F becomes Q R.  cf is the friction parameter.
// ---------------------------------------------------------------------------
struct RMState {
    QRState qr;
    M3<float> Rm;       // mapped R
    int branch;         // 0 unchanged, 1 ks == 0, 2 R22 > 1, 3 R22 < 0, 4 shear scaled
    float k, sq, scale;
};

__device__ __forceinline__ M3<float> return_map_forward(const M3<float>& F, float ks, float kc, float cf,
                                                        RMState& S) {
    qr_forward(F, S.qr);
    M3<float> R = S.qr.R;
    const float R02 = R(0, 2), R12 = R(1, 2), R22 = R(2, 2);
    S.branch = 0;
    S.k = S.sq = S.scale = 0.f;
    if (ks == 0.f) {
        R(0, 2) = 1e-7f; R(1, 2) = 1e-7f; R(2, 2) = fminf(R22, 1.f);
        S.branch = 1;
    } else if (R22 > 1.f) {
        R(0, 2) = 1e-7f; R(1, 2) = 1e-7f; R(2, 2) = 1.f;
        S.branch = 2;
    } else if (R22 < 0.f) {
        R(0, 2) = 1e-7f; R(1, 2) = 1e-7f; R(2, 2) = fmaxf(R22, -1.f);
        S.branch = 3;
    } else if (R22 <= 1.f) {
        const float shear = R02 * R02 + R12 * R12;
        const float k = ks / kc;
        const float om = R22 - 1.f;
        const float c2 = cf * (om * om);
        if ((k * k) * shear - c2 * c2 > 0.f) {
            S.sq = sqrtf(shear);
            S.k = k;
            S.scale = c2 / (k * S.sq);
            R(0, 2) = R02 * S.scale;
            R(1, 2) = R12 * S.scale;
            S.branch = 4;
        }
    }
    S.Rm = R;
    return mmul(S.qr.Q, R);
}

// Given dL/dFm of Fm = Q Rm, accumulate dL/dF and the gradients of ks, kc, cf.
__device__ __forceinline__ void return_map_adjoint(const M3<float>& F, const RMState& S, float ks, float kc, float cf,
                                                   const M3<float>& gFm, M3<float>& gF, float& gks, float& gkc,
                                                   float& gcf) {
    const M3<float> gQ = mmul(gFm, mT(S.Rm));
    M3<float> gR = mmul(mT(S.qr.Q), gFm);    // identity on the entries the mapping keeps
    const float R02 = S.qr.R(0, 2), R12 = S.qr.R(1, 2), R22 = S.qr.R(2, 2);
    if (S.branch == 1) {
        gR(0, 2) = 0.f; gR(1, 2) = 0.f;
        if (!(R22 < 1.f)) gR(2, 2) = 0.f;
    } else if (S.branch == 2) {
        gR(0, 2) = 0.f; gR(1, 2) = 0.f; gR(2, 2) = 0.f;
    } else if (S.branch == 3) {
        gR(0, 2) = 0.f; gR(1, 2) = 0.f;
        if (!(R22 > -1.f)) gR(2, 2) = 0.f;
    } else if (S.branch == 4) {
        // n = scale u, u = (R02, R12), scale = cf (R22-1)^2 / (k |u|)
        const float g0 = gR(0, 2), g1 = gR(1, 2);
        const float gu = g0 * R02 + g1 * R12;               // g . u
        const float inv_sq2 = 1.f / (S.sq * S.sq);
        gR(0, 2) = S.scale * (g0 - R02 * gu * inv_sq2);
        gR(1, 2) = S.scale * (g1 - R12 * gu * inv_sq2);
        const float om = R22 - 1.f;
        const float ksq = S.k * S.sq;
        gR(2, 2) += gu * (2.f * cf * om) / ksq;
        gcf += gu * (om * om) / ksq;
        const float gk = -(gu * S.scale) / S.k;            // d n / d k = -n / k
        gks += gk / kc;
        gkc += -gk * ks / (kc * kc);
    }
    qr_adjoint(F, S.qr, gQ, gR, gF);
}
