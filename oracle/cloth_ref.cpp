// FP32 C++/OpenMP restatement of DiffCP's cloth substep and loss, with the
// gradient w.r.t. E and nu by forward-mode dual numbers.
//
// TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline): only tests/
// and bench.py's cpu_baseline leg load it.  The product library never links
// or calls it.
//
// It follows the reference kernel by kernel, in FP32 (the reference's
// default_fp, initialization.py:20), compiled with -ffp-contract=off:
//   clear_grid                 simulation.py:9-16   (the box of cells this substep's stencils reach)
//   compute_inplane_F_tmp      :24-28               (Gram-Schmidt QR, eps 1e-9)
//   inplane_svd                :39-42               (polar rotation U V^T, closed form)
//   p2g                        :137-271             (four classes, per-thread copies of the box)
//   grid_op                    :274-302             (over the box, walls; cells outside it stay empty)
//   g2p                        :304-383
//   update_particle_state      :386-430             (return mapping dead: left out)
//   compute_loss               parameter_fitting_functions.py:3-15 (first argmin)
// Every state value carries two tangents, d/dE and d/dnu (the only inputs
// seeded), so one forward pass yields the loss and both gradient components
// (the reference gets them from its tape; the values are the same quantity).
// The grid mass depends on E and nu through the particle positions in the
// weights, so it carries tangents too.
//
// Build: oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off -shared).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

// value + tangents (d/dE, d/dnu), FP32
struct D {
    float v, a, b;
    D() : v(0.f), a(0.f), b(0.f) {}
    D(float x) : v(x), a(0.f), b(0.f) {}
    D(float x, float da, float db) : v(x), a(da), b(db) {}
};
inline D operator+(D x, D y) { return {x.v + y.v, x.a + y.a, x.b + y.b}; }
inline D operator-(D x, D y) { return {x.v - y.v, x.a - y.a, x.b - y.b}; }
inline D operator-(D x) { return {-x.v, -x.a, -x.b}; }
inline D operator*(D x, D y) { return {x.v * y.v, x.a * y.v + x.v * y.a, x.b * y.v + x.v * y.b}; }
inline D operator/(D x, D y) {
    const float q = x.v / y.v;
    return {q, (x.a - q * y.a) / y.v, (x.b - q * y.b) / y.v};
}
inline D dsqrt(D x) {
    const float r = std::sqrt(x.v);
    return {r, x.a / (2.f * r), x.b / (2.f * r)};
}
inline D& operator+=(D& x, D y) { x = x + y; return x; }

struct V { D x, y, z; };
inline V operator+(V p, V q) { return {p.x + q.x, p.y + q.y, p.z + q.z}; }
inline V operator-(V p, V q) { return {p.x - q.x, p.y - q.y, p.z - q.z}; }
inline V operator*(D s, V p) { return {s * p.x, s * p.y, s * p.z}; }
inline D dot(V p, V q) { return p.x * q.x + p.y * q.y + p.z * q.z; }
inline D& comp(V& p, int i) { return i == 0 ? p.x : (i == 1 ? p.y : p.z); }
inline D comp(const V& p, int i) { return i == 0 ? p.x : (i == 1 ? p.y : p.z); }

struct M { D m[3][3]; };
inline M mmul(const M& a, const M& b) {
    M r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}
inline V mv(const M& a, const V& p) {
    return {a.m[0][0] * p.x + a.m[0][1] * p.y + a.m[0][2] * p.z, a.m[1][0] * p.x + a.m[1][1] * p.y + a.m[1][2] * p.z,
            a.m[2][0] * p.x + a.m[2][1] * p.y + a.m[2][2] * p.z};
}
inline M mT(const M& a) {
    M r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}
// general 3x3 inverse by cofactors (Taichi Matrix.inverse)
inline M minv(const M& a) {
    const auto& m = a.m;
    D c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    D c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    D c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    D det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    D id = D(1.f) / det;
    M r;
    r.m[0][0] = c00 * id;
    r.m[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    r.m[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    r.m[1][0] = c01 * id;
    r.m[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    r.m[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
    r.m[2][0] = c02 * id;
    r.m[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
    r.m[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
    return r;
}
inline V col(const M& a, int j) { return {a.m[0][j], a.m[1][j], a.m[2][j]}; }
inline void set_col(M& a, int j, const V& c) { a.m[0][j] = c.x; a.m[1][j] = c.y; a.m[2][j] = c.z; }

// Gram-Schmidt with norm(eps) = sqrt(v.v + 1e-9), R = inverse(Q) F (simulation.py:123-135)
inline void qr3(const M& F, M& Q, M& R) {
    V c1 = col(F, 0), c2 = col(F, 1), c3 = col(F, 2);
    V q1 = (D(1.f) / dsqrt(dot(c1, c1) + D(1e-9f))) * c1;
    V h2 = c2 - dot(c2, q1) * q1;
    V q2 = (D(1.f) / dsqrt(dot(h2, h2) + D(1e-9f))) * h2;
    V h3 = c3 - dot(c3, q1) * q1 - dot(c3, q2) * q2;
    V q3 = (D(1.f) / dsqrt(dot(h3, h3) + D(1e-9f))) * h3;
    set_col(Q, 0, q1); set_col(Q, 1, q2); set_col(Q, 2, q3);
    R = mmul(minv(Q), F);
}

struct Stencil {
    int base[3];
    float fx[3];            // plain: positions' tangents enter through D fxd
    D fxd[3];
    D w[3][3];              // [node][dim]
    D dw[3][3];             // d w / d fx
};
// set when a stencil leaves the grid (clamped, so that no access leaves the arrays)
int g_out_of_grid = 0;
int g_grid_side = 0;

// base = trunc(x * inv_dx - 0.5) (ti.cast to i32), simulation.py:140
inline void stencil(const V& x, float inv_dx, Stencil& s) {
    for (int d = 0; d < 3; ++d) {
        D xs = comp(x, d) * D(inv_dx);
        int b = (int)(xs.v - 0.5f);
        if (b < 0 || b > g_grid_side - 3 || !(xs.v == xs.v)) {
            g_out_of_grid = 1;
            b = b < 0 || !(xs.v == xs.v) ? 0 : g_grid_side - 3;
        }
        s.base[d] = b;
        D f = xs - D((float)b);
        s.fxd[d] = f;
        s.fx[d] = f.v;
        D t0 = D(1.5f) - f, t1 = f - D(1.f), t2 = f - D(0.5f);
        s.w[0][d] = D(0.5f) * t0 * t0;
        s.w[1][d] = D(0.75f) - t1 * t1;
        s.w[2][d] = D(0.5f) * t2 * t2;
        s.dw[0][d] = f - D(1.5f);
        s.dw[1][d] = D(2.f) * (D(1.f) - f);
        s.dw[2][d] = f - D(0.5f);
    }
}

struct Scene {
    int G, nv, ne, nc, no, n_grasp, n_actions, steps_per_frame, frames_per_window;
    float dt, dx, inv_dx, gravity, m_v, m_q, vol_q, m_c, m_o, damping;
    std::vector<int> tri;
    std::vector<float> traj;               // [n_actions][n_grasp][3]
    std::vector<M> dinv, dinvinv;          // rest-shape inverse and its inverse (the d <- F Dinv^-1 round trip)
};

struct Particles {
    std::vector<V> x, v;
    std::vector<M> C;
    void resize(int n) { x.assign(n, V{}); v.assign(n, V{}); C.assign(n, M{}); }
};

struct State {
    Particles vert, grasp, cube, elem;
    std::vector<M> F, d;
};

struct Grid {
    int G;
    std::vector<float> m;                  // [cell][3]: value, d/dE, d/dnu
    std::vector<float> mv;                 // [cell][3][3]: value, d/dE, d/dnu per component
    std::vector<D> vout;                   // [cell][3]
    void init(int g) {
        G = g;
        m.assign((size_t)g * g * g * 3, 0.f);
        mv.assign((size_t)g * g * g * 9, 0.f);
        vout.assign((size_t)g * g * g * 3, D());
    }
    size_t idx(int i, int j, int k) const { return ((size_t)i * G + j) * G + k; }
    // P2G accumulates into one private copy per thread of the box of cells
    // the particles' stencils reach (12 floats per cell: m, then mv), summed
    // into the dense grid in thread order afterwards: no atomics, and the
    // same sums for the same thread count
    int lo[3] = {0, 0, 0}, ext[3] = {1, 1, 1};
    std::vector<std::vector<float>> priv;
    void begin(const int l[3], const int h[3], int threads) {
        for (int d = 0; d < 3; ++d) { lo[d] = l[d]; ext[d] = h[d] - l[d] + 1; }
        const size_t n = (size_t)ext[0] * ext[1] * ext[2] * 12;
        priv.resize(threads);
#pragma omp parallel for schedule(static)
        for (int t = 0; t < threads; ++t) priv[t].assign(n, 0.f);
    }
    float* slot(int i, int j, int k) {
        const size_t l = (((size_t)(i - lo[0]) * ext[1] + (j - lo[1])) * ext[2] + (k - lo[2])) * 12;
        return &priv[omp_get_thread_num()][l];
    }
    static void add_d(float* q, const D& x) { q[0] += x.v; q[1] += x.a; q[2] += x.b; }
    void add(int i, int j, int k, const V& p) {
        float* q = slot(i, j, k) + 3;
        add_d(q, p.x);
        add_d(q + 3, p.y);
        add_d(q + 6, p.z);
    }
    void add(int i, int j, int k, const V& p, const D& mass) {
        float* q = slot(i, j, k);
        add_d(q, mass);
        add_d(q + 3, p.x);
        add_d(q + 6, p.y);
        add_d(q + 9, p.z);
    }
    void merge() {
        const int n = ext[0] * ext[1] * ext[2];
#pragma omp parallel for schedule(static)
        for (int l = 0; l < n; ++l) {
            const int k = l % ext[2], t = l / ext[2], j = t % ext[1], i = t / ext[1];
            const size_t c = idx(lo[0] + i, lo[1] + j, lo[2] + k);
            float s[12] = {};
            for (size_t th = 0; th < priv.size(); ++th) {
                const float* q = &priv[th][(size_t)l * 12];
                for (int a = 0; a < 12; ++a) s[a] += q[a];
            }
            for (int a = 0; a < 3; ++a) m[3 * c + a] = s[a];
            for (int a = 0; a < 9; ++a) mv[9 * c + a] = s[3 + a];
        }
    }
};

// one particle's scatter: w (m vel + affine dpos) and w m (simulation.py:139-191)
void scatter_point(Grid& g, const Scene& S, const V& x, const V* vel, const M& C, float mass) {
    Stencil st;
    stencil(x, S.inv_dx, st);
    M affine;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) affine.m[i][j] = D(mass) * C.m[i][j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                const int o[3] = {i, j, k};
                V dpos;
                for (int d = 0; d < 3; ++d) comp(dpos, d) = (D((float)o[d]) - st.fxd[d]) * D(S.dx);
                D weight = st.w[i][0] * st.w[j][1] * st.w[k][2];
                V mom = mv(affine, dpos);
                if (vel) mom = (D(mass) * *vel) + mom;
                g.add(st.base[0] + i, st.base[1] + j, st.base[2] + k, weight * mom, weight * D(mass));
            }
}

void substep(const Scene& S, const State& in, State& out, Grid& g, int f, int gw, D E, D nu, float ks, float kc) {
    const float dt = S.dt, dx = S.dx, inv_dx = S.inv_dx;
    // clear_grid (:9-16): merge() below assigns every cell of the box, the
    // only cells P2G writes and G2P reads this substep
    // the box of cells every stencil of this substep reaches
    {
        int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-1, -1, -1};
        auto take = [&](const std::vector<V>& xs, int n) {
            for (int p = 0; p < n; ++p) {
                Stencil st;
                stencil(xs[p], inv_dx, st);
                for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], st.base[d]); hi[d] = std::max(hi[d], st.base[d] + 2); }
            }
        };
        take(in.cube.x, S.nc); take(in.grasp.x, S.no); take(in.vert.x, S.nv); take(in.elem.x, S.ne);
        g.begin(lo, hi, omp_get_max_threads());
    }
    // p2g: cube (:139-152)
#pragma omp parallel for schedule(static)
    for (int p = 0; p < S.nc; ++p) scatter_point(g, S, in.cube.x[p], nullptr, in.cube.C[p], S.m_c);
    // grasper (:154-174), frame index as the integer f / steps_per_frame
    const int frame = f / S.steps_per_frame + S.frames_per_window * gw;
    const int group = S.n_grasp > 0 ? S.no / S.n_grasp : 1;
    const float interval = (float)S.steps_per_frame * dt;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < S.no; ++p) {
        V vel{};
        if (frame + 1 < S.n_actions) {
            const int gid = p / group;
            const float* a = &S.traj[3 * ((size_t)(frame + 1) * S.n_grasp + gid)];
            const float* b = &S.traj[3 * ((size_t)frame * S.n_grasp + gid)];
            vel = {D((a[0] - b[0]) / interval), D((a[1] - b[1]) / interval), D((a[2] - b[2]) / interval)};
        }
        scatter_point(g, S, in.grasp.x[p], &vel, in.grasp.C[p], S.m_o);
    }
    // vertices (:176-191)
#pragma omp parallel for schedule(static)
    for (int p = 0; p < S.nv; ++p) {
        V vel = in.vert.v[p];
        vel.y = vel.y + D(dt * S.gravity);
        scatter_point(g, S, in.vert.x[p], &vel, in.vert.C[p], S.m_v);
    }
    // elements (:193-271): stress, own scatter, normal impulse, three vertex stencils
    const D mu = E / (D(2.f) * (D(1.f) + nu));
    const D la = E * nu / ((D(1.f) + nu) * (D(1.f) - D(2.f) * nu));
#pragma omp parallel for schedule(static)
    for (int e = 0; e < S.ne; ++e) {
        M Q, R;
        qr3(in.F[e], Q, R);
        // in-plane F and its polar rotation U V^T (closed form of the 2x2 SVD)
        D a00 = R.m[0][0], a01 = R.m[0][1], a10 = R.m[1][0], a11 = R.m[1][1];
        D px = a00 + a11, py = a10 - a01;
        D rr = dsqrt(px * px + py * py);
        D c = px / rr, s = py / rr;
        // A (:215-229): the E term of A[0,0] is overwritten
        M A;
        D R02 = R.m[0][2], R12 = R.m[1][2], R22 = R.m[2][2];
        A.m[0][0] = D(ks) * R02 * R02;
        A.m[0][1] = D(ks) * R02 * R12;
        A.m[0][2] = D(ks) * R02 * R22;
        A.m[1][1] = D(ks) * R12 * R12;
        A.m[1][2] = D(ks) * R12 * R22;
        D om = D(1.f) - R22;
        A.m[2][2] = (D(R22.v <= 1.f ? 1.f : 0.f) * D(-kc) * (om * om)) * R22;
        A.m[1][0] = A.m[0][1]; A.m[2][0] = A.m[0][2]; A.m[2][1] = A.m[1][2];
        // P_fc = 2 mu (Ft - R) + la (J - 1) J inverse(Ft^T) (:234-237)
        D J = a00 * a11 - a01 * a10;
        D jj = la * (J - D(1.f)) * J;
        D det = a00 * a11 - a10 * a01;
        M Pf;
        Pf.m[0][0] = D(2.f) * mu * (a00 - c) + jj * (a11 / det);
        Pf.m[0][1] = D(2.f) * mu * (a01 + s) + jj * (-a10 / det);
        Pf.m[1][0] = D(2.f) * mu * (a10 - s) + jj * (-a01 / det);
        Pf.m[1][1] = D(2.f) * mu * (a11 - c) + jj * (a00 / det);
        M P = mmul(Q, Pf);
        M QART = mmul(mmul(Q, A), mT(minv(R)));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) P.m[i][j] = P.m[i][j] + QART.m[i][j];
        const M& dd = in.d[e];
        V Pc3 = col(P, 2), dc3 = col(dd, 2);
        const int* t = &S.tri[3 * (size_t)e];
        // raw vectors mix the rest-shape inverse with the current d (:247-248)
        V raw1{S.dinv[e].m[0][0], dd.m[0][1], dd.m[0][2]};
        V raw2{S.dinv[e].m[1][0], dd.m[1][1], dd.m[1][2]};
        V f1 = mv(P, raw1 + raw2), f2 = mv(P, raw1), f3 = mv(P, raw2);
        Stencil st;
        stencil(in.elem.x[e], inv_dx, st);
        Stencil vs[3];
        for (int q = 0; q < 3; ++q) stencil(in.vert.x[t[q]], inv_dx, vs[q]);
        M affine;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) affine.m[i][j] = in.elem.C[e].m[i][j] * D(S.m_q);
        const D vol(S.vol_q), tdt(dt);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    const int o[3] = {i, j, k};
                    V gw{st.dw[i][0] * D(inv_dx) * st.w[j][1] * st.w[k][2],
                         st.w[i][0] * st.dw[j][1] * D(inv_dx) * st.w[k][2],
                         st.w[i][0] * st.w[j][1] * st.dw[k][2] * D(inv_dx)};
                    D weight = st.w[i][0] * st.w[j][1] * st.w[k][2];
                    V dpos;
                    for (int d = 0; d < 3; ++d) comp(dpos, d) = (D((float)o[d]) - st.fxd[d]) * D(dx);
                    const int ci = st.base[0] + i, cj = st.base[1] + j, ck = st.base[2] + k;
                    g.add(ci, cj, ck, weight * ((D(S.m_q) * in.elem.v[e]) + mv(affine, dpos)), weight * D(S.m_q));
                    g.add(ci, cj, ck, ((-vol) * dot(gw, dc3) * tdt) * Pc3);
                    D w1 = vs[0].w[i][0] * vs[0].w[j][1] * vs[0].w[k][2];
                    D w2 = vs[1].w[i][0] * vs[1].w[j][1] * vs[1].w[k][2];
                    D w3 = vs[2].w[i][0] * vs[2].w[j][1] * vs[2].w[k][2];
                    g.add(vs[0].base[0] + i, vs[0].base[1] + j, vs[0].base[2] + k, (vol * w1 * tdt) * f1);
                    g.add(vs[1].base[0] + i, vs[1].base[1] + j, vs[1].base[2] + k, ((-vol) * w2 * tdt) * f2);
                    g.add(vs[2].base[0] + i, vs[2].base[1] + j, vs[2].base[2] + k, ((-vol) * w3 * tdt) * f3);
                }
    }
    g.merge();
    // grid_op (:274-302) over the box
    const int G = S.G, bound = 3;
    const float damp = 1.0f - S.damping * S.dt;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = g.lo[0]; i < g.lo[0] + g.ext[0]; ++i)
        for (int j = g.lo[1]; j < g.lo[1] + g.ext[1]; ++j)
            for (int k = g.lo[2]; k < g.lo[2] + g.ext[2]; ++k) {
                const size_t c = g.idx(i, j, k);
                const float* q = &g.mv[c * 9];
                const float* mm = &g.m[c * 3];
                D inv_m = D(1.f) / (D(mm[0], mm[1], mm[2]) + D(1e-9f));
                D v[3];
                for (int a = 0; a < 3; ++a) v[a] = (inv_m * D(q[3 * a], q[3 * a + 1], q[3 * a + 2])) * D(damp);
                if (i < bound && v[0].v < 0) v[0] = v[1] = v[2] = D();
                if (i > G - bound && v[0].v > 0) v[0] = v[1] = v[2] = D();
                if (k < bound && v[2].v < 0) v[0] = v[1] = v[2] = D();
                if (k > G - bound && v[2].v > 0) v[0] = v[1] = v[2] = D();
                if (j < bound && v[1].v < 0) v[1] = D();
                if (j > G - bound && v[1].v > 0) v[0] = v[1] = v[2] = D();
                for (int a = 0; a < 3; ++a) g.vout[3 * c + a] = v[a];
            }
    // g2p (:304-383)
    auto gather = [&](const V& x, V& nv, M& nC) {
        Stencil st;
        stencil(x, inv_dx, st);
        nv = V{};
        nC = M{};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    const int o[3] = {i, j, k};
                    const size_t c = g.idx(st.base[0] + i, st.base[1] + j, st.base[2] + k);
                    V gv{g.vout[3 * c], g.vout[3 * c + 1], g.vout[3 * c + 2]};
                    D weight = st.w[i][0] * st.w[j][1] * st.w[k][2];
                    nv = nv + weight * gv;
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b)
                            nC.m[a][b] = nC.m[a][b] + D(4.f) * weight * comp(gv, a) * (D((float)o[b]) - st.fxd[b]) * D(inv_dx);
                }
    };
    auto g2p_class = [&](const Particles& pin, Particles& pout, int n) {
#pragma omp parallel for schedule(static)
        for (int p = 0; p < n; ++p) {
            V nv;
            M nC;
            gather(pin.x[p], nv, nC);
            pout.v[p] = nv;
            pout.x[p] = pin.x[p] + D(dt) * nv;
            pout.C[p] = nC;
        }
    };
    g2p_class(in.cube, out.cube, S.nc);
    g2p_class(in.grasp, out.grasp, S.no);
    g2p_class(in.vert, out.vert, S.nv);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < S.ne; ++e) {
        V nv;
        M nC;
        gather(in.elem.x[e], nv, nC);
        M Cn;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                D skew = D(0.5f) * (nC.m[a][b] - nC.m[b][a]);
                D sym = D(0.5f) * (nC.m[a][b] + nC.m[b][a]);
                Cn.m[a][b] = skew + D((float)(1.0 - 0.95)) * sym;
            }
        out.elem.C[e] = Cn;
    }
    // update_particle_state (:386-430)
#pragma omp parallel for schedule(static)
    for (int e = 0; e < S.ne; ++e) {
        const int* t = &S.tri[3 * (size_t)e];
        const D third(1.f / 3.f);
        out.elem.v[e] = third * (out.vert.v[t[0]] + out.vert.v[t[1]] + out.vert.v[t[2]]);
        out.elem.x[e] = third * (out.vert.x[t[0]] + out.vert.x[t[1]] + out.vert.x[t[2]]);
        V c3 = col(in.d[e], 2);
        const M& C = out.elem.C[e];
        M Mx;
        const M CC = mmul(C, C);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) Mx.m[a][b] = D(dt) * C.m[a][b] + D(0.5f * dt * dt) * CC.m[a][b];
        c3 = c3 + mv(Mx, c3);
        M nd;
        set_col(nd, 0, out.vert.x[t[1]] - out.vert.x[t[0]]);
        set_col(nd, 1, out.vert.x[t[2]] - out.vert.x[t[0]]);
        set_col(nd, 2, c3);
        M Fn = mmul(nd, S.dinv[e]);
        out.F[e] = Fn;
        out.d[e] = mmul(Fn, S.dinvinv[e]);
    }
}

}  // namespace

extern "C" {

// One scene (the layout of the arrays of the HIP library's instance
// description, natural order, FP64 host values).
typedef struct cref_scene {
    int32_t n_grid, n_vertices, n_elements, n_cube, n_grasper, n_grasp_points, n_actions, steps_per_frame,
        frames_per_window, reserved;
    double dt, gravity, vertex_p_mass, quadrature_p_mass, quadrature_p_vol, cube_p_mass, operator_p_mass;
    double damping;                        // grid velocity damping (0: the reference's grid_op)
    const double *vertex_x;
    const int32_t* triangles;
    const double *cube_x, *grasper_x, *trajectory;
} cref_scene;

// Loss on the state after n_substeps - 2 substeps of window 0 from the rest
// state (parameter_fitting_functions.py:17-38) and its gradient w.r.t. E and
// nu; params = (E, nu, shearing, contact).  threads <= 0: OpenMP default.
// Returns 0, -1 on bad arguments, -2 when a stencil left the grid (clamped).
int cref_loss_grad(const cref_scene* d, const double* params, const double* targets, int32_t n_targets,
                   int32_t n_substeps, int32_t threads, double* loss_out, double* grad_out, double* seconds,
                   float* final_vertex_x) {
    if (!d || !params || !targets || n_targets <= 0 || n_substeps < 3) return -1;
    if (threads > 0) omp_set_num_threads(threads);
    Scene S;
    S.G = d->n_grid; S.nv = d->n_vertices; S.ne = d->n_elements; S.nc = d->n_cube; S.no = d->n_grasper;
    S.n_grasp = d->n_grasp_points; S.n_actions = d->n_actions; S.steps_per_frame = d->steps_per_frame;
    S.frames_per_window = d->frames_per_window;
    S.dt = (float)d->dt; S.dx = (float)(1.0 / d->n_grid); S.inv_dx = (float)d->n_grid; S.gravity = (float)d->gravity;
    S.m_v = (float)d->vertex_p_mass; S.m_q = (float)d->quadrature_p_mass; S.vol_q = (float)d->quadrature_p_vol;
    S.m_c = (float)d->cube_p_mass; S.m_o = (float)d->operator_p_mass; S.damping = (float)d->damping;
    S.tri.assign(d->triangles, d->triangles + 3 * (size_t)S.ne);
    S.traj.resize(3 * (size_t)S.n_actions * S.n_grasp);
    for (size_t i = 0; i < S.traj.size(); ++i) S.traj[i] = (float)d->trajectory[i];
    State st[2];
    for (auto& s : st) {
        s.vert.resize(S.nv); s.grasp.resize(S.no); s.cube.resize(S.nc); s.elem.resize(S.ne);
        s.F.assign(S.ne, M{}); s.d.assign(S.ne, M{});
    }
    auto load = [](Particles& P, const double* x, int n) {
        for (int i = 0; i < n; ++i) P.x[i] = {D((float)x[3 * i]), D((float)x[3 * i + 1]), D((float)x[3 * i + 2])};
    };
    load(st[0].vert, d->vertex_x, S.nv);
    load(st[0].grasp, d->grasper_x, S.no);
    load(st[0].cube, d->cube_x, S.nc);
    // initialize_ti_field (simulation.py:447-477)
    std::vector<M>& Dinv = S.dinv;
    std::vector<M>& Dinvinv = S.dinvinv;
    Dinv.resize(S.ne);
    Dinvinv.resize(S.ne);
    for (int e = 0; e < S.ne; ++e) {
        const int* t = &S.tri[3 * (size_t)e];
        V x0 = st[0].vert.x[t[0]], x1 = st[0].vert.x[t[1]], x2 = st[0].vert.x[t[2]];
        V d1 = x1 - x0, d2 = x2 - x0;
        V n1 = (D(1.f) / dsqrt(dot(d1, d1) + D(1e-6f))) * d1;
        V n2 = (D(1.f) / dsqrt(dot(d2, d2) + D(1e-6f))) * d2;
        V c3{n1.y * n2.z - n1.z * n2.y, n1.z * n2.x - n1.x * n2.z, n1.x * n2.y - n1.y * n2.x};
        V n3 = (D(1.f) / dsqrt(dot(c3, c3) + D(1e-6f))) * c3;
        M dd;
        set_col(dd, 0, d1); set_col(dd, 1, d2); set_col(dd, 2, n3);
        st[0].d[e] = dd;
        Dinv[e] = minv(dd);
        Dinvinv[e] = minv(Dinv[e]);
        M I{};
        I.m[0][0] = I.m[1][1] = I.m[2][2] = D(1.f);
        st[0].F[e] = I;
        st[0].elem.x[e] = D(1.f / 3.f) * (x0 + x1 + x2);
    }
    const D E((float)params[0], 1.f, 0.f), nu((float)params[1], 0.f, 1.f);
    Grid g;
    g.init(S.G);
    g_grid_side = S.G;
    g_out_of_grid = 0;
    const int L = n_substeps - 2;
    auto t0 = std::chrono::steady_clock::now();
    int cur = 0;
    for (int f = 0; f < L; ++f) {
        substep(S, st[cur], st[1 - cur], g, f, 0, E, nu, (float)params[2], (float)params[3]);
        cur = 1 - cur;
    }
    // compute_loss (parameter_fitting_functions.py:3-15): first argmin, strict <
    const auto& X = st[cur].vert.x;
    double loss = 0.0, gE = 0.0, gnu = 0.0;
#pragma omp parallel for reduction(+ : loss, gE, gnu) schedule(static)
    for (int q = 0; q < n_targets; ++q) {
        const float t[3] = {(float)targets[3 * q], (float)targets[3 * q + 1], (float)targets[3 * q + 2]};
        D best;
        for (int p = 0; p < S.nv; ++p) {
            D dist = ((X[p].x - D(t[0])) * (X[p].x - D(t[0])) + (X[p].y - D(t[1])) * (X[p].y - D(t[1])) +
                      (X[p].z - D(t[2])) * (X[p].z - D(t[2]))) / D(3.f);
            if (p == 0 || dist.v < best.v) best = dist;
        }
        loss += (double)best.v / n_targets;
        gE += (double)best.a / n_targets;
        gnu += (double)best.b / n_targets;
    }
    auto t1 = std::chrono::steady_clock::now();
    *loss_out = loss;
    grad_out[0] = gE;
    grad_out[1] = gnu;
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (final_vertex_x)
        for (int p = 0; p < S.nv; ++p) {
            final_vertex_x[3 * p] = X[p].x.v; final_vertex_x[3 * p + 1] = X[p].y.v; final_vertex_x[3 * p + 2] = X[p].z.v;
        }
    return g_out_of_grid ? -2 : 0;
}

int cref_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
