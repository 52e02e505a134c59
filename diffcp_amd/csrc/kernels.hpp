// Device helpers shared by the forward and adjoint kernels, and the kernel
// declarations used by the host engine.
#pragma once
#include "sim_common.hpp"
#include "element_math.hpp"
#include "diffcp.h"

// (1 - 0.95) evaluated in double then rounded, as the reference's constant
// folding does (simulation.py:383)
#define DCP_C_SYM_KEEP ((float)(1.0 - 0.95))
#define DCP_BOUND 3        // initialization.py:22

// An index that is the same for every lane of the workgroup (a patch of a
// workgroup-strided loop, its instance), said so to the compiler: the records
// it selects are then scalar loads held in scalar registers, not per-lane
// copies in vector registers.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
// add channels c0 .. c0+n-1 of a tile node into the fixed-point scatter sums
// q of a grid cell (words 2c, 2c+1 of channel c); a term out of range marks
// the instance's poison flag instead
template <int N>
__device__ __forceinline__ void fx_add_node(unsigned long long* q, long long cell, const float* v, int* poison) {
    bool ok = true;
#pragma unroll
    for (int c = 0; c < N; ++c)
        ok &= fx_atomic_add(fx_cell_word(q, cell, 2 * c), fx_cell_word(q, cell, 2 * c + 1), (double)v[c]);
    if (!ok) atomicOr(poison, 1);
}
// the (m, m v) or d/d grid_v_out channels of a cell from its fixed-point
// sums; NaN when the instance's poison flag is set; the words are cleared
template <int N>
__device__ __forceinline__ void fx_take_cell(unsigned long long* q, long long cell, bool poisoned, float* v) {
#pragma unroll
    for (int c = 0; c < N; ++c) {
        unsigned long long* w0 = fx_cell_word(q, cell, 2 * c);
        unsigned long long* w1 = fx_cell_word(q, cell, 2 * c + 1);
        v[c] = poisoned ? __int_as_float(0x7fc00000) : (float)fx_join(*w0, *w1);
        *w0 = 0ull;
        *w1 = 0ull;
    }
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

struct ElemKin {
    int i1, i2, i3;
    V3<float> np, xe;
    M3<float> C, M, dn, Dinv, F, d;
    M3<float> Fr;          // F before the live return mapping (apply_return_map)
};

__device__ __forceinline__ ElemParams elem_params(const SimDev& sim, const InstDev& I, int e) {
    ElemParams p{I.E, I.nu, I.ks, I.kc};
    if (sim.per_elem) {
        p.E = sim.emod[e];
        p.ks = sim.emod[(size_t)sim.NE + e];
        p.kc = sim.emod[(size_t)2 * sim.NE + e];
    }
    return p;
}

// d and F of element e at the state in S (fresh: F = I, d = rest shape).
__device__ __forceinline__ void element_kinematics(const SimDev& sim, const InstDev& I, const StateView& S, int e,
                                                   int fresh, ElemKin& K) {
    const int NE = sim.NE;
    K.i1 = sim.tri[e]; K.i2 = sim.tri[NE + e]; K.i3 = sim.tri[2 * NE + e];
    V3<float> x1 = S.x(K.i1), x2 = S.x(K.i2), x3 = S.x(K.i3);
    K.np = S.n(e);
    K.C = S.Ce(e);
    V3<float> col3 = K.np;
    if (!fresh) {
        // col3 += (dt C + 0.5 dt^2 C C) col3   (simulation.py:397-400)
        M3<float> hC = mscale(I.half_dt2, K.C);
        K.M = madd(mscale(I.dt, K.C), mmul(hC, K.C));
        col3 = K.np + mv(K.M, K.np);
    }
    V3<float> e1 = x2 - x1, e2 = x3 - x1;
    for (int i = 0; i < 3; ++i) { K.dn(i, 0) = e1[i]; K.dn(i, 1) = e2[i]; K.dn(i, 2) = col3[i]; }
    K.Dinv = load_m9(sim.Dinv, NE, e);
    if (fresh) {
        K.F = meye<float>();
        K.d = K.dn;
    } else {
        K.F = mmul(K.dn, K.Dinv);                 // new_F = new_d @ Dinv       (:407)
        K.d = mmul(K.F, minv(K.Dinv));            // d = new_F @ inverse(Dinv)  (:429)
    }
    if (fresh) {
        K.xe = {(x1.x + x2.x + x3.x) / 3.0f, (x1.y + x2.y + x3.y) / 3.0f, (x1.z + x2.z + x3.z) / 3.0f};
    } else {
        const float third = 1.0f / 3.0f;
        K.xe = {third * (x1.x + x2.x + x3.x), third * (x1.y + x2.y + x3.y), third * (x1.z + x2.z + x3.z)};
    }
}

// Live return mapping of a non-fresh element's F (off by default; see
// return_map_forward): F = Q return_map(R) with cf = the instance's friction,
// d = F inverse(Dinv); the unmapped F is kept in K.Fr for the adjoint.
__device__ __forceinline__ void apply_return_map(const SimDev& sim, const InstDev& I, int e, const ElemParams& prm,
                                                 ElemKin& K) {
    RMState rs;
    K.Fr = K.F;
    K.F = return_map_forward(K.Fr, prm.ks, prm.kc, I.friction, rs);
    K.d = mmul(K.F, minv(K.Dinv));
}

__device__ __forceinline__ V3<float> particle_x(const SimDev& sim, const StateView& S, const Patch& P, int q) {
    if (P.cls == CLS_ELEMENT) return load_v3(sim.xe, sim.NE, P.start + q);
    return S.x(P.start + q);
}

// prescribed grasper velocity (simulation.py:160-165); the frame index is the
// integer s / steps_per_frame + frames_per_window * g
__device__ __forceinline__ int grasper_frame(const InstDev& I, int s, int g) {
    return s / I.steps_per_frame + I.frames_per_window * g;
}
__device__ __forceinline__ V3<float> grasper_velocity(const SimDev& sim, const InstDev& I, int local, int s, int g) {
    int frame = grasper_frame(I, s, g);
    int gid = local / I.group_size;
    if (frame + 1 < I.n_actions) {
        const float* a = sim.traj + 3 * ((size_t)I.traj_off + (size_t)(frame + 1) * I.n_grasp + gid);
        const float* b = sim.traj + 3 * ((size_t)I.traj_off + (size_t)frame * I.n_grasp + gid);
        return {(a[0] - b[0]) / I.interval, (a[1] - b[1]) / I.interval, (a[2] - b[2]) / I.interval};
    }
    return {0.f, 0.f, 0.f};
}

struct P2GRec {
    V3<float> x, b, n, k;
    M3<float> A;
    float m;
    bool elem;
};

// Per-element corner vectors are stored element-major, [element][corner][3]
// (and per-element vectors [element][3]): a vertex's gather reads one 12-byte
// run per incident (element, corner) entry, whose CSR code ec = 3 element +
// corner is also its vector's index.
__device__ __forceinline__ V3<float> load3(const float* p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ void store3(float* p, V3<float> v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// Sums, in CSR order, of the per-corner vectors a[element][corner][.] (and of
// the per-element vectors b[element][.], when TWO) over the (element, corner)
// entries incident to vertex p.  Entries go eight at a time with their list and
// vector loads issued together: one entry after the other would chain two load
// latencies per entry.
template <bool TWO>
__device__ __forceinline__ void incident_sum(const SimDev& sim, int p, const float* a, const float* b,
                                             V3<float>& sa, V3<float>& sb) {
    const int t0 = sim.inc_off[p], t1 = sim.inc_off[p + 1];
    for (int t = t0; t < t1; t += 8) {
        int ec[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) ec[u] = t + u < t1 ? sim.inc_list[t + u] : -1;
        V3<float> va[8], vb[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (ec[u] >= 0) {
                va[u] = load3(a + (size_t)3 * ec[u]);
                if (TWO) vb[u] = load3(b + (size_t)3 * (ec[u] / 3));
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (ec[u] >= 0) {
                sa = sa + va[u];
                if (TWO) sb = sb + vb[u];
            }
        }
    }
}

// everything a particle scatters: per cell w*(b + A dpos) + (grad w . n) k, mass w*m
__device__ __forceinline__ void p2g_record(const SimDev& sim, const InstDev& I, const StateView& S, const Patch& P,
                                           int q, int s, int g, P2GRec& r) {
    // locals assigned once at the end: a struct filled field by field in
    // branches with an early return ends up in scratch memory
    V3<float> x, b, n = {0.f, 0.f, 0.f}, k = {0.f, 0.f, 0.f};
    M3<float> C;
    float m;
    const bool elem = P.cls == CLS_ELEMENT;
    if (elem) {
        const int e = P.start + q;
        x = load_v3(sim.xe, sim.NE, e);
        m = I.m_quad;
        b = m * load_v3(sim.ve, sim.NE, e);
        C = S.Ce(e);
        n = load_v3(sim.nd, sim.NE, e);
        k = load_v3(sim.ke, sim.NE, e);
    } else {
        const int p = P.start + q;
        x = S.x(p);
        C = S.C(p);
        if (P.cls == CLS_VERTEX) {
            m = I.m_vertex;
            V3<float> v = S.v(p);
            v.y += I.gdt;                                   // simulation.py:181-182
            V3<float> f = {0.f, 0.f, 0.f}, unused = {0.f, 0.f, 0.f};
            incident_sum<false>(sim, p, sim.fc, nullptr, f, unused);
            b = m * v + f;
        } else if (P.cls == CLS_GRASPER) {
            m = I.m_op;
            b = m * grasper_velocity(sim, I, p - I.o_start, s, g);
        } else {
            m = I.m_cube;
            b = {0.f, 0.f, 0.f};                            // simulation.py:150-151 (no velocity term)
        }
    }
    r.x = x; r.b = b; r.n = n; r.k = k;
    r.A = mscale(m, C);
    r.m = m;
    r.elem = elem;
}

// y + n x for a stencil offset n in {0, 1, 2} known at compile time after
// unrolling (no multiply by 0, which IEEE arithmetic cannot fold away)
__device__ __forceinline__ float add_times(float y, int n, float x) {
    return n == 0 ? y : (n == 1 ? y + x : y + 2.0f * x);
}

// A P2G record in the separable form of its node terms.  With the node offset
// o = (i, j, k) and dpos = (o - fx) dx, the momentum term
//   w (b + A dpos) = w (u + i ad0 + j ad1 + k ad2),  u = b - dx A fx,
// where ad_c is column c of dx A, and the element normal-impulse term
//   (grad w . n) kimp = (W2_k alpha_ij + dW2_k beta_ij) kimp
// with alpha_ij = dW0_i W1_j nn0 + W0_i dW1_j nn1, beta_ij = W0_i W1_j nn2 and
// nn = inv_dx n: a node costs a few adds and multiplies instead of the affine
// product and the weight gradient.  The weights are rebuilt from fx per layer
// (a few operations) rather than kept in registers across the layer passes.
// Points carry nn = kimp = 0, so the element term adds zeros and a wave that
// holds both vertices and elements does not diverge.
struct P2GSep {
    float u[3], ad[3][3];   // ad[c][a] = dx A(a, c)
    float nn[3], kimp[3];
    float fx[3], m;
};

// quadratic B-spline weights and their d/dfx along one axis (simulation.py:142,197)
__device__ __forceinline__ void bspline_axis(float f, float w[3], float dw[3]) {
    w[0] = 0.5f * (1.5f - f) * (1.5f - f);
    w[1] = 0.75f - (f - 1.0f) * (f - 1.0f);
    w[2] = 0.5f * (f - 0.5f) * (f - 0.5f);
    dw[0] = f - 1.5f;
    dw[1] = 2.0f * (1.0f - f);
    dw[2] = f - 0.5f;
}

// fx = x inv_dx - base for a (clamped) base cell, as the stencil builds it
__device__ __forceinline__ float stencil_fx(float x, float inv_dx, int base) {
    return scaled_position(x, inv_dx) - (float)base;
}

__device__ __forceinline__ void p2g_separate(const InstDev& I, const P2GRec& r, int b0, int b1, int b2, P2GSep& q) {
    q.fx[0] = stencil_fx(r.x.x, I.inv_dx, b0);
    q.fx[1] = stencil_fx(r.x.y, I.inv_dx, b1);
    q.fx[2] = stencil_fx(r.x.z, I.inv_dx, b2);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int c = 0; c < 3; ++c) q.ad[c][a] = I.dx * r.A(a, c);
        q.u[a] = r.b[a] - (q.ad[0][a] * q.fx[0] + q.ad[1][a] * q.fx[1] + q.ad[2][a] * q.fx[2]);
    }
    q.nn[0] = I.inv_dx * r.n.x; q.nn[1] = I.inv_dx * r.n.y; q.nn[2] = I.inv_dx * r.n.z;
    q.kimp[0] = r.k.x; q.kimp[1] = r.k.y; q.kimp[2] = r.k.z;
    q.m = r.m;
}

// The nine nodes (i, j, k), j, k = 0..2, of stencil layer i: emit(j, k, mass,
// momentum x, y, z) with the weights w = W0_i W1_j W2_k (simulation.py:142-152,
// 172-174, 186-199, 262-265: mass, APIC momentum, element normal impulse).
template <typename Emit>
__device__ __forceinline__ void p2g_layer(const P2GSep& q, int i, Emit emit) {
    float w0[3], d0[3], w1[3], d1[3], w2[3], d2[3];
    bspline_axis(q.fx[0], w0, d0);
    bspline_axis(q.fx[1], w1, d1);
    bspline_axis(q.fx[2], w2, d2);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const float w01 = w0[i] * w1[j];
        const float al = d0[i] * w1[j] * q.nn[0] + w0[i] * d1[j] * q.nn[1];
        const float be = w01 * q.nn[2];
        float tij[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) tij[a] = add_times(add_times(q.u[a], i, q.ad[0][a]), j, q.ad[1][a]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float w = w01 * w2[k];
            const float sd = w2[k] * al + d2[k] * be;
            emit(j, k, w * q.m, w * add_times(tij[0], k, q.ad[2][0]) + sd * q.kimp[0],
                 w * add_times(tij[1], k, q.ad[2][1]) + sd * q.kimp[1],
                 w * add_times(tij[2], k, q.ad[2][2]) + sd * q.kimp[2]);
        }
    }
}

__device__ __forceinline__ void block_cell_coords(const InstDev& I, int blk, int lane, int& i, int& j, int& k) {
    int bk = blk % I.nb, t = blk / I.nb;
    int bj = t % I.nb, bi = t / I.nb;
    i = bi * 4 + (lane >> 4);
    j = bj * 4 + ((lane >> 2) & 3);
    k = bk * 4 + (lane & 3);
}

// Walk of the per-instance active block lists by a persistent grid of
// B x `per` workgroups: workgroup (b, k) takes every (per * waves)-th entry of
// instance b, one 4^3 block (64 cells) per wave.
struct ActiveWalk {
    int b, first, stride, n;
};
__device__ __forceinline__ bool active_walk(const SimDev& sim, const InstDev* inst, ActiveWalk& w) {
    const int per = max(1, (int)gridDim.x / sim.B);
    w.b = blockIdx.x / per;
    if (w.b >= sim.B) return false;
    const int nw = DCP_WG / 64;
    w.first = (blockIdx.x % per) * nw + (threadIdx.x >> 6);
    w.stride = per * nw;
    const int nb = inst[w.b].nb;
    w.n = min(sim.n_active[w.b], nb * nb * nb);
    return true;
}

// DCP_WALK_U consecutive entries of a wave's walk (it0, it0 + stride, ...)
// with their first cells, loaded together: a wave's few blocks then cost one
// round of dependent loads (list entry, pool slot, cell data) instead of one
// round per block.  blk < 0 past the end of the list.
#define DCP_WALK_U 4
struct WalkBatch {
    int blk[DCP_WALK_U];
    long long c0[DCP_WALK_U];
};
__device__ __forceinline__ void walk_batch(const int* list, const InstDev& I, int it0, int stride, int n,
                                           WalkBatch& wb) {
#pragma unroll
    for (int u = 0; u < DCP_WALK_U; ++u) {
        const int it = it0 + u * stride;
        wb.blk[u] = it < n ? list[it] : -1;
    }
#pragma unroll
    for (int u = 0; u < DCP_WALK_U; ++u) wb.c0[u] = wb.blk[u] >= 0 ? block_cell0(I, wb.blk[u]) : 0;
}

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// v_out of one cell (simulation.py:277-302) with the damping extension;
// mask[c] = 1 where the component passed the wall conditions
__device__ __forceinline__ void grid_velocity(const InstDev& I, float4 c, int i, int j, int k, float v[3], float* mask,
                                              float* vraw = nullptr) {
    float inv_m = 1.0f / (c.x + 1e-9f);
    v[0] = inv_m * c.y; v[1] = inv_m * c.z; v[2] = inv_m * c.w;
    if (vraw) { vraw[0] = v[0]; vraw[1] = v[1]; vraw[2] = v[2]; }
    float damp = 1.0f - I.damping * I.dt;
    v[0] *= damp; v[1] *= damp; v[2] *= damp;
    bool all = false, yonly = false;
    const int G = I.G, b = DCP_BOUND;
    if (i < b && v[0] < 0) { v[0] = v[1] = v[2] = 0.f; all = true; }
    if (i > G - b && v[0] > 0) { v[0] = v[1] = v[2] = 0.f; all = true; }
    if (k < b && v[2] < 0) { v[0] = v[1] = v[2] = 0.f; all = true; }
    if (k > G - b && v[2] > 0) { v[0] = v[1] = v[2] = 0.f; all = true; }
    if (j < b && v[1] < 0) { v[1] = 0.f; yonly = true; }
    if (j > G - b && v[1] > 0) { v[0] = v[1] = v[2] = 0.f; all = true; }
    if (mask) {
        mask[0] = all ? 0.f : 1.f;
        mask[1] = (all || yonly) ? 0.f : 1.f;
        mask[2] = all ? 0.f : 1.f;
    }
}

// kernels -------------------------------------------------------------------
template <typename PT, bool RM> __global__ void k_prep(SimDev sim, StateView S, StateView Snext, int fresh);
// caller arrays of a state slot in device memory (k_state_export / k_state_import)
struct StateIO {
    float* a[11];      // vertex x v C, grasper x v C, cube x v C, element C, element normal column
    const int* perm;   // natural particle (vertices, graspers, cubes, elements) -> slot index
    int n[4];          // vertices, graspers, cubes, elements (all instances)
};
__global__ void k_state_export(const float* slot, int NP, int NE, StateIO io);
__global__ void k_state_import(float* slot, int NP, int NE, StateIO io);
template <int GF> __global__ void k_p2g(SimDev sim, StateView S, int s, int g, const int* skip);
template <int GF> __global__ void k_p2g_big(SimDev sim, StateView S, int s, int g, const int* skip);
__global__ void k_active_list(SimDev sim, const int* skip);
__global__ void k_pool_alloc(SimDev sim, StateView S, const int* skip);
__global__ void k_pool_reset(int* btab, int n, int* pool_n);
__global__ void k_gridop(SimDev sim, int keep, GridSnap save, const int* skip);
__global__ void k_grid_restore(SimDev sim, GridSnap snap);
template <bool POOL> __global__ void k_g2p(SimDev sim, StateView S, StateView Sn);
__global__ void k_loss_min(SimDev sim, StateView S, const float* targets, const int4* tchunks, int n_chunks,
                           int n_parts, int n_targets, const int* vnat, float* pmin, int* pnat, int* parg);
__global__ void k_loss_sum(const float* pmin, const int* pnat, const int* parg, int n_parts, int n_targets,
                           float* tmin, int* targ, const int* toff, const int* tcount, double* loss, int B);
__global__ void k_binning_keys(SimDev sim, StateView S, int* keys);
__global__ void k_patch_order(SimDev sim, StateView S, int* out);
__global__ void k_rebase(SimDev sim, StateView S, float* f_old, float* Dinv);

template <int GF> __global__ void k_g2p_adj(SimDev sim, StateView S, StateView G1, StateView Gout, int pending);
template <int GF> __global__ void k_g2p_adj_big(SimDev sim, StateView S, StateView G1, StateView Gout, int pending);
__global__ void k_gridop_adj(SimDev sim, int big_left_out);
template <bool POOL> __global__ void k_p2g_adj(SimDev sim, StateView S, StateView Gout, int s, int g);
template <typename PT, bool RM, bool POOL>
__global__ void k_elem_adj(SimDev sim, StateView S, StateView G1, StateView Gout, int fresh, int s, int g);
__global__ void k_vertex_final(SimDev sim, StateView Gout);
template <typename PT> __global__ void k_polar_adj(const float* Ft, const float* gR, int n, int mode, float* gF, float* R);
__global__ void k_loss_adj(SimDev sim, StateView S, StateView Gout, const float* targets, const int* toff,
                           const int* tcount, const int* targ, const double* seed);
__global__ void k_fill_f(float* p, size_t n, float v);
