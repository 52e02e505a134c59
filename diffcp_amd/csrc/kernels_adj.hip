// Adjoint substep kernels: the hand-written reverse of k_prep, k_p2g, k_gridop
// and k_g2p (what Taichi's tape generates for simulation.py:500-506, with the
// manual SVD backward of :71-94).  The adjoint of substep s runs on its grid:
// restored from the forward's snapshot (k_grid_restore), or recomputed
// (k_prep, k_p2g, k_gridop with keep=1):
//   k_g2p_adj       per patch: dL/d grid_v_out scattered through an LDS tile
//                   (the same tiling as P2G), dL/dx through the weights
//   k_gridop_adj    per active block: wall masks, damping, normalisation
//   k_p2g_adj       per vertex / grasper / cube particle: 27-cell gather of
//                   dL/d(m, mv) straight from the grid
//   k_elem_adj      per element gather patch: the same gather for the element
//                   particles feeding, in registers, the element prep adjoint
//                   (stress, QR, polar (FP64), kinematics); parameter
//                   gradients reduced per patch in FP64, fixed order
//   k_vertex_final  per vertex: incident-element corner gradients (CSR order)
#include "kernels.hpp"
#include "patch_sort.hpp"

__global__ void k_fill_f(float* p, size_t n, float v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// G2P adjoint
// ---------------------------------------------------------------------------
// adjoint inputs of one particle for the G2P reverse: dL/dv_new (incl. the
// x_new = x + dt v_new term), dL/dC_new (element: through the damping map) and
// the direct dL/dx_new.  `pending`: the vertex entries of G1 still lack their
// incident elements' corner gradients (the previous adjoint substep left them
// in ec_gx / ec_gv instead of gathering them), which are added here.
__device__ __forceinline__ void g2p_adj_inputs(const SimDev& sim, const InstDev& I, const StateView& G1, const Patch& P,
                                               int q, int pending, V3<float>& gv, V3<float>& gx1, M3<float>& gC) {
    const float csym = 0.5f * DCP_C_SYM_KEEP;
    if (P.cls == CLS_ELEMENT) {
        M3<float> go = G1.Ce(P.start + q);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) gC(a, b) = 0.5f * (go(a, b) - go(b, a)) + csym * (go(a, b) + go(b, a));
        gv = {0.f, 0.f, 0.f};
        gx1 = {0.f, 0.f, 0.f};
    } else {
        const int p = P.start + q;
        gx1 = G1.x(p);
        V3<float> g1v = G1.v(p);
        if (pending && P.cls == CLS_VERTEX) incident_sum<true>(sim, p, sim.ec_gx, sim.ec_gv, gx1, g1v);
        gv = {g1v.x + I.dt * gx1.x, g1v.y + I.dt * gx1.y, g1v.z + I.dt * gx1.z};
        gC = G1.C(p);
    }
}

// The G2P reverse of one particle in separable form.  With the node offset o
// and t = gv + 4 inv_dx gC (o - fx) = u + o_0 A_0 + o_1 A_1 + o_2 A_2, where
// A_b = 4 inv_dx gC(:, b) and u = gv - fx_0 A_0 - fx_1 A_1 - fx_2 A_2:
//   dL/d grid_v_out(node) = w t,
//   dL/dfx_d = sum_nodes (d w / d fx_d) (g . t) - A_d . (sum_nodes w g),
// g the node's grid_v_out.  The first sum runs over k inside each (i, j) row
// (X = sum_k W2_k g.t, Y = sum_k dW2_k g.t), so a node costs about fifteen
// operations; the weights are rebuilt from fx per layer.
struct G2PAdjSep {
    float u[3], A[3][3];    // A[b][a] = 4 inv_dx gC(a, b)
    float fx[3];
    float gw[3], gfx[3];    // sum of w g; dL/dfx without its A . gw term
};

__device__ __forceinline__ void g2p_adj_separate(const InstDev& I, const Stencil& st, const V3<float>& gv,
                                                 const M3<float>& gC, G2PAdjSep& q) {
    const float c4 = 4.0f * I.inv_dx;
    const float g0[3] = {gv.x, gv.y, gv.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int b = 0; b < 3; ++b) q.A[b][a] = c4 * gC(a, b);
        q.u[a] = g0[a] - (q.A[0][a] * st.fx[0] + q.A[1][a] * st.fx[1] + q.A[2][a] * st.fx[2]);
        q.fx[a] = st.fx[a];
        q.gw[a] = 0.f;
        q.gfx[a] = 0.f;
    }
}

// Layer i: the nine nodes' grid_v_out from fetch(j, k) (a float4, xyz used),
// their dL/d grid_v_out handed to emit(j, k, g0, g1, g2).
template <typename Fetch, typename Emit>
__device__ __forceinline__ void g2p_adj_layer(G2PAdjSep& q, int i, Fetch fetch, Emit emit) {
    float w0[3], d0[3], w1[3], d1[3], w2[3], d2[3];
    bspline_axis(q.fx[0], w0, d0);
    bspline_axis(q.fx[1], w1, d1);
    bspline_axis(q.fx[2], w2, d2);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const float w01 = w0[i] * w1[j];
        float tij[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) tij[a] = add_times(add_times(q.u[a], i, q.A[0][a]), j, q.A[1][a]);
        float X = 0.f, Y = 0.f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float4 g = fetch(j, k);
            const float t0 = add_times(tij[0], k, q.A[2][0]);
            const float t1 = add_times(tij[1], k, q.A[2][1]);
            const float t2 = add_times(tij[2], k, q.A[2][2]);
            const float w = w01 * w2[k];
            emit(j, k, w * t0, w * t1, w * t2);
            const float scal = g.x * t0 + g.y * t1 + g.z * t2;
            X += w2[k] * scal;
            Y += d2[k] * scal;
            q.gw[0] += w * g.x; q.gw[1] += w * g.y; q.gw[2] += w * g.z;
        }
        q.gfx[0] += d0[i] * w1[j] * X;
        q.gfx[1] += w0[i] * d1[j] * X;
        q.gfx[2] += w01 * Y;
    }
}

__device__ __forceinline__ void g2p_adj_finish(const G2PAdjSep& q, float gfx[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) gfx[d] = q.gfx[d] - (q.A[d][0] * q.gw[0] + q.A[d][1] * q.gw[1] + q.A[d][2] * q.gw[2]);
}

__device__ __forceinline__ void g2p_adj_store(const SimDev& sim, const InstDev& I, const StateView& Gout, const Patch& P,
                                              int q, const V3<float>& gx1, const float gfx[3]) {
    V3<float> gxt = {I.inv_dx * gfx[0], I.inv_dx * gfx[1], I.inv_dx * gfx[2]};
    if (P.cls == CLS_ELEMENT) store_v3(sim.g_xe, sim.NE, P.start + q, gxt);
    else store_v3(Gout.px(0), sim.NP, P.start + q, gx1 + gxt);
}

template <int GF>
__device__ __forceinline__ void g2p_adj_spill(const SimDev& sim, const StateView& S, const StateView& G1,
                                           const StateView& Gout, const Patch& P, const InstDev& I, int pending) {
    [[maybe_unused]] constexpr bool POOL = (GF & DCP_GF_POOL) != 0, DET = (GF & DCP_GF_DET) != 0;
    for (int q = threadIdx.x; q < P.count; q += DCP_WG) {
        Stencil st;
        make_stencil(particle_x(sim, S, P, q), I.inv_dx, I.G, st, nullptr);
        V3<float> gv, gx1;
        M3<float> gC;
        g2p_adj_inputs(sim, I, G1, P, q, pending, gv, gx1, gC);
        G2PAdjSep sep;
        g2p_adj_separate(I, st, gv, gC, sep);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            g2p_adj_layer(
                sep, i,
                [&](int j, int k) { return sim.vout[cell_index<POOL>(I, st.base[0] + i, st.base[1] + j, st.base[2] + k)]; },
                [&](int j, int k, float g0, float g1, float g2) {
                    long long gi = cell_index<POOL>(I, st.base[0] + i, st.base[1] + j, st.base[2] + k);
                    if (DET) {
                        const float v[3] = {g0, g1, g2};
                        fx_add_node<3>(sim.gaccq, gi, v, sim.poison + sim.B + P.inst);
                    } else {
                        atomicAdd(&sim.gacc[gi].y, g0); atomicAdd(&sim.gacc[gi].z, g1); atomicAdd(&sim.gacc[gi].w, g2);
                    }
                });
        float gfx[3];
        g2p_adj_finish(sep, gfx);
        g2p_adj_store(sim, I, Gout, P, q, gx1, gfx);
    }
}

// Compute-then-gather form (the structure of k_p2g): one particle per lane with
// its G2P-reverse inputs in registers; per x-layer i each lane computes its
// particle's dL/d grid_v_out contributions to the layer's nine nodes into LDS
// slots, the runs of each occupied base cell are summed, and the lane owning a
// tile node adds its nine run sums.  dL/dx of each particle accumulates over
// the three layers and is stored after the last.  grid_v_out of the tile is staged in
// LDS once per patch, so every node read after that is an LDS read.
// G2P reverse, through one LDS tile over box B, of the patch particles whose
// base cells lie in B (`act` lanes; B is the patch's box, or one sub-box of it
// in k_g2p_adj_big): node terms, dL/dx stores, and the flush of dL/d grid_v_out.
template <int GF>
__device__ __forceinline__ void g2p_adj_tile(const SimDev& sim, const InstDev& I, const StateView& Gout,
                                             const PatchBox& B, bool act, const V3<float>& x, const V3<float>& gv,
                                             const V3<float>& gx1, const M3<float>& gC, const Patch& Pl, int ql,
                                             int b0, int b1, int b2, BinLDS& L, float2* cA, float* cB, float4* vt) {
    [[maybe_unused]] constexpr bool POOL = (GF & DCP_GF_POOL) != 0, DET = (GF & DCP_GF_DET) != 0;
    const int tid = threadIdx.x;
    const int pos = patch_bin(act ? tile_cell(B, b0, b1, b2) : -1, B, L);
    // this lane's tile nodes: tid + h * DCP_WG (vol <= DCP_NODES * DCP_WG), and their planes
    int nci[DCP_NODES];
    bool has[DCP_NODES];
    float4 a[DCP_NODES];
#pragma unroll
    for (int h = 0; h < DCP_NODES; ++h) {
        const int c = tid + h * DCP_WG;
        has[h] = c < B.vol;
        const int ck = c % B.e2, t = c / B.e2;
        nci[h] = t / B.e1;
        a[h] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (has[h]) vt[c] = sim.vout[cell_index<POOL>(I, nci[h] + B.l0, t - nci[h] * B.e1 + B.l1, ck + B.l2)];
    }
    __syncthreads();
    // tile index of the particle's base cell; node (i, j, k) is key + (i * e1 + j) * e2 + k
    const int key = act ? tile_cell(B, b0, b1, b2) : 0;
    G2PAdjSep sep;
    if (act) {
        Stencil st;
        make_stencil(x, I.inv_dx, I.G, st, nullptr);
        g2p_adj_separate(I, st, gv, gC, sep);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (act) {
            g2p_adj_layer(
                sep, i, [&](int j, int k) { return vt[key + (i * B.e1 + j) * B.e2 + k]; },
                [&](int j, int k, float g0, float g1, float g2) {
                    cA[(j * 3 + k) * DCP_SLOT_STRIDE + pos] = make_float2(g0, g1);
                    cB[(j * 3 + k) * DCP_SLOT_STRIDE + pos] = g2;
                });
            if (i == 2) {
                float gfx[3];
                g2p_adj_finish(sep, gfx);
                g2p_adj_store(sim, I, Gout, Pl, ql, gx1, gfx);
            }
        }
        __syncthreads();
        run_sums3(L, cA, cB);
        __syncthreads();
#pragma unroll
        for (int h = 0; h < DCP_NODES; ++h) {
            const float3 v = node_sum3(L, B, cA, cB, tid + h * DCP_WG, nci[h], i, has[h]);
            a[h].y += v.x; a[h].z += v.y; a[h].w += v.z;
        }
        __syncthreads();
    }
#pragma unroll
    for (int h = 0; h < DCP_NODES; ++h) {
        if (has[h] && (a[h].y != 0.f || a[h].z != 0.f || a[h].w != 0.f)) {
            const int c = tid + h * DCP_WG, t = c / B.e2;
            long long gidx = cell_index<POOL>(I, nci[h] + B.l0, t - nci[h] * B.e1 + B.l1, c - t * B.e2 + B.l2);
            if (DET) {
                const float v[3] = {a[h].y, a[h].z, a[h].w};
                fx_add_node<3>(sim.gaccq, gidx, v, sim.poison + sim.B + Pl.inst);
            } else {
                atomicAdd(&sim.gacc[gidx].y, a[h].y); atomicAdd(&sim.gacc[gidx].z, a[h].z);
                atomicAdd(&sim.gacc[gidx].w, a[h].w);
            }
        }
    }
}

// Lane q's inputs of a scatter patch: position, G2P-reverse inputs, base cell.
__device__ __forceinline__ void g2p_adj_patch_inputs(const SimDev& sim, const InstDev& I, const StateView& S,
                                                     const StateView& G1, const SPatch& SP, int q, int pending,
                                                     Patch& Pl, int& ql, V3<float>& x, V3<float>& gv,
                                                     V3<float>& gx1, M3<float>& gC, int& b0, int& b1, int& b2) {
    SP.lane(q, Pl, ql);
    x = particle_x(sim, S, Pl, ql);
    g2p_adj_inputs(sim, I, G1, Pl, ql, pending, gv, gx1, gC);
    b0 = stencil_base(x.x, I.inv_dx, I.G);
    b1 = stencil_base(x.y, I.inv_dx, I.G);
    b2 = stencil_base(x.z, I.inv_dx, I.G);
}

// As in k_p2g, a patch whose tile overflows is listed for k_g2p_adj_big.  The
// tile reverse is written out here rather than called as g2p_adj_tile
// (k_g2p_adj_big's): the shared form compiled to a schedule measured 3 %
// slower on this kernel.
template <int GF>
__global__ __launch_bounds__(DCP_WG, 3) void k_g2p_adj(SimDev sim, StateView S, StateView G1, StateView Gout, int pending) {
    [[maybe_unused]] constexpr bool POOL = (GF & DCP_GF_POOL) != 0, DET = (GF & DCP_GF_DET) != 0;
    // a recomputed grid update, just before on the stream, has read the
    // scatter's poison flags
    if (blockIdx.x == 0)
        for (int b = threadIdx.x; b < sim.B; b += DCP_WG) sim.poison[b] = 0;
    __shared__ BinLDS L;
    __shared__ float2 cA[DCP_CONTRIB_ZERO + 1];   // contributions, channels 0 and 1
    __shared__ float cB[DCP_CONTRIB_ZERO + 1];    // channel 2
    __shared__ float4 vt[DCP_STILE_CAP];        // grid_v_out over the patch tile
    const int tid = threadIdx.x;
    bin_clear(L);
    if (tid < 3) { L.box[tid] = 1 << 30; L.box[3 + tid] = -(1 << 30); }
    if (tid == 0) { cA[DCP_CONTRIB_ZERO] = make_float2(0.f, 0.f); cB[DCP_CONTRIB_ZERO] = 0.f; }   // never overwritten
    __syncthreads();
    for (int pi = blockIdx.x; pi < sim.n_spatches; pi += gridDim.x) {
        const SPatch SP = sim.spatches[pi];
        const InstDev I = sim.inst[SP.inst];
        bool tiled = SP.count() <= DCP_WG;
        if (tiled) {
            const bool act = tid < SP.count();
            V3<float> x = {0.f, 0.f, 0.f}, gv, gx1;
            M3<float> gC;
            int b0 = 0, b1 = 0, b2 = 0;
            Patch Pl{SP.inst, 0, 0, 0};
            int ql = 0;
            if (act) {
                SP.lane(tid, Pl, ql);
                x = particle_x(sim, S, Pl, ql);
                g2p_adj_inputs(sim, I, G1, Pl, ql, pending, gv, gx1, gC);
                b0 = stencil_base(x.x, I.inv_dx, I.G);
                b1 = stencil_base(x.y, I.inv_dx, I.G);
                b2 = stencil_base(x.z, I.inv_dx, I.G);
            }
            PatchBox B;
            box_reduce(b0, b1, b2, act, L.box, B);
            tiled = B.vol <= DCP_STILE_CAP;
            if (tiled) {
                const int pos = patch_bin(act ? tile_cell(B, b0, b1, b2) : -1, B, L);
                // this lane's tile nodes: tid + h * DCP_WG (vol <= DCP_NODES * DCP_WG), and their planes
                int nci[DCP_NODES];
                bool has[DCP_NODES];
                float4 a[DCP_NODES];
#pragma unroll
                for (int h = 0; h < DCP_NODES; ++h) {
                    const int c = tid + h * DCP_WG;
                    has[h] = c < B.vol;
                    const int ck = c % B.e2, t = c / B.e2;
                    nci[h] = t / B.e1;
                    a[h] = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (has[h]) vt[c] = sim.vout[cell_index<POOL>(I, nci[h] + B.l0, t - nci[h] * B.e1 + B.l1, ck + B.l2)];
                }
                __syncthreads();
                // tile index of the particle's base cell; node (i, j, k) is key + (i * e1 + j) * e2 + k
                const int key = act ? tile_cell(B, b0, b1, b2) : 0;
                G2PAdjSep sep;
                if (act) {
                    Stencil st;
                    make_stencil(x, I.inv_dx, I.G, st, nullptr);
                    g2p_adj_separate(I, st, gv, gC, sep);
                }
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    if (act) {
                        g2p_adj_layer(
                            sep, i, [&](int j, int k) { return vt[key + (i * B.e1 + j) * B.e2 + k]; },
                            [&](int j, int k, float g0, float g1, float g2) {
                                cA[(j * 3 + k) * DCP_SLOT_STRIDE + pos] = make_float2(g0, g1);
                                cB[(j * 3 + k) * DCP_SLOT_STRIDE + pos] = g2;
                            });
                        if (i == 2) {
                            float gfx[3];
                            g2p_adj_finish(sep, gfx);
                            g2p_adj_store(sim, I, Gout, Pl, ql, gx1, gfx);
                        }
                    }
                    __syncthreads();
                    run_sums3(L, cA, cB);
                    __syncthreads();
#pragma unroll
                    for (int h = 0; h < DCP_NODES; ++h) {
                        const float3 v = node_sum3(L, B, cA, cB, tid + h * DCP_WG, nci[h], i, has[h]);
                        a[h].y += v.x; a[h].z += v.y; a[h].w += v.z;
                    }
                    __syncthreads();
                }
#pragma unroll
                for (int h = 0; h < DCP_NODES; ++h) {
                    if (has[h] && (a[h].y != 0.f || a[h].z != 0.f || a[h].w != 0.f)) {
                        const int c = tid + h * DCP_WG, t = c / B.e2;
                        long long gidx = cell_index<POOL>(I, nci[h] + B.l0, t - nci[h] * B.e1 + B.l1, c - t * B.e2 + B.l2);
                        if (DET) {
                            const float v[3] = {a[h].y, a[h].z, a[h].w};
                            fx_add_node<3>(sim.gaccq, gidx, v, sim.poison + sim.B + SP.inst);
                        } else {
                            atomicAdd(&sim.gacc[gidx].y, a[h].y); atomicAdd(&sim.gacc[gidx].z, a[h].z);
                            atomicAdd(&sim.gacc[gidx].w, a[h].w);
                        }
                    }
                }
            }
        }
        if (!tiled && tid == 0)   // left to k_g2p_adj_big
            sim.big_list[sim.n_spatches + atomicAdd(&sim.big_n[DCP_BIG_G2P_ADJ], 1)] = pi;
        bin_clear(L);
        if (tid < 3) { L.box[tid] = 1 << 30; L.box[3 + tid] = -(1 << 30); }
        __syncthreads();
    }
}

// The patches k_g2p_adj left: the same tile reverse over the sub-boxes of the
// patch box (as k_p2g_big).  k_gridop_adj, next on the stream, empties the
// list.
template <int GF>
__global__ __launch_bounds__(DCP_WG, 2) void k_g2p_adj_big(SimDev sim, StateView S, StateView G1, StateView Gout,
                                                           int pending) {
    [[maybe_unused]] constexpr bool POOL = (GF & DCP_GF_POOL) != 0, DET = (GF & DCP_GF_DET) != 0;
    const int n_items = sim.big_n[DCP_BIG_G2P_ADJ] * DCP_BIG_SPLIT;
    if ((int)blockIdx.x >= n_items) return;   // the usual case: an empty list
    __shared__ BinLDS L;
    __shared__ float2 cA[DCP_CONTRIB_ZERO + 1];
    __shared__ float cB[DCP_CONTRIB_ZERO + 1];
    __shared__ float4 vt[DCP_STILE_CAP];
    const int tid = threadIdx.x;
    bin_clear(L);
    if (tid < 3) { L.box[tid] = 1 << 30; L.box[3 + tid] = -(1 << 30); }
    if (tid == 0) { cA[DCP_CONTRIB_ZERO] = make_float2(0.f, 0.f); cB[DCP_CONTRIB_ZERO] = 0.f; }
    __syncthreads();
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const int bi = it / DCP_BIG_SPLIT, part = it - bi * DCP_BIG_SPLIT;
        const int pi = sim.big_list[sim.n_spatches + bi];
        const SPatch SP = sim.spatches[pi];
        const InstDev I = sim.inst[SP.inst];
        bool tiled = SP.count() <= DCP_WG;
        if (tiled) {
            const bool act = tid < SP.count();
            V3<float> x = {0.f, 0.f, 0.f}, gv, gx1;
            M3<float> gC;
            int b0 = 0, b1 = 0, b2 = 0;
            Patch Pl{SP.inst, 0, 0, 0};
            int ql = 0;
            if (act) g2p_adj_patch_inputs(sim, I, S, G1, SP, tid, pending, Pl, ql, x, gv, gx1, gC, b0, b1, b2);
            PatchBox PB;
            box_reduce(b0, b1, b2, act, L.box, PB);
            const SubPlan SPL = sub_plan(PB);
            bool first = true;
            int k = 0;
            for (int x0 = 0; x0 < SPL.n[0]; x0 += SPL.w[0])
                for (int y0 = 0; y0 < SPL.n[1]; y0 += SPL.w[1])
                    for (int z0 = 0; z0 < SPL.n[2]; z0 += SPL.w[2]) {
                        if (k++ % DCP_BIG_SPLIT != part) continue;   // another item's sub-box
                        const PatchBox B = sub_box(PB, SPL, x0, y0, z0);
                        const bool in = act && in_sub_box(B, b0, b1, b2);
                        if (!__syncthreads_or(in)) continue;   // no particle in this sub-box
                        if (!first) {   // the previous sub-box's bins and tile are read
                            bin_clear(L);
                            __syncthreads();
                        }
                        first = false;
                        g2p_adj_tile<GF>(sim, I, Gout, B, in, x, gv, gx1, gC, Pl, ql, b0, b1, b2, L, cA, cB, vt);
                    }
        }
        if (!tiled && part == 0) {
            g2p_adj_spill<GF>(sim, S, G1, Gout, SP.part(0), I, pending);
            g2p_adj_spill<GF>(sim, S, G1, Gout, SP.part(1), I, pending);
        }
        bin_clear(L);
        if (tid < 3) { L.box[tid] = 1 << 30; L.box[3 + tid] = -(1 << 30); }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// grid update adjoint
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(DCP_WG) void k_gridop_adj(SimDev sim, int big_left_out) {
    // k_g2p_adj_big, just before on the stream, has read the overflow list;
    // when it was left out (the forward listed no patch) the list is empty
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (big_left_out && sim.big_n[DCP_BIG_G2P_ADJ] != 0) atomicOr(sim.status, (unsigned)DCP_STATUS_INTERNAL);
        sim.big_n[DCP_BIG_G2P_ADJ] = 0;
    }
    const int lane = threadIdx.x & 63;
    ActiveWalk w;
    if (!active_walk(sim, sim.inst, w)) return;
    const InstDev I = sim.inst[w.b];
    const bool poisoned = sim.poison[sim.B + w.b] != 0;
    long long dhi = 0, dlo = 0;   // deterministic sums: this wave's dL/d damping, fixed point
    bool dok = true;
    double gdamp_wave = 0.0;      // float-atomic sums: this lane's part of it
    for (int it0 = w.first; it0 < w.n; it0 += DCP_WALK_U * w.stride) {
        WalkBatch wb;
        walk_batch(sim.active + I.flag_base, I, it0, w.stride, w.n, wb);
        float4 cs[DCP_WALK_U], gs[DCP_WALK_U];
#pragma unroll
        for (int u = 0; u < DCP_WALK_U; ++u) {
            const bool in = wb.blk[u] >= 0;
            cs[u] = in ? sim.acc[wb.c0[u] + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
            gs[u] = (in && !sim.det) ? sim.gacc[wb.c0[u] + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < DCP_WALK_U; ++u) {
            const int blk = wb.blk[u];
            if (blk < 0) break;
            const long long cell = wb.c0[u] + lane;
            int ci, cj, ck;
            block_cell_coords(I, blk, lane, ci, cj, ck);
            const float4 c = cs[u];
            float4 ga = gs[u];
            if (sim.det) {
                float gv[3];
                fx_take_cell<3>(sim.gaccq, cell, poisoned, gv);
                ga = make_float4(0.f, gv[0], gv[1], gv[2]);
            } else {
                sim.gacc[cell] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            float v[3], mask[3], vr[3];
            grid_velocity(I, c, ci, cj, ck, v, mask, vr);
            float gd0 = mask[0] * ga.y, gd1 = mask[1] * ga.z, gd2 = mask[2] * ga.w;
            double gdamp = -(double)I.dt * ((double)gd0 * vr[0] + (double)gd1 * vr[1] + (double)gd2 * vr[2]);
            float damp = 1.0f - I.damping * I.dt;
            float g0 = damp * gd0, g1 = damp * gd1, g2 = damp * gd2;
            float inv_m = 1.0f / (c.x + 1e-9f);
            float gm = -inv_m * (g0 * vr[0] + g1 * vr[1] + g2 * vr[2]);
            sim.gmv[cell] = make_float4(gm, inv_m * g0, inv_m * g1, inv_m * g2);
            sim.acc[cell] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (sim.det) {
                // the block's sum in lane order, then fixed point: the same bits
                // whichever wave takes the block and in whatever order
                const double sb = wave_sum_d(gdamp);
                long long h, l;
                if (fx_split(sb, h, l)) { dhi += h; dlo += l; } else { dok = false; }
            } else {
                gdamp_wave += gdamp;
            }
            if (lane == 0) sim.flags[I.flag_base + blk] = 0;
        }
    }
    if (sim.det) {
        if (lane == 0) {
            if (dhi) atomicAdd(sim.damp_fx + 2 * w.b, (unsigned long long)dhi);
            if (dlo) atomicAdd(sim.damp_fx + 2 * w.b + 1, (unsigned long long)dlo);
            if (!dok) atomicOr(sim.status, (unsigned)DCP_STATUS_NONFINITE);
        }
    } else {
        const double s = wave_sum_d(gdamp_wave);
        if (lane == 0 && s != 0.0) atomicAdd(&sim.pgrad_inst[w.b * DCP_N_PARAMS + 4], s);
    }
}

// ---------------------------------------------------------------------------
// P2G adjoint (gather)
// ---------------------------------------------------------------------------
// Reverse of one particle's P2G scatter: dL/d(m, m v) of its 27 stencil nodes
// (fetched by `node(i, j, k)`) to the gradients of the particle's record.
//
// Per node N = (i, j, k) the scatter adds w (b + A dp) + (grad w . n) k and
// w m, with w = W0(i) W1(j) W2(k), dp = ((i, j, k) - fx) dx.  Every sum over
// the 27 nodes that its reverse needs is of the form
//   sum_N f0(i) f1(j) f2(k) g_c(N),   f in {W, D = dW/dfx, H = d2W/dfx2, W dp, D dp},
// with g_c a component of dL/d(m, m v) at N, so it is contracted one axis at
// a time: over k into T (per (i, j)), over j into U (per i), over i into the
// record gradients.  U(f1, f2) below holds sum_{j,k} f1(j) f2(k) g(i, j, k).
// The terms and their grouping follow the per-node form
//   gb += w g,  gA += w g dp^T,  gfx += dW (g_m m + g . (b + A dp)) - w dx A^T g
//   elements: gk += (dW . n) idx g,  gn += (g . k) idx dW,
//             gfx += (g . k) idx (d2W n)
// Results: gx = dL/dx, gb = dL/db / m (b = m v), gA = dL/dA / m (A = m C),
// gn = dL/dn, gk = dL/dk (the last two for elements only).
struct P2GAdj {
    V3<float> gx;
    float gb[3], gn[3], gk[3];
    M3<float> gA;
};

// ELEM: the particle is an element (normal-force terms); known per kernel.
template <bool ELEM, typename NodeFetch>
__device__ __forceinline__ void p2g_adj_gather(const InstDev& I, const P2GRec& r, const Stencil& st, NodeFetch node,
                                               P2GAdj& out) {
        const float hf[3] = {1.0f, -2.0f, 1.0f};   // d2W/dfx2 of the quadratic B-spline nodes
        const float dx = I.dx, idx = I.inv_dx;
        constexpr bool elem = ELEM;
        float* gb = out.gb;
        float* gn = out.gn;
        float* gk = out.gk;
        M3<float>& gA = out.gA;
        float gfx[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int a = 0; a < 3; ++a) gb[a] = gn[a] = gk[a] = 0.f;
        gA = mzero<float>();
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            // U(f1, f2): .x = mass channel, .yzw = momentum channels
            float4 uWW = {0.f, 0.f, 0.f, 0.f}, uDW = {0.f, 0.f, 0.f, 0.f}, uWD = {0.f, 0.f, 0.f, 0.f};
            float uPW[3] = {0.f, 0.f, 0.f}, uQW[3] = {0.f, 0.f, 0.f}, uHW[3] = {0.f, 0.f, 0.f};   // P = W dp, Q = D dp
            float uWP[3] = {0.f, 0.f, 0.f}, uDP[3] = {0.f, 0.f, 0.f}, uPD[3] = {0.f, 0.f, 0.f};
            float uDD[3] = {0.f, 0.f, 0.f}, uWQ[3] = {0.f, 0.f, 0.f}, uWH[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float4 tW = {0.f, 0.f, 0.f, 0.f}, tD = {0.f, 0.f, 0.f, 0.f};
                float tP[3] = {0.f, 0.f, 0.f}, tQ[3] = {0.f, 0.f, 0.f}, tH[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float4 gg = node(i, j, k);
                    const float w2 = st.w[k][2], d2 = st.dw[k][2], p2 = ((float)k - st.fx[2]) * dx;
                    const float wp2 = w2 * p2, dp2 = d2 * p2;
                    tW.x += w2 * gg.x; tW.y += w2 * gg.y; tW.z += w2 * gg.z; tW.w += w2 * gg.w;
                    tD.x += d2 * gg.x; tD.y += d2 * gg.y; tD.z += d2 * gg.z; tD.w += d2 * gg.w;
                    tP[0] += wp2 * gg.y; tP[1] += wp2 * gg.z; tP[2] += wp2 * gg.w;
                    tQ[0] += dp2 * gg.y; tQ[1] += dp2 * gg.z; tQ[2] += dp2 * gg.w;
                    if (elem) { tH[0] += hf[k] * gg.y; tH[1] += hf[k] * gg.z; tH[2] += hf[k] * gg.w; }
                }
                const float w1 = st.w[j][1], d1 = st.dw[j][1], p1 = ((float)j - st.fx[1]) * dx;
                const float wp1 = w1 * p1, dp1 = d1 * p1;
                uWW.x += w1 * tW.x; uWW.y += w1 * tW.y; uWW.z += w1 * tW.z; uWW.w += w1 * tW.w;
                uDW.x += d1 * tW.x; uDW.y += d1 * tW.y; uDW.z += d1 * tW.z; uDW.w += d1 * tW.w;
                uWD.x += w1 * tD.x; uWD.y += w1 * tD.y; uWD.z += w1 * tD.z; uWD.w += w1 * tD.w;
                const float tWa[3] = {tW.y, tW.z, tW.w}, tDa[3] = {tD.y, tD.z, tD.w};
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    uPW[a] += wp1 * tWa[a];
                    uQW[a] += dp1 * tWa[a];
                    uWP[a] += w1 * tP[a];
                    uDP[a] += d1 * tP[a];
                    uPD[a] += wp1 * tDa[a];
                    uWQ[a] += w1 * tQ[a];
                    if (elem) {
                        uHW[a] += hf[j] * tWa[a];
                        uDD[a] += d1 * tDa[a];
                        uWH[a] += w1 * tH[a];
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            const float w0 = st.w[i][0], d0 = st.dw[i][0], p0 = ((float)i - st.fx[0]) * dx;
            const float wp0 = w0 * p0, dp0 = d0 * p0;
            const float aWW[3] = {uWW.y, uWW.z, uWW.w}, aDW[3] = {uDW.y, uDW.z, uDW.w}, aWD[3] = {uWD.y, uWD.z, uWD.w};
            float s0 = r.m * uWW.x, s1 = r.m * uDW.x, s2 = r.m * uWD.x;       // g_m m + g . b
            float t00 = 0.f, t01 = 0.f, t02 = 0.f, t10 = 0.f, t11 = 0.f, t12 = 0.f, t20 = 0.f, t21 = 0.f, t22 = 0.f;
            float c1 = 0.f, c2 = 0.f;                                          // (A^T sum w g)_{1,2}
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float ba = (&r.b.x)[a];
                gb[a] += w0 * aWW[a];
                gA(a, 0) += wp0 * aWW[a];
                gA(a, 1) += w0 * uPW[a];
                gA(a, 2) += w0 * uWP[a];
                s0 += ba * aWW[a]; s1 += ba * aDW[a]; s2 += ba * aWD[a];
                t00 += r.A(a, 0) * aWW[a]; t01 += r.A(a, 1) * uPW[a]; t02 += r.A(a, 2) * uWP[a];
                t10 += r.A(a, 0) * aDW[a]; t11 += r.A(a, 1) * uQW[a]; t12 += r.A(a, 2) * uDP[a];
                t20 += r.A(a, 0) * aWD[a]; t21 += r.A(a, 1) * uPD[a]; t22 += r.A(a, 2) * uWQ[a];
                c1 += r.A(a, 1) * aWW[a]; c2 += r.A(a, 2) * aWW[a];
            }
            gfx[0] += d0 * (s0 + t01 + t02) + (dp0 - dx * w0) * t00;
            gfx[1] += w0 * (s1 + t11 + t12 - dx * c1) + wp0 * t10;
            gfx[2] += w0 * (s2 + t21 + t22 - dx * c2) + wp0 * t20;
            if (elem) {
                const float n0 = r.n.x, n1 = r.n.y, n2 = r.n.z;
                float kWW = 0.f, kDW = 0.f, kWD = 0.f, kHW = 0.f, kDD = 0.f, kWH = 0.f;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float ka = (&r.k.x)[a];
                    gk[a] += idx * (n0 * d0 * aWW[a] + n1 * w0 * aDW[a] + n2 * w0 * aWD[a]);
                    kWW += ka * aWW[a]; kDW += ka * aDW[a]; kWD += ka * aWD[a];
                    kHW += ka * uHW[a]; kDD += ka * uDD[a]; kWH += ka * uWH[a];
                }
                gn[0] += idx * d0 * kWW; gn[1] += idx * w0 * kDW; gn[2] += idx * w0 * kWD;
                gfx[0] += idx * (n0 * hf[i] * kWW + n1 * d0 * kDW + n2 * d0 * kWD);
                gfx[1] += idx * (n0 * d0 * kDW + n1 * w0 * kHW + n2 * w0 * kDD);
                gfx[2] += idx * (n0 * d0 * kWD + n1 * w0 * kDD + n2 * w0 * kWH);
            }
            // keep the next plane's node reads from being hoisted above this
            // one's contraction (they would all be live at once)
            __builtin_amdgcn_sched_barrier(0);
        }
        out.gx = {idx * gfx[0], idx * gfx[1], idx * gfx[2]};
}

// Store the P2G-reverse gradients of a vertex, grasper or cube particle: dL/dx
// adds to the G2P reverse's, dL/dC and dL/dv (vertices; 0 for the prescribed
// classes) are written and the vertex force-impulse gradient dL/df = dL/db is
// kept for the elements' adjoint.  A grasper's dL/dv reaches the two
// trajectory frames its velocity is the difference of (simulation.py:160-165)
// through grasper_traj_grad, summed over the wave first.
__device__ __forceinline__ void p2g_adj_store_point(const SimDev& sim, const InstDev& I, const StateView& Gout,
                                                    const Patch& P, int q, const P2GRec& r, const P2GAdj& o) {
    const int p = P.start + q;
    V3<float> prev = Gout.x(p);
    store_v3(Gout.px(0), sim.NP, p, prev + o.gx);
    store_m9(Gout.pC(0), sim.NP, p, mscale(r.m, o.gA));
    if (P.cls == CLS_VERTEX) {
        store_v3(Gout.pv(0), sim.NP, p, {r.m * o.gb[0], r.m * o.gb[1], r.m * o.gb[2]});
        store_v3(sim.g_fv, sim.NP, p, {o.gb[0], o.gb[1], o.gb[2]});
    } else {
        store_v3(Gout.pv(0), sim.NP, p, {0.f, 0.f, 0.f});
    }
}

// Trajectory gradient of a grasper patch (all of whose particles belong to one
// grasp point: patches are cut per grasp point): each lane's dL/dv = m gb
// enters frame f + 1 with +1/interval and frame f with -1/interval.  Summed
// over the wave in FP64, one pair of atomics per wave and component (all of a
// grasp point's particles add to the same six values).  Called by every lane
// of the workgroup; `act` lanes contribute.
__device__ __forceinline__ void grasper_traj_grad(const SimDev& sim, const InstDev& I, const Patch& P, int s, int g,
                                                  bool act, const P2GRec& r, const P2GAdj& o) {
    const int frame = grasper_frame(I, s, g);
    if (!sim.traj_grad || frame + 1 >= I.n_actions) return;
    const int gid = (P.start - I.o_start) / I.group_size;
    const size_t f1 = (size_t)I.traj_off + (size_t)(frame + 1) * I.n_grasp + gid;
    const size_t f0 = (size_t)I.traj_off + (size_t)frame * I.n_grasp + gid;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double gvp = wave_sum_d(act ? (double)r.m * (double)o.gb[a] / (double)I.interval : 0.0);
        if ((threadIdx.x & 63) == 0 && gvp != 0.0) {
            if (sim.det) {
                // fixed point: the wave sums of a frame add up to the same bits in any order
                unsigned long long* t1 = sim.traj_fx + 6 * f1 + 2 * a;
                unsigned long long* t0 = sim.traj_fx + 6 * f0 + 2 * a;
                const bool ok = fx_atomic_add(t1, t1 + 1, gvp) && fx_atomic_add(t0, t0 + 1, -gvp);
                if (!ok) atomicOr(sim.status, (unsigned)DCP_STATUS_NONFINITE);
            } else {
                atomicAdd(sim.traj_grad + 3 * f1 + a, gvp);
                atomicAdd(sim.traj_grad + 3 * f0 + a, -gvp);
            }
        }
    }
}

// One patch's particle records, their base-cell box and the patch tile of
// dL/d(m, m v) in LDS (when the box fits the tile cap).  Returns whether the
// tile was staged; the caller's lane holds particle tid when `act`.
// `next_box`, when given, is reset here for the next patch (after every lane
// has read this patch's box at the barrier of box_reduce).
template <bool POOL>
__device__ __forceinline__ bool p2g_adj_stage(const SimDev& sim, const InstDev& I, const StateView& S,
                                              const Patch& P, int s, int g, bool act, P2GRec& r, Stencil& st,
                                              int* box, float4* gt, PatchBox& B, int* next_box = nullptr) {
    const int tid = threadIdx.x;
    if (act) {
        p2g_record(sim, I, S, P, tid, s, g, r);
        make_stencil(r.x, I.inv_dx, I.G, st, nullptr);
    } else {
        st.base[0] = st.base[1] = st.base[2] = 0;
    }
    box_reduce(st.base[0], st.base[1], st.base[2], act, box, B);
    if (next_box && tid < 3) { next_box[tid] = 1 << 30; next_box[3 + tid] = -(1 << 30); }
    const bool tiled = B.vol <= DCP_STILE_CAP;
    if (tiled) {
#pragma unroll
        for (int h = 0; h < DCP_NODES; ++h) {
            const int c = tid + h * DCP_WG;
            if (c < B.vol) {
                const int ck = c % B.e2, t = c / B.e2;
                gt[c] = sim.gmv[cell_index<POOL>(I, t / B.e1 + B.l0, t % B.e1 + B.l1, ck + B.l2)];
            }
        }
    }
    __syncthreads();
    return tiled;
}

// P2G adjoint of the vertex, grasper and cube particles, one particle per lane
// over the point gather patches: each lane reads its stencil's 27 nodes of
// dL/d(m, m v) from the grid (no tile: no box, no barrier; the reads hit the
// L2).  Element particles are left to k_elem_adj, which needs the vertex
// force-impulse gradients stored here.
template <bool POOL>
__global__ __launch_bounds__(DCP_WG) void k_p2g_adj(SimDev sim, StateView S, StateView Gout, int s, int g) {
    // the grid-update adjoint, just before on the stream, has read the
    // adjoint scatter's poison flags
    if (blockIdx.x == 0)
        for (int b = threadIdx.x; b < sim.B; b += DCP_WG) sim.poison[sim.B + b] = 0;
    const int tid = threadIdx.x;
    for (int pi = blockIdx.x; pi < sim.n_gpoint; pi += gridDim.x) {
        const Patch P = sim.gpatches[uniform(pi)];
        const InstDev I = sim.inst[uniform(P.inst)];
        const bool act = tid < P.count;   // patches hold at most DCP_WG particles
        P2GRec r;
        P2GAdj o;
        if (act) {
            p2g_record(sim, I, S, P, tid, s, g, r);
            Stencil st;
            make_stencil(r.x, I.inv_dx, I.G, st, nullptr);
            if (POOL) {
                p2g_adj_gather<false>(I, r, st, [&](int i, int j, int k) {
                    return sim.gmv[cell_index<true>(I, st.base[0] + i, st.base[1] + j, st.base[2] + k)];
                }, o);
            } else {
                StencilOffsets so;
                stencil_offsets(I, st.base, so);
                const float4* gg = sim.gmv + I.grid_base;
                p2g_adj_gather<false>(I, r, st, [&](int i, int j, int k) { return cell_at(gg, so.x[i] + so.y[j] + so.z[k]); }, o);
            }
            p2g_adj_store_point(sim, I, Gout, P, tid, r, o);
        }
        if (P.cls == CLS_GRASPER) grasper_traj_grad(sim, I, P, s, g, act, r, o);
    }
}

// ---------------------------------------------------------------------------
// element adjoint: P2G reverse of the element particles + element prep adjoint
// ---------------------------------------------------------------------------
// Adjoint of one element's prep (kinematics, stress, QR, polar; the reverse of
// k_prep) given the gradients of its prep outputs: dL/dx_e and dL/dv_e (the
// G2P and P2G reverses), dL/dC_e from the P2G reverse, dL/dn from the P2G
// reverse plus the next state's normal, dL/dk from the P2G reverse, and the
// three vertices' force-impulse gradients (stored by k_p2g_adj).  Writes the
// element's dL/dC and dL/dn of the state, its corner gradients for the vertex
// gather, and adds its parameter gradients into acc (or the per-element
// moduli gradients).
template <typename PT, bool RM>
__device__ __forceinline__ void elem_adjoint(const SimDev& sim, const InstDev& I, const StateView& S,
                                             const StateView& G1, const StateView& Gout, int e, int fresh,
                                             const V3<float>& gxe, const V3<float>& gve, const M3<float>& gC0,
                                             const V3<float>& gn0, const V3<float>& gkv,
                                             double acc[DCP_PATCH_GRADS]) {
    const int NE = sim.NE;
    ElemKin K;
    element_kinematics(sim, I, S, e, fresh, K);
    ElemParams prm{I.E, I.nu, I.ks, I.kc};
    if (sim.per_elem) {
        prm.E = sim.emod[e]; prm.ks = sim.emod[(size_t)NE + e]; prm.kc = sim.emod[(size_t)2 * NE + e];
    }
    if (RM && !fresh) apply_return_map(sim, I, e, prm, K);
    StressState SS;
    M3<float> P = stress_forward<PT>(K.F, prm, SS);
    const float vd = I.vol_quad * I.dt;
    V3<float> r1 = {K.Dinv(0, 0), K.d(0, 1), K.d(0, 2)};
    V3<float> r2 = {K.Dinv(1, 0), K.d(1, 1), K.d(1, 2)};
    V3<float> gf1 = load_v3(sim.g_fv, sim.NP, K.i1);
    V3<float> gf2 = load_v3(sim.g_fv, sim.NP, K.i2);
    V3<float> gf3 = load_v3(sim.g_fv, sim.NP, K.i3);
    M3<float> gP = mzero<float>();
    gP(0, 2) += -vd * gkv.x; gP(1, 2) += -vd * gkv.y; gP(2, 2) += -vd * gkv.z;
    gP = madd(gP, mscale(vd, outer(gf1, r1 + r2)));
    gP = madd(gP, mscale(-vd, outer(gf2, r1)));
    gP = madd(gP, mscale(-vd, outer(gf3, r2)));
    V3<float> gr12 = vd * mtv(P, gf1);
    V3<float> gr1 = gr12 - vd * mtv(P, gf2);
    V3<float> gr2 = gr12 - vd * mtv(P, gf3);
    M3<float> gd = mzero<float>();
    gd(0, 1) += gr1.y; gd(0, 2) += gr1.z;
    gd(1, 1) += gr2.y; gd(1, 2) += gr2.z;
    V3<float> gnt = gn0 + G1.n(e);
    gd(0, 2) += gnt.x; gd(1, 2) += gnt.y; gd(2, 2) += gnt.z;
    M3<float> gF = mzero<float>();
    float gp[DCP_PATCH_GRADS] = {0.f, 0.f, 0.f, 0.f, 0.f};
    stress_adjoint<PT>(K.F, prm, SS, gP, sim.svd_grad, gF, gp);
    M3<float> gdn;
    if (!fresh) {
        gF = madd(gF, mmul(gd, mT(minv(K.Dinv))));     // d = F Dinv^-1
        if (RM) {
            // F = Q return_map(R) of the unmapped F (its QR recomputed here)
            RMState rs;
            (void)return_map_forward(K.Fr, prm.ks, prm.kc, I.friction, rs);
            M3<float> gFr = mzero<float>();
            return_map_adjoint(K.Fr, rs, prm.ks, prm.kc, I.friction, gF, gFr, gp[2], gp[3], gp[4]);
            gF = gFr;
        }
        gdn = mmul(gF, mT(K.Dinv));           // F = d_new Dinv
    } else {
        gdn = gd;                             // F = I is a constant of the fresh state
    }
    const float third = 1.0f / 3.0f;
    V3<float> c0 = {gdn(0, 0), gdn(1, 0), gdn(2, 0)};
    V3<float> c1 = {gdn(0, 1), gdn(1, 1), gdn(2, 1)};
    V3<float> gcol3 = {gdn(0, 2), gdn(1, 2), gdn(2, 2)};
    V3<float> gx1 = third * gxe - c0 - c1;
    V3<float> gx2 = third * gxe + c0;
    V3<float> gx3 = third * gxe + c1;
    // the element velocity of a fresh state is the constant 0 (simulation.py:466)
    V3<float> gvc = fresh ? V3<float>{0.f, 0.f, 0.f} : third * gve;
    M3<float> gC = gC0;
    V3<float> gnp = gcol3;
    if (!fresh) {
        gnp = gcol3 + mtv(K.M, gcol3);
        M3<float> gM = outer(gcol3, K.np);
        gC = madd(gC, madd(mscale(I.dt, gM), mscale(I.half_dt2, madd(mmul(gM, mT(K.C)), mmul(mT(K.C), gM)))));
    }
    store_m9(Gout.eC(0), NE, e, gC);
    store_v3(Gout.en(0), NE, e, gnp);
    store3(sim.ec_gx + (size_t)9 * e, gx1);        // [element][corner][3]
    store3(sim.ec_gx + (size_t)9 * e + 3, gx2);
    store3(sim.ec_gx + (size_t)9 * e + 6, gx3);
    store3(sim.ec_gv + (size_t)3 * e, gvc);        // the same for the three corners
    if (sim.per_elem) {
        sim.emod_grad[e] += (double)gp[0];
        sim.emod_grad[(size_t)NE + e] += (double)gp[2];
        sim.emod_grad[(size_t)2 * NE + e] += (double)gp[3];
        acc[1] += (double)gp[1];
        acc[4] += (double)gp[4];
    } else {
        for (int t = 0; t < DCP_PATCH_GRADS; ++t) acc[t] += (double)gp[t];
    }
}

// Element gather patches (full 256-element chunks of an instance's elements in
// their bin order), one element per lane: the P2G reverse of the element
// particle (gather of dL/d(m, m v) over the patch tile, or from the grid when
// the chunk's box exceeds the tile cap) feeds the element's prep adjoint in
// registers.  Parameter gradients are reduced per patch in FP64 in a fixed
// order.  Runs after k_p2g_adj (the vertex force-impulse gradients).
template <typename PT, bool RM, bool POOL>
__global__ __launch_bounds__(DCP_WG, 2) void k_elem_adj(SimDev sim, StateView S, StateView G1, StateView Gout,
                                                        int fresh, int s, int g) {
    // two boxes, alternating by patch: the one of the next patch is reset
    // while this patch's tile is staged, so that no barrier is needed at the
    // end of a patch (the next patch's box barrier orders the tile's reuse)
    __shared__ int box[2][6];
    __shared__ float4 gt[DCP_STILE_CAP];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid < 6) { box[tid / 3][tid % 3] = 1 << 30; box[tid / 3][3 + tid % 3] = -(1 << 30); }
    __syncthreads();
    int par = 0;
    for (int pi = blockIdx.x; pi < sim.n_gpatches - sim.n_gpoint; pi += gridDim.x) {
        const Patch P = sim.gpatches[sim.n_gpoint + pi];
        const InstDev I = sim.inst[P.inst];
        const bool act = tid < P.count;   // patches hold at most DCP_WG particles
        P2GRec r;
        Stencil st;
        PatchBox B;
        const bool tiled = p2g_adj_stage<POOL>(sim, I, S, P, s, g, act, r, st, box[par], gt, B, box[par ^ 1]);
        par ^= 1;
        double acc[DCP_PATCH_GRADS] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (act) {
            P2GAdj o;
            if (tiled) {
                const int key = tile_cell(B, st.base[0], st.base[1], st.base[2]);
                p2g_adj_gather<true>(I, r, st, [&](int i, int j, int k) { return gt[key + (i * B.e1 + j) * B.e2 + k]; }, o);
            } else {
                p2g_adj_gather<true>(I, r, st, [&](int i, int j, int k) {
                    return sim.gmv[cell_index<POOL>(I, st.base[0] + i, st.base[1] + j, st.base[2] + k)];
                }, o);
            }
            // the gather's live values end here: keep the prep adjoint's loads below
            __builtin_amdgcn_sched_barrier(0);
            const int e = P.start + tid;
            const V3<float> gxe = load_v3(sim.g_xe, sim.NE, e) + o.gx;   // G2P reverse + P2G reverse
            const V3<float> gve = {r.m * o.gb[0], r.m * o.gb[1], r.m * o.gb[2]};
            const V3<float> gn = {o.gn[0], o.gn[1], o.gn[2]};
            const V3<float> gk = {o.gk[0], o.gk[1], o.gk[2]};
            elem_adjoint<PT, RM>(sim, I, S, G1, Gout, e, fresh, gxe, gve, mscale(r.m, o.gA), gn, gk, acc);
        }
        // per (patch, wave) sums, one add per launch to each slot: the same
        // order every run (the host sums the waves and patches in order)
        double* pg = sim.pgrad_patch + ((size_t)pi * (DCP_WG / 64) + wv) * DCP_PATCH_GRADS;
        for (int t = 0; t < DCP_PATCH_GRADS; ++t) {
            if (!RM && t == 4) continue;   // no friction gradient without the return mapping
            const double v = wave_sum_d(acc[t]);
            if (lane == 0) atomicAdd(pg + t, v);
        }
    }
}

// ---------------------------------------------------------------------------
// vertex gather of the element corner gradients
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(DCP_WG) void k_vertex_final(SimDev sim, StateView Gout) {
    const int NE = sim.NE;
    for (int pi = blockIdx.x; pi < sim.n_patches; pi += gridDim.x) {
        const Patch P = sim.patches[pi];
        if (P.cls != CLS_VERTEX) continue;
        for (int q = threadIdx.x; q < P.count; q += DCP_WG) {
            const int p = P.start + q;
            V3<float> gx = Gout.x(p), gv = Gout.v(p);
            incident_sum<true>(sim, p, sim.ec_gx, sim.ec_gv, gx, gv);
            store_v3(Gout.px(0), sim.NP, p, gx);
            store_v3(Gout.pv(0), sim.NP, p, gv);
        }
    }
}

// ---------------------------------------------------------------------------
// loss adjoint: every target sends (2/3)(x - t)/N_t to its first-minimum vertex
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(DCP_WG) void k_loss_adj(SimDev sim, StateView S, StateView Gout, const float* targets,
                                                      const int* toff, const int* tcount, const int* targ,
                                                      const double* seed) {
    for (int pi = blockIdx.x; pi < sim.n_patches; pi += gridDim.x) {
        const Patch P = sim.patches[pi];
        if (P.cls != CLS_VERTEX) continue;
        const int b = P.inst;
        const int nt = tcount[b], o = toff[b];
        if (nt == 0) continue;
        const float sd = (float)seed[b];
        for (int q = threadIdx.x; q < P.count; q += DCP_WG) {
            const int p = P.start + q;
            V3<float> x = S.x(p);
            float g0 = 0.f, g1 = 0.f, g2 = 0.f;
            for (int t = 0; t < nt; ++t) {
                if (targ[o + t] == p) {
                    const float* tp = targets + 3 * (size_t)(o + t);
                    g0 += 2.0f * (x.x - tp[0]) / 3.0f;
                    g1 += 2.0f * (x.y - tp[1]) / 3.0f;
                    g2 += 2.0f * (x.z - tp[2]) / 3.0f;
                }
            }
            if (g0 != 0.f || g1 != 0.f || g2 != 0.f) {
                float sc = sd / (float)nt;
                V3<float> prev = Gout.x(p);
                store_v3(Gout.px(0), sim.NP, p, {prev.x + sc * g0, prev.y + sc * g1, prev.z + sc * g2});
            }
        }
    }
}

// The 2x2 in-plane polar rotation and its adjoint on a batch of matrices,
// one per thread (the unit entry dcp_polar_adjoint): Ft and gR row-major
// [n][4]; R = U V^T, gF = d<gR, R>/dFt (mode 0: the reference's clamped SVD
// backward, 1: the exact polar derivative).
template <typename PT>
__global__ void k_polar_adj(const float* Ft, const float* gR, int n, int mode, float* gF, float* R) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float* a = Ft + 4 * (size_t)i;
        float c, s;
        polar2<PT>(a[0], a[1], a[2], a[3], c, s);
        float g[4] = {gR[4 * (size_t)i], gR[4 * (size_t)i + 1], gR[4 * (size_t)i + 2], gR[4 * (size_t)i + 3]};
        float out[4] = {0.f, 0.f, 0.f, 0.f};
        polar2_adj<PT>(a[0], a[1], a[2], a[3], g, mode, out);
        for (int k = 0; k < 4; ++k) gF[4 * (size_t)i + k] = out[k];
        R[4 * (size_t)i] = c; R[4 * (size_t)i + 1] = -s; R[4 * (size_t)i + 2] = s; R[4 * (size_t)i + 3] = c;
    }
}
template __global__ void k_polar_adj<double>(const float*, const float*, int, int, float*, float*);
template __global__ void k_polar_adj<float>(const float*, const float*, int, int, float*, float*);

#define DCP_ELEM_ADJ(PT, RM, POOL) \
    template __global__ void k_elem_adj<PT, RM, POOL>(SimDev, StateView, StateView, StateView, int, int, int);
DCP_ELEM_ADJ(double, false, false) DCP_ELEM_ADJ(float, false, false)
DCP_ELEM_ADJ(double, true, false) DCP_ELEM_ADJ(float, true, false)
DCP_ELEM_ADJ(double, false, true) DCP_ELEM_ADJ(float, false, true)
DCP_ELEM_ADJ(double, true, true) DCP_ELEM_ADJ(float, true, true)
#undef DCP_ELEM_ADJ
#define DCP_G2P_ADJ_FORMS(GF)                                                           \
    template __global__ void k_g2p_adj<GF>(SimDev, StateView, StateView, StateView, int); \
    template __global__ void k_g2p_adj_big<GF>(SimDev, StateView, StateView, StateView, int);
DCP_G2P_ADJ_FORMS(0) DCP_G2P_ADJ_FORMS(1) DCP_G2P_ADJ_FORMS(2) DCP_G2P_ADJ_FORMS(3)
#undef DCP_G2P_ADJ_FORMS
template __global__ void k_p2g_adj<false>(SimDev, StateView, StateView, int, int);
template __global__ void k_p2g_adj<true>(SimDev, StateView, StateView, int, int);
