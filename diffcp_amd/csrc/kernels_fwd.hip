// Forward substep kernels (simulation.py:9-430) and the point-cloud loss
// (parameter_fitting_functions.py:3-15), batched over instances.
//
// One forward substep = 4 main launches (and the list kernels between them:
// k_p2g_big for patches whose tile overflows, k_active_list for the grid
// update's block lists; k_pool_alloc before k_p2g on a pooled grid):
//   k_prep     per element: x_e, v_e (vertex means), d and F rebuilt from the
//              vertex positions and the stored normal column, QR + FP64 polar,
//              stress P, the element's normal impulse vector and the three
//              vertex force impulses (simulation.py:24-42, 193-248, 386-430)
//   k_p2g      per patch: APIC scatter into an LDS tile over the patch's
//              bounding box, then one flush of the touched cells into the
//              instance grid; the element in-plane forces reach the grid with
//              the vertex particles (sum of incident-element impulses, the same
//              weights, simulation.py:266-271)
//   k_gridop   per active 4^3 block: normalise, damping, wall conditions
//              (simulation.py:274-302)
//   k_g2p      per particle of full-size gather patches: APIC gather, x/v/C
//              update, element C damping (simulation.py:304-383)
#include "kernels.hpp"
#include "patch_sort.hpp"

// ---------------------------------------------------------------------------
// element prep
// ---------------------------------------------------------------------------
template <typename PT, bool RM>
__global__ __launch_bounds__(DCP_WG) void k_prep(SimDev sim, StateView S, StateView Snext, int fresh) {
    if (blockIdx.x == 0)
        for (int b = threadIdx.x; b < sim.B; b += blockDim.x) sim.n_active[b] = 0;
    const int NE = sim.NE;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < NE; e += gridDim.x * blockDim.x) {
        const InstDev I = sim.inst[sim.elem_inst[e]];
        ElemKin K;
        element_kinematics(sim, I, S, e, fresh, K);
        ElemParams prm = elem_params(sim, I, e);
        if (RM && !fresh) apply_return_map(sim, I, e, prm, K);
        V3<float> v1 = S.v(K.i1), v2 = S.v(K.i2), v3 = S.v(K.i3);
        const float third = 1.0f / 3.0f;
        V3<float> ve = {third * (v1.x + v2.x + v3.x), third * (v1.y + v2.y + v3.y), third * (v1.z + v2.z + v3.z)};
        V3<float> n = {K.d(0, 2), K.d(1, 2), K.d(2, 2)};
        StressState SS;
        // after a rest-shape re-base the first substep's stress keeps the F of
        // the old rest shape (the reference re-bases d and Dinv but not F,
        // simulation_without_grad.py:472-486)
        const M3<float> Fs = sim.f_override ? load_m9(sim.f_override, NE, e) : K.F;
        M3<float> P = stress_forward<PT>(Fs, prm, SS);
        float vd = I.vol_quad * I.dt;
        V3<float> k = {-vd * P(0, 2), -vd * P(1, 2), -vd * P(2, 2)};
        V3<float> r1 = {K.Dinv(0, 0), K.d(0, 1), K.d(0, 2)};   // simulation.py:247
        V3<float> r2 = {K.Dinv(1, 0), K.d(1, 1), K.d(1, 2)};   // simulation.py:248
        V3<float> f1 = vd * mv(P, r1 + r2);
        V3<float> f2 = (-vd) * mv(P, r1);
        V3<float> f3 = (-vd) * mv(P, r2);
        store_v3(sim.xe, NE, e, K.xe);
        store_v3(sim.ve, NE, e, ve);
        store_v3(sim.ke, NE, e, k);
        store_v3(sim.nd, NE, e, n);
        store3(sim.fc + (size_t)9 * e, f1);          // [element][corner][3]
        store3(sim.fc + (size_t)9 * e + 3, f2);
        store3(sim.fc + (size_t)9 * e + 6, f3);
        if (Snext.p) store_v3(Snext.en(0), NE, e, n);
    }
}

// ---------------------------------------------------------------------------
// P2G
// ---------------------------------------------------------------------------
// Mark the 4^3 block holding cell (i, j, k) active: a plain store that no
// lane waits for (k_active_list, after the scatter, turns the flags into the
// per-instance lists of active blocks).
__device__ __forceinline__ void mark_block(const SimDev& sim, const InstDev& I, int i, int j, int k) {
    const int blk = ((i >> 2) * I.nb + (j >> 2)) * I.nb + (k >> 2);
    sim.flags[I.flag_base + blk] = 1;
}

// tile blocks of a patch box (the 4^3 grid blocks it overlaps)
struct TileBlocks {
    int bx0, by0, bz0, nby, nbz, n;
};
__device__ __forceinline__ TileBlocks tile_blocks(int l0, int l1, int l2, int e0, int e1, int e2) {
    TileBlocks t;
    t.bx0 = l0 >> 2; t.by0 = l1 >> 2; t.bz0 = l2 >> 2;
    int nbx = ((l0 + e0 - 1) >> 2) - t.bx0 + 1;
    t.nby = ((l1 + e1 - 1) >> 2) - t.by0 + 1;
    t.nbz = ((l2 + e2 - 1) >> 2) - t.bz0 + 1;
    t.n = nbx * t.nby * t.nbz;
    return t;
}
#define DCP_TB_CAP 256

// Particles of a patch that does not fit the sorted tile scatter straight into
// the grid with float atomics (slow path, flagged in the status word).
template <int GF>
__device__ __forceinline__ void p2g_spill(const SimDev& sim, const StateView& S, const Patch& P, const InstDev& I, int s,
                                       int g) {
    [[maybe_unused]] constexpr bool POOL = (GF & DCP_GF_POOL) != 0, DET = (GF & DCP_GF_DET) != 0;
    for (int q = threadIdx.x; q < P.count; q += DCP_WG) {
        P2GRec r;
        p2g_record(sim, I, S, P, q, s, g, r);
        Stencil st;
        make_stencil(r.x, I.inv_dx, I.G, st, sim.status);
        P2GSep sq;
        p2g_separate(I, r, st.base[0], st.base[1], st.base[2], sq);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            p2g_layer(sq, i, [&](int j, int k, float mm, float m0, float m1, float m2) {
                long long gi = cell_index<POOL>(I, st.base[0] + i, st.base[1] + j, st.base[2] + k);
                if (DET) {
                    const float v[4] = {mm, m0, m1, m2};
                    fx_add_node<4>(sim.accq, gi, v, sim.poison + P.inst);
                } else {
                    atomicAdd(&sim.acc[gi].x, mm); atomicAdd(&sim.acc[gi].y, m0);
                    atomicAdd(&sim.acc[gi].z, m1); atomicAdd(&sim.acc[gi].w, m2);
                }
            });
        // the 4^3 blocks of the 27 nodes: one or two per axis
        for (int bx = st.base[0] >> 2; bx <= (st.base[0] + 2) >> 2; ++bx)
            for (int by = st.base[1] >> 2; by <= (st.base[1] + 2) >> 2; ++by)
                for (int bz = st.base[2] >> 2; bz <= (st.base[2] + 2) >> 2; ++bz)
                    mark_block(sim, I, bx * 4, by * 4, bz * 4);
    }
}

// Add one tile node's (m, m v) into the instance grid and note its 4^3 block.
template <int GF>
__device__ __forceinline__ void p2g_flush_node(const SimDev& sim, const InstDev& I, const Patch& P, const PatchBox& B,
                                               const TileBlocks& TB, int* tb_touched, int ci, int cj, int ck, float4 a) {
    [[maybe_unused]] constexpr bool POOL = (GF & DCP_GF_POOL) != 0, DET = (GF & DCP_GF_DET) != 0;
    if (a.x == 0.f && a.y == 0.f && a.z == 0.f && a.w == 0.f) return;
    ci += B.l0; cj += B.l1; ck += B.l2;
    long long gi = cell_index<POOL>(I, ci, cj, ck);
    if (DET) {
        const float v[4] = {a.x, a.y, a.z, a.w};
        fx_add_node<4>(sim.accq, gi, v, sim.poison + P.inst);
    } else {
        atomicAdd(&sim.acc[gi].x, a.x); atomicAdd(&sim.acc[gi].y, a.y);
        atomicAdd(&sim.acc[gi].z, a.z); atomicAdd(&sim.acc[gi].w, a.w);
    }
    if (TB.n <= DCP_TB_CAP)
        tb_touched[((ci >> 2) - TB.bx0) * TB.nby * TB.nbz + ((cj >> 2) - TB.by0) * TB.nbz + ((ck >> 2) - TB.bz0)] = 1;
    else
        mark_block(sim, I, ci, cj, ck);
}

// one activation per touched grid block of the tile
__device__ __forceinline__ void p2g_mark_blocks(const SimDev& sim, const InstDev& I, const Patch& P,
                                                const TileBlocks& TB, const int* tb_touched) {
    if (TB.n > DCP_TB_CAP) return;
    for (int t = threadIdx.x; t < TB.n; t += DCP_WG) {
        if (tb_touched[t]) {
            int bz = t % TB.nbz, r = t / TB.nbz;
            int by = r % TB.nby, bx = r / TB.nby;
            mark_block(sim, I, (TB.bx0 + bx) * 4, (TB.by0 + by) * 4, (TB.bz0 + bz) * 4);
        }
    }
}

// P2G of one scatter patch (a bin's vertices and elements together, or a
// grasper or cube chunk), compute-then-gather.  One particle per lane: the lane
// builds its particle's record in registers, the patch is binned by base cell
// (patch_bin: LDS-atomic ranks, one packed scan), then three passes over the
// stencil's x-layer i: (1) each lane computes its particle's nine
// contributions to layer i and writes them to LDS slots [(j, k)][position];
// (2) one task per (occupied cell, slot) sums the cell's run; (3) the lane
// owning tile node n adds the nine run sums of the base cells n - (i, j, k).
// Node sums stay in registers across the passes and are flushed once into the
// instance grid with float atomics (tiles of neighbouring patches overlap).
// Lane q's particle record of a scatter patch and its base cell.
__device__ __forceinline__ void p2g_patch_record(const SimDev& sim, const InstDev& I, const StateView& S,
                                                 const SPatch& SP, int q, int s, int g, P2GRec& r, int& b0, int& b1,
                                                 int& b2) {
    Patch Pl;
    int ql;
    SP.lane(q, Pl, ql);
    p2g_record(sim, I, S, Pl, ql, s, g, r);
    b0 = stencil_base(r.x.x, I.inv_dx, I.G);
    b1 = stencil_base(r.x.y, I.inv_dx, I.G);
    b2 = stencil_base(r.x.z, I.inv_dx, I.G);
}

// Scatter, through one LDS tile over box B, of the patch particles whose base
// cells lie in B (`act` lanes; B is the patch's box, or one sub-box of it in
// k_p2g_big).
template <int GF>
__device__ __forceinline__ void p2g_tile(const SimDev& sim, const InstDev& I, const Patch& P, const PatchBox& B,
                                         bool act, const P2GRec& r, int b0, int b1, int b2, BinLDS& L,
                                         float4* contrib, int* tb_touched) {
    [[maybe_unused]] constexpr bool POOL = (GF & DCP_GF_POOL) != 0, DET = (GF & DCP_GF_DET) != 0;
    const int tid = threadIdx.x;
    const TileBlocks TB = tile_blocks(B.l0, B.l1, B.l2, B.e0, B.e1, B.e2);
    for (int t = tid; t < TB.n && t < DCP_TB_CAP; t += DCP_WG) tb_touched[t] = 0;
    const int pos = patch_bin(act ? tile_cell(B, b0, b1, b2) : -1, B, L);
    // this lane's tile nodes: tid + h * DCP_WG (vol <= DCP_NODES * DCP_WG), and their planes
    int nci[DCP_NODES];
    bool has[DCP_NODES];
#pragma unroll
    for (int h = 0; h < DCP_NODES; ++h) {
        const int c = tid + h * DCP_WG;
        has[h] = c < B.vol;
        nci[h] = c / (B.e1 * B.e2);
    }
    float4 a[DCP_NODES];
#pragma unroll
    for (int h = 0; h < DCP_NODES; ++h) a[h] = make_float4(0.f, 0.f, 0.f, 0.f);
    P2GSep sq;
    if (act) p2g_separate(I, r, b0, b1, b2, sq);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (act)
            p2g_layer(sq, i, [&](int j, int k, float mm, float m0, float m1, float m2) {
                contrib[(j * 3 + k) * DCP_SLOT_STRIDE + pos] = make_float4(mm, m0, m1, m2);
            });
        __syncthreads();
        run_sums(L, contrib);
        __syncthreads();
#pragma unroll
        for (int h = 0; h < DCP_NODES; ++h) {
            const float4 v = node_sum(L, B, contrib, tid + h * DCP_WG, nci[h], i, has[h]);
            a[h].x += v.x; a[h].y += v.y; a[h].z += v.z; a[h].w += v.w;
        }
        __syncthreads();
    }
#pragma unroll
    for (int h = 0; h < DCP_NODES; ++h)
        if (has[h]) {
            const int t = (tid + h * DCP_WG) / B.e2;
            p2g_flush_node<GF>(sim, I, P, B, TB, tb_touched, nci[h], t - nci[h] * B.e1,
                           tid + h * DCP_WG - t * B.e2, a[h]);
        }
    __syncthreads();
    p2g_mark_blocks(sim, I, P, TB, tb_touched);
}

// A patch whose tile overflows is listed for k_p2g_big.  The tile scatter is
// written out here rather than called as p2g_tile (k_p2g_big's): the shared
// form compiled to a schedule measured 2 % slower on this kernel.
template <int GF>
__global__ __launch_bounds__(DCP_WG, (GF & DCP_GF_DET) ? 3 : 2) void k_p2g(SimDev sim, StateView S, int s, int g, const int* skip) {
    [[maybe_unused]] constexpr bool POOL = (GF & DCP_GF_POOL) != 0, DET = (GF & DCP_GF_DET) != 0;
    if (skip && *skip) return;   // adjoint substep whose grid is restored from a snapshot
    __shared__ BinLDS L;
    __shared__ float4 contrib[DCP_CONTRIB_ZERO + 1];
    __shared__ int tb_touched[DCP_TB_CAP];
    const int tid = threadIdx.x;
    bin_clear(L);
    if (tid < 3) { L.box[tid] = 1 << 30; L.box[3 + tid] = -(1 << 30); }
    if (tid == 0) contrib[DCP_CONTRIB_ZERO] = make_float4(0.f, 0.f, 0.f, 0.f);   // never overwritten
    __syncthreads();
    for (int pi = blockIdx.x; pi < sim.n_spatches; pi += gridDim.x) {
        const SPatch SP = sim.spatches[pi];
        const Patch P = SP.part(0);      // instance and block marks
        const InstDev I = sim.inst[SP.inst];
        bool tiled = SP.count() <= DCP_WG;
        if (tiled) {
            const bool act = tid < SP.count();
            P2GRec r;
            int b0 = 0, b1 = 0, b2 = 0;
            if (act) {
                Patch Pl;
                int ql;
                SP.lane(tid, Pl, ql);
                p2g_record(sim, I, S, Pl, ql, s, g, r);
                b0 = stencil_base(r.x.x, I.inv_dx, I.G);
                b1 = stencil_base(r.x.y, I.inv_dx, I.G);
                b2 = stencil_base(r.x.z, I.inv_dx, I.G);
            }
            PatchBox B;
            box_reduce(b0, b1, b2, act, L.box, B);
            tiled = B.vol <= DCP_STILE_CAP;
            if (tiled) {
                const TileBlocks TB = tile_blocks(B.l0, B.l1, B.l2, B.e0, B.e1, B.e2);
                for (int t = tid; t < TB.n && t < DCP_TB_CAP; t += DCP_WG) tb_touched[t] = 0;
                const int pos = patch_bin(act ? tile_cell(B, b0, b1, b2) : -1, B, L);
                // this lane's tile nodes: tid + h * DCP_WG (vol <= DCP_NODES * DCP_WG), and their planes
                int nci[DCP_NODES];
                bool has[DCP_NODES];
#pragma unroll
                for (int h = 0; h < DCP_NODES; ++h) {
                    const int c = tid + h * DCP_WG;
                    has[h] = c < B.vol;
                    nci[h] = c / (B.e1 * B.e2);
                }
                float4 a[DCP_NODES];
#pragma unroll
                for (int h = 0; h < DCP_NODES; ++h) a[h] = make_float4(0.f, 0.f, 0.f, 0.f);
                P2GSep sq;
                if (act) p2g_separate(I, r, b0, b1, b2, sq);
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    if (act)
                        p2g_layer(sq, i, [&](int j, int k, float mm, float m0, float m1, float m2) {
                            contrib[(j * 3 + k) * DCP_SLOT_STRIDE + pos] = make_float4(mm, m0, m1, m2);
                        });
                    __syncthreads();
                    run_sums(L, contrib);
                    __syncthreads();
#pragma unroll
                    for (int h = 0; h < DCP_NODES; ++h) {
                        const float4 v = node_sum(L, B, contrib, tid + h * DCP_WG, nci[h], i, has[h]);
                        a[h].x += v.x; a[h].y += v.y; a[h].z += v.z; a[h].w += v.w;
                    }
                    __syncthreads();
                }
#pragma unroll
                for (int h = 0; h < DCP_NODES; ++h)
                    if (has[h]) {
                        const int t = (tid + h * DCP_WG) / B.e2;
                        p2g_flush_node<GF>(sim, I, P, B, TB, tb_touched, nci[h], t - nci[h] * B.e1,
                                       tid + h * DCP_WG - t * B.e2, a[h]);
                    }
                __syncthreads();
                p2g_mark_blocks(sim, I, P, TB, tb_touched);
            }
        }
        if (!tiled && tid == 0) {   // left to k_p2g_big
            sim.big_list[atomicAdd(&sim.big_n[DCP_BIG_P2G], 1)] = pi;
            atomicAdd(&sim.big_n[DCP_BIG_SEEN], 1);
        }
        // next patch: empty counts and box
        bin_clear(L);
        if (tid < 3) { L.box[tid] = 1 << 30; L.box[3 + tid] = -(1 << 30); }
        __syncthreads();
    }
}

// The patches k_p2g left (their tile exceeds the LDS cap: a cloth patch
// spread by folding, a wide rigid body): the same tile scatter over sub-boxes
// of the patch box whose tiles fit (sub_plan), skipping empty ones.  A patch
// of more than DCP_WG particles (none by construction) would scatter per
// particle with global atomics (status bit DCP_STATUS_TILE_SPILL).  k_gridop,
// after the list scan, empties the list.
template <int GF>
__global__ __launch_bounds__(DCP_WG, 2) void k_p2g_big(SimDev sim, StateView S, int s, int g, const int* skip) {
    [[maybe_unused]] constexpr bool POOL = (GF & DCP_GF_POOL) != 0, DET = (GF & DCP_GF_DET) != 0;
    if (skip && *skip) return;
    const int n_items = sim.big_n[DCP_BIG_P2G] * DCP_BIG_SPLIT;
    if ((int)blockIdx.x >= n_items) return;   // the usual case: an empty list
    __shared__ BinLDS L;
    __shared__ float4 contrib[DCP_CONTRIB_ZERO + 1];
    __shared__ int tb_touched[DCP_TB_CAP];
    const int tid = threadIdx.x;
    bin_clear(L);
    if (tid < 3) { L.box[tid] = 1 << 30; L.box[3 + tid] = -(1 << 30); }
    if (tid == 0) contrib[DCP_CONTRIB_ZERO] = make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const int bi = it / DCP_BIG_SPLIT, part = it - bi * DCP_BIG_SPLIT;
        const int pi = sim.big_list[bi];
        const SPatch SP = sim.spatches[pi];
        const Patch P = SP.part(0);
        const InstDev I = sim.inst[SP.inst];
        bool tiled = SP.count() <= DCP_WG;
        if (tiled) {
            const bool act = tid < SP.count();
            P2GRec r;
            int b0 = 0, b1 = 0, b2 = 0;
            if (act) p2g_patch_record(sim, I, S, SP, tid, s, g, r, b0, b1, b2);
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
                        if (!first) {   // the previous sub-box's bins and block marks are read
                            bin_clear(L);
                            __syncthreads();
                        }
                        first = false;
                        p2g_tile<GF>(sim, I, P, B, in, r, b0, b1, b2, L, contrib, tb_touched);
                    }
        }
        if (!tiled && part == 0) {
            if (tid == 0) atomicOr(sim.status, (unsigned)DCP_STATUS_TILE_SPILL);
            p2g_spill<GF>(sim, S, SP.part(0), I, s, g);
            p2g_spill<GF>(sim, S, SP.part(1), I, s, g);
        }
        bin_clear(L);
        if (tid < 3) { L.box[tid] = 1 << 30; L.box[3 + tid] = -(1 << 30); }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// grid block pool
// ---------------------------------------------------------------------------
// Before each scatter (pooled grids only): every 4^3 block a particle's
// stencil reaches (its base cell b .. b + 2 per axis: one or two blocks per
// axis) gets a pool slot if it has none.  The first lane to move a table
// entry from -1 to -2 takes the next slot and publishes it; the others go on
// (the scatter, the next launch, reads the published slots).  When the pool
// is full the block is given the trash slot and the status word says so.
__global__ __launch_bounds__(DCP_WG) void k_pool_alloc(SimDev sim, StateView S, const int* skip) {
    if (skip && *skip) return;
    for (int pi = blockIdx.x; pi < sim.n_patches; pi += gridDim.x) {
        const Patch P = sim.patches[pi];
        const InstDev I = sim.inst[P.inst];
        for (int q = threadIdx.x; q < P.count; q += DCP_WG) {
            const V3<float> x = particle_x(sim, S, P, q);
            int lo[3], hi[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int b = stencil_base(x[d], I.inv_dx, I.G);
                lo[d] = b >> 2;
                hi[d] = (b + 2) >> 2;
            }
            for (int bx = lo[0]; bx <= hi[0]; ++bx)
                for (int by = lo[1]; by <= hi[1]; ++by)
                    for (int bz = lo[2]; bz <= hi[2]; ++bz) {
                        int* t = I.btab + (bx * I.nb + by) * I.nb + bz;
                        if (*t == -1 && atomicCAS(t, -1, -2) == -1) {
                            int slot = atomicAdd(sim.pool_n, 1);
                            if (slot >= sim.pool_cap) {
                                slot = I.pool_trash;
                                atomicOr(sim.status, (unsigned)DCP_STATUS_POOL_FULL);
                            }
                            atomicExch(t, slot);
                        }
                    }
        }
    }
}

// Empty the block table (every entry -1) and hand the pool out afresh.
__global__ void k_pool_reset(int* btab, int n, int* pool_n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *pool_n = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) btab[i] = -1;
}

// ---------------------------------------------------------------------------
// grid update over the active blocks
// ---------------------------------------------------------------------------
// Per-instance lists of the blocks k_p2g flagged, for the grid update (and
// the snapshot and adjoint that follow it): workgroup (b, k) of B x per scans
// runs of DCP_LIST_RUN flags of instance b, one run per wave and pass, with
// the run's loads issued together and one count atomic per run that holds a
// flag.  n_active[b] is zero on entry (cleared by the element prep, or by the
// grid restore before a recomputed scatter).  `skip` != 0: the adjoint
// restores this substep's grid (and its list) from a snapshot instead.
// (Measured on the batched workload: 6.0 us per launch, against 13.0 us for
// one flag per lane over a chunks x B grid and 17.5 us for 32 flags per lane.)
#define DCP_LIST_U 8
#define DCP_LIST_RUN (64 * DCP_LIST_U)
__global__ __launch_bounds__(DCP_WG) void k_active_list(SimDev sim, const int* skip) {
    if (skip && *skip) return;
    const int lane = threadIdx.x & 63;
    const int per = max(1, (int)gridDim.x / sim.B);
    const int b = blockIdx.x / per;
    if (b >= sim.B) return;
    const int nw = DCP_WG / 64;
    const int first = (blockIdx.x % per) * nw + (threadIdx.x >> 6), stride = per * nw;
    const InstDev I = sim.inst[b];
    const int nbl = I.nb * I.nb * I.nb;
    const int* flags = sim.flags + I.flag_base;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int base = first * DCP_LIST_RUN; base < nbl; base += stride * DCP_LIST_RUN) {
        int f[DCP_LIST_U];
#pragma unroll
        for (int u = 0; u < DCP_LIST_U; ++u) {
            const int blk = base + u * 64 + lane;
            f[u] = blk < nbl ? flags[blk] : 0;
        }
        unsigned long long m[DCP_LIST_U];
        int n = 0;
#pragma unroll
        for (int u = 0; u < DCP_LIST_U; ++u) {
            m[u] = __ballot(f[u] != 0);
            n += __popcll(m[u]);
        }
        if (n == 0) continue;
        int slot = 0;
        if (lane == 0) slot = atomicAdd(sim.n_active + b, n);
        slot = __shfl(slot, 0, 64);
#pragma unroll
        for (int u = 0; u < DCP_LIST_U; ++u) {
            if (f[u] != 0) sim.active[I.flag_base + slot + __popcll(m[u] & below)] = base + u * 64 + lane;
            slot += __popcll(m[u]);
        }
    }
}

// Offset of instance b's blocks in a snapshot and the total, from per-instance
// block counts: lane l reads instances l, l + 64, ... (one round of loads for
// up to 64 instances), then two wave sums.  Called by every lane of the wave.
__device__ __forceinline__ void snap_offsets(const SimDev& sim, const int* counts, int b, int& base, int& total) {
    int pb = 0, pt = 0;
    for (int q = threadIdx.x & 63; q < sim.B; q += 64) {
        const int nb = sim.inst[q].nb;
        const int n = min(counts[q], nb * nb * nb);
        if (q < b) pb += n;
        pt += n;
    }
    base = wave_sum_i(pb);
    total = wave_sum_i(pt);
}

// Grid update; with `save`, the (m, m v) of the active blocks are copied into
// the snapshot before being cleared.  `skip` != 0: the adjoint restores this
// substep's grid from its snapshot instead.
__global__ __launch_bounds__(DCP_WG) void k_gridop(SimDev sim, int keep, GridSnap save, const int* skip) {
    // k_p2g_big, just before on the stream, has read the overflow list
    if (blockIdx.x == 0 && threadIdx.x == 0) sim.big_n[DCP_BIG_P2G] = 0;
    if (skip && *skip) return;
    const int lane = threadIdx.x & 63;
    ActiveWalk w;
    if (!active_walk(sim, sim.inst, w)) return;
    const InstDev I = sim.inst[w.b];
    int base = 0, total = 0;
    if (save.acc) {
        snap_offsets(sim, sim.n_active, w.b, base, total);
        if (w.first == 0 && lane == 0) save.cnt[w.b] = w.n;
        if (blockIdx.x == 0 && threadIdx.x == 0) *save.ok = total <= save.cap;
    }
    const bool saving = save.acc && total <= save.cap;
    const bool poisoned = sim.poison[w.b] != 0;
    for (int it0 = w.first; it0 < w.n; it0 += DCP_WALK_U * w.stride) {
        WalkBatch wb;
        walk_batch(sim.active + I.flag_base, I, it0, w.stride, w.n, wb);
        float4 cs[DCP_WALK_U];
#pragma unroll
        for (int u = 0; u < DCP_WALK_U; ++u)
            cs[u] = (wb.blk[u] >= 0 && !sim.det) ? sim.acc[wb.c0[u] + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int u = 0; u < DCP_WALK_U; ++u) {
            const int blk = wb.blk[u], it = it0 + u * w.stride;
            if (blk < 0) break;
            const long long cell = wb.c0[u] + lane;
            int ci, cj, ck;
            block_cell_coords(I, blk, lane, ci, cj, ck);
            float4 c = cs[u];
            if (sim.det) {
                float cv[4];
                fx_take_cell<4>(sim.accq, cell, poisoned, cv);
                c = make_float4(cv[0], cv[1], cv[2], cv[3]);
            }
            if (saving) {
                save.acc[(size_t)(base + it) * 64 + lane] = c;
                if (lane == 0) save.blk[base + it] = blk;
            }
            float v[3];
            grid_velocity(I, c, ci, cj, ck, v, nullptr);
            sim.vout[cell] = make_float4(v[0], v[1], v[2], 0.f);
            if (!(isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]))) atomicOr(sim.status, (unsigned)DCP_STATUS_NONFINITE);
            if (keep) {
                if (sim.det) sim.acc[cell] = c;    // for the grid-update adjoint, which clears it
            } else {
                if (!sim.det) sim.acc[cell] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (lane == 0) sim.flags[I.flag_base + blk] = 0;
            }
        }
    }
}

// Rebuild a substep's grid from its snapshot: (m, m v), v_out, block flags and
// the per-instance active lists (what P2G + the grid update with keep = 1
// leave behind).  When the snapshot overflowed it only clears the active-block
// counts, for the P2G that then runs.
__global__ __launch_bounds__(DCP_WG) void k_grid_restore(SimDev sim, GridSnap snap) {
    if (!*snap.ok) {
        if (blockIdx.x == 0)
            for (int b = threadIdx.x; b < sim.B; b += blockDim.x) sim.n_active[b] = 0;
        return;
    }
    const int lane = threadIdx.x & 63;
    const int per = max(1, (int)gridDim.x / sim.B);
    const int b = blockIdx.x / per;
    if (b >= sim.B) return;
    const int nw = DCP_WG / 64;
    const int first = (blockIdx.x % per) * nw + (threadIdx.x >> 6), stride = per * nw;
    const InstDev I = sim.inst[b];
    int base, total;
    snap_offsets(sim, snap.cnt, b, base, total);
    const int n = min(snap.cnt[b], I.nb * I.nb * I.nb);
    if (first == 0 && lane == 0) sim.n_active[b] = n;
    for (int it0 = first; it0 < n; it0 += DCP_WALK_U * stride) {
        WalkBatch wb;
        walk_batch(snap.blk + base, I, it0, stride, n, wb);
        float4 cs[DCP_WALK_U];
#pragma unroll
        for (int u = 0; u < DCP_WALK_U; ++u)
            cs[u] = wb.blk[u] >= 0 ? snap.acc[(size_t)(base + it0 + u * stride) * 64 + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int u = 0; u < DCP_WALK_U; ++u) {
            const int blk = wb.blk[u], it = it0 + u * stride;
            if (blk < 0) break;
            const long long cell = wb.c0[u] + lane;
            int ci, cj, ck;
            block_cell_coords(I, blk, lane, ci, cj, ck);
            const float4 c = cs[u];
            sim.acc[cell] = c;
            float v[3];
            grid_velocity(I, c, ci, cj, ck, v, nullptr);
            sim.vout[cell] = make_float4(v[0], v[1], v[2], 0.f);
            if (lane == 0) {
                sim.flags[I.flag_base + blk] = 1;
                sim.active[I.flag_base + it] = blk;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// G2P
// ---------------------------------------------------------------------------
template <bool POOL>
__global__ __launch_bounds__(DCP_WG) void k_g2p(SimDev sim, StateView S, StateView Sn) {
    const int tid = threadIdx.x;
    // the grid update, just before on the stream, has read the poison flags
    if (blockIdx.x == 0)
        for (int b = tid; b < sim.B; b += DCP_WG) sim.poison[b] = 0;
    for (int pi = blockIdx.x; pi < sim.n_gpatches; pi += gridDim.x) {
        const Patch P = sim.gpatches[uniform(pi)];
        const InstDev I = sim.inst[uniform(P.inst)];
        for (int q = tid; q < P.count; q += DCP_WG) {
            V3<float> x = particle_x(sim, S, P, q);
            Stencil st;
            // every particle passes here once per substep with the position the
            // substep's P2G scattered from, so this is where a stencil clamped
            // at the grid edge (by P2G and G2P alike) raises the status bit
            make_stencil(x, I.inv_dx, I.G, st, sim.status);
            // new_v = sum w g, new_C = 4 inv_dx sum w g dp^T (simulation.py:317-330),
            // dp = (i, j, k) - fx in cells; w = W0(i) W1(j) W2(k) is separable, so
            // the sums are contracted over k, then j, then i:
            // u*(f1, f2) = sum_{j,k} f1(j) f2(k) g with f in {W, W dp}
            V3<float> nv = {0.f, 0.f, 0.f};
            M3<float> nC = mzero<float>();
            StencilOffsets so;
            if (!POOL) stencil_offsets(I, st.base, so);
            const float4* vg = sim.vout + I.grid_base;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                float uWW[3] = {0.f, 0.f, 0.f}, uPW[3] = {0.f, 0.f, 0.f}, uWP[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    float tW[3] = {0.f, 0.f, 0.f}, tP[3] = {0.f, 0.f, 0.f};
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float4 gv = POOL ? sim.vout[cell_index<true>(I, st.base[0] + i, st.base[1] + j, st.base[2] + k)]
                                               : cell_at(vg, so.x[i] + so.y[j] + so.z[k]);
                        const float w2 = st.w[k][2], wp2 = w2 * ((float)k - st.fx[2]);
                        tW[0] += w2 * gv.x; tW[1] += w2 * gv.y; tW[2] += w2 * gv.z;
                        tP[0] += wp2 * gv.x; tP[1] += wp2 * gv.y; tP[2] += wp2 * gv.z;
                    }
                    const float w1 = st.w[j][1], wp1 = w1 * ((float)j - st.fx[1]);
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        uWW[a] += w1 * tW[a];
                        uPW[a] += wp1 * tW[a];
                        uWP[a] += w1 * tP[a];
                    }
                }
                const float w0 = st.w[i][0], wp0 = w0 * ((float)i - st.fx[0]);
                nv.x += w0 * uWW[0]; nv.y += w0 * uWW[1]; nv.z += w0 * uWW[2];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    nC(a, 0) += wp0 * uWW[a];
                    nC(a, 1) += w0 * uPW[a];
                    nC(a, 2) += w0 * uWP[a];
                }
            }
            nC = mscale(4.0f * I.inv_dx, nC);
            if (P.cls == CLS_ELEMENT) {
                const int e = P.start + q;
                M3<float> out;
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        float sk = 0.5f * (nC(a, b) - nC(b, a));
                        float sy = 0.5f * (nC(a, b) + nC(b, a));
                        out(a, b) = sk + DCP_C_SYM_KEEP * sy;        // simulation.py:381-383
                    }
                store_m9(Sn.eC(0), sim.NE, e, out);
            } else {
                const int p = P.start + q;
                bool ok = isfinite(nv.x) && isfinite(nv.y) && isfinite(nv.z);
                if (!ok) atomicOr(sim.status, (unsigned)DCP_STATUS_NONFINITE);
                store_v3(Sn.pv(0), sim.NP, p, nv);
                store_v3(Sn.px(0), sim.NP, p, {x.x + I.dt * nv.x, x.y + I.dt * nv.y, x.z + I.dt * nv.z});
                store_m9(Sn.pC(0), sim.NP, p, nC);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// loss: per target the first-minimum vertex (strict <, lowest natural index)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(DCP_WG) void k_loss_min(SimDev sim, StateView S, const float* __restrict__ targets,
                                                       const int4* __restrict__ tchunks, int n_chunks, int n_parts,
                                                       int n_targets, const int* __restrict__ vnat, float* pmin,
                                                       int* pnat, int* parg) {
    __shared__ float4 sp[DCP_WG];   // vertex x, y, z and natural index (bits), one 16-byte read per candidate
    const int tid = threadIdx.x;
    // work item = (target chunk, vertex part): part v scans the instance's
    // vertices [v n / P, (v + 1) n / P); k_loss_sum merges the parts
    for (int wi = blockIdx.x; wi < n_chunks * n_parts; wi += gridDim.x) {
        const int ci = wi / n_parts, part = wi - ci * n_parts;
        const int4 ch = tchunks[ci];            // (instance, first target, count, total targets)
        const InstDev I = sim.inst[ch.x];
        const int v0 = (int)((long long)I.v_count * part / n_parts);
        const int v1 = (int)((long long)I.v_count * (part + 1) / n_parts);
        const int t = ch.y + tid;
        const bool act = tid < ch.z;
        float t0 = 0.f, t1 = 0.f, t2 = 0.f;
        if (act) { t0 = targets[3 * t]; t1 = targets[3 * t + 1]; t2 = targets[3 * t + 2]; }
        // The first part takes its first candidate whatever its distance (the
        // reference's loop starts from the first vertex); the others start
        // from +inf, so that a NaN distance is never taken there, as it never
        // is after the first vertex.
        float best = part == 0 ? 0.f : INFINITY;
        int bnat = 0x7fffffff, bidx = -1;
        // thr: squared distances above it give a quotient by 3 strictly above
        // best (at least 4 ulps above the best's, more than the rounding of
        // the quotient can close), so only those at or below it are divided
        // and compared; after a best that small quotients could flush to zero,
        // every candidate is compared
        float thr = INFINITY;
        for (int base = v0; base < v1; base += DCP_WG) {
            __syncthreads();
            int p = base + tid;
            if (p < v1) {
                int gp = I.v_start + p;
                sp[tid] = make_float4(S.px(0)[gp], S.px(1)[gp], S.px(2)[gp], __int_as_float(vnat[gp]));
            }
            __syncthreads();
            int m = min(DCP_WG, v1 - base);
            if (act) {
                for (int u = 0; u < m; ++u) {
                    const float4 q = sp[u];
                    float d0 = q.x - t0, d1 = q.y - t1, d2 = q.z - t2;
                    const float sq = d0 * d0 + d1 * d1 + d2 * d2;
                    if (!(sq <= thr) && bidx >= 0) continue;
                    float dn = sq / 3.0f;
                    int nat = __float_as_int(q.w);
                    if ((part == 0 && bidx < 0) || dn < best || (dn == best && nat < bnat)) {
                        best = dn; bnat = nat; bidx = I.v_start + base + u;
                        thr = sq < 1e-30f ? INFINITY : sq * (1.0f + 0x1p-21f);
                    }
                }
            }
        }
        if (act) {
            const size_t o = (size_t)part * n_targets + t;
            pmin[o] = best; pnat[o] = bnat; parg[o] = bidx;
        }
    }
}

// per-instance FP64 sum in a fixed order
__global__ __launch_bounds__(DCP_WG) void k_loss_sum(const float* pmin, const int* pnat, const int* parg, int n_parts,
                                                      int n_targets, float* tmin, int* targ, const int* toff,
                                                      const int* tcount, double* loss, int B) {
    __shared__ double red[DCP_WG];
    for (int b = blockIdx.x; b < B; b += gridDim.x) {
        double acc = 0.0;
        const int n = tcount[b], o = toff[b];
        for (int q = threadIdx.x; q < n; q += DCP_WG) {
            // parts in vertex order: the first minimum of (distance, natural index)
            const int t = o + q;
            float best = pmin[t];
            int bnat = pnat[t], bidx = parg[t];
            for (int v = 1; v < n_parts; ++v) {
                const size_t i = (size_t)v * n_targets + t;
                const float dn = pmin[i];
                const int nat = pnat[i];
                if (parg[i] >= 0 && (bidx < 0 || dn < best || (dn == best && nat < bnat))) {
                    best = dn; bnat = nat; bidx = parg[i];
                }
            }
            tmin[t] = best;
            targ[t] = bidx;
            acc += (double)best;
        }
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int st = DCP_WG / 2; st > 0; st >>= 1) {
            if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) loss[b] = n > 0 ? red[0] / (double)n : 0.0;
        __syncthreads();
    }
}

// Re-derive every element's rest shape from the vertex positions of state S
// (reset_quadrature_particles, simulation_without_grad.py:472-486): d = [x1 -
// x0 | x2 - x0 | n] with n the unit normal of the normalised edges (norms with
// eps 1e-6, as initialize_ti_field) and Dinv = d^-1.  Velocities and affine
// matrices are kept.  The state carries the normal column before this
// substep's (I + dt C + dt^2/2 C C) update, which the prep applies; the
// reference overwrites d after it, so the stored column is (I + M)^-1 n.
// f_old receives the F each element has under the old rest shape, which the
// reference still uses for the stress of the next substep.
__global__ __launch_bounds__(DCP_WG) void k_rebase(SimDev sim, StateView S, float* f_old, float* Dinv) {
    const int NE = sim.NE;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < NE; e += gridDim.x * blockDim.x) {
        const InstDev I = sim.inst[sim.elem_inst[e]];
        ElemKin K;
        element_kinematics(sim, I, S, e, 0, K);
        if (sim.return_mapping) apply_return_map(sim, I, e, elem_params(sim, I, e), K);
        store_m9(f_old, NE, e, K.F);
        const V3<float> x0 = S.x(K.i1), x1 = S.x(K.i2), x2 = S.x(K.i3);
        const V3<float> d1 = x1 - x0, d2 = x2 - x0;
        const float l1 = sqrtf(d1.x * d1.x + d1.y * d1.y + d1.z * d1.z + 1e-6f);
        const float l2 = sqrtf(d2.x * d2.x + d2.y * d2.y + d2.z * d2.z + 1e-6f);
        const V3<float> n1 = {d1.x / l1, d1.y / l1, d1.z / l1}, n2 = {d2.x / l2, d2.y / l2, d2.z / l2};
        const V3<float> c3 = {n1.y * n2.z - n1.z * n2.y, n1.z * n2.x - n1.x * n2.z, n1.x * n2.y - n1.y * n2.x};
        const float l3 = sqrtf(c3.x * c3.x + c3.y * c3.y + c3.z * c3.z + 1e-6f);
        const V3<float> n3 = {c3.x / l3, c3.y / l3, c3.z / l3};
        M3<float> d;
        d(0, 0) = d1.x; d(1, 0) = d1.y; d(2, 0) = d1.z;
        d(0, 1) = d2.x; d(1, 1) = d2.y; d(2, 1) = d2.z;
        d(0, 2) = n3.x; d(1, 2) = n3.y; d(2, 2) = n3.z;
        const M3<float> di = minv(d);
        store_m9(Dinv, NE, e, di);
        store_v3(S.en(0), NE, e, mv(minv(madd(meye<float>(), K.M)), n3));
    }
}

// Base cells and sorted positions of every scatter patch's binning (the
// stable counting sort patch_bin runs inside k_p2g and k_g2p_adj), for the
// tests: out[4 (pi DCP_WG + q) + 0..3] = (b0, b1, b2, position) of lane q of
// scatter patch pi; position -1 when the patch's tile exceeds the cap.
__global__ __launch_bounds__(DCP_WG) void k_patch_order(SimDev sim, StateView S, int* out) {
    __shared__ BinLDS L;
    const int tid = threadIdx.x;
    for (int pi = blockIdx.x; pi < sim.n_spatches; pi += gridDim.x) {
        bin_clear(L);
        if (tid < 3) { L.box[tid] = 1 << 30; L.box[3 + tid] = -(1 << 30); }
        __syncthreads();
        const SPatch SP = sim.spatches[pi];
        if (SP.count() > DCP_WG) continue;   // none by construction (uniform per workgroup)
        const InstDev I = sim.inst[SP.inst];
        const bool act = tid < SP.count();
        int b0 = 0, b1 = 0, b2 = 0;
        if (act) {
            Patch Pl;
            int ql;
            SP.lane(tid, Pl, ql);
            const V3<float> x = particle_x(sim, S, Pl, ql);
            b0 = stencil_base(x.x, I.inv_dx, I.G);
            b1 = stencil_base(x.y, I.inv_dx, I.G);
            b2 = stencil_base(x.z, I.inv_dx, I.G);
        }
        PatchBox B;
        box_reduce(b0, b1, b2, act, L.box, B);
        int pos = -1;
        if (B.vol <= DCP_STILE_CAP) pos = patch_bin(act ? tile_cell(B, b0, b1, b2) : -1, B, L);
        if (act) {
            int* o = out + 4 * ((size_t)pi * DCP_WG + tid);
            o[0] = b0; o[1] = b1; o[2] = b2; o[3] = pos;
        }
        __syncthreads();
    }
}

// binning keys of a slot (integer, for the bit-exact tests)
__global__ void k_binning_keys(SimDev sim, StateView S, int* keys) {
    for (int pi = blockIdx.x; pi < sim.n_patches; pi += gridDim.x) {
        const Patch P = sim.patches[pi];
        const InstDev I = sim.inst[P.inst];
        for (int q = threadIdx.x; q < P.count; q += blockDim.x) {
            V3<float> x = particle_x(sim, S, P, q);
            int idx = (P.cls == CLS_ELEMENT) ? sim.NP + P.start + q : P.start + q;
#pragma unroll
            for (int d = 0; d < 3; ++d) keys[3 * idx + d] = base_key(x[d], I.inv_dx);
        }
    }
}

#define DCP_P2G_FORMS(GF)                                                          \
    template __global__ void k_p2g<GF>(SimDev, StateView, int, int, const int*); \
    template __global__ void k_p2g_big<GF>(SimDev, StateView, int, int, const int*);
DCP_P2G_FORMS(0) DCP_P2G_FORMS(1) DCP_P2G_FORMS(2) DCP_P2G_FORMS(3)
#undef DCP_P2G_FORMS
template __global__ void k_g2p<false>(SimDev, StateView, StateView);
template __global__ void k_g2p<true>(SimDev, StateView, StateView);
template __global__ void k_prep<double, false>(SimDev, StateView, StateView, int);
template __global__ void k_prep<float, false>(SimDev, StateView, StateView, int);
template __global__ void k_prep<double, true>(SimDev, StateView, StateView, int);
template __global__ void k_prep<float, true>(SimDev, StateView, StateView, int);


// ---------------------------------------------------------------------------
// state slot <-> caller arrays in device memory
// ---------------------------------------------------------------------------
// One thread per particle in natural order (vertices, graspers, cubes, then
// elements of every instance); perm holds the slot index of each.  The
// caller's arrays are row-major per particle: x, v [n][3], C [n][3][3],
// element C [ne][3][3] and normal column [ne][3]; a null array is skipped
// (export) or left zero (import, after the slot is cleared).
__global__ void k_state_export(const float* slot, int NP, int NE, StateIO io) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n_pts = io.n[0] + io.n[1] + io.n[2];
    if (t >= n_pts + io.n[3]) return;
    if (t < n_pts) {
        const int cls = t < io.n[0] ? 0 : (t < io.n[0] + io.n[1] ? 1 : 2);
        const int i = t - (cls > 0 ? io.n[0] : 0) - (cls > 1 ? io.n[1] : 0);
        const int p = io.perm[t];
        float* x = io.a[3 * cls];
        float* v = io.a[3 * cls + 1];
        float* C = io.a[3 * cls + 2];
        for (int k = 0; k < 3; ++k) {
            if (x) x[3 * (size_t)i + k] = slot[(size_t)k * NP + p];
            if (v) v[3 * (size_t)i + k] = slot[(size_t)(3 + k) * NP + p];
        }
        if (C) for (int k = 0; k < 9; ++k) C[9 * (size_t)i + k] = slot[(size_t)(6 + k) * NP + p];
    } else {
        const int i = t - n_pts, e = io.perm[t];
        const float* base = slot + 15 * (size_t)NP;
        if (io.a[9]) for (int k = 0; k < 9; ++k) io.a[9][9 * (size_t)i + k] = base[(size_t)k * NE + e];
        if (io.a[10]) for (int k = 0; k < 3; ++k) io.a[10][3 * (size_t)i + k] = base[(size_t)(9 + k) * NE + e];
    }
}

__global__ void k_state_import(float* slot, int NP, int NE, StateIO io) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n_pts = io.n[0] + io.n[1] + io.n[2];
    if (t >= n_pts + io.n[3]) return;
    if (t < n_pts) {
        const int cls = t < io.n[0] ? 0 : (t < io.n[0] + io.n[1] ? 1 : 2);
        const int i = t - (cls > 0 ? io.n[0] : 0) - (cls > 1 ? io.n[1] : 0);
        const int p = io.perm[t];
        const float* x = io.a[3 * cls];
        const float* v = io.a[3 * cls + 1];
        const float* C = io.a[3 * cls + 2];
        for (int k = 0; k < 3; ++k) {
            if (x) slot[(size_t)k * NP + p] = x[3 * (size_t)i + k];
            if (v) slot[(size_t)(3 + k) * NP + p] = v[3 * (size_t)i + k];
        }
        if (C) for (int k = 0; k < 9; ++k) slot[(size_t)(6 + k) * NP + p] = C[9 * (size_t)i + k];
    } else {
        const int i = t - n_pts, e = io.perm[t];
        float* base = slot + 15 * (size_t)NP;
        if (io.a[9]) for (int k = 0; k < 9; ++k) base[(size_t)k * NE + e] = io.a[9][9 * (size_t)i + k];
        if (io.a[10]) for (int k = 0; k < 3; ++k) base[(size_t)(9 + k) * NE + e] = io.a[10][3 * (size_t)i + k];
    }
}
