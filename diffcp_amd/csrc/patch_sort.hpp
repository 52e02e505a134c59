// Per-patch binning of particles by base cell for the compute-then-gather
// scatters (P2G and the G2P adjoint), and the two-stage reduction of stencil
// contributions over the bins.
//
// The base cell of a particle is trunc(x * inv_dx - 0.5) per axis
// (simulation.py:140); its key is the linear index of that cell inside the
// patch's bounding box of base cells.  The tile of a patch spans the box plus
// the 2-cell stencil reach.
#pragma once
#include "kernels.hpp"

#define DCP_SORT_MAX 256           // particles per binned patch (one per lane)
// Stride of one stencil slot in the contribution arrays: one entry of padding
// puts the nine slots of a run in different LDS banks (run_sums reads them in
// adjacent lanes; with a stride of exactly 256 entries all nine hit one bank).
#define DCP_SLOT_STRIDE (DCP_SORT_MAX + 1)
#define DCP_STILE_CAP 1024         // cells of a patch tile
#define DCP_NODES (DCP_STILE_CAP / DCP_WG)   // tile nodes per lane
static_assert(DCP_STILE_CAP <= 4 * DCP_WG, "patch_bin scans four cells per lane");

struct PatchBox {
    int l0, l1, l2, e0, e1, e2, vol;
};

__device__ __forceinline__ int stencil_base(float x, float inv_dx, int G) {
    int b = base_key(x, inv_dx);
    if (!(x == x) || b < 0) b = 0;
    if (b > G - 3) b = G - 3;
    return b;
}

// key of a base cell inside the patch tile (cells index the tile incl. the +2 halo)
__device__ __forceinline__ int tile_cell(const PatchBox& B, int i, int j, int k) {
    return ((i - B.l0) * B.e1 + (j - B.l1)) * B.e2 + (k - B.l2);
}

// ---------------------------------------------------------------------------
// Binning of a patch for the compute-then-gather scatters (one particle per
// lane, P.count <= DCP_WG), stable: a particle's sorted position is the
// offset of its cell plus the number of lanes before it (in lane order, i.e.
// particle order) whose particle has the same cell, so the permutation is a
// stable argsort of the cell keys.
//   1. per wave, ten ballots over the key bits give each lane the mask of the
//      wave's lanes with its key; its rank in the wave is the population count
//      of that mask below the lane;
//   2. the first lane of each distinct key adds the key's count in its wave
//      into byte `wave` of the cell's count word (LDS atomics; at most 64 per
//      byte), so the word holds the four per-wave counts in wave order;
//   3. after a barrier each lane adds the counts of the earlier waves (the
//      bytes below its wave) to its rank, and one block scan of packed
//      (count, occupied) words gives every cell's offset and every occupied
//      cell's run index, and the run table.
// No permutation is stored: a particle's sorted position is offset(cell) + rank.
//
// The offset table is indexed by tile cell and framed for the node sums: a
// margin of DCP_OFF_MARGIN zero entries before cell 0 and the particle count
// after cell vol - 1, so that cell c's run is [off(c), off(c + 1)) for every
// c in [-DCP_OFF_MARGIN, vol).
// ---------------------------------------------------------------------------
#define DCP_OFF_MARGIN 256         // >= 2 e2 + 2 for every tile (e2 <= DCP_STILE_CAP / 9)
static_assert(DCP_OFF_MARGIN >= 2 * (DCP_STILE_CAP / 9) + 2, "node_sum reaches 2 e2 + 2 cells before cell 0");
static_assert(DCP_WG == 256, "patch_bin packs one count byte per wave of a 256-lane workgroup");
static_assert(DCP_STILE_CAP <= 1024, "patch_bin matches 10-bit keys");
struct BinLDS {
    int box[6];
    int wsum[DCP_WG / 64];
    int n_runs;
    int off_mem[DCP_OFF_MARGIN + DCP_STILE_CAP + 1];   // margin | per-wave counts, then exclusive offsets | count
    unsigned short run_start[DCP_SORT_MAX];    // run (occupied cell) -> first position
    unsigned short run_len[DCP_SORT_MAX];
    __device__ int* off() { return off_mem + DCP_OFF_MARGIN; }
    __device__ const int* off() const { return off_mem + DCP_OFF_MARGIN; }
};

// Zero the margin, the cell counts and the count slot; the box barrier that
// follows orders this before use.
__device__ __forceinline__ void bin_clear(BinLDS& L) {
    for (int c = threadIdx.x; c < DCP_OFF_MARGIN + DCP_STILE_CAP + 1; c += DCP_WG) L.off_mem[c] = 0;
}

// Lanes of this wave whose key equals this lane's (keys in [0, 1024), or < 0
// for lanes without a particle, which match each other only).  Called by
// every lane of the wave.
__device__ __forceinline__ unsigned long long same_key_lanes(int key) {
    unsigned long long m = __ballot(key >= 0);
    if (key < 0) m = ~m;
#pragma unroll
    for (int b = 0; b < 10; ++b) {
        const unsigned long long set = __ballot((key >> b) & 1);
        m &= ((key >> b) & 1) ? set : ~set;
    }
    return m;
}

// sum of the count bytes of waves 0 .. w-1 in a packed count word
__device__ __forceinline__ int bytes_below(unsigned word, int w) {
    int s = 0;
#pragma unroll
    for (int q = 0; q < DCP_WG / 64 - 1; ++q)
        if (q < w) s += (word >> (8 * q)) & 0xff;
    return s;
}
__device__ __forceinline__ int bytes_sum(unsigned word) {
    return (int)((word & 0xff) + ((word >> 8) & 0xff) + ((word >> 16) & 0xff) + (word >> 24));
}

// `key` < 0 for a lane without a particle.  Returns the sorted position.
// Called by every lane of the workgroup.
__device__ __forceinline__ int patch_bin(int key, const PatchBox& B, BinLDS& L) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int* off = L.off();
    const unsigned long long same = same_key_lanes(key);
    int rank = __popcll(same & ((1ull << lane) - 1ull));
    if (key >= 0 && rank == 0) atomicAdd(&off[key], __popcll(same) << (8 * wv));
    __syncthreads();
    if (key >= 0) rank += bytes_below((unsigned)off[key], wv);
    int v[4], sum = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int c = tid * 4 + t;
        const int n = c < B.vol ? bytes_sum((unsigned)off[c]) : 0;
        v[t] = n | (n > 0 ? 1 << 16 : 0);
        sum += v[t];
    }
    int incl = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) L.wsum[wv] = incl;
    __syncthreads();
    int run = incl - sum, total = 0;
#pragma unroll
    for (int w = 0; w < DCP_WG / 64; ++w) {
        if (w < wv) run += L.wsum[w];
        total += L.wsum[w];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int c = tid * 4 + t;
        if (c < B.vol) {
            const int n = v[t] & 0xffff;
            off[c] = run & 0xffff;
            if (n) {
                L.run_start[run >> 16] = (unsigned short)(run & 0xffff);
                L.run_len[run >> 16] = (unsigned short)n;
            }
        }
        run += v[t];
    }
    if (tid == 0) {
        L.n_runs = total >> 16;
        off[B.vol] = total & 0xffff;
    }
    __syncthreads();
    return key >= 0 ? off[key] + rank : -1;
}

// Two-stage reduction of one stencil x-layer's contributions.  `contrib`
// holds one float4 per (slot, sorted position) at [slot * DCP_SLOT_STRIDE +
// position], slot = (j, k) of the node relative to the particle's base cell.
//
// Stage 1 (run_sums): one task per (run, slot) sums the run's entries of that
// slot and leaves the sum in the run's first entry.  Loads go eight at a time
// (the dense cells of a fine cloth hold tens of particles).
__device__ __forceinline__ void add4(float4& a, const float4& b) { a.x += b.x; a.y += b.y; a.z += b.z; a.w += b.w; }
__device__ __forceinline__ void run_sums(const BinLDS& L, float4* contrib) {
    const int ntask = 9 * L.n_runs;
    for (int t = threadIdx.x; t < ntask; t += DCP_WG) {
        const int slot = t / L.n_runs, r = t - slot * L.n_runs;
        float4* c = contrib + slot * DCP_SLOT_STRIDE + L.run_start[r];
        const int len = L.run_len[r];
        float4 a = c[0];
        int u = 1;
        for (; u + 7 < len; u += 8) {
            float4 b[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) b[q] = c[u + q];
            add4(b[0], b[1]); add4(b[2], b[3]); add4(b[4], b[5]); add4(b[6], b[7]);
            add4(b[0], b[2]); add4(b[4], b[6]); add4(b[0], b[4]);
            add4(a, b[0]);
        }
        for (; u + 3 < len; u += 4) {
            const float4 b0 = c[u], b1 = c[u + 1], b2 = c[u + 2], b3 = c[u + 3];
            a.x += (b0.x + b1.x) + (b2.x + b3.x); a.y += (b0.y + b1.y) + (b2.y + b3.y);
            a.z += (b0.z + b1.z) + (b2.z + b3.z); a.w += (b0.w + b1.w) + (b2.w + b3.w);
        }
        for (; u < len; ++u) {
            const float4 b = c[u];
            a.x += b.x; a.y += b.y; a.z += b.z; a.w += b.w;
        }
        c[0] = a;
    }
}

// Stage 2 (node_sum): the tile node c = (ci, cj, ck) adds, for its nine base
// cells c - (i, dj, dk), the run sum of slot (dj, dk).  No bounds tests: the
// tile's halo cells (the last two planes, rows and columns of the tile) never
// hold a base cell, so an offset that leaves the base-cell range in j or k
// lands on a halo cell of the row or plane before, or in the table's margin,
// and reads as an empty run; only the plane (ci >= i) is tested.  Per row
// (dj) the four offsets off(c' - 2 .. c' + 1) bound the three cells' runs.  An
// empty cell reads the zero entry at DCP_CONTRIB_ZERO.
#define DCP_CONTRIB_ZERO (9 * DCP_SLOT_STRIDE)    // contrib arrays hold DCP_CONTRIB_ZERO + 1 entries
__device__ __forceinline__ void node_runs(const BinLDS& L, const PatchBox& B, int c, int i, int idx[9]) {
    const int* o0 = L.off() + c - i * B.e1 * B.e2;
#pragma unroll
    for (int dj = 0; dj < 3; ++dj) {
        const int* r = o0 - dj * B.e2;
        const int q[4] = {r[-2], r[-1], r[0], r[1]};
#pragma unroll
        for (int dk = 0; dk < 3; ++dk) {
            const int a = q[2 - dk], e = q[3 - dk];
            idx[dj * 3 + dk] = e > a ? (dj * 3 + dk) * DCP_SLOT_STRIDE + a : DCP_CONTRIB_ZERO;
        }
    }
}
__device__ __forceinline__ float4 node_sum(const BinLDS& L, const PatchBox& B, const float4* contrib, int c, int ci,
                                           int i, bool has) {
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!has || ci < i) return acc;
    int idx[9];
    node_runs(L, B, c, i, idx);
    float4 x[9];
#pragma unroll
    for (int o = 0; o < 9; ++o) x[o] = contrib[idx[o]];
    add4(x[0], x[1]); add4(x[2], x[3]); add4(x[4], x[5]); add4(x[6], x[7]);
    add4(x[0], x[2]); add4(x[4], x[6]); add4(x[0], x[4]); add4(x[0], x[8]);
    return x[0];
}

// The same two stages for three-channel contributions (the G2P adjoint's
// dL/d grid_v_out), kept as a float2 array (channels 0, 1) and a float array
// (channel 2): 12 B per entry instead of 16, so the workgroup's LDS leaves
// room for a third workgroup per CU.
__device__ __forceinline__ void run_sums3(const BinLDS& L, float2* ca, float* cb) {
    const int ntask = 9 * L.n_runs;
    for (int t = threadIdx.x; t < ntask; t += DCP_WG) {
        const int slot = t / L.n_runs, r = t - slot * L.n_runs;
        const int base = slot * DCP_SLOT_STRIDE + L.run_start[r];
        const int len = L.run_len[r];
        float2* a2 = ca + base;
        float* b1 = cb + base;
        float2 a = a2[0];
        float c = b1[0];
        int u = 1;
        for (; u + 7 < len; u += 8) {
            float2 x[8];
            float y[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) { x[q] = a2[u + q]; y[q] = b1[u + q]; }
            a.x += ((x[0].x + x[1].x) + (x[2].x + x[3].x)) + ((x[4].x + x[5].x) + (x[6].x + x[7].x));
            a.y += ((x[0].y + x[1].y) + (x[2].y + x[3].y)) + ((x[4].y + x[5].y) + (x[6].y + x[7].y));
            c += ((y[0] + y[1]) + (y[2] + y[3])) + ((y[4] + y[5]) + (y[6] + y[7]));
        }
        for (; u + 3 < len; u += 4) {
            const float2 x0 = a2[u], x1 = a2[u + 1], x2 = a2[u + 2], x3 = a2[u + 3];
            const float y0 = b1[u], y1 = b1[u + 1], y2 = b1[u + 2], y3 = b1[u + 3];
            a.x += (x0.x + x1.x) + (x2.x + x3.x);
            a.y += (x0.y + x1.y) + (x2.y + x3.y);
            c += (y0 + y1) + (y2 + y3);
        }
        for (; u < len; ++u) {
            const float2 x = a2[u];
            a.x += x.x; a.y += x.y; c += b1[u];
        }
        a2[0] = a;
        b1[0] = c;
    }
}

__device__ __forceinline__ float3 node_sum3(const BinLDS& L, const PatchBox& B, const float2* ca, const float* cb,
                                            int c, int ci, int i, bool has) {
    float3 acc = make_float3(0.f, 0.f, 0.f);
    if (!has || ci < i) return acc;
    int idx[9];
    node_runs(L, B, c, i, idx);
    float2 x[9];
    float y[9];
#pragma unroll
    for (int o = 0; o < 9; ++o) { x[o] = ca[idx[o]]; y[o] = cb[idx[o]]; }
    acc.x = (((x[0].x + x[1].x) + (x[2].x + x[3].x)) + ((x[4].x + x[5].x) + (x[6].x + x[7].x))) + x[8].x;
    acc.y = (((x[0].y + x[1].y) + (x[2].y + x[3].y)) + ((x[4].y + x[5].y) + (x[6].y + x[7].y))) + x[8].y;
    acc.z = (((y[0] + y[1]) + (y[2] + y[3])) + ((y[4] + y[5]) + (y[6] + y[7]))) + y[8];
    return acc;
}

// A patch whose tile exceeds DCP_STILE_CAP cells is scattered block by block:
// its box of base cells is cut into sub-boxes of at most w0 x w1 x w2 base
// cells whose tiles ((w + 2) per axis) fit the cap, k extents first (whole
// rows of the box when they fit), then j, then i.  Since 3^3 <= the cap, every
// box can be cut so.
struct SubPlan {
    int n[3], w[3];   // base-cell extents of the box, sub-box extents
};
__device__ __forceinline__ SubPlan sub_plan(const PatchBox& B) {
    SubPlan p;
    p.n[0] = B.e0 - 2; p.n[1] = B.e1 - 2; p.n[2] = B.e2 - 2;
    p.w[2] = min(p.n[2], DCP_STILE_CAP / 9 - 2);
    p.w[1] = min(p.n[1], DCP_STILE_CAP / (3 * (p.w[2] + 2)) - 2);
    p.w[0] = min(p.n[0], DCP_STILE_CAP / ((p.w[1] + 2) * (p.w[2] + 2)) - 2);
    return p;
}
// sub-box with base-cell offset (x0, y0, z0) inside the patch box
__device__ __forceinline__ PatchBox sub_box(const PatchBox& B, const SubPlan& p, int x0, int y0, int z0) {
    PatchBox S = B;
    S.l0 = B.l0 + x0; S.l1 = B.l1 + y0; S.l2 = B.l2 + z0;
    S.e0 = min(p.w[0], p.n[0] - x0) + 2;
    S.e1 = min(p.w[1], p.n[1] - y0) + 2;
    S.e2 = min(p.w[2], p.n[2] - z0) + 2;
    S.vol = S.e0 * S.e1 * S.e2;
    return S;
}
__device__ __forceinline__ bool in_sub_box(const PatchBox& S, int b0, int b1, int b2) {
    return b0 >= S.l0 && b0 < S.l0 + S.e0 - 2 && b1 >= S.l1 && b1 < S.l1 + S.e1 - 2 && b2 >= S.l2 &&
           b2 < S.l2 + S.e2 - 2;
}

// The sub-boxes of a listed patch are shared out over DCP_BIG_SPLIT work
// items (sub-box k goes to item k mod DCP_BIG_SPLIT), so that the few patches
// of a folded cloth that overflow do not each run all their sub-boxes in one
// workgroup; every item of a patch loads its records and box again.
#define DCP_BIG_SPLIT 8

// Lists of the patches whose tile exceeds the cap, filled by k_p2g /
// k_g2p_adj, worked off by k_p2g_big / k_g2p_adj_big and emptied by the grid
// update that follows each (k_gridop / k_gridop_adj): lengths in big_n[0..1],
// list l at big_list[l * n_spatches ...].
#define DCP_BIG_P2G 0
#define DCP_BIG_G2P_ADJ 1
// big_n[DCP_BIG_SEEN]: patches k_p2g listed since the host last cleared it
// (after each forward or recompute segment): zero means no adjoint substep
// of that segment lists one either (same states, same patches, same test),
// so its adjoint graph can leave k_g2p_adj_big out
#define DCP_BIG_SEEN 2

// Box of the patch's base cells from one candidate per lane.
__device__ __forceinline__ void box_reduce(int b0, int b1, int b2, bool act, int* box, PatchBox& B) {
    const int big = 1 << 30;
    const int lo0 = wave_min(act ? b0 : big), lo1 = wave_min(act ? b1 : big), lo2 = wave_min(act ? b2 : big);
    const int hi0 = wave_max(act ? b0 : -big), hi1 = wave_max(act ? b1 : -big), hi2 = wave_max(act ? b2 : -big);
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&box[0], lo0); atomicMin(&box[1], lo1); atomicMin(&box[2], lo2);
        atomicMax(&box[3], hi0); atomicMax(&box[4], hi1); atomicMax(&box[5], hi2);
    }
    __syncthreads();
    B.l0 = box[0]; B.l1 = box[1]; B.l2 = box[2];
    B.e0 = box[3] - B.l0 + 3; B.e1 = box[4] - B.l1 + 3; B.e2 = box[5] - B.l2 + 3;
    B.vol = B.e0 * B.e1 * B.e2;
}
