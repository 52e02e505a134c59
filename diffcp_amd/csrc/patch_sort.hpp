// Per-patch binning: bounding box of the base cells, stable LDS counting sort
// of the patch's particles by base cell, and the wave-level segmented
// reduction used by the P2G-type scatters.
//
// The base cell of a particle is trunc(x * inv_dx - 0.5) per axis
// (simulation.py:140); its key is the linear index of that cell inside the
// patch's bounding box.  The sort is a counting sort (histogram, exclusive
// scan, scatter) whose ranks are stable: inside a wave the rank of a lane is
// the number of lower lanes with the same key (ballot match), and the waves of
// a round add their counts in wave order.  The resulting order equals a stable
// argsort of the keys, so lanes holding the same base cell are contiguous.
#pragma once
#include "kernels.hpp"

#define DCP_SORT_MAX 256           // particles per sorted patch
#define DCP_STILE_CAP 512          // cells of a sorted-scatter tile (wave-private copies)
#define DCP_STAGE_MAX 256          // particles of a patch staged in LDS by the scatters

struct PatchBox {
    int l0, l1, l2, e0, e1, e2, vol;
};

// Per-lane segment description of a wave of sorted keys: bit s of `same` says
// that the lane 2^s below holds the same key; `tail` marks the last lane of a
// run; `steps` is the wave-uniform number of scan steps needed.
struct SegInfo {
    unsigned same;
    bool tail;
    int steps;
};

struct SortLDS {
    int box[6];
    int cnt[DCP_STILE_CAP];                    // per-key counts, then exclusive offsets
    unsigned short key[DCP_SORT_MAX];          // key of particle q (input order)
    unsigned short rank[DCP_SORT_MAX];         // stable rank of q among its key
    unsigned short perm[DCP_SORT_MAX];         // sorted position -> q
    unsigned short skey[DCP_SORT_MAX];         // sorted position -> key
    unsigned short ch_start[DCP_SORT_MAX];     // chunk -> first sorted position
    unsigned short ch_len[DCP_SORT_MAX];       // chunk -> particles (<= K)
    unsigned short ch_key[DCP_SORT_MAX];       // chunk -> key
    int wsum[DCP_WG / 64];
    int n_chunks;
};

__device__ __forceinline__ int stencil_base(float x, float inv_dx, int G) {
    int b = (int)(x * inv_dx - 0.5f);
    if (!(x == x) || b < 0) b = 0;
    if (b > G - 3) b = G - 3;
    return b;
}

__device__ __forceinline__ void patch_bbox(const SimDev& sim, const StateView& S, const Patch& P, const InstDev& I,
                                           SortLDS& L, PatchBox& B) {
    const int tid = threadIdx.x;
    if (tid < 3) { L.box[tid] = 1 << 30; L.box[3 + tid] = -(1 << 30); }
    __syncthreads();
    int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    for (int q = tid; q < P.count; q += DCP_WG) {
        V3<float> x = particle_x(sim, S, P, q);
        int b0 = stencil_base(x.x, I.inv_dx, I.G), b1 = stencil_base(x.y, I.inv_dx, I.G),
            b2 = stencil_base(x.z, I.inv_dx, I.G);
        lo[0] = min(lo[0], b0); lo[1] = min(lo[1], b1); lo[2] = min(lo[2], b2);
        hi[0] = max(hi[0], b0); hi[1] = max(hi[1], b1); hi[2] = max(hi[2], b2);
    }
    for (int d = 0; d < 3; ++d) {
        int a = wave_min(lo[d]), b = wave_max(hi[d]);
        if ((tid & 63) == 0) { atomicMin(&L.box[d], a); atomicMax(&L.box[3 + d], b); }
    }
    __syncthreads();
    B.l0 = L.box[0]; B.l1 = L.box[1]; B.l2 = L.box[2];
    B.e0 = L.box[3] - B.l0 + 3; B.e1 = L.box[4] - B.l1 + 3; B.e2 = L.box[5] - B.l2 + 3;
    B.vol = B.e0 * B.e1 * B.e2;
    __syncthreads();
}

// Bounding box from per-thread min/max base cells (block reduction).
__device__ __forceinline__ void bbox_reduce(int lo0, int lo1, int lo2, int hi0, int hi1, int hi2, SortLDS& L,
                                            PatchBox& B) {
    const int tid = threadIdx.x;
    lo0 = wave_min(lo0); lo1 = wave_min(lo1); lo2 = wave_min(lo2);
    hi0 = wave_max(hi0); hi1 = wave_max(hi1); hi2 = wave_max(hi2);
    if ((tid & 63) == 0) {
        atomicMin(&L.box[0], lo0); atomicMin(&L.box[1], lo1); atomicMin(&L.box[2], lo2);
        atomicMax(&L.box[3], hi0); atomicMax(&L.box[4], hi1); atomicMax(&L.box[5], hi2);
    }
    __syncthreads();
    B.l0 = L.box[0]; B.l1 = L.box[1]; B.l2 = L.box[2];
    B.e0 = L.box[3] - B.l0 + 3; B.e1 = L.box[4] - B.l1 + 3; B.e2 = L.box[5] - B.l2 + 3;
    B.vol = B.e0 * B.e1 * B.e2;
    __syncthreads();
}

__device__ __forceinline__ int pack_base(int b0, int b1, int b2) { return (b0 << 20) | (b1 << 10) | b2; }
__device__ __forceinline__ void unpack_base(int p, int& b0, int& b1, int& b2) {
    b0 = p >> 20; b1 = (p >> 10) & 1023; b2 = p & 1023;
}

// key of a base cell inside the patch tile (cells index the tile incl. the +2 halo)
__device__ __forceinline__ int tile_cell(const PatchBox& B, int i, int j, int k) {
    return ((i - B.l0) * B.e1 + (j - B.l1)) * B.e2 + (k - B.l2);
}

// Stable counting sort of the patch's particles by tile cell of their base.
// Requires B.vol <= DCP_STILE_CAP and P.count <= DCP_SORT_MAX.  With
// keys_ready the caller has already written L.key[q] for every particle.
__device__ __forceinline__ void patch_sort(const SimDev& sim, const StateView& S, const Patch& P, const InstDev& I,
                                           const PatchBox& B, SortLDS& L, bool keys_ready = false) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int c = tid; c < B.vol; c += DCP_WG) L.cnt[c] = 0;
    __syncthreads();
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r0 = 0; r0 < P.count; r0 += DCP_WG) {
        const int q = r0 + tid;
        const bool act = q < P.count;
        int key = -1;
        if (act) {
            if (keys_ready) {
                key = L.key[q];
            } else {
                V3<float> x = particle_x(sim, S, P, q);
                key = tile_cell(B, stencil_base(x.x, I.inv_dx, I.G), stencil_base(x.y, I.inv_dx, I.G),
                                stencil_base(x.z, I.inv_dx, I.G));
                L.key[q] = (unsigned short)key;
            }
        }
        // rank among lower lanes of this wave with the same key, and the group size
        unsigned long long todo = __ballot(act);
        int wrank = 0, wcount = 0;
        bool leader = false;
        while (todo) {
            int first = __ffsll((long long)todo) - 1;
            int k0 = __shfl(key, first, 64);
            unsigned long long grp = __ballot(act && key == k0);
            if (act && key == k0) {
                wrank = __popcll(grp & lt_mask);
                wcount = __popcll(grp);
                leader = (lane == first);
            }
            todo &= ~grp;
        }
        // waves of the round add their counts in wave order
#pragma unroll
        for (int w = 0; w < DCP_WG / 64; ++w) {
            if (wv == w && act) {
                int old = L.cnt[key];
                L.rank[q] = (unsigned short)(old + wrank);
                if (leader) L.cnt[key] = old + wcount;
            }
            __syncthreads();
        }
    }
    // exclusive scan of the counts (vol <= DCP_STILE_CAP <= 4 per thread)
    int v[4], s = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        int c = tid * 4 + t;
        v[t] = c < B.vol ? L.cnt[c] : 0;
        s += v[t];
    }
    int incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) L.wsum[wv] = incl;
    __syncthreads();
    int wbase = 0;
    for (int w = 0; w < wv; ++w) wbase += L.wsum[w];
    int run = wbase + incl - s;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        int c = tid * 4 + t;
        if (c < B.vol) L.cnt[c] = run;
        run += v[t];
    }
    __syncthreads();
    for (int q = tid; q < P.count; q += DCP_WG) {
        int k = L.key[q];
        int pos = L.cnt[k] + L.rank[q];
        L.perm[pos] = (unsigned short)q;
        L.skey[pos] = (unsigned short)k;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Binning of a patch for the compute-then-gather scatters (one particle per
// lane, P.count <= DCP_WG): per-cell counts with LDS atomics (the returned old
// count is the particle's rank inside its cell; the order among particles of
// one cell follows the LDS atomic order), then one block scan of packed
// (count, occupied) words giving every cell's offset and every occupied
// cell's run index, and the run table.  No permutation is stored: a particle's
// sorted position is offset(cell) + rank.
// ---------------------------------------------------------------------------
struct BinLDS {
    int box[6];
    int wsum[DCP_WG / 64];
    int n_runs;
    int off[DCP_STILE_CAP];                    // counts, then exclusive offsets
    unsigned short run_start[DCP_SORT_MAX];    // run (occupied cell) -> first position
    unsigned short run_len[DCP_SORT_MAX];
};

// Zero the cell counts; the box barrier that follows orders this before use.
__device__ __forceinline__ void bin_clear(BinLDS& L) {
    for (int c = threadIdx.x; c < DCP_STILE_CAP; c += DCP_WG) L.off[c] = 0;
}

// `key` < 0 for a lane without a particle.  Returns the sorted position.
__device__ __forceinline__ int patch_bin(int key, const PatchBox& B, BinLDS& L) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int rank = key >= 0 ? atomicAdd(&L.off[key], 1) : 0;
    __syncthreads();
    int v[4], sum = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int c = tid * 4 + t;
        const int n = c < B.vol ? L.off[c] : 0;
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
            L.off[c] = run & 0xffff;
            if (n) {
                L.run_start[run >> 16] = (unsigned short)(run & 0xffff);
                L.run_len[run >> 16] = (unsigned short)n;
            }
        }
        run += v[t];
    }
    if (tid == 0) L.n_runs = total >> 16;
    __syncthreads();
    return key >= 0 ? L.off[key] + rank : -1;
}

// Two-stage reduction of one stencil x-layer's contributions.  `contrib`
// holds one float4 per (slot, sorted position) at [slot * DCP_SORT_MAX +
// position], slot = (j, k) of the node relative to the particle's base cell.
//
// Stage 1 (run_sums): one task per (run, slot) sums the run's entries of that
// slot and leaves the sum in the run's first entry.  Loads go four at a time.
__device__ __forceinline__ void run_sums(const BinLDS& L, float4* contrib) {
    const int ntask = 9 * L.n_runs;
    for (int t = threadIdx.x; t < ntask; t += DCP_WG) {
        const int r = t / 9, slot = t - 9 * r;
        float4* c = contrib + slot * DCP_SORT_MAX + L.run_start[r];
        const int len = L.run_len[r];
        float4 a = c[0];
        int u = 1;
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

// Stage 2 (node_sum): the tile node (ci, cj, ck) adds, for its nine base cells
// node - (i, dj, dk), the run sum of slot (dj, dk) — nine independent loads,
// the same count for every node.  `has` = false contributes nothing.
__device__ __forceinline__ float4 node_sum(const BinLDS& L, const PatchBox& B, int count, const float4* contrib,
                                           int ci, int cj, int ck, int i, bool has) {
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const int bi = ci - i;
    if (!has || bi < 0 || bi > B.e0 - 3) return acc;
#pragma unroll
    for (int o = 0; o < 9; ++o) {
        const int dj = o / 3, dk = o % 3;
        const int bj = cj - dj, bk = ck - dk;
        if (bj >= 0 && bj <= B.e1 - 3 && bk >= 0 && bk <= B.e2 - 3) {
            const int kb = (bi * B.e1 + bj) * B.e2 + bk;
            const int a = L.off[kb];
            const int e = kb + 1 < B.vol ? L.off[kb + 1] : count;
            if (e > a) {
                const float4 v = contrib[o * DCP_SORT_MAX + a];
                acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
            }
        }
    }
    return acc;
}

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

// Block-wide exclusive scan of one int per key (vol <= DCP_STILE_CAP, 4 keys
// per thread); returns the total.  `vals` are this thread's 4 values.
__device__ __forceinline__ int block_exscan4(int vals[4], int out[4], SortLDS& L) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int s = vals[0] + vals[1] + vals[2] + vals[3];
    int incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    __syncthreads();
    if (lane == 63) L.wsum[wv] = incl;
    __syncthreads();
    int wbase = 0, total = 0;
#pragma unroll
    for (int w = 0; w < DCP_WG / 64; ++w) {
        if (w < wv) wbase += L.wsum[w];
        total += L.wsum[w];
    }
    int run = wbase + incl - s;
#pragma unroll
    for (int t = 0; t < 4; ++t) { out[t] = run; run += vals[t]; }
    return total;
}

// Cut every key's run of sorted particles into chunks of at most K; chunks of
// one key are consecutive.  Uses the exclusive offsets left in L.cnt.
__device__ __forceinline__ void patch_chunks(const Patch& P, const PatchBox& B, SortLDS& L, int K) {
    const int tid = threadIdx.x;
    int nch[4], off[4], cnt[4], choff[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        int c = tid * 4 + t;
        off[t] = c < B.vol ? L.cnt[c] : 0;
        int nxt = (c + 1 < B.vol) ? L.cnt[c + 1] : P.count;
        cnt[t] = c < B.vol ? nxt - off[t] : 0;
        nch[t] = (cnt[t] + K - 1) / K;
    }
    int total = block_exscan4(nch, choff, L);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        int c = tid * 4 + t;
        for (int j = 0; j < nch[t]; ++j) {
            int ci = choff[t] + j;
            L.ch_start[ci] = (unsigned short)(off[t] + j * K);
            L.ch_len[ci] = (unsigned short)min(K, cnt[t] - j * K);
            L.ch_key[ci] = (unsigned short)c;
        }
    }
    if (tid == 0) L.n_chunks = total;
    __syncthreads();
}

// segment description over chunk keys (same as segment_info, keyed by chunk)
__device__ __forceinline__ SegInfo segment_info_key(int key, bool act) {
    const int lane = threadIdx.x & 63;
    const int k = act ? key : -1 - lane;
    SegInfo g;
    g.same = 0;
    unsigned any = 0;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        int d = 1 << s;
        int kd = __shfl_up(k, d, 64);
        bool same = act && lane >= d && kd == k;
        if (same) g.same |= 1u << s;
        if (__ballot(same)) any |= 1u << s;
    }
    int nxt = __shfl_down(k, 1, 64);
    g.tail = act && (lane == 63 || nxt != k);
    g.steps = 0;
    while (g.steps < 6 && ((any >> g.steps) & 1u)) ++g.steps;
    return g;
}

__device__ __forceinline__ SegInfo segment_info(const SortLDS& L, int sp, int count) {
    const int lane = threadIdx.x & 63;
    const bool act = sp < count;
    const int key = act ? L.skey[sp] : -1 - lane;
    SegInfo g;
    g.same = 0;
    unsigned any = 0;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        int d = 1 << s;
        int kd = __shfl_up(key, d, 64);
        bool same = act && lane >= d && kd == key;
        if (same) g.same |= 1u << s;
        if (__ballot(same)) any |= 1u << s;
    }
    int nxt = __shfl_down(key, 1, 64);
    g.tail = act && (lane == 63 || nxt != key);
    g.steps = 0;
    while (g.steps < 6 && ((any >> g.steps) & 1u)) ++g.steps;
    return g;
}

// inclusive segmented scan; after it the tail lane of each run holds the run sum
// Segmented inclusive sum of N values at once: every shuffle of one step is
// issued before any result is used, so the LDS-crossbar latencies overlap
// instead of serialising per value.  The step count is wave-uniform.
template <int N>
__device__ __forceinline__ void seg_sum_n(float (&v)[N], const SegInfo& g) {
    const int steps = __builtin_amdgcn_readfirstlane(g.steps);
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        if (s < steps) {
            float t[N];
#pragma unroll
            for (int o = 0; o < N; ++o) t[o] = __shfl_up(v[o], 1 << s, 64);
            const bool take = (g.same >> s) & 1u;
#pragma unroll
            for (int o = 0; o < N; ++o) v[o] = take ? v[o] + t[o] : v[o];
        }
    }
}

// Compiler-only ordering point between read-modify-writes of a wave-private
// LDS tile.  Lanes of a wave update distinct cells for one stencil offset, but
// a cell written at one offset may be read by another lane at the next; the
// LDS executes a wave's instructions in issue order, so keeping the compiler
// from hoisting the next read above this write is all that is needed.
__device__ __forceinline__ void tile_order() { asm volatile("" ::: "memory"); }

__device__ __forceinline__ float seg_sum(float v, const SegInfo& g) {
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        if (s < g.steps) {
            float t = __shfl_up(v, 1 << s, 64);
            if ((g.same >> s) & 1u) v += t;
        }
    }
    return v;
}
