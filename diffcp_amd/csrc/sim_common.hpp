// Shared device layout and small-matrix math for the codim-MPM cloth kernels.
//
// Layout in HBM (all instances of a batch concatenated):
//   points   = vertex, grasper and cube particles, SoA, one index space [0, NP)
//   elements = one particle per triangle, SoA, [0, NE)
//   a state slot = x[3][NP] v[3][NP] C[9][NP] | Ce[9][NE] n[3][NE]   (floats)
//     n = third column of d at the previous substep (the rest-shape normal
//     at a fresh state); F and d are rebuilt from the vertex positions.
//   grid     = cells in 4x4x4 blocks (64 cells = 1 KB of float4 contiguous):
//     either one dense G^3 region per instance (only active blocks touched),
//     or one shared pool of blocks with a per-instance block table (a block
//     gets a pool slot the first substep a particle stencil reaches it and
//     keeps it until the table is reset at the start of a rollout).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DCP_TILE_CAP 2048          // LDS tile cells of a P2G patch (4 floats each)
#define DCP_PATCH_MAX 1024         // particles per patch
#define DCP_WG 256                 // threads per particle workgroup
#define DCP_PATCH_GRADS 5          // per-patch parameter gradients: E, nu, shear, contact, friction

enum : int { CLS_VERTEX = 0, CLS_GRASPER = 1, CLS_CUBE = 2, CLS_ELEMENT = 3 };

// template forms of the scatter kernels (bit flags)
#define DCP_GF_POOL 1              // grid from the block pool
#define DCP_GF_DET 2               // fixed-point scatter sums

struct InstDev {
    int G, nb;                     // grid side, blocks per side (ceil(G/4))
    long long grid_base;           // first cell of this instance's dense region (0 with a block pool)
    int* btab;                     // block pool: this instance's block -> pool slot (-1: none); null: dense
    int pool_trash;                // block pool: slot that stands in for a block without one
    int flag_base;                 // first block flag of this instance
    int steps_per_frame, frames_per_window, group_size, n_actions, n_grasp, traj_off;
    int v_start, v_count, o_start, o_count, c_start, c_count, e_start, e_count;
    float dt, dx, inv_dx, gravity;
    float gdt;                     // dt * scene_scaling * gravity   (simulation.py:182)
    float half_dt2;                // 0.5 * dt * dt                   (simulation.py:399)
    float m_vertex, m_quad, vol_quad, m_cube, m_op, interval;
    float E, nu, ks, kc, damping, friction;
};

struct Patch {                     // a chunk of one particle class of one instance
    int inst, cls, start, count;
};

// A scatter patch: up to two chunks of one instance that share a spatial bin,
// lanes [0, count_a) taking chunk a and the next count_b lanes chunk b (the
// vertices and the elements of one bin together, so that one LDS tile sums
// both classes' contributions before the flush).  count_b = 0: one chunk.
struct SPatch {
    int inst, cls_a, start_a, count_a, cls_b, start_b, count_b;
    __device__ int count() const { return count_a + count_b; }
    // the chunk and in-chunk index of lane q (q < count())
    __device__ void lane(int q, Patch& P, int& ql) const {
        const bool a = q < count_a;
        P.inst = inst;
        P.cls = a ? cls_a : cls_b;
        P.start = a ? start_a : start_b;
        P.count = a ? count_a : count_b;
        ql = a ? q : q - count_a;
    }
    __device__ Patch part(int which) const {
        return which == 0 ? Patch{inst, cls_a, start_a, count_a} : Patch{inst, cls_b, start_b, count_b};
    }
};

// Compact copy of one substep's grid for the adjoint: the (m, m v) of every
// active 4^3 block, instances in order, block ids, blocks per instance, and
// ok = 1 when every instance fitted in `cap` blocks.  acc == nullptr: none.
struct GridSnap {
    float4* acc;
    int* blk;
    int* cnt;
    int* ok;
    int cap;
};

struct SimDev {
    int B, NP, NE, n_patches;
    int precision, svd_grad, per_elem, max_active;
    int return_mapping;            // 1: live element return mapping with cf = friction
    const InstDev* inst;
    const Patch* patches;          // point patches then element patches
    const SPatch* spatches;        // scatter patches (P2G and the G2P adjoint)
    int n_spatches;
    const Patch* gpatches;         // gather patches (G2P, P2G adjoint): full chunks, point patches first
    int n_gpatches, n_gpoint;
    int* big_list;                 // [2][n_spatches]: scatter patches whose tile exceeds the cap
    int* big_n;                    // [4]: list lengths, patches listed since the host last looked
    const int* tri;                // [3][NE] point indices
    const float* Dinv;             // [9][NE] inverse of the rest d
    const float* emod;             // [3][NE] per-element (E, shear, contact) or null
    const float* f_override;       // [9][NE] stress F of the first substep after a rest-shape re-base, or null
    const int* inc_off;            // [NP+1] vertex -> incident (element, corner)
    const int* inc_list;           // [3 NE]   e*3 + corner
    const float* traj;             // trajectory pool [frames*grasp][3]
    const int* elem_inst;          // [NE] instance of each element
    // forward temporaries (element prep outputs)
    float* xe; float* ve; float* ke; float* fc;          // [3][NE] [3][NE] [3][NE] [NE][3][3]
    float* nd;                                            // [3][NE] n of the current d
    // grid
    int det;                       // 1: scatter sums in fixed point (bitwise reproducible), 0: float atomics
    unsigned long long* accq;      // det: (m, mv) scatter sums, two-word fixed point, per block [8 words][64 cells]
    float4* acc;                   // (m, mv): float-atomic scatter sums, or the grid update's (kept for the adjoint)
    float4* vout;                  // grid_v_out
    unsigned long long* gaccq;     // det: d/d grid_v_out scatter sums, as accq (words 0..5)
    float4* gacc;                  // not det: d/d grid_v_out float-atomic scatter sums (y, z, w)
    float4* gmv;                   // (d/dm, d/dmv)
    int* poison;                   // [2][B]: a non-finite or out-of-range term reached the (adjoint) scatter sums
    int* flags;                    // active block flags (per instance at flag_base)
    int* active;                   // active block ids, per instance at flag_base
    int* n_active;                 // [B] active blocks per instance
    unsigned* status;
    int* pool_n;                   // block pool: slots handed out (null: dense grids)
    int pool_cap;                  // block pool: slots available (the trash slot follows them)
    // adjoint temporaries
    float* g_xe;                   // [3][NE] dL/dx_e of the G2P reverse
    float* g_fv;                   // [3][NP] d/d (vertex force impulse)
    float* ec_gx; float* ec_gv;    // [NE][3][3] per-corner d/dx; [NE][3] d/dv (one for the three corners)
    double* pgrad_patch;           // [element gather patches][waves][DCP_PATCH_GRADS]  E, nu, shear, contact, friction
    double* pgrad_inst;            // [B][DCP_N_PARAMS]
    unsigned long long* damp_fx;   // [B][2] dL/d damping, two-word fixed point
    double* emod_grad;             // [3][NE]
    double* traj_grad;             // [frames*grasp][3] dL/d trajectory (float-atomic form)
    unsigned long long* traj_fx;   // [frames*grasp][3][2] the same, two-word fixed point (det)
};

// ---------------------------------------------------------------------------
// small fixed-size math
// ---------------------------------------------------------------------------
template <typename T> struct V3 {
    T x, y, z;
    __host__ __device__ V3() {}
    __host__ __device__ V3(T a, T b, T c) : x(a), y(b), z(c) {}
    __host__ __device__ T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    __host__ __device__ T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
template <typename T> __device__ __forceinline__ V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> __device__ __forceinline__ V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> __device__ __forceinline__ V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> __device__ __forceinline__ T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

template <typename T> struct M3 {
    T m[9];                                    // row-major
    __device__ T& operator()(int i, int j) { return m[3 * i + j]; }
    __device__ T operator()(int i, int j) const { return m[3 * i + j]; }
};
template <typename T> __device__ __forceinline__ M3<T> mzero() { M3<T> r; for (int i = 0; i < 9; ++i) r.m[i] = T(0); return r; }
template <typename T> __device__ __forceinline__ M3<T> meye() { M3<T> r = mzero<T>(); r.m[0] = r.m[4] = r.m[8] = T(1); return r; }
template <typename T> __device__ __forceinline__ M3<T> mmul(const M3<T>& a, const M3<T>& b) {
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r(i, j) = a(i, 0) * b(0, j) + a(i, 1) * b(1, j) + a(i, 2) * b(2, j);
    return r;
}
template <typename T> __device__ __forceinline__ M3<T> mT(const M3<T>& a) {
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r(i, j) = a(j, i);
    return r;
}
template <typename T> __device__ __forceinline__ M3<T> madd(const M3<T>& a, const M3<T>& b) { M3<T> r; for (int i = 0; i < 9; ++i) r.m[i] = a.m[i] + b.m[i]; return r; }
template <typename T> __device__ __forceinline__ M3<T> mscale(T s, const M3<T>& a) { M3<T> r; for (int i = 0; i < 9; ++i) r.m[i] = s * a.m[i]; return r; }
template <typename T> __device__ __forceinline__ V3<T> mv(const M3<T>& a, V3<T> v) {
    return {a(0, 0) * v.x + a(0, 1) * v.y + a(0, 2) * v.z, a(1, 0) * v.x + a(1, 1) * v.y + a(1, 2) * v.z,
            a(2, 0) * v.x + a(2, 1) * v.y + a(2, 2) * v.z};
}
template <typename T> __device__ __forceinline__ V3<T> mtv(const M3<T>& a, V3<T> v) {  // a^T v
    return {a(0, 0) * v.x + a(1, 0) * v.y + a(2, 0) * v.z, a(0, 1) * v.x + a(1, 1) * v.y + a(2, 1) * v.z,
            a(0, 2) * v.x + a(1, 2) * v.y + a(2, 2) * v.z};
}
template <typename T> __device__ __forceinline__ M3<T> outer(V3<T> a, V3<T> b) {
    M3<T> r;
    r(0, 0) = a.x * b.x; r(0, 1) = a.x * b.y; r(0, 2) = a.x * b.z;
    r(1, 0) = a.y * b.x; r(1, 1) = a.y * b.y; r(1, 2) = a.y * b.z;
    r(2, 0) = a.z * b.x; r(2, 1) = a.z * b.y; r(2, 2) = a.z * b.z;
    return r;
}
// General 3x3 inverse by cofactors (the same formula as Taichi Matrix.inverse).
template <typename T> __device__ __forceinline__ M3<T> minv(const M3<T>& a) {
    T c00 = a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1);
    T c01 = a(1, 2) * a(2, 0) - a(1, 0) * a(2, 2);
    T c02 = a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0);
    T det = a(0, 0) * c00 + a(0, 1) * c01 + a(0, 2) * c02;
    T id = T(1) / det;
    M3<T> r;
    r(0, 0) = c00 * id;
    r(0, 1) = (a(0, 2) * a(2, 1) - a(0, 1) * a(2, 2)) * id;
    r(0, 2) = (a(0, 1) * a(1, 2) - a(0, 2) * a(1, 1)) * id;
    r(1, 0) = c01 * id;
    r(1, 1) = (a(0, 0) * a(2, 2) - a(0, 2) * a(2, 0)) * id;
    r(1, 2) = (a(0, 2) * a(1, 0) - a(0, 0) * a(1, 2)) * id;
    r(2, 0) = c02 * id;
    r(2, 1) = (a(0, 1) * a(2, 0) - a(0, 0) * a(2, 1)) * id;
    r(2, 2) = (a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0)) * id;
    return r;
}

// ---------------------------------------------------------------------------
// state and static accessors
// ---------------------------------------------------------------------------
struct StateView {
    float* p; int NP, NE;
    __device__ float* px(int c) const { return p + (size_t)c * NP; }               // c in 0..2
    __device__ float* pv(int c) const { return p + (size_t)(3 + c) * NP; }
    __device__ float* pC(int c) const { return p + (size_t)(6 + c) * NP; }         // c in 0..8
    __device__ float* eC(int c) const { return p + (size_t)15 * NP + (size_t)c * NE; }
    __device__ float* en(int c) const { return p + (size_t)15 * NP + (size_t)(9 + c) * NE; }
    __device__ V3<float> x(int i) const { return {px(0)[i], px(1)[i], px(2)[i]}; }
    __device__ V3<float> v(int i) const { return {pv(0)[i], pv(1)[i], pv(2)[i]}; }
    __device__ M3<float> C(int i) const { M3<float> r; for (int c = 0; c < 9; ++c) r.m[c] = pC(c)[i]; return r; }
    __device__ M3<float> Ce(int e) const { M3<float> r; for (int c = 0; c < 9; ++c) r.m[c] = eC(c)[e]; return r; }
    __device__ V3<float> n(int e) const { return {en(0)[e], en(1)[e], en(2)[e]}; }
};
__device__ __forceinline__ M3<float> load_m9(const float* base, int stride, int i) {
    M3<float> r; for (int c = 0; c < 9; ++c) r.m[c] = base[(size_t)c * stride + i]; return r;
}
__device__ __forceinline__ void store_m9(float* base, int stride, int i, const M3<float>& a) {
    for (int c = 0; c < 9; ++c) base[(size_t)c * stride + i] = a.m[c];
}
__device__ __forceinline__ V3<float> load_v3(const float* base, int stride, int i) {
    return {base[i], base[(size_t)stride + i], base[(size_t)2 * stride + i]};
}
__device__ __forceinline__ void store_v3(float* base, int stride, int i, V3<float> a) {
    base[i] = a.x; base[(size_t)stride + i] = a.y; base[(size_t)2 * stride + i] = a.z;
}

// ---------------------------------------------------------------------------
// B-spline stencil (simulation.py:140-142, 196-197)
// ---------------------------------------------------------------------------
struct Stencil {
    int base[3];
    float fx[3];
    float w[3][3];    // [node][dim]
    float dw[3][3];   // d w / d fx
};

// Integer base cell trunc(x*inv_dx - 0.5) (ti.cast to i32, simulation.py:140)
// from a rounded FP32 product and a rounded FP32 difference, as the closed
// form states it: the product is kept out of a fused multiply-add, which
// could round across an integer when inv_dx is not a power of two.
__device__ __forceinline__ float scaled_position(float x, float inv_dx) {
#pragma clang fp contract(off)
    return x * inv_dx;
}
__device__ __forceinline__ int base_key(float x, float inv_dx) {
#pragma clang fp contract(off)
    return (int)(scaled_position(x, inv_dx) - 0.5f);
}

// base = trunc(x*inv_dx - 0.5) (ti.cast to i32); clamped into [0, G-3] with a status bit.
__device__ __forceinline__ void make_stencil(V3<float> x, float inv_dx, int G, Stencil& s, unsigned* status) {
    bool out = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float xs = scaled_position(x[d], inv_dx);
        int b = base_key(x[d], inv_dx);
        if (b < 0 || b > G - 3 || !(xs == xs)) { out = true; b = b < 0 || !(xs == xs) ? 0 : G - 3; }
        s.base[d] = b;
        float f = xs - (float)b;
        s.fx[d] = f;
        s.w[0][d] = 0.5f * (1.5f - f) * (1.5f - f);
        s.w[1][d] = 0.75f - (f - 1.0f) * (f - 1.0f);
        s.w[2][d] = 0.5f * (f - 0.5f) * (f - 0.5f);
        s.dw[0][d] = f - 1.5f;
        s.dw[1][d] = 2.0f * (1.0f - f);
        s.dw[2][d] = f - 0.5f;
    }
    if (out && status) atomicOr(status, 2u);
}

// Two-word fixed point for sums whose value must not depend on the order of
// their terms (atomic scatters, reductions over waves): x = hi 2^-24 +
// lo 2^-62 with hi, lo 64-bit integers, which add exactly in any order.  A
// term must satisfy |x| < 2^38; below 2^-62 it rounds to a multiple of 2^-62.
__device__ __forceinline__ bool fx_split(double d, long long& hi, long long& lo) {
    if (!(fabs(d) < 0x1p38)) { hi = lo = 0; return false; }
    const double h = rint(d * 0x1p24);
    hi = (long long)h;
    lo = (long long)rint((d - h * 0x1p-24) * 0x1p62);
    return true;
}
__device__ __forceinline__ double fx_join(unsigned long long hi, unsigned long long lo) {
    return (double)(long long)hi * 0x1p-24 + (double)(long long)lo * 0x1p-62;
}
// atomically add x to the fixed-point pair (w0: hi, w1: lo); false (nothing
// added) when x is not finite or out of range
__device__ __forceinline__ bool fx_atomic_add(unsigned long long* w0, unsigned long long* w1, double x) {
    long long hi, lo;
    if (!fx_split(x, hi, lo)) return false;
    if (hi) atomicAdd(w0, (unsigned long long)hi);
    if (lo) atomicAdd(w1, (unsigned long long)lo);
    return true;
}
// words of cell `cell` in a per-block [8 words][64 cells] fixed-point array
__device__ __forceinline__ unsigned long long* fx_cell_word(unsigned long long* q, long long cell, int word) {
    return q + ((cell >> 6) << 9) + (word << 6) + (cell & 63);
}

// First cell of 4^3 block blk of an instance grid: its place in the dense
// region, or the pool slot its table entry names.  A block without a slot
// (never reached by a stencil: a tile corner staged but never read) reads
// and writes the trash slot, so every access stays inside the pool.  Per-node
// kernels take the form as a template parameter (POOL), so that the dense
// kernels carry no table code; per-block kernels test I.btab.
__device__ __forceinline__ long long block_cell0(const InstDev& I, int blk) {
    if (I.btab) {
        const int t = I.btab[blk];
        return (long long)(t >= 0 ? t : I.pool_trash) * 64;
    }
    return I.grid_base + (long long)blk * 64;
}
// blocked cell index inside an instance grid
template <bool POOL>
__device__ __forceinline__ long long cell_index(const InstDev& I, int i, int j, int k) {
    const int blk = ((i >> 2) * I.nb + (j >> 2)) * I.nb + (k >> 2);
    long long c0;
    if (POOL) {
        const int t = I.btab[blk];
        c0 = (long long)(t >= 0 ? t : I.pool_trash) * 64;
    } else {
        c0 = I.grid_base + (long long)blk * 64;
    }
    return c0 + ((i & 3) << 4) + ((j & 3) << 2) + (k & 3);
}
// Dense grids: the blocked index of cell (i, j, k) past the instance's first
// cell is a sum of per-axis terms, so the 27 nodes of a stencil with base b
// are at byte offsets x[i] + y[j] + z[k] of a float4 cell array from the
// instance's first cell: a uniform base pointer plus one 32-bit offset per
// node, which the loads take as they are (a grid of at most DCP_MAX_GRID^3
// cells spans less than 2^31 bytes of float4).
#define DCP_MAX_GRID 512
struct StencilOffsets {
    unsigned x[3], y[3], z[3];
};
__device__ __forceinline__ void stencil_offsets(const InstDev& I, const int base[3], StencilOffsets& o) {
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const unsigned ci = base[0] + t, cj = base[1] + t, ck = base[2] + t, nb = (unsigned)I.nb;
        o.x[t] = ((((ci >> 2) * nb * nb) << 6) + ((ci & 3) << 4)) << 4;
        o.y[t] = ((((cj >> 2) * nb) << 6) + ((cj & 3) << 2)) << 4;
        o.z[t] = (((ck >> 2) << 6) + (ck & 3)) << 4;
    }
}
__device__ __forceinline__ float4 cell_at(const float4* base, unsigned byte_off) {
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(base) + byte_off);
}
