// Host engine behind the C-ABI (include/diffcp.h): device memory, internal
// particle order, patches, launches, checkpointed rollouts and hipGraphs.
#include "kernels.hpp"
#include "patch_sort.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

static thread_local std::string g_err;
static int set_err(const std::string& m) { g_err = m; return -1; }
// for the other sources of the library (noalloc.hip)
int dcp_record_error(const std::string& m) { return set_err(m); }
#define HIP_OK(x)                                                                             \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) return set_err(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define LAUNCH_OK()                                                                                \
    do {                                                                                           \
        hipError_t e_ = hipGetLastError();                                                         \
        if (e_ != hipSuccess) return set_err(std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

namespace {

struct HostInst {
    int G, nv, ne, nc, no, ngp, nact;
    std::vector<int> vperm;   // internal vertex -> natural vertex
    std::vector<int> vinv;    // natural -> internal
    std::vector<int> eperm;   // internal element -> natural element
    std::vector<int> einv;
    std::vector<float> x0v, x0o, x0c;   // creation positions (natural order)
    std::vector<int> tri;     // natural triangles
    std::vector<uint64_t> vbin, ebin;   // spatial bin of each internal vertex / element
    std::vector<int> cperm, cinv;       // cube particles: internal -> natural, natural -> internal
    std::vector<uint64_t> cbin;         // spatial bin of each internal cube particle
};

// Base cell of a position (trunc(x G - 0.5) per axis, as the scatters key it),
// row-major in 30 bits: i, then j, then k.
uint32_t base_cell_key(float x, float y, float z, int G) {
    auto q = [&](float a) { int t = (int)(a * (float)G - 0.5f); return (uint32_t)std::min(1023, std::max(0, t)); };
    return (q(x) << 20) | (q(y) << 10) | q(z);
}

// Spatial bin (cubes of `cells` grid cells) in the high bits, the row-major
// base cell in the low bits: sorting by this key groups particles by bin
// (patches never straddle a bin, so a patch's tile stays compact) and orders
// a bin's particles as the scatters' per-patch binning orders the tile cells,
// so a particle's sorted position in its patch stays close to its lane and the
// scatters' LDS stores at sorted positions spread over the banks.
uint64_t bin_of(float x, float y, float z, int G, int cells) {
    auto b = [&](float a) { int t = (int)std::floor(a * G / cells); return (uint64_t)std::min(127, std::max(0, t)); };
    return (b(x) * 128 + b(y)) * 128 + b(z);
}
uint64_t spatial_key(float x, float y, float z, int G, int cells) {
    return (bin_of(x, y, z, G, cells) << 32) | base_cell_key(x, y, z, G);
}

// Cells of the scatter tile over a point set: the box of base cells
// trunc(x G - 0.5) plus the 2-cell stencil reach per axis.
long long tile_cells(const std::vector<float>& x, int G) {
    if (x.size() < 3) return 0;
    int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    for (size_t i = 0; i + 2 < x.size(); i += 3)
        for (int d = 0; d < 3; ++d) {
            const int b = (int)(x[i + d] * (float)G - 0.5f);
            lo[d] = std::min(lo[d], b);
            hi[d] = std::max(hi[d], b);
        }
    return (long long)(hi[0] - lo[0] + 3) * (hi[1] - lo[1] + 3) * (hi[2] - lo[2] + 3);
}

// Bin side (grid cells) for which one bin of a cloth sheet holds about one
// patch: particles per occupied cell rho, a bin of s x s occupied cells holds
// s^2 rho.  Small bins keep a patch's box small when the sheet bends.
int bin_cells(const std::vector<float>& xv, const std::vector<float>& xe, int G, int patch) {
    std::vector<uint64_t> cells;
    auto add = [&](const std::vector<float>& x) {
        for (size_t i = 0; i + 2 < x.size(); i += 3) {
            uint64_t a = (uint64_t)std::min(G - 1, std::max(0, (int)(x[i] * G)));
            uint64_t b = (uint64_t)std::min(G - 1, std::max(0, (int)(x[i + 1] * G)));
            uint64_t c = (uint64_t)std::min(G - 1, std::max(0, (int)(x[i + 2] * G)));
            cells.push_back((a * G + b) * G + c);
        }
    };
    add(xv);
    add(xe);
    if (cells.empty()) return 6;
    std::sort(cells.begin(), cells.end());
    const double occupied = (double)(std::unique(cells.begin(), cells.end()) - cells.begin());
    const double rho = (xv.size() + xe.size()) / 3.0 / occupied;
    return std::max(2, std::min(8, (int)std::floor(std::sqrt(patch / rho))));
}

// host FP32 cofactor inverse (same formula as the device minv)
void inv3h(const float* a, float* r) {
    float c00 = a[4] * a[8] - a[5] * a[7];
    float c01 = a[5] * a[6] - a[3] * a[8];
    float c02 = a[3] * a[7] - a[4] * a[6];
    float det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    float id = 1.0f / det;
    r[0] = c00 * id;
    r[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    r[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    r[3] = c01 * id;
    r[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    r[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    r[6] = c02 * id;
    r[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    r[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

template <typename T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
};

}  // namespace

#define DCP_PREP_FLOATS 21   // x_e, v_e, k, n (3 each) and three corner forces (9) per element

struct dcp_ctx {
    int B = 0, NP = 0, NE = 0, n_patches = 0, n_spatches = 0, n_gpatches = 0, n_gpoint = 0, n_cu = 256;
    dcp_options opt{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<HostInst> hi;
    std::vector<InstDev> inst_h;
    std::vector<Patch> patches_h;
    std::vector<SPatch> spatches_h;
    std::vector<Patch> gpatches_h;
    std::vector<int> tri_h;            // internal [3][NE]
    std::vector<int> elem_inst_h;
    std::vector<int> vnat_h;           // per point: natural vertex index (vertices) or index within class
    std::vector<int> toff_h, tcount_h;
    std::vector<int4> tchunk_h;
    size_t slot_floats = 0;
    int max_states = 0, n_adj = 4;
    long long grid_cells = 0;
    int n_flags = 0;
    SimDev sim{};
    // owned device buffers and their sizes
    std::vector<std::pair<void*, size_t>> allocs;
    float* states = nullptr;
    float* adj = nullptr;
    float* scratch = nullptr;
    float* targets = nullptr;
    int* toff = nullptr; int* tcount = nullptr; int4* tchunks = nullptr;
    int* vnat = nullptr;
    float* tmin = nullptr; int* targ = nullptr;
    int loss_parts = 1;                       // vertex parts of the loss minimum (k_loss_min)
    size_t part_cap = 0;                      // entries of the per-part buffers
    float* pmin = nullptr; int* pnat = nullptr; int* parg = nullptr;
    double* loss_d = nullptr; double* seed_d = nullptr;
    std::vector<double> ones_h;        // host source of the unit loss seeds
    std::vector<float> traj_h;         // staging of dcp_set_trajectory
    // per-substep grid snapshots of one checkpoint segment (slot = substep - segment base)
    float4* snap_acc = nullptr;
    int* snap_blk = nullptr;
    int* snap_cnt = nullptr;
    int* snap_ok = nullptr;
    int snap_cap = 0, snap_slots = 0;
    size_t snap_bytes = 0;         // device bytes of the snapshots
    size_t alloc_bytes = 0;        // device bytes of everything else the context allocated
    long long init_blocks = 0;     // 4^3 blocks the initial particle stencils touch
    // element prep outputs (x_e, v_e, k, n, corner forces: DCP_PREP_FLOATS per
    // element): the working block, and one block per snapshot slot when memory
    // allows, so that the adjoint reuses the forward's instead of recomputing
    float* prep_work = nullptr;
    float* snap_prep = nullptr;
    // rest-shape re-base (dcp_rebase_elements): the old-rest-shape F of every
    // element, used once by the next forward substep's stress
    float* f_old = nullptr;
    bool rebase_pending = false;
    // natural particle -> slot index for caller state arrays in device memory
    int* io_perm = nullptr;
    void use_prep(int slot) {
        float* b = slot >= 0 && snap_prep ? snap_prep + (size_t)slot * DCP_PREP_FLOATS * NE : prep_work;
        sim.xe = b; sim.ve = b + (size_t)3 * NE; sim.ke = b + (size_t)6 * NE; sim.nd = b + (size_t)9 * NE;
        sim.fc = b + (size_t)12 * NE;
    }
    void free_snapshots() {
        for (void* q : {(void*)snap_acc, (void*)snap_blk, (void*)snap_cnt, (void*)snap_ok, (void*)snap_prep})
            if (q) (void)hipFree(q);
        snap_acc = nullptr; snap_blk = snap_cnt = snap_ok = nullptr;
        snap_prep = nullptr;
        snap_cap = snap_slots = 0;
        snap_bytes = 0;
    }
    GridSnap snap(int slot) const {
        if (slot < 0 || slot >= snap_slots) return GridSnap{};
        return GridSnap{snap_acc + (size_t)slot * snap_cap * 64, snap_blk + (size_t)slot * snap_cap,
                        snap_cnt + (size_t)slot * B, snap_ok + slot, snap_cap};
    }
    int* keys_d = nullptr;
    int n_targets_total = 0;
    InstDev* inst_d = nullptr;
    // graphs
    std::map<std::tuple<int, int, int, int, int>, hipGraphExec_t> graphs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // second stream (default priority) for an adjoint substep's grid restore
    // beside the previous substep's tail, and its fork / join events
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    double last_ms = 0.0;
    int last_final_slot = -1;
    // per-launch kernel timing
    int time_kid = 0;
    std::vector<hipEvent_t> tev;
    int tev_used = 0;
    bool capturing = false;
    // lean adjoint: every grid snapshot of the segment being walked is known to
    // hold its substep (read back after the segment's forward), so the adjoint
    // substeps restore their grids without launching P2G and the grid update
    bool lean_adjoint = false;
    // ... and no patch of the segment's forward overflowed its tile, so none of
    // its adjoint substeps does: k_g2p_adj_big is left out
    bool adj_big_left_out = false;
    // forward segments of dcp_rollout_grad leave k_p2g_big out until a patch
    // has overflowed its tile in this context (the window is then run again
    // with it, and it stays in)
    bool fwd_big_seen = false, fwd_big_left_out = false;
    void drop_timing_events() {
        for (hipEvent_t e : tev) (void)hipEventDestroy(e);
        tev.clear();
        tev_used = 0;
    }

    StateView state(int slot) const { return StateView{states + (size_t)slot * slot_floats, NP, NE}; }
    StateView adjv(int g) const { return StateView{adj + (size_t)g * slot_floats, NP, NE}; }
    StateView scratch_view() const { return StateView{scratch, NP, NE}; }
    StateView none() const { return StateView{nullptr, NP, NE}; }

    template <typename T> int alloc(T** p, size_t n) {
        void* q = nullptr;
        size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) return set_err(std::string("hipMalloc ") + std::to_string(bytes) + " B: " + hipGetErrorString(e));
        e = hipMemset(q, 0, bytes);
        if (e != hipSuccess) return set_err(std::string("hipMemset: ") + hipGetErrorString(e));
        allocs.emplace_back(q, bytes);
        alloc_bytes += bytes;
        *p = (T*)q;
        return 0;
    }
    // free one buffer from alloc() (null: nothing); the stream must be idle
    template <typename T> void release(T*& p) {
        if (!p) return;
        for (size_t i = 0; i < allocs.size(); ++i)
            if (allocs[i].first == (void*)p) {
                (void)hipFree(p);
                alloc_bytes -= allocs[i].second;
                allocs.erase(allocs.begin() + (long)i);
                break;
            }
        p = nullptr;
    }
    // cached rollout graphs hold kernel arguments (buffer pointers, sizes):
    // drop them before any of those change
    void drop_graphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
    template <typename T> int upload(T** p, const std::vector<T>& h) {
        if (alloc(p, h.size())) return -1;
        if (!h.empty()) HIP_OK(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return 0;
    }
    int grid_patches() const {
        return std::max(1, std::min(n_patches, opt.scatter_wgs > 0 ? opt.scatter_wgs : n_cu * 16));
    }
    // gather patches: all of them (G2P), the point ones (P2G adjoint) or the
    // element ones (element adjoint)
    int n_gelem() const { return n_gpatches - n_gpoint; }
    int grid_gpatches(bool points) const {
        return std::max(1, std::min(points ? n_gpoint : n_gpatches, opt.scatter_wgs > 0 ? opt.scatter_wgs : n_cu * 16));
    }
    int grid_gelem() const { return std::max(1, std::min(n_gelem(), opt.scatter_wgs > 0 ? opt.scatter_wgs : n_cu * 16)); }
    int grid_spatches() const {
        return std::max(1, std::min(n_spatches, opt.scatter_wgs > 0 ? opt.scatter_wgs : n_cu * 16));
    }
    int grid_big() const { return 2 * n_cu; }
    // block pool of the grids (grid_pool_blocks > 0)
    int* btab = nullptr;
    int pool_cap = 0;
    bool pooled() const { return btab != nullptr; }
    // template form of the scatter kernels: DCP_GF_POOL | DCP_GF_DET bits
    int grid_form() const { return (pooled() ? DCP_GF_POOL : 0) | (sim.det ? DCP_GF_DET : 0); }
    // hand the pool out afresh (a new rollout or reset state)
    int reset_pool() {
        if (!pooled()) return 0;
        hipLaunchKernelGGL(k_pool_reset, dim3(std::max(1, std::min((n_flags + 255) / 256, n_cu * 4))), dim3(256), 0,
                           stream, btab, n_flags, sim.pool_n);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    int grid_elems() const { return std::max(1, std::min((NE + DCP_WG - 1) / DCP_WG, n_cu * 16)); }
    // B x per workgroups walking the per-instance active block lists
    int grid_gridop() const {
        int total = opt.grid_wgs > 0 ? opt.grid_wgs : n_cu * 4;
        return B * std::max(1, total / std::max(1, B));
    }
};

static bool slot_ok(const dcp_ctx* c, int s) { return s >= 0 && s < c->max_states; }
static bool gslot_ok(const dcp_ctx* c, int s) { return s >= 0 && s < c->n_adj; }

extern "C" const char* dcp_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// launch sequences
// ---------------------------------------------------------------------------
// Optional per-launch HIP events around one kernel kind (dcp_set_kernel_timing).
// Inside a stream capture the events become event-record nodes of the graph,
// so every replay re-records them.
static hipError_t time_mark(dcp_ctx* c) {
    hipEvent_t e = c->tev[c->tev_used++];
    return c->capturing ? hipEventRecordWithFlags(e, c->stream, hipEventRecordExternal) : hipEventRecord(e, c->stream);
}
#define TIMED_LAUNCH(kid, ...)                                                          \
    do {                                                                                \
        const bool on_ = (kid) == c->time_kid && c->tev_used + 2 <= (int)c->tev.size(); \
        if (on_) HIP_OK(time_mark(c));                                                  \
        hipLaunchKernelGGL(__VA_ARGS__);                                                \
        LAUNCH_OK();                                                                    \
        if (on_) HIP_OK(time_mark(c));                                                  \
    } while (0)

// A launch with a skip flag is left out of the kernel timing: it exits at once
// whenever the grid snapshot holds the substep.
// Overflowing patch tiles (a patch spread beyond DCP_STILE_CAP cells): the
// scatters list such patches, and k_p2g_big / k_g2p_adj_big, launched after
// each, scatter them through the LDS tile over sub-boxes that fit it (with an
// empty list they exit at once).
template <int GF> static int launch_p2g_form(dcp_ctx* c, StateView S, int s, int g, const int* skip);
static int launch_p2g(dcp_ctx* c, StateView S, int s, int g, const int* skip) {
    if (c->pooled()) {
        // pool slots for the blocks this substep's stencils reach
        hipLaunchKernelGGL(k_pool_alloc, dim3(c->grid_patches()), dim3(DCP_WG), 0, c->stream, c->sim, S, skip);
        LAUNCH_OK();
    }
    switch (c->grid_form()) {
        case 0: return launch_p2g_form<0>(c, S, s, g, skip);
        case 1: return launch_p2g_form<1>(c, S, s, g, skip);
        case 2: return launch_p2g_form<2>(c, S, s, g, skip);
        default: return launch_p2g_form<3>(c, S, s, g, skip);
    }
}
template <int GF>
static int launch_p2g_form(dcp_ctx* c, StateView S, int s, int g, const int* skip) {
    if (skip) {
        hipLaunchKernelGGL(k_p2g<GF>, dim3(c->grid_spatches()), dim3(DCP_WG), 0, c->stream, c->sim, S, s, g, skip);
        LAUNCH_OK();
    } else {
        TIMED_LAUNCH(DCP_KERNEL_P2G, k_p2g<GF>, dim3(c->grid_spatches()), dim3(DCP_WG), 0, c->stream, c->sim, S, s, g,
                     skip);
    }
    if (!c->fwd_big_left_out) {
        hipLaunchKernelGGL(k_p2g_big<GF>, dim3(c->grid_big()), dim3(DCP_WG), 0, c->stream, c->sim, S, s, g, skip);
        LAUNCH_OK();
    }
    hipLaunchKernelGGL(k_active_list, dim3(c->grid_gridop()), dim3(DCP_WG), 0, c->stream, c->sim, skip);
    LAUNCH_OK();
    return 0;
}

template <typename PT, bool RM>
static int launch_prep_form(dcp_ctx* c, StateView S, StateView Sn, int fresh) {
    TIMED_LAUNCH(DCP_KERNEL_PREP, (k_prep<PT, RM>), dim3(c->grid_elems()), dim3(DCP_WG), 0, c->stream, c->sim, S, Sn,
                 fresh);
    return 0;
}
static int launch_prep(dcp_ctx* c, StateView S, StateView Sn, int fresh) {
    const bool rm = c->opt.return_mapping != 0;
    if (c->opt.precision == 1)
        return rm ? launch_prep_form<float, true>(c, S, Sn, fresh) : launch_prep_form<float, false>(c, S, Sn, fresh);
    return rm ? launch_prep_form<double, true>(c, S, Sn, fresh) : launch_prep_form<double, false>(c, S, Sn, fresh);
}

// Points the kernels' element prep arrays at snapshot slot `slot` for the
// lifetime of the guard (no slot, or no prep snapshots: the working arrays).
struct PrepSlot {
    dcp_ctx* c;
    PrepSlot(dcp_ctx* c_, int slot) : c(c_) { c->use_prep(slot); }
    ~PrepSlot() { c->use_prep(-1); }
};

// snap >= 0: save this substep's grid (and its element prep outputs) into
// snapshot slot `snap`
static int launch_forward(dcp_ctx* c, int s, int g, StateView S, StateView Sn, int snap = -1) {
    PrepSlot ps(c, snap >= 0 && snap < c->snap_slots ? snap : -1);
    if (c->rebase_pending) {
        // the first forward substep after a re-base takes its stress F from f_old
        c->sim.f_override = c->f_old;
        c->rebase_pending = false;
        const int r = launch_prep(c, S, Sn, s == 0);
        c->sim.f_override = nullptr;
        if (r) return -1;
    } else if (launch_prep(c, S, Sn, s == 0)) {
        return -1;
    }
    if (launch_p2g(c, S, s, g, nullptr)) return -1;
    TIMED_LAUNCH(DCP_KERNEL_GRIDOP, k_gridop, dim3(c->grid_gridop()), dim3(DCP_WG), 0, c->stream, c->sim, 0, c->snap(snap),
                 nullptr);
    if (c->pooled())
        TIMED_LAUNCH(DCP_KERNEL_G2P, k_g2p<true>, dim3(c->grid_gpatches(false)), dim3(DCP_WG), 0, c->stream, c->sim, S, Sn);
    else
        TIMED_LAUNCH(DCP_KERNEL_G2P, k_g2p<false>, dim3(c->grid_gpatches(false)), dim3(DCP_WG), 0, c->stream, c->sim, S, Sn);
    return 0;
}

// snap >= 0: the grid of substep s was saved in snapshot slot `snap` by the
// forward; it is restored instead of recomputed (P2G and the grid update exit
// at once unless the snapshot overflowed), and the element prep outputs the
// forward left in that slot are used as they are.
// pending: G1's vertex entries still lack the element corner gradients of the
// previous adjoint substep (k_g2p_adj adds them); finalize: gather this
// substep's corner gradients into Gout now (k_vertex_final) instead of leaving
// them to the next adjoint substep.
static bool prep_saved(const dcp_ctx* c, int snap) { return c->snap(snap).acc && c->snap_prep; }

// The adjoint substep's grid: restored from the snapshot (P2G and the grid
// update then exit at once), or recomputed with the element prep.
static int launch_adjoint_grid(dcp_ctx* c, int s, int g, StateView S, int snap) {
    const GridSnap sn = c->snap(snap);
    const bool saved = prep_saved(c, snap);
    PrepSlot ps(c, saved ? snap : -1);
    if (!saved && launch_prep(c, S, c->scratch_view(), s == 0)) return -1;
    if (sn.acc) {
        // before the P2G: an overflowed snapshot leaves the active-block counts
        // cleared for it
        hipLaunchKernelGGL(k_grid_restore, dim3(c->grid_gridop()), dim3(DCP_WG), 0, c->stream, c->sim, sn);
        LAUNCH_OK();
        if (c->lean_adjoint && saved) return 0;   // the snapshot is known to hold: nothing to recompute
    }
    if (launch_p2g(c, S, s, g, sn.ok)) return -1;
    if (sn.ok) {
        hipLaunchKernelGGL(k_gridop, dim3(c->grid_gridop()), dim3(DCP_WG), 0, c->stream, c->sim, 1, GridSnap{}, sn.ok);
        LAUNCH_OK();
    } else {
        TIMED_LAUNCH(DCP_KERNEL_GRIDOP, k_gridop, dim3(c->grid_gridop()), dim3(DCP_WG), 0, c->stream, c->sim, 1, GridSnap{},
                     nullptr);
    }
    return 0;
}

template <int GF>
static int launch_g2p_adj_form(dcp_ctx* c, StateView S, StateView G1, StateView Gout, int pending) {
    TIMED_LAUNCH(DCP_KERNEL_G2P_ADJ, k_g2p_adj<GF>, dim3(c->grid_spatches()), dim3(DCP_WG), 0, c->stream, c->sim, S,
                 G1, Gout, pending);
    if (!c->adj_big_left_out) {
        hipLaunchKernelGGL(k_g2p_adj_big<GF>, dim3(c->grid_big()), dim3(DCP_WG), 0, c->stream, c->sim, S, G1, Gout,
                           pending);
        LAUNCH_OK();
    }
    return 0;
}

// G2P adjoint and grid-update adjoint (they need the substep's grid).
static int launch_adjoint_head(dcp_ctx* c, StateView S, StateView G1, StateView Gout, int snap, int pending) {
    PrepSlot ps(c, prep_saved(c, snap) ? snap : -1);
    int r;
    switch (c->grid_form()) {
        case 0: r = launch_g2p_adj_form<0>(c, S, G1, Gout, pending); break;
        case 1: r = launch_g2p_adj_form<1>(c, S, G1, Gout, pending); break;
        case 2: r = launch_g2p_adj_form<2>(c, S, G1, Gout, pending); break;
        default: r = launch_g2p_adj_form<3>(c, S, G1, Gout, pending); break;
    }
    if (r) return -1;
    TIMED_LAUNCH(DCP_KERNEL_GRIDOP_ADJ, k_gridop_adj, dim3(c->grid_gridop()), dim3(DCP_WG), 0, c->stream, c->sim,
                 c->adj_big_left_out ? 1 : 0);
    return 0;
}

template <typename PT, bool RM>
static int launch_elem_adj_form(dcp_ctx* c, StateView S, StateView G1, StateView Gout, int fresh, int s, int g) {
    if (c->pooled())
        TIMED_LAUNCH(DCP_KERNEL_PREP_ADJ, (k_elem_adj<PT, RM, true>), dim3(c->grid_gelem()), dim3(DCP_WG), 0,
                     c->stream, c->sim, S, G1, Gout, fresh, s, g);
    else
        TIMED_LAUNCH(DCP_KERNEL_PREP_ADJ, (k_elem_adj<PT, RM, false>), dim3(c->grid_gelem()), dim3(DCP_WG), 0,
                     c->stream, c->sim, S, G1, Gout, fresh, s, g);
    return 0;
}
static int launch_elem_adj(dcp_ctx* c, StateView S, StateView G1, StateView Gout, int fresh, int s, int g) {
    const bool rm = c->opt.return_mapping != 0;
    if (c->opt.precision == 1)
        return rm ? launch_elem_adj_form<float, true>(c, S, G1, Gout, fresh, s, g)
                  : launch_elem_adj_form<float, false>(c, S, G1, Gout, fresh, s, g);
    return rm ? launch_elem_adj_form<double, true>(c, S, G1, Gout, fresh, s, g)
              : launch_elem_adj_form<double, false>(c, S, G1, Gout, fresh, s, g);
}

// P2G adjoint of the point classes, the fused element adjoint and, with
// finalize, the vertex gather of the element corner gradients.  These no
// longer touch the grid accumulators, so the next adjoint substep's grid can
// be rebuilt beside them.
static int launch_adjoint_tail(dcp_ctx* c, int s, int g, StateView S, StateView G1, StateView Gout, int snap,
                               bool finalize) {
    PrepSlot ps(c, prep_saved(c, snap) ? snap : -1);
    const int fresh = (s == 0);
    if (c->pooled())
        TIMED_LAUNCH(DCP_KERNEL_P2G_ADJ, k_p2g_adj<true>, dim3(c->grid_gpatches(true)), dim3(DCP_WG), 0, c->stream, c->sim, S,
                     Gout, s, g);
    else
        TIMED_LAUNCH(DCP_KERNEL_P2G_ADJ, k_p2g_adj<false>, dim3(c->grid_gpatches(true)), dim3(DCP_WG), 0, c->stream, c->sim,
                     S, Gout, s, g);
    if (launch_elem_adj(c, S, G1, Gout, fresh, s, g)) return -1;
    if (finalize)
        TIMED_LAUNCH(DCP_KERNEL_VERTEX_FINAL, k_vertex_final, dim3(c->grid_patches()), dim3(DCP_WG), 0, c->stream, c->sim,
                     Gout);
    return 0;
}

static int launch_adjoint(dcp_ctx* c, int s, int g, StateView S, StateView G1, StateView Gout, int snap = -1,
                          int pending = 0, bool finalize = true) {
    if (launch_adjoint_grid(c, s, g, S, snap)) return -1;
    if (launch_adjoint_head(c, S, G1, Gout, snap, pending)) return -1;
    return launch_adjoint_tail(c, s, g, S, G1, Gout, snap, finalize);
}

static int launch_fill(dcp_ctx* c, float* p, size_t n) {
    int blocks = (int)std::min<size_t>((n + 255) / 256, (size_t)c->n_cu * 16);
    hipLaunchKernelGGL(k_fill_f, dim3(std::max(1, blocks)), dim3(256), 0, c->stream, p, n, 0.0f);
    LAUNCH_OK();
    return 0;
}

static int launch_loss(dcp_ctx* c, StateView S) {
    if (c->n_targets_total == 0) return set_err("no targets set (dcp_set_targets)");
    int nch = (int)c->tchunk_h.size();
    const int P = c->loss_parts, T = c->n_targets_total;
    hipLaunchKernelGGL(k_loss_min, dim3(std::max(1, std::min(nch * P, c->n_cu * 16))), dim3(DCP_WG), 0, c->stream,
                       c->sim, S, c->targets, c->tchunks, nch, P, T, c->vnat, c->pmin, c->pnat, c->parg);
    LAUNCH_OK();
    hipLaunchKernelGGL(k_loss_sum, dim3(std::min(c->B, 1024)), dim3(DCP_WG), 0, c->stream, c->pmin, c->pnat, c->parg,
                       P, T, c->tmin, c->targ, c->toff, c->tcount, c->loss_d, c->B);
    LAUNCH_OK();
    return 0;
}

static int launch_loss_adj(dcp_ctx* c, StateView S, StateView G) {
    hipLaunchKernelGGL(k_loss_adj, dim3(c->grid_patches()), dim3(DCP_WG), 0, c->stream, c->sim, S, G, c->targets,
                       c->toff, c->tcount, c->targ, c->seed_d);
    LAUNCH_OK();
    return 0;
}

// ---------------------------------------------------------------------------
// creation
// ---------------------------------------------------------------------------
static int build(dcp_ctx* c, const dcp_instance_desc* d, int B) {
    const int patch_size = c->opt.patch_size > 0 ? std::min(c->opt.patch_size, 256) : 256;
    c->B = B;
    c->hi.resize(B);
    c->inst_h.resize(B);
    int np = 0, ne = 0;
    for (int b = 0; b < B; ++b) {
        const dcp_instance_desc& D = d[b];
        if (D.n_grid < 4 || D.n_vertices <= 0 || D.n_elements < 0 || D.n_cube < 0 || D.n_grasper < 0)
            return set_err("instance " + std::to_string(b) + ": invalid sizes");
        if (D.n_grid > DCP_MAX_GRID)
            return set_err("instance " + std::to_string(b) + ": n_grid above " + std::to_string(DCP_MAX_GRID));
        if (D.n_grasper > 0 && (D.n_grasp_points <= 0 || D.n_grasper % D.n_grasp_points != 0))
            return set_err("instance " + std::to_string(b) + ": n_grasper must be a multiple of n_grasp_points");
        if (D.steps_per_frame <= 0) return set_err("steps_per_frame must be > 0");
        if (!D.vertex_x || (D.n_elements && !D.triangles)) return set_err("missing vertex/triangle arrays");
        HostInst& H = c->hi[b];
        H.G = D.n_grid; H.nv = D.n_vertices; H.ne = D.n_elements; H.nc = D.n_cube; H.no = D.n_grasper;
        H.ngp = D.n_grasp_points; H.nact = D.n_actions;
        H.x0v.assign(3 * H.nv, 0.f); H.x0o.assign(3 * H.no, 0.f); H.x0c.assign(3 * H.nc, 0.f);
        for (int i = 0; i < 3 * H.nv; ++i) H.x0v[i] = (float)D.vertex_x[i];
        for (int i = 0; i < 3 * H.no; ++i) H.x0o[i] = (float)D.grasper_x[i];
        for (int i = 0; i < 3 * H.nc; ++i) H.x0c[i] = (float)D.cube_x[i];
        H.tri.assign(D.triangles, D.triangles + 3 * H.ne);
        for (int t : H.tri)
            if (t < 0 || t >= H.nv) return set_err("triangle index out of range");
        // spatial order (bin, then row-major base cell) of vertices and element centroids
        std::vector<float> xc(3 * (size_t)H.ne);
        for (int e = 0; e < H.ne; ++e)
            for (int d = 0; d < 3; ++d)
                xc[3 * e + d] = (H.x0v[3 * H.tri[3 * e] + d] + H.x0v[3 * H.tri[3 * e + 1] + d] +
                                 H.x0v[3 * H.tri[3 * e + 2] + d]) / 3;
        const int bcells = c->opt.bin_cells > 0 ? std::max(2, std::min(16, (int)c->opt.bin_cells))
                           : c->opt.bin_cells < 0 ? bin_cells(H.x0v, xc, H.G, patch_size) : 6;
        std::vector<uint64_t> code(H.nv);
        for (int i = 0; i < H.nv; ++i) code[i] = spatial_key(H.x0v[3 * i], H.x0v[3 * i + 1], H.x0v[3 * i + 2], H.G, bcells);
        H.vperm.resize(H.nv);
        std::iota(H.vperm.begin(), H.vperm.end(), 0);
        std::stable_sort(H.vperm.begin(), H.vperm.end(), [&](int a, int bb) { return code[a] < code[bb]; });
        H.vinv.resize(H.nv);
        for (int i = 0; i < H.nv; ++i) H.vinv[H.vperm[i]] = i;
        H.vbin.resize(H.nv);
        for (int i = 0; i < H.nv; ++i) H.vbin[i] = code[H.vperm[i]] >> 32;
        std::vector<uint64_t> ecode(H.ne);
        for (int e = 0; e < H.ne; ++e) ecode[e] = spatial_key(xc[3 * e], xc[3 * e + 1], xc[3 * e + 2], H.G, bcells);
        H.eperm.resize(H.ne);
        std::iota(H.eperm.begin(), H.eperm.end(), 0);
        std::stable_sort(H.eperm.begin(), H.eperm.end(), [&](int a, int bb) { return ecode[a] < ecode[bb]; });
        H.einv.resize(H.ne);
        for (int i = 0; i < H.ne; ++i) H.einv[H.eperm[i]] = i;
        H.ebin.resize(H.ne);
        for (int i = 0; i < H.ne; ++i) H.ebin[i] = ecode[H.eperm[i]] >> 32;
        // the cube's particles in the same spatial order: a cube wider than one
        // patch tile is cut into per-bin patches instead of one patch whose
        // tile would overflow
        std::vector<uint64_t> ccode(H.nc);
        for (int i = 0; i < H.nc; ++i) ccode[i] = spatial_key(H.x0c[3 * i], H.x0c[3 * i + 1], H.x0c[3 * i + 2], H.G, bcells);
        H.cperm.resize(H.nc);
        std::iota(H.cperm.begin(), H.cperm.end(), 0);
        std::stable_sort(H.cperm.begin(), H.cperm.end(), [&](int a, int bb) { return ccode[a] < ccode[bb]; });
        H.cinv.resize(H.nc);
        for (int i = 0; i < H.nc; ++i) H.cinv[H.cperm[i]] = i;
        H.cbin.resize(H.nc);
        for (int i = 0; i < H.nc; ++i) H.cbin[i] = ccode[H.cperm[i]] >> 32;

        InstDev& I = c->inst_h[b];
        memset(&I, 0, sizeof(I));
        I.G = D.n_grid; I.nb = (D.n_grid + 3) / 4;
        I.steps_per_frame = D.steps_per_frame; I.frames_per_window = D.frames_per_window;
        I.group_size = D.n_grasp_points > 0 ? D.n_grasper / D.n_grasp_points : 1;
        if (I.group_size <= 0) I.group_size = 1;
        I.n_actions = D.n_actions; I.n_grasp = D.n_grasp_points;
        I.v_start = np; I.v_count = H.nv; np += H.nv;
        I.o_start = np; I.o_count = H.no; np += H.no;
        I.c_start = np; I.c_count = H.nc; np += H.nc;
        I.e_start = ne; I.e_count = H.ne; ne += H.ne;
        I.dt = (float)D.dt; I.dx = (float)(1.0 / D.n_grid); I.inv_dx = (float)D.n_grid; I.gravity = (float)D.gravity;
        I.gdt = (float)(D.dt * D.gravity);
        I.half_dt2 = (float)(0.5 * D.dt * D.dt);
        I.m_vertex = (float)D.vertex_p_mass; I.m_quad = (float)D.quadrature_p_mass; I.vol_quad = (float)D.quadrature_p_vol;
        I.m_cube = (float)D.cube_p_mass; I.m_op = (float)D.operator_p_mass;
        I.interval = (float)(D.steps_per_frame * D.dt);
        I.E = 5e3f; I.nu = 0.3f; I.ks = 100.f; I.kc = 100.f;
    }
    c->NP = np; c->NE = ne;
    // grids: dense regions per instance, or one shared block pool (its
    // capacity plus the trash slot) with per-instance block tables
    long long cells = 0; int flags = 0, traj = 0;
    const bool pooled = c->opt.grid_pool_blocks > 0;
    for (int b = 0; b < B; ++b) {
        InstDev& I = c->inst_h[b];
        I.grid_base = pooled ? 0 : cells; I.flag_base = flags; I.traj_off = traj;
        I.btab = nullptr; I.pool_trash = 0;
        long long nbl = (long long)I.nb * I.nb * I.nb;
        cells += nbl * 64; flags += (int)nbl; traj += I.n_actions * I.n_grasp;
    }
    if (pooled) {
        c->pool_cap = c->opt.grid_pool_blocks;
        cells = ((long long)c->pool_cap + 1) * 64;
        if (c->alloc(&c->btab, (size_t)flags) || c->alloc(&c->sim.pool_n, 1)) return -1;
        HIP_OK(hipMemset(c->btab, 0xff, (size_t)flags * sizeof(int)));   // every entry -1
        c->sim.pool_cap = c->pool_cap;
        for (int b = 0; b < B; ++b) {
            c->inst_h[b].btab = c->btab + c->inst_h[b].flag_base;
            c->inst_h[b].pool_trash = c->pool_cap;
        }
    }
    c->grid_cells = cells; c->n_flags = flags;
    // static element data in internal order
    c->tri_h.assign(3 * (size_t)ne, 0);
    c->elem_inst_h.assign(ne, 0);
    std::vector<int> inc_cnt(np + 1, 0);
    for (int b = 0; b < B; ++b) {
        const HostInst& H = c->hi[b];
        const InstDev& I = c->inst_h[b];
        for (int ei = 0; ei < H.ne; ++ei) {
            int en = H.eperm[ei];
            int e = I.e_start + ei;
            c->elem_inst_h[e] = b;
            for (int k = 0; k < 3; ++k) {
                int p = I.v_start + H.vinv[H.tri[3 * en + k]];
                c->tri_h[(size_t)k * ne + e] = p;
                inc_cnt[p + 1]++;
            }
        }
    }
    std::vector<int> inc_off(np + 1, 0);
    for (int p = 0; p < np; ++p) inc_off[p + 1] = inc_off[p] + inc_cnt[p + 1];
    std::vector<int> inc_list(std::max(1, inc_off[np])), fillp(inc_off.begin(), inc_off.end() - 1);
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 3; ++k) {
            int p = c->tri_h[(size_t)k * ne + e];
            inc_list[fillp[p]++] = e * 3 + k;
        }
    // natural index of each point (vertex natural index; grasper/cube index in class)
    c->vnat_h.assign(np, 0);
    for (int b = 0; b < B; ++b) {
        const HostInst& H = c->hi[b];
        const InstDev& I = c->inst_h[b];
        for (int i = 0; i < H.nv; ++i) c->vnat_h[I.v_start + i] = H.vperm[i];
        for (int i = 0; i < H.no; ++i) c->vnat_h[I.o_start + i] = i;
        for (int i = 0; i < H.nc; ++i) c->vnat_h[I.c_start + i] = H.cperm[i];
    }
    // patches: point patches (vertex, grasper, cube) then element patches
    // particles per patch: the scatters stage up to 256 particle records in LDS
    const int ps = patch_size;
    auto chunk = [&](int b, int cls, int start, int count) {
        for (int o = 0; o < count; o += ps) c->patches_h.push_back(Patch{b, cls, start + o, std::min(ps, count - o)});
    };
    // runs of equal spatial bin, each cut into patches of at most ps particles
    auto chunk_bins = [&](int b, int cls, int start, const std::vector<uint64_t>& bins) {
        int n = (int)bins.size(), r0 = 0;
        for (int i = 1; i <= n; ++i)
            if (i == n || bins[i] != bins[r0]) { chunk(b, cls, start + r0, i - r0); r0 = i; }
    };
    for (int b = 0; b < B; ++b) {
        const InstDev& I = c->inst_h[b];
        chunk_bins(b, CLS_VERTEX, I.v_start, c->hi[b].vbin);
        for (int gp = 0; gp < std::max(1, I.n_grasp); ++gp) {    // one grasp point per patch
            int g0 = gp * I.group_size, g1 = std::min(I.o_count, g0 + I.group_size);
            if (I.n_grasp == 0) { g0 = 0; g1 = I.o_count; }
            if (g1 > g0) chunk(b, CLS_GRASPER, I.o_start + g0, g1 - g0);
        }
        // one cube patch when its tile fits (the cube moves rigidly), else one per bin
        if (tile_cells(c->hi[b].x0c, c->hi[b].G) <= DCP_STILE_CAP) chunk(b, CLS_CUBE, I.c_start, I.c_count);
        else chunk_bins(b, CLS_CUBE, I.c_start, c->hi[b].cbin);
    }
    for (int b = 0; b < B; ++b) {
        const InstDev& I = c->inst_h[b];
        chunk_bins(b, CLS_ELEMENT, I.e_start, c->hi[b].ebin);
    }
    c->n_patches = (int)c->patches_h.size();
    // scatter patches: per spatial bin, the bin's vertices and elements
    // together, k patches of at most ps particles taking equal shares of both
    // runs (consecutive in the bin order); grasper and cube patches as
    // above
    for (int b = 0; b < B; ++b) {
        const HostInst& H = c->hi[b];
        const InstDev& I = c->inst_h[b];
        int vi = 0, ei = 0;
        while (vi < H.nv || ei < H.ne) {
            const uint64_t bin = vi < H.nv && (ei >= H.ne || H.vbin[vi] <= H.ebin[ei]) ? H.vbin[vi] : H.ebin[ei];
            int v1 = vi, e1 = ei;
            while (v1 < H.nv && H.vbin[v1] == bin) ++v1;
            while (e1 < H.ne && H.ebin[e1] == bin) ++e1;
            const int nv = v1 - vi, nel = e1 - ei;
            const int k = std::max(1, (nv + nel + ps - 3) / (ps - 2));   // shares of floor boundaries stay <= ps
            for (int j = 0; j < k; ++j) {
                const int va = vi + (int)((long long)nv * j / k), vb = vi + (int)((long long)nv * (j + 1) / k);
                const int ea = ei + (int)((long long)nel * j / k), eb = ei + (int)((long long)nel * (j + 1) / k);
                if (vb - va + eb - ea == 0) continue;
                c->spatches_h.push_back(SPatch{b, CLS_VERTEX, I.v_start + va, vb - va, CLS_ELEMENT, I.e_start + ea, eb - ea});
            }
            vi = v1; ei = e1;
        }
    }
    for (const Patch& P : c->patches_h)
        if (P.cls == CLS_GRASPER || P.cls == CLS_CUBE)
            c->spatches_h.push_back(SPatch{P.inst, P.cls, P.start, P.count, CLS_ELEMENT, 0, 0});
    // largest first: the persistent scatter grid works them off in index order,
    // so the last round holds the small patches (grasper, cube, sparse bins)
    std::stable_sort(c->spatches_h.begin(), c->spatches_h.end(), [](const SPatch& x, const SPatch& y) {
        return x.count_a + x.count_b > y.count_a + y.count_b;
    });
    c->n_spatches = (int)c->spatches_h.size();
    // gather patches (G2P, P2G adjoint: per-particle stencil reads, no tile):
    // each class of an instance cut into full patches of ps particles in
    // their bin order, graspers one grasp point per patch; point patches first
    for (int b = 0; b < B; ++b) {
        const InstDev& I = c->inst_h[b];
        auto gchunk = [&](int cls, int start, int count) {
            for (int o = 0; o < count; o += ps) c->gpatches_h.push_back(Patch{b, cls, start + o, std::min(ps, count - o)});
        };
        gchunk(CLS_VERTEX, I.v_start, I.v_count);
        for (const Patch& P : c->patches_h)
            if (P.inst == b && P.cls == CLS_GRASPER) c->gpatches_h.push_back(P);
        gchunk(CLS_CUBE, I.c_start, I.c_count);
    }
    c->n_gpoint = (int)c->gpatches_h.size();
    for (int b = 0; b < B; ++b)
        for (int o = 0; o < c->inst_h[b].e_count; o += ps)
            c->gpatches_h.push_back(Patch{b, CLS_ELEMENT, c->inst_h[b].e_start + o, std::min(ps, c->inst_h[b].e_count - o)});
    c->n_gpatches = (int)c->gpatches_h.size();
    // the partial chunks last in each of the two ranges (as the scatter patches)
    auto fuller = [](const Patch& x, const Patch& y) { return x.count > y.count; };
    std::stable_sort(c->gpatches_h.begin(), c->gpatches_h.begin() + c->n_gpoint, fuller);
    std::stable_sort(c->gpatches_h.begin() + c->n_gpoint, c->gpatches_h.end(), fuller);
    // 4^3 blocks the initial stencils touch (sizes the grid snapshots)
    c->init_blocks = 0;
    for (int b = 0; b < B; ++b) {
        const HostInst& H = c->hi[b];
        const int G = H.G, nb = (G + 3) / 4;
        std::vector<int> blk;
        auto add = [&](const std::vector<float>& x) {
            for (size_t i = 0; i + 2 < x.size(); i += 3) {
                int base[3];
                for (int d = 0; d < 3; ++d) base[d] = std::min(G - 3, std::max(0, (int)(x[i + d] * G - 0.5f)));
                for (int a = base[0] >> 2; a <= (base[0] + 2) >> 2; ++a)
                    for (int bb = base[1] >> 2; bb <= (base[1] + 2) >> 2; ++bb)
                        for (int cc = base[2] >> 2; cc <= (base[2] + 2) >> 2; ++cc) blk.push_back((a * nb + bb) * nb + cc);
            }
        };
        add(H.x0v);
        add(H.x0o);
        add(H.x0c);
        std::sort(blk.begin(), blk.end());
        c->init_blocks += (long long)(std::unique(blk.begin(), blk.end()) - blk.begin());
    }
    // trajectory pool
    std::vector<float> trajh(3 * (size_t)std::max(1, traj), 0.f);
    for (int b = 0; b < B; ++b) {
        const InstDev& I = c->inst_h[b];
        size_t n = (size_t)I.n_actions * I.n_grasp * 3;
        for (size_t i = 0; i < n; ++i) trajh[3 * (size_t)I.traj_off + i] = (float)d[b].trajectory[i];
    }

    // ---- device allocations ----
    SimDev& S = c->sim;
    S.B = B; S.NP = np; S.NE = ne; S.n_patches = c->n_patches;
    S.precision = c->opt.precision; S.svd_grad = c->opt.svd_grad; S.per_elem = c->opt.per_element_moduli;
    S.return_mapping = c->opt.return_mapping != 0;
    S.max_active = flags;
    c->slot_floats = 15 * (size_t)np + 12 * (size_t)ne;
    c->max_states = std::max(2, c->opt.max_states);
    Patch *pd, *gpd; SPatch* spd; int *trid, *eid, *iod, *ild; float* trd;
    if (c->upload(&pd, c->patches_h) || c->upload(&spd, c->spatches_h) || c->upload(&gpd, c->gpatches_h) ||
        c->upload(&trid, c->tri_h) ||
        c->upload(&eid, c->elem_inst_h) ||
        c->upload(&iod, inc_off) || c->upload(&ild, inc_list) || c->upload(&trd, trajh) ||
        c->upload(&c->inst_d, c->inst_h) || c->upload(&c->vnat, c->vnat_h))
        return -1;
    S.inst = c->inst_d; S.patches = pd; S.spatches = spd; S.n_spatches = c->n_spatches;
    S.gpatches = gpd; S.n_gpatches = c->n_gpatches; S.n_gpoint = c->n_gpoint; S.tri = trid; S.elem_inst = eid; S.inc_off = iod; S.inc_list = ild; S.traj = trd;
    // overflow lists of the two scatters; the counts start at zero and the
    // *_big kernels leave them at zero
    if (c->alloc(&S.big_list, 2 * (size_t)std::max(1, c->n_spatches)) || c->upload(&S.big_n, std::vector<int>(4, 0)))
        return -1;
    float *Dinv, *emod;
    if (c->alloc(&Dinv, 9 * (size_t)ne)) return -1;
    S.Dinv = Dinv;
    if (c->opt.per_element_moduli) {
        if (c->alloc(&emod, 3 * (size_t)ne)) return -1;
        S.emod = emod;
    }
    if (c->alloc(&c->prep_work, DCP_PREP_FLOATS * (size_t)ne)) return -1;
    c->use_prep(-1);
    // deterministic: scatter sums in two-word fixed point (8 words per cell),
    // so that a substep's grid does not depend on the order of the atomic
    // adds; otherwise float atomics into acc / gacc
    S.det = c->opt.deterministic != 0;
    if (S.det ? (c->alloc(&S.accq, 8 * (size_t)cells) || c->alloc(&S.gaccq, 8 * (size_t)cells))
              : c->alloc(&S.gacc, (size_t)cells))
        return -1;
    if (c->alloc(&S.poison, 2 * (size_t)B) || c->alloc(&S.damp_fx, 2 * (size_t)B) ||
        c->alloc(&S.traj_fx, 6 * (size_t)std::max(1, traj)) || c->alloc(&S.traj_grad, 3 * (size_t)std::max(1, traj)))
        return -1;
    if (c->alloc(&S.acc, (size_t)cells) ||
        c->alloc(&S.vout, (size_t)cells) || c->alloc(&S.gmv, (size_t)cells) ||
        c->alloc(&S.flags, (size_t)flags) || c->alloc(&S.active, (size_t)flags) || c->alloc(&S.n_active, (size_t)B) ||
        c->alloc(&S.status, 1) || c->alloc(&S.g_xe, 3 * (size_t)ne) ||
        c->alloc(&S.g_fv, 3 * (size_t)np) ||
        c->alloc(&S.ec_gx, 9 * (size_t)ne) || c->alloc(&S.ec_gv, 3 * (size_t)ne) ||
        c->alloc(&S.pgrad_patch, (size_t)DCP_PATCH_GRADS * (DCP_WG / 64) * std::max(1, c->n_gelem())) || c->alloc(&S.pgrad_inst, (size_t)DCP_N_PARAMS * B) ||
        c->alloc(&S.emod_grad, 3 * (size_t)ne))
        return -1;
    if (c->alloc(&c->states, c->slot_floats * c->max_states) || c->alloc(&c->adj, c->slot_floats * c->n_adj) ||
        c->alloc(&c->scratch, c->slot_floats) || c->alloc(&c->loss_d, B) || c->alloc(&c->seed_d, B) ||
        c->alloc(&c->keys_d, 3 * (size_t)(np + ne)))
        return -1;
    c->ones_h.assign(B, 1.0);
    HIP_OK(hipMemcpy(c->seed_d, c->ones_h.data(), B * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipEventCreate(&c->ev0));
    HIP_OK(hipEventCreate(&c->ev1));
    HIP_OK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    HIP_OK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    return 0;
}

// ---------------------------------------------------------------------------
// caller arrays in host or device memory
// ---------------------------------------------------------------------------
// Every array argument of the C-ABI may be host memory or device memory of
// the context's GPU (a torch tensor's data_ptr, say), told apart by HIP's
// unified addressing.  Device arrays of a state slot go through the
// k_state_export / k_state_import kernels; the other inputs are read on the
// host (parameters, targets, positions for the rest shapes: small, and
// reordered or preprocessed there), so device inputs are first copied in;
// results computed on the host or in device buffers are copied to the
// caller's memory, whichever it is.
static bool is_device_ptr(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();   // plain host memory the runtime does not know
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

// the caller's input of n elements as host memory (tmp holds a copy of a device array)
template <typename T>
static int host_in(dcp_ctx* c, const T*& p, size_t n, std::vector<T>& tmp) {
    if (!p || !is_device_ptr(p)) return 0;
    tmp.resize(std::max<size_t>(n, 1));
    HIP_OK(hipMemcpyAsync(tmp.data(), p, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    p = tmp.data();
    return 0;
}

// host result -> the caller's memory (host or device)
static int put_out(dcp_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!dst || !bytes) return 0;
    if (is_device_ptr(dst)) {
        HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    } else {
        std::memcpy(dst, src, bytes);
    }
    return 0;
}

extern "C" int dcp_create(const dcp_instance_desc* instances, int32_t n, const dcp_options* options, void* stream,
                          dcp_ctx** out) {
    if (!instances || n <= 0 || !out) return set_err("dcp_create: bad arguments");
    dcp_ctx* c = new dcp_ctx();
    if (options) c->opt = *options;
    if (c->opt.max_states <= 0) c->opt.max_states = 2;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { delete c; return set_err("no HIP device"); }
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cu > 0) c->n_cu = cu;
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return set_err("stream"); }
        c->own_stream = true;
    }
    if (build(c, instances, n)) {
        std::string e = g_err;
        dcp_destroy(c);
        return set_err(e);
    }
    *out = c;
    return 0;
}

extern "C" int dcp_destroy(dcp_ctx* c) {
    if (!c) return 0;
    c->drop_graphs();
    for (auto& a : c->allocs) hipFree(a.first);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->side) hipStreamDestroy(c->side);
    c->drop_timing_events();
    c->free_snapshots();
    if (c->own_stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

extern "C" int dcp_sizes(dcp_ctx* c, int64_t* o) {
    if (!c || !o) return set_err("bad args");
    o[0] = c->B; o[1] = c->NP; o[2] = c->NE; o[3] = (int64_t)c->slot_floats; o[4] = c->max_states;
    return 0;
}

extern "C" int dcp_pool_blocks(dcp_ctx* c, int32_t* used, int32_t* capacity) {
    if (!c || !used || !capacity) return set_err("bad args");
    *used = 0;
    *capacity = c->pool_cap;
    if (c->pooled()) {
        HIP_OK(hipMemcpyAsync(used, c->sim.pool_n, sizeof(int32_t), hipMemcpyDefault, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    }
    return 0;
}

extern "C" int dcp_device_bytes(dcp_ctx* c, int64_t* out) {
    if (!c || !out) return set_err("bad args");
    *out = (int64_t)(c->alloc_bytes + c->snap_bytes);
    return 0;
}

extern "C" int dcp_set_params(dcp_ctx* c, const double* p) {
    if (!c || !p) return set_err("bad args");
    std::vector<double> tmp;
    if (host_in(c, p, (size_t)c->B * DCP_N_PARAMS, tmp)) return -1;
    for (int b = 0; b < c->B; ++b) {
        const double* q = p + (size_t)b * DCP_N_PARAMS;
        if (!(q[1] > -1.0 && q[1] < 0.5)) return set_err("nu must lie in (-1, 0.5)");
        InstDev& I = c->inst_h[b];
        I.E = (float)q[0]; I.nu = (float)q[1]; I.ks = (float)q[2]; I.kc = (float)q[3];
        I.damping = (float)q[4]; I.friction = (float)q[5];
    }
    HIP_OK(hipMemcpyAsync(c->inst_d, c->inst_h.data(), c->inst_h.size() * sizeof(InstDev), hipMemcpyHostToDevice,
                          c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int dcp_set_element_moduli(dcp_ctx* c, const double* m) {
    if (!c || !m) return set_err("bad args");
    if (!c->opt.per_element_moduli) return set_err("context built without per_element_moduli");
    std::vector<double> tmp;
    if (host_in(c, m, 3 * (size_t)c->NE, tmp)) return -1;
    std::vector<float> h(3 * (size_t)c->NE);
    size_t nat = 0;
    for (int b = 0; b < c->B; ++b) {
        const HostInst& H = c->hi[b];
        const InstDev& I = c->inst_h[b];
        for (int en = 0; en < H.ne; ++en) {
            int e = I.e_start + H.einv[en];
            for (int k = 0; k < 3; ++k) h[(size_t)k * c->NE + e] = (float)m[3 * (nat + en) + k];
        }
        nat += H.ne;
    }
    HIP_OK(hipMemcpy((void*)c->sim.emod, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int dcp_set_targets(dcp_ctx* c, const double* t, const int32_t* counts) {
    if (!c || !counts) return set_err("bad args");
    std::vector<int32_t> ctmp;
    if (host_in(c, counts, (size_t)c->B, ctmp)) return -1;
    int tot = 0;
    for (int b = 0; b < c->B; ++b) {
        if (counts[b] < 0) return set_err("negative target count");
        tot += counts[b];
    }
    if (tot > 0 && !t) return set_err("bad args");
    std::vector<double> ttmp;
    if (host_in(c, t, 3 * (size_t)tot, ttmp)) return -1;
    std::vector<float> h(3 * (size_t)std::max(1, tot), 0.f);
    for (int i = 0; i < 3 * tot; ++i) h[i] = (float)t[i];
    // same counts as the targets in place: only the coordinates change, so
    // the buffers, chunks and every cached graph stay valid
    const bool same = c->targets && (int)c->tcount_h.size() == c->B &&
                      std::equal(c->tcount_h.begin(), c->tcount_h.end(), counts);
    if (same) {
        HIP_OK(hipMemcpyAsync(c->targets, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        return 0;
    }
    // new counts: the loss kernels' buffers, chunk table and launch sizes
    // change, and cached graphs hold the old ones
    HIP_OK(hipStreamSynchronize(c->stream));
    c->drop_graphs();
    c->release(c->targets); c->release(c->tmin); c->release(c->targ); c->release(c->tchunks);
    c->release(c->pmin); c->release(c->pnat); c->release(c->parg);
    c->part_cap = 0;
    c->n_targets_total = 0;
    c->toff_h.assign(c->B, 0);
    c->tcount_h.assign(c->B, 0);
    c->tchunk_h.clear();
    int off = 0;
    for (int b = 0; b < c->B; ++b) {
        c->toff_h[b] = off; c->tcount_h[b] = counts[b];
        for (int o = 0; o < counts[b]; o += DCP_WG) c->tchunk_h.push_back(make_int4(b, off + o, std::min(DCP_WG, counts[b] - o), counts[b]));
        off += counts[b];
    }
    if (c->alloc(&c->targets, h.size()) || c->alloc(&c->tmin, (size_t)std::max(1, tot)) ||
        c->alloc(&c->targ, (size_t)std::max(1, tot)))
        return -1;
    if (!c->toff) {
        if (c->alloc(&c->toff, c->B) || c->alloc(&c->tcount, c->B)) return -1;
    }
    if (c->upload(&c->tchunks, c->tchunk_h)) return -1;
    // vertex parts of the loss minimum: enough work items for 16 per CU, at
    // most 8 parts, every part of every instance with targets at least 128
    // vertices (so the first part of an instance is never empty)
    int minv = 1 << 30;
    for (int b = 0; b < c->B; ++b)
        if (counts[b] > 0) minv = std::min(minv, c->inst_h[b].v_count);
    const int nch = std::max(1, (int)c->tchunk_h.size());
    c->loss_parts = std::max(1, std::min({8, (c->n_cu * 16 + nch - 1) / nch, minv / 128}));
    const size_t need = (size_t)c->loss_parts * std::max(1, tot);
    if (c->alloc(&c->pmin, need) || c->alloc(&c->pnat, need) || c->alloc(&c->parg, need)) return -1;
    c->part_cap = need;
    HIP_OK(hipMemcpy(c->targets, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->toff, c->toff_h.data(), c->B * sizeof(int), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->tcount, c->tcount_h.data(), c->B * sizeof(int), hipMemcpyHostToDevice));
    c->n_targets_total = tot;
    return 0;
}

// initialize_ti_field (simulation.py:447-477) in FP32 on the host
extern "C" int dcp_reset(dcp_ctx* c, int32_t slot, const double* vx, const double* ox, const double* cx) {
    if (!c || !slot_ok(c, slot)) return set_err("dcp_reset: bad slot");
    std::vector<double> vx_h, ox_h, cx_h;   // copies of device inputs
    {
        size_t tv = 0, to = 0, tc = 0;
        for (const HostInst& H : c->hi) { tv += H.nv; to += H.no; tc += H.nc; }
        if (host_in(c, vx, 3 * tv, vx_h) || host_in(c, ox, 3 * to, ox_h) || host_in(c, cx, 3 * tc, cx_h)) return -1;
    }
    const int NP = c->NP, NE = c->NE;
    std::vector<float> st(c->slot_floats, 0.f);
    std::vector<float> Dinv(9 * (size_t)NE);
    size_t nv = 0, no = 0, nc = 0;
    std::vector<float> xv, xo, xc;   // this reset's positions (the creation ones unless given)
    for (int b = 0; b < c->B; ++b) {
        const HostInst& H = c->hi[b];
        const InstDev& I = c->inst_h[b];
        xv = H.x0v; xo = H.x0o; xc = H.x0c;
        if (vx) for (int i = 0; i < 3 * H.nv; ++i) xv[i] = (float)vx[3 * nv + i];
        if (ox) for (int i = 0; i < 3 * H.no; ++i) xo[i] = (float)ox[3 * no + i];
        if (cx) for (int i = 0; i < 3 * H.nc; ++i) xc[i] = (float)cx[3 * nc + i];
        nv += H.nv; no += H.no; nc += H.nc;
        for (int i = 0; i < H.nv; ++i) {
            int p = I.v_start + H.vinv[i];
            for (int k = 0; k < 3; ++k) st[(size_t)k * NP + p] = xv[3 * i + k];
        }
        for (int i = 0; i < H.no; ++i)
            for (int k = 0; k < 3; ++k) st[(size_t)k * NP + I.o_start + i] = xo[3 * i + k];
        for (int i = 0; i < H.nc; ++i)
            for (int k = 0; k < 3; ++k) st[(size_t)k * NP + I.c_start + H.cinv[i]] = xc[3 * i + k];
        for (int ei = 0; ei < H.ne; ++ei) {
            int en = H.eperm[ei], e = I.e_start + ei;
            const float* a = &xv[3 * H.tri[3 * en]];
            const float* bb = &xv[3 * H.tri[3 * en + 1]];
            const float* cc = &xv[3 * H.tri[3 * en + 2]];
            float d1[3], d2[3], n1[3], n2[3], c3[3], n3[3];
            for (int k = 0; k < 3; ++k) { d1[k] = bb[k] - a[k]; d2[k] = cc[k] - a[k]; }
            float l1 = std::sqrt(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2] + 1e-6f);
            float l2 = std::sqrt(d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2] + 1e-6f);
            for (int k = 0; k < 3; ++k) { n1[k] = d1[k] / l1; n2[k] = d2[k] / l2; }
            c3[0] = n1[1] * n2[2] - n1[2] * n2[1];
            c3[1] = n1[2] * n2[0] - n1[0] * n2[2];
            c3[2] = n1[0] * n2[1] - n1[1] * n2[0];
            float l3 = std::sqrt(c3[0] * c3[0] + c3[1] * c3[1] + c3[2] * c3[2] + 1e-6f);
            for (int k = 0; k < 3; ++k) n3[k] = c3[k] / l3;
            float D[9] = {d1[0], d2[0], n3[0], d1[1], d2[1], n3[1], d1[2], d2[2], n3[2]};
            float Di[9];
            inv3h(D, Di);
            for (int k = 0; k < 9; ++k) Dinv[(size_t)k * NE + e] = Di[k];
            for (int k = 0; k < 3; ++k) st[15 * (size_t)NP + (size_t)(9 + k) * NE + e] = n3[k];
        }
    }
    HIP_OK(hipMemcpyAsync((void*)c->sim.Dinv, Dinv.data(), Dinv.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(c->states + (size_t)slot * c->slot_floats, st.data(), st.size() * 4, hipMemcpyHostToDevice,
                          c->stream));
    if (c->reset_pool()) return set_err("dcp_reset: pool reset launch failed");
    HIP_OK(hipStreamSynchronize(c->stream));
    c->rebase_pending = false;
    return 0;
}

extern "C" int dcp_rebase_elements(dcp_ctx* c, int32_t slot) {
    if (!c || !slot_ok(c, slot)) return set_err("dcp_rebase_elements: bad slot");
    if (!c->f_old && c->alloc(&c->f_old, 9 * (size_t)std::max(1, c->NE))) return -1;
    hipLaunchKernelGGL(k_rebase, dim3(c->grid_elems()), dim3(DCP_WG), 0, c->stream, c->sim, c->state(slot), c->f_old,
                       (float*)c->sim.Dinv);
    LAUNCH_OK();
    c->rebase_pending = true;
    return 0;
}

extern "C" int dcp_substep(dcp_ctx* c, int32_t s, int32_t g, int32_t si, int32_t so) {
    if (!c || !slot_ok(c, si) || !slot_ok(c, so) || si == so || s < 0) return set_err("dcp_substep: bad arguments");
    return launch_forward(c, s, g, c->state(si), c->state(so));
}

extern "C" int dcp_substep_adjoint(dcp_ctx* c, int32_t s, int32_t g, int32_t si, int32_t gi, int32_t go) {
    if (!c || !slot_ok(c, si) || !gslot_ok(c, gi) || !gslot_ok(c, go) || gi == go || s < 0)
        return set_err("dcp_substep_adjoint: bad arguments");
    return launch_adjoint(c, s, g, c->state(si), c->adjv(gi), c->adjv(go));
}

extern "C" int dcp_zero_adjoint(dcp_ctx* c, int32_t gs) {
    if (!c || !gslot_ok(c, gs)) return set_err("bad adjoint slot");
    return launch_fill(c, c->adj + (size_t)gs * c->slot_floats, c->slot_floats);
}

extern "C" int dcp_zero_param_grads(dcp_ctx* c) {
    if (!c) return set_err("bad args");
    HIP_OK(hipMemsetAsync(c->sim.pgrad_patch, 0, DCP_PATCH_GRADS * (DCP_WG / 64) * sizeof(double) * c->n_gelem(),
                          c->stream));
    HIP_OK(hipMemsetAsync(c->sim.pgrad_inst, 0, DCP_N_PARAMS * sizeof(double) * c->B, c->stream));
    HIP_OK(hipMemsetAsync(c->sim.emod_grad, 0, 3 * sizeof(double) * c->NE, c->stream));
    HIP_OK(hipMemsetAsync(c->sim.damp_fx, 0, 2 * sizeof(unsigned long long) * c->B, c->stream));
    size_t nt = 0;
    for (auto& I : c->inst_h) nt += (size_t)I.n_actions * I.n_grasp;
    HIP_OK(hipMemsetAsync(c->sim.traj_fx, 0, 6 * sizeof(unsigned long long) * std::max<size_t>(1, nt), c->stream));
    HIP_OK(hipMemsetAsync(c->sim.traj_grad, 0, 3 * sizeof(double) * std::max<size_t>(1, nt), c->stream));
    return 0;
}

// value of a two-word fixed-point sum (hi 2^-24 + lo 2^-62), as the kernels
// add them
static double fx_value(unsigned long long hi, unsigned long long lo) {
    return std::ldexp((double)(long long)hi, -24) + std::ldexp((double)(long long)lo, -62);
}

extern "C" int dcp_loss(dcp_ctx* c, int32_t slot, double* out) {
    if (!c || !slot_ok(c, slot)) return set_err("dcp_loss: bad slot");
    if (launch_loss(c, c->state(slot))) return -1;
    if (out) {
        HIP_OK(hipMemcpyAsync(out, c->loss_d, c->B * sizeof(double), hipMemcpyDefault, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    }
    return 0;
}

extern "C" int dcp_loss_adjoint(dcp_ctx* c, int32_t slot, int32_t gs, const double* seed) {
    if (!c || !slot_ok(c, slot) || !gslot_ok(c, gs)) return set_err("dcp_loss_adjoint: bad slot");
    if (seed) HIP_OK(hipMemcpyAsync(c->seed_d, seed, c->B * sizeof(double), hipMemcpyDefault, c->stream));
    if (launch_loss(c, c->state(slot))) return -1;
    if (launch_loss_adj(c, c->state(slot), c->adjv(gs))) return -1;
    HIP_OK(hipStreamSynchronize(c->stream));
    return 0;
}

static int reduce_param_grads(dcp_ctx* c, double* out_caller) {
    std::vector<double> outv((size_t)DCP_N_PARAMS * c->B);
    double* out = outv.data();
    const int nw = DCP_WG / 64;
    std::vector<double> pp((size_t)DCP_PATCH_GRADS * nw * c->n_gelem()), pi((size_t)DCP_N_PARAMS * c->B);
    HIP_OK(hipMemcpyAsync(pp.data(), c->sim.pgrad_patch, pp.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipMemcpyAsync(pi.data(), c->sim.pgrad_inst, pi.size() * 8, hipMemcpyDeviceToHost, c->stream));
    std::vector<unsigned long long> dfx(2 * (size_t)c->B);
    HIP_OK(hipMemcpyAsync(dfx.data(), c->sim.damp_fx, dfx.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    for (int b = 0; b < c->B; ++b)
        for (int k = 0; k < DCP_N_PARAMS; ++k) out[b * DCP_N_PARAMS + k] = pi[b * DCP_N_PARAMS + k];
    for (int b = 0; b < c->B; ++b) out[b * DCP_N_PARAMS + 4] += fx_value(dfx[2 * b], dfx[2 * b + 1]);
    for (int p = 0; p < c->n_gelem(); ++p) {
        const Patch& P = c->gpatches_h[c->n_gpoint + p];
        // E, nu, shearing, contact, friction (parameter slots 0-3 and 5), waves in order
        for (int w = 0; w < nw; ++w) {
            const double* q = &pp[DCP_PATCH_GRADS * ((size_t)p * nw + w)];
            for (int k = 0; k < 4; ++k) out[P.inst * DCP_N_PARAMS + k] += q[k];
            out[P.inst * DCP_N_PARAMS + 5] += q[4];
        }
    }
    return put_out(c, out_caller, out, outv.size() * sizeof(double));
}

extern "C" int dcp_param_grads(dcp_ctx* c, double* out) {
    if (!c || !out) return set_err("bad args");
    return reduce_param_grads(c, out);
}

extern "C" int dcp_element_moduli_grads(dcp_ctx* c, double* out) {
    if (!c || !out) return set_err("bad args");
    std::vector<double> h(3 * (size_t)c->NE), o(3 * (size_t)c->NE);
    HIP_OK(hipMemcpyAsync(h.data(), c->sim.emod_grad, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    size_t nat = 0;
    for (int b = 0; b < c->B; ++b) {
        const HostInst& H = c->hi[b];
        const InstDev& I = c->inst_h[b];
        for (int en = 0; en < H.ne; ++en) {
            int e = I.e_start + H.einv[en];
            for (int k = 0; k < 3; ++k) o[3 * (nat + en) + k] = h[(size_t)k * c->NE + e];
        }
        nat += H.ne;
    }
    return put_out(c, out, o.data(), o.size() * sizeof(double));
}

extern "C" int dcp_set_trajectory(dcp_ctx* c, const double* traj) {
    if (!c || !traj) return set_err("bad args");
    size_t nt = 0;
    for (auto& I : c->inst_h) nt += (size_t)I.n_actions * I.n_grasp;
    std::vector<double> tmp;
    if (host_in(c, traj, 3 * nt, tmp)) return -1;
    c->traj_h.resize(3 * nt);
    for (size_t i = 0; i < 3 * nt; ++i) c->traj_h[i] = (float)traj[i];
    HIP_OK(hipMemcpyAsync((void*)c->sim.traj, c->traj_h.data(), 3 * nt * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int dcp_trajectory_grads(dcp_ctx* c, double* out) {
    if (!c || !out) return set_err("bad args");
    size_t nt = 0;
    for (auto& I : c->inst_h) nt += (size_t)I.n_actions * I.n_grasp;
    std::vector<unsigned long long> q(6 * std::max<size_t>(1, nt));
    std::vector<double> g(3 * std::max<size_t>(1, nt));
    HIP_OK(hipMemcpyAsync(q.data(), c->sim.traj_fx, q.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipMemcpyAsync(g.data(), c->sim.traj_grad, g.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    g.resize(3 * nt);   // one of the two forms holds zeros
    for (size_t i = 0; i < 3 * nt; ++i) g[i] += fx_value(q[2 * i], q[2 * i + 1]);
    HIP_OK(hipMemcpy(out, g.data(), g.size() * 8, hipMemcpyDefault));
    return 0;
}

// natural-order host arrays <-> one device slot (11 arrays, any may be null)
struct HostArrays {
    float* a[11];
};

static void slot_to_host(const dcp_ctx* c, const std::vector<float>& h, const HostArrays& A) {
    const int NP = c->NP, NE = c->NE;
    size_t nv = 0, no = 0, nc = 0, ne = 0;
    auto put_point = [&](int p, size_t idx, float* x, float* v, float* C) {
        for (int k = 0; k < 3; ++k) {
            if (x) x[3 * idx + k] = h[(size_t)k * NP + p];
            if (v) v[3 * idx + k] = h[(size_t)(3 + k) * NP + p];
        }
        if (C) for (int k = 0; k < 9; ++k) C[9 * idx + k] = h[(size_t)(6 + k) * NP + p];
    };
    for (int b = 0; b < c->B; ++b) {
        const HostInst& H = c->hi[b];
        const InstDev& I = c->inst_h[b];
        for (int i = 0; i < H.nv; ++i) put_point(I.v_start + H.vinv[i], nv + i, A.a[0], A.a[1], A.a[2]);
        for (int i = 0; i < H.no; ++i) put_point(I.o_start + i, no + i, A.a[3], A.a[4], A.a[5]);
        for (int i = 0; i < H.nc; ++i) put_point(I.c_start + H.cinv[i], nc + i, A.a[6], A.a[7], A.a[8]);
        for (int en0 = 0; en0 < H.ne; ++en0) {
            int e = I.e_start + H.einv[en0];
            if (A.a[9]) for (int k = 0; k < 9; ++k) A.a[9][9 * (ne + en0) + k] = h[15 * (size_t)NP + (size_t)k * NE + e];
            if (A.a[10]) for (int k = 0; k < 3; ++k) A.a[10][3 * (ne + en0) + k] = h[15 * (size_t)NP + (size_t)(9 + k) * NE + e];
        }
        nv += H.nv; no += H.no; nc += H.nc; ne += H.ne;
    }
}

static void host_to_slot(const dcp_ctx* c, std::vector<float>& h, const HostArrays& A) {
    const int NP = c->NP, NE = c->NE;
    size_t nv = 0, no = 0, nc = 0, ne = 0;
    auto get_point = [&](int p, size_t idx, const float* x, const float* v, const float* C) {
        for (int k = 0; k < 3; ++k) {
            if (x) h[(size_t)k * NP + p] = x[3 * idx + k];
            if (v) h[(size_t)(3 + k) * NP + p] = v[3 * idx + k];
        }
        if (C) for (int k = 0; k < 9; ++k) h[(size_t)(6 + k) * NP + p] = C[9 * idx + k];
    };
    for (int b = 0; b < c->B; ++b) {
        const HostInst& H = c->hi[b];
        const InstDev& I = c->inst_h[b];
        for (int i = 0; i < H.nv; ++i) get_point(I.v_start + H.vinv[i], nv + i, A.a[0], A.a[1], A.a[2]);
        for (int i = 0; i < H.no; ++i) get_point(I.o_start + i, no + i, A.a[3], A.a[4], A.a[5]);
        for (int i = 0; i < H.nc; ++i) get_point(I.c_start + H.cinv[i], nc + i, A.a[6], A.a[7], A.a[8]);
        for (int en0 = 0; en0 < H.ne; ++en0) {
            int e = I.e_start + H.einv[en0];
            if (A.a[9]) for (int k = 0; k < 9; ++k) h[15 * (size_t)NP + (size_t)k * NE + e] = A.a[9][9 * (ne + en0) + k];
            if (A.a[10]) for (int k = 0; k < 3; ++k) h[15 * (size_t)NP + (size_t)(9 + k) * NE + e] = A.a[10][3 * (ne + en0) + k];
        }
        nv += H.nv; no += H.no; nc += H.nc; ne += H.ne;
    }
}

// 1: every given array is device memory, 0: every one host memory, -1: a mix
static int arrays_kind(const HostArrays& A) {
    int dev = 0, host = 0;
    for (float* a : A.a) {
        if (!a) continue;
        if (is_device_ptr(a)) ++dev; else ++host;
    }
    return dev && host ? -1 : (dev ? 1 : 0);
}

static int state_io(dcp_ctx* c, const HostArrays& A, StateIO& io) {
    if (!c->io_perm) {   // natural particle -> slot index, built on first use
        std::vector<int> perm;
        for (int cls = 0; cls < 4; ++cls)
            for (int b = 0; b < c->B; ++b) {
                const HostInst& H = c->hi[b];
                const InstDev& I = c->inst_h[b];
                if (cls == 0) for (int i = 0; i < H.nv; ++i) perm.push_back(I.v_start + H.vinv[i]);
                if (cls == 1) for (int i = 0; i < H.no; ++i) perm.push_back(I.o_start + i);
                if (cls == 2) for (int i = 0; i < H.nc; ++i) perm.push_back(I.c_start + H.cinv[i]);
                if (cls == 3) for (int i = 0; i < H.ne; ++i) perm.push_back(I.e_start + H.einv[i]);
            }
        if (c->upload(&c->io_perm, perm)) return -1;
    }
    for (int k = 0; k < 11; ++k) io.a[k] = A.a[k];
    io.perm = c->io_perm;
    io.n[0] = io.n[1] = io.n[2] = io.n[3] = 0;
    for (const HostInst& H : c->hi) { io.n[0] += H.nv; io.n[1] += H.no; io.n[2] += H.nc; io.n[3] += H.ne; }
    return 0;
}

static int download_to(dcp_ctx* c, const float* dev, const HostArrays& A) {
    const int kind = arrays_kind(A);
    if (kind < 0) return set_err("state arrays: all host or all device memory");
    if (kind == 1) {
        StateIO io;
        if (state_io(c, A, io)) return -1;
        const int n = io.n[0] + io.n[1] + io.n[2] + io.n[3];
        hipLaunchKernelGGL(k_state_export, dim3((n + 255) / 256), dim3(256), 0, c->stream, dev, c->NP, c->NE, io);
        LAUNCH_OK();
        HIP_OK(hipStreamSynchronize(c->stream));
        return 0;
    }
    std::vector<float> h(c->slot_floats);
    HIP_OK(hipMemcpyAsync(h.data(), dev, h.size() * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    slot_to_host(c, h, A);
    return 0;
}

static int upload_from(dcp_ctx* c, float* dev, const HostArrays& A) {
    const int kind = arrays_kind(A);
    if (kind < 0) return set_err("state arrays: all host or all device memory");
    if (kind == 1) {
        StateIO io;
        if (state_io(c, A, io) || launch_fill(c, dev, c->slot_floats)) return -1;
        const int n = io.n[0] + io.n[1] + io.n[2] + io.n[3];
        hipLaunchKernelGGL(k_state_import, dim3((n + 255) / 256), dim3(256), 0, c->stream, dev, c->NP, c->NE, io);
        LAUNCH_OK();
        HIP_OK(hipStreamSynchronize(c->stream));
        return 0;
    }
    std::vector<float> h(c->slot_floats, 0.f);
    host_to_slot(c, h, A);
    HIP_OK(hipMemcpyAsync(dev, h.data(), h.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return 0;
}

#define ARRS11 float *vx, float *vv, float *vC, float *ox, float *ov, float *oC, float *cx, float *cv, float *cC, float *eC, float *en
#define CARRS11 const float *vx, const float *vv, const float *vC, const float *ox, const float *ov, const float *oC, \
    const float *cx, const float *cv, const float *cC, const float *eC, const float *en
#define PACK11 HostArrays{{(float*)vx, (float*)vv, (float*)vC, (float*)ox, (float*)ov, (float*)oC, (float*)cx, (float*)cv, (float*)cC, (float*)eC, (float*)en}}

extern "C" int dcp_download_state(dcp_ctx* c, int32_t slot, ARRS11) {
    if (!c || !slot_ok(c, slot)) return set_err("dcp_download_state: bad slot");
    return download_to(c, c->states + (size_t)slot * c->slot_floats, PACK11);
}

extern "C" int dcp_upload_state(dcp_ctx* c, int32_t slot, CARRS11) {
    if (!c || !slot_ok(c, slot)) return set_err("dcp_upload_state: bad slot");
    return upload_from(c, c->states + (size_t)slot * c->slot_floats, PACK11);
}

extern "C" int dcp_download_adjoint(dcp_ctx* c, int32_t gs, ARRS11) {
    if (!c || !gslot_ok(c, gs)) return set_err("dcp_download_adjoint: bad slot");
    return download_to(c, c->adj + (size_t)gs * c->slot_floats, PACK11);
}

extern "C" int dcp_upload_adjoint(dcp_ctx* c, int32_t gs, CARRS11) {
    if (!c || !gslot_ok(c, gs)) return set_err("dcp_upload_adjoint: bad slot");
    return upload_from(c, c->adj + (size_t)gs * c->slot_floats, PACK11);
}

extern "C" int dcp_binning_keys(dcp_ctx* c, int32_t slot, int32_t* out) {
    if (!c || !slot_ok(c, slot) || !out) return set_err("dcp_binning_keys: bad args");
    // element positions are the vertex means computed by the prep kernel; a
    // slot is "fresh" only for slot-0 states right after dcp_reset, which the
    // caller signals through the sign of nothing: use the non-fresh form.
    if (launch_prep(c, c->state(slot), c->scratch_view(), 0)) return -1;
    hipLaunchKernelGGL(k_binning_keys, dim3(c->grid_patches()), dim3(DCP_WG), 0, c->stream, c->sim, c->state(slot), c->keys_d);
    LAUNCH_OK();
    std::vector<int> h(3 * (size_t)(c->NP + c->NE));
    HIP_OK(hipMemcpyAsync(h.data(), c->keys_d, h.size() * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    // natural order: per instance, vertices, grasper, cube, then elements
    size_t o = 0;
    for (int b = 0; b < c->B; ++b) {
        const HostInst& H = c->hi[b];
        const InstDev& I = c->inst_h[b];
        for (int i = 0; i < H.nv; ++i, ++o) for (int k = 0; k < 3; ++k) out[3 * o + k] = h[3 * (size_t)(I.v_start + H.vinv[i]) + k];
        for (int i = 0; i < H.no; ++i, ++o) for (int k = 0; k < 3; ++k) out[3 * o + k] = h[3 * (size_t)(I.o_start + i) + k];
        for (int i = 0; i < H.nc; ++i, ++o) for (int k = 0; k < 3; ++k) out[3 * o + k] = h[3 * (size_t)(I.c_start + H.cinv[i]) + k];
        for (int en = 0; en < H.ne; ++en, ++o)
            for (int k = 0; k < 3; ++k) out[3 * o + k] = h[3 * (size_t)(c->NP + I.e_start + H.einv[en]) + k];
    }
    return 0;
}

extern "C" int dcp_patch_order(dcp_ctx* c, int32_t slot, int32_t* out) {
    if (!c || !slot_ok(c, slot) || !out) return set_err("dcp_patch_order: bad args");
    // element positions are the vertex means of the prep (non-fresh form, as dcp_binning_keys)
    if (launch_prep(c, c->state(slot), c->scratch_view(), 0)) return -1;
    int* d = nullptr;
    const size_t n = 4 * (size_t)c->n_spatches * DCP_WG;
    if (c->alloc(&d, n)) return -1;
    hipLaunchKernelGGL(k_patch_order, dim3(c->grid_spatches()), dim3(DCP_WG), 0, c->stream, c->sim, c->state(slot), d);
    hipError_t le = hipGetLastError();
    std::vector<int> h(n);
    hipError_t ce = le == hipSuccess ? hipMemcpyAsync(h.data(), d, n * sizeof(int), hipMemcpyDeviceToHost, c->stream) : le;
    hipError_t se = hipStreamSynchronize(c->stream);
    c->release(d);
    if (ce != hipSuccess) return set_err(std::string("dcp_patch_order: ") + hipGetErrorString(ce));
    if (se != hipSuccess) return set_err(std::string("dcp_patch_order: ") + hipGetErrorString(se));
    size_t row = 0;
    for (int pi = 0; pi < c->n_spatches; ++pi) {
        const SPatch& P = c->spatches_h[pi];
        const HostInst& H = c->hi[P.inst];
        const InstDev& I = c->inst_h[P.inst];
        for (int q = 0; q < P.count_a + P.count_b; ++q, ++row) {
            const bool a = q < P.count_a;
            const int cls = a ? P.cls_a : P.cls_b, idx = a ? P.start_a + q : P.start_b + (q - P.count_a);
            int nat = 0;
            if (cls == CLS_ELEMENT) nat = H.eperm[idx - I.e_start];
            else nat = c->vnat_h[idx];
            int32_t* o = out + 9 * row;
            const int* k = &h[4 * ((size_t)pi * DCP_WG + q)];
            o[0] = pi; o[1] = q; o[2] = P.inst; o[3] = cls; o[4] = nat;
            o[5] = k[0]; o[6] = k[1]; o[7] = k[2]; o[8] = k[3];
        }
    }
    return 0;
}

extern "C" int dcp_status(dcp_ctx* c, uint32_t* out, int32_t clear) {
    if (!c || !out) return set_err("bad args");
    HIP_OK(hipMemcpyAsync(out, c->sim.status, 4, hipMemcpyDeviceToHost, c->stream));
    if (clear) HIP_OK(hipMemsetAsync(c->sim.status, 0, 4, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// checkpointed rollouts, one hipGraph per segment
// ---------------------------------------------------------------------------
// slot map: checkpoints at [0, nck], segment buffer at [nck+1, nck+1+K]
struct Plan {
    int S, K, L, nseg, nck, seg0;
    int ck(int c) const { return c; }
    int seg(int j) const { return seg0 + j; }
};

static int plan_rollout(dcp_ctx* c, int S, int K, Plan& p) {
    if (S < 3) return set_err("n_substeps must be >= 3 (the loss reads state n_substeps-2)");
    if (K <= 0 || K > S) K = S;
    p.S = S; p.K = K; p.L = S - 2;
    p.nseg = (S + K - 1) / K;
    p.nck = p.nseg;
    p.seg0 = p.nck + 1;
    if (p.seg0 + K + 1 > c->max_states)
        return set_err("max_states too small for checkpoint_every=" + std::to_string(K) + ": need " +
                       std::to_string(p.seg0 + K + 1));
    return 0;
}

// slot holding state t inside segment cseg (states cseg*K .. )
static int seg_slot(const Plan& p, int cseg, int t) {
    int base = cseg * p.K;
    if (t == base) return p.ck(cseg);
    int end = std::min(base + p.K, p.S);
    if (t == end) return p.ck(cseg + 1);
    return p.seg(t - base);
}

static int record_forward_segment(dcp_ctx* c, const Plan& p, int cseg, int g) {
    int base = cseg * p.K, end = std::min(base + p.K, p.S);
    // the last segment's states stay resident for the backward: keep its grids too
    const bool keep_grids = cseg == p.nseg - 1 && c->snap_slots >= p.K;
    for (int f = base; f < end; ++f) {
        const int snap = keep_grids && f < p.L ? f - base : -1;
        if (launch_forward(c, f, g, c->state(seg_slot(p, cseg, f)), c->state(seg_slot(p, cseg, f + 1)), snap)) return -1;
        if (f + 1 == p.L) {
            if (launch_loss(c, c->state(seg_slot(p, cseg, f + 1)))) return -1;
            if (launch_fill(c, c->adj, c->slot_floats)) return -1;
            if (launch_loss_adj(c, c->state(seg_slot(p, cseg, f + 1)), c->adjv(0))) return -1;
        }
    }
    return 0;
}

// adjoint slots ping-pong: gradient of state t lives in adj[(L - t) & 1]
// Recompute of a segment's states base+1 .. end-1 before its adjoint walk
// (every segment but the last forward one, whose states are still resident),
// saving the grids of all its substeps but the last.
static int record_recompute_segment(dcp_ctx* c, const Plan& p, int cseg, int g) {
    const int base = cseg * p.K, end = std::min(base + p.K, p.L);
    const bool grids = c->snap_slots >= p.K;
    for (int f = base; f + 1 < end; ++f)
        if (launch_forward(c, f, g, c->state(seg_slot(p, cseg, f)), c->state(seg_slot(p, cseg, f + 1)),
                           grids ? f - base : -1))
            return -1;
    return 0;
}

// Snapshot slot of adjoint substep f of segment cseg (-1: none): every
// substep of a resident segment, all but the last one otherwise.
static int segment_snap(const dcp_ctx* c, const Plan& p, int cseg, int f) {
    const int base = cseg * p.K, end = std::min(base + p.K, p.L);
    const bool resident = cseg == p.nseg - 1;
    return c->snap_slots >= p.K && (resident || f + 1 < end) ? f - base : -1;
}

// Whether every grid snapshot the adjoint walk of segment cseg reads holds its
// substep, from the flags the segment's forward left (waits for it), and
// whether no scatter patch overflowed its tile since the last look (the count
// k_p2g keeps in big_n[DCP_BIG_SEEN], cleared here).
static int segment_snapshots_hold(dcp_ctx* c, const Plan& p, int cseg, bool& hold, bool& no_big) {
    hold = false;
    no_big = false;
    const int base = cseg * p.K, end = std::min(base + p.K, p.L);
    if (end <= base || !c->snap_acc || !c->snap_prep) return 0;
    std::vector<int> ok((size_t)c->snap_slots, 0);
    int seen = 1;
    HIP_OK(hipMemcpyAsync(ok.data(), c->snap_ok, ok.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipMemcpyAsync(&seen, c->sim.big_n + DCP_BIG_SEEN, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipMemsetAsync(c->sim.big_n + DCP_BIG_SEEN, 0, sizeof(int), c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    for (int f = base; f < end; ++f) {
        const int sn = segment_snap(c, p, cseg, f);
        if (sn >= 0 && !ok[sn]) return 0;
    }
    hold = true;
    no_big = seen == 0;
    return 0;
}

static int record_backward_segment(dcp_ctx* c, const Plan& p, int cseg, int g) {
    int base = cseg * p.K;
    int end = std::min(base + p.K, p.L);   // adjoint substeps base .. end-1
    if (end <= base) return 0;
    auto snap_of = [&](int f) { return segment_snap(c, p, cseg, f); };
    bool grid_ready = false;   // the grid of substep f was rebuilt on the side stream
    for (int f = end - 1; f >= base; --f) {
        int gin = (p.L - (f + 1)) & 1, gout = (p.L - f) & 1;
        const int snap = snap_of(f);
        const StateView S = c->state(seg_slot(p, cseg, f));
        // the window's first adjoint substep reads the loss adjoint (complete);
        // its last one completes the gradient of the initial state
        const int pending = f + 1 < p.L;
        if (!grid_ready && launch_adjoint_grid(c, f, g, S, snap)) return -1;
        if (launch_adjoint_head(c, S, c->adjv(gin), c->adjv(gout), snap, pending)) return -1;
        // Substep f-1's grid restore needs only this substep's grid-update
        // adjoint to have run (it rewrites the accumulators, flags and lists
        // that one cleared): it goes on the side stream beside this substep's
        // tail, when both substeps use saved element prep (no shared arrays).
        grid_ready = f - 1 >= base && prep_saved(c, snap) && prep_saved(c, snap_of(f - 1));
        if (grid_ready) {
            HIP_OK(hipEventRecord(c->ev_fork, c->stream));
            HIP_OK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
            hipStream_t main_stream = c->stream;
            c->stream = c->side;
            const int r = launch_adjoint_grid(c, f - 1, g, c->state(seg_slot(p, cseg, f - 1)), snap_of(f - 1));
            c->stream = main_stream;
            if (r) return -1;
            HIP_OK(hipEventRecord(c->ev_join, c->side));
        }
        if (launch_adjoint_tail(c, f, g, S, c->adjv(gin), c->adjv(gout), snap, f == 0)) return -1;
        if (grid_ready) HIP_OK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    }
    return 0;
}

// Graph kinds: forward segment, adjoint walk (full: P2G and the grid update
// launched on every substep, skipping themselves where the snapshot holds),
// recompute of a segment's states, lean adjoint walk (no such launches), and
// the lean walk without the G2P adjoint's sub-box kernel (no tile overflowed
// in the segment's forward).
enum { REC_FORWARD = 0, REC_ADJOINT = 1, REC_RECOMPUTE = 2, REC_ADJOINT_LEAN = 3, REC_ADJOINT_LEAN_NOBIG = 4,
       REC_FORWARD_NOBIG = 5 };
static int run_recorded(dcp_ctx* c, int kind, int cseg, int g, const Plan& p, int use_graph) {
    auto rec = [&]() {
        if (kind == REC_FORWARD || kind == REC_FORWARD_NOBIG) {
            c->fwd_big_left_out = kind == REC_FORWARD_NOBIG;
            const int r = record_forward_segment(c, p, cseg, g);
            c->fwd_big_left_out = false;
            return r;
        }
        if (kind == REC_RECOMPUTE) return record_recompute_segment(c, p, cseg, g);
        c->lean_adjoint = kind == REC_ADJOINT_LEAN || kind == REC_ADJOINT_LEAN_NOBIG;
        c->adj_big_left_out = kind == REC_ADJOINT_LEAN_NOBIG;
        const int r = record_backward_segment(c, p, cseg, g);
        c->lean_adjoint = false;
        c->adj_big_left_out = false;
        return r;
    };
    if (!use_graph) return rec();
    auto key = std::make_tuple(kind, cseg, g, p.S, p.K);
    auto it = c->graphs.find(key);
    if (it == c->graphs.end()) {
        hipGraph_t graph;
        HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        c->capturing = true;
        int r = rec();
        c->capturing = false;
        hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (r) return r;
        if (e != hipSuccess) return set_err(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        hipGraphExec_t exec;
        HIP_OK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        hipGraphDestroy(graph);
        it = c->graphs.emplace(key, exec).first;
    }
    HIP_OK(hipGraphLaunch(it->second, c->stream));
    return 0;
}

// Grid snapshot slots for a plan with checkpoint interval K (one per substep of
// a segment).  Capacity per slot: 4x the blocks the initial stencils touch
// (deformation headroom), within the instances' blocks, and all slots within a
// quarter of the free device memory; below 256 blocks per slot, none.
static int ensure_snapshots(dcp_ctx* c, int K) {
    if (c->opt.grid_snapshots < 0 || c->snap_slots >= K) return 0;
    long long all = 0;
    for (auto& I : c->inst_h) all += (long long)I.nb * I.nb * I.nb;
    size_t free_b = 0, total_b = 0;
    HIP_OK(hipMemGetInfo(&free_b, &total_b));
    // cached graphs hold the old buffers
    HIP_OK(hipStreamSynchronize(c->stream));
    c->drop_graphs();
    c->free_snapshots();
    const size_t per_block = 64 * sizeof(float4) + sizeof(int);
    // element prep slots when they fit in an eighth of the free memory
    const size_t prep_b = (size_t)DCP_PREP_FLOATS * c->NE * sizeof(float) * (size_t)K;
    const bool prep = prep_b <= free_b / 8;
    const size_t grid_budget = free_b / 4 - (prep ? prep_b : 0);
    long long cap = std::min(all, std::max<long long>(1024, 4 * c->init_blocks));
    cap = std::min<long long>(cap, (long long)(grid_budget / per_block / (size_t)K));
    if (c->opt.grid_snapshots > 0)
        cap = std::min<long long>(cap, c->opt.grid_snapshots);   // the caller's bound on blocks per slot
    else if (cap < 256)
        return 0;
    if (cap < 1) return 0;
    if (hipMalloc(&c->snap_acc, (size_t)K * cap * 64 * sizeof(float4)) != hipSuccess ||
        hipMalloc(&c->snap_blk, (size_t)K * cap * sizeof(int)) != hipSuccess ||
        hipMalloc(&c->snap_cnt, (size_t)K * c->B * sizeof(int)) != hipSuccess ||
        hipMalloc(&c->snap_ok, (size_t)K * sizeof(int)) != hipSuccess) {
        (void)hipGetLastError();
        c->free_snapshots();
        return 0;   // no snapshots: the adjoint recomputes its grids
    }
    if (prep && hipMalloc(&c->snap_prep, prep_b) != hipSuccess) {
        (void)hipGetLastError();
        c->snap_prep = nullptr;   // grids only: the adjoint recomputes the element prep
    }
    c->snap_cap = (int)cap;
    c->snap_slots = K;
    c->snap_bytes = (size_t)K * cap * (64 * sizeof(float4) + sizeof(int)) + (size_t)K * (c->B + 1) * sizeof(int) +
                    (c->snap_prep ? prep_b : 0);
    return 0;
}

extern "C" int dcp_rollout_grad(dcp_ctx* c, int32_t S, int32_t g, int32_t K, int32_t use_graph, double* loss_out,
                                double* grads_out) {
    if (!c) return set_err("bad args");
    if (c->rebase_pending) return set_err("a rest-shape re-base is pending: advance it with dcp_substep");
    Plan p;
    if (plan_rollout(c, S, K, p)) return -1;
    if (ensure_snapshots(c, p.K)) return -1;
    if (c->n_targets_total == 0) return set_err("no targets set (dcp_set_targets)");
    for (;;) {
        const int fkind = c->fwd_big_seen ? REC_FORWARD : REC_FORWARD_NOBIG;
        // the reset state must already be in slot 0 (dcp_reset); pooled grids
        // start the rollout with an empty block table
        if (c->reset_pool()) return set_err("pool reset launch failed");
        // the overflow count the adjoint walks consult starts with this rollout
        HIP_OK(hipMemsetAsync(c->sim.big_n + DCP_BIG_SEEN, 0, sizeof(int), c->stream));
        // timing events restart with an eager rollout or with a plan's first capture
        if (!use_graph || !c->graphs.count(std::make_tuple(fkind, 0, g, p.S, p.K))) c->tev_used = 0;
        HIP_OK(hipEventRecord(c->ev0, c->stream));
        // unit loss seeds (dcp_loss_adjoint may have left others)
        HIP_OK(hipMemcpyAsync(c->seed_d, c->ones_h.data(), c->B * sizeof(double), hipMemcpyHostToDevice, c->stream));
        if (dcp_zero_param_grads(c)) return -1;
        for (int cs = 0; cs < p.nseg; ++cs)
            if (run_recorded(c, fkind, cs, g, p, use_graph)) return -1;
        if (fkind == REC_FORWARD) break;
        // a forward without the sub-box kernel is valid only if no patch was
        // listed for it; otherwise run the window again with it (slot 0 still
        // holds the reset state), and keep it from now on
        int seen = 0;
        HIP_OK(hipMemcpyAsync(&seen, c->sim.big_n + DCP_BIG_SEEN, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        if (seen == 0) break;
        c->fwd_big_seen = true;
    }
    HIP_OK(hipMemcpyAsync(loss_out ? loss_out : nullptr, c->loss_d, loss_out ? c->B * sizeof(double) : 0,
                          hipMemcpyDefault, c->stream));
    int last = (p.L - 1) / p.K;
    for (int cs = last; cs >= 0; --cs) {
        if (cs != p.nseg - 1 && run_recorded(c, REC_RECOMPUTE, cs, g, p, use_graph)) return -1;
        bool hold = false, no_big = false;
        if (segment_snapshots_hold(c, p, cs, hold, no_big)) return -1;
        const int kind = !hold ? REC_ADJOINT : no_big ? REC_ADJOINT_LEAN_NOBIG : REC_ADJOINT_LEAN;
        if (run_recorded(c, kind, cs, g, p, use_graph)) return -1;
    }
    HIP_OK(hipEventRecord(c->ev1, c->stream));
    HIP_OK(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_final_slot = seg_slot(p, p.nseg - 1, p.S);
    if (grads_out && reduce_param_grads(c, grads_out)) return -1;
    return 0;
}

extern "C" int dcp_rollout(dcp_ctx* c, int32_t S, int32_t g, int32_t use_graph, int32_t* final_slot) {
    if (!c) return set_err("bad args");
    if (c->rebase_pending) return set_err("a rest-shape re-base is pending: advance it with dcp_substep");
    Plan p;
    if (S < 1) return set_err("n_substeps must be >= 1");
    p.S = S; p.K = S; p.L = -1; p.nseg = 1; p.nck = 1; p.seg0 = 2;
    // forward-only: ping-pong between slots 0 and 1 would lose the reset state;
    // use slots 1 and 2 after the reset slot 0
    if (c->max_states < 3) return set_err("dcp_rollout needs max_states >= 3");
    c->tev_used = 0;
    if (c->reset_pool()) return set_err("pool reset launch failed");
    HIP_OK(hipEventRecord(c->ev0, c->stream));
    int cur = 0;
    for (int f = 0; f < S; ++f) {
        int nxt = (f % 2 == 0) ? 1 : 2;
        if (launch_forward(c, f, g, c->state(cur), c->state(nxt))) return -1;
        cur = nxt;
    }
    (void)use_graph;
    HIP_OK(hipEventRecord(c->ev1, c->stream));
    HIP_OK(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_final_slot = cur;
    if (final_slot) *final_slot = cur;
    return 0;
}

extern "C" int dcp_polar_adjoint(const float* Ft, const float* gR, int32_t n, int32_t precision, int32_t svd_grad,
                                 float* gF, float* R) {
    if (!Ft || !gR || !gF || !R || n < 0 || (precision != 0 && precision != 1) || (svd_grad != 0 && svd_grad != 1))
        return set_err("dcp_polar_adjoint: bad arguments");
    if (n == 0) return 0;
    const size_t bytes = 16 * (size_t)n;
    float* d = nullptr;
    HIP_OK(hipMalloc(&d, 4 * bytes));
    float *dF = d, *dg = d + 4 * (size_t)n, *dgF = d + 8 * (size_t)n, *dR = d + 12 * (size_t)n;
    hipError_t e = hipMemcpy(dF, Ft, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dg, gR, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
        if (precision == 1)
            hipLaunchKernelGGL(k_polar_adj<float>, dim3(blocks), dim3(256), 0, nullptr, dF, dg, n, svd_grad == 0 ? 0 : 1, dgF, dR);
        else
            hipLaunchKernelGGL(k_polar_adj<double>, dim3(blocks), dim3(256), 0, nullptr, dF, dg, n, svd_grad == 0 ? 0 : 1, dgF, dR);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(gF, dgF, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(R, dR, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return set_err(std::string("dcp_polar_adjoint: ") + hipGetErrorString(e));
    return 0;
}

extern "C" int dcp_synchronize(dcp_ctx* c) {
    if (!c) return set_err("bad args");
    HIP_OK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int dcp_set_kernel_timing(dcp_ctx* c, int32_t kernel, int32_t max_launches) {
    if (!c || kernel < 0 || kernel > DCP_KERNEL_VERTEX_FINAL || max_launches < 0) return set_err("bad args");
    HIP_OK(hipStreamSynchronize(c->stream));
    // cached graphs hold the previous event nodes (or none): rebuild them
    c->drop_graphs();
    c->drop_timing_events();
    c->time_kid = kernel;
    if (kernel == 0) return 0;
    c->tev.resize((size_t)2 * max_launches, nullptr);
    for (auto& e : c->tev) HIP_OK(hipEventCreate(&e));
    // a capture rejects an event that was never recorded: record each once
    for (auto& e : c->tev) HIP_OK(hipEventRecord(e, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int dcp_kernel_timing(dcp_ctx* c, double* total_ms, int32_t* launches) {
    if (!c || !total_ms || !launches) return set_err("bad args");
    HIP_OK(hipStreamSynchronize(c->stream));
    double tot = 0.0;
    int n = c->tev_used / 2;
    for (int i = 0; i < n; ++i) {
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, c->tev[2 * i], c->tev[2 * i + 1]));
        tot += ms;
    }
    *total_ms = tot;
    *launches = n;
    return 0;
}

extern "C" int dcp_last_rollout_final_slot(dcp_ctx* c, int32_t* slot) {
    if (!c || !slot) return set_err("bad args");
    if (c->last_final_slot < 0) return set_err("no rollout has run");
    *slot = c->last_final_slot;
    return 0;
}

extern "C" int dcp_last_rollout_ms(dcp_ctx* c, double* ms) {
    if (!c || !ms) return set_err("bad args");
    *ms = c->last_ms;
    return 0;
}
