/*
 * diffcp.h - C-ABI of the MI355X-native differentiable codimensional-MPM
 * cloth substep (forward + adjoint).  Plain pointers and sizes only.
 *
 * The reference has no native ABI: its boundary is the Taichi runtime
 * (module-global fields, kernels called with ti.i32 scalars, the tape).
 * Each entry point below names the reference interface it replaces.
 *
 * Conventions
 *   - every call returns 0 on success and a negative code on error; the
 *     message is in dcp_last_error() (thread-local);
 *   - a context owns all device memory for B cloth instances and runs on one
 *     HIP stream; a context is not thread-safe;
 *   - launches are asynchronous except the calls that return host data
 *     (dcp_download_state, dcp_loss, dcp_param_grads, dcp_status,
 *     dcp_rollout_grad), which synchronise the context stream;
 *   - caller arrays are in the caller's (natural) particle order; the library
 *     reorders particles internally and restores the order on download;
 *   - every array argument may be host memory or device memory of the
 *     context's GPU (e.g. a torch tensor's data_ptr), told apart by HIP's
 *     unified addressing: state-slot arrays in device memory are copied by
 *     device kernels (all arrays of one call in the same kind of memory),
 *     inputs the library preprocesses on the host (parameters, targets, rest
 *     positions, trajectories) are first copied in, and results are written
 *     to whichever memory the pointer names.  Work the caller queued on
 *     another stream into a device array must be complete before the call;
 *     the library has finished with every array when the call returns.
 */
#ifndef DIFFCP_H
#define DIFFCP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dcp_ctx dcp_ctx;

/* number of per-instance scalar parameters: E, nu, shearing_stiffness,
 * contact_stiffness, damping, friction */
#define DCP_N_PARAMS 6

/* status bits reported by dcp_status() */
#define DCP_STATUS_NONFINITE   1u   /* NaN/Inf in a particle velocity or grid value */
#define DCP_STATUS_OUT_OF_GRID 2u   /* a particle stencil left [0, n_grid) (clamped) */
#define DCP_STATUS_TILE_SPILL  4u   /* a patch of more than 256 particles used global atomics */
#define DCP_STATUS_POOL_FULL   8u   /* the grid block pool ran out of slots (results invalid) */
#define DCP_STATUS_INTERNAL   16u   /* an internal consistency check failed (results invalid) */

/* One cloth scene.  Replaces the module globals of
 * src/initialization.py:21-193 and the fields of
 * src/taichi_variable_definition.py:5-52 (sizes, masses, trajectory). */
typedef struct dcp_instance_desc {
    int32_t n_grid;              /* grid side G (dx = 1/G), 4 .. 1024              */
    int32_t n_vertices;          /* cloth vertex particles                         */
    int32_t n_elements;          /* triangles = element (quadrature) particles     */
    int32_t n_cube;              /* rigid cube particles                           */
    int32_t n_grasper;           /* grasper (operator) particles                   */
    int32_t n_grasp_points;      /* trajectory width                               */
    int32_t n_actions;           /* trajectory frames                              */
    int32_t steps_per_frame;     /* operator_update_interval / dt                  */
    int32_t frames_per_window;   /* int(grad_update_interval / interval)           */
    int32_t reserved;
    double dt;
    double gravity;              /* scene_scaling * gravity (applied on y)          */
    double vertex_p_mass;
    double quadrature_p_mass;
    double quadrature_p_vol;
    double cube_p_mass;
    double operator_p_mass;
    const double*  vertex_x;     /* [n_vertices][3]                                */
    const int32_t* triangles;    /* [n_elements][3] vertex indices                 */
    const double*  cube_x;       /* [n_cube][3]                                    */
    const double*  grasper_x;    /* [n_grasper][3]                                 */
    const double*  trajectory;   /* [n_actions][n_grasp_points][3]                 */
} dcp_instance_desc;

/* Context options.  Zero-initialise the whole struct (memset, `= {0}`) and
 * set the fields you need: every field's 0 is its default, including fields
 * a later version appends. */
typedef struct dcp_options {
    int32_t max_states;          /* state slots kept on device (>= 2)              */
    int32_t precision;           /* 0: FP32 bulk + FP64 2x2 polar/SVD-grad; 1: all FP32 */
    int32_t svd_grad;            /* 0: reference clamped SVD backward; 1: exact polar derivative */
    int32_t per_element_moduli;  /* 1: per-triangle (E, shear, contact) arrays     */
    int32_t patch_size;          /* particles per P2G/G2P patch (<= 1024); 0 = default */
    int32_t scatter_wgs;         /* persistent workgroups of the per-patch kernels; 0 = default (16 per CU) */
    int32_t grid_wgs;            /* persistent grid-update workgroups; 0 = default */
    int32_t bin_cells;           /* side of the spatial bins patches are cut from (grid cells, 2..16); 0 = 6, -1 = per-instance from particle density */
    int32_t grid_snapshots;      /* 0: dcp_rollout_grad keeps each substep's grid (and element prep outputs) of a
                                    checkpoint segment for the adjoint instead of recomputing them (memory
                                    permitting); > 0: the same with at most this many 4^3 blocks per substep
                                    (a substep whose grid overflows it is recomputed); -1: always recompute */
    int32_t return_mapping;      /* 0: as the reference, whose element return mapping (simulation.py:408-428)
                                    is computed and discarded; 1: live: F = Q return_map(R) with cf = the
                                    friction parameter, and its adjoint (gradients w.r.t. friction) */
    int32_t grid_pool_blocks;    /* 0: each instance has a dense G^3 grid; > 0: the grids share a pool of this
                                    many 4^3 blocks, handed out to the blocks particle stencils reach (a block
                                    keeps its slot until the next dcp_reset / rollout start); a full pool sets
                                    DCP_STATUS_POOL_FULL */
    int32_t deterministic;       /* 0: the scatters add into the grid with float atomics, whose order varies from
                                    run to run; 1: they add two-word fixed-point sums (and the damping and
                                    trajectory gradients too), so that a window gives the same bits every run */
} dcp_options;

/* Build a context: upload B instances, build the internal layout (spatial
 * particle order, patches, vertex->element CSR) and allocate every buffer.
 * Replaces ti.init + field placement (initialization.py:18-20,
 * taichi_variable_definition.py:5-52).  `stream` is a hipStream_t or NULL. */
int dcp_create(const dcp_instance_desc* instances, int32_t n_instances,
               const dcp_options* options, void* stream, dcp_ctx** out);

/* Free a context and all its device memory. */
int dcp_destroy(dcp_ctx* ctx);

/* Thread-local message of the last failing call. */
const char* dcp_last_error(void);

/* Per-instance parameters [B][DCP_N_PARAMS].  Replaces the scalar fields
 * garment_E, garment_nu, shearing_stiffness, contact_stiffness
 * (taichi_variable_definition.py:5-8) and adds damping and friction. */
int dcp_set_params(dcp_ctx* ctx, const double* params);

/* Per-element moduli [sum(n_elements)][3] = (E, shear, contact), natural
 * order; only with options.per_element_moduli = 1. */
int dcp_set_element_moduli(dcp_ctx* ctx, const double* moduli);

/* Loss targets: concatenated [sum(counts)][3], counts[B].  Replaces
 * target_pos.from_numpy (parameter_fitting_functions.py:22). */
int dcp_set_targets(dcp_ctx* ctx, const double* targets, const int32_t* counts);

/* Initial state into `slot` (F = I, v = C = 0, rest shape).  Replaces
 * initialize_objects / initialize_ti_field (simulation.py:432-477).
 * `vertex_x` may be NULL (use the creation positions) or [sum(n_vertices)][3]
 * to re-base the rest shape (window carry-over, parameter_fitting_functions.py:40-49). */
int dcp_reset(dcp_ctx* ctx, int32_t slot, const double* vertex_x,
              const double* grasper_x, const double* cube_x);

/* Re-derive every element's rest shape from the vertex positions in `slot`
 * (centroid, d = [x1-x0 | x2-x0 | unit normal], Dinv = d^-1; the slot's normal
 * column becomes the unit normal), keeping all velocities and affine
 * matrices.  As in the reference, the next forward substep still computes its
 * stress from each element's F under the old rest shape.  The rest shape is
 * context-wide until the next dcp_reset.  Forward-only: dcp_rollout and
 * dcp_rollout_grad refuse to start while such a substep is pending.
 * Replaces reset_quadrature_particles (simulation_without_grad.py:472-486). */
int dcp_rebase_elements(dcp_ctx* ctx, int32_t slot);

/* One forward substep: state s in slot_in -> state s+1 in slot_out, window g.
 * Replaces forward(s, grad_update) (simulation.py:480-488). */
int dcp_substep(dcp_ctx* ctx, int32_t s, int32_t g, int32_t slot_in, int32_t slot_out);

/* Adjoint of substep s: recomputes the substep's grid from slot_in, then maps
 * the state gradient in gslot_in (d loss / d state s+1) to gslot_out
 * (d loss / d state s) and accumulates parameter gradients.
 * Replaces backward(s, grad_update) (simulation.py:490-506). */
int dcp_substep_adjoint(dcp_ctx* ctx, int32_t s, int32_t g, int32_t slot_in,
                        int32_t gslot_in, int32_t gslot_out);

/* Zero an adjoint slot. */
int dcp_zero_adjoint(dcp_ctx* ctx, int32_t gslot);

/* Zero the accumulated parameter gradients (the .grad of the scalar fields). */
int dcp_zero_param_grads(dcp_ctx* ctx);

/* Point-cloud loss on the vertices of `slot` per instance (FP64 sum of FP32
 * minima), out[B].  Replaces compute_loss(f) + clear_distances
 * (parameter_fitting_functions.py:3-15, simulation.py:3-7). */
int dcp_loss(dcp_ctx* ctx, int32_t slot, double* out);

/* Adjoint of the loss: adds seed[b] * d loss_b / d vertex_x into gslot. */
int dcp_loss_adjoint(dcp_ctx* ctx, int32_t slot, int32_t gslot, const double* seed);

/* Accumulated parameter gradients [B][DCP_N_PARAMS] (FP64).  Replaces
 * garment_E.grad[None] etc. (parameter_fitting_functions.py:32-37). */
int dcp_param_grads(dcp_ctx* ctx, double* out);

/* Accumulated per-element moduli gradients [sum(n_elements)][3], natural order. */
int dcp_element_moduli_grads(dcp_ctx* ctx, double* out);

/* Replace the grasp trajectories [sum(n_actions*n_grasp_points)][3] (same
 * layout as dcp_instance_desc.trajectory, instances concatenated).  Replaces
 * grasp_tracking_seq.from_numpy (simulation.py:434). */
int dcp_set_trajectory(dcp_ctx* ctx, const double* trajectory);

/* Accumulated trajectory gradient [sum(n_actions*n_grasp_points)][3]
 * (grasp_tracking_seq.grad, gradient_descent_for_traj_generation.py:24). */
int dcp_trajectory_grads(dcp_ctx* ctx, double* out);

/* Copy a state slot to host in natural order.  Each pointer may be NULL.
 * Points: [n][3] x, [n][3] v, [n][9] C (row-major) for vertices, grasper and
 * cube; elements: [n][9] C and [n][3] the normal column of d at the previous
 * substep.  Replaces field.to_numpy() (parameter_fitting_functions.py:40-43). */
int dcp_download_state(dcp_ctx* ctx, int32_t slot, float* vertex_x, float* vertex_v,
                       float* vertex_C, float* grasper_x, float* grasper_v, float* grasper_C,
                       float* cube_x, float* cube_v, float* cube_C, float* element_C,
                       float* element_n);

/* Overwrite a state slot from host arrays (same layout; NULL = zero).
 * Replaces field.from_numpy() of the state fields. */
int dcp_upload_state(dcp_ctx* ctx, int32_t slot, const float* vertex_x, const float* vertex_v,
                     const float* vertex_C, const float* grasper_x, const float* grasper_v,
                     const float* grasper_C, const float* cube_x, const float* cube_v,
                     const float* cube_C, const float* element_C, const float* element_n);

/* Copy an adjoint slot to host (same layout as dcp_download_state). */
int dcp_download_adjoint(dcp_ctx* ctx, int32_t gslot, float* vertex_x, float* vertex_v,
                         float* vertex_C, float* grasper_x, float* grasper_v, float* grasper_C,
                         float* cube_x, float* cube_v, float* cube_C, float* element_C,
                         float* element_n);

/* Overwrite an adjoint slot from host arrays (NULL = zero). */
int dcp_upload_adjoint(dcp_ctx* ctx, int32_t gslot, const float* vertex_x,
                       const float* vertex_v, const float* vertex_C, const float* grasper_x,
                       const float* grasper_v, const float* grasper_C, const float* cube_x,
                       const float* cube_v, const float* cube_C, const float* element_C,
                       const float* element_n);

/* Integer binning of a slot: base cell trunc(x*inv_dx - 0.5) of every point
 * and element particle, natural order, [n][3] int32 (points then elements per
 * instance, instances concatenated).  Used by the bit-exact binning tests. */
int dcp_binning_keys(dcp_ctx* ctx, int32_t slot, int32_t* keys_out);

/* The scatters' per-patch binning of a slot, for the permutation tests:
 * out[NP + NE][9] (every point and element particle once, scatter-patch
 * order) = (patch, lane, instance, class 0..3 = vertex/grasper/cube/element,
 * natural index within the class and instance, base cell b0, b1, b2, sorted
 * position in the patch or -1 when the patch's tile exceeds the LDS cap).
 * The position is a stable counting sort of the base cells in lane order.
 * Replaces nothing in the reference, which never sorts: its base cells are
 * the keys of simulation.py:140,155,177,194. */
int dcp_patch_order(dcp_ctx* ctx, int32_t slot, int32_t* out);

/* Block pool of the grids (dcp_options.grid_pool_blocks): slots handed out
 * since the last reset (used) and available (capacity); 0 and 0 for dense
 * grids.  No counterpart in the reference, whose grids are dense fields
 * (taichi_variable_definition.py:33-36).  Synchronises the context stream. */
int dcp_pool_blocks(dcp_ctx* ctx, int32_t* used, int32_t* capacity);

/* Sticky status bits (DCP_STATUS_*); `clear` resets them. */
int dcp_status(dcp_ctx* ctx, uint32_t* out, int32_t clear);

/* Whole forward + adjoint rollout on device: n_substeps forward substeps of
 * window g from the reset state, loss on state n_substeps-2, reverse pass
 * with checkpoints every `checkpoint_every` substeps (segment recompute),
 * each segment replayed from a captured hipGraph when use_graph = 1.
 * loss_out[B], grads_out[B][DCP_N_PARAMS] (may be NULL).  Replaces
 * compute_grad's tape window (parameter_fitting_functions.py:20-38). */
int dcp_rollout_grad(dcp_ctx* ctx, int32_t n_substeps, int32_t g, int32_t checkpoint_every,
                     int32_t use_graph, double* loss_out, double* grads_out);

/* Forward-only rollout from the reset state into the last slot used; returns
 * the slot holding state n_substeps in *final_slot. */
int dcp_rollout(dcp_ctx* ctx, int32_t n_substeps, int32_t g, int32_t use_graph,
                int32_t* final_slot);

/* Device time of the last dcp_rollout_grad / dcp_rollout call in ms. */
int dcp_last_rollout_ms(dcp_ctx* ctx, double* ms);

/* State slot holding the final state (n_substeps) of the last
 * dcp_rollout_grad / dcp_rollout call: what vertex_x.to_numpy()[max_steps-1]
 * reads for the window carry-over (parameter_fitting_functions.py:40-49). */
int dcp_last_rollout_final_slot(dcp_ctx* ctx, int32_t* slot);

/* The 2x2 in-plane polar rotation and its adjoint on host arrays, without a
 * context: for each of n matrices Ft [n][2][2] (row-major) R = U V^T of
 * its SVD and gF = d<gR, R>/dFt for the given gR [n][2][2], with the
 * reference's clamped SVD backward (svd_grad 0: exactly zero at
 * sigma1 = sigma2) or the exact polar derivative (svd_grad 1), computed in
 * FP64 (precision 0) or FP32 (precision 1) from the FP32 inputs.  Replaces
 * inplane_svd + backward_inplane_svd / inplane_svd_grad
 * (simulation.py:39-42, 44-94) for one batch. */
int dcp_polar_adjoint(const float* Ft, const float* gR, int32_t n, int32_t precision, int32_t svd_grad,
                      float* gF, float* R);

/* dcp_polar_adjoint without any device allocation: no hipMalloc, hipFree or
 * allocating library call.  No counterpart in the reference; it is for
 * callers that cannot allocate (tight serving loops, harnesses that forbid
 * allocation).  The device staging of inputs and outputs is a static array
 * of at most 16384 matrices (1 MiB, held by each GPU the library runs on
 * until the process ends); larger n is rejected with an error.  All calls on
 * one GPU share that array: calls that overlap on one GPU race. */
int dcp_polar_adjoint_noalloc(const float* Ft, const float* gR, int32_t n, int32_t precision, int32_t svd_grad,
                              float* gF, float* R);

/* Block until all work queued on the context's stream has finished. */
int dcp_synchronize(dcp_ctx* ctx);

/* Kernel kinds for dcp_set_kernel_timing. */
#define DCP_KERNEL_NONE         0
#define DCP_KERNEL_PREP         1
#define DCP_KERNEL_P2G          2
#define DCP_KERNEL_GRIDOP       3
#define DCP_KERNEL_G2P          4
#define DCP_KERNEL_G2P_ADJ      5
#define DCP_KERNEL_GRIDOP_ADJ   6
#define DCP_KERNEL_P2G_ADJ      7   /* P2G adjoint of the vertex, grasper and cube patches */
#define DCP_KERNEL_PREP_ADJ     8   /* element patches: P2G adjoint + element prep adjoint */
#define DCP_KERNEL_VERTEX_FINAL 9

/* Record a pair of HIP events on the context's stream around each launch of
 * one kernel kind, for at most max_launches launches (DCP_KERNEL_NONE turns it
 * off).  Cached rollout graphs are dropped and recaptured with the events as
 * event-record nodes.  Profiling aid with no counterpart in the reference. */
int dcp_set_kernel_timing(dcp_ctx* ctx, int32_t kernel, int32_t max_launches);

/* Sum of the recorded launch durations (ms) and their number, for the most
 * recent eager rollout or, with graphs, the most recent replay of each graph. */
int dcp_kernel_timing(dcp_ctx* ctx, double* total_ms, int32_t* launches);

/* Sizes: [B, total points, total elements, slot floats, max_states]. */
int dcp_sizes(dcp_ctx* ctx, int64_t* out5);

/* Device memory the context holds (bytes): every buffer it allocated,
 * including the grid and element-prep snapshots of dcp_rollout_grad.  No
 * counterpart in the reference (Taichi's device_memory_GB pool,
 * initialization.py:19, is a fixed reservation). */
int dcp_device_bytes(dcp_ctx* ctx, int64_t* out);

#ifdef __cplusplus
}
#endif
#endif /* DIFFCP_H */
