/* fvk.h — C ABI of the MI355X explicit unstructured FV residual-and-update engine.
 *
 * The library replaces the reference's solver library seam: the free functions
 * of the src/include headers that src/base/main.cpp calls every iteration
 * (ssp_rk3 -> res -> compute_gradient / squeeze / reconstruction / rusanov),
 * together with the per-run set-up that main.cpp does around them.  Plain C
 * types only; all calls are synchronous with respect to host buffers; one
 * context per host thread; the device memory belongs to the context.
 *
 * Numbering: every array crossing this boundary is in the caller's (grid
 * file) numbering.  The library renumbers cells internally.
 *
 * Errors: every call returns FVK_OK (0) or a negative FVK_ERR_* code and never
 * exits; fvk_last_error() gives the message.
 */
#ifndef FVK_H
#define FVK_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FVK_OK 0
#define FVK_ERR_ARG (-1)          /* NULL pointer, bad size */
#define FVK_ERR_MESH (-2)         /* index out of range, degenerate cell */
#define FVK_ERR_OPTION (-3)       /* option value the reference does not define */
#define FVK_ERR_HIP (-4)          /* HIP runtime error (no device, launch failure) */
#define FVK_ERR_NONFINITE (-5)    /* non-finite residual norm */
#define FVK_ERR_UNSUPPORTED (-6)  /* reserved: no option returns it (ROE, which the
                                     reference declares but never ran, is an extension) */

/* BC kinds, read_inputs.cpp:171-205 (struct_BC values) */
#define FVK_BC_FREESTREAM (-1)
#define FVK_BC_SLIP_WALL (-2)
#define FVK_BC_EXTRAPOLATION (-3)
#define FVK_BC_NO_SLIP_WALL (-4)

/* precision forms */
#define FVK_FP64 0         /* plain FP64: reproduces the reference */
#define FVK_FP32_MIXED 1   /* state stored as the FP32 perturbation Q - Q_inf; face
                              states, fluxes and the update in FP32 perturbation
                              form; geometry, LSQ inverses and the dt maximum FP64.
                              No counterpart in the reference (it is FP64 only). */

typedef struct fvk_ctx fvk_ctx;

/* The mesh as src/base/readmesh.cpp:31-85 reads it (class_mesh.hpp:7-14),
 * 0-based.  face_cell2 < 0 is the file marker minus one, before the BC kind is
 * substituted: -1 = BC1 group, -2 = BC2 group, -3 = EXTRAPOLATION and
 * -4 = NO_SLIP_WALL (readmesh.cpp:59-69).  Cells may have any number (>= 3) of
 * faces. */
typedef struct fvk_mesh {
  int32_t num_points;
  int32_t num_faces;
  int32_t num_cells;
  int32_t reserved_;
  const double* x;
  const double* y;
  const int32_t* face_point1;
  const int32_t* face_point2;
  const int32_t* face_cell1;
  const int32_t* face_cell2;
} fvk_mesh;

/* The input.in options (struct_inputs.hpp, struct_time.hpp, struct_BC.hpp,
 * class_flow.hpp) with read_inputs.cpp's enum codes. */
typedef struct fvk_options {
  int32_t order;           /* solution_poly_order + 1: 1 or 2 */
  int32_t limiter;         /* 0 NONE, 1 MINMOD (a no-op, res.cpp:36-38), 2 SQUEEZE;
                              3 Barth-Jespersen, an extension (the reference's minmod.cpp
                              is a draft it never builds): the multidimensional
                              minmod-type limiter on the current reconstruction */
  int32_t eqn;             /* 1 EULER, 2 NS */
  int32_t flux_solver;     /* 1 RUSANOV; 2 ROE, an extension: the reference reads it
                              (read_inputs.cpp:89-91) but its Roe flux is an unfinished
                              draft (res.cpp:66-68 leaves the flux at -101); here Roe's
                              flux with Harten's entropy fix, in the LDS-staged gather
                              pass B only (kernel_variant 0, 3 or 4, accumulate 0) */
  int32_t time_scheme;     /* 1 SSP_RK2 (runs SSP-RK3, main.cpp:123-126), 2 SSP_RK3;
                              3 two-stage SSP-RK2, an extension (the reference's
                              ssp_rk2.cpp is a draft it never builds): Q* = Q + dt R(Q),
                              Q_n+1 = Q/2 + (Q* + dt R(Q*))/2 */
  int32_t time_step_type;  /* 1 GLOBAL, 2 LOCAL; 3 UNIFORM, an extension: one dt for
                              every cell, 2*CFL / max_i(Sigma_i / V_i), so that a run
                              is time-accurate (GLOBAL's dt_i = 2*CFL*V_i / max Sigma,
                              ssp_rk3.cpp:38-42, differs per cell) */
  int32_t bc1;             /* kind of the BC1 group (FVK_BC_*) */
  int32_t bc2;             /* kind of the BC2 group (FVK_BC_*) */
  double cfl;
  double gamma, R, mu, Pr; /* [fluid] */
  double P, rho, u, v;     /* [init] freestream */
  int32_t precision;       /* FVK_FP64 (default) or FVK_FP32_MIXED */
  int32_t renumber;        /* cell blocks: 1 recursive coordinate bisection (default),
                              2 reverse Cuthill-McKee ranges, 0 file-order ranges */
  int32_t block_size;      /* cells per block: 64, 128, 192 or 256; 0 (default) =
                              128 for FP64 or above 2M cells in the solve, else 256 */
  int32_t xcd_remap;       /* XCD block order: 1 contiguous block ranges per XCD;
                              G in 2..4096 runs of G consecutive blocks dealt to
                              the XCDs in turn; 0 off; -1 (default) = 1 up to 4096
                              blocks, else 8 */
  int32_t kernel_variant;  /* 0: default (3 when the blocks fit in LDS, else 1);
                              1: one thread per cell; 2: one thread per face slot;
                              3: LDS-staged blocks, one workgroup per block;
                              4: the same blocks on a persistent grid (as many
                              workgroups as are resident at once, each walking a
                              stride of its XCD's blocks; results identical to 3) */
  int32_t accumulate;      /* face-to-cell residual accumulation: 0 (default) owner
                              gather in face order (bit-exact); 1 atomic scatter of
                              each face's terms into the block's LDS accumulators
                              (LDS-staged shapes only; sums in run-time order, so
                              not bit-exact) */
} fvk_options;

/* Fill *opt with the defaults above for the given BC kinds. */
void fvk_default_options(fvk_options* opt);

/* Set-up: geometry of readmesh.cpp:113-411 (bit-identical), device layout and
 * memory.  Replaces readmesh's geometry half and main.cpp:75-80. */
int fvk_create(const fvk_mesh* mesh, const fvk_options* opt, fvk_ctx** out);
int fvk_destroy(fvk_ctx* ctx);
/* Message of the last failing call on ctx (or of the last fvk_create when
 * ctx is NULL). */
const char* fvk_last_error(const fvk_ctx* ctx);

/* Uniform freestream start: Qbar.init + updateQ and the face-state
 * initialisation of main.cpp:107-118 (limiter sees the -101 sentinel, the
 * first time step uses |V_inf| and c_inf on every face). */
int fvk_init_freestream(fvk_ctx* ctx);
/* Restart start, main.cpp:82-106 and 114-118: conserved state q[4*N_c]
 * (p1 block, p2 block, p3 block, p4 block), primitives recomputed, the face
 * states re-initialised as on a fresh start. */
int fvk_set_state(fvk_ctx* ctx, const double* q);

/* nsteps calls of ssp_rk3 (ssp_rk3.cpp:18-77).  The stage-3 residual of the
 * last step is kept for fvk_get_residual / fvk_residual_l2.  Steps that are
 * not the first after a start, begin on an even step count and leave a step
 * for the end of the call are replayed from a captured HIP graph of G steps
 * (G = 2 from 262,144 padded cells up, 8 from 16,384, else 32); the first
 * call long enough to use it captures and instantiates it, later calls (also
 * after fvk_init_freestream, fvk_set_state or a checkpoint load) replay it. */
int fvk_step(fvk_ctx* ctx, int32_t nsteps);
/* 1 (default): replay the step graph as above; 0: launch every step's
 * kernels one by one (the graph is released).  Both give the same bits.
 * No counterpart in the reference. */
int fvk_use_graph(fvk_ctx* ctx, int32_t on);

/* Conserved state q[4*N_c]; grad[8*N_c] (nullable) = the stage-1 limited
 * gradients Qxp1..4, Qyp1..4 that Qbar holds after ssp_rk3 and that
 * save_sol.cpp:34-37,123-142 reads (zeros at first order).  FVK_FP32_MIXED
 * returns Q_inf + (double)dQ and the FP32 gradients widened. */
int fvk_get_state(fvk_ctx* ctx, double* q, double* grad);
/* Stage-3 residual r[4*N_c] of the last step (class_residual after ssp_rk3). */
int fvk_get_residual(fvk_ctx* ctx, double* r);
/* sqrt(sum_i R_k,i^2), k = 1..4, of that residual (save_sol.cpp:79-89);
 * FVK_ERR_NONFINITE if any is not finite (out[] is still written). */
int fvk_residual_l2(fvk_ctx* ctx, double out[4]);
/* fvk_residual_l2 without any allocation, for callers that cannot allocate
 * (serving loops and harnesses that forbid hipMalloc and host-heap use on
 * their path).  No counterpart in the reference.  The same norms, bit for
 * bit: the per-block partial sums and their compensated sum go through a
 * static device array of 2 MiB per GPU (held until the process ends) instead
 * of the context's buffers.  Meshes of at most 65,536 blocks
 * (8,388,608 cells at 128-cell blocks); larger ones get FVK_ERR_ARG.  The
 * context's device must be current.  Every call on a GPU shares that GPU's
 * array, so calls that overlap on one GPU (from several host threads) race. */
int fvk_residual_l2_noalloc(fvk_ctx* ctx, double out[4]);

/* Global conservation sums of the current state: out[k] = sum_i V_i Q_k,i,
 * k = 1..4 (mass, x and y momentum, energy), with compensated summation
 * (per-block TwoSum pairs on the device, a Neumaier sum of the pairs on the
 * host; the FP32-mixed form adds Q_inf,k * sum(V) to sum(V dQ_k)).  No
 * counterpart in the reference, which checks no conservation. */
int fvk_conservation_sums(fvk_ctx* ctx, double out[4]);

/* Geometry arrays of class_mesh, original numbering, for checks:
 * "face_area", "face_nx", "face_ny", "face_centerx", "face_centery" (N_f);
 * "cell_vol", "cell_centerx", "cell_centery", "Ixx", "Iyy", "Ixy" (N_c). */
int fvk_get_geometry(fvk_ctx* ctx, const char* name, double* out);

/* Host-only (no device needed): the same geometry straight from a mesh, i.e.
 * the geometry half of readmesh.cpp:113-320. */
int fvk_geometry(const fvk_mesh* mesh, const fvk_options* opt, const char* name, double* out);

/* Sizes: [num_cells, num_faces, num_boundary_faces, num_slots, max_faces_per_cell,
 * num_wall_faces (readmesh.cpp:356-360)]. */
int fvk_get_info(fvk_ctx* ctx, int64_t info[6]);

/* Physical time reached (sum of the steps' dt) with time_step_type 3, reset by
 * fvk_init_freestream and fvk_set_state; FVK_ERR_OPTION for GLOBAL or LOCAL,
 * whose steps differ per cell.  No counterpart in the reference. */
int fvk_get_time(fvk_ctx* ctx, double* t);

/* The options in force, with the automatic choices (block_size 0,
 * xcd_remap -1) resolved.  No counterpart in the reference, whose options
 * are fixed by input.in. */
int fvk_get_options(fvk_ctx* ctx, fvk_options* out);

/* Timing: run nsteps with a HIP event pair around every launch on the
 * context's stream.  out_ms[0] = wall time of the nsteps (event-timed),
 * out_ms[1..] = per kernel kind total time: [1] pass A (gradient+limiter),
 * [2] pass B (flux+residual+update), [3] other.  counts[0..3] launches. */
int fvk_time_steps(fvk_ctx* ctx, int32_t nsteps, double out_ms[4], int64_t counts[4]);

/* Checkpoints (an extension: the reference's restart file, post_process.cpp:13-28,
 * holds 6-digit conserved variables and a restart re-arms the lagged limiter
 * and wave speeds, main.cpp:82-118).  A checkpoint holds the complete device
 * state -- the three stage states, the three limited gradients, the spectral
 * sums and dt maxima, the residual and the lagged-data flags -- so that
 * loading it into a context created from the same mesh and options and
 * stepping on gives exactly the results of never having stopped.  The buffer
 * is opaque host memory of fvk_checkpoint_size bytes (the device layout,
 * so it is only valid for the same mesh, options and library build). */
int fvk_checkpoint_size(fvk_ctx* ctx, int64_t* bytes);
int fvk_checkpoint_save(fvk_ctx* ctx, void* buf, int64_t bytes);
int fvk_checkpoint_load(fvk_ctx* ctx, const void* buf, int64_t bytes);

/* Per-kernel timing: `reps` back-to-back launches of the stage-2 pass A
 * kernel (gradient + limiter) and of the stage-2 pass B kernel (flux +
 * residual + update), each between one HIP event pair on the context's
 * stream.  Both read Q1/Qn/G1 and write G2/Q2 only, so repeating them leaves
 * the solver state as it was.  avg_us[0] = pass A (0 at first order),
 * avg_us[1] = pass B, in microseconds per launch. */
int fvk_time_kernels(fvk_ctx* ctx, int32_t reps, double avg_us[2]);

/* ------------------------------------------------------------------------
 * Batches: many independent instances (own mesh, freestream, CFL, BC kinds,
 * limiter switch) advanced together, one launch per kernel and stage for the
 * whole batch.  No counterpart in the reference, which solves one case per
 * process; each instance's results are those of a single fvk_ctx run of the
 * same mesh and options, bit for bit.  Instances must agree on order, eqn,
 * time_step_type, precision, block_size, kernel_variant (0, 3 or 4) and
 * xcd_remap.  Steps are replayed from a captured HIP graph as fvk_step
 * describes, counting the cells of all instances; the graph is captured
 * again when the number of active blocks changes (fvk_batch_use_graph).
 * Every call is synchronous.
 * ------------------------------------------------------------------------ */
typedef struct fvk_batch fvk_batch;

int fvk_batch_create(int32_t n, const fvk_mesh* meshes, const fvk_options* opts, fvk_batch** out);
int fvk_batch_destroy(fvk_batch* b);
/* message of the last failing call on b (or of the last fvk_batch_create
 * when b is NULL) */
const char* fvk_batch_last_error(const fvk_batch* b);
/* Start every instance: q == NULL, or q[i] == NULL, for instance i's
 * freestream start of fvk_init_freestream; else q[i] = instance i's
 * conserved state as for fvk_set_state (a state taken from a batch of the
 * other precision included).  All instances become active. */
int fvk_batch_init(fvk_batch* b, const double* const* q);
/* nsteps steps of every active instance */
int fvk_batch_step(fvk_batch* b, int32_t nsteps);
/* active[i] != 0: instance i takes part in the following steps; a retired
 * instance keeps its state and residual */
int fvk_batch_set_active(fvk_batch* b, const uint8_t* active);
/* out[4*i + k] = sqrt(sum R_k^2) of instance i's last stage-3 residual */
int fvk_batch_residual_l2(fvk_batch* b, double* out);
/* out[4*i + k] = instance i's global conservation sums, as fvk_conservation_sums */
int fvk_batch_conservation_sums(fvk_batch* b, double* out);
/* Step the active instances, checking every check_every steps (and after
 * the first step, whose density residual norm is instance i's reference
 * res0[i]), and retire each instance whose density residual norm has fallen
 * to drop * res0[i] or is not finite, or (fvk_batch_set_stall) has stalled;
 * stop when none is active or after max_steps steps.  steps[i] = steps
 * instance i ran; res_end[i] = its last checked density residual norm (res0,
 * res_end nullable). */
int fvk_batch_run_to_drop(fvk_batch* b, double drop, int32_t max_steps, int32_t check_every,
                          int32_t* steps, double* res0, double* res_end);
/* Stall retirement in fvk_batch_run_to_drop (off by default, checks = 0): an
 * instance also retires after `checks` checks in a row in which its density
 * residual norm has not fallen below factor (0 < factor < 1) times the lowest
 * norm that counted (res0, then each norm that fell below factor times the
 * previous one).  A run stalls where its residual stops falling: the SQUEEZE
 * limiter's cycling, or the FP32-mixed state's resolution (a converged FP32
 * perturbation state carries a residual of about its rounding).  No
 * counterpart in the reference, which runs a fixed nmax. */
int fvk_batch_set_stall(fvk_batch* b, int32_t checks, double factor);
/* Reference norms for the following fvk_batch_run_to_drop calls: res0[i] > 0
 * replaces instance i's first-step norm as the norm its drop is measured
 * from (and is returned as its res0); res0 == NULL, or res0[i] <= 0, goes
 * back to the first-step norm.  This lets a run that continues another
 * one, such as an FP64 batch finishing the instances an FP32-mixed batch
 * retired by a stall (their states passed to fvk_batch_init), keep the
 * first run's target.  No counterpart in the reference. */
int fvk_batch_set_reference(fvk_batch* b, const double* res0);
/* why[i] = why instance i is not active: one of FVK_RETIRED_* */
#define FVK_RETIRED_NONE 0       /* active */
#define FVK_RETIRED_DROP 1       /* residual norm fell to the drop */
#define FVK_RETIRED_NONFINITE 2  /* residual norm not finite */
#define FVK_RETIRED_STALL 3      /* residual norm stalled (fvk_batch_set_stall) */
#define FVK_RETIRED_CALLER 4     /* fvk_batch_set_active */
int fvk_batch_retired(fvk_batch* b, int32_t* why);
/* instance i's state and stage-1 limited gradients, as fvk_get_state */
int fvk_batch_get_state(fvk_batch* b, int32_t i, double* q, double* grad);
/* [instances, cells, faces, mesh blocks, active blocks, active instances] */
int fvk_batch_get_info(fvk_batch* b, int64_t info[6]);
/* t[i] = instance i's physical time, as fvk_get_time */
int fvk_batch_get_time(fvk_batch* b, double* t);
/* instance i's options in force, as fvk_get_options */
int fvk_batch_get_options(fvk_batch* b, int32_t i, fvk_options* out);
/* 0: launch every step's kernels one by one; 1 (default): replay the graph */
int fvk_batch_use_graph(fvk_batch* b, int32_t on);
/* checkpoints of a whole batch (with the active set and step counters); a
 * checkpoint records a hash of every instance's flow options (freestream, gas
 * constants, CFL, scheme switches) and is refused by a batch, or context,
 * whose options differ */
int fvk_batch_checkpoint_size(fvk_batch* b, int64_t* bytes);
int fvk_batch_checkpoint_save(fvk_batch* b, void* buf, int64_t bytes);
int fvk_batch_checkpoint_load(fvk_batch* b, const void* buf, int64_t bytes);
/* fvk_time_kernels over the active blocks of the whole batch */
int fvk_batch_time_kernels(fvk_batch* b, int32_t reps, double avg_us[2]);

#ifdef __cplusplus
}
#endif
#endif
