/* mfcfd.h -- C ABI of the MI355X (gfx950) LSKUM fixed-point iteration.
 *
 * The reference (Nischay-Pro/mfcfd, src_mpi_primal) has no FFI: its hot path is a set
 * of argument-less Fortran module procedures that mutate module state
 * (data_structure.F90:13-49).  This ABI replaces that module state with an opaque
 * context and exports one entry point per reference procedure, plus a fused,
 * graph-captured driver for the whole q_lskum loop.  A Fortran host binds it through
 * ISO_C_BINDING (interfaces in mfcfd_amd/fortran/mfcfd_api.f90).
 *
 * Conventions
 *  - Every function returns MF_OK (0) or a negative error code; mf_last_error()
 *    returns the message of the last failure on the calling thread.
 *  - Host arrays are owned by the caller and use the reference's layouts and its
 *    original point ordering: prim (4,N), dq/qm (2,4,N) column-major, 1-based
 *    neighbour indices in conn(N,W).
 *  - The context owns all device memory and one HIP stream; a context is not
 *    thread-safe.  Only one GPU is used (the current HIP device at mf_create).
 */
#ifndef MFCFD_H
#define MFCFD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MF_OK 0
#define MF_ERR_HIP (-1)      /* a HIP runtime call failed                        */
#define MF_ERR_INVALID (-2)  /* invalid argument or cloud                         */
#define MF_ERR_STATE (-3)    /* call not valid in the context's current state     */

#define MF_FP64 0   /* every step in FP64, as the reference                       */
#define MF_MIXED 1  /* FP32 state, flux, limiter, update; FP64 LSQ, norms, forces */

typedef struct mf_ctx mf_ctx;

/* One point cloud, as read by point_preprocessor.F90:177-546 (single rank). */
typedef struct mf_geometry {
  int32_t n_points;        /* max_points = local_points                              */
  int32_t max_nbhs;        /* W, leading dimension of conn (reference: 20)           */
  const double *x, *y;     /* [N] coordinates                                        */
  const double *nx, *ny;   /* [N] file normals, read only when format == 2; may be NULL */
  const double *min_dist;  /* [N]                                                     */
  const int32_t *left;     /* [N] 1-based left neighbour of boundary points (else 0)  */
  const int32_t *right;    /* [N] 1-based right neighbour                             */
  const int32_t *flag_1;   /* [N] 0 wall, 1 interior, 2 outer                         */
  const int32_t *flag_2;   /* [N] shape id (1..shapes) of wall points, else 0         */
  const int32_t *nbhs;     /* [N] neighbour counts                                    */
  const int32_t *conn;     /* conn(N,W) column-major, 1-based: conn[(k-1)*N + (i-1)] */
} mf_geometry;

/* Run options (readcase.F90:10-125 names and defaults). */
typedef struct mf_params {
  double cfl;              /* -cfl                                                   */
  double vl_const;         /* -vl_const (Venkatakrishnan K), default 150             */
  double power;            /* -power, LSQ weights dist**power, default 0             */
  double mach;             /* Mach (parameter.F90:5)                                 */
  double aoa;              /* angle of attack, degrees (parameter.F90:6)             */
  double euler;            /* 1 for ssprk43, 2 for first (readcase.F90:50-60)        */
  double res_old;          /* residual reference when restart == 1                   */
  int32_t rks;             /* RK stages: 4 (ssprk43) or 1 (first)                    */
  int32_t inner_iterations;/* -inner_iterations                                      */
  int32_t shapes;          /* -shapes                                                */
  int32_t interior_points_normal_flag; /* 0: (0,1), 1: (1,0)                         */
  int32_t format;          /* 1 legacy, 2 quadtree (file normals for interior points) */
  int32_t restart;         /* 0: res_old taken at iterations 1-2; 1: given           */
  int32_t precision;       /* MF_FP64 or MF_MIXED                                    */
  int32_t reorder;         /* 1: Hilbert-curve point order on the device (default)   */
  int32_t block_size;      /* threads per block; 0 = built-in choice                 */
  int32_t first_order;     /* 1: -solution_accuracy first: q-tilde = q at both ends of
                              every edge (no reconstruction, no limiter); 0: second order
                              (default).  readcase.F90:107-115 parses the option but the
                              reference's primal kernels never read it; this is the
                              first-order scheme that option names. */
} mf_params;

/* Fills `p` with readcase.F90 defaults (cfl 0, vl_const 150, power 0, ssprk43, ...). */
void mf_default_params(mf_params *p);

/* Builds the context: normals (point_normals.F90:8-105), stencil split
 * (generate_connectivity.F90:8-256), device layout, freestream initial state
 * (initial_conditions.F90:9-24).  Replaces readcase/read_hdf5input_point_data/
 * allocate_soln/initial_conditions/compute_normals/generate_connectivity as the
 * module-state setup done in meshfree_solver.F90:42-70 and q_lskum.F90:33-34. */
int mf_create(const mf_geometry *geom, const mf_params *params, mf_ctx **ctx);
int mf_destroy(mf_ctx *ctx);
const char *mf_last_error(void);

/* point%prim in/out, (4,N) column-major, original order (initial_conditions.F90:36-70
 * restart read; post_processing.F90:11-39 writer). */
int mf_set_prim(mf_ctx *ctx, const double *prim);
int mf_get_prim(mf_ctx *ctx, double *prim);

/* Fields of the module state, original order, reference layouts. */
#define MF_FIELD_Q 1         /* q(4,N)                                */
#define MF_FIELD_DQ 2        /* dq(2,4,N); after a first-order run that of the last
                              * iteration's first stage (first order does not use dq,
                              * so mf_run computes it only where the time step needs it) */
#define MF_FIELD_QM 3        /* qm(2,4,N)                             */
#define MF_FIELD_FLUX_RES 4  /* flux_res(4,N) (last cal_flux_residual) */
#define MF_FIELD_DELTA 5     /* delta(N)                              */
#define MF_FIELD_PRIM_OLD 6  /* prim_old(4,N)                         */
#define MF_FIELD_NX 7        /* nx(N)                                 */
#define MF_FIELD_NY 8        /* ny(N)                                 */
int mf_get_field(mf_ctx *ctx, int32_t field, double *out);
/* Accepts MF_FIELD_Q, MF_FIELD_DQ, MF_FIELD_QM, MF_FIELD_DELTA, MF_FIELD_PRIM_OLD. */
int mf_set_field(mf_ctx *ctx, int32_t field, const double *in);

/* One entry point per reference procedure, acting on the context's state. */
int mf_func_delta(mf_ctx *ctx);            /* state_update.F90:190-236   */
int mf_eval_q_variables(mf_ctx *ctx);      /* q_variables.F90:8-37       */
int mf_eval_q_derivatives(mf_ctx *ctx);    /* q_variables.F90:41-112     */
/* eval_q_inner_loop followed by eval_update_innerloop (q_variables.F90:115-227) */
int mf_eval_q_inner_loop(mf_ctx *ctx);
int mf_cal_flux_residual(mf_ctx *ctx);     /* flux_residual.F90:14-73    */
/* state_update.F90:10-141; rk is 1-based; returns sum_res_sqr of this call */
int mf_state_update(mf_ctx *ctx, int32_t rk, double *sum_res_sqr);
/* compute_force_coeffs.F90:11-102 on the current prim: per-shape Cl, Cd, Cm
 * (arrays of `shapes`) and cp of every shape point in original point order. */
int mf_compute_cl_cd_cm(mf_ctx *ctx, double *cl, double *cd, double *cm, double *cp);
int32_t mf_shape_points(mf_ctx *ctx);

/* fpi_solver.F90:13-83 repeated `iters` times from iteration `first_it`, i.e. the
 * q_lskum.F90:59-73 loop, fused and replayed as a HIP graph.  Per iteration it stores
 * residue (fpi_solver.F90:63-70) and res_new, and Cd of shape 1 (the reference's
 * vector_cost_func).  Stops after the first NaN residue, as q_lskum.F90:71 does;
 * *iters_done reports how many ran.  Any history pointer may be NULL. */
int mf_run(mf_ctx *ctx, int32_t first_it, int32_t iters, double *residue_hist,
           double *res_new_hist, double *cd_hist, int32_t *iters_done);
/* Same iterations, queued without any host synchronisation or history copy-out:
 * for timing.  mf_sync waits for the context's stream. */
int mf_run_async(mf_ctx *ctx, int32_t first_it, int32_t iters);
int mf_sync(mf_ctx *ctx);

/* Kernel timing: runs `iters` fused iterations launched one kernel at a time with HIP
 * events around every launch on the context's stream, and returns the mean duration
 * in ms of each kernel kind: [0] lsq, [1] inner, [2] flux+update, [3] finalize,
 * [4] forces.  Count of launches of each kind per iteration in `launches`. */
#define MF_NUM_KERNEL_KINDS 5
int mf_profile_kernels(mf_ctx *ctx, int32_t iters, double *mean_ms, int32_t *launches);
/* Steady launch time: the LSQ, inner-sweep and flux+update kernels at the context's
 * current shapes (mf_set_tuning / mf_autotune), each launched `reps` times back to back
 * between two HIP events on the context's stream, as they run inside an iteration;
 * us[0..2] receives the mean microseconds per launch.  Writes only what an iteration
 * recomputes before reading (dq, qm and the next state buffers), so the state and the
 * next mf_run are unchanged.  Single-case contexts.  No counterpart in the reference:
 * a measurement entry point beside mf_profile_kernels. */
int mf_time_kernels(mf_ctx *ctx, int32_t reps, double *us);

/* A batch: several independent flow cases (each its own cloud, of any size, with its
 * own Mach number and angle of attack; the other options shared) advanced together by
 * every launch of one context.  Per-point arrays of the context (mf_get_prim,
 * mf_set_prim, mf_get_field, dJ/dx and dJ/dy) are the cases' points concatenated in
 * case order; mf_compute_cl_cd_cm returns ncases * shapes entries, case-major, and cp
 * of every case's shape points in that order; the adjoint's J is the sum of every
 * case's Cd (each point's gradient is its own case's dCd/dx, the cases being
 * independent).  No counterpart in the reference, which runs one case per
 * process. */
typedef struct mf_case {
  const mf_geometry *geometry;
  double mach, aoa;  /* parameter.F90:5-6 of this case */
} mf_case;
int mf_create_batch(const mf_case *cases, int32_t ncases, const mf_params *params, mf_ctx **ctx);
/* `iters` iterations of every case (fpi_solver.F90:13-83 per case, residual norms per
 * case; a NaN in one case does not stop the others): residue[it * ncases + c],
 * res_new likewise, cd[(it * ncases + c) * shapes + s].  Any pointer may be NULL. */
int mf_batch_run(mf_ctx *ctx, int32_t first_it, int32_t iters, double *residue, double *res_new, double *cd);
/* Number of cases and the offset of each case's first point (first[ncases] = N). */
int mf_batch_cases(mf_ctx *ctx, int32_t *ncases, int32_t *first);

/* Launch shape of the stencil kernels: lanes per point (1, 2, 4 or 8) of the LSQ
 * derivative kernel (with func_delta), the inner-iteration kernel and the fused flux
 * residual + state update.  A point's neighbours are split over that many lanes of a
 * wavefront and the partial sums combined with shuffles; 1 (the default) keeps the
 * reference's summation order.  Replaces the thread-block shape option of the CUDA
 * build (src_cuda data_structure.F90:104 blockx/blocky/blockz, q_lskum.cuf:51).
 * Staged shapes, MF_LPP_STAGED | lanes: blocks of consecutive points first copy their
 * neighbours' records (each once) into LDS and walk the stencils there; same per-lane
 * summation order, bit-identical results.  LSQ and inner-iteration kernels: 64-point
 * blocks, 1, 2 or 4 lanes; flux + update: 32-point blocks, 2, 4 or 8 lanes.
 * mf_set_tuning rejects a staged shape whose halos do not fit in LDS.
 * Resident inner sweeps, lpp_inner = MF_LPP_RESIDENT | lanes (1, 2 or 4; MF_MIXED
 * contexts): every inner iteration of a stage in one launch of the staged inner kernel's
 * blocks, all resident on the GPU at once, with a grid barrier between sweeps; the
 * staged shape's results bit for bit.  Offered when the occupancy calculator fits every
 * block on the device (about 180,000 points).  A wait at the barrier gives up after
 * barrier_polls polls (0: the default, about a second): the run then stops,
 * mf_run/mf_sync return MF_ERR_HIP and the context falls back to the staged shape with
 * the same lanes.  us_inner is per sweep. */
#define MF_LPP_STAGED 16
#define MF_LPP_RESIDENT 32
/* Adjoint LSQ and inner-iteration reverses, lpp_lsq_rev / lpp_inner_rev = MF_LPP_PULL |
 * lanes (1, 2 or 4): the pull form, one launch per reversed step.  Each point gathers
 * what it receives as a neighbour over the transposed connectivity (from its owners'
 * records and solve adjoints staged in LDS), instead of the owners writing it to an edge
 * buffer that a second kernel gathers; same terms, summed in the push form's orders.
 * The LSQ reverse's qm adjoints (at most 8 per point) reach the selected neighbours by
 * atomic adds, so the gradient's last bits can vary between runs. */
#define MF_LPP_PULL 64
typedef struct mf_tuning {
  int32_t lpp_lsq, lpp_inner, lpp_flux;
  int32_t barrier_polls;
  double us_lsq, us_inner, us_flux;  /* mf_autotune: mean launch time of the chosen shape */
  /* reverse kernels of the adjoint (flux, LSQ, inner iteration, edge gathers): 0 lets
   * the first mf_adjoint_forward time the shapes (the LSQ and inner reverses: their push
   * forms with the gather against their pull forms) and keep the fastest */
  int32_t lpp_flux_rev, lpp_lsq_rev, lpp_inner_rev, lpp_gather;
} mf_tuning;
/* Times every shape on the context's current state (`reps` launches each, default 10),
 * keeps the fastest of each kernel and reports them.  Leaves prim untouched.  Like
 * mf_set_tuning, it ends by capturing and instantiating the HIP graphs that mf_run
 * replays for the chosen shapes (a few ms), so the next run starts without that cost;
 * neither may be called while the context's stream is being captured. */
int mf_autotune(mf_ctx *ctx, int32_t reps, mf_tuning *out);
int mf_get_tuning(mf_ctx *ctx, mf_tuning *out);
int mf_set_tuning(mf_ctx *ctx, const mf_tuning *t);

/* Device layout facts for bandwidth accounting: bytes of each device array class. */
typedef struct mf_layout_info {
  int64_t n_points, n_wall, n_outer, n_interior, n_shape;
  int64_t max_nbhs, total_nbhs;  /* sum of nbhs                                   */
  int64_t state_bytes;           /* bytes per scalar of the state arrays (8 or 4) */
} mf_layout_info;
int mf_layout(mf_ctx *ctx, mf_layout_info *info);

/* Adjoint of the iteration (replaces the Tapenade q_lskum_chkpts_b driver of
 * src_mpi_adjoint, sensitivity_files/q_lskum_chkpts_b.F90:23-300).
 * J = sum over shapes of Cd after `iters` iterations from the context's current state
 * (the reference seeds vector_cost_funcb = 1 per shape).  mf_adjoint_forward runs the
 * primal with a binomial checkpoint schedule of `snaps` prim snapshots in HBM and
 * returns J; mf_adjoint_backward runs the reverse sweep and returns dJ/dx, dJ/dy per
 * point in original order (the reference's pointb%x, pointb%y, written to
 * sensitivity/sensitivity.dat by src_mpi_adjoint/post_processing.F90:41-66).  The
 * context is left holding the final primal state.  Any power (the LSQ weights'
 * derivatives included, q_variables_b.F90).  On an
 * MF_MIXED context the primal sweep and the taped re-runs run the mixed path's FP32
 * kernels; the flux reverse runs in FP32 on their records, as the mixed primal's
 * fluxes do, the other reverse kernels in FP64 on the records widened to FP64, and
 * every adjoint sum is FP64 (the derivative of the iteration along the mixed
 * trajectory).
 * snaps = MF_ADJ_TAPE_ALL tapes every iteration's stage states during the primal
 * sweep instead (about 1.3 kB per point per iteration at 2 inner iterations; refused
 * when that exceeds the free device memory), so the reverse sweep re-runs nothing. */
#define MF_ADJ_TAPE_ALL (-1)
typedef struct mf_adjoint_stats {
  int64_t forward_steps;  /* iterations advanced: the primal sweep plus recomputation */
  int64_t taped_steps;    /* iterations re-run with their stage intermediates taped  */
  int32_t snapshots;      /* prim snapshots held at the peak                         */
  int32_t reserved;
  double primal_ms;       /* the primal sweep (mf_adjoint_forward)                   */
  double adjoint_ms;      /* the reverse sweep (mf_adjoint_backward)                 */
  /* lanes per point of the reverse kernels (flux, LSQ, inner iteration, edge
   * gathers), timed and chosen by the first mf_adjoint_forward on the context */
  int32_t lpp_flux_rev, lpp_lsq_rev, lpp_inner_rev, lpp_gather;
} mf_adjoint_stats;
int mf_adjoint_forward(mf_ctx *ctx, int32_t first_it, int32_t iters, int32_t snaps, double *J);
int mf_adjoint_backward(mf_ctx *ctx, double *dJdx, double *dJdy, mf_adjoint_stats *stats);
int mf_adjoint(mf_ctx *ctx, int32_t first_it, int32_t iters, int32_t snaps, double *J,
               double *dJdx, double *dJdy, mf_adjoint_stats *stats);

/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc): the
 * force coefficients of mf_compute_cl_cd_cm (compute_force_coeffs.F90:11-102) on
 * caller-owned DEVICE arrays, with no context and no allocation.  Inputs in the
 * reference's layouts and original point order: x, y, nx, ny [n_points]; prim (4,N)
 * column-major; left, right [n_points] 0-based; flag_2 [n_points] (shape id 1..shapes
 * of wall points); shape_idx [n_shape], 0-based ids of the shape points in the order
 * cp is wanted.  Outputs: clcdcm[3 * shapes] (Cl, Cd, Cm of each shape) and
 * cp[n_shape].  Work is queued on `stream` (a hipStream_t; NULL: the null stream)
 * and the call returns without waiting.  Partial sums go to a fixed device array of
 * the library (256 blocks x MF_NOALLOC_MAX_SHAPES shapes x 3 doubles per GPU, held
 * until the process ends); calls that overlap on one GPU share it and race.  Rejects
 * shapes above MF_NOALLOC_MAX_SHAPES.  A shape point whose ids are out of range gets
 * cp = NaN and adds nothing. */
#define MF_NOALLOC_MAX_SHAPES 8
int mf_compute_cl_cd_cm_noalloc(int32_t n_points, int32_t n_shape, int32_t shapes, double mach,
                                double aoa, const double *x, const double *y, const double *nx,
                                const double *ny, const double *prim, const int32_t *left,
                                const int32_t *right, const int32_t *flag_2,
                                const int32_t *shape_idx, double *clcdcm, double *cp, void *stream);

/* Tangent (forward) mode of the iteration (replaces the Tapenade forward mode of
 * src_mpi_forward: q_lskum_d.F90:28-60 seeds pointd%x/y, fpi_solver_d.F90:27-85 carries
 * the tangents through each iteration).  From the context's current state, `iters`
 * iterations with the point coordinates moving along (vx, vy) (original order; the
 * reference seeds one coordinate, pointd%x(1) = 1): *J = sum over shapes of Cd after
 * them and *dJ its derivative along (vx, vy); prim_d (4,N), if not NULL, receives the
 * tangent of the final prim.  FP64 dual numbers on the device (a mixed context's state
 * is widened first); any power; single-case contexts.  The context's state is left as
 * it was. */
int mf_tangent(mf_ctx *ctx, int32_t iters, const double *vx, const double *vy, double *J, double *dJ,
               double *prim_d);

/* Measurement aid with no counterpart in the reference: a plain device copy of
 * n_bytes (each byte read once and written once), `reps` times, timed with HIP
 * events; the known-size kernel against which HBM traffic counters are calibrated. */
int mf_copy_probe(int64_t n_bytes, int32_t reps, double *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif /* MFCFD_H */
