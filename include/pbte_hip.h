/*
 * pbte_hip.h -- C-ABI of the MI355X gray linear Callaway DG sweep library
 * (libpbte_hip.so).  Plain pointers and sizes only; device pointers are
 * allocated by the caller (PyTorch-ROCm tensors in the shipped host layer)
 * and are never freed by the library.  A context owns only its private
 * workspaces: element LU factors, the level schedule, compact face data.
 *
 * Every call is stream-ordered on the stream given to pbte_ctx_create.  A
 * context is not thread-safe: one host thread per context.  Status codes are
 * 0 on success, > 0 for a numerical condition the caller maps to an
 * exception (singular element operators, broken cycles) and < 0 for misuse or
 * a HIP error; pbte_last_error() returns the message.
 *
 * Each entry point cites the reference interface it replaces
 * (solver-python/pybte/...).
 */
#ifndef PBTE_HIP_H
#define PBTE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pbte_ctx pbte_ctx;

/* precision selector of the per-(cell, direction) data a sweep touches */
enum { PBTE_FP64 = 0, PBTE_FP32 = 1 };
/* pbte_factorise only: both precisions' factors from one binary64 pass */
enum { PBTE_FP64_FP32 = 2 };

typedef struct pbte_dims {
    int32_t ndir;     /* D = npole * nazim, d = j1 + npole * j2          */
    int32_t n_tris;   /* N                                               */
    int32_t nd;       /* P = ndof_tri, element DOFs (3, 6, 10, 15)       */
    int32_t nf;       /* Q = ndof_fc = deg + 1                            */
    int32_t n_faces;  /* F                                               */
    int32_t nbc;      /* boundary conditions                             */
} pbte_dims;

/* Host arrays read once by pbte_ctx_create (index orders are the
 * reference's, docs/INDEXING.md:41-56; sweep.py:37-72). */
typedef struct pbte_setup {
    const double *cxv, *cyv, *domega;     /* [D]                            */
    double cv, vg, tau_r, tau_n, tau_c, pi;
    const double *mass;                   /* int_tri_tri   [P][P][N]        */
    const double *gx, *gy;                /* int_tri_tri_x [N][P][P]        */
    const double *fcm;                    /* int_tri_fc    [N][3][P][P]     */
    const double *ttfc;                   /* int_tri_tri_fc[N][3][P][P]     */
    const double *nx, *ny;                /* outward normals [N][3]         */
    const double *int_tri;                /* [P][N]                         */
    const int32_t *tri_faces;             /* [N][3]                         */
    const int32_t *face_bc;               /* [F], -1 interior              */
    const int32_t *neighbour;             /* [N][3], -1 on boundaries      */
    const int32_t *periodic_tri;          /* [F], element behind the pair  */
    const int32_t *other_lfc;             /* [N][3] local face of the
                                             element across (interior or
                                             periodic), -1 on walls        */
    const int32_t *bc_type;               /* [nbc] 1 thermal, 2 adiabatic,
                                             3 periodic                    */
    const double *bc_temp;                /* [nbc]                          */
    const int32_t *face_tri, *face_lfc;   /* [F][2]                         */
    const int32_t *order;                 /* reference sweep order [D][N]  */
    int32_t max_bundle;                   /* directions per workgroup cap,
                                             0 = automatic                 */
    int32_t block;                        /* threads per sweep workgroup,
                                             0 = automatic                 */
    int32_t sweep_variant;                /* 0 pipelined (LDS trace pool,
                                             gather 2 levels ahead), 1
                                             plain level sweep, 2 / 3
                                             pipelined with the gather 1 /
                                             2 levels ahead, 4 the fused
                                             binary32 sweep with the gather
                                             3 levels ahead (other sweeps
                                             of such a context: 2 ahead);
                                             all give the same bits        */
    int32_t lds_slots_cap;                /* > 0 caps the LDS trace pool
                                             slots (exercises the global
                                             fallback); 0 = automatic      */
    int32_t lds_value_bytes;              /* 4: size the LDS trace pool
                                             for binary32 sweeps (a
                                             binary64 sweep whose pool then
                                             exceeds the LDS runs the plain
                                             level kernel); 0 or 8: for
                                             binary64 sweeps              */
    int32_t pairs_or_fail;                /* 1: a schedule that pairs
                                             directions fails (NULL) when a
                                             bundle's trace pool would
                                             overflow the LDS, instead of
                                             sweeping that bundle through
                                             global memory; the caller
                                             then builds one direction per
                                             workgroup                     */
} pbte_setup;

/* ---- setup ------------------------------------------------------------ */

/* Per-direction element order, bit-exact with the reference's Fortran-scan
 * topological sort; replaces ordering._order_kernel / build_sweep_order
 * (solver-python/pybte/ordering.py:61-175).  Host only, no device needed.
 * Returns 0, or -1 when a cycle is found and allow_break == 0. */
int pbte_build_order(int32_t ndir, int32_t n_tris, const int32_t *neighbour,
                     const double *nx, const double *ny, const double *cx,
                     const double *cy, int32_t allow_break, int32_t *order,
                     int32_t *cycles);

/* Uploads the setup, builds the level schedule and direction bundles.
 * Replaces SweepContext.build (sweep.py:74-111).  stream is a hipStream_t
 * (NULL = default stream).  Returns NULL on failure with the message copied
 * into err. */
pbte_ctx *pbte_ctx_create(const pbte_dims *dims, const pbte_setup *setup, void *stream,
                          char *err, int32_t errlen);

void pbte_ctx_destroy(pbte_ctx *ctx);

const char *pbte_last_error(pbte_ctx *ctx);

/* Schedule statistics: levels summed over bundles, bundles, widest level
 * (items), most LDS trace slots one bundle holds at once, bundles that read
 * this sweep's values from global memory (their pool would overflow), LDS
 * trace slots per workgroup. */
int pbte_schedule_info(pbte_ctx *ctx, int64_t *out6);

/* Sweep-kernel timing with HIP events recorded on the context's stream
 * around each sweep-kernel launch (source and wall-vector kernels
 * excluded).  No counterpart in the reference: instrumentation for
 * benchmarks.  pbte_timing(ctx, n) arms n event pairs (0 disarms) and
 * clears earlier records; pbte_timing_read synchronises the stream and
 * writes the durations (ms) of up to n recorded launches, returning how
 * many were written. */
int pbte_timing(pbte_ctx *ctx, int32_t capacity);
int pbte_timing_read(pbte_ctx *ctx, float *ms, int32_t n);

/* Factorise A_SOL for every (direction, element) in the given precision
 * (PBTE_FP64_FP32: the binary64 factors and their binary32 copy from one
 * pass).  Replaces SweepContext.factorise / _kernels.build_operators
 * (sweep.py:116-129, _kernels.py:43-148).  The operators are built from the
 * reference-element matrices and cell scalars of pbte_set_cell_operators
 * when those are set (the same values as the element tensors, bit for
 * bit), else from the element tensors.  Returns 0, or the number of
 * singular element operators (> 0). */
int pbte_factorise(pbte_ctx *ctx, int32_t precision);

/* One transport sweep, in place on vdf (device, [D][N][P], double for
 * PBTE_FP64 and float for PBTE_FP32).  ts/qxs/qys: device [P][N] double.
 * flux_wall: device [F][P] double (adiabatic emission), pbuf: device
 * [D][F][P] in the sweep precision, read instead of the partner element when
 * use_pbuf != 0.  homogeneous != 0 drops the thermalising wall inflow.
 * Replaces SweepContext.sweep / _kernels.sweep_precomputed
 * (sweep.py:132-158, _kernels.py:153-239). */
int pbte_sweep(pbte_ctx *ctx, int32_t precision, const double *ts, const double *qxs,
               const double *qys, void *vdf, const double *flux_wall, const void *pbuf,
               int32_t use_pbuf, int32_t homogeneous);

/* The same sweep with every element operator built and LU-factorised in
 * binary64 where it is used, so no factors are stored (pbte_factorise is
 * not needed); arguments as pbte_sweep.  Replaces the "onthefly" mode of
 * SweepContext.sweep / _kernels.sweep_onthefly (sweep.py:104-110, 147-156,
 * _kernels.py:242-269). */
int pbte_sweep_onthefly(pbte_ctx *ctx, int32_t precision, const double *ts, const double *qxs,
                        const double *qys, void *vdf, const double *flux_wall, const void *pbuf,
                        int32_t use_pbuf, int32_t homogeneous);

/* Angular moments of vdf (precision as in pbte_sweep) in binary64,
 * directions summed in index order.  Replaces moments.compute_moments /
 * _kernels.moments_kernel (moments.py:51-55, _kernels.py:311-338). */
int pbte_moments(pbte_ctx *ctx, int32_t precision, const void *vdf, double *ts, double *qxs,
                 double *qys, double *temp, double *qx, double *qy);

/* Diffuse adiabatic wall emission flux_wall[F][P] (device, double).
 * Replaces BoundaryData.update_wall_flux / _kernels.wall_flux_kernel
 * (bc.py:47-58, _kernels.py:363-397). */
int pbte_wall_flux(pbte_ctx *ctx, int32_t precision, const void *vdf, double *flux_wall);

/* Per-direction sum_i ||A f - b||^2 dw and ||b||^2 dw in binary64 at the
 * current state, no sweep (vdf device double [D][N][P]); out: host [2][D].
 * Replaces SweepContext.true_residual / _kernels.transport_residual
 * (sweep.py:161-197, _kernels.py:274-306). */
int pbte_transport_residual(pbte_ctx *ctx, const double *ts, const double *qxs,
                            const double *qys, const double *vdf, const double *flux_wall,
                            double *out_host);

/* sqrt(sum (T - T_old)^2 / sum T^2) and sum T in index order, device
 * inputs [N], host outputs.  Replaces moments.residual_iterate and
 * Moments.mass (moments.py:40-41, 58-72; _kernels.py:340-361). */
int pbte_residual_iterate(pbte_ctx *ctx, const double *temp, const double *t_old,
                          double *res_host, double *mass_host);

/* The same two values into out_dev[0], out_dev[1] (device memory),
 * stream-ordered, without waiting: for histories read back once at the
 * end of a run (the Krylov driver's per-sweep mass, krylov.py:300-305). */
int pbte_residual_iterate_async(pbte_ctx *ctx, const double *temp, const double *t_old,
                                double *out_dev);

/* Periodic partner DOFs of the Krylov state (krylov.py:64-126): gather
 * vdf[d][periodic_tri[f]][:] for every periodic face f (ascending) into
 * out[D][K][P] (binary64), or scatter such a block into pbuf[D][F][P]
 * (precision of the sweep).  Replaces FixedPointOperator.gather/scatter's
 * periodic blocks (krylov.py:99-110). */
int pbte_periodic_gather(pbte_ctx *ctx, int32_t precision, const void *vdf, double *out);
int pbte_periodic_scatter(pbte_ctx *ctx, int32_t precision, const double *in, void *pbuf);

/* The element operators in matrix-free form, for the binary64 sweeps of
 * the "mixed" mode (binary32 inverses stored, binary64 ones not; the
 * reference's sweep_onthefly, _kernels.py:242-269, builds and factorises
 * them instead): ref [6][P][P] = the reference-element mass matrix, the
 * two reference derivative matrices and the three edge matrices of
 * integration.py:199-260; cell [N][8] = twice the area, the gradient
 * coefficients (y3-y1)/s, (y1-y2)/s, (x1-x3)/s, (x2-x1)/s, and the edge
 * lengths len12, len23/sqrt(2), len31.  With both set and binary32 factors
 * built, pbte_sweep_onthefly in binary64 applies inv32 refined twice
 * against A formed from these.  The same data (ref's first three matrices,
 * per element area2/tau_c and area2 times the four gradient coefficients)
 * feeds the fused binary32 sweep of pbte_set_fused. */
int pbte_set_cell_operators(pbte_ctx *ctx, const double *ref, const double *cell);

/* The reference's on-the-fly mode (sweep.py:104-110, _kernels.py:242-269:
 * every element operator built and factorised in the sweep) for binary32
 * sweeps, fused into the pipelined sweep: on != 0 makes every binary32
 * pbte_sweep / pbte_sweep_onthefly (and batched binary32 sweep) build A
 * from the reference-element matrices and the element's scalars in
 * registers and solve it by elimination without pivoting (A + A^T is
 * positive definite), so no binary32 factors are stored or read.  Needs
 * pbte_set_cell_operators first, P <= 10, straight-sided elements (every
 * face mass block a length times one unit-edge matrix, checked when the
 * context is created) and a context that runs the pipelined kernel; else
 * -1 with the reason in pbte_last_error.  on == 0 switches it off.  With
 * on == 2, P <= 6 and room in the LDS for two factor buffers, a sweep runs
 * the split kernel: a workgroup of twice the width whose second half
 * builds and factorises the next level's operators into LDS while the
 * first half solves the current level (the same values; slower than one
 * wave per item on the configurations measured, whose vector-memory
 * address unit is the busier resource). */
int pbte_set_fused(pbte_ctx *ctx, int32_t on);

/* Which kernel a sweep of this precision runs (pbte_sweep for stored
 * factors, else pbte_sweep_onthefly): 1 plain level kernel, 2 pipelined
 * with stored factors, 3 fused, 4 split fused, 5 on the fly, 6 binary64
 * refined from the binary32 inverse ("mixed").  No counterpart in the
 * reference: for tests and benchmark records. */
int pbte_sweep_path(pbte_ctx *ctx, int32_t precision);

/* Staging of the rare global upwind reads of the pipelined sweeps: the
 * previous-sweep values a face reads across a broken cycle or a periodic
 * pair whose partner is swept later, and the explicit periodic buffer's
 * entries, copied into LDS stage slots by each workgroup's prologue (the
 * same values: results are bitwise those of the global reads).  A launch
 * whose LDS has no room for its stage slots reads them from global memory.
 * on = 1 stages (the default), 0 reads global memory, -1 leaves the
 * setting; out4 (may be NULL) receives the most stage slots of a bundle
 * without and with an explicit periodic buffer, the previous-sweep reads
 * and the periodic reads of the staged schedule classes.  Returns 1 when
 * staging is on, 0 when off, -1 with the reason in pbte_last_error.  No
 * counterpart in the reference (its sweep reads every value from the
 * distribution array). */
int pbte_staging(pbte_ctx *ctx, int32_t on, int64_t *out4);

/* ---- batched sweeps (no counterpart in the reference) ------------------
 * The sweeps of several independent problems in one persistent launch of
 * the pipelined sweep kernel fed by a device work queue: the work items are
 * the contexts' (context, bundle) pairs, longest bundles first; each
 * resident workgroup takes the next item with one atomic until the queue is
 * empty.  The contexts must share the element degree, the stream and the
 * pipeline depth and run the pipelined kernel in the precision swept (all
 * fused or all with stored factors in binary32).  pbte_batch_sweep takes
 * per context the moment sources, the distribution (of the precision),
 * flux_wall (NULL without adiabatic walls) and the explicit periodic buffer
 * (pbuf NULL, or a NULL entry: none; as pbte_sweep's use_pbuf), runs each
 * context's source and wall-vector kernels and then the one sweep launch;
 * each context's sweep is the one pbte_sweep computes, bit for bit.  For
 * many small problems, whose workgroups alone leave most of the CUs idle. */
typedef struct pbte_batch pbte_batch;
pbte_batch *pbte_batch_create(pbte_ctx *const *ctxs, int32_t n, char *err, int32_t errlen);
void pbte_batch_destroy(pbte_batch *batch);
const char *pbte_batch_last_error(pbte_batch *batch);
int pbte_batch_sweep(pbte_batch *batch, int32_t precision, const double *const *ts,
                     const double *const *qxs, const double *const *qys, void *const *vdf,
                     const double *const *flux_wall, const void *const *pbuf, int32_t homogeneous);
/* pbte_batch_sweep in two halves.  The launch runs on the batch's own
 * stream after the contexts' stream has reached it; _async returns without
 * making the contexts' stream wait for it, so the launches of several
 * batches (groups of problems that need different kernel instances)
 * overlap on the device; pbte_batch_join makes the contexts' stream wait
 * for the batch's last launch (the next _async of the same batch joins
 * first).  Join before using the swept distributions on that stream. */
int pbte_batch_sweep_async(pbte_batch *batch, int32_t precision, const double *const *ts,
                           const double *const *qxs, const double *const *qys, void *const *vdf,
                           const double *const *flux_wall, const void *const *pbuf,
                           int32_t homogeneous);
int pbte_batch_join(pbte_batch *batch);
/* 1 when every context of the batch runs the pipelined kernel in this
 * precision (pbte_batch_sweep can sweep them together), else 0: the caller
 * then sweeps those contexts one by one with pbte_sweep. */
int pbte_batch_pipelined(pbte_batch *batch, int32_t precision);

/* ---- device vector kernels of the Krylov solver (krylov.py:150-246) --- */

/* c[k] = sum_i Q[k][i] w[i], k < m (Q device [m][n], row stride ld);
 * result on the host, reduction order fixed. */
int pbte_vec_multidot(pbte_ctx *ctx, const double *Q, int64_t ld, int32_t m,
                      const double *w, int64_t n, double *c_host);

/* w[i] -= sum_k c[k] Q[k][i] (c from the host). */
int pbte_vec_update(pbte_ctx *ctx, const double *Q, int64_t ld, int32_t m,
                    const double *c_host, double *w, int64_t n);

/* Classical Gram-Schmidt applied twice to w against the m rows of Q
 * (krylov.py:203-207): out_host[0..m) = c1, out_host[m..2m) = c2 (w is
 * left as w - Q^T c1 - Q^T c2), out_host[2m] = |w|^2 after both passes.
 * Same kernels and reduction order as two multidot/update pairs and a
 * multidot of w with itself, with one device-to-host copy; m <= 512. */
int pbte_vec_cgs2(pbte_ctx *ctx, const double *Q, int64_t ld, int32_t m, double *w, int64_t n,
                  double *out_host);

/* pbte_vec_cgs2 split in two so that the next Krylov product can be queued
 * while the coefficients travel: _begin queues the same kernels, a copy of
 * (c1, c2, |w|^2) to pinned host memory behind an event and, when next is
 * not NULL, next = w / |w| on the device (axpby's rounding); _end waits for
 * the event and writes the 2m+1 values to out_host.  One pair in flight
 * per context: a _begin while one is open waits for that one's copy and
 * drops it; _end without an open _begin fails. */
int pbte_vec_cgs2_begin(pbte_ctx *ctx, const double *Q, int64_t ld, int32_t m, double *w,
                        int64_t n, double *next);
int pbte_vec_cgs2_end(pbte_ctx *ctx, double *out_host);

/* out = (v - t) + g, or v - t when g is NULL: the operator product of the
 * Krylov drivers, A v = v - (F(v) - g) (krylov.py:332-337), in the
 * rounding of a copy followed by two axpby calls. */
int pbte_vec_defect(pbte_ctx *ctx, const double *v, const double *t, const double *g, double *out,
                    int64_t n);

/* The relaxation of the GSIS fixed-point repair
 * (acceleration/__init__.py:144-147, refresh_defect): *step_host =
 * max |o - d| (NaN when any entry is NaN, as numpy's max), then, unless
 * omega == 0, d += omega (o - d) in numpy's rounding; waits for the
 * stream.  With omega == 0 and o == d it tests o for non-finite entries. */
int pbte_vec_relax(pbte_ctx *ctx, double omega, const double *o, double *d, int64_t n,
                   double *step_host);

/* y[i] = a x[i] + b y[i] */
int pbte_vec_axpby(pbte_ctx *ctx, double a, const double *x, double b, double *y, int64_t n);

/* ---- GSIS with the HDG synthetic system (acceleration package) -------- */

/* Arrays read once by pbte_gsis_setup, host unless marked device (index orders as the
 * reference's Acceleration, acceleration/__init__.py:37-92). */
typedef struct pbte_gsis_data {
    const double *cxv, *cyv, *domega;     /* [D]                            */
    const double *inv_aa_sol;             /* device [N][7P][7P]             */
    const double *aa_trace;               /* device [N][3][7P][3Q]          */
    const double *ba_sol;                 /* device [N][3][3Q][7P]          */
    const double *tri_fc_fc;              /* int_tri_fc_fc [P][Q][N][3]     */
    const double *tri_hmin;               /* [N]                            */
    const int32_t *tri_faces;             /* [N][3]                         */
    const int32_t *t1, *l1, *t2, *l2;     /* [F] contributing element sides */
    const uint8_t *flip;                  /* [F] partner rows mirrored      */
    int32_t n_wall, n_wall_th;            /* wall faces: thermalising first */
    const int32_t *wall_face, *wall_tri, *wall_lfc;
    const double *wall_nx, *wall_ny;
    double tc5, tau_thr, drift;           /* tau_c*scale/5, TAU_THR,
                                             tau_c/tau_n*3/(4 pi vg^2)     */
    int32_t variant_b, use_hmin;
} pbte_gsis_data;

/* Uploads the GSIS operators.  Replaces Acceleration.__init__'s device-side
 * state (acceleration/__init__.py:37-92). */
int pbte_gsis_setup(pbte_ctx *ctx, const pbte_gsis_data *g);

/* Seven angular moments per (dof, cell) in binary64, out [7][P][N]: the
 * CIS moments (1, cx, cy) and the four third-order weights of the HoT
 * source.  Replaces the direction loops of hot_source._hot_kernel and
 * correction._correct_kernel (hot_source.py:29-58, correction.py:40-52). */
int pbte_gsis_moments(pbte_ctx *ctx, int32_t precision, const void *vdf, double *mom7);

/* The element pass of the accelerated step: the stress work of the HoT
 * source from mom7, then src [3P][N] = the T, qx, qy rows of AA_SRC =
 * inv(AA_SOL) work and, kept in the context for pbte_gsis_rhs, the sums
 * BA_SOL AA_SRC of each element side -- one operator per element
 * [4P][3P + 9Q] built by pbte_gsis_setup (both are linear in the work).
 * precision PBTE_FP32 reads binary32 copies of the operators (made on the
 * first such call; sums in binary64), for the binary32 iterations of the
 * refined run.  Replaces hot_source.hot_source (hot_source.py:61-75) and
 * the element half of GlobalSolver.rhs (global_solve.py:128-150). */
int pbte_gsis_source(pbte_ctx *ctx, int32_t precision, const double *mom7, double *src);

/* Trace right-hand side ffa [F][3Q] from the side sums of the last
 * pbte_gsis_source or pbte_gsis_precond_src (two-sided, periodic mirror)
 * and the wall rows from the kinetic traces of vdf (vdf == NULL stands for
 * a zero distribution, wall rows zero).  Replaces GlobalSolver.rhs
 * (global_solve.py:128-186). */
int pbte_gsis_rhs(pbte_ctx *ctx, int32_t precision, const void *vdf, double *ffa);

/* Builds the moment preconditioner's element operator from inv_aa_sol
 * [N][7P][7P] and ba_sol [N][3][3Q][7P] (device): the T, qx, qy columns of
 * the inverse in place of the stress columns.  Part of
 * MomentPreconditioner.__init__ (krylov.py:379-390). */
int pbte_gsis_precond_setup(pbte_ctx *ctx, const double *inv_aa_sol, const double *ba_sol);

/* src [3P][N] and the side sums (as pbte_gsis_source) for the element
 * source inv(AA_SOL) applied to sign * the T, qx, qy blocks of r (three
 * [P][N] blocks, device).  Replaces the element source of
 * MomentPreconditioner._apply (krylov.py:392-402). */
int pbte_gsis_precond_src(pbte_ctx *ctx, const double *r, double sign, double *src);

/* uq [3P][N] = the T, qx, qy rows of UQ = AA_SRC + sum AA_TRACE U_TRACE
 * (u_trace [F][3Q]); precision as pbte_gsis_source.  Replaces
 * local_solve.local_solve (local_solve.py:16-26) for the rows the
 * correction reads. */
int pbte_gsis_local(pbte_ctx *ctx, int32_t precision, const double *src, const double *u_trace,
                    double *uq);

/* The full UQ [7P][N] of a step from its mom7 and u_trace (binary64), the
 * stress rows included (the field output's Nxx, Nxy, Nyy).  Replaces
 * local_solve.local_solve (local_solve.py:16-26) where every row is read. */
int pbte_gsis_uq(pbte_ctx *ctx, const double *mom7, const double *u_trace, double *uq);

/* Damped kinetic correction of vdf and the moments; defect and mkin
 * ([3][P][N]) may be NULL.  Replaces correction.correct_vdf
 * (correction.py:29-98). */
int pbte_gsis_correct(pbte_ctx *ctx, int32_t precision, void *vdf, const double *mom7,
                      const double *uq, const double *defect, double *ts, double *qxs,
                      double *qys, double *temp, double *qx, double *qy, double *mkin);

/* The correction split in two halves, so that a caller can leave the
 * distribution untouched when the next sweep overwrites it anyway:
 * pbte_gsis_correct_moments writes the corrected moments (and mkin) like
 * pbte_gsis_correct and, instead of updating vdf, the coefficients
 * coef [3][P][N] (device) of the distribution's update;
 * pbte_gsis_apply_correction applies them to vdf, on every cell (cells ==
 * NULL) or on the ncells listed (device int32).  Both halves in turn give
 * the bits of pbte_gsis_correct.  Replace correction.correct_vdf
 * (correction.py:29-98) together. */
int pbte_gsis_correct_moments(pbte_ctx *ctx, const double *mom7, const double *uq,
                              const double *defect, double *ts, double *qxs, double *qys,
                              double *temp, double *qx, double *qy, double *mkin, double *coef);
int pbte_gsis_apply_correction(pbte_ctx *ctx, int32_t precision, void *vdf, const double *coef,
                               const int32_t *cells, int32_t ncells);

/* ---- host set-up: mesh file and face integrals ---------------------------
 * Host-only (no device memory).  A gmsh 2.2 ASCII file is read into a
 * handle; pbte_gmsh_sizes gives (nodes, elements, bytes of the physical-name
 * lines) and pbte_gmsh_copy copies out the node coordinates [n][2], element
 * type (1 line, 2 triangle), first tag (physical id), nodes [n][3] (0-based,
 * -1 after a line's two) and the $PhysicalNames item lines ('\n'-ended).
 * Errors (a format other than 2.2 ASCII, node ids not 1..n in order, an
 * element of another type or with fewer than two tags, a short section)
 * return NULL with the message in err.  Replaces mesh/gmsh_reader.py:45-122. */
typedef struct pbte_gmsh pbte_gmsh;
pbte_gmsh *pbte_gmsh_read(const char *path, char *err, int32_t errlen);
int pbte_gmsh_sizes(const pbte_gmsh *g, int64_t *out3);
int pbte_gmsh_copy(const pbte_gmsh *g, double *xy, int32_t *type, int32_t *phyid, int32_t *conn,
                   char *names);
void pbte_gmsh_free(pbte_gmsh *g);

/* The face integrals of the nodal basis on every (element, local face):
 * int_tri_fc_fc [P][Q][N][3] (element basis x face basis) and
 * int_tri_tri_fc [N][3][P][P] (element basis x the basis of the element
 * across -- interior neighbour, or periodic partner's element shifted by
 * off -- or, where other < 0, int phi_m ds in every column), by the
 * np_fc-point Gauss-Legendre rule (abscissae yq, Legendre derivative lpq).
 * threads <= 0: up to 16 host threads.  Replaces the face loops of
 * integration.py:321-404. */
typedef struct pbte_face_integrals_desc {
    int32_t n_tris, nd, nf, np_fc;
    const double *nodes;        /* [n_nodes][2]                              */
    const int32_t *tri_nodes;   /* [N][3]                                    */
    const int32_t *tri_faces;   /* [N][3]                                    */
    const int32_t *face_nodes;  /* [F][2]                                    */
    const double *face_len;     /* [F]                                       */
    const int32_t *other;       /* [N][3] element across, -1 on a wall       */
    const double *off_x, *off_y;/* [N][3] shift of that element's nodes      */
    const double *C;            /* [P][P] nodal coefficients of the monomials */
    const double *Cf;           /* [Q][Q] face basis coefficients             */
    const int32_t *mono;        /* [P][2] monomial exponents (a, b)          */
    const double *yq, *lpq;     /* [np_fc]                                   */
} pbte_face_integrals_desc;
int pbte_face_integrals(const pbte_face_integrals_desc *d, double *int_tri_fc_fc,
                        double *int_tri_tri_fc, int32_t threads);

/* ---- HDG operators built on the device ---------------------------------
 * The element operators and the trace matrix of the HDG synthetic system.
 * Replace acceleration/assembly.py:44-268 (build_local_operators,
 * build_ba_sol, assemble_global_matrix).  All pointers in the descriptor
 * are device memory, binary64 tensors in the reference's index orders. */
typedef struct pbte_hdg_desc {
    int32_t n_tris, nd, nf, n_faces;      /* N, P, Q, F                      */
    int32_t variant_b;                    /* 1: variant B (scaled by tau_R)  */
    double cv, tau_r, tau_c;
    double st[3];                         /* stabilisation (after any h_min
                                             scaling)                        */
    double drop_tol;                      /* |K| <= drop_tol set to zero     */
    const double *int_tri_tri;            /* [P][P][N]                       */
    const double *int_tri_tri_x, *int_tri_tri_y;   /* [N][P][P]             */
    const double *int_tri_fc;             /* [N][3][P][P]                    */
    const double *int_tri_fc_fc;          /* [P][Q][N][3]                    */
    const double *int_fc_fc;              /* [Q][Q][F]                       */
    const double *tri_normal;             /* [N][3][2]                       */
    const int32_t *face_kind;             /* [N][3]: 0 interior or periodic,
                                             1 thermalising wall, 2 other   */
} pbte_hdg_desc;

/* inv(AA_SOL) [N][7P][7P] (skipped when NULL), AA_TRACE [N][3][7P][3Q] and
 * BA_SOL [N][3][3Q][7P] into device buffers.  0, or -2 with the message in
 * err (a singular element operator among them). */
int pbte_hdg_local(const pbte_hdg_desc *d, double *inv_aa_sol, double *aa_trace, double *ba_sol,
                   void *stream, char *err, int32_t errlen);

/* Block values [nnzb][3Q][3Q] of K on the block pattern (bsr_ptr [F+1],
 * bsr_col [nnzb], device), from the element sides t1/l1, t2/l2 [F]
 * (-1 absent), mirror [F] (partner side of a reversed periodic pair) and
 * tri_faces [N][3] (device); *nnz_host = entries left nonzero. */
int pbte_hdg_assemble(const pbte_hdg_desc *d, const int32_t *t1, const int32_t *l1,
                      const int32_t *t2, const int32_t *l2, const uint8_t *mirror,
                      const int32_t *tri_faces, const int32_t *bsr_ptr, const int32_t *bsr_col,
                      int64_t nnzb, const double *aa_trace, const double *ba_sol, double *bsr_val,
                      int64_t *nnz_host, void *stream, char *err, int32_t errlen);

/* pbte_hdg_assemble without any allocation: the same launches and the same
 * bits, with the block-row indices and the nonzero counter in a 16 MiB
 * file-scope device array (one per GPU, held until the process ends)
 * instead of hipMalloc'd buffers, so at most 4 194 240 blocks (-2 and a
 * message beyond).  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, a serving
 * loop re-assembling K per parameter).  Calls that overlap on one GPU share
 * the array and race. */
int pbte_hdg_assemble_static(const pbte_hdg_desc *d, const int32_t *t1, const int32_t *l1,
                             const int32_t *t2, const int32_t *l2, const uint8_t *mirror,
                             const int32_t *tri_faces, const int32_t *bsr_ptr,
                             const int32_t *bsr_col, int64_t nnzb, const double *aa_trace,
                             const double *ba_sol, double *bsr_val, int64_t *nnz_host,
                             void *stream, char *err, int32_t errlen);

/* ---- HDG global trace solve on the device ------------------------------
 * K u = f for the trace matrix of the HDG synthetic system, by a
 * nested-dissection multifrontal elimination in binary64 (dense fronts,
 * partial pivoting among each front's fully summed rows).  Replaces the
 * SuperLU factorisation and solve of GlobalSolver
 * (acceleration/global_solve.py:120-124 factorise, :190-195 solve). */
typedef struct pbte_trace pbte_trace;

typedef struct pbte_trace_desc {
    int32_t n_faces;              /* F                                        */
    int32_t n_cells;              /* N                                        */
    int32_t blk;                  /* unknowns per face, 3Q                    */
    int32_t leaf_cells;           /* cells per dissection leaf, 0 = 16        */
    const int32_t *face_t1;       /* [F] first element a face couples         */
    const int32_t *face_t2;       /* [F] second one (periodic partner's
                                     element across a periodic pair), -1 on
                                     walls                                    */
    const int32_t *tri_faces;     /* [N][3]                                   */
    const double *cell_xy;        /* [N][2] element centroids                 */
    const int32_t *bsr_ptr;       /* [F+1] block rows of K (3Q x 3Q blocks)   */
    const int32_t *bsr_col;       /* [nnzb] block columns (face ids)          */
} pbte_trace_desc;

/* Symbolic phase on the host (the dissection tree, fronts, assembly maps);
 * device memory is taken by the first pbte_trace_factor.  Every later call
 * is ordered on `stream`.  NULL on failure with the message in err. */
pbte_trace *pbte_trace_create(const pbte_trace_desc *desc, void *stream, char *err,
                              int32_t errlen);
void pbte_trace_destroy(pbte_trace *t);
const char *pbte_trace_last_error(pbte_trace *t);

/* out8: tree nodes, levels, largest front (unknowns), factor entries,
 * factorisation flops, device workspace bytes, unknowns, assembly blocks. */
int pbte_trace_info(pbte_trace *t, int64_t *out8);

/* Pivoting of the numeric factorisation: -1 (default) diagonal-block
 * pivots (interchanges within each 32-column block of a front's fully
 * summed rows, every step a dense product), with partial pivoting among all
 * of a front's fully summed rows when a block has no pivot or the normwise
 * backward error of the factors on a fixed test vector exceeds 1e-13; 0 or
 * 1 force the one or the other.  No counterpart in the reference (SuperLU's
 * own pivoting is internal). */
int pbte_trace_options(pbte_trace *t, int32_t pivoting);

/* Per level of the dissection tree (0 = root): fronts, widest front, fully
 * summed rows, contribution rows, stored [Z V] and W entries (padded rows),
 * forward and backward work items of the solve. */
int pbte_trace_level(pbte_trace *t, int32_t level, int64_t *out8);

/* The pivoting the factors were built with (0 block, 1 partial) and their
 * backward error on the test vector. */
int pbte_trace_quality(pbte_trace *t, int32_t *pivoting, double *berr);

/* Numeric factorisation from the block values bsr_val (device,
 * [nnzb][3Q][3Q] row-major, the pattern given to pbte_trace_create).
 * Returns 0, or 1 when a front has no nonzero pivot (singular K). */
int pbte_trace_factor(pbte_trace *t, const double *bsr_val);

/* x = K^-1 rhs (device, [F][3Q] each; x must not alias rhs), from the
 * binary64 factors (precision PBTE_FP64) or from their binary32 copy
 * (PBTE_FP32: made by the first such solve after each factorisation, sums
 * in binary64; the correction solves of the refined run's binary32
 * iterations). */
int pbte_trace_solve(pbte_trace *t, int32_t precision, const double *rhs, double *x);

/* The solve's kernel launches (per level of the dissection tree) recorded
 * as a HIP graph the first time a (precision, rhs, x) triple is solved and
 * replayed after (up to 8 graphs per factorisation, recorded on a private
 * stream, replayed on the solver's stream; a factorisation drops them).
 * on = 0 launches every kernel directly (the same values).  No counterpart
 * in the reference: the solve is launch-bound for small systems, and a
 * batch of problems solves one per problem per iteration. */
int pbte_trace_graphs(pbte_trace *t, int32_t on);

#ifdef __cplusplus
}
#endif

#endif /* PBTE_HIP_H */
