/*
 * ptd.h -- C ABI of libptd, the MI355X (gfx950) pseudo-transient diffusion library.
 *
 * The reference (PTsolvers/PseudoTransientDiffusion.jl) has no native library: each script defines its
 * kernels as ParallelStencil functions and launches them with `@parallel f(args...)`, then runs the
 * pseudo-transient (PT) loop on the host. This header is the boundary that replaces those launches and
 * that loop. A Julia host binds it with `ccall`; the Python package in this repository binds it with
 * ctypes. Signatures use plain C types only.
 *
 * Conventions
 *  - Arrays cross the boundary as host `double*` in the reference's layout: column-major (x fastest)
 *    with the reference's staggered shapes (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:77-83):
 *        H, Hold, D : (nx, ny, nz)
 *        qx         : (nx-1, ny-2, nz-2)   qy : (nx-2, ny-1, nz-2)   qz : (nx-2, ny-2, nz-1)
 *        ResH, dtau : (nx-2, ny-2, nz-2)
 *    In 1D ny = nz = 1 and in 2D nz = 1; an absent axis contributes a factor 1 (not n-2) to every shape.
 *  - The library owns the device copies inside an opaque context (one HIP stream per context).
 *    Every call is synchronous from the caller's view except where noted. A context is not
 *    thread-safe; separate contexts are independent.
 *  - Every function returns a ptd_status; ptd_last_error() describes the most recent failure of the
 *    calling thread.
 */
#ifndef PTD_H
#define PTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTD_ABI_VERSION 1

typedef enum ptd_status {
    PTD_OK = 0,
    PTD_ERR_ARG = 1,      /* invalid argument (shape, enum, NULL pointer, size mismatch) */
    PTD_ERR_HIP = 2,      /* HIP runtime error, message in ptd_last_error() */
    PTD_ERR_NAN = 3,      /* the residual norm became NaN: the reference's error("NaN") */
    PTD_ERR_NOMEM = 4,    /* device allocation failed */
    PTD_ERR_STATE = 5,    /* call not valid for this context (e.g. field absent for this case) */
    PTD_ERR_NODEV = 6     /* no usable gfx950 device */
} ptd_status;

/* Diffusion coefficient: D = d0 (scalar), a step field D (nx,ny,nz), or D = H^3. */
typedef enum ptd_dcase { PTD_LIN = 0, PTD_STEP = 1, PTD_NONLIN = 2 } ptd_dcase;

/* Expression form of the kernels.
 *  PTD_PLAIN: the converging scripts (diff_*_{lin,linstep,nonlin}.jl): /dx, /dt, and the residual
 *             flux qH2 taken from the H *before* the last update (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:41-43).
 *  PTD_PERF : the perf scripts (diff_*_nonlin_perf.jl): multiplications by _dx = 1/dx, _dt = 1/dt and
 *             the residual flux recomputed from the current H (scripts/diff_3D/diff_3D_nonlin_perf.jl:44-49). */
typedef enum ptd_form { PTD_PLAIN = 0, PTD_PERF = 1 } ptd_form;

/* Arithmetic of the PT sweeps.
 *  PTD_FP64 : as the reference (@init_parallel_stencil(..., Float64, ...)).
 *  PTD_FP32 : fields and sweeps in float32.
 *  PTD_MIXED: float32 sweeps inside a float64 defect-correction loop: the residual of the float64
 *             backward-Euler system and the accumulated solution stay float64 on the device (3D
 *             only). ptd_pt_solve then stops on the float64 residual with the flux taken from the
 *             current iterate (the consistent residual), and its iteration count is the number of
 *             float32 correction sweeps. */
typedef enum ptd_precision { PTD_FP64 = 0, PTD_FP32 = 1, PTD_MIXED = 2 } ptd_precision;

/* Arithmetic of the sweeps.
 *  PTD_ARITH_EXACT: every expression in the reference's operation order with IEEE division and
 *                   square root (bit-identical to the reference kernels' Float64 results).
 *  PTD_ARITH_FAST : the same update rules rearranged algebraically, q' = q - (q + g)*Re/(Re + c1) and
 *                   H' = H + R*dtau/(1 + dtau/dt) with R the residual, so that the damping
 *                   coefficients (computed from hardware reciprocal / reciprocal square root with one
 *                   Newton step, ~4e-15 relative) do not move the fixed point; agrees with EXACT to
 *                   ~1e-14 per iteration and to < 1e-10 relative on converged solutions. Requires
 *                   H > 0 in the nonlinear case. */
typedef enum ptd_arith { PTD_ARITH_EXACT = 0, PTD_ARITH_FAST = 1 } ptd_arith;

/* Norm that stops the PT loop. The reference stops on RMS = norm(ResH)/sqrt(length(ResH))
 * (scripts/diff_3D/diff_3D_nonlin_perf.jl:106); PTD_NORM_MAX stops on max|ResH| instead. Both are
 * always computed. */
typedef enum ptd_norm { PTD_NORM_RMS = 0, PTD_NORM_MAX = 1 } ptd_norm;

typedef enum ptd_field {
    PTD_F_H = 0, PTD_F_HOLD = 1, PTD_F_D = 2,
    PTD_F_QX = 3, PTD_F_QY = 4, PTD_F_QZ = 5,
    PTD_F_QX2 = 6, PTD_F_QY2 = 7, PTD_F_QZ2 = 8,
    PTD_F_RESH = 9, PTD_F_DTAU = 10
} ptd_field;

/* Problem description. The scalar block carries the values the scripts compute in their
 * "Derived numerics" section (e.g. scripts/diff_3D/diff_3D_nonlin_perf.jl:73-78,
 * scripts/diff_3D/diff_3D_lin_multixpu.jl:58-62); the library uses them as given, exactly as the
 * reference kernels receive them as arguments. */
typedef struct ptd_params {
    int32_t ndim;       /* 1, 2 or 3 */
    int32_t dcase;      /* ptd_dcase */
    int32_t form;       /* ptd_form */
    int32_t precision;  /* ptd_precision */
    int64_t nx, ny, nz; /* cells per axis, >= 3 on every active axis; 1 on absent axes */
    int64_t itmax;      /* PT iteration cap per physical step (itMax) */
    int32_t nout;       /* residual check interval (nout) */
    int32_t stop_norm;  /* ptd_norm */
    int32_t arith;      /* ptd_arith of the fused sweeps (the kernel-level entry points are always EXACT) */
    int32_t reserved;   /* 0 */
    double dx, dy, dz;  /* cell sizes */
    double dt;          /* physical time step (theta_kin for the reaction scripts) */
    double vpdt;        /* Vpdtau = CFL*min(dx,dy,dz) */
    double max_l;       /* max(lx,ly,lz) (lx in 1D) */
    double max_l2;      /* max_l^2 */
    double resc;        /* Resc (1/1.2 for 2D/3D nonlin, 1 otherwise) */
    double d0;          /* lin: scalar D */
    double re;          /* lin/step: Re */
    double theta_r;     /* lin/step: theta_r_dtau */
    double dtau_rho;    /* lin: scalar dtau_rho */
    double tol;         /* PT tolerance on the stopping norm */
    double norm_len;    /* err = ||ResH|| / sqrt(norm_len); 0 selects length(ResH) */
} ptd_params;

typedef struct ptd_ctx ptd_ctx;

/* ---- library / context --------------------------------------------------------------------- */

int32_t     ptd_abi_version(void);
const char* ptd_last_error(void);
/* Number of visible devices and whether device `dev` is gfx950 (returns PTD_ERR_NODEV otherwise). */
int32_t     ptd_device_check(int32_t dev, char* name_out, int32_t name_len);

/* Allocates every device field for `params` (H and q are double-buffered) and zeroes them, like
 * the scripts' @zeros allocations (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:77-83). */
int32_t ptd_ctx_create(const ptd_params* params, int32_t device, ptd_ctx** out);
int32_t ptd_ctx_destroy(ptd_ctx* ctx);
/* Replaces the scalar block (e.g. to run the reaction scripts' theta_kin loop); shapes must match. */
int32_t ptd_ctx_set_params(ptd_ctx* ctx, const ptd_params* params);
/* Device bytes held by the context. */
int64_t ptd_ctx_device_bytes(const ptd_ctx* ctx);
/* PTD_MIXED: float64 defect-correction (outer) iterations of the last PT solve (0 otherwise). No
 * reference counterpart: the reference has no mixed-precision solver. */
int64_t ptd_ctx_outer_iterations(const ptd_ctx* ctx);
/* The context's HIP stream (hipStream_t), for callers that time or order work on it. */
void*   ptd_ctx_stream(ptd_ctx* ctx);

/* Host <-> device copies of one field in the reference layout (count = number of doubles). Setting H
 * writes both H buffers, so the Dirichlet boundary cells hold the given values in either one. */
int32_t ptd_set_field(ptd_ctx* ctx, int32_t field, const double* host, int64_t count);
int32_t ptd_get_field(ptd_ctx* ctx, int32_t field, double* host, int64_t count);
/* Number of doubles of a field for this context's shape (0 if absent for the case). */
int64_t ptd_field_count(const ptd_ctx* ctx, int32_t field);

/* ---- kernel-level entry points: one reference kernel launch each --------------------------- */

/* compute_iter_params!: dtau_rho = Vpdtau*max_l/@maxloc(D)/Re (step case)
 * (scripts/diff_3D/diff_3D_linstep_multixpu.jl:24-27, diff_2D_linstep.jl:18-21, diff_1D_linstep.jl:17-20). */
int32_t ptd_compute_iter_params(ptd_ctx* ctx);
/* compute_flux!: damped fluxes qx, qy, qz from H; the plain form also writes qx2, qy2, qz2
 * (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:37-45, diff_3D_nonlin_perf.jl:32-37, and the 1D/2D
 * scripts' compute_flux!). */
int32_t ptd_compute_flux(ptd_ctx* ctx);
/* compute_update!: H update on the inner cells (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:47-50,
 * diff_3D_nonlin_perf.jl:39-42). */
int32_t ptd_compute_update(ptd_ctx* ctx);
/* compute_flux_res!: qx2, qy2, qz2 from the current H (scripts/diff_3D/diff_3D_nonlin_perf.jl:44-49). */
int32_t ptd_compute_flux_res(ptd_ctx* ctx);
/* check_res! and the error: writes ResH and returns err_rms = norm(ResH)/sqrt(norm_len) and
 * err_max = max|ResH| (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:52-55,113; diff_2D_nonlin_perf.jl:44-47). */
int32_t ptd_check_res(ptd_ctx* ctx, double* err_rms, double* err_max);
/* assign! / Hold .= H (scripts/diff_3D/diff_3D_nonlin_perf.jl:56-59,110). */
int32_t ptd_assign(ptd_ctx* ctx);

/* ---- loop-level entry points: the PT loop, device resident --------------------------------- */

/* `niter` PT iterations (flux + update, fused into one launch each) with no residual check: the
 * perf scripts' inner loop body (scripts/diff_3D/diff_3D_nonlin_perf.jl:100-102). */
int32_t ptd_iterate(ptd_ctx* ctx, int64_t niter);
/* Same as ptd_iterate, bracketed by HIP events on the context's stream; *ms = elapsed time. */
int32_t ptd_iterate_timed(ptd_ctx* ctx, int64_t niter, double* ms);
/* PT iterations per launch of the context's unchecked iterations (ptd_iterate and the runs between checks):
 * 3 (three-level passes), 2 (two-level passes), 1 (single sweeps) or, for small 2D grids, the LDS-resident
 * launch's up to 10. No counterpart in the reference: a query for measurement tools (which kernel a run
 * times), following the library's own choice for these parameters and PTD_T2 / PTD_T3 / PTD_LDS2D. */
int32_t ptd_ctx_iterations_per_launch(const ptd_ctx* ctx, int32_t* n);
/* One PT solve at fixed Hold: `while err > tol && iter < itmax` with a check every nout iterations
 * (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:102-115). Returns PTD_ERR_NAN if the error is NaN. */
int32_t ptd_pt_solve(ptd_ctx* ctx, int64_t* iters, double* err_rms, double* err_max);
/* The physical-time loop: nt x { ptd_pt_solve; Hold .= H } (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:101-119).
 * iters[nt] and errs[nt] (may be NULL) receive per-step iteration counts and final errors. */
int32_t ptd_solve(ptd_ctx* ctx, int32_t nt, int64_t* iters, double* errs);

/* ---- a PT solve that allocates nothing ----------------------------------------------------------
 *
 * No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): ptd_pt_solve on caller-owned device
 * buffers, with no context. The residual reduction's scratch (per-tile partials of the checking sweeps
 * and the reduced sum) lives in a fixed library array of PTD_NOALLOC_MAX_PARTIALS + 1 (sum, max) pairs
 * (2 MiB) per device, held until the process ends; grids whose checks need more partials (cubes
 * beyond about 1260^3, squares beyond about 10600^2) are refused with PTD_ERR_ARG. The call makes no
 * allocation (no hipMalloc / hipHostMalloc / hipFree). Calls that overlap on one device share that
 * array and race: serialise them (one stream, or one host thread per device).
 *
 *  params   : ndim 2 or 3, form PTD_PLAIN, precision PTD_FP64 or PTD_FP32 (the device buffers hold
 *             double or float accordingly), arith as for contexts.
 *  stream   : hipStream_t to run on (NULL: the null stream); the call returns after it has synchronised.
 *  H0, H1   : the two H buffers (nx*ny*nz each); BOTH must hold the Dirichlet boundary values, which no
 *             sweep writes. Hold (nx*ny*nz), D (the step case, nx*ny*nz; NULL otherwise).
 *  qx0..qz1 : the two buffers of each flux array, reference shapes (qz NULL in 2D).
 *  parity   : in: which buffer (0 or 1) of H and of the fluxes holds the current iterate; out: the
 *             buffer holding the result.
 * Iteration counts, errors and results are bit-identical to ptd_pt_solve on a context holding the same
 * fields. */
#define PTD_NOALLOC_MAX_PARTIALS 131072
int32_t ptd_pt_solve_noalloc(const ptd_params* params, int32_t device, void* stream, void* H0, void* H1,
                             const void* Hold, const void* D, void* qx0, void* qx1, void* qy0, void* qy1,
                             void* qz0, void* qz1, int32_t* parity, int64_t* iters, double* err_rms,
                             double* err_max);

/* ---- gradients --------------------------------------------------------------------------------
 *
 * No reference counterpart: the reference has no derivative. Runs ptd_solve(nt) from the context's
 * current state while keeping H^0..H^nt on the device, then the adjoint of the nt backward-Euler steps,
 * and returns
 *   *loss    = L = 1/2 sum over the inner cells of (H^nt - H_obs)^2,
 *   grad_d   = dL/dD per cell (nx*ny*nz; in the lin case the sensitivity to a per-cell D at d0; must be NULL
 *              for PTD_NONLIN, where D = H^3 is no parameter),
 *   grad_h0  = dL/dH0 per cell (inner cells: lam_1/dt; boundary cells: their Dirichlet coupling),
 *   iters_fwd[nt], iters_adj[nt] = PT iterations of each forward / adjoint solve.
 * lin / step: the step operator A = I/dt - div(av(D) grad .) is symmetric, so each adjoint solve A mu = r is
 * the same PT solve with homogeneous boundary values. D = H^3: the step Jacobian J is not symmetric; each
 * adjoint solve J^T lam = r runs the transpose of the PT iteration linearised at that step's H (same
 * damping, same convergence rate). Every adjoint solve stops at the reduction relative to its right-hand
 * side that the forward solves reach (tol * dt / rms(H^nt) relative). Any output may be NULL except as
 * stated. On success the context ends as after ptd_solve(nt); after an error (NaN, allocation) its fields
 * are unspecified and should be set again. Allocates (nt + 4) grid-sized float64 buffers and a copy of the
 * fluxes for the call (D = H^3: 3 + 2*ndim more), PTD_ERR_NOMEM if they do not fit. float64 only. */
int32_t ptd_grad_solve(ptd_ctx* ctx, int32_t nt, const double* h_obs, int64_t count, double* loss,
                       double* grad_d, double* grad_h0, int64_t* iters_fwd, int64_t* iters_adj);

/* ---- batches of independent solves ------------------------------------------------------------
 *
 * No reference counterpart: the reference solves one problem per script run. A batch holds n
 * independent float64 problems (2D or 3D, lin / step / nonlin, each with its own grid and scalars,
 * the converging scripts' PTD_PLAIN form, one nout for all, itmax a multiple of nout, nout even) and
 * solves them together: one launch per (sweep kind, dcase, arith) group covers every member's tiles, each
 * member's PT loop (checks, stopping, Hold .= H, physical steps) runs on the device, and each group's check
 * block is a HIP graph replayed on the group's own stream (the groups advance independently). Every
 * member's result, iteration counts included, is bit-identical to ptd_solve on that problem alone. */
typedef struct ptd_batch ptd_batch;

/* params[n], 1 <= n <= 1024. Allocates every member's fields (zeroed) on `device`. */
int32_t ptd_batch_create(const ptd_params* params, int32_t n, int32_t device, ptd_batch** out);
int32_t ptd_batch_destroy(ptd_batch* batch);
int32_t ptd_batch_size(const ptd_batch* batch);
int64_t ptd_batch_device_bytes(const ptd_batch* batch);
/* Fields of one member, reference layout as ptd_set_field / ptd_get_field: H (set writes both
 * buffers), HOLD and D (step case) can be set; H, HOLD, D, QX..QZ can be read. */
int32_t ptd_batch_set_field(ptd_batch* batch, int32_t member, int32_t field, const double* host, int64_t count);
int32_t ptd_batch_get_field(ptd_batch* batch, int32_t member, int32_t field, double* host, int64_t count);
/* nt physical steps of every member (ptd_solve of each, together). iters[n*nt] / errs[n*nt] receive
 * per-member, per-step iteration counts and final errors (member-major), latency_ms[n] the time from
 * the start of the call until each member finished, status[n] PTD_OK or PTD_ERR_NAN, *total_ms the
 * wall time of the call; any output may be NULL. The first call also picks the tiling of the
 * non-checking launches by timing check blocks with each candidate in turn, three rounds, keeping the least
 * median (results do not depend on it). */
int32_t ptd_batch_solve(ptd_batch* batch, int32_t nt, int64_t* iters, double* errs, double* latency_ms,
                        int32_t* status, double* total_ms);
/* The tiling chosen by the first solve: z chunk depth of the 3D launches and rows per block of the 2D
 * launches, and the measured time of one check block per candidate (PTD_ERR_STATE before a solve). */
int32_t ptd_batch_tiling(const ptd_batch* batch, int32_t* tz3d, int32_t* rows2d, double* candidate_ms, int32_t ncand);
/* Where the last ptd_batch_solve spent its time, per member group (members sharing a sweep kind, D case and
 * arithmetic run as one launch per sweep): *ngroups, and for the first `cap` groups the sweep kind (3: 3D
 * two- / three-level, 22: 2D two-level, 20: 2D LDS-resident, 2: 2D single-level), the D case, the member count and the device time the
 * group's stream took, summed over the replays (ms; the groups run concurrently, so the times overlap). */
int32_t ptd_batch_group_times(const ptd_batch* batch, int32_t* ngroups, int32_t* kind, int32_t* dcase,
                              int32_t* members, double* ms, int32_t cap);

#ifdef __cplusplus
}
#endif
#endif /* PTD_H */
