/*
 * ptd_oracle.h -- CPU restatement of the reference's PT diffusion solvers.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product path: only the test
 * suite, the smoke check and the benchmark's CPU-baseline leg may load it. The product library
 * (libptd) never links or calls it.
 *
 * It restates, in scalar C with IEEE double arithmetic and no FMA contraction (-ffp-contract=off),
 * the kernels and the pseudo-transient loop of these reference scripts, keeping each expression's
 * operation order (Julia evaluates a/b/c as (a/b)/c and a*b*c as (a*b)*c):
 *   scripts/diff_1D_{lin,linstep,nonlin,react}.jl
 *   scripts/diff_2D_{lin,linstep,nonlin,nonlin_perf,react}.jl
 *   scripts/diff_3D/diff_3D_{lin,linstep,nonlin}_multixpu.jl
 *   scripts/diff_3D/diff_3D_nonlin_perf.jl, scripts/diff_3D/diff_3D_nonlin_multixpu_perf.jl
 * Array convention (as the reference's Data.Array): column-major, x fastest, 0-based here.
 *   H, Hold, D : (nx, ny, nz)
 *   qx, qx2    : (nx-1, my, mz)      qy, qy2 : (mx, ny-1, mz)      qz, qz2 : (mx, my, nz-1)
 *   ResH, dtau : (mx, my, mz)        with m = n-2 on an active axis and 1 on an absent one
 *                                    (1D: ny = nz = 1; 2D: nz = 1).
 */
#ifndef PTD_ORACLE_H
#define PTD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PTDO_LIN = 0, PTDO_STEP = 1, PTDO_NONLIN = 2 };
enum { PTDO_PLAIN = 0, PTDO_PERF = 1 };

typedef struct ptdo_case {
    int32_t ndim;   /* 1, 2 or 3 */
    int32_t dcase;  /* PTDO_LIN, PTDO_STEP, PTDO_NONLIN */
    int32_t form;   /* PTDO_PLAIN: /dx, /dt, residual flux from the pre-update H (diff_*_{lin,..}.jl)
                       PTDO_PERF : *_dx, *_dt, residual flux from the current H (diff_*_perf.jl) */
    int32_t nout;   /* residual check interval */
    int64_t nx, ny, nz;
    int64_t itmax;  /* PT iteration cap per physical step */
    int32_t nt;     /* physical steps */
    int32_t pad_;
    double lx, ly, lz;
    double dt;      /* physical step (theta_kin for the reaction scripts) */
    double tol;
    double cfl;
    double resc;    /* iteration parameter scaling (1/1.2 for 2D/3D nonlin, 1 elsewhere) */
    double d0;      /* scalar D (lin) */
    double dmax;    /* max(D1, D2), used by the step case's Re */
    double norm_len;/* err = ||ResH|| / sqrt(norm_len); 0 selects length(ResH) */
} ptdo_case;

/* Scalars the scripts derive before the loop ("Derived numerics"). */
typedef struct ptdo_derived {
    double dx, dy, dz, _dx, _dy, _dz, _dt;
    double vpdt;        /* Vpdtau = CFL*min(dx,dy,dz) */
    double max_l, max_l2;
    double re;          /* lin/step scalar Re */
    double theta_r;     /* lin/step scalar theta_r_dtau */
    double dtau_rho;    /* lin scalar dtau_rho */
} ptdo_derived;

void ptdo_derive(const ptdo_case* c, ptdo_derived* d);

/* Julia LinRange(a, b, n)[i] for i = 0..n-1 (Base lerpi: (1-t)*a + t*b with t = i/(n-1)). */
void ptdo_linrange(double a, double b, int64_t n, double* out);

/* Initial conditions.
 * kind 0: 1D/2D scripts: exp.(-(xc .- lx/2).^2 .- ((yc .- ly/2)').^2); the shifted coordinate is the
 *         LinRange(start - lx/2, stop - lx/2) that Julia's range broadcasting returns.
 * kind 1: 3D perf script comprehension exp(-(xc[ix]-0.5*lx)^2 - ...).
 * kind 2: 3D multixpu scripts with ImplicitGlobalGrid x_g(ix) = (ix-1)*dx on one rank:
 *         exp(-(x_g-0.5*lx+dx/2)*(x_g-0.5*lx+dx/2) - ...).
 * kind 3: benchmark configuration: 1 + amp*exp(-r^2) inside, H = bval on the boundary cells.
 * ndim selects which of ly, lz are used. H0 has nx*ny*nz entries. */
void ptdo_ic(int32_t kind, int32_t ndim, int64_t nx, int64_t ny, int64_t nz,
             double lx, double ly, double lz, double amp, double bval, double* H0);

/* Step-function D fields.
 * kind 0: 1D/2D scripts: D[1:ceil(n/2.2)] = D2 along x (and along y in 2D), D1 elsewhere.
 * kind 1: 3D scripts: D2 where x_g < lx/2.2 or y_g < ly/2.2.
 * kind 2: benchmark configuration: D2 where the cell centre x < lx/2. */
void ptdo_dstep(int32_t kind, int32_t ndim, int64_t nx, int64_t ny, int64_t nz,
                double lx, double ly, double lz, double d1, double d2, double* D);

/* Kernel-level restatements (one call = one reference kernel launch). */
void ptdo_iter_params(const ptdo_case* c, const double* D, double* dtau);          /* compute_iter_params! */
void ptdo_flux(const ptdo_case* c, const double* H, const double* D,
               double* qx, double* qy, double* qz,
               double* qx2, double* qy2, double* qz2);                             /* compute_flux! (plain also writes q2) */
void ptdo_update(const ptdo_case* c, double* H, const double* Hold, const double* dtau,
                 const double* qx, const double* qy, const double* qz);            /* compute_update! */
void ptdo_flux_res(const ptdo_case* c, const double* H, const double* D,
                   double* qx2, double* qy2, double* qz2);                         /* compute_flux_res! (perf) */
double ptdo_check_res(const ptdo_case* c, const double* H, const double* Hold,
                      const double* qx2, const double* qy2, const double* qz2,
                      double* ResH, double* err_max);                              /* check_res! + norm */

/* The whole script loop. H (in: initial H incl. Dirichlet boundary values, out: final H), Hold (in:
 * initial Hold), D (step case only; may be NULL otherwise). q arrays start at zero and persist over
 * physical steps as in the reference. iters[nt], errs[nt] receive per-step counts and final errors.
 * Returns 0, or 1 if the error became NaN (the reference's error("NaN")). */
int ptdo_solve(const ptdo_case* c, double* H, double* Hold, const double* D,
               int64_t* iters, double* errs);

/* Fixed number of PT iterations within one physical step (no residual check, Hold fixed), with
 * caller-owned q arrays: one unit of the benchmark's CPU baseline. */
void ptdo_iterate(const ptdo_case* c, double* H, const double* Hold, const double* D,
                  const double* dtau, double* qx, double* qy, double* qz, int64_t niter);

#ifdef __cplusplus
}
#endif
#endif
