/*
 * ptd_oracle.c -- CPU restatement of the reference PT diffusion solvers (test infrastructure only;
 * the product path never links this file). Build: oracle/Makefile, with -ffp-contract=off so that
 * every expression rounds exactly as the reference's Julia source does.
 *
 * Every function cites the reference lines it restates. Macro semantics of ParallelStencil's
 * FiniteDifferences{1,2,3}D (@all, @inn, @d_xa, @d_xi, @av_xi, @maxloc) are the ones the reference's
 * own perf kernels spell out with explicit indices (scripts/diff_3D/diff_3D_nonlin_perf.jl:19-49).
 */
#include "ptd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PTDO_PI 3.141592653589793115997963468544185161590576171875 /* Float64(pi) */

/* ------------------------------------------------------------------------------------------- */
/* geometry                                                                                    */
/* ------------------------------------------------------------------------------------------- */

typedef struct geo {
    int64_t nx, ny, nz;        /* cell counts (1 on absent axes) */
    int64_t mx, my, mz;        /* inner counts: n-2 on active axes, 1 on absent ones */
    int64_t ox, oy, oz;        /* offset of the first inner cell: 1 on active axes, 0 otherwise */
    int hy, hz;                /* y / z axes active */
} geo;

static geo make_geo(const ptdo_case* c) {
    geo g;
    g.nx = c->nx;
    g.ny = c->ndim >= 2 ? c->ny : 1;
    g.nz = c->ndim >= 3 ? c->nz : 1;
    g.hy = c->ndim >= 2;
    g.hz = c->ndim >= 3;
    g.mx = g.nx - 2;
    g.my = g.hy ? g.ny - 2 : 1;
    g.mz = g.hz ? g.nz - 2 : 1;
    g.ox = 1;
    g.oy = g.hy ? 1 : 0;
    g.oz = g.hz ? 1 : 0;
    return g;
}

#define IH(g, i, j, k) ((i) + (g).nx * ((j) + (g).ny * (k)))
/* face arrays, indexed by face f along their own axis and inner cell index elsewhere */
#define IQX(g, f, jj, kk) ((f) + ((g).nx - 1) * ((jj) + (g).my * (kk)))
#define IQY(g, ii, f, kk) ((ii) + (g).mx * ((f) + ((g).ny - 1) * (kk)))
#define IQZ(g, ii, jj, f) ((ii) + (g).mx * ((jj) + (g).my * (f)))
#define IR(g, ii, jj, kk) ((ii) + (g).mx * ((jj) + (g).my * (kk)))

/* ------------------------------------------------------------------------------------------- */
/* derived numerics                                                                            */
/* ------------------------------------------------------------------------------------------- */

/* "Derived numerics" blocks: diff_1D_lin.jl:42-50, diff_2D_lin.jl:47-53, diff_2D_linstep.jl:65-71,
 * diff_2D_nonlin.jl:52-58, diff_2D_nonlin_perf.jl:64-71, diff_3D_lin_multixpu.jl:54-62,
 * diff_3D_linstep_multixpu.jl:78-86, diff_3D_nonlin_multixpu.jl:66-75, diff_3D_nonlin_perf.jl:70-78. */
void ptdo_derive(const ptdo_case* c, ptdo_derived* d) {
    memset(d, 0, sizeof(*d));
    d->dx = c->lx / (double)c->nx;
    d->dy = c->ndim >= 2 ? c->ly / (double)c->ny : 0.0;
    d->dz = c->ndim >= 3 ? c->lz / (double)c->nz : 0.0;
    d->_dx = 1.0 / d->dx;
    d->_dy = c->ndim >= 2 ? 1.0 / d->dy : 0.0;
    d->_dz = c->ndim >= 3 ? 1.0 / d->dz : 0.0;
    d->_dt = 1.0 / c->dt;
    double mind = d->dx;
    double maxl = c->lx;
    if (c->ndim >= 2) { mind = fmin(mind, d->dy); maxl = fmax(maxl, c->ly); }
    if (c->ndim >= 3) { mind = fmin(mind, d->dz); maxl = fmax(maxl, c->lz); }
    d->vpdt = c->cfl * mind;
    d->max_l = maxl;
    d->max_l2 = maxl * maxl;
    /* Re = pi + sqrt(pi^2 + (max_l^2 / D / dt)) (lin), (max_l^2 / max(D1,D2)) / dt (step);
     * both are ((max_l^2 / D) / dt). */
    double Dre = c->dcase == PTDO_STEP ? c->dmax : c->d0;
    d->re = PTDO_PI + sqrt(PTDO_PI * PTDO_PI + (d->max_l2 / Dre) / c->dt);
    d->theta_r = d->max_l / d->vpdt / d->re;
    d->dtau_rho = d->vpdt * d->max_l / c->d0 / d->re;
}

/* ------------------------------------------------------------------------------------------- */
/* initial conditions                                                                          */
/* ------------------------------------------------------------------------------------------- */

/* Base.LinRange getindex: lerpi(j, d, a, b) = (1-t)*a + t*b, t = j/d, d = n-1. */
void ptdo_linrange(double a, double b, int64_t n, double* out) {
    int64_t d = n > 1 ? n - 1 : 1;
    for (int64_t i = 0; i < n; ++i) {
        double t = (double)i / (double)d;
        out[i] = (1.0 - t) * a + t * b;
    }
}

static double* axis_shifted_linrange(int64_t n, double l) {
    /* (LinRange(dx/2, l-dx/2, n) .- l/2) is LinRange(dx/2 - l/2, (l-dx/2) - l/2, n). */
    double dxl = l / (double)n;
    double* v = (double*)malloc(sizeof(double) * (size_t)n);
    ptdo_linrange(dxl / 2.0 - l / 2.0, (l - dxl / 2.0) - l / 2.0, n, v);
    return v;
}

static double* axis_linrange_minus(int64_t n, double l) {
    /* xc[ix] - 0.5*l, xc = LinRange(dx/2, l-dx/2, n) (diff_3D_nonlin_perf.jl:77,89). */
    double dxl = l / (double)n;
    double* v = (double*)malloc(sizeof(double) * (size_t)n);
    ptdo_linrange(dxl / 2.0, l - dxl / 2.0, n, v);
    for (int64_t i = 0; i < n; ++i) v[i] = v[i] - 0.5 * l;
    return v;
}

static double* axis_xg(int64_t n, double l) {
    /* x_g(ix,dx,A) - 0.5*l + dx/2 with x_g = (ix-1)*dx on a single rank (diff_3D_*_multixpu.jl:86). */
    double dxl = l / (double)n;
    double* v = (double*)malloc(sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        double xg = (double)i * dxl;
        v[i] = (xg - 0.5 * l) + dxl / 2.0;
    }
    return v;
}

void ptdo_ic(int32_t kind, int32_t ndim, int64_t nx, int64_t ny, int64_t nz,
             double lx, double ly, double lz, double amp, double bval, double* H0) {
    if (ndim < 2) ny = 1;
    if (ndim < 3) nz = 1;
    double *ax, *ay = NULL, *az = NULL;
    if (kind == 0) {
        ax = axis_shifted_linrange(nx, lx);
        if (ndim >= 2) ay = axis_shifted_linrange(ny, ly);
        if (ndim >= 3) az = axis_shifted_linrange(nz, lz);
    } else if (kind == 2) {
        ax = axis_xg(nx, lx);
        if (ndim >= 2) ay = axis_xg(ny, ly);
        if (ndim >= 3) az = axis_xg(nz, lz);
    } else {
        ax = axis_linrange_minus(nx, lx);
        if (ndim >= 2) ay = axis_linrange_minus(ny, ly);
        if (ndim >= 3) az = axis_linrange_minus(nz, lz);
    }
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                double a = ax[i];
                double e = -(a * a);  /* -(x)^2 == (-x)*x exactly */
                if (ndim >= 2) e = e - ay[j] * ay[j];
                if (ndim >= 3) e = e - az[k] * az[k];
                double h = exp(e);
                if (kind == 3) {
                    int bnd = (i == 0 || i == nx - 1);
                    if (ndim >= 2) bnd |= (j == 0 || j == ny - 1);
                    if (ndim >= 3) bnd |= (k == 0 || k == nz - 1);
                    h = bnd ? bval : 1.0 + amp * h;
                }
                H0[i + nx * (j + ny * k)] = h;
            }
    free(ax);
    free(ay);
    free(az);
}

void ptdo_dstep(int32_t kind, int32_t ndim, int64_t nx, int64_t ny, int64_t nz,
                double lx, double ly, double lz, double d1, double d2, double* D) {
    if (ndim < 2) ny = 1;
    if (ndim < 3) nz = 1;
    double dx = lx / (double)nx, dy = ndim >= 2 ? ly / (double)ny : 0.0;
    int64_t cx = (int64_t)ceil((double)nx / 2.2), cy = (int64_t)ceil((double)ny / 2.2);
    double* xc = (double*)malloc(sizeof(double) * (size_t)nx);
    ptdo_linrange(dx / 2.0, lx - dx / 2.0, nx, xc);
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                int low = 0;
                if (kind == 0) {          /* diff_1D_linstep.jl:69-70, diff_2D_linstep.jl:81-83 */
                    low = i < cx;
                    if (ndim >= 2) low |= j < cy;
                } else if (kind == 1) {   /* diff_3D_linstep_multixpu.jl:99-104 */
                    low = ((double)i * dx < lx / 2.2) || (ndim >= 2 && (double)j * dy < ly / 2.2);
                } else {                  /* step at x = lx/2 on cell centres */
                    low = xc[i] < lx / 2.0;
                }
                D[i + nx * (j + ny * k)] = low ? d2 : d1;
            }
    free(xc);
}

/* ------------------------------------------------------------------------------------------- */
/* kernels                                                                                     */
/* ------------------------------------------------------------------------------------------- */

/* Face coefficient between cells a and b:
 *  lin   : D                                                (diff_3D_lin_multixpu.jl:25)
 *  step  : @av_xi(D) = (D[a] + D[b])*0.5                       (diff_3D_linstep_multixpu.jl:30)
 *  nonlin: plain @av_xi(H)*@av_xi(H)*@av_xi(H)                 (diff_3D_nonlin_multixpu.jl:25)
 *          perf  0.5*(Ha+Hb) * 0.5*(Ha+Hb) * 0.5*(Ha+Hb), a left fold (diff_3D_nonlin_perf.jl:20) */
static inline double face_coef(const ptdo_case* c, const double* H, const double* D, int64_t a, int64_t b) {
    if (c->dcase == PTDO_LIN) return c->d0;
    if (c->dcase == PTDO_STEP) return (D[a] + D[b]) * 0.5;
    double s = H[a] + H[b];
    if (c->form == PTDO_PERF) return 0.5 * s * 0.5 * s * 0.5 * s;
    double av = s * 0.5;
    return av * av * av;
}

/* Face damping theta_r_dtau:
 *  lin/step: scalar max_l/Vpdt/Re                               (diff_3D_lin_multixpu.jl:61)
 *  nonlin  : max_l / Vpdt / (pi + sqrt(pi*pi + max_l2 / H3 / dt)) * Resc   (plain, :28,:32)
 *                                     ... max_l2 / H3 * _dt ...            (perf, diff_3D_nonlin_perf.jl:23,27) */
static inline double face_theta(const ptdo_case* c, const ptdo_derived* d, double coef) {
    if (c->dcase != PTDO_NONLIN) return d->theta_r;
    double re;
    if (c->form == PTDO_PERF) re = PTDO_PI + sqrt(PTDO_PI * PTDO_PI + d->max_l2 / coef * d->_dt);
    else re = PTDO_PI + sqrt(PTDO_PI * PTDO_PI + d->max_l2 / coef / c->dt);
    return d->max_l / d->vpdt / re * c->resc;
}

/* Cell dtau_rho:
 *  lin   : scalar Vpdt*max_l/D/Re                               (diff_3D_lin_multixpu.jl:62)
 *  step  : array from compute_iter_params!                      (diff_3D_linstep_multixpu.jl:24-27)
 *  nonlin: Vpdt * max_l / H3 / (pi + sqrt(pi*pi + max_l2/H3/dt)) * Resc  (diff_3D_nonlin_multixpu.jl:31,35;
 *          perf uses *_dt, diff_3D_nonlin_perf.jl:26,30) */
static inline double cell_dtau(const ptdo_case* c, const ptdo_derived* d, const double* H, const double* dtau,
                               int64_t h, int64_t r) {
    if (c->dcase == PTDO_LIN) return d->dtau_rho;
    if (c->dcase == PTDO_STEP) return dtau[r];
    double h3 = H[h] * H[h] * H[h];
    double re;
    if (c->form == PTDO_PERF) re = PTDO_PI + sqrt(PTDO_PI * PTDO_PI + d->max_l2 / h3 * d->_dt);
    else re = PTDO_PI + sqrt(PTDO_PI * PTDO_PI + d->max_l2 / h3 / c->dt);
    return d->vpdt * d->max_l / h3 / re * c->resc;
}

/* Damped flux: plain (q*theta - coef*dH/dx)/(1+theta) (diff_3D_nonlin_multixpu.jl:38);
 * perf (q*theta - coef*_dx*dH)/(1+theta) (diff_3D_nonlin_perf.jl:33). */
static inline double damped_flux(const ptdo_case* c, double q, double theta, double coef, double dH,
                                 double h, double _h) {
    if (c->form == PTDO_PERF) return (q * theta - coef * _h * dH) / (1.0 + theta);
    return (q * theta - coef * dH / h) / (1.0 + theta);
}

/* Residual flux: plain -coef*dH/dx (diff_3D_nonlin_multixpu.jl:41); perf -coef*_dx*dH
 * (diff_3D_nonlin_perf.jl:45). */
static inline double plain_flux(const ptdo_case* c, double coef, double dH, double h, double _h) {
    if (c->form == PTDO_PERF) return -coef * _h * dH;
    return -coef * dH / h;
}

/* compute_iter_params!: dtau_rho = Vpdt*max_l/@maxloc(D)/Re over the 3 / 5 / 7-point neighbourhood
 * (diff_1D_linstep.jl:17-20, diff_2D_linstep.jl:18-21, diff_3D_linstep_multixpu.jl:24-27). */
void ptdo_iter_params(const ptdo_case* c, const double* D, double* dtau) {
    geo g = make_geo(c);
    ptdo_derived d;
    ptdo_derive(c, &d);
#pragma omp parallel for schedule(static)
    for (int64_t kk = 0; kk < g.mz; ++kk)
        for (int64_t jj = 0; jj < g.my; ++jj)
            for (int64_t ii = 0; ii < g.mx; ++ii) {
                int64_t i = ii + g.ox, j = jj + g.oy, k = kk + g.oz;
                int64_t h = IH(g, i, j, k);
                double m = fmax(fmax(D[h - 1], D[h + 1]), D[h]);
                if (g.hy) m = fmax(m, fmax(D[h - g.nx], D[h + g.nx]));
                if (g.hz) m = fmax(m, fmax(D[h - g.nx * g.ny], D[h + g.nx * g.ny]));
                dtau[IR(g, ii, jj, kk)] = d.vpdt * d.max_l / m / d.re;
            }
}

/* compute_flux!: every face of qx (nx-1, my, mz), qy (mx, ny-1, mz), qz (mx, my, nz-1)
 * (@all over each q array: diff_3D_nonlin_multixpu.jl:37-45, diff_3D_nonlin_perf.jl:32-37,
 * diff_2D_*.jl compute_flux!, diff_1D_*.jl compute_flux!). In the plain form the residual fluxes q2 are
 * written from the same (pre-update) H. */
void ptdo_flux(const ptdo_case* c, const double* H, const double* D,
               double* qx, double* qy, double* qz, double* qx2, double* qy2, double* qz2) {
    geo g = make_geo(c);
    ptdo_derived d;
    ptdo_derive(c, &d);
    const int plain = c->form == PTDO_PLAIN;
    /* x faces */
#pragma omp parallel for schedule(static)
    for (int64_t kk = 0; kk < g.mz; ++kk)
        for (int64_t jj = 0; jj < g.my; ++jj)
            for (int64_t f = 0; f < g.nx - 1; ++f) {
                int64_t a = IH(g, f, jj + g.oy, kk + g.oz), b = a + 1;
                double coef = face_coef(c, H, D, a, b);
                double th = face_theta(c, &d, coef);
                double dH = H[b] - H[a];
                int64_t q = IQX(g, f, jj, kk);
                qx[q] = damped_flux(c, qx[q], th, coef, dH, d.dx, d._dx);
                if (plain && qx2) qx2[q] = plain_flux(c, coef, dH, d.dx, d._dx);
            }
    if (g.hy) {
#pragma omp parallel for schedule(static)
        for (int64_t kk = 0; kk < g.mz; ++kk)
            for (int64_t f = 0; f < g.ny - 1; ++f)
                for (int64_t ii = 0; ii < g.mx; ++ii) {
                    int64_t a = IH(g, ii + 1, f, kk + g.oz), b = a + g.nx;
                    double coef = face_coef(c, H, D, a, b);
                    double th = face_theta(c, &d, coef);
                    double dH = H[b] - H[a];
                    int64_t q = IQY(g, ii, f, kk);
                    qy[q] = damped_flux(c, qy[q], th, coef, dH, d.dy, d._dy);
                    if (plain && qy2) qy2[q] = plain_flux(c, coef, dH, d.dy, d._dy);
                }
    }
    if (g.hz) {
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < g.nz - 1; ++f)
            for (int64_t jj = 0; jj < g.my; ++jj)
                for (int64_t ii = 0; ii < g.mx; ++ii) {
                    int64_t a = IH(g, ii + 1, jj + 1, f), b = a + g.nx * g.ny;
                    double coef = face_coef(c, H, D, a, b);
                    double th = face_theta(c, &d, coef);
                    double dH = H[b] - H[a];
                    int64_t q = IQZ(g, ii, jj, f);
                    qz[q] = damped_flux(c, qz[q], th, coef, dH, d.dz, d._dz);
                    if (plain && qz2) qz2[q] = plain_flux(c, coef, dH, d.dz, d._dz);
                }
    }
}

/* compute_flux_res! (perf scripts): q2 from the current H (diff_3D_nonlin_perf.jl:44-49,
 * diff_2D_nonlin_perf.jl:38-42). */
void ptdo_flux_res(const ptdo_case* c, const double* H, const double* D,
                   double* qx2, double* qy2, double* qz2) {
    geo g = make_geo(c);
    ptdo_derived d;
    ptdo_derive(c, &d);
#pragma omp parallel for schedule(static)
    for (int64_t kk = 0; kk < g.mz; ++kk)
        for (int64_t jj = 0; jj < g.my; ++jj)
            for (int64_t f = 0; f < g.nx - 1; ++f) {
                int64_t a = IH(g, f, jj + g.oy, kk + g.oz), b = a + 1;
                qx2[IQX(g, f, jj, kk)] = plain_flux(c, face_coef(c, H, D, a, b), H[b] - H[a], d.dx, d._dx);
            }
    if (g.hy) {
#pragma omp parallel for schedule(static)
        for (int64_t kk = 0; kk < g.mz; ++kk)
            for (int64_t f = 0; f < g.ny - 1; ++f)
                for (int64_t ii = 0; ii < g.mx; ++ii) {
                    int64_t a = IH(g, ii + 1, f, kk + g.oz), b = a + g.nx;
                    qy2[IQY(g, ii, f, kk)] = plain_flux(c, face_coef(c, H, D, a, b), H[b] - H[a], d.dy, d._dy);
                }
    }
    if (g.hz) {
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < g.nz - 1; ++f)
            for (int64_t jj = 0; jj < g.my; ++jj)
                for (int64_t ii = 0; ii < g.mx; ++ii) {
                    int64_t a = IH(g, ii + 1, jj + 1, f), b = a + g.nx * g.ny;
                    qz2[IQZ(g, ii, jj, f)] = plain_flux(c, face_coef(c, H, D, a, b), H[b] - H[a], d.dz, d._dz);
                }
    }
}

/* Flux divergence at inner cell (ii,jj,kk):
 *  plain: @d_xa(qx)/dx + @d_ya(qy)/dy + @d_za(qz)/dz   (diff_3D_nonlin_multixpu.jl:48)
 *  perf : _dx*(qx[i+1]-qx[i]) + _dy*(...) + _dz*(...)   (diff_3D_nonlin_perf.jl:40)
 * summed left to right; 1D/2D drop the absent terms. */
static inline double divergence(const ptdo_case* c, const geo* g, const ptdo_derived* d,
                                const double* qx, const double* qy, const double* qz,
                                int64_t ii, int64_t jj, int64_t kk) {
    const int perf = c->form == PTDO_PERF;
    double ddx = qx[IQX(*g, ii + 1, jj, kk)] - qx[IQX(*g, ii, jj, kk)];
    double s = perf ? d->_dx * ddx : ddx / d->dx;
    if (g->hy) {
        double ddy = qy[IQY(*g, ii, jj + 1, kk)] - qy[IQY(*g, ii, jj, kk)];
        s = s + (perf ? d->_dy * ddy : ddy / d->dy);
    }
    if (g->hz) {
        double ddz = qz[IQZ(*g, ii, jj, kk + 1)] - qz[IQZ(*g, ii, jj, kk)];
        s = s + (perf ? d->_dz * ddz : ddz / d->dz);
    }
    return s;
}

/* compute_update! on @inn(H), in place:
 *  plain: (H + dtau*(Hold/dt - div)) / (1.0 + dtau/dt)        (diff_3D_nonlin_multixpu.jl:48)
 *  perf : (H + dtau*(_dt*Hold - div)) / (1.0 + _dt*dtau)      (diff_3D_nonlin_perf.jl:40) */
void ptdo_update(const ptdo_case* c, double* H, const double* Hold, const double* dtau,
                 const double* qx, const double* qy, const double* qz) {
    geo g = make_geo(c);
    ptdo_derived d;
    ptdo_derive(c, &d);
    const int perf = c->form == PTDO_PERF;
#pragma omp parallel for schedule(static)
    for (int64_t kk = 0; kk < g.mz; ++kk)
        for (int64_t jj = 0; jj < g.my; ++jj)
            for (int64_t ii = 0; ii < g.mx; ++ii) {
                int64_t h = IH(g, ii + g.ox, jj + g.oy, kk + g.oz);
                double dt_rho = cell_dtau(c, &d, H, dtau, h, IR(g, ii, jj, kk));
                double div = divergence(c, &g, &d, qx, qy, qz, ii, jj, kk);
                if (perf) H[h] = (H[h] + dt_rho * (d._dt * Hold[h] - div)) / (1.0 + d._dt * dt_rho);
                else H[h] = (H[h] + dt_rho * (Hold[h] / c->dt - div)) / (1.0 + dt_rho / c->dt);
            }
}

/* check_res! + err:
 *  plain: ResH = -(H - Hold)/dt - div(q2)                   (diff_3D_nonlin_multixpu.jl:53)
 *  perf : ResH = -_dt*(H - Hold) - div(q2)                  (diff_2D_nonlin_perf.jl:45; the 3D perf
 *         check_res! reads qHx instead of qHx2 but is unreachable there, diff_3D_nonlin_perf.jl:52,103)
 *  err  = norm(ResH)/sqrt(length(ResH))                     (diff_3D_nonlin_perf.jl:106, norm_g :20)
 * Sum of squares is accumulated per (jj,kk) row and the rows are then summed in order, so the result
 * does not depend on the thread count. */
double ptdo_check_res(const ptdo_case* c, const double* H, const double* Hold,
                      const double* qx2, const double* qy2, const double* qz2,
                      double* ResH, double* err_max) {
    geo g = make_geo(c);
    ptdo_derived d;
    ptdo_derive(c, &d);
    const int perf = c->form == PTDO_PERF;
    int64_t rows = g.my * g.mz;
    double* part = (double*)calloc((size_t)rows, sizeof(double));
    double* pmax = (double*)calloc((size_t)rows, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < rows; ++row) {
        int64_t jj = row % g.my, kk = row / g.my;
        double s = 0.0, m = 0.0;
        for (int64_t ii = 0; ii < g.mx; ++ii) {
            int64_t h = IH(g, ii + g.ox, jj + g.oy, kk + g.oz);
            double div = divergence(c, &g, &d, qx2, qy2, qz2, ii, jj, kk);
            double r = perf ? -d._dt * (H[h] - Hold[h]) - div : -(H[h] - Hold[h]) / c->dt - div;
            if (ResH) ResH[IR(g, ii, jj, kk)] = r;
            s += r * r;
            double a = fabs(r);
            if (!(a <= m)) m = a; /* propagates NaN */
        }
        part[row] = s;
        pmax[row] = m;
    }
    double s = 0.0, m = 0.0;
    for (int64_t row = 0; row < rows; ++row) {
        s += part[row];
        if (!(pmax[row] <= m)) m = pmax[row];
    }
    free(part);
    free(pmax);
    if (err_max) *err_max = m;
    double len = c->norm_len > 0.0 ? c->norm_len : (double)(g.mx * g.my * g.mz);
    return sqrt(s) / sqrt(len);
}

static size_t nq(const geo* g, int axis) {
    if (axis == 0) return (size_t)((g->nx - 1) * g->my * g->mz);
    if (axis == 1) return g->hy ? (size_t)(g->mx * (g->ny - 1) * g->mz) : 0;
    return g->hz ? (size_t)(g->mx * g->my * (g->nz - 1)) : 0;
}

void ptdo_iterate(const ptdo_case* c, double* H, const double* Hold, const double* D,
                  const double* dtau, double* qx, double* qy, double* qz, int64_t niter) {
    for (int64_t it = 0; it < niter; ++it) {
        ptdo_flux(c, H, D, qx, qy, qz, NULL, NULL, NULL);
        ptdo_update(c, H, Hold, dtau, qx, qy, qz);
    }
}

/* Physical-time loop with the PT loop inside (diff_3D_nonlin_multixpu.jl:99-119; perf variant
 * diff_3D_nonlin_perf.jl:93-112; reaction variant diff_1D_react.jl:55-67 with nt = 1). */
int ptdo_solve(const ptdo_case* c, double* H, double* Hold, const double* D,
               int64_t* iters, double* errs) {
    geo g = make_geo(c);
    size_t nqx = nq(&g, 0), nqy = nq(&g, 1), nqz = nq(&g, 2);
    size_t nr = (size_t)(g.mx * g.my * g.mz);
    size_t nh = (size_t)(g.nx * g.ny * g.nz);
    double* qx = (double*)calloc(nqx, sizeof(double));
    double* qx2 = (double*)calloc(nqx, sizeof(double));
    double* qy = nqy ? (double*)calloc(nqy, sizeof(double)) : NULL;
    double* qy2 = nqy ? (double*)calloc(nqy, sizeof(double)) : NULL;
    double* qz = nqz ? (double*)calloc(nqz, sizeof(double)) : NULL;
    double* qz2 = nqz ? (double*)calloc(nqz, sizeof(double)) : NULL;
    double* ResH = (double*)calloc(nr, sizeof(double));
    double* dtau = NULL;
    if (c->dcase == PTDO_STEP) {
        dtau = (double*)calloc(nr, sizeof(double));
        ptdo_iter_params(c, D, dtau);
    }
    int status = 0;
    for (int32_t it = 0; it < c->nt; ++it) {
        int64_t iter = 0;
        double err = 2.0 * c->tol;
        while (err > c->tol && iter < c->itmax) {
            ptdo_flux(c, H, D, qx, qy, qz, qx2, qy2, qz2);
            ptdo_update(c, H, Hold, dtau, qx, qy, qz);
            iter += 1;
            if (iter % c->nout == 0) {
                if (c->form == PTDO_PERF) ptdo_flux_res(c, H, D, qx2, qy2, qz2);
                err = ptdo_check_res(c, H, Hold, qx2, qy2, qz2, ResH, NULL);
            }
        }
        if (iters) iters[it] = iter;
        if (errs) errs[it] = err;
        memcpy(Hold, H, nh * sizeof(double));
        if (isnan(err)) { status = 1; break; }
    }
    free(qx); free(qx2); free(qy); free(qy2); free(qz); free(qz2); free(ResH); free(dtau);
    return status;
}
