// ptd_grad.hip -- gradient of a loss on the final iterate of the lin / step solves (include/ptd.h:
// ptd_grad_solve): L = 1/2 sum over inner cells (H(T) - H_obs)^2 with respect to the per-cell D (step
// case; the scalar d0 spread over the cells in the lin case) and to the initial condition H0.
//
// For D independent of H every physical step solves A(D) H^n = H^{n-1}/dt with
// A = I/dt - div(av(D) grad .) on the inner cells (scripts/diff_3D/diff_3D_linstep_multixpu.jl:29-47),
// and A is symmetric. The adjoint of the nt steps is therefore
//     A mu_nt = H^nt - H_obs,     A mu_n = mu_{n+1}/dt   (n = nt-1 .. 1),
// i.e. the same PT loop run on mu with homogeneous Dirichlet values: Hold = (H^nt - H_obs)*dt for the
// first solve, then Hold .= H between solves. The forward iterates H^0..H^nt are kept on the device
// (checkpoints); after each adjoint solve one kernel gathers the step's face sensitivities
// dL/dD_f = (H_b - H_a)(mu_a - mu_b)/h^2 into the cells (half to each side), and the boundary coupling
// into dL/dH0; dL/dH0 = mu_1/dt on the inner cells. The context ends as after ptd_solve(nt).
// Each adjoint solve stops at the reduction relative to its right-hand side that the forward solves
// reach: tol_adj = tol * rms((H^nt - H_obs)*dt) / rms(H^nt) over the inner cells (the forward's right-hand
// side is Hold ~ H^nt, the adjoint's (H^nt - H_obs)*dt), so a small misfit does not end the adjoint
// before it starts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "ptd_internal.hpp"

extern "C" int32_t ptd_pt_solve(ptd_ctx* ctx, int64_t* iters, double* err_rms, double* err_max);
extern "C" int32_t ptd_assign(ptd_ctx* ctx);

namespace {

#define PTD_HIPG(expr)                                                                             \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            status = ptd::set_error(PTD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
            goto done;                                                                             \
        }                                                                                          \
    } while (0)

#define PTD_ALLOCG(ptr, bytes)                                                                     \
    do {                                                                                           \
        const size_t b_ = (bytes);                                                                 \
        if (hipMalloc((void**)(ptr), b_) != hipSuccess) {                                          \
            *(ptr) = nullptr;                                                                      \
            (void)hipGetLastError(); /* later launch checks read the thread's last error */        \
            status = ptd::set_error(PTD_ERR_NOMEM, "hipMalloc of " + std::to_string(b_) + " bytes failed"); \
            goto done;                                                                             \
        }                                                                                          \
    } while (0)

}  // namespace

extern "C" int32_t ptd_grad_solve(ptd_ctx* c, int32_t nt, const double* h_obs, int64_t count, double* loss,
                                  double* grad_d, double* grad_h0, int64_t* iters_fwd, int64_t* iters_adj) {
    if (!c) return ptd::set_error(PTD_ERR_ARG, "ctx is NULL");
    if (nt < 1 || !h_obs) return ptd::set_error(PTD_ERR_ARG, "need nt >= 1 and H_obs");
    if (count != c->nH) return ptd::set_error(PTD_ERR_ARG, "H_obs must have nx*ny*nz values");
    const bool nonlin = c->p.dcase == PTD_NONLIN;
    if (nonlin && grad_d) return ptd::set_error(PTD_ERR_ARG, "dL/dD is not defined for D = H^3: pass grad_d = NULL");
    if (c->p.precision != PTD_FP64) return ptd::set_error(PTD_ERR_STATE, "gradients are float64");
    if (hipSetDevice(c->device) != hipSuccess) return ptd::set_error(PTD_ERR_HIP, "hipSetDevice failed");

    const size_t nH = (size_t)c->nH;
    if ((size_t)nt + 1 > SIZE_MAX / (nH * 8))
        return ptd::set_error(PTD_ERR_NOMEM, "nt + 1 checkpoints of the grid exceed the address space");
    int status = PTD_OK;
    double *ckpt = nullptr, *hobs = nullptr, *gD = nullptr, *gH0 = nullptr;
    double* qbak[3] = {nullptr, nullptr, nullptr};
    double red[2] = {0, 0}, hsq[2] = {0, 0};
    const double tol_fwd = c->p.tol;
    double *v = nullptr, *lam[2] = {nullptr, nullptr}, *u[3] = {nullptr, nullptr, nullptr};
    double* z[3] = {nullptr, nullptr, nullptr};
    {
        if (nonlin) {  // the transposed iteration's cell and face state, lam of two consecutive steps
            PTD_ALLOCG(&v, nH * 8);
            for (int b = 0; b < 2; ++b) PTD_ALLOCG(&lam[b], nH * 8);
            for (int a = 0; a < c->p.ndim; ++a) {
                PTD_ALLOCG(&u[a], nH * 8);
                PTD_ALLOCG(&z[a], nH * 8);
            }
            for (int a = c->p.ndim; a < 3; ++a) u[a] = z[a] = u[0];  // absent axes: never read
        }
        PTD_ALLOCG(&ckpt, (size_t)(nt + 1) * nH * 8);
        PTD_ALLOCG(&hobs, nH * 8);
        PTD_ALLOCG(&gD, nH * 8);
        PTD_ALLOCG(&gH0, nH * 8);
        for (int a = 0; a < 3; ++a)
            if (c->nQ[a] > 0) PTD_ALLOCG(&qbak[a], (size_t)c->nQ[a] * 8);
        PTD_HIPG(hipMemsetAsync(gD, 0, nH * 8, c->stream));
        PTD_HIPG(hipMemsetAsync(gH0, 0, nH * 8, c->stream));
        PTD_HIPG(hipMemcpyAsync(hobs, h_obs, nH * 8, hipMemcpyHostToDevice, c->stream));

        // forward: ptd_solve(nt), keeping H^0..H^nt
        PTD_HIPG(hipMemcpyAsync(ckpt, c->H[c->hcur], nH * 8, hipMemcpyDeviceToDevice, c->stream));
        for (int n = 1; n <= nt; ++n) {
            int64_t it = 0;
            double rms = 0, mx = 0;
            status = ptd_pt_solve(c, &it, &rms, &mx);
            if (iters_fwd) iters_fwd[n - 1] = it;
            if (status != PTD_OK) goto done;
            PTD_HIPG(hipMemcpyAsync(ckpt + (size_t)n * nH, c->H[c->hcur], nH * 8, hipMemcpyDeviceToDevice, c->stream));
            status = ptd_assign(c);
            if (status != PTD_OK) goto done;
        }
        for (int a = 0; a < 3; ++a)
            if (qbak[a])
                PTD_HIPG(hipMemcpyAsync(qbak[a], c->q[a][c->qcur], (size_t)c->nQ[a] * 8, hipMemcpyDeviceToDevice, c->stream));

        // scale of the forward right-hand side: sum of (H^nt)^2 over the inner cells
        PTD_HIPG(ptd::launch_sumsq_inner(c, (const double*)c->H[c->hcur]));
        PTD_HIPG(hipMemcpyAsync(hsq, c->dred, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        // loss and the adjoint source: Hold = H = (H^nt - H_obs)*dt (0 on the boundary), fluxes from 0
        PTD_HIPG(ptd::launch_adj_source(c, (const double*)c->H[c->hcur], hobs, (double*)c->Hold));
        PTD_HIPG(hipMemcpyAsync(red, c->dred, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        for (int b = 0; b < 2; ++b) PTD_HIPG(hipMemcpyAsync(c->H[b], c->Hold, nH * 8, hipMemcpyDeviceToDevice, c->stream));
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 2; ++b)
                if (c->q[a][b]) PTD_HIPG(hipMemsetAsync(c->q[a][b], 0, (size_t)c->nQ[a] * 8, c->stream));
        PTD_HIPG(hipStreamSynchronize(c->stream));
        if (loss) *loss = red[0];
        // rms(r*dt)/rms(H^nt) = sqrt(2L)*dt/sqrt(sum H^2) (both over the inner cells); a zero misfit gives
        // tol 0 and an adjoint of 0, whose first residual is already 0
        c->p.tol = hsq[0] > 0.0 ? tol_fwd * std::sqrt(2.0 * red[0]) * c->p.dt / std::sqrt(hsq[0]) : tol_fwd;

        if (nonlin) {
            // J_n^T lam_n = rhs_n/dt with rhs_nt = (H^nt - H_obs)*dt (in Hold), rhs_n = lam_{n+1}; each solve
            // stops at the forward's relative reduction: rms(residual) <= tol*dt/rms(H^nt) * rms(rhs_n/dt).
            // v and u carry over from one step to the next (warm start).
            const double nR = (double)c->nR;
            const double tol_rel = hsq[0] > 0.0 ? tol_fwd * c->p.dt / std::sqrt(hsq[0] / nR) : 0.0;
            PTD_HIPG(hipMemsetAsync(v, 0, nH * 8, c->stream));
            for (int a = 0; a < c->p.ndim; ++a) PTD_HIPG(hipMemsetAsync(u[a], 0, nH * 8, c->stream));
            const double* rhs = (const double*)c->Hold;
            double rhs_rms = std::sqrt(2.0 * red[0] / nR);  // rms(rhs/dt) = rms(H^nt - H_obs)
            for (int n = nt; n >= 1; --n) {
                const double* Hn = ckpt + (size_t)n * nH;
                double* lamn = lam[n & 1];
                if (n < nt) {  // rms(rhs_n/dt) of rhs_n = lam_{n+1}
                    PTD_HIPG(ptd::launch_sumsq_inner(c, rhs));
                    PTD_HIPG(hipMemcpyAsync(hsq, c->dred, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
                    PTD_HIPG(hipStreamSynchronize(c->stream));
                    rhs_rms = std::sqrt(hsq[0] / nR) / c->p.dt;
                }
                const double tol_n = tol_rel * rhs_rms;
                int64_t it = 0;
                double err = HUGE_VAL;
                while (it < c->p.itmax) {
                    const int64_t blk = std::min<int64_t>(c->p.nout, c->p.itmax - it);
                    for (int64_t b = 0; b < blk; ++b) PTD_HIPG(ptd::launch_adjnl_iteration(c, Hn, v, u, z, rhs));
                    it += blk;
                    PTD_HIPG(ptd::launch_adjnl_lambda(c, Hn, v, lamn, rhs, true));
                    PTD_HIPG(hipMemcpyAsync(red, c->dred, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
                    PTD_HIPG(hipStreamSynchronize(c->stream));
                    err = std::sqrt(red[0] / nR);
                    if (std::isnan(err) || err <= tol_n) break;
                }
                if (iters_adj) iters_adj[n - 1] = it;
                if (std::isnan(err)) {
                    status = ptd::set_error(PTD_ERR_NAN, "NaN in the adjoint solve");
                    goto done;
                }
                if (it == 0) PTD_HIPG(ptd::launch_adjnl_lambda(c, Hn, v, lamn, rhs, false));
                PTD_HIPG(ptd::launch_adjnl_grad(c, Hn, lamn, gH0, n == 1));
                rhs = lamn;
            }
        }

        // adjoint: mu_nt, ..., mu_1, each followed by its gradient contribution
        for (int n = nt; n >= 1 && !nonlin; --n) {
            int64_t it = 0;
            double rms = 0, mx = 0;
            status = ptd_pt_solve(c, &it, &rms, &mx);
            if (iters_adj) iters_adj[n - 1] = it;
            if (status != PTD_OK) goto done;
            PTD_HIPG(ptd::launch_grad_accum(c, (const double*)c->H[c->hcur], ckpt + (size_t)n * nH, gD, gH0, n == 1));
            if (n > 1) {
                status = ptd_assign(c);  // Hold = mu_n: the next solve is A mu_{n-1} = mu_n/dt
                if (status != PTD_OK) goto done;
            }
        }
        if (grad_d) PTD_HIPG(hipMemcpyAsync(grad_d, gD, nH * 8, hipMemcpyDeviceToHost, c->stream));
        if (grad_h0) PTD_HIPG(hipMemcpyAsync(grad_h0, gH0, nH * 8, hipMemcpyDeviceToHost, c->stream));

        // the context as after ptd_solve(nt): H = Hold = H^nt, the forward fluxes
        for (int b = 0; b < 2; ++b)
            PTD_HIPG(hipMemcpyAsync(c->H[b], ckpt + (size_t)nt * nH, nH * 8, hipMemcpyDeviceToDevice, c->stream));
        PTD_HIPG(hipMemcpyAsync(c->Hold, ckpt + (size_t)nt * nH, nH * 8, hipMemcpyDeviceToDevice, c->stream));
        for (int a = 0; a < 3; ++a)
            if (qbak[a])
                PTD_HIPG(hipMemcpyAsync(c->q[a][c->qcur], qbak[a], (size_t)c->nQ[a] * 8, hipMemcpyDeviceToDevice, c->stream));
        PTD_HIPG(hipStreamSynchronize(c->stream));
    }
done:
    c->p.tol = tol_fwd;
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(v);
    for (int b = 0; b < 2; ++b) (void)hipFree(lam[b]);
    for (int a = 0; a < 3; ++a)
        if (a < c->p.ndim) {
            (void)hipFree(u[a]);
            (void)hipFree(z[a]);
        }
    (void)hipFree(ckpt);
    (void)hipFree(hobs);
    (void)hipFree(gD);
    (void)hipFree(gH0);
    for (int a = 0; a < 3; ++a) (void)hipFree(qbak[a]);
    return status;
}
