// ptd_api.hip -- the C ABI of libptd (include/ptd.h): context and field management, the
// kernel-level entry points (one reference kernel each) and the device-resident PT loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ptd_internal.hpp"

namespace ptd {
int64_t partial_capacity(const ptd_params& p, int64_t nR);
}

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define PTD_HIP(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(PTD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define PTD_REQUIRE(cond, code, msg) \
    do {                             \
        if (!(cond)) return fail((code), (msg)); \
    } while (0)

int check_params(const ptd_params* p) {
    PTD_REQUIRE(p != nullptr, PTD_ERR_ARG, "params is NULL");
    PTD_REQUIRE(p->ndim >= 1 && p->ndim <= 3, PTD_ERR_ARG, "ndim must be 1, 2 or 3");
    PTD_REQUIRE(p->dcase >= PTD_LIN && p->dcase <= PTD_NONLIN, PTD_ERR_ARG, "dcase must be PTD_LIN, PTD_STEP or PTD_NONLIN");
    PTD_REQUIRE(p->form == PTD_PLAIN || p->form == PTD_PERF, PTD_ERR_ARG, "form must be PTD_PLAIN or PTD_PERF");
    PTD_REQUIRE(p->precision == PTD_FP64 || p->precision == PTD_FP32 || p->precision == PTD_MIXED, PTD_ERR_ARG,
                "precision must be PTD_FP64, PTD_FP32 or PTD_MIXED");
    PTD_REQUIRE(p->precision != PTD_MIXED || p->ndim == 3, PTD_ERR_ARG,
                "PTD_MIXED runs the fused 3D correction sweeps: ndim must be 3");
    PTD_REQUIRE(p->stop_norm == PTD_NORM_RMS || p->stop_norm == PTD_NORM_MAX, PTD_ERR_ARG, "stop_norm must be PTD_NORM_RMS or PTD_NORM_MAX");
    PTD_REQUIRE(p->arith == PTD_ARITH_EXACT || p->arith == PTD_ARITH_FAST, PTD_ERR_ARG, "arith must be PTD_ARITH_EXACT or PTD_ARITH_FAST");
    // the FAST / float32 nonlinear Re = pi + a*rsq(a*h3) keeps a*h3 normal at the H^3 floor only for
    // K = max_l^2/dt >= 1e-6 (the reference's scripts have K ~ 500)
    PTD_REQUIRE(p->dcase != PTD_NONLIN || (p->arith == PTD_ARITH_EXACT && p->precision == PTD_FP64) ||
                    p->max_l2 / p->dt >= 1e-6,
                PTD_ERR_ARG, "FAST / float32 / mixed nonlinear sweeps need max_l^2/dt >= 1e-6");
    PTD_REQUIRE(p->nx >= 3, PTD_ERR_ARG, "nx must be >= 3");
    PTD_REQUIRE(p->ndim < 2 || p->ny >= 3, PTD_ERR_ARG, "ny must be >= 3");
    PTD_REQUIRE(p->ndim < 3 || p->nz >= 3, PTD_ERR_ARG, "nz must be >= 3");
    PTD_REQUIRE(p->nout >= 1, PTD_ERR_ARG, "nout must be >= 1");
    PTD_REQUIRE(p->itmax >= 0, PTD_ERR_ARG, "itmax must be >= 0");
    PTD_REQUIRE(p->dx > 0 && (p->ndim < 2 || p->dy > 0) && (p->ndim < 3 || p->dz > 0), PTD_ERR_ARG, "cell sizes must be > 0");
    PTD_REQUIRE(p->dt > 0 && p->vpdt > 0 && p->max_l > 0, PTD_ERR_ARG, "dt, vpdt and max_l must be > 0");
    const int64_t ny = p->ndim >= 2 ? p->ny : 1, nz = p->ndim >= 3 ? p->nz : 1;
    const double cells = (double)p->nx * (double)ny * (double)nz;
    PTD_REQUIRE(cells < 9.0e18 / 8.0, PTD_ERR_ARG, "grid too large");
    PTD_REQUIRE(p->precision != PTD_MIXED || ptd::max_chunk_tz(*p, 8) >= 1, PTD_ERR_ARG,
                "PTD_MIXED: an x-y plane of the grid must stay under 4 GiB / 6");
    return PTD_OK;
}

void shapes(const ptd_params& p, int64_t& nH, int64_t nQ[3], int64_t& nR) {
    const int64_t nx = p.nx, ny = p.ndim >= 2 ? p.ny : 1, nz = p.ndim >= 3 ? p.nz : 1;
    const int64_t mx = nx - 2, my = p.ndim >= 2 ? ny - 2 : 1, mz = p.ndim >= 3 ? nz - 2 : 1;
    nH = nx * ny * nz;
    nQ[0] = (nx - 1) * my * mz;
    nQ[1] = p.ndim >= 2 ? mx * (ny - 1) * mz : 0;
    nQ[2] = p.ndim >= 3 ? mx * my * (nz - 1) : 0;
    nR = mx * my * mz;
}

void* field_ptr(ptd_ctx* c, int field, int64_t* count) {
    switch (field) {
        case PTD_F_H: *count = c->nH; return c->H[c->hcur];
        case PTD_F_HOLD: *count = c->nH; return c->Hold;
        case PTD_F_D: *count = c->D ? c->nH : 0; return c->D;
        case PTD_F_QX: case PTD_F_QY: case PTD_F_QZ: {
            const int a = field - PTD_F_QX;
            *count = c->nQ[a];
            return c->q[a][c->qcur];
        }
        case PTD_F_QX2: case PTD_F_QY2: case PTD_F_QZ2: {
            const int a = field - PTD_F_QX2;
            *count = c->nQ[a];
            return c->q2[a];
        }
        case PTD_F_RESH: *count = c->nR; return c->resh;
        case PTD_F_DTAU: *count = c->dtau ? c->nR : 0; return c->dtau;
        default: *count = -1; return nullptr;
    }
}

// Device allocation of a context field, zeroed. Each field starts at its own offset into its allocation
// (a different multiple of 17 x 256 bytes below 64 KiB): the fields a sweep streams are read at equal
// element offsets, and bases that all sit on large power-of-two boundaries send those concurrent
// streams to the same HBM channels. Measured: the mixed solve at 511^3 0.4% faster (two runs each);
// at 1023^3 the run-to-run spread of the box (43-48 s) hides any difference.
int dev_alloc(ptd_ctx* c, void** ptr, int64_t n, size_t esz) {
    *ptr = nullptr;
    if (n <= 0) return PTD_OK;
    if (c->nraw >= (int)(sizeof(c->raw) / sizeof(c->raw[0]))) return fail(PTD_ERR_ARG, "too many allocations");
    const size_t bytes = (size_t)n * esz;
    constexpr size_t kSpan = 65536, kStep = 17 * 256;
    const size_t off = ((size_t)c->nraw * kStep) % kSpan;
    void* raw = nullptr;
    hipError_t e = hipMalloc(&raw, bytes + kSpan);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // clear it: later launch checks read the thread's last error
        return fail(PTD_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    c->raw[c->nraw++] = raw;
    *ptr = (char*)raw + off;
    e = hipMemsetAsync(*ptr, 0, bytes, c->stream);
    if (e != hipSuccess) return fail(PTD_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    c->bytes += (int64_t)bytes;
    return PTD_OK;
}

void free_ctx(ptd_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int k = 0; k < c->nraw; ++k) (void)hipFree(c->raw[k]);
    if (c->hred) (void)hipHostFree(c->hred);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (int k = 0; k < 2; ++k)
        if (c->evc[k]) (void)hipEventDestroy(c->evc[k]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// Copies between a host double array and a device array of the sweep precision.
int upload(ptd_ctx* c, void* dst, const double* src, int64_t n) {
    if (c->esz == 8) {
        PTD_HIP(hipMemcpyAsync(dst, src, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
        PTD_HIP(hipStreamSynchronize(c->stream));
        return PTD_OK;
    }
    const int64_t chunk = 1 << 24;
    std::vector<float> tmp((size_t)std::min<int64_t>(n, chunk));
    for (int64_t o = 0; o < n; o += chunk) {
        const int64_t m = std::min<int64_t>(chunk, n - o);
        for (int64_t t = 0; t < m; ++t) tmp[(size_t)t] = (float)src[o + t];
        PTD_HIP(hipMemcpyAsync((float*)dst + o, tmp.data(), (size_t)m * 4, hipMemcpyHostToDevice, c->stream));
        PTD_HIP(hipStreamSynchronize(c->stream));
    }
    return PTD_OK;
}

int download(ptd_ctx* c, double* dst, const void* src, int64_t n) {
    if (c->esz == 8) {
        PTD_HIP(hipMemcpyAsync(dst, src, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
        PTD_HIP(hipStreamSynchronize(c->stream));
        return PTD_OK;
    }
    const int64_t chunk = 1 << 24;
    std::vector<float> tmp((size_t)std::min<int64_t>(n, chunk));
    for (int64_t o = 0; o < n; o += chunk) {
        const int64_t m = std::min<int64_t>(chunk, n - o);
        PTD_HIP(hipMemcpyAsync(tmp.data(), (const float*)src + o, (size_t)m * 4, hipMemcpyDeviceToHost, c->stream));
        PTD_HIP(hipStreamSynchronize(c->stream));
        for (int64_t t = 0; t < m; ++t) dst[o + t] = (double)tmp[(size_t)t];
    }
    return PTD_OK;
}

int device_guard(ptd_ctx* c) {
    PTD_REQUIRE(c != nullptr, PTD_ERR_ARG, "ctx is NULL");
    PTD_HIP(hipSetDevice(c->device));
    return PTD_OK;
}

// Reads (sum r^2, max|r|) of the last reduction and forms the two norms (into the context's pinned
// mirror, or a stack buffer for contexts without one).
int fetch_errors(ptd_ctx* c, double* rms, double* mx) {
    double local[2] = {0.0, 0.0};
    double* h = c->hred ? c->hred : local;
    PTD_HIP(hipMemcpyAsync(h, c->dred, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PTD_HIP(hipStreamSynchronize(c->stream));
    const double len = c->p.norm_len > 0 ? c->p.norm_len : (double)c->nR;
    *rms = std::sqrt(h[0]) / std::sqrt(len);
    *mx = h[1];
    return PTD_OK;
}

// One PT iteration; on check iterations the residual reduction is left in dred.
int one_iteration(ptd_ctx* c, bool check) {
    const bool plain = c->p.form == PTD_PLAIN;
    if (ptd::fused_available(c)) {
        PTD_HIP(ptd::launch_fused(c, check && plain));
        if (check && !plain) {
            PTD_HIP(ptd::launch_flux(c, 2));
            PTD_HIP(ptd::launch_residual(c, false));
        }
        return PTD_OK;
    }
    // reference structure: compute_flux! (plain: residual fluxes from the same, pre-update H), compute_update!
    PTD_HIP(ptd::launch_flux(c, plain && check ? 1 : 0));
    PTD_HIP(ptd::launch_update(c));
    if (check) {
        if (!plain) PTD_HIP(ptd::launch_flux(c, 2));
        PTD_HIP(ptd::launch_residual(c, false));
    }
    return PTD_OK;
}

// n PT iterations; with `check_last` the last one also leaves the residual reduction in dred. The 3D
// path runs pairs of iterations through the two-level sweep (one HBM pass per pair) and a single
// sweep for an odd remainder; PTD_T2=0 in the environment forces single sweeps. Small 2D grids run
// their iterations in LDS-resident launches of up to lds2d_max_iters() each, the last one checking (plain
// form; the other forms check in a single sweep after them).
int run_iterations(ptd_ctx* c, int64_t n, bool check_last) {
    const char* t2e = getenv("PTD_T2");
    const bool t2_env = !(t2e && t2e[0] == '0');
    const bool plain = c->p.form == PTD_PLAIN;
    const bool t2 = t2_env && ptd::fused2_available(c);
    const bool t3 = t2_env && ptd::fused3_available(c);
    const bool lds = ptd::use_2d_lds(c);
    const int64_t lmax = ptd::lds2d_max_iters();
    while (n > 0) {
        if (lds && check_last && plain && n <= lmax) {  // the checking launch, residual reduced in it
            PTD_HIP(ptd::launch_lds2d(c, (int)n, true));
            n = 0;
        } else if (lds && n - (check_last ? 1 : 0) >= 1) {
            const int m = (int)std::min<int64_t>(n - (check_last ? 1 : 0), lmax);
            PTD_HIP(ptd::launch_lds2d(c, m, false));
            n -= m;
        } else if (ptd::next_launch_iters(n, check_last, t2, t3) == 3) {  // checks end a two-level pass
            PTD_HIP(ptd::launch_fused3(c));
            n -= 3;
        } else if (t2 && n >= 2) {
            const bool chk = check_last && n == 2;
            PTD_HIP(ptd::launch_fused2(c, chk && plain));
            if (chk && !plain) {
                PTD_HIP(ptd::launch_flux(c, 2));
                PTD_HIP(ptd::launch_residual(c, false));
            }
            n -= 2;
        } else {
            const int st = one_iteration(c, check_last && n == 1);
            if (st != PTD_OK) return st;
            n -= 1;
        }
    }
    return PTD_OK;
}

// float32 correction iterations (three-level passes for linear / step D on large grids, pairs through the
// two-level sweep), the last one optionally checked, on a two-level pass as in run_iterations
int run_delta(ptd_ctx* c, int64_t n, bool check_last) {
    const bool t3 = ptd::delta_t3(c);
    while (n > 0) {
        if (ptd::next_launch_iters(n, check_last, true, t3) == 3) {
            PTD_HIP(ptd::launch_delta_sweep3(c));
            n -= 3;
            continue;
        }
        const bool two = n >= 2;
        const bool chk = check_last && n == (two ? 2 : 1);
        PTD_HIP(ptd::launch_delta_sweep(c, two, chk));
        n -= two ? 2 : 1;
    }
    return PTD_OK;
}

// Defect-correction PT solve (PTD_MIXED). Outer loop in float64: R = residual of the backward-Euler
// system at the iterate Hc (flux from Hc itself); inner loop in float32: PT sweeps on the correction
// delta with source R (same Re / theta / dtau rules, evaluated at Hc + delta), checked on the float32
// correction residual; then Hc += delta. The inner sweeps are the reference PT iteration written in
// shifted variables (delta = H - Hc, p = q + g(Hc)); they start from the true fluxes q carried over from
// the previous solve, as the reference's q arrays are (starting from the equilibrium flux of a
// non-equilibrium H instead overshoots), and the flux state moves with Hc at every correction, so the
// corrections continue one PT iteration. An inner solve stops once it has reduced its residual
// kInner-fold, or reached kLast*tol (the last correction), or stalls. The float32 floor of one
// correction solve (rounding of delta, amplified by D/dx^2) lies ~3.5 decades below its starting
// residual at 511^3 and closer at larger n, and sweeps near it converge markedly slower than the
// float64 iteration; kInner keeps every correction clear of it. Checks are placed by prediction: the
// first after 70% of the sweeps the last complete correction took (scaled by the decades to go), the
// next ones after the sweeps the observed decay rate predicts. Measured (5 steps): 3910 float32 sweeps
// at 511^3 vs 3860 float64 iterations checked every 10; 7830 at 1023^3 vs 7670.
int pt_solve_mixed(ptd_ctx* c, int64_t* iters, double* err_rms, double* err_max) {
    const ptd_params& p = c->p;
    // reduction asked of one correction solve: its float32 floor lies a fixed factor times the
    // condition number (~ n^2) below its start, so the reduction per correction shrinks as n grows
    // (measured optimum: 1e-2 .. 3e-2 at 511^3, 5e-2 .. 1e-1 at 1023^3)
    const double nmax = (double)std::max(p.nx, std::max(p.ny, p.nz));
    const double kInner = std::min(1e-1, std::max(1e-2, 2.5e-2 * (nmax / 511.0) * (nmax / 511.0)));
    constexpr double kLast = 0.85;
    if (p.dcase == PTD_STEP) PTD_HIP(ptd::launch_f64_to_f32(c, (float*)c->Df, (const double*)c->D, c->nH));
    // residual of the float64 system at Hc, with the next correction's source, float32(Hc), delta = 0
    auto outer_residual = [&](double* rms, double* mx) -> int {
        PTD_HIP(ptd::launch_outer_residual(c));
        return fetch_errors(c, rms, mx);
    };
    double rms = 0, mx = 0;
    int st = outer_residual(&rms, &mx);
    if (st != PTD_OK) return st;
    double err = p.stop_norm == PTD_NORM_MAX ? mx : rms;
    int64_t total = 0, outer = 0;
    bool in_p = false;  // the correction fluxes p hold the flux state (q is stale)
    while (err > p.tol && total < p.itmax && !std::isnan(err)) {
        if (!in_p) {
            // start from the true fluxes (as the reference keeps qHx.. across steps): p = q + g(Hc)
            PTD_HIP(ptd::launch_flux_state(c, true));
            in_p = true;
        }
        // the last correction of the solve aims just below tol: the float64 residual after the update
        // lands within a few percent of the float32 one. A correction whose regular target would leave
        // less than kMerge-fold to go becomes the last one (saving a short final correction) where the
        // float32 floor leaves room for the larger reduction (up to 2.2 decades at 511^3; none merged
        // from kInner = 0.1, n >= 1023).
        const double kMerge = std::min(4.0, std::max(1.0, 0.1 / kInner));
        const bool last = kInner * err <= kMerge * kLast * p.tol;
        const double target = last ? kLast * p.tol : kInner * err;
        double best = HUGE_VAL;
        int stall = 0;
        int64_t it = 0;
        // A correction solve takes about as many sweeps per decade of reduction as the last one that
        // met the kInner target: the first check comes after 70% of the predicted count (in whole nout
        // blocks). (These checks are the float32 inner solve's own; the float64 stopping test above is
        // evaluated after every correction.)
        const double decades = std::log10(err / target) / std::log10(1.0 / kInner);
        const int64_t first = (int64_t)(0.7 * (double)c->inner_hint * std::min(1.0, decades)) / p.nout * p.nout;
        if (first > 0 && total + first < p.itmax) {
            st = run_delta(c, first, false);
            if (st != PTD_OK) return st;
            it = first;
        }
        bool met = false;
        int64_t gap = p.nout;             // sweeps before the next check
        double r_prev = err;
        int64_t it_prev = 0;
        while (total + it < p.itmax) {
            const int64_t n = std::min<int64_t>(gap, p.itmax - total - it);
            st = run_delta(c, n, true);
            if (st != PTD_OK) return st;
            it += n;
            double irms = 0, imx = 0;
            st = fetch_errors(c, &irms, &imx);
            if (st != PTD_OK) return st;
            const double r = p.stop_norm == PTD_NORM_MAX ? imx : irms;
            if (std::isnan(r)) break;
            if (r <= target) {
                met = !last;
                break;
            }
            if (r < 0.95 * best) {
                best = r;
                stall = 0;
            } else if (++stall >= 3) {
                break;  // at the float32 floor of this correction
            }
            // next check: the sweeps the decay rate since the previous check predicts for the rest of the
            // reduction (80% of them for the last correction, whose overshoot is not used), at most half
            // the sweeps run so far (whole nout blocks, at least one)
            gap = p.nout;
            if (r < r_prev && it > it_prev) {
                const double per_sweep = std::log(r / r_prev) / (double)(it - it_prev);
                const double need = std::min((last ? 0.8 : 1.0) * std::log(target / r) / per_sweep, 0.5 * (double)it);
                gap = std::max<int64_t>(p.nout, (int64_t)need / p.nout * p.nout);
            }
            r_prev = r;
            it_prev = it;
        }
        total += it;
        if (met) c->inner_hint = it;
        // one pass: p moves to Hc + delta, Hc += delta, and the next correction's residual, source,
        // float32(Hc) and delta = 0
        PTD_HIP(ptd::launch_outer_step(c));
        st = fetch_errors(c, &rms, &mx);
        if (st != PTD_OK) return st;
        err = p.stop_norm == PTD_NORM_MAX ? mx : rms;
        ++outer;
        if (it == 0) break;
    }
    if (in_p) PTD_HIP(ptd::launch_flux_state(c, false));  // q = p - g(Hc): the flux state in float64
    PTD_HIP(hipStreamSynchronize(c->stream));
    c->outer_last = outer;
    if (iters) *iters = total;
    if (err_rms) *err_rms = rms;
    if (err_max) *err_max = mx;
    if (std::isnan(err)) return fail(PTD_ERR_NAN, "NaN");
    return PTD_OK;
}

// The solve loop below for small 2D grids, whose check blocks are one LDS-resident launch each (plain
// form, nout <= lds2d_max_iters()) and last a few tens of microseconds: the next block is enqueued before
// the previous block's residual is read, so the GPU does not wait for the host between blocks. A block
// reads one half of the double-buffered H / q and writes the other, so the state a check accepts stays
// intact under the block enqueued after it; accepting it restores the buffer parity, and the extra
// block's results are never read. Iterations, residuals and the final state are those of the loop below.
int pt_solve_pipelined(ptd_ctx* c, int64_t* iters, double* err_rms, double* err_max) {
    const ptd_params& p = c->p;
    const int64_t nout = p.nout;
    const double len = p.norm_len > 0 ? p.norm_len : (double)c->nR;
    double* slot[2] = {c->hred, c->hred + 2};
    auto enqueue = [&](int s) -> int {
        const int st = run_iterations(c, nout, true);
        if (st != PTD_OK) return st;
        PTD_HIP(hipMemcpyAsync(slot[s], c->dred, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        PTD_HIP(hipEventRecord(c->evc[s], c->stream));
        return PTD_OK;
    };
    int64_t iter = 0, queued = 0;  // iterations checked / enqueued
    double rms = 2.0 * p.tol, mx = rms, err = rms;
    int s = 0;
    if (nout <= p.itmax) {
        const int st = enqueue(s);
        if (st != PTD_OK) return st;
        queued = nout;
    }
    while (queued > iter) {
        const int hcur = c->hcur, qcur = c->qcur;  // parity of the state after the block in flight
        const bool spec = queued + nout <= p.itmax;
        if (spec) {
            const int st = enqueue(1 - s);
            if (st != PTD_OK) return st;
            queued += nout;
        }
        PTD_HIP(hipEventSynchronize(c->evc[s]));
        iter += nout;
        rms = std::sqrt(slot[s][0]) / std::sqrt(len);
        mx = slot[s][1];
        err = p.stop_norm == PTD_NORM_MAX ? mx : rms;
        if (!(err > p.tol)) {  // converged (or NaN): keep this block's state
            c->hcur = hcur;
            c->qcur = qcur;
            break;
        }
        s = 1 - s;
    }
    if (err > p.tol && iter < p.itmax) {  // the iterations up to itMax, fewer than nout: no check
        const int st = run_iterations(c, p.itmax - iter, false);
        if (st != PTD_OK) return st;
        iter = p.itmax;
    }
    PTD_HIP(hipStreamSynchronize(c->stream));
    if (iters) *iters = iter;
    if (err_rms) *err_rms = rms;
    if (err_max) *err_max = mx;
    if (std::isnan(err)) return fail(PTD_ERR_NAN, "NaN");
    return PTD_OK;
}

// One PT solve at fixed Hold (float64 / float32 sweeps): `while err > tol && iter < itmax` with a check
// every nout iterations (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:102-115). err only changes at checks,
// so the iterations up to the next check (or itMax) run as one block.
int pt_solve_loop(ptd_ctx* c, int64_t* iters, double* err_rms, double* err_max) {
    const ptd_params& p = c->p;
    const char* pipe = getenv("PTD_PIPELINE");
    if (c->hred && c->evc[0] && c->evc[1] && ptd::use_2d_lds(c) && p.form == PTD_PLAIN &&
        p.nout <= ptd::lds2d_max_iters() && !(pipe && pipe[0] == '0'))
        return pt_solve_pipelined(c, iters, err_rms, err_max);
    int64_t iter = 0;
    double err = 2.0 * p.tol, rms = err, mx = err;
    while (err > p.tol && iter < p.itmax) {
        const int64_t next_check = (iter / p.nout + 1) * (int64_t)p.nout;
        const int64_t stop = next_check < p.itmax ? next_check : p.itmax;
        const bool check = (stop % p.nout) == 0;
        int st = run_iterations(c, stop - iter, check);
        if (st != PTD_OK) return st;
        iter = stop;
        if (check) {
            st = fetch_errors(c, &rms, &mx);
            if (st != PTD_OK) return st;
            err = p.stop_norm == PTD_NORM_MAX ? mx : rms;
        }
    }
    PTD_HIP(hipStreamSynchronize(c->stream));
    if (iters) *iters = iter;
    if (err_rms) *err_rms = rms;
    if (err_max) *err_max = mx;
    if (std::isnan(err)) return fail(PTD_ERR_NAN, "NaN");
    return PTD_OK;
}

// ---- the solve without allocation (ptd_pt_solve_noalloc) --------------------------------------
//
// Scratch of the residual reduction: per-tile (sum r^2, max|r|) partials of the checking launch and the
// reduced pair after them. Every device holds its own copy until the process ends; its device address is
// looked up once per device and kept in g_scratch (the array's name is a host-side stand-in, never passed
// to a kernel). Nothing in it needs clearing: every tile of a checking launch writes its partial before
// the reduction reads it. Calls that overlap on one device share it and race.
constexpr int64_t kNoAllocPartials = PTD_NOALLOC_MAX_PARTIALS;
__device__ __align__(256) double2 scratch[kNoAllocPartials + 1];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double2* g_scratch[kMaxDevices] = {};

int scratch_for(int device, double2** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(PTD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = (double2*)ptr;
    }
    *out = g_scratch[device];
    return PTD_OK;
}

}  // namespace

namespace ptd {
// shared with the batch entry points (ptd_batch.hip)
int validate_params(const ptd_params* p) { return check_params(p); }
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace ptd

extern "C" {

int32_t ptd_abi_version(void) { return PTD_ABI_VERSION; }

const char* ptd_last_error(void) { return g_err.c_str(); }

int32_t ptd_device_check(int32_t dev, char* name_out, int32_t name_len) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return fail(PTD_ERR_NODEV, "no HIP device visible");
    if (dev < 0 || dev >= n) return fail(PTD_ERR_NODEV, "device index out of range");
    hipDeviceProp_t prop;
    PTD_HIP(hipGetDeviceProperties(&prop, dev));
    if (name_out && name_len > 0) {
        std::snprintf(name_out, (size_t)name_len, "%s", prop.gcnArchName);
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(PTD_ERR_NODEV, std::string("device is ") + prop.gcnArchName + ", libptd is built for gfx950");
    return PTD_OK;
}

int32_t ptd_ctx_create(const ptd_params* params, int32_t device, ptd_ctx** out) {
    PTD_REQUIRE(out != nullptr, PTD_ERR_ARG, "out is NULL");
    *out = nullptr;
    int st = check_params(params);
    if (st != PTD_OK) return st;
    st = ptd_device_check(device, nullptr, 0);
    if (st != PTD_OK) return st;
    PTD_HIP(hipSetDevice(device));
    ptd_ctx* c = new ptd_ctx();
    std::memset((void*)c, 0, sizeof(*c));
    c->p = *params;
    if (c->p.ndim < 2) c->p.ny = 1;
    if (c->p.ndim < 3) c->p.nz = 1;
    c->device = device;
    c->esz = params->precision == PTD_FP32 ? 4 : 8;
    c->mixed = params->precision == PTD_MIXED;
    shapes(c->p, c->nH, c->nQ, c->nR);
    hipError_t e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) {
        free_ctx(c);
        return fail(PTD_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
#define PTD_ALLOC(ptr, n, esz)                 \
    do {                                       \
        int s_ = dev_alloc(c, (ptr), (n), (esz)); \
        if (s_ != PTD_OK) {                    \
            free_ctx(c);                       \
            return s_;                         \
        }                                      \
    } while (0)
    for (int b = 0; b < 2; ++b) {
        PTD_ALLOC(&c->H[b], c->nH, c->esz);
        for (int a = 0; a < 3; ++a) PTD_ALLOC(&c->q[a][b], c->nQ[a], c->esz);
    }
    for (int a = 0; a < 3; ++a) PTD_ALLOC(&c->q2[a], c->nQ[a], c->esz);
    PTD_ALLOC(&c->Hold, c->nH, c->esz);
    if (c->p.dcase == PTD_STEP) {
        PTD_ALLOC(&c->D, c->nH, c->esz);
        PTD_ALLOC(&c->dtau, c->nR, c->esz);
    }
    PTD_ALLOC(&c->resh, c->nR, c->esz);
    if (c->mixed) {
        for (int b = 0; b < 2; ++b) {
            PTD_ALLOC(&c->dl[b], c->nH, sizeof(float));
            for (int a = 0; a < 3; ++a) PTD_ALLOC(&c->pq[a][b], c->nQ[a], sizeof(float));
        }
        PTD_ALLOC(&c->src, c->nH, sizeof(float));
        if (c->p.dcase == PTD_STEP) PTD_ALLOC(&c->Df, c->nH, sizeof(float));
        if (c->p.dcase == PTD_NONLIN) PTD_ALLOC(&c->hcf, c->nH, sizeof(float));
    }
    c->npartial = ptd::partial_capacity(c->p, c->nH);  // per-cell reductions (adjoint source) cover nH
    PTD_ALLOC((void**)&c->partial, 2 * c->npartial, sizeof(double));
    PTD_ALLOC((void**)&c->dred, 2, sizeof(double));
#undef PTD_ALLOC
    e = hipHostMalloc((void**)&c->hred, 4 * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipEventCreateWithFlags(&c->evc[k], hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        free_ctx(c);
        return fail(PTD_ERR_HIP, std::string("context setup: ") + hipGetErrorString(e));
    }
    *out = c;
    return PTD_OK;
}

int32_t ptd_ctx_destroy(ptd_ctx* ctx) {
    free_ctx(ctx);
    return PTD_OK;
}

int32_t ptd_ctx_set_params(ptd_ctx* ctx, const ptd_params* params) {
    int st = device_guard(ctx);
    if (st != PTD_OK) return st;
    st = check_params(params);
    if (st != PTD_OK) return st;
    const ptd_params& o = ctx->p;
    PTD_REQUIRE(params->ndim == o.ndim && params->nx == o.nx && (o.ndim < 2 || params->ny == o.ny) &&
                    (o.ndim < 3 || params->nz == o.nz) && params->dcase == o.dcase &&
                    params->precision == o.precision,
                PTD_ERR_ARG, "set_params may not change shape, dcase or precision");
    ptd_params np = *params;
    np.ny = o.ny;
    np.nz = o.nz;
    ctx->p = np;
    return PTD_OK;
}

int64_t ptd_ctx_device_bytes(const ptd_ctx* ctx) { return ctx ? ctx->bytes : 0; }

int64_t ptd_ctx_outer_iterations(const ptd_ctx* ctx) { return ctx ? ctx->outer_last : -1; }

void* ptd_ctx_stream(ptd_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int64_t ptd_field_count(const ptd_ctx* ctx, int32_t field) {
    if (!ctx) return -1;
    int64_t n = -1;
    field_ptr(const_cast<ptd_ctx*>(ctx), field, &n);
    return n;
}

int32_t ptd_set_field(ptd_ctx* ctx, int32_t field, const double* host, int64_t count) {
    int st = device_guard(ctx);
    if (st != PTD_OK) return st;
    int64_t n = 0;
    void* dst = field_ptr(ctx, field, &n);
    PTD_REQUIRE(n >= 0, PTD_ERR_ARG, "unknown field");
    PTD_REQUIRE(n > 0 && dst != nullptr, PTD_ERR_STATE, "field is absent for this case");
    PTD_REQUIRE(host != nullptr, PTD_ERR_ARG, "host pointer is NULL");
    PTD_REQUIRE(count == n, PTD_ERR_ARG, "count " + std::to_string(count) + " != field size " + std::to_string(n));
    st = upload(ctx, dst, host, n);
    if (st != PTD_OK) return st;
    if (field == PTD_F_H) {
        PTD_HIP(hipMemcpyAsync(ctx->H[1 - ctx->hcur], dst, (size_t)n * ctx->esz, hipMemcpyDeviceToDevice, ctx->stream));
        PTD_HIP(hipStreamSynchronize(ctx->stream));
    }
    return PTD_OK;
}

int32_t ptd_get_field(ptd_ctx* ctx, int32_t field, double* host, int64_t count) {
    int st = device_guard(ctx);
    if (st != PTD_OK) return st;
    int64_t n = 0;
    void* src = field_ptr(ctx, field, &n);
    PTD_REQUIRE(n >= 0, PTD_ERR_ARG, "unknown field");
    PTD_REQUIRE(n > 0 && src != nullptr, PTD_ERR_STATE, "field is absent for this case");
    PTD_REQUIRE(host != nullptr, PTD_ERR_ARG, "host pointer is NULL");
    PTD_REQUIRE(count == n, PTD_ERR_ARG, "count " + std::to_string(count) + " != field size " + std::to_string(n));
    return download(ctx, host, src, n);
}

int32_t ptd_compute_iter_params(ptd_ctx* ctx) {
    int st = device_guard(ctx);
    if (st != PTD_OK) return st;
    PTD_REQUIRE(ctx->p.dcase == PTD_STEP, PTD_ERR_STATE, "compute_iter_params needs the step case");
    PTD_HIP(ptd::launch_iter_params(ctx));
    PTD_HIP(hipStreamSynchronize(ctx->stream));
    return PTD_OK;
}

int32_t ptd_compute_flux(ptd_ctx* ctx) {
    int st = device_guard(ctx);
    if (st != PTD_OK) return st;
    PTD_HIP(ptd::launch_flux(ctx, ctx->p.form == PTD_PLAIN ? 1 : 0));
    PTD_HIP(hipStreamSynchronize(ctx->stream));
    return PTD_OK;
}

int32_t ptd_compute_update(ptd_ctx* ctx) {
    int st = device_guard(ctx);
    if (st != PTD_OK) return st;
    PTD_HIP(ptd::launch_update(ctx));
    PTD_HIP(hipStreamSynchronize(ctx->stream));
    return PTD_OK;
}

int32_t ptd_compute_flux_res(ptd_ctx* ctx) {
    int st = device_guard(ctx);
    if (st != PTD_OK) return st;
    PTD_HIP(ptd::launch_flux(ctx, 2));
    PTD_HIP(hipStreamSynchronize(ctx->stream));
    return PTD_OK;
}

int32_t ptd_check_res(ptd_ctx* ctx, double* err_rms, double* err_max) {
    int st = device_guard(ctx);
    if (st != PTD_OK) return st;
    PTD_HIP(ptd::launch_residual(ctx, true));
    double rms = 0, mx = 0;
    st = fetch_errors(ctx, &rms, &mx);
    if (st != PTD_OK) return st;
    if (err_rms) *err_rms = rms;
    if (err_max) *err_max = mx;
    return PTD_OK;
}

int32_t ptd_assign(ptd_ctx* ctx) {
    int st = device_guard(ctx);
    if (st != PTD_OK) return st;
    PTD_HIP(hipMemcpyAsync(ctx->Hold, ctx->H[ctx->hcur], (size_t)ctx->nH * ctx->esz, hipMemcpyDeviceToDevice, ctx->stream));
    PTD_HIP(hipStreamSynchronize(ctx->stream));
    return PTD_OK;
}

int32_t ptd_iterate(ptd_ctx* ctx, int64_t niter) {
    int st = device_guard(ctx);
    if (st != PTD_OK) return st;
    PTD_REQUIRE(niter >= 0, PTD_ERR_ARG, "niter must be >= 0");
    st = run_iterations(ctx, niter, false);
    if (st != PTD_OK) return st;
    PTD_HIP(hipStreamSynchronize(ctx->stream));
    return PTD_OK;
}

int32_t ptd_iterate_timed(ptd_ctx* ctx, int64_t niter, double* ms) {
    int st = device_guard(ctx);
    if (st != PTD_OK) return st;
    PTD_REQUIRE(niter >= 0, PTD_ERR_ARG, "niter must be >= 0");
    PTD_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    st = run_iterations(ctx, niter, false);
    if (st != PTD_OK) return st;
    PTD_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    PTD_HIP(hipEventSynchronize(ctx->ev1));
    float f = 0.0f;
    PTD_HIP(hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
    if (ms) *ms = (double)f;
    return PTD_OK;
}

int32_t ptd_ctx_iterations_per_launch(const ptd_ctx* ctx, int32_t* n) {
    if (!ctx || !n) return ptd::set_error(PTD_ERR_ARG, "ctx or n is NULL");
    // run_iterations' choice for a long unchecked run
    const char* t2e = getenv("PTD_T2");
    const bool t2_env = !(t2e && t2e[0] == '0');
    if (ptd::use_2d_lds(ctx)) *n = (int32_t)ptd::lds2d_max_iters();
    else *n = ptd::next_launch_iters(60, false, t2_env && ptd::fused2_available(ctx), t2_env && ptd::fused3_available(ctx));
    return PTD_OK;
}

int32_t ptd_pt_solve(ptd_ctx* ctx, int64_t* iters, double* err_rms, double* err_max) {
    int st = device_guard(ctx);
    if (st != PTD_OK) return st;
    if (ctx->mixed) return pt_solve_mixed(ctx, iters, err_rms, err_max);
    return pt_solve_loop(ctx, iters, err_rms, err_max);
}

int32_t ptd_pt_solve_noalloc(const ptd_params* params, int32_t device, void* stream, void* H0, void* H1,
                             const void* Hold, const void* D, void* qx0, void* qx1, void* qy0, void* qy1,
                             void* qz0, void* qz1, int32_t* parity, int64_t* iters, double* err_rms,
                             double* err_max) {
    int st = check_params(params);
    if (st != PTD_OK) return st;
    PTD_REQUIRE(params->ndim >= 2, PTD_ERR_ARG, "ptd_pt_solve_noalloc runs the fused 2D / 3D sweeps: ndim must be 2 or 3");
    PTD_REQUIRE(params->form == PTD_PLAIN, PTD_ERR_ARG,
                "ptd_pt_solve_noalloc checks inside the fused sweep: form must be PTD_PLAIN");
    PTD_REQUIRE(params->precision == PTD_FP64 || params->precision == PTD_FP32, PTD_ERR_ARG,
                "ptd_pt_solve_noalloc: precision must be PTD_FP64 or PTD_FP32");
    PTD_REQUIRE(parity != nullptr && (*parity == 0 || *parity == 1), PTD_ERR_ARG, "parity must point to 0 or 1");
    PTD_REQUIRE(H0 && H1 && Hold && qx0 && qx1 && qy0 && qy1, PTD_ERR_ARG, "a field pointer is NULL");
    PTD_REQUIRE(params->ndim < 3 || (qz0 && qz1), PTD_ERR_ARG, "qz buffers are NULL");
    PTD_REQUIRE(params->dcase != PTD_STEP || D, PTD_ERR_ARG, "the step case needs D");
    PTD_REQUIRE(device >= 0 && device < kMaxDevices, PTD_ERR_ARG, "device index out of range");
    st = ptd_device_check(device, nullptr, 0);
    if (st != PTD_OK) return st;
    PTD_HIP(hipSetDevice(device));
    double2* scr = nullptr;
    st = scratch_for(device, &scr);
    if (st != PTD_OK) return st;
    ptd_ctx c;
    std::memset((void*)&c, 0, sizeof(c));
    c.p = *params;
    if (c.p.ndim < 3) c.p.nz = 1;
    c.device = device;
    c.stream = (hipStream_t)stream;
    c.esz = params->precision == PTD_FP32 ? 4 : 8;
    shapes(c.p, c.nH, c.nQ, c.nR);
    PTD_REQUIRE(ptd::fused_available(&c), PTD_ERR_ARG, "grid not supported by the fused sweeps");
    const int64_t need = ptd::fused_check_tiles(&c);
    PTD_REQUIRE(need <= kNoAllocPartials, PTD_ERR_ARG,
                "grid too large for ptd_pt_solve_noalloc: its checks reduce " + std::to_string(need) +
                    " tile partials, the scratch holds " + std::to_string(kNoAllocPartials));
    c.H[0] = H0;
    c.H[1] = H1;
    void* q[3][2] = {{qx0, qx1}, {qy0, qy1}, {qz0, qz1}};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 2; ++b) c.q[a][b] = c.p.ndim == 2 && a == 2 ? nullptr : q[a][b];
    c.hcur = c.qcur = *parity;
    c.Hold = const_cast<void*>(Hold);
    c.D = const_cast<void*>(D);
    c.partial = (double*)scr;
    c.npartial = kNoAllocPartials;
    c.dred = (double*)(scr + kNoAllocPartials);
    st = pt_solve_loop(&c, iters, err_rms, err_max);
    *parity = c.hcur;
    return st;
}

int32_t ptd_solve(ptd_ctx* ctx, int32_t nt, int64_t* iters, double* errs) {
    int st = device_guard(ctx);
    if (st != PTD_OK) return st;
    PTD_REQUIRE(nt >= 0, PTD_ERR_ARG, "nt must be >= 0");
    for (int32_t it = 0; it < nt; ++it) {
        int64_t n = 0;
        double rms = 0, mx = 0;
        const int s = ptd_pt_solve(ctx, &n, &rms, &mx);
        if (iters) iters[it] = n;
        if (errs) errs[it] = ctx->p.stop_norm == PTD_NORM_MAX ? mx : rms;
        if (s != PTD_OK && s != PTD_ERR_NAN) return s;
        st = ptd_assign(ctx);  // Hold .= H
        if (st != PTD_OK) return st;
        if (s == PTD_ERR_NAN) return s;
    }
    return PTD_OK;
}

}  // extern "C"
