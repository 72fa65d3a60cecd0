// ptd_internal.hpp -- shared definitions of libptd (not part of the C ABI).
//
// Device-side restatement of the reference's per-face and per-cell expressions. Every template keeps
// the operation order of the reference source it cites, and the library is compiled with
// -ffp-contract=off, so the float64 path rounds exactly as the reference's Julia code does.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <stdint.h>

#include "../../include/ptd.h"

namespace ptd {

// Float64(pi) and pi*pi as Julia evaluates them for Irrational pi (both operands rounded first).
constexpr double kPi = 3.141592653589793115997963468544185161590576171875;

// Scalars of one problem, in the sweep precision T.
template <typename T>
struct KP {
    int64_t nx, ny, nz;   // cells (1 on absent axes)
    int64_t mx, my, mz;   // inner cells (n-2 on active axes, 1 on absent ones)
    int32_t hy, hz;       // y / z axes present
    T dx, dy, dz, _dx, _dy, _dz, dt, _dt;
    T pi, pi2;            // Float64(pi), pi*pi
    T c_lv;               // max_l / Vpdtau      (first division of max_l / Vpdtau / Re * Resc)
    T c_vl;               // Vpdtau * max_l      (first product of Vpdtau * max_l / H3 / Re * Resc)
    T max_l2, resc, d0, theta_r, dtau_rho, re;
    // constants of the FAST arithmetic (rearranged update rules)
    T K;                  // max_l2 * _dt
    T K8;                 // 8 * K (nonlinear faces work with u = (Ha + Hb)^3 = 8 * Hface^3)
    T _dx8, _dy8, _dz8;   // _dx / 8, _dy / 8, _dz / 8
    T _dxs, _dys, _dzs;   // _dx * _dx / 8 ...: divergence weights of the scaled correction fluxes (kDelta)
    T c1;                 // max_l / Vpdtau * Resc           (theta = c1 / Re)
    T c2;                 // Vpdtau * max_l * Resc           (nonlinear dtau_rho = c2 / (H^3 Re))
    T c3;                 // Vpdtau * max_l / Re             (step dtau_rho = c3 / maxloc(D))
    T inv1p_theta;        // 1 / (1 + theta_r_dtau)          (lin/step)
    T inv1p_dtau;         // 1 / (1 + _dt * dtau_rho)        (lin)
};

template <typename T>
KP<T> make_kp(const ptd_params& p) {
    KP<T> k;
    k.nx = p.nx;
    k.ny = p.ndim >= 2 ? p.ny : 1;
    k.nz = p.ndim >= 3 ? p.nz : 1;
    k.hy = p.ndim >= 2;
    k.hz = p.ndim >= 3;
    k.mx = k.nx - 2;
    k.my = k.hy ? k.ny - 2 : 1;
    k.mz = k.hz ? k.nz - 2 : 1;
    k.dx = (T)p.dx;
    k.dy = (T)(k.hy ? p.dy : 1.0);
    k.dz = (T)(k.hz ? p.dz : 1.0);
    k._dx = (T)(1.0 / p.dx);  // _dx = 1.0/dx (scripts/diff_3D/diff_3D_nonlin_perf.jl:78)
    k._dy = (T)(k.hy ? 1.0 / p.dy : 0.0);
    k._dz = (T)(k.hz ? 1.0 / p.dz : 0.0);
    k.dt = (T)p.dt;
    k._dt = (T)(1.0 / p.dt);
    k.pi = (T)kPi;
    k.pi2 = (T)(kPi * kPi);
    k.c_lv = (T)(p.max_l / p.vpdt);
    k.c_vl = (T)(p.vpdt * p.max_l);
    k.max_l2 = (T)p.max_l2;
    k.resc = (T)p.resc;
    k.d0 = (T)p.d0;
    k.theta_r = (T)p.theta_r;
    k.dtau_rho = (T)p.dtau_rho;
    k.re = (T)p.re;
    const double _dt = 1.0 / p.dt;
    k.K = (T)(p.max_l2 * _dt);
    k.K8 = (T)(8.0 * (p.max_l2 * _dt));
    k._dx8 = (T)(0.125 / p.dx);
    k._dy8 = (T)(k.hy ? 0.125 / p.dy : 0.0);
    k._dz8 = (T)(k.hz ? 0.125 / p.dz : 0.0);
    k._dxs = (T)(0.125 / (p.dx * p.dx));
    k._dys = (T)(k.hy ? 0.125 / (p.dy * p.dy) : 0.0);
    k._dzs = (T)(k.hz ? 0.125 / (p.dz * p.dz) : 0.0);
    k.c1 = (T)(p.max_l / p.vpdt * p.resc);
    k.c2 = (T)(p.vpdt * p.max_l * p.resc);
    k.c3 = (T)(p.re > 0 ? p.vpdt * p.max_l / p.re : 0.0);
    k.inv1p_theta = (T)(1.0 / (1.0 + p.theta_r));
    k.inv1p_dtau = (T)(1.0 / (1.0 + _dt * p.dtau_rho));
    return k;
}

// ---------------------------------------------------------------------------------------------
// physics, one face or one cell
// ---------------------------------------------------------------------------------------------

// Face coefficient between the left/lower cell a and the right/upper cell b.
//  lin   : D                                        (scripts/diff_3D/diff_3D_lin_multixpu.jl:25)
//  step  : @av_xi(D) = (D[a] + D[b])*0.5               (scripts/diff_3D/diff_3D_linstep_multixpu.jl:30)
//  nonlin: plain @av_xi(H)*@av_xi(H)*@av_xi(H)         (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:25)
//          perf  0.5*(a+b) * 0.5*(a+b) * 0.5*(a+b)     (scripts/diff_3D/diff_3D_nonlin_perf.jl:20)
template <typename T, int DC, int FORM>
__device__ __forceinline__ T face_coef(const KP<T>& p, T Ha, T Hb, T Da, T Db) {
    if constexpr (DC == PTD_LIN) {
        return p.d0;
    } else if constexpr (DC == PTD_STEP) {
        return (Da + Db) * (T)0.5;
    } else {
        T s = Ha + Hb;
        if constexpr (FORM == PTD_PERF) {
            return (T)0.5 * s * (T)0.5 * s * (T)0.5 * s;
        } else {
            T av = s * (T)0.5;
            return av * av * av;
        }
    }
}

// theta_r_dtau at a face: scalar for lin/step; for nonlin
//   max_l / Vpdtau / (pi + sqrt(pi*pi + max_l2 / H3 / dt)) * Resc    (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:28,32)
//   ... max_l2 / H3 * _dt ...                                        (scripts/diff_3D/diff_3D_nonlin_perf.jl:23,27)
template <typename T, int DC, int FORM>
__device__ __forceinline__ T face_theta(const KP<T>& p, T coef) {
    if constexpr (DC != PTD_NONLIN) {
        return p.theta_r;
    } else {
        T re;
        if constexpr (FORM == PTD_PERF) re = p.pi + sqrt(p.pi2 + p.max_l2 / coef * p._dt);
        else re = p.pi + sqrt(p.pi2 + p.max_l2 / coef / p.dt);
        return p.c_lv / re * p.resc;
    }
}

// Damped flux (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:38, diff_3D_nonlin_perf.jl:33).
template <typename T, int FORM>
__device__ __forceinline__ T damped_flux(T q, T theta, T coef, T dH, T h, T ih) {
    if constexpr (FORM == PTD_PERF) return (q * theta - coef * ih * dH) / ((T)1.0 + theta);
    else return (q * theta - coef * dH / h) / ((T)1.0 + theta);
}

// Residual flux -coef*dH/dx (plain, diff_3D_nonlin_multixpu.jl:41) or -coef*_dx*dH (perf, diff_3D_nonlin_perf.jl:45).
template <typename T, int FORM>
__device__ __forceinline__ T res_flux(T coef, T dH, T h, T ih) {
    if constexpr (FORM == PTD_PERF) return -coef * ih * dH;
    else return -coef * dH / h;
}

// Nonlinear cell dtau_rho = Vpdtau * max_l / H3 / Re * Resc with Re from the cell H3
// (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:31,35; diff_3D_nonlin_perf.jl:26,30).
template <typename T, int FORM>
__device__ __forceinline__ T nonlin_dtau(const KP<T>& p, T H) {
    T h3 = H * H * H;
    T re;
    if constexpr (FORM == PTD_PERF) re = p.pi + sqrt(p.pi2 + p.max_l2 / h3 * p._dt);
    else re = p.pi + sqrt(p.pi2 + p.max_l2 / h3 / p.dt);
    return p.c_vl / h3 / re * p.resc;
}

// Step dtau_rho from the neighbourhood maximum of D (compute_iter_params!,
// scripts/diff_3D/diff_3D_linstep_multixpu.jl:24-27: Vpdtau * max_l / @maxloc(D) / Re).
template <typename T>
__device__ __forceinline__ T step_dtau(const KP<T>& p, T dmax) {
    return p.c_vl / dmax / p.re;
}

// Divergence term from the three flux differences (left to right sum, absent axes dropped).
//  plain: dqx/dx + dqy/dy + dqz/dz       (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:48)
//  perf : _dx*dqx + _dy*dqy + _dz*dqz    (scripts/diff_3D/diff_3D_nonlin_perf.jl:40)
template <typename T, int FORM, int NDIM>
__device__ __forceinline__ T divergence(const KP<T>& p, T ddx, T ddy, T ddz) {
    T s;
    if constexpr (FORM == PTD_PERF) {
        s = p._dx * ddx;
        if constexpr (NDIM >= 2) s = s + p._dy * ddy;
        if constexpr (NDIM >= 3) s = s + p._dz * ddz;
    } else {
        s = ddx / p.dx;
        if constexpr (NDIM >= 2) s = s + ddy / p.dy;
        if constexpr (NDIM >= 3) s = s + ddz / p.dz;
    }
    return s;
}

// H update (compute_update!): plain (H + dtau*(Hold/dt - div)) / (1.0 + dtau/dt)
// (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:48); perf (H + dtau*(_dt*Hold - div)) / (1.0 + _dt*dtau)
// (scripts/diff_3D/diff_3D_nonlin_perf.jl:40).
template <typename T, int FORM>
__device__ __forceinline__ T h_update(const KP<T>& p, T H, T Hold, T dtau, T div) {
    if constexpr (FORM == PTD_PERF) return (H + dtau * (p._dt * Hold - div)) / ((T)1.0 + p._dt * dtau);
    else return (H + dtau * (Hold / p.dt - div)) / ((T)1.0 + dtau / p.dt);
}

// Residual (check_res!): plain -(H - Hold)/dt - div (diff_3D_nonlin_multixpu.jl:53);
// perf -_dt*(H - Hold) - div (diff_2D_nonlin_perf.jl:45).
template <typename T, int FORM>
__device__ __forceinline__ T residual(const KP<T>& p, T H, T Hold, T div) {
    if constexpr (FORM == PTD_PERF) return -p._dt * (H - Hold) - div;
    else return -(H - Hold) / p.dt - div;
}

// ---------------------------------------------------------------------------------------------
// FAST arithmetic: Newton-refined hardware reciprocal / reciprocal square root
// ---------------------------------------------------------------------------------------------

// One Newton step on v_rcp_f64 / v_rsq_f64 (raw relative error ~5e-8, after the step ~4e-15). Used
// only for the damping coefficients of the FAST update rules below, whose fixed point does not depend
// on them; float32 sweeps use the hardware value (within 1 ulp of float32).
__device__ __forceinline__ double rcp_1nr(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ float rcp_1nr(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rsq_1nr(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    return fma(0.5 * y, fma(-(x * y), y, 1.0), y);
}
__device__ __forceinline__ float rsq_1nr(float x) { return __builtin_amdgcn_rsqf(x); }
// hardware square root (about 1 ulp), for coefficients whose rounding only affects the convergence rate
__device__ __forceinline__ float sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }

template <typename T>
__device__ __forceinline__ T fmaT(T a, T b, T c) { return fma(a, b, c); }

// H^3 floored (float32 1e-30, float64 1e-300): with the reference's initial condition H reaches 1e-31
// near the corners and H^3 underflows to 0 in float32 (0 * inf in the Re terms). The floor only acts
// where D = H^3 is negligible either way; it keeps a*h3 below normal for K = max_l^2/dt >= 1e-6 (the
// host rejects smaller K on these paths), so every reciprocal square root below is finite.
__device__ __forceinline__ float floor_h3(float x) { return fmaxf(x, 1e-30f); }
__device__ __forceinline__ double floor_h3(double x) { return fmax(x, 1e-300); }
// the same floor for u = 8*h3
__device__ __forceinline__ float floor_h3x8(float x) { return fmaxf(x, 8e-30f); }
__device__ __forceinline__ double floor_h3x8(double x) { return fmax(x, 8e-300); }

// FAST Re = pi + sqrt(pi^2 + K/h3) = pi + a*rsq(a*h3) with a = pi^2*h3 + K: one reciprocal square root
// instead of a reciprocal and a square root.
template <typename T>
__device__ __forceinline__ T re_fast(const KP<T>& p, T h3) {
    const T a = fmaT(p.pi2, h3, p.K);
    return p.pi + a * rsq_1nr(a * h3);
}
// The same from u = 8*h3 (faces: u = (Ha + Hb)^3 needs no 0.5): sqrt(pi^2 + 8K/u) = c*rsq(c*u), c = pi^2*u + 8K.
template <typename T>
__device__ __forceinline__ T re_fast8(const KP<T>& p, T u) {
    const T c = fmaT(p.pi2, u, p.K8);
    return p.pi + c * rsq_1nr(c * u);
}

// ---------------------------------------------------------------------------------------------
// one face / one cell, EXACT (reference order) or FAST (rearranged)
// ---------------------------------------------------------------------------------------------

enum { kExact = PTD_ARITH_EXACT, kFast = PTD_ARITH_FAST };

// Internal D case of the mixed-precision correction sweep for D = H^3: the sweep iterates on the
// correction delta (float32) around the float64 iterate Hc, whose float32 rounding travels through the
// kernels in the per-cell "aux" channel (the channel the step case uses for D). The coefficients only
// need float32 accuracy: the fixed point of the defect correction is set by the float64 residual.
// Its correction fluxes are stored scaled by 8*h of their axis (P = 8h*p), which takes the 1/(8h) of
// the face flux out of every face evaluation and into the divergence weights _dxs, _dys, _dzs.
constexpr int kDelta = 3;
// scale of the stored correction fluxes of axis a for D case DC (kDelta: 8*h; others: 1)
inline double corr_flux_scale(const ptd_params& p, int dc, int axis) {
    if (dc != PTD_NONLIN) return 1.0;
    return 8.0 * (axis == 0 ? p.dx : axis == 1 ? p.dy : p.dz);
}

template <typename T, int DC>
struct AuxOf {
    using type = T;
};
template <int DC>
constexpr bool has_aux() { return DC == PTD_STEP || DC == kDelta; }

template <typename T, int DC, int FORM, int ARITH>
struct Phys {
    using Ta = typename AuxOf<T, DC>::type;

    // New damped flux on the face (a | b) from the old flux q; q2 receives the residual flux
    // (-coef*dH/h, the q2 of the reference) when wanted. Da, Db: aux values (D, or Hc for kDelta).
    template <bool Q2>
    __device__ static __forceinline__ T face(const KP<T>& p, T q, T Ha, T Hb, Ta Da, Ta Db, T h, T ih, T ih8,
                                             T& q2) {
        const T dH = Hb - Ha;
        if constexpr (DC == kDelta) {
            // correction flux around Hc: g(Hc+delta) - g(Hc) = dD*dHc + (hbar+dbar)^3*ddelta with
            // dD = (hbar+dbar)^3 - hbar^3 = dbar*(3*hbar*(hbar+dbar) + dbar^2), no cancellation. In the
            // sums S = 2*hbar, Tt = 2*dbar, U = S + Tt: dD = Tt*(3*S*U + Tt^2)/8, (hbar+dbar)^3 = U^3/8.
            // q, g and the returned flux are the scaled ones (8h times the flux).
            const T S = Da + Db;
            const T Tt = Ha + Hb;
            const T U = S + Tt;
            const T u = floor_h3x8(U * U * U);
            const T dHc = Db - Da;
            const T g = fmaT(u, dH, Tt * fmaT((T)3 * S, U, Tt * Tt) * dHc);
            const T re = re_fast8(p, u);
            if constexpr (Q2) q2 = -g;
            return fmaT(-(q + g), re * rcp_1nr(re + p.c1), q);
        } else if constexpr (ARITH == kExact) {
            const T coef = face_coef<T, DC, FORM>(p, Ha, Hb, Da, Db);
            if constexpr (Q2) q2 = res_flux<T, FORM>(coef, dH, h, ih);
            return damped_flux<T, FORM>(q, face_theta<T, DC, FORM>(p, coef), coef, dH, h, ih);
        } else {
            if constexpr (DC == PTD_NONLIN) {
                // q' = (q*theta - g)/(1+theta) with theta = c1/Re is q' = q - (q + g)*w, w = Re/(Re + c1):
                // the fixed point q = -g does not depend on the rounding of w
                const T U = Ha + Hb;  // 2*Hface: u = U^3 = 8*Hface^3, g = u*(ih/8)*dH
                const T u = floor_h3x8(U * U * U);
                const T re = re_fast8(p, u);
                const T g = u * (ih8 * dH);
                if constexpr (Q2) q2 = -g;
                return fmaT(-(q + g), re * rcp_1nr(re + p.c1), q);
            } else {
                T coef;
                if constexpr (DC == PTD_LIN) coef = p.d0;
                else coef = (Da + Db) * (T)0.5;
                const T g = coef * (ih * dH);
                if constexpr (Q2) q2 = -g;
                return fmaT(q, p.theta_r, -g) * p.inv1p_theta;
            }
        }
    }

    // Residual flux only (compute_flux_res!).
    __device__ static __forceinline__ T face_res(const KP<T>& p, T Ha, T Hb, T Da, T Db, T h, T ih) {
        const T dH = Hb - Ha;
        if constexpr (ARITH == kExact) {
            return res_flux<T, FORM>(face_coef<T, DC, FORM>(p, Ha, Hb, Da, Db), dH, h, ih);
        } else {
            T coef;
            if constexpr (DC == PTD_NONLIN) {
                const T hf = (T)0.5 * (Ha + Hb);
                coef = hf * hf * hf;
            } else if constexpr (DC == PTD_LIN) {
                coef = p.d0;
            } else {
                coef = (Da + Db) * (T)0.5;
            }
            return -(coef * (ih * dH));
        }
    }

    template <int NDIM>
    __device__ static __forceinline__ T div(const KP<T>& p, T ddx, T ddy, T ddz) {
        if constexpr (ARITH == kExact && DC != kDelta) {
            return divergence<T, FORM, NDIM>(p, ddx, ddy, ddz);
        } else if constexpr (DC == kDelta) {
            T s = p._dxs * ddx;
            if constexpr (NDIM >= 2) s = fmaT(p._dys, ddy, s);
            if constexpr (NDIM >= 3) s = fmaT(p._dzs, ddz, s);
            return s;
        } else {
            T s = p._dx * ddx;
            if constexpr (NDIM >= 2) s = fmaT(p._dy, ddy, s);
            if constexpr (NDIM >= 3) s = fmaT(p._dz, ddz, s);
            return s;
        }
    }

    // H update of one cell from the flux divergence; dmax = @maxloc(D) (step case only); Ac = Hc
    // (kDelta only).
    __device__ static __forceinline__ T cell(const KP<T>& p, T H, T Hold, T div, T dmax, Ta Ac = Ta{}) {
        if constexpr (DC == kDelta) {
            // correction update with dtau from H = Hc + delta; Hold carries R*dt (source R)
            // P = h3*Re = pi*h3 + sqrt(h3*(pi^2*h3 + K))
            const T src = fmaT(p._dt, Hold, -div);
            const T Hf = Ac + H;
            const T h3 = floor_h3(Hf * Hf * Hf);
            const T P = fmaT(p.pi, h3, sqrt_hw(h3 * fmaT(p.pi2, h3, p.K)));
            return fmaT(fmaT(-H, p._dt, src), p.c2 * rcp_1nr(fmaT(p.c2, p._dt, P)), H);
        } else if constexpr (ARITH == kExact) {
            T dtr;
            if constexpr (DC == PTD_LIN) dtr = p.dtau_rho;
            else if constexpr (DC == PTD_STEP) dtr = step_dtau(p, dmax);
            else dtr = nonlin_dtau<T, FORM>(p, H);
            return h_update<T, FORM>(p, H, Hold, dtr, div);
        } else {
            const T src = fmaT(p._dt, Hold, -div);
            // (H + dtau*src)/(1 + _dt*dtau) = H + (src - _dt*H)*v, v = dtau/(1 + _dt*dtau): src - _dt*H is
            // the residual, so the fixed point does not depend on the rounding of v
            if constexpr (DC == PTD_LIN) {
                return fmaT(p.dtau_rho, src, H) * p.inv1p_dtau;
            } else if constexpr (DC == PTD_STEP) {
                // dtau = c3/m  ==>  v = c3/(m + _dt*c3)
                return fmaT(fmaT(-H, p._dt, src), p.c3 * rcp_1nr(fmaT(p.c3, p._dt, dmax)), H);
            } else {
                // dtau = c2/(H^3 Re) = c2/P  ==>  v = c2/(P + _dt*c2)
                const T h3 = floor_h3(H * H * H);
                const T P = h3 * re_fast(p, h3);
                return fmaT(fmaT(-H, p._dt, src), p.c2 * rcp_1nr(fmaT(p.c2, p._dt, P)), H);
            }
        }
    }

    // kDelta with damping carried between the two levels of the two-level sweep: the level-1 face /
    // cell computes its damping factor (w, v) and returns it; the level-2 evaluation of the same face /
    // cell one iteration later reuses it. The damping factors only set the convergence rate of the
    // correction sweep (its fixed point q = -g, residual = 0 does not depend on them).
    __device__ static __forceinline__ T face_w(const KP<T>& p, T q, T Ha, T Hb, Ta Da, Ta Db, T& w) {
        static_assert(DC == kDelta, "damping carry is a correction-sweep rule");
        const T dH = Hb - Ha;
        const T S = Da + Db;
        const T Tt = Ha + Hb;
        const T U = S + Tt;
        const T u = floor_h3x8(U * U * U);
        const T dHc = Db - Da;
        const T g = fmaT(u, dH, Tt * fmaT((T)3 * S, U, Tt * Tt) * dHc);
        const T re = re_fast8(p, u);
        w = re * rcp_1nr(re + p.c1);
        return fmaT(-(q + g), w, q);
    }
    template <bool Q2>
    __device__ static __forceinline__ T face_lag(T q, T Ha, T Hb, Ta Da, Ta Db, T w, T& q2) {
        static_assert(DC == kDelta, "damping carry is a correction-sweep rule");
        const T dH = Hb - Ha;
        const T S = Da + Db;
        const T Tt = Ha + Hb;
        const T U = S + Tt;
        const T dHc = Db - Da;
        const T g = fmaT(U * U * U, dH, Tt * fmaT((T)3 * S, U, Tt * Tt) * dHc);
        if constexpr (Q2) q2 = -g;
        return fmaT(-(q + g), w, q);
    }
    __device__ static __forceinline__ T cell_w(const KP<T>& p, T H, T Hold, T div, Ta Ac, T& v) {
        static_assert(DC == kDelta, "damping carry is a correction-sweep rule");
        const T src = fmaT(p._dt, Hold, -div);
        const T Hf = Ac + H;
        const T h3 = floor_h3(Hf * Hf * Hf);
        const T P = fmaT(p.pi, h3, sqrt_hw(h3 * fmaT(p.pi2, h3, p.K)));
        v = p.c2 * rcp_1nr(fmaT(p.c2, p._dt, P));
        return fmaT(fmaT(-H, p._dt, src), v, H);
    }
    __device__ static __forceinline__ T cell_lag(const KP<T>& p, T H, T Hold, T div, T v) {
        const T src = fmaT(p._dt, Hold, -div);
        return fmaT(fmaT(-H, p._dt, src), v, H);
    }

    __device__ static __forceinline__ T resid(const KP<T>& p, T H, T Hold, T div) {
        if constexpr (ARITH == kExact && DC != kDelta) return residual<T, FORM>(p, H, Hold, div);
        else return fmaT(-p._dt, H - Hold, -div);
    }
};

}  // namespace ptd

// ---------------------------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------------------------

struct ptd_ctx {
    ptd_params p;
    int device;
    hipStream_t stream;
    size_t esz;                 // bytes per sweep element (8 for FP64, 4 for FP32)
    int64_t nH, nQ[3], nR;      // element counts (nQ[a] = 0 for absent axes)
    void* H[2];                 // double-buffered H (fused sweeps write the other buffer)
    void* q[3][2];              // double-buffered damped fluxes
    int hcur, qcur;
    void* Hold;
    void* D;                    // step field (nullptr otherwise)
    void* dtau;                 // step dtau_rho (nullptr otherwise)
    void* q2[3];                // residual fluxes (kernel-level API and 1D/2D plain loop)
    void* resh;                 // ResH (kernel-level API)
    double* partial;            // per-block (sum, max) of the residual
    int64_t npartial;
    double* dred;               // device (sum, max) after the final reduction
    double* hred;               // pinned host mirror of dred (two (sum, max) slots)
    hipEvent_t ev0, ev1;        // timing (ptd_iterate_timed)
    hipEvent_t evc[2];          // completion of the checks in flight (pipelined 2D solves)
    void* raw[48];              // the context's device allocations (fields may start at an offset)
    int nraw;
    int64_t bytes;
    // PTD_MIXED: float32 correction sweeps around the float64 iterate H[hcur]
    bool mixed;
    void* dl[2];                // correction delta (float, nH)
    void* pq[3][2];             // correction fluxes (float)
    void* src;                  // R*dt (float, nH): the Hold slot of the correction sweep
    void* Df;                   // float copy of D (step case)
    void* hcf;                  // float copy of the float64 iterate (aux of the D = H^3 correction sweep)
    int dcur, pcur;
    int64_t outer_last;         // outer (defect-correction) iterations of the last PT solve
    int64_t inner_hint;         // float32 sweeps of the last correction solve that met its reduction target
};

namespace ptd {

// launchers (ptd_kernels.hip)
hipError_t launch_flux(ptd_ctx* c, int q2mode);       // 0: q only, 1: q and q2, 2: q2 only
hipError_t launch_update(ptd_ctx* c);
hipError_t launch_iter_params(ptd_ctx* c);
hipError_t launch_residual(ptd_ctx* c, bool write_resh);  // partial sums -> dred
hipError_t launch_fused(ptd_ctx* c, bool check);          // one fused PT iteration (2D / 3D)
hipError_t launch_fused2(ptd_ctx* c, bool check);         // two fused PT iterations in one pass (3D)
int64_t max_chunk_tz(const ptd_params& p, int64_t esz);
bool fused_available(const ptd_ctx* c);
bool fused2_available(const ptd_ctx* c);
hipError_t launch_fused3(ptd_ctx* c);  // three fused PT iterations in one pass (3D, no residual)
bool fused3_available(const ptd_ctx* c);
// whether the 3D sweeps of a problem with these parameters and element size take three-level passes
// (the FAST linear and step sweeps: float64 on every grid, float32 on grids of at least kT3MinCellsF32
// inner cells; PTD_T3=0 / 1 in the environment overrides)
bool fused3_wanted(const ptd_params& p, int64_t esz);
// PT iterations of the next launch of a 3D run of n iterations (3: a three-level pass, 2: a two-level
// pass, 1: a single sweep), given which passes are available. With check_last the run ends on a checking
// two-level pass whenever n >= 2: three-level passes are taken while they leave at least two iterations
// and never leave exactly three (which would end on a single checking sweep). Single-problem solves and
// batches both follow it, so their checking launches have the same tiling.
int next_launch_iters(int64_t n, bool check_last, bool t2, bool t3);
// 2D: niter (1..lds2d_max_iters()) PT iterations in one LDS-resident launch (check: the last one's
// plain-form residual reduced into dred), and whether a single-problem 2D solve runs its iterations that
// way (small grids)
hipError_t launch_lds2d(ptd_ctx* c, int niter, bool check);
int lds2d_max_iters();
bool use_2d_lds(const ptd_ctx* c);
int64_t lds2d_tiles(const ptd_params& p);
// residual partials (tiles) the checking launches of a fused 2D / 3D solve write at most
int64_t fused_check_tiles(const ptd_ctx* c);
// mixed precision: float32 correction sweeps (1 or 2 iterations), source and correction kernels
hipError_t launch_delta_sweep(ptd_ctx* c, bool two, bool check);
// three float32 correction iterations in one three-level pass (linear and step D), and whether a mixed
// solve takes them (fused3_wanted for its float32 FAST sweeps)
hipError_t launch_delta_sweep3(ptd_ctx* c);
bool delta_t3(const ptd_ctx* c);
hipError_t launch_f64_to_f32(ptd_ctx* c, float* dst, const double* src, int64_t n);
// correction fluxes <-> true fluxes: to_p: p = float(q + g(H)); !to_p: q = double(p) - g(H), where g(H)
// is the equilibrium flux magnitude coef*dH/h of the float64 iterate H on each face
hipError_t launch_flux_state(ptd_ctx* c, bool to_p);
// correction boundaries. outer_residual (start of a PT solve): residual of the float64 system at Hc +
// source + float32(Hc) + delta = 0 (partials -> dred). outer_step (after a correction, one pass):
// p <- p + g(Hc + delta) - g(Hc), Hc + delta into the other H buffer (flips hcur), then the same outputs
// at the new Hc, with delta = 0 in the other delta buffer (flips dcur)
hipError_t launch_outer_residual(ptd_ctx* c);
hipError_t launch_outer_step(ptd_ctx* c);

// gradients of the lin / step solves (ptd_grad.hip): adjoint source (H - Hobs)*dt with L's partials
// reduced into c->dred, and one step's gradient contribution
hipError_t launch_adj_source(ptd_ctx* c, const double* H, const double* Hobs, double* out);
hipError_t launch_grad_accum(ptd_ctx* c, const double* mu, const double* Hn, double* gD, double* gH0, bool first);
// sum of H^2 over the inner cells into c->dred
hipError_t launch_sumsq_inner(ptd_ctx* c, const double* H);
// adjoint of the nonlinear steps at the converged H of one step (ptd_grad.hip): one transposed PT
// iteration on (u faces, v cells) with right-hand side rhs/dt; lam = rho v (and with `resid` the residual
// rhs/dt - J^T lam reduced into c->dred); the step's contribution to dL/dH0
hipError_t launch_adjnl_iteration(ptd_ctx* c, const double* H, double* v, double* const u[3], double* const z[3],
                                  const double* rhs);
hipError_t launch_adjnl_lambda(ptd_ctx* c, const double* H, const double* v, double* lam, const double* rhs,
                               bool resid);
hipError_t launch_adjnl_grad(ptd_ctx* c, const double* H, const double* lam, double* gH0, bool first);

// ---- batches of independent float64 solves (ptd_batch.hip) -----------------------------------

// Device-resident state of one member of a batch. The sweep kernels read the member's fields and
// buffer parities from it; the controller kernel advances its PT loop after every check block.
struct BatchMember {
    KP<double> p;
    double* H[2];
    double* Hold;
    const double* D;
    double* q[3][2];
    double2* partial;     // per-tile residual partials of the checking launch
    int64_t nH;           // cells
    int64_t iter;         // PT iterations of the current physical step
    int64_t itmax;
    int32_t ndim, dcase, arith, nout;
    int32_t lpb;          // sweep launches per check block (each flips the H / q buffers)
    int32_t active;       // 0 once every physical step is done (or the error is NaN)
    int32_t hpar, qpar;   // H / q buffer holding the current iterate at the start of a check block
    int32_t assign;       // set by the controller: Hold <- H before the next block (the step ended)
    int32_t status;       // PTD_OK or PTD_ERR_NAN
    int32_t step, nt;     // physical steps done / requested
    int32_t stop_norm, ntiles_check;
    double tol, norm_len;
    double2 red;          // (sum of squares, max) of the last check
    uint64_t t_done;      // device wall clock when the member finished
};

// One block of a batched sweep launch: the member it works on, its tile and z chunk depth.
struct BatchTile {
    int32_t member, tile, tz, pad;
};

enum { kBatchMaxMembers = 1024 };

int validate_params(const ptd_params* p);          // the single-problem checks (ptd_api.hip)
int set_error(int code, const std::string& msg);   // sets ptd_last_error(), returns code

// The sweep a problem runs, as ptd_solve picks it: kSweep3D (two- / three-level 3D), kSweepLds2D (2D grids
// of at most 1024 32 x 32 cores, LDS-resident iterations), kSweep2D2 (two-level 2D, grids with enough
// tiles to fill the GPU) or kSweep2D (single-level 2D).
enum { kSweep2D = 2, kSweep3D = 3, kSweep2D2 = 4, kSweepLds2D = 5 };
bool use_2d_lds_params(const ptd_params& p);  // whether 2D solves of p run LDS-resident iterations
bool use_2d_t2(const ptd_params& p);
int batch_kind(const ptd_params& p);
int default_chunk_tz(const ptd_params& p);  // z chunk depth of the single-problem 3D sweeps
int check_chunk_tz(const ptd_params& p);    // ... of their checking launches (the batched ones follow)
// A batched sweep launch over the tiles of every member of one (kind, dcase, arith) group: launch
// `li` of the check block (buffers selected by the members' parities + li), with rows R per block
// (two-level 2D: R = rows per tile chunk).
hipError_t launch_batch_sweep(hipStream_t s, int kind, int dcase, int arith, int rows, bool check,
                              BatchMember* mem, const BatchTile* tiles, int ntiles, int li);
int batch_rows_nocheck(int kind, int dcase, int arith, int variant);  // rows of the non-checking launch
// three-level passes over the tiles of a group's members (float64 FAST 3D groups, batch_t3),
// rows fused3_rows() per block
hipError_t launch_batch_sweep3(hipStream_t s, int dcase, int arith, BatchMember* mem, const BatchTile* tiles,
                               int ntiles, int li);
bool batch_t3(const ptd_params& p);
int fused3_rows();
// niter (1..lds2d_max_iters()) LDS-resident PT iterations over the 32 x 32 cores of a 2D group's members
// (the checking launch also writes each core's residual partial)
hipError_t launch_batch_lds2d(hipStream_t s, int dcase, int arith, BatchMember* mem, const BatchTile* tiles,
                              int ntiles, int li, int niter, bool check);
int batch_rows_check(int kind, int dcase, int arith);                 // rows of the checking launch
hipError_t launch_batch_reduce(hipStream_t s, BatchMember* mem, const int32_t* members, int n);
// the reference loop's rules for the members of one group after a check block; *active_count (may be null)
// receives how many of them are still running. Hold .= H for the members whose step ended.
hipError_t launch_batch_control_list(hipStream_t s, BatchMember* mem, const int32_t* members, int n, int64_t* iters,
                                     double* errs, int32_t* active_count, int nt_cap);
hipError_t launch_batch_assign_list(hipStream_t s, BatchMember* mem, const int32_t* members, int n);
hipError_t launch_batch_clock(hipStream_t s, uint64_t* t);
// span[0] = device clock (start) / span[1] += clock - span[0] (stop)
hipError_t launch_batch_span(hipStream_t s, uint64_t* span, bool stop);

}  // namespace ptd
