// ptd_kernels.hip -- gfx950 kernels of libptd.
//
// Two families:
//  * reference-structured kernels, one launch per reference kernel (compute_flux!, compute_update!,
//    compute_flux_res!, check_res!, compute_iter_params!), any ndim, one thread per face or cell;
//  * the fused 3D PT sweep: flux and H update of one PT iteration in one launch. A block is a
//    64 x R tile of (x, y) columns that marches along z: x neighbours come from wavefront shuffles,
//    y neighbours from one LDS plane, z neighbours from registers carried along the march. H and the
//    fluxes are double-buffered so every iteration reads only the previous iterate (Jacobi order, the
//    same values the reference's two in-place kernels produce).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ptd_internal.hpp"

namespace ptd {

static constexpr int kBlock = 256;

__device__ __forceinline__ double nanmax(double a, double b) {
    return (a != a || a > b) ? a : b;
}
__device__ __forceinline__ float nanmax(float a, float b) {
    return (a != a || a > b) ? a : b;
}

template <typename T>
__device__ __forceinline__ T load_or0(const T* __restrict__ a, int64_t idx, bool ok) {
    return ok ? a[idx] : (T)0;
}

// Sum and NaN-propagating max over a block of `nthreads` threads (a multiple of 64), written by
// thread 0 to *out. Fixed order: wavefront butterfly, then waves in index order.
__device__ __forceinline__ void block_sum_max(double s, double m, int tid, int nthreads, double2* out) {
    __shared__ double ws[16], wm[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        m = nanmax(m, __shfl_xor(m, off));
    }
    const int wid = tid >> 6;
    if ((tid & 63) == 0) {
        ws[wid] = s;
        wm[wid] = m;
    }
    __syncthreads();
    if (tid == 0) {
        double S = 0.0, M = 0.0;
        for (int w = 0; w < (nthreads >> 6); ++w) {
            S += ws[w];
            M = nanmax(wm[w], M);
        }
        *out = make_double2(S, M);
    }
}

// ---------------------------------------------------------------------------------------------
// reference-structured kernels
// ---------------------------------------------------------------------------------------------

// compute_flux! / compute_flux_res! along one axis: thread t owns face t of the q array of that axis
// (column-major: qx (nx-1,my,mz), qy (mx,ny-1,mz), qz (mx,my,nz-1)).
// Q2MODE 0: damped flux only (perf compute_flux!); 1: damped and residual flux from the same H
// (plain compute_flux!, scripts/diff_3D/diff_3D_nonlin_multixpu.jl:37-45); 2: residual flux only
// (compute_flux_res!, scripts/diff_3D/diff_3D_nonlin_perf.jl:44-49).
// t = a + n1*(b + n2*c) -> (a, b, c): 32-bit divisions while t and the sizes fit in 32 bits (every
// grid below 2^32 elements; a 64-bit division costs several times more instructions), 64-bit otherwise.
__device__ __forceinline__ void split3(int64_t t, int64_t n1, int64_t n2, int64_t& a, int64_t& b, int64_t& c) {
    if (t <= 0xffffffffLL && n1 <= 0xffffffffLL && n2 <= 0xffffffffLL) {
        const uint32_t tu = (uint32_t)t, n1u = (uint32_t)n1, n2u = (uint32_t)n2;
        const uint32_t r = tu / n1u, cu = r / n2u;
        a = (int64_t)(tu - r * n1u);
        b = (int64_t)(r - cu * n2u);
        c = (int64_t)cu;
    } else {
        const int64_t r = t / n1;
        a = t - r * n1;
        c = r / n2;
        b = r - c * n2;
    }
}

template <typename T, int DC, int FORM, int AXIS, int Q2MODE>
__global__ void __launch_bounds__(kBlock)
k_flux(const KP<T> p, const T* __restrict__ H, const T* __restrict__ D, T* __restrict__ q,
       T* __restrict__ q2, int64_t count) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= count) return;
    int64_t a, step;
    T h, ih;
    if constexpr (AXIS == 0) {
        int64_t f, jj, kk;
        split3(t, p.nx - 1, p.my, f, jj, kk);
        a = f + p.nx * ((jj + p.hy) + p.ny * (kk + p.hz));
        step = 1;
        h = p.dx;
        ih = p._dx;
    } else if constexpr (AXIS == 1) {
        int64_t ii, f, kk;
        split3(t, p.mx, p.ny - 1, ii, f, kk);
        a = (ii + 1) + p.nx * (f + p.ny * (kk + p.hz));
        step = p.nx;
        h = p.dy;
        ih = p._dy;
    } else {
        int64_t ii, jj, f;
        split3(t, p.mx, p.my, ii, jj, f);
        a = (ii + 1) + p.nx * ((jj + 1) + p.ny * f);
        step = p.nx * p.ny;
        h = p.dz;
        ih = p._dz;
    }
    const int64_t b = a + step;
    const T Ha = H[a], Hb = H[b];
    T Da = 0, Db = 0;
    if constexpr (DC == PTD_STEP) {
        Da = D[a];
        Db = D[b];
    }
    const T coef = face_coef<T, DC, FORM>(p, Ha, Hb, Da, Db);
    const T dH = Hb - Ha;
    if constexpr (Q2MODE != 2) {
        const T th = face_theta<T, DC, FORM>(p, coef);
        q[t] = damped_flux<T, FORM>(q[t], th, coef, dH, h, ih);
    }
    if constexpr (Q2MODE >= 1) q2[t] = res_flux<T, FORM>(coef, dH, h, ih);
}

struct CellIdx {
    int64_t h;             // H index
    int64_t x0, x1;        // qx faces left/right
    int64_t y0, y1;        // qy faces
    int64_t z0, z1;        // qz faces
};

template <typename T>
__device__ __forceinline__ CellIdx cell_index(const KP<T>& p, int64_t t) {
    int64_t ii, jj, kk;
    split3(t, p.mx, p.my, ii, jj, kk);
    CellIdx c;
    c.h = (ii + 1) + p.nx * ((jj + p.hy) + p.ny * (kk + p.hz));
    c.x0 = ii + (p.nx - 1) * (jj + p.my * kk);
    c.x1 = c.x0 + 1;
    c.y0 = ii + p.mx * (jj + (p.ny - 1) * kk);
    c.y1 = c.y0 + p.mx;
    c.z0 = ii + p.mx * (jj + p.my * kk);
    c.z1 = c.z0 + p.mx * p.my;
    return c;
}

// compute_update! on @inn(H), in place (reads H only at the updated cell).
template <typename T, int DC, int FORM, int NDIM>
__global__ void __launch_bounds__(kBlock)
k_update(const KP<T> p, T* __restrict__ H, const T* __restrict__ Hold, const T* __restrict__ dtau,
         const T* __restrict__ qx, const T* __restrict__ qy, const T* __restrict__ qz, int64_t count) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= count) return;
    const CellIdx c = cell_index(p, t);
    const T Hc = H[c.h];
    T dtr;
    if constexpr (DC == PTD_LIN) dtr = p.dtau_rho;
    else if constexpr (DC == PTD_STEP) dtr = dtau[t];
    else dtr = nonlin_dtau<T, FORM>(p, Hc);
    const T ddx = qx[c.x1] - qx[c.x0];
    T ddy = 0, ddz = 0;
    if constexpr (NDIM >= 2) ddy = qy[c.y1] - qy[c.y0];
    if constexpr (NDIM >= 3) ddz = qz[c.z1] - qz[c.z0];
    const T div = divergence<T, FORM, NDIM>(p, ddx, ddy, ddz);
    H[c.h] = h_update<T, FORM>(p, Hc, Hold[c.h], dtr, div);
}

// check_res!: ResH (optional) and per-block (sum r^2, max |r|).
template <typename T, int FORM, int NDIM>
__global__ void __launch_bounds__(kBlock)
k_residual(const KP<T> p, const T* __restrict__ H, const T* __restrict__ Hold, const T* __restrict__ qx2,
           const T* __restrict__ qy2, const T* __restrict__ qz2, T* __restrict__ resh,
           double2* __restrict__ partial, int64_t count) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double s = 0.0, m = 0.0;
    if (t < count) {
        const CellIdx c = cell_index(p, t);
        const T ddx = qx2[c.x1] - qx2[c.x0];
        T ddy = 0, ddz = 0;
        if constexpr (NDIM >= 2) ddy = qy2[c.y1] - qy2[c.y0];
        if constexpr (NDIM >= 3) ddz = qz2[c.z1] - qz2[c.z0];
        const T div = divergence<T, FORM, NDIM>(p, ddx, ddy, ddz);
        const T r = residual<T, FORM>(p, H[c.h], Hold[c.h], div);
        if (resh) resh[t] = r;
        const double rd = (double)r;
        s = rd * rd;
        m = fabs(rd);
    }
    block_sum_max(s, m, threadIdx.x, kBlock, &partial[blockIdx.x]);
}

// compute_iter_params!: dtau_rho from the 3/5/7-point @maxloc(D).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_iter_params(const KP<T> p, const T* __restrict__ D, T* __restrict__ dtau, int64_t count) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= count) return;
    const CellIdx c = cell_index(p, t);
    const int64_t h = c.h;
    T m = fmax(fmax(D[h - 1], D[h + 1]), D[h]);
    if (p.hy) m = fmax(m, fmax(D[h - p.nx], D[h + p.nx]));
    if (p.hz) m = fmax(m, fmax(D[h - p.nx * p.ny], D[h + p.nx * p.ny]));
    dtau[t] = step_dtau(p, m);
}

// Sum of the per-block partials in a fixed order (deterministic): out = (sum, max).
__device__ __forceinline__ double2 reduce_partials_block(const double2* __restrict__ partial, int64_t n) {
    double s = 0.0, m = 0.0;
    for (int64_t t = threadIdx.x; t < n; t += kBlock) {
        const double2 v = partial[t];
        s += v.x;
        m = nanmax(v.y, m);
    }
    __shared__ double2 res;
    block_sum_max(s, m, threadIdx.x, kBlock, &res);
    __syncthreads();
    return res;
}

__global__ void __launch_bounds__(kBlock)
k_reduce_partials(const double2* __restrict__ partial, int64_t n, double* __restrict__ out) {
    const double2 res = reduce_partials_block(partial, n);
    if (threadIdx.x == 0) {
        out[0] = res.x;
        out[1] = res.y;
    }
}

// ---------------------------------------------------------------------------------------------
// fused 3D PT iteration
// ---------------------------------------------------------------------------------------------
//
// Block = 64 lanes (x) x R rows (y); lane L of tile bx holds cell i = 62*bx + L and row r of tile by
// holds j = (R-2)*by + r, so lanes 0/63 and rows 0/R-1 are one-cell overlaps with the neighbouring
// tiles. Every active lane computes the flux on its own lower x face (i-1|i) and lower y face (j-1|j);
// the upper faces of a cell are its neighbours' lower faces (shuffle down in x, LDS row+1 in y).
// Updated cells: lanes 1..62, rows 1..R-2. z: the block marches k = k0..k1-1; the upper z face flux of
// plane k is carried in a register as the lower z face of plane k+1. Loads for plane k+1 are issued
// before plane k is computed. Each face and each cell is written by exactly one thread.
// Addressing: the plane base of every array is block-uniform (scalar registers); each thread keeps a
// 32-bit byte offset into the plane.
// CHECK also evaluates the plain form's residual (residual fluxes from the pre-update H, transient
// term from the updated H: scripts/diff_3D/diff_3D_nonlin_multixpu.jl:41-43,53) and reduces it.

template <typename T>
__device__ __forceinline__ T ldp(const T* __restrict__ plane, uint32_t off, bool ok) {
    return ok ? *reinterpret_cast<const T*>(reinterpret_cast<const char*>(plane) + off) : (T)0;
}
template <typename T>
__device__ __forceinline__ void stp(T* __restrict__ plane, uint32_t off, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(plane) + off) = v;
}
// The same from a block-uniform chunk base and a 32-bit byte offset that already holds the plane: the
// address is (scalar base) + (vector offset), one global access with a scalar base and no 64-bit
// address arithmetic or per-array 64-bit address registers.
template <typename T>
__device__ __forceinline__ T ldc(const char* __restrict__ base, uint32_t off) {
    return *reinterpret_cast<const T*>(base + off);
}
template <typename T>
__device__ __forceinline__ void stc(char* __restrict__ base, uint32_t off, T v) {
    *reinterpret_cast<T*>(base + off) = v;
}

// x-neighbour exchange by one lane (one wavefront is one row of 64 cells) with DPP wave shifts
// (v_mov_b32_dpp wave_shr:1 / wave_shl:1) instead of ds_bpermute through the LDS crossbar: lane 0 (up) and
// lane 63 (down) keep their own value, as __shfl_up / __shfl_down would. Only at points where the whole wave
// is active (every call sits outside divergent branches).
template <typename V>
__device__ __forceinline__ V dpp_up1(V v) {
    if constexpr (sizeof(V) == 4) {
        const int i = __builtin_bit_cast(int, v);
        return __builtin_bit_cast(V, __builtin_amdgcn_update_dpp(i, i, 0x138, 0xf, 0xf, false));
    } else {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        const int lo = (int)(unsigned)u, hi = (int)(unsigned)(u >> 32);
        const unsigned a = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false);
        const unsigned b = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false);
        return __builtin_bit_cast(V, ((unsigned long long)b << 32) | a);
    }
}
template <typename V>
__device__ __forceinline__ V dpp_down1(V v) {
    if constexpr (sizeof(V) == 4) {
        const int i = __builtin_bit_cast(int, v);
        return __builtin_bit_cast(V, __builtin_amdgcn_update_dpp(i, i, 0x130, 0xf, 0xf, false));
    } else {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        const int lo = (int)(unsigned)u, hi = (int)(unsigned)(u >> 32);
        const unsigned a = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xf, 0xf, false);
        const unsigned b = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false);
        return __builtin_bit_cast(V, ((unsigned long long)b << 32) | a);
    }
}

// Operands of one plane p of the march (loaded two planes ahead).
template <typename T, typename Ta>
struct PlaneOps {
    T Hn;      // H(p+1)
    Ta Dn;     // aux(p+1): D (step) or Hc (correction sweep)
    T Ho;      // Hold(p)
    T qx, qy;  // old fluxes on this thread's lower x / y faces at plane p
    T qz;      // old flux on the upper z face (p | p+1)
};

template <typename T, int DC, int FORM, int ARITH, int R, bool CHECK>
__global__ void __launch_bounds__(64 * R)
k_fused3d(const KP<T> p, const T* __restrict__ Hin, T* __restrict__ Hout, const T* __restrict__ Hold,
          const typename AuxOf<T, DC>::type* __restrict__ D, const T* __restrict__ qxi, const T* __restrict__ qyi,
          const T* __restrict__ qzi, T* __restrict__ qxo, T* __restrict__ qyo, T* __restrict__ qzo,
          double2* __restrict__ partial, int tz, int ntx, int nty, int ntiles, int per_xcd) {
    using Ph = Phys<T, DC, FORM, ARITH>;
    using Ta = typename AuxOf<T, DC>::type;
    constexpr bool STEP = DC == PTD_STEP;
    constexpr bool AUX = has_aux<DC>();
    __shared__ T sH[R][64];
    __shared__ T sQ[R][64];
    __shared__ Ta sD[AUX ? R : 1][64];
    __shared__ T sQ2[CHECK ? R : 1][64];

    // XCD-aware tile order: consecutive blocks go round-robin to the 8 XCDs, so block b takes tile
    // (b % 8) * per_xcd + b / 8 and each XCD walks one contiguous range of tiles (shared halos in L2).
    const int b = blockIdx.x;
    const int tile = (b & 7) * per_xcd + (b >> 3);
    if (tile >= ntiles) return;
    const int bx = tile % ntx;
    const int rest = tile / ntx;
    const int by = rest % nty;
    const int bz = rest / nty;

    const int lane = threadIdx.x, row = threadIdx.y;
    const int nx = (int)p.nx, ny = (int)p.ny, nz = (int)p.nz;
    const int i = 62 * bx + lane;
    const int j = (R - 2) * by + row;
    const int k0 = 1 + tz * bz;
    const int k1 = (k0 + tz < nz - 1) ? k0 + tz : nz - 1;

    const bool xin = (i >= 1) && (i <= nx - 2);
    const bool yin = (j >= 1) && (j <= ny - 2);
    const bool rmid = (row >= 1) && (row <= R - 2);
    const bool lmid = (lane >= 1) && (lane <= 62);
    const bool upd = lmid && rmid && xin && yin;
    const bool fx = (lane >= 1) && (i <= nx - 1) && rmid && yin;
    const bool fy = (row >= 1) && (j <= ny - 1) && lmid && xin;
    const bool wx = fx && ((lane <= 62 && i <= nx - 2) || i == nx - 1);
    const bool wy = fy && ((row <= R - 2 && j <= ny - 2) || j == ny - 1);

    // Plane strides (elements) and per-thread byte offsets within a plane. Offsets are clamped into
    // each array's index range so every load is in bounds; values of lanes outside the computed set
    // are never stored and never reach a computed cell.
    const size_t psH = (size_t)nx * ny;
    const size_t psQx = (size_t)(nx - 1) * (ny - 2);
    const size_t psQy = (size_t)(nx - 2) * (ny - 1);
    const size_t psQz = (size_t)(nx - 2) * (ny - 2);
    const int ic = min(i, nx - 1), jc = min(j, ny - 1);
    const int ixq = max(1, min(i, nx - 1)), jxq = max(1, min(j, ny - 2));
    const int iyq = max(1, min(i, nx - 2)), jyq = max(1, min(j, ny - 1));
    const int jzq = max(1, min(j, ny - 2));
    constexpr uint32_t es = (uint32_t)sizeof(T);
    const uint32_t oH = (uint32_t)(ic + nx * jc) * es;
    const uint32_t oA = (uint32_t)(ic + nx * jc) * (uint32_t)sizeof(Ta);
    const uint32_t oQx = (uint32_t)((ixq - 1) + (nx - 1) * (jxq - 1)) * es;
    const uint32_t oQy = (uint32_t)((iyq - 1) + (nx - 2) * (jyq - 1)) * es;
    const uint32_t oQz = (uint32_t)((iyq - 1) + (nx - 2) * (jzq - 1)) * es;

    auto load_plane = [&](int pl) {  // operands of plane pl (clamped to the chunk's last plane)
        pl = pl < k1 ? pl : k1 - 1;
        PlaneOps<T, Ta> o;
        o.Hn = ldp(Hin + psH * (pl + 1), oH, true);
        if constexpr (AUX) o.Dn = ldp(D + psH * (pl + 1), oA, true);
        else o.Dn = Ta{};
        o.Ho = ldp(Hold + psH * pl, oH, true);
        o.qx = ldp(qxi + psQx * (pl - 1), oQx, true);
        o.qy = ldp(qyi + psQy * (pl - 1), oQy, true);
        o.qz = ldp(qzi + psQz * pl, oQz, true);
        return o;
    };

    // lower z face of the first plane
    T Hc = ldp(Hin + psH * k0, oH, true);
    const T Hkm = ldp(Hin + psH * (k0 - 1), oH, true);
    Ta Dc{}, Dkm{};
    if constexpr (AUX) {
        Dc = ldp(D + psH * k0, oA, true);
        Dkm = ldp(D + psH * (k0 - 1), oA, true);
    }
    T q2zL = 0;
    T qzL = Ph::template face<CHECK>(p, ldp(qzi + psQz * (k0 - 1), oQz, true), Hkm, Hc, Dkm, Dc, p.dz, p._dz, p._dz8, q2zL);
    if (upd && k0 == 1) stp(qzo, oQz, qzL);

    PlaneOps<T, Ta> A = load_plane(k0);
    PlaneOps<T, Ta> B = load_plane(k0 + 1);

    double acc = 0.0, amax = 0.0;
    for (int k = k0; k < k1; ++k) {
        sH[row][lane] = Hc;
        if constexpr (AUX) sD[row][lane] = Dc;
        __syncthreads();

        const T Hw = dpp_up1(Hc);
        const T Hs = sH[row > 0 ? row - 1 : 0][lane];
        Ta Dw{}, De{}, Ds{}, Dnn{};
        if constexpr (AUX) {
            Dw = dpp_up1(Dc);
            De = dpp_down1(Dc);
            Ds = sD[row > 0 ? row - 1 : 0][lane];
            Dnn = sD[row < R - 1 ? row + 1 : R - 1][lane];
        }
        T q2x = 0, q2y = 0, q2z = 0;
        const T qxn = Ph::template face<CHECK>(p, A.qx, Hw, Hc, Dw, Dc, p.dx, p._dx, p._dx8, q2x);
        const T qyn = Ph::template face<CHECK>(p, A.qy, Hs, Hc, Ds, Dc, p.dy, p._dy, p._dy8, q2y);
        const T qzn = Ph::template face<CHECK>(p, A.qz, Hc, A.Hn, Dc, A.Dn, p.dz, p._dz, p._dz8, q2z);
        sQ[row][lane] = qyn;
        if constexpr (CHECK) sQ2[row][lane] = q2y;
        __syncthreads();

        const T qxE = dpp_down1(qxn);
        const T qyN = sQ[row < R - 1 ? row + 1 : R - 1][lane];
        T dmax = 0;
        if constexpr (STEP) {
            dmax = fmax(fmax(Dw, De), Dc);
            dmax = fmax(dmax, fmax(Ds, Dnn));
            dmax = fmax(dmax, fmax(Dkm, A.Dn));
        }
        const T dv = Ph::template div<3>(p, qxE - qxn, qyN - qyn, qzn - qzL);
        const T Hnew = Ph::cell(p, Hc, A.Ho, dv, dmax, Dc);
        if constexpr (CHECK) {
            const T q2xE = dpp_down1(q2x);
            const T q2yN = sQ2[row < R - 1 ? row + 1 : R - 1][lane];
            const T dv2 = Ph::template div<3>(p, q2xE - q2x, q2yN - q2y, q2z - q2zL);
            const double r = upd ? (double)Ph::resid(p, Hnew, A.Ho, dv2) : 0.0;
            acc += r * r;
            amax = nanmax(fabs(r), amax);
        }
        // stores of plane k, then the operands of plane k+2 (so waiting for plane k+1's operands
        // never waits for these stores)
        if (upd) {
            stp(Hout + psH * k, oH, Hnew);
            stp(qzo + psQz * k, oQz, qzn);
        }
        if (wx) stp(qxo + psQx * (k - 1), oQx, qxn);
        if (wy) stp(qyo + psQy * (k - 1), oQy, qyn);
        qzL = qzn;
        q2zL = q2z;
        Hc = A.Hn;
        if constexpr (AUX) {
            Dkm = Dc;
            Dc = A.Dn;
        }
        A = B;
        B = load_plane(k + 2);
    }
    if constexpr (CHECK) block_sum_max(acc, amax, row * 64 + lane, 64 * R, &partial[tile]);
}

// ---------------------------------------------------------------------------------------------
// fused 2D PT iteration
// ---------------------------------------------------------------------------------------------
//
// The 3D sweep's tile on a single plane: block = 64 lanes (x) x R rows (y); lane L of tile bx holds
// i = 62*bx + L and row r of tile by holds j = (R-2)*by + r. Each thread computes the flux on its lower
// x face (shuffle for the upper one) and lower y face (LDS row+1 for the upper one), then updates its
// cell (lanes 1..62, rows 1..R-2). Same expressions, same ownership of every face and cell as the 3D
// sweep, so EXACT arithmetic reproduces compute_flux! + compute_update! bit for bit. CHECK also
// evaluates the plain form's residual (residual fluxes from the pre-update H, transient term from the
// updated H: scripts/diff_2D_nonlin.jl:28-37) and reduces it per tile.
template <typename T, int DC, int FORM, int ARITH, int R, bool CHECK>
__device__ __forceinline__ void fused2d_tile(const KP<T>& p, const T* __restrict__ Hin, T* __restrict__ Hout,
                                             const T* __restrict__ Hold, const T* __restrict__ D,
                                             const T* __restrict__ qxi, const T* __restrict__ qyi,
                                             T* __restrict__ qxo, T* __restrict__ qyo,
                                             double2* __restrict__ part, int bx, int by) {
    using Ph = Phys<T, DC, FORM, ARITH>;
    constexpr bool STEP = DC == PTD_STEP;
    __shared__ T sH[R][64];
    __shared__ T sQ[R][64];
    __shared__ T sD[STEP ? R : 1][64];
    __shared__ T sQ2[CHECK ? R : 1][64];

    const int lane = threadIdx.x, row = threadIdx.y;
    const int nx = (int)p.nx, ny = (int)p.ny;
    const int i = 62 * bx + lane;
    const int j = (R - 2) * by + row;
    const bool xin = (i >= 1) && (i <= nx - 2);
    const bool yin = (j >= 1) && (j <= ny - 2);
    const bool rmid = (row >= 1) && (row <= R - 2);
    const bool lmid = (lane >= 1) && (lane <= 62);
    const bool upd = lmid && rmid && xin && yin;
    const bool fx = (lane >= 1) && (i <= nx - 1) && rmid && yin;
    const bool fy = (row >= 1) && (j <= ny - 1) && lmid && xin;
    const bool wx = fx && ((lane <= 62 && i <= nx - 2) || i == nx - 1);
    const bool wy = fy && ((row <= R - 2 && j <= ny - 2) || j == ny - 1);

    // clamped indices: every load is in bounds; values outside the computed set are never stored
    const int ic = min(i, nx - 1), jc = min(j, ny - 1);
    const int ixq = max(1, min(i, nx - 1)), jxq = max(1, min(j, ny - 2));
    const int iyq = max(1, min(i, nx - 2)), jyq = max(1, min(j, ny - 1));
    const int64_t oH = (int64_t)ic + (int64_t)nx * jc;
    const int64_t oQx = (int64_t)(ixq - 1) + (int64_t)(nx - 1) * (jxq - 1);
    const int64_t oQy = (int64_t)(iyq - 1) + (int64_t)(nx - 2) * (jyq - 1);

    const T Hc = Hin[oH];
    const T Ho = Hold[oH];
    T Dc = 0;
    if constexpr (STEP) Dc = D[oH];
    sH[row][lane] = Hc;
    if constexpr (STEP) sD[row][lane] = Dc;
    __syncthreads();

    const T Hw = dpp_up1(Hc);
    const T Hs = sH[row > 0 ? row - 1 : 0][lane];
    T Dw = 0, De = 0, Ds = 0, Dn = 0;
    if constexpr (STEP) {
        Dw = dpp_up1(Dc);
        De = dpp_down1(Dc);
        Ds = sD[row > 0 ? row - 1 : 0][lane];
        Dn = sD[row < R - 1 ? row + 1 : R - 1][lane];
    }
    T q2x = 0, q2y = 0;
    const T qxn = Ph::template face<CHECK>(p, qxi[oQx], Hw, Hc, Dw, Dc, p.dx, p._dx, p._dx8, q2x);
    const T qyn = Ph::template face<CHECK>(p, qyi[oQy], Hs, Hc, Ds, Dc, p.dy, p._dy, p._dy8, q2y);
    sQ[row][lane] = qyn;
    if constexpr (CHECK) sQ2[row][lane] = q2y;
    __syncthreads();

    const T qxE = dpp_down1(qxn);
    const T qyN = sQ[row < R - 1 ? row + 1 : R - 1][lane];
    T dmax = 0;
    if constexpr (STEP) {
        dmax = fmax(fmax(Dw, De), Dc);
        dmax = fmax(dmax, fmax(Ds, Dn));
    }
    const T dv = Ph::template div<2>(p, qxE - qxn, qyN - qyn, (T)0);
    const T Hnew = Ph::cell(p, Hc, Ho, dv, dmax, Dc);
    if (upd) Hout[oH] = Hnew;
    if (wx) qxo[oQx] = qxn;
    if (wy) qyo[oQy] = qyn;
    if constexpr (CHECK) {
        const T q2xE = dpp_down1(q2x);
        const T q2yN = sQ2[row < R - 1 ? row + 1 : R - 1][lane];
        const T dv2 = Ph::template div<2>(p, q2xE - q2x, q2yN - q2y, (T)0);
        const double r = upd ? (double)Ph::resid(p, Hnew, Ho, dv2) : 0.0;
        block_sum_max(r * r, upd ? fabs(r) : 0.0, row * 64 + lane, 64 * R, part);
    }
}

template <typename T, int DC, int FORM, int ARITH, int R, bool CHECK>
__global__ void __launch_bounds__(64 * R)
k_fused2d(const KP<T> p, const T* __restrict__ Hin, T* __restrict__ Hout, const T* __restrict__ Hold,
          const T* __restrict__ D, const T* __restrict__ qxi, const T* __restrict__ qyi, T* __restrict__ qxo,
          T* __restrict__ qyo, double2* __restrict__ partial, int ntx, int ntiles, int per_xcd) {
    const int b = blockIdx.x;
    const int tile = (b & 7) * per_xcd + (b >> 3);  // XCD-aware: each XCD walks a contiguous tile range
    if (tile >= ntiles) return;
    fused2d_tile<T, DC, FORM, ARITH, R, CHECK>(p, Hin, Hout, Hold, D, qxi, qyi, qxo, qyo, &partial[tile],
                                               tile % ntx, tile / ntx);
}

// ---------------------------------------------------------------------------------------------
// fused 2D sweep over two PT iterations (temporal blocking, one wavefront per tile)
// ---------------------------------------------------------------------------------------------
//
// A tile is a strip of 64 columns (lane L holds i = 60*bx - 1 + L) marching along y over the rows of
// its chunk; the block is that one wavefront, so the march needs no LDS and no barrier: x neighbours
// are lane shuffles, y neighbours are the rows carried in registers. Level 1 (iteration n+1) runs one
// row ahead of level 2 (iteration n+2): at step jj level 1 updates row jj and level 2 row jj-1, each
// face and cell with the same expressions as the single-level 2D sweep, so two iterations give the
// same bits as two k_fused2d launches. Level 1 is valid on lanes 1..62 and rows j0-1..j1, level 2 on
// lanes 2..61 and rows j0..j1-1 (j0 = 1 + ty*by, j1 = min(j0 + ty, ny - 1)); each face and cell is
// written by exactly one tile. CHECK also reduces the plain form's residual of iteration n+2.
template <typename T, int DC, int FORM, int ARITH, bool CHECK>
__device__ __forceinline__ void fused2d_t2_tile(const KP<T>& p, const T* __restrict__ Hin, T* __restrict__ Hout,
                                                const T* __restrict__ Hold, const T* __restrict__ D,
                                                const T* __restrict__ qxi, const T* __restrict__ qyi,
                                                T* __restrict__ qxo, T* __restrict__ qyo,
                                                double2* __restrict__ part, int bx, int by, int ty) {
    using Ph = Phys<T, DC, FORM, ARITH>;
    constexpr bool STEP = DC == PTD_STEP;
    const int lane = threadIdx.x;
    const int nx = (int)p.nx, ny = (int)p.ny;
    const int i = 60 * bx - 1 + lane;
    const int j0 = 1 + ty * by;
    const int j1 = j0 + ty < ny - 1 ? j0 + ty : ny - 1;  // level 2 updates rows j0..j1-1
    const int ja = j0 - 1;                               // level 1 runs rows ja..j1
    const bool xin = i >= 1 && i <= nx - 2;
    const bool l1 = lane >= 1 && lane <= 62 && xin;      // level-1 update lanes
    const bool l2 = lane >= 2 && lane <= 61 && xin;      // level-2 update lanes (owned cells)
    // owned x faces: lanes 2..61 inside the domain, and the face (nx-2 | nx-1) at the domain end
    const bool wx = lane >= 2 && lane <= 62 && ((lane <= 61 && i <= nx - 2) || i == nx - 1);
    const bool wy = l2;                                  // owned y faces: the lower face of each owned cell

    const int ic = max(0, min(i, nx - 1));
    const int ixq = max(1, min(i, nx - 1)), iyq = max(1, min(i, nx - 2));
    auto row_h = [&](int r) { return (int64_t)ic + (int64_t)nx * r; };            // H, Hold, D of row r
    auto row_qx = [&](int r) {                                                   // x face (i-1 | i) of row r
        const int rc = max(1, min(r, ny - 2));
        return (int64_t)(ixq - 1) + (int64_t)(nx - 1) * (rc - 1);
    };
    auto row_qy = [&](int f) {                                                   // y face (f | f+1)
        const int fc = max(0, min(f, ny - 2));
        return (int64_t)(iyq - 1) + (int64_t)(nx - 2) * fc;
    };

    // rows ja-1 .. : H(jj+1) is loaded one row ahead
    const int rb = ja > 0 ? ja - 1 : 0;
    T Hc = Hin[row_h(ja)];
    T Hn = Hin[row_h(ja + 1 < ny ? ja + 1 : ny - 1)];
    T Dp = 0, Dc = 0, Dn = 0, Dpp = 0;  // D at rows jj-1, jj, jj+1, jj-2
    if constexpr (STEP) {
        Dp = D[row_h(rb)];
        Dc = D[row_h(ja)];
        Dn = D[row_h(ja + 1 < ny ? ja + 1 : ny - 1)];
    }
    T unused = 0;
    // level-1 lower y face of row ja: face (ja-1 | ja), from H(ja-1), H(ja) (needed when ja is inner)
    T q1yL = Ph::template face<false>(p, qyi[row_qy(ja - 1)], Hin[row_h(rb)], Hc, Dp, Dc, p.dy, p._dy, p._dy8,
                                      unused);
    T H1p = 0, Hop = 0, q1xp = 0, q2yL = 0, r2yL = 0;
    T acc = 0, amax = 0;

    for (int jj = ja; jj <= j1; ++jj) {
        const bool rin = jj >= 1 && jj <= ny - 2;
        // operands of row jj (and of the next row for the upper y face)
        const T Ho = Hold[row_h(jj)];
        const T qxo_ = qxi[row_qx(jj)];
        const T qyo_ = qyi[row_qy(jj)];
        const T Hnn = Hin[row_h(jj + 2 < ny ? jj + 2 : ny - 1)];  // prefetch for the next step
        T Dnn = 0;
        if constexpr (STEP) Dnn = D[row_h(jj + 2 < ny ? jj + 2 : ny - 1)];

        // level 1, row jj
        const T Hw = dpp_up1(Hc);
        T Dw = 0, De = 0;
        if constexpr (STEP) {
            Dw = dpp_up1(Dc);
            De = dpp_down1(Dc);
        }
        const T q1x = Ph::template face<false>(p, qxo_, Hw, Hc, Dw, Dc, p.dx, p._dx, p._dx8, unused);
        const T q1yU = Ph::template face<false>(p, qyo_, Hc, Hn, Dc, Dn, p.dy, p._dy, p._dy8, unused);
        const T q1xE = dpp_down1(q1x);
        T dmax1 = 0;
        if constexpr (STEP) {
            dmax1 = fmax(fmax(Dw, De), Dc);
            dmax1 = fmax(dmax1, fmax(Dp, Dn));
        }
        const T dv1 = Ph::template div<2>(p, q1xE - q1x, q1yU - q1yL, (T)0);
        const T H1c = Ph::cell(p, Hc, Ho, dv1, dmax1, Dc);
        const T H1 = (l1 && rin) ? H1c : Hc;

        // level 2: upper y face of row jj-1 (face (jj-1 | jj)) and row jj-1 itself
        T r2yU = 0;
        const T q2yU = Ph::template face<CHECK>(p, q1yL, H1p, H1, Dp, Dc, p.dy, p._dy, p._dy8, r2yU);
        const int m = jj - 1;
        if (m >= j0) {
            const T H1w = dpp_up1(H1p);
            T Dpw = 0, Dpe = 0;
            if constexpr (STEP) {
                Dpw = dpp_up1(Dp);
                Dpe = dpp_down1(Dp);
            }
            T r2x = 0;
            const T q2x = Ph::template face<CHECK>(p, q1xp, H1w, H1p, Dpw, Dp, p.dx, p._dx, p._dx8, r2x);
            const T q2xE = dpp_down1(q2x);
            T dmax2 = 0;
            if constexpr (STEP) {
                dmax2 = fmax(fmax(Dpw, Dpe), Dp);
                dmax2 = fmax(dmax2, fmax(Dpp, Dc));
            }
            const T dv2 = Ph::template div<2>(p, q2xE - q2x, q2yU - q2yL, (T)0);
            const T H2 = Ph::cell(p, H1p, Hop, dv2, dmax2, Dp);
            if constexpr (CHECK) {
                const T r2xE = dpp_down1(r2x);
                const T dvr = Ph::template div<2>(p, r2xE - r2x, r2yU - r2yL, (T)0);
                const T rr = l2 ? Ph::resid(p, H2, Hop, dvr) : (T)0;
                acc += rr * rr;
                amax = nanmax(fabs(rr), amax);
            }
            if (l2) Hout[row_h(m)] = H2;
            if (wx) qxo[row_qx(m)] = q2x;
            if (wy) qyo[row_qy(m - 1)] = q2yL;
            if (wy && m == ny - 2) qyo[row_qy(m)] = q2yU;  // face (ny-2 | ny-1): the domain end
        }
        // carry to row jj+1
        q1yL = q1yU;
        q2yL = q2yU;
        r2yL = r2yU;
        H1p = H1;
        Hop = Ho;
        q1xp = q1x;
        Hc = Hn;
        Hn = Hnn;
        if constexpr (STEP) {
            Dpp = Dp;
            Dp = Dc;
            Dc = Dn;
            Dn = Dnn;
        }
    }
    if constexpr (CHECK) block_sum_max((double)acc, (double)amax, lane, 64, part);
}

template <typename T, int DC, int FORM, int ARITH, bool CHECK>
__global__ void __launch_bounds__(64)
k_fused2d_t2(const KP<T> p, const T* __restrict__ Hin, T* __restrict__ Hout, const T* __restrict__ Hold,
             const T* __restrict__ D, const T* __restrict__ qxi, const T* __restrict__ qyi, T* __restrict__ qxo,
             T* __restrict__ qyo, double2* __restrict__ partial, int ty, int ntx, int ntiles, int per_xcd) {
    const int b = blockIdx.x;
    const int tile = (b & 7) * per_xcd + (b >> 3);  // XCD-aware: each XCD walks a contiguous tile range
    if (tile >= ntiles) return;
    fused2d_t2_tile<T, DC, FORM, ARITH, CHECK>(p, Hin, Hout, Hold, D, qxi, qyi, qxo, qyo, &partial[tile],
                                               tile % ntx, tile / ntx, ty);
}

// ---------------------------------------------------------------------------------------------
// 2D PT iterations resident in LDS (small grids)
// ---------------------------------------------------------------------------------------------
//
// On small 2D grids one PT iteration is a few microseconds of load -> flux -> update latency plus a
// launch gap, whatever its bandwidth. This sweep runs up to kLdsHalo iterations per launch: a block owns
// a kLdsCore^2 core of cells and loads the core plus a kLdsHalo-cell frame once. Every iteration's
// results are exact only on a region one cell narrower per side than the last (a cell needs its
// neighbours' previous values), so after m <= kLdsHalo iterations the core holds exactly what m single
// sweeps give; the block then writes its core cells and the faces it owns (the lower x / y face of each
// core cell) and nothing else. Iteration k computes only the region its remaining iterations need.
//
// Frame cell c = y*W + x (W = kLdsCore + 2*kLdsHalo) is global cell (kLdsCore*bx - kLdsHalo + x,
// kLdsCore*by - kLdsHalo + y); its state (H, Hold, the flux on its lower x and y face, D and the step
// case's neighbourhood maximum) lives in the registers of thread c % kLdsThreads, slot c / kLdsThreads.
// An iteration publishes H to LDS, computes each owned cell's lower faces from its west / south
// neighbour's H and publishes the fluxes, then updates each owned cell from its east / north neighbours'
// fluxes: the same expressions on the same operands as the single sweep (compute_flux! then
// compute_update!, scripts/diff_2D_nonlin.jl:18-44), so EXACT arithmetic gives the same bits.
//
// CHECK (plain form): the last iteration also evaluates the residual on the core (residual fluxes from
// the pre-update H, transient term from the updated H: scripts/diff_2D_nonlin.jl:28-37) and every block
// writes its (sum r^2, max|r|) partial. (Finishing the reduction in the last block to arrive needs an
// agent-scope fence per block, which writes back the XCD's L2: measured 2-13 ms slower per 512^2 -
// 1024^2 solve than the separate one-block reduction.)
static constexpr int kLdsCore = 32;
static constexpr int kLdsHalo = 10;
static constexpr int kLdsW = kLdsCore + 2 * kLdsHalo;  // 52
static constexpr int kLdsThreads = 1024;

// One 32 x 32 core (tx, ty) and its frame through niter PT iterations; `part` receives the core's residual
// partial (CHECK). Called by the single-problem kernel below and by the batched kernel.
template <typename T, int DC, int FORM, int ARITH, bool CHECK>
__device__ __forceinline__ void lds2d_tile(const KP<T>& p, const T* __restrict__ Hin, T* __restrict__ Hout,
                                           const T* __restrict__ Hold, const T* __restrict__ D,
                                           const T* __restrict__ qxi, const T* __restrict__ qyi,
                                           T* __restrict__ qxo, T* __restrict__ qyo, int niter, int tx, int ty,
                                           double2* __restrict__ part) {
    using Ph = Phys<T, DC, FORM, ARITH>;
    constexpr bool STEP = DC == PTD_STEP;
    constexpr int W = kLdsW, NC = W * W;
    constexpr int NS = (NC + kLdsThreads - 1) / kLdsThreads;
    __shared__ T sH[NC];
    __shared__ T sQx[NC];
    __shared__ T sQy[NC];
    __shared__ T sRx[CHECK ? NC : 1];  // residual fluxes of the checking iteration
    __shared__ T sRy[CHECK ? NC : 1];

    const int nx = (int)p.nx, ny = (int)p.ny;
    const int x0 = kLdsCore * tx - kLdsHalo;
    const int y0 = kLdsCore * ty - kLdsHalo;
    const int t = threadIdx.x;

    T H[NS], Ho[NS], qx[NS], qy[NS], Dc[NS], Dw[NS], Ds[NS], dmax[NS];
    int xs[NS], ys[NS];
    bool fxs[NS], fys[NS], upds[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int c = t + kLdsThreads * s;
        const bool valid = c < NC;
        const int y = c / W, x = c - y * W;
        const int i = x0 + x, j = y0 + y;
        const bool cin = valid && i >= 0 && i < nx && j >= 0 && j < ny;
        const int64_t o = (int64_t)i + (int64_t)nx * j;
        xs[s] = valid ? x : -W;  // invalid slots fall outside every active region
        ys[s] = valid ? y : -W;
        fxs[s] = valid && i >= 1 && i <= nx - 1 && j >= 1 && j <= ny - 2;
        fys[s] = valid && j >= 1 && j <= ny - 1 && i >= 1 && i <= nx - 2;
        upds[s] = valid && i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2;
        H[s] = cin ? Hin[o] : (T)0;
        Ho[s] = cin ? Hold[o] : (T)0;
        Dc[s] = (STEP && cin) ? D[o] : (T)0;
        qx[s] = fxs[s] ? qxi[(int64_t)(i - 1) + (int64_t)(nx - 1) * (j - 1)] : (T)0;
        qy[s] = fys[s] ? qyi[(int64_t)(i - 1) + (int64_t)(nx - 2) * (j - 1)] : (T)0;
        Dw[s] = Ds[s] = dmax[s] = (T)0;
    }
    if constexpr (STEP) {
        // D is constant over the launch: neighbours and the neighbourhood maximum once, through sH
#pragma unroll
        for (int s = 0; s < NS; ++s)
            if (t + kLdsThreads * s < NC) sH[t + kLdsThreads * s] = Dc[s];
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int c = t + kLdsThreads * s;
            if (xs[s] >= 1) Dw[s] = sH[c - 1];  // faces reach column / row W-1
            if (ys[s] >= 1) Ds[s] = sH[c - W];
            if (xs[s] >= 1 && xs[s] <= W - 2 && ys[s] >= 1 && ys[s] <= W - 2) {
                dmax[s] = fmax(fmax(Dw[s], sH[c + 1]), Dc[s]);
                dmax[s] = fmax(dmax[s], fmax(Ds[s], sH[c + W]));
            }
        }
        __syncthreads();
    }

    double rs = 0.0, rm = 0.0;  // residual (sum r^2, max|r|) of this thread's core cells
    for (int k = 1; k <= niter; ++k) {
        const bool chk = CHECK && k == niter;
        // after this iteration cells [lo, W-1-lo] and faces [lo, W-lo] along each axis must be exact
        const int lo = k + kLdsHalo - niter;
#pragma unroll
        for (int s = 0; s < NS; ++s)
            if (t + kLdsThreads * s < NC) sH[t + kLdsThreads * s] = H[s];
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int c = t + kLdsThreads * s;
            const bool yrow = ys[s] >= lo && ys[s] <= W - 1 - lo, xcol = xs[s] >= lo && xs[s] <= W - 1 - lo;
            T rx = 0, ry = 0;
            // qx faces [lo, W-lo] of rows [lo, W-1-lo]; qy faces [lo, W-lo] of columns [lo, W-1-lo]
            if (fxs[s] && yrow && xs[s] >= lo && xs[s] <= W - lo) {
                if (chk)
                    qx[s] = Ph::template face<true>(p, qx[s], sH[c - 1], H[s], Dw[s], Dc[s], p.dx, p._dx, p._dx8, rx);
                else
                    qx[s] = Ph::template face<false>(p, qx[s], sH[c - 1], H[s], Dw[s], Dc[s], p.dx, p._dx, p._dx8, rx);
            }
            if (fys[s] && xcol && ys[s] >= lo && ys[s] <= W - lo) {
                if (chk)
                    qy[s] = Ph::template face<true>(p, qy[s], sH[c - W], H[s], Ds[s], Dc[s], p.dy, p._dy, p._dy8, ry);
                else
                    qy[s] = Ph::template face<false>(p, qy[s], sH[c - W], H[s], Ds[s], Dc[s], p.dy, p._dy, p._dy8, ry);
            }
            if (c < NC) {
                sQx[c] = qx[s];
                sQy[c] = qy[s];
                if constexpr (CHECK) {
                    if (chk) {
                        sRx[c] = rx;
                        sRy[c] = ry;
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int c = t + kLdsThreads * s;
            if (upds[s] && xs[s] >= lo && xs[s] <= W - 1 - lo && ys[s] >= lo && ys[s] <= W - 1 - lo) {
                const T dv = Ph::template div<2>(p, sQx[c + 1] - qx[s], sQy[c + W] - qy[s], (T)0);
                H[s] = Ph::cell(p, H[s], Ho[s], dv, dmax[s], Dc[s]);
                if constexpr (CHECK) {
                    if (chk) {  // lo == kLdsHalo: exactly the core
                        const T dv2 = Ph::template div<2>(p, sRx[c + 1] - sRx[c], sRy[c + W] - sRy[c], (T)0);
                        const double r = (double)Ph::resid(p, H[s], Ho[s], dv2);
                        rs += r * r;
                        rm = nanmax(fabs(r), rm);
                    }
                }
            }
        }
    }

    // the core: inner cells and the lower faces of its cells
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int x = xs[s], y = ys[s];
        if (x < kLdsHalo || x >= kLdsHalo + kLdsCore || y < kLdsHalo || y >= kLdsHalo + kLdsCore) continue;
        const int i = x0 + x, j = y0 + y;
        if (upds[s]) Hout[(int64_t)i + (int64_t)nx * j] = H[s];
        if (fxs[s]) qxo[(int64_t)(i - 1) + (int64_t)(nx - 1) * (j - 1)] = qx[s];
        if (fys[s]) qyo[(int64_t)(i - 1) + (int64_t)(nx - 2) * (j - 1)] = qy[s];
    }

    if constexpr (CHECK) block_sum_max(rs, rm, t, kLdsThreads, part);
}

template <typename T, int DC, int FORM, int ARITH, bool CHECK>
__global__ void __launch_bounds__(kLdsThreads)
k_lds2d(const KP<T> p, const T* __restrict__ Hin, T* __restrict__ Hout, const T* __restrict__ Hold,
        const T* __restrict__ D, const T* __restrict__ qxi, const T* __restrict__ qyi, T* __restrict__ qxo,
        T* __restrict__ qyo, int niter, int ntx, int ntiles, int per_xcd, double2* __restrict__ partial) {
    const int b = blockIdx.x;
    const int tile = (b & 7) * per_xcd + (b >> 3);  // XCD-aware: each XCD walks a contiguous tile range
    if (tile >= ntiles) return;
    lds2d_tile<T, DC, FORM, ARITH, CHECK>(p, Hin, Hout, Hold, D, qxi, qyi, qxo, qyo, niter, tile % ntx, tile / ntx,
                                          &partial[tile]);
}

// ---------------------------------------------------------------------------------------------
// fused 3D sweep over two PT iterations (temporal blocking)
// ---------------------------------------------------------------------------------------------
//
// Same z-marching block as k_fused3d, but every plane goes through two PT iterations before it is
// written: level 1 (iteration n+1) runs one plane ahead of level 2 (iteration n+2). Level-1 values
// (H1 and the level-1 fluxes) live only in registers and LDS, so HBM sees one read of the iterate-n
// fields and one write of the iterate-(n+2) fields for two iterations. The tile halo shrinks by one
// cell per level: lane L holds i = 60*bx - 1 + L and row r holds j = (R-4)*by - 1 + r; level 1 updates
// lanes 1..62 / rows 1..R-2, level 2 updates lanes 2..61 / rows 2..R-3. Per plane step k: level 1 does
// plane k, level 2 does plane k-1 (its upper z face needs H1 at plane k). Level 1 also runs the planes
// k0-1 and k1 around the chunk. Cells on the domain boundary keep H (Dirichlet) at both levels.
// Each level evaluates exactly the expressions of the one-iteration sweep, so two levels here give the
// same bits as two k_fused3d launches.

template <typename T, typename Ta>
struct PlaneOps2 {
    T Hn;      // H(k+1)
    Ta Dn;     // aux(k+1): D (step) or Hc (correction sweep)
    T Ho;      // Hold(k)
    T qx, qy;  // old fluxes on the lower x / y faces at plane k
    T qz;      // old flux on the upper z face (k | k+1)
};

// Occupancy (measured at 511^3): two blocks per CU without spills is what pays. float32: 16 rows,
// operands prefetched one plane ahead, <= 64 VGPRs (8 waves per SIMD); its checking variant 12 rows
// (6 waves per SIMD). float64 lin / nonlin: 12 rows, <= 80 VGPRs (6 waves per SIMD), operands two
// planes ahead for the FAST sweep and one for the checking and EXACT ones. The float64 variants with
// more live values (step D, EXACT checking) would spill under that cap and run 16 rows, one block per CU.
template <typename T, int DC, int ARITH, bool CHECK>
constexpr bool t2_compact() {
    return sizeof(T) == 8 && DC != PTD_STEP && (ARITH == kFast || !CHECK);
}
// float64 linear FAST (non-checking) has few live values: 16 rows at 8 waves per SIMD fit in 64 VGPRs
// without spills (1.27 -> 1.235 ms per iteration at 511^3 against the 12-row tiling)
template <typename T, int DC, int ARITH, bool CHECK>
constexpr bool t2_lin16() {
    return sizeof(T) == 8 && DC == PTD_LIN && ARITH == kFast && !CHECK;
}
template <typename T, int DC, int ARITH, int R, bool CHECK>
constexpr int t2_waves_per_simd() {
    return sizeof(T) == 4 || t2_compact<T, DC, ARITH, CHECK>() || t2_lin16<T, DC, ARITH, CHECK>() ? 2 * (R / 4) : 1;
}
template <typename T, int DC, int ARITH, int R, bool CHECK>
constexpr int t2_prefetch_depth() {
    return sizeof(T) == 4 || CHECK || ARITH == kExact || (DC == PTD_LIN && R == 16) ? 1 : 2;
}
template <typename T, int DC, int ARITH, bool CHECK>
constexpr int t2_rows() {
    if constexpr (sizeof(T) == 4) return CHECK ? 12 : 16;
    else return t2_compact<T, DC, ARITH, CHECK>() && !t2_lin16<T, DC, ARITH, CHECK>() ? 12 : 16;
}
constexpr int kMinRows2 = 12;  // smallest two-level block (sizes the partials buffer)

// One tile (bx, by) x z chunk bz of the two-level sweep; `part` receives the tile's residual partial
// (CHECK). Called by the single-problem kernel below and by the batched kernel.
template <typename T, int DC, int FORM, int ARITH, int R, bool CHECK>
__device__ __forceinline__ void fused3d_t2_tile(const KP<T>& p, const T* __restrict__ Hin, T* __restrict__ Hout,
                                                const T* __restrict__ Hold,
                                                const typename AuxOf<T, DC>::type* __restrict__ D,
                                                const T* __restrict__ qxi, const T* __restrict__ qyi,
                                                const T* __restrict__ qzi, T* __restrict__ qxo,
                                                T* __restrict__ qyo, T* __restrict__ qzo,
                                                double2* __restrict__ part, int bx, int by, int bz, int tz) {
    using Ph = Phys<T, DC, FORM, ARITH>;
    using Ta = typename AuxOf<T, DC>::type;
    constexpr bool STEP = DC == PTD_STEP;
    constexpr bool AUX = has_aux<DC>();
    __shared__ T sH[R][64];                      // H at plane k (level-1 y neighbours)
    __shared__ T sH1[R][64];                     // H1 at plane k-1 (level-2 y neighbours)
    __shared__ T sQ1[R][64];                     // level-1 y flux at plane k
    __shared__ T sQ2[R][64];                     // level-2 y flux at plane k-1
    __shared__ Ta sD[AUX ? R : 1][64];           // aux at plane k
    __shared__ Ta sDp[AUX ? R : 1][64];          // aux at plane k-1
    __shared__ T sR2[CHECK ? R : 1][64];         // level-2 residual y flux

    const int lane = threadIdx.x, row = threadIdx.y;
    const int nx = (int)p.nx, ny = (int)p.ny, nz = (int)p.nz;
    // One row per wavefront (blockDim.x == 64): rows whose results no computed cell needs skip the
    // arithmetic with wave-uniform branches. Level 1: rows 1..R-2 update (faces + cell), row R-1 only
    // evaluates its lower y face (row R-2's upper one). Level 2: rows 2..R-3 update, row R-2 only
    // evaluates its lower y face. Row 0 only loads.
    const int wrow = __builtin_amdgcn_readfirstlane(row);
    const bool l1all = wrow >= 1 && wrow <= R - 2;
    const bool l1y = wrow >= 1;
    const bool l2all = wrow >= 2 && wrow <= R - 3;
    const bool l2y = wrow >= 2 && wrow <= R - 2;
    const int i = 60 * bx - 1 + lane;
    const int j = (R - 4) * by - 1 + row;
    const int k0 = 1 + tz * bz;
    const int k1 = (k0 + tz < nz - 1) ? k0 + tz : nz - 1;  // level 2 updates planes k0..k1-1
    const int ka = k0 - 1;                                  // level 1 runs planes ka..k1

    const bool xin = (i >= 1) && (i <= nx - 2);
    const bool yin = (j >= 1) && (j <= ny - 2);
    const bool u1 = (lane >= 1) && (lane <= 62) && (row >= 1) && (row <= R - 2) && xin && yin;
    const bool r2mid = (row >= 2) && (row <= R - 3);
    const bool l2mid = (lane >= 2) && (lane <= 61);
    const bool u2 = l2mid && r2mid && xin && yin;
    const bool f2x = (lane >= 2) && (lane <= 62) && (i <= nx - 1) && r2mid && yin;
    const bool f2y = (row >= 2) && (row <= R - 2) && (j <= ny - 1) && l2mid && xin;
    const bool w2x = f2x && ((lane <= 61 && i <= nx - 2) || i == nx - 1);
    const bool w2y = f2y && ((row <= R - 3 && j <= ny - 2) || j == ny - 1);

    // Byte offsets, clamped into each array's range (every load is in bounds). Values of lanes outside
    // a level's computed set are never stored and never reach a computed cell.
    const size_t psH = (size_t)nx * ny;
    const size_t psQx = (size_t)(nx - 1) * (ny - 2);
    const size_t psQy = (size_t)(nx - 2) * (ny - 1);
    const size_t psQz = (size_t)(nx - 2) * (ny - 2);
    // Hold, qx and qz are used only on the rows that update a level-1 cell (1..R-2) and qy only on the rows
    // that evaluate their lower y face (1..R-1): the other rows load them from the nearest such row, so the
    // tile's halo rows fetch no bytes beyond H (the same instructions, cache hits instead of new lines).
    const int jr = (R - 4) * by - 1 + max(1, min(row, R - 2));
    const int jq = (R - 4) * by - 1 + max(1, row);
    const int ic = max(0, min(i, nx - 1)), jc = max(0, min(j, ny - 1)), jo = max(0, min(jr, ny - 1));
    const int ixq = max(1, min(i, nx - 1)), jxq = max(1, min(jr, ny - 2));
    const int iyq = max(1, min(i, nx - 2)), jyq = max(1, min(jq, ny - 1));
    const int jzq = max(1, min(jr, ny - 2));
    constexpr uint32_t es = (uint32_t)sizeof(T);
    const uint32_t oH = (uint32_t)(ic + nx * jc) * es;
    const uint32_t oHo = (uint32_t)(ic + nx * jo) * es;
    const uint32_t oA = (uint32_t)(ic + nx * jc) * (uint32_t)sizeof(Ta);
    const uint32_t oQx = (uint32_t)((ixq - 1) + (nx - 1) * (jxq - 1)) * es;
    const uint32_t oQy = (uint32_t)((iyq - 1) + (nx - 2) * (jyq - 1)) * es;
    const uint32_t oQz = (uint32_t)((iyq - 1) + (nx - 2) * (jzq - 1)) * es;

    // Every plane this tile touches lies in kb .. kb + tz + 4 (kb = ka - 1, or 0): the arrays are addressed
    // from their plane kb (block-uniform bases) with 32-bit byte offsets (the host keeps (tz + 5) planes
    // of H under 4 GiB).
    const int kb = ka > 0 ? ka - 1 : 0;
    const char* const bHin = reinterpret_cast<const char*>(Hin + psH * kb);
    const char* const bD = reinterpret_cast<const char*>(D + (AUX ? psH * kb : 0));
    const char* const bHo = reinterpret_cast<const char*>(Hold + psH * kb);
    const char* const bqx = reinterpret_cast<const char*>(qxi + psQx * kb);
    const char* const bqy = reinterpret_cast<const char*>(qyi + psQy * kb);
    const char* const bqz = reinterpret_cast<const char*>(qzi + psQz * kb);
    char* const bHout = reinterpret_cast<char*>(Hout + psH * kb);
    char* const bqxo = reinterpret_cast<char*>(qxo + psQx * kb);
    char* const bqyo = reinterpret_cast<char*>(qyo + psQy * kb);
    char* const bqzo = reinterpret_cast<char*>(qzo + psQz * kb);
    const uint32_t pbH = (uint32_t)psH * es, pbA = (uint32_t)psH * (uint32_t)sizeof(Ta);
    const uint32_t pbQx = (uint32_t)psQx * es, pbQy = (uint32_t)psQy * es, pbQz = (uint32_t)psQz * es;

    auto load_plane = [&](int pl) {  // level-1 operands of plane pl, pl in [0, nz-1]
        PlaneOps2<T, Ta> o;
        const uint32_t ph = (uint32_t)((pl + 1 < nz ? pl + 1 : nz - 1) - kb);
        const uint32_t pq = (uint32_t)((pl - 1 < 0 ? 0 : (pl - 1 > nz - 3 ? nz - 3 : pl - 1)) - kb);
        const uint32_t pz = (uint32_t)((pl > nz - 2 ? nz - 2 : pl) - kb);
        o.Hn = ldc<T>(bHin, oH + ph * pbH);
        if constexpr (AUX) o.Dn = ldc<Ta>(bD, oA + ph * pbA);
        else o.Dn = Ta{};
        // float32: Hold and the old fluxes only where level 1 evaluates them (wave-uniform: rows 0 and
        // R-1 skip these loads). Measured: -1% on the float32 correction sweep, but +4% on the float64
        // nonlinear and +18% on the float64 step sweeps, whose loads then lose their scheduling.
        constexpr bool kSkip = sizeof(T) == 4;
        o.Ho = (!kSkip || l1all) ? ldc<T>(bHo, oHo + (uint32_t)(pl - kb) * pbH) : T{};
        o.qx = (!kSkip || l1all) ? ldc<T>(bqx, oQx + pq * pbQx) : T{};
        o.qy = (!kSkip || l1y) ? ldc<T>(bqy, oQy + pq * pbQy) : T{};
        o.qz = (!kSkip || l1all) ? ldc<T>(bqz, oQz + pz * pbQz) : T{};
        return o;
    };

    // level-1 lower z face of plane ka (face (ka-1 | ka)); needed only when plane ka is inner
    T Hc = ldc<T>(bHin, oH + (uint32_t)(ka - kb) * pbH);
    Ta Dc{}, Dm1{}, Dm2{};  // aux at planes k, k-1, k-2
    if constexpr (AUX) {
        Dc = ldc<Ta>(bD, oA + (uint32_t)(ka - kb) * pbA);
        Dm1 = ldc<Ta>(bD, oA);
    }
    T unused = 0;
    T q1zL = Ph::template face<false>(p, ldc<T>(bqz, oQz), ldc<T>(bHin, oH), Hc, Dm1, Dc, p.dz, p._dz, p._dz8,
                                      unused);

    constexpr bool kAhead2 = t2_prefetch_depth<T, DC, ARITH, R, CHECK>() == 2;
    PlaneOps2<T, Ta> A = load_plane(ka);
    PlaneOps2<T, Ta> B;
    if constexpr (kAhead2) B = load_plane(ka + 1 <= nz - 1 ? ka + 1 : nz - 1);

    // correction sweep: level-1 damping factors of the faces / cell that level 2 evaluates next
    constexpr bool LAG = DC == kDelta;
    T w1x = 0, w1y = 0, w1zU = 0, v1 = 0;          // this iteration's level-1 factors (plane k)
    T w1xp = 0, w1yp = 0, w1zL = 0, v1p = 0;       // carried: plane k-1, face (k-1 | k)
    T H1p = 0;             // H1 at plane k-1
    T Hop = 0;             // Hold at plane k-1
    T q1xp = 0, q1yp = 0;  // level-1 fluxes on the lower x / y faces at plane k-1
    T q2zL = 0, r2zL = 0;  // level-2 flux (and residual flux) on the lower z face of plane k-1
    T acc = 0, amax = 0;   // per-thread sum of squares / max of the residual (in the sweep precision)

    for (int k = ka; k <= k1; ++k) {
        // phase A: planes into LDS
        sH[row][lane] = Hc;
        sH1[row][lane] = H1p;
        if constexpr (AUX) {
            sD[row][lane] = Dc;
            sDp[row][lane] = Dm1;
        }
        __syncthreads();

        // phase B: level-1 faces of plane k, level-2 x / y faces of plane k-1
        const int rs = row > 0 ? row - 1 : 0, rn = row < R - 1 ? row + 1 : R - 1;
        const T Hw = dpp_up1(Hc);
        const T Hs = sH[rs][lane];
        Ta Dw{}, De{}, Ds{}, Dnn{}, Dpw{}, Dpe{}, Dps{}, Dpn{};
        if constexpr (AUX) {
            Dw = dpp_up1(Dc);
            De = dpp_down1(Dc);
            Ds = sD[rs][lane];
            Dnn = sD[rn][lane];
            Dpw = dpp_up1(Dm1);
            Dpe = dpp_down1(Dm1);
            Dps = sDp[rs][lane];
            Dpn = sDp[rn][lane];
        }
        T q1x = 0, q1y = 0, q1zU = 0;
        if constexpr (LAG) {
            if (l1y) q1y = Ph::face_w(p, A.qy, Hs, Hc, Ds, Dc, w1y);
            if (l1all) {
                q1x = Ph::face_w(p, A.qx, Hw, Hc, Dw, Dc, w1x);
                q1zU = Ph::face_w(p, A.qz, Hc, A.Hn, Dc, A.Dn, w1zU);
            }
        } else {
            if (l1y) q1y = Ph::template face<false>(p, A.qy, Hs, Hc, Ds, Dc, p.dy, p._dy, p._dy8, unused);
            if (l1all) {
                q1x = Ph::template face<false>(p, A.qx, Hw, Hc, Dw, Dc, p.dx, p._dx, p._dx8, unused);
                q1zU = Ph::template face<false>(p, A.qz, Hc, A.Hn, Dc, A.Dn, p.dz, p._dz, p._dz8, unused);
            }
        }
        const T H1w = dpp_up1(H1p);
        const T H1s = sH1[rs][lane];
        T r2x = 0, r2y = 0, q2x = 0, q2y = 0;
        if constexpr (LAG) {
            if (l2y) q2y = Ph::template face_lag<CHECK>(q1yp, H1s, H1p, Dps, Dm1, w1yp, r2y);
            if (l2all) q2x = Ph::template face_lag<CHECK>(q1xp, H1w, H1p, Dpw, Dm1, w1xp, r2x);
        } else {
            if (l2y) q2y = Ph::template face<CHECK>(p, q1yp, H1s, H1p, Dps, Dm1, p.dy, p._dy, p._dy8, r2y);
            if (l2all) q2x = Ph::template face<CHECK>(p, q1xp, H1w, H1p, Dpw, Dm1, p.dx, p._dx, p._dx8, r2x);
        }
        sQ1[row][lane] = q1y;
        sQ2[row][lane] = q2y;
        if constexpr (CHECK) sR2[row][lane] = r2y;
        __syncthreads();

        // phase C: level-1 update of plane k; level-2 z face (k-1 | k) and update of plane k-1
        const T q1xE = dpp_down1(q1x);
        const T q1yN = sQ1[rn][lane];
        T dmax1 = 0, dmax2 = 0;
        if constexpr (STEP) {
            dmax1 = fmax(fmax(Dw, De), Dc);
            dmax1 = fmax(dmax1, fmax(Ds, Dnn));
            dmax1 = fmax(dmax1, fmax(Dm1, A.Dn));
            dmax2 = fmax(fmax(Dpw, Dpe), Dm1);
            dmax2 = fmax(dmax2, fmax(Dps, Dpn));
            dmax2 = fmax(dmax2, fmax(Dm2, Dc));
        }
        T H1 = Hc;  // boundary cells and rows without a level-1 update keep H
        if (l1all) {
            const T dv1 = Ph::template div<3>(p, q1xE - q1x, q1yN - q1y, q1zU - q1zL);
            T H1c;
            if constexpr (LAG) H1c = Ph::cell_w(p, Hc, A.Ho, dv1, Dc, v1);
            else H1c = Ph::cell(p, Hc, A.Ho, dv1, dmax1, Dc);
            if (u1 && k >= 1 && k <= nz - 2) H1 = H1c;
        }

        T r2z = 0, q2zU = 0;
        if constexpr (LAG) {
            if (l2all) q2zU = Ph::template face_lag<CHECK>(q1zL, H1p, H1, Dm1, Dc, w1zL, r2z);
        } else {
            if (l2all) q2zU = Ph::template face<CHECK>(p, q1zL, H1p, H1, Dm1, Dc, p.dz, p._dz, p._dz8, r2z);
        }
        if (k > k0) {
            const int m = k - 1;
            if (l2all) {
                const T q2xE = dpp_down1(q2x);
                const T q2yN = sQ2[rn][lane];
                const T dv2 = Ph::template div<3>(p, q2xE - q2x, q2yN - q2y, q2zU - q2zL);
                T H2;
                if constexpr (LAG) H2 = Ph::cell_lag(p, H1p, Hop, dv2, v1p);
                else H2 = Ph::cell(p, H1p, Hop, dv2, dmax2, Dm1);
                if constexpr (CHECK) {
                    const T r2xE = dpp_down1(r2x);
                    const T r2yN = sR2[rn][lane];
                    const T dvr = Ph::template div<3>(p, r2xE - r2x, r2yN - r2y, r2z - r2zL);
                    const T rr = u2 ? Ph::resid(p, H2, Hop, dvr) : (T)0;
                    acc += rr * rr;
                    amax = nanmax(fabs(rr), amax);
                }
                const uint32_t rm = (uint32_t)(m - kb);
                if (u2) {
                    stc<T>(bHout, oH + rm * pbH, H2);
                    stc<T>(bqzo, oQz + rm * pbQz, q2zU);
                }
                if (w2x) stc<T>(bqxo, oQx + (rm - 1) * pbQx, q2x);
            }
            if (w2y) stc<T>(bqyo, oQy + (uint32_t)(k - 2 - kb) * pbQy, q2y);
        } else if (k == k0 && k0 == 1) {
            if (u2) stc<T>(bqzo, oQz, q2zU);  // face (0 | 1): lower face of the first inner plane (kb = 0)
        }
        // carry to plane k+1
        if constexpr (LAG) {
            w1xp = w1x;
            w1yp = w1y;
            w1zL = w1zU;
            v1p = v1;
        }
        q2zL = q2zU;
        r2zL = r2z;
        q1zL = q1zU;
        H1p = H1;
        Hop = A.Ho;
        q1xp = q1x;
        q1yp = q1y;
        if constexpr (AUX) {
            Dm2 = Dm1;
            Dm1 = Dc;
            Dc = A.Dn;
        }
        Hc = A.Hn;
        if constexpr (kAhead2) {
            A = B;
            B = load_plane(k + 2 <= nz - 1 ? k + 2 : nz - 1);
        } else {
            A = load_plane(k + 1 <= nz - 1 ? k + 1 : nz - 1);
        }
    }
    if constexpr (CHECK) block_sum_max((double)acc, (double)amax, row * 64 + lane, 64 * R, part);
}

template <typename T, int DC, int FORM, int ARITH, int R, bool CHECK>
__global__ void __launch_bounds__(64 * R, (t2_waves_per_simd<T, DC, ARITH, R, CHECK>()))
k_fused3d_t2(const KP<T> p, const T* __restrict__ Hin, T* __restrict__ Hout, const T* __restrict__ Hold,
             const typename AuxOf<T, DC>::type* __restrict__ D, const T* __restrict__ qxi, const T* __restrict__ qyi,
             const T* __restrict__ qzi, T* __restrict__ qxo, T* __restrict__ qyo, T* __restrict__ qzo,
             double2* __restrict__ partial, int tz, int ntx, int nty, int ntiles, int per_xcd) {
    const int b = blockIdx.x;
    const int tile = (b & 7) * per_xcd + (b >> 3);  // XCD-aware: each XCD walks a contiguous tile range
    if (tile >= ntiles) return;
    const int bx = tile % ntx;
    const int rest = tile / ntx;
    fused3d_t2_tile<T, DC, FORM, ARITH, R, CHECK>(p, Hin, Hout, Hold, D, qxi, qyi, qzi, qxo, qyo, qzo,
                                                  &partial[tile], bx, rest % nty, rest / nty, tz);
}

// ---------------------------------------------------------------------------------------------
// fused 3D sweep over three PT iterations (temporal blocking, three levels)
// ---------------------------------------------------------------------------------------------
//
// The two-level z-marching block extended by a third level: per plane step k, level 1 (iteration n+1)
// does plane k, level 2 (n+2) plane k-1 and level 3 (n+3) plane k-2, so HBM sees one read of the
// iterate-n fields and one write of the iterate-(n+3) fields for three iterations. The halo shrinks by
// one cell per level: lane L holds i = 58*bx - 2 + L and row r holds j = (R-6)*by - 2 + r; level 1 updates
// lanes 1..62 / rows 1..R-2, level 2 lanes 2..61 / rows 2..R-3, level 3 lanes 3..60 / rows 3..R-4. Level 1
// runs the planes k0-2 .. k1+1 around the chunk, level 2 k0-1 .. k1, level 3 updates k0 .. k1-1. Cells on the
// domain boundary keep H at every level. Each level evaluates exactly the expressions of the one-iteration
// sweep, so a pass gives the same bits as three k_fused3d launches. (Plain / perf forms of the lin, step
// and nonlinear D cases; no residual: the checking iteration of a solve stays a two-level or single sweep.)
template <typename T, int DC, int FORM, int ARITH, int R>
__device__ __forceinline__ void fused3d_t3_tile(const KP<T>& p, const T* __restrict__ Hin, T* __restrict__ Hout,
                                                const T* __restrict__ Hold, const T* __restrict__ D,
                                                const T* __restrict__ qxi, const T* __restrict__ qyi,
                                                const T* __restrict__ qzi, T* __restrict__ qxo,
                                                T* __restrict__ qyo, T* __restrict__ qzo, int bx, int by, int bz,
                                                int tz) {
    static_assert(DC == PTD_LIN || DC == PTD_STEP || DC == PTD_NONLIN, "three-level sweep: lin / step / nonlin");
    using Ph = Phys<T, DC, FORM, ARITH>;
    constexpr bool STEP = DC == PTD_STEP;
    __shared__ T sH[R][64];    // H at plane k (level-1 y neighbours)
    __shared__ T sH1[R][64];   // H1 at plane k-1 (level-2 y neighbours)
    __shared__ T sH2[R][64];   // H2 at plane k-2 (level-3 y neighbours)
    __shared__ T sQ1[R][64];   // level-1 y flux at plane k
    __shared__ T sQ2[R][64];   // level-2 y flux at plane k-1
    __shared__ T sQ3[R][64];   // level-3 y flux at plane k-2
    __shared__ T sD0[STEP ? R : 1][64];  // D at planes k, k-1, k-2
    __shared__ T sD1[STEP ? R : 1][64];
    __shared__ T sD2[STEP ? R : 1][64];

    const int lane = threadIdx.x, row = threadIdx.y;
    const int nx = (int)p.nx, ny = (int)p.ny, nz = (int)p.nz;
    const int wrow = __builtin_amdgcn_readfirstlane(row);
    // wave-uniform row roles: level L updates rows L..R-1-L and evaluates lower y faces on rows L..R-L
    const bool l1all = wrow >= 1 && wrow <= R - 2, l1y = wrow >= 1;
    const bool l2all = wrow >= 2 && wrow <= R - 3, l2y = wrow >= 2 && wrow <= R - 2;
    const bool l3all = wrow >= 3 && wrow <= R - 4, l3y = wrow >= 3 && wrow <= R - 3;
    const int i = 58 * bx - 2 + lane;
    const int j = (R - 6) * by - 2 + row;
    const int k0 = 1 + tz * bz;
    const int k1 = (k0 + tz < nz - 1) ? k0 + tz : nz - 1;  // level 3 updates planes k0..k1-1
    const int ka = k0 - 2 > 0 ? k0 - 2 : 0;                 // level 1 runs planes ka..k1+1
    const int kb = ka > 0 ? ka - 1 : 0;

    const bool xin = (i >= 1) && (i <= nx - 2);
    const bool yin = (j >= 1) && (j <= ny - 2);
    const bool u1 = (lane >= 1) && (lane <= 62) && (row >= 1) && (row <= R - 2) && xin && yin;
    const bool u2 = (lane >= 2) && (lane <= 61) && (row >= 2) && (row <= R - 3) && xin && yin;
    const bool r3mid = (row >= 3) && (row <= R - 4);
    const bool l3mid = (lane >= 3) && (lane <= 60);
    const bool u3 = l3mid && r3mid && xin && yin;
    const bool f3x = (lane >= 3) && (lane <= 61) && (i <= nx - 1) && r3mid && yin;
    const bool f3y = (row >= 3) && (row <= R - 3) && (j <= ny - 1) && l3mid && xin;
    const bool w3x = f3x && ((lane <= 60 && i <= nx - 2) || i == nx - 1);
    const bool w3y = f3y && ((row <= R - 4 && j <= ny - 2) || j == ny - 1);

    // clamped byte offsets (every load in bounds); Hold, qx, qz load on rows 1..R-2 and qy on rows 1..R-1
    // (the rows level 1 uses them on), other rows repeat the nearest such row's address
    const size_t psH = (size_t)nx * ny;
    const size_t psQx = (size_t)(nx - 1) * (ny - 2);
    const size_t psQy = (size_t)(nx - 2) * (ny - 1);
    const size_t psQz = (size_t)(nx - 2) * (ny - 2);
    const int jr = (R - 6) * by - 2 + max(1, min(row, R - 2));
    const int jq = (R - 6) * by - 2 + max(1, row);
    const int ic = max(0, min(i, nx - 1)), jc = max(0, min(j, ny - 1)), jo = max(0, min(jr, ny - 1));
    const int ixq = max(1, min(i, nx - 1)), jxq = max(1, min(jr, ny - 2));
    const int iyq = max(1, min(i, nx - 2)), jyq = max(1, min(jq, ny - 1));
    const int jzq = max(1, min(jr, ny - 2));
    constexpr uint32_t es = (uint32_t)sizeof(T);
    const uint32_t oH = (uint32_t)(ic + nx * jc) * es;
    const uint32_t oHo = (uint32_t)(ic + nx * jo) * es;
    const uint32_t oQx = (uint32_t)((ixq - 1) + (nx - 1) * (jxq - 1)) * es;
    const uint32_t oQy = (uint32_t)((iyq - 1) + (nx - 2) * (jyq - 1)) * es;
    const uint32_t oQz = (uint32_t)((iyq - 1) + (nx - 2) * (jzq - 1)) * es;

    // every plane the tile touches lies in kb .. kb + tz + 5: block-uniform plane-kb bases, 32-bit offsets
    const char* const bHin = reinterpret_cast<const char*>(Hin + psH * kb);
    const char* const bD = reinterpret_cast<const char*>(D + (STEP ? psH * kb : 0));
    const char* const bHo = reinterpret_cast<const char*>(Hold + psH * kb);
    const char* const bqx = reinterpret_cast<const char*>(qxi + psQx * kb);
    const char* const bqy = reinterpret_cast<const char*>(qyi + psQy * kb);
    const char* const bqz = reinterpret_cast<const char*>(qzi + psQz * kb);
    char* const bHout = reinterpret_cast<char*>(Hout + psH * kb);
    char* const bqxo = reinterpret_cast<char*>(qxo + psQx * kb);
    char* const bqyo = reinterpret_cast<char*>(qyo + psQy * kb);
    char* const bqzo = reinterpret_cast<char*>(qzo + psQz * kb);
    const uint32_t pbH = (uint32_t)psH * es;
    const uint32_t pbQx = (uint32_t)psQx * es, pbQy = (uint32_t)psQy * es, pbQz = (uint32_t)psQz * es;

    auto load_plane = [&](int pl) {  // level-1 operands of plane pl, pl in [0, nz-1]
        PlaneOps2<T, T> o;
        const uint32_t ph = (uint32_t)((pl + 1 < nz ? pl + 1 : nz - 1) - kb);
        const uint32_t pq = (uint32_t)((pl - 1 < 0 ? 0 : (pl - 1 > nz - 3 ? nz - 3 : pl - 1)) - kb);
        const uint32_t pz = (uint32_t)((pl > nz - 2 ? nz - 2 : pl) - kb);
        o.Hn = ldc<T>(bHin, oH + ph * pbH);
        if constexpr (STEP) o.Dn = ldc<T>(bD, oH + ph * pbH);
        else o.Dn = T{};
        o.Ho = ldc<T>(bHo, oHo + (uint32_t)(pl - kb) * pbH);
        o.qx = ldc<T>(bqx, oQx + pq * pbQx);
        o.qy = ldc<T>(bqy, oQy + pq * pbQy);
        o.qz = ldc<T>(bqz, oQz + pz * pbQz);
        return o;
    };

    // level-1 lower z face of plane ka (face (ka-1 | ka)); needed only when plane ka is inner
    T Hc = ldc<T>(bHin, oH + (uint32_t)(ka - kb) * pbH);
    T Dc{}, Dm1{}, Dm2{}, Dm3{};  // D at planes k, k-1, k-2, k-3
    if constexpr (STEP) {
        Dc = ldc<T>(bD, oH + (uint32_t)(ka - kb) * pbH);
        Dm1 = ldc<T>(bD, oH);
    }
    T unused = 0;
    T q1zL = Ph::template face<false>(p, ldc<T>(bqz, oQz), ldc<T>(bHin, oH), Hc, Dm1, Dc, p.dz, p._dz, p._dz8, unused);
    // operands two planes ahead: a three-level plane step does half as much again of arithmetic as a
    // two-level one between its loads, and one plane of lookahead left them exposed (float64 FAST at 511^3,
    // ms per iteration: linear 0.998 -> 0.836, step 1.236 -> 1.200, nonlinear 1.370 -> 1.284; three planes
    // ahead were slower again). The float64 EXACT step pass keeps one plane: it would spill.
    constexpr bool kAhead2 = !(sizeof(T) == 8 && DC == PTD_STEP && ARITH == kExact);
    // The float64 FAST linear / nonlinear passes and the float32 FAST step pass rotate three operand sets
    // (steps ka, ka+3, ... on A, the next ones on B, then C; each set is reloaded three planes ahead right
    // after its step): a copy B -> A at the end of every step would wait for the loads just issued into B,
    // leaving them one plane step to return. They have the register room (one 16-wave block per CU either
    // way: float64 nonlinear 90 -> 109 VGPRs, float32 step 84 -> 98). Float64 nonlinear at 511^3, ms per
    // iteration: copy 1.171-1.178, two alternating sets 1.124-1.137; 1023^3 9.42-9.54 -> 9.07-9.08; three
    // sets a further 0.8% (511^3) and 1.3% (1023^3). Float32 step: 511^3 0.597 -> 0.527, 767^3 2.082 ->
    // 1.840. The float32 linear pass would drop to one block per CU (56 -> 72 VGPRs; 511^3 2-3% slower,
    // 767^3 10% faster), and the float64 step pass would spill 64 VGPRs.
    constexpr bool kAlt = ARITH == kFast && (sizeof(T) == 8 ? DC != PTD_STEP : DC == PTD_STEP);
    PlaneOps2<T, T> A = load_plane(ka);
    PlaneOps2<T, T> B;
    if constexpr (kAhead2) B = load_plane(ka + 1 <= nz - 1 ? ka + 1 : nz - 1);
    PlaneOps2<T, T> C;
    if constexpr (kAlt) C = load_plane(ka + 2 <= nz - 1 ? ka + 2 : nz - 1);

    T H1p = 0, H2p = 0;          // H1 at plane k-1, H2 at plane k-2
    T Hop = 0, Hopp = 0;         // Hold at planes k-1, k-2
    T q1xp = 0, q1yp = 0;        // level-1 fluxes on the lower x / y faces at plane k-1
    T q2xp = 0, q2yp = 0;        // level-2 fluxes on the lower x / y faces at plane k-2
    T q2zL = 0, q3zL = 0;        // level-2 flux on face (k-2 | k-1), level-3 flux on face (k-3 | k-2)

    auto step = [&](const int k, PlaneOps2<T, T>& A) {
        // phase A: planes into LDS
        sH[row][lane] = Hc;
        sH1[row][lane] = H1p;
        sH2[row][lane] = H2p;
        if constexpr (STEP) {
            sD0[row][lane] = Dc;
            sD1[row][lane] = Dm1;
            sD2[row][lane] = Dm2;
        }
        __syncthreads();

        // phase B: x / y faces of level 1 (plane k), level 2 (plane k-1), level 3 (plane k-2)
        const int rs = row > 0 ? row - 1 : 0, rn = row < R - 1 ? row + 1 : R - 1;
        T Dw0{}, De0{}, Ds0{}, Dn0{}, Dw1{}, De1{}, Ds1{}, Dn1{}, Dw2{}, De2{}, Ds2{}, Dn2{};
        if constexpr (STEP) {
            Dw0 = dpp_up1(Dc);
            De0 = dpp_down1(Dc);
            Ds0 = sD0[rs][lane];
            Dn0 = sD0[rn][lane];
            Dw1 = dpp_up1(Dm1);
            De1 = dpp_down1(Dm1);
            Ds1 = sD1[rs][lane];
            Dn1 = sD1[rn][lane];
            Dw2 = dpp_up1(Dm2);
            De2 = dpp_down1(Dm2);
            Ds2 = sD2[rs][lane];
            Dn2 = sD2[rn][lane];
        }
        const T Hw = dpp_up1(Hc);
        const T Hs = sH[rs][lane];
        T q1x = 0, q1y = 0, q1zU = 0;
        if (l1y) q1y = Ph::template face<false>(p, A.qy, Hs, Hc, Ds0, Dc, p.dy, p._dy, p._dy8, unused);
        if (l1all) {
            q1x = Ph::template face<false>(p, A.qx, Hw, Hc, Dw0, Dc, p.dx, p._dx, p._dx8, unused);
            q1zU = Ph::template face<false>(p, A.qz, Hc, A.Hn, Dc, A.Dn, p.dz, p._dz, p._dz8, unused);
        }
        const T H1w = dpp_up1(H1p);
        const T H1s = sH1[rs][lane];
        T q2x = 0, q2y = 0;
        if (l2y) q2y = Ph::template face<false>(p, q1yp, H1s, H1p, Ds1, Dm1, p.dy, p._dy, p._dy8, unused);
        if (l2all) q2x = Ph::template face<false>(p, q1xp, H1w, H1p, Dw1, Dm1, p.dx, p._dx, p._dx8, unused);
        const T H2w = dpp_up1(H2p);
        const T H2s = sH2[rs][lane];
        T q3x = 0, q3y = 0;
        if (l3y) q3y = Ph::template face<false>(p, q2yp, H2s, H2p, Ds2, Dm2, p.dy, p._dy, p._dy8, unused);
        if (l3all) q3x = Ph::template face<false>(p, q2xp, H2w, H2p, Dw2, Dm2, p.dx, p._dx, p._dx8, unused);
        sQ1[row][lane] = q1y;
        sQ2[row][lane] = q2y;
        sQ3[row][lane] = q3y;
        __syncthreads();

        // phase C: level-1 update of plane k; level-2 z face (k-1 | k) and update of plane k-1; level-3 z
        // face (k-2 | k-1) and update of plane k-2
        T dmax1 = 0, dmax2 = 0, dmax3 = 0;
        if constexpr (STEP) {
            dmax1 = fmax(fmax(Dw0, De0), Dc);
            dmax1 = fmax(dmax1, fmax(Ds0, Dn0));
            dmax1 = fmax(dmax1, fmax(Dm1, A.Dn));
            dmax2 = fmax(fmax(Dw1, De1), Dm1);
            dmax2 = fmax(dmax2, fmax(Ds1, Dn1));
            dmax2 = fmax(dmax2, fmax(Dm2, Dc));
            dmax3 = fmax(fmax(Dw2, De2), Dm2);
            dmax3 = fmax(dmax3, fmax(Ds2, Dn2));
            dmax3 = fmax(dmax3, fmax(Dm3, Dm1));
        }
        T H1 = Hc;  // boundary cells, boundary planes and rows without a level-1 update keep H
        if (l1all) {
            const T q1xE = dpp_down1(q1x);
            const T q1yN = sQ1[rn][lane];
            const T dv1 = Ph::template div<3>(p, q1xE - q1x, q1yN - q1y, q1zU - q1zL);
            const T H1c = Ph::cell(p, Hc, A.Ho, dv1, dmax1, Dc);
            if (u1 && k >= 1 && k <= nz - 2) H1 = H1c;
        }
        T q2zU = 0, H2 = H1p;
        if (l2all) {
            q2zU = Ph::template face<false>(p, q1zL, H1p, H1, Dm1, Dc, p.dz, p._dz, p._dz8, unused);
            const T q2xE = dpp_down1(q2x);
            const T q2yN = sQ2[rn][lane];
            const T dv2 = Ph::template div<3>(p, q2xE - q2x, q2yN - q2y, q2zU - q2zL);
            const T H2c = Ph::cell(p, H1p, Hop, dv2, dmax2, Dm1);
            if (u2 && k - 1 >= 1 && k - 1 <= nz - 2) H2 = H2c;
        }
        if (l3all) {
            const T q3zU = Ph::template face<false>(p, q2zL, H2p, H2, Dm2, Dm1, p.dz, p._dz, p._dz8, unused);
            const int m = k - 2;
            if (m >= k0 && m <= k1 - 1) {
                const T q3xE = dpp_down1(q3x);
                const T q3yN = sQ3[rn][lane];
                const T dv3 = Ph::template div<3>(p, q3xE - q3x, q3yN - q3y, q3zU - q3zL);
                const T H3 = Ph::cell(p, H2p, Hopp, dv3, dmax3, Dm2);
                const uint32_t rm = (uint32_t)(m - kb);
                if (u3) {
                    stc<T>(bHout, oH + rm * pbH, H3);
                    stc<T>(bqzo, oQz + rm * pbQz, q3zU);
                }
                if (w3x) stc<T>(bqxo, oQx + (rm - 1) * pbQx, q3x);
            } else if (m == 0 && k0 == 1) {
                if (u3) stc<T>(bqzo, oQz, q3zU);  // face (0 | 1): lower face of the first inner plane (kb = 0)
            }
            q3zL = q3zU;
        }
        if (w3y && k - 2 >= k0 && k - 2 <= k1 - 1) stc<T>(bqyo, oQy + (uint32_t)(k - 3 - kb) * pbQy, q3y);
        // carry to plane k+1
        q2zL = q2zU;
        q1zL = q1zU;
        H2p = H2;
        H1p = H1;
        Hopp = Hop;
        Hop = A.Ho;
        q2xp = q2x;
        q2yp = q2y;
        q1xp = q1x;
        q1yp = q1y;
        if constexpr (STEP) {
            Dm3 = Dm2;
            Dm2 = Dm1;
            Dm1 = Dc;
            Dc = A.Dn;
        }
        Hc = A.Hn;
        if constexpr (kAlt) A = load_plane(k + 3 <= nz - 1 ? k + 3 : nz - 1);
    };
    if constexpr (kAlt) {
        for (int k = ka; k <= k1 + 1; k += 3) {
            step(k, A);
            if (k + 1 > k1 + 1) break;
            step(k + 1, B);
            if (k + 2 > k1 + 1) break;
            step(k + 2, C);
        }
    } else {
        for (int k = ka; k <= k1 + 1; ++k) {
            step(k, A);
            if constexpr (kAhead2) {
                A = B;
                B = load_plane(k + 2 <= nz - 1 ? k + 2 : nz - 1);
            } else {
                A = load_plane(k + 1 <= nz - 1 ? k + 1 : nz - 1);
            }
        }
    }
}

static constexpr int kRows3 = 16;  // rows (waves) per block of the three-level sweep

template <typename T, int DC, int FORM, int ARITH, int R>
__global__ void __launch_bounds__(64 * R)
k_fused3d_t3(const KP<T> p, const T* __restrict__ Hin, T* __restrict__ Hout, const T* __restrict__ Hold,
             const T* __restrict__ D, const T* __restrict__ qxi, const T* __restrict__ qyi,
             const T* __restrict__ qzi, T* __restrict__ qxo, T* __restrict__ qyo, T* __restrict__ qzo, int tz,
             int ntx, int nty, int ntiles, int per_xcd) {
    const int b = blockIdx.x;
    const int tile = (b & 7) * per_xcd + (b >> 3);  // XCD-aware: each XCD walks a contiguous tile range
    if (tile >= ntiles) return;
    const int bx = tile % ntx;
    const int rest = tile / ntx;
    fused3d_t3_tile<T, DC, FORM, ARITH, R>(p, Hin, Hout, Hold, D, qxi, qyi, qzi, qxo, qyo, qzo, bx, rest % nty,
                                           rest / nty, tz);
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------

static inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <typename T, int DC, int FORM>
static hipError_t flux_t(ptd_ctx* c, int q2mode) {
    const KP<T> kp = make_kp<T>(c->p);
    const T* H = (const T*)c->H[c->hcur];
    const T* D = (const T*)c->D;
    for (int a = 0; a < 3; ++a) {
        const int64_t n = c->nQ[a];
        if (n == 0) continue;
        T* q = (T*)c->q[a][c->qcur];
        T* q2 = (T*)c->q2[a];
        const dim3 g(grid_for(n)), bl(kBlock);
#define PTD_FLUX_AXIS(A)                                                                                      \
    if (a == A) {                                                                                             \
        if (q2mode == 0) hipLaunchKernelGGL((k_flux<T, DC, FORM, A, 0>), g, bl, 0, c->stream, kp, H, D, q, q2, n); \
        else if (q2mode == 1) hipLaunchKernelGGL((k_flux<T, DC, FORM, A, 1>), g, bl, 0, c->stream, kp, H, D, q, q2, n); \
        else hipLaunchKernelGGL((k_flux<T, DC, FORM, A, 2>), g, bl, 0, c->stream, kp, H, D, q, q2, n);         \
    }
        PTD_FLUX_AXIS(0)
        PTD_FLUX_AXIS(1)
        PTD_FLUX_AXIS(2)
#undef PTD_FLUX_AXIS
    }
    return hipGetLastError();
}

template <typename T, int DC, int FORM>
static hipError_t update_t(ptd_ctx* c) {
    const KP<T> kp = make_kp<T>(c->p);
    const dim3 g(grid_for(c->nR)), bl(kBlock);
    T* H = (T*)c->H[c->hcur];
    const T* Ho = (const T*)c->Hold;
    const T* dt = (const T*)c->dtau;
    const T* qx = (const T*)c->q[0][c->qcur];
    const T* qy = (const T*)c->q[1][c->qcur];
    const T* qz = (const T*)c->q[2][c->qcur];
    switch (c->p.ndim) {
        case 1: hipLaunchKernelGGL((k_update<T, DC, FORM, 1>), g, bl, 0, c->stream, kp, H, Ho, dt, qx, qy, qz, c->nR); break;
        case 2: hipLaunchKernelGGL((k_update<T, DC, FORM, 2>), g, bl, 0, c->stream, kp, H, Ho, dt, qx, qy, qz, c->nR); break;
        default: hipLaunchKernelGGL((k_update<T, DC, FORM, 3>), g, bl, 0, c->stream, kp, H, Ho, dt, qx, qy, qz, c->nR); break;
    }
    return hipGetLastError();
}

template <typename T, int FORM>
static hipError_t residual_t(ptd_ctx* c, bool write_resh) {
    const KP<T> kp = make_kp<T>(c->p);
    const unsigned nb = grid_for(c->nR);
    const dim3 g(nb), bl(kBlock);
    const T* H = (const T*)c->H[c->hcur];
    const T* Ho = (const T*)c->Hold;
    const T* q0 = (const T*)c->q2[0];
    const T* q1 = (const T*)c->q2[1];
    const T* q2 = (const T*)c->q2[2];
    T* r = write_resh ? (T*)c->resh : nullptr;
    switch (c->p.ndim) {
        case 1: hipLaunchKernelGGL((k_residual<T, FORM, 1>), g, bl, 0, c->stream, kp, H, Ho, q0, q1, q2, r, (double2*)c->partial, c->nR); break;
        case 2: hipLaunchKernelGGL((k_residual<T, FORM, 2>), g, bl, 0, c->stream, kp, H, Ho, q0, q1, q2, r, (double2*)c->partial, c->nR); break;
        default: hipLaunchKernelGGL((k_residual<T, FORM, 3>), g, bl, 0, c->stream, kp, H, Ho, q0, q1, q2, r, (double2*)c->partial, c->nR); break;
    }
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, (const double2*)c->partial,
                       (int64_t)nb, c->dred);
    return hipGetLastError();
}

static constexpr int kRows = 16;  // rows (waves) per block of the one-iteration 3D sweep

// Planes per z chunk. Measured at 511^3 FP64 (FAST): 8 -> 2.20 ms, 16 -> 2.17, 32 -> 2.21, 64 -> 2.30,
// 128 -> 2.45, 509 -> 3.01 ms per sweep: short chunks give many more blocks than CUs (small tail) and
// the per-chunk prologue (one extra z face) stays cheap.
// Deepest z chunk whose planes the fused 3D sweeps can address with 32-bit byte offsets from the
// chunk's lowest plane: (tz + 5) planes of H under 4 GiB (0: none).
int64_t max_chunk_tz(const ptd_params& p, int64_t esz) {
    const int64_t plane = p.nx * p.ny * esz;
    return std::max<int64_t>(0, ((int64_t)1 << 32) / std::max<int64_t>(plane, 1) - 5);
}

// z chunk depth of the fused 3D sweeps: a chunk re-reads and re-computes 2 halo planes, so deeper chunks
// cost less per plane, while fewer chunks leave fewer blocks to fill 256 CUs. Measured at 511^3 (float64
// FAST sweep, T_eff GB/s): 8 7296, 16 8054, 24 8159, 28 8259, 32 8450, 36 8223, 40 8242, 48 8207, 64 8317;
// 32 also gives the fastest float32 correction and EXACT sweeps. The time of a sweep is modelled as the
// rounds of blocks the GPU runs times the planes a block marches (its chunk plus 4): 256 blocks per round
// for the float64 step-D sweeps (one 16-row block per CU), 512 for the others (two per CU).
//  - Grids whose 32-plane tiling fills fewer than two rounds take the depth of least modelled time among
//    32, 24, 16, 12, 8, 4 for their non-checking launches (measured 5-step solves, ms: 64^3 nonlinear
//    7.0 -> 3.0 at 4 planes; 128^3 nonlinear 12.7 -> 11.8 at 16, linear 25.6 -> 19.3 at 12, step
//    35.6 -> 28.1 at 24; 192^3 nonlinear 46.3 -> 41.1 at 12).
//  - The step-D sweeps also choose between 24 and 32 on larger grids, where their last round of blocks
//    can leave most CUs idle (t_it: 255^3 0.1965 vs 0.2029 ms, 383^3 0.5832 vs 0.5913 ms; 511^3 keeps 32,
//    1.346 vs 1.386 ms).
// The checking launches keep 32 planes (24 for step D where the model prefers it) on small grids too:
// their tiling fixes the order in which the residual partials are summed, which batched solves, where
// the many members fill the GPU anyway, reproduce bit for bit (check_chunk_tz).
namespace {
struct ChunkModel {
    int64_t mz, slots, tiles;
    bool step64;
    explicit ChunkModel(const ptd_params& p)
        : mz(p.nz - 2), slots(p.dcase == PTD_STEP && p.precision == PTD_FP64 ? 256 : 512), tiles(0),
          step64(p.dcase == PTD_STEP && p.precision == PTD_FP64) {
        const int64_t rows = step64 || p.dcase == PTD_LIN || p.precision != PTD_FP64 ? 12 : 8;  // new rows per tile
        tiles = (p.nx - 2 + 59) / 60 * ((p.ny - 2 + rows - 1) / rows);
    }
    int64_t blocks(int64_t d) const { return tiles * ((mz + d - 1) / d); }
    int64_t cost(int64_t d) const { return (blocks(d) + slots - 1) / slots * (d + 4); }
};
}  // namespace

int check_chunk_tz(const ptd_params& p) {
    const int64_t mz = p.nz - 2;
    const int tz = (int)std::min<int64_t>(32, mz);
    if (p.ndim != 3 || mz <= 32) return tz;
    const ChunkModel m(p);
    return m.step64 && m.cost(24) < m.cost(32) ? 24 : tz;
}

int default_chunk_tz(const ptd_params& p) {
    const int tz = check_chunk_tz(p);
    if (p.ndim != 3) return tz;
    const ChunkModel m(p);
    const int64_t t32 = std::min<int64_t>(32, m.mz);
    if (m.blocks(t32) >= 2 * m.slots) return tz;
    int64_t best = t32;
    for (int64_t d : {24, 16, 12, 8, 4})
        if (d < t32 && m.cost(d) < m.cost(best)) best = d;
    return (int)best;
}

static int fused_tz(const ptd_ctx* c, bool check) {
    const char* env = getenv("PTD_TZ");
    const int64_t want = env && atoi(env) > 0 ? atoi(env) : (check ? check_chunk_tz(c->p) : default_chunk_tz(c->p));
    return (int)std::max<int64_t>(1, std::min(want, max_chunk_tz(c->p, (int64_t)c->esz)));
}

// Buffers of one sweep: iterate n in, iterate n+1 (or n+2) out.
struct SweepBufs {
    const void* Hin;
    void* Hout;
    const void* Hold;
    const void* aux;       // D (step), Hc (correction sweep of D = H^3), unused otherwise
    const void* qin[3];
    void* qout[3];
};

// One fused sweep (two = false: one iteration, k_fused3d) or a two-level pass (two iterations,
// k_fused3d_t2); with `check` the residual of the last iteration is reduced into c->dred.
template <typename T, int DC, int FORM, int ARITH>
static hipError_t sweep_run(ptd_ctx* c, const SweepBufs& b, bool two, bool check) {
    using Ta = typename AuxOf<T, DC>::type;
    const KP<T> kp = make_kp<T>(c->p);
    const int tz = fused_tz(c, check);
    constexpr int R2 = t2_rows<T, DC, ARITH, false>(), R2C = t2_rows<T, DC, ARITH, true>();
    const int brows = two ? (check ? R2C : R2) : kRows;
    const int w = two ? 60 : 62, rows = two ? brows - 4 : brows - 2;
    const int ntx = (int)((c->p.nx - 2 + w - 1) / w);
    const int nty = (int)((c->p.ny - 2 + rows - 1) / rows);
    const int nbz = (int)((c->p.nz - 2 + tz - 1) / tz);
    const int ntiles = ntx * nty * nbz;
    const int per = (ntiles + 7) / 8;
    const dim3 g(8 * per), bl(64, brows);
    const T* Hin = (const T*)b.Hin;
    T* Hout = (T*)b.Hout;
    const T* Ho = (const T*)b.Hold;
    const Ta* A = (const Ta*)b.aux;
    const T* q0 = (const T*)b.qin[0];
    const T* q1 = (const T*)b.qin[1];
    const T* q2 = (const T*)b.qin[2];
    T* o0 = (T*)b.qout[0];
    T* o1 = (T*)b.qout[1];
    T* o2 = (T*)b.qout[2];
    double2* part = (double2*)c->partial;
    if (two) {
        if (check)
            hipLaunchKernelGGL((k_fused3d_t2<T, DC, FORM, ARITH, R2C, true>), g, bl, 0, c->stream, kp, Hin, Hout,
                               Ho, A, q0, q1, q2, o0, o1, o2, part, tz, ntx, nty, ntiles, per);
        else
            hipLaunchKernelGGL((k_fused3d_t2<T, DC, FORM, ARITH, R2, false>), g, bl, 0, c->stream, kp, Hin, Hout,
                               Ho, A, q0, q1, q2, o0, o1, o2, part, tz, ntx, nty, ntiles, per);
    } else {
        if (check)
            hipLaunchKernelGGL((k_fused3d<T, DC, FORM, ARITH, kRows, true>), g, bl, 0, c->stream, kp, Hin, Hout, Ho,
                               A, q0, q1, q2, o0, o1, o2, part, tz, ntx, nty, ntiles, per);
        else
            hipLaunchKernelGGL((k_fused3d<T, DC, FORM, ARITH, kRows, false>), g, bl, 0, c->stream, kp, Hin, Hout, Ho,
                               A, q0, q1, q2, o0, o1, o2, part, tz, ntx, nty, ntiles, per);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (check)
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, (const double2*)c->partial,
                           (int64_t)ntiles, c->dred);
    return hipGetLastError();
}

static constexpr int kRows2D = 16;  // rows (waves) per block of the 2D sweep

// One fused 2D sweep (one iteration); with `check` the plain-form residual is reduced into c->dred.
template <typename T, int DC, int FORM, int ARITH>
static hipError_t sweep2d_run(ptd_ctx* c, const SweepBufs& b, bool check) {
    const KP<T> kp = make_kp<T>(c->p);
    const int ntx = (int)((c->p.nx - 2 + 61) / 62);
    const int nty = (int)((c->p.ny - 2 + kRows2D - 3) / (kRows2D - 2));
    const int ntiles = ntx * nty;
    const int per = (ntiles + 7) / 8;
    const dim3 g(8 * per), bl(64, kRows2D);
    const T* Hin = (const T*)b.Hin;
    T* Hout = (T*)b.Hout;
    const T* Ho = (const T*)b.Hold;
    const T* A = (const T*)b.aux;
    const T* q0 = (const T*)b.qin[0];
    const T* q1 = (const T*)b.qin[1];
    T* o0 = (T*)b.qout[0];
    T* o1 = (T*)b.qout[1];
    double2* part = (double2*)c->partial;
    if (check)
        hipLaunchKernelGGL((k_fused2d<T, DC, FORM, ARITH, kRows2D, true>), g, bl, 0, c->stream, kp, Hin, Hout, Ho, A,
                           q0, q1, o0, o1, part, ntx, ntiles, per);
    else
        hipLaunchKernelGGL((k_fused2d<T, DC, FORM, ARITH, kRows2D, false>), g, bl, 0, c->stream, kp, Hin, Hout, Ho,
                           A, q0, q1, o0, o1, part, ntx, ntiles, per);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (check)
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, (const double2*)c->partial,
                           (int64_t)ntiles, c->dred);
    return hipGetLastError();
}

static constexpr int kTy2D = 16;  // rows per tile (y chunk) of the two-level 2D sweep

// One two-level 2D pass (two iterations); with `check` the plain-form residual of the second is
// reduced into c->dred.
template <typename T, int DC, int FORM, int ARITH>
static hipError_t sweep2d_t2_run(ptd_ctx* c, const SweepBufs& b, bool check) {
    const KP<T> kp = make_kp<T>(c->p);
    const int ntx = (int)((c->p.nx - 2 + 59) / 60);
    const int nty = (int)((c->p.ny - 2 + kTy2D - 1) / kTy2D);
    const int ntiles = ntx * nty;
    const int per = (ntiles + 7) / 8;
    const dim3 g(8 * per), bl(64);
    const T* Hin = (const T*)b.Hin;
    T* Hout = (T*)b.Hout;
    const T* Ho = (const T*)b.Hold;
    const T* A = (const T*)b.aux;
    const T* q0 = (const T*)b.qin[0];
    const T* q1 = (const T*)b.qin[1];
    T* o0 = (T*)b.qout[0];
    T* o1 = (T*)b.qout[1];
    double2* part = (double2*)c->partial;
    if (check)
        hipLaunchKernelGGL((k_fused2d_t2<T, DC, FORM, ARITH, true>), g, bl, 0, c->stream, kp, Hin, Hout, Ho, A, q0,
                           q1, o0, o1, part, kTy2D, ntx, ntiles, per);
    else
        hipLaunchKernelGGL((k_fused2d_t2<T, DC, FORM, ARITH, false>), g, bl, 0, c->stream, kp, Hin, Hout, Ho, A,
                           q0, q1, o0, o1, part, kTy2D, ntx, ntiles, per);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (check)
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, (const double2*)c->partial,
                           (int64_t)ntiles, c->dred);
    return hipGetLastError();
}

// niter (1..kLdsHalo) 2D iterations in one LDS-resident launch; with `check` the plain-form residual of
// the last one is reduced into c->dred.
template <typename T, int DC, int FORM, int ARITH>
static hipError_t sweep2d_lds_run(ptd_ctx* c, const SweepBufs& b, int niter, bool check) {
    const KP<T> kp = make_kp<T>(c->p);
    const int ntx = (int)((c->p.nx + kLdsCore - 1) / kLdsCore);
    const int nty = (int)((c->p.ny + kLdsCore - 1) / kLdsCore);
    const int ntiles = ntx * nty;
    const int per = (ntiles + 7) / 8;
    double2* part = (double2*)c->partial;
    if (check)
        hipLaunchKernelGGL((k_lds2d<T, DC, FORM, ARITH, true>), dim3(8 * per), dim3(kLdsThreads), 0, c->stream, kp,
                           (const T*)b.Hin, (T*)b.Hout, (const T*)b.Hold, (const T*)b.aux, (const T*)b.qin[0],
                           (const T*)b.qin[1], (T*)b.qout[0], (T*)b.qout[1], niter, ntx, ntiles, per, part);
    else
        hipLaunchKernelGGL((k_lds2d<T, DC, FORM, ARITH, false>), dim3(8 * per), dim3(kLdsThreads), 0, c->stream, kp,
                           (const T*)b.Hin, (T*)b.Hout, (const T*)b.Hold, (const T*)b.aux, (const T*)b.qin[0],
                           (const T*)b.qin[1], (T*)b.qout[0], (T*)b.qout[1], niter, ntx, ntiles, per, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !check) return e;
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, (const double2*)part, (int64_t)ntiles,
                       c->dred);
    return hipGetLastError();
}

// z chunk depth of the three-level passes. A block marches tz + 4 planes (level 1 runs two planes beyond
// each end of its chunk), `slots` blocks run at once (the kernel's 16-wave blocks per CU times the CUs:
// one per CU for the float64 passes and the float32 step pass, two for the float32 linear one), and the
// blocks go in rounds: the depth of least (rounds) x (tz + 6) up to the 32-bit offset limit of a chunk
// (tz + 6 planes). Measured at 511^3 (float64 FAST, ms per iteration for tz 24 / 32 / 64 / 85 / 102 / 128 /
// 170): linear 0.883 / 0.842 / 0.828 / 0.820 / 0.803 / 0.864 / 0.867, step 1.310 / 1.199 / 1.149 / 1.127 /
// 1.106 / 1.169 / 1.183 (the model's choice: 102, five chunks of the 509 inner planes).
int t3_chunk_tz(const ptd_params& p, int64_t esz, int64_t slots) {
    const int64_t mz = p.nz - 2;
    // (deeper than 512 planes a chunk saves nothing measurable; the bound keeps the search short per launch)
    const int64_t lim = std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(mz, 512), max_chunk_tz(p, esz) - 1));
    const int64_t tiles = (p.nx - 2 + 57) / 58 * ((p.ny - 2 + kRows3 - 7) / (kRows3 - 6));
    int64_t best = lim, best_cost = -1;
    for (int64_t d = std::min<int64_t>(4, lim); d <= lim; ++d) {  // at least 4 planes (or the whole grid)
        const int64_t cost = (tiles * ((mz + d - 1) / d) + slots - 1) / slots * (d + 6);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = d;
        }
    }
    return (int)best;
}

// Blocks of `kernel` (block size `threads`) that run at once on the device: its occupancy per CU (registers,
// LDS) times the CUs; at least one per CU.
static int64_t resident_blocks(const void* kernel, int threads, int device) {
    int per_cu = 0, ncu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ncu < 1) {
        (void)hipGetLastError();
        ncu = 256;
    }
    return (int64_t)per_cu * ncu;
}

// One three-level pass (three iterations, no residual).
template <typename T, int DC, int FORM, int ARITH>
static hipError_t sweep3_run(ptd_ctx* c, const SweepBufs& b) {
    const KP<T> kp = make_kp<T>(c->p);
    constexpr int R = kRows3;
    static const int64_t slots =
        resident_blocks(reinterpret_cast<const void*>(&k_fused3d_t3<T, DC, FORM, ARITH, R>), 64 * R, c->device);
    // a chunk touches tz + 6 planes from its base (32-bit offsets): one plane fewer than the two-level limit
    const char* env = getenv("PTD_TZ");
    const int64_t want = env && atoi(env) > 0 ? atoi(env) : t3_chunk_tz(c->p, (int64_t)sizeof(T), slots);
    const int tz = (int)std::max<int64_t>(1, std::min<int64_t>(want, max_chunk_tz(c->p, (int64_t)sizeof(T)) - 1));
    const int ntx = (int)((c->p.nx - 2 + 57) / 58);
    const int nty = (int)((c->p.ny - 2 + R - 7) / (R - 6));
    const int nbz = (int)((c->p.nz - 2 + tz - 1) / tz);
    const int ntiles = ntx * nty * nbz;
    const int per = (ntiles + 7) / 8;
    hipLaunchKernelGGL((k_fused3d_t3<T, DC, FORM, ARITH, R>), dim3(8 * per), dim3(64, R), 0, c->stream, kp,
                       (const T*)b.Hin, (T*)b.Hout, (const T*)b.Hold, (const T*)b.aux, (const T*)b.qin[0],
                       (const T*)b.qin[1], (const T*)b.qin[2], (T*)b.qout[0], (T*)b.qout[1], (T*)b.qout[2], tz, ntx,
                       nty, ntiles, per);
    return hipGetLastError();
}

// The context's own fields (H, q double-buffered; the sweep flips them).
static SweepBufs ctx_bufs(ptd_ctx* c) {
    const int h = c->hcur, q = c->qcur;
    SweepBufs b;
    b.Hin = c->H[h];
    b.Hout = c->H[1 - h];
    b.Hold = c->Hold;
    b.aux = c->D;
    for (int a = 0; a < 3; ++a) {
        b.qin[a] = c->q[a][q];
        b.qout[a] = c->q[a][1 - q];
    }
    return b;
}

template <typename T, int DC, int FORM, bool TWO>
static hipError_t ctx_sweep_t(ptd_ctx* c, bool check) {
    const SweepBufs b = ctx_bufs(c);
    hipError_t e;
    if (c->p.ndim == 2 && TWO)
        e = c->p.arith == PTD_ARITH_FAST ? sweep2d_t2_run<T, DC, FORM, kFast>(c, b, check)
                                         : sweep2d_t2_run<T, DC, FORM, kExact>(c, b, check);
    else if (c->p.ndim == 2)
        e = c->p.arith == PTD_ARITH_FAST ? sweep2d_run<T, DC, FORM, kFast>(c, b, check)
                                         : sweep2d_run<T, DC, FORM, kExact>(c, b, check);
    else
        e = c->p.arith == PTD_ARITH_FAST ? sweep_run<T, DC, FORM, kFast>(c, b, TWO, check)
                                         : sweep_run<T, DC, FORM, kExact>(c, b, TWO, check);
    if (e != hipSuccess) return e;
    c->hcur = 1 - c->hcur;
    c->qcur = 1 - c->qcur;
    return hipSuccess;
}

template <typename T, int DC, int FORM>
static hipError_t fused_t(ptd_ctx* c, bool check) { return ctx_sweep_t<T, DC, FORM, false>(c, check); }
template <typename T, int DC, int FORM>
static hipError_t fused2_t(ptd_ctx* c, bool check) { return ctx_sweep_t<T, DC, FORM, true>(c, check); }
template <typename T, int DC, int FORM>
static hipError_t fused3_t(ptd_ctx* c) {
    const SweepBufs b = ctx_bufs(c);
    const hipError_t e = c->p.arith == PTD_ARITH_FAST ? sweep3_run<T, DC, FORM, kFast>(c, b)
                                                      : sweep3_run<T, DC, FORM, kExact>(c, b);
    if (e != hipSuccess) return e;
    c->hcur = 1 - c->hcur;
    c->qcur = 1 - c->qcur;
    return hipSuccess;
}
template <typename T, int DC, int FORM>
static hipError_t lds2d_t(ptd_ctx* c, int niter, bool check) {
    const SweepBufs b = ctx_bufs(c);
    const hipError_t e = c->p.arith == PTD_ARITH_FAST ? sweep2d_lds_run<T, DC, FORM, kFast>(c, b, niter, check)
                                                      : sweep2d_lds_run<T, DC, FORM, kExact>(c, b, niter, check);
    if (e != hipSuccess) return e;
    c->hcur = 1 - c->hcur;
    c->qcur = 1 - c->qcur;
    return hipSuccess;
}

// dispatch on (precision, dcase, form)
#define PTD_DISPATCH(FN, c, ...)                                                                        \
    do {                                                                                                \
        const bool f32 = (c)->esz == 4;                                                                 \
        const int dc = (c)->p.dcase, fm = (c)->p.form;                                                  \
        if (f32) {                                                                                      \
            if (dc == PTD_LIN) return fm == PTD_PERF ? FN<float, PTD_LIN, PTD_PERF>(c, ##__VA_ARGS__)    \
                                                     : FN<float, PTD_LIN, PTD_PLAIN>(c, ##__VA_ARGS__);  \
            if (dc == PTD_STEP) return fm == PTD_PERF ? FN<float, PTD_STEP, PTD_PERF>(c, ##__VA_ARGS__)  \
                                                      : FN<float, PTD_STEP, PTD_PLAIN>(c, ##__VA_ARGS__); \
            return fm == PTD_PERF ? FN<float, PTD_NONLIN, PTD_PERF>(c, ##__VA_ARGS__)                     \
                                  : FN<float, PTD_NONLIN, PTD_PLAIN>(c, ##__VA_ARGS__);                   \
        }                                                                                               \
        if (dc == PTD_LIN) return fm == PTD_PERF ? FN<double, PTD_LIN, PTD_PERF>(c, ##__VA_ARGS__)       \
                                                 : FN<double, PTD_LIN, PTD_PLAIN>(c, ##__VA_ARGS__);     \
        if (dc == PTD_STEP) return fm == PTD_PERF ? FN<double, PTD_STEP, PTD_PERF>(c, ##__VA_ARGS__)     \
                                                  : FN<double, PTD_STEP, PTD_PLAIN>(c, ##__VA_ARGS__);   \
        return fm == PTD_PERF ? FN<double, PTD_NONLIN, PTD_PERF>(c, ##__VA_ARGS__)                        \
                              : FN<double, PTD_NONLIN, PTD_PLAIN>(c, ##__VA_ARGS__);                      \
    } while (0)

hipError_t launch_flux(ptd_ctx* c, int q2mode) { PTD_DISPATCH(flux_t, c, q2mode); }
hipError_t launch_update(ptd_ctx* c) { PTD_DISPATCH(update_t, c); }
hipError_t launch_fused(ptd_ctx* c, bool check) { PTD_DISPATCH(fused_t, c, check); }
hipError_t launch_fused2(ptd_ctx* c, bool check) { PTD_DISPATCH(fused2_t, c, check); }
hipError_t launch_lds2d(ptd_ctx* c, int niter, bool check) {
    if (niter < 1 || niter > kLdsHalo || c->p.ndim != 2 || (check && c->p.form != PTD_PLAIN))
        return hipErrorInvalidValue;
    PTD_DISPATCH(lds2d_t, c, niter, check);
}
hipError_t launch_fused3(ptd_ctx* c) {
    if (c->p.ndim != 3) return hipErrorInvalidValue;
    PTD_DISPATCH(fused3_t, c);
}
// Three-level passes (3D grids whose chunk keeps its planes under 32-bit offsets) pay where the two-level
// sweep is bound by its bytes with VALU to spare: the FAST linear and step sweeps. Measured at 511^3, ms per
// iteration two- / three-level (operands two planes ahead, chunk model): float64 FAST lin 1.212 / 0.838,
// step 1.315 / 1.205, nonlinear 1.235 / 1.284 (level with the model's chunks, not taken); EXACT lin 1.44 /
// 1.66, step 2.021 / 2.019; float32 lin 0.621 / 0.541, step 0.719 / 0.645, nonlinear 0.653 / 0.779. The
// float32 passes lose on small grids, where their blocks fill the GPU worse (same context, lin / step:
// 191^3 0.0243 / 0.0359 and 0.0422 / 0.0467, 255^3 0.085 / 0.085 and 0.101 / 0.107) and win from ~300^3 on
// (319^3 0.1728 / 0.1412 and 0.1975 / 0.1761, 383^3 0.2754 / 0.2515 and 0.3187 / 0.3126, 767^3 2.054 / 1.859
// and 2.439 / 2.282, 1023^3 lin 5.09 / 4.25). PTD_T3=0 / 1 in the environment switches them off / on for
// every 3D grid.
constexpr int64_t kT3MinCellsF32 = int64_t(24) << 20;  // inner cells, ~294^3
// float64 FAST D = H^3: the three-level pass (one 16-wave block per CU) and the two-level sweep (two
// 12-wave blocks per CU) are close, and which one wins on a grid depends on how their blocks fill the rounds.
// The model: cost per iteration = (rounds of blocks) x (planes a block marches) / (iterations per launch),
// each with its own chunk depth. A three-level plane step costs less than the model's equal-cost assumption
// (measured time ratio about 0.9 x the model's). Below ~347^3 inner cells three-level passes are taken where
// they model below 97% of the two-level sweep (at least ~77^3: smaller grids are launch-bound); from ~347^3
// on, below 113%. Measured with the operand rotation, same context (ms per iteration two- / three-level,
// model ratio): 159^3 0.0502 / 0.0393 (0.905), 223^3 0.1288 / 0.1089 (0.944), 255^3 0.1719 / 0.1806 (1.056),
// 287^3 0.2331 / 0.2049 (0.870), 319^3 0.3368 / 0.2871 (0.956), 351^3 0.4057 / 0.3952 (1.096), 383^3 0.5398 /
// 0.5275 (1.125), 415^3 0.6513 / 0.6416 (1.067), 447^3 0.8529 / 0.7730 (0.979), 511^3 1.2229 / 1.0957
// (1.000), 639^3 2.3791 / 2.1000 (0.966). Before the rotation (lane shifts): 80^3 0.0128 / 0.0114 (0.917),
// 112^3 0.0202 / 0.0175 (0.889), 191^3 0.0846 / 0.0846 (0.958), 767^3 4.175 / 4.034 (1.056), 1023^3 10.18 /
// 9.24 (0.977); 63^3 0.0086 / 0.0106 (0.833, below the size floor).
constexpr int64_t kT3MinCellsNonlin = int64_t(448) << 10;  // inner cells, ~77^3
constexpr int64_t kT3LargeCellsNonlin = int64_t(40) << 20;  // inner cells, ~347^3
static bool nonlin_t3_model(const ptd_params& p) {
    const int64_t cells = (p.nx - 2) * (p.ny - 2) * (p.nz - 2);
    if (cells < kT3MinCellsNonlin) return false;
    const ChunkModel m(p);
    const double c2 = (double)m.cost(default_chunk_tz(p)) / 2.0;
    constexpr int64_t slots3 = 256;  // the float64 nonlinear three-level pass: one block per CU
    const int64_t tz3 = t3_chunk_tz(p, 8, slots3);
    const int64_t mz = p.nz - 2;
    const int64_t tiles3 = (p.nx - 2 + 57) / 58 * ((p.ny - 2 + kRows3 - 7) / (kRows3 - 6));
    const double c3 = (double)((tiles3 * ((mz + tz3 - 1) / tz3) + slots3 - 1) / slots3 * (tz3 + 6)) / 3.0;
    return c3 < (cells >= kT3LargeCellsNonlin ? 1.13 : 0.97) * c2;
}
bool fused3_wanted(const ptd_params& p, int64_t esz) {
    if (!(p.ndim == 3 && p.nx >= 3 && p.ny >= 3 && p.nz >= 3 && max_chunk_tz(p, esz) >= 2)) return false;
    const char* env = getenv("PTD_T3");
    if (env && env[0] == '0') return false;
    if (env && env[0] == '1') return true;
    if (p.arith != PTD_ARITH_FAST) return false;
    if (p.dcase == PTD_NONLIN) return esz == 8 && nonlin_t3_model(p);
    if (p.dcase != PTD_LIN && p.dcase != PTD_STEP) return false;
    return esz == 8 || (p.nx - 2) * (p.ny - 2) * (p.nz - 2) >= kT3MinCellsF32;
}
bool fused3_available(const ptd_ctx* c) {
    return fused_available(c) && fused3_wanted(c->p, (int64_t)c->esz);
}
int next_launch_iters(int64_t n, bool check_last, bool t2, bool t3) {
    if (t3 && n - (check_last ? 2 : 0) >= 3 && !(check_last && n == 6)) return 3;
    if (t2 && n >= 2) return 2;
    return 1;
}
int lds2d_max_iters() { return kLdsHalo; }

hipError_t launch_residual(ptd_ctx* c, bool write_resh) {
    const bool f32 = c->esz == 4;
    if (c->p.form == PTD_PERF) return f32 ? residual_t<float, PTD_PERF>(c, write_resh) : residual_t<double, PTD_PERF>(c, write_resh);
    return f32 ? residual_t<float, PTD_PLAIN>(c, write_resh) : residual_t<double, PTD_PLAIN>(c, write_resh);
}

hipError_t launch_iter_params(ptd_ctx* c) {
    const dim3 g(grid_for(c->nR)), bl(kBlock);
    if (c->esz == 4) {
        const KP<float> kp = make_kp<float>(c->p);
        hipLaunchKernelGGL(k_iter_params<float>, g, bl, 0, c->stream, kp, (const float*)c->D, (float*)c->dtau, c->nR);
    } else {
        const KP<double> kp = make_kp<double>(c->p);
        hipLaunchKernelGGL(k_iter_params<double>, g, bl, 0, c->stream, kp, (const double*)c->D, (double*)c->dtau, c->nR);
    }
    return hipGetLastError();
}

bool fused_available(const ptd_ctx* c) {
    return (c->p.ndim == 3 && c->p.nx >= 3 && c->p.ny >= 3 && c->p.nz >= 3 && max_chunk_tz(c->p, (int64_t)c->esz) >= 1) ||
           (c->p.ndim == 2 && c->p.nx >= 3 && c->p.ny >= 3);
}
// Two-level 2D sweeps march one wavefront per 60-column x 16-row tile: below ~1024 tiles they cannot
// fill the GPU and the single-level sweep (16-wave blocks) is faster (measured: 512^2 7.4 vs 9.6 us per
// iteration single / two-level, 1024^2 14.4 vs 11.2, 2048^2 40.6 vs 26.0).
bool use_2d_t2(const ptd_params& p) {
    if (p.ndim != 2) return false;
    const int64_t ntx = (p.nx - 2 + 59) / 60, nty = (p.ny - 2 + kTy2D - 1) / kTy2D;
    return ntx * nty >= 1024;
}
// the sweep a single solve runs for p: LDS-resident 2D iterations take precedence (run_iterations)
int batch_kind(const ptd_params& p) {
    if (p.ndim == 3) return kSweep3D;
    if (use_2d_lds_params(p)) return kSweepLds2D;
    return use_2d_t2(p) ? kSweep2D2 : kSweep2D;
}

// LDS-resident 2D iterations pay off while a launch of single sweeps is latency-bound: grids of at most
// kLdsMaxTiles 32x32 cores. PTD_LDS2D=0 / 1 in the environment switches them off / on for every 2D grid.
static constexpr int64_t kLdsMaxTiles = 1024;
int64_t lds2d_tiles(const ptd_params& p) {
    return ((p.nx + kLdsCore - 1) / kLdsCore) * ((p.ny + kLdsCore - 1) / kLdsCore);
}
bool use_2d_lds_params(const ptd_params& p) {
    if (p.ndim != 2 || p.nx < 3 || p.ny < 3) return false;
    const char* env = getenv("PTD_LDS2D");
    if (env && env[0] == '0') return false;
    if (env && env[0] == '1') return true;
    return lds2d_tiles(p) <= kLdsMaxTiles;
}
bool use_2d_lds(const ptd_ctx* c) { return fused_available(c) && use_2d_lds_params(c->p); }
// PTD_T2=2 in the environment also runs the single-problem 2D solves on two-level sweeps below that
// threshold (the same bits as single sweeps; lets the tests compare them with the oracle on small grids).
// Batches keep the threshold.
bool fused2_available(const ptd_ctx* c) {
    const char* env = getenv("PTD_T2");
    const bool force2d = env && env[0] == '2';
    return fused_available(c) && (c->p.ndim == 3 || use_2d_t2(c->p) || (c->p.ndim == 2 && force2d));
}

// ---------------------------------------------------------------------------------------------
// mixed precision: float32 correction sweeps around a float64 iterate
// ---------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock)
k_f64_to_f32(float* __restrict__ dst, const double* __restrict__ src, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t < n) dst[t] = (float)src[t];
}

hipError_t launch_f64_to_f32(ptd_ctx* c, float* dst, const double* src, int64_t n) {
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, dst, src, n);
    return hipGetLastError();
}

// Face state conversion between the true fluxes q (float64, they persist over physical steps as the
// reference's qHx/qHy/qHz do) and the correction fluxes p = q + g(Hc) (float32, stored times sc, the
// axis' scale of the correction sweep) of the correction sweep, g(Hc) = coef * dHc / h the equilibrium
// flux magnitude of the float64 iterate.
template <int DC, int AXIS, bool TO_P>
__global__ void __launch_bounds__(kBlock)
k_flux_state(const KP<double> p, const double* __restrict__ H, const double* __restrict__ D,
             double* __restrict__ q, float* __restrict__ pf, int64_t count, double sc, double isc) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= count) return;
    int64_t a, step;
    double ih;
    if constexpr (AXIS == 0) {
        int64_t f, jj, kk;
        split3(t, p.nx - 1, p.my, f, jj, kk);
        a = f + p.nx * ((jj + p.hy) + p.ny * (kk + p.hz));
        step = 1;
        ih = p._dx;
    } else if constexpr (AXIS == 1) {
        int64_t ii, f, kk;
        split3(t, p.mx, p.ny - 1, ii, f, kk);
        a = (ii + 1) + p.nx * (f + p.ny * (kk + p.hz));
        step = p.nx;
        ih = p._dy;
    } else {
        int64_t ii, jj, f;
        split3(t, p.mx, p.my, ii, jj, f);
        a = (ii + 1) + p.nx * ((jj + 1) + p.ny * f);
        step = p.nx * p.ny;
        ih = p._dz;
    }
    const double Ha = H[a], Hb = H[a + step];
    double coef;
    if constexpr (DC == PTD_NONLIN) {
        const double hf = 0.5 * (Ha + Hb);
        coef = hf * hf * hf;
    } else if constexpr (DC == PTD_STEP) {
        coef = (D[a] + D[a + step]) * 0.5;
    } else {
        coef = p.d0;
    }
    const double g = coef * (ih * (Hb - Ha));
    if constexpr (TO_P) pf[t] = (float)((q[t] + g) * sc);
    else q[t] = (double)pf[t] * isc - g;
}

template <int DC, bool TO_P>
static hipError_t flux_state_t(ptd_ctx* c) {
    const KP<double> kp = make_kp<double>(c->p);
    const double* H = (const double*)c->H[c->hcur];
    const double* D = (const double*)c->D;
    for (int a = 0; a < 3; ++a) {
        const int64_t n = c->nQ[a];
        if (n == 0) continue;
        double* q = (double*)c->q[a][c->qcur];
        float* pf = (float*)c->pq[a][c->pcur];
        const dim3 g(grid_for(n)), bl(kBlock);
        const double sc = corr_flux_scale(c->p, DC, a), isc = 1.0 / sc;
        if (a == 0)
            hipLaunchKernelGGL((k_flux_state<DC, 0, TO_P>), g, bl, 0, c->stream, kp, H, D, q, pf, n, sc, isc);
        else if (a == 1)
            hipLaunchKernelGGL((k_flux_state<DC, 1, TO_P>), g, bl, 0, c->stream, kp, H, D, q, pf, n, sc, isc);
        else
            hipLaunchKernelGGL((k_flux_state<DC, 2, TO_P>), g, bl, 0, c->stream, kp, H, D, q, pf, n, sc, isc);
    }
    return hipGetLastError();
}

hipError_t launch_flux_state(ptd_ctx* c, bool to_p) {
    if (c->p.dcase == PTD_NONLIN) return to_p ? flux_state_t<PTD_NONLIN, true>(c) : flux_state_t<PTD_NONLIN, false>(c);
    if (c->p.dcase == PTD_STEP) return to_p ? flux_state_t<PTD_STEP, true>(c) : flux_state_t<PTD_STEP, false>(c);
    return to_p ? flux_state_t<PTD_LIN, true>(c) : flux_state_t<PTD_LIN, false>(c);
}

// Start of a correction (one pass over every cell of the float64 iterate Hc): the consistent residual
// R of the float64 system on the inner cells (residual fluxes from Hc itself: compute_flux_res! +
// check_res!, scripts/diff_3D/diff_3D_nonlin_perf.jl:44-53, in the context's form and arithmetic),
// written to ResH, with its per-block (sum r^2, max |r|); the float32 source src = R*dt; float32(Hc) for
// the correction sweep's aux channel (D = H^3); and delta = 0.
template <int DC, int FORM, int ARITH>
__global__ void __launch_bounds__(kBlock)
k_outer_residual(const KP<double> p, const double* __restrict__ H, const double* __restrict__ Hold,
                 const double* __restrict__ D, double* __restrict__ resh, float* __restrict__ src,
                 float* __restrict__ hcf, float* __restrict__ dl, double2* __restrict__ partial, int64_t n) {
    using Ph = Phys<double, DC, FORM, ARITH>;
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double s = 0.0, m = 0.0;
    if (g < n) {
        const double Hc = H[g];
        if (hcf) hcf[g] = (float)Hc;
        dl[g] = 0.0f;
        const int64_t nx = p.nx, ny = p.ny, pl = nx * ny;
        int64_t i, j, k;
        split3(g, nx, ny, i, j, k);
        if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 && k <= p.nz - 2) {
            auto Dv = [&](int64_t o) { return DC == PTD_STEP ? D[o] : 0.0; };
            const double Dc = Dv(g);
            const double fxw = Ph::face_res(p, H[g - 1], Hc, Dv(g - 1), Dc, p.dx, p._dx);
            const double fxe = Ph::face_res(p, Hc, H[g + 1], Dc, Dv(g + 1), p.dx, p._dx);
            const double fys = Ph::face_res(p, H[g - nx], Hc, Dv(g - nx), Dc, p.dy, p._dy);
            const double fyn = Ph::face_res(p, Hc, H[g + nx], Dc, Dv(g + nx), p.dy, p._dy);
            const double fzd = Ph::face_res(p, H[g - pl], Hc, Dv(g - pl), Dc, p.dz, p._dz);
            const double fzu = Ph::face_res(p, Hc, H[g + pl], Dc, Dv(g + pl), p.dz, p._dz);
            const double dv = Ph::template div<3>(p, fxe - fxw, fyn - fys, fzu - fzd);
            const double r = Ph::resid(p, Hc, Hold[g], dv);
            resh[(i - 1) + p.mx * ((j - 1) + p.my * (k - 1))] = r;
            src[g] = (float)(r * p.dt);
            s = r * r;
            m = fabs(r);
        } else {
            src[g] = 0.0f;
        }
    }
    block_sum_max(s, m, threadIdx.x, kBlock, &partial[blockIdx.x]);
}

// Correction boundary in one pass (flux-state move, Hc += delta and k_outer_residual's outputs): with the new
// iterate Hn = Hc + delta evaluated on the fly (delta is 0 on the boundary cells), the correction fluxes
// move from Hc to Hn, Hn goes to the other H buffer, and the residual of the float64 system at Hn gives
// ResH, the next source R*dt, float32(Hn) and the partials. The next correction's delta = 0 goes to the
// other delta buffer, so no thread writes what another reads. A block owns a 64 x 4 (x, y) tile of a z
// chunk, a thread one (i, j) column of it, marching along z with the z neighbours in registers. The y
// neighbours come from the other waves of the tile or, at its edges, from the tiles beside it, which
// the XCD-aware numbering places on the same XCD at about the same time, so they are L2 hits.
constexpr int kOuterTx = 64, kOuterTy = kBlock / kOuterTx, kOuterTz = 32;  // tile and z chunk (planes per thread)
template <int DC, int FORM, int ARITH>
__global__ void __launch_bounds__(kBlock)
k_outer_step(const KP<double> p, const double* __restrict__ Hin, double* __restrict__ Hout,
             const float* __restrict__ dl, float* __restrict__ dl_next, const double* __restrict__ Hold,
             const double* __restrict__ D, double* __restrict__ resh, float* __restrict__ src,
             float* __restrict__ hcf, float* __restrict__ px, float* __restrict__ py, float* __restrict__ pz,
             double3 sc, double3 isc, double2* __restrict__ partial, int tz, int ntx, int ntiles, int per_xcd) {
    using Ph = Phys<double, DC, FORM, ARITH>;
    const int64_t nx = p.nx, ny = p.ny, nz = p.nz, pl = nx * ny;
    const int b = blockIdx.x;
    const int tile = (b & 7) * per_xcd + (b >> 3);  // XCD-aware: each XCD walks a contiguous tile range
    const int tyl = tile / ntx;
    const int64_t i = (int64_t)(tile - tyl * ntx) * kOuterTx + threadIdx.x;
    const int64_t j = (int64_t)tyl * kOuterTy + threadIdx.y;
    const int64_t k0 = (int64_t)blockIdx.y * tz, k1 = k0 + tz < nz ? k0 + tz : nz;
    const int64_t part = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    double s = 0.0, m = 0.0;
    if (tile < ntiles && i < nx && j < ny) {
        const bool ii = i >= 1 && i <= nx - 2, ji = j >= 1 && j <= ny - 2;
        auto Dv = [&](int64_t o) { return DC == PTD_STEP ? D[o] : 0.0; };
        auto gflux = [&](double Ha, double Hb, double Da, double Db, double ih) {
            double coef;
            if constexpr (DC == PTD_NONLIN) {
                const double hf = 0.5 * (Ha + Hb);
                coef = hf * hf * hf;
            } else if constexpr (DC == PTD_STEP) {
                coef = (Da + Db) * 0.5;
            } else {
                coef = p.d0;
            }
            return coef * (ih * (Hb - Ha));
        };
        // move of the face (a | this cell) with the old flux state pv: old Hc and new Hn at both cells
        auto moved = [&](float pv, double Hoa, double Hob, double Hna, double Hnb, double Da, double Db, double ih,
                         double sa, double isa) {
            const double go = gflux(Hoa, Hob, Da, Db, ih);
            const double gn = gflux(Hna, Hnb, Da, Db, ih);
            return (float)(((double)pv * isa - go + gn) * sa);
        };
        const int64_t col = i + nx * j;
        int64_t g = col + pl * k0;
        // flux / ResH offsets of plane k (face arrays and ResH are indexed from the first inner plane; only
        // used where the face / cell exists), advanced by one plane of each array per step
        const int64_t sx = (nx - 1) * (ny - 2), sy = (nx - 2) * (ny - 1), sz = (nx - 2) * (ny - 2);
        int64_t ox = (i - 1) + (nx - 1) * (j - 1) + sx * (k0 - 1);
        int64_t oy = (i - 1) + (nx - 2) * (j - 1) + sy * (k0 - 1);
        int64_t oz = (i - 1) + (nx - 2) * (j - 1) + sz * (k0 - 1);  // also ResH: (nx-2, ny-2, nz-2)
        // x / y neighbour columns, clamped to this column on the boundary (values then unused)
        const int64_t dw = i >= 1 ? 1 : 0, de = i + 1 < nx ? 1 : 0, ds = j >= 1 ? nx : 0, dn = j + 1 < ny ? nx : 0;
        // z window: old Hc, new Hn and D of planes k-1 and k
        double Hom = 0.0, Hnm = 0.0, Dm = 0.0;
        if (k0 >= 1) {
            Hom = Hin[g - pl];
            Hnm = Hom + (double)dl[g - pl];
            Dm = Dv(g - pl);
        }
        double Hoc = Hin[g];
        double Hnc = Hoc + (double)dl[g];
        double Dc = Dv(g);
        // The loads of a plane step are unconditional and in range (for a face or cell that does not
        // exist they read a valid address, the value then unused), so they issue together instead of
        // each behind its own branch and wait.
        for (int64_t k = k0; k < k1; ++k, g += pl, ox += sx, oy += sy, oz += sz) {
            const bool ki = k >= 1 && k <= nz - 2;
            const bool vx = i >= 1 && ji && ki, vy = j >= 1 && ii && ki, vz = k >= 1 && ii && ji;
            const bool inner = ii && ji && ki;
            const int64_t gp = k + 1 < nz ? g + pl : g;  // plane k+1 (unused past the last plane)
            const double Hop = Hin[gp];
            const float dlp = dl[gp];
            const double Dp = Dv(gp);
            const double Hw = Hin[g - dw], He = Hin[g + de], Hs = Hin[g - ds], Hn_ = Hin[g + dn];
            const float dlw = dl[g - dw], dle = dl[g + de], dls = dl[g - ds], dln = dl[g + dn];
            const double Dw = Dv(g - dw), De = Dv(g + de), Ds = Dv(g - ds), Dn = Dv(g + dn);
            const float pxv = px[vx ? ox : 0], pyv = py[vy ? oy : 0], pzv = pz[vz ? oz : 0];
            const double Ho = Hold[g];
            const double Hnp = Hop + (double)dlp;
            const double Hnw = Hw + (double)dlw, Hne = He + (double)dle, Hns = Hs + (double)dls, Hnn = Hn_ + (double)dln;
            if (vx) px[ox] = moved(pxv, Hw, Hoc, Hnw, Hnc, Dw, Dc, p._dx, sc.x, isc.x);
            if (vy) py[oy] = moved(pyv, Hs, Hoc, Hns, Hnc, Ds, Dc, p._dy, sc.y, isc.y);
            if (vz) pz[oz] = moved(pzv, Hom, Hoc, Hnm, Hnc, Dm, Dc, p._dz, sc.z, isc.z);
            if (hcf) hcf[g] = (float)Hnc;
            dl_next[g] = 0.0f;
            if (inner) {
                Hout[g] = Hnc;
                const double fxw = Ph::face_res(p, Hnw, Hnc, Dw, Dc, p.dx, p._dx);
                const double fxe = Ph::face_res(p, Hnc, Hne, Dc, De, p.dx, p._dx);
                const double fys = Ph::face_res(p, Hns, Hnc, Ds, Dc, p.dy, p._dy);
                const double fyn = Ph::face_res(p, Hnc, Hnn, Dc, Dn, p.dy, p._dy);
                const double fzd = Ph::face_res(p, Hnm, Hnc, Dm, Dc, p.dz, p._dz);
                const double fzu = Ph::face_res(p, Hnc, Hnp, Dc, Dp, p.dz, p._dz);
                const double dv = Ph::template div<3>(p, fxe - fxw, fyn - fys, fzu - fzd);
                const double r = Ph::resid(p, Hnc, Ho, dv);
                resh[oz] = r;
                src[g] = (float)(r * p.dt);
                s += r * r;
                m = nanmax(fabs(r), m);
            } else {
                src[g] = 0.0f;
            }
            Hom = Hoc;
            Hnm = Hnc;
            Dm = Dc;
            Hoc = Hop;
            Hnc = Hnp;
            Dc = Dp;
        }
    }
    block_sum_max(s, m, threadIdx.y * kOuterTx + threadIdx.x, kBlock, &partial[part]);
}

template <int DC, int FORM, int ARITH>
static hipError_t outer_step_t(ptd_ctx* c) {
    const KP<double> kp = make_kp<double>(c->p);
    const int64_t ntx = (c->p.nx + kOuterTx - 1) / kOuterTx, nty = (c->p.ny + kOuterTy - 1) / kOuterTy;
    const int64_t ntiles = ntx * nty, per = (ntiles + 7) / 8;
    if (8 * per > (int64_t)INT32_MAX) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(8 * per), (unsigned)((c->p.nz + kOuterTz - 1) / kOuterTz));
    const int64_t gb = (int64_t)grid.x * grid.y * grid.z;
    if (gb > c->npartial) return hipErrorInvalidValue;
    const int dc = c->p.dcase;
    const double3 sc = make_double3(corr_flux_scale(c->p, dc, 0), corr_flux_scale(c->p, dc, 1),
                                    corr_flux_scale(c->p, dc, 2));
    const double3 isc = make_double3(1.0 / sc.x, 1.0 / sc.y, 1.0 / sc.z);
    hipLaunchKernelGGL((k_outer_step<DC, FORM, ARITH>), grid, dim3(kOuterTx, kOuterTy), 0, c->stream, kp,
                       (const double*)c->H[c->hcur], (double*)c->H[1 - c->hcur], (const float*)c->dl[c->dcur],
                       (float*)c->dl[1 - c->dcur], (const double*)c->Hold, (const double*)c->D, (double*)c->resh,
                       (float*)c->src, (float*)c->hcf, (float*)c->pq[0][c->pcur], (float*)c->pq[1][c->pcur],
                       (float*)c->pq[2][c->pcur], sc, isc, (double2*)c->partial, kOuterTz, (int)ntx, (int)ntiles,
                       (int)per);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, (const double2*)c->partial,
                       (int64_t)gb, c->dred);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        c->hcur = 1 - c->hcur;
        c->dcur = 1 - c->dcur;
    }
    return e;
}

// the residual of the float64 system in the context's arithmetic (EXACT: the reference's check_res!
// expressions; FAST: the rearranged ones, no divisions)
template <int ARITH>
static hipError_t outer_step_a(ptd_ctx* c) {
    const bool perf = c->p.form == PTD_PERF;
    switch (c->p.dcase) {
        case PTD_LIN: return perf ? outer_step_t<PTD_LIN, PTD_PERF, ARITH>(c) : outer_step_t<PTD_LIN, PTD_PLAIN, ARITH>(c);
        case PTD_STEP: return perf ? outer_step_t<PTD_STEP, PTD_PERF, ARITH>(c) : outer_step_t<PTD_STEP, PTD_PLAIN, ARITH>(c);
        default: return perf ? outer_step_t<PTD_NONLIN, PTD_PERF, ARITH>(c) : outer_step_t<PTD_NONLIN, PTD_PLAIN, ARITH>(c);
    }
}

hipError_t launch_outer_step(ptd_ctx* c) {
    return c->p.arith == PTD_ARITH_FAST ? outer_step_a<kFast>(c) : outer_step_a<kExact>(c);
}

template <int DC, int FORM, int ARITH>
static hipError_t outer_residual_t(ptd_ctx* c) {
    const KP<double> kp = make_kp<double>(c->p);
    const unsigned gb = grid_for(c->nH);
    if ((int64_t)gb > c->npartial) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_outer_residual<DC, FORM, ARITH>), dim3(gb), dim3(kBlock), 0, c->stream, kp,
                       (const double*)c->H[c->hcur], (const double*)c->Hold, (const double*)c->D, (double*)c->resh,
                       (float*)c->src, (float*)c->hcf, (float*)c->dl[c->dcur], (double2*)c->partial, c->nH);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, (const double2*)c->partial,
                       (int64_t)gb, c->dred);
    return hipGetLastError();
}

template <int ARITH>
static hipError_t outer_residual_a(ptd_ctx* c) {
    const bool perf = c->p.form == PTD_PERF;
    switch (c->p.dcase) {
        case PTD_LIN: return perf ? outer_residual_t<PTD_LIN, PTD_PERF, ARITH>(c) : outer_residual_t<PTD_LIN, PTD_PLAIN, ARITH>(c);
        case PTD_STEP: return perf ? outer_residual_t<PTD_STEP, PTD_PERF, ARITH>(c) : outer_residual_t<PTD_STEP, PTD_PLAIN, ARITH>(c);
        default: return perf ? outer_residual_t<PTD_NONLIN, PTD_PERF, ARITH>(c) : outer_residual_t<PTD_NONLIN, PTD_PLAIN, ARITH>(c);
    }
}

hipError_t launch_outer_residual(ptd_ctx* c) {
    return c->p.arith == PTD_ARITH_FAST ? outer_residual_a<kFast>(c) : outer_residual_a<kExact>(c);
}

// One or two float32 correction iterations on (delta, correction fluxes) with source src = R*dt in
// the Hold slot; for D = H^3 float32(Hc) rides in the aux channel.
hipError_t launch_delta_sweep(ptd_ctx* c, bool two, bool check) {
    SweepBufs b;
    const int d = c->dcur, q = c->pcur;
    b.Hin = c->dl[d];
    b.Hout = c->dl[1 - d];
    b.Hold = c->src;
    for (int a = 0; a < 3; ++a) {
        b.qin[a] = c->pq[a][q];
        b.qout[a] = c->pq[a][1 - q];
    }
    hipError_t e;
    if (c->p.dcase == PTD_NONLIN) {
        b.aux = c->hcf;
        e = sweep_run<float, kDelta, PTD_PERF, kFast>(c, b, two, check);
    } else if (c->p.dcase == PTD_STEP) {
        b.aux = c->Df;
        e = sweep_run<float, PTD_STEP, PTD_PERF, kFast>(c, b, two, check);
    } else {
        b.aux = nullptr;
        e = sweep_run<float, PTD_LIN, PTD_PERF, kFast>(c, b, two, check);
    }
    if (e != hipSuccess) return e;
    c->dcur = 1 - d;
    c->pcur = 1 - q;
    return hipSuccess;
}

// Three float32 correction iterations in one three-level pass (linear and step D; D = H^3 keeps its
// two-level correction sweep, whose flux is the correction flux around Hc).
hipError_t launch_delta_sweep3(ptd_ctx* c) {
    if (c->p.ndim != 3 || (c->p.dcase != PTD_LIN && c->p.dcase != PTD_STEP)) return hipErrorInvalidValue;
    SweepBufs b;
    const int d = c->dcur, q = c->pcur;
    b.Hin = c->dl[d];
    b.Hout = c->dl[1 - d];
    b.Hold = c->src;
    for (int a = 0; a < 3; ++a) {
        b.qin[a] = c->pq[a][q];
        b.qout[a] = c->pq[a][1 - q];
    }
    hipError_t e;
    if (c->p.dcase == PTD_STEP) {
        b.aux = c->Df;
        e = sweep3_run<float, PTD_STEP, PTD_PERF, kFast>(c, b);
    } else {
        b.aux = nullptr;
        e = sweep3_run<float, PTD_LIN, PTD_PERF, kFast>(c, b);
    }
    if (e != hipSuccess) return e;
    c->dcur = 1 - d;
    c->pcur = 1 - q;
    return hipSuccess;
}
bool delta_t3(const ptd_ctx* c) {
    if (c->p.dcase != PTD_LIN && c->p.dcase != PTD_STEP) return false;
    ptd_params p = c->p;
    p.arith = PTD_ARITH_FAST;  // the correction sweeps are FAST whatever the float64 passes use
    return fused3_wanted(p, 4);
}

// Number of (sum,max) partials the reductions can produce for this context.
// Residual partials the checking launches of a fused float64 / float32 PT solve (2D / 3D) write: the
// largest tile count over the single-level and two-level checking tilings (the two-level one counted with
// its smallest block, 12 rows), at the z chunk depth the checking launches use.
int64_t fused_check_tiles(const ptd_ctx* c) {
    const ptd_params& p = c->p;
    if (p.ndim == 2) {
        const int64_t ntx = (p.nx - 2 + 61) / 62, nty = (p.ny - 2 + kRows2D - 3) / (kRows2D - 2);
        const int64_t ntx2 = (p.nx - 2 + 59) / 60, nty2 = (p.ny - 2 + kTy2D - 1) / kTy2D;
        return std::max<int64_t>(std::max<int64_t>(ntx * nty, ntx2 * nty2), lds2d_tiles(p));
    }
    const int64_t tz = fused_tz(c, true);
    const int64_t nbz = (p.nz - 2 + tz - 1) / tz;
    const int64_t ntx = (p.nx - 2 + 61) / 62, nty = (p.ny - 2 + kRows - 3) / (kRows - 2);
    const int64_t ntx2 = (p.nx - 2 + 59) / 60, nty2 = (p.ny - 2 + kMinRows2 - 5) / (kMinRows2 - 4);
    return std::max<int64_t>(ntx * nty, ntx2 * nty2) * nbz;
}

int64_t partial_capacity(const ptd_params& p, int64_t nR) {
    int64_t nb = (nR + kBlock - 1) / kBlock;
    if (p.ndim == 2) {
        const int64_t ntx = (p.nx - 2 + 61) / 62, nty = (p.ny - 2 + kRows2D - 3) / (kRows2D - 2);
        nb = std::max<int64_t>(nb, ntx * nty);
        const int64_t ntx2 = (p.nx - 2 + 59) / 60, nty2 = (p.ny - 2 + kTy2D - 1) / kTy2D;
        nb = std::max<int64_t>(nb, ntx2 * nty2);
        nb = std::max<int64_t>(nb, lds2d_tiles(p));
    }
    if (p.ndim == 3) {
        const int64_t ntx = (p.nx - 2 + 61) / 62, nty = (p.ny - 2 + kRows - 3) / (kRows - 2);
        // smallest chunk is one plane
        nb = std::max<int64_t>(nb, ntx * nty * (p.nz - 2));
        const int64_t ntx2 = (p.nx - 2 + 59) / 60, nty2 = (p.ny - 2 + kMinRows2 - 5) / (kMinRows2 - 4);
        nb = std::max<int64_t>(nb, ntx2 * nty2 * (p.nz - 2));
        // k_outer_step: one partial per block, its tiles padded to a multiple of 8 (XCD mapping)
        const int64_t to = ((p.nx + kOuterTx - 1) / kOuterTx) * ((p.ny + kOuterTy - 1) / kOuterTy);
        nb = std::max<int64_t>(nb, (to + 7) / 8 * 8 * ((p.nz + kOuterTz - 1) / kOuterTz));
    }
    return std::max<int64_t>(nb, 1);
}

// ---------------------------------------------------------------------------------------------
// gradients of the lin / step solves (adjoint)
// ---------------------------------------------------------------------------------------------

// Adjoint source of L = 1/2 sum_inner (H - Hobs)^2: out = (H - Hobs)*dt on the inner cells, 0 on the
// boundary cells (the adjoint has homogeneous Dirichlet values); per-block partials of r^2/2 for L.
__global__ void __launch_bounds__(kBlock)
k_adj_source(KP<double> p, const double* __restrict__ H, const double* __restrict__ Hobs, double* __restrict__ out,
             double2* __restrict__ partial, int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double l = 0.0;
    if (c < n) {
        const int64_t i = c % p.nx, rest = c / p.nx, j = rest % p.ny, k = rest / p.ny;
        const bool in = i >= 1 && i <= p.nx - 2 && (p.hy == 0 || (j >= 1 && j <= p.ny - 2)) &&
                        (p.hz == 0 || (k >= 1 && k <= p.nz - 2));
        const double r = in ? H[c] - Hobs[c] : 0.0;
        out[c] = r * p.dt;
        l = 0.5 * r * r;
    }
    block_sum_max(l, 0.0, threadIdx.x, kBlock, &partial[blockIdx.x]);
}

// Per-block partials of sum H^2 over the inner cells (the scale of a forward solve's right-hand side).
__global__ void __launch_bounds__(kBlock)
k_sumsq_inner(KP<double> p, const double* __restrict__ H, double2* __restrict__ partial, int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double l = 0.0;
    if (c < n) {
        const int64_t i = c % p.nx, rest = c / p.nx, j = rest % p.ny, k = rest / p.ny;
        const bool in = i >= 1 && i <= p.nx - 2 && (p.hy == 0 || (j >= 1 && j <= p.ny - 2)) &&
                        (p.hz == 0 || (k >= 1 && k <= p.nz - 2));
        l = in ? H[c] * H[c] : 0.0;
    }
    block_sum_max(l, 0.0, threadIdx.x, kBlock, &partial[blockIdx.x]);
}

// Accumulates one step's contribution to the gradient, per cell (gather over the faces at the cell):
//   gD[c]  += 1/2 sum_{faces f at c} (H_b - H_a)(mu_a - mu_b)/h^2        (face f between a < b)
//   gH0[c] += sum_{faces f between boundary c and inner i} mu_i D_f/h^2  (c a boundary cell)
// plus gH0 += mu/dt on the inner cells for the first step. A face exists where the reference's flux
// array has it: transverse indices inner. mu is 0 on the boundary cells.
template <int NDIM, bool STEP>
__global__ void __launch_bounds__(kBlock)
k_grad_accum(KP<double> p, const double* __restrict__ mu, const double* __restrict__ Hn, const double* __restrict__ D,
             double* __restrict__ gD, double* __restrict__ gH0, int first, int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const int64_t nx = p.nx, ny = p.ny, nz = p.nz;
    int64_t i, j, k;
    split3(c, nx, ny, i, j, k);
    const int64_t idx[3] = {i, j, k};
    const int64_t nn[3] = {nx, ny, nz};
    const int64_t st[3] = {1, nx, nx * ny};
    const double w[3] = {p._dx * p._dx, p._dy * p._dy, p._dz * p._dz};
    auto inner = [&](int ax) { return idx[ax] >= 1 && idx[ax] <= nn[ax] - 2; };
    const bool cin = inner(0) && (NDIM < 2 || inner(1)) && (NDIM < 3 || inner(2));
    double g = 0.0, gb = 0.0;
#pragma unroll
    for (int ax = 0; ax < NDIM; ++ax) {
        bool trans = true;  // transverse indices inner: the flux array has faces on this line
#pragma unroll
        for (int o = 0; o < NDIM; ++o)
            if (o != ax) trans = trans && inner(o);
        if (!trans) continue;
#pragma unroll
        for (int side = 0; side < 2; ++side) {  // 0: face (c-1 | c), 1: face (c | c+1)
            const bool has = side == 0 ? idx[ax] >= 1 : idx[ax] <= nn[ax] - 2;
            if (!has) continue;
            const int64_t a = side == 0 ? c - st[ax] : c, b = side == 0 ? c : c + st[ax];
            g += 0.5 * (Hn[b] - Hn[a]) * (mu[a] - mu[b]) * w[ax];
            if (!cin) {
                const int64_t o = side == 0 ? a : b;  // the neighbour across the face
                const double Df = STEP ? (D[a] + D[b]) * 0.5 : p.d0;
                gb += mu[o] * Df * w[ax];
            }
        }
    }
    gD[c] += g;
    if (cin) {
        if (first) gH0[c] += mu[c] * p._dt;
    } else {
        gH0[c] += gb;
    }
}

// ---- adjoint of the nonlinear (D = H^3) steps ---------------------------------------------------
//
// One backward-Euler step solves F(H) = (H - Hold)/dt + div q(H) = 0 on the inner cells with
// q_f = -G_f, G_f = c_f (H_b - H_a)/h, c_f = av^3, av = (H_a + H_b)*0.5 (scripts/diff_3D/
// diff_3D_nonlin_multixpu.jl:37-50). Its Jacobian J = I/dt - S J_G/h (S: cell-face incidence, J_G = dG/dH)
// is not symmetric: dG_f/dH_a = (1.5 av^2 dH - c_f)/h, dG_f/dH_b = (1.5 av^2 dH + c_f)/h. The adjoint
// J^T lam = rhs is solved by the transpose of the forward PT iteration linearised at the converged H
// (theta_f and dtau frozen there; the same spectrum, so the same convergence rate):
//   face f:  w = u - (rho_a v_a - rho_b v_b)/h,  z = w/(1 + theta),  u <- theta z
//   cell j:  v <- v/(1 + dtau/dt) - sum_f dG_f/dH_j z_f + rhs_j/dt        (inner cells; v = 0 elsewhere)
// with rho = dtau/(1 + dtau/dt); its fixed point gives lam = rho v. Face values are stored in grid-sized
// arrays at the index of the face's upper cell b (a = b - stride); a face exists where the reference's
// flux array has one (transverse indices inner). theta and dtau are the plain-form expressions
// (face_theta, nonlin_dtau).

__device__ __forceinline__ double adj_rho(const KP<double>& p, double H) {
    const double dtau = nonlin_dtau<double, PTD_PLAIN>(p, H);
    return dtau / (1.0 + dtau / p.dt);
}

// (dG/dH_a, dG/dH_b) and theta of the face between cells with values Ha (lower) and Hb (upper)
__device__ __forceinline__ void adj_face_coefs(const KP<double>& p, double Ha, double Hb, double ih, double& ga,
                                               double& gb, double& theta) {
    const double av = (Ha + Hb) * 0.5;
    const double c = av * av * av;
    const double k = 1.5 * av * av * (Hb - Ha);
    ga = (k - c) * ih;
    gb = (k + c) * ih;
    theta = face_theta<double, PTD_NONLIN, PTD_PLAIN>(p, c);
}

struct AdjGeo {
    int64_t i, j, k, st[3], n[3];
    int ndim;
    __device__ AdjGeo(const KP<double>& p, int64_t c, int nd) : ndim(nd) {
        split3(c, p.nx, p.ny, i, j, k);
        st[0] = 1;
        st[1] = p.nx;
        st[2] = p.nx * p.ny;
        n[0] = p.nx;
        n[1] = p.ny;
        n[2] = p.nz;
    }
    __device__ int64_t idx(int ax) const { return ax == 0 ? i : ax == 1 ? j : k; }
    __device__ bool inner_ax(int ax) const { return idx(ax) >= 1 && idx(ax) <= n[ax] - 2; }
    __device__ bool inner() const {
        bool in = inner_ax(0);
        if (ndim >= 2) in = in && inner_ax(1);
        if (ndim >= 3) in = in && inner_ax(2);
        return in;
    }
    // the face (c - st | c) along ax exists: transverse indices inner and c not on the lower edge
    __device__ bool lower_face(int ax) const {
        if (idx(ax) < 1) return false;
        for (int o = 0; o < ndim; ++o)
            if (o != ax && !inner_ax(o)) return false;
        return true;
    }
};

template <int NDIM>
__global__ void __launch_bounds__(kBlock)
k_adjnl_face(KP<double> p, const double* __restrict__ H, const double* __restrict__ v, double* __restrict__ u0,
             double* __restrict__ u1, double* __restrict__ u2, double* __restrict__ z0, double* __restrict__ z1,
             double* __restrict__ z2, int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const AdjGeo g(p, c, NDIM);
    double* const u[3] = {u0, u1, u2};
    double* const z[3] = {z0, z1, z2};
    const double ih[3] = {p._dx, p._dy, p._dz};
    const bool cin = g.inner();
    const double rb = cin ? adj_rho(p, H[c]) * v[c] : 0.0;
#pragma unroll
    for (int ax = 0; ax < NDIM; ++ax) {
        if (!g.lower_face(ax)) continue;
        const int64_t a = c - g.st[ax];
        double ga, gb, th;
        adj_face_coefs(p, H[a], H[c], ih[ax], ga, gb, th);
        bool ain = true;
        for (int o = 0; o < NDIM; ++o) ain = ain && (o == ax ? g.idx(o) - 1 >= 1 : g.inner_ax(o));
        const double ra = ain ? adj_rho(p, H[a]) * v[a] : 0.0;
        const double w = u[ax][c] - (ra - rb) * ih[ax];
        const double zz = w / (1.0 + th);
        z[ax][c] = zz;
        u[ax][c] = th * zz;
    }
}

template <int NDIM>
__global__ void __launch_bounds__(kBlock)
k_adjnl_cell(KP<double> p, const double* __restrict__ H, double* __restrict__ v, const double* __restrict__ z0,
             const double* __restrict__ z1, const double* __restrict__ z2, const double* __restrict__ rhs, int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const AdjGeo g(p, c, NDIM);
    if (!g.inner()) return;
    const double* const z[3] = {z0, z1, z2};
    const double ih[3] = {p._dx, p._dy, p._dz};
    const double Hc = H[c];
    double acc = 0.0;
#pragma unroll
    for (int ax = 0; ax < NDIM; ++ax) {
        const int64_t lo = c - g.st[ax], up = c + g.st[ax];
        double ga, gb, th;
        adj_face_coefs(p, H[lo], Hc, ih[ax], ga, gb, th);  // lower face: c is its upper cell b
        acc += gb * z[ax][c];
        adj_face_coefs(p, Hc, H[up], ih[ax], ga, gb, th);  // upper face: c is its lower cell a
        acc += ga * z[ax][up];
    }
    const double dtau = nonlin_dtau<double, PTD_PLAIN>(p, Hc);
    v[c] = v[c] / (1.0 + dtau / p.dt) - acc + rhs[c] * p._dt;
}

// lam = rho v on the inner cells (0 on the boundary); with `resid` also the residual
// rhs/dt - J^T lam, per-block (sum r^2, max |r|).
template <int NDIM>
__global__ void __launch_bounds__(kBlock)
k_adjnl_lambda(KP<double> p, const double* __restrict__ H, const double* __restrict__ v, double* __restrict__ lam,
               const double* __restrict__ rhs, double2* __restrict__ partial, int resid, int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double s = 0.0, m = 0.0;
    if (c < n) {
        const AdjGeo g(p, c, NDIM);
        if (g.inner()) {
            const double lc = adj_rho(p, H[c]) * v[c];
            lam[c] = lc;
            if (resid) {
                const double ih[3] = {p._dx, p._dy, p._dz};
                double jt = lc * p._dt;
#pragma unroll
                for (int ax = 0; ax < NDIM; ++ax) {
                    const int64_t lo = c - g.st[ax], up = c + g.st[ax];
                    const AdjGeo gl(p, lo, NDIM), gu(p, up, NDIM);
                    const double ll = gl.inner() ? adj_rho(p, H[lo]) * v[lo] : 0.0;
                    const double lu = gu.inner() ? adj_rho(p, H[up]) * v[up] : 0.0;
                    double ga, gb, th;
                    adj_face_coefs(p, H[lo], H[c], ih[ax], ga, gb, th);
                    jt += gb * (lc - ll) * ih[ax];  // face (lo | c): dG/dH_b (lam_b - lam_a)/h
                    adj_face_coefs(p, H[c], H[up], ih[ax], ga, gb, th);
                    jt += ga * (lu - lc) * ih[ax];  // face (c | up): dG/dH_a (lam_b - lam_a)/h
                }
                const double r = rhs[c] * p._dt - jt;
                s = r * r;
                m = fabs(r);
            }
        } else {
            lam[c] = 0.0;
        }
    }
    if (resid) block_sum_max(s, m, threadIdx.x, kBlock, &partial[blockIdx.x]);
}

// Boundary cells: gH0 += sum over their faces of dG_f/dH_c (lam_a - lam_b)/h (lam = 0 on the boundary); inner
// cells on the first step: gH0 += lam/dt.
template <int NDIM>
__global__ void __launch_bounds__(kBlock)
k_adjnl_grad(KP<double> p, const double* __restrict__ H, const double* __restrict__ lam, double* __restrict__ gH0,
             int first, int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const AdjGeo g(p, c, NDIM);
    if (g.inner()) {
        if (first) gH0[c] += lam[c] * p._dt;
        return;
    }
    const double ih[3] = {p._dx, p._dy, p._dz};
    double acc = 0.0;
#pragma unroll
    for (int ax = 0; ax < NDIM; ++ax) {
        bool trans = true;
        for (int o = 0; o < NDIM; ++o)
            if (o != ax) trans = trans && g.inner_ax(o);
        if (!trans) continue;
        double ga, gb, th;
        if (g.idx(ax) >= 1) {  // face (c - st | c): c is b
            const int64_t a = c - g.st[ax];
            adj_face_coefs(p, H[a], H[c], ih[ax], ga, gb, th);
            acc += gb * lam[a] * ih[ax];
        }
        if (g.idx(ax) <= g.n[ax] - 2) {  // face (c | c + st): c is a
            const int64_t b = c + g.st[ax];
            adj_face_coefs(p, H[c], H[b], ih[ax], ga, gb, th);
            acc -= ga * lam[b] * ih[ax];
        }
    }
    gH0[c] += acc;
}

#define PTD_ADJ_ND(KERNEL, ...)                                                                           \
    do {                                                                                                  \
        const KP<double> kp = make_kp<double>(c->p);                                                      \
        const dim3 g(grid_for(c->nH)), bl(kBlock);                                                        \
        if (c->p.ndim == 1) hipLaunchKernelGGL((KERNEL<1>), g, bl, 0, c->stream, kp, __VA_ARGS__, c->nH); \
        else if (c->p.ndim == 2) hipLaunchKernelGGL((KERNEL<2>), g, bl, 0, c->stream, kp, __VA_ARGS__, c->nH); \
        else hipLaunchKernelGGL((KERNEL<3>), g, bl, 0, c->stream, kp, __VA_ARGS__, c->nH);                \
    } while (0)

hipError_t launch_adjnl_iteration(ptd_ctx* c, const double* H, double* v, double* const u[3], double* const z[3],
                                  const double* rhs) {
    PTD_ADJ_ND(k_adjnl_face, H, (const double*)v, u[0], u[1], u[2], z[0], z[1], z[2]);
    PTD_ADJ_ND(k_adjnl_cell, H, v, (const double*)z[0], (const double*)z[1], (const double*)z[2], rhs);
    return hipGetLastError();
}

hipError_t launch_adjnl_lambda(ptd_ctx* c, const double* H, const double* v, double* lam, const double* rhs,
                               bool resid) {
    const unsigned g = grid_for(c->nH);
    if ((int64_t)g > c->npartial) return hipErrorInvalidValue;
    PTD_ADJ_ND(k_adjnl_lambda, H, v, lam, rhs, (double2*)c->partial, (int)resid);
    if (resid)
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, (const double2*)c->partial,
                           (int64_t)g, c->dred);
    return hipGetLastError();
}

hipError_t launch_adjnl_grad(ptd_ctx* c, const double* H, const double* lam, double* gH0, bool first) {
    PTD_ADJ_ND(k_adjnl_grad, H, lam, gH0, (int)first);
    return hipGetLastError();
}
#undef PTD_ADJ_ND

hipError_t launch_adj_source(ptd_ctx* c, const double* H, const double* Hobs, double* out) {
    const KP<double> kp = make_kp<double>(c->p);
    const unsigned g = grid_for(c->nH);
    if ((int64_t)g > c->npartial) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_adj_source, dim3(g), dim3(kBlock), 0, c->stream, kp, H, Hobs, out, (double2*)c->partial, c->nH);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, (const double2*)c->partial, (int64_t)g,
                       c->dred);
    return hipGetLastError();
}

hipError_t launch_sumsq_inner(ptd_ctx* c, const double* H) {
    const KP<double> kp = make_kp<double>(c->p);
    const unsigned g = grid_for(c->nH);
    if ((int64_t)g > c->npartial) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sumsq_inner, dim3(g), dim3(kBlock), 0, c->stream, kp, H, (double2*)c->partial, c->nH);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, (const double2*)c->partial, (int64_t)g,
                       c->dred);
    return hipGetLastError();
}

hipError_t launch_grad_accum(ptd_ctx* c, const double* mu, const double* Hn, double* gD, double* gH0, bool first) {
    const KP<double> kp = make_kp<double>(c->p);
    const dim3 g(grid_for(c->nH)), bl(kBlock);
    const double* D = (const double*)c->D;
    const bool step = c->p.dcase == PTD_STEP;
#define PTD_GRAD(ND)                                                                                          \
    if (step) hipLaunchKernelGGL((k_grad_accum<ND, true>), g, bl, 0, c->stream, kp, mu, Hn, D, gD, gH0, (int)first, c->nH); \
    else hipLaunchKernelGGL((k_grad_accum<ND, false>), g, bl, 0, c->stream, kp, mu, Hn, D, gD, gH0, (int)first, c->nH);
    if (c->p.ndim == 1) { PTD_GRAD(1) } else if (c->p.ndim == 2) { PTD_GRAD(2) } else { PTD_GRAD(3) }
#undef PTD_GRAD
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// batches of independent float64 solves
// ---------------------------------------------------------------------------------------------
//
// One launch covers the tiles of every member of a (ndim, dcase, arith) group: block b takes table
// entry (b mod 8)*per + b/8 (XCD-aware), i.e. (member, tile, z chunk depth), and runs the same tile
// function as the single-problem sweep on that member's fields. Buffers come from the member's
// parities + the launch index li in the check block; members that are done exit at once.

// A pointer read from device memory (the member table) is a generic pointer to the compiler, which
// then emits flat accesses with 64-bit vector addresses; re-deriving it in the global address space
// gives global accesses with scalar bases. Every member array is device memory from hipMalloc.
template <typename T>
__device__ __forceinline__ T* as_global(T* ptr) {
    return (T*)(__attribute__((address_space(1))) T*)(uintptr_t)ptr;
}

template <int DC, int ARITH, int R, bool CHECK>
__global__ void __launch_bounds__(64 * R, (t2_waves_per_simd<double, DC, ARITH, R, CHECK>()))
k_batch3d_t2(BatchMember* __restrict__ mem, const BatchTile* __restrict__ tiles, int ntot, int per_xcd, int li) {
    const int b = blockIdx.x;
    const int t = (b & 7) * per_xcd + (b >> 3);
    if (t >= ntot) return;
    const BatchTile e = tiles[t];
    const int mi = __builtin_amdgcn_readfirstlane(e.member);
    const int tile = __builtin_amdgcn_readfirstlane(e.tile);
    const int tz = __builtin_amdgcn_readfirstlane(e.tz);
    const BatchMember& M = mem[mi];
    if (!M.active) return;
    const int ntx = ((int)M.p.nx - 2 + 59) / 60;
    const int nty = ((int)M.p.ny - 2 + R - 5) / (R - 4);
    const int hp = (M.hpar + li) & 1, qp = (M.qpar + li) & 1;
    const int rest = tile / ntx;
    fused3d_t2_tile<double, DC, PTD_PLAIN, ARITH, R, CHECK>(
        M.p, as_global(M.H[hp]), as_global(M.H[hp ^ 1]), as_global(M.Hold), as_global(M.D), as_global(M.q[0][qp]),
        as_global(M.q[1][qp]), as_global(M.q[2][qp]), as_global(M.q[0][qp ^ 1]), as_global(M.q[1][qp ^ 1]),
        as_global(M.q[2][qp ^ 1]), as_global(M.partial + tile), tile % ntx, rest % nty, rest / nty, tz);
}

// three-level passes of the members of a float64 FAST group (no residual)
template <int DC, int ARITH>
__global__ void __launch_bounds__(64 * kRows3)
k_batch3d_t3(BatchMember* __restrict__ mem, const BatchTile* __restrict__ tiles, int ntot, int per_xcd, int li) {
    constexpr int R = kRows3;
    const int b = blockIdx.x;
    const int t = (b & 7) * per_xcd + (b >> 3);
    if (t >= ntot) return;
    const BatchTile e = tiles[t];
    const int mi = __builtin_amdgcn_readfirstlane(e.member);
    const int tile = __builtin_amdgcn_readfirstlane(e.tile);
    const int tz = __builtin_amdgcn_readfirstlane(e.tz);
    const BatchMember& M = mem[mi];
    if (!M.active) return;
    const int ntx = ((int)M.p.nx - 2 + 57) / 58;
    const int nty = ((int)M.p.ny - 2 + R - 7) / (R - 6);
    const int hp = (M.hpar + li) & 1, qp = (M.qpar + li) & 1;
    const int rest = tile / ntx;
    fused3d_t3_tile<double, DC, PTD_PLAIN, ARITH, R>(
        M.p, as_global(M.H[hp]), as_global(M.H[hp ^ 1]), as_global(M.Hold), as_global(M.D), as_global(M.q[0][qp]),
        as_global(M.q[1][qp]), as_global(M.q[2][qp]), as_global(M.q[0][qp ^ 1]), as_global(M.q[1][qp ^ 1]),
        as_global(M.q[2][qp ^ 1]), tile % ntx, rest % nty, rest / nty, tz);
}

template <int DC, int ARITH, int R, bool CHECK>
__global__ void __launch_bounds__(64 * R)
k_batch2d(BatchMember* __restrict__ mem, const BatchTile* __restrict__ tiles, int ntot, int per_xcd, int li) {
    const int b = blockIdx.x;
    const int t = (b & 7) * per_xcd + (b >> 3);
    if (t >= ntot) return;
    const BatchTile e = tiles[t];
    const int mi = __builtin_amdgcn_readfirstlane(e.member);
    const int tile = __builtin_amdgcn_readfirstlane(e.tile);
    const BatchMember& M = mem[mi];
    if (!M.active) return;
    const int ntx = ((int)M.p.nx - 2 + 61) / 62;
    const int hp = (M.hpar + li) & 1, qp = (M.qpar + li) & 1;
    fused2d_tile<double, DC, PTD_PLAIN, ARITH, R, CHECK>(
        M.p, as_global(M.H[hp]), as_global(M.H[hp ^ 1]), as_global(M.Hold), as_global(M.D), as_global(M.q[0][qp]),
        as_global(M.q[1][qp]), as_global(M.q[0][qp ^ 1]), as_global(M.q[1][qp ^ 1]), as_global(M.partial + tile),
        tile % ntx, tile / ntx);
}

template <int DC, int ARITH, bool CHECK>
__global__ void __launch_bounds__(64)
k_batch2d_t2(BatchMember* __restrict__ mem, const BatchTile* __restrict__ tiles, int ntot, int per_xcd, int li) {
    const int b = blockIdx.x;
    const int t = (b & 7) * per_xcd + (b >> 3);
    if (t >= ntot) return;
    const BatchTile e = tiles[t];
    const int mi = __builtin_amdgcn_readfirstlane(e.member);
    const int tile = __builtin_amdgcn_readfirstlane(e.tile);
    const int ty = __builtin_amdgcn_readfirstlane(e.tz);
    const BatchMember& M = mem[mi];
    if (!M.active) return;
    const int ntx = ((int)M.p.nx - 2 + 59) / 60;
    const int hp = (M.hpar + li) & 1, qp = (M.qpar + li) & 1;
    fused2d_t2_tile<double, DC, PTD_PLAIN, ARITH, CHECK>(
        M.p, as_global(M.H[hp]), as_global(M.H[hp ^ 1]), as_global(M.Hold), as_global(M.D), as_global(M.q[0][qp]),
        as_global(M.q[1][qp]), as_global(M.q[0][qp ^ 1]), as_global(M.q[1][qp ^ 1]), as_global(M.partial + tile),
        tile % ntx, tile / ntx, ty);
}

// up to kLdsHalo PT iterations of the members of a 2D group whose grids run LDS-resident in their single
// solves, 32 x 32 cores per block; the checking launch also writes each core's residual partial
template <int DC, int ARITH, bool CHECK>
__global__ void __launch_bounds__(kLdsThreads)
k_batch_lds2d(BatchMember* __restrict__ mem, const BatchTile* __restrict__ tiles, int ntot, int per_xcd, int li,
              int niter) {
    const int b = blockIdx.x;
    const int t = (b & 7) * per_xcd + (b >> 3);
    if (t >= ntot) return;
    const BatchTile e = tiles[t];
    const int mi = __builtin_amdgcn_readfirstlane(e.member);
    const int tile = __builtin_amdgcn_readfirstlane(e.tile);
    const BatchMember& M = mem[mi];
    if (!M.active) return;
    const int ntx = ((int)M.p.nx + kLdsCore - 1) / kLdsCore;
    const int hp = (M.hpar + li) & 1, qp = (M.qpar + li) & 1;
    lds2d_tile<double, DC, PTD_PLAIN, ARITH, CHECK>(
        M.p, as_global(M.H[hp]), as_global(M.H[hp ^ 1]), as_global(M.Hold), as_global(M.D), as_global(M.q[0][qp]),
        as_global(M.q[1][qp]), as_global(M.q[0][qp ^ 1]), as_global(M.q[1][qp ^ 1]), niter, tile % ntx, tile / ntx,
        as_global(M.partial + tile));
}

hipError_t launch_batch_lds2d(hipStream_t s, int dcase, int arith, BatchMember* mem, const BatchTile* tiles,
                              int ntiles, int li, int niter, bool check) {
    if (ntiles <= 0) return hipSuccess;
    if (niter < 1 || niter > kLdsHalo) return hipErrorInvalidValue;
    const int per = (ntiles + 7) / 8;
    const dim3 g(8 * per), bl(kLdsThreads);
    const bool fast = arith == PTD_ARITH_FAST;
#define PTD_BATCH_LDS(DC)                                                                                           \
    if (fast && check) hipLaunchKernelGGL((k_batch_lds2d<DC, kFast, true>), g, bl, 0, s, mem, tiles, ntiles, per, li, niter); \
    else if (fast) hipLaunchKernelGGL((k_batch_lds2d<DC, kFast, false>), g, bl, 0, s, mem, tiles, ntiles, per, li, niter);   \
    else if (check) hipLaunchKernelGGL((k_batch_lds2d<DC, kExact, true>), g, bl, 0, s, mem, tiles, ntiles, per, li, niter);  \
    else hipLaunchKernelGGL((k_batch_lds2d<DC, kExact, false>), g, bl, 0, s, mem, tiles, ntiles, per, li, niter);
    if (dcase == PTD_LIN) {
        PTD_BATCH_LDS(PTD_LIN)
    } else if (dcase == PTD_STEP) {
        PTD_BATCH_LDS(PTD_STEP)
    } else {
        PTD_BATCH_LDS(PTD_NONLIN)
    }
#undef PTD_BATCH_LDS
    return hipGetLastError();
}

// rows of the batched launches: the checking launch uses the single-problem checking tiling (so the
// residual partials, and hence every stopping decision, are those of a single-problem solve); the
// other launches may use any tiling (every cell is computed by its owner with the same expressions).
int batch_rows_check(int kind, int dcase, int arith) {
    if (kind == kSweepLds2D) return kLdsCore;
    if (kind == kSweep2D2) return kTy2D;
    if (kind != kSweep3D) return kRows2D;
    const bool fast = arith == PTD_ARITH_FAST;
    if (dcase == PTD_STEP) return fast ? t2_rows<double, PTD_STEP, kFast, true>() : t2_rows<double, PTD_STEP, kExact, true>();
    if (dcase == PTD_LIN) return fast ? t2_rows<double, PTD_LIN, kFast, true>() : t2_rows<double, PTD_LIN, kExact, true>();
    return fast ? t2_rows<double, PTD_NONLIN, kFast, true>() : t2_rows<double, PTD_NONLIN, kExact, true>();
}
int batch_rows_nocheck(int kind, int dcase, int arith, int variant) {
    if (kind == kSweepLds2D) return kLdsCore;
    if (kind == kSweep2D2) return kTy2D;
    if (kind == kSweep2D) return variant == 1 ? 8 : 16;
    const bool fast = arith == PTD_ARITH_FAST;
    if (dcase == PTD_STEP) return fast ? t2_rows<double, PTD_STEP, kFast, false>() : t2_rows<double, PTD_STEP, kExact, false>();
    // linear FAST: the 12-row tiling (the batched kernel's extra registers would spill at 16 rows)
    if (dcase == PTD_LIN) return fast ? 12 : t2_rows<double, PTD_LIN, kExact, false>();
    return fast ? t2_rows<double, PTD_NONLIN, kFast, false>() : t2_rows<double, PTD_NONLIN, kExact, false>();
}

template <int DC, int ARITH>
static hipError_t batch_sweep_t(hipStream_t s, int kind, int rows, bool check, BatchMember* mem,
                                const BatchTile* tiles, int ntiles, int li) {
    const int per = (ntiles + 7) / 8;
    const dim3 g(8 * per);
    if (kind == kSweep2D2) {
        if (rows != kTy2D) return hipErrorInvalidValue;
        if (check)
            hipLaunchKernelGGL((k_batch2d_t2<DC, ARITH, true>), g, dim3(64), 0, s, mem, tiles, ntiles, per, li);
        else
            hipLaunchKernelGGL((k_batch2d_t2<DC, ARITH, false>), g, dim3(64), 0, s, mem, tiles, ntiles, per, li);
    } else if (kind == kSweep3D) {
        constexpr int RN = t2_lin16<double, DC, ARITH, false>() ? 12 : t2_rows<double, DC, ARITH, false>();
        constexpr int RC = t2_rows<double, DC, ARITH, true>();
        if (check) {
            if (rows != RC) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_batch3d_t2<DC, ARITH, RC, true>), g, dim3(64, RC), 0, s, mem, tiles, ntiles, per, li);
        } else {
            if (rows != RN) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_batch3d_t2<DC, ARITH, RN, false>), g, dim3(64, RN), 0, s, mem, tiles, ntiles, per, li);
        }
    } else {
        if (check) {
            if (rows != kRows2D) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_batch2d<DC, ARITH, kRows2D, true>), g, dim3(64, kRows2D), 0, s, mem, tiles, ntiles, per, li);
        } else if (rows == 8) {
            hipLaunchKernelGGL((k_batch2d<DC, ARITH, 8, false>), g, dim3(64, 8), 0, s, mem, tiles, ntiles, per, li);
        } else if (rows == 16) {
            hipLaunchKernelGGL((k_batch2d<DC, ARITH, 16, false>), g, dim3(64, 16), 0, s, mem, tiles, ntiles, per, li);
        } else {
            return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

hipError_t launch_batch_sweep(hipStream_t s, int kind, int dcase, int arith, int rows, bool check,
                              BatchMember* mem, const BatchTile* tiles, int ntiles, int li) {
    if (ntiles <= 0) return hipSuccess;
    const bool fast = arith == PTD_ARITH_FAST;
    switch (dcase) {
        case PTD_LIN:
            return fast ? batch_sweep_t<PTD_LIN, kFast>(s, kind, rows, check, mem, tiles, ntiles, li)
                        : batch_sweep_t<PTD_LIN, kExact>(s, kind, rows, check, mem, tiles, ntiles, li);
        case PTD_STEP:
            return fast ? batch_sweep_t<PTD_STEP, kFast>(s, kind, rows, check, mem, tiles, ntiles, li)
                        : batch_sweep_t<PTD_STEP, kExact>(s, kind, rows, check, mem, tiles, ntiles, li);
        default:
            return fast ? batch_sweep_t<PTD_NONLIN, kFast>(s, kind, rows, check, mem, tiles, ntiles, li)
                        : batch_sweep_t<PTD_NONLIN, kExact>(s, kind, rows, check, mem, tiles, ntiles, li);
    }
}

// Three-level passes of a batched group: the float64 FAST 3D groups. Linear and step members take them as
// their single solves do; nonlinear members take them too, as a batch's non-checking launches may use any
// pass (the checking one keeps the single-problem tiling): at the batch's grid sizes the nonlinear
// three-level pass is the faster one up to ~224^3 (ms per iteration, two- / three-level: 128^3 0.0356 /
// 0.0312, 224^3 0.133 / 0.121; 256^3 0.175 / 0.200), and the 96-member wall went 4.234 -> 4.212 s
// (medians, same box). PTD_T3=0 switches them off.
bool batch_t3(const ptd_params& p) {
    if (p.ndim != 3 || p.arith != PTD_ARITH_FAST) return false;
    if (p.dcase == PTD_LIN || p.dcase == PTD_STEP) return fused3_wanted(p, 8);
    const char* env = getenv("PTD_T3");
    return !(env && env[0] == '0') && p.nx >= 3 && p.ny >= 3 && p.nz >= 3 && max_chunk_tz(p, 8) >= 2;
}
int fused3_rows() { return kRows3; }

hipError_t launch_batch_sweep3(hipStream_t s, int dcase, int arith, BatchMember* mem, const BatchTile* tiles,
                               int ntiles, int li) {
    if (ntiles <= 0) return hipSuccess;
    if (arith != PTD_ARITH_FAST) return hipErrorInvalidValue;
    const int per = (ntiles + 7) / 8;
    const dim3 g(8 * per), bl(64, kRows3);
    if (dcase == PTD_LIN)
        hipLaunchKernelGGL((k_batch3d_t3<PTD_LIN, kFast>), g, bl, 0, s, mem, tiles, ntiles, per, li);
    else if (dcase == PTD_STEP)
        hipLaunchKernelGGL((k_batch3d_t3<PTD_STEP, kFast>), g, bl, 0, s, mem, tiles, ntiles, per, li);
    else
        hipLaunchKernelGGL((k_batch3d_t3<PTD_NONLIN, kFast>), g, bl, 0, s, mem, tiles, ntiles, per, li);
    return hipGetLastError();
}

// per member of a group: (sum, max) of its checking launch's tile partials, in the fixed order of the
// single-problem reduction
__global__ void __launch_bounds__(kBlock)
k_batch_reduce(BatchMember* __restrict__ mem, const int32_t* __restrict__ members) {
    BatchMember& M = mem[members[blockIdx.x]];
    if (!M.active) return;
    const double2 res = reduce_partials_block(M.partial, M.ntiles_check);
    if (threadIdx.x == 0) M.red = res;
}

hipError_t launch_batch_reduce(hipStream_t s, BatchMember* mem, const int32_t* members, int n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_reduce, dim3(n), dim3(kBlock), 0, s, mem, members);
    return hipGetLastError();
}

// The PT loop of one member after one check block of nout iterations
// (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:104-118): iter += nout; err from the check; the step
// ends when !(err > tol) or iter >= itmax (then Hold .= H, next step); the member is done after nt
// steps or on NaN. Returns whether the member is still running.
__device__ __forceinline__ int control_member(BatchMember& M, int t, int64_t* __restrict__ iters,
                                              double* __restrict__ errs, int nt_cap) {
    M.assign = 0;
    if (M.active) {
        M.iter += M.nout;
        M.hpar = (M.hpar + M.lpb) & 1;
        M.qpar = (M.qpar + M.lpb) & 1;
        const double rms = sqrt(M.red.x) / sqrt(M.norm_len);
        const double err = M.stop_norm == PTD_NORM_MAX ? M.red.y : rms;
        if (!(err > M.tol) || M.iter >= M.itmax) {
            if (M.step < nt_cap) {
                iters[(int64_t)t * nt_cap + M.step] = M.iter;
                errs[(int64_t)t * nt_cap + M.step] = err;
            }
            M.step += 1;
            M.iter = 0;
            if (isnan(err)) {
                M.status = PTD_ERR_NAN;
                M.active = 0;
                M.t_done = (uint64_t)wall_clock64();
            } else {
                M.assign = 1;
                if (M.step >= M.nt) {
                    M.active = 0;
                    M.t_done = (uint64_t)wall_clock64();
                }
            }
        }
    }
    return M.active;
}

// The members of one group, one thread each (a check block of that group); with `active_count` the number
// of its members still running goes there.
__global__ void __launch_bounds__(kBatchMaxMembers)
k_batch_control_list(BatchMember* __restrict__ mem, const int32_t* __restrict__ members, int n,
                     int64_t* __restrict__ iters, double* __restrict__ errs, int32_t* __restrict__ active_count,
                     int nt_cap) {
    const int t = threadIdx.x;
    const int running = t < n ? control_member(mem[members[t]], members[t], iters, errs, nt_cap) : 0;
    const int cnt = __syncthreads_count(running);
    if (t == 0 && active_count) *active_count = cnt;
}

hipError_t launch_batch_control_list(hipStream_t s, BatchMember* mem, const int32_t* members, int n, int64_t* iters,
                                     double* errs, int32_t* active_count, int nt_cap) {
    if (n <= 0 || n > kBatchMaxMembers) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_batch_control_list, dim3(1), dim3(kBatchMaxMembers), 0, s, mem, members, n, iters, errs,
                       active_count, nt_cap);
    return hipGetLastError();
}
// Hold .= H for the members whose step just ended (kBatchAssignBlocks blocks per member; `members`
// lists them, or null for members 0..n-1)
constexpr int kBatchAssignBlocks = 64;
__global__ void __launch_bounds__(kBlock)
k_batch_assign(BatchMember* __restrict__ mem, const int32_t* __restrict__ members) {
    const int slot = blockIdx.x / kBatchAssignBlocks;
    const BatchMember& M = mem[members ? members[slot] : slot];
    if (!M.assign) return;
    const double* __restrict__ H = M.H[M.hpar];
    double* __restrict__ Ho = M.Hold;
    const int64_t stride = (int64_t)kBatchAssignBlocks * kBlock;
    for (int64_t i = (int64_t)(blockIdx.x % kBatchAssignBlocks) * kBlock + threadIdx.x; i < M.nH; i += stride)
        Ho[i] = H[i];
}

hipError_t launch_batch_assign_list(hipStream_t s, BatchMember* mem, const int32_t* members, int n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_assign, dim3(n * kBatchAssignBlocks), dim3(kBlock), 0, s, mem, members);
    return hipGetLastError();
}

__global__ void k_batch_clock(uint64_t* t) {
    if (threadIdx.x == 0) *t = (uint64_t)wall_clock64();
}

hipError_t launch_batch_clock(hipStream_t s, uint64_t* t) {
    hipLaunchKernelGGL(k_batch_clock, dim3(1), dim3(64), 0, s, t);
    return hipGetLastError();
}

// span[0] = now (start), or span[1] += now - span[0] (stop): the device time between two points of a
// stream, summed over graph replays (kernel nodes only: nothing to clear inside the graph)
__global__ void k_batch_span(uint64_t* span, int stop) {
    if (threadIdx.x == 0) {
        const uint64_t now = (uint64_t)wall_clock64();
        if (stop) span[1] += now - span[0];
        else span[0] = now;
    }
}

hipError_t launch_batch_span(hipStream_t s, uint64_t* span, bool stop) {
    hipLaunchKernelGGL(k_batch_span, dim3(1), dim3(64), 0, s, span, (int)stop);
    return hipGetLastError();
}

}  // namespace ptd
