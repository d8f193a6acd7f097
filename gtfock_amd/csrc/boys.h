// boys.h -- Boys function F_m(T) = int_0^1 t^(2m) exp(-T t^2) dt on the GPU.
//
// T < BOYS_TMAX: Taylor expansion of the highest order around the nearest
//   grid point, F_M(T) = sum_k F_{M+k}(T0) (T0 - T)^k / k!
//   (dF_m/dT = -F_{m+1}), then the stable downward recursion
//   F_m = (2T F_{m+1} + exp(-T)) / (2m + 1).
// T >= BOYS_TMAX: F_0 = sqrt(pi/T)/2 (erfc(6) ~ 2e-17) and the upward
//   recursion F_m = ((2m - 1) F_{m-1} - exp(-T)) / (2T).
// The grid values are made on the host from the convergent series
//   F_m(T) = exp(-T) sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1))
// in long double (gtf_boys_table in basis.cpp).
#pragma once
#include "gtf_internal.h"

// F_0..F_M(T) for the ERI kernels from the compact Taylor table: slab holds,
// for grid points T0 = i * BOYS8_DELTA (i = 0 .. BOYS8_NGRID - 1), the
// BOYS8_NTERM = 8 coefficients c_k = F_{M+k}(T0) / k! of order M, one 64-byte
// row per point, so that
//   F_M(T) = sum_k c_k (T0 - T)^k,  |T0 - T| <= BOYS8_DELTA / 2
// (truncation below 1e-15 relative).  FP64 in every kernel, the FP32 tiles
// included.  A class touches one slab (23 KB), small enough to stay in the
// vector L1 cache or in LDS.
template <int M>
__device__ __forceinline__ void boys_eval8(double T, double* F, const double* __restrict__ slab)
{
    // exp(-T) serves both regimes; computed once before the branch, so a
    // wavefront with lanes on both sides does not run it twice
    double e = 0.0;
    if constexpr (M > 0) e = exp(-T);
    if (T < BOYS_TMAX) {
        const int i = (int)(T * BOYS8_INV_DELTA + 0.5);
        const double dt = i * BOYS8_DELTA - T;
        const double2* r = reinterpret_cast<const double2*>(slab + i * BOYS8_NTERM);
        const double2 c01 = r[0], c23 = r[1], c45 = r[2], c67 = r[3];
        double f = fma(c67.y, dt, c67.x);
        f = fma(f, dt, c45.y);
        f = fma(f, dt, c45.x);
        f = fma(f, dt, c23.y);
        f = fma(f, dt, c23.x);
        f = fma(f, dt, c01.y);
        f = fma(f, dt, c01.x);
        F[M] = f;
        if constexpr (M > 0) {
            const double t2 = T + T;
            static_for<0, M>([&](auto k) {
                constexpr int m = M - 1 - decltype(k)::value;
                F[m] = (t2 * F[m + 1] + e) * (1.0 / (2 * m + 1));
            });
        }
    } else {
        const double rt = rsqrt_nr(T);
        F[0] = 0.88622692545275801365 * rt;  // sqrt(pi)/2
        if constexpr (M > 0) {
            const double h = 0.5 * rt * rt;
            static_for<1, M + 1>([&](auto k) {
                constexpr int m = decltype(k)::value;
                F[m] = ((2 * m - 1) * F[m - 1] - e) * h;
            });
        }
    }
}

// runtime-order variant for the one-electron kernels (m <= mmax <= 8)
__device__ inline void boys_eval_rt(int mmax, double T, double* F, const double* __restrict__ tab)
{
    if (T < BOYS_TMAX) {
        int i = (int)(T * BOYS_INV_DELTA + 0.5);
        double dt = i * BOYS_DELTA - T;
        const double* row = tab + i * BOYS_NCOL + mmax;
        double f = row[6];
        f = row[5] + dt * (1.0 / 6.0) * f;
        f = row[4] + dt * (1.0 / 5.0) * f;
        f = row[3] + dt * (1.0 / 4.0) * f;
        f = row[2] + dt * (1.0 / 3.0) * f;
        f = row[1] + dt * (1.0 / 2.0) * f;
        f = row[0] + dt * f;
        F[mmax] = f;
        double e = exp(-T);
        for (int m = mmax - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
    } else {
        F[0] = 0.88622692545275801365 / sqrt(T);
        double e = exp(-T);
        for (int m = 1; m <= mmax; ++m) F[m] = ((2 * m - 1) * F[m - 1] - e) / (2.0 * T);
    }
}
