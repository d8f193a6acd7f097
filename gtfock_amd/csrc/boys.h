// boys.h -- Boys function F_m(T) = int_0^1 t^(2m) exp(-T t^2) dt on the GPU.
//
// T < BOYS_TMAX: 7-term Taylor expansion of the highest order around the
//   nearest grid point, F_M(T) = sum_k F_{M+k}(T0) (T0 - T)^k / k!
//   (dF_m/dT = -F_{m+1}), then the stable downward recursion
//   F_m = (2T F_{m+1} + exp(-T)) / (2m + 1).
// T >= BOYS_TMAX: F_0 = sqrt(pi/T)/2 (erfc(6) ~ 2e-17) and the upward
//   recursion F_m = ((2m - 1) F_{m-1} - exp(-T)) / (2T).
// The grid values are made on the host from the convergent series
//   F_m(T) = exp(-T) sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1))
// in long double (gtf_boys_table in basis.cpp).
#pragma once
#include "gtf_internal.h"

// tab: the grid with row stride STRIDE; STRIDE == BOYS_NCOL is the full
// table (columns m = 0..15), STRIDE == BOYS_NTAYLOR the 7 columns M..M+6
// staged in LDS for one class.
template <int M, typename Real, int STRIDE = BOYS_NCOL>
__device__ __forceinline__ void boys_eval(Real T, Real* F, const double* __restrict__ tab)
{
    if (T < Real(BOYS_TMAX)) {
        int i = (int)(T * Real(BOYS_INV_DELTA) + Real(0.5));
        Real dt = Real(i) * Real(BOYS_DELTA) - T;
        const double* row = tab + i * STRIDE + (STRIDE == BOYS_NCOL ? M : 0);
        Real f = Real(row[6]);
        f = Real(row[5]) + dt * (Real(1.0 / 6.0)) * f;
        f = Real(row[4]) + dt * (Real(1.0 / 5.0)) * f;
        f = Real(row[3]) + dt * (Real(1.0 / 4.0)) * f;
        f = Real(row[2]) + dt * (Real(1.0 / 3.0)) * f;
        f = Real(row[1]) + dt * (Real(1.0 / 2.0)) * f;
        f = Real(row[0]) + dt * f;
        F[M] = f;
        if constexpr (M > 0) {
            Real e = exp(-T);
            Real t2 = T + T;
            static_for<0, M>([&](auto k) {
                constexpr int m = M - 1 - decltype(k)::value;
                F[m] = (t2 * F[m + 1] + e) * Real(1.0 / (2 * m + 1));
            });
        }
    } else {
        Real rt = rsqrt_nr(T);
        F[0] = Real(0.88622692545275801365) * rt;  // sqrt(pi)/2
        if constexpr (M > 0) {
            Real e = exp(-T);
            Real h = Real(0.5) * rt * rt;
            static_for<1, M + 1>([&](auto k) {
                constexpr int m = decltype(k)::value;
                F[m] = (Real(2 * m - 1) * F[m - 1] - e) * h;
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
