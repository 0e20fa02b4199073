// Boys function F_m(x), m = 0..L, on the device.
//
// Same scheme as simint's boys_F_split (skel/simint/boys/boys.h:14-92):
//   x < 36.5 : 7th-order Taylor expansion of F_L around the nearest grid
//              point x_i = 0.1 i of a 366 x 32 table of F_m(x_i)
//              (boys_taylor.h:14-40), then downward recursion
//              F_m = (2x F_{m+1} + e^-x) / (2m+1) for the lower orders
//              (boys.h:69-80);
//   x >= 36.5: F_m = longfac[m] x^-(m+1/2), longfac[m] = sqrt(pi)(2m-1)!!/2^(m+1)
//              (boys_long.h:14-27) for every order when L < 4; for L >= 4
//              only F_L, then the same downward recursion (boys.h:49-80).
//              The asymptotic form drops Gamma(m+1/2, x), so near the cutoff
//              the reference's F_m carry a relative error that grows with m
//              (about 4e-12 for m = 4 at x = 36.5); the device follows the
//              reference's branch structure so both give the same values.
// The reference evaluates every order by Taylor when L < 4; here the
// recursion is used for every L >= 1, which trades 8(L) table loads for one
// exp(), and agrees to a few ulp.  The table is built by the package's own
// generator from mpmath (emitter/boys_table.py).
#pragma once
#include <hip/hip_runtime.h>
#include "gen/boys_longfac.inc"

#define SIMINT_BOYS_NPT 366
#define SIMINT_BOYS_NCOL 32
#define SIMINT_BOYS_MAXX 36.5

template <int L>
__device__ __forceinline__ void boys_eval(const double *__restrict__ grid, double x, double *F)
{
    // sqrt(pi)/2 * (2m-1)!! / 2^m, m = 0..31 (generated, correctly rounded)
    constexpr double longfac[] = SIMINT_BOYS_LONGFAC;
    if (x < SIMINT_BOYS_MAXX) {
        const int idx = (int)(10.0 * (x + 0.05));
        const double dx = (double)idx * 0.1 - x;
        const double *__restrict__ g = grid + idx * SIMINT_BOYS_NCOL + L;
        const double g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];
        const double g4 = g[4], g5 = g[5], g6 = g[6], g7 = g[7];
        F[L] = g0 + dx * (g1 + dx * ((1.0 / 2.0) * g2 + dx * ((1.0 / 6.0) * g3
               + dx * ((1.0 / 24.0) * g4 + dx * ((1.0 / 120.0) * g5
               + dx * ((1.0 / 720.0) * g6 + dx * ((1.0 / 5040.0) * g7)))))));
        if (L > 0) {
            const double x2 = 2.0 * x;
            const double ex = exp(-x);
#pragma unroll
            for (int n = L - 1; n >= 0; n--)
                F[n] = (1.0 / (2 * n + 1)) * fma(x2, F[n + 1], ex);
        }
    } else if (L < 4) {
        // every order from the long-range form (boys.h:18-25, boys_long.h:14-27)
        const double x1 = 1.0 / x;
        double x2 = sqrt(x1);
#pragma unroll
        for (int n = 0; n <= L; n++) {
            F[n] = longfac[n] * x2;
            x2 *= x1;
        }
    } else {
        // top order from the long-range form, then down (boys.h:49-80,
        // boys_long.h:29-35): this keeps the reference's values, whose
        // e^-x tail enters through the recursion.  pow, not a product of
        // powers of x^-1/2: the latter carries ~(2L+1) ulp, and high-AM
        // classes amplify input errors far above the reference's own.
        F[L] = longfac[L] * sqrt(pow(x, -(double)(2 * L + 1)));
        const double x2 = 2.0 * x;
        const double ex = exp(-x);
#pragma unroll
        for (int n = L - 1; n >= 0; n--)
            F[n] = (1.0 / (2 * n + 1)) * fma(x2, F[n + 1], ex);
    }
}
