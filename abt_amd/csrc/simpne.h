// simpne.h -- quadrature on uneven nodes (host + device).
//
// simpne_eval restates simpne (misc_submod.f90:2392-2531): 3-point Lagrange
// panels over x[m-1], x[m], x[m+1] for m = 1, 3, 5, ..., plus, for an even node
// count, the integral of the last three points' parabola over the final
// interval.  The result is linear in y, so simpne_weights gives the weights w
// with simpne(y) = sum_i w_i y_i (used to fold the phi and theta integrals of
// the orientation-averaged far field into per-bin weights).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace abt {

__host__ __device__ inline double simpne_eval(int n, const double* __restrict__ x,
                                              const double* __restrict__ y, int64_t stride)
{
#pragma clang fp contract(off)
    double result = 0.0;
    if (n <= 2) return __builtin_nan("");
    int m = 1;
    for (;;) {
        const double x1 = x[m - 1], x2 = x[m], x3 = x[m + 1];
        const double e = x3 * x3 - x1 * x1;
        const double f = x3 * x3 * x3 - x1 * x1 * x1;
        const double fe = x3 - x1;
        const double del[3] = {x3 - x2, x1 - x3, x2 - x1};
        const double g[3] = {x2 + x3, x1 + x3, x1 + x2};
        const double p[3] = {x2 * x3, x1 * x3, x1 * x2};
        double sum1 = 0.0;
        for (int i = 0; i < 3; i++)
            sum1 = sum1 + y[(int64_t)(m - 1 + i) * stride] * del[i] * (f / 3.0 - g[i] * 0.5 * e + p[i] * fe);
        result = result - sum1 / (del[0] * del[1] * del[2]);
        m = m + 2;
        if (n <= m + 1) break;
    }
    if (n % 2 != 0) return result;
    m = n - 2;
    const double x3 = x[n - 1], x2 = x[n - 2], x1 = x[n - 3];
    const double e = x3 * x3 - x2 * x2;
    const double f = x3 * x3 * x3 - x2 * x2 * x2;
    const double fe = x3 - x2;
    const double del[3] = {x3 - x2, x1 - x3, x2 - x1};
    const double g[3] = {x2 + x3, x1 + x3, x1 + x2};
    const double p[3] = {x2 * x3, x1 * x3, x1 * x2};
    double sum1 = 0.0;
    for (int i = 0; i < 3; i++)
        sum1 = sum1 + y[(int64_t)(m - 1 + i) * stride] * del[i] * (f / 3.0 - g[i] * 0.5 * e + p[i] * fe);
    return result - sum1 / (del[0] * del[1] * del[2]);
}

// w[i] = simpne_eval(n, x, e_i): the quadrature weights of the nodes x (n >= 3).
inline void simpne_weights(int n, const double* x, double* w)
{
    double* y = new double[n];
    for (int i = 0; i < n; i++) y[i] = 0.0;
    for (int i = 0; i < n; i++) {
        y[i] = 1.0;
        w[i] = simpne_eval(n, x, y, 1);
        y[i] = 0.0;
    }
    delete[] y;
}

}  // namespace abt
