// Exponential integral E_n(x) = int_1^inf exp(-x t) t^-n dt for integer n >= 1
// and x > 0, in FP64, for the host and the device.
//
//   x < 1:  E_1 from its power series (Abramowitz & Stegun 5.1.11),
//             E_1(x) = -gamma - ln x + sum_{k>=1} (-1)^(k+1) x^k / (k k!),
//           then E_2 ... E_n by the upward recurrence (A&S 5.1.14)
//             E_{m+1}(x) = (exp(-x) - x E_m(x)) / m,
//           which shrinks an error of E_m by the factor x / m < 1 per step.
//   x >= 1: the continued fraction of A&S 5.1.22 in its even (contracted)
//           form, exp(x) E_n(x) = 1 / (b_0 + a_1 / (b_1 + a_2 / (b_2 + ...)))
//           with b_k = x + n + 2k and a_k = -k (n + k - 1), summed from the
//           bottom up at a depth that doubles until two depths agree.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace issdev {

// exp(x) E_n(x) from the even continued fraction truncated after `depth`
// levels (the tail below the last level is dropped)
__host__ __device__ inline double expint_cf_tail(int n, double x, int depth) {
    double t = x + n + 2.0 * depth;
    for (int k = depth; k >= 1; k--) {
        const double ak = -static_cast<double>(k) * static_cast<double>(n + k - 1);
        t = (x + n + 2.0 * (k - 1)) + ak / t;
    }
    return 1.0 / t;
}

__host__ __device__ inline double expint_en(int n, double x) {
    if (x >= 1.0) {
        double prev = expint_cf_tail(n, x, 8);
        for (int depth = 16; depth <= 8192; depth *= 2) {
            const double cur = expint_cf_tail(n, x, depth);
            const bool agree = fabs(cur - prev) <= 1e-16 * fabs(cur);
            prev = cur;
            if (agree) break;
        }
        return prev * exp(-x);
    }
    // E_1 by its series: term_k = (-1)^(k+1) x^k / k!, added as term_k / k
    const double euler_gamma = 0.57721566490153286061;
    double e1 = -euler_gamma - log(x);
    double term = 1.0;
    for (int k = 1; k < 64; k++) {
        term *= -x / k;
        const double add = -term / k;
        e1 += add;
        if (fabs(add) < 1e-17 * fabs(e1)) break;
    }
    const double emx = exp(-x);
    double em = e1;
    for (int m = 1; m < n; m++) em = (emx - x * em) / m;
    return em;
}

}  // namespace issdev
