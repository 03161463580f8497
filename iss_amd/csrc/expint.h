// Exponential integral E_n(x) = int_1^inf exp(-x t) t^-n dt for integer n >= 1
// and x > 0, in FP64, for the host and the device.
//   x >= 1: the continued fraction (Abramowitz & Stegun 5.1.22)
//           E_n(x) = exp(-x) / (x + n - 1 n / (x + n + 2 - 2 (n + 1) / (x + n + 4 - ...)))
//           evaluated with the modified Lentz method;
//   x < 1:  the power series (A&S 5.1.12)
//           E_n(x) = (-x)^(n-1) / (n-1)! (psi(n) - ln x) - sum_{k != n-1} (-x)^k / ((k - n + 1) k!),
//           psi(n) = -gamma + sum_{m < n} 1 / m.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace issdev {

__host__ __device__ inline double expint_en(int n, double x) {
    const double eps = 1e-16, tiny = 1e-300;
    if (x >= 1.0) {
        double b = x + n, c = 1.0 / tiny, d = 1.0 / b, h = d;
        for (int i = 1; i < 1000; i++) {
            const double a = -static_cast<double>(i) * (n - 1 + i);
            b += 2.0;
            d = 1.0 / (a * d + b);
            c = b + a / c;
            const double del = c * d;
            h *= del;
            if (fabs(del - 1.0) < eps) break;
        }
        return h * exp(-x);
    }
    const int nm1 = n - 1;
    double ans = (nm1 != 0) ? 1.0 / nm1 : -log(x) - 0.5772156649015329;
    double fact = 1.0;
    for (int i = 1; i < 1000; i++) {
        fact *= -x / i;
        double del;
        if (i != nm1) {
            del = -fact / (i - nm1);
        } else {
            double psi = -0.5772156649015329;
            for (int m = 1; m <= nm1; m++) psi += 1.0 / m;
            del = fact * (-log(x) + psi);
        }
        ans += del;
        if (fabs(del) < fabs(ans) * eps) break;
    }
    return ans;
}

}  // namespace issdev
