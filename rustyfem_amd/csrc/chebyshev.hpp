// Coefficients of the Chebyshev smoother of the AMG V-cycle: the polynomial
// in D^-1 A that minimizes the largest error factor on [hi / ratio, hi], run
// as the three-term recurrence
//   d_0 = c0 D^-1 b,                 x_1 = d_0          (from x = 0)
//   d_k = c1_k d_{k-1} + c2_k D^-1 r_k,  x_{k+1} = x_k + d_k,  r_k = b - A x_k,
// whose error after `degree` steps is T_deg((theta - lambda) / delta) /
// T_deg(theta / delta) for each eigenvalue lambda of D^-1 A (theta, delta:
// centre and half width of the interval; Saad, Iterative Methods for Sparse
// Linear Systems, 2nd ed., Algorithm 12.1).
#pragma once
#include <utility>
#include <vector>

namespace rf {

template <typename T>
void chebyshev(double hi, double ratio, int degree, T& c0, std::vector<std::pair<T, T>>& steps) {
  const double lo = hi / ratio;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
  const double sigma = theta / delta;
  double rho_old = 1.0 / sigma;
  c0 = static_cast<T>(1.0 / theta);
  steps.clear();
  for (int k = 1; k < degree; ++k) {
    const double rho_new = 1.0 / (2.0 * sigma - rho_old);
    steps.push_back({static_cast<T>(rho_new * rho_old), static_cast<T>(2.0 * rho_new / delta)});
    rho_old = rho_new;
  }
}

}  // namespace rf
