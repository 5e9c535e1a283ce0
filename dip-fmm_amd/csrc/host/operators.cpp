#include "operators.hpp"

#include <cmath>
#include <cstdlib>
#include <numbers>

namespace cdfmm_hip {

double monomial_over_factorial(const std::array<double, 3>& r, const MultiIndex& g) {
  return std::pow(r[0], g[0]) * std::pow(r[1], g[1]) * std::pow(r[2], g[2]) /
         (factorial(g[0]) * factorial(g[1]) * factorial(g[2]));
}

std::vector<double> laplace_derivatives(int order, const std::array<double, 3>& r) {
  const Basis basis(order);
  const int n = basis.size();
  const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  std::vector<double> a(static_cast<std::size_t>(n), 0.0);
  a[0] = 1.0 / std::sqrt(r2);
  for (int i = 1; i < n; ++i) {
    const MultiIndex k = basis[i];
    const int kk = degree(k);
    double first = 0.0, second = 0.0;
    for (int axis = 0; axis < 3; ++axis) {
      if (k[axis] >= 1) {
        MultiIndex m = k;
        m[axis] -= 1;
        first += r[axis] * a[static_cast<std::size_t>(basis.find(m[0], m[1], m[2]))];
      }
      if (k[axis] >= 2) {
        MultiIndex m = k;
        m[axis] -= 2;
        second += a[static_cast<std::size_t>(basis.find(m[0], m[1], m[2]))];
      }
    }
    a[static_cast<std::size_t>(i)] = -((2.0 * kk - 1.0) * first + (kk - 1.0) * second) / (kk * r2);
  }
  const double c = 1.0 / (4.0 * std::numbers::pi);
  for (int i = 0; i < n; ++i) {
    const MultiIndex k = basis[i];
    a[static_cast<std::size_t>(i)] *= c * factorial(k[0]) * factorial(k[1]) * factorial(k[2]);
  }
  return a;
}

std::vector<double> m2l_class_matrix(const Basis& basis, const std::array<int, 3>& offset) {
  const int n = basis.size();
  const Basis deriv_basis(2 * basis.order());
  const auto D = laplace_derivatives(2 * basis.order(), {double(offset[0]), double(offset[1]), double(offset[2])});
  std::vector<double> m(static_cast<std::size_t>(n) * n);
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      const int g = deriv_basis.find(basis[a][0] + basis[b][0], basis[a][1] + basis[b][1], basis[a][2] + basis[b][2]);
      m[static_cast<std::size_t>(b) + static_cast<std::size_t>(n) * a] = D[static_cast<std::size_t>(g)];
    }
  return m;
}

std::vector<std::array<int, 3>> list2_offsets() {
  std::vector<std::array<int, 3>> out;
  for (int x = -3; x <= 3; ++x)
    for (int y = -3; y <= 3; ++y)
      for (int z = -3; z <= 3; ++z)
        if (std::abs(x) >= 2 || std::abs(y) >= 2 || std::abs(z) >= 2) out.push_back({x, y, z});
  return out;
}

std::vector<double> m2m_dense(const Basis& basis, const std::array<double, 3>& d) {
  const int n = basis.size();
  std::vector<double> m(static_cast<std::size_t>(n) * n, 0.0);
  for (int a = 0; a < n; ++a)
    for (int g = 0; g < n; ++g) {
      const MultiIndex A = basis[a], G = basis[g];
      if (G[0] > A[0] || G[1] > A[1] || G[2] > A[2]) continue;
      const int in = basis.find(A[0] - G[0], A[1] - G[1], A[2] - G[2]);
      m[static_cast<std::size_t>(a) * n + in] += monomial_over_factorial(d, G);
    }
  return m;
}

std::vector<double> l2l_dense(const Basis& basis, const std::array<double, 3>& d) {
  const int n = basis.size();
  std::vector<double> m(static_cast<std::size_t>(n) * n, 0.0);
  for (int b = 0; b < n; ++b)
    for (int g = 0; g < n; ++g) {
      const MultiIndex B = basis[b], G = basis[g];
      const int in = basis.find(B[0] + G[0], B[1] + G[1], B[2] + G[2]);
      if (in < 0) continue;
      m[static_cast<std::size_t>(b) * n + in] += monomial_over_factorial(d, G);
    }
  return m;
}

}  // namespace cdfmm_hip
