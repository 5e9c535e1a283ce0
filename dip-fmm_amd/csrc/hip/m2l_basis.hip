// Harmonic subspace of the M2L operators (host, FP64, once per plan).
//
// Every M2L class matrix T_o[beta][alpha] (src/static_operators.cpp:159-184)
// is built from the derivatives D^(alpha+beta) G of a harmonic kernel.  Within
// one degree block (|alpha| = k, |beta| = j) the block is traceless in both
// indices (the trace of D^(alpha+beta) G is D^(...) laplace(G) = 0), so for
// every offset o the columns lie in one fixed (2j+1)-dimensional space and
// the rows in one fixed (2k+1)-dimensional space.  With orthonormal bases Q_j
// (C(j+2,2) x (2j+1)) and P_k ((2k+1) x C(k+2,2)) of those spaces
//   T_o = Q C_o P,   C_o = Q^T T_o P^T,
// exactly (to FP64 rounding), with C_o of size (p+1)^2 x (p+1)^2 instead of
// C(p+3,3)^2 -- 81 x 81 instead of 165 x 165 at p = 8.  The M2L then runs as
//   L = ls (.) Q sum_o C_o P (ms (.) M),
// the same linear map, 4.2x fewer products at p = 8.
//
// The bases are computed from the caller's matrices (one-sided Jacobi SVD of
// each degree block stacked over a few offsets), and the factorisation is
// checked on every class; matrices that are not of this form keep the full
// M2L (harmonic_m2l_basis returns nc = 0).
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"

namespace hipfmm {

namespace {

// Right singular vectors of A (m x c, column-major, overwritten) by one-sided
// Jacobi rotations; V (c x c, column-major) sorted by singular value, sv
// descending.
void jacobi_svd(std::vector<double>& A, int m, int c, std::vector<double>& V, std::vector<double>& sv) {
  V.assign(size_t(c) * c, 0.0);
  for (int i = 0; i < c; ++i) V[size_t(i) * c + i] = 1.0;
  for (int sweep = 0; sweep < 80; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < c - 1; ++i)
      for (int j = i + 1; j < c; ++j) {
        double* ai = &A[size_t(i) * m];
        double* aj = &A[size_t(j) * m];
        double a = 0.0, b = 0.0, g = 0.0;
        for (int r = 0; r < m; ++r) {
          a += ai[r] * ai[r];
          b += aj[r] * aj[r];
          g += ai[r] * aj[r];
        }
        if (a == 0.0 || b == 0.0) continue;
        const double rel = std::fabs(g) / std::sqrt(a * b);
        if (rel <= 1e-15) continue;
        off = std::max(off, rel);
        const double zeta = (b - a) / (2.0 * g);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
        for (int r = 0; r < m; ++r) {
          const double x = ai[r], y = aj[r];
          ai[r] = cs * x - sn * y;
          aj[r] = sn * x + cs * y;
        }
        double* vi = &V[size_t(i) * c];
        double* vj = &V[size_t(j) * c];
        for (int r = 0; r < c; ++r) {
          const double x = vi[r], y = vj[r];
          vi[r] = cs * x - sn * y;
          vj[r] = sn * x + cs * y;
        }
      }
    if (off <= 1e-15) break;
  }
  std::vector<double> norms(c);
  for (int i = 0; i < c; ++i) {
    double s = 0.0;
    for (int r = 0; r < m; ++r) s += A[size_t(i) * m + r] * A[size_t(i) * m + r];
    norms[i] = std::sqrt(s);
  }
  std::vector<int> order(c);
  for (int i = 0; i < c; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return norms[x] > norms[y]; });
  std::vector<double> Vs(size_t(c) * c);
  sv.resize(c);
  for (int i = 0; i < c; ++i) {
    sv[i] = norms[order[i]];
    std::copy(&V[size_t(order[i]) * c], &V[size_t(order[i]) * c] + c, &Vs[size_t(i) * c]);
  }
  V.swap(Vs);
}

}  // namespace

HarmonicBasis harmonic_m2l_basis(int p, const double* T, int classes) {
  HarmonicBasis h;
  const int n = coefficient_count(p), nc = (p + 1) * (p + 1);
  if (classes <= 0 || p < 1) return h;
  // Degree blocks of the basis (degree-major order): [first, first + size).
  std::vector<int> first(p + 2, 0);
  for (int k = 0; k <= p; ++k) first[k + 1] = first[k] + (k + 1) * (k + 2) / 2;
  // A few offsets span the fixed spaces; spread them over the class list.
  std::vector<int> pick;
  const int want = std::min(classes, 6);
  for (int i = 0; i < want; ++i) pick.push_back(int((int64_t(i) * classes) / want));
  auto t = [&](int cls, int beta, int alpha) { return T[size_t(cls) * n * n + beta + size_t(n) * alpha]; };
  std::vector<double> P(size_t(nc) * n, 0.0), Q(size_t(n) * nc, 0.0);  // row-major
  for (int k = 0; k <= p; ++k) {
    const int ck = first[k + 1] - first[k], rk = 2 * k + 1;
    // Row space in alpha (degree k): stack rows (offset, beta).
    {
      const int m = int(pick.size()) * n;
      std::vector<double> A(size_t(m) * ck), V, sv;
      for (int col = 0; col < ck; ++col)
        for (size_t q = 0; q < pick.size(); ++q)
          for (int beta = 0; beta < n; ++beta) A[size_t(col) * m + q * n + beta] = t(pick[q], beta, first[k] + col);
      jacobi_svd(A, m, ck, V, sv);
      if (sv[0] == 0.0 || (rk < ck && sv[rk] > 1e-12 * sv[0]) || sv[rk - 1] < 1e-10 * sv[0]) return h;
      for (int r = 0; r < rk; ++r)
        for (int a = 0; a < ck; ++a) P[size_t(k * k + r) * n + first[k] + a] = V[size_t(r) * ck + a];
    }
    // Column space in beta (degree k): stack columns (offset, alpha).
    {
      const int m = int(pick.size()) * n;
      std::vector<double> A(size_t(m) * ck), V, sv;
      for (int row = 0; row < ck; ++row)
        for (size_t q = 0; q < pick.size(); ++q)
          for (int alpha = 0; alpha < n; ++alpha) A[size_t(row) * m + q * n + alpha] = t(pick[q], first[k] + row, alpha);
      jacobi_svd(A, m, ck, V, sv);
      if (sv[0] == 0.0 || (rk < ck && sv[rk] > 1e-12 * sv[0]) || sv[rk - 1] < 1e-10 * sv[0]) return h;
      for (int r = 0; r < rk; ++r)
        for (int b = 0; b < ck; ++b) Q[size_t(first[k] + b) * nc + k * k + r] = V[size_t(r) * ck + b];
    }
  }
  // C_o = Q^T T_o P^T per class (block-diagonal bases), then the check
  // T_o = Q C_o P on every class.
  std::vector<double> C(size_t(classes) * nc * nc, 0.0);
  std::vector<double> W(size_t(n) * nc);  // T_o P^T, n x nc row-major
  double worst = 0.0;
  for (int cls = 0; cls < classes; ++cls) {
    double tmax = 0.0;
    for (int beta = 0; beta < n; ++beta)
      for (int k = 0; k <= p; ++k)
        for (int r = k * k; r < (k + 1) * (k + 1); ++r) {
          double s = 0.0;
          for (int a = first[k]; a < first[k + 1]; ++a) s += t(cls, beta, a) * P[size_t(r) * n + a];
          W[size_t(beta) * nc + r] = s;
        }
    double* Co = &C[size_t(cls) * nc * nc];
    for (int j = 0; j <= p; ++j)
      for (int rb = j * j; rb < (j + 1) * (j + 1); ++rb)
        for (int ra = 0; ra < nc; ++ra) {
          double s = 0.0;
          for (int b = first[j]; b < first[j + 1]; ++b) s += Q[size_t(b) * nc + rb] * W[size_t(b) * nc + ra];
          Co[rb + size_t(nc) * ra] = s;
        }
    // U = C_o P (nc x n), then the reconstruction Q U against T_o.
    std::vector<double> U(size_t(nc) * n, 0.0);
    for (int rb = 0; rb < nc; ++rb)
      for (int k = 0; k <= p; ++k)
        for (int alpha = first[k]; alpha < first[k + 1]; ++alpha) {
          double s = 0.0;
          for (int ra = k * k; ra < (k + 1) * (k + 1); ++ra) s += Co[rb + size_t(nc) * ra] * P[size_t(ra) * n + alpha];
          U[size_t(rb) * n + alpha] = s;
        }
    double err = 0.0;
    for (int j = 0; j <= p; ++j)
      for (int beta = first[j]; beta < first[j + 1]; ++beta)
        for (int alpha = 0; alpha < n; ++alpha) {
          double s = 0.0;
          for (int rb = j * j; rb < (j + 1) * (j + 1); ++rb) s += Q[size_t(beta) * nc + rb] * U[size_t(rb) * n + alpha];
          const double v = t(cls, beta, alpha);
          tmax = std::max(tmax, std::fabs(v));
          err = std::max(err, std::fabs(v - s));
        }
    if (tmax > 0.0) worst = std::max(worst, err / tmax);
  }
  if (!(worst <= 1e-11)) return h;
  // Degree 0 goes last: compressed index r holds harmonic row (r + 1) mod nc.
  // Multipoles of dipole sources have no monopole (P2M writes none, M2M maps
  // it only to itself), so the last compressed multipole is exactly zero and
  // the matrix-core M2L can stop its depth at nc - 1.
  std::vector<double> Pr(P.size()), Qr(Q.size()), Cr(C.size());
  for (int r = 0; r < nc; ++r) {
    const int o = (r + 1) % nc;
    for (int a = 0; a < n; ++a) Pr[size_t(r) * n + a] = P[size_t(o) * n + a];
    for (int b = 0; b < n; ++b) Qr[size_t(b) * nc + r] = Q[size_t(b) * nc + o];
  }
  for (int cls = 0; cls < classes; ++cls)
    for (int a = 0; a < nc; ++a)
      for (int b = 0; b < nc; ++b)
        Cr[size_t(cls) * nc * nc + b + size_t(nc) * a] =
            C[size_t(cls) * nc * nc + (b + 1) % nc + size_t(nc) * ((a + 1) % nc)];
  h.nc = nc;
  h.P = std::move(Pr);
  h.Q = std::move(Qr);
  h.C = std::move(Cr);
  h.max_reconstruction_error = worst;
  return h;
}

}  // namespace hipfmm
