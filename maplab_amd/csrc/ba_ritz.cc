// ba_ritz.cc -- host eigen-solvers for the recycled PCG deflation space.
//
// tridiag_smallest_eigenpairs: the Lanczos tridiagonal of a finished PCG
// solve is small (at most the stored iterations); only its k smallest
// eigenpairs are wanted, so eigenvalues come from Sturm-sequence bisection
// (the count of negative pivots of the LDL' factorisation of T - x I is the
// number of eigenvalues below x) and eigenvectors from inverse iteration with
// a partially pivoted tridiagonal LU (Golub & Van Loan, Matrix Computations,
// 4th ed., 8.4.1 and 8.2.2), orthogonalised in order by modified Gram-Schmidt.
//
// generalized_smallest_eigenpairs: A c = theta B c for the Rayleigh-Ritz merge
// of two Ritz blocks (n <= a few tens): Cholesky B = L L', the standard
// problem L^-1 A L^-T by the cyclic Jacobi method (Golub & Van Loan 8.5.2),
// c = L^-T v.
#include "ba_ritz.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace mlba {

namespace {

// Eigenvalues of T below x.
int sturm_count(int m, const double* d, const double* e, double x) {
  const double tiny = std::numeric_limits<double>::min() * 4.0;
  int cnt = 0;
  double q = d[0] - x;
  if (q < 0.0) ++cnt;
  for (int j = 1; j < m; ++j) {
    if (std::fabs(q) < tiny) q = q < 0.0 ? -tiny : tiny;
    q = d[j] - x - e[j - 1] * e[j - 1] / q;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

// Solves (T - s I) x = b in place with a partially pivoted tridiagonal LU.
void shifted_solve(int m, const double* d, const double* e, double s, std::vector<double>& x) {
  // Rows of U: diagonal u0, first and second super-diagonals u1, u2.
  std::vector<double> u0(m), u1(m, 0.0), u2(m, 0.0), mult(m, 0.0);
  std::vector<char> swp(m, 0);
  const double eps = std::numeric_limits<double>::epsilon();
  double norm = 0.0;
  for (int j = 0; j < m; ++j) norm = std::max(norm, std::fabs(d[j] - s) + (j ? std::fabs(e[j - 1]) : 0.0) +
                                                     (j + 1 < m ? std::fabs(e[j]) : 0.0));
  const double floor_piv = eps * std::max(norm, 1e-300);
  // Current row j: (a, b, c) on columns j, j + 1, j + 2.
  double a = d[0] - s, b = m > 1 ? e[0] : 0.0, c = 0.0;
  for (int j = 0; j < m - 1; ++j) {
    const double lo = e[j];                      // sub-diagonal entry of row j + 1
    const double na = d[j + 1] - s, nb = j + 2 < m ? e[j + 1] : 0.0;
    if (std::fabs(lo) > std::fabs(a)) {
      // Swap rows j and j + 1.
      swp[j] = 1;
      const double f = a / lo;
      mult[j] = f;
      u0[j] = lo;
      u1[j] = na;
      u2[j] = nb;
      a = b - f * na;
      b = c - f * nb;
      c = 0.0;
    } else {
      if (std::fabs(a) < floor_piv) a = floor_piv;
      const double f = lo / a;
      mult[j] = f;
      u0[j] = a;
      u1[j] = b;
      u2[j] = c;
      a = na - f * b;
      b = nb - f * c;
      c = 0.0;
    }
  }
  if (std::fabs(a) < floor_piv) a = floor_piv;
  u0[m - 1] = a;
  // Forward: apply the row swaps and multipliers to x.
  for (int j = 0; j < m - 1; ++j) {
    if (swp[j]) std::swap(x[j], x[j + 1]);
    x[j + 1] -= mult[j] * x[j];
  }
  // Back substitution with U.
  for (int j = m - 1; j >= 0; --j) {
    double v = x[j];
    if (j + 1 < m) v -= u1[j] * x[j + 1];
    if (j + 2 < m) v -= u2[j] * x[j + 2];
    x[j] = v / u0[j];
  }
}

double normalize(std::vector<double>& x) {
  double n = 0.0;
  for (double v : x) n += v * v;
  n = std::sqrt(n);
  if (n > 0.0)
    for (double& v : x) v /= n;
  return n;
}

}  // namespace

int32_t tridiag_smallest_eigenpairs(int32_t m, const double* d, const double* e, int32_t k, double* vals,
                                    double* vecs) {
  if (m <= 0 || k <= 0 || k > m) return -1;
  double lo = d[0], hi = d[0];
  for (int j = 0; j < m; ++j) {
    const double r = (j ? std::fabs(e[j - 1]) : 0.0) + (j + 1 < m ? std::fabs(e[j]) : 0.0);
    lo = std::min(lo, d[j] - r);
    hi = std::max(hi, d[j] + r);
  }
  const double scale = std::max(std::fabs(lo), std::fabs(hi));
  const double eps = std::numeric_limits<double>::epsilon();
  for (int c = 0; c < k; ++c) {
    // The (c + 1)-th smallest eigenvalue: the least x with count(x) >= c + 1.
    double a = lo, b = hi;
    for (int it = 0; it < 200 && b - a > 4.0 * eps * scale; ++it) {
      const double mid = 0.5 * (a + b);
      if (sturm_count(m, d, e, mid) >= c + 1) b = mid; else a = mid;
    }
    vals[c] = 0.5 * (a + b);
  }
  std::vector<std::vector<double>> V(k, std::vector<double>(m));
  for (int c = 0; c < k; ++c) {
    std::vector<double>& x = V[c];
    for (int j = 0; j < m; ++j) x[j] = 1.0 + 0.01 * ((j * 7 + c * 13) % 17);
    normalize(x);
    // Shift just off the eigenvalue so that T - s I stays non-singular.
    const double s = vals[c] + 8.0 * eps * std::max(scale, 1e-300);
    for (int it = 0; it < 4; ++it) {
      shifted_solve(m, d, e, s, x);
      for (int p = 0; p < c; ++p) {  // modified Gram-Schmidt against the earlier vectors
        double dot = 0.0;
        for (int j = 0; j < m; ++j) dot += V[p][j] * x[j];
        for (int j = 0; j < m; ++j) x[j] -= dot * V[p][j];
      }
      if (!(normalize(x) > 0.0)) {
        for (int j = 0; j < m; ++j) x[j] = (j == c) ? 1.0 : 0.0;
      }
    }
  }
  for (int j = 0; j < m; ++j)
    for (int c = 0; c < k; ++c) vecs[(size_t)j * k + c] = V[c][j];
  return 0;
}

int32_t generalized_smallest_eigenpairs(int32_t n, const double* A, const double* B, int32_t k, double* vals,
                                        double* C) {
  if (n <= 0 || k <= 0 || k > n) return -1;
  // Cholesky B = L L'.
  std::vector<double> L((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    double s = B[(size_t)j * n + j];
    for (int p = 0; p < j; ++p) s -= L[(size_t)j * n + p] * L[(size_t)j * n + p];
    if (!(s > 0.0)) return -1;
    const double ljj = std::sqrt(s);
    L[(size_t)j * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      double t = B[(size_t)i * n + j];
      for (int p = 0; p < j; ++p) t -= L[(size_t)i * n + p] * L[(size_t)j * n + p];
      L[(size_t)i * n + j] = t / ljj;
    }
  }
  // M = L^-1 A L^-T: X = L^-1 A column by column, then M = L^-1 X'.
  std::vector<double> X((size_t)n * n), M((size_t)n * n);
  for (int col = 0; col < n; ++col)
    for (int i = 0; i < n; ++i) {
      double t = A[(size_t)i * n + col];
      for (int p = 0; p < i; ++p) t -= L[(size_t)i * n + p] * X[(size_t)p * n + col];
      X[(size_t)i * n + col] = t / L[(size_t)i * n + i];
    }
  for (int col = 0; col < n; ++col)  // column col of M = L^-1 (row col of X)
    for (int i = 0; i < n; ++i) {
      double t = X[(size_t)col * n + i];
      for (int p = 0; p < i; ++p) t -= L[(size_t)i * n + p] * M[(size_t)p * n + col];
      M[(size_t)i * n + col] = t / L[(size_t)i * n + i];
    }
  for (int i = 0; i < n; ++i)  // symmetrise
    for (int j = i + 1; j < n; ++j) {
      const double v = 0.5 * (M[(size_t)i * n + j] + M[(size_t)j * n + i]);
      M[(size_t)i * n + j] = M[(size_t)j * n + i] = v;
    }
  // Cyclic Jacobi: M = V diag V'.
  std::vector<double> V((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < n; ++i) {
      diag += M[(size_t)i * n + i] * M[(size_t)i * n + i];
      for (int j = i + 1; j < n; ++j) off += M[(size_t)i * n + j] * M[(size_t)i * n + j];
    }
    if (off <= 1e-30 * diag || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = M[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double app = M[(size_t)p * n + p], aqq = M[(size_t)q * n + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int r = 0; r < n; ++r) {  // columns p, q
          const double mp = M[(size_t)r * n + p], mq = M[(size_t)r * n + q];
          M[(size_t)r * n + p] = c * mp - s * mq;
          M[(size_t)r * n + q] = s * mp + c * mq;
        }
        for (int r = 0; r < n; ++r) {  // rows p, q
          const double mp = M[(size_t)p * n + r], mq = M[(size_t)q * n + r];
          M[(size_t)p * n + r] = c * mp - s * mq;
          M[(size_t)q * n + r] = s * mp + c * mq;
        }
        for (int r = 0; r < n; ++r) {
          const double vp = V[(size_t)r * n + p], vq = V[(size_t)r * n + q];
          V[(size_t)r * n + p] = c * vp - s * vq;
          V[(size_t)r * n + q] = s * vp + c * vq;
        }
      }
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return M[(size_t)x * n + x] < M[(size_t)y * n + y]; });
  for (int c = 0; c < k; ++c) {
    const int col = order[c];
    vals[c] = M[(size_t)col * n + col];
    // c = L^-T v.
    std::vector<double> v(n);
    for (int i = 0; i < n; ++i) v[i] = V[(size_t)i * n + col];
    for (int i = n - 1; i >= 0; --i) {
      double t = v[i];
      for (int p = i + 1; p < n; ++p) t -= L[(size_t)p * n + i] * v[p];
      v[i] = t / L[(size_t)i * n + i];
    }
    for (int i = 0; i < n; ++i) C[(size_t)i * k + c] = v[i];
  }
  return 0;
}

}  // namespace mlba
