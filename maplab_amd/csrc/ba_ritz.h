// ba_ritz.h -- small dense eigenproblems on the host for the recycled
// deflation space of the PCG solve (Ritz vectors of a finished solve).
#ifndef MAPLAB_AMD_BA_RITZ_H_
#define MAPLAB_AMD_BA_RITZ_H_

#include <cstdint>

namespace mlba {

// The k smallest eigenpairs of the symmetric tridiagonal m x m matrix with
// diagonal d (m) and off-diagonal e (m - 1): eigenvalues by Sturm-sequence
// bisection, eigenvectors by inverse iteration, orthonormalised.
// vals (k) ascending, vecs (m x k, column c = eigenvector of vals[c],
// row-major: vecs[j * k + c]). Returns 0, or -1 for bad sizes.
int32_t tridiag_smallest_eigenpairs(int32_t m, const double* d, const double* e, int32_t k, double* vals,
                                    double* vecs);

// The k smallest eigenpairs of the symmetric-definite pencil (A, B), n x n
// row-major: A c = theta B c with c' B c = 1. vals (k) ascending, C (n x k,
// row-major: C[i * k + c]). Returns 0, or -1 when B is not positive definite
// (or bad sizes).
int32_t generalized_smallest_eigenpairs(int32_t n, const double* A, const double* B, int32_t k, double* vals,
                                        double* C);

}  // namespace mlba

#endif  // MAPLAB_AMD_BA_RITZ_H_
