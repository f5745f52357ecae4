// C entry points over the host eigen-solvers of the recycled PCG deflation
// space (maplab_amd/csrc/ba_ritz.cc), compiled with it by the CPU tests.
#include "../../maplab_amd/csrc/ba_ritz.h"

extern "C" int ritz_tridiag(int m, const double* d, const double* e, int k, double* vals, double* vecs) {
  return mlba::tridiag_smallest_eigenpairs(m, d, e, k, vals, vecs);
}

extern "C" int ritz_generalized(int n, const double* A, const double* B, int k, double* vals, double* C) {
  return mlba::generalized_smallest_eigenpairs(n, A, B, k, vals, C);
}
