// Krylov-basis kernels of the restarted GMRES driving Newton: the solution update from a basis
// stored in FP32 or FP64, accumulated in FP64.
//
// Reference: the GMRES update x <- x + Z y of lineax's restarted GMRES used by solve_linear
// (aquaflux/solve/linear.py:199-298); there the basis is FP64.  Here the outer basis may be FP32
// (it only carries search directions to 1e-3), so the update reads it once, widens each entry and
// sums the k terms in FP64 in a fixed order: the correction keeps FP64 relative accuracy in the
// FP32 basis' own directions.
#include <type_traits>

#include "afx_common.h"
#include "afx_profile.h"
#include "aquaflux_amd.h"

using namespace afx;

namespace {

// x[i] += sum_j y[j] basis[j * ld + i], j = 0, 1, ..., k - 1 in that order.
template <class T>
__global__ void __launch_bounds__(256) k_basis_combine(int64_t n, int32_t k, const T* __restrict__ basis, int64_t ld,
                                                       const double* __restrict__ y, double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int32_t j = 0; j < k; ++j) acc += y[j] * (double)basis[(int64_t)j * ld + i];
  x[i] += acc;
}

}  // namespace

extern "C" int afx_basis_combine(int64_t n, int32_t k, int precision, const void* basis, int64_t ld,
                                 const double* y, double* x, hipStream_t s) {
  AFX_REQUIRE(basis && y && x && n > 0 && k >= 0 && ld >= n, "afx_basis_combine: bad argument");
  if (k == 0) return 0;
  if (precision == 0) {
    AFX_LAUNCH(kKKrylov, k_basis_combine<float>, grid_for(n), dim3(kBlock), 0, s, n, k, (const float*)basis, ld, y, x);
  } else if (precision == 1) {
    AFX_LAUNCH(kKKrylov, k_basis_combine<double>, grid_for(n), dim3(kBlock), 0, s, n, k, (const double*)basis, ld, y,
               x);
  } else {
    afx_set_error("afx_basis_combine: precision must be 0 (fp32) or 1 (fp64)");
    return 2;
  }
  AFX_CHECK_LAUNCH();
  return 0;
}
