// Aggregation-multigrid kernels for a scalar ELL operator (the pressure Schur of the block
// preconditioner): a damped Jacobi sweep fused with its operator product, the residual, the
// restriction that sums each aggregate's members, and the prolongation of a coarse correction.
//
// ELL layout as the block-ELL with block 1: vals[s * n + i], cols[s * n + i] for slot s of row i,
// slot 0 the diagonal, padding slots zero with column i.  One thread per row; the restriction reads
// an aggregate's members from a (max_members x n_coarse) table padded with -1, so every sum has one
// writer and a fixed order (no atomics, bitwise reproducible).
//
// Reference: the smoothed-aggregation hierarchies the reference builds with scipy for its Schur
// and velocity blocks (aquaflux/flow/block_preconditioner.py:266-343, aquaflux/solve/multigrid.py:
// 1248-1700); this hierarchy is an unsmoothed, geometrically aggregated one run on the device.
#include <type_traits>

#include "afx_common.h"
#include "afx_profile.h"

using namespace afx;

namespace {

template <class T>
AFX_DEV T ell_row_product(int32_t n, int32_t nslot, const T* __restrict__ vals, const int32_t* __restrict__ cols,
                          const T* __restrict__ x, int32_t i) {
  T acc = T(0);
  for (int32_t s = 0; s < nslot; ++s) {
    const int64_t e = (int64_t)s * n + i;
    acc += vals[e] * x[cols[e]];
  }
  return acc;
}

// y = x + omega (b - A x) / diag(A); from x = 0 when x is null: y = omega b / diag(A).
template <class T>
__global__ void __launch_bounds__(256) k_ell_jacobi(int32_t n, int32_t nslot, const T* __restrict__ vals,
                                                    const int32_t* __restrict__ cols, const T* __restrict__ b,
                                                    const T* __restrict__ x, T* __restrict__ y, T omega) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T d = vals[i];
  if (x == nullptr) {
    y[i] = omega * b[i] / d;
    return;
  }
  y[i] = x[i] + omega * (b[i] - ell_row_product<T>(n, nslot, vals, cols, x, i)) / d;
}

template <class T>
__global__ void __launch_bounds__(256) k_ell_residual(int32_t n, int32_t nslot, const T* __restrict__ vals,
                                                      const int32_t* __restrict__ cols, const T* __restrict__ b,
                                                      const T* __restrict__ x, T* __restrict__ r) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  r[i] = b[i] - ell_row_product<T>(n, nslot, vals, cols, x, i);
}

template <class T>
__global__ void __launch_bounds__(256) k_aggregate_restrict(int32_t nc, int32_t mmax,
                                                            const int32_t* __restrict__ members,
                                                            const T* __restrict__ r, T* __restrict__ rc) {
  const int32_t I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nc) return;
  T acc = T(0);
  for (int32_t k = 0; k < mmax; ++k) {
    const int32_t m = members[(int64_t)k * nc + I];
    if (m >= 0) acc += r[m];
  }
  rc[I] = acc;
}

template <class T>
__global__ void __launch_bounds__(256) k_aggregate_prolong(int32_t n, const int32_t* __restrict__ agg,
                                                           const T* __restrict__ xc, T alpha, T* __restrict__ x) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] += alpha * xc[agg[i]];
}

// Dot product in two passes with no atomics and no cleared memory: kDotBlocks partial sums
// (grid-stride, FP64 accumulation, LDS tree per block), then one block sums the partials.
constexpr int kDotBlocks = 256;

AFX_DEV double block_sum(double v) {
  __shared__ double lds[256];
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
    __syncthreads();
  }
  return lds[0];
}

template <class T>
__global__ void __launch_bounds__(256) k_dot_partial(int64_t n, const T* __restrict__ a, const T* __restrict__ b,
                                                     double* __restrict__ partial) {
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc += (double)a[i] * (double)b[i];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <class T>
__global__ void __launch_bounds__(256) k_dot_final(const double* __restrict__ partial, T* __restrict__ out) {
  const double s = block_sum(partial[threadIdx.x]);
  if (threadIdx.x == 0) out[0] = (T)s;
}

// y = A x for a small dense row-major matrix (the coarsest level's pseudo-inverse): one 64-lane
// wavefront per row, lanes striding the row (coalesced), a shuffle tree for the sum.
template <class T>
__global__ void __launch_bounds__(256) k_dense_matvec(int32_t n, const T* __restrict__ A, const T* __restrict__ x,
                                                      T* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int32_t i = (int32_t)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= n) return;  // i is uniform across the wavefront
  const T* row = A + (int64_t)i * n;
  double acc = 0.0;
  for (int32_t j = lane; j < n; j += 64) acc += (double)row[j] * (double)x[j];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) y[i] = (T)acc;
}

template <class Fn>
int by_precision(int precision, Fn&& fn) {
  if (precision == 0) return fn((float*)nullptr);
  if (precision == 1) return fn((double*)nullptr);
  afx_set_error("precision must be 0 (fp32) or 1 (fp64)");
  return 2;
}

}  // namespace

extern "C" int afx_ell_jacobi(int64_t n, int32_t nslot, int precision, const void* vals, const int32_t* cols,
                              const void* b, const void* x, void* y, double omega, hipStream_t s) {
  AFX_REQUIRE(vals && cols && b && y && n > 0 && n < INT32_MAX && nslot > 0, "afx_ell_jacobi: bad argument");
  AFX_REQUIRE(x != y, "afx_ell_jacobi: x and y must be distinct buffers");
  return by_precision(precision, [&](auto* tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    AFX_LAUNCH(kKAmg, k_ell_jacobi<T>, grid_for(n), dim3(kBlock), 0, s, (int32_t)n, nslot, (const T*)vals, cols,
               (const T*)b, (const T*)x, (T*)y, (T)omega);
    AFX_CHECK_LAUNCH();
    return 0;
  });
}

extern "C" int afx_ell_residual(int64_t n, int32_t nslot, int precision, const void* vals, const int32_t* cols,
                                const void* b, const void* x, void* r, hipStream_t s) {
  AFX_REQUIRE(vals && cols && b && x && r && n > 0 && n < INT32_MAX && nslot > 0, "afx_ell_residual: bad argument");
  AFX_REQUIRE(x != r, "afx_ell_residual: x and r must be distinct buffers");
  return by_precision(precision, [&](auto* tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    AFX_LAUNCH(kKAmg, k_ell_residual<T>, grid_for(n), dim3(kBlock), 0, s, (int32_t)n, nslot, (const T*)vals, cols,
               (const T*)b, (const T*)x, (T*)r);
    AFX_CHECK_LAUNCH();
    return 0;
  });
}

extern "C" int afx_aggregate_restrict(int64_t n_coarse, int32_t max_members, int precision, const int32_t* members,
                                      const void* r, void* rc, hipStream_t s) {
  AFX_REQUIRE(members && r && rc && n_coarse > 0 && n_coarse < INT32_MAX && max_members > 0,
              "afx_aggregate_restrict: bad argument");
  return by_precision(precision, [&](auto* tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    AFX_LAUNCH(kKAmg, k_aggregate_restrict<T>, grid_for(n_coarse), dim3(kBlock), 0, s, (int32_t)n_coarse,
               max_members, members, (const T*)r, (T*)rc);
    AFX_CHECK_LAUNCH();
    return 0;
  });
}

extern "C" int afx_aggregate_prolong(int64_t n, int precision, const int32_t* agg, const void* xc, double alpha,
                                     void* x, hipStream_t s) {
  AFX_REQUIRE(agg && xc && x && n > 0 && n < INT32_MAX, "afx_aggregate_prolong: bad argument");
  return by_precision(precision, [&](auto* tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    AFX_LAUNCH(kKAmg, k_aggregate_prolong<T>, grid_for(n), dim3(kBlock), 0, s, (int32_t)n, agg, (const T*)xc,
               (T)alpha, (T*)x);
    AFX_CHECK_LAUNCH();
    return 0;
  });
}

extern "C" int afx_dot(int64_t n, int precision, const void* a, const void* b, double* partial, void* out,
                       hipStream_t s) {
  AFX_REQUIRE(a && b && partial && out && n > 0, "afx_dot: bad argument");
  return by_precision(precision, [&](auto* tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    AFX_LAUNCH(kKAmg, k_dot_partial<T>, dim3(kDotBlocks), dim3(kBlock), 0, s, n, (const T*)a, (const T*)b, partial);
    AFX_CHECK_LAUNCH();
    AFX_LAUNCH(kKAmg, k_dot_final<T>, dim3(1), dim3(kBlock), 0, s, (const double*)partial, (T*)out);
    AFX_CHECK_LAUNCH();
    return 0;
  });
}

extern "C" int afx_dense_matvec(int64_t n, int precision, const void* A, const void* x, void* y, hipStream_t s) {
  AFX_REQUIRE(A && x && y && n > 0 && n < INT32_MAX && x != y, "afx_dense_matvec: bad argument");
  return by_precision(precision, [&](auto* tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    AFX_LAUNCH(kKAmg, k_dense_matvec<T>, grid_for(n * 64), dim3(kBlock), 0, s, (int32_t)n, (const T*)A, (const T*)x,
               (T*)y);
    AFX_CHECK_LAUNCH();
    return 0;
  });
}
