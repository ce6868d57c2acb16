// Krylov kernels of the restarted GMRES driving Newton, with the whole Arnoldi recurrence on the
// device: Gram-Schmidt, the Hessenberg column, its Givens rotations, the residual estimate, the
// stopping test, the small triangular solve and the solution update.  The host only enqueues work
// and reads a few status doubles when it decides to look, so a cycle needs no host round trip per
// step.
//
// Reference: lineax's restarted GMRES used by solve_linear (aquaflux/solve/linear.py:199-298):
// Arnoldi with Givens rotations, stop when |g_{k}| <= atol + rtol ||b||, x <- x + Z y at restart.
// The basis may be FP32 (the outer basis only carries search directions to 1e-3): every dot
// product, norm and update reads it once, widens each entry and sums in FP64 in a fixed order, so
// every result is bitwise reproducible run to run.
//
// Orthogonalisation is classical Gram-Schmidt, with a second pass when the first removed more than
// 1 - 1/sqrt(2) of the new vector's norm (Daniel, Gragg, Kaufman and Stewart 1976, "twice is
// enough"), or on every step.  One pass is two sweeps over the basis: the dot products of all j + 1
// basis vectors with w (w held in registers, the basis read once), then w - V^T h (the basis read
// once more) with the new norm.  The decision for a second pass is taken on the device: the pass-2
// kernels are always enqueued and return at once when it is not needed.
#include <type_traits>

#include "afx_common.h"
#include "afx_profile.h"
#include "aquaflux_amd.h"

using namespace afx;

namespace {

constexpr int kEPT = 8;                  // vector entries per thread held in registers
constexpr int kChunk = kBlock * kEPT;    // entries per workgroup of a Gram-Schmidt sweep
constexpr double kDgksEta = 0.70710678118654752440;  // 1 / sqrt(2)

// dstate layout (FP64): status scalars, the residual estimate after each step, then the cycle's
// small dense state.
enum : int64_t { kSK = 0, kSDone = 1, kSBreak = 2, kSPass2 = 3, kSTol = 4, kSBeta = 5, kSRes = 6, kSHn = 7,
                 kSHead = 8 };
struct Lay {
  int64_t m;
  AFX_HD int64_t resid() const { return kSHead; }
  AFX_HD int64_t g() const { return kSHead + m; }
  AFX_HD int64_t cs() const { return g() + m + 1; }
  AFX_HD int64_t sn() const { return cs() + m; }
  AFX_HD int64_t y() const { return sn() + m; }
  AFX_HD int64_t hp() const { return y() + m; }   // this pass's coefficients, then w0^2 at [m + 1]
  AFX_HD int64_t hs() const { return hp() + m + 2; }  // the column summed over the passes
  AFX_HD int64_t H() const { return hs() + m + 1; }   // column j at H + j (m + 1)
  AFX_HD int64_t total() const { return H() + (m + 1) * m; }
};

inline int64_t gs_blocks(int64_t n) { return (n + kChunk - 1) / kChunk; }

// Sum over the workgroup of NV values per thread, in a fixed order (wave butterflies, then the four
// wave sums in order).  Valid in every thread of wave 0; thread 0 holds the result.
template <int NV>
AFX_DEV void block_sum_n(double (&a)[NV]) {
  __shared__ double lds[NV][kBlock / 64];
#pragma unroll
  for (int u = 0; u < NV; ++u) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a[u] += __shfl_xor(a[u], off, 64);
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
#pragma unroll
    for (int u = 0; u < NV; ++u) lds[u][wave] = a[u];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int u = 0; u < NV; ++u) a[u] = ((lds[u][0] + lds[u][1]) + lds[u][2]) + lds[u][3];
  }
  __syncthreads();
}

AFX_DEV bool gated(const double* st, int pass) {
  if (st[kSDone] != 0.0) return true;
  return pass == 2 && st[kSPass2] == 0.0;
}

// Sum of partial[0 .. nb) in a fixed order by one workgroup (result in thread 0).
AFX_DEV double block_fixed_sum(const double* __restrict__ partial, int64_t nb) {
  double a[1] = {0.0};
  for (int64_t b = threadIdx.x; b < nb; b += kBlock) a[0] += partial[b];
  block_sum_n<1>(a);
  return a[0];
}

// ---- cycle start: beta = ||r||, V_0 = r / beta ---------------------------------------------------

template <class TR>
__global__ void __launch_bounds__(256) k_norm_partial(int64_t n, const TR* __restrict__ r, double* __restrict__ partial) {
  double a[1] = {0.0};
  const int64_t base = (int64_t)blockIdx.x * kChunk + threadIdx.x;
#pragma unroll
  for (int s = 0; s < kEPT; ++s) {
    const int64_t e = base + (int64_t)s * kBlock;
    if (e < n) {
      const double v = (double)r[e];
      a[0] += v * v;
    }
  }
  block_sum_n<1>(a);
  if (threadIdx.x == 0) partial[blockIdx.x] = a[0];
}

__global__ void __launch_bounds__(256) k_start_finalize(int64_t m, int64_t nb, const double* __restrict__ partial,
                                                        double* __restrict__ st, double rtol, double atol, int first) {
  const double s = block_fixed_sum(partial, nb);
  if (threadIdx.x != 0) return;
  const Lay L{m};
  const double beta = sqrt(s);
  if (first) st[kSTol] = atol + rtol * beta;
  st[kSK] = 0.0;
  st[kSBreak] = 0.0;
  st[kSPass2] = 0.0;
  st[kSBeta] = beta;
  st[kSRes] = beta;
  st[kSHn] = 0.0;
  st[L.g()] = beta;
  st[kSDone] = beta <= st[kSTol] ? 1.0 : 0.0;
}

template <class TB, class TR>
__global__ void __launch_bounds__(256) k_start_scale(int64_t n, const TR* __restrict__ r, TB* __restrict__ v0,
                                                     const double* __restrict__ st) {
  if (st[kSDone] != 0.0) return;
  const double beta = st[kSBeta];
  const int64_t base = (int64_t)blockIdx.x * kChunk + threadIdx.x;
#pragma unroll
  for (int s = 0; s < kEPT; ++s) {
    const int64_t e = base + (int64_t)s * kBlock;
    if (e < n) v0[e] = (TB)((double)r[e] / beta);
  }
}

// ---- one Gram-Schmidt pass ----------------------------------------------------------------------

// partial[v * nb + b] = sum over workgroup b's entries of V_v . w for v < nv; in pass 1 also
// partial[nv * nb + b] = w . w.  w is read as stored in the basis precision.
template <class TB, class TW>
__global__ void __launch_bounds__(256) k_gs_dots(int64_t n, int32_t nv, const TB* __restrict__ basis, int64_t ld,
                                                 const TW* __restrict__ w, double* __restrict__ partial, int64_t nb,
                                                 const double* __restrict__ st, int pass) {
  if (gated(st, pass)) return;
  const int64_t base = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  double wv[kEPT];
#pragma unroll
  for (int s = 0; s < kEPT; ++s) {
    const int64_t e = base + (int64_t)s * kBlock;
    wv[s] = e < n ? (double)(TB)w[e] : 0.0;
  }
  const int32_t nt = pass == 1 ? nv + 1 : nv;
  for (int32_t v0 = 0; v0 < nt; v0 += 4) {
    double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int32_t v = v0 + u;
      if (v < nv) {
        const TB* __restrict__ row = basis + (int64_t)v * ld;
#pragma unroll
        for (int s = 0; s < kEPT; ++s) {
          const int64_t e = base + (int64_t)s * kBlock;
          if (e < n) a[u] += (double)row[e] * wv[s];
        }
      } else if (v == nv && pass == 1) {
#pragma unroll
        for (int s = 0; s < kEPT; ++s) a[u] += wv[s] * wv[s];
      }
    }
    block_sum_n<4>(a);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (v0 + u < nt) partial[(int64_t)(v0 + u) * nb + blockIdx.x] = a[u];
    }
  }
}

// Workgroup v sums partial[v * nb ..] in a fixed order into hp[v] (pass 1: also hs[v] = hp[v], and
// w0^2 into hp[m + 1]; pass 2: hs[v] += hp[v]).
__global__ void __launch_bounds__(256) k_gs_reduce(int64_t m, int32_t nv, int64_t nb, const double* __restrict__ partial,
                                                   double* __restrict__ st, int pass) {
  if (gated(st, pass)) return;
  const Lay L{m};
  const int32_t v = blockIdx.x;
  const double s = block_fixed_sum(partial + (int64_t)v * nb, nb);
  if (threadIdx.x != 0) return;
  if (v == nv) {
    st[L.hp() + m + 1] = s;
    return;
  }
  st[L.hp() + v] = s;
  if (pass == 1)
    st[L.hs() + v] = s;
  else
    st[L.hs() + v] += s;
}

// out = w - sum_v hp[v] V_v (FP64, v in order), stored in the basis precision; partial[b] = the
// workgroup's sum of the stored entries squared.  out may alias w (pass 2).
template <class TB, class TW>
__global__ void __launch_bounds__(256) k_gs_update(int64_t n, int64_t m, int32_t nv, const TB* __restrict__ basis,
                                                   int64_t ld, const TW* w, TB* out, double* __restrict__ partial,
                                                   const double* __restrict__ st, int pass) {
  if (gated(st, pass)) return;
  const Lay L{m};
  const double* __restrict__ hp = st + L.hp();
  const int64_t base = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  double acc[kEPT];
#pragma unroll
  for (int s = 0; s < kEPT; ++s) {
    const int64_t e = base + (int64_t)s * kBlock;
    acc[s] = e < n ? (double)(TB)w[e] : 0.0;
  }
  for (int32_t v = 0; v < nv; ++v) {
    const double h = hp[v];
    const TB* __restrict__ row = basis + (int64_t)v * ld;
#pragma unroll
    for (int s = 0; s < kEPT; ++s) {
      const int64_t e = base + (int64_t)s * kBlock;
      if (e < n) acc[s] -= h * (double)row[e];
    }
  }
  double a[1] = {0.0};
#pragma unroll
  for (int s = 0; s < kEPT; ++s) {
    const int64_t e = base + (int64_t)s * kBlock;
    if (e < n) {
      const TB q = (TB)acc[s];
      out[e] = q;
      a[0] += (double)q * (double)q;
    }
  }
  block_sum_n<1>(a);
  if (threadIdx.x == 0) partial[blockIdx.x] = a[0];
}

// After a pass over step j: hn = ||w||.  Pass 1 decides whether a second pass is needed; the last
// pass of the step writes the Hessenberg column, applies the earlier rotations, forms the new one,
// updates g and the residual estimate |g_{j+1}|, and sets done when it reaches tol or on breakdown
// (hn = 0).
__global__ void __launch_bounds__(256) k_gs_finalize(int64_t m, int32_t j, int64_t nb, const double* __restrict__ partial,
                                                     double* __restrict__ st, int pass, int always) {
  if (gated(st, pass)) return;
  const double s = block_fixed_sum(partial, nb);
  if (threadIdx.x != 0) return;
  const Lay L{m};
  const double hn = sqrt(s);
  if (pass == 1) {
    const double w0 = sqrt(st[L.hp() + m + 1]);
    if (always || !(hn > kDgksEta * w0)) {
      st[kSPass2] = 1.0;
      return;
    }
  } else {
    st[kSPass2] = 0.0;
  }
  double* col = st + L.H() + (int64_t)j * (m + 1);
  double* cs = st + L.cs();
  double* sn = st + L.sn();
  double* g = st + L.g();
  for (int32_t i = 0; i <= j; ++i) col[i] = st[L.hs() + i];
  col[j + 1] = hn;
  for (int32_t i = 0; i < j; ++i) {
    const double t = cs[i] * col[i] + sn[i] * col[i + 1];
    col[i + 1] = -sn[i] * col[i] + cs[i] * col[i + 1];
    col[i] = t;
  }
  const double den = hypot(col[j], col[j + 1]);
  if (den == 0.0) {
    cs[j] = 1.0;
    sn[j] = 0.0;
  } else {
    cs[j] = col[j] / den;
    sn[j] = col[j + 1] / den;
  }
  col[j] = den;
  col[j + 1] = 0.0;
  g[j + 1] = -sn[j] * g[j];
  g[j] = cs[j] * g[j];
  const double res = fabs(g[j + 1]);
  st[kSK] = (double)(j + 1);
  st[kSRes] = res;
  st[L.resid() + j] = res;
  st[kSHn] = hn;
  if (hn == 0.0) st[kSBreak] = 1.0;
  if (res <= st[kSTol] || hn == 0.0) st[kSDone] = 1.0;
}

// V_{j+1} = w / hn (skipped once the cycle is done: that vector is never used).
template <class TB>
__global__ void __launch_bounds__(256) k_gs_scale(int64_t n, TB* __restrict__ out, const double* __restrict__ st) {
  if (st[kSDone] != 0.0) return;
  const double hn = st[kSHn];
  const int64_t base = (int64_t)blockIdx.x * kChunk + threadIdx.x;
#pragma unroll
  for (int s = 0; s < kEPT; ++s) {
    const int64_t e = base + (int64_t)s * kBlock;
    if (e < n) out[e] = (TB)((double)out[e] / hn);
  }
}

// ---- end of cycle: y = R^-1 g (k from the state), x += basis[:k]^T y ----------------------------

// Back substitution on the k x k upper triangle (a zero diagonal entry is taken as 1, the guard of
// the host solve it replaces).
__global__ void __launch_bounds__(64) k_gmres_y(int64_t m, double* __restrict__ st) {
  if (threadIdx.x != 0) return;
  const Lay L{m};
  const int32_t k = (int32_t)st[kSK];
  const double* H = st + L.H();
  const double* g = st + L.g();
  double* y = st + L.y();
  for (int32_t i = k - 1; i >= 0; --i) {
    double s = g[i];
    for (int32_t l = i + 1; l < k; ++l) s -= H[(int64_t)l * (m + 1) + i] * y[l];
    const double d = H[(int64_t)i * (m + 1) + i];
    y[i] = s / (d == 0.0 ? 1.0 : d);
  }
}

// x[i] += sum_j y[j] basis[j * ld + i], j = 0, 1, ..., k - 1 in that order (k given, or read from
// the GMRES state when kdev is set).
template <class T, class TX>
__global__ void __launch_bounds__(256) k_basis_combine(int64_t n, int32_t k, const double* __restrict__ kdev,
                                                       const T* __restrict__ basis, int64_t ld,
                                                       const double* __restrict__ y, TX* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (kdev) k = (int32_t)kdev[kSK];
  double acc = 0.0;
  for (int32_t j = 0; j < k; ++j) acc += y[j] * (double)basis[(int64_t)j * ld + i];
  x[i] = (TX)((double)x[i] + acc);
}

template <class F>
int by_precision(int p, const char* what, F&& f) {
  if (p == 0) return f(float{});
  if (p == 1) return f(double{});
  afx_set_error(std::string(what) + ": precision must be 0 (fp32) or 1 (fp64)");
  return 2;
}

}  // namespace

extern "C" int afx_basis_combine(int64_t n, int32_t k, int precision, const void* basis, int64_t ld,
                                 const double* y, double* x, hipStream_t s) {
  AFX_REQUIRE(basis && y && x && n > 0 && k >= 0 && ld >= n, "afx_basis_combine: bad argument");
  if (k == 0) return 0;
  return by_precision(precision, "afx_basis_combine", [&](auto tb) {
    using TB = decltype(tb);
    AFX_LAUNCH(kKKrylov, (k_basis_combine<TB, double>), grid_for(n), dim3(kBlock), 0, s, n, k, nullptr,
               (const TB*)basis, ld, y, x);
    AFX_CHECK_LAUNCH();
    return 0;
  });
}

extern "C" int afx_gmres_sizes(int64_t n, int32_t m, int64_t* state_doubles, int64_t* work_doubles) {
  AFX_REQUIRE(n > 0 && m > 0 && state_doubles && work_doubles, "afx_gmres_sizes: bad argument");
  *state_doubles = Lay{m}.total();
  *work_doubles = gs_blocks(n) * (int64_t)(m + 2);
  return 0;
}

extern "C" int afx_gmres_start(int64_t n, int32_t m, int precision, int r_precision, const void* r, double rtol,
                               double atol, int first, void* basis, int64_t ld, double* state, double* work,
                               hipStream_t s) {
  AFX_REQUIRE(n > 0 && m > 0 && r && basis && state && work && ld >= n, "afx_gmres_start: bad argument");
  const int64_t nb = gs_blocks(n);
  return by_precision(precision, "afx_gmres_start", [&](auto tb) {
    using TB = decltype(tb);
    return by_precision(r_precision, "afx_gmres_start (r)", [&](auto tr) {
      using TR = decltype(tr);
      AFX_LAUNCH(kKKrylov, k_norm_partial<TR>, dim3((unsigned)nb), dim3(kBlock), 0, s, n, (const TR*)r, work);
      AFX_CHECK_LAUNCH();
      AFX_LAUNCH(kKKrylov, k_start_finalize, dim3(1), dim3(kBlock), 0, s, (int64_t)m, nb, (const double*)work, state,
                 rtol, atol, first);
      AFX_CHECK_LAUNCH();
      AFX_LAUNCH(kKKrylov, (k_start_scale<TB, TR>), dim3((unsigned)nb), dim3(kBlock), 0, s, n, (const TR*)r,
                 (TB*)basis, (const double*)state);
      AFX_CHECK_LAUNCH();
      return 0;
    });
  });
}

extern "C" int afx_gmres_arnoldi(int64_t n, int32_t m, int32_t j, int precision, int w_precision, const void* w,
                                 void* basis, int64_t ld, int reorthogonalize, double* state, double* work,
                                 hipStream_t s) {
  AFX_REQUIRE(n > 0 && m > 0 && j >= 0 && j < m && w && basis && state && work && ld >= n,
              "afx_gmres_arnoldi: bad argument");
  AFX_REQUIRE(reorthogonalize == 0 || reorthogonalize == 1, "afx_gmres_arnoldi: reorthogonalize must be 0 or 1");
  const int64_t nb = gs_blocks(n);
  const int32_t nv = j + 1;
  return by_precision(precision, "afx_gmres_arnoldi", [&](auto tb) {
    using TB = decltype(tb);
    TB* V = (TB*)basis;
    TB* out = V + (int64_t)(j + 1) * ld;
    return by_precision(w_precision, "afx_gmres_arnoldi (w)", [&](auto tw) {
      using TW = decltype(tw);
      // pass 1 from the product w
      AFX_LAUNCH(kKKrylov, (k_gs_dots<TB, TW>), dim3((unsigned)nb), dim3(kBlock), 0, s, n, nv, (const TB*)V, ld,
                 (const TW*)w, work, nb, (const double*)state, 1);
      AFX_CHECK_LAUNCH();
      AFX_LAUNCH(kKKrylov, k_gs_reduce, dim3((unsigned)(nv + 1)), dim3(kBlock), 0, s, (int64_t)m, nv, nb,
                 (const double*)work, state, 1);
      AFX_CHECK_LAUNCH();
      AFX_LAUNCH(kKKrylov, (k_gs_update<TB, TW>), dim3((unsigned)nb), dim3(kBlock), 0, s, n, (int64_t)m, nv,
                 (const TB*)V, ld, (const TW*)w, out, work, (const double*)state, 1);
      AFX_CHECK_LAUNCH();
      AFX_LAUNCH(kKKrylov, k_gs_finalize, dim3(1), dim3(kBlock), 0, s, (int64_t)m, j, nb, (const double*)work, state,
                 1, reorthogonalize);
      AFX_CHECK_LAUNCH();
      // pass 2 on the stored vector, when pass 1 asked for it
      AFX_LAUNCH(kKKrylov, (k_gs_dots<TB, TB>), dim3((unsigned)nb), dim3(kBlock), 0, s, n, nv, (const TB*)V, ld,
                 (const TB*)out, work, nb, (const double*)state, 2);
      AFX_CHECK_LAUNCH();
      AFX_LAUNCH(kKKrylov, k_gs_reduce, dim3((unsigned)nv), dim3(kBlock), 0, s, (int64_t)m, nv, nb,
                 (const double*)work, state, 2);
      AFX_CHECK_LAUNCH();
      AFX_LAUNCH(kKKrylov, (k_gs_update<TB, TB>), dim3((unsigned)nb), dim3(kBlock), 0, s, n, (int64_t)m, nv,
                 (const TB*)V, ld, (const TB*)out, out, work, (const double*)state, 2);
      AFX_CHECK_LAUNCH();
      AFX_LAUNCH(kKKrylov, k_gs_finalize, dim3(1), dim3(kBlock), 0, s, (int64_t)m, j, nb, (const double*)work, state,
                 2, reorthogonalize);
      AFX_CHECK_LAUNCH();
      AFX_LAUNCH(kKKrylov, k_gs_scale<TB>, dim3((unsigned)nb), dim3(kBlock), 0, s, n, out, (const double*)state);
      AFX_CHECK_LAUNCH();
      return 0;
    });
  });
}

extern "C" int afx_gmres_update(int64_t n, int32_t m, int precision, const void* basis, int64_t ld, double* state,
                                int x_precision, void* x, hipStream_t s) {
  AFX_REQUIRE(n > 0 && m > 0 && basis && state && x && ld >= n, "afx_gmres_update: bad argument");
  AFX_LAUNCH(kKKrylov, k_gmres_y, dim3(1), dim3(64), 0, s, (int64_t)m, state);
  AFX_CHECK_LAUNCH();
  const double* y = state + Lay{m}.y();
  return by_precision(precision, "afx_gmres_update", [&](auto tb) {
    using TB = decltype(tb);
    return by_precision(x_precision, "afx_gmres_update (x)", [&](auto tx) {
      using TX = decltype(tx);
      AFX_LAUNCH(kKKrylov, (k_basis_combine<TB, TX>), grid_for(n), dim3(kBlock), 0, s, n, m, (const double*)state,
                 (const TB*)basis, ld, y, (TX*)x);
      AFX_CHECK_LAUNCH();
      return 0;
    });
  });
}
