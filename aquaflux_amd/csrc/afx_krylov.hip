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
#include <algorithm>
#include <type_traits>

#include "afx_common.h"
#include "afx_profile.h"
#include "aquaflux_amd.h"

using namespace afx;

namespace {

constexpr int kEPT = 8;                  // vector entries per thread of a Gram-Schmidt sweep
constexpr int kChunk = kBlock * kEPT;    // entries per workgroup of a sweep
constexpr int kMaxRows = 256;            // restart length limit (rows of the per-workgroup sums)
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

constexpr int kMaxSweepBlocks = 4096;  // partial sums per row of a sweep, at most

inline int64_t gs_chunks(int64_t n) { return (n + kChunk - 1) / kChunk; }

// VW consecutive entries, one 16-byte (or narrower) load when the address is VW-aligned.
template <class T, int VW>
struct alignas(sizeof(T) * VW) Vec {
  T v[VW];
};

// p[e0 .. e0 + VW) widened to FP64.  FULL: the whole range is below n and VW-aligned (one load);
// otherwise entries past n read as zero.
template <class T, int VW, bool FULL>
AFX_DEV void vload(const T* p, int64_t e0, int64_t n, double (&out)[VW]) {
  if constexpr (FULL) {
    const Vec<T, VW> q = *reinterpret_cast<const Vec<T, VW>*>(p + e0);
#pragma unroll
    for (int u = 0; u < VW; ++u) out[u] = (double)q.v[u];
  } else {
#pragma unroll
    for (int u = 0; u < VW; ++u) out[u] = e0 + u < n ? (double)p[e0 + u] : 0.0;
  }
}

template <class T, int VW, bool FULL>
AFX_DEV void vstore(T* p, int64_t e0, int64_t n, const double (&x)[VW]) {
  if constexpr (FULL) {
    Vec<T, VW> q;
#pragma unroll
    for (int u = 0; u < VW; ++u) q.v[u] = (T)x[u];
    *reinterpret_cast<Vec<T, VW>*>(p + e0) = q;
  } else {
#pragma unroll
    for (int u = 0; u < VW; ++u)
      if (e0 + u < n) p[e0 + u] = (T)x[u];
  }
}

AFX_DEV double wave_sum(double a) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
  return a;
}

// Orders this wave's LDS accesses around the transposed reductions (lanes of one wave exchange
// values through LDS; no other wave touches the region).
AFX_DEV void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sums eight values per lane over the wave's 64 lanes, for eight rows at once: each lane writes its
// a[0..7] to this wave's 8 x 64 LDS buffer, lane l then adds the eight entries 8 (l % 8) .. 8 (l % 8)
// + 7 of row l / 8 in order, and three butterflies over the groups of eight lanes finish the row.
// Lane l returns the sum of row l / 8 (a fixed order throughout).
AFX_DEV double wave_sum8(const double (&a)[8], double* __restrict__ buf) {
  const int lane = threadIdx.x & 63;
  wave_sync();
#pragma unroll
  for (int r = 0; r < 8; ++r) buf[r * 64 + lane] = a[r];
  wave_sync();
  const int r = lane >> 3, seg = lane & 7;
  double sum = 0.0;
#pragma unroll
  for (int i = 0; i < 8; ++i) sum += buf[r * 64 + seg * 8 + i];
  sum += __shfl_xor(sum, 1, 64);
  sum += __shfl_xor(sum, 2, 64);
  sum += __shfl_xor(sum, 4, 64);
  return sum;
}

// Sum over the workgroup of NV values per thread, in a fixed order (wave butterflies, then the four
// wave sums in order).  Thread 0 holds the result.
template <int NV>
AFX_DEV void block_sum_n(double (&a)[NV]) {
  __shared__ double lds[NV][kBlock / 64];
#pragma unroll
  for (int u = 0; u < NV; ++u) a[u] = wave_sum(a[u]);
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

// Per-row wave sums collected in LDS (red[row][wave]), written out as the workgroup's partial sums
// partial[row * nb + blockIdx.x] after a single barrier.
AFX_DEV void write_row_partials(double (*red)[kBlock / 64], int32_t nt, double* __restrict__ partial, int64_t nb) {
  __syncthreads();
  for (int32_t t = threadIdx.x; t < nt; t += kBlock)
    partial[(int64_t)t * nb + blockIdx.x] = ((red[t][0] + red[t][1]) + red[t][2]) + red[t][3];
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
__global__ void __launch_bounds__(256) k_norm_partial(int64_t n, int64_t cpb, const TR* __restrict__ r,
                                                      double* __restrict__ partial) {
  double a[1] = {0.0};
  const int64_t nchunks = (n + kChunk - 1) / kChunk;
  const int64_t c1 = std::min<int64_t>(nchunks, ((int64_t)blockIdx.x + 1) * cpb);
  for (int64_t c = (int64_t)blockIdx.x * cpb; c < c1; ++c) {
    const int64_t base = c * kChunk + threadIdx.x;
#pragma unroll
    for (int s = 0; s < kEPT; ++s) {
      const int64_t e = base + (int64_t)s * kBlock;
      if (e < n) {
        const double v = (double)r[e];
        a[0] += v * v;
      }
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

// ---- Gram-Schmidt sweeps ------------------------------------------------------------------------

// Zeroes the per-row wave sums red[0 .. nt)[wave] before a grid-stride sweep.
AFX_DEV void clear_rows(double (*red)[kBlock / 64], int32_t nt) {
  for (int32_t t = threadIdx.x; t < nt * (kBlock / 64); t += kBlock) red[t / (kBlock / 64)][t % (kBlock / 64)] = 0.0;
  __syncthreads();
}

// One chunk of kChunk entries: adds to red[v][wave] the wave's sums of V_v . w for v < nv (and, in
// pass 1, w . w for v = nv).  w is read as stored in the basis precision; each thread holds 8 of its
// entries (VW per load).
template <class TB, class TW, int VW, bool FULL>
AFX_DEV void gs_dots_chunk(int64_t c, int64_t n, int32_t nv, int32_t nt, const TB* __restrict__ basis, int64_t ld,
                           const TW* __restrict__ w, double (*red)[kBlock / 64], double* __restrict__ buf) {
  constexpr int S = kEPT / VW;
  const int64_t cb = c * kChunk;
  double wv[S][VW];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    vload<TW, VW, FULL>(w, cb + ((int64_t)s * kBlock + threadIdx.x) * VW, n, wv[s]);
#pragma unroll
    for (int q = 0; q < VW; ++q) wv[s][q] = (double)(TB)wv[s][q];
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int32_t v0 = 0; v0 < nt; v0 += 8) {
    double a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a[u] = 0.0;
      const int32_t v = v0 + u;
      if (v < nv) {
        const TB* row = basis + (int64_t)v * ld;
#pragma unroll
        for (int s = 0; s < S; ++s) {
          double r[VW];
          vload<TB, VW, FULL>(row, cb + ((int64_t)s * kBlock + threadIdx.x) * VW, n, r);
#pragma unroll
          for (int q = 0; q < VW; ++q) a[u] += r[q] * wv[s][q];
        }
      } else if (v == nv) {
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
          for (int q = 0; q < VW; ++q) a[u] += wv[s][q] * wv[s][q];
      }
    }
    const double rs = wave_sum8(a, buf);
    if ((lane & 7) == 0 && v0 + (lane >> 3) < nt) red[v0 + (lane >> 3)][wave] += rs;
  }
}

// partial[v * gridDim.x + b] = workgroup b's sum of V_v . w over its chunks (b cpb, ..., b cpb + cpb
// - 1) for v < nv; in pass 1 also row nv = w . w.
template <class TB, class TW, int VW>
__global__ void __launch_bounds__(256) k_gs_dots(int64_t n, int64_t cpb, int32_t nv, const TB* __restrict__ basis,
                                                 int64_t ld, const TW* __restrict__ w, double* __restrict__ partial,
                                                 const double* __restrict__ st, int pass) {
  if (gated(st, pass)) return;
  __shared__ double red[kMaxRows + 8][kBlock / 64];
  __shared__ double wbuf[kBlock / 64 * 8 * 64];
  const int32_t nt = pass == 1 ? nv + 1 : nv;
  clear_rows(red, nt);
  double* buf = wbuf + (threadIdx.x >> 6) * 8 * 64;
  const int64_t nchunks = (n + kChunk - 1) / kChunk;
  const int64_t c1 = std::min<int64_t>(nchunks, ((int64_t)blockIdx.x + 1) * cpb);
  for (int64_t c = (int64_t)blockIdx.x * cpb; c < c1; ++c) {
    if ((c + 1) * kChunk <= n)
      gs_dots_chunk<TB, TW, VW, true>(c, n, nv, nt, basis, ld, w, red, buf);
    else
      gs_dots_chunk<TB, TW, VW, false>(c, n, nv, nt, basis, ld, w, red, buf);
  }
  write_row_partials(red, nt, partial, gridDim.x);
}

// Workgroup v sums partial[v * nb ..] in a fixed order into hp[v] (pass 1: also hs[v] = hp[v], and
// row nv = w0^2 into hp[m + 1]; pass 2: hs[v] += hp[v]).
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

// One chunk: out = w - sum_v hp[v] V_v (FP64, v in order), stored in the basis precision; returns the
// thread's sum of the stored entries squared.  out may alias w (pass 2).
template <class TB, class TW, int VW, bool FULL>
AFX_DEV double gs_update_chunk(int64_t c, int64_t n, int32_t nv, const TB* __restrict__ basis, int64_t ld, const TW* w,
                               TB* out, const double* __restrict__ hp) {
  constexpr int S = kEPT / VW;
  const int64_t cb = c * kChunk;
  double acc[S][VW];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    vload<TW, VW, FULL>(w, cb + ((int64_t)s * kBlock + threadIdx.x) * VW, n, acc[s]);
#pragma unroll
    for (int q = 0; q < VW; ++q) acc[s][q] = (double)(TB)acc[s][q];
  }
  for (int32_t v = 0; v < nv; ++v) {
    const double h = hp[v];
    const TB* row = basis + (int64_t)v * ld;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      double r[VW];
      vload<TB, VW, FULL>(row, cb + ((int64_t)s * kBlock + threadIdx.x) * VW, n, r);
#pragma unroll
      for (int q = 0; q < VW; ++q) acc[s][q] -= h * r[q];
    }
  }
  double a = 0.0;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int64_t e0 = cb + ((int64_t)s * kBlock + threadIdx.x) * VW;
#pragma unroll
    for (int q = 0; q < VW; ++q) {
      acc[s][q] = (double)(TB)acc[s][q];
      if (FULL || e0 + q < n) a += acc[s][q] * acc[s][q];
    }
    vstore<TB, VW, FULL>(out, e0, n, acc[s]);
  }
  return a;
}

// out = w - sum_v hp[v] V_v over the workgroup's chunks (b cpb, ..., b cpb + cpb - 1); partial[b] =
// its sum of the stored entries squared.
template <class TB, class TW, int VW>
__global__ void __launch_bounds__(256) k_gs_update(int64_t n, int64_t cpb, int64_t m, int32_t nv,
                                                   const TB* __restrict__ basis, int64_t ld, const TW* w, TB* out,
                                                   double* __restrict__ partial, const double* __restrict__ st,
                                                   int pass) {
  if (gated(st, pass)) return;
  const double* __restrict__ hp = st + Lay{m}.hp();
  double a[1] = {0.0};
  const int64_t nchunks = (n + kChunk - 1) / kChunk;
  const int64_t c1 = std::min<int64_t>(nchunks, ((int64_t)blockIdx.x + 1) * cpb);
  for (int64_t c = (int64_t)blockIdx.x * cpb; c < c1; ++c) {
    if ((c + 1) * kChunk <= n)
      a[0] += gs_update_chunk<TB, TW, VW, true>(c, n, nv, basis, ld, w, out, hp);
    else
      a[0] += gs_update_chunk<TB, TW, VW, false>(c, n, nv, basis, ld, w, out, hp);
  }
  block_sum_n<1>(a);
  if (threadIdx.x == 0) partial[blockIdx.x] = a[0];
}

// Pass 1's update fused with pass 2's dot products; workgroup b takes entries [b kBlock VW, (b + 1)
// kBlock VW): out = w - sum_v hp[v] V_v as in k_gs_update, with the nv basis entries of each of the
// thread's VW entries held in registers (in the basis precision), then partial[v * gridDim.x + b] =
// the workgroup's sums of V_v . out for v < nv and of out . out for v = nv.  One sweep over the basis
// instead of two.
template <class TB, class TW, int VW, int NVB, bool FULL>
AFX_DEV void gs_update_dots(int64_t n, int32_t nv, const TB* __restrict__ basis, int64_t ld, const TW* __restrict__ w,
                            TB* __restrict__ out, const double* __restrict__ hp, double (*red)[kBlock / 64],
                            double* __restrict__ buf) {
  const int64_t e0 = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * VW;
  double x[VW];
  vload<TW, VW, FULL>(w, e0, n, x);
#pragma unroll
  for (int q = 0; q < VW; ++q) x[q] = (double)(TB)x[q];
  // every row slot is loaded unconditionally (slots past nv repeat row nv - 1, a cache hit, and are
  // never used), so that all NVB loads are in flight together
  Vec<TB, VW> rows[NVB];
#pragma unroll
  for (int v = 0; v < NVB; ++v) {
    const TB* row = basis + (int64_t)(v < nv ? v : nv - 1) * ld;
    if constexpr (FULL) {
      rows[v] = *reinterpret_cast<const Vec<TB, VW>*>(row + e0);
    } else {
#pragma unroll
      for (int q = 0; q < VW; ++q) rows[v].v[q] = e0 + q < n ? row[e0 + q] : (TB)0;
    }
  }
#pragma unroll
  for (int v = 0; v < NVB; ++v) {
    if (v < nv) {
      const double h = hp[v];
#pragma unroll
      for (int q = 0; q < VW; ++q) x[q] -= h * (double)rows[v].v[q];
    }
  }
#pragma unroll
  for (int q = 0; q < VW; ++q) x[q] = (double)(TB)x[q];
  vstore<TB, VW, FULL>(out, e0, n, x);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int v0 = 0; v0 < NVB + 1; v0 += 8) {
    if (v0 <= nv) {
      double a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        a[u] = 0.0;
        const int v = v0 + u;
        if (v < NVB && v < nv) {
#pragma unroll
          for (int q = 0; q < VW; ++q) a[u] += (double)rows[v < NVB ? v : 0].v[q] * x[q];
        }
        if (v == nv) {
#pragma unroll
          for (int q = 0; q < VW; ++q) a[u] += x[q] * x[q];
        }
      }
      const double rs = wave_sum8(a, buf);
      if ((lane & 7) == 0 && v0 + (lane >> 3) <= nv) red[v0 + (lane >> 3)][wave] = rs;
    }
  }
}

template <class TB, class TW, int VW, int NVB>
__global__ void __launch_bounds__(256) k_gs_update_dots(int64_t n, int64_t m, int32_t nv, const TB* __restrict__ basis,
                                                        int64_t ld, const TW* __restrict__ w, TB* __restrict__ out,
                                                        double* __restrict__ partial, const double* __restrict__ st) {
  if (gated(st, 1)) return;
  __shared__ double red[NVB + 8][kBlock / 64];
  __shared__ double wbuf[kBlock / 64 * 8 * 64];
  const double* __restrict__ hp = st + Lay{m}.hp();
  double* buf = wbuf + (threadIdx.x >> 6) * 8 * 64;
  if (((int64_t)blockIdx.x + 1) * kBlock * VW <= n)
    gs_update_dots<TB, TW, VW, NVB, true>(n, nv, basis, ld, w, out, hp, red, buf);
  else
    gs_update_dots<TB, TW, VW, NVB, false>(n, nv, basis, ld, w, out, hp, red, buf);
  write_row_partials(red, nv + 1, partial, gridDim.x);
}

// folded[r * nbo + i] = sum_{k < F} partial[r * nb + i F + k] (k in order, entries past nb are zero):
// shortens rows of partial sums before the single-workgroup reductions.
__global__ void __launch_bounds__(256) k_fold_partials(int32_t nrows, int64_t nb, int64_t nbo, int32_t F,
                                                       const double* __restrict__ partial, double* __restrict__ folded,
                                                       const double* __restrict__ st, int pass) {
  if (gated(st, pass)) return;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)nrows * nbo) return;
  const int64_t r = t / nbo, i = t % nbo;
  const double* src = partial + r * nb + i * F;
  const int64_t cnt = std::min<int64_t>(F, nb - i * F);
  double s = 0.0;
  for (int64_t k = 0; k < cnt; ++k) s += src[k];
  folded[t] = s;
}

// After a pass over step j: hn = ||w|| from partial row `row` (nb sums).  Pass 1 decides whether a
// second pass is needed; the last pass of the step writes the Hessenberg column, applies the
// earlier rotations, forms the new one, updates g and the residual estimate |g_{j+1}|, and sets
// done when it reaches tol or on breakdown (hn = 0).  The column and rotations are staged in LDS.
__global__ void __launch_bounds__(256) k_gs_finalize(int64_t m, int32_t j, int64_t nb, const double* __restrict__ partial,
                                                     double* __restrict__ st, int pass, int always) {
  if (gated(st, pass)) return;
  __shared__ double col[kMaxRows + 1], cs[kMaxRows], sn[kMaxRows];
  __shared__ int stop;
  const Lay L{m};
  const double s = block_fixed_sum(partial, nb);
  if (threadIdx.x == 0) {
    const double hn = sqrt(s);
    stop = 0;
    if (pass == 1) {
      const double w0 = sqrt(st[L.hp() + m + 1]);
      if (always || !(hn > kDgksEta * w0)) {
        st[kSPass2] = 1.0;
        stop = 1;
      }
    } else {
      st[kSPass2] = 0.0;
    }
    col[j + 1] = hn;
  }
  __syncthreads();
  if (stop) return;
  for (int32_t i = threadIdx.x; i <= j; i += kBlock) {
    col[i] = st[L.hs() + i];
    if (i < j) {
      cs[i] = st[L.cs() + i];
      sn[i] = st[L.sn() + i];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double hn = col[j + 1];
    for (int32_t i = 0; i < j; ++i) {
      const double t = cs[i] * col[i] + sn[i] * col[i + 1];
      col[i + 1] = -sn[i] * col[i] + cs[i] * col[i + 1];
      col[i] = t;
    }
    const double den = hypot(col[j], col[j + 1]);
    double c, sj;
    if (den == 0.0) {
      c = 1.0;
      sj = 0.0;
    } else {
      c = col[j] / den;
      sj = col[j + 1] / den;
    }
    st[L.cs() + j] = c;
    st[L.sn() + j] = sj;
    col[j] = den;
    col[j + 1] = 0.0;
    double* g = st + L.g();
    const double gj = g[j];
    g[j + 1] = -sj * gj;
    g[j] = c * gj;
    const double res = fabs(-sj * gj);
    st[kSK] = (double)(j + 1);
    st[kSRes] = res;
    st[L.resid() + j] = res;
    st[kSHn] = hn;
    if (hn == 0.0) st[kSBreak] = 1.0;
    if (res <= st[kSTol] || hn == 0.0) st[kSDone] = 1.0;
  }
  __syncthreads();
  double* H = st + L.H() + (int64_t)j * (m + 1);
  for (int32_t i = threadIdx.x; i <= j + 1; i += kBlock) H[i] = col[i];
}

// V_{j+1} = w / hn (skipped once the cycle is done: that vector is never used).
template <class TB, int VW, bool FULL>
AFX_DEV void gs_scale(int64_t n, TB* __restrict__ out, double hn) {
  constexpr int S = kEPT / VW;
  const int64_t cb = (int64_t)blockIdx.x * kChunk;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int64_t e0 = cb + ((int64_t)s * kBlock + threadIdx.x) * VW;
    double x[VW];
    vload<TB, VW, FULL>(out, e0, n, x);
#pragma unroll
    for (int q = 0; q < VW; ++q) x[q] /= hn;
    vstore<TB, VW, FULL>(out, e0, n, x);
  }
}

template <class TB, int VW>
__global__ void __launch_bounds__(256) k_gs_scale(int64_t n, TB* __restrict__ out, const double* __restrict__ st) {
  if (st[kSDone] != 0.0) return;
  const double hn = st[kSHn];
  if (((int64_t)blockIdx.x + 1) * kChunk <= n)
    gs_scale<TB, VW, true>(n, out, hn);
  else
    gs_scale<TB, VW, false>(n, out, hn);
}

// ---- end of cycle: y = R^-1 g (k from the state), x += basis[:k]^T y ----------------------------

// Back substitution on the k x k upper triangle held in LDS (a zero diagonal entry is taken as 1,
// the guard of the host solve it replaces): y_i = (g_i - sum_{l > i} R_il y_l) / R_ii, the sum
// taken over l = i + 1, ..., k - 1 in order.
__global__ void __launch_bounds__(256) k_gmres_y(int64_t m, double* __restrict__ st) {
  constexpr int kLdsRows = 64;  // triangles up to 64 x 64 are staged (32 KB); larger ones read the state
  __shared__ double R[kLdsRows * kLdsRows];
  __shared__ double y[kMaxRows], g[kMaxRows];
  const Lay L{m};
  const int32_t k = (int32_t)st[kSK];
  const bool lds = k <= kLdsRows;
  if (lds) {
    for (int32_t t = threadIdx.x; t < k * k; t += kBlock) {
      const int32_t i = t % k, l = t / k;
      R[t] = st[L.H() + (int64_t)l * (m + 1) + i];
    }
  }
  for (int32_t t = threadIdx.x; t < k; t += kBlock) g[t] = st[L.g() + t];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int32_t i = k - 1; i >= 0; --i) {
      double s = g[i];
      for (int32_t l = i + 1; l < k; ++l)
        s -= (lds ? R[l * k + i] : st[L.H() + (int64_t)l * (m + 1) + i]) * y[l];
      const double d = lds ? R[i * k + i] : st[L.H() + (int64_t)i * (m + 1) + i];
      y[i] = s / (d == 0.0 ? 1.0 : d);
    }
  }
  __syncthreads();
  for (int32_t t = threadIdx.x; t < k; t += kBlock) st[L.y() + t] = y[t];
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

namespace {

// Launch shape of a sweep over nchunks chunks: each workgroup takes cpb consecutive chunks, with cpb
// the smallest count that keeps the workgroups (the rows of partial sums) to kMaxSweepBlocks.
struct Sweep {
  int64_t grid, cpb;
};

Sweep sweep_grid(int64_t nchunks) {
  const int64_t cpb = std::max<int64_t>(1, (nchunks + kMaxSweepBlocks - 1) / kMaxSweepBlocks);
  return Sweep{(nchunks + cpb - 1) / cpb, cpb};
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
  AFX_REQUIRE(m < kMaxRows, "afx_gmres_sizes: restart length must be below 256");
  *state_doubles = Lay{m}.total();
  // partial sums: one row of up to max(kMaxSweepBlocks, one per kBlock entries) sums per basis vector,
  // then the folded rows
  *work_doubles = (std::max<int64_t>((n + kBlock - 1) / kBlock, kMaxSweepBlocks) + kMaxSweepBlocks) * (m + 2);
  return 0;
}

extern "C" int afx_gmres_start(int64_t n, int32_t m, int precision, int r_precision, const void* r, double rtol,
                               double atol, int first, void* basis, int64_t ld, double* state, double* work,
                               hipStream_t s) {
  AFX_REQUIRE(n > 0 && m > 0 && m < kMaxRows && r && basis && state && work && ld >= n,
              "afx_gmres_start: bad argument");
  const int64_t nc = gs_chunks(n);
  return by_precision(precision, "afx_gmres_start", [&](auto tb) {
    using TB = decltype(tb);
    return by_precision(r_precision, "afx_gmres_start (r)", [&](auto tr) {
      using TR = decltype(tr);
      const Sweep sw = sweep_grid(nc);
      const int64_t nb = sw.grid;
      AFX_LAUNCH(kKKrylov, k_norm_partial<TR>, dim3((unsigned)nb), dim3(kBlock), 0, s, n, sw.cpb, (const TR*)r, work);
      AFX_CHECK_LAUNCH();
      AFX_LAUNCH(kKKrylov, k_start_finalize, dim3(1), dim3(kBlock), 0, s, (int64_t)m, nb, (const double*)work, state,
                 rtol, atol, first);
      AFX_CHECK_LAUNCH();
      AFX_LAUNCH(kKKrylov, (k_start_scale<TB, TR>), dim3((unsigned)nc), dim3(kBlock), 0, s, n, (const TR*)r,
                 (TB*)basis, (const double*)state);
      AFX_CHECK_LAUNCH();
      return 0;
    });
  });
}

namespace {

inline bool aligned_to(const void* p, size_t a) { return ((uintptr_t)p % a) == 0; }

// The sweeps of one Arnoldi step with VW basis entries per load (VW = 1: any alignment).
template <class TB, class TW, int VW>
int arnoldi_sweeps(int64_t n, int32_t m, int32_t j, const TW* w, TB* V, int64_t ld, int reo, double* state,
                   double* work, hipStream_t s) {
  const int64_t nc = gs_chunks(n);
  const int32_t nv = j + 1;
  TB* out = V + (int64_t)(j + 1) * ld;
  // pass 1 dot products and ||w||^2
  const Sweep swd = sweep_grid(nc);
  const int64_t nbd = swd.grid;
  AFX_LAUNCH(kKKrylov, (k_gs_dots<TB, TW, VW>), dim3((unsigned)nbd), dim3(kBlock), 0, s, n, swd.cpb, nv, (const TB*)V,
             ld, w, work, (const double*)state, 1);
  AFX_CHECK_LAUNCH();
  AFX_LAUNCH(kKKrylov, k_gs_reduce, dim3((unsigned)(nv + 1)), dim3(kBlock), 0, s, (int64_t)m, nv, nbd,
             (const double*)work, state, 1);
  AFX_CHECK_LAUNCH();
  // pass 1 update with pass 2's dot products in the same sweep (the basis held in registers, up to 64
  // vectors)
  auto fused = [&](auto nvb_tag, auto vwf_tag) {
    constexpr int NVB = decltype(nvb_tag)::value, VWF = decltype(vwf_tag)::value;
    const int64_t cw = (int64_t)kBlock * VWF;
    const int64_t nbf = (n + cw - 1) / cw;  // one workgroup per kBlock x VWF entries
    AFX_LAUNCH(kKKrylov, (k_gs_update_dots<TB, TW, VWF, NVB>), dim3((unsigned)nbf), dim3(kBlock), 0, s, n, (int64_t)m,
               nv, (const TB*)V, ld, w, out, work, (const double*)state);
    AFX_CHECK_LAUNCH();
    // rows of more than kMaxSweepBlocks sums are folded first (into the work area after the partials)
    const double* rows = work;
    int64_t nr = nbf;
    if (nbf > kMaxSweepBlocks) {
      const int32_t F = (int32_t)((nbf + kMaxSweepBlocks - 1) / kMaxSweepBlocks);
      nr = (nbf + F - 1) / F;
      double* folded = work + (int64_t)(m + 2) * nbf;
      AFX_LAUNCH(kKKrylov, k_fold_partials, grid_for((int64_t)(nv + 1) * nr), dim3(kBlock), 0, s, nv + 1, nbf, nr, F,
                 (const double*)work, folded, (const double*)state, 1);
      AFX_CHECK_LAUNCH();
      rows = folded;
    }
    // ||out||^2 is row nv of the fused sums
    AFX_LAUNCH(kKKrylov, k_gs_finalize, dim3(1), dim3(kBlock), 0, s, (int64_t)m, j, nr, rows + (int64_t)nv * nr, state,
               1, reo);
    AFX_CHECK_LAUNCH();
    AFX_LAUNCH(kKKrylov, k_gs_reduce, dim3((unsigned)nv), dim3(kBlock), 0, s, (int64_t)m, nv, nr, rows, state, 2);
    AFX_CHECK_LAUNCH();
    return 0;
  };
  using I8 = std::integral_constant<int, 8>;
  using I16 = std::integral_constant<int, 16>;
  using I24 = std::integral_constant<int, 24>;
  using I32 = std::integral_constant<int, 32>;
  using I48 = std::integral_constant<int, 48>;
  using I64 = std::integral_constant<int, 64>;
  // entries per thread of the fused sweep: 16 bytes per basis vector, at most 256 bytes (64 VGPRs) of
  // basis entries per thread
  auto vwf = [](int nvb) { return std::max(1, std::min((int)(16 / sizeof(TB)), (int)(256 / (nvb * sizeof(TB))))); };
  using W8 = std::integral_constant<int, VW == 1 ? 1 : vwf(8)>;
  using W16 = std::integral_constant<int, VW == 1 ? 1 : vwf(16)>;
  using W24 = std::integral_constant<int, VW == 1 ? 1 : vwf(24)>;
  using W32 = std::integral_constant<int, VW == 1 ? 1 : vwf(32)>;
  using W48 = std::integral_constant<int, VW == 1 ? 1 : vwf(48)>;
  using W64 = std::integral_constant<int, VW == 1 ? 1 : vwf(64)>;
  const Sweep swu = sweep_grid(nc);
  const int64_t nbu = swu.grid;
  int rc = 0;
  if (nv <= 8)
    rc = fused(I8{}, W8{});
  else if (nv <= 16)
    rc = fused(I16{}, W16{});
  else if (nv <= 24)
    rc = fused(I24{}, W24{});
  else if (nv <= 32)
    rc = fused(I32{}, W32{});
  else if (nv <= 48)
    rc = fused(I48{}, W48{});
  else if (nv <= 64)
    rc = fused(I64{}, W64{});
  else {
    // longer bases: pass 1 update, then pass 2 dot products as a sweep of their own
    const Sweep sw1 = sweep_grid(nc);
    const int64_t nb1 = sw1.grid;
    AFX_LAUNCH(kKKrylov, (k_gs_update<TB, TW, VW>), dim3((unsigned)nb1), dim3(kBlock), 0, s, n, sw1.cpb, (int64_t)m,
               nv, (const TB*)V, ld, w, out, work, (const double*)state, 1);
    AFX_CHECK_LAUNCH();
    AFX_LAUNCH(kKKrylov, k_gs_finalize, dim3(1), dim3(kBlock), 0, s, (int64_t)m, j, nb1, (const double*)work, state, 1,
               reo);
    AFX_CHECK_LAUNCH();
    const Sweep sw2 = sweep_grid(nc);
    const int64_t nb2 = sw2.grid;
    AFX_LAUNCH(kKKrylov, (k_gs_dots<TB, TB, VW>), dim3((unsigned)nb2), dim3(kBlock), 0, s, n, sw2.cpb, nv, (const TB*)V,
               ld, (const TB*)out, work, (const double*)state, 2);
    AFX_CHECK_LAUNCH();
    AFX_LAUNCH(kKKrylov, k_gs_reduce, dim3((unsigned)nv), dim3(kBlock), 0, s, (int64_t)m, nv, nb2, (const double*)work,
               state, 2);
    AFX_CHECK_LAUNCH();
  }
  if (rc) return rc;
  // pass 2 update, when pass 1 asked for it
  AFX_LAUNCH(kKKrylov, (k_gs_update<TB, TB, VW>), dim3((unsigned)nbu), dim3(kBlock), 0, s, n, swu.cpb, (int64_t)m, nv,
             (const TB*)V, ld, (const TB*)out, out, work, (const double*)state, 2);
  AFX_CHECK_LAUNCH();
  AFX_LAUNCH(kKKrylov, k_gs_finalize, dim3(1), dim3(kBlock), 0, s, (int64_t)m, j, nbu, (const double*)work, state, 2,
             reo);
  AFX_CHECK_LAUNCH();
  AFX_LAUNCH(kKKrylov, (k_gs_scale<TB, VW>), dim3((unsigned)nc), dim3(kBlock), 0, s, n, out, (const double*)state);
  AFX_CHECK_LAUNCH();
  return 0;
}

}  // namespace

extern "C" int afx_gmres_arnoldi(int64_t n, int32_t m, int32_t j, int precision, int w_precision, const void* w,
                                 void* basis, int64_t ld, int reorthogonalize, double* state, double* work,
                                 hipStream_t s) {
  AFX_REQUIRE(n > 0 && m > 0 && m < kMaxRows && j >= 0 && j < m && w && basis && state && work && ld >= n,
              "afx_gmres_arnoldi: bad argument");
  AFX_REQUIRE(reorthogonalize == 0 || reorthogonalize == 1, "afx_gmres_arnoldi: reorthogonalize must be 0 or 1");
  return by_precision(precision, "afx_gmres_arnoldi", [&](auto tb) {
    using TB = decltype(tb);
    return by_precision(w_precision, "afx_gmres_arnoldi (w)", [&](auto tw) {
      using TW = decltype(tw);
      constexpr int VW = (int)(16 / sizeof(TB));
      // 16-byte loads need every row and w aligned to VW entries
      const bool vec = ld % VW == 0 && aligned_to(basis, 16) && aligned_to(w, sizeof(TW) * VW);
      if (vec) return arnoldi_sweeps<TB, TW, VW>(n, m, j, (const TW*)w, (TB*)basis, ld, reorthogonalize, state, work, s);
      return arnoldi_sweeps<TB, TW, 1>(n, m, j, (const TW*)w, (TB*)basis, ld, reorthogonalize, state, work, s);
    });
  });
}

extern "C" int afx_gmres_update(int64_t n, int32_t m, int precision, const void* basis, int64_t ld, double* state,
                                int x_precision, void* x, hipStream_t s) {
  AFX_REQUIRE(n > 0 && m > 0 && m < kMaxRows && basis && state && x && ld >= n, "afx_gmres_update: bad argument");
  AFX_LAUNCH(kKKrylov, k_gmres_y, dim3(1), dim3(kBlock), 0, s, (int64_t)m, state);
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
