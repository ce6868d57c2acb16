// Compact point-block Jacobian of the coupled-flow residual, its SpMV and a point-block Jacobi
// preconditioner, for the inner solves of the Newton-Krylov driver.
//
// Storage is block-ELL, cell-major: row c has `nslot = 1 + max_faces_per_cell` block slots; slot 0
// is the diagonal block, slot 1 + k the coupling through the cell's k-th face (a zero block with
// column c on a boundary face).  Block (B x B, B = d + 1) entries are stored slot-major and
// cell-fastest, vals[(slot * B*B + r*B + k) * C + c], so consecutive threads (cells) read
// consecutive addresses for every entry.  Vectors are cell-major: x[c * B + v].
//
// The Jacobian is the reach-1 ("7-point" on hexes) approximation of dR/dx: each face flux is
// differentiated with respect to the owner's and neighbour's (u, p) with gradients, a_P and the
// limiter frozen -- exact for the direct dependence, omitting the 2-3 ring couplings carried by
// gradients / Rhie-Chow / the limiter.  The reference materialises the same operator by coloured
// JVP probing (aquaflux/solve/sparse_jacobian.py:338-927); here each block is obtained by seeding
// one variable at a time through the face-flux routine on `dual`.
#include <type_traits>

#include "afx_common.h"
#include "afx_mesh.h"
#include "afx_flow_face.h"
#include "afx_profile.h"

using namespace afx;

namespace {

struct JacArgs {
  const int8_t* kind;
  const double *x, *ub, *pb, *gB, *mu, *rho, *G, *aP, *psi;
  int32_t pin;
};

template <int D>
AFX_DEV void load_flow_cell(const JacArgs& a, const MeshView& m, int32_t c, bool psi, FlowCell<D, dual>& X) {
  constexpr int NV = D + 1;
#pragma unroll
  for (int i = 0; i < D; ++i) X.u[i] = dual{a.x[(int64_t)i * m.C + c], 0.0};
  X.p = dual{a.x[(int64_t)D * m.C + c], 0.0};
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) X.Gu[i][j] = dual{a.G[(int64_t)(i * D + j) * m.C + c], 0.0};
#pragma unroll
  for (int j = 0; j < D; ++j) X.Gp[j] = dual{a.G[(int64_t)((NV - 1) * D + j) * m.C + c], 0.0};
  X.aP = dual{a.aP[c], 0.0};
  X.mu = dual{a.mu[c], 0.0};
  X.rho = dual{a.rho[c], 0.0};
  X.V = m.vol[c];
#pragma unroll
  for (int i = 0; i < D; ++i) X.psi[i] = dual{psi ? a.psi[(int64_t)i * m.C + c] : 1.0, 0.0};
}

template <int D>
AFX_DEV void seed(FlowCell<D, dual>& X, int v, double t) {
  if (v < D)
    X.u[v].d = t;
  else
    X.p.d = t;
}

template <int D, int ADV, class T>
__global__ void __launch_bounds__(256) k_flow_jac(MeshView m, JacArgs a, int32_t nslot, T* __restrict__ vals,
                                                  int32_t* __restrict__ cols, int32_t* __restrict__ tslot) {
  constexpr int NV = D + 1, BB = NV * NV;
  const int32_t c = xcd_thread_index();
  if (c >= m.C) return;
  FlowCell<D, dual> X;
  load_flow_cell<D>(a, m, c, ADV == kAdvLimited, X);
  double diag[NV][NV];
#pragma unroll
  for (int r = 0; r < NV; ++r)
#pragma unroll
    for (int k = 0; k < NV; ++k) diag[r][k] = 0.0;
  const int32_t kb = m.cptr[c], ke = m.cptr[c + 1];
  for (int32_t k = kb; k < ke; ++k) {
    FaceRef r = decode_face(m, k, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D> q;
    load_face_geo<D>(m, r.f, r.P, r.N, q);
    const double sg = r.own ? 1.0 : -1.0;
    double off[NV][NV];
#pragma unroll
    for (int rr = 0; rr < NV; ++rr)
#pragma unroll
      for (int kk = 0; kk < NV; ++kk) off[rr][kk] = 0.0;
    if (o >= 0) {
      FlowCell<D, dual> O;
      load_flow_cell<D>(a, m, o, ADV == kAdvLimited, O);
      // fully unrolled so that `side` and `v` are constants: a runtime choice of X or O (or a
      // runtime seed index) would place both cells in per-lane scratch memory
#pragma unroll
      for (int side = 0; side < 2; ++side) {     // 0: seed own cell, 1: seed the other cell
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          FlowCell<D, dual>& S = side == 0 ? X : O;
          seed<D>(S, v, 1.0);
          dual out[NV];
          if (r.own)
            flow_flux_interior<D, ADV, dual>(q, X, O, out);
          else
            flow_flux_interior<D, ADV, dual>(q, O, X, out);
          seed<D>(S, v, 0.0);
#pragma unroll
          for (int rr = 0; rr < NV; ++rr) {
            if (side == 0)
              diag[rr][v] += sg * out[rr].d;
            else
              off[rr][v] += sg * out[rr].d;
          }
        }
      }
    } else {
      dual ub[D];
#pragma unroll
      for (int i = 0; i < D; ++i) ub[i] = dual{a.ub[(int64_t)i * m.F + r.fu], 0.0};
      const dual pb{a.pb[r.fu], 0.0};
      const dual gb = a.gB ? dual{a.gB[r.fu], 0.0} : X.mu;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        seed<D>(X, v, 1.0);
        dual out[NV];
        flow_flux_boundary<D, ADV, dual>(q, a.kind[r.fu], X, ub, pb, a.gB ? gb : X.mu, out);
        seed<D>(X, v, 0.0);
#pragma unroll
        for (int rr = 0; rr < NV; ++rr) diag[rr][v] += out[rr].d;
      }
    }
    if (c == a.pin) {
#pragma unroll
      for (int kk = 0; kk < NV; ++kk) off[NV - 1][kk] = 0.0;
    }
    const int32_t slot = 1 + (k - kb);
    cols[(int64_t)slot * m.C + c] = o >= 0 ? o : c;
    if (tslot) {  // the slot of this face in the other cell's row (for the transposed product)
      int32_t back = 0;
      if (o >= 0)
        for (int32_t kk = m.cptr[o]; kk < m.cptr[o + 1]; ++kk) {
          const int32_t e = m.cface[kk];
          if ((e >= 0 ? e : ~e) == r.f) back = 1 + (kk - m.cptr[o]);
        }
      tslot[(int64_t)slot * m.C + c] = back;
    }
#pragma unroll
    for (int rr = 0; rr < NV; ++rr)
#pragma unroll
      for (int kk = 0; kk < NV; ++kk) vals[((int64_t)slot * BB + rr * NV + kk) * m.C + c] = (T)off[rr][kk];
  }
  if (c == a.pin) {
#pragma unroll
    for (int kk = 0; kk < NV; ++kk) diag[NV - 1][kk] = kk == NV - 1 ? 1.0 : 0.0;
  }
  cols[c] = c;
  if (tslot) tslot[c] = 0;
#pragma unroll
  for (int rr = 0; rr < NV; ++rr)
#pragma unroll
    for (int kk = 0; kk < NV; ++kk) vals[((int64_t)rr * NV + kk) * m.C + c] = (T)diag[rr][kk];
  for (int32_t s = 1 + (ke - kb); s < nslot; ++s) {  // padding slots
    cols[(int64_t)s * m.C + c] = c;
    if (tslot) tslot[(int64_t)s * m.C + c] = 0;
#pragma unroll
    for (int e = 0; e < BB; ++e) vals[((int64_t)s * BB + e) * m.C + c] = (T)0.0;
  }
}

// y = A x (block-ELL, cell-major vectors).
template <int B, class T>
__global__ void __launch_bounds__(256) k_bell_spmv(int32_t C, int32_t nslot, const T* __restrict__ vals,
                                                   const int32_t* __restrict__ cols, const T* __restrict__ x,
                                                   T* __restrict__ y) {
  const int32_t c = xcd_thread_index();
  if (c >= C) return;
  T acc[B];
#pragma unroll
  for (int r = 0; r < B; ++r) acc[r] = (T)0;
  for (int32_t s = 0; s < nslot; ++s) {
    const int32_t col = cols[(int64_t)s * C + c];
    T xv[B];
#pragma unroll
    for (int k = 0; k < B; ++k) xv[k] = x[(int64_t)col * B + k];
    // the block values are read exactly once per product: non-temporal loads keep them from
    // evicting the gathered x rows from L2 (FP32 K5 product 194 -> 171 us)
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int k = 0; k < B; ++k)
        acc[r] += __builtin_nontemporal_load(&vals[((int64_t)s * B * B + r * B + k) * C + c]) * xv[k];
  }
#pragma unroll
  for (int r = 0; r < B; ++r) y[(int64_t)c * B + r] = acc[r];
}

// y = A^T x: row c of A^T gathers block (col, c) of A from the neighbour's row via `tslot`.
template <int B, class T>
__global__ void __launch_bounds__(256) k_bell_spmv_t(int32_t C, int32_t nslot, const T* __restrict__ vals,
                                                     const int32_t* __restrict__ cols,
                                                     const int32_t* __restrict__ tslot, const T* __restrict__ x,
                                                     T* __restrict__ y) {
  const int32_t c = xcd_thread_index();
  if (c >= C) return;
  T acc[B];
#pragma unroll
  for (int r = 0; r < B; ++r) acc[r] = (T)0;
  for (int32_t s = 0; s < nslot; ++s) {
    const int32_t col = cols[(int64_t)s * C + c];
    if (s > 0 && col == c) continue;  // boundary / padding slot
    const int32_t src = s == 0 ? 0 : tslot[(int64_t)s * C + c];
    T xv[B];
#pragma unroll
    for (int k = 0; k < B; ++k) xv[k] = x[(int64_t)col * B + k];
    // (A^T)_{c, col} = (A_{col, c})^T: entry [r][k] of the transpose is A_{col,c}[k][r]
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int k = 0; k < B; ++k) acc[r] += vals[((int64_t)src * B * B + k * B + r) * C + col] * xv[k];
  }
#pragma unroll
  for (int r = 0; r < B; ++r) y[(int64_t)c * B + r] = acc[r];
}

// Inverse of each diagonal block (Gauss-Jordan with partial pivoting, in FP64).
template <int B, class T>
__global__ void __launch_bounds__(256) k_pbj_setup(int32_t C, const T* __restrict__ vals, T* __restrict__ inv,
                                                   int32_t* __restrict__ singular) {
  const int32_t c = xcd_thread_index();
  if (c >= C) return;
  double a[B][2 * B];
#pragma unroll
  for (int r = 0; r < B; ++r)
#pragma unroll
    for (int k = 0; k < B; ++k) {
      a[r][k] = (double)vals[((int64_t)r * B + k) * C + c];
      a[r][B + k] = r == k ? 1.0 : 0.0;
    }
  bool bad = false;
#pragma unroll
  for (int col = 0; col < B; ++col) {
    int piv = col;
    double best = fabs(a[col][col]);
#pragma unroll
    for (int r = col + 1; r < B; ++r)
      if (fabs(a[r][col]) > best) {
        best = fabs(a[r][col]);
        piv = r;
      }
    if (best == 0.0) {
      bad = true;
      break;
    }
    // row swap over constant indices (a runtime row index would place `a` in scratch memory)
#pragma unroll
    for (int r = col + 1; r < B; ++r) {
      const bool sw = r == piv;
#pragma unroll
      for (int k = 0; k < 2 * B; ++k) {
        const double t = a[col][k];
        a[col][k] = sw ? a[r][k] : t;
        a[r][k] = sw ? t : a[r][k];
      }
    }
    const double ip = 1.0 / a[col][col];
#pragma unroll
    for (int k = 0; k < 2 * B; ++k) a[col][k] *= ip;
#pragma unroll
    for (int r = 0; r < B; ++r) {
      if (r == col) continue;
      const double f = a[r][col];
#pragma unroll
      for (int k = 0; k < 2 * B; ++k) a[r][k] -= f * a[col][k];
    }
  }
  if (bad) {
    singular[0] = 1;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int k = 0; k < B; ++k) inv[((int64_t)r * B + k) * C + c] = (T)(r == k ? 1.0 : 0.0);
    return;
  }
#pragma unroll
  for (int r = 0; r < B; ++r)
#pragma unroll
    for (int k = 0; k < B; ++k) inv[((int64_t)r * B + k) * C + c] = (T)a[r][B + k];
}

template <int B, class T>
__global__ void __launch_bounds__(256) k_pbj_apply(int32_t C, const T* __restrict__ inv, const T* __restrict__ x,
                                                   T* __restrict__ y) {
  const int32_t c = xcd_thread_index();
  if (c >= C) return;
  T xv[B];
#pragma unroll
  for (int k = 0; k < B; ++k) xv[k] = x[(int64_t)c * B + k];
#pragma unroll
  for (int r = 0; r < B; ++r) {
    T s = (T)0;
#pragma unroll
    for (int k = 0; k < B; ++k) s += inv[((int64_t)r * B + k) * C + c] * xv[k];
    y[(int64_t)c * B + r] = s;
  }
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// host entry points

extern "C" int afx_bell_slots(const afx_mesh* m, int32_t* nslot) {
  AFX_REQUIRE(m && nslot, "afx_bell_slots: null argument");
  *nslot = 1 + m->max_faces;
  return 0;
}

extern "C" int afx_flow_jacobian(const afx_mesh* m, const afx_geometry* g, const afx_flow_config* cfg,
                                 const afx_flow_inputs* in, int precision, void* vals, int32_t* cols,
                                 int32_t* tslot, double* fields, hipStream_t s) {
  AFX_REQUIRE(m && g && cfg && in && vals && cols && fields, "afx_flow_jacobian: null argument");
  AFX_REQUIRE(precision == 0 || precision == 1, "afx_flow_jacobian: precision must be 0 (fp32) or 1 (fp64)");
  const int D = m->dim, NV = D + 1;
  double* G = fields;
  double* aP = G + (int64_t)NV * D * m->C;
  double* psi = aP + m->C;
  int rc = afx_flow_fields(m, g, cfg, in, G, aP, psi, s);
  if (rc) return rc;
  JacArgs a{cfg->face_kind, in->state, in->u_b, in->p_b, in->gamma_b, in->mu, in->rho, G, aP, psi,
            cfg->pressure_pin >= 0 ? (int32_t)cfg->pressure_pin : -1};
  MeshView v;
  if (int rv_ = afx_rview(m, g, s, &v)) return rv_;
  const int32_t nslot = 1 + m->max_faces;
#define AFX_JAC(DD, AA, TT) \
  AFX_LAUNCH(kKBcsrAssemble, (k_flow_jac<DD, AA, TT>), grid_for(m->C), dim3(kBlock), 0, s, v, a, nslot, (TT*)vals, cols, tslot)
  const int adv = cfg->advection;
  if (D == 2) {
    if (precision == 0) {
      if (adv == 0) AFX_JAC(2, 0, float); else if (adv == 1) AFX_JAC(2, 1, float); else AFX_JAC(2, 2, float);
    } else {
      if (adv == 0) AFX_JAC(2, 0, double); else if (adv == 1) AFX_JAC(2, 1, double); else AFX_JAC(2, 2, double);
    }
  } else {
    if (precision == 0) {
      if (adv == 0) AFX_JAC(3, 0, float); else if (adv == 1) AFX_JAC(3, 1, float); else AFX_JAC(3, 2, float);
    } else {
      if (adv == 0) AFX_JAC(3, 0, double); else if (adv == 1) AFX_JAC(3, 1, double); else AFX_JAC(3, 2, double);
    }
  }
#undef AFX_JAC
  AFX_CHECK_LAUNCH();
  return 0;
}

template <class Fn>
static int by_block(int B, int precision, Fn&& fn) {
#define AFX_BB(BB)                                                                                    \
  if (B == BB && precision == 0) return fn(std::integral_constant<int, BB>{}, (float*)nullptr);       \
  if (B == BB && precision == 1) return fn(std::integral_constant<int, BB>{}, (double*)nullptr);
  AFX_BB(1) AFX_BB(2) AFX_BB(3) AFX_BB(4) AFX_BB(5) AFX_BB(6)
#undef AFX_BB
  afx_set_error("block size must be 1 to 6 and precision 0 or 1");
  return 2;
}

extern "C" int afx_bell_spmv(int64_t n_rows, int32_t block, int32_t nslot, int precision, const void* vals,
                             const int32_t* cols, const void* x, void* y, hipStream_t s) {
  AFX_REQUIRE(vals && cols && x && y && n_rows > 0 && nslot > 0, "afx_bell_spmv: bad argument");
  return by_block(block, precision, [&](auto Bc, auto* tp) {
    constexpr int B = decltype(Bc)::value;
    using T = std::remove_pointer_t<decltype(tp)>;
    AFX_LAUNCH(kKBcsrSpmv, (k_bell_spmv<B, T>), grid_for(n_rows), dim3(kBlock), 0, s, (int32_t)n_rows, nslot,
               (const T*)vals, cols, (const T*)x, (T*)y);
    AFX_CHECK_LAUNCH();
    return 0;
  });
}

extern "C" int afx_bell_spmv_t(int64_t n_rows, int32_t block, int32_t nslot, int precision, const void* vals,
                               const int32_t* cols, const int32_t* tslot, const void* x, void* y, hipStream_t s) {
  AFX_REQUIRE(vals && cols && tslot && x && y && n_rows > 0 && nslot > 0, "afx_bell_spmv_t: bad argument");
  return by_block(block, precision, [&](auto Bc, auto* tp) {
    constexpr int B = decltype(Bc)::value;
    using T = std::remove_pointer_t<decltype(tp)>;
    AFX_LAUNCH(kKBcsrSpmv, (k_bell_spmv_t<B, T>), grid_for(n_rows), dim3(kBlock), 0, s, (int32_t)n_rows, nslot,
               (const T*)vals, cols, tslot, (const T*)x, (T*)y);
    AFX_CHECK_LAUNCH();
    return 0;
  });
}

extern "C" int afx_pbj_setup(int64_t n_rows, int32_t block, int precision, const void* vals, void* inv,
                             int32_t* singular, hipStream_t s) {
  AFX_REQUIRE(vals && inv && singular && n_rows > 0, "afx_pbj_setup: bad argument");
  return by_block(block, precision, [&](auto Bc, auto* tp) {
    constexpr int B = decltype(Bc)::value;
    using T = std::remove_pointer_t<decltype(tp)>;
    AFX_LAUNCH(kKPbjSetup, (k_pbj_setup<B, T>), grid_for(n_rows), dim3(kBlock), 0, s, (int32_t)n_rows,
               (const T*)vals, (T*)inv, singular);
    AFX_CHECK_LAUNCH();
    return 0;
  });
}

extern "C" int afx_pbj_apply(int64_t n_rows, int32_t block, int precision, const void* inv, const void* x, void* y,
                             hipStream_t s) {
  AFX_REQUIRE(inv && x && y && n_rows > 0, "afx_pbj_apply: bad argument");
  return by_block(block, precision, [&](auto Bc, auto* tp) {
    constexpr int B = decltype(Bc)::value;
    using T = std::remove_pointer_t<decltype(tp)>;
    AFX_LAUNCH(kKPbjApply, (k_pbj_apply<B, T>), grid_for(n_rows), dim3(kBlock), 0, s, (int32_t)n_rows,
               (const T*)inv, (const T*)x, (T*)y);
    AFX_CHECK_LAUNCH();
    return 0;
  });
}
