// Coupled (u, v[, w], p) residual R(state), its tangent (JVP) and its adjoint (VJP).
//
// Three cell-centric passes, one thread per cell walking its face list (fixed order, no atomics):
//   G  gradients of u_i and p (compact Green-Gauss or least squares) + Venkatakrishnan limiter
//   D  momentum diagonal a_P (flow/momentum.py:596-635, flow/rhie_chow.py:81-175)
//   R  face fluxes (diffusion, pressure, advection, Rhie-Chow mass flux) summed per cell
// An interior face is evaluated by both of its cells with identical inputs, so the owner's +flux
// and the neighbour's -flux are the same number: conservation holds and no face array is stored.
// The JVP instantiates the same passes on `dual`.  The VJP recomputes G and D, then runs the
// hand-written reverse passes R', D', G' (each again one thread per cell, gather form).
#include <type_traits>

#include "afx_common.h"
#include "afx_mesh.h"
#include "afx_flow_face.h"
#include "afx_profile.h"
#include "afx_cgg.h"

using namespace afx;

namespace {

template <class S>
struct FlowArgs {
  const int8_t* kind;
  In<S> x, ub, pb, gB, mu, rho;  // state [(D+1)][C], ub [D][F], pb [F], gB [F] (optional)
  In<S> G, aP, psi;              // pass products (read in later passes)
  int32_t pin;
  double pin_value;
  double bf[3];
  double lim_k;
  GeoTan gt;  // geometry tangents (read only by the GEO instantiations of the forward passes)
};

template <int D, class S, class GT>
AFX_DEV void load_state(const FlowArgs<S>& a, int32_t C, int32_t c, FlowCell<D, S, GT>& X) {
#pragma unroll
  for (int i = 0; i < D; ++i) X.u[i] = a.x[(int64_t)i * C + c];
  X.p = a.x[(int64_t)D * C + c];
}

template <int D, class S, class GT>
AFX_DEV void load_grad(const FlowArgs<S>& a, int32_t C, int32_t c, FlowCell<D, S, GT>& X) {
  constexpr int NV = D + 1;
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) X.Gu[i][j] = a.G[(int64_t)(i * D + j) * C + c];
#pragma unroll
  for (int j = 0; j < D; ++j) X.Gp[j] = a.G[(int64_t)((NV - 1) * D + j) * C + c];
}

template <int D, class S, bool GEO = false, class GT = typename GeoRead<D, GEO>::T>
AFX_DEV void load_cell_full(const FlowArgs<S>& a, const MeshView& m, int32_t c, bool need_ap, bool need_psi,
                            FlowCell<D, S, GT>& X) {
  load_state<D, S>(a, m.C, c, X);
  load_grad<D, S>(a, m.C, c, X);
  X.mu = a.mu[c];
  X.rho = a.rho[c];
  X.V = GeoRead<D, GEO>::vol(m, a.gt, c);
  X.aP = need_ap ? a.aP[c] : sconst<S>(1.0);
#pragma unroll
  for (int i = 0; i < D; ++i) X.psi[i] = need_psi ? a.psi[(int64_t)i * m.C + c] : sconst<S>(1.0);
}

// Flow boundary value of variable v (u_0..u_{D-1}, p) on boundary face f owned by cell P.
template <int D, class S>
AFX_DEV S flow_bval(const FlowArgs<S>& a, int kind, int v, int32_t F, int32_t f, const S* xP) {
  if (v < D) {
    if (kind == kFlowWall || kind == kFlowInlet) return a.ub[(int64_t)v * F + f];
    if (kind == kFlowOutlet) return xP[v];
    return sconst<S>(0.0);
  }
  if (kind == kFlowWall || kind == kFlowInlet) return xP[D];
  if (kind == kFlowOutlet) return a.pb[f];
  return sconst<S>(0.0);
}

// ---------------------------------------------------------------------------------------------
// pass G: gradients (+ limiter)
// With GEO the geometry is read as (value, tangent) pairs (the node-coordinate JVP; S = dual).
template <int D, class S, int GRAD, bool LIM, bool GEO = false>
__global__ void __launch_bounds__(256) k_flow_grad(MeshView m, FlowArgs<S> a, Out<S> G, Out<S> psi_out,
                                                   int32_t* __restrict__ ext) {
  constexpr int NV = D + 1;
  using GR = GeoRead<D, GEO>;
  using GT = typename GR::T;
  int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  S xc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) xc[v] = a.x[(int64_t)v * m.C + c];
  S acc[NV][D];
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int j = 0; j < D; ++j) acc[v][j] = sconst<S>(0.0);
  GT M[D][D];
  if (GRAD == 1) {
#pragma unroll
    for (int p = 0; p < D; ++p)
#pragma unroll
      for (int q = 0; q < D; ++q) M[p][q] = sconst<GT>(0.0);
  }
  S mx[D], mn[D];
  int32_t amx[D], amn[D];
  if (LIM) {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      mx[i] = xc[i];
      mn[i] = xc[i];
      amx[i] = c;
      amn[i] = c;
    }
  }
  const int32_t kb = m.cptr[c], ke = m.cptr[c + 1];
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    if (o >= 0) {
      S xo[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) xo[v] = a.x[(int64_t)v * m.C + o];
      if (LIM) {
#pragma unroll
        for (int i = 0; i < D; ++i) {
          if (val(xo[i]) > val(mx[i])) {
            mx[i] = xo[i];
            amx[i] = o;
          }
          if (val(xo[i]) < val(mn[i])) {
            mn[i] = xo[i];
            amn[i] = o;
          }
        }
      }
      if (GRAD == 0) {
        FaceGeo<D, GT> q;
        GR::face(m, a.gt, r.f, r.P, r.N, q);
        const GT w = 1.0 - q.g;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          S phiP = r.own ? xc[v] : xo[v];
          S phiN = r.own ? xo[v] : xc[v];
          S phi = w * phiP + q.g * phiN;
#pragma unroll
          for (int j = 0; j < D; ++j) {
            S t = phi * (q.n[j] * q.A);
            acc[v][j] = r.own ? acc[v][j] + t : acc[v][j] - t;
          }
        }
      } else {
        GT dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = GR::cc(m, a.gt, j, o) - GR::cc(m, a.gt, j, c);
        const GT wgt = 1.0 / vdotg<D, GT>(dx, dx);
#pragma unroll
        for (int p = 0; p < D; ++p)
#pragma unroll
          for (int q2 = 0; q2 < D; ++q2) M[p][q2] += wgt * dx[p] * dx[q2];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          S dphi = xo[v] - xc[v];
#pragma unroll
          for (int j = 0; j < D; ++j) acc[v][j] += (wgt * dx[j]) * dphi;
        }
      }
    } else {
      const int kind = a.kind[r.fu];
      if (GRAD == 0) {
        const GT A = GR::area(m, a.gt, r.f);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          S phi = flow_bval<D, S>(a, kind, v, m.F, r.fu, xc);
#pragma unroll
          for (int j = 0; j < D; ++j) acc[v][j] += phi * (GR::nrm(m, a.gt, j, r.f) * A);
        }
      } else {
        GT dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = GR::fc(m, a.gt, j, r.f) - GR::cc(m, a.gt, j, c);
        const GT wgt = 1.0 / vdotg<D, GT>(dx, dx);
#pragma unroll
        for (int p = 0; p < D; ++p)
#pragma unroll
          for (int q2 = 0; q2 < D; ++q2) M[p][q2] += wgt * dx[p] * dx[q2];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          S dphi = flow_bval<D, S>(a, kind, v, m.F, r.fu, xc) - xc[v];
#pragma unroll
          for (int j = 0; j < D; ++j) acc[v][j] += (wgt * dx[j]) * dphi;
        }
      }
    }
  }
  S grad[NV][D];
  if (GRAD == 0) {
    const GT iv = 1.0 / GR::vol(m, a.gt, c);
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int j = 0; j < D; ++j) grad[v][j] = acc[v][j] * iv;
  } else {
    GT I[D][D];
    sym_inverse<D, GT>(M, I);
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int p = 0; p < D; ++p) {
        S s = I[p][0] * acc[v][0];
#pragma unroll
        for (int q2 = 1; q2 < D; ++q2) s = s + I[p][q2] * acc[v][q2];
        grad[v][p] = s;
      }
  }
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int j = 0; j < D; ++j) G.set((int64_t)(v * D + j) * m.C + c, grad[v][j]);
  if (LIM) {
    const GT e2 = GR::vol(m, a.gt, c) * (a.lim_k * a.lim_k * a.lim_k);
    GT xcen[D];
#pragma unroll
    for (int j = 0; j < D; ++j) xcen[j] = GR::cc(m, a.gt, j, c);
    S lo[D];
    bool first = true;
    for (int32_t k = kb; k < ke; ++k) {
      int32_t e = m.cface[k];
      int32_t f = e >= 0 ? e : ~e;
      GT dxf[D];
#pragma unroll
      for (int j = 0; j < D; ++j) dxf[j] = GR::fc(m, a.gt, j, f) - xcen[j];
#pragma unroll
      for (int i = 0; i < D; ++i) {
        S dm = grad[i][0] * dxf[0];
#pragma unroll
        for (int j = 1; j < D; ++j) dm = dm + grad[i][j] * dxf[j];
        S vf = venkat_face<S, GT>(dm, mx[i] - xc[i], mn[i] - xc[i], e2);
        if (first || val(vf) < val(lo[i])) lo[i] = vf;
      }
      first = false;
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
      S ps = dmin(dmax(lo[i], 0.0), 1.0);
      psi_out.set((int64_t)i * m.C + c, ps);
      if (ext) {
        ext[(int64_t)i * m.C + c] = amx[i];
        ext[(int64_t)(D + i) * m.C + c] = amn[i];
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// pass D: momentum diagonal a_P
template <int D, class S, int ADV, bool GEO = false>
__global__ void __launch_bounds__(256) k_flow_ap(MeshView m, FlowArgs<S> a, Out<S> aP) {
  using GR = GeoRead<D, GEO>;
  using GT = typename GR::T;
  int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  FlowCell<D, S, GT> X;
  load_state<D, S>(a, m.C, c, X);
  load_grad<D, S>(a, m.C, c, X);
  X.mu = a.mu[c];
  X.rho = a.rho[c];
  X.V = GR::vol(m, a.gt, c);
  S acc = sconst<S>(0.0);
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D, GT> q;
    GR::face(m, a.gt, r.f, r.P, r.N, q);
    if (o >= 0) {
      FlowCell<D, S, GT> O;
      load_state<D, S>(a, m.C, o, O);
      load_grad<D, S>(a, m.C, o, O);
      O.mu = a.mu[o];
      O.rho = a.rho[o];
      O.V = GR::vol(m, a.gt, o);
      S oc, nc;
      if (r.own)
        ap_interior<D, ADV, S>(q, X, O, oc, nc);
      else
        ap_interior<D, ADV, S>(q, O, X, oc, nc);
      acc += r.own ? oc : nc;
    } else {
      S gb = a.gB.v ? a.gB[r.fu] : X.mu;
      acc += ap_boundary<D, ADV, S>(q, a.kind[r.fu], X, gb);
    }
  }
  aP.set(c, acc);
}

// ---------------------------------------------------------------------------------------------
// pass R: residual
template <int D, class S, int ADV, bool GEO = false>
__global__ void __launch_bounds__(256) k_flow_res(MeshView m, FlowArgs<S> a, Out<S> R) {
  constexpr int NV = D + 1;
  using GR = GeoRead<D, GEO>;
  using GT = typename GR::T;
  int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  FlowCell<D, S, GT> X;
  load_cell_full<D, S, GEO>(a, m, c, true, ADV == kAdvLimited, X);
  S acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = sconst<S>(0.0);
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D, GT> q;
    GR::face(m, a.gt, r.f, r.P, r.N, q);
    S out[NV];
    if (o >= 0) {
      FlowCell<D, S, GT> O;
      load_cell_full<D, S, GEO>(a, m, o, true, ADV == kAdvLimited, O);
      if (r.own)
        flow_flux_interior<D, ADV, S>(q, X, O, out);
      else
        flow_flux_interior<D, ADV, S>(q, O, X, out);
    } else {
      S ub[D];
#pragma unroll
      for (int i = 0; i < D; ++i) ub[i] = a.ub[(int64_t)i * m.F + r.fu];
      S gb = a.gB.v ? a.gB[r.fu] : X.mu;
      flow_flux_boundary<D, ADV, S>(q, a.kind[r.fu], X, ub, a.pb[r.fu], gb, out);
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = r.own ? acc[v] + out[v] : acc[v] - out[v];
  }
#pragma unroll
  for (int i = 0; i < D; ++i) R.set((int64_t)i * m.C + c, acc[i] - a.bf[i] * X.V);
  S cont = acc[D];
  if (c == a.pin) cont = X.p - a.pin_value;
  R.set((int64_t)D * m.C + c, cont);
}

// Rhie-Chow face mass flux of every face, written by the face's owner in the caller's face order
// (the mdot that continuity sums and the advection fluxes carry).
template <int D, class S, int ADV>
__global__ void __launch_bounds__(256) k_flow_mdot(MeshView m, FlowArgs<S> a, Out<S> mdot) {
  int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  FlowCell<D, S> X;
  load_cell_full<D, S>(a, m, c, true, ADV == kAdvLimited, X);
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    if (!r.own) continue;
    FaceGeo<D> q;
    load_face_geo<D>(m, r.f, r.P, r.N, q);
    S out[D + 1];
    if (r.N >= 0) {
      FlowCell<D, S> O;
      load_cell_full<D, S>(a, m, r.N, true, ADV == kAdvLimited, O);
      flow_flux_interior<D, ADV, S>(q, X, O, out);
    } else {
      S ub[D];
#pragma unroll
      for (int i = 0; i < D; ++i) ub[i] = a.ub[(int64_t)i * m.F + r.fu];
      S gb = a.gB.v ? a.gB[r.fu] : X.mu;
      flow_flux_boundary<D, ADV, S>(q, a.kind[r.fu], X, ub, a.pb[r.fu], gb, out);
    }
    mdot.set(r.fu, out[D]);
  }
}

// ---------------------------------------------------------------------------------------------
// reverse pass R': cotangent of the residual -> cell-local cotangents of state, gradients, a_P,
// mu, rho, psi (written), and of the boundary-face data of faces this cell owns.
struct FlowBars {
  double* x;      // [(D+1)][C]
  double* G;      // [(D+1)D][C]
  double* aP;     // [C]
  double* psi;    // [D][C]
  double* mu;     // [C]
  double* rho;    // [C]
  double* ub;     // [D][F]
  double* pb;     // [F]
  double* gB;     // [F] (only with a boundary-coefficient override)
  double* ext;    // [2D][C] cotangents of the limiter stencil extrema
};

template <int D, int ADV>
__global__ void __launch_bounds__(256) k_flow_res_adj(MeshView m, FlowArgs<double> a, const double* __restrict__ lam,
                                                      FlowBars bar) {
  constexpr int NV = D + 1;
  int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  FlowCell<D, double> X;
  load_cell_full<D, double>(a, m, c, true, ADV == kAdvLimited, X);
  FlowCellBar<D> B;
  B.zero();
  double lc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) lc[v] = lam[(int64_t)v * m.C + c];
  const double lc_cont = (c == a.pin) ? 0.0 : lc[D];
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D> q;
    load_face_geo<D>(m, r.f, r.P, r.N, q);
    if (o >= 0) {
      FlowCell<D, double> O;
      load_cell_full<D, double>(a, m, o, true, ADV == kAdvLimited, O);
      double lf[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        double lo = lam[(int64_t)v * m.C + o];
        if (v == D && o == a.pin) lo = 0.0;
        double lcv = v == D ? lc_cont : lc[v];
        lf[v] = r.own ? lcv - lo : lo - lcv;
      }
      if (r.own)
        adj_flux_interior<D, ADV>(q, X, O, lf, 0, B);
      else
        adj_flux_interior<D, ADV>(q, O, X, lf, 1, B);
    } else {
      double lf[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) lf[v] = v == D ? lc_cont : lc[v];
      double ub[D], ubb[D], pbb, gbb;
#pragma unroll
      for (int i = 0; i < D; ++i) ub[i] = a.ub[(int64_t)i * m.F + r.fu];
      const double gb = a.gB.v ? a.gB[r.fu] : X.mu;
      adj_flux_boundary<D, ADV>(q, a.kind[r.fu], X, ub, a.pb[r.fu], gb, lf, B, ubb, pbb, gbb);
#pragma unroll
      for (int i = 0; i < D; ++i) bar.ub[(int64_t)i * m.F + r.fu] = ubb[i];
      bar.pb[r.fu] = pbb;
      if (a.gB.v) {
        if (bar.gB) bar.gB[r.fu] = gbb;
      } else {
        B.mu += gbb;
      }
    }
  }
  if (c == a.pin) B.p += lc[D];
#pragma unroll
  for (int i = 0; i < D; ++i) bar.x[(int64_t)i * m.C + c] = B.u[i];
  bar.x[(int64_t)D * m.C + c] = B.p;
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) bar.G[(int64_t)(i * D + j) * m.C + c] = B.Gu[i][j];
#pragma unroll
  for (int j = 0; j < D; ++j) bar.G[(int64_t)(D * D + j) * m.C + c] = B.Gp[j];
  bar.aP[c] = B.aP;
  bar.mu[c] = B.mu;
  bar.rho[c] = B.rho;
  if (ADV == kAdvLimited) {
#pragma unroll
    for (int i = 0; i < D; ++i) bar.psi[(int64_t)i * m.C + c] = B.psi[i];
  }
}

// reverse pass D': a_P cotangent (+ limiter) -> state / gradient / mu / rho cotangents; then the
// gradient cotangent is pre-multiplied by the cell's reconstruction operator (1/V or M^-1).
template <int D, int ADV, int GRAD>
__global__ void __launch_bounds__(256) k_flow_ap_adj(MeshView m, FlowArgs<double> a, const int32_t* __restrict__ ext,
                                                     FlowBars bar) {
  constexpr int NV = D + 1;
  int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  FlowCell<D, double> X;
  load_state<D, double>(a, m.C, c, X);
  load_grad<D, double>(a, m.C, c, X);
  X.mu = a.mu[c];
  X.rho = a.rho[c];
  X.V = m.vol[c];
  FlowCellBar<D> B;
  B.zero();
  const double apb_c = bar.aP[c];
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D> q;
    load_face_geo<D>(m, r.f, r.P, r.N, q);
    if (o >= 0) {
      FlowCell<D, double> O;
      load_state<D, double>(a, m.C, o, O);
      load_grad<D, double>(a, m.C, o, O);
      O.mu = a.mu[o];
      O.rho = a.rho[o];
      O.V = m.vol[o];
      const double apb_o = bar.aP[o];
      if (r.own)
        adj_ap_interior<D, ADV>(q, X, O, apb_c, apb_o, 0, B);
      else
        adj_ap_interior<D, ADV>(q, O, X, apb_o, apb_c, 1, B);
    } else {
      const double gb = a.gB.v ? a.gB[r.fu] : X.mu;
      double gbb;
      adj_ap_boundary<D, ADV>(q, a.kind[r.fu], X, gb, apb_c, B, gbb);
      if (a.gB.v) {
        if (bar.gB) bar.gB[r.fu] += gbb;
      } else {
        B.mu += gbb;
      }
    }
  }
  // limiter: psi = clip(min_f venkat(dm_f, head_f, e2), 0, 1)
  if (ADV == kAdvLimited) {
    const double e2 = X.V * (a.lim_k * a.lim_k * a.lim_k);
    double xcen[D];
#pragma unroll
    for (int j = 0; j < D; ++j) xcen[j] = m.cc[(int64_t)j * m.C + c];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const double pbar = bar.psi[(int64_t)i * m.C + c];
      const int32_t amx = ext[(int64_t)i * m.C + c], amn = ext[(int64_t)(D + i) * m.C + c];
      const double mxv = a.x[(int64_t)i * m.C + amx], mnv = a.x[(int64_t)i * m.C + amn];
      // forward recompute: minimum over faces and its first argmin
      double lo = 0.0;
      int32_t kmin = -1;
      for (int32_t k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
        int32_t e = m.cface[k];
        int32_t f = e >= 0 ? e : ~e;
        // recompute with the same association as the forward pass
        double dmf = X.Gu[i][0] * (m.fc[f] - xcen[0]);
#pragma unroll
        for (int j = 1; j < D; ++j) dmf = dmf + X.Gu[i][j] * (m.fc[(int64_t)j * m.F + f] - xcen[j]);
        double vf = venkat_face<double>(dmf, mxv - X.u[i], mnv - X.u[i], e2);
        if (kmin < 0 || vf < lo) {
          lo = vf;
          kmin = k;
        }
      }
      // d clip / d lo  (JAX minimum(maximum(x, 0), 1) tie convention)
      double dclip = (lo > 0.0 && lo < 1.0) ? 1.0 : ((lo == 0.0 || lo == 1.0) ? 0.5 : 0.0);
      const double vbar = pbar * dclip;
      double mxb = 0.0, mnb = 0.0;
      if (vbar != 0.0 && kmin >= 0) {
        int32_t e = m.cface[kmin];
        int32_t f = e >= 0 ? e : ~e;
        double dxf[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dxf[j] = m.fc[(int64_t)j * m.F + f] - xcen[j];
        double dmf = X.Gu[i][0] * dxf[0];
#pragma unroll
        for (int j = 1; j < D; ++j) dmf = dmf + X.Gu[i][j] * dxf[j];
        // partial derivatives by two seeded tangent evaluations
        dual r_dm = venkat_face<dual>(dual{dmf, 1.0}, dual{mxv - X.u[i], 0.0}, dual{mnv - X.u[i], 0.0}, e2);
        dual r_hx = venkat_face<dual>(dual{dmf, 0.0}, dual{mxv - X.u[i], 1.0}, dual{mnv - X.u[i], 0.0}, e2);
        dual r_hn = venkat_face<dual>(dual{dmf, 0.0}, dual{mxv - X.u[i], 0.0}, dual{mnv - X.u[i], 1.0}, e2);
        const double dmb = vbar * r_dm.d;
#pragma unroll
        for (int j = 0; j < D; ++j) B.Gu[i][j] += dmb * dxf[j];
        mxb = vbar * r_hx.d;
        mnb = vbar * r_hn.d;
        B.u[i] -= mxb + mnb;
        if (amx == c) {
          B.u[i] += mxb;
          mxb = 0.0;
        }
        if (amn == c) {
          B.u[i] += mnb;
          mnb = 0.0;
        }
      }
      bar.ext[(int64_t)i * m.C + c] = mxb;
      bar.ext[(int64_t)(D + i) * m.C + c] = mnb;
    }
  }
  // accumulate onto the R' cotangents
#pragma unroll
  for (int i = 0; i < D; ++i) bar.x[(int64_t)i * m.C + c] += B.u[i];
  bar.mu[c] += B.mu;
  bar.rho[c] += B.rho;
  // total gradient cotangent, pre-multiplied by the reconstruction operator
  double gb[NV][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) gb[i][j] = bar.G[(int64_t)(i * D + j) * m.C + c] + B.Gu[i][j];
#pragma unroll
  for (int j = 0; j < D; ++j) gb[D][j] = bar.G[(int64_t)(D * D + j) * m.C + c] + B.Gp[j];
  if (GRAD == kGradCorrected) {  // the swept solve's reverse pass forms the cotangent of b
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int j = 0; j < D; ++j) bar.G[(int64_t)(v * D + j) * m.C + c] = gb[v][j];
  } else if (GRAD == 0) {
    const double iv = 1.0 / X.V;
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int j = 0; j < D; ++j) bar.G[(int64_t)(v * D + j) * m.C + c] = gb[v][j] * iv;
  } else {
    double M[D][D], I[D][D];
    lsq_matrix<D>(m, c, M);
    sym_inverse<D>(M, I);
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int p = 0; p < D; ++p) {
        double s = 0.0;
#pragma unroll
        for (int q2 = 0; q2 < D; ++q2) s += I[p][q2] * gb[v][q2];
        bar.G[(int64_t)(v * D + p) * m.C + c] = s;
      }
  }
}

// reverse pass G': scaled gradient cotangents (and limiter extrema cotangents) -> state and
// boundary-value cotangents.
template <int D, int GRAD, bool LIM>
__global__ void __launch_bounds__(256) k_flow_grad_adj(MeshView m, FlowArgs<double> a, const int32_t* __restrict__ ext,
                                                       FlowBars bar) {
  constexpr int NV = D + 1;
  int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  double xb[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) xb[v] = 0.0;
  double Bc[NV][D];
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int j = 0; j < D; ++j) Bc[v][j] = bar.G[(int64_t)(v * D + j) * m.C + c];
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    if (o >= 0) {
      if (GRAD == 0) {
        FaceGeo<D> q;
        load_face_geo<D>(m, r.f, r.P, r.N, q);
        const double ws = r.own ? 1.0 - q.g : q.g;
        const int32_t P = r.P, N = r.N;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          double pb_ = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j)
            pb_ += (q.n[j] * q.A) * (bar.G[(int64_t)(v * D + j) * m.C + P] - bar.G[(int64_t)(v * D + j) * m.C + N]);
          xb[v] += ws * pb_;
        }
      } else {
        double dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = m.cc[(int64_t)j * m.C + o] - m.cc[(int64_t)j * m.C + c];
        const double wgt = 1.0 / vdot<D>(dx, dx);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          double own_dot = 0.0, oth_dot = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) {
            own_dot += dx[j] * Bc[v][j];
            oth_dot += dx[j] * bar.G[(int64_t)(v * D + j) * m.C + o];
          }
          // own gradient: b_c += w dx (x_o - x_c);  other's: b_o += w (-dx) (x_c - x_o)
          xb[v] += -wgt * own_dot - wgt * oth_dot;
        }
      }
      if (LIM) {  // all extremum cells and cotangents loaded before any comparison
        int32_t amx[D], amn[D];
        double bmx[D], bmn[D];
#pragma unroll
        for (int i = 0; i < D; ++i) {
          amx[i] = ext[(int64_t)i * m.C + o];
          amn[i] = ext[(int64_t)(D + i) * m.C + o];
          bmx[i] = bar.ext[(int64_t)i * m.C + o];
          bmn[i] = bar.ext[(int64_t)(D + i) * m.C + o];
        }
#pragma unroll
        for (int i = 0; i < D; ++i) {
          if (amx[i] == c) xb[i] += bmx[i];
          if (amn[i] == c) xb[i] += bmn[i];
        }
      }
    } else {
      const int kind = a.kind[r.fu];
      double phib[NV];
      if (GRAD == 0) {
        const double A = m.area[r.f];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          double s = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) s += (m.nrm[(int64_t)j * m.F + r.f] * A) * Bc[v][j];
          phib[v] = s;
        }
      } else {
        double dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = m.fc[(int64_t)j * m.F + r.f] - m.cc[(int64_t)j * m.C + c];
        const double wgt = 1.0 / vdot<D>(dx, dx);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          double s = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) s += dx[j] * Bc[v][j];
          phib[v] = wgt * s;
          xb[v] -= wgt * s;
        }
      }
      // route the boundary-value cotangents through the closures
#pragma unroll
      for (int i = 0; i < D; ++i) {
        if (kind == kFlowWall || kind == kFlowInlet)
          bar.ub[(int64_t)i * m.F + r.fu] += phib[i];
        else if (kind == kFlowOutlet)
          xb[i] += phib[i];
      }
      if (kind == kFlowWall || kind == kFlowInlet)
        xb[D] += phib[D];
      else if (kind == kFlowOutlet)
        bar.pb[r.fu] += phib[D];
    }
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) bar.x[(int64_t)v * m.C + c] += xb[v];
}

// ---------------------------------------------------------------------------------------------
// geometry cotangents (node-coordinate VJP).  Run after the state adjoint: they read its a_P and
// limiter cotangents and its gradient cotangents pre-multiplied by the reconstruction operator
// (y = V^-1 Gbar for Green-Gauss, M^-1 Gbar for least squares).  One thread per cell again; a
// face's cotangent is evaluated by both of its cells, the owner writes the face arrays and each
// cell keeps its own centroid and volume parts, so no two threads write one location.

struct FlowGeoBars {
  double *area, *nrm, *fc;  // [F], [D][F], [D][F] in the caller's face order (written)
  double *vol, *cc;         // [C], [D][C] (written)
  double* limx;             // [D][D][C] scratch: limiter face-centroid cotangent at its argmin face
  int32_t* limf;            // [D][C]    scratch: that face (view index, -1 none)
  CggGeoBars cgg;           // corrected Green-Gauss: the swept solve's geometry cotangents
};

// Limiter part: psi_i = clip(min_f venkat(G_i . (xf - xc), heads, V K^3), 0, 1) at its first
// argmin face.  Writes the cell's volume / centroid partials and the argmin face's centroid part.
template <int D>
__global__ void __launch_bounds__(256) k_flow_geo_lim(MeshView m, FlowArgs<double> a, const int32_t* __restrict__ ext,
                                                      const double* __restrict__ psib, FlowGeoBars gb) {
  int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  const double V = m.vol[c], k3 = a.lim_k * a.lim_k * a.lim_k;
  const double e2 = V * k3;
  double xcen[D], ccb[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    xcen[j] = m.cc[(int64_t)j * m.C + c];
    ccb[j] = 0.0;
  }
  double Vb = 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double Gi[D];
#pragma unroll
    for (int j = 0; j < D; ++j) Gi[j] = a.G[(int64_t)(i * D + j) * m.C + c];
    const double xi = a.x[(int64_t)i * m.C + c];
    const double mxv = a.x[(int64_t)i * m.C + ext[(int64_t)i * m.C + c]];
    const double mnv = a.x[(int64_t)i * m.C + ext[(int64_t)(D + i) * m.C + c]];
    double lo = 0.0;
    int32_t fmin = -1;
    for (int32_t k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
      const int32_t e = m.cface[k];
      const int32_t f = e >= 0 ? e : ~e;
      double dmf = Gi[0] * (m.fc[f] - xcen[0]);
#pragma unroll
      for (int j = 1; j < D; ++j) dmf = dmf + Gi[j] * (m.fc[(int64_t)j * m.F + f] - xcen[j]);
      const double vf = venkat_face<double>(dmf, mxv - xi, mnv - xi, e2);
      if (fmin < 0 || vf < lo) {
        lo = vf;
        fmin = f;
      }
    }
    const double dclip = (lo > 0.0 && lo < 1.0) ? 1.0 : ((lo == 0.0 || lo == 1.0) ? 0.5 : 0.0);
    const double vbar = psib[(int64_t)i * m.C + c] * dclip;
    double xb[D];
#pragma unroll
    for (int j = 0; j < D; ++j) xb[j] = 0.0;
    if (vbar != 0.0 && fmin >= 0) {
      double dxf[D];
#pragma unroll
      for (int j = 0; j < D; ++j) dxf[j] = m.fc[(int64_t)j * m.F + fmin] - xcen[j];
      double dmf = Gi[0] * dxf[0];
#pragma unroll
      for (int j = 1; j < D; ++j) dmf = dmf + Gi[j] * dxf[j];
      const dual r_dm = venkat_face<dual>(dual{dmf, 1.0}, dual{mxv - xi, 0.0}, dual{mnv - xi, 0.0}, e2);
      const dual r_e2 =
          venkat_face<dual, dual>(dual{dmf, 0.0}, dual{mxv - xi, 0.0}, dual{mnv - xi, 0.0}, dual{e2, 1.0});
      const double dmb = vbar * r_dm.d;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        xb[j] = dmb * Gi[j];
        ccb[j] -= xb[j];
      }
      Vb += vbar * r_e2.d * k3;
    }
    gb.limf[(int64_t)i * m.C + c] = vbar != 0.0 ? fmin : -1;
#pragma unroll
    for (int j = 0; j < D; ++j) gb.limx[(int64_t)(i * D + j) * m.C + c] = xb[j];
  }
  gb.vol[c] = Vb;
#pragma unroll
  for (int j = 0; j < D; ++j) gb.cc[(int64_t)j * m.C + c] = ccb[j];
}

template <int D, int ADV, int GRAD>
__global__ void __launch_bounds__(256) k_flow_geo_adj(MeshView m, FlowArgs<double> a, const double* __restrict__ lam,
                                                      const double* __restrict__ apb, const double* __restrict__ yb,
                                                      FlowGeoBars gb) {
  constexpr int NV = D + 1;
  constexpr bool LIM = ADV == kAdvLimited;
  int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  FlowCell<D, double> X;
  load_cell_full<D, double>(a, m, c, true, LIM, X);
  double xc[NV], lc[NV], yc[NV][D], Gc[NV][D];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    xc[v] = a.x[(int64_t)v * m.C + c];
    lc[v] = lam[(int64_t)v * m.C + c];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      yc[v][j] = yb[(int64_t)(v * D + j) * m.C + c];
      Gc[v][j] = a.G[(int64_t)(v * D + j) * m.C + c];
    }
  }
  if (c == a.pin) lc[D] = 0.0;
  const double apb_c = apb[c];
  double Vb = LIM ? gb.vol[c] : 0.0, ccb[D];
#pragma unroll
  for (int j = 0; j < D; ++j) ccb[j] = LIM ? gb.cc[(int64_t)j * m.C + c] : 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i) Vb -= a.bf[i] * lc[i];  // body force - bf V
  if (GRAD == 0) {  // G = acc / V
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int j = 0; j < D; ++j) Vb -= yc[v][j] * Gc[v][j];
  }
  if (GRAD == kGradCorrected) Vb += gb.cgg.vol[c];  // the sweeps' V^-1
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D> q;
    load_face_geo<D>(m, r.f, r.P, r.N, q);
    FaceGeoBarF<D> qb;
    qb.zero();
    double xfb[D];  // face-centroid parts outside the face functions (fit directions, limiter)
#pragma unroll
    for (int j = 0; j < D; ++j) xfb[j] = 0.0;
    if (o >= 0) {
      FlowCell<D, double> O;
      load_cell_full<D, double>(a, m, o, true, LIM, O);
      double xo[NV], lo[NV], yo[NV][D], Go[NV][D];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        xo[v] = a.x[(int64_t)v * m.C + o];
        lo[v] = lam[(int64_t)v * m.C + o];
#pragma unroll
        for (int j = 0; j < D; ++j) {
          yo[v][j] = yb[(int64_t)(v * D + j) * m.C + o];
          Go[v][j] = a.G[(int64_t)(v * D + j) * m.C + o];
        }
      }
      if (o == a.pin) lo[D] = 0.0;
      const double apb_o = apb[o];
      const FlowCell<D, double>& P = r.own ? X : O;
      const FlowCell<D, double>& N = r.own ? O : X;
      double lf[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) lf[v] = r.own ? lc[v] - lo[v] : lo[v] - lc[v];
      double VPb = 0.0, VNb = 0.0;
      geo_adj_flux_interior<D, ADV>(q, P, N, lf, qb, VPb, VNb);
      geo_adj_ap_interior<D, ADV>(q, P, N, r.own ? apb_c : apb_o, r.own ? apb_o : apb_c, qb);
      if (GRAD == kGradCorrected) {  // the sweeps' face terms, folded with the rest below
        double xPb[D], xNb[D];
#pragma unroll
        for (int j = 0; j < D; ++j) xPb[j] = xNb[j] = 0.0;
        cgg_face_geo_bar<D>(gb.cgg, m.F, r.f, q, qb.A, qb.n, qb.xf, xPb, xNb);
#pragma unroll
        for (int j = 0; j < D; ++j) ccb[j] += r.own ? xPb[j] : xNb[j];
      }
      if (GRAD != 1) {  // acc_P += phi_f n A, acc_N -= phi_f n A (compact, or the corrected scheme's b)
        const double w = 1.0 - q.g;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          const double phP = r.own ? xc[v] : xo[v], phN = r.own ? xo[v] : xc[v];
          const double phf = w * phP + q.g * phN;
          double zn = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) {
            const double z = r.own ? yc[v][j] - yo[v][j] : yo[v][j] - yc[v][j];
            qb.n[j] += z * phf * q.A;
            zn += z * q.n[j];
          }
          qb.A += zn * phf;
          qb.g += zn * q.A * (phN - phP);
        }
      } else {  // both cells' fits use the direction between the centroids
        double d[D], dphi[NV], db[D];
#pragma unroll
        for (int j = 0; j < D; ++j) d[j] = m.cc[(int64_t)j * m.C + o] - m.cc[(int64_t)j * m.C + c];
#pragma unroll
        for (int v = 0; v < NV; ++v) dphi[v] = xo[v] - xc[v];
        lsq_dir_bar<D, NV>(d, dphi, Gc, yc, db);  // own fit: d = x_o - x_c
#pragma unroll
        for (int j = 0; j < D; ++j) {
          ccb[j] -= db[j];
          d[j] = -d[j];
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) dphi[v] = -dphi[v];
        lsq_dir_bar<D, NV>(d, dphi, Go, yo, db);  // the other cell's fit: d = x_c - x_o
#pragma unroll
        for (int j = 0; j < D; ++j) ccb[j] += db[j];
      }
      fold_face_geo_bar<D>(q, true, qb);
      Vb += r.own ? VPb : VNb;
#pragma unroll
      for (int j = 0; j < D; ++j) ccb[j] += r.own ? qb.xP[j] : qb.xN[j];
      if (r.own && LIM) {
#pragma unroll
        for (int i = 0; i < D; ++i) {
          const bool hc = gb.limf[(int64_t)i * m.C + c] == r.f, ho = gb.limf[(int64_t)i * m.C + o] == r.f;
#pragma unroll
          for (int j = 0; j < D; ++j) {
            if (hc) xfb[j] += gb.limx[(int64_t)(i * D + j) * m.C + c];
            if (ho) xfb[j] += gb.limx[(int64_t)(i * D + j) * m.C + o];
          }
        }
      }
    } else {
      const int kind = a.kind[r.fu];
      double lf[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) lf[v] = lc[v];
      double ub[D];
#pragma unroll
      for (int i = 0; i < D; ++i) ub[i] = a.ub[(int64_t)i * m.F + r.fu];
      const double gB = a.gB.v ? a.gB[r.fu] : X.mu;
      double VPb = 0.0;
      geo_adj_flux_boundary<D, ADV>(q, kind, X, ub, a.pb[r.fu], gB, lf, qb, VPb);
      geo_adj_ap_boundary<D, ADV>(q, kind, X, gB, apb_c, qb);
      double phb[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) phb[v] = flow_bval<D, double>(a, kind, v, m.F, r.fu, xc);
      if (GRAD != 1) {  // acc_P += phi_b n A
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          double zn = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) {
            qb.n[j] += yc[v][j] * phb[v] * q.A;
            zn += yc[v][j] * q.n[j];
          }
          qb.A += zn * phb[v];
        }
      } else {  // fit direction d = xf - x_c
        double d[D], dphi[NV], db[D];
#pragma unroll
        for (int j = 0; j < D; ++j) d[j] = q.xf[j] - m.cc[(int64_t)j * m.C + c];
#pragma unroll
        for (int v = 0; v < NV; ++v) dphi[v] = phb[v] - xc[v];
        lsq_dir_bar<D, NV>(d, dphi, Gc, yc, db);
#pragma unroll
        for (int j = 0; j < D; ++j) {
          xfb[j] += db[j];
          ccb[j] -= db[j];
        }
      }
      fold_face_geo_bar<D>(q, false, qb);
      Vb += VPb;
#pragma unroll
      for (int j = 0; j < D; ++j) ccb[j] += qb.xP[j];
      if (LIM) {
#pragma unroll
        for (int i = 0; i < D; ++i) {
          if (gb.limf[(int64_t)i * m.C + c] == r.f) {
#pragma unroll
            for (int j = 0; j < D; ++j) xfb[j] += gb.limx[(int64_t)(i * D + j) * m.C + c];
          }
        }
      }
    }
    if (r.own) {
      gb.area[r.fu] = qb.A;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        gb.nrm[(int64_t)j * m.F + r.fu] = qb.n[j];
        gb.fc[(int64_t)j * m.F + r.fu] = qb.xf[j] + xfb[j];
      }
    }
  }
  gb.vol[c] = Vb;
#pragma unroll
  for (int j = 0; j < D; ++j) gb.cc[(int64_t)j * m.C + c] = ccb[j];
}

// ---------------------------------------------------------------------------------------------
// host side

struct Work {
  double *G, *aP, *psi;
  double *Gt, *aPt, *psit;
  double *Gb, *aPb, *psib, *extb;
  int32_t* ext;
  double* cgg;  // corrected Green-Gauss sweep scratch [2 (D+1) D][C]
  double* limx;  // mode 3: limiter face-centroid cotangents [D][D][C]
  int32_t* limf; // mode 3: their faces [D][C]
};

int64_t work_bytes(int D, int64_t C, int mode) {
  const int NV = D + 1;
  int64_t n = (int64_t)(NV * D + 1 + D) * C;  // G, aP, psi
  if (mode == 1) n *= 2;
  if (mode >= 2) n += (int64_t)(NV * D + 1 + D + 2 * D) * C + (int64_t)D * C;  // bars + ext (2D int32)
  n += (int64_t)2 * NV * D * C;  // corrected Green-Gauss sweep scratch
  if (mode == 3) n += (int64_t)(D * D + D) * C;  // geometry cotangents: limiter scratch
  return n * 8 + 256;
}

Work carve(void* base, int D, int64_t C, int mode) {
  const int NV = D + 1;
  double* p = (double*)base;
  Work w{};
  w.G = p;
  p += (int64_t)NV * D * C;
  w.aP = p;
  p += C;
  w.psi = p;
  p += (int64_t)D * C;
  if (mode == 1) {
    w.Gt = p;
    p += (int64_t)NV * D * C;
    w.aPt = p;
    p += C;
    w.psit = p;
    p += (int64_t)D * C;
  }
  if (mode >= 2) {
    w.Gb = p;
    p += (int64_t)NV * D * C;
    w.aPb = p;
    p += C;
    w.psib = p;
    p += (int64_t)D * C;
    w.extb = p;
    p += (int64_t)2 * D * C;
    w.ext = (int32_t*)p;
    p += (int64_t)D * C;
  }
  w.cgg = p;
  if (mode == 3) {
    p += (int64_t)2 * NV * D * C;
    w.limx = p;
    p += (int64_t)D * D * C;
    w.limf = (int32_t*)p;
  }
  return w;
}

template <class S>
FlowArgs<S> make_args(const afx_flow_config* cfg, const afx_flow_inputs* in, const afx_flow_inputs* tan) {
  FlowArgs<S> a{};
  a.kind = cfg->face_kind;
  a.x.v = in->state;
  a.ub.v = in->u_b;
  a.pb.v = in->p_b;
  a.gB.v = in->gamma_b;
  a.mu.v = in->mu;
  a.rho.v = in->rho;
  if (tan) {
    a.x.t = tan->state;
    a.ub.t = tan->u_b;
    a.pb.t = tan->p_b;
    a.gB.t = tan->gamma_b;
    a.mu.t = tan->mu;
    a.rho.t = tan->rho;
  }
  a.pin = cfg->pressure_pin >= 0 ? (int32_t)cfg->pressure_pin : -1;
  a.pin_value = cfg->pressure_pin_value;
  for (int i = 0; i < 3; ++i) a.bf[i] = cfg->body_force[i];
  a.lim_k = cfg->limiter_k;
  return a;
}

// GEO: the geometry carries tangents in a.gt (S = dual).  iters (corrected Green-Gauss, nullable):
// keep the sweeps' iterates for the geometry adjoint.
template <int D, class S, int ADV, int GRAD, bool GEO = false>
int run_forward(const afx_mesh* mh, const afx_geometry* g, FlowArgs<S> a, Out<S> R, double* G, double* Gt,
                double* aP, double* aPt, double* psi, double* psit, int32_t* ext, double* cgg, int sweeps,
                double relax, hipStream_t s, Out<S> mdot = Out<S>{}, double* iters = nullptr) {
  static_assert(!GEO || std::is_same<S, dual>::value, "geometry tangents need a dual state");
  MeshView m;
  if (int rv_ = afx_rview(mh, g, s, &m)) return rv_;
  Out<S> oG{G, Gt}, oP{psi, psit}, oA{aP, aPt};
  constexpr int GK = GRAD == kGradCorrected ? 0 : GRAD;  // the compact pass supplies V^-1 b
  if (ADV == kAdvLimited && GRAD != kGradCorrected)
    AFX_LAUNCH(kKFlowGrad, (k_flow_grad<D, S, GK, true, GEO>), grid_for(m.C), dim3(kBlock), 0, s, m, a, oG, oP,
               ext);
  else
    AFX_LAUNCH(kKFlowGrad, (k_flow_grad<D, S, GK, false, GEO>), grid_for(m.C), dim3(kBlock), 0, s, m, a, oG, oP,
               ext);
  AFX_CHECK_LAUNCH();
  a.G.v = G;
  a.G.t = Gt;
  if constexpr (GRAD == kGradCorrected) {
    LimiterSlots sl{};
    sl.nl = D;
    sl.ext_stride = D;
    for (int i = 0; i < D; ++i) sl.var[i] = sl.slot[i] = i;
    if constexpr (GEO) {  // the sweeps and the limiter with the geometry's tangents
      double* scr = afx_mesh_scratch(const_cast<afx_mesh*>(mh), (int64_t)4 * (D + 1) * D * m.C);
      AFX_REQUIRE(scr, "flow: could not allocate the corrected-gradient sweep scratch");
      int rc = cgg_sweeps_tangent(m, a.gt, D, D + 1, G, Gt, scr, sweeps, relax, s);
      if (rc) return rc;
      if (ADV == kAdvLimited) {
        rc = venkat_pass_tangent(m, a.gt, D, sl, a.x, a.G, a.lim_k, oP, s);
        if (rc) return rc;
      }
    } else {
      AFX_REQUIRE(cgg, "flow: corrected Green-Gauss needs sweep scratch");
      int rc = iters ? cgg_sweeps_store(m, D, D + 1, G, iters, sweeps, relax, s)
                     : cgg_sweeps(m, D, D + 1, G, cgg, sweeps, relax, s);
      if (rc) return rc;
      if (Gt) {
        rc = cgg_sweeps(m, D, D + 1, Gt, cgg, sweeps, relax, s);
        if (rc) return rc;
      }
      if (ADV == kAdvLimited) {
        rc = venkat_pass(m, D, sl, a.x, a.G, a.lim_k, oP, ext, s);
        if (rc) return rc;
      }
    }
  }
  AFX_LAUNCH(kKFlowAp, (k_flow_ap<D, S, ADV, GEO>), grid_for(m.C), dim3(kBlock), 0, s, m, a, oA);
  AFX_CHECK_LAUNCH();
  if (R.v == nullptr && mdot.v == nullptr) return 0;
  a.aP.v = aP;
  a.aP.t = aPt;
  a.psi.v = psi;
  a.psi.t = psit;
  if constexpr (!GEO) {
    if (mdot.v != nullptr) {
      AFX_LAUNCH(kKFlowRes, (k_flow_mdot<D, S, ADV>), grid_for(m.C), dim3(kBlock), 0, s, m, a, mdot);
      AFX_CHECK_LAUNCH();
    }
  }
  if (R.v == nullptr) return 0;
  AFX_LAUNCH(kKFlowRes, (k_flow_res<D, S, ADV, GEO>), grid_for(m.C), dim3(kBlock), 0, s, m, a, R);
  AFX_CHECK_LAUNCH();
  return 0;
}

// iters / cb (corrected Green-Gauss, geometry adjoint): the sweeps' iterates are kept and their
// geometry cotangents written to cb.
template <int D, int ADV, int GRAD>
int run_vjp(const afx_mesh* mh, const afx_geometry* g, FlowArgs<double> a, const double* lam, FlowBars bar,
            const Work& w, int sweeps, double relax, hipStream_t s, double* iters = nullptr,
            CggGeoBars cb = CggGeoBars{nullptr, nullptr}) {
  MeshView m;
  if (int rv_ = afx_rview(mh, g, s, &m)) return rv_;
  int rc = run_forward<D, double, ADV, GRAD>(mh, g, a, Out<double>{nullptr, nullptr}, w.G, nullptr, w.aP, nullptr,
                                             w.psi, nullptr, w.ext, w.cgg, sweeps, relax, s, Out<double>{}, iters);
  if (rc) return rc;
  a.G.v = w.G;
  a.aP.v = w.aP;
  a.psi.v = w.psi;
  // boundary-value cotangents: zero over all faces here, boundary entries written by the kernel
  if (int rz = afx_fill_zero(bar.ub, (int64_t)D * m.F, s)) return rz;
  if (int rz = afx_fill_zero(bar.pb, m.F, s)) return rz;
  if (int rz = afx_fill_zero(bar.gB, m.F, s)) return rz;
  AFX_LAUNCH(kKFlowResAdj, (k_flow_res_adj<D, ADV>), grid_for(m.C), dim3(kBlock), 0, s, m, a, lam, bar);
  AFX_CHECK_LAUNCH();
  AFX_LAUNCH(kKFlowApAdj, (k_flow_ap_adj<D, ADV, GRAD>), grid_for(m.C), dim3(kBlock), 0, s, m, a, w.ext, bar);
  AFX_CHECK_LAUNCH();
  constexpr int GK = GRAD == kGradCorrected ? 0 : GRAD;
  if (GRAD == kGradCorrected) {
    rc = cb.face ? cgg_sweeps_adjoint_geometry(m, D, D + 1, bar.G, w.G, iters, w.cgg, sweeps, relax, cb, s)
                 : cgg_sweeps_adjoint(m, D, D + 1, bar.G, w.cgg, sweeps, relax, s);
    if (rc) return rc;
  }
  if (ADV == kAdvLimited)
    AFX_LAUNCH(kKFlowGradAdj, (k_flow_grad_adj<D, GK, true>), grid_for(m.C), dim3(kBlock), 0, s, m, a, w.ext, bar);
  else
    AFX_LAUNCH(kKFlowGradAdj, (k_flow_grad_adj<D, GK, false>), grid_for(m.C), dim3(kBlock), 0, s, m, a, w.ext, bar);
  AFX_CHECK_LAUNCH();
  return 0;
}

// The state adjoint, then the geometry cotangents from its intermediates.
template <int D, int ADV, int GRAD>
int run_vjp_geometry(const afx_mesh* mh, const afx_geometry* g, FlowArgs<double> a, const double* lam, FlowBars bar,
                     FlowGeoBars gb, const Work& w, int sweeps, double relax, hipStream_t s) {
  double* iters = nullptr;
  if (GRAD == kGradCorrected) {  // iterates and the sweeps' geometry cotangents in the handle's scratch
    const int64_t ni = cgg_iterate_doubles(D, D + 1, mh->C, sweeps);
    double* scr = afx_mesh_scratch(const_cast<afx_mesh*>(mh), ni + cgg_geo_bar_doubles(D, mh->C, mh->F));
    AFX_REQUIRE(scr, "afx_flow_vjp_geometry: could not allocate the corrected-gradient scratch");
    iters = ni ? scr : nullptr;
    gb.cgg.face = scr + ni;
    gb.cgg.vol = gb.cgg.face + (int64_t)(2 * D + 1) * mh->F;
  }
  int rc = run_vjp<D, ADV, GRAD>(mh, g, a, lam, bar, w, sweeps, relax, s, iters, gb.cgg);
  if (rc) return rc;
  MeshView m;
  if (int rv_ = afx_rview(mh, g, s, &m)) return rv_;
  a.G.v = w.G;
  a.aP.v = w.aP;
  a.psi.v = w.psi;
  if (ADV == kAdvLimited) {
    AFX_LAUNCH(kKFlowGradAdj, (k_flow_geo_lim<D>), grid_for(m.C), dim3(kBlock), 0, s, m, a, w.ext, bar.psi, gb);
    AFX_CHECK_LAUNCH();
  }
  AFX_LAUNCH(kKFlowResAdj, (k_flow_geo_adj<D, ADV, GRAD>), grid_for(m.C), dim3(kBlock), 0, s, m, a, lam, bar.aP,
             bar.G, gb);
  AFX_CHECK_LAUNCH();
  return 0;
}

// Dispatch on (dim, advection, gradient) to a templated body.
template <class Fn>
int dispatch(int D, int adv, int grad, Fn&& fn) {
#define AFX_CASE(DD, AA, GG) \
  if (D == DD && adv == AA && grad == GG) return fn(std::integral_constant<int, DD>{}, std::integral_constant<int, AA>{}, std::integral_constant<int, GG>{});
  AFX_CASE(2, 0, 0) AFX_CASE(2, 1, 0) AFX_CASE(2, 2, 0) AFX_CASE(2, 0, 1) AFX_CASE(2, 1, 1) AFX_CASE(2, 2, 1)
  AFX_CASE(3, 0, 0) AFX_CASE(3, 1, 0) AFX_CASE(3, 2, 0) AFX_CASE(3, 0, 1) AFX_CASE(3, 1, 1) AFX_CASE(3, 2, 1)
  AFX_CASE(2, 0, 3) AFX_CASE(2, 1, 3) AFX_CASE(2, 2, 3) AFX_CASE(3, 0, 3) AFX_CASE(3, 1, 3) AFX_CASE(3, 2, 3)
#undef AFX_CASE
  afx_set_error("unsupported (dim, advection, gradient) combination");
  return 2;
}

int check_common(const afx_mesh* m, const afx_geometry* g, const afx_flow_config* cfg, const afx_flow_inputs* in,
                 void* work, int64_t work_size, int mode) {
  AFX_REQUIRE(m && g && cfg && in, "flow: null argument");
  AFX_REQUIRE(in->state && in->mu && in->rho && in->u_b && in->p_b && cfg->face_kind,
              "flow: a required input pointer is null");
  AFX_REQUIRE(cfg->advection >= 0 && cfg->advection <= 2, "flow: advection must be 0, 1 or 2");
  AFX_REQUIRE(cfg->gradient == 0 || cfg->gradient == 1 || cfg->gradient == 3, "flow: gradient must be 0, 1 or 3");
  AFX_REQUIRE(cfg->gradient != 3 || cfg->gradient_relaxation > 0.0, "flow: gradient relaxation must be positive");
  AFX_REQUIRE(cfg->pressure_pin < (int64_t)m->C, "flow: pressure_pin out of range");
  AFX_REQUIRE(work != nullptr && work_size >= work_bytes(m->dim, m->C, mode), "flow: workspace too small");
  return 0;
}

}  // namespace

extern "C" int afx_flow_workspace_bytes(const afx_mesh* m, int mode, int64_t* bytes) {
  AFX_REQUIRE(m && bytes, "afx_flow_workspace_bytes: null argument");
  AFX_REQUIRE(mode >= 0 && mode <= 3, "afx_flow_workspace_bytes: mode must be 0, 1, 2 or 3");
  *bytes = work_bytes(m->dim, m->C, mode);
  return 0;
}

extern "C" int afx_flow_residual(const afx_mesh* m, const afx_geometry* g, const afx_flow_config* cfg,
                                 const afx_flow_inputs* in, double* residual, void* work, int64_t work_size,
                                 hipStream_t s) {
  int rc = check_common(m, g, cfg, in, work, work_size, 0);
  if (rc) return rc;
  AFX_REQUIRE(residual, "afx_flow_residual: null residual");
  Work w = carve(work, m->dim, m->C, 0);
  FlowArgs<double> a = make_args<double>(cfg, in, nullptr);
  return dispatch(m->dim, cfg->advection, cfg->gradient, [&](auto Dc, auto Ac, auto Gc) {
    return run_forward<decltype(Dc)::value, double, decltype(Ac)::value, decltype(Gc)::value>(
        m, g, a, Out<double>{residual, nullptr}, w.G, nullptr, w.aP, nullptr, w.psi, nullptr, nullptr, w.cgg,
        cfg->gradient_sweeps, cfg->gradient_relaxation, s);
  });
}

extern "C" int afx_flow_jvp(const afx_mesh* m, const afx_geometry* g, const afx_flow_config* cfg,
                            const afx_flow_inputs* in, const afx_flow_inputs* tangent, double* residual,
                            double* residual_tangent, void* work, int64_t work_size, hipStream_t s) {
  int rc = check_common(m, g, cfg, in, work, work_size, 1);
  if (rc) return rc;
  AFX_REQUIRE(tangent && residual && residual_tangent, "afx_flow_jvp: null argument");
  Work w = carve(work, m->dim, m->C, 1);
  FlowArgs<dual> a = make_args<dual>(cfg, in, tangent);
  return dispatch(m->dim, cfg->advection, cfg->gradient, [&](auto Dc, auto Ac, auto Gc) {
    return run_forward<decltype(Dc)::value, dual, decltype(Ac)::value, decltype(Gc)::value>(
        m, g, a, Out<dual>{residual, residual_tangent}, w.G, w.Gt, w.aP, w.aPt, w.psi, w.psit, nullptr, w.cgg,
        cfg->gradient_sweeps, cfg->gradient_relaxation, s);
  });
}

extern "C" int afx_flow_jvp_geometry(const afx_mesh* m, const afx_geometry* g, const afx_geometry* geom_tangent,
                                     const afx_flow_config* cfg, const afx_flow_inputs* in,
                                     const afx_flow_inputs* tangent, double* residual, double* residual_tangent,
                                     void* work, int64_t work_size, hipStream_t s) {
  int rc = check_common(m, g, cfg, in, work, work_size, 1);
  if (rc) return rc;
  AFX_REQUIRE(geom_tangent && tangent && residual && residual_tangent, "afx_flow_jvp_geometry: null argument");
  AFX_REQUIRE(geom_tangent->area && geom_tangent->normal && geom_tangent->face_centroid && geom_tangent->volume &&
                  geom_tangent->cell_centroid,
              "afx_flow_jvp_geometry: the geometry tangent needs all five arrays");
  AFX_REQUIRE(m->fu == nullptr || geom_tangent->face_pack,
              "afx_flow_jvp_geometry: this mesh's faces are reordered: geom_tangent->face_pack must hold the face "
              "tangents in the internal order (afx_geometry_pack of the tangent arrays)");
  Work w = carve(work, m->dim, m->C, 1);
  FlowArgs<dual> a = make_args<dual>(cfg, in, tangent);
  const double* pack = geom_tangent->face_pack;
  a.gt.area = m->fu ? pack : geom_tangent->area;
  a.gt.nrm = m->fu ? pack + m->F : geom_tangent->normal;
  a.gt.fc = m->fu ? pack + (int64_t)(1 + m->dim) * m->F : geom_tangent->face_centroid;
  a.gt.vol = geom_tangent->volume;
  a.gt.cc = geom_tangent->cell_centroid;
  return dispatch(m->dim, cfg->advection, cfg->gradient, [&](auto Dc, auto Ac, auto Gc) {
    constexpr int D = decltype(Dc)::value, ADV = decltype(Ac)::value, GR = decltype(Gc)::value;
    return run_forward<D, dual, ADV, GR, true>(m, g, a, Out<dual>{residual, residual_tangent}, w.G, w.Gt, w.aP,
                                               w.aPt, w.psi, w.psit, nullptr, w.cgg, cfg->gradient_sweeps,
                                               cfg->gradient_relaxation, s);
  });
}

extern "C" int afx_flow_vjp(const afx_mesh* m, const afx_geometry* g, const afx_flow_config* cfg,
                            const afx_flow_inputs* in, const double* cotangent, afx_flow_cotangents* out,
                            void* work, int64_t work_size, hipStream_t s) {
  int rc = check_common(m, g, cfg, in, work, work_size, 2);
  if (rc) return rc;
  AFX_REQUIRE(cotangent && out && out->state && out->mu && out->rho && out->u_b && out->p_b,
              "afx_flow_vjp: null cotangent argument");
  AFX_REQUIRE(in->gamma_b == nullptr || out->gamma_b != nullptr,
              "afx_flow_vjp: gamma_b given without a gamma_b cotangent buffer");
  Work w = carve(work, m->dim, m->C, 2);
  FlowArgs<double> a = make_args<double>(cfg, in, nullptr);
  FlowBars bar{out->state, w.Gb, w.aPb, w.psib, out->mu, out->rho, out->u_b, out->p_b, out->gamma_b, w.extb};
  return dispatch(m->dim, cfg->advection, cfg->gradient, [&](auto Dc, auto Ac, auto Gc) {
    return run_vjp<decltype(Dc)::value, decltype(Ac)::value, decltype(Gc)::value>(
        m, g, a, cotangent, bar, w, cfg->gradient_sweeps, cfg->gradient_relaxation, s);
  });
}

extern "C" int afx_flow_vjp_geometry(const afx_mesh* m, const afx_geometry* g, const afx_flow_config* cfg,
                                     const afx_flow_inputs* in, const double* cotangent, afx_flow_cotangents* out,
                                     afx_geometry_cotangents* geom_out, void* work, int64_t work_size,
                                     hipStream_t s) {
  int rc = check_common(m, g, cfg, in, work, work_size, 3);
  if (rc) return rc;
  AFX_REQUIRE(cotangent && out && out->state && out->mu && out->rho && out->u_b && out->p_b && geom_out,
              "afx_flow_vjp_geometry: null cotangent argument");
  AFX_REQUIRE(in->gamma_b == nullptr || out->gamma_b != nullptr,
              "afx_flow_vjp_geometry: gamma_b given without a gamma_b cotangent buffer");
  AFX_REQUIRE(geom_out->area && geom_out->normal && geom_out->face_centroid && geom_out->volume &&
                  geom_out->cell_centroid,
              "afx_flow_vjp_geometry: all five geometry cotangent buffers are needed");
  Work w = carve(work, m->dim, m->C, 3);
  FlowArgs<double> a = make_args<double>(cfg, in, nullptr);
  FlowBars bar{out->state, w.Gb, w.aPb, w.psib, out->mu, out->rho, out->u_b, out->p_b, out->gamma_b, w.extb};
  FlowGeoBars gb{geom_out->area, geom_out->normal, geom_out->face_centroid, geom_out->volume,
                 geom_out->cell_centroid, w.limx, w.limf, CggGeoBars{nullptr, nullptr}};
  return dispatch(m->dim, cfg->advection, cfg->gradient, [&](auto Dc, auto Ac, auto Gc) {
    constexpr int D = decltype(Dc)::value, ADV = decltype(Ac)::value, GR = decltype(Gc)::value;
    return run_vjp_geometry<D, ADV, GR>(m, g, a, cotangent, bar, gb, w, cfg->gradient_sweeps,
                                        cfg->gradient_relaxation, s);
  });
}

extern "C" int afx_flow_fields(const afx_mesh* m, const afx_geometry* g, const afx_flow_config* cfg,
                               const afx_flow_inputs* in, double* gradients, double* a_p, double* psi,
                               hipStream_t s) {
  AFX_REQUIRE(m && g && cfg && in && gradients && a_p, "afx_flow_fields: null argument");
  AFX_REQUIRE(cfg->advection != 2 || psi, "afx_flow_fields: limited advection needs a psi buffer");
  FlowArgs<double> a = make_args<double>(cfg, in, nullptr);
  double* cgg = nullptr;
  if (cfg->gradient == kGradCorrected) {
    cgg = afx_mesh_scratch(const_cast<afx_mesh*>(m), (int64_t)2 * (m->dim + 1) * m->dim * m->C);
    AFX_REQUIRE(cgg, "afx_flow_fields: could not allocate the corrected-gradient sweep scratch");
  }
  return dispatch(m->dim, cfg->advection, cfg->gradient, [&](auto Dc, auto Ac, auto Gc) {
    return run_forward<decltype(Dc)::value, double, decltype(Ac)::value, decltype(Gc)::value>(
        m, g, a, Out<double>{nullptr, nullptr}, gradients, nullptr, a_p, nullptr, psi, nullptr, nullptr, cgg,
        cfg->gradient_sweeps, cfg->gradient_relaxation, s);
  });
}

extern "C" int afx_flow_mass_flux(const afx_mesh* m, const afx_geometry* g, const afx_flow_config* cfg,
                                  const afx_flow_inputs* in, const afx_flow_inputs* tangent, double* mdot,
                                  double* mdot_tangent, void* work, int64_t work_size, hipStream_t s) {
  const int mode = tangent ? 1 : 0;
  int rc = check_common(m, g, cfg, in, work, work_size, mode);
  if (rc) return rc;
  AFX_REQUIRE(mdot, "afx_flow_mass_flux: null mdot");
  AFX_REQUIRE(!tangent || mdot_tangent, "afx_flow_mass_flux: a tangent needs an mdot_tangent buffer");
  Work w = carve(work, m->dim, m->C, mode);
  if (!tangent) {
    FlowArgs<double> a = make_args<double>(cfg, in, nullptr);
    return dispatch(m->dim, cfg->advection, cfg->gradient, [&](auto Dc, auto Ac, auto Gc) {
      return run_forward<decltype(Dc)::value, double, decltype(Ac)::value, decltype(Gc)::value>(
          m, g, a, Out<double>{nullptr, nullptr}, w.G, nullptr, w.aP, nullptr, w.psi, nullptr, nullptr, w.cgg,
          cfg->gradient_sweeps, cfg->gradient_relaxation, s, Out<double>{mdot, nullptr});
    });
  }
  FlowArgs<dual> a = make_args<dual>(cfg, in, tangent);
  return dispatch(m->dim, cfg->advection, cfg->gradient, [&](auto Dc, auto Ac, auto Gc) {
    return run_forward<decltype(Dc)::value, dual, decltype(Ac)::value, decltype(Gc)::value>(
        m, g, a, Out<dual>{nullptr, nullptr}, w.G, w.Gt, w.aP, w.aPt, w.psi, w.psit, nullptr, w.cgg,
        cfg->gradient_sweeps, cfg->gradient_relaxation, s, Out<dual>{mdot, mdot_tangent});
  });
}
