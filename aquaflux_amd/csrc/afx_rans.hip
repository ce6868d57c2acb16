// Coupled RANS residual R(u, v[, w], p, k, omega) with the k-omega SST closure: forward, tangent
// (dual instantiation) and adjoint (hand-written reverse passes).
//
// Reference: CoupledRANS.residual (turbulence/coupled.py:437-473) = closure_fields
// (turbulence/transport.py:508-546) -> momentum.with_eddy_viscosity(nu_t, nu_t,wall) ->
// flow_fields / residual_from_fields (flow/momentum.py:822-878) -> k_residual / omega_residual
// (turbulence/transport.py:581-718) on the Rhie-Chow mass flux over the density.
//
// Passes (one thread per cell, gather over the cell's faces, no atomics):
//   G  gradients of u_i, p, k, omega (+ Venkatakrishnan psi of u_i, k, omega)
//   C  cell-local SST closure: nu_t, F1, strain, cross-diffusion, wall-function nu_t, wall shear rate
//   D  momentum diagonal with mu_eff = mu + rho nu_t and the wall-function boundary viscosity
//   R  momentum / continuity / k / omega fluxes, SST sources, omega wall-cell fixation
// The adjoint runs G, C, D forward, then R', D', C' (closure and limiter adjoints by seeded duals,
// face fluxes by hand), G'.
#include <memory>
#include <type_traits>
#include <vector>

#include <omp.h>

#include "afx_common.h"
#include "afx_mesh.h"
#include "afx_flow_face.h"
#include "afx_scalar_face.h"
#include "afx_scalar_adj.h"
#include "afx_sst.h"
#include "afx_profile.h"
#include "afx_cgg.h"

using namespace afx;

namespace {

template <class S>
struct RansArgs {
  const int8_t* kind;   // flow boundary kind per face
  const int8_t* kbc;    // k closure per face (Dirichlet / ZeroGradient)
  const int8_t* wbc;    // omega closure per face
  const int8_t* wallf;  // 1 on the SST wall faces
  const int8_t* wallc;  // 1 on wall-adjacent cells
  In<S> x, ub, pb, mu, rho, nu, kval, wval;
  const double* dwall;
  const double* dgrad;
  const double* dwall_t;  // tangents of the wall distance and its gradient (GEO instantiations only)
  const double* dgrad_t;
  In<S> G, psi, nut, f1, Sr, cd, nutw, tau, mueff, aP;
  int32_t pin;
  double pin_value;
  double bf[3];
  double lim_k;
  SstConst sc;
  int sweeps;    // corrected Green-Gauss sweep count
  double relax;  // ... and relaxation
  int log_w;     // omega solved as log(omega): the wall fixation row is log(omega / omega_wall)
  GeoTan gt;     // geometry tangents (read only by the GEO instantiations of the forward passes)
};

// closure arrays (written by pass C)
struct ClosureOut {
  double *nut, *f1, *Sr, *cd, *nutw, *tau, *mueff;
  double *nut_t, *f1_t, *Sr_t, *cd_t, *nutw_t, *tau_t, *mueff_t;
  double* lsqI;  // nullable [D][D][C]: the least-squares fit's inverse matrix, kept for the closure adjoint
};

template <int D, class S>
AFX_DEV S rans_bval_lead(const RansArgs<S>& a, int v, int32_t F, int32_t f, const S* xc) {
  // leading-order boundary value (zero owner gradient) of variable v
  if (v <= D) {
    const int kind = a.kind[f];
    if (v < D) {
      if (kind == kFlowWall || kind == kFlowInlet) return a.ub[(int64_t)v * F + f];
      if (kind == kFlowOutlet) return xc[v];
      return sconst<S>(0.0);
    }
    if (kind == kFlowWall || kind == kFlowInlet) return xc[D];
    if (kind == kFlowOutlet) return a.pb[f];
    return sconst<S>(0.0);
  }
  const int8_t bk = v == D + 1 ? a.kbc[f] : a.wbc[f];
  if (bk == kBcDirichlet) return v == D + 1 ? a.kval[f] : a.wval[f];
  if (bk == kBcZeroGradient) return xc[v];
  return sconst<S>(0.0);
}

// limiter slot of variable v (-1 when not limited): u_i -> i, k -> D, omega -> D + 1
template <int D>
AFX_DEV int lim_slot(int v, bool limf, bool lims) {
  if (v < D) return limf ? v : -1;
  if (v == D) return -1;
  return lims ? v - 1 : -1;
}

// Wall distance and its gradient at cell c: values, or (GEO) (value, tangent) pairs.
template <bool GEO, class S>
AFX_DEV typename std::conditional<GEO, dual, double>::type wall_dist(const RansArgs<S>& a, int32_t c) {
  if constexpr (GEO)
    return dual{a.dwall[c], a.dwall_t[c]};
  else
    return a.dwall[c];
}
template <bool GEO, class S>
AFX_DEV typename std::conditional<GEO, dual, double>::type wall_dist_grad(const RansArgs<S>& a, int32_t C, int j,
                                                                          int32_t c) {
  const int64_t i = (int64_t)j * C + c;
  if constexpr (GEO)
    return dual{a.dgrad[i], a.dgrad_t[i]};
  else
    return a.dgrad[i];
}

// The SST closure of cell c from its k, omega and gradients (pass C, defined below).
template <int D, class S, bool GEO = false>
AFX_DEV void closure_cell(const MeshView& m, const RansArgs<S>& a, int32_t c, S k, S w, const S (&Gu)[D][D],
                          const S (&Gk)[D], const S (&Gw)[D], const ClosureOut& out);

// ---------------------------------------------------------------------------------------------
// pass G.  With CLO the cell-local closure (pass C) follows in the same thread from the gradient
// still in registers, saving the closure pass's re-read of k, omega and the gradients.  With GEO
// the geometry is read as (value, tangent) pairs (the node-coordinate JVP; S = dual), centroids
// from global memory rather than the tile.
// LDS staging of k_rans_grad: the block's states and centroids.
template <int D, class S>
struct RansGradSh {
  S tx[D + 3][kBlock];
  double tcc[D][kBlock];
};

template <int D, class S, int GRAD, bool LIMF, bool LIMS, bool CLO = false, bool GEO = false, class Exec, class Sh>
AFX_DEV void k_rans_grad_body(const Exec& ex, Sh& sh, MeshView m, RansArgs<S> a, Out<S> G, Out<S> psi_out, int32_t* __restrict__ ext, int32_t* __restrict__ kmin_out, ClosureOut co) {
  constexpr int NV = D + 3, NL = D + 2;
  using GR = GeoRead<D, GEO>;
  using GT = typename GR::T;
  auto& tx = sh.tx;
  auto& tcc = sh.tcc;
  const int32_t base = ex.base;
  const int t = ex.t;
  const int32_t c = base + t;
  const bool valid = c < m.C;
  S xc[NV];
  GT ccc[D];
  if (valid) {
#pragma unroll
    for (int v = 0; v < NV; ++v) tx[v][t] = xc[v] = a.x[(int64_t)v * m.C + c];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      ccc[j] = GR::cc(m, a.gt, j, c);
      tcc[j][t] = val(ccc[j]);
    }
  }
  ex.sync();
  if (!valid) return;
  S acc[NV][D];
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int j = 0; j < D; ++j) acc[v][j] = sconst<S>(0.0);
  GT M[D][D];
#pragma unroll
  for (int p = 0; p < D; ++p)
#pragma unroll
    for (int q = 0; q < D; ++q) M[p][q] = sconst<GT>(0.0);
  S mx[NL], mn[NL];
  int32_t amx[NL], amn[NL];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int l = lim_slot<D>(v, LIMF, LIMS);
    if (l >= 0) {
      mx[l] = xc[v];
      mn[l] = xc[v];
      amx[l] = c;
      amn[l] = c;
    }
  }
  const int32_t kb = m.cptr[c], ke = m.cptr[c + 1];
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    if (o >= 0) {
      S xo[NV];
      GT cco[D];
      const int ls = ex.slot(o);
      if (ls >= 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) xo[v] = tx[v][ls];
        if constexpr (GEO) {
#pragma unroll
          for (int j = 0; j < D; ++j) cco[j] = GR::cc(m, a.gt, j, o);
        } else {
#pragma unroll
          for (int j = 0; j < D; ++j) cco[j] = tcc[j][ls];
        }
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) xo[v] = a.x[(int64_t)v * m.C + o];
#pragma unroll
        for (int j = 0; j < D; ++j) cco[j] = GR::cc(m, a.gt, j, o);
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int l = lim_slot<D>(v, LIMF, LIMS);
        if (l >= 0) {
          if (val(xo[v]) > val(mx[l])) {
            mx[l] = xo[v];
            amx[l] = o;
          }
          if (val(xo[v]) < val(mn[l])) {
            mn[l] = xo[v];
            amn[l] = o;
          }
        }
      }
      if (GRAD == 0) {
        FaceGeo<D, GT> q;
        GR::face_pts(m, a.gt, r.f, true, r.own ? ccc : cco, r.own ? cco : ccc, q);
        const GT w = 1.0 - q.g;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          S phi = w * (r.own ? xc[v] : xo[v]) + q.g * (r.own ? xo[v] : xc[v]);
#pragma unroll
          for (int j = 0; j < D; ++j) {
            S t = phi * (q.n[j] * q.A);
            acc[v][j] = r.own ? acc[v][j] + t : acc[v][j] - t;
          }
        }
      } else {
        GT dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = cco[j] - ccc[j];
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
      if (GRAD == 0) {
        const GT A = GR::area(m, a.gt, r.f);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          S phi = rans_bval_lead<D, S>(a, v, m.F, r.fu, xc);
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
          S dphi = rans_bval_lead<D, S>(a, v, m.F, r.fu, xc) - xc[v];
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
    if constexpr (!GEO) {
      if (co.lsqI) {
#pragma unroll
        for (int p = 0; p < D; ++p)
#pragma unroll
          for (int q2 = 0; q2 < D; ++q2) co.lsqI[(int64_t)(p * D + q2) * m.C + c] = I[p][q2];
      }
    }
  }
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int j = 0; j < D; ++j) G.set((int64_t)(v * D + j) * m.C + c, grad[v][j]);
  if (LIMF || LIMS) {
    const GT e2 = GR::vol(m, a.gt, c) * (a.lim_k * a.lim_k * a.lim_k);
    GT xcen[D];
#pragma unroll
    for (int j = 0; j < D; ++j) xcen[j] = GR::cc(m, a.gt, j, c);
    S lo[NL];
    int32_t ks[NL];
    bool first = true;
    for (int32_t k = kb; k < ke; ++k) {
      const int32_t e = m.cface[k];
      const int32_t f = e >= 0 ? e : ~e;
      GT dxf[D];
#pragma unroll
      for (int j = 0; j < D; ++j) dxf[j] = GR::fc(m, a.gt, j, f) - xcen[j];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int l = lim_slot<D>(v, LIMF, LIMS);
        if (l < 0) continue;
        S dm = grad[v][0] * dxf[0];
#pragma unroll
        for (int j = 1; j < D; ++j) dm = dm + grad[v][j] * dxf[j];
        S vf = venkat_face<S, GT>(dm, mx[l] - xc[v], mn[l] - xc[v], e2);
        if (first || val(vf) < val(lo[l])) {
          lo[l] = vf;
          ks[l] = f;
        }
      }
      first = false;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int l = lim_slot<D>(v, LIMF, LIMS);
      if (l < 0) continue;
      psi_out.set((int64_t)l * m.C + c, dmin(dmax(lo[l], 0.0), 1.0));
      if (ext) {
        ext[(int64_t)l * m.C + c] = amx[l];
        ext[(int64_t)(NL + l) * m.C + c] = amn[l];
      }
      if (kmin_out) kmin_out[(int64_t)l * m.C + c] = ks[l];
    }
  }
  if constexpr (CLO) {
    S Gu[D][D], Gk[D], Gw[D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) Gu[i][j] = grad[i][j];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      Gk[j] = grad[D + 1][j];
      Gw[j] = grad[D + 2][j];
    }
    closure_cell<D, S, GEO>(m, a, c, xc[D + 1], xc[D + 2], Gu, Gk, Gw, co);
  }
}

template <int D, class S, int GRAD, bool LIMF, bool LIMS, bool CLO = false, bool GEO = false>
__global__ void __launch_bounds__(256) k_rans_grad(MeshView m, RansArgs<S> a, Out<S> G, Out<S> psi_out,
                                                   int32_t* __restrict__ ext, int32_t* __restrict__ kmin_out,
                                                   ClosureOut co = ClosureOut{}) {
  __shared__ RansGradSh<D, S> sh;
  k_rans_grad_body<D, S, GRAD, LIMF, LIMS, CLO, GEO>(DevExec(m), sh, m, a, G, psi_out, ext, kmin_out, co);
}

// ---------------------------------------------------------------------------------------------
// pass C: closure
template <int D, class S>
AFX_DEV void load_closure_inputs(const RansArgs<S>& a, const MeshView& m, int32_t c, S& k, S& w, S (&Gu)[D][D],
                                 S (&Gk)[D], S (&Gw)[D]) {
  k = a.x[(int64_t)(D + 1) * m.C + c];
  w = a.x[(int64_t)(D + 2) * m.C + c];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) Gu[i][j] = a.G[(int64_t)(i * D + j) * m.C + c];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    Gk[j] = a.G[(int64_t)((D + 1) * D + j) * m.C + c];
    Gw[j] = a.G[(int64_t)((D + 2) * D + j) * m.C + c];
  }
}

// Wall shear rate of a wall cell: area-weighted |u_P - u_b| / d over its SST wall faces.
template <int D, class S, bool GEO = false>
AFX_DEV S wall_shear_rate(const RansArgs<S>& a, const MeshView& m, int32_t c, typename GeoRead<D, GEO>::T d) {
  using GT = typename GeoRead<D, GEO>::T;
  S num = sconst<S>(0.0);
  GT den = sconst<GT>(0.0);
  S xc[D + 1];
#pragma unroll
  for (int v = 0; v <= D; ++v) xc[v] = a.x[(int64_t)v * m.C + c];
  for (int32_t k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
    const int32_t e = m.cface[k];
    if (e < 0) continue;
    const int32_t eu = user_face(m, e);
    if (!a.wallf[eu]) continue;
    S s2 = sconst<S>(0.0);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      S sl = xc[j] - rans_bval_lead<D, S>(a, j, m.F, eu, xc);
      s2 = j == 0 ? sl * sl : s2 + sl * sl;
    }
    const GT A = GeoRead<D, GEO>::area(m, a.gt, e);
    num = num + (safe_sqrt(s2) / d) * A;
    den += A;
  }
  return val(den) > 0.0 ? num / den : sconst<S>(0.0);
}

template <int D, class S, bool GEO>
AFX_DEV void closure_cell(const MeshView& m, const RansArgs<S>& a, int32_t c, S k, S w, const S (&Gu)[D][D],
                          const S (&Gk)[D], const S (&Gw)[D], const ClosureOut& out) {
  const bool wall = a.wallc[c] != 0;
  using GT = typename GeoRead<D, GEO>::T;
  const GT d = wall_dist<GEO>(a, c);
  GT dg[D];
#pragma unroll
  for (int j = 0; j < D; ++j) dg[j] = wall ? wall_dist_grad<GEO>(a, m.C, j, c) : sconst<GT>(0.0);
  const S nu = a.nu[c];
  SstClosure<S> o = sst_closure<D, S, GT>(a.sc, k, w, Gu, Gk, Gw, nu, d, dg, wall);
  S tau = wall ? wall_shear_rate<D, S, GEO>(a, m, c, d) : sconst<S>(0.0);
  S mueff = a.mu[c] + a.rho[c] * o.nut;
  Out<S>{out.nut, out.nut_t}.set(c, o.nut);
  Out<S>{out.f1, out.f1_t}.set(c, o.f1);
  Out<S>{out.Sr, out.Sr_t}.set(c, o.strain);
  Out<S>{out.cd, out.cd_t}.set(c, o.cd);
  Out<S>{out.nutw, out.nutw_t}.set(c, o.nutw);
  Out<S>{out.tau, out.tau_t}.set(c, tau);
  Out<S>{out.mueff, out.mueff_t}.set(c, mueff);
}

template <int D, class S, bool GEO = false>
AFX_DEV void k_rans_closure_body(int32_t c, MeshView m, RansArgs<S> a, ClosureOut out) {
  if (c >= m.C) return;
  S k, w, Gu[D][D], Gk[D], Gw[D];
  load_closure_inputs<D, S>(a, m, c, k, w, Gu, Gk, Gw);
  closure_cell<D, S, GEO>(m, a, c, k, w, Gu, Gk, Gw, out);
}

template <int D, class S, bool GEO = false>
__global__ void __launch_bounds__(256) k_rans_closure(MeshView m, RansArgs<S> a, ClosureOut out) {
  k_rans_closure_body<D, S, GEO>(xcd_thread_index(m), m, a, out);
}

// ---------------------------------------------------------------------------------------------
// shared loaders for passes D / R
template <int D, class S, bool GEO = false, class GT = typename GeoRead<D, GEO>::T>
AFX_DEV void load_flow(const RansArgs<S>& a, const MeshView& m, int32_t c, bool with_ap, bool with_psi,
                       FlowCell<D, S, GT>& X) {
#pragma unroll
  for (int i = 0; i < D; ++i) X.u[i] = a.x[(int64_t)i * m.C + c];
  X.p = a.x[(int64_t)D * m.C + c];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) X.Gu[i][j] = a.G[(int64_t)(i * D + j) * m.C + c];
#pragma unroll
  for (int j = 0; j < D; ++j) X.Gp[j] = a.G[(int64_t)(D * D + j) * m.C + c];
  X.mu = a.mueff[c];
  X.rho = a.rho[c];
  X.V = GeoRead<D, GEO>::vol(m, a.gt, c);
  X.aP = with_ap ? a.aP[c] : sconst<S>(1.0);
#pragma unroll
  for (int i = 0; i < D; ++i) X.psi[i] = with_psi ? a.psi[(int64_t)i * m.C + c] : sconst<S>(1.0);
}

// The momentum boundary-face viscosity: mu + rho nu_t,wall on shearing walls (flow/momentum.py:351-418).
template <class S>
AFX_DEV S rans_gamma_b(const RansArgs<S>& a, int kind, int32_t f, int32_t P, S mueffP) {
  if (kind == kFlowWall) return a.mu[P] + a.rho[P] * (a.wallf[f] ? a.nutw[P] : sconst<S>(0.0));
  return mueffP;
}

template <int D, class S>
struct ScalarCell {
  S phi, G[D], psi, gam;
};

template <int D, class S>
AFX_DEV void load_scalar(const RansArgs<S>& a, const MeshView& m, int32_t c, int var, bool with_psi, double s1,
                         double s2, ScalarCell<D, S>& Z) {
  Z.phi = a.x[(int64_t)var * m.C + c];
#pragma unroll
  for (int j = 0; j < D; ++j) Z.G[j] = a.G[(int64_t)(var * D + j) * m.C + c];
  Z.psi = with_psi ? a.psi[(int64_t)(var - 1) * m.C + c] : sconst<S>(1.0);
  Z.gam = a.nu[c] + sst_blend(a.f1[c], s1, s2) * a.nut[c];
}

// LDS copy of a workgroup's k and omega cells (value, gradient, limiter, diffusivity), centroids and,
// with LAM, the k row's cotangent and the omega row's flux cotangent (zero on wall cells).
template <int D, class S, bool LAM>
struct ScalarTile {
  S f[2 * (D + 3)][256];
  double cc[D][256];
  double lam[LAM ? 2 : 1][LAM ? 256 : 1];
  AFX_DEV static void put1(S (*f)[256], int t, const ScalarCell<D, S>& Z) {
    f[0][t] = Z.phi;
#pragma unroll
    for (int j = 0; j < D; ++j) f[1 + j][t] = Z.G[j];
    f[D + 1][t] = Z.psi;
    f[D + 2][t] = Z.gam;
  }
  AFX_DEV static void get1(const S (*f)[256], int t, ScalarCell<D, S>& Z) {
    Z.phi = f[0][t];
#pragma unroll
    for (int j = 0; j < D; ++j) Z.G[j] = f[1 + j][t];
    Z.psi = f[D + 1][t];
    Z.gam = f[D + 2][t];
  }
  AFX_DEV void put(int t, const ScalarCell<D, S>& K, const ScalarCell<D, S>& W, const double* xc) {
    put1(f, t, K);
    put1(f + D + 3, t, W);
#pragma unroll
    for (int j = 0; j < D; ++j) cc[j][t] = xc[j];
  }
  AFX_DEV void get(int t, ScalarCell<D, S>& K, ScalarCell<D, S>& W, double* xo) const {
    get1(f, t, K);
    get1(f + D + 3, t, W);
#pragma unroll
    for (int j = 0; j < D; ++j) xo[j] = cc[j][t];
  }
};

// ---------------------------------------------------------------------------------------------
// pass D
// With GEO the geometry is read as (value, tangent) pairs and every cell from global memory.
template <int D, class S>
struct RansApSh {
  FlowTile<D, S, false> tile;
};

template <int D, class S, int ADV, bool GEO = false, class Exec, class Sh>
AFX_DEV void k_rans_ap_body(const Exec& ex, Sh& sh, MeshView m, RansArgs<S> a, Out<S> aP) {
  using GR = GeoRead<D, GEO>;
  using GT = typename GR::T;
  auto& tile = sh.tile;
  const int32_t base = ex.base;
  const int t = ex.t;
  const int32_t c = base + t;
  const bool valid = c < m.C;
  FlowCell<D, S, GT> X;
  GT xc[D];
  if (valid) {
    load_flow<D, S, GEO>(a, m, c, false, false, X);
#pragma unroll
    for (int j = 0; j < D; ++j) xc[j] = GR::cc(m, a.gt, j, c);
    if constexpr (!GEO) tile.put(t, X, xc);
  }
  if constexpr (!GEO) ex.sync();
  if (!valid) return;
  S acc = sconst<S>(0.0);
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D, GT> q;
    if (o >= 0) {
      FlowCell<D, S, GT> O;
      GT xo[D];
      int lo = -1;
      if constexpr (!GEO) lo = ex.slot(o);
      if (lo >= 0) {
        if constexpr (!GEO) tile.get(lo, O, xo);
      } else {
        load_flow<D, S, GEO>(a, m, o, false, false, O);
#pragma unroll
        for (int j = 0; j < D; ++j) xo[j] = GR::cc(m, a.gt, j, o);
      }
      GR::face_pts(m, a.gt, r.f, true, r.own ? xc : xo, r.own ? xo : xc, q);
      S oc, nc;
      if (r.own)
        ap_interior<D, ADV, S>(q, X, O, oc, nc);
      else
        ap_interior<D, ADV, S>(q, O, X, oc, nc);
      acc += r.own ? oc : nc;
    } else {
      GR::face_pts(m, a.gt, r.f, false, xc, xc, q);
      const int kind = a.kind[r.fu];
      acc += ap_boundary<D, ADV, S>(q, kind, X, rans_gamma_b<S>(a, kind, r.fu, c, X.mu));
    }
  }
  aP.set(c, acc);
}

template <int D, class S, int ADV, bool GEO = false>
__global__ void __launch_bounds__(256) k_rans_ap(MeshView m, RansArgs<S> a, Out<S> aP) {
  __shared__ RansApSh<D, S> sh;
  k_rans_ap_body<D, S, ADV, GEO>(DevExec(m), sh, m, a, aP);
}

// ---------------------------------------------------------------------------------------------
// pass R, flow part: momentum and continuity rows, and the Rhie-Chow face mass flux (written by
// each face's owner) that the k / omega fluxes advect on.
template <int D>
struct RansFresSh {
  FlowTile<D, double, true> tile;
};

template <int D, class S, int ADVF, bool GEO = false, class Exec, class Sh>
AFX_DEV void k_rans_fres_body(const Exec& ex, Sh& sh, MeshView m, RansArgs<S> a, Out<S> R, Out<S> mdot) {
  constexpr int NV = D + 1;
  using GR = GeoRead<D, GEO>;
  using GT = typename GR::T;
  // the FP64 pass stages its block's cells (53 KB of LDS in 3D); the dual-number (tangent) pass,
  // whose tile would be twice that, gathers from global memory
  constexpr bool kStage = std::is_same<S, double>::value && !GEO;
  auto& tile = sh.tile;
  const int32_t base = ex.base;
  const int t = ex.t;
  const int32_t c = base + t;
  const bool valid = c < m.C;
  FlowCell<D, S, GT> X;
  GT xc[D];
  if (valid) {
    load_flow<D, S, GEO>(a, m, c, true, ADVF == kAdvLimited, X);
#pragma unroll
    for (int j = 0; j < D; ++j) xc[j] = GR::cc(m, a.gt, j, c);
    if constexpr (kStage) tile.put(t, X, xc);
  }
  if constexpr (kStage) ex.sync();
  if (!valid) return;
  S acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = sconst<S>(0.0);
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D, GT> q;
    S out[NV];
    if (o >= 0) {
      FlowCell<D, S, GT> O;
      GT xo[D];
      int lo = -1;
      if constexpr (kStage) lo = ex.slot(o);
      if (lo >= 0) {
        if constexpr (kStage) tile.get(lo, O, xo);
      } else {
        load_flow<D, S, GEO>(a, m, o, true, ADVF == kAdvLimited, O);
#pragma unroll
        for (int j = 0; j < D; ++j) xo[j] = GR::cc(m, a.gt, j, o);
      }
      GR::face_pts(m, a.gt, r.f, true, r.own ? xc : xo, r.own ? xo : xc, q);
      if (r.own)
        flow_flux_interior<D, ADVF, S>(q, X, O, out);
      else
        flow_flux_interior<D, ADVF, S>(q, O, X, out);
    } else {
      GR::face_pts(m, a.gt, r.f, false, xc, xc, q);
      S ub[D];
#pragma unroll
      for (int i = 0; i < D; ++i) ub[i] = a.ub[(int64_t)i * m.F + r.fu];
      const int kind = a.kind[r.fu];
      flow_flux_boundary<D, ADVF, S>(q, kind, X, ub, a.pb[r.fu], rans_gamma_b<S>(a, kind, r.fu, c, X.mu), out);
    }
    if (r.own) mdot.set(r.f, out[D]);
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = r.own ? acc[v] + out[v] : acc[v] - out[v];
  }
  if (R.v == nullptr) return;
#pragma unroll
  for (int i = 0; i < D; ++i) R.set((int64_t)i * m.C + c, acc[i] - a.bf[i] * X.V);
  S cont = acc[D];
  if (c == a.pin) cont = X.p - a.pin_value;
  R.set((int64_t)D * m.C + c, cont);
}

template <int D, class S, int ADVF, bool GEO = false>
__global__ void __launch_bounds__(256) k_rans_fres(MeshView m, RansArgs<S> a, Out<S> R, Out<S> mdot) {
  __shared__ RansFresSh<D> sh;
  k_rans_fres_body<D, S, ADVF, GEO>(DevExec(m), sh, m, a, R, mdot);
}

// k and omega face fluxes of one face seen from cell c (all scalar operands in registers).
template <int D, class S, int ADVS, class GT = double>
AFX_DEV void rans_scalar_face(const RansArgs<S>& a, const MeshView& m, const FaceRef& r, int32_t o,
                              const FaceGeo<D, GT>& q, S Qf, const ScalarCell<D, S>& Kc, const ScalarCell<D, S>& Wc,
                              const ScalarCell<D, S>& Ko, const ScalarCell<D, S>& Wo, GT dP, S& fk, S& fw) {
  const SstConst& sc = a.sc;
  if (o >= 0) {
    if (r.own) {
      fk = scalar_flux_interior<D, ADVS, S>(q, Qf, Kc.phi, Ko.phi, Kc.G, Ko.G, Kc.psi, Ko.psi, Kc.gam, Ko.gam);
      fw = scalar_flux_interior<D, ADVS, S>(q, Qf, Wc.phi, Wo.phi, Wc.G, Wo.G, Wc.psi, Wo.psi, Wc.gam, Wo.gam);
    } else {
      fk = scalar_flux_interior<D, ADVS, S>(q, Qf, Ko.phi, Kc.phi, Ko.G, Kc.G, Ko.psi, Kc.psi, Ko.gam, Kc.gam);
      fw = scalar_flux_interior<D, ADVS, S>(q, Qf, Wo.phi, Wc.phi, Wo.G, Wc.G, Wo.psi, Wc.psi, Wo.gam, Wc.gam);
    }
  } else {
    const S zero = sconst<S>(0.0);
    const S bk = scalar_bval<D, S>(a.kbc[r.fu], q, Kc.phi, Kc.G, a.kval[r.fu], zero, zero);
    S gk = Kc.gam;
    if (a.wallf[r.fu]) gk = (1.0 - wall_function_weight<S>(sc, a.nu[r.P], dP, Kc.phi)) * Kc.gam;
    fk = scalar_flux_boundary<D, ADVS, S>(q, Qf, Kc.phi, Kc.G, Kc.psi, bk, gk);
    const S bw = scalar_bval<D, S>(a.wbc[r.fu], q, Wc.phi, Wc.G, a.wval[r.fu], zero, zero);
    fw = scalar_flux_boundary<D, ADVS, S>(q, Qf, Wc.phi, Wc.G, Wc.psi, bw, Wc.gam);
  }
}

// pass R, k / omega part: transport on mdot / density, SST sources, omega wall fixation.
template <int D, class S>
struct RansSresSh {
  ScalarTile<D, S, false> tile;
};

template <int D, class S, int ADVS, bool GEO = false, class Exec, class Sh>
AFX_DEV void k_rans_sres_body(const Exec& ex, Sh& sh, MeshView m, RansArgs<S> a, In<S> mdot, Out<S> R) {
  using GR = GeoRead<D, GEO>;
  using GT = typename GR::T;
  const SstConst& sc = a.sc;
  auto& tile = sh.tile;
  const int32_t base = ex.base;
  const int t = ex.t;
  const int32_t c = base + t;
  const bool valid = c < m.C;
  ScalarCell<D, S> Kc, Wc;
  GT xc[D];
  if (valid) {
    load_scalar<D, S>(a, m, c, D + 1, ADVS == kAdvLimited, sc.sk1, sc.sk2, Kc);
    load_scalar<D, S>(a, m, c, D + 2, ADVS == kAdvLimited, sc.sw1, sc.sw2, Wc);
#pragma unroll
    for (int j = 0; j < D; ++j) xc[j] = GR::cc(m, a.gt, j, c);
    if constexpr (!GEO) tile.put(t, Kc, Wc, xc);
  }
  if constexpr (!GEO) ex.sync();
  if (!valid) return;
  S acck = sconst<S>(0.0), accw = sconst<S>(0.0);
  const double irho = 1.0 / sc.density;
  const GT dP = wall_dist<GEO>(a, c);
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D, GT> q;
    ScalarCell<D, S> Ko, Wo;
    if (o >= 0) {
      GT xo[D];
      int lo = -1;
      if constexpr (!GEO) lo = ex.slot(o);
      if (lo >= 0) {
        if constexpr (!GEO) tile.get(lo, Ko, Wo, xo);
      } else {
        load_scalar<D, S>(a, m, o, D + 1, ADVS == kAdvLimited, sc.sk1, sc.sk2, Ko);
        load_scalar<D, S>(a, m, o, D + 2, ADVS == kAdvLimited, sc.sw1, sc.sw2, Wo);
#pragma unroll
        for (int j = 0; j < D; ++j) xo[j] = GR::cc(m, a.gt, j, o);
      }
      GR::face_pts(m, a.gt, r.f, true, r.own ? xc : xo, r.own ? xo : xc, q);
    } else {
      GR::face_pts(m, a.gt, r.f, false, xc, xc, q);
    }
    const S Qf = mdot[r.f] * irho;
    S fk, fw;
    rans_scalar_face<D, S, ADVS>(a, m, r, o, q, Qf, Kc, Wc, Ko, Wo, dP, fk, fw);
    acck = r.own ? acck + fk : acck - fk;
    accw = r.own ? accw + fw : accw - fw;
  }
  const bool wall = a.wallc[c] != 0;
  S pk, dk, pw, dw, cw;
  sst_sources<S>(sc, Kc.phi, Wc.phi, a.nut[c], a.Sr[c], a.f1[c], a.cd[c], a.tau[c], a.nu[c], dP, wall,
                 GR::vol(m, a.gt, c), pk, dk, pw, dw, cw);
  R.set((int64_t)(D + 1) * m.C + c, (acck - pk) - dk);
  S rw = ((accw - pw) - dw) - cw;
  if (wall) {
    const S t2 = omega_wall_k<S>(sc, a.nu[c], dP, Kc.phi);
    rw = a.log_w ? dlog(Wc.phi / t2) : Wc.phi - t2;
  }
  R.set((int64_t)(D + 2) * m.C + c, rw);
}

template <int D, class S, int ADVS, bool GEO = false>
__global__ void __launch_bounds__(256) k_rans_sres(MeshView m, RansArgs<S> a, In<S> mdot, Out<S> R) {
  __shared__ RansSresSh<D, S> sh;
  k_rans_sres_body<D, S, ADVS, GEO>(DevExec(m), sh, m, a, mdot, R);
}

// ---------------------------------------------------------------------------------------------
// adjoint storage
struct RansBars {
  double* x;      // [(D+3)][C]   (output)
  double* mu;     // [C]          (output: molecular dynamic viscosity)
  double* rho;    // [C]          (output)
  double* nu;     // [C]          (output: molecular kinematic viscosity)
  double* ub;     // [D][F]       (output)
  double* pb;     // [F]          (output)
  double* G;      // [(D+3)D][C]  scratch
  double* aP;     // [C]
  double* psi;    // [D+2][C]
  double* ext;    // [2(D+2)][C]
  double* mueff;  // [C]
  double* nut;    // [C]
  double* f1;
  double* Sr;
  double* cd;
  double* tau;
  double* nutw;
};

template <int D>
struct RansCellBar {
  FlowCellBar<D> fl;  // fl.mu = mu_eff cotangent
  ScalarBar<D> k, w;
  double nut, f1, Sr, cd, tau, nutw, mum, nu;
  AFX_DEV void zero() {
    fl.zero();
    k.zero();
    w.zero();
    nut = f1 = Sr = cd = tau = nutw = mum = nu = 0.0;
  }
};

// gamma = nu + blend(f1, s1, s2) nu_t
AFX_DEV void adj_gamma(double gam_bar, double f1, double nut, double s1, double s2, double& nu_b, double& f1_b,
                       double& nut_b) {
  nu_b += gam_bar;
  f1_b += gam_bar * (s1 - s2) * nut;
  nut_b += gam_bar * (f1 * s1 + (1.0 - f1) * s2);
}

// Routes the cotangent of the momentum boundary viscosity (B has fl, mum, nutw).
template <int D, class BarT>
AFX_DEV void adj_gamma_b(const RansArgs<double>& a, int kind, int32_t f, int32_t P, double gbb, BarT& B) {
  if (kind == kFlowWall) {
    B.mum += gbb;
    const double nw = a.wallf[f] ? a.nutw[P] : 0.0;
    B.fl.rho += gbb * nw;
    if (a.wallf[f]) B.nutw += gbb * a.rho[P];
  } else {
    B.fl.mu += gbb;
  }
}

template <int D>
struct FlowBarExt {
  FlowCellBar<D> fl;
  double mum, nutw;
};

template <int D>
struct ScalarBarExt {
  ScalarBar<D> k, w;
  double nut, f1, Sr, cd, tau, nu;
  AFX_DEV void zero() {
    k.zero();
    w.zero();
    nut = f1 = Sr = cd = tau = nu = 0.0;
  }
};

// reverse pass R', k / omega part: cotangents of the scalar fluxes, sources and fixation; each
// face's owner writes the mass-flux cotangent mdot_bar[f] the flow part consumes.  With R set, the
// k / omega residual rows come out of the same face evaluations.
template <int D>
struct RansSresAdjSh {
  ScalarTile<D, double, true> tile;
};

template <int D, int ADVS, class Exec, class Sh>
AFX_DEV void k_rans_sres_adj_body(const Exec& ex, Sh& sh, MeshView m, RansArgs<double> a, const double* __restrict__ mdot, const double* __restrict__ lam, RansBars bar, double* __restrict__ mdot_bar, double* __restrict__ R) {
  const SstConst& sc = a.sc;
  auto& tile = sh.tile;
  const int32_t base = ex.base;
  const int t = ex.t;
  const int32_t c = base + t;
  const bool valid = c < m.C;
  ScalarCell<D, double> Kc, Wc;
  double xc[D];
  bool wall = false;
  double lck = 0.0, lcw = 0.0;
  if (valid) {
    load_scalar<D, double>(a, m, c, D + 1, ADVS == kAdvLimited, sc.sk1, sc.sk2, Kc);
    load_scalar<D, double>(a, m, c, D + 2, ADVS == kAdvLimited, sc.sw1, sc.sw2, Wc);
#pragma unroll
    for (int j = 0; j < D; ++j) xc[j] = m.cc[(int64_t)j * m.C + c];
    tile.put(t, Kc, Wc, xc);
    wall = a.wallc[c] != 0;
    lck = lam[(int64_t)(D + 1) * m.C + c];
    lcw = wall ? 0.0 : lam[(int64_t)(D + 2) * m.C + c];  // flux / source part of the omega row
    tile.lam[0][t] = lck;
    tile.lam[1][t] = lcw;
  }
  ex.sync();
  if (!valid) return;
  ScalarBarExt<D> B;
  B.zero();
  const double irho = 1.0 / sc.density;
  const double dP = a.dwall[c];
  const double V = m.vol[c];
  double acck = 0.0, accw = 0.0;
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D> q;
    const double Qf = mdot[r.f] * irho;
    double Qb;
    if (o >= 0) {
      ScalarCell<D, double> Ko, Wo;
      double xo[D], lk_o, lw_o;
      const int ls = ex.slot(o);
      if (ls >= 0) {
        tile.get(ls, Ko, Wo, xo);
        lk_o = tile.lam[0][ls];
        lw_o = tile.lam[1][ls];
      } else {
        load_scalar<D, double>(a, m, o, D + 1, ADVS == kAdvLimited, sc.sk1, sc.sk2, Ko);
        load_scalar<D, double>(a, m, o, D + 2, ADVS == kAdvLimited, sc.sw1, sc.sw2, Wo);
#pragma unroll
        for (int j = 0; j < D; ++j) xo[j] = m.cc[(int64_t)j * m.C + o];
        lk_o = lam[(int64_t)(D + 1) * m.C + o];
        lw_o = a.wallc[o] ? 0.0 : lam[(int64_t)(D + 2) * m.C + o];
      }
      load_face_geo_pts<D>(m, r.f, true, r.own ? xc : xo, r.own ? xo : xc, q);
      const double lkf = r.own ? lck - lk_o : lk_o - lck;
      const double lwf = r.own ? lcw - lw_o : lw_o - lcw;
      if (r.own) {
        Qb = adj_scalar_flux_interior<D, ADVS>(q, Qf, Kc.phi, Ko.phi, Kc.G, Ko.G, Kc.psi, Ko.psi, Kc.gam, Ko.gam, lkf,
                                               0, B.k);
        Qb += adj_scalar_flux_interior<D, ADVS>(q, Qf, Wc.phi, Wo.phi, Wc.G, Wo.G, Wc.psi, Wo.psi, Wc.gam, Wo.gam,
                                                lwf, 0, B.w);
        if (R) {
          acck += scalar_flux_interior<D, ADVS, double>(q, Qf, Kc.phi, Ko.phi, Kc.G, Ko.G, Kc.psi, Ko.psi, Kc.gam,
                                                        Ko.gam);
          accw += scalar_flux_interior<D, ADVS, double>(q, Qf, Wc.phi, Wo.phi, Wc.G, Wo.G, Wc.psi, Wo.psi, Wc.gam,
                                                        Wo.gam);
        }
      } else {
        Qb = adj_scalar_flux_interior<D, ADVS>(q, Qf, Ko.phi, Kc.phi, Ko.G, Kc.G, Ko.psi, Kc.psi, Ko.gam, Kc.gam, lkf,
                                               1, B.k);
        Qb += adj_scalar_flux_interior<D, ADVS>(q, Qf, Wo.phi, Wc.phi, Wo.G, Wc.G, Wo.psi, Wc.psi, Wo.gam, Wc.gam,
                                                lwf, 1, B.w);
        if (R) {
          acck -= scalar_flux_interior<D, ADVS, double>(q, Qf, Ko.phi, Kc.phi, Ko.G, Kc.G, Ko.psi, Kc.psi, Ko.gam,
                                                        Kc.gam);
          accw -= scalar_flux_interior<D, ADVS, double>(q, Qf, Wo.phi, Wc.phi, Wo.G, Wc.G, Wo.psi, Wc.psi, Wo.gam,
                                                        Wc.gam);
        }
      }
    } else {
      load_face_geo_pts<D>(m, r.f, false, xc, xc, q);
      // k face
      const double bk = scalar_bval<D, double>(a.kbc[r.fu], q, Kc.phi, Kc.G, a.kval[r.fu], 0.0, 0.0);
      double wgt = 0.0, gk = Kc.gam;
      if (a.wallf[r.fu]) {
        wgt = wall_function_weight<double>(sc, a.nu[c], dP, Kc.phi);
        gk = (1.0 - wgt) * Kc.gam;
      }
      double bkb, gkb, vb;
      Qb = adj_scalar_flux_boundary<D, ADVS>(q, Qf, Kc.phi, Kc.G, Kc.psi, bk, gk, lck, B.k, bkb, gkb);
      adj_scalar_bval<D>(a.kbc[r.fu], q, bkb, B.k, vb);
      if (a.wallf[r.fu]) {
        B.k.gam += (1.0 - wgt) * gkb;
        const double wb = -Kc.gam * gkb;
        dual wk = wall_function_weight<dual>(sc, dual{a.nu[c], 0.0}, dP, dual{Kc.phi, 1.0});
        dual wn = wall_function_weight<dual>(sc, dual{a.nu[c], 1.0}, dP, dual{Kc.phi, 0.0});
        B.k.phi += wb * wk.d;
        B.nu += wb * wn.d;
      } else {
        B.k.gam += gkb;
      }
      // omega face
      const double bw = scalar_bval<D, double>(a.wbc[r.fu], q, Wc.phi, Wc.G, a.wval[r.fu], 0.0, 0.0);
      double bwb, gwb;
      Qb += adj_scalar_flux_boundary<D, ADVS>(q, Qf, Wc.phi, Wc.G, Wc.psi, bw, Wc.gam, lcw, B.w, bwb, gwb);
      adj_scalar_bval<D>(a.wbc[r.fu], q, bwb, B.w, vb);
      B.w.gam += gwb;
      if (R) {
        acck += scalar_flux_boundary<D, ADVS, double>(q, Qf, Kc.phi, Kc.G, Kc.psi, bk, gk);
        accw += scalar_flux_boundary<D, ADVS, double>(q, Qf, Wc.phi, Wc.G, Wc.psi, bw, Wc.gam);
      }
    }
    if (r.own) mdot_bar[r.f] = Qb * irho;
  }
  // sources (cell-local): R_k -= prod_k + dest_k, R_w -= prod_w + dest_w + cross_w (non-wall)
  {
    const double lk = -lck, lw = -lcw;
    if (wall) {  // the near-wall blends: seeded tangents of the source function
      double inp[8] = {Kc.phi, Wc.phi, a.nut[c], a.Sr[c], a.f1[c], a.cd[c], a.tau[c], a.nu[c]};
      double grad[8];
      for (int s = 0; s < 8; ++s) {
        dual in[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) in[t] = dual{inp[t], t == s ? 1.0 : 0.0};
        dual pk, dk, pw, dw, cw;
        sst_sources<dual>(sc, in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], dP, wall, V, pk, dk, pw, dw, cw);
        grad[s] = lk * (pk.d + dk.d) + lw * (pw.d + dw.d + cw.d);
      }
      B.k.phi += grad[0];
      B.w.phi += grad[1];
      B.nut += grad[2];
      B.Sr += grad[3];
      B.f1 += grad[4];
      B.cd += grad[5];
      B.tau += grad[6];
      B.nu += grad[7];
    } else {
      adj_sst_sources_open(sc, Kc.phi, Wc.phi, a.nut[c], a.Sr[c], a.f1[c], V, lk, lw, B.k.phi, B.w.phi, B.nut, B.Sr,
                           B.f1, B.cd);
    }
  }
  if (wall) {  // R_w = w - omega_wall(nu, d, k), or log(w / omega_wall(nu, d, k))
    const double l = lam[(int64_t)(D + 2) * m.C + c];
    dual ok = omega_wall_k<dual>(sc, dual{a.nu[c], 0.0}, dP, dual{Kc.phi, 1.0});
    dual on = omega_wall_k<dual>(sc, dual{a.nu[c], 1.0}, dP, dual{Kc.phi, 0.0});
    const double lw_ = a.log_w ? l / Wc.phi : l, lt = a.log_w ? l / ok.v : l;
    B.w.phi += lw_;
    B.k.phi -= lt * ok.d;
    B.nu -= lt * on.d;
  }
  if (R) {
    double pk, dk, pw, dw, cw;
    sst_sources<double>(sc, Kc.phi, Wc.phi, a.nut[c], a.Sr[c], a.f1[c], a.cd[c], a.tau[c], a.nu[c], dP, wall, V, pk,
                        dk, pw, dw, cw);
    R[(int64_t)(D + 1) * m.C + c] = (acck - pk) - dk;
    double rw = ((accw - pw) - dw) - cw;
    if (wall) {
      const double t = omega_wall_k<double>(sc, a.nu[c], dP, Kc.phi);
      rw = a.log_w ? log(Wc.phi / t) : Wc.phi - t;
    }
    R[(int64_t)(D + 2) * m.C + c] = rw;
  }
  // diffusivities gamma_k, gamma_w of this cell
  adj_gamma(B.k.gam, a.f1[c], a.nut[c], sc.sk1, sc.sk2, B.nu, B.f1, B.nut);
  adj_gamma(B.w.gam, a.f1[c], a.nut[c], sc.sw1, sc.sw2, B.nu, B.f1, B.nut);
  bar.x[(int64_t)(D + 1) * m.C + c] = B.k.phi;
  bar.x[(int64_t)(D + 2) * m.C + c] = B.w.phi;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    bar.G[(int64_t)((D + 1) * D + j) * m.C + c] = B.k.G[j];
    bar.G[(int64_t)((D + 2) * D + j) * m.C + c] = B.w.G[j];
  }
  if (ADVS == kAdvLimited) {
    bar.psi[(int64_t)D * m.C + c] = B.k.psi;
    bar.psi[(int64_t)(D + 1) * m.C + c] = B.w.psi;
  }
  bar.nu[c] = B.nu;
  bar.nut[c] = B.nut;
  bar.f1[c] = B.f1;
  bar.Sr[c] = B.Sr;
  bar.cd[c] = B.cd;
  bar.tau[c] = B.tau;
}

template <int D, int ADVS>
__global__ void __launch_bounds__(256) k_rans_sres_adj(MeshView m, RansArgs<double> a,
                                                       const double* __restrict__ mdot,
                                                       const double* __restrict__ lam, RansBars bar,
                                                       double* __restrict__ mdot_bar, double* __restrict__ R) {
  __shared__ RansSresAdjSh<D> sh;
  k_rans_sres_adj_body<D, ADVS>(DevExec(m), sh, m, a, mdot, lam, bar, mdot_bar, R);
}

// reverse pass R', flow part: momentum / continuity flux cotangents plus the scalar part's
// mass-flux cotangent.  Two resident waves per SIMD (a 256-register cap): the unconstrained 3D
// build needs ~300 registers and runs one wave per SIMD; the cap spills 24 values and the gather
// latency hidden by the second wave more than pays for them (270 -> 239 us on the 128x64x64 SST
// channel).
template <int D>
struct RansFresAdjSh {
  FlowTile<D, double, true> tile;
  double tlam[D + 1][kBlock];
};

template <int D, int ADVF, class Exec, class Sh>
AFX_DEV void k_rans_fres_adj_body(const Exec& ex, Sh& sh, MeshView m, RansArgs<double> a, const double* __restrict__ lam, const double* __restrict__ mdot_bar, RansBars bar) {
  constexpr int NV = D + 1;
  auto& tile = sh.tile;
  auto& tlam = sh.tlam;
  const int32_t base = ex.base;
  const int t = ex.t;
  const int32_t c = base + t;
  const bool valid = c < m.C;
  FlowCell<D, double> X;
  double xc[D];
  double lc[NV];
  if (valid) {
    load_flow<D, double>(a, m, c, true, ADVF == kAdvLimited, X);
#pragma unroll
    for (int j = 0; j < D; ++j) xc[j] = m.cc[(int64_t)j * m.C + c];
    tile.put(t, X, xc);
#pragma unroll
    for (int v = 0; v < NV; ++v) tlam[v][t] = lc[v] = lam[(int64_t)v * m.C + c];
  }
  ex.sync();
  if (!valid) return;
  FlowBarExt<D> B;
  B.fl.zero();
  B.mum = B.nutw = 0.0;
  const double lc_cont = c == a.pin ? 0.0 : lc[D];
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D> q;
    double lf[NV];
    if (o >= 0) {
      FlowCell<D, double> O;
      double xo[D];
      const int ls = ex.slot(o);
      if (ls >= 0) {
        tile.get(ls, O, xo);
      } else {
        load_flow<D, double>(a, m, o, true, ADVF == kAdvLimited, O);
#pragma unroll
        for (int j = 0; j < D; ++j) xo[j] = m.cc[(int64_t)j * m.C + o];
      }
      load_face_geo_pts<D>(m, r.f, true, r.own ? xc : xo, r.own ? xo : xc, q);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        double lo = ls >= 0 ? tlam[v][ls] : lam[(int64_t)v * m.C + o];
        if (v == D && o == a.pin) lo = 0.0;
        const double lcv = v == D ? lc_cont : lc[v];
        lf[v] = r.own ? lcv - lo : lo - lcv;
      }
      lf[D] += mdot_bar[r.f];
      if (r.own)
        adj_flux_interior<D, ADVF>(q, X, O, lf, 0, B.fl);
      else
        adj_flux_interior<D, ADVF>(q, O, X, lf, 1, B.fl);
    } else {
      load_face_geo_pts<D>(m, r.f, false, xc, xc, q);
      const int kind = a.kind[r.fu];
      double ub[D];
#pragma unroll
      for (int i = 0; i < D; ++i) ub[i] = a.ub[(int64_t)i * m.F + r.fu];
      const double gB = rans_gamma_b<double>(a, kind, r.fu, c, X.mu);
#pragma unroll
      for (int v = 0; v < NV; ++v) lf[v] = v == D ? lc_cont : lc[v];
      lf[D] += mdot_bar[r.f];
      double ubb[D], pbb, gbb;
      adj_flux_boundary<D, ADVF>(q, kind, X, ub, a.pb[r.fu], gB, lf, B.fl, ubb, pbb, gbb);
      adj_gamma_b<D>(a, kind, r.fu, c, gbb, B);
#pragma unroll
      for (int i = 0; i < D; ++i) bar.ub[(int64_t)i * m.F + r.fu] = ubb[i];
      bar.pb[r.fu] = pbb;
    }
  }
  if (c == a.pin) B.fl.p += lc[D];
#pragma unroll
  for (int i = 0; i < D; ++i) bar.x[(int64_t)i * m.C + c] = B.fl.u[i];
  bar.x[(int64_t)D * m.C + c] = B.fl.p;
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) bar.G[(int64_t)(i * D + j) * m.C + c] = B.fl.Gu[i][j];
#pragma unroll
  for (int j = 0; j < D; ++j) bar.G[(int64_t)(D * D + j) * m.C + c] = B.fl.Gp[j];
  bar.aP[c] = B.fl.aP;
  if (ADVF == kAdvLimited) {
#pragma unroll
    for (int i = 0; i < D; ++i) bar.psi[(int64_t)i * m.C + c] = B.fl.psi[i];
  }
  bar.mueff[c] = B.fl.mu;
  bar.rho[c] = B.fl.rho;
  bar.mu[c] = B.mum;
  bar.nutw[c] = B.nutw;
}

template <int D, int ADVF>
__global__ void __launch_bounds__(256, 2) k_rans_fres_adj(MeshView m, RansArgs<double> a, const double* __restrict__ lam,
                                                       const double* __restrict__ mdot_bar, RansBars bar) {
  __shared__ RansFresAdjSh<D> sh;
  k_rans_fres_adj_body<D, ADVF>(DevExec(m), sh, m, a, lam, mdot_bar, bar);
}

// reverse pass D'.  Three waves per SIMD (168 registers; the 3D build needs 174 unconstrained,
// the cap spills two): 152 -> 145 us on the 128x64x64 SST channel.
template <int D>
struct RansApAdjSh {
  FlowTile<D, double, false> tile;
  double tapb[kBlock];
};

template <int D, int ADV, class Exec, class Sh>
AFX_DEV void k_rans_ap_adj_body(const Exec& ex, Sh& sh, MeshView m, RansArgs<double> a, RansBars bar) {
  auto& tile = sh.tile;
  auto& tapb = sh.tapb;
  const int32_t base = ex.base;
  const int t = ex.t;
  const int32_t c = base + t;
  const bool valid = c < m.C;
  FlowCell<D, double> X;
  double xc[D];
  double apb_c = 0.0;
  if (valid) {
    load_flow<D, double>(a, m, c, false, false, X);
#pragma unroll
    for (int j = 0; j < D; ++j) xc[j] = m.cc[(int64_t)j * m.C + c];
    tile.put(t, X, xc);
    tapb[t] = apb_c = bar.aP[c];
  }
  ex.sync();
  if (!valid) return;
  RansCellBar<D> B;
  B.zero();
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D> q;
    if (o >= 0) {
      FlowCell<D, double> O;
      double xo[D];
      double apb_o;
      const int ls = ex.slot(o);
      if (ls >= 0) {
        tile.get(ls, O, xo);
        apb_o = tapb[ls];
      } else {
        load_flow<D, double>(a, m, o, false, false, O);
#pragma unroll
        for (int j = 0; j < D; ++j) xo[j] = m.cc[(int64_t)j * m.C + o];
        apb_o = bar.aP[o];
      }
      load_face_geo_pts<D>(m, r.f, true, r.own ? xc : xo, r.own ? xo : xc, q);
      if (r.own)
        adj_ap_interior<D, ADV>(q, X, O, apb_c, apb_o, 0, B.fl);
      else
        adj_ap_interior<D, ADV>(q, O, X, apb_o, apb_c, 1, B.fl);
    } else {
      load_face_geo_pts<D>(m, r.f, false, xc, xc, q);
      const int kind = a.kind[r.fu];
      const double gB = rans_gamma_b<double>(a, kind, r.fu, c, X.mu);
      double gbb;
      adj_ap_boundary<D, ADV>(q, kind, X, gB, apb_c, B.fl, gbb);
      adj_gamma_b<D>(a, kind, r.fu, c, gbb, B);
    }
  }
#pragma unroll
  for (int i = 0; i < D; ++i) bar.x[(int64_t)i * m.C + c] += B.fl.u[i];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) bar.G[(int64_t)(i * D + j) * m.C + c] += B.fl.Gu[i][j];
  bar.mueff[c] += B.fl.mu;
  bar.rho[c] += B.fl.rho;
  bar.mu[c] += B.mum;
  bar.nutw[c] += B.nutw;
}

template <int D, int ADV>
__global__ void __launch_bounds__(256, 3) k_rans_ap_adj(MeshView m, RansArgs<double> a, RansBars bar) {
  __shared__ RansApAdjSh<D> sh;
  k_rans_ap_adj_body<D, ADV>(DevExec(m), sh, m, a, bar);
}

// reverse pass C': closure (seeded duals), wall shear rate, limiters; then the gradient cotangent
// is pre-multiplied by the cell's reconstruction operator.
template <int D, int ADVF, int ADVS, int GRAD>
AFX_DEV void k_rans_closure_adj_body(int32_t c, MeshView m, RansArgs<double> a, const int32_t* __restrict__ ext, const int32_t* __restrict__ kmin_face, const double* __restrict__ lsqI, RansBars bar) {
  constexpr int NA = D + 3, NL = D + 2;
  const SstConst& sc = a.sc;
  if (c >= m.C) return;
  const bool wall = a.wallc[c] != 0;
  const double d = a.dwall[c];
  double dg[D];
#pragma unroll
  for (int j = 0; j < D; ++j) dg[j] = wall ? a.dgrad[(int64_t)j * m.C + c] : 0.0;
  // total closure-output cotangents (mu_eff = mu + rho nu_t)
  const double me_b = bar.mueff[c];
  const double nut_b = bar.nut[c] + a.rho[c] * me_b;
  double rho_b = bar.rho[c] + a.nut[c] * me_b;
  double mu_b = bar.mu[c] + me_b;
  const double f1_b = bar.f1[c], Sr_b = bar.Sr[c], cd_b = bar.cd[c], nutw_b = bar.nutw[c];
  double nu_b = bar.nu[c];
  double k0, w0, Gu0[D][D], Gk0[D], Gw0[D];
  load_closure_inputs<D, double>(a, m, c, k0, w0, Gu0, Gk0, Gw0);
  const double nu0 = a.nu[c];
  // The closure depends on the gradients only through the strain magnitude and the products
  // (A, B) of sst_closure_core: seed those six scalars, then chain to the gradients by hand.
  const double strain0 = strain_magnitude<D, double>(Gu0);
  double A0, B0;
  closure_products<D, double>(Gk0, Gw0, dg, wall, A0, B0);
  // one evaluation carrying the six tangents (seed s on input s of k, omega, nu, strain0, A, B)
  using D6 = dualn<6>;
  const SstClosure<D6> o = sst_closure_core<D6>(sc, seeded<6>(k0, 0), seeded<6>(w0, 1), seeded<6>(strain0, 3),
                                                seeded<6>(A0, 4), seeded<6>(B0, 5), seeded<6>(nu0, 2), d, wall);
  double gin[6];
  const int nseed = wall ? 6 : 5;
#pragma unroll
  for (int s = 0; s < 6; ++s)
    gin[s] = s < nseed ? nut_b * o.nut.d[s] + f1_b * o.f1.d[s] + Sr_b * o.strain.d[s] + cd_b * o.cd.d[s] +
                             nutw_b * o.nutw.d[s]
                       : 0.0;
  double xb_k = gin[0], xb_w = gin[1];
  nu_b += gin[2];
  double Gb[NA][D];
#pragma unroll
  for (int v = 0; v < NA; ++v)
#pragma unroll
    for (int j = 0; j < D; ++j) Gb[v][j] = bar.G[(int64_t)(v * D + j) * m.C + c];
  // strain0 = sqrt(2 S:S): d strain0 / d Gu_ij = 2 S_ij / strain0 (zero at strain0 = 0)
  if (strain0 > 0.0) {
    const double sb = gin[3] / strain0;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) Gb[i][j] += sb * (Gu0[i][j] + Gu0[j][i]);
  }
  if (wall) {  // A = grad k . grad d, B = |grad k|^2
#pragma unroll
    for (int j = 0; j < D; ++j) Gb[D + 1][j] += gin[4] * dg[j] + 2.0 * gin[5] * Gk0[j];
  } else {  // A = grad k . grad omega
#pragma unroll
    for (int j = 0; j < D; ++j) {
      Gb[D + 1][j] += gin[4] * Gw0[j];
      Gb[D + 2][j] += gin[4] * Gk0[j];
    }
  }
  double xb[NA];
#pragma unroll
  for (int v = 0; v < NA; ++v) xb[v] = 0.0;
  xb[D + 1] = xb_k;
  xb[D + 2] = xb_w;
  // wall shear rate: tau = sum_f (|u - u_b| / d) A_f / sum_f A_f
  const double tau_b = bar.tau[c];
  if (wall && tau_b != 0.0) {
    double xcl[D + 1];
#pragma unroll
    for (int v = 0; v <= D; ++v) xcl[v] = a.x[(int64_t)v * m.C + c];
    double den = 0.0;
    for (int32_t k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
      const int32_t e = m.cface[k];
      if (e >= 0 && a.wallf[user_face(m, e)]) den += m.area[e];
    }
    for (int32_t k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
      const int32_t e = m.cface[k];
      if (e < 0) continue;
      const int32_t eu = user_face(m, e);
      if (!a.wallf[eu]) continue;
      double sl[D], s2 = 0.0;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        sl[j] = xcl[j] - rans_bval_lead<D, double>(a, j, m.F, eu, xcl);
        s2 += sl[j] * sl[j];
      }
      if (s2 <= 0.0) continue;
      const double rate_b = tau_b * m.area[e] / den;
      const double mag = sqrt(s2);
      const int kind = a.kind[eu];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const double sb = rate_b / d * sl[j] / mag;
        xb[j] += sb;  // d slip / d u_P
        if (kind == kFlowWall || kind == kFlowInlet)
          bar.ub[(int64_t)j * m.F + eu] -= sb;
        else if (kind == kFlowOutlet)
          xb[j] -= sb;
      }
    }
  }
  // limiters (u_i if the momentum advection is limited, k and omega if theirs is).  The extremum
  // cotangents are stored after all of them, so no store sits between one variable's chain of
  // gathers and the next one's.
  constexpr bool LFV = ADVF == kAdvLimited, LSV = ADVS == kAdvLimited;
  if constexpr (LFV || LSV) {
    const double e2 = m.vol[c] * (a.lim_k * a.lim_k * a.lim_k);
    double xcen[D];
#pragma unroll
    for (int j = 0; j < D; ++j) xcen[j] = m.cc[(int64_t)j * m.C + c];
    double mxbs[NL], mnbs[NL];
#pragma unroll
    for (int v = 0; v < NA; ++v) {
      const int l = lim_slot<D>(v, LFV, LSV);
      if (l < 0) continue;
      const double pbar = bar.psi[(int64_t)l * m.C + c];
      const int32_t amx = ext[(int64_t)l * m.C + c], amn = ext[(int64_t)(NL + l) * m.C + c];
      const double xv = a.x[(int64_t)v * m.C + c];
      const double mxv = a.x[(int64_t)v * m.C + amx], mnv = a.x[(int64_t)v * m.C + amn];
      double Gv[D];
#pragma unroll
      for (int j = 0; j < D; ++j) Gv[j] = a.G[(int64_t)(v * D + j) * m.C + c];
      // the first face attaining the forward minimum, recorded by the forward limiter pass
      const int32_t f = kmin_face[(int64_t)l * m.C + c];
      double dxf[D];
#pragma unroll
      for (int j = 0; j < D; ++j) dxf[j] = m.fc[(int64_t)j * m.F + f] - xcen[j];
      double dmf = Gv[0] * dxf[0];
#pragma unroll
      for (int j = 1; j < D; ++j) dmf = dmf + Gv[j] * dxf[j];
      const double lo = venkat_face<double>(dmf, mxv - xv, mnv - xv, e2);
      const double dclip = (lo > 0.0 && lo < 1.0) ? 1.0 : ((lo == 0.0 || lo == 1.0) ? 0.5 : 0.0);
      const double vbar = pbar * dclip;
      double mxb = 0.0, mnb = 0.0;
      if (vbar != 0.0) {
        dual r_dm = venkat_face<dual>(dual{dmf, 1.0}, dual{mxv - xv, 0.0}, dual{mnv - xv, 0.0}, e2);
        dual r_hx = venkat_face<dual>(dual{dmf, 0.0}, dual{mxv - xv, 1.0}, dual{mnv - xv, 0.0}, e2);
        dual r_hn = venkat_face<dual>(dual{dmf, 0.0}, dual{mxv - xv, 0.0}, dual{mnv - xv, 1.0}, e2);
#pragma unroll
        for (int j = 0; j < D; ++j) Gb[v][j] += vbar * r_dm.d * dxf[j];
        mxb = vbar * r_hx.d;
        mnb = vbar * r_hn.d;
        xb[v] -= mxb + mnb;
        if (amx == c) {
          xb[v] += mxb;
          mxb = 0.0;
        }
        if (amn == c) {
          xb[v] += mnb;
          mnb = 0.0;
        }
      }
      mxbs[l] = mxb;
      mnbs[l] = mnb;
    }
#pragma unroll
    for (int v = 0; v < NA; ++v) {
      const int l = lim_slot<D>(v, LFV, LSV);
      if (l < 0) continue;
      bar.ext[(int64_t)l * m.C + c] = mxbs[l];
      bar.ext[(int64_t)(NL + l) * m.C + c] = mnbs[l];
    }
  }
#pragma unroll
  for (int v = 0; v < NA; ++v) bar.x[(int64_t)v * m.C + c] += xb[v];
  bar.nu[c] = nu_b;
  bar.mu[c] = mu_b;
  bar.rho[c] = rho_b;
  // pre-multiply the gradient cotangent by the reconstruction operator (the corrected scheme's
  // swept-solve reverse pass does it separately)
  if (GRAD == kGradCorrected) {
#pragma unroll
    for (int v = 0; v < NA; ++v)
#pragma unroll
      for (int j = 0; j < D; ++j) bar.G[(int64_t)(v * D + j) * m.C + c] = Gb[v][j];
  } else if (GRAD == 0) {
    const double iv = 1.0 / m.vol[c];
#pragma unroll
    for (int v = 0; v < NA; ++v)
#pragma unroll
      for (int j = 0; j < D; ++j) bar.G[(int64_t)(v * D + j) * m.C + c] = Gb[v][j] * iv;
  } else {
    double I[D][D];
    if (lsqI) {  // the forward gradient pass's inverse
#pragma unroll
      for (int p = 0; p < D; ++p)
#pragma unroll
        for (int q2 = 0; q2 < D; ++q2) I[p][q2] = lsqI[(int64_t)(p * D + q2) * m.C + c];
    } else {
      double M[D][D];
      lsq_matrix<D>(m, c, M);
      sym_inverse<D>(M, I);
    }
#pragma unroll
    for (int v = 0; v < NA; ++v)
#pragma unroll
      for (int p = 0; p < D; ++p) {
        double s = 0.0;
#pragma unroll
        for (int q2 = 0; q2 < D; ++q2) s += I[p][q2] * Gb[v][q2];
        bar.G[(int64_t)(v * D + p) * m.C + c] = s;
      }
  }
}

template <int D, int ADVF, int ADVS, int GRAD>
__global__ void __launch_bounds__(256) k_rans_closure_adj(MeshView m, RansArgs<double> a,
                                                          const int32_t* __restrict__ ext,
                                                          const int32_t* __restrict__ kmin_face,
                                                          const double* __restrict__ lsqI, RansBars bar) {
  k_rans_closure_adj_body<D, ADVF, ADVS, GRAD>(xcd_thread_index(m), m, a, ext, kmin_face, lsqI, bar);
}

// reverse pass G'
template <int D, int NE>
struct RansGradAdjSh {
  double tB[(D + 3) * D][kBlock];
  double tcc[D][kBlock];
  int32_t tex[2 * (NE > 0 ? NE : 1)][kBlock];
  double tbe[2 * (NE > 0 ? NE : 1)][kBlock];
};

template <int D, int GRAD, bool LIMF, bool LIMS, class Exec, class Sh>
AFX_DEV void k_rans_grad_adj_body(const Exec& ex, Sh& sh, MeshView m, RansArgs<double> a, const int32_t* __restrict__ ext, RansBars bar) {
  constexpr int NA = D + 3, NL = D + 2;
  constexpr int NE = (LIMF ? D : 0) + (LIMS ? 2 : 0);  // limited variables
  auto& tB = sh.tB;
  auto& tcc = sh.tcc;
  auto& tex = sh.tex;
  auto& tbe = sh.tbe;  // the limiter extremum cotangents
  const int32_t base = ex.base;
  const int t = ex.t;
  const int32_t c = base + t;
  const bool valid = c < m.C;
  double Bc[NA][D];
  double ccc[D];
  if (valid) {
#pragma unroll
    for (int v = 0; v < NA; ++v)
#pragma unroll
      for (int j = 0; j < D; ++j) tB[v * D + j][t] = Bc[v][j] = bar.G[(int64_t)(v * D + j) * m.C + c];
#pragma unroll
    for (int j = 0; j < D; ++j) tcc[j][t] = ccc[j] = m.cc[(int64_t)j * m.C + c];
    int e = 0;
#pragma unroll
    for (int v = 0; v < NA; ++v) {
      const int l = lim_slot<D>(v, LIMF, LIMS);
      if (l < 0) continue;
      tex[2 * e][t] = ext[(int64_t)l * m.C + c];
      tex[2 * e + 1][t] = ext[(int64_t)(NL + l) * m.C + c];
      tbe[2 * e][t] = bar.ext[(int64_t)l * m.C + c];
      tbe[2 * e + 1][t] = bar.ext[(int64_t)(NL + l) * m.C + c];
      ++e;
    }
  }
  ex.sync();
  if (!valid) return;
  double xb[NA];
#pragma unroll
  for (int v = 0; v < NA; ++v) xb[v] = 0.0;
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    if (o >= 0) {
      const int ls = ex.slot(o);
      // the neighbour's gradient cotangent, centroid and limiter extrema (which cells attain them and
      // their cotangents), all fetched together: no load waits on a comparison
      double Bo[NA][D], cco[D];
      int32_t oam[2 * (NE > 0 ? NE : 1)];
      double obe[2 * (NE > 0 ? NE : 1)];
      if (ls >= 0) {
#pragma unroll
        for (int v = 0; v < NA; ++v)
#pragma unroll
          for (int j = 0; j < D; ++j) Bo[v][j] = tB[v * D + j][ls];
#pragma unroll
        for (int j = 0; j < D; ++j) cco[j] = tcc[j][ls];
#pragma unroll
        for (int e = 0; e < 2 * NE; ++e) {
          oam[e] = tex[e][ls];
          obe[e] = tbe[e][ls];
        }
      } else {
#pragma unroll
        for (int v = 0; v < NA; ++v)
#pragma unroll
          for (int j = 0; j < D; ++j) Bo[v][j] = bar.G[(int64_t)(v * D + j) * m.C + o];
#pragma unroll
        for (int j = 0; j < D; ++j) cco[j] = m.cc[(int64_t)j * m.C + o];
        int e = 0;
#pragma unroll
        for (int v = 0; v < NA; ++v) {
          const int l = lim_slot<D>(v, LIMF, LIMS);
          if (l < 0) continue;
          oam[2 * e] = ext[(int64_t)l * m.C + o];
          oam[2 * e + 1] = ext[(int64_t)(NL + l) * m.C + o];
          obe[2 * e] = bar.ext[(int64_t)l * m.C + o];
          obe[2 * e + 1] = bar.ext[(int64_t)(NL + l) * m.C + o];
          ++e;
        }
      }
      if (GRAD == 0) {
        FaceGeo<D> q;
        load_face_geo_pts<D>(m, r.f, true, r.own ? ccc : cco, r.own ? cco : ccc, q);
        const double ws = r.own ? 1.0 - q.g : q.g;
#pragma unroll
        for (int v = 0; v < NA; ++v) {
          double pb_ = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) pb_ += (q.n[j] * q.A) * (r.own ? Bc[v][j] - Bo[v][j] : Bo[v][j] - Bc[v][j]);
          xb[v] += ws * pb_;
        }
      } else {
        double dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = cco[j] - ccc[j];
        const double wgt = 1.0 / vdot<D>(dx, dx);
#pragma unroll
        for (int v = 0; v < NA; ++v) {
          double own_dot = 0.0, oth_dot = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) {
            own_dot += dx[j] * Bc[v][j];
            oth_dot += dx[j] * Bo[v][j];
          }
          xb[v] += -wgt * own_dot - wgt * oth_dot;
        }
      }
      int e = 0;
#pragma unroll
      for (int v = 0; v < NA; ++v) {
        const int l = lim_slot<D>(v, LIMF, LIMS);
        if (l < 0) continue;
        if (oam[2 * e] == c) xb[v] += obe[2 * e];
        if (oam[2 * e + 1] == c) xb[v] += obe[2 * e + 1];
        ++e;
      }
    } else {
      const int kind = a.kind[r.fu];
      double phib[NA];
      if (GRAD == 0) {
        const double A = m.area[r.f];
#pragma unroll
        for (int v = 0; v < NA; ++v) {
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
        for (int v = 0; v < NA; ++v) {
          double s = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) s += dx[j] * Bc[v][j];
          phib[v] = wgt * s;
          xb[v] -= wgt * s;
        }
      }
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
      if (a.kbc[r.fu] == kBcZeroGradient) xb[D + 1] += phib[D + 1];
      if (a.wbc[r.fu] == kBcZeroGradient) xb[D + 2] += phib[D + 2];
    }
  }
#pragma unroll
  for (int v = 0; v < NA; ++v) bar.x[(int64_t)v * m.C + c] += xb[v];
}

template <int D, int GRAD, bool LIMF, bool LIMS>
__global__ void __launch_bounds__(256) k_rans_grad_adj(MeshView m, RansArgs<double> a, const int32_t* __restrict__ ext,
                                                       RansBars bar) {
  __shared__ RansGradAdjSh<D, (LIMF ? D : 0) + (LIMS ? 2 : 0)> sh;
  k_rans_grad_adj_body<D, GRAD, LIMF, LIMS>(DevExec(m), sh, m, a, ext, bar);
}

// ---------------------------------------------------------------------------------------------
// geometry cotangents (node-coordinate VJP).  Three passes after the state adjoint, reading its
// intermediates: the closure-field, a_P, limiter and face mass-flux cotangents, and the gradient
// cotangents pre-multiplied by the reconstruction operator (y).  Each location has one writer per
// pass: the cell pass writes (=) the cell arrays, the face and gradient passes add to them, and a
// face's arrays are written by its owner.  Besides the five geometry arrays the passes produce the
// cotangents of the wall distance and of its gradient, which the caller carries through the
// distance and gradient set-up to the geometry.
struct RansGeoBars {
  double *area, *nrm, *fc;  // [F], [D][F], [D][F] caller face order
  double *vol, *cc;         // [C], [D][C]
  double *dwall, *dgrad;    // [C], [D][C] wall distance and its gradient
  double* limx;             // [NL][D][C] scratch: limiter face-centroid cotangent at its argmin face
  int32_t* limf;            // [NL][C]    scratch: that face (view index, -1 none)
  double* tauA;             // [C]        scratch: tau cotangent / the cell's wall area
  CggGeoBars cgg;           // corrected Green-Gauss: the swept solve's geometry cotangents
};

// Cell-local terms: source and body-force volumes, Green-Gauss 1/V, the closure's, sources' and
// wall fixation's wall-distance derivatives, the wall shear rate's, and the limiters.
template <int D, bool LFV, bool LSV, int GRAD>
__global__ void __launch_bounds__(256) k_rans_geo_cell(MeshView m, RansArgs<double> a, const double* __restrict__ lam,
                                                       const int32_t* __restrict__ ext,
                                                       const int32_t* __restrict__ kmin_face, RansBars bar,
                                                       RansGeoBars gb) {
  constexpr int NA = D + 3, NL = D + 2;
  const SstConst& sc = a.sc;
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  const bool wall = a.wallc[c] != 0;
  const double V = m.vol[c], d = a.dwall[c];
  double Vb = 0.0, db = 0.0, dgb[D], ccb[D];
#pragma unroll
  for (int j = 0; j < D; ++j) dgb[j] = ccb[j] = 0.0;
  const double lk = lam[(int64_t)(D + 1) * m.C + c];
  const double lw_row = lam[(int64_t)(D + 2) * m.C + c];
  const double lw = wall ? 0.0 : lw_row;  // the omega row's transport / source part
#pragma unroll
  for (int i = 0; i < D; ++i) Vb -= a.bf[i] * lam[(int64_t)i * m.C + c];
  if (GRAD == 0) {  // G = acc / V
#pragma unroll
    for (int v = 0; v < NA; ++v)
#pragma unroll
      for (int j = 0; j < D; ++j) Vb -= bar.G[(int64_t)(v * D + j) * m.C + c] * a.G[(int64_t)(v * D + j) * m.C + c];
  }
  const double kc = a.x[(int64_t)(D + 1) * m.C + c], wc = a.x[(int64_t)(D + 2) * m.C + c];
  const double nu = a.nu[c];
  {  // sources: linear in V; on wall cells their near-wall blends also read d
    double pk, dk, pw, dw, cw;
    sst_sources<double>(sc, kc, wc, a.nut[c], a.Sr[c], a.f1[c], a.cd[c], a.tau[c], nu, d, wall, 1.0, pk, dk, pw, dw,
                        cw);
    Vb -= lk * (pk + dk) + lw * ((pw + dw) + cw);
    if (wall) {
      dual sk, sdk, sw, sdw, scw;
      sst_sources<dual>(sc, dual{kc, 0.0}, dual{wc, 0.0}, dual{a.nut[c], 0.0}, dual{a.Sr[c], 0.0},
                        dual{a.f1[c], 0.0}, dual{a.cd[c], 0.0}, dual{a.tau[c], 0.0}, dual{nu, 0.0}, dual{d, 1.0},
                        wall, V, sk, sdk, sw, sdw, scw);
      db -= lk * (sk.d + sdk.d) + lw * ((sw.d + sdw.d) + scw.d);
      // omega wall fixation: w - omega_wall(nu, d, k), or log(w / omega_wall)
      const dual t2 = omega_wall_k<dual>(sc, dual{nu, 0.0}, dual{d, 1.0}, dual{kc, 0.0});
      db -= a.log_w ? lw_row * t2.d / t2.v : lw_row * t2.d;
    }
  }
  {  // closure (nu_t, F1, strain, CD, wall nu_t) and the wall shear rate
    dual Gu[D][D], Gk[D], Gw[D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) Gu[i][j] = dual{a.G[(int64_t)(i * D + j) * m.C + c], 0.0};
#pragma unroll
    for (int j = 0; j < D; ++j) {
      Gk[j] = dual{a.G[(int64_t)((D + 1) * D + j) * m.C + c], 0.0};
      Gw[j] = dual{a.G[(int64_t)((D + 2) * D + j) * m.C + c], 0.0};
    }
    const double nut_b = bar.nut[c] + a.rho[c] * bar.mueff[c];
    const double f1_b = bar.f1[c], Sr_b = bar.Sr[c], cd_b = bar.cd[c], nutw_b = bar.nutw[c];
    // seed 0: d; seeds 1..D: the wall-distance gradient (read on wall cells only)
    for (int s = 0; s <= (wall ? D : 0); ++s) {
      dual dd = dual{d, s == 0 ? 1.0 : 0.0}, dg[D];
#pragma unroll
      for (int j = 0; j < D; ++j) dg[j] = dual{wall ? a.dgrad[(int64_t)j * m.C + c] : 0.0, s == j + 1 ? 1.0 : 0.0};
      const SstClosure<dual> o =
          sst_closure<D, dual, dual>(sc, dual{kc, 0.0}, dual{wc, 0.0}, Gu, Gk, Gw, dual{nu, 0.0}, dd, dg, wall);
      const double g = nut_b * o.nut.d + f1_b * o.f1.d + Sr_b * o.strain.d + cd_b * o.cd.d + nutw_b * o.nutw.d;
      if (s == 0)
        db += g;
      else
        dgb[s - 1] += g;
    }
    if (wall) {  // tau = sum_f (|slip| / d) A_f / sum_f A_f
      const double tau_b = bar.tau[c];
      double den = 0.0;
      for (int32_t k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
        const int32_t e = m.cface[k];
        if (e >= 0 && a.wallf[user_face(m, e)]) den += m.area[e];
      }
      db -= tau_b * a.tau[c] / d;
      gb.tauA[c] = den > 0.0 ? tau_b / den : 0.0;
    } else {
      gb.tauA[c] = 0.0;
    }
  }
  if constexpr (LFV || LSV) {  // limiters at their recorded argmin faces
    const double k3 = a.lim_k * a.lim_k * a.lim_k, e2 = V * k3;
    double xcen[D];
#pragma unroll
    for (int j = 0; j < D; ++j) xcen[j] = m.cc[(int64_t)j * m.C + c];
#pragma unroll
    for (int v = 0; v < NA; ++v) {
      const int l = lim_slot<D>(v, LFV, LSV);
      if (l < 0) continue;
      const double pbar = bar.psi[(int64_t)l * m.C + c];
      const int32_t amx = ext[(int64_t)l * m.C + c], amn = ext[(int64_t)(NL + l) * m.C + c];
      const double xv = a.x[(int64_t)v * m.C + c];
      const double mxv = a.x[(int64_t)v * m.C + amx], mnv = a.x[(int64_t)v * m.C + amn];
      double Gv[D];
#pragma unroll
      for (int j = 0; j < D; ++j) Gv[j] = a.G[(int64_t)(v * D + j) * m.C + c];
      const int32_t f = kmin_face[(int64_t)l * m.C + c];
      double dxf[D];
#pragma unroll
      for (int j = 0; j < D; ++j) dxf[j] = m.fc[(int64_t)j * m.F + f] - xcen[j];
      double dmf = Gv[0] * dxf[0];
#pragma unroll
      for (int j = 1; j < D; ++j) dmf = dmf + Gv[j] * dxf[j];
      const double lo = venkat_face<double>(dmf, mxv - xv, mnv - xv, e2);
      const double dclip = (lo > 0.0 && lo < 1.0) ? 1.0 : ((lo == 0.0 || lo == 1.0) ? 0.5 : 0.0);
      const double vbar = pbar * dclip;
      double xb[D];
#pragma unroll
      for (int j = 0; j < D; ++j) xb[j] = 0.0;
      if (vbar != 0.0) {
        const dual r_dm = venkat_face<dual>(dual{dmf, 1.0}, dual{mxv - xv, 0.0}, dual{mnv - xv, 0.0}, e2);
        const dual r_e2 =
            venkat_face<dual, dual>(dual{dmf, 0.0}, dual{mxv - xv, 0.0}, dual{mnv - xv, 0.0}, dual{e2, 1.0});
#pragma unroll
        for (int j = 0; j < D; ++j) {
          xb[j] = vbar * r_dm.d * Gv[j];
          ccb[j] -= xb[j];
        }
        Vb += vbar * r_e2.d * k3;
      }
      gb.limf[(int64_t)l * m.C + c] = vbar != 0.0 ? f : -1;
#pragma unroll
      for (int j = 0; j < D; ++j) gb.limx[(int64_t)(l * D + j) * m.C + c] = xb[j];
    }
  }
  gb.vol[c] = Vb;
  gb.dwall[c] = db;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    gb.cc[(int64_t)j * m.C + c] = ccb[j];
    gb.dgrad[(int64_t)j * m.C + c] = dgb[j];
  }
}

// Face fluxes: momentum / continuity with their face mass flux (whose cotangent also carries the
// k / omega transport's), the momentum diagonal, the k / omega fluxes, and the wall shear rate's
// face areas.  Writes the owned faces' arrays (=) and adds the cell's centroid / volume / distance
// parts.
template <int D, int ADVF, int ADVS>
__global__ void __launch_bounds__(256) k_rans_geo_face(MeshView m, RansArgs<double> a, const double* __restrict__ lam,
                                                       const double* __restrict__ mdot,
                                                       const double* __restrict__ mdot_bar, RansBars bar,
                                                       RansGeoBars gb) {
  constexpr int NF = D + 1;
  const SstConst& sc = a.sc;
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  const bool wall = a.wallc[c] != 0;
  FlowCell<D, double> X;
  load_flow<D, double>(a, m, c, true, ADVF == kAdvLimited, X);
  ScalarCell<D, double> Kc, Wc;
  load_scalar<D, double>(a, m, c, D + 1, ADVS == kAdvLimited, sc.sk1, sc.sk2, Kc);
  load_scalar<D, double>(a, m, c, D + 2, ADVS == kAdvLimited, sc.sw1, sc.sw2, Wc);
  double lc[NF];
#pragma unroll
  for (int v = 0; v < NF; ++v) lc[v] = lam[(int64_t)v * m.C + c];
  if (c == a.pin) lc[D] = 0.0;
  const double lck = lam[(int64_t)(D + 1) * m.C + c];
  const double lcw = wall ? 0.0 : lam[(int64_t)(D + 2) * m.C + c];
  const double apb_c = bar.aP[c], irho = 1.0 / sc.density, dP = a.dwall[c], tauA = gb.tauA[c];
  double Vb = 0.0, db = 0.0, ccb[D];
#pragma unroll
  for (int j = 0; j < D; ++j) ccb[j] = 0.0;
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D> q;
    load_face_geo<D>(m, r.f, r.P, r.N, q);
    FaceGeoBarF<D> qf;
    qf.zero();
    FaceGeoBar<D> qs;
    qs.zero();
    ScalarBar<D> Bd;
    Bd.zero();
    const double Qf = mdot[r.f] * irho;
    double xfb[D];  // face-centroid part the folds do not see (none here; kept for symmetry)
#pragma unroll
    for (int j = 0; j < D; ++j) xfb[j] = 0.0;
    if (o >= 0) {
      FlowCell<D, double> O;
      load_flow<D, double>(a, m, o, true, ADVF == kAdvLimited, O);
      double lf[NF];
#pragma unroll
      for (int v = 0; v < NF; ++v) {
        double lo = lam[(int64_t)v * m.C + o];
        if (v == D && o == a.pin) lo = 0.0;
        lf[v] = r.own ? lc[v] - lo : lo - lc[v];
      }
      lf[D] += mdot_bar[r.f];
      const FlowCell<D, double>& P = r.own ? X : O;
      const FlowCell<D, double>& N = r.own ? O : X;
      double VPb = 0.0, VNb = 0.0;
      geo_adj_flux_interior<D, ADVF>(q, P, N, lf, qf, VPb, VNb);
      const double apb_o = bar.aP[o];
      geo_adj_ap_interior<D, ADVF>(q, P, N, r.own ? apb_c : apb_o, r.own ? apb_o : apb_c, qf);
      Vb += r.own ? VPb : VNb;
      ScalarCell<D, double> Ko, Wo;
      load_scalar<D, double>(a, m, o, D + 1, ADVS == kAdvLimited, sc.sk1, sc.sk2, Ko);
      load_scalar<D, double>(a, m, o, D + 2, ADVS == kAdvLimited, sc.sw1, sc.sw2, Wo);
      const double lk_o = lam[(int64_t)(D + 1) * m.C + o];
      const double lw_o = a.wallc[o] ? 0.0 : lam[(int64_t)(D + 2) * m.C + o];
      const double lkf = r.own ? lck - lk_o : lk_o - lck, lwf = r.own ? lcw - lw_o : lw_o - lcw;
      const ScalarCell<D, double>& KP = r.own ? Kc : Ko;
      const ScalarCell<D, double>& KN = r.own ? Ko : Kc;
      const ScalarCell<D, double>& WP = r.own ? Wc : Wo;
      const ScalarCell<D, double>& WN = r.own ? Wo : Wc;
      adj_sc_interior<D, ADVS, true>(q, Qf, KP.phi, KN.phi, KP.G, KN.G, KP.psi, KN.psi, KP.gam, KN.gam, lkf, 0, Bd,
                                     qs);
      adj_sc_interior<D, ADVS, true>(q, Qf, WP.phi, WN.phi, WP.G, WN.G, WP.psi, WN.psi, WP.gam, WN.gam, lwf, 0, Bd,
                                     qs);
      fold_face_geo_bar<D>(q, true, qf);
      adj_face_geo<D>(q, true, qs);
#pragma unroll
      for (int j = 0; j < D; ++j) ccb[j] += r.own ? qf.xP[j] + qs.xP[j] : qf.xN[j] + qs.xN[j];
    } else {
      const int kind = a.kind[r.fu];
      double lf[NF], ub[D];
#pragma unroll
      for (int v = 0; v < NF; ++v) lf[v] = lc[v];
      lf[D] += mdot_bar[r.f];
#pragma unroll
      for (int i = 0; i < D; ++i) ub[i] = a.ub[(int64_t)i * m.F + r.fu];
      const double gB = rans_gamma_b<double>(a, kind, r.fu, c, X.mu);
      double VPb = 0.0;
      geo_adj_flux_boundary<D, ADVF>(q, kind, X, ub, a.pb[r.fu], gB, lf, qf, VPb);
      geo_adj_ap_boundary<D, ADVF>(q, kind, X, gB, apb_c, qf);
      Vb += VPb;
      // k face (wall-function-reduced diffusivity on the SST wall faces)
      const bool wf = a.wallf[r.fu] != 0;
      double gk = Kc.gam, vb, tb;
      if (wf) gk = (1.0 - wall_function_weight<double>(sc, a.nu[c], dP, Kc.phi)) * Kc.gam;
      adj_sc_boundary<D, ADVS, true>(a.kbc[r.fu], q, Qf, Kc.phi, Kc.G, Kc.psi, gk, a.kval[r.fu], 0.0, lck, Bd, vb, tb,
                                     qs);
      if (wf) {  // d gk / d d = -gamma_k d w / d d
        const dual wd = wall_function_weight<dual>(sc, dual{a.nu[c], 0.0}, dual{dP, 1.0}, dual{Kc.phi, 0.0});
        db -= Bd.gam * Kc.gam * wd.d;
      }
      adj_sc_boundary<D, ADVS, true>(a.wbc[r.fu], q, Qf, Wc.phi, Wc.G, Wc.psi, Wc.gam, a.wval[r.fu], 0.0, lcw, Bd,
                                     vb, tb, qs);
      fold_face_geo_bar<D>(q, false, qf);
      adj_face_geo<D>(q, false, qs);
#pragma unroll
      for (int j = 0; j < D; ++j) ccb[j] += qf.xP[j] + qs.xP[j];
      if (wall && wf && tauA != 0.0) {  // tau = sum_f r_f A_f / sum_f A_f: d tau / d A_f = (r_f - tau) / sum A
        double xcl[D + 1], s2 = 0.0;
#pragma unroll
        for (int v = 0; v <= D; ++v) xcl[v] = a.x[(int64_t)v * m.C + c];
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const double sl = xcl[j] - rans_bval_lead<D, double>(a, j, m.F, r.fu, xcl);
          s2 += sl * sl;
        }
        qf.A += tauA * (safe_sqrt(s2) / dP - a.tau[c]);
      }
    }
    if (r.own) {
      gb.area[r.fu] = qf.A + qs.A;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        gb.nrm[(int64_t)j * m.F + r.fu] = qf.n[j] + qs.n[j];
        gb.fc[(int64_t)j * m.F + r.fu] = (qf.xf[j] + qs.xf[j]) + xfb[j];
      }
    }
  }
  gb.vol[c] += Vb;
  gb.dwall[c] += db;
#pragma unroll
  for (int j = 0; j < D; ++j) gb.cc[(int64_t)j * m.C + c] += ccb[j];
}

// Gradients of all D + 3 fields (Green-Gauss: the face sums and 1/V was in the cell pass; least
// squares: the fit directions) and the limiters' face-centroid parts.  Adds to the arrays.
template <int D, int GRAD, bool LFV, bool LSV>
__global__ void __launch_bounds__(256) k_rans_geo_grad(MeshView m, RansArgs<double> a, RansBars bar, RansGeoBars gb) {
  constexpr int NA = D + 3;
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  double ccb[D];
#pragma unroll
  for (int j = 0; j < D; ++j) ccb[j] = 0.0;
  double xc[NA];
#pragma unroll
  for (int v = 0; v < NA; ++v) xc[v] = a.x[(int64_t)v * m.C + c];
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    double Ab = 0.0, nb[D], xfb[D], gbar = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) nb[j] = xfb[j] = 0.0;
    FaceGeo<D> q;
    load_face_geo<D>(m, r.f, r.P, r.N, q);
    if (o >= 0) {
      if (GRAD != 1) {  // acc_P += phi_f n A, acc_N -= phi_f n A (compact, or the corrected scheme's b)
        const double w = 1.0 - q.g;
#pragma unroll
        for (int v = 0; v < NA; ++v) {
          const double xo = a.x[(int64_t)v * m.C + o];
          const double phP = r.own ? xc[v] : xo, phN = r.own ? xo : xc[v];
          const double phf = w * phP + q.g * phN;
          double zn = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) {
            const double yc = bar.G[(int64_t)(v * D + j) * m.C + c], yo = bar.G[(int64_t)(v * D + j) * m.C + o];
            const double z = r.own ? yc - yo : yo - yc;
            nb[j] += z * phf * q.A;
            zn += z * q.n[j];
          }
          Ab += zn * phf;
          gbar += zn * q.A * (phN - phP);
        }
      } else {
        double d[D], dphi[NA], db[D], Gc[NA][D], yc[NA][D];
#pragma unroll
        for (int j = 0; j < D; ++j) d[j] = m.cc[(int64_t)j * m.C + o] - m.cc[(int64_t)j * m.C + c];
#pragma unroll
        for (int v = 0; v < NA; ++v) {
          dphi[v] = a.x[(int64_t)v * m.C + o] - xc[v];
#pragma unroll
          for (int j = 0; j < D; ++j) {
            Gc[v][j] = a.G[(int64_t)(v * D + j) * m.C + c];
            yc[v][j] = bar.G[(int64_t)(v * D + j) * m.C + c];
          }
        }
        lsq_dir_bar<D, NA>(d, dphi, Gc, yc, db);  // own fit: d = x_o - x_c
#pragma unroll
        for (int j = 0; j < D; ++j) {
          ccb[j] -= db[j];
          d[j] = -d[j];
        }
#pragma unroll
        for (int v = 0; v < NA; ++v) {
          dphi[v] = -dphi[v];
#pragma unroll
          for (int j = 0; j < D; ++j) {
            Gc[v][j] = a.G[(int64_t)(v * D + j) * m.C + o];
            yc[v][j] = bar.G[(int64_t)(v * D + j) * m.C + o];
          }
        }
        lsq_dir_bar<D, NA>(d, dphi, Gc, yc, db);  // the other cell's fit: d = x_c - x_o
#pragma unroll
        for (int j = 0; j < D; ++j) ccb[j] += db[j];
      }
    } else {
      double phb[NA];
#pragma unroll
      for (int v = 0; v < NA; ++v) phb[v] = rans_bval_lead<D, double>(a, v, m.F, r.fu, xc);
      if (GRAD != 1) {
#pragma unroll
        for (int v = 0; v < NA; ++v) {
          double zn = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) {
            const double y = bar.G[(int64_t)(v * D + j) * m.C + c];
            nb[j] += y * phb[v] * q.A;
            zn += y * q.n[j];
          }
          Ab += zn * phb[v];
        }
      } else {
        double d[D], dphi[NA], db[D], Gc[NA][D], yc[NA][D];
#pragma unroll
        for (int j = 0; j < D; ++j) d[j] = q.xf[j] - m.cc[(int64_t)j * m.C + c];
#pragma unroll
        for (int v = 0; v < NA; ++v) {
          dphi[v] = phb[v] - xc[v];
#pragma unroll
          for (int j = 0; j < D; ++j) {
            Gc[v][j] = a.G[(int64_t)(v * D + j) * m.C + c];
            yc[v][j] = bar.G[(int64_t)(v * D + j) * m.C + c];
          }
        }
        lsq_dir_bar<D, NA>(d, dphi, Gc, yc, db);
#pragma unroll
        for (int j = 0; j < D; ++j) {
          xfb[j] += db[j];
          ccb[j] -= db[j];
        }
      }
    }
    if (GRAD != 1) {  // fold g (interior) onto the centroids
      FaceGeoBarF<D> qf;
      qf.zero();
      qf.g = gbar;
      fold_face_geo_bar<D>(q, o >= 0, qf);
#pragma unroll
      for (int j = 0; j < D; ++j) {
        xfb[j] += qf.xf[j];
        ccb[j] += r.own ? qf.xP[j] : qf.xN[j];
      }
    }
    if (GRAD == kGradCorrected && o >= 0) {  // the sweeps' own face terms
      double xPb[D], xNb[D];
#pragma unroll
      for (int j = 0; j < D; ++j) xPb[j] = xNb[j] = 0.0;
      cgg_face_geo_bar<D>(gb.cgg, m.F, r.f, q, Ab, nb, xfb, xPb, xNb);
#pragma unroll
      for (int j = 0; j < D; ++j) ccb[j] += r.own ? xPb[j] : xNb[j];
    }
    if (LFV || LSV) {
      if (r.own || o < 0) {
#pragma unroll
        for (int v = 0; v < NA; ++v) {  // the limited fields' slots only (the others are never written)
          const int l = lim_slot<D>(v, LFV, LSV);
          if (l < 0) continue;
          if (gb.limf[(int64_t)l * m.C + c] == r.f) {
#pragma unroll
            for (int j = 0; j < D; ++j) xfb[j] += gb.limx[(int64_t)(l * D + j) * m.C + c];
          }
          if (o >= 0 && gb.limf[(int64_t)l * m.C + o] == r.f) {
#pragma unroll
            for (int j = 0; j < D; ++j) xfb[j] += gb.limx[(int64_t)(l * D + j) * m.C + o];
          }
        }
      }
    }
    if (r.own) {
      gb.area[r.fu] += Ab;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        gb.nrm[(int64_t)j * m.F + r.fu] += nb[j];
        gb.fc[(int64_t)j * m.F + r.fu] += xfb[j];
      }
    }
  }
#pragma unroll
  for (int j = 0; j < D; ++j) gb.cc[(int64_t)j * m.C + c] += ccb[j];
  if (GRAD == kGradCorrected) gb.vol[c] += gb.cgg.vol[c];  // the sweeps' V^-1
}

// ---------------------------------------------------------------------------------------------
// host side

int64_t rans_work_doubles(int D, int64_t C, int64_t F, int mode) {
  const int NA = D + 3, NL = D + 2;
  int64_t fwd = (int64_t)(NA * D + NL + 7 + 1) * C + F;  // G, psi, closure(7), aP, face mass flux
  const int64_t cgg = (int64_t)2 * NA * D * C;            // corrected Green-Gauss sweep scratch
  const int64_t phys = (int64_t)NA * C;                   // physical state of a transformed k / omega
  if (mode == 1) return 2 * fwd + cgg + 2 * phys;
  // mode 2 adds the cotangents, the stencil extremum owners (2 NL int32) and the limiter argmin
  // face positions (NL int32)
  if (mode >= 2)
    return fwd + (int64_t)(NA * D + 1 + NL + 2 * NL + 7) * C + (int64_t)NL * C + ((int64_t)NL * C + 1) / 2 + F + cgg +
           phys + (mode == 3 ? (int64_t)(NL * D + NL + 1) * C : 0);  // mode 3: geometry-cotangent scratch
  return fwd + cgg + phys;
}

struct RWork {
  double *G, *psi, *aP, *cl, *mdot;  // cl: 7 closure arrays
  double *Gt, *psit, *aPt, *clt, *mdott;  // tangents
  double *Gb, *aPb, *psib, *extb, *mueffb, *nutb, *f1b, *Srb, *cdb, *taub, *nutwb, *mdotb;
  int32_t* ext;
  int32_t* kmin;  // [NL][C] limiter argmin faces (internal face index; VJP only)
  double* cgg;
  double *xs, *xst;  // physical state (and tangent) when k or omega is solved in log form
  double *limx, *tauA;  // mode 3
  int32_t* limf;
};

RWork rcarve(void* base, int D, int64_t C, int64_t F, int mode) {
  const int NA = D + 3, NL = D + 2;
  double* p = (double*)base;
  RWork w{};
  auto take = [&](int64_t n) {
    double* r = p;
    p += n;
    return r;
  };
  w.G = take((int64_t)NA * D * C);
  w.psi = take((int64_t)NL * C);
  w.aP = take(C);
  w.cl = take(7 * C);
  w.mdot = take(F);
  if (mode == 1) {
    w.Gt = take((int64_t)NA * D * C);
    w.psit = take((int64_t)NL * C);
    w.aPt = take(C);
    w.clt = take(7 * C);
    w.mdott = take(F);
  }
  if (mode >= 2) {
    w.Gb = take((int64_t)NA * D * C);
    w.aPb = take(C);
    w.psib = take((int64_t)NL * C);
    w.extb = take((int64_t)2 * NL * C);
    w.mueffb = take(C);
    w.nutb = take(C);
    w.f1b = take(C);
    w.Srb = take(C);
    w.cdb = take(C);
    w.taub = take(C);
    w.nutwb = take(C);
    w.ext = (int32_t*)take((int64_t)NL * C);
    w.kmin = (int32_t*)take(((int64_t)NL * C + 1) / 2);
    w.mdotb = take(F);
  }
  w.cgg = take((int64_t)2 * NA * D * C);
  w.xs = take((int64_t)NA * C);
  if (mode == 1) w.xst = take((int64_t)NA * C);
  if (mode == 3) {
    w.limx = take((int64_t)NL * D * C);
    w.limf = (int32_t*)take((int64_t)NL * C);
    w.tauA = take(C);
  }
  return w;
}

SstConst make_sst(const afx_rans_config* cfg) {
  SstConst s;
  const double* k = cfg->sst;
  s.a1 = k[0];
  s.beta_star = k[1];
  s.alpha1 = k[2];
  s.beta1 = k[3];
  s.sk1 = k[4];
  s.sw1 = k[5];
  s.alpha2 = k[6];
  s.beta2 = k[7];
  s.sk2 = k[8];
  s.sw2 = k[9];
  s.kappa = k[10];
  s.e_wall = k[11];
  s.p_exp = k[12];
  s.c_vis = k[13];
  s.bsq = k[14];
  s.ylam = k[15];
  s.density = cfg->density;
  return s;
}

template <class S>
RansArgs<S> rargs(const afx_rans_config* cfg, const afx_rans_inputs* in, const afx_rans_inputs* tan) {
  RansArgs<S> a{};
  a.kind = cfg->flow.face_kind;
  a.kbc = cfg->k_kind;
  a.wbc = cfg->omega_kind;
  a.wallf = cfg->wall_face;
  a.wallc = cfg->wall_cell;
  a.x.v = in->state;
  a.ub.v = in->u_b;
  a.pb.v = in->p_b;
  a.mu.v = in->mu;
  a.rho.v = in->rho;
  a.nu.v = in->nu;
  a.kval.v = in->k_value;
  a.wval.v = in->omega_value;
  if (tan) {
    a.x.t = tan->state;
    a.ub.t = tan->u_b;
    a.pb.t = tan->p_b;
    a.mu.t = tan->mu;
    a.rho.t = tan->rho;
    a.nu.t = tan->nu;
    a.kval.t = tan->k_value;
    a.wval.t = tan->omega_value;
  }
  a.dwall = cfg->wall_distance;
  a.dgrad = cfg->wall_distance_gradient;
  a.pin = cfg->flow.pressure_pin >= 0 ? (int32_t)cfg->flow.pressure_pin : -1;
  a.pin_value = cfg->flow.pressure_pin_value;
  for (int i = 0; i < 3; ++i) a.bf[i] = cfg->flow.body_force[i];
  a.lim_k = cfg->flow.limiter_k;
  a.sc = make_sst(cfg);
  a.sweeps = cfg->flow.gradient_sweeps;
  a.relax = cfg->flow.gradient_relaxation;
  a.log_w = cfg->omega_transform == AFX_SCALAR_LOG;
  return a;
}

// Solved -> physical state (ScalarVariableTransform.to_physical, turbulence/coupled.py:379-392):
// k (row D+1) and omega (row D+2) are e^w when solved in log form, every other row is copied; the
// tangent is e^w dw.
__global__ void __launch_bounds__(256) k_rans_to_physical(int64_t C, int NA, int logk_row, int logw_row,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ xt, double* __restrict__ xs,
                                                          double* __restrict__ xst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)NA * C) return;
  const int v = (int)(i / C);
  const double w = x[i];
  const bool lg = v == logk_row || v == logw_row;
  const double e = lg ? exp(w) : w;
  xs[i] = e;
  if (xst) xst[i] = xt ? (lg ? e * xt[i] : xt[i]) : 0.0;
}

// Cotangent of the solved log-form rows: wbar = e^w * phibar (in place).
__global__ void __launch_bounds__(256) k_rans_log_bar(int64_t C, int row, const double* __restrict__ xs,
                                                      double* __restrict__ xb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C) return;
  const int64_t j = (int64_t)row * C + i;
  xb[j] *= xs[j];
}

bool rans_transformed(const afx_rans_config* cfg) {
  return cfg->k_transform == AFX_SCALAR_LOG || cfg->omega_transform == AFX_SCALAR_LOG;
}

// Points `a` at the physical state; `xs` / `xst` receive it (`xst` only with a tangent).
template <class S>
int rans_bind_physical(const afx_mesh* m, const afx_rans_config* cfg, const afx_rans_inputs* in,
                       const afx_rans_inputs* tan, double* xs, double* xst, RansArgs<S>& a, hipStream_t s) {
  if (!rans_transformed(cfg)) return 0;
  const int D = m->dim, NA = D + 3;
  const int64_t C = m->C;
  const double* xt = tan ? tan->state : nullptr;
  k_rans_to_physical<<<grid_for((int64_t)NA * C), dim3(kBlock), 0, s>>>(
      C, NA, cfg->k_transform == AFX_SCALAR_LOG ? D + 1 : -1, cfg->omega_transform == AFX_SCALAR_LOG ? D + 2 : -1,
      in->state, xt, xs, tan ? xst : nullptr);
  AFX_CHECK_LAUNCH();
  a.x.v = xs;
  a.x.t = tan ? xst : nullptr;
  return 0;
}

template <class S>
void bind_closure(RansArgs<S>& a, double* cl, double* clt, int64_t C) {
  In<S>* dst[7] = {&a.nut, &a.f1, &a.Sr, &a.cd, &a.nutw, &a.tau, &a.mueff};
  for (int i = 0; i < 7; ++i) {
    dst[i]->v = cl + i * C;
    dst[i]->t = clt ? clt + i * C : nullptr;
  }
}

// Forward passes: G, closure, a_P; then, for `R` set, the flow and k / omega residual rows; for
// `mdot_only`, just the face mass flux (and the flow rows when R is set) for the reverse passes.
// GEO: the geometry carries tangents in a.gt (S = dual).  iters (corrected Green-Gauss, nullable):
// keep the sweeps' iterates for the geometry adjoint.
template <int D, class S, int ADVF, int ADVS, int GRAD, bool GEO = false>
int rans_forward(const afx_mesh* mh, const afx_geometry* g, RansArgs<S> a, Out<S> R, const RWork& w, double* Gt,
                 double* psit, double* aPt, double* clt, double* mdott, int32_t* ext, bool mdot_only, hipStream_t s,
                 double* iters = nullptr) {
  static_assert(!GEO || std::is_same<S, dual>::value, "geometry tangents need a dual state");
  MeshView m;
  if (int rv_ = afx_rview(mh, g, s, &m)) return rv_;
  const int64_t C = m.C;
  Out<S> oG{w.G, Gt}, oP{w.psi, psit};
  constexpr bool CG = GRAD == kGradCorrected;
  constexpr int GK = CG ? 0 : GRAD;
  constexpr bool LF = ADVF == kAdvLimited, LS = ADVS == kAdvLimited;
  ClosureOut co{};
  {
    double** vv[7] = {&co.nut, &co.f1, &co.Sr, &co.cd, &co.nutw, &co.tau, &co.mueff};
    double** tt[7] = {&co.nut_t, &co.f1_t, &co.Sr_t, &co.cd_t, &co.nutw_t, &co.tau_t, &co.mueff_t};
    for (int i = 0; i < 7; ++i) {
      *vv[i] = w.cl + i * C;
      *tt[i] = clt ? clt + i * C : nullptr;
    }
  }
  // the reverse pass (ext given) keeps the least-squares inverse in the corrected scheme's sweep
  // scratch, which that scheme alone uses
  if constexpr (GRAD == 1 && !GEO) co.lsqI = ext ? w.cgg : nullptr;
  // the closure rides on the gradient pass unless corrected-GG sweeps still change the gradient
  // (the dual-number build keeps them apart: fused, it would drop to one wave per SIMD)
  constexpr bool FUSE = !CG && std::is_same<S, double>::value;
  AFX_LAUNCH(kKRansGrad, (k_rans_grad<D, S, GK, LF && !CG, LS && !CG, FUSE, GEO>), grid_for(C), dim3(kBlock), 0, s,
             m, a, oG, oP, ext, ext ? w.kmin : nullptr, co);
  AFX_CHECK_LAUNCH();
  a.G.v = w.G;
  a.G.t = Gt;
  if constexpr (CG) {
    const int sw = a.sweeps;
    const double rl = a.relax;
    int rc = 0;
    if constexpr (GEO) {  // the sweeps with the geometry's tangents
      double* scr = afx_mesh_scratch(const_cast<afx_mesh*>(mh), (int64_t)4 * (D + 3) * D * C);
      AFX_REQUIRE(scr, "rans: could not allocate the corrected-gradient sweep scratch");
      rc = cgg_sweeps_tangent(m, a.gt, D, D + 3, w.G, Gt, scr, sw, rl, s);
      if (rc) return rc;
    } else {
      rc = iters ? cgg_sweeps_store(m, D, D + 3, w.G, iters, sw, rl, s) : cgg_sweeps(m, D, D + 3, w.G, w.cgg, sw, rl, s);
      if (rc) return rc;
      if (Gt) {
        rc = cgg_sweeps(m, D, D + 3, Gt, w.cgg, sw, rl, s);
        if (rc) return rc;
      }
    }
    LimiterSlots sl{};
    sl.ext_stride = D + 2;
    sl.kmin = ext ? w.kmin : nullptr;
    if (LF)
      for (int i = 0; i < D; ++i) {
        sl.var[sl.nl] = i;
        sl.slot[sl.nl++] = i;
      }
    if (LS) {
      sl.var[sl.nl] = D + 1;
      sl.slot[sl.nl++] = D;
      sl.var[sl.nl] = D + 2;
      sl.slot[sl.nl++] = D + 1;
    }
    if constexpr (GEO)
      rc = venkat_pass_tangent(m, a.gt, D, sl, a.x, a.G, a.lim_k, oP, s);
    else
      rc = venkat_pass(m, D, sl, a.x, a.G, a.lim_k, oP, ext, s);
    if (rc) return rc;
  }
  a.psi.v = w.psi;
  a.psi.t = psit;
  if (!FUSE) {
    AFX_LAUNCH(kKSstClosure, (k_rans_closure<D, S, GEO>), grid_for(C), dim3(kBlock), 0, s, m, a, co);
    AFX_CHECK_LAUNCH();
  }
  bind_closure<S>(a, w.cl, clt, C);
  AFX_LAUNCH(kKRansAp, (k_rans_ap<D, S, ADVF, GEO>), grid_for(C), dim3(kBlock), 0, s, m, a, Out<S>{w.aP, aPt});
  AFX_CHECK_LAUNCH();
  if (R.v == nullptr && !mdot_only) return 0;
  a.aP.v = w.aP;
  a.aP.t = aPt;
  AFX_LAUNCH(kKSstRes, (k_rans_fres<D, S, ADVF, GEO>), grid_for(C), dim3(kBlock), 0, s, m, a, R,
             Out<S>{w.mdot, mdott});
  AFX_CHECK_LAUNCH();
  if (mdot_only) return 0;
  In<S> md;
  md.v = w.mdot;
  md.t = mdott;
  AFX_LAUNCH(kKRansScalarRes, (k_rans_sres<D, S, ADVS, GEO>), grid_for(C), dim3(kBlock), 0, s, m, a, md, R);
  AFX_CHECK_LAUNCH();
  return 0;
}

template <class Fn>
int rdispatch(int D, int advf, int advs, int grad, Fn&& fn) {
#define AFX_R(DD, AF, AS, GG)                                                                            \
  if (D == DD && advf == AF && advs == AS && grad == GG)                                                 \
    return fn(std::integral_constant<int, DD>{}, std::integral_constant<int, AF>{},                       \
              std::integral_constant<int, AS>{}, std::integral_constant<int, GG>{});
  AFX_R(2, 1, 1, 0) AFX_R(2, 2, 2, 0) AFX_R(2, 1, 1, 1) AFX_R(2, 2, 2, 1) AFX_R(2, 1, 2, 1) AFX_R(2, 2, 1, 1)
  AFX_R(3, 1, 1, 0) AFX_R(3, 2, 2, 0) AFX_R(3, 1, 1, 1) AFX_R(3, 2, 2, 1) AFX_R(3, 1, 2, 1) AFX_R(3, 2, 1, 1)
  AFX_R(2, 1, 1, 3) AFX_R(2, 2, 2, 3) AFX_R(3, 1, 1, 3) AFX_R(3, 2, 2, 3)
#undef AFX_R
  afx_set_error("rans: unsupported (dim, advection, scalar advection, gradient) combination");
  return 2;
}

int rcheck(const afx_mesh* m, const afx_geometry* g, const afx_rans_config* cfg, const afx_rans_inputs* in,
           void* work, int64_t work_size, int mode) {
  AFX_REQUIRE(m && g && cfg && in, "rans: null argument");
  AFX_REQUIRE(in->state && in->mu && in->rho && in->nu && in->u_b && in->p_b && in->k_value && in->omega_value,
              "rans: a required input pointer is null");
  AFX_REQUIRE(cfg->flow.face_kind && cfg->k_kind && cfg->omega_kind && cfg->wall_face && cfg->wall_cell &&
                  cfg->wall_distance && cfg->wall_distance_gradient,
              "rans: a required configuration pointer is null");
  AFX_REQUIRE(cfg->density > 0.0, "rans: density must be positive");
  AFX_REQUIRE(cfg->k_transform >= 0 && cfg->k_transform <= 1 && cfg->omega_transform >= 0 &&
                  cfg->omega_transform <= 1,
              "rans: k_transform / omega_transform must be AFX_SCALAR_DIRECT or AFX_SCALAR_LOG");
  AFX_REQUIRE(cfg->flow.gradient != kGradCorrected || cfg->flow.gradient_relaxation > 0.0,
              "rans: gradient relaxation must be positive");
  AFX_REQUIRE(work && work_size >= rans_work_doubles(m->dim, m->C, m->F, mode) * 8, "rans: workspace too small");
  return 0;
}

}  // namespace

extern "C" int afx_rans_workspace_bytes(const afx_mesh* m, int mode, int64_t* bytes) {
  AFX_REQUIRE(m && bytes && mode >= 0 && mode <= 3, "afx_rans_workspace_bytes: bad argument");
  *bytes = rans_work_doubles(m->dim, m->C, m->F, mode) * 8 + 256;
  return 0;
}

extern "C" int afx_rans_residual(const afx_mesh* m, const afx_geometry* g, const afx_rans_config* cfg,
                                 const afx_rans_inputs* in, double* residual, void* work, int64_t work_size,
                                 hipStream_t s) {
  int rc = rcheck(m, g, cfg, in, work, work_size, 0);
  if (rc) return rc;
  AFX_REQUIRE(residual, "afx_rans_residual: null residual");
  RWork w = rcarve(work, m->dim, m->C, m->F, 0);
  RansArgs<double> a = rargs<double>(cfg, in, nullptr);
  rc = rans_bind_physical<double>(m, cfg, in, nullptr, w.xs, nullptr, a, s);
  if (rc) return rc;
  return rdispatch(m->dim, cfg->flow.advection, cfg->scalar_advection, cfg->flow.gradient,
                   [&](auto Dc, auto Fc, auto Sc, auto Gc) {
                     return rans_forward<decltype(Dc)::value, double, decltype(Fc)::value, decltype(Sc)::value,
                                         decltype(Gc)::value>(m, g, a, Out<double>{residual, nullptr}, w, nullptr,
                                                              nullptr, nullptr, nullptr, nullptr, nullptr, false, s);
                   });
}

extern "C" int afx_rans_jvp(const afx_mesh* m, const afx_geometry* g, const afx_rans_config* cfg,
                            const afx_rans_inputs* in, const afx_rans_inputs* tangent, double* residual,
                            double* residual_tangent, void* work, int64_t work_size, hipStream_t s) {
  int rc = rcheck(m, g, cfg, in, work, work_size, 1);
  if (rc) return rc;
  AFX_REQUIRE(tangent && residual && residual_tangent, "afx_rans_jvp: null argument");
  RWork w = rcarve(work, m->dim, m->C, m->F, 1);
  RansArgs<dual> a = rargs<dual>(cfg, in, tangent);
  rc = rans_bind_physical<dual>(m, cfg, in, tangent, w.xs, w.xst, a, s);
  if (rc) return rc;
  return rdispatch(m->dim, cfg->flow.advection, cfg->scalar_advection, cfg->flow.gradient,
                   [&](auto Dc, auto Fc, auto Sc, auto Gc) {
                     return rans_forward<decltype(Dc)::value, dual, decltype(Fc)::value, decltype(Sc)::value,
                                         decltype(Gc)::value>(m, g, a, Out<dual>{residual, residual_tangent}, w,
                                                              w.Gt, w.psit, w.aPt, w.clt, w.mdott, nullptr, false, s);
                   });
}

extern "C" int afx_rans_jvp_geometry(const afx_mesh* m, const afx_geometry* g, const afx_geometry* geom_tangent,
                                     const double* wall_distance_tangent,
                                     const double* wall_distance_gradient_tangent, const afx_rans_config* cfg,
                                     const afx_rans_inputs* in, const afx_rans_inputs* tangent, double* residual,
                                     double* residual_tangent, void* work, int64_t work_size, hipStream_t s) {
  int rc = rcheck(m, g, cfg, in, work, work_size, 1);
  if (rc) return rc;
  AFX_REQUIRE(geom_tangent && wall_distance_tangent && wall_distance_gradient_tangent && tangent && residual &&
                  residual_tangent,
              "afx_rans_jvp_geometry: null argument");
  AFX_REQUIRE(geom_tangent->area && geom_tangent->normal && geom_tangent->face_centroid && geom_tangent->volume &&
                  geom_tangent->cell_centroid,
              "afx_rans_jvp_geometry: the geometry tangent needs all five arrays");
  AFX_REQUIRE(m->fu == nullptr || geom_tangent->face_pack,
              "afx_rans_jvp_geometry: this mesh's faces are reordered: geom_tangent->face_pack must hold the face "
              "tangents in the internal order (afx_geometry_pack of the tangent arrays)");
  RWork w = rcarve(work, m->dim, m->C, m->F, 1);
  RansArgs<dual> a = rargs<dual>(cfg, in, tangent);
  const double* pack = geom_tangent->face_pack;
  a.gt.area = m->fu ? pack : geom_tangent->area;
  a.gt.nrm = m->fu ? pack + m->F : geom_tangent->normal;
  a.gt.fc = m->fu ? pack + (int64_t)(1 + m->dim) * m->F : geom_tangent->face_centroid;
  a.gt.vol = geom_tangent->volume;
  a.gt.cc = geom_tangent->cell_centroid;
  a.dwall_t = wall_distance_tangent;
  a.dgrad_t = wall_distance_gradient_tangent;
  rc = rans_bind_physical<dual>(m, cfg, in, tangent, w.xs, w.xst, a, s);
  if (rc) return rc;
  return rdispatch(m->dim, cfg->flow.advection, cfg->scalar_advection, cfg->flow.gradient,
                   [&](auto Dc, auto Fc, auto Sc, auto Gc) {
                     return rans_forward<decltype(Dc)::value, dual, decltype(Fc)::value, decltype(Sc)::value,
                                         decltype(Gc)::value, true>(m, g, a, Out<dual>{residual, residual_tangent}, w,
                                                                    w.Gt, w.psit, w.aPt, w.clt, w.mdott, nullptr,
                                                                    false, s);
                   });
}

namespace {

// The reverse passes, after one forward recompute of G, the closure and a_P; `residual` (nullable)
// also receives R from the same forward products.
int rans_vjp_impl(const afx_mesh* m, const afx_geometry* g, const afx_rans_config* cfg, const afx_rans_inputs* in,
                  const double* cotangent, afx_rans_cotangents* out, double* residual, void* work, int64_t work_size,
                  hipStream_t s, const afx_geometry_cotangents* geom_out = nullptr, double* dwall_bar = nullptr,
                  double* dgrad_bar = nullptr) {
  const int mode = geom_out ? 3 : 2;
  int rc = rcheck(m, g, cfg, in, work, work_size, mode);
  if (rc) return rc;
  AFX_REQUIRE(cotangent && out && out->state && out->mu && out->rho && out->nu && out->u_b && out->p_b,
              "afx_rans_vjp: null cotangent argument");
  RWork w = rcarve(work, m->dim, m->C, m->F, mode);
  RansArgs<double> a = rargs<double>(cfg, in, nullptr);
  rc = rans_bind_physical<double>(m, cfg, in, nullptr, w.xs, nullptr, a, s);
  if (rc) return rc;
  RansBars bar{out->state, out->mu,  out->rho, out->nu, out->u_b, out->p_b, w.Gb,  w.aPb, w.psib,
               w.extb,     w.mueffb, w.nutb,   w.f1b,   w.Srb,    w.cdb,    w.taub, w.nutwb};
  return rdispatch(m->dim, cfg->flow.advection, cfg->scalar_advection, cfg->flow.gradient,
                   [&](auto Dc, auto Fc, auto Sc, auto Gc) {
                     constexpr int D = decltype(Dc)::value, AF = decltype(Fc)::value, AS = decltype(Sc)::value,
                                   GR = decltype(Gc)::value;
                     constexpr bool LF = AF == kAdvLimited, LS = AS == kAdvLimited;
                     // corrected Green-Gauss with geometry cotangents: the sweeps' iterates and their
                     // own geometry cotangents live in the handle's scratch
                     double* iters = nullptr;
                     CggGeoBars cb{nullptr, nullptr};
                     if (GR == kGradCorrected && geom_out) {
                       const int64_t ni = cgg_iterate_doubles(D, D + 3, m->C, a.sweeps);
                       double* scr =
                           afx_mesh_scratch(const_cast<afx_mesh*>(m), ni + cgg_geo_bar_doubles(D, m->C, m->F));
                       AFX_REQUIRE(scr, "afx_rans_vjp_geometry: could not allocate the corrected-gradient scratch");
                       iters = ni ? scr : nullptr;
                       cb.face = scr + ni;
                       cb.vol = cb.face + (int64_t)(2 * D + 1) * m->F;
                     }
                     int r2 = rans_forward<D, double, AF, AS, GR>(m, g, a, Out<double>{residual, nullptr}, w,
                                                                  nullptr, nullptr, nullptr, nullptr, nullptr, w.ext,
                                                                  true, s, iters);
                     if (r2) return r2;
                     MeshView v;
                     if (int rv_ = afx_rview(m, g, s, &v)) return rv_;
                     RansArgs<double> b = a;
                     b.G.v = w.G;
                     b.psi.v = w.psi;
                     b.aP.v = w.aP;
                     bind_closure<double>(b, w.cl, nullptr, m->C);
                     AFX_LAUNCH(kKSstResAdj, (k_rans_sres_adj<D, AS>), grid_for(m->C), dim3(kBlock), 0, s, v, b,
                                w.mdot, cotangent, bar, w.mdotb, residual);
                     AFX_CHECK_LAUNCH();
                     // boundary-value cotangents: zero over all faces, boundary entries written below
                     if (int rz = afx_fill_zero(bar.ub, (int64_t)D * m->F, s)) return rz;
                     if (int rz = afx_fill_zero(bar.pb, m->F, s)) return rz;
                     AFX_LAUNCH(kKRansFlowResAdj, (k_rans_fres_adj<D, AF>), grid_for(m->C), dim3(kBlock), 0, s, v, b,
                                cotangent, w.mdotb, bar);
                     AFX_CHECK_LAUNCH();
                     AFX_LAUNCH(kKRansApAdj, (k_rans_ap_adj<D, AF>), grid_for(m->C), dim3(kBlock), 0, s, v, b, bar);
                     AFX_CHECK_LAUNCH();
                     AFX_LAUNCH(kKSstClosureAdj, (k_rans_closure_adj<D, AF, AS, GR>), grid_for(m->C), dim3(kBlock), 0,
                                s, v, b, w.ext, w.kmin, GR == 1 ? w.cgg : nullptr, bar);
                     AFX_CHECK_LAUNCH();
                     constexpr int GK = GR == kGradCorrected ? 0 : GR;
                     if (GR == kGradCorrected) {
                       int r3 = cb.face ? cgg_sweeps_adjoint_geometry(v, D, D + 3, bar.G, w.G, iters, w.cgg, a.sweeps,
                                                                      a.relax, cb, s)
                                        : cgg_sweeps_adjoint(v, D, D + 3, bar.G, w.cgg, a.sweeps, a.relax, s);
                       if (r3) return r3;
                     }
                     AFX_LAUNCH(kKRansGradAdj, (k_rans_grad_adj<D, GK, LF, LS>), grid_for(m->C), dim3(kBlock), 0, s,
                                v, b, w.ext, bar);
                     AFX_CHECK_LAUNCH();
                     if (geom_out) {  // the geometry's cotangents from the adjoint's intermediates
                       RansGeoBars gb{geom_out->area, geom_out->normal, geom_out->face_centroid,
                                      geom_out->volume, geom_out->cell_centroid, dwall_bar, dgrad_bar,
                                      w.limx, w.limf, w.tauA, cb};
                       AFX_LAUNCH(kKSstClosureAdj, (k_rans_geo_cell<D, LF, LS, GR>), grid_for(m->C), dim3(kBlock),
                                  0, s, v, b, cotangent, w.ext, w.kmin, bar, gb);
                       AFX_CHECK_LAUNCH();
                       AFX_LAUNCH(kKRansFlowResAdj, (k_rans_geo_face<D, AF, AS>), grid_for(m->C), dim3(kBlock), 0,
                                  s, v, b, cotangent, w.mdot, w.mdotb, bar, gb);
                       AFX_CHECK_LAUNCH();
                       AFX_LAUNCH(kKRansGradAdj, (k_rans_geo_grad<D, GR, LF, LS>), grid_for(m->C), dim3(kBlock), 0,
                                  s, v, b, bar, gb);
                       AFX_CHECK_LAUNCH();
                     }
                     for (int row = D + 1; row <= D + 2; ++row) {
                       const int tr = row == D + 1 ? cfg->k_transform : cfg->omega_transform;
                       if (tr != AFX_SCALAR_LOG) continue;
                       k_rans_log_bar<<<grid_for(m->C), dim3(kBlock), 0, s>>>(m->C, row, w.xs, bar.x);
                       AFX_CHECK_LAUNCH();
                     }
                     return 0;
                   });
}

}  // namespace

extern "C" int afx_rans_vjp_geometry(const afx_mesh* m, const afx_geometry* g, const afx_rans_config* cfg,
                                     const afx_rans_inputs* in, const double* cotangent, afx_rans_cotangents* out,
                                     const afx_geometry_cotangents* geom_out, double* wall_distance_bar,
                                     double* wall_distance_gradient_bar, void* work, int64_t work_size,
                                     hipStream_t s) {
  AFX_REQUIRE(m && cfg && geom_out && wall_distance_bar && wall_distance_gradient_bar,
              "afx_rans_vjp_geometry: null argument");
  AFX_REQUIRE(geom_out->area && geom_out->normal && geom_out->face_centroid && geom_out->volume &&
                  geom_out->cell_centroid,
              "afx_rans_vjp_geometry: all five geometry cotangent buffers are needed");
  return rans_vjp_impl(m, g, cfg, in, cotangent, out, nullptr, work, work_size, s, geom_out, wall_distance_bar,
                       wall_distance_gradient_bar);
}

extern "C" int afx_rans_vjp(const afx_mesh* m, const afx_geometry* g, const afx_rans_config* cfg,
                            const afx_rans_inputs* in, const double* cotangent, afx_rans_cotangents* out, void* work,
                            int64_t work_size, hipStream_t s) {
  return rans_vjp_impl(m, g, cfg, in, cotangent, out, nullptr, work, work_size, s);
}

extern "C" int afx_rans_residual_vjp(const afx_mesh* m, const afx_geometry* g, const afx_rans_config* cfg,
                                     const afx_rans_inputs* in, const double* cotangent, double* residual,
                                     afx_rans_cotangents* out, void* work, int64_t work_size, hipStream_t s) {
  AFX_REQUIRE(residual, "afx_rans_residual_vjp: null residual");
  return rans_vjp_impl(m, g, cfg, in, cotangent, out, residual, work, work_size, s);
}

extern "C" int afx_rans_fields(const afx_mesh* m, const afx_geometry* g, const afx_rans_config* cfg,
                               const afx_rans_inputs* in, double* fields, hipStream_t s) {
  AFX_REQUIRE(m && g && cfg && in && fields, "afx_rans_fields: null argument");
  AFX_REQUIRE(cfg->k_transform >= 0 && cfg->k_transform <= 1 && cfg->omega_transform >= 0 &&
                  cfg->omega_transform <= 1,
              "afx_rans_fields: k_transform / omega_transform must be AFX_SCALAR_DIRECT or AFX_SCALAR_LOG");
  // `fields` holds only the forward products (gradients, limiters, a_P, closure); the sweep
  // scratch of the corrected gradient and the physical state of a log-form k / omega live in the
  // mesh handle's scratch.
  RWork w = rcarve(fields, m->dim, m->C, m->F, 0);
  const int NA = m->dim + 3;
  const int64_t nscr = (int64_t)2 * NA * m->dim * m->C + (int64_t)NA * m->C;
  double* scr = afx_mesh_scratch(const_cast<afx_mesh*>(m), nscr);
  AFX_REQUIRE(scr, "afx_rans_fields: could not allocate scratch");
  w.mdot = nullptr;
  w.cgg = scr;
  w.xs = scr + (int64_t)2 * NA * m->dim * m->C;
  RansArgs<double> a = rargs<double>(cfg, in, nullptr);
  int rc = rans_bind_physical<double>(m, cfg, in, nullptr, w.xs, nullptr, a, s);
  if (rc) return rc;
  return rdispatch(m->dim, cfg->flow.advection, cfg->scalar_advection, cfg->flow.gradient,
                   [&](auto Dc, auto Fc, auto Sc, auto Gc) {
                     return rans_forward<decltype(Dc)::value, double, decltype(Fc)::value, decltype(Sc)::value,
                                         decltype(Gc)::value>(m, g, a, Out<double>{nullptr, nullptr}, w, nullptr,
                                                              nullptr, nullptr, nullptr, nullptr, nullptr, false, s);
                   });
}

// ---------------------------------------------------------------------------------------------
// The OpenMP CPU line: afx_rans_residual_vjp on host arrays.  The passes are the HIP kernels' own
// cell bodies, run one cell at a time by OpenMP threads (HostExec: no LDS tile, every neighbour
// read from memory), in the caller's face order.  A measured baseline, not a fallback: nothing in
// the device path calls it.
namespace {

// Cell -> face lists of a host mesh in face order: entry f where the cell owns f, ~f where it is
// f's neighbour, with the other cell alongside (-1 on a boundary face).
struct HostCells {
  std::vector<int32_t> cptr, cface, cnbr;
};

HostCells host_cell_faces(int64_t C, int64_t F, const int32_t* owner, const int32_t* nb) {
  HostCells h;
  h.cptr.assign(C + 1, 0);
  for (int64_t f = 0; f < F; ++f) {
    ++h.cptr[owner[f] + 1];
    if (nb[f] >= 0) ++h.cptr[nb[f] + 1];
  }
  for (int64_t c = 0; c < C; ++c) h.cptr[c + 1] += h.cptr[c];
  h.cface.resize(h.cptr[C]);
  h.cnbr.resize(h.cptr[C]);
  std::vector<int32_t> fill(h.cptr.begin(), h.cptr.end() - 1);
  for (int64_t f = 0; f < F; ++f) {
    const int32_t P = owner[f], N = nb[f];
    h.cface[fill[P]] = (int32_t)f;
    h.cnbr[fill[P]++] = N;
    if (N >= 0) {
      h.cface[fill[N]] = ~(int32_t)f;
      h.cnbr[fill[N]++] = P;
    }
  }
  return h;
}

// Runs a staged body over every cell; each thread owns one staging block, written and never read.
template <class Sh, class Fn>
void host_staged(int64_t C, int threads, Fn&& fn) {
#pragma omp parallel num_threads(threads)
  {
    std::unique_ptr<Sh> sh(new Sh);
#pragma omp for schedule(static)
    for (int64_t c = 0; c < C; ++c) fn(HostExec((int32_t)c), *sh);
  }
}

template <class Fn>
void host_cells(int64_t C, int threads, Fn&& fn) {
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t c = 0; c < C; ++c) fn((int32_t)c);
}

// The pass sequence of rans_vjp_impl with R requested.
template <int D, int AF, int AS, int GR>
int rans_residual_vjp_host_run(const MeshView& m, RansArgs<double> a, const double* lam, double* R, RansBars bar,
                               const RWork& w, const afx_rans_config* cfg, int threads) {
  constexpr bool LF = AF == kAdvLimited, LS = AS == kAdvLimited;
  const int64_t C = m.C, F = m.F;
  ClosureOut co{};
  double** vv[7] = {&co.nut, &co.f1, &co.Sr, &co.cd, &co.nutw, &co.tau, &co.mueff};
  for (int i = 0; i < 7; ++i) *vv[i] = w.cl + i * C;
  co.lsqI = GR == 1 ? w.cgg : nullptr;
  const Out<double> oG{w.G, nullptr}, oP{w.psi, nullptr};
  host_staged<RansGradSh<D, double>>(C, threads, [&](const HostExec& ex, RansGradSh<D, double>& sh) {
    k_rans_grad_body<D, double, GR, LF, LS, true, false>(ex, sh, m, a, oG, oP, w.ext, w.kmin, co);
  });
  a.G.v = w.G;
  a.psi.v = w.psi;
  bind_closure<double>(a, w.cl, nullptr, C);
  host_staged<RansApSh<D, double>>(C, threads, [&](const HostExec& ex, RansApSh<D, double>& sh) {
    k_rans_ap_body<D, double, AF, false>(ex, sh, m, a, Out<double>{w.aP, nullptr});
  });
  a.aP.v = w.aP;
  host_staged<RansFresSh<D>>(C, threads, [&](const HostExec& ex, RansFresSh<D>& sh) {
    k_rans_fres_body<D, double, AF, false>(ex, sh, m, a, Out<double>{R, nullptr}, Out<double>{w.mdot, nullptr});
  });
  host_staged<RansSresAdjSh<D>>(C, threads, [&](const HostExec& ex, RansSresAdjSh<D>& sh) {
    k_rans_sres_adj_body<D, AS>(ex, sh, m, a, w.mdot, lam, bar, w.mdotb, R);
  });
  std::fill(bar.ub, bar.ub + (int64_t)D * F, 0.0);
  std::fill(bar.pb, bar.pb + F, 0.0);
  host_staged<RansFresAdjSh<D>>(C, threads, [&](const HostExec& ex, RansFresAdjSh<D>& sh) {
    k_rans_fres_adj_body<D, AF>(ex, sh, m, a, lam, w.mdotb, bar);
  });
  host_staged<RansApAdjSh<D>>(C, threads, [&](const HostExec& ex, RansApAdjSh<D>& sh) {
    k_rans_ap_adj_body<D, AF>(ex, sh, m, a, bar);
  });
  host_cells(C, threads, [&](int32_t c) { k_rans_closure_adj_body<D, AF, AS, GR>(c, m, a, w.ext, w.kmin, GR == 1 ? w.cgg : nullptr, bar); });
  using GASh = RansGradAdjSh<D, (LF ? D : 0) + (LS ? 2 : 0)>;
  host_staged<GASh>(C, threads, [&](const HostExec& ex, GASh& sh) {
    k_rans_grad_adj_body<D, GR, LF, LS>(ex, sh, m, a, w.ext, bar);
  });
  for (int row = D + 1; row <= D + 2; ++row) {  // log-form rows: wbar = e^w phibar
    const int tr = row == D + 1 ? cfg->k_transform : cfg->omega_transform;
    if (tr != AFX_SCALAR_LOG) continue;
    double* xb = bar.x + (int64_t)row * C;
    const double* xs = a.x.v + (int64_t)row * C;
    host_cells(C, threads, [&](int32_t c) { xb[c] *= xs[c]; });
  }
  return 0;
}

}  // namespace

extern "C" int afx_rans_residual_vjp_host(int dim, int64_t n_cells, int64_t n_faces, const int32_t* owner,
                                          const int32_t* neighbour, const afx_geometry* g,
                                          const afx_rans_config* cfg, const afx_rans_inputs* in,
                                          const double* cotangent, double* residual, afx_rans_cotangents* out,
                                          int threads) {
  AFX_REQUIRE(owner && neighbour && g && cfg && in && cotangent && residual && out,
              "afx_rans_residual_vjp_host: null argument");
  AFX_REQUIRE(dim == 2 || dim == 3, "afx_rans_residual_vjp_host: dim must be 2 or 3");
  AFX_REQUIRE(n_cells > 0 && n_cells < (int64_t)1 << 31 && n_faces > 0 && n_faces < (int64_t)1 << 31,
              "afx_rans_residual_vjp_host: mesh size out of range");
  AFX_REQUIRE(g->area && g->normal && g->face_centroid && g->volume && g->cell_centroid,
              "afx_rans_residual_vjp_host: the geometry needs all five arrays");
  AFX_REQUIRE(in->state && in->mu && in->rho && in->nu && in->u_b && in->p_b && in->k_value && in->omega_value,
              "afx_rans_residual_vjp_host: a required input pointer is null");
  AFX_REQUIRE(cfg->flow.face_kind && cfg->k_kind && cfg->omega_kind && cfg->wall_face && cfg->wall_cell &&
                  cfg->wall_distance && cfg->wall_distance_gradient,
              "afx_rans_residual_vjp_host: a required configuration pointer is null");
  AFX_REQUIRE(out->state && out->mu && out->rho && out->nu && out->u_b && out->p_b,
              "afx_rans_residual_vjp_host: null cotangent argument");
  AFX_REQUIRE(cfg->flow.gradient == 0 || cfg->flow.gradient == 1,
              "afx_rans_residual_vjp_host: compact Green-Gauss or least-squares gradients");
  AFX_REQUIRE(cfg->density > 0.0, "afx_rans_residual_vjp_host: density must be positive");
  AFX_REQUIRE(cfg->k_transform >= 0 && cfg->k_transform <= 1 && cfg->omega_transform >= 0 &&
                  cfg->omega_transform <= 1,
              "afx_rans_residual_vjp_host: k_transform / omega_transform must be AFX_SCALAR_DIRECT or "
              "AFX_SCALAR_LOG");
  if (threads <= 0) threads = omp_get_max_threads();
  const int D = dim, NA = D + 3;
  const int64_t C = n_cells, F = n_faces;
  const HostCells hc = host_cell_faces(C, F, owner, neighbour);
  MeshView m{};
  m.C = (int32_t)C;
  m.F = (int32_t)F;
  m.owner = owner;
  m.nb = neighbour;
  m.cptr = hc.cptr.data();
  m.cface = hc.cface.data();
  m.cnbr = hc.cnbr.data();
  m.area = g->area;
  m.nrm = g->normal;
  m.fc = g->face_centroid;
  m.vol = g->volume;
  m.cc = g->cell_centroid;
  m.fu = nullptr;
  m.bsched = nullptr;
  m.nbs = 0;
  // every workspace array is written by a pass before a later one reads it
  std::unique_ptr<double[]> work(new double[(size_t)rans_work_doubles(D, C, F, 2)]);
  RWork w = rcarve(work.get(), D, C, F, 2);
  RansArgs<double> a = rargs<double>(cfg, in, nullptr);
  if (rans_transformed(cfg)) {  // solved -> physical state
    const int lk = cfg->k_transform == AFX_SCALAR_LOG ? D + 1 : -1;
    const int lw = cfg->omega_transform == AFX_SCALAR_LOG ? D + 2 : -1;
    const double* x = in->state;
    double* xs = w.xs;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t i = 0; i < (int64_t)NA * C; ++i) {
      const int v = (int)(i / C);
      xs[i] = (v == lk || v == lw) ? exp(x[i]) : x[i];
    }
    a.x.v = xs;
  }
  RansBars bar{out->state, out->mu,  out->rho, out->nu, out->u_b, out->p_b, w.Gb,  w.aPb, w.psib,
               w.extb,     w.mueffb, w.nutb,   w.f1b,   w.Srb,    w.cdb,    w.taub, w.nutwb};
  return rdispatch(D, cfg->flow.advection, cfg->scalar_advection, cfg->flow.gradient,
                   [&](auto Dc, auto Fc, auto Sc, auto Gc) {
                     constexpr int GR = decltype(Gc)::value;
                     if constexpr (GR == kGradCorrected) {
                       return 2;  // refused above
                     } else {
                       return rans_residual_vjp_host_run<decltype(Dc)::value, decltype(Fc)::value,
                                                         decltype(Sc)::value, GR>(m, a, cotangent, residual, bar, w,
                                                                                  cfg, threads);
                     }
                   });
}
