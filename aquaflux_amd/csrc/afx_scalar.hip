// Scalar advection-diffusion residual R(phi) of ScalarTransport on a given face flux: forward,
// tangent (dual instantiation) and adjoint -- with respect to phi, the diffusivity, the face flux,
// the boundary values and the geometry (face areas, normals, centroids, cell volumes and
// centroids), so that a node-coordinate adjoint follows from afx_geometry_vjp.
//
// Reference: ScalarTransport.assembler / residual (transport/scalar.py:185-223) ->
// ResidualAssembler.residual (discretization/residual.py:252-412) with (AdvectionFlux, DiffusionFlux)
// summed in that order; CompactGreenGauss (schemes/gradient.py:129-153); VenkatakrishnanLimiter
// (schemes/limiter.py:78-117); boundary closures (boundary/conditions.py:46-201).
//
// Passes (one thread per cell, gather over the cell's faces, no atomics):
//   G  leading-order boundary values -> cell gradient (+ limiter psi and stencil-extremum owners)
//   R  face fluxes Q phi_f - Gamma grad_n phi A summed per cell
// Adjoint: G forward, then R' (state, flux, boundary-value and geometry cotangents of the face
// fluxes), L' (limiter; gradient cotangent premultiplied by the reconstruction operator), G'.
// Face-array cotangents are written by the face's owner thread only; each cell keeps its own
// centroid / volume cotangent, so every output element has exactly one writer per pass.
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "afx_common.h"
#include "afx_mesh.h"
#include "afx_flow_face.h"
#include "afx_scalar_face.h"
#include "afx_scalar_adj.h"
#include "afx_profile.h"
#include "afx_cgg.h"

using namespace afx;

namespace {

constexpr int kGradNone = 2;

template <class S>
struct ScArgs {
  const int8_t* kind;  // boundary closure per face (AFX_BC_*)
  In<S> phi, gam, flux, val, tinf;
  In<S> G, psi;
  In<S> phi_old, phi_older;  // previous time levels (transient term)
  double lim_k;
  int bdf;     // 0 steady, 1 backward Euler, 2 BDF2
  double dt;
};

// Accumulation V (phi - phi_old) / dt (BDF1) or V (1.5 phi - 2 phi_old + 0.5 phi_older) / dt
// (transient.py:35-62).
template <class S, class VT = double>
AFX_DEV S accumulation(const ScArgs<S>& a, int32_t c, S phi, VT V) {
  if (a.bdf == 1) return V * (phi - a.phi_old[c]) / a.dt;
  return V * ((1.5 * phi - 2.0 * a.phi_old[c]) + 0.5 * a.phi_older[c]) / a.dt;
}

template <int D, class S>
AFX_DEV void load_grad_row(const In<S>& G, int32_t C, int32_t c, S (&g)[D]) {
#pragma unroll
  for (int j = 0; j < D; ++j) g[j] = G[(int64_t)j * C + c];
}

// ---------------------------------------------------------------------------------------------
// pass G
// LDS copy of a workgroup's cells as the gradient pass reads its neighbours: value and centroid.
template <int D, class S>
struct ScGradTile {
  S x[kBlock];
  double cc[D][kBlock];
};

template <int D, class S, int GRAD, bool LIM>
__global__ void __launch_bounds__(256) k_sc_grad(MeshView m, ScArgs<S> a, Out<S> G, Out<S> psi_out,
                                                 int32_t* __restrict__ ext, int32_t* __restrict__ kmin_out) {
  __shared__ ScGradTile<D, S> tile;
  const int32_t base = block_base(m);
  const int t = threadIdx.x;
  const int32_t c = base + t;
  const bool valid = c < m.C;
  S xc = sconst<S>(0.0);
  double ccc[D];  // this cell's centroid, for every face
#pragma unroll
  for (int j = 0; j < D; ++j) ccc[j] = 0.0;
  if (valid) {
    xc = a.phi[c];
    tile.x[t] = xc;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      ccc[j] = m.cc[(int64_t)j * m.C + c];
      tile.cc[j][t] = ccc[j];
    }
  }
  __syncthreads();
  if (!valid) return;
  S acc[D];
#pragma unroll
  for (int j = 0; j < D; ++j) acc[j] = sconst<S>(0.0);
  double M[D][D];
#pragma unroll
  for (int p = 0; p < D; ++p)
#pragma unroll
    for (int q = 0; q < D; ++q) M[p][q] = 0.0;
  S mx = xc, mn = xc;
  int32_t amx = c, amn = c;
  const int32_t kb = m.cptr[c], ke = m.cptr[c + 1];
  S zero[D];
#pragma unroll
  for (int j = 0; j < D; ++j) zero[j] = sconst<S>(0.0);
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    if (o >= 0) {
      S xo;
      double cco[D];
      const int ls = tile_slot(o, base);
      if (ls >= 0) {
        xo = tile.x[ls];
#pragma unroll
        for (int j = 0; j < D; ++j) cco[j] = tile.cc[j][ls];
      } else {
        xo = a.phi[o];
#pragma unroll
        for (int j = 0; j < D; ++j) cco[j] = m.cc[(int64_t)j * m.C + o];
      }
      if (LIM) {
        if (val(xo) > val(mx)) {
          mx = xo;
          amx = o;
        }
        if (val(xo) < val(mn)) {
          mn = xo;
          amn = o;
        }
      }
      if (GRAD == 0) {
        FaceGeo<D> q;
        double xPa[D], xNa[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
          xPa[j] = r.own ? ccc[j] : cco[j];
          xNa[j] = r.own ? cco[j] : ccc[j];
        }
        load_face_geo_pts<D>(m, r.f, true, xPa, xNa, q);
        const double w = 1.0 - q.g;
        const S pf = w * (r.own ? xc : xo) + q.g * (r.own ? xo : xc);
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const S t = pf * (q.n[j] * q.A);
          acc[j] = r.own ? acc[j] + t : acc[j] - t;
        }
      } else if (GRAD == 1) {
        double dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = cco[j] - ccc[j];
        const double wgt = 1.0 / vdot<D>(dx, dx);
#pragma unroll
        for (int p = 0; p < D; ++p)
#pragma unroll
          for (int q2 = 0; q2 < D; ++q2) M[p][q2] += wgt * dx[p] * dx[q2];
        const S dphi = xo - xc;
#pragma unroll
        for (int j = 0; j < D; ++j) acc[j] += (wgt * dx[j]) * dphi;
      }
    } else if (GRAD != kGradNone) {
      FaceGeo<D> q;
      load_face_geo_pts<D>(m, r.f, false, ccc, ccc, q);
      const S pb = scalar_bval<D, S>(a.kind[r.fu], q, xc, zero, a.val[r.fu], a.gam[c], a.tinf[r.fu]);
      if (GRAD == 0) {
#pragma unroll
        for (int j = 0; j < D; ++j) acc[j] += pb * (q.n[j] * q.A);
      } else {
        double dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = q.dp[j];
        const double wgt = 1.0 / vdot<D>(dx, dx);
#pragma unroll
        for (int p = 0; p < D; ++p)
#pragma unroll
          for (int q2 = 0; q2 < D; ++q2) M[p][q2] += wgt * dx[p] * dx[q2];
        const S dphi = pb - xc;
#pragma unroll
        for (int j = 0; j < D; ++j) acc[j] += (wgt * dx[j]) * dphi;
      }
    }
  }
  S grad[D];
  if (GRAD == 0) {
    const double iv = 1.0 / m.vol[c];
#pragma unroll
    for (int j = 0; j < D; ++j) grad[j] = acc[j] * iv;
  } else if (GRAD == 1) {
    double I[D][D];
    sym_inverse<D>(M, I);
#pragma unroll
    for (int p = 0; p < D; ++p) {
      S s = I[p][0] * acc[0];
#pragma unroll
      for (int q2 = 1; q2 < D; ++q2) s = s + I[p][q2] * acc[q2];
      grad[p] = s;
    }
  } else {
#pragma unroll
    for (int j = 0; j < D; ++j) grad[j] = sconst<S>(0.0);
  }
#pragma unroll
  for (int j = 0; j < D; ++j) G.set((int64_t)j * m.C + c, grad[j]);
  if (LIM) {
    const double e2 = m.vol[c] * (a.lim_k * a.lim_k * a.lim_k);
    const double* xcen = ccc;
    S lo = sconst<S>(0.0);
    bool first = true;
    int32_t ks = 0;  // the first face attaining the minimum
    for (int32_t k = kb; k < ke; ++k) {
      const int32_t e = m.cface[k];
      const int32_t f = e >= 0 ? e : ~e;
      S dm = grad[0] * (m.fc[f] - xcen[0]);
#pragma unroll
      for (int j = 1; j < D; ++j) dm = dm + grad[j] * (m.fc[(int64_t)j * m.F + f] - xcen[j]);
      const S vf = venkat_face<S>(dm, mx - xc, mn - xc, e2);
      if (first || val(vf) < val(lo)) {
        lo = vf;
        ks = f;
      }
      first = false;
    }
    psi_out.set(c, dmin(dmax(lo, 0.0), 1.0));
    if (ext) {
      ext[c] = amx;
      ext[(int64_t)m.C + c] = amn;
      if (kmin_out) kmin_out[c] = ks;
    }
  }
}

// LDS copy of a workgroup's cells as the face passes read their neighbours: value, diffusivity,
// limiter, gradient, centroid and (adjoint) the residual cotangent.
template <int D, class S>
struct ScTile {
  S f[D + 3][kBlock];
  double cc[D][kBlock];
  double lam[kBlock];
  AFX_DEV void put(int t, S x, S g, S p, const S* G, const double* xc) {
    f[0][t] = x;
    f[1][t] = g;
    f[2][t] = p;
#pragma unroll
    for (int j = 0; j < D; ++j) f[3 + j][t] = G[j];
#pragma unroll
    for (int j = 0; j < D; ++j) cc[j][t] = xc[j];
  }
  AFX_DEV void get(int t, S& x, S& g, S& p, S* G, double* xo) const {
    x = f[0][t];
    g = f[1][t];
    p = f[2][t];
#pragma unroll
    for (int j = 0; j < D; ++j) G[j] = f[3 + j][t];
#pragma unroll
    for (int j = 0; j < D; ++j) xo[j] = cc[j][t];
  }
};

// ---------------------------------------------------------------------------------------------
// pass R
template <int D, class S, int ADV>
__global__ void __launch_bounds__(256) k_sc_res(MeshView m, ScArgs<S> a, Out<S> R) {
  __shared__ ScTile<D, S> tile;
  const int32_t base = block_base(m);
  const int t = threadIdx.x;
  const int32_t c = base + t;
  const bool valid = c < m.C;
  S xc, gc, pc;
  S Gc[D];
  double ccc[D];
  if (valid) {
    xc = a.phi[c];
    gc = a.gam[c];
    pc = ADV == kAdvLimited ? a.psi[c] : sconst<S>(1.0);
    load_grad_row<D, S>(a.G, m.C, c, Gc);
#pragma unroll
    for (int j = 0; j < D; ++j) ccc[j] = m.cc[(int64_t)j * m.C + c];
    tile.put(t, xc, gc, pc, Gc, ccc);
  }
  __syncthreads();
  if (!valid) return;
  S acc = sconst<S>(0.0);
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D> q;
    const S Qf = a.flux[r.fu];
    S fl;
    if (o >= 0) {
      S xo, go, po;
      S Go[D];
      double cco[D];
      const int ls = tile_slot(o, base);
      if (ls >= 0) {
        tile.get(ls, xo, go, po, Go, cco);
      } else {
        xo = a.phi[o];
        go = a.gam[o];
        po = ADV == kAdvLimited ? a.psi[o] : sconst<S>(1.0);
        load_grad_row<D, S>(a.G, m.C, o, Go);
#pragma unroll
        for (int j = 0; j < D; ++j) cco[j] = m.cc[(int64_t)j * m.C + o];
      }
      load_face_geo_pts<D>(m, r.f, true, r.own ? ccc : cco, r.own ? cco : ccc, q);
      if (r.own)
        fl = scalar_flux_interior<D, ADV, S>(q, Qf, xc, xo, Gc, Go, pc, po, gc, go);
      else
        fl = scalar_flux_interior<D, ADV, S>(q, Qf, xo, xc, Go, Gc, po, pc, go, gc);
    } else {
      load_face_geo_pts<D>(m, r.f, false, ccc, ccc, q);
      const S bv = scalar_bval<D, S>(a.kind[r.fu], q, xc, Gc, a.val[r.fu], gc, a.tinf[r.fu]);
      fl = scalar_flux_boundary<D, ADV, S>(q, Qf, xc, Gc, pc, bv, gc);
    }
    acc = r.own ? acc + fl : acc - fl;
  }
  if (a.bdf) acc = acc + accumulation<S>(a, c, xc, m.vol[c]);
  R.set(c, acc);
}

// ---------------------------------------------------------------------------------------------
// passes G and R with a geometry tangent (the node-coordinate JVP): the same face physics on `dual`
// geometry.  Used only when a geometry tangent is requested; no LDS tiles.

template <int D, int GRAD, bool LIM>
__global__ void __launch_bounds__(256) k_sc_grad_g(MeshView m, GeoTan gt, ScArgs<dual> a, Out<dual> G,
                                                   Out<dual> psi_out) {
  const int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= m.C) return;
  const dual xc = a.phi[c];
  dual ccc[D];
  cell_centroid_g<D>(m, gt, c, ccc);
  dual acc[D], M[D][D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    acc[j] = dual{0.0, 0.0};
#pragma unroll
    for (int q2 = 0; q2 < D; ++q2) M[j][q2] = dual{0.0, 0.0};
  }
  dual mx = xc, mn = xc, zero[D];
#pragma unroll
  for (int j = 0; j < D; ++j) zero[j] = dual{0.0, 0.0};
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    if (o >= 0) {
      const dual xo = a.phi[o];
      if (LIM) {
        if (xo.v > mx.v) mx = xo;
        if (xo.v < mn.v) mn = xo;
      }
      dual cco[D];
      cell_centroid_g<D>(m, gt, o, cco);
      if (GRAD == 0) {
        FaceGeo<D, dual> q;
        load_face_geo_g<D>(m, gt, r.f, true, r.own ? ccc : cco, r.own ? cco : ccc, q);
        const dual w = 1.0 - q.g;
        const dual pf = w * (r.own ? xc : xo) + q.g * (r.own ? xo : xc);
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const dual t = pf * (q.n[j] * q.A);
          acc[j] = r.own ? acc[j] + t : acc[j] - t;
        }
      } else if (GRAD == 1) {
        dual dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = cco[j] - ccc[j];
        const dual wgt = 1.0 / vdotg<D, dual>(dx, dx);
#pragma unroll
        for (int p = 0; p < D; ++p)
#pragma unroll
          for (int q2 = 0; q2 < D; ++q2) M[p][q2] += wgt * dx[p] * dx[q2];
        const dual dphi = xo - xc;
#pragma unroll
        for (int j = 0; j < D; ++j) acc[j] += (wgt * dx[j]) * dphi;
      }
    } else if (GRAD != kGradNone) {
      FaceGeo<D, dual> q;
      load_face_geo_g<D>(m, gt, r.f, false, ccc, ccc, q);
      const dual pb = scalar_bval<D, dual, dual>(a.kind[r.fu], q, xc, zero, a.val[r.fu], a.gam[c], a.tinf[r.fu]);
      if (GRAD == 0) {
#pragma unroll
        for (int j = 0; j < D; ++j) acc[j] += pb * (q.n[j] * q.A);
      } else {
        const dual wgt = 1.0 / vdotg<D, dual>(q.dp, q.dp);
#pragma unroll
        for (int p = 0; p < D; ++p)
#pragma unroll
          for (int q2 = 0; q2 < D; ++q2) M[p][q2] += wgt * q.dp[p] * q.dp[q2];
        const dual dphi = pb - xc;
#pragma unroll
        for (int j = 0; j < D; ++j) acc[j] += (wgt * q.dp[j]) * dphi;
      }
    }
  }
  const dual V = dual{m.vol[c], gt.vol[c]};
  dual grad[D];
  if (GRAD == 0) {
    const dual iv = 1.0 / V;
#pragma unroll
    for (int j = 0; j < D; ++j) grad[j] = acc[j] * iv;
  } else if (GRAD == 1) {
    dual I[D][D];
    sym_inverse<D, dual>(M, I);
#pragma unroll
    for (int p = 0; p < D; ++p) {
      dual sum = I[p][0] * acc[0];
#pragma unroll
      for (int q2 = 1; q2 < D; ++q2) sum = sum + I[p][q2] * acc[q2];
      grad[p] = sum;
    }
  } else {
#pragma unroll
    for (int j = 0; j < D; ++j) grad[j] = dual{0.0, 0.0};
  }
#pragma unroll
  for (int j = 0; j < D; ++j) G.set((int64_t)j * m.C + c, grad[j]);
  if (LIM) {
    const dual e2 = V * (a.lim_k * a.lim_k * a.lim_k);
    dual lo = dual{0.0, 0.0};
    bool first = true;
    for (FaceWalk walk(m, c); walk.more();) {
      FaceRef r = walk.next(m, c);
      dual dm = grad[0] * (dual{m.fc[r.f], gt.fc[r.f]} - ccc[0]);
#pragma unroll
      for (int j = 1; j < D; ++j)
        dm = dm + grad[j] * (dual{m.fc[(int64_t)j * m.F + r.f], gt.fc[(int64_t)j * m.F + r.f]} - ccc[j]);
      const dual vf = venkat_face<dual, dual>(dm, mx - xc, mn - xc, e2);
      if (first || vf.v < lo.v) lo = vf;
      first = false;
    }
    psi_out.set(c, dmin(dmax(lo, 0.0), 1.0));
  }
}

template <int D, int ADV>
__global__ void __launch_bounds__(256) k_sc_res_g(MeshView m, GeoTan gt, ScArgs<dual> a, Out<dual> R) {
  const int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= m.C) return;
  const dual xc = a.phi[c], gc = a.gam[c];
  const dual pc = ADV == kAdvLimited ? a.psi[c] : dual{1.0, 0.0};
  dual Gc[D], ccc[D];
  load_grad_row<D, dual>(a.G, m.C, c, Gc);
  cell_centroid_g<D>(m, gt, c, ccc);
  dual acc = dual{0.0, 0.0};
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D, dual> q;
    const dual Qf = a.flux[r.fu];
    dual fl;
    if (o >= 0) {
      const dual xo = a.phi[o], go = a.gam[o];
      const dual po = ADV == kAdvLimited ? a.psi[o] : dual{1.0, 0.0};
      dual Go[D], cco[D];
      load_grad_row<D, dual>(a.G, m.C, o, Go);
      cell_centroid_g<D>(m, gt, o, cco);
      load_face_geo_g<D>(m, gt, r.f, true, r.own ? ccc : cco, r.own ? cco : ccc, q);
      if (r.own)
        fl = scalar_flux_interior<D, ADV, dual, dual>(q, Qf, xc, xo, Gc, Go, pc, po, gc, go);
      else
        fl = scalar_flux_interior<D, ADV, dual, dual>(q, Qf, xo, xc, Go, Gc, po, pc, go, gc);
    } else {
      load_face_geo_g<D>(m, gt, r.f, false, ccc, ccc, q);
      const dual bv = scalar_bval<D, dual, dual>(a.kind[r.fu], q, xc, Gc, a.val[r.fu], gc, a.tinf[r.fu]);
      fl = scalar_flux_boundary<D, ADV, dual, dual>(q, Qf, xc, Gc, pc, bv, gc);
    }
    acc = r.own ? acc + fl : acc - fl;
  }
  if (a.bdf) acc = acc + accumulation<dual, dual>(a, c, xc, dual{m.vol[c], gt.vol[c]});
  R.set(c, acc);
}

// ---------------------------------------------------------------------------------------------
// adjoint helpers with geometry cotangents

struct ScBars {
  double* phi_old;    // [C] (transient only)
  double* phi_older;  // [C] (BDF2 only)
  double* phi;   // [C]
  double* gam;   // [C]
  double* flux;  // [F]
  double* val;   // [F]
  double* tinf;  // [F]
  double* G;     // [D][C] scratch
  double* psi;   // [C]    scratch
  double* ext;   // [2][C] scratch: cotangents of the stencil extrema
  double* limx;  // [D][C] scratch: limiter's face-centroid cotangent (at its argmin face)
  int32_t* limf; // [C]    scratch: limiter argmin face (-1 none)
  // geometry cotangents (null when not requested)
  double *area, *nrm, *fc, *vol, *cc;
  // corrected Green-Gauss: the swept solve's geometry cotangents (null for the other schemes)
  double *cggf, *cggv;
  // with geometry cotangents: per internal face one 64-byte record (area, normal, centroid and
  // face-flux cotangents) that the residual adjoint fills and the gradient adjoint completes and
  // writes out
  double* frec;
};

// Face record layout: [0] area, [1, 1 + D) normal, [1 + D, 1 + 2D) centroid, [7] face flux.  A
// face's cotangents are stored as one contiguous record (four 16-byte stores) rather than one
// double in each of 2D + 2 arrays, so a wave's stores fill whole lines at once.
constexpr int kFaceRec = 8;

template <int D>
AFX_DEV void store_face_rec(double* __restrict__ frec, int32_t f, const FaceGeoBar<D>& qb, double flux_b) {
  double r[kFaceRec];
  r[0] = qb.A;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    r[1 + j] = qb.n[j];
    r[1 + D + j] = qb.xf[j];
  }
#pragma unroll
  for (int i = 1 + 2 * D; i < kFaceRec - 1; ++i) r[i] = 0.0;
  r[kFaceRec - 1] = flux_b;
  double2* q = reinterpret_cast<double2*>(frec + (int64_t)kFaceRec * f);
#pragma unroll
  for (int i = 0; i < kFaceRec / 2; ++i) q[i] = make_double2(r[2 * i], r[2 * i + 1]);
}

// The owner's last visit of a face (the gradient adjoint): the record plus this pass's part, out to
// the caller's arrays at the caller's face index u.
template <int D>
AFX_DEV void finish_face_rec(const ScBars& bar, int32_t F, const double* __restrict__ frec, int32_t f, int32_t u,
                             const FaceGeoBar<D>& qb) {
  const double2* q = reinterpret_cast<const double2*>(frec + (int64_t)kFaceRec * f);
  double r[kFaceRec];
#pragma unroll
  for (int i = 0; i < kFaceRec / 2; ++i) {
    const double2 v = q[i];
    r[2 * i] = v.x;
    r[2 * i + 1] = v.y;
  }
  bar.area[u] = r[0] + qb.A;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    bar.nrm[(int64_t)j * F + u] = r[1 + j] + qb.n[j];
    bar.fc[(int64_t)j * F + u] = r[1 + D + j] + qb.xf[j];
  }
  bar.flux[u] = r[kFaceRec - 1];
}

// ---------------------------------------------------------------------------------------------
// reverse pass R'.  Three waves per SIMD: the 3D limited build with geometry cotangents needs
// 174 registers unconstrained and fits 168 without spills (1.59 -> 1.55 ms on 6 x 96^3 tets).
template <int D, int ADV, bool GEO>
__global__ void __launch_bounds__(256, 3) k_sc_res_adj(MeshView m, ScArgs<double> a, const double* __restrict__ lam,
                                                    ScBars bar) {
  __shared__ ScTile<D, double> tile;
  const int32_t base = block_base(m);
  const int t = threadIdx.x;
  const int32_t c = base + t;
  const bool valid = c < m.C;
  double xc = 0.0, gc = 0.0, pc = 0.0, lc = 0.0;
  double Gc[D], ccc[D];
  if (valid) {
    xc = a.phi[c];
    gc = a.gam[c];
    pc = ADV == kAdvLimited ? a.psi[c] : 1.0;
    load_grad_row<D, double>(a.G, m.C, c, Gc);
#pragma unroll
    for (int j = 0; j < D; ++j) ccc[j] = m.cc[(int64_t)j * m.C + c];
    tile.put(t, xc, gc, pc, Gc, ccc);
    tile.lam[t] = lc = lam[c];
  }
  __syncthreads();
  if (!valid) return;
  ScalarBar<D> B;
  B.zero();
  double ccb[D];
#pragma unroll
  for (int j = 0; j < D; ++j) ccb[j] = 0.0;
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D> q;
    FaceGeoBar<D> qb;
    qb.zero();
    const double Qf = a.flux[r.fu];
    if (o >= 0) {
      double xo, go, po, lo;
      double Go[D], cco[D];
      const int ls = tile_slot(o, base);
      if (ls >= 0) {
        tile.get(ls, xo, go, po, Go, cco);
        lo = tile.lam[ls];
      } else {
        xo = a.phi[o];
        go = a.gam[o];
        po = ADV == kAdvLimited ? a.psi[o] : 1.0;
        load_grad_row<D, double>(a.G, m.C, o, Go);
#pragma unroll
        for (int j = 0; j < D; ++j) cco[j] = m.cc[(int64_t)j * m.C + o];
        lo = lam[o];
      }
      load_face_geo_pts<D>(m, r.f, true, r.own ? ccc : cco, r.own ? cco : ccc, q);
      const double lf = r.own ? lc - lo : lo - lc;
      double Qb;
      if (r.own)
        Qb = adj_sc_interior<D, ADV, GEO>(q, Qf, xc, xo, Gc, Go, pc, po, gc, go, lf, 0, B, qb);
      else
        Qb = adj_sc_interior<D, ADV, GEO>(q, Qf, xo, xc, Go, Gc, po, pc, go, gc, lf, 1, B, qb);
      if (GEO) adj_face_geo<D>(q, true, qb);
      if (r.own) {
        // (the boundary-value cotangents of interior faces are zero-filled before)
        if (GEO)
          store_face_rec<D>(bar.frec, r.f, qb, Qb);
        else
          bar.flux[r.fu] = Qb;
      }
#pragma unroll
      for (int j = 0; j < D; ++j) ccb[j] += (r.own ? qb.xP[j] : 0.0) + (r.own ? 0.0 : qb.xN[j]);
    } else {
      load_face_geo_pts<D>(m, r.f, false, ccc, ccc, q);
      double vb = 0.0, tb = 0.0;
      const double Qb = adj_sc_boundary<D, ADV, GEO>(a.kind[r.fu], q, Qf, xc, Gc, pc, gc, a.val[r.fu], a.tinf[r.fu],
                                                     lc, B, vb, tb, qb);
      if (GEO) adj_face_geo<D>(q, false, qb);
      bar.val[r.fu] = vb;
      bar.tinf[r.fu] = tb;
      if (GEO)
        store_face_rec<D>(bar.frec, r.f, qb, Qb);
      else
        bar.flux[r.fu] = Qb;
#pragma unroll
      for (int j = 0; j < D; ++j) ccb[j] += qb.xP[j];
    }
  }
  double volb = 0.0;
  if (a.bdf) {  // accumulation term
    const double V = m.vol[c], s_ = lc / a.dt;
    if (a.bdf == 1) {
      B.phi += s_ * V;
      bar.phi_old[c] = -s_ * V;
      volb = s_ * (xc - a.phi_old[c]);
    } else {
      B.phi += s_ * V * 1.5;
      bar.phi_old[c] = -s_ * V * 2.0;
      bar.phi_older[c] = s_ * V * 0.5;
      volb = s_ * ((1.5 * xc - 2.0 * a.phi_old[c]) + 0.5 * a.phi_older[c]);
    }
  }
  bar.phi[c] = B.phi;
  bar.gam[c] = B.gam;
  bar.psi[c] = B.psi;
#pragma unroll
  for (int j = 0; j < D; ++j) bar.G[(int64_t)j * m.C + c] = B.G[j];
  if (GEO) {
#pragma unroll
    for (int j = 0; j < D; ++j) bar.cc[(int64_t)j * m.C + c] = ccb[j];
    bar.vol[c] = volb;
  }
}

// Venkatakrishnan face value with a differentiable e2 (for the volume cotangent).
AFX_DEV dual venkat_face_de2(double dm0, double hmax, double hmin, double e2) {
  const double sgn = dm0 >= 0.0 ? 1.0 : -1.0;
  const double dm = sgn * (fabs(dm0) + 1e-12);
  const double head = dm > 0.0 ? hmax : hmin;
  const dual E{e2, 1.0};
  const double h2 = head * head, dm2 = dm * dm;
  const dual num = (h2 + E) * dm + 2.0 * dm2 * head;
  const dual den = (h2 + 2.0 * dm2 + head * dm + E) * dm;
  return num / den;
}

// reverse pass L': limiter, then the gradient cotangent is premultiplied by the reconstruction
// operator (GG: 1/V, LSQ: M^-1); GG also yields the volume cotangent -Gbar.G / V.
template <int D, int GRAD, bool LIM, bool GEO>
__global__ void __launch_bounds__(256) k_sc_lim_adj(MeshView m, ScArgs<double> a, const int32_t* __restrict__ ext,
                                                    const int32_t* __restrict__ kmin_face, ScBars bar) {
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  double Gb[D], Gc[D];
#pragma unroll
  for (int j = 0; j < D; ++j) Gb[j] = bar.G[(int64_t)j * m.C + c];
  load_grad_row<D, double>(a.G, m.C, c, Gc);
  double xb = 0.0, volb = 0.0, xcb[D];
#pragma unroll
  for (int j = 0; j < D; ++j) xcb[j] = 0.0;
  if (LIM) {
    const double V = m.vol[c];
    const double K3 = a.lim_k * a.lim_k * a.lim_k;
    const double e2 = V * K3;
    double xcen[D];
#pragma unroll
    for (int j = 0; j < D; ++j) xcen[j] = m.cc[(int64_t)j * m.C + c];
    const double pbar = bar.psi[c];
    const int32_t amx = ext[c], amn = ext[(int64_t)m.C + c];
    const double xv = a.phi[c], mxv = a.phi[amx], mnv = a.phi[amn];
    // the first face attaining the forward minimum, recorded by the forward limiter pass: only it
    // is evaluated again
    const int32_t f = kmin_face[c];
    double dxf[D];
#pragma unroll
    for (int j = 0; j < D; ++j) dxf[j] = m.fc[(int64_t)j * m.F + f] - xcen[j];
    double dmf = Gc[0] * dxf[0];
#pragma unroll
    for (int j = 1; j < D; ++j) dmf = dmf + Gc[j] * dxf[j];
    const double lo = venkat_face<double>(dmf, mxv - xv, mnv - xv, e2);
    const double dclip = (lo > 0.0 && lo < 1.0) ? 1.0 : ((lo == 0.0 || lo == 1.0) ? 0.5 : 0.0);
    const double vbar = pbar * dclip;
    double mxb = 0.0, mnb = 0.0;
    int32_t fmin = -1;
    double xfb[D];
#pragma unroll
    for (int j = 0; j < D; ++j) xfb[j] = 0.0;
    if (vbar != 0.0) {
      const dual r_dm = venkat_face<dual>(dual{dmf, 1.0}, dual{mxv - xv, 0.0}, dual{mnv - xv, 0.0}, e2);
      const dual r_hx = venkat_face<dual>(dual{dmf, 0.0}, dual{mxv - xv, 1.0}, dual{mnv - xv, 0.0}, e2);
      const dual r_hn = venkat_face<dual>(dual{dmf, 0.0}, dual{mxv - xv, 0.0}, dual{mnv - xv, 1.0}, e2);
      const double dmb = vbar * r_dm.d;
#pragma unroll
      for (int j = 0; j < D; ++j) Gb[j] += dmb * dxf[j];
      if (GEO) {
        fmin = f;
#pragma unroll
        for (int j = 0; j < D; ++j) {
          xfb[j] = dmb * Gc[j];
          xcb[j] -= dmb * Gc[j];
        }
        volb += vbar * venkat_face_de2(dmf, mxv - xv, mnv - xv, e2).d * K3;
      }
      mxb = vbar * r_hx.d;
      mnb = vbar * r_hn.d;
      xb -= mxb + mnb;
      if (amx == c) {
        xb += mxb;
        mxb = 0.0;
      }
      if (amn == c) {
        xb += mnb;
        mnb = 0.0;
      }
    }
    bar.ext[c] = mxb;
    bar.ext[(int64_t)m.C + c] = mnb;
    if (GEO) {
      bar.limf[c] = fmin;
#pragma unroll
      for (int j = 0; j < D; ++j) bar.limx[(int64_t)j * m.C + c] = xfb[j];
    }
  }
  bar.phi[c] += xb;
  if (GRAD == kGradCorrected) {  // the swept solve's reverse pass forms the cotangent of b
#pragma unroll
    for (int j = 0; j < D; ++j) bar.G[(int64_t)j * m.C + c] = Gb[j];
  } else if (GRAD == 0) {
    const double iv = 1.0 / m.vol[c];
    if (GEO) {
      double gg = 0.0;
#pragma unroll
      for (int j = 0; j < D; ++j) gg += Gb[j] * Gc[j];
      volb -= gg * iv;
    }
#pragma unroll
    for (int j = 0; j < D; ++j) bar.G[(int64_t)j * m.C + c] = Gb[j] * iv;
  } else if (GRAD == 1) {
    double M[D][D], I[D][D];
    lsq_matrix<D>(m, c, M);
    sym_inverse<D>(M, I);
#pragma unroll
    for (int p = 0; p < D; ++p) {
      double s = 0.0;
#pragma unroll
      for (int q2 = 0; q2 < D; ++q2) s += I[p][q2] * Gb[q2];
      bar.G[(int64_t)p * m.C + c] = s;
    }
  } else {
#pragma unroll
    for (int j = 0; j < D; ++j) bar.G[(int64_t)j * m.C + c] = 0.0;
  }
  if (GEO) {
    bar.vol[c] += volb;
#pragma unroll
    for (int j = 0; j < D; ++j) bar.cc[(int64_t)j * m.C + c] += xcb[j];
  }
}

// reverse pass G'.  CG: the gradient is corrected Green-Gauss (GRAD 0 here, the swept solve having
// formed the cotangent of b), whose own geometry cotangents are folded in with the face sums.
// What a cell of the workgroup's block offers its neighbours in the reverse gradient pass: the
// gradient cotangent, centroid and value, the forward gradient (least-squares geometry cotangents)
// and the limiter extrema with their cotangents.
template <int D, int GRAD, bool LIM, bool GEO>
struct ScGradAdjTile {
  static constexpr int NG = (GRAD == 1 && GEO) ? D : 1;
  static constexpr int NE = LIM ? 2 : 1;
  double B[D][kBlock];
  double cc[D][kBlock];
  double phi[kBlock];
  double G[NG][kBlock];
  double bex[NE][kBlock];
  int32_t ex[NE][kBlock];
};

template <int D, int GRAD, bool LIM, bool GEO, bool CG = false>
__global__ void __launch_bounds__(256) k_sc_grad_adj(MeshView m, ScArgs<double> a, const int32_t* __restrict__ ext,
                                                     ScBars bar) {
  __shared__ ScGradAdjTile<D, GRAD, LIM, GEO> tile;
  const int32_t base = block_base(m);
  const int t = threadIdx.x;
  const int32_t c = base + t;
  const bool valid = c < m.C;
  double Bc[D], ccc[D];  // this cell's gradient cotangent and centroid, for every face
  double Gc[D];          // the forward gradient (least-squares geometry cotangents)
  double xc = 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) Gc[j] = 0.0;
  if (valid) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      Bc[j] = bar.G[(int64_t)j * m.C + c];
      ccc[j] = m.cc[(int64_t)j * m.C + c];
      tile.B[j][t] = Bc[j];
      tile.cc[j][t] = ccc[j];
    }
    xc = a.phi[c];
    tile.phi[t] = xc;
    if (GRAD == 1 && GEO) {
      load_grad_row<D, double>(a.G, m.C, c, Gc);
#pragma unroll
      for (int j = 0; j < D; ++j) tile.G[j < ScGradAdjTile<D, GRAD, LIM, GEO>::NG ? j : 0][t] = Gc[j];
    }
    if (LIM) {
      tile.ex[0][t] = ext[c];
      tile.ex[LIM ? 1 : 0][t] = ext[(int64_t)m.C + c];
      tile.bex[0][t] = bar.ext[c];
      tile.bex[LIM ? 1 : 0][t] = bar.ext[(int64_t)m.C + c];
    }
  }
  __syncthreads();
  if (!valid) return;
  double xb = 0.0, gamb = 0.0, ccb[D];
#pragma unroll
  for (int j = 0; j < D; ++j) ccb[j] = 0.0;
  double zero[D];
#pragma unroll
  for (int j = 0; j < D; ++j) zero[j] = 0.0;
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeoBar<D> qb;
    qb.zero();
    bool have_geo = false;
    if (o >= 0) {
      // the other cell's record, from the tile when it is in this block: its gradient cotangent,
      // centroid and value, forward gradient, and limiter extrema (which cells attain them, and
      // their cotangents), all fetched before use, so no load waits on a comparison
      int32_t oamx = 0, oamn = 0;
      double obmx = 0.0, obmn = 0.0, po = 0.0;
      double Bo[D], cco[D], Go[D];
#pragma unroll
      for (int j = 0; j < D; ++j) Go[j] = 0.0;
      const int ls = tile_slot(o, base);
      if (ls >= 0) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
          Bo[j] = tile.B[j][ls];
          cco[j] = tile.cc[j][ls];
        }
        po = tile.phi[ls];
        if (GRAD == 1 && GEO) {
#pragma unroll
          for (int j = 0; j < D; ++j) Go[j] = tile.G[j < ScGradAdjTile<D, GRAD, LIM, GEO>::NG ? j : 0][ls];
        }
        if (LIM) {
          oamx = tile.ex[0][ls];
          oamn = tile.ex[LIM ? 1 : 0][ls];
          obmx = tile.bex[0][ls];
          obmn = tile.bex[LIM ? 1 : 0][ls];
        }
      } else {
#pragma unroll
        for (int j = 0; j < D; ++j) {
          Bo[j] = bar.G[(int64_t)j * m.C + o];
          cco[j] = m.cc[(int64_t)j * m.C + o];
        }
        po = a.phi[o];
        if (GRAD == 1 && GEO) load_grad_row<D, double>(a.G, m.C, o, Go);
        if (LIM) {
          oamx = ext[o];
          oamn = ext[(int64_t)m.C + o];
          obmx = bar.ext[o];
          obmn = bar.ext[(int64_t)m.C + o];
        }
      }
      if (GRAD == 0) {
        FaceGeo<D> q;
        double xPa[D], xNa[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
          xPa[j] = r.own ? ccc[j] : cco[j];
          xNa[j] = r.own ? cco[j] : ccc[j];
        }
        load_face_geo_pts<D>(m, r.f, true, xPa, xNa, q);
        double BP[D], BN[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
          BP[j] = r.own ? Bc[j] : Bo[j];
          BN[j] = r.own ? Bo[j] : Bc[j];
        }
        double pb_ = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) pb_ += (q.n[j] * q.A) * (BP[j] - BN[j]);
        xb += (r.own ? 1.0 - q.g : q.g) * pb_;
        if (GEO) {
          const double phP = r.own ? xc : po, phN = r.own ? po : xc;
          const double phf = (1.0 - q.g) * phP + q.g * phN;
          double dn = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) {
            qb.n[j] += phf * q.A * (BP[j] - BN[j]);
            dn += q.n[j] * (BP[j] - BN[j]);
          }
          qb.A += phf * dn;
          qb.g += pb_ * (phN - phP);
          adj_face_geo<D>(q, true, qb);
          if constexpr (CG)  // the sweeps' own geometry dependence
            cgg_face_geo_bar<D>(CggGeoBars{bar.cggf, bar.cggv}, m.F, r.f, q, qb.A, qb.n, qb.xf, qb.xP, qb.xN);
          have_geo = true;
        }
      } else if (GRAD == 1) {
        double dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = cco[j] - ccc[j];
        const double wgt = 1.0 / vdot<D>(dx, dx);
        double own_dot = 0.0, oth_dot = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) {
          own_dot += dx[j] * Bc[j];
          oth_dot += dx[j] * Bo[j];
        }
        xb += -wgt * own_dot - wgt * oth_dot;
        if (GEO) {
          // G = M^-1 b with M = sum w d d^T, b = sum w d dphi, w = 1 / d.d; with y = M^-1 Gbar
          // (bar.G here) and the fit residual r = dphi - G.d of the face, the cotangent of its d is
          // w (r y - (y.d) G) - 2 w^2 (y.d) r d.  c's own fit runs along d = x_o - x_c, o's along -d;
          // both ends are cell centroids, of which this thread owns x_c.
          double gd_c = 0.0, gd_o = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) {
            gd_c += Gc[j] * dx[j];
            gd_o -= Go[j] * dx[j];
          }
          const double rc = (po - xc) - gd_c, ro = (xc - po) - gd_o;
          const double yd_c = own_dot, yd_o = -oth_dot;
#pragma unroll
          for (int j = 0; j < D; ++j) {
            const double db_c = wgt * (rc * Bc[j] - yd_c * Gc[j]) - 2.0 * wgt * wgt * yd_c * rc * dx[j];
            const double db_o = wgt * (ro * Bo[j] - yd_o * Go[j]) + 2.0 * wgt * wgt * yd_o * ro * dx[j];
            ccb[j] += db_o - db_c;
          }
        }
      }
      if (LIM) {
        if (oamx == c) xb += obmx;
        if (oamn == c) xb += obmn;
      }
    } else if (GRAD != kGradNone) {
      FaceGeo<D> q;
      load_face_geo_pts<D>(m, r.f, false, ccc, ccc, q);
      double phib;
      if (GRAD == 0) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) s += (q.n[j] * q.A) * Bc[j];
        phib = s;
        if (GEO) {
          const double pbv = scalar_bval<D, double>(a.kind[r.fu], q, xc, zero, a.val[r.fu], a.gam[c], a.tinf[r.fu]);
#pragma unroll
          for (int j = 0; j < D; ++j) qb.n[j] += pbv * q.A * Bc[j];
          double dn = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) dn += q.n[j] * Bc[j];
          qb.A += pbv * dn;
        }
      } else {
        const double wgt = 1.0 / vdot<D>(q.dp, q.dp);
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) s += q.dp[j] * Bc[j];
        phib = wgt * s;
        xb -= wgt * s;
        if (GEO) {  // the fit's direction d = x_f - x_c (folded into x_f, x_c below)
          const double pbv = scalar_bval<D, double>(a.kind[r.fu], q, xc, zero, a.val[r.fu], a.gam[c], a.tinf[r.fu]);
          double gd = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) gd += Gc[j] * q.dp[j];
          const double rr = (pbv - xc) - gd;
#pragma unroll
          for (int j = 0; j < D; ++j) qb.dp[j] += wgt * (rr * Bc[j] - s * Gc[j]) - 2.0 * wgt * wgt * s * rr * q.dp[j];
        }
      }
      ScalarBar<D> Bl;
      Bl.zero();
      double vb = 0.0, tb = 0.0;
      adj_sc_bval<D, GEO>(a.kind[r.fu], q, xc, zero, a.val[r.fu], a.gam[c], a.tinf[r.fu], phib, Bl, vb, tb, qb);
      xb += Bl.phi;
      gamb += Bl.gam;
      bar.val[r.fu] += vb;
      bar.tinf[r.fu] += tb;
      if (GEO) {
        adj_face_geo<D>(q, false, qb);
        have_geo = true;
      }
    }
    if (GEO) {
      // the limiters of this face's cells that took their minimum here
      if (r.own) {
        if (LIM && bar.limf[r.P] == r.f) {
#pragma unroll
          for (int j = 0; j < D; ++j) qb.xf[j] += bar.limx[(int64_t)j * m.C + r.P];
        }
        if (LIM && r.N >= 0 && bar.limf[r.N] == r.f) {
#pragma unroll
          for (int j = 0; j < D; ++j) qb.xf[j] += bar.limx[(int64_t)j * m.C + r.N];
        }
        if (GEO) finish_face_rec<D>(bar, m.F, bar.frec, r.f, r.fu, qb);
      }
      if (have_geo) {
#pragma unroll
        for (int j = 0; j < D; ++j) ccb[j] += (r.own ? qb.xP[j] : 0.0) + (r.own ? 0.0 : qb.xN[j]);
      }
    }
  }
  bar.phi[c] += xb;
  bar.gam[c] += gamb;
  if (GEO) {
#pragma unroll
    for (int j = 0; j < D; ++j) bar.cc[(int64_t)j * m.C + c] += ccb[j];
    if constexpr (CG) bar.vol[c] += bar.cggv[c];
  }
}

// ---------------------------------------------------------------------------------------------
// gradient of a scalar with Dirichlet / zero-gradient leading-order boundary values (used for the
// SST wall-distance gradient)
template <class S>
AFX_DEV S lead_bval(int8_t kind, double v, S phiP) {
  if (kind == kBcDirichlet) return sconst<S>(v);
  if (kind == kBcNone) return sconst<S>(0.0);
  return phiP;
}

// With GEO the geometry and phi are (value, tangent) pairs (the node-coordinate JVP of the
// gradient); the boundary values stay constant.
template <int D, int GRAD, bool GEO = false>
__global__ void __launch_bounds__(256) k_scalar_grad(MeshView m, GeoTan gt, const int8_t* __restrict__ kind,
                                                     const double* __restrict__ bval,
                                                     In<typename GeoRead<D, GEO>::T> phi,
                                                     Out<typename GeoRead<D, GEO>::T> G) {
  using GR = GeoRead<D, GEO>;
  using T = typename GR::T;
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  const T xc = phi[c];
  T acc[D];
#pragma unroll
  for (int j = 0; j < D; ++j) acc[j] = sconst<T>(0.0);
  T M[D][D];
#pragma unroll
  for (int p = 0; p < D; ++p)
#pragma unroll
    for (int q = 0; q < D; ++q) M[p][q] = sconst<T>(0.0);
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    if (o >= 0) {
      const T xo = phi[o];
      if (GRAD == 0) {
        FaceGeo<D, T> q;
        GR::face(m, gt, r.f, r.P, r.N, q);
        const T w = 1.0 - q.g;
        const T pf = w * (r.own ? xc : xo) + q.g * (r.own ? xo : xc);
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const T t = pf * (q.n[j] * q.A);
          acc[j] = r.own ? acc[j] + t : acc[j] - t;
        }
      } else {
        T dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = GR::cc(m, gt, j, o) - GR::cc(m, gt, j, c);
        const T wgt = 1.0 / vdotg<D, T>(dx, dx);
#pragma unroll
        for (int p = 0; p < D; ++p)
#pragma unroll
          for (int q2 = 0; q2 < D; ++q2) M[p][q2] += wgt * dx[p] * dx[q2];
        const T dphi = xo - xc;
#pragma unroll
        for (int j = 0; j < D; ++j) acc[j] += (wgt * dx[j]) * dphi;
      }
    } else {
      const T pb = lead_bval<T>(kind[r.fu], bval[r.fu], xc);
      if (GRAD == 0) {
        const T A = GR::area(m, gt, r.f);
#pragma unroll
        for (int j = 0; j < D; ++j) acc[j] += pb * (GR::nrm(m, gt, j, r.f) * A);
      } else {
        T dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = GR::fc(m, gt, j, r.f) - GR::cc(m, gt, j, c);
        const T wgt = 1.0 / vdotg<D, T>(dx, dx);
#pragma unroll
        for (int p = 0; p < D; ++p)
#pragma unroll
          for (int q2 = 0; q2 < D; ++q2) M[p][q2] += wgt * dx[p] * dx[q2];
        const T dphi = pb - xc;
#pragma unroll
        for (int j = 0; j < D; ++j) acc[j] += (wgt * dx[j]) * dphi;
      }
    }
  }
  if (GRAD == 0) {
    const T iv = 1.0 / GR::vol(m, gt, c);
#pragma unroll
    for (int j = 0; j < D; ++j) G.set((int64_t)j * m.C + c, acc[j] * iv);
  } else {
    T I[D][D];
    sym_inverse<D, T>(M, I);
#pragma unroll
    for (int p = 0; p < D; ++p) {
      T s = I[p][0] * acc[0];
#pragma unroll
      for (int q2 = 1; q2 < D; ++q2) s = s + I[p][q2] * acc[q2];
      G.set((int64_t)p * m.C + c, s);
    }
  }
}

// Reverse of k_scalar_grad.  Green-Gauss (GRAD 0): `bbar` is the cotangent of the face-value sum
// b_c = sum_f +-phi_f n_f A_f (compact: gradient cotangent / V, applied here when COMPACT; corrected:
// the sweeps' reverse already produced it); an interior face passes (1 - g) resp. g of
// n A . (bbar_P - bbar_N) to its owner resp. neighbour.  Least squares (GRAD 1): `bbar` holds
// y = M^-T Gbar per cell (k_lsq_solve_t) and a face adds w dx . y of each side with the sign of its
// difference.  A Dirichlet boundary face's value receives its share in value_bar (caller order);
// zero-gradient-type closures read the owner value, whose difference to the owner has zero
// derivative in least squares and which feeds the owner in Green-Gauss.
template <int D, int GRAD, bool COMPACT>
__global__ void __launch_bounds__(256) k_scalar_grad_vjp(MeshView m, const int8_t* __restrict__ kind,
                                                         const double* __restrict__ bbar,
                                                         double* __restrict__ phi_bar, double* __restrict__ value_bar) {
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  double yc[D];
  const double ivc = COMPACT ? 1.0 / m.vol[c] : 1.0;
#pragma unroll
  for (int j = 0; j < D; ++j) yc[j] = bbar[(int64_t)j * m.C + c] * ivc;
  double acc = 0.0;
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    if (o >= 0) {
      const double ivo = COMPACT ? 1.0 / m.vol[o] : 1.0;
      double yo[D];
#pragma unroll
      for (int j = 0; j < D; ++j) yo[j] = bbar[(int64_t)j * m.C + o] * ivo;
      if (GRAD == 0) {
        FaceGeo<D> q;
        load_face_geo<D>(m, r.f, r.P, r.N, q);
        double sz = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) sz += (q.n[j] * q.A) * (r.own ? yc[j] - yo[j] : yo[j] - yc[j]);
        acc += (r.own ? 1.0 - q.g : q.g) * sz;
      } else {
        double dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = m.cc[(int64_t)j * m.C + o] - m.cc[(int64_t)j * m.C + c];
        const double wgt = 1.0 / vdot<D>(dx, dx);
        double s2 = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) s2 += dx[j] * (yc[j] + yo[j]);
        acc -= wgt * s2;
      }
    } else {
      const int8_t kd = kind[r.fu];
      double t = 0.0;
      if (GRAD == 0) {
        const double A = m.area[r.f];
#pragma unroll
        for (int j = 0; j < D; ++j) t += (m.nrm[(int64_t)j * m.F + r.f] * A) * yc[j];
        if (kd != kBcDirichlet && kd != kBcNone) acc += t;
      } else {
        double dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = m.fc[(int64_t)j * m.F + r.f] - m.cc[(int64_t)j * m.C + c];
        const double wgt = 1.0 / vdot<D>(dx, dx);
        t = wgt * vdot<D>(dx, yc);
        if (kd == kBcDirichlet || kd == kBcNone) acc -= t;  // the face value is fixed (or 0), phi_P enters as -phi_P
      }
      if (value_bar && kd == kBcDirichlet) value_bar[r.fu] = t;
    }
  }
  phi_bar[c] = acc;
}

// y_c = M_c^-T gbar_c, with M the least-squares normal matrix and its inverse as k_scalar_grad forms it.
// Geometry cotangents of k_scalar_grad (node-coordinate VJP of a standalone gradient, e.g. the SST
// wall-distance gradient): `yb` holds y = Gbar / V (Green-Gauss, divided here) or M^-1 Gbar (least
// squares, from k_lsq_solve_t), G the forward gradient.  Writes (=) the owned faces' area, normal
// and centroid cotangents (caller order) and each cell's volume and centroid cotangents.
// Corrected Green-Gauss (GRAD 3): `yb` is the cotangent of the Green-Gauss sum b from the swept
// solve's reverse pass, whose own geometry cotangents come in `cb`.
template <int D, int GRAD>
__global__ void __launch_bounds__(256) k_scalar_grad_geo_vjp(MeshView m, const int8_t* __restrict__ kind,
                                                             const double* __restrict__ bval,
                                                             const double* __restrict__ phi,
                                                             const double* __restrict__ G,
                                                             const double* __restrict__ yb, afx_geometry_cotangents out,
                                                             CggGeoBars cb) {
  constexpr bool GG = GRAD != 1;  // Green-Gauss face sums (compact, or the corrected scheme's b)
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  const double xc = phi[c];
  const double ivc = GRAD == 0 ? 1.0 / m.vol[c] : 1.0;
  double yc[D], Gc[D], ccb[D], Vb = GRAD == kGradCorrected ? cb.vol[c] : 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    yc[j] = yb[(int64_t)j * m.C + c] * ivc;
    Gc[j] = G[(int64_t)j * m.C + c];
    ccb[j] = 0.0;
    if (GRAD == 0) Vb -= yc[j] * Gc[j];  // G = acc / V
  }
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    const int32_t o = r.own ? r.N : r.P;
    FaceGeo<D> q;
    load_face_geo<D>(m, r.f, r.P, r.N, q);
    double Ab = 0.0, nb[D], xfb[D];
#pragma unroll
    for (int j = 0; j < D; ++j) nb[j] = xfb[j] = 0.0;
    if (o >= 0) {
      const double xo = phi[o];
      if (GG) {
        const double ivo = GRAD == 0 ? 1.0 / m.vol[o] : 1.0;
        const double phP = r.own ? xc : xo, phN = r.own ? xo : xc;
        const double phf = (1.0 - q.g) * phP + q.g * phN;
        double zn = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const double yo = yb[(int64_t)j * m.C + o] * ivo;
          const double z = r.own ? yc[j] - yo : yo - yc[j];
          nb[j] += z * phf * q.A;
          zn += z * q.n[j];
        }
        Ab += zn * phf;
        FaceGeoBarF<D> qf;
        qf.zero();
        qf.g = zn * q.A * (phN - phP);
        fold_face_geo_bar<D>(q, true, qf);
#pragma unroll
        for (int j = 0; j < D; ++j) {
          xfb[j] += qf.xf[j];
          ccb[j] += r.own ? qf.xP[j] : qf.xN[j];
        }
        if (GRAD == kGradCorrected) {
          double xPb[D], xNb[D];
#pragma unroll
          for (int j = 0; j < D; ++j) xPb[j] = xNb[j] = 0.0;
          cgg_face_geo_bar<D>(cb, m.F, r.f, q, Ab, nb, xfb, xPb, xNb);
#pragma unroll
          for (int j = 0; j < D; ++j) ccb[j] += r.own ? xPb[j] : xNb[j];
        }
      } else {
        double d[D], dphi[1] = {xo - xc}, db[D], Gv[1][D], yv[1][D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
          d[j] = m.cc[(int64_t)j * m.C + o] - m.cc[(int64_t)j * m.C + c];
          Gv[0][j] = Gc[j];
          yv[0][j] = yc[j];
        }
        lsq_dir_bar<D, 1>(d, dphi, Gv, yv, db);  // own fit: d = x_o - x_c
#pragma unroll
        for (int j = 0; j < D; ++j) {
          ccb[j] -= db[j];
          d[j] = -d[j];
          Gv[0][j] = G[(int64_t)j * m.C + o];
          yv[0][j] = yb[(int64_t)j * m.C + o];
        }
        dphi[0] = -dphi[0];
        lsq_dir_bar<D, 1>(d, dphi, Gv, yv, db);  // the other cell's fit: d = x_c - x_o
#pragma unroll
        for (int j = 0; j < D; ++j) ccb[j] += db[j];
      }
    } else {
      const double pb = lead_bval<double>(kind[r.fu], bval[r.fu], xc);
      if (GG) {
        double zn = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) {
          nb[j] += yc[j] * pb * q.A;
          zn += yc[j] * q.n[j];
        }
        Ab += zn * pb;
      } else {
        double d[D], dphi[1] = {pb - xc}, db[D], Gv[1][D], yv[1][D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
          d[j] = q.xf[j] - m.cc[(int64_t)j * m.C + c];
          Gv[0][j] = Gc[j];
          yv[0][j] = yc[j];
        }
        lsq_dir_bar<D, 1>(d, dphi, Gv, yv, db);
#pragma unroll
        for (int j = 0; j < D; ++j) {
          xfb[j] += db[j];
          ccb[j] -= db[j];
        }
      }
    }
    if (r.own) {
      out.area[r.fu] = Ab;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        out.normal[(int64_t)j * m.F + r.fu] = nb[j];
        out.face_centroid[(int64_t)j * m.F + r.fu] = xfb[j];
      }
    }
  }
  out.volume[c] = Vb;
#pragma unroll
  for (int j = 0; j < D; ++j) out.cell_centroid[(int64_t)j * m.C + c] = ccb[j];
}

template <int D>
__global__ void __launch_bounds__(256) k_lsq_solve_t(MeshView m, const double* __restrict__ gbar,
                                                     double* __restrict__ y) {
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  double M[D][D], I[D][D];
  lsq_matrix<D>(m, c, M);
  sym_inverse<D>(M, I);
#pragma unroll
  for (int p = 0; p < D; ++p) {
    double s = I[0][p] * gbar[c];
#pragma unroll
    for (int q = 1; q < D; ++q) s = s + I[q][p] * gbar[(int64_t)q * m.C + c];
    y[(int64_t)p * m.C + c] = s;
  }
}

// ---------------------------------------------------------------------------------------------
// host side

// Workspace in doubles: mode 0 = G [D][C], psi [C], ext (2 int32 per cell); mode 1 adds tangents of
// G and psi; mode 2 adds the adjoint scratch (G, psi, ext, limiter face cotangent, argmin face, and
// the forward's minimum face).
int64_t sc_work_doubles(int D, int64_t C, int mode, int64_t F = 0) {
  int64_t n = (int64_t)(D + 1 + 1) * C + (int64_t)2 * D * C;  // ... + corrected Green-Gauss sweep scratch
  if (mode == 1) n += (int64_t)(D + 1) * C;
  if (mode == 2) n += (int64_t)(D + 1 + 2 + D + 1) * C + kFaceRec * F + kFaceRec;  // + face records (aligned)
  return n;
}

struct ScWork {
  double *G, *psi, *Gt, *psit;
  int32_t* ext;
  double *Gb, *psib, *extb, *limx;
  int32_t *limf, *kmin;  // kmin: the limiter's minimum face (internal face index)
  double* frec;          // [F][kFaceRec] face records, 64-byte aligned (mode 2)
  double* cgg;
};

ScWork sc_carve(void* base, int D, int64_t C, int mode, int64_t F = 0) {
  double* p = (double*)base;
  ScWork w{};
  w.G = p;
  p += (int64_t)D * C;
  w.psi = p;
  p += C;
  w.ext = (int32_t*)p;
  p += C;
  if (mode == 1) {
    w.Gt = p;
    p += (int64_t)D * C;
    w.psit = p;
    p += C;
  }
  if (mode == 2) {
    w.Gb = p;
    p += (int64_t)D * C;
    w.psib = p;
    p += C;
    w.extb = p;
    p += 2 * C;
    w.limx = p;
    p += (int64_t)D * C;
    w.limf = (int32_t*)p;
    w.kmin = w.limf + C;
    p += C;
    w.frec = (double*)(((uintptr_t)p + 63) & ~(uintptr_t)63);
    p += kFaceRec * F + kFaceRec;
  }
  w.cgg = p;
  return w;
}

template <class S>
ScArgs<S> sc_args(const afx_scalar_config* cfg, const afx_scalar_inputs* in, const afx_scalar_inputs* tan) {
  ScArgs<S> a{};
  a.kind = cfg->bc_kind;
  a.phi.v = in->phi;
  a.gam.v = in->gamma;
  a.flux.v = in->flux;
  a.val.v = in->value;
  a.tinf.v = in->t_inf;
  if (tan) {
    a.phi.t = tan->phi;
    a.gam.t = tan->gamma;
    a.flux.t = tan->flux;
    a.val.t = tan->value;
    a.tinf.t = tan->t_inf;
  }
  a.phi_old.v = in->phi_old;
  a.phi_older.v = in->phi_older;
  if (tan) {
    a.phi_old.t = tan->phi_old;
    a.phi_older.t = tan->phi_older;
  }
  a.lim_k = cfg->limiter_k;
  a.bdf = cfg->transient;
  a.dt = cfg->dt;
  return a;
}

template <class Fn>
int sc_dispatch(int D, int adv, int grad, Fn&& fn) {
#define AFX_S(DD, AA, GG)                                                                               \
  if (D == DD && adv == AA && grad == GG)                                                               \
    return fn(std::integral_constant<int, DD>{}, std::integral_constant<int, AA>{},                      \
              std::integral_constant<int, GG>{});
  AFX_S(2, 0, 0) AFX_S(2, 1, 0) AFX_S(2, 2, 0) AFX_S(2, 0, 1) AFX_S(2, 1, 1) AFX_S(2, 2, 1)
  AFX_S(2, 0, 2) AFX_S(2, 1, 2) AFX_S(2, 2, 2)
  AFX_S(3, 0, 0) AFX_S(3, 1, 0) AFX_S(3, 2, 0) AFX_S(3, 0, 1) AFX_S(3, 1, 1) AFX_S(3, 2, 1)
  AFX_S(3, 0, 2) AFX_S(3, 1, 2) AFX_S(3, 2, 2)
  AFX_S(2, 0, 3) AFX_S(2, 1, 3) AFX_S(2, 2, 3) AFX_S(3, 0, 3) AFX_S(3, 1, 3) AFX_S(3, 2, 3)
#undef AFX_S
  afx_set_error("scalar: unsupported (dim, advection, gradient) combination");
  return 2;
}

template <int D, class S, int ADV, int GRAD>
int sc_forward(const afx_mesh* mh, const afx_geometry* g, ScArgs<S> a, Out<S> R, const ScWork& w, double* Gt,
               double* psit, int32_t* ext, int sweeps, double relax, hipStream_t s, double* iters = nullptr) {
  MeshView m;
  if (int rv_ = afx_rview(mh, g, s, &m)) return rv_;
  constexpr bool LIM = ADV == kAdvLimited;
  constexpr int GK = GRAD == kGradCorrected ? 0 : GRAD;
  AFX_LAUNCH(kKScalarGrad, (k_sc_grad<D, S, GK, LIM && GRAD != kGradCorrected>), grid_for(m.C), dim3(kBlock), 0, s,
             m, a, Out<S>{w.G, Gt}, Out<S>{w.psi, psit}, ext, ext ? w.kmin : nullptr);
  AFX_CHECK_LAUNCH();
  a.G.v = w.G;
  a.G.t = Gt;
  if (GRAD == kGradCorrected) {
    int rc = iters ? cgg_sweeps_store(m, D, 1, w.G, iters, sweeps, relax, s)
                   : cgg_sweeps(m, D, 1, w.G, w.cgg, sweeps, relax, s);
    if (rc) return rc;
    if (Gt) {
      rc = cgg_sweeps(m, D, 1, Gt, w.cgg, sweeps, relax, s);
      if (rc) return rc;
    }
    if (LIM) {
      LimiterSlots sl{};
      sl.nl = 1;
      sl.ext_stride = 1;
      sl.kmin = ext ? w.kmin : nullptr;
      rc = venkat_pass(m, D, sl, a.phi, a.G, a.lim_k, Out<S>{w.psi, psit}, ext, s);
      if (rc) return rc;
    }
  }
  if (R.v == nullptr) return 0;
  a.psi.v = w.psi;
  a.psi.t = psit;
  AFX_LAUNCH(kKScalarRes, (k_sc_res<D, S, ADV>), grid_for(m.C), dim3(kBlock), 0, s, m, a, R);
  AFX_CHECK_LAUNCH();
  return 0;
}

int sc_check(const afx_mesh* m, const afx_geometry* g, const afx_scalar_config* cfg, const afx_scalar_inputs* in,
             void* work, int64_t work_size, int mode) {
  AFX_REQUIRE(m && g && cfg && in, "scalar: null argument");
  AFX_REQUIRE(in->phi && in->gamma && in->flux && in->value && in->t_inf, "scalar: a required input pointer is null");
  AFX_REQUIRE(cfg->bc_kind, "scalar: bc_kind is null");
  AFX_REQUIRE(cfg->advection >= 0 && cfg->advection <= 2, "scalar: unknown advection scheme");
  AFX_REQUIRE(cfg->gradient >= 0 && cfg->gradient <= 3, "scalar: unknown gradient scheme");
  AFX_REQUIRE(cfg->gradient != 3 || cfg->gradient_relaxation > 0.0, "scalar: gradient relaxation must be positive");
  AFX_REQUIRE(m->C > 0, "scalar: empty mesh");
  AFX_REQUIRE(cfg->transient >= 0 && cfg->transient <= 2, "scalar: transient must be 0 (steady), 1 or 2 (BDF order)");
  AFX_REQUIRE(cfg->transient == 0 || (cfg->dt > 0.0 && in->phi_old), "scalar: the transient term needs dt > 0 and phi_old");
  AFX_REQUIRE(cfg->transient != 2 || in->phi_older, "scalar: BDF2 needs phi_older");
  AFX_REQUIRE(mode < 0 || (work && work_size >= sc_work_doubles(m->dim, m->C, mode, m->F) * 8),
              "scalar: workspace too small");
  return 0;
}

}  // namespace

extern "C" int afx_scalar_gradient(const afx_mesh* mh, const afx_geometry* g, int gradient, int sweeps,
                                   double relaxation, const int8_t* kind, const double* value, const double* phi,
                                   double* grad, hipStream_t s) {
  AFX_REQUIRE(mh && g && kind && value && phi && grad, "afx_scalar_gradient: null argument");
  AFX_REQUIRE(gradient == 0 || gradient == 1 || gradient == 3, "afx_scalar_gradient: unknown gradient scheme");
  AFX_REQUIRE(gradient != 3 || relaxation > 0.0, "afx_scalar_gradient: relaxation must be positive");
  MeshView m;
  if (int rv_ = afx_rview(mh, g, s, &m)) return rv_;
  if (m.C == 0) return 0;
  const int gk = gradient == 3 ? 0 : gradient;
#define AFX_SG(DD, GG)                                                                                  \
  if (mh->dim == DD && gk == GG) {                                                                      \
    AFX_LAUNCH(kKScalarGrad, (k_scalar_grad<DD, GG>), grid_for(m.C), dim3(kBlock), 0, s, m, GeoTan{}, kind, \
               value, In<double>{phi, nullptr}, Out<double>{grad, nullptr});                             \
    AFX_CHECK_LAUNCH();                                                                                 \
    if (gradient != 3) return 0;                                                                        \
    double* scratch = afx_mesh_scratch(const_cast<afx_mesh*>(mh), (int64_t)2 * DD * m.C);              \
    AFX_REQUIRE(scratch, "afx_scalar_gradient: could not allocate the sweep scratch");                \
    return cgg_sweeps(m, DD, 1, grad, scratch, sweeps, relaxation, s);                                  \
  }
  AFX_SG(2, 0) AFX_SG(2, 1) AFX_SG(3, 0) AFX_SG(3, 1)
#undef AFX_SG
  afx_set_error("afx_scalar_gradient: unsupported dimension");
  return 2;
}

extern "C" int afx_scalar_gradient_jvp_geometry(const afx_mesh* mh, const afx_geometry* g,
                                                const afx_geometry* geom_tangent, int gradient, int sweeps,
                                                double relaxation, const int8_t* kind, const double* value,
                                                const double* phi, const double* phi_tangent, double* grad,
                                                double* grad_tangent, hipStream_t s) {
  AFX_REQUIRE(mh && g && geom_tangent && kind && value && phi && phi_tangent && grad && grad_tangent,
              "afx_scalar_gradient_jvp_geometry: null argument");
  AFX_REQUIRE(gradient == 0 || gradient == 1 || gradient == 3,
              "afx_scalar_gradient_jvp_geometry: unknown gradient scheme");
  AFX_REQUIRE(gradient != 3 || relaxation > 0.0, "afx_scalar_gradient_jvp_geometry: relaxation must be positive");
  AFX_REQUIRE(geom_tangent->area && geom_tangent->normal && geom_tangent->face_centroid && geom_tangent->volume &&
                  geom_tangent->cell_centroid,
              "afx_scalar_gradient_jvp_geometry: the geometry tangent needs all five arrays");
  AFX_REQUIRE(mh->fu == nullptr || geom_tangent->face_pack,
              "afx_scalar_gradient_jvp_geometry: this mesh's faces are reordered: geom_tangent->face_pack must hold "
              "the face tangents in the internal order");
  MeshView m;
  if (int rv_ = afx_rview(mh, g, s, &m)) return rv_;
  if (m.C == 0) return 0;
  GeoTan gt;
  const double* pack = geom_tangent->face_pack;
  gt.area = mh->fu ? pack : geom_tangent->area;
  gt.nrm = mh->fu ? pack + mh->F : geom_tangent->normal;
  gt.fc = mh->fu ? pack + (int64_t)(1 + mh->dim) * mh->F : geom_tangent->face_centroid;
  gt.vol = geom_tangent->volume;
  gt.cc = geom_tangent->cell_centroid;
  const In<dual> ph{phi, phi_tangent};
  const Out<dual> og{grad, grad_tangent};
  const int gk = gradient == 3 ? 0 : gradient;
#define AFX_SGJ(DD, GG)                                                                                   \
  if (mh->dim == DD && gk == GG) {                                                                        \
    AFX_LAUNCH(kKScalarGrad, (k_scalar_grad<DD, GG, true>), grid_for(m.C), dim3(kBlock), 0, s, m, gt, kind, \
               value, ph, og);                                                                            \
    AFX_CHECK_LAUNCH();                                                                                   \
    if (gradient != 3) return 0;                                                                          \
    double* scr = afx_mesh_scratch(const_cast<afx_mesh*>(mh), (int64_t)4 * DD * m.C);                    \
    AFX_REQUIRE(scr, "afx_scalar_gradient_jvp_geometry: could not allocate the sweep scratch");          \
    return cgg_sweeps_tangent(m, gt, DD, 1, grad, grad_tangent, scr, sweeps, relaxation, s);             \
  }
  AFX_SGJ(2, 0) AFX_SGJ(2, 1) AFX_SGJ(3, 0) AFX_SGJ(3, 1)
#undef AFX_SGJ
  afx_set_error("afx_scalar_gradient_jvp_geometry: unsupported dimension");
  return 2;
}

extern "C" int afx_scalar_gradient_vjp_geometry(const afx_mesh* mh, const afx_geometry* g, int gradient, int sweeps,
                                                double relaxation, const int8_t* kind, const double* value,
                                                const double* phi, const double* grad_bar,
                                                afx_geometry_cotangents* out, hipStream_t s) {
  AFX_REQUIRE(mh && g && kind && value && phi && grad_bar && out, "afx_scalar_gradient_vjp_geometry: null argument");
  AFX_REQUIRE(out->area && out->normal && out->face_centroid && out->volume && out->cell_centroid,
              "afx_scalar_gradient_vjp_geometry: all five geometry cotangent buffers are needed");
  AFX_REQUIRE(gradient == 0 || gradient == 1 || gradient == 3,
              "afx_scalar_gradient_vjp_geometry: unknown gradient scheme");
  AFX_REQUIRE(gradient != 3 || relaxation > 0.0, "afx_scalar_gradient_vjp_geometry: relaxation must be positive");
  MeshView m;
  if (int rv_ = afx_rview(mh, g, s, &m)) return rv_;
  if (m.C == 0) return 0;
  const int D = mh->dim;
  const int64_t n = (int64_t)D * m.C;
  // G, y; corrected Green-Gauss adds the sweep scratch (2n), the iterates and the sweeps' geometry
  // cotangents
  const int64_t ni = gradient == 3 ? cgg_iterate_doubles(D, 1, m.C, sweeps) : 0;
  const int64_t extra = gradient == 3 ? 2 * n + ni + cgg_geo_bar_doubles(D, m.C, m.F) : 0;
  double* scratch = afx_mesh_scratch(const_cast<afx_mesh*>(mh), 2 * n + extra);
  AFX_REQUIRE(scratch, "afx_scalar_gradient_vjp_geometry: could not allocate the scratch");
  double* G = scratch;
  double* y = scratch + n;
  double* sw = scratch + 2 * n;
  double* iters = ni ? sw + 2 * n : nullptr;
  CggGeoBars cb{sw + 2 * n + ni, sw + 2 * n + ni + (int64_t)(2 * D + 1) * m.F};
#define AFX_SGG(DD, GG)                                                                                        \
  if (D == DD && gradient == GG) {                                                                             \
    constexpr int GK = GG == 3 ? 0 : GG;                                                                       \
    AFX_LAUNCH(kKScalarGrad, (k_scalar_grad<DD, GK>), grid_for(m.C), dim3(kBlock), 0, s, m, GeoTan{}, kind,     \
               value, In<double>{phi, nullptr}, Out<double>{G, nullptr});                                      \
    AFX_CHECK_LAUNCH();                                                                                        \
    const double* yb = grad_bar;                                                                               \
    if (GG == 1) {                                                                                             \
      AFX_LAUNCH(kKScalarGradAdj, k_lsq_solve_t<DD>, grid_for(m.C), dim3(kBlock), 0, s, m, grad_bar, y);        \
      AFX_CHECK_LAUNCH();                                                                                      \
      yb = y;                                                                                                  \
    } else if (GG == 3) {                                                                                      \
      if (int rc = cgg_sweeps_store(m, DD, 1, G, iters, sweeps, relaxation, s)) return rc;                     \
      AFX_CHECK_HIP(hipMemcpyAsync(y, grad_bar, n * 8, hipMemcpyDeviceToDevice, s));                           \
      if (int rc = cgg_sweeps_adjoint_geometry(m, DD, 1, y, G, iters, sw, sweeps, relaxation, cb, s)) return rc; \
      yb = y;                                                                                                  \
    }                                                                                                          \
    AFX_LAUNCH(kKScalarGradAdj, (k_scalar_grad_geo_vjp<DD, GG>), grid_for(m.C), dim3(kBlock), 0, s, m, kind,    \
               value, phi, G, yb, *out, cb);                                                                   \
    AFX_CHECK_LAUNCH();                                                                                        \
    return 0;                                                                                                  \
  }
  AFX_SGG(2, 0) AFX_SGG(2, 1) AFX_SGG(2, 3) AFX_SGG(3, 0) AFX_SGG(3, 1) AFX_SGG(3, 3)
#undef AFX_SGG
  afx_set_error("afx_scalar_gradient_vjp_geometry: unsupported dimension");
  return 2;
}

extern "C" int afx_scalar_gradient_vjp(const afx_mesh* mh, const afx_geometry* g, int gradient, int sweeps,
                                       double relaxation, const int8_t* kind, const double* grad_bar,
                                       double* phi_bar, double* value_bar, hipStream_t s) {
  AFX_REQUIRE(mh && g && kind && grad_bar && phi_bar, "afx_scalar_gradient_vjp: null argument");
  AFX_REQUIRE(gradient == 0 || gradient == 1 || gradient == 3, "afx_scalar_gradient_vjp: unknown gradient scheme");
  AFX_REQUIRE(gradient != 3 || relaxation > 0.0, "afx_scalar_gradient_vjp: relaxation must be positive");
  MeshView m;
  if (int rv_ = afx_rview(mh, g, s, &m)) return rv_;
  if (m.C == 0) return 0;
  if (value_bar)
    if (int rz = afx_fill_zero(value_bar, m.F, s)) return rz;
  const int D = mh->dim;
  const int64_t n = (int64_t)D * m.C;
  double* scratch = afx_mesh_scratch(const_cast<afx_mesh*>(mh), (gradient == 3 ? 3 : 1) * n);
  AFX_REQUIRE(scratch, "afx_scalar_gradient_vjp: could not allocate the scratch");
#define AFX_SGV(DD)                                                                                          \
  if (D == DD) {                                                                                             \
    if (gradient == 1) {                                                                                     \
      AFX_LAUNCH(kKScalarGradAdj, k_lsq_solve_t<DD>, grid_for(m.C), dim3(kBlock), 0, s, m, grad_bar, scratch);\
      AFX_CHECK_LAUNCH();                                                                                    \
      AFX_LAUNCH(kKScalarGradAdj, (k_scalar_grad_vjp<DD, 1, false>), grid_for(m.C), dim3(kBlock), 0, s, m,    \
                 kind, scratch, phi_bar, value_bar);                                                         \
    } else if (gradient == 0) {                                                                              \
      AFX_LAUNCH(kKScalarGradAdj, (k_scalar_grad_vjp<DD, 0, true>), grid_for(m.C), dim3(kBlock), 0, s, m,     \
                 kind, grad_bar, phi_bar, value_bar);                                                        \
    } else {                                                                                                 \
      AFX_CHECK_HIP(hipMemcpyAsync(scratch, grad_bar, n * 8, hipMemcpyDeviceToDevice, s));                   \
      if (int rc = cgg_sweeps_adjoint(m, DD, 1, scratch, scratch + n, sweeps, relaxation, s)) return rc;     \
      AFX_LAUNCH(kKScalarGradAdj, (k_scalar_grad_vjp<DD, 0, false>), grid_for(m.C), dim3(kBlock), 0, s, m,    \
                 kind, scratch, phi_bar, value_bar);                                                         \
    }                                                                                                        \
    AFX_CHECK_LAUNCH();                                                                                      \
    return 0;                                                                                                \
  }
  AFX_SGV(2) AFX_SGV(3)
#undef AFX_SGV
  afx_set_error("afx_scalar_gradient_vjp: unsupported dimension");
  return 2;
}

extern "C" int afx_scalar_workspace_bytes(const afx_mesh* m, int mode, int64_t* bytes) {
  AFX_REQUIRE(m && bytes && mode >= 0 && mode <= 2, "afx_scalar_workspace_bytes: bad argument");
  *bytes = sc_work_doubles(m->dim, m->C, mode, m->F) * 8 + 256;
  return 0;
}

extern "C" int afx_scalar_residual(const afx_mesh* m, const afx_geometry* g, const afx_scalar_config* cfg,
                                   const afx_scalar_inputs* in, double* residual, void* work, int64_t work_size,
                                   hipStream_t s) {
  int rc = sc_check(m, g, cfg, in, work, work_size, 0);
  if (rc) return rc;
  AFX_REQUIRE(residual, "afx_scalar_residual: null residual");
  ScWork w = sc_carve(work, m->dim, m->C, 0);
  ScArgs<double> a = sc_args<double>(cfg, in, nullptr);
  return sc_dispatch(m->dim, cfg->advection, cfg->gradient, [&](auto Dc, auto Ac, auto Gc) {
    return sc_forward<decltype(Dc)::value, double, decltype(Ac)::value, decltype(Gc)::value>(
        m, g, a, Out<double>{residual, nullptr}, w, nullptr, nullptr, nullptr, cfg->gradient_sweeps,
        cfg->gradient_relaxation, s);
  });
}

// ---------------------------------------------------------------------------------------------
// Allocation-free residual: the workspace lives in a file-scope device array, so a caller that may
// not allocate (graph capture, a tight serving loop) passes only inputs and outputs.  Every call on
// one GPU shares that GPU's copy; calls that overlap on one GPU race.  Nothing here needs zeroed
// memory: every workspace array is written by the gradient pass before any kernel reads it.
namespace {

constexpr int64_t kNoAllocScratchBytes = 64ll << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // per device: the address of `scratch`, taken once

// Device address of `scratch` on the current device (looked up on the first call on that device).
int noalloc_scratch(void** out) {
  int dev = 0;
  AFX_CHECK_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    AFX_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    AFX_REQUIRE(ptr != nullptr, "afx_scalar_residual_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" int afx_scalar_noalloc_max_bytes(int64_t* bytes) {
  AFX_REQUIRE(bytes, "afx_scalar_noalloc_max_bytes: null argument");
  *bytes = kNoAllocScratchBytes;
  return 0;
}

extern "C" int afx_scalar_residual_noalloc(const afx_mesh* m, const afx_geometry* g, const afx_scalar_config* cfg,
                                           const afx_scalar_inputs* in, double* residual, hipStream_t s) {
  AFX_REQUIRE(m && g, "afx_scalar_residual_noalloc: null argument");
  int dev = 0;
  AFX_CHECK_HIP(hipGetDevice(&dev));
  AFX_REQUIRE(dev == m->device, "afx_scalar_residual_noalloc: the mesh handle's device is not the current device");
  // with a permuted face order the residual kernels read the packed face geometry; gathering it
  // would need the handle's own (allocated) pack buffer
  AFX_REQUIRE(m->fu == nullptr || g->face_pack != nullptr,
              "afx_scalar_residual_noalloc: this mesh needs geometry.face_pack (afx_geometry_pack)");
  const int64_t need = sc_work_doubles(m->dim, m->C, 0) * 8;
  if (need > kNoAllocScratchBytes) {
    afx_set_error("afx_scalar_residual_noalloc: the mesh needs " + std::to_string(need) +
                  " scratch bytes, more than the " + std::to_string(kNoAllocScratchBytes) + " available");
    return 2;
  }
  void* work = nullptr;
  if (int rc = noalloc_scratch(&work)) return rc;
  return afx_scalar_residual(m, g, cfg, in, residual, work, kNoAllocScratchBytes, s);
}

extern "C" int afx_scalar_jvp(const afx_mesh* m, const afx_geometry* g, const afx_scalar_config* cfg,
                              const afx_scalar_inputs* in, const afx_scalar_inputs* tangent, double* residual,
                              double* residual_tangent, void* work, int64_t work_size, hipStream_t s) {
  int rc = sc_check(m, g, cfg, in, work, work_size, 1);
  if (rc) return rc;
  AFX_REQUIRE(tangent && residual && residual_tangent, "afx_scalar_jvp: null argument");
  ScWork w = sc_carve(work, m->dim, m->C, 1);
  ScArgs<dual> a = sc_args<dual>(cfg, in, tangent);
  return sc_dispatch(m->dim, cfg->advection, cfg->gradient, [&](auto Dc, auto Ac, auto Gc) {
    return sc_forward<decltype(Dc)::value, dual, decltype(Ac)::value, decltype(Gc)::value>(
        m, g, a, Out<dual>{residual, residual_tangent}, w, w.Gt, w.psit, nullptr, cfg->gradient_sweeps,
        cfg->gradient_relaxation, s);
  });
}

extern "C" int afx_scalar_jvp_geometry(const afx_mesh* m, const afx_geometry* g, const afx_geometry* geom_tangent,
                                       const afx_scalar_config* cfg, const afx_scalar_inputs* in,
                                       const afx_scalar_inputs* tangent, double* residual, double* residual_tangent,
                                       void* work, int64_t work_size, hipStream_t s) {
  int rc = sc_check(m, g, cfg, in, work, work_size, 1);
  if (rc) return rc;
  AFX_REQUIRE(geom_tangent && tangent && residual && residual_tangent, "afx_scalar_jvp_geometry: null argument");
  AFX_REQUIRE(geom_tangent->area && geom_tangent->normal && geom_tangent->face_centroid && geom_tangent->volume &&
                  geom_tangent->cell_centroid,
              "afx_scalar_jvp_geometry: the geometry tangent needs all five arrays");
  AFX_REQUIRE(m->fu == nullptr || geom_tangent->face_pack,
              "afx_scalar_jvp_geometry: this mesh's faces are reordered: geom_tangent->face_pack must hold the face "
              "tangents in the internal order (afx_geometry_pack of the tangent arrays)");
  ScWork w = sc_carve(work, m->dim, m->C, 1);
  ScArgs<dual> a = sc_args<dual>(cfg, in, tangent);
  MeshView v;
  if (int rv_ = afx_rview(m, g, s, &v)) return rv_;
  GeoTan gt;
  if (m->fu) {
    gt.area = geom_tangent->face_pack;
    gt.nrm = geom_tangent->face_pack + m->F;
    gt.fc = geom_tangent->face_pack + (int64_t)(1 + m->dim) * m->F;
  } else {
    gt.area = geom_tangent->area;
    gt.nrm = geom_tangent->normal;
    gt.fc = geom_tangent->face_centroid;
  }
  gt.vol = geom_tangent->volume;
  gt.cc = geom_tangent->cell_centroid;
  return sc_dispatch(m->dim, cfg->advection, cfg->gradient, [&](auto Dc, auto Ac, auto Gc) {
    constexpr int D = decltype(Dc)::value, ADV = decltype(Ac)::value, GR = decltype(Gc)::value;
    constexpr bool LIM = ADV == kAdvLimited;
    constexpr bool CG = GR == kGradCorrected;
    // corrected Green-Gauss: the compact pass supplies V^-1 b, the sweeps and the limiter follow
    AFX_LAUNCH(kKScalarGrad, (k_sc_grad_g<D, CG ? 0 : GR, LIM && !CG>), grid_for(v.C), dim3(kBlock), 0, s, v, gt, a,
               Out<dual>{w.G, w.Gt}, Out<dual>{w.psi, w.psit});
    AFX_CHECK_LAUNCH();
    a.G.v = w.G;
    a.G.t = w.Gt;
    if (CG) {
      double* scr = afx_mesh_scratch(const_cast<afx_mesh*>(m), (int64_t)4 * D * m->C);
      AFX_REQUIRE(scr, "afx_scalar_jvp_geometry: could not allocate the sweep scratch");
      int r2 = cgg_sweeps_tangent(v, gt, D, 1, w.G, w.Gt, scr, cfg->gradient_sweeps, cfg->gradient_relaxation, s);
      if (r2) return r2;
      if (LIM) {
        LimiterSlots sl{};
        sl.nl = 1;
        sl.ext_stride = 1;
        r2 = venkat_pass_tangent(v, gt, D, sl, a.phi, In<dual>{w.G, w.Gt}, a.lim_k, Out<dual>{w.psi, w.psit}, s);
        if (r2) return r2;
      }
    }
    a.psi.v = w.psi;
    a.psi.t = w.psit;
    AFX_LAUNCH(kKScalarRes, (k_sc_res_g<D, ADV>), grid_for(v.C), dim3(kBlock), 0, s, v, gt, a,
               Out<dual>{residual, residual_tangent});
    AFX_CHECK_LAUNCH();
    return 0;
  });
}

// ---------------------------------------------------------------------------------------------
// Forward record: the gradient pass's outputs (gradient, limiter, the limiter's extremum cells) kept
// from the residual evaluation, so the reverse pass at the same inputs does not run it again.
namespace {

int64_t sc_record_doubles(int D, int64_t C) { return (int64_t)(D + 2) * C + (C + 1) / 2; }

ScWork sc_record_view(const void* record, int D, int64_t C) {
  ScWork w{};
  double* p = (double*)const_cast<void*>(record);
  w.G = p;
  w.psi = p + (int64_t)D * C;
  w.ext = (int32_t*)(w.psi + C);
  w.kmin = w.ext + 2 * C;
  return w;
}

int sc_vjp(const afx_mesh* m, const afx_geometry* g, const afx_scalar_config* cfg, const afx_scalar_inputs* in,
           const double* cotangent, const void* record, afx_scalar_cotangents* out, void* work, int64_t work_size,
           hipStream_t s);

}  // namespace

extern "C" int afx_scalar_record_bytes(const afx_mesh* m, int64_t* bytes) {
  AFX_REQUIRE(m && bytes, "afx_scalar_record_bytes: null argument");
  *bytes = sc_record_doubles(m->dim, m->C) * 8;
  return 0;
}

extern "C" int afx_scalar_residual_record(const afx_mesh* m, const afx_geometry* g, const afx_scalar_config* cfg,
                                          const afx_scalar_inputs* in, double* residual, void* record,
                                          int64_t record_size, hipStream_t s) {
  AFX_REQUIRE(m && g && cfg && in, "afx_scalar_residual_record: null argument");
  AFX_REQUIRE(cfg->gradient != 3,
              "afx_scalar_residual_record: the corrected Green-Gauss reverse pass needs its sweep iterates; "
              "use afx_scalar_residual and afx_scalar_vjp");
  AFX_REQUIRE(residual, "afx_scalar_residual_record: null residual");
  AFX_REQUIRE(record && record_size >= sc_record_doubles(m->dim, m->C) * 8,
              "afx_scalar_residual_record: record too small");
  // no workspace: the record holds the gradient pass's arrays, the only ones a forward pass without
  // the corrected scheme writes
  int rc = sc_check(m, g, cfg, in, nullptr, 0, -1);
  if (rc) return rc;
  ScWork w = sc_record_view(record, m->dim, m->C);
  ScArgs<double> a = sc_args<double>(cfg, in, nullptr);
  return sc_dispatch(m->dim, cfg->advection, cfg->gradient, [&](auto Dc, auto Ac, auto Gc) {
    return sc_forward<decltype(Dc)::value, double, decltype(Ac)::value, decltype(Gc)::value>(
        m, g, a, Out<double>{residual, nullptr}, w, nullptr, nullptr, w.ext, cfg->gradient_sweeps,
        cfg->gradient_relaxation, s);
  });
}

extern "C" int afx_scalar_vjp_record(const afx_mesh* m, const afx_geometry* g, const afx_scalar_config* cfg,
                                     const afx_scalar_inputs* in, const double* cotangent, const void* record,
                                     int64_t record_size, afx_scalar_cotangents* out, void* work,
                                     int64_t work_size, hipStream_t s) {
  AFX_REQUIRE(m && cfg, "afx_scalar_vjp_record: null argument");
  AFX_REQUIRE(cfg->gradient != 3, "afx_scalar_vjp_record: the corrected Green-Gauss scheme has no record");
  AFX_REQUIRE(record && record_size >= sc_record_doubles(m->dim, m->C) * 8,
              "afx_scalar_vjp_record: record too small");
  return sc_vjp(m, g, cfg, in, cotangent, record, out, work, work_size, s);
}

extern "C" int afx_scalar_vjp(const afx_mesh* m, const afx_geometry* g, const afx_scalar_config* cfg,
                              const afx_scalar_inputs* in, const double* cotangent, afx_scalar_cotangents* out,
                              void* work, int64_t work_size, hipStream_t s) {
  return sc_vjp(m, g, cfg, in, cotangent, nullptr, out, work, work_size, s);
}

namespace {

// record: null (the gradient pass runs here) or what afx_scalar_residual_record kept at these inputs.
int sc_vjp(const afx_mesh* m, const afx_geometry* g, const afx_scalar_config* cfg, const afx_scalar_inputs* in,
           const double* cotangent, const void* record, afx_scalar_cotangents* out, void* work, int64_t work_size,
           hipStream_t s) {
  int rc = sc_check(m, g, cfg, in, work, work_size, 2);
  if (rc) return rc;
  AFX_REQUIRE(cotangent && out && out->phi && out->gamma && out->flux && out->value && out->t_inf,
              "afx_scalar_vjp: null cotangent argument");
  const afx_geometry_cotangents& gb = out->geometry;
  const bool geo = gb.area || gb.normal || gb.face_centroid || gb.volume || gb.cell_centroid;
  AFX_REQUIRE(!geo || (gb.area && gb.normal && gb.face_centroid && gb.volume && gb.cell_centroid),
              "afx_scalar_vjp: geometry cotangents need all five arrays");
  ScWork w = sc_carve(work, m->dim, m->C, 2, m->F);
  if (record) {
    const ScWork r = sc_record_view(record, m->dim, m->C);
    w.G = r.G;
    w.psi = r.psi;
    w.ext = r.ext;
    w.kmin = r.kmin;
  }
  ScArgs<double> a = sc_args<double>(cfg, in, nullptr);
  AFX_REQUIRE(cfg->transient == 0 || out->phi_old, "afx_scalar_vjp: the transient term needs a phi_old cotangent");
  AFX_REQUIRE(cfg->transient != 2 || out->phi_older, "afx_scalar_vjp: BDF2 needs a phi_older cotangent");
  ScBars bar{out->phi_old, out->phi_older, out->phi, out->gamma, out->flux, out->value, out->t_inf, w.Gb, w.psib,
             w.extb, w.limx, w.limf, gb.area, gb.normal, gb.face_centroid, gb.volume, gb.cell_centroid,
             nullptr, nullptr, geo ? w.frec : nullptr};
  return sc_dispatch(m->dim, cfg->advection, cfg->gradient, [&](auto Dc, auto Ac, auto Gc) {
    constexpr int D = decltype(Dc)::value, ADV = decltype(Ac)::value, GR = decltype(Gc)::value;
    constexpr bool LIM = ADV == kAdvLimited;
    // corrected Green-Gauss with geometry cotangents: the forward iterates and the sweeps' own
    // geometry cotangents live in the mesh handle's scratch
    double* iters = nullptr;
    CggGeoBars cb{nullptr, nullptr};
    if (GR == kGradCorrected && geo) {
      const int64_t ni = cgg_iterate_doubles(D, 1, m->C, cfg->gradient_sweeps);
      double* scr = afx_mesh_scratch(const_cast<afx_mesh*>(m), ni + cgg_geo_bar_doubles(D, m->C, m->F));
      AFX_REQUIRE(scr, "afx_scalar_vjp: could not allocate the corrected-gradient scratch");
      iters = ni ? scr : nullptr;
      cb.face = scr + ni;
      cb.vol = cb.face + (int64_t)(2 * D + 1) * m->F;
      bar.cggf = cb.face;
      bar.cggv = cb.vol;
    }
    if (!record) {
      int r2 = sc_forward<D, double, ADV, GR>(m, g, a, Out<double>{nullptr, nullptr}, w, nullptr, nullptr, w.ext,
                                              cfg->gradient_sweeps, cfg->gradient_relaxation, s, iters);
      if (r2) return r2;
    }
    MeshView v;
    if (int rv_ = afx_rview(m, g, s, &v)) return rv_;
    ScArgs<double> b = a;
    b.G.v = w.G;
    b.psi.v = w.psi;
    constexpr int GK = GR == kGradCorrected ? 0 : GR;
    auto run = [&](auto Geo) {
      constexpr bool GEO = decltype(Geo)::value;
      // boundary-value cotangents: zero over all faces (one coalesced fill), boundary entries
      // written by the face passes
      if (int rz = afx_fill_zero(bar.val, m->F, s)) return rz;
      if (int rz = afx_fill_zero(bar.tinf, m->F, s)) return rz;
      AFX_LAUNCH(kKScalarResAdj, (k_sc_res_adj<D, ADV, GEO>), grid_for(m->C), dim3(kBlock), 0, s, v, b, cotangent,
                 bar);
      AFX_CHECK_LAUNCH();
      AFX_LAUNCH(kKScalarLimAdj, (k_sc_lim_adj<D, GR, LIM, GEO>), grid_for(m->C), dim3(kBlock), 0, s, v, b, w.ext,
                 w.kmin, bar);
      AFX_CHECK_LAUNCH();
      if (GR == kGradCorrected) {
        int r3 = GEO ? cgg_sweeps_adjoint_geometry(v, D, 1, bar.G, w.G, iters, w.cgg, cfg->gradient_sweeps,
                                                   cfg->gradient_relaxation, cb, s)
                     : cgg_sweeps_adjoint(v, D, 1, bar.G, w.cgg, cfg->gradient_sweeps, cfg->gradient_relaxation, s);
        if (r3) return r3;
      }
      AFX_LAUNCH(kKScalarGradAdj, (k_sc_grad_adj<D, GK, LIM, GEO, GEO && GR == kGradCorrected>), grid_for(m->C),
                 dim3(kBlock), 0, s, v, b, w.ext, bar);
      AFX_CHECK_LAUNCH();
      return 0;
    };
    return geo ? run(std::true_type{}) : run(std::false_type{});
  });
}

}  // namespace
