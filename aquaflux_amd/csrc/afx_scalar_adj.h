// Hand-written adjoints of the scalar face fluxes that also return the cotangent of the face
// geometry (area, normal, face centroid and the two cell centroids), for the node-coordinate VJPs
// of the scalar-transport and RANS residuals.  The derived quantities of FaceGeo (g, d.n, the
// fit / correction vectors) are folded back onto the primitive ones by adj_face_geo.
//
// Reference (file:line): DiffusionFlux.face_flux / flux_continuous_denominator
// (discretization/diffusion.py:59-199), _upwind_value / LimitedUpwind.face_value / AdvectionFlux
// (discretization/advection.py:41-187), boundary closures (boundary/conditions.py:46-201).
#pragma once

#include "afx_common.h"
#include "afx_scalar_face.h"

namespace afx {

// Cotangents of the quantities a face's physics reads from its FaceGeo.
template <int D>
struct FaceGeoBar {
  double A, n[D], xf[D], xP[D], xN[D], g, dp[D], dn[D], dpn, dnn;
  AFX_DEV void zero() {
    A = g = dpn = dnn = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) n[j] = xf[j] = xP[j] = xN[j] = dp[j] = dn[j] = 0.0;
  }
};

// Folds the derived cotangents (g, dp, dn, dpn, dnn) into the primitive ones (A, n, xf, xP, xN),
// reversing load_face_geo.  On a boundary face xN aliases xP.
template <int D>
AFX_DEV void adj_face_geo(const FaceGeo<D>& q, bool interior, FaceGeoBar<D>& b) {
#pragma unroll
  for (int j = 0; j < D; ++j) {
    b.dp[j] += b.dpn * q.n[j];
    b.n[j] += b.dpn * q.dp[j] + b.dnn * q.dn[j];
    b.dn[j] += b.dnn * q.n[j];
  }
  if (interior) {
    const double dd = vdot<D>(q.dvec, q.dvec);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      b.dp[j] += b.g * q.dvec[j] / dd;
      const double dvb = b.g * (q.dp[j] - 2.0 * q.g * q.dvec[j]) / dd;
      b.xN[j] += dvb;
      b.xP[j] -= dvb;
    }
  }
#pragma unroll
  for (int j = 0; j < D; ++j) {
    b.xf[j] += b.dp[j] + b.dn[j];
    b.xP[j] -= b.dp[j];
    b.xN[j] -= b.dn[j];
    if (!interior) {
      b.xP[j] += b.xN[j];
      b.xN[j] = 0.0;
    }
    b.dp[j] = b.dn[j] = 0.0;
  }
  b.dpn = b.dnn = b.g = 0.0;
}

// corr = G . (d - n (d.n)) for d = dp (side 0) or dn (side 1), cotangent cb: the gradient part ...
template <int D>
AFX_DEV void adj_corr_state(const FaceGeo<D>& q, int side, double cb, double* Gb) {
  const double* d = side == 0 ? q.dp : q.dn;
  const double dn_ = side == 0 ? q.dpn : q.dnn;
#pragma unroll
  for (int j = 0; j < D; ++j) Gb[j] += cb * (d[j] - q.n[j] * dn_);
}

// ... and the geometry part.
template <int D>
AFX_DEV void adj_corr_geo(const FaceGeo<D>& q, int side, const double* G, double cb, FaceGeoBar<D>& qb) {
  const double dn_ = side == 0 ? q.dpn : q.dnn;
  double gn = 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) gn += G[j] * q.n[j];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    if (side == 0)
      qb.dp[j] += cb * G[j];
    else
      qb.dn[j] += cb * G[j];
    qb.n[j] -= cb * G[j] * dn_;
  }
  if (side == 0)
    qb.dpn -= cb * gn;
  else
    qb.dnn -= cb * gn;
}

// Adjoint of scalar_bval (owner P, owner gradient G); bb = cotangent of the boundary value.
template <int D, bool GEO>
AFX_DEV void adj_sc_bval(int kind, const FaceGeo<D>& q, double phiP, const double* G, double v, double gam,
                         double tinf, double bb, ScalarBar<D>& B, double& val_b, double& tinf_b,
                         FaceGeoBar<D>& qb) {
  if (kind == kBcDirichlet) {
    val_b += bb;
    return;
  }
  if (kind == kBcNone) return;
  double corr = 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) corr += G[j] * (q.dp[j] - q.n[j] * q.dpn);
  double ub = bb;  // cotangent of phiP + corr
  if (kind == kBcNeumann) {
    // phiP + corr - v / gam * dpn
    val_b -= bb * q.dpn / gam;
    B.gam += bb * v * q.dpn / (gam * gam);
    if (GEO) qb.dpn -= bb * v / gam;
  } else if (kind == kBcConvective) {
    const double beta = v / gam * q.dpn;
    const double u = phiP + corr + beta * tinf;
    const double b = u / (1.0 + beta);
    ub = bb / (1.0 + beta);
    const double beta_b = ub * (tinf - b);
    tinf_b += ub * beta;
    val_b += beta_b * q.dpn / gam;
    B.gam -= beta_b * v * q.dpn / (gam * gam);
    if (GEO) qb.dpn += beta_b * v / gam;
  }
  B.phi += ub;
  adj_corr_state<D>(q, 0, ub, B.G);
  if (GEO) adj_corr_geo<D>(q, 0, G, ub, qb);
}

// Adjoint of the interior-face flux for the side `side` (0 = owner P, 1 = neighbour N); the face
// cotangent is lb.  Accumulates that side's state cotangents into B and, with GEO, the full face
// geometry cotangent into qb (identical on both sides).  Returns the cotangent of Q.
template <int D, int ADV, bool GEO>
AFX_DEV double adj_sc_interior(const FaceGeo<D>& q, double Qf, double phiP, double phiN, const double* GP,
                               const double* GN, double psiP, double psiN, double gamP, double gamN, double lb,
                               int side, ScalarBar<D>& B, FaceGeoBar<D>& qb) {
  double Q_bar = 0.0;
  if (ADV != kAdvNone) {
    const bool outflow = Qf >= 0.0;
    const double fv = scalar_face_value<D, ADV, double>(q, outflow, phiP, phiN, GP, GN, psiP, psiN);
    Q_bar = lb * fv;
    const double fb = lb * Qf;
    const bool mine = (side == 0) == outflow;
    if (mine) B.phi += fb;
    if (ADV == kAdvLimited) {
      // element-wise selects: a runtime choice between two arrays' addresses would place both
      // in per-lane scratch memory
      double xu[D], Gu[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        xu[j] = outflow ? q.xP[j] : q.xN[j];
        Gu[j] = outflow ? GP[j] : GN[j];
      }
      const double ps = outflow ? psiP : psiN;
      double t = 0.0;
#pragma unroll
      for (int j = 0; j < D; ++j) t += Gu[j] * (q.xf[j] - xu[j]);
      if (mine) {
        B.psi += fb * t;
#pragma unroll
        for (int j = 0; j < D; ++j) B.G[j] += fb * ps * (q.xf[j] - xu[j]);
      }
      if (GEO) {
        const double tb = fb * ps;
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const double v = tb * Gu[j];
          qb.xf[j] += v;
          // both fields updated, one by an exact 0.0: a branch per field would be merged into a
          // store through a selected address, which keeps qb out of registers
          qb.xP[j] -= outflow ? v : 0.0;
          qb.xN[j] -= outflow ? 0.0 : v;
        }
      }
    }
  }
  double cp = 0.0, cn = 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    cp += GP[j] * (q.dp[j] - q.n[j] * q.dpn);
    cn += GN[j] * (q.dn[j] - q.n[j] * q.dnn);
  }
  const double ratio = gamP / gamN;
  const double denom = q.dpn - ratio * q.dnn;
  const double ng = (((phiN - phiP) + cn) - cp) / denom;
  const double ng_bar = lb * (-gamP) * q.A;
  const double num_bar = ng_bar / denom;
  const double denom_bar = -ng_bar * ng / denom;
  const double ratio_bar = -denom_bar * q.dnn;
  if (side == 0) {
    B.gam += -lb * ng * q.A + ratio_bar / gamN;
    B.phi -= num_bar;
    adj_corr_state<D>(q, 0, -num_bar, B.G);
  } else {
    B.gam -= ratio_bar * ratio / gamN;
    B.phi += num_bar;
    adj_corr_state<D>(q, 1, num_bar, B.G);
  }
  if (GEO) {
    qb.A += -lb * gamP * ng;
    qb.dpn += denom_bar;
    qb.dnn -= denom_bar * ratio;
    adj_corr_geo<D>(q, 0, GP, -num_bar, qb);
    adj_corr_geo<D>(q, 1, GN, num_bar, qb);
  }
  return Q_bar;
}

// Adjoint of the boundary-face flux (owner side) including its boundary value.  Returns Q's cotangent.
template <int D, int ADV, bool GEO>
AFX_DEV double adj_sc_boundary(int kind, const FaceGeo<D>& q, double Qf, double phiP, const double* GP, double psiP,
                               double gam, double v, double tinf, double lb, ScalarBar<D>& B, double& val_b,
                               double& tinf_b, FaceGeoBar<D>& qb) {
  const double bval = scalar_bval<D, double>(kind, q, phiP, GP, v, gam, tinf);
  double bval_b = 0.0, Q_bar = 0.0;
  if (ADV != kAdvNone) {
    const bool outflow = Qf >= 0.0;
    const double fb = lb * Qf;
    if (outflow) {
      if (ADV == kAdvLimited) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) t += GP[j] * (q.xf[j] - q.xP[j]);
        Q_bar = lb * (phiP + psiP * t);
        B.phi += fb;
        B.psi += fb * t;
#pragma unroll
        for (int j = 0; j < D; ++j) B.G[j] += fb * psiP * (q.xf[j] - q.xP[j]);
        if (GEO) {
#pragma unroll
          for (int j = 0; j < D; ++j) {
            qb.xf[j] += fb * psiP * GP[j];
            qb.xP[j] -= fb * psiP * GP[j];
          }
        }
      } else {
        Q_bar = lb * phiP;
        B.phi += fb;
      }
    } else {
      Q_bar = lb * bval;
      bval_b += fb;
    }
  }
  double cp = 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) cp += GP[j] * (q.dp[j] - q.n[j] * q.dpn);
  const double ng = ((bval - phiP) - cp) / q.dpn;
  B.gam += -lb * ng * q.A;
  const double ng_bar = lb * (-gam) * q.A;
  const double num_bar = ng_bar / q.dpn;
  bval_b += num_bar;
  B.phi -= num_bar;
  adj_corr_state<D>(q, 0, -num_bar, B.G);
  if (GEO) {
    adj_corr_geo<D>(q, 0, GP, -num_bar, qb);
    qb.A += -lb * gam * ng;
    qb.dpn -= ng_bar * ng / q.dpn;
  }
  adj_sc_bval<D, GEO>(kind, q, phiP, GP, v, gam, tinf, bval_b, B, val_b, tinf_b, qb);
  return Q_bar;
}

}  // namespace afx
