// Per-face physics of the coupled (u, v[, w], p) residual, forward (templated on the scalar type,
// so `dual` gives the tangent) and hand-written adjoints.
//
// Reference functions restated here (file:line in the reference):
//   interpolate_to_face / blend            schemes/interpolation.py:84-176
//   advective_momentum_flux                flow/rhie_chow.py:220-271
//   interior_mass_flux (Rhie-Chow)         flow/rhie_chow.py:274-353
//   momentum_diagonal (boundary corrected) flow/rhie_chow.py:81-175, flow/momentum.py:556-635
//   flux_continuous_conductance            discretization/diffusion.py:85-134
//   DiffusionFlux.face_flux                discretization/diffusion.py:163-199
//   PressureForce.face_flux                flow/momentum.py:143-155
//   FirstOrderUpwind / LimitedUpwind       discretization/advection.py:96-158
//   flow boundary closures                 flow/boundary.py:278-446
//   VenkatakrishnanLimiter.limit           schemes/limiter.py:78-117
// Sums over spatial components run left to right (aquaflux/vectors.py:28-65); the momentum face
// flux adds diffusion, pressure, advection in that order (flow/momentum.py:793-799).
#pragma once

#include "afx_common.h"

namespace afx {

enum FlowKind : int8_t { kFlowNone = 0, kFlowWall = 1, kFlowInlet = 2, kFlowOutlet = 3 };
enum AdvKind : int { kAdvNone = 0, kAdvUpwind = 1, kAdvLimited = 2 };

// Flow data of one cell as seen by a face; the volume has the geometry's number type GT (dual
// when the geometry carries a node-coordinate tangent).
template <int D, class S, class GT = double>
struct FlowCell {
  S u[D], p, Gu[D][D], Gp[D], aP, mu, rho;
  S psi[D];
  GT V;
};

// LDS copy of the flow data of a workgroup's 256 cells (and their centroids).  FULL adds the
// pressure, pressure gradient, a_P, limiters and volume to the velocity, velocity gradient,
// viscosity and density that the momentum-diagonal pass reads.
template <int D, class S, bool FULL>
struct FlowTile {
  static constexpr int NF = FULL ? 3 * D + D * D + 4 : D + D * D + 2;
  S f[NF][256];
  double cc[D][256];
  double V[FULL ? 256 : 1];
  AFX_DEV void put(int t, const FlowCell<D, S>& X, const double* xc) {
    int i = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) f[i++][t] = X.u[a];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) f[i++][t] = X.Gu[a][b];
    f[i++][t] = X.mu;
    f[i++][t] = X.rho;
    if (FULL) {
      f[i++][t] = X.p;
#pragma unroll
      for (int a = 0; a < D; ++a) f[i++][t] = X.Gp[a];
      f[i++][t] = X.aP;
#pragma unroll
      for (int a = 0; a < D; ++a) f[i++][t] = X.psi[a];
      V[t] = X.V;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) cc[a][t] = xc[a];
  }
  AFX_DEV void get(int t, FlowCell<D, S>& O, double* xo) const {
    int i = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) O.u[a] = f[i++][t];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) O.Gu[a][b] = f[i++][t];
    O.mu = f[i++][t];
    O.rho = f[i++][t];
    if (FULL) {
      O.p = f[i++][t];
#pragma unroll
      for (int a = 0; a < D; ++a) O.Gp[a] = f[i++][t];
      O.aP = f[i++][t];
#pragma unroll
      for (int a = 0; a < D; ++a) O.psi[a] = f[i++][t];
      O.V = V[t];
    } else {
      O.p = sconst<S>(0.0);
#pragma unroll
      for (int a = 0; a < D; ++a) {
        O.Gp[a] = sconst<S>(0.0);
        O.psi[a] = sconst<S>(1.0);
      }
      O.aP = sconst<S>(1.0);
      O.V = 0.0;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) xo[a] = cc[a][t];
  }
};

// Cotangent accumulators of one cell's flow data.
template <int D>
struct FlowCellBar {
  double u[D], p, Gu[D][D], Gp[D], aP, mu, rho, psi[D];
  AFX_DEV void zero() {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      u[i] = 0.0;
      psi[i] = 0.0;
      Gp[i] = 0.0;
#pragma unroll
      for (int j = 0; j < D; ++j) Gu[i][j] = 0.0;
    }
    p = aP = mu = rho = 0.0;
  }
};

// ---------------------------------------------------------------------------------------------
// forward pieces

// (rho u)_f . n at the integration point (interior face, gradient-reconstructed).  The forward
// pieces take the geometry's number type GT as a template parameter (double, or dual for the
// node-coordinate JVP, where S is dual too).
template <int D, class S, class GT = double>
AFX_DEV S mom_normal_interior(const FaceGeo<D, GT>& q, const FlowCell<D, S, GT>& P, const FlowCell<D, S, GT>& N) {
  const GT w = 1.0 - q.g, g = q.g;
  S mn;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    S mf = w * (P.rho * P.u[i]) + g * (N.rho * N.u[i]);
    S corr = (w * (P.rho * P.Gu[i][0]) + g * (N.rho * N.Gu[i][0])) * q.skew[0];
#pragma unroll
    for (int j = 1; j < D; ++j) corr = corr + (w * (P.rho * P.Gu[i][j]) + g * (N.rho * N.Gu[i][j])) * q.skew[j];
    mf = mf + corr;
    mn = (i == 0) ? mf * q.n[0] : mn + mf * q.n[i];
  }
  return mn;
}

// Boundary face: owner value extrapolated to the face (g = 0, skew = xf - xP).
template <int D, class S, class GT = double>
AFX_DEV S mom_normal_boundary(const FaceGeo<D, GT>& q, const FlowCell<D, S, GT>& P) {
  S mn;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    S mf = P.rho * P.u[i];
    S corr = (P.rho * P.Gu[i][0]) * q.skew[0];
#pragma unroll
    for (int j = 1; j < D; ++j) corr = corr + (P.rho * P.Gu[i][j]) * q.skew[j];
    mf = mf + corr;
    mn = (i == 0) ? mf * q.n[0] : mn + mf * q.n[i];
  }
  return mn;
}

template <int D>
AFX_DEV double nsq(const FaceGeo<D>& q) {
  double s = q.n[0] * q.n[0];
#pragma unroll
  for (int j = 1; j < D; ++j) s = s + q.n[j] * q.n[j];
  return s;
}

// Interior face flux: out[i] = owner-outward momentum flux of u_i, out[D] = mass flux.
template <int D, int ADV, class S, class GT = double>
AFX_DEV void flow_flux_interior(const FaceGeo<D, GT>& q, const FlowCell<D, S, GT>& P, const FlowCell<D, S, GT>& N,
                                S out[D + 1]) {
  const GT w = 1.0 - q.g, g = q.g;
  S mn = mom_normal_interior<D, S, GT>(q, P, N);
  S rho_f = w * P.rho + g * N.rho;
  S gpf[D];
#pragma unroll
  for (int j = 0; j < D; ++j) gpf[j] = w * P.Gp[j] + g * N.Gp[j];
  S dcP = P.V / P.aP, dcN = N.V / N.aP;
  S dcf = w * dcP + g * dcN;
  S dhat = (q.n[0] * q.n[0]) * dcf;
#pragma unroll
  for (int j = 1; j < D; ++j) dhat = dhat + (q.n[j] * q.n[j]) * dcf;
  S gd = gpf[0] * q.dvec[0];
#pragma unroll
  for (int j = 1; j < D; ++j) gd = gd + gpf[j] * q.dvec[j];
  S damping = ((N.p - P.p) - gd) / q.dist;
  S mdot = (mn - rho_f * dhat * damping) * q.A;
  S gs = gpf[0] * q.skew[0];
#pragma unroll
  for (int j = 1; j < D; ++j) gs = gs + gpf[j] * q.skew[j];
  S pf = (w * P.p + g * N.p) + gs;
  S ratio = P.mu / N.mu;
  S denom = q.dpn - ratio * q.dnn;
  const bool outflow = val(mdot) >= 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    S cp = P.Gu[i][0] * (q.dp[0] - q.n[0] * q.dpn);
    S cn = N.Gu[i][0] * (q.dn[0] - q.n[0] * q.dnn);
#pragma unroll
    for (int j = 1; j < D; ++j) {
      cp = cp + P.Gu[i][j] * (q.dp[j] - q.n[j] * q.dpn);
      cn = cn + N.Gu[i][j] * (q.dn[j] - q.n[j] * q.dnn);
    }
    S qd = (((N.u[i] - P.u[i]) + cn) - cp) / denom;
    S flux = (-P.mu) * qd * q.A;
    flux = flux + (pf * q.n[i]) * q.A;
    if (ADV == kAdvUpwind) {
      flux = flux + mdot * (outflow ? P.u[i] : N.u[i]);
    } else if (ADV == kAdvLimited) {
      S rec;
      if (outflow) {
        S t = P.Gu[i][0] * (q.xf[0] - q.xP[0]);
#pragma unroll
        for (int j = 1; j < D; ++j) t = t + P.Gu[i][j] * (q.xf[j] - q.xP[j]);
        rec = P.u[i] + P.psi[i] * t;
      } else {
        S t = N.Gu[i][0] * (q.xf[0] - q.xN[0]);
#pragma unroll
        for (int j = 1; j < D; ++j) t = t + N.Gu[i][j] * (q.xf[j] - q.xN[j]);
        rec = N.u[i] + N.psi[i] * t;
      }
      flux = flux + mdot * rec;
    }
    out[i] = flux;
  }
  out[D] = mdot;
}

// Boundary face flux.  ub / pb are the closure's prescribed velocity / pressure on this face,
// gB the face's momentum diffusion coefficient (owner mu_eff, or a wall-function override).
template <int D, int ADV, class S, class GT = double>
AFX_DEV void flow_flux_boundary(const FaceGeo<D, GT>& q, int kind, const FlowCell<D, S, GT>& P, const S ub[D],
                                S pb, S gB, S out[D + 1]) {
  S ubf[D];
  S pf;
  S mdot = sconst<S>(0.0);
  if (kind == kFlowWall || kind == kFlowInlet) {
#pragma unroll
    for (int i = 0; i < D; ++i) ubf[i] = ub[i];
    pf = P.p;
    if (kind == kFlowInlet) {
      S un = ub[0] * q.n[0];
#pragma unroll
      for (int j = 1; j < D; ++j) un = un + ub[j] * q.n[j];
      mdot = P.rho * un * q.A;
    }
  } else if (kind == kFlowOutlet) {
#pragma unroll
    for (int i = 0; i < D; ++i) ubf[i] = P.u[i];
    pf = pb;
    S un = P.u[0] * q.n[0];
    S gn = P.Gp[0] * q.n[0];
#pragma unroll
    for (int j = 1; j < D; ++j) {
      un = un + P.u[j] * q.n[j];
      gn = gn + P.Gp[j] * q.n[j];
    }
    S dcP = P.V / P.aP;
    S dh = (q.n[0] * q.n[0]) * dcP;
#pragma unroll
    for (int j = 1; j < D; ++j) dh = dh + (q.n[j] * q.n[j]) * dcP;
    S compact = (pb - P.p) / q.dist;
    mdot = P.rho * (un - dh * (compact - gn)) * q.A;
  } else {
#pragma unroll
    for (int i = 0; i < D; ++i) ubf[i] = sconst<S>(0.0);
    pf = sconst<S>(0.0);
  }
  const bool outflow = val(mdot) >= 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    S cp = P.Gu[i][0] * (q.dp[0] - q.n[0] * q.dpn);
#pragma unroll
    for (int j = 1; j < D; ++j) cp = cp + P.Gu[i][j] * (q.dp[j] - q.n[j] * q.dpn);
    S qd = ((ubf[i] - P.u[i]) - cp) / q.dpn;
    S flux = (-gB) * qd * q.A;
    flux = flux + (pf * q.n[i]) * q.A;
    if (ADV == kAdvUpwind) {
      flux = flux + mdot * (outflow ? P.u[i] : ubf[i]);
    } else if (ADV == kAdvLimited) {
      S rec;
      if (outflow) {
        S t = P.Gu[i][0] * (q.xf[0] - q.xP[0]);
#pragma unroll
        for (int j = 1; j < D; ++j) t = t + P.Gu[i][j] * (q.xf[j] - q.xP[j]);
        rec = P.u[i] + P.psi[i] * t;
      } else {
        rec = ubf[i];
      }
      flux = flux + mdot * rec;
    }
    out[i] = flux;
  }
  out[D] = mdot;
}

// Momentum-diagonal contributions of an interior face to its owner (oc) and neighbour (nc).
template <int D, int ADV, class S, class GT = double>
AFX_DEV void ap_interior(const FaceGeo<D, GT>& q, const FlowCell<D, S, GT>& P, const FlowCell<D, S, GT>& N, S& oc,
                         S& nc) {
  S denom = q.dpn - (P.mu / N.mu) * q.dnn;
  S visc = P.mu * q.A / denom;
  if (ADV != kAdvNone) {
    S mest = mom_normal_interior<D, S, GT>(q, P, N) * q.A;
    oc = visc + dmax(mest, 0.0);
    nc = visc + dmax(-mest, 0.0);
  } else {
    oc = visc;
    nc = visc;
  }
}

// Owner momentum-diagonal contribution of a boundary face (per-patch coefficient).
template <int D, int ADV, class S, class GT = double>
AFX_DEV S ap_boundary(const FaceGeo<D, GT>& q, int kind, const FlowCell<D, S, GT>& P, S gB) {
  if (kind == kFlowNone) return sconst<S>(0.0);
  S visc = gB * q.A / q.dist;
  S conv = sconst<S>(0.0);
  if (ADV != kAdvNone) conv = dmax(mom_normal_boundary<D, S, GT>(q, P) * q.A, 0.0);
  if (kind == kFlowWall) return visc;
  if (kind == kFlowInlet) return visc + conv;
  return conv;  // outlet
}

// Venkatakrishnan face value seen from a cell (schemes/limiter.py:88-113).
template <class S, class E = double>
AFX_DEV S venkat_face(S dm0, S head_max, S head_min, E e2) {
  double sgn = val(dm0) >= 0.0 ? 1.0 : -1.0;
  S dm = sgn * (dabs(dm0) + 1e-12);
  S head = val(dm) > 0.0 ? head_max : head_min;
  S h2 = head * head;
  S dm2 = dm * dm;
  S num = (h2 + e2) * dm + 2.0 * dm2 * head;
  S den = (h2 + 2.0 * dm2 + head * dm + e2) * dm;
  return num / den;
}

// ---------------------------------------------------------------------------------------------
// adjoints (hand-written).  `side` selects whose cotangents are accumulated: 0 = owner P,
// 1 = neighbour N.  Each routine recomputes the forward intermediates it needs.

template <int D>
AFX_DEV void adj_mom_normal_interior(const FaceGeo<D>& q, const FlowCell<D, double>& P,
                                     const FlowCell<D, double>& N, double mn_bar, int side,
                                     FlowCellBar<D>& B) {
  const double w = side == 0 ? 1.0 - q.g : q.g;
  const FlowCell<D, double>& X = side == 0 ? P : N;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double mfb = mn_bar * q.n[i];
    double wm = w * mfb;
    B.u[i] += wm * X.rho;
    double acc = X.u[i];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      B.Gu[i][j] += wm * X.rho * q.skew[j];
      acc += X.Gu[i][j] * q.skew[j];
    }
    B.rho += wm * acc;
  }
}

template <int D>
AFX_DEV void adj_mom_normal_boundary(const FaceGeo<D>& q, const FlowCell<D, double>& P, double mn_bar,
                                     FlowCellBar<D>& B) {
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double mfb = mn_bar * q.n[i];
    B.u[i] += mfb * P.rho;
    double acc = P.u[i];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      B.Gu[i][j] += mfb * P.rho * q.skew[j];
      acc += P.Gu[i][j] * q.skew[j];
    }
    B.rho += mfb * acc;
  }
}

// Adjoint of flow_flux_interior: lam[D+1] are the cotangents of out[]; accumulates side's bars.
template <int D, int ADV>
AFX_DEV void adj_flux_interior(const FaceGeo<D>& q, const FlowCell<D, double>& P, const FlowCell<D, double>& N,
                               const double lam[D + 1], int side, FlowCellBar<D>& B) {
  const double w = 1.0 - q.g, g = q.g;
  const double ws = side == 0 ? w : g;
  const double mn = mom_normal_interior<D, double>(q, P, N);
  const double rho_f = w * P.rho + g * N.rho;
  double gpf[D];
#pragma unroll
  for (int j = 0; j < D; ++j) gpf[j] = w * P.Gp[j] + g * N.Gp[j];
  const double dcP = P.V / P.aP, dcN = N.V / N.aP;
  const double dcf = w * dcP + g * dcN;
  const double ns = nsq<D>(q);
  double dhat = (q.n[0] * q.n[0]) * dcf;
#pragma unroll
  for (int j = 1; j < D; ++j) dhat = dhat + (q.n[j] * q.n[j]) * dcf;
  double gd = gpf[0] * q.dvec[0];
#pragma unroll
  for (int j = 1; j < D; ++j) gd = gd + gpf[j] * q.dvec[j];
  const double damping = ((N.p - P.p) - gd) / q.dist;
  const double mdot = (mn - rho_f * dhat * damping) * q.A;
  const double ratio = P.mu / N.mu;
  const double denom = q.dpn - ratio * q.dnn;
  const bool outflow = mdot >= 0.0;

  double mdot_bar = lam[D];
  double pf_bar = 0.0, denom_bar = 0.0, muP_bar = 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    const double li = lam[i];
    // advection
    if (ADV == kAdvUpwind) {
      const double up = outflow ? P.u[i] : N.u[i];
      mdot_bar += li * up;
      if ((side == 0) == outflow) B.u[i] += li * mdot;
    } else if (ADV == kAdvLimited) {
      const FlowCell<D, double>& U = outflow ? P : N;
      double t = 0.0, dx[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        dx[j] = q.xf[j] - (outflow ? q.xP[j] : q.xN[j]);
        t += U.Gu[i][j] * dx[j];
      }
      const double rec = U.u[i] + U.psi[i] * t;
      mdot_bar += li * rec;
      if ((side == 0) == outflow) {
        const double rb = li * mdot;
        B.u[i] += rb;
        B.psi[i] += rb * t;
#pragma unroll
        for (int j = 0; j < D; ++j) B.Gu[i][j] += rb * U.psi[i] * dx[j];
      }
    }
    // pressure force
    pf_bar += li * q.n[i] * q.A;
    // diffusion
    double cp = 0.0, cn = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      cp += P.Gu[i][j] * (q.dp[j] - q.n[j] * q.dpn);
      cn += N.Gu[i][j] * (q.dn[j] - q.n[j] * q.dnn);
    }
    const double qd = (((N.u[i] - P.u[i]) + cn) - cp) / denom;
    const double qd_bar = li * (-P.mu) * q.A;
    muP_bar -= li * qd * q.A;
    const double num_bar = qd_bar / denom;
    denom_bar -= qd_bar * qd / denom;
    if (side == 0) {
      B.u[i] -= num_bar;
#pragma unroll
      for (int j = 0; j < D; ++j) B.Gu[i][j] -= num_bar * (q.dp[j] - q.n[j] * q.dpn);
    } else {
      B.u[i] += num_bar;
#pragma unroll
      for (int j = 0; j < D; ++j) B.Gu[i][j] += num_bar * (q.dn[j] - q.n[j] * q.dnn);
    }
  }
  // denom = dpn - (muP/muN) dnn
  const double ratio_bar = -denom_bar * q.dnn;
  if (side == 0) {
    B.mu += muP_bar + ratio_bar / N.mu;
  } else {
    B.mu -= ratio_bar * ratio / N.mu;
  }
  // face pressure pf = w pP + g pN + gpf . skew
  double gpf_bar[D];
#pragma unroll
  for (int j = 0; j < D; ++j) gpf_bar[j] = pf_bar * q.skew[j];
  B.p += ws * pf_bar;
  // mass flux
  const double t = mdot_bar * q.A;
  const double rho_f_bar = -t * dhat * damping;
  const double dhat_bar = -t * rho_f * damping;
  const double damping_bar = -t * rho_f * dhat;
  const double dpp_bar = damping_bar / q.dist;
  B.p += side == 0 ? -dpp_bar : dpp_bar;
#pragma unroll
  for (int j = 0; j < D; ++j) gpf_bar[j] -= dpp_bar * q.dvec[j];
  const double dcf_bar = dhat_bar * ns;
  const double dcs_bar = ws * dcf_bar;
  const FlowCell<D, double>& X = side == 0 ? P : N;
  const double dcX = side == 0 ? dcP : dcN;
  B.aP -= dcs_bar * dcX / X.aP;
#pragma unroll
  for (int j = 0; j < D; ++j) B.Gp[j] += ws * gpf_bar[j];
  B.rho += ws * rho_f_bar;
  adj_mom_normal_interior<D>(q, P, N, t, side, B);
}

// Adjoint of flow_flux_boundary (owner side).  Writes the cotangents of ub / pb / gB.
template <int D, int ADV>
AFX_DEV void adj_flux_boundary(const FaceGeo<D>& q, int kind, const FlowCell<D, double>& P, const double ub[D],
                               double pb, double gB, const double lam[D + 1], FlowCellBar<D>& B, double ub_bar[D],
                               double& pb_bar, double& gB_bar) {
  double ubf[D], mdot = 0.0;
  double dcP = P.V / P.aP, dh = 0.0, compact = 0.0, gn = 0.0, un = 0.0;
  if (kind == kFlowWall || kind == kFlowInlet) {
#pragma unroll
    for (int i = 0; i < D; ++i) ubf[i] = ub[i];
    if (kind == kFlowInlet) {
      un = ub[0] * q.n[0];
#pragma unroll
      for (int j = 1; j < D; ++j) un = un + ub[j] * q.n[j];
      mdot = P.rho * un * q.A;
    }
  } else if (kind == kFlowOutlet) {
#pragma unroll
    for (int i = 0; i < D; ++i) ubf[i] = P.u[i];
    un = P.u[0] * q.n[0];
    gn = P.Gp[0] * q.n[0];
#pragma unroll
    for (int j = 1; j < D; ++j) {
      un = un + P.u[j] * q.n[j];
      gn = gn + P.Gp[j] * q.n[j];
    }
    dh = (q.n[0] * q.n[0]) * dcP;
#pragma unroll
    for (int j = 1; j < D; ++j) dh = dh + (q.n[j] * q.n[j]) * dcP;
    compact = (pb - P.p) / q.dist;
    mdot = P.rho * (un - dh * (compact - gn)) * q.A;
  } else {
#pragma unroll
    for (int i = 0; i < D; ++i) ubf[i] = 0.0;
  }
  const bool outflow = mdot >= 0.0;
  double mdot_bar = lam[D], pf_bar = 0.0, ubf_bar[D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    ubf_bar[i] = 0.0;
    ub_bar[i] = 0.0;
  }
  pb_bar = 0.0;
  gB_bar = 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    const double li = lam[i];
    if (ADV == kAdvUpwind) {
      mdot_bar += li * (outflow ? P.u[i] : ubf[i]);
      if (outflow)
        B.u[i] += li * mdot;
      else
        ubf_bar[i] += li * mdot;
    } else if (ADV == kAdvLimited) {
      if (outflow) {
        double t = 0.0, dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
          dx[j] = q.xf[j] - q.xP[j];
          t += P.Gu[i][j] * dx[j];
        }
        mdot_bar += li * (P.u[i] + P.psi[i] * t);
        const double rb = li * mdot;
        B.u[i] += rb;
        B.psi[i] += rb * t;
#pragma unroll
        for (int j = 0; j < D; ++j) B.Gu[i][j] += rb * P.psi[i] * dx[j];
      } else {
        mdot_bar += li * ubf[i];
        ubf_bar[i] += li * mdot;
      }
    }
    pf_bar += li * q.n[i] * q.A;
    double cp = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) cp += P.Gu[i][j] * (q.dp[j] - q.n[j] * q.dpn);
    const double qd = ((ubf[i] - P.u[i]) - cp) / q.dpn;
    const double qd_bar = li * (-gB) * q.A;
    gB_bar -= li * qd * q.A;
    const double nb_ = qd_bar / q.dpn;
    ubf_bar[i] += nb_;
    B.u[i] -= nb_;
#pragma unroll
    for (int j = 0; j < D; ++j) B.Gu[i][j] -= nb_ * (q.dp[j] - q.n[j] * q.dpn);
  }
  if (kind == kFlowWall || kind == kFlowInlet) {
#pragma unroll
    for (int i = 0; i < D; ++i) ub_bar[i] += ubf_bar[i];
    B.p += pf_bar;
    if (kind == kFlowInlet) {
      const double t = mdot_bar * q.A;
      B.rho += t * un;
      const double un_bar = t * P.rho;
#pragma unroll
      for (int j = 0; j < D; ++j) ub_bar[j] += un_bar * q.n[j];
    }
  } else if (kind == kFlowOutlet) {
#pragma unroll
    for (int i = 0; i < D; ++i) B.u[i] += ubf_bar[i];
    pb_bar += pf_bar;
    const double t = mdot_bar * q.A;
    const double inner = un - dh * (compact - gn);
    B.rho += t * inner;
    const double inner_bar = t * P.rho;
    const double dh_bar = -inner_bar * (compact - gn);
    const double cg_bar = -inner_bar * dh;  // d/d(compact - gn)
    const double compact_bar = cg_bar;
    const double gn_bar = -cg_bar;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      B.u[j] += inner_bar * q.n[j];
      B.Gp[j] += gn_bar * q.n[j];
    }
    pb_bar += compact_bar / q.dist;
    B.p -= compact_bar / q.dist;
    const double dcP_bar = dh_bar * nsq<D>(q);
    B.aP -= dcP_bar * dcP / P.aP;
  }
}

// Adjoint of ap_interior given the cotangents of the owner's and neighbour's a_P.
template <int D, int ADV>
AFX_DEV void adj_ap_interior(const FaceGeo<D>& q, const FlowCell<D, double>& P, const FlowCell<D, double>& N,
                             double aPbarP, double aPbarN, int side, FlowCellBar<D>& B) {
  const double ratio = P.mu / N.mu;
  const double denom = q.dpn - ratio * q.dnn;
  const double visc = P.mu * q.A / denom;
  const double visc_bar = aPbarP + aPbarN;
  const double denom_bar = -visc_bar * visc / denom;
  const double ratio_bar = -denom_bar * q.dnn;
  if (side == 0)
    B.mu += visc_bar * q.A / denom + ratio_bar / N.mu;
  else
    B.mu -= ratio_bar * ratio / N.mu;
  if (ADV != kAdvNone) {
    const double mest = mom_normal_interior<D, double>(q, P, N) * q.A;
    const double mest_bar = aPbarP * dmax0_dx(mest) - aPbarN * dmax0_dx(-mest);
    adj_mom_normal_interior<D>(q, P, N, mest_bar * q.A, side, B);
  }
}

template <int D, int ADV>
AFX_DEV void adj_ap_boundary(const FaceGeo<D>& q, int kind, const FlowCell<D, double>& P, double gB,
                             double aPbar, FlowCellBar<D>& B, double& gB_bar) {
  gB_bar = 0.0;
  if (kind == kFlowNone) return;
  if (kind == kFlowWall || kind == kFlowInlet) gB_bar = aPbar * q.A / q.dist;
  if (ADV != kAdvNone && (kind == kFlowInlet || kind == kFlowOutlet)) {
    const double mest = mom_normal_boundary<D, double>(q, P) * q.A;
    adj_mom_normal_boundary<D>(q, P, aPbar * dmax0_dx(mest) * q.A, B);
  }
}

// ---------------------------------------------------------------------------------------------
// geometry cotangents (node-coordinate VJP).  Each routine returns the cotangent of the whole face
// geometry (the same on both sides of the face) given the cotangents of the face's outputs; cell
// volumes come back separately.  fold_face_geo_bar then reverses finish_face_geo.

template <int D>
struct FaceGeoBarF {
  double A, n[D], xf[D], xP[D], xN[D], g, dvec[D], dist, skew[D], dp[D], dn[D], dpn, dnn;
  AFX_DEV void zero() {
    A = g = dist = dpn = dnn = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) n[j] = xf[j] = xP[j] = xN[j] = dvec[j] = skew[j] = dp[j] = dn[j] = 0.0;
  }
};

// Folds the cotangents of the derived quantities into A, n, xf, xP, xN (xN aliases xP on a
// boundary face, where its cotangent is moved onto xP).
template <int D>
AFX_DEV void fold_face_geo_bar(const FaceGeo<D>& q, bool interior, FaceGeoBarF<D>& b) {
#pragma unroll
  for (int j = 0; j < D; ++j) {  // skew = xf - (xP + dvec g)
    b.xf[j] += b.skew[j];
    b.xP[j] -= b.skew[j];
    b.dvec[j] -= b.skew[j] * q.g;
    b.g -= b.skew[j] * q.dvec[j];
  }
  if (interior) {
    const double dd = vdot<D>(q.dvec, q.dvec);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      b.dvec[j] += b.dist * q.n[j];  // dist = dvec . n
      b.n[j] += b.dist * q.dvec[j];
      b.dp[j] += b.g * q.dvec[j] / dd;  // g = (dp . dvec) / (dvec . dvec)
      b.dvec[j] += b.g * (q.dp[j] - 2.0 * q.g * q.dvec[j]) / dd;
    }
  } else {
    b.dpn += b.dist;  // dist = dpn; g = 0
  }
#pragma unroll
  for (int j = 0; j < D; ++j) {
    b.dp[j] += b.dpn * q.n[j];
    b.n[j] += b.dpn * q.dp[j] + b.dnn * q.dn[j];
    b.dn[j] += b.dnn * q.n[j];
    b.xf[j] += b.dp[j] + b.dn[j];
    b.xP[j] -= b.dp[j] + b.dvec[j];
    b.xN[j] += b.dvec[j] - b.dn[j];
    if (!interior) {
      b.xP[j] += b.xN[j];
      b.xN[j] = 0.0;
    }
  }
}

// mn = sum_i n_i mf_i, mf_i = w rhoP uP_i + g rhoN uN_i + sum_j (w rhoP GuP_ij + g rhoN GuN_ij) skew_j
template <int D>
AFX_DEV void geo_adj_mom_normal_interior(const FaceGeo<D>& q, const FlowCell<D, double>& P,
                                         const FlowCell<D, double>& N, double mn_bar, FaceGeoBarF<D>& b) {
  const double w = 1.0 - q.g, g = q.g;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double mf = w * (P.rho * P.u[i]) + g * (N.rho * N.u[i]);
    double dmf = N.rho * N.u[i] - P.rho * P.u[i];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      mf += (w * (P.rho * P.Gu[i][j]) + g * (N.rho * N.Gu[i][j])) * q.skew[j];
      dmf += (N.rho * N.Gu[i][j] - P.rho * P.Gu[i][j]) * q.skew[j];
    }
    b.n[i] += mn_bar * mf;
    const double mfb = mn_bar * q.n[i];
    b.g += mfb * dmf;
#pragma unroll
    for (int j = 0; j < D; ++j) b.skew[j] += mfb * (w * (P.rho * P.Gu[i][j]) + g * (N.rho * N.Gu[i][j]));
  }
}

template <int D>
AFX_DEV void geo_adj_mom_normal_boundary(const FaceGeo<D>& q, const FlowCell<D, double>& P, double mn_bar,
                                         FaceGeoBarF<D>& b) {
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double mf = P.rho * P.u[i];
#pragma unroll
    for (int j = 0; j < D; ++j) mf += (P.rho * P.Gu[i][j]) * q.skew[j];
    b.n[i] += mn_bar * mf;
    const double mfb = mn_bar * q.n[i];
#pragma unroll
    for (int j = 0; j < D; ++j) b.skew[j] += mfb * (P.rho * P.Gu[i][j]);
  }
}

// Geometry cotangent of flow_flux_interior for output cotangents lam[D+1]; VP_bar / VN_bar receive
// the cell-volume cotangents (the Rhie-Chow V / a_P).
template <int D, int ADV>
AFX_DEV void geo_adj_flux_interior(const FaceGeo<D>& q, const FlowCell<D, double>& P, const FlowCell<D, double>& N,
                                   const double lam[D + 1], FaceGeoBarF<D>& b, double& VP_bar, double& VN_bar) {
  const double w = 1.0 - q.g, g = q.g;
  const double mn = mom_normal_interior<D, double>(q, P, N);
  const double rho_f = w * P.rho + g * N.rho;
  double gpf[D];
#pragma unroll
  for (int j = 0; j < D; ++j) gpf[j] = w * P.Gp[j] + g * N.Gp[j];
  const double dcP = P.V / P.aP, dcN = N.V / N.aP;
  const double dcf = w * dcP + g * dcN;
  const double ns = nsq<D>(q);
  const double dhat = ns * dcf;
  double gd = 0.0, gs = 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    gd += gpf[j] * q.dvec[j];
    gs += gpf[j] * q.skew[j];
  }
  const double damping = ((N.p - P.p) - gd) / q.dist;
  const double mdot = (mn - rho_f * dhat * damping) * q.A;
  const double pf = (w * P.p + g * N.p) + gs;
  const double ratio = P.mu / N.mu;
  const double denom = q.dpn - ratio * q.dnn;
  const bool outflow = mdot >= 0.0;
  double mdot_bar = lam[D], pf_bar = 0.0, denom_bar = 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    const double li = lam[i];
    if (ADV == kAdvUpwind) {
      mdot_bar += li * (outflow ? P.u[i] : N.u[i]);
    } else if (ADV == kAdvLimited) {
      const FlowCell<D, double>& U = outflow ? P : N;
      double t = 0.0;
#pragma unroll
      for (int j = 0; j < D; ++j) t += U.Gu[i][j] * (q.xf[j] - (outflow ? q.xP[j] : q.xN[j]));
      mdot_bar += li * (U.u[i] + U.psi[i] * t);
      const double tb = li * mdot * U.psi[i];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        b.xf[j] += tb * U.Gu[i][j];
        if (outflow)
          b.xP[j] -= tb * U.Gu[i][j];
        else
          b.xN[j] -= tb * U.Gu[i][j];
      }
    }
    double cp = 0.0, cn = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      cp += P.Gu[i][j] * (q.dp[j] - q.n[j] * q.dpn);
      cn += N.Gu[i][j] * (q.dn[j] - q.n[j] * q.dnn);
    }
    const double qd = (((N.u[i] - P.u[i]) + cn) - cp) / denom;
    b.A += li * ((-P.mu) * qd + pf * q.n[i]);
    b.n[i] += li * pf * q.A;
    pf_bar += li * q.n[i] * q.A;
    const double qd_bar = li * (-P.mu) * q.A;
    denom_bar -= qd_bar * qd / denom;
    const double cb = qd_bar / denom;  // d qd / d cn = 1 / denom, d qd / d cp = -1 / denom
#pragma unroll
    for (int j = 0; j < D; ++j) {
      b.dp[j] -= cb * P.Gu[i][j];
      b.n[j] += cb * P.Gu[i][j] * q.dpn;
      b.dpn += cb * P.Gu[i][j] * q.n[j];
      b.dn[j] += cb * N.Gu[i][j];
      b.n[j] -= cb * N.Gu[i][j] * q.dnn;
      b.dnn -= cb * N.Gu[i][j] * q.n[j];
    }
  }
  b.dpn += denom_bar;
  b.dnn -= denom_bar * ratio;
  // pf = w pP + g pN + gpf . skew
  double gpf_bar[D];
  b.g += pf_bar * (N.p - P.p);
#pragma unroll
  for (int j = 0; j < D; ++j) {
    gpf_bar[j] = pf_bar * q.skew[j];
    b.skew[j] += pf_bar * gpf[j];
  }
  // mdot = (mn - rho_f dhat damping) A
  b.A += mdot_bar * (mn - rho_f * dhat * damping);
  const double t = mdot_bar * q.A;
  const double rho_f_bar = -t * dhat * damping;
  const double dhat_bar = -t * rho_f * damping;
  const double damping_bar = -t * rho_f * dhat;
  b.dist -= damping_bar * damping / q.dist;
  const double gd_bar = -damping_bar / q.dist;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    gpf_bar[j] += gd_bar * q.dvec[j];
    b.dvec[j] += gd_bar * gpf[j];
    b.n[j] += dhat_bar * 2.0 * q.n[j] * dcf;
  }
  const double dcf_bar = dhat_bar * ns;
  b.g += dcf_bar * (dcN - dcP);
  VP_bar += w * dcf_bar / P.aP;
  VN_bar += g * dcf_bar / N.aP;
  double gg = rho_f_bar * (N.rho - P.rho);
#pragma unroll
  for (int j = 0; j < D; ++j) gg += gpf_bar[j] * (N.Gp[j] - P.Gp[j]);
  b.g += gg;
  geo_adj_mom_normal_interior<D>(q, P, N, t, b);
}

// Geometry cotangent of flow_flux_boundary (owner side); VP_bar receives the owner volume's.
template <int D, int ADV>
AFX_DEV void geo_adj_flux_boundary(const FaceGeo<D>& q, int kind, const FlowCell<D, double>& P, const double ub[D],
                                   double pb, double gB, const double lam[D + 1], FaceGeoBarF<D>& b,
                                   double& VP_bar) {
  double ubf[D], pf = 0.0, mdot = 0.0, un = 0.0, gn = 0.0, dh = 0.0, compact = 0.0;
  const double dcP = P.V / P.aP;
  if (kind == kFlowWall || kind == kFlowInlet) {
#pragma unroll
    for (int i = 0; i < D; ++i) ubf[i] = ub[i];
    pf = P.p;
    if (kind == kFlowInlet) {
#pragma unroll
      for (int j = 0; j < D; ++j) un += ub[j] * q.n[j];
      mdot = P.rho * un * q.A;
    }
  } else if (kind == kFlowOutlet) {
#pragma unroll
    for (int i = 0; i < D; ++i) ubf[i] = P.u[i];
    pf = pb;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      un += P.u[j] * q.n[j];
      gn += P.Gp[j] * q.n[j];
    }
    dh = nsq<D>(q) * dcP;
    compact = (pb - P.p) / q.dist;
    mdot = P.rho * (un - dh * (compact - gn)) * q.A;
  } else {
#pragma unroll
    for (int i = 0; i < D; ++i) ubf[i] = 0.0;
  }
  const bool outflow = mdot >= 0.0;
  double mdot_bar = lam[D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    const double li = lam[i];
    if (ADV == kAdvUpwind) {
      mdot_bar += li * (outflow ? P.u[i] : ubf[i]);
    } else if (ADV == kAdvLimited) {
      if (outflow) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) t += P.Gu[i][j] * (q.xf[j] - q.xP[j]);
        mdot_bar += li * (P.u[i] + P.psi[i] * t);
        const double tb = li * mdot * P.psi[i];
#pragma unroll
        for (int j = 0; j < D; ++j) {
          b.xf[j] += tb * P.Gu[i][j];
          b.xP[j] -= tb * P.Gu[i][j];
        }
      } else {
        mdot_bar += li * ubf[i];
      }
    }
    double cp = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) cp += P.Gu[i][j] * (q.dp[j] - q.n[j] * q.dpn);
    const double qd = ((ubf[i] - P.u[i]) - cp) / q.dpn;
    b.A += li * ((-gB) * qd + pf * q.n[i]);
    b.n[i] += li * pf * q.A;
    const double qd_bar = li * (-gB) * q.A;
    b.dpn -= qd_bar * qd / q.dpn;
    const double cb = -qd_bar / q.dpn;  // cotangent of cp
#pragma unroll
    for (int j = 0; j < D; ++j) {
      b.dp[j] += cb * P.Gu[i][j];
      b.n[j] -= cb * P.Gu[i][j] * q.dpn;
      b.dpn -= cb * P.Gu[i][j] * q.n[j];
    }
  }
  if (kind == kFlowInlet) {
    b.A += mdot_bar * P.rho * un;
    const double un_bar = mdot_bar * P.rho * q.A;
#pragma unroll
    for (int j = 0; j < D; ++j) b.n[j] += un_bar * ub[j];
  } else if (kind == kFlowOutlet) {
    const double inner = un - dh * (compact - gn);
    b.A += mdot_bar * P.rho * inner;
    const double inner_bar = mdot_bar * P.rho * q.A;
    const double dh_bar = -inner_bar * (compact - gn);
    const double cg_bar = -inner_bar * dh;
    b.dist -= cg_bar * compact / q.dist;
#pragma unroll
    for (int j = 0; j < D; ++j) b.n[j] += inner_bar * P.u[j] - cg_bar * P.Gp[j] + dh_bar * 2.0 * q.n[j] * dcP;
    VP_bar += dh_bar * nsq<D>(q) / P.aP;
  }
}

// Geometry cotangent of ap_interior given the cotangents of the owner's and neighbour's a_P.
template <int D, int ADV>
AFX_DEV void geo_adj_ap_interior(const FaceGeo<D>& q, const FlowCell<D, double>& P, const FlowCell<D, double>& N,
                                 double aPbarP, double aPbarN, FaceGeoBarF<D>& b) {
  const double ratio = P.mu / N.mu;
  const double denom = q.dpn - ratio * q.dnn;
  const double visc_bar = aPbarP + aPbarN;
  b.A += visc_bar * P.mu / denom;
  const double denom_bar = -visc_bar * P.mu * q.A / (denom * denom);
  b.dpn += denom_bar;
  b.dnn -= denom_bar * ratio;
  if (ADV != kAdvNone) {
    const double mn = mom_normal_interior<D, double>(q, P, N);
    const double mest = mn * q.A;
    const double mest_bar = aPbarP * dmax0_dx(mest) - aPbarN * dmax0_dx(-mest);
    b.A += mest_bar * mn;
    geo_adj_mom_normal_interior<D>(q, P, N, mest_bar * q.A, b);
  }
}

template <int D, int ADV>
AFX_DEV void geo_adj_ap_boundary(const FaceGeo<D>& q, int kind, const FlowCell<D, double>& P, double gB,
                                 double aPbar, FaceGeoBarF<D>& b) {
  if (kind == kFlowNone) return;
  if (kind == kFlowWall || kind == kFlowInlet) {  // gB A / dist
    b.A += aPbar * gB / q.dist;
    b.dist -= aPbar * gB * q.A / (q.dist * q.dist);
  }
  if (ADV != kAdvNone && (kind == kFlowInlet || kind == kFlowOutlet)) {
    const double mn = mom_normal_boundary<D, double>(q, P);
    const double mest_bar = aPbar * dmax0_dx(mn * q.A);
    b.A += mest_bar * mn;
    geo_adj_mom_normal_boundary<D>(q, P, mest_bar * q.A, b);
  }
}

}  // namespace afx
