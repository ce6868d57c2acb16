// Per-face physics of a transported scalar: the face flux of ScalarTransport / the SST k and omega
// equations, forward (templated) and hand-written adjoint.
//
// Reference: ResidualAssembler.residual with flux operators (AdvectionFlux(Q, scheme),
// DiffusionFlux) summed in that order (aquaflux/transport/scalar.py:185-208,
// discretization/residual.py:101-148, :358-412); boundary closures Dirichlet / ZeroGradient /
// Neumann / Convective (boundary/conditions.py:46-201).
#pragma once

#include "afx_common.h"
#include "afx_flow_face.h"

namespace afx {

enum ScalarBC : int8_t { kBcNone = 0, kBcDirichlet = 1, kBcZeroGradient = 2, kBcNeumann = 3, kBcConvective = 4 };

// Weak boundary value phi_ip of a closure (``grad`` = owner gradient; zero for leading order).
// Neumann / Convective read ``gam`` (owner diffusivity), ``val`` (q or h) and ``tinf``.
template <int D, class S, class GT = double>
AFX_DEV S scalar_bval(int kind, const FaceGeo<D, GT>& q, S phiP, const S* G, S val, S gam, S tinf) {
  if (kind == kBcDirichlet) return val;
  if (kind == kBcNone) return sconst<S>(0.0);
  S corr = G[0] * (q.dp[0] - q.n[0] * q.dpn);
#pragma unroll
  for (int j = 1; j < D; ++j) corr = corr + G[j] * (q.dp[j] - q.n[j] * q.dpn);
  if (kind == kBcZeroGradient) return phiP + corr;
  if (kind == kBcNeumann) return phiP + corr - val / gam * q.dpn;
  S beta = val / gam * q.dpn;  // convective
  return (phiP + corr + beta * tinf) / (1.0 + beta);
}

// Upwind / limited face value (interior face).
template <int D, int ADV, class S, class GT = double>
AFX_DEV S scalar_face_value(const FaceGeo<D, GT>& q, bool outflow, S phiP, S phiN, const S* GP, const S* GN, S psiP,
                            S psiN) {
  if (ADV == kAdvLimited) {
    const GT* xu = outflow ? q.xP : q.xN;
    const S* Gu = outflow ? GP : GN;
    S t = Gu[0] * (q.xf[0] - xu[0]);
#pragma unroll
    for (int j = 1; j < D; ++j) t = t + Gu[j] * (q.xf[j] - xu[j]);
    return outflow ? phiP + psiP * t : phiN + psiN * t;
  }
  return outflow ? phiP : phiN;
}

// Interior-face flux: Q phi_f + (-Gamma_P ((phiN - phiP) + cN - cP) / denom) A.
template <int D, int ADV, class S, class GT = double>
AFX_DEV S scalar_flux_interior(const FaceGeo<D, GT>& q, S Qf, S phiP, S phiN, const S* GP, const S* GN, S psiP, S psiN,
                               S gamP, S gamN) {
  S flux = sconst<S>(0.0);
  if (ADV != kAdvNone) {
    const bool outflow = val(Qf) >= 0.0;
    flux = Qf * scalar_face_value<D, ADV, S, GT>(q, outflow, phiP, phiN, GP, GN, psiP, psiN);
  }
  S cp = GP[0] * (q.dp[0] - q.n[0] * q.dpn);
  S cn = GN[0] * (q.dn[0] - q.n[0] * q.dnn);
#pragma unroll
  for (int j = 1; j < D; ++j) {
    cp = cp + GP[j] * (q.dp[j] - q.n[j] * q.dpn);
    cn = cn + GN[j] * (q.dn[j] - q.n[j] * q.dnn);
  }
  S denom = q.dpn - (gamP / gamN) * q.dnn;
  S ng = (((phiN - phiP) + cn) - cp) / denom;
  return flux + (-gamP) * ng * q.A;
}

// Boundary-face flux with weak value ``bval`` and boundary diffusion coefficient ``gB``.
template <int D, int ADV, class S, class GT = double>
AFX_DEV S scalar_flux_boundary(const FaceGeo<D, GT>& q, S Qf, S phiP, const S* GP, S psiP, S bval, S gB) {
  S flux = sconst<S>(0.0);
  if (ADV != kAdvNone) {
    const bool outflow = val(Qf) >= 0.0;
    S fv;
    if (outflow) {
      if (ADV == kAdvLimited) {
        S t = GP[0] * (q.xf[0] - q.xP[0]);
#pragma unroll
        for (int j = 1; j < D; ++j) t = t + GP[j] * (q.xf[j] - q.xP[j]);
        fv = phiP + psiP * t;
      } else {
        fv = phiP;
      }
    } else {
      fv = bval;
    }
    flux = Qf * fv;
  }
  S cp = GP[0] * (q.dp[0] - q.n[0] * q.dpn);
#pragma unroll
  for (int j = 1; j < D; ++j) cp = cp + GP[j] * (q.dp[j] - q.n[j] * q.dpn);
  S ng = ((bval - phiP) - cp) / q.dpn;
  return flux + (-gB) * ng * q.A;
}

// Cotangents of one side of a scalar face.
template <int D>
struct ScalarBar {
  double phi, G[D], psi, gam;
  AFX_DEV void zero() {
    phi = psi = gam = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) G[j] = 0.0;
  }
};

// Adjoint of scalar_flux_interior for cotangent ``lb`` of the flux; ``side`` 0 = owner, 1 = neighbour.
// Returns the cotangent of Q (shared by both sides; only the caller's side adds it once).
template <int D, int ADV>
AFX_DEV double adj_scalar_flux_interior(const FaceGeo<D>& q, double Qf, double phiP, double phiN, const double* GP,
                                        const double* GN, double psiP, double psiN, double gamP, double gamN,
                                        double lb, int side, ScalarBar<D>& B) {
  double Q_bar = 0.0;
  if (ADV != kAdvNone) {
    const bool outflow = Qf >= 0.0;
    const double fv = scalar_face_value<D, ADV, double>(q, outflow, phiP, phiN, GP, GN, psiP, psiN);
    Q_bar = lb * fv;
    const double fb = lb * Qf;
    if ((side == 0) == outflow) {
      B.phi += fb;
      if (ADV == kAdvLimited) {
        const double* xu = outflow ? q.xP : q.xN;
        const double* Gu = outflow ? GP : GN;
        const double ps = outflow ? psiP : psiN;
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) t += Gu[j] * (q.xf[j] - xu[j]);
        B.psi += fb * t;
#pragma unroll
        for (int j = 0; j < D; ++j) B.G[j] += fb * ps * (q.xf[j] - xu[j]);
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
#pragma unroll
    for (int j = 0; j < D; ++j) B.G[j] -= num_bar * (q.dp[j] - q.n[j] * q.dpn);
  } else {
    B.gam -= ratio_bar * ratio / gamN;
    B.phi += num_bar;
#pragma unroll
    for (int j = 0; j < D; ++j) B.G[j] += num_bar * (q.dn[j] - q.n[j] * q.dnn);
  }
  return Q_bar;
}

// Adjoint of scalar_flux_boundary; returns Q's cotangent, writes bval's and gB's.
template <int D, int ADV>
AFX_DEV double adj_scalar_flux_boundary(const FaceGeo<D>& q, double Qf, double phiP, const double* GP, double psiP,
                                        double bval, double gB, double lb, ScalarBar<D>& B, double& bval_bar,
                                        double& gB_bar) {
  double Q_bar = 0.0;
  bval_bar = 0.0;
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
      } else {
        Q_bar = lb * phiP;
        B.phi += fb;
      }
    } else {
      Q_bar = lb * bval;
      bval_bar += fb;
    }
  }
  double cp = 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) cp += GP[j] * (q.dp[j] - q.n[j] * q.dpn);
  const double ng = ((bval - phiP) - cp) / q.dpn;
  gB_bar = -lb * ng * q.A;
  const double num_bar = lb * (-gB) * q.A / q.dpn;
  bval_bar += num_bar;
  B.phi -= num_bar;
#pragma unroll
  for (int j = 0; j < D; ++j) B.G[j] -= num_bar * (q.dp[j] - q.n[j] * q.dpn);
  return Q_bar;
}

// Adjoint of scalar_bval for Dirichlet / ZeroGradient (the closures the turbulence fields use):
// routes bval's cotangent to the owner value / gradient, or to the prescribed value.
template <int D>
AFX_DEV void adj_scalar_bval(int kind, const FaceGeo<D>& q, double bb, ScalarBar<D>& B, double& val_bar) {
  val_bar = 0.0;
  if (kind == kBcDirichlet) {
    val_bar = bb;
  } else if (kind == kBcZeroGradient) {
    B.phi += bb;
#pragma unroll
    for (int j = 0; j < D; ++j) B.G[j] += bb * (q.dp[j] - q.n[j] * q.dpn);
  }
}

}  // namespace afx
