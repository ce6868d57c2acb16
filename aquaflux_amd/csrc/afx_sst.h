// k-omega SST closure, wall treatment and sources as cell-local device functions, templated on the
// scalar type (so `dual` gives tangents; the adjoints seed them one input at a time).
//
// Reference (file:line): SSTModel.f1 / f2 / eddy_viscosity / blend (turbulence/sst.py:84-239);
// safe_sqrt, strain_rate_magnitude (turbulence/strain.py:32-66); omega_wall, _power_mean, nut_wall,
// wall_y_star, wall_function_weight, log_layer_shear_rate, k_wall_production, omega_wall_gradient
// (turbulence/boundary.py:45-589); KProduction, KDestruction, NearWallKClosure, OmegaProduction,
// OmegaDestruction, OmegaCrossDiffusion (turbulence/sources.py:60-412); SSTTurbulence.strain_rate,
// closure_fields, k_residual, omega_residual (turbulence/transport.py:369-718).
#pragma once

#include "afx_common.h"

namespace afx {

struct SstConst {
  double a1, beta_star, alpha1, beta1, sk1, sw1, alpha2, beta2, sk2, sw2, kappa, e_wall, p_exp, c_vis;
  double bsq;   // beta_star ** 0.25
  double ylam;  // laminar / log crossover y*_lam
  double density;
};

constexpr double kCrossDiffusionFloor = 1e-10;
constexpr double kProductionLimitRatio = 10.0;
constexpr double kEddyViscosityFloor = 1e-30;

// safe_sqrt: zero value and zero derivative at a non-positive argument.
template <class S>
AFX_DEV S safe_sqrt(S x) {
  return val(x) > 0.0 ? dsqrt(x) : sconst<S>(0.0);
}

template <class S>
AFX_DEV S kpos(S k) {
  return dmax(k, 0.0);
}

template <class S>
AFX_DEV S sst_blend(S f1, double inner, double outer) {
  return f1 * inner + (1.0 - f1) * outer;
}

// The wall distance d has its own number type DT: dual when the node coordinates carry a tangent.
template <class S>
AFX_DEV S lift(double x) {
  return sconst<S>(x);
}
template <class S>
AFX_DEV S lift(dual x) {
  return x;
}

template <class S, class DT = double>
AFX_DEV S wall_y_star(const SstConst& c, S nu, DT d, S k) {
  return c.bsq * safe_sqrt(kpos(k)) * d / nu;
}

template <class S, class DT = double>
AFX_DEV S wall_function_weight(const SstConst& c, S nu, DT d, S k) {
  S r = wall_y_star(c, nu, d, k) / c.ylam;
  S r2 = r * r;
  return dtanh(r2 * r2);
}

template <class S, class DT = double>
AFX_DEV S log_layer_shear_rate(const SstConst& c, DT d, S k) {
  return c.bsq * safe_sqrt(kpos(k)) / (c.kappa * d);
}

template <class S, class DT = double>
AFX_DEV S nut_wall(const SstConst& c, S nu, DT d, S k) {
  S ys = wall_y_star(c, nu, d, k);
  S arg = c.e_wall * dmax(ys, c.ylam);
  S r = ys * c.kappa / dlog(arg) - 1.0;
  return val(ys) > c.ylam ? nu * r : sconst<S>(0.0);
}

template <class S>
AFX_DEV S pow_p(S x, double p) {
  if (p == 2.0) return x * x;
  return dexp(p * dlog(x));
}

// omega_wall and its partial derivatives d/dd, d/dk (the imposed wall gradient of omega).
template <class S>
AFX_DEV S omega_wall(const SstConst& c, S nu, S d, S k, S* dodd = nullptr, S* dodk = nullptr) {
  const double p = c.p_exp;
  S a = c.c_vis * (6.0 * nu / (c.beta1 * (d * d)));
  S sk = safe_sqrt(kpos(k));
  S b = sk / (c.bsq * c.kappa * d);
  S m = dmax(a, b);
  S ra = a / m, rb = b / m;
  S bt = val(rb) > 0.0 ? pow_p(rb, p) : sconst<S>(0.0);
  S T = pow_p(ra, p) + bt;
  S om = p == 2.0 ? m * dsqrt(T) : m * dexp(dlog(T) / p);
  if (dodd) {
    // (a/om)^(p-1) da + (b/om)^(p-1) db, with da/dd = -2a/d, db/dd = -b/d, db/dk = b/(2k) (k > 0)
    S wa = p == 2.0 ? a / om : dexp((p - 1.0) * dlog(a / om));
    S wb = val(b) > 0.0 ? (p == 2.0 ? b / om : dexp((p - 1.0) * dlog(b / om))) : sconst<S>(0.0);
    *dodd = wa * (-2.0 * a / d) + wb * (-b / d);
    S dbdk = val(k) > 0.0 ? 0.5 / (sk * (c.bsq * c.kappa * d)) : sconst<S>(0.0);
    *dodk = wb * dbdk;
  }
  return om;
}

template <class S, class DT = double>
AFX_DEV S omega_wall_k(const SstConst& c, S nu, DT d, S k) {
  return omega_wall<S>(c, nu, lift<S>(d), k);
}

// Outputs of the cell-local closure.
template <class S>
struct SstClosure {
  S nut, f1, strain, cd, nutw;
};

// Strain-rate magnitude sqrt(2 S_ij S_ij) of a velocity gradient (strain.py:46-66).
template <int D, class S>
AFX_DEV S strain_magnitude(const S (&Gu)[D][D]) {
  S ss = sconst<S>(0.0);
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      S sij = 0.5 * (Gu[i][j] + Gu[j][i]);
      ss = ss + sij * sij;
    }
  return safe_sqrt(2.0 * ss);
}

// Closure of one cell given the raw strain magnitude and grad k . grad omega in factored form:
// away from walls kw = A (= grad k . grad omega); at a wall cell the omega gradient is the imposed
// d omega_wall/dd grad d + d omega_wall/dk grad k, so kw = dodd A + dodk B with A = grad k . grad d,
// B = |grad k|^2.  (transport.py:369-546, sst.py:125-239)
template <class S, class DT = double>
AFX_DEV SstClosure<S> sst_closure_core(const SstConst& c, S k, S w, S strain0, S A, S B, S nu, DT d,
                                       bool wall) {
  SstClosure<S> o;
  S strain = strain0;
  S kw = A;
  if (wall) {
    S wgt = wall_function_weight(c, nu, d, k);
    strain = (1.0 - wgt) * strain0 + wgt * log_layer_shear_rate(c, d, k);
    S dodd, dodk;
    omega_wall<S>(c, nu, lift<S>(d), k, &dodd, &dodk);
    kw = dodd * A + dodk * B;
  }
  // F1 (sst.py:152-195)
  S cdk = dmax(2.0 * c.sw2 * kw / w, kCrossDiffusionFloor);
  S pk = kpos(k);
  S sq = safe_sqrt(pk);
  S arg1 = dmin(dmax(sq / (c.beta_star * w * d), 500.0 * nu / ((d * d) * w)), 4.0 * c.sw2 * pk / (cdk * (d * d)));
  S a2 = arg1 * arg1;
  o.f1 = dtanh(a2 * a2);
  // F2 and eddy viscosity (sst.py:125-150, 197-239)
  S arg2 = dmax(2.0 * sq / (c.beta_star * w * d), 500.0 * nu / ((d * d) * w));
  S f2 = dtanh(arg2 * arg2);
  o.nut = c.a1 * pk / dmax(c.a1 * w, strain * f2);
  o.strain = strain;
  o.cd = 2.0 * (1.0 - o.f1) * c.sw2 * kw / w;
  o.nutw = wall ? nut_wall(c, nu, d, k) : sconst<S>(0.0);
  return o;
}

// The factored inputs (A, B) of sst_closure_core.
template <int D, class S, class DT = double>
AFX_DEV void closure_products(const S (&Gk)[D], const S (&Gw_rec)[D], const DT (&dgrad)[D], bool wall, S& A,
                              S& B) {
  if (wall) {
    A = Gk[0] * dgrad[0];
    B = Gk[0] * Gk[0];
#pragma unroll
    for (int j = 1; j < D; ++j) {
      A = A + Gk[j] * dgrad[j];
      B = B + Gk[j] * Gk[j];
    }
  } else {
    A = Gk[0] * Gw_rec[0];
#pragma unroll
    for (int j = 1; j < D; ++j) A = A + Gk[j] * Gw_rec[j];
    B = sconst<S>(0.0);
  }
}

// Closure fields of one cell (transport.py:369-546): strain rate (with the wall log-layer blend),
// F1 from the (wall-imposed) omega gradient, eddy viscosity with the F2 limiter, the cross-diffusion
// value 2 (1 - F1) sigma_w2 grad k . grad w / w, and the wall-function eddy viscosity of its wall faces.
template <int D, class S, class DT = double>
AFX_DEV SstClosure<S> sst_closure(const SstConst& c, S k, S w, const S (&Gu)[D][D], const S (&Gk)[D],
                                  const S (&Gw_rec)[D], S nu, DT d, const DT (&dgrad)[D], bool wall) {
  S A, B;
  closure_products<D, S, DT>(Gk, Gw_rec, dgrad, wall, A, B);
  return sst_closure_core<S, DT>(c, k, w, strain_magnitude<D, S>(Gu), A, B, nu, d, wall);
}

// Volume-integrated sources of one cell: out_k = KProduction + KDestruction, out_w = OmegaProduction
// + OmegaDestruction + OmegaCrossDiffusion (each production positive; the residual subtracts them
// in this order).  ``tau`` is the wall shear rate of a wall cell.
template <class S, class VT = double, class DT = double>
AFX_DEV void sst_sources(const SstConst& c, S k, S w, S nut, S strain, S f1, S cd, S tau, S nu, DT d,
                         bool wall, VT V, S& prod_k, S& dest_k, S& prod_w, S& dest_w, S& cross_w) {
  S limited = dmin(nut * (strain * strain), kProductionLimitRatio * c.beta_star * kpos(k) * w);
  S wdis = w;
  if (wall) {
    S lw = wall_function_weight(c, nu, d, k);
    S wv = (nu + nut_wall(c, nu, d, k)) * tau * log_layer_shear_rate(c, d, k);
    limited = (1.0 - lw) * limited + lw * wv;
    wdis = omega_wall_k(c, nu, d, k);
  }
  prod_k = limited * V;
  dest_k = -c.beta_star * k * wdis * V;
  S alpha = sst_blend(f1, c.alpha1, c.alpha2);
  S cap = kProductionLimitRatio * c.beta_star * kpos(k) * w / dmax(nut, kEddyViscosityFloor);
  prod_w = alpha * dmin(strain * strain, cap) * V;
  S beta = sst_blend(f1, c.beta1, c.beta2);
  dest_w = -beta * (w * w) * V;
  cross_w = cd * V;
}

// d min(a, b) / d a with the JAX tie convention (half on a tie).
AFX_DEV double dmin_da(double a, double b) { return a < b ? 1.0 : (a > b ? 0.0 : 0.5); }

// Hand-written adjoint of sst_sources away from walls, for cotangents lk of (prod_k + dest_k) and
// lw of (prod_w + dest_w + cross_w); accumulates into the k, omega, nu_t, strain, F1 and
// cross-diffusion cotangents.
AFX_DEV void adj_sst_sources_open(const SstConst& c, double k, double w, double nut, double strain, double f1,
                                  double V, double lk, double lw, double& k_b, double& w_b, double& nut_b,
                                  double& strain_b, double& f1_b, double& cd_b) {
  const double pk = k > 0.0 ? k : 0.0;
  const double dpk = k > 0.0 ? 1.0 : (k < 0.0 ? 0.0 : 0.5);
  const double s2 = strain * strain;
  const double r = kProductionLimitRatio * c.beta_star;
  double pk_b = 0.0;
  // prod_k = min(nut S^2, r k+ w) V
  {
    const double a = nut * s2, b = r * pk * w;
    const double lb = lk * V;
    const double sa = dmin_da(a, b);
    const double ab = lb * sa, bb = lb * (1.0 - sa);
    nut_b += ab * s2;
    strain_b += ab * 2.0 * nut * strain;
    pk_b += bb * r * w;
    w_b += bb * r * pk;
  }
  // dest_k = -beta* k w V
  k_b += lk * (-c.beta_star * w * V);
  w_b += lk * (-c.beta_star * k * V);
  // prod_w = alpha(F1) min(S^2, cap) V,  cap = r k+ w / max(nut, floor)
  {
    const double alpha = sst_blend(f1, c.alpha1, c.alpha2);
    const double den = nut > kEddyViscosityFloor ? nut : kEddyViscosityFloor;
    const double cap = r * pk * w / den;
    const double m2 = s2 < cap ? s2 : cap;
    f1_b += lw * m2 * V * (c.alpha1 - c.alpha2);
    const double m2b = lw * alpha * V;
    const double ss = dmin_da(s2, cap);
    strain_b += m2b * ss * 2.0 * strain;
    const double capb = m2b * (1.0 - ss);
    const double numb = capb / den;
    const double denb = -capb * cap / den;
    nut_b += denb * (nut > kEddyViscosityFloor ? 1.0 : (nut < kEddyViscosityFloor ? 0.0 : 0.5));
    pk_b += numb * r * w;
    w_b += numb * r * pk;
  }
  // dest_w = -beta(F1) w^2 V
  {
    const double beta = sst_blend(f1, c.beta1, c.beta2);
    f1_b += lw * (-(w * w) * V) * (c.beta1 - c.beta2);
    w_b += lw * (-beta * 2.0 * w * V);
  }
  // cross_w = cd V
  cd_b += lw * V;
  k_b += pk_b * dpk;
}

}  // namespace afx
