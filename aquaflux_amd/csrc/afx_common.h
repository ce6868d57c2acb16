// Shared device/host helpers for the aquaflux MI355X kernels (gfx950, wave64).
//
// * `dual`  -- a value/tangent pair.  The forward kernels are templates over the scalar type, so
//   instantiating them with `dual` gives the exact tangent (JVP) of the same arithmetic.  The
//   non-smooth primitives follow JAX's conventions, because the reference differentiates with JAX:
//   `maximum`/`minimum` split the tangent 1/2-1/2 on a tie, `where` sends it to the taken branch,
//   `abs` has derivative sign(x) (0 at 0).
// * `In<S>` / `Out<S>` -- structure-of-arrays views: a value pointer plus an optional tangent
//   pointer (null = zero tangent).
// * error plumbing -- every C entry point returns an int status; the message of the last failure
//   on the calling thread is kept for `afx_last_error`.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include <string>

#ifndef AFX_DEV
#define AFX_DEV __host__ __device__ __forceinline__
#endif
#define AFX_HD __host__ __device__ __forceinline__

namespace afx {

// ---------------------------------------------------------------------------------------------
// dual numbers
struct dual {
  double v, d;
};

AFX_HD dual mkd(double v, double d) { return dual{v, d}; }
AFX_HD double val(double x) { return x; }
AFX_HD double val(const dual& x) { return x.v; }
AFX_HD double tan_of(double) { return 0.0; }
AFX_HD double tan_of(const dual& x) { return x.d; }

AFX_HD dual operator+(dual a, dual b) { return {a.v + b.v, a.d + b.d}; }
AFX_HD dual operator-(dual a, dual b) { return {a.v - b.v, a.d - b.d}; }
AFX_HD dual operator*(dual a, dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
AFX_HD dual operator/(dual a, dual b) {
  double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
AFX_HD dual operator-(dual a) { return {-a.v, -a.d}; }
AFX_HD dual operator+(dual a, double b) { return {a.v + b, a.d}; }
AFX_HD dual operator+(double a, dual b) { return {a + b.v, b.d}; }
AFX_HD dual operator-(dual a, double b) { return {a.v - b, a.d}; }
AFX_HD dual operator-(double a, dual b) { return {a - b.v, -b.d}; }
AFX_HD dual operator*(dual a, double b) { return {a.v * b, a.d * b}; }
AFX_HD dual operator*(double a, dual b) { return {a * b.v, a * b.d}; }
AFX_HD dual operator/(dual a, double b) { return {a.v / b, a.d / b}; }
AFX_HD dual operator/(double a, dual b) {
  double q = a / b.v;
  return {q, -q * b.d / b.v};
}
AFX_HD dual& operator+=(dual& a, dual b) { a.v += b.v; a.d += b.d; return a; }
AFX_HD dual& operator-=(dual& a, dual b) { a.v -= b.v; a.d -= b.d; return a; }

AFX_HD double dsqrt(double x) { return sqrt(x); }
AFX_HD dual dsqrt(dual x) {
  double s = sqrt(x.v);
  return {s, x.d * 0.5 / s};
}
AFX_HD double dtanh(double x) { return tanh(x); }
AFX_HD dual dtanh(dual x) {
  double t = tanh(x.v);
  return {t, x.d * (1.0 - t * t)};
}
AFX_HD double dlog(double x) { return log(x); }
AFX_HD dual dlog(dual x) { return {log(x.v), x.d / x.v}; }
AFX_HD double dexp(double x) { return exp(x); }
AFX_HD dual dexp(dual x) {
  double e = exp(x.v);
  return {e, x.d * e};
}
AFX_HD double dabs(double x) { return fabs(x); }
AFX_HD dual dabs(dual x) {
  double s = x.v > 0.0 ? 1.0 : (x.v < 0.0 ? -1.0 : 0.0);
  return {fabs(x.v), s * x.d};
}

// JAX `maximum` / `minimum`: tangent split evenly on a tie.
AFX_HD double dmax(double a, double b) { return a > b ? a : b; }
AFX_HD double dmin(double a, double b) { return a < b ? a : b; }
AFX_HD dual dmax(dual a, dual b) {
  if (a.v > b.v) return a;
  if (a.v < b.v) return b;
  return {a.v, 0.5 * (a.d + b.d)};
}
AFX_HD dual dmin(dual a, dual b) {
  if (a.v < b.v) return a;
  if (a.v > b.v) return b;
  return {a.v, 0.5 * (a.d + b.d)};
}
AFX_HD dual dmax(dual a, double b) { return dmax(a, dual{b, 0.0}); }
AFX_HD dual dmin(dual a, double b) { return dmin(a, dual{b, 0.0}); }
AFX_HD dual dmax(double a, dual b) { return dmax(dual{a, 0.0}, b); }
AFX_HD dual dmin(double a, dual b) { return dmin(dual{a, 0.0}, b); }

// d max(x, 0) / dx with the tie convention (used by the hand-written adjoints).
AFX_HD double dmax0_dx(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? 0.0 : 0.5); }

template <class S>
AFX_HD S sconst(double v) { return S(v); }
template <>
AFX_HD dual sconst<dual>(double v) { return dual{v, 0.0}; }

// Multi-tangent dual number: one value and N directional derivatives, so a cell-local function is
// evaluated once for N seeds (each transcendental once).  Every tangent component follows exactly
// the operations `dual` applies to its single tangent.
template <int N>
struct dualn {
  double v;
  double d[N];
  AFX_HD dualn() {}
  AFX_HD dualn(double x) : v(x) {
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = 0.0;
  }
};
template <int N>
AFX_HD dualn<N> seeded(double x, int s) {
  dualn<N> r;
  r.v = x;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = i == s ? 1.0 : 0.0;
  return r;
}
template <int N>
AFX_HD double val(const dualn<N>& x) { return x.v; }
#define AFX_DN_MAP(expr_v, expr_d)      \
  dualn<N> r;                            \
  r.v = (expr_v);                        \
  _Pragma("unroll") for (int i = 0; i < N; ++i) r.d[i] = (expr_d); \
  return r
template <int N> AFX_HD dualn<N> operator+(const dualn<N>& a, const dualn<N>& b) { AFX_DN_MAP(a.v + b.v, a.d[i] + b.d[i]); }
template <int N> AFX_HD dualn<N> operator-(const dualn<N>& a, const dualn<N>& b) { AFX_DN_MAP(a.v - b.v, a.d[i] - b.d[i]); }
template <int N> AFX_HD dualn<N> operator*(const dualn<N>& a, const dualn<N>& b) {
  AFX_DN_MAP(a.v * b.v, a.d[i] * b.v + a.v * b.d[i]);
}
template <int N> AFX_HD dualn<N> operator/(const dualn<N>& a, const dualn<N>& b) {
  const double q = a.v / b.v;
  AFX_DN_MAP(q, (a.d[i] - q * b.d[i]) / b.v);
}
template <int N> AFX_HD dualn<N> operator-(const dualn<N>& a) { AFX_DN_MAP(-a.v, -a.d[i]); }
template <int N> AFX_HD dualn<N> operator+(const dualn<N>& a, double b) { AFX_DN_MAP(a.v + b, a.d[i]); }
template <int N> AFX_HD dualn<N> operator+(double a, const dualn<N>& b) { AFX_DN_MAP(a + b.v, b.d[i]); }
template <int N> AFX_HD dualn<N> operator-(const dualn<N>& a, double b) { AFX_DN_MAP(a.v - b, a.d[i]); }
template <int N> AFX_HD dualn<N> operator-(double a, const dualn<N>& b) { AFX_DN_MAP(a - b.v, -b.d[i]); }
template <int N> AFX_HD dualn<N> operator*(const dualn<N>& a, double b) { AFX_DN_MAP(a.v * b, a.d[i] * b); }
template <int N> AFX_HD dualn<N> operator*(double a, const dualn<N>& b) { AFX_DN_MAP(a * b.v, a * b.d[i]); }
template <int N> AFX_HD dualn<N> operator/(const dualn<N>& a, double b) { AFX_DN_MAP(a.v / b, a.d[i] / b); }
template <int N> AFX_HD dualn<N> operator/(double a, const dualn<N>& b) {
  const double q = a / b.v;
  AFX_DN_MAP(q, -q * b.d[i] / b.v);
}
template <int N> AFX_HD dualn<N> dsqrt(const dualn<N>& x) {
  const double s = sqrt(x.v);
  AFX_DN_MAP(s, x.d[i] * 0.5 / s);
}
template <int N> AFX_HD dualn<N> dtanh(const dualn<N>& x) {
  const double t = tanh(x.v);
  AFX_DN_MAP(t, x.d[i] * (1.0 - t * t));
}
template <int N> AFX_HD dualn<N> dlog(const dualn<N>& x) { AFX_DN_MAP(log(x.v), x.d[i] / x.v); }
template <int N> AFX_HD dualn<N> dexp(const dualn<N>& x) {
  const double e = exp(x.v);
  AFX_DN_MAP(e, x.d[i] * e);
}
template <int N> AFX_HD dualn<N> dmax(const dualn<N>& a, const dualn<N>& b) {
  if (a.v > b.v) return a;
  if (a.v < b.v) return b;
  AFX_DN_MAP(a.v, 0.5 * (a.d[i] + b.d[i]));
}
template <int N> AFX_HD dualn<N> dmin(const dualn<N>& a, const dualn<N>& b) {
  if (a.v < b.v) return a;
  if (a.v > b.v) return b;
  AFX_DN_MAP(a.v, 0.5 * (a.d[i] + b.d[i]));
}
template <int N> AFX_HD dualn<N> dmax(const dualn<N>& a, double b) { return dmax(a, dualn<N>(b)); }
template <int N> AFX_HD dualn<N> dmin(const dualn<N>& a, double b) { return dmin(a, dualn<N>(b)); }
template <int N> AFX_HD dualn<N> dmax(double a, const dualn<N>& b) { return dmax(dualn<N>(a), b); }
template <int N> AFX_HD dualn<N> dmin(double a, const dualn<N>& b) { return dmin(dualn<N>(a), b); }
#undef AFX_DN_MAP

// ---------------------------------------------------------------------------------------------
// structure-of-arrays views
template <class S>
struct In;
template <>
struct In<double> {
  const double* v = nullptr;
  const double* t = nullptr;
  AFX_DEV double operator[](int64_t i) const { return v[i]; }
};
template <>
struct In<dual> {
  const double* v = nullptr;
  const double* t = nullptr;
  AFX_DEV dual operator[](int64_t i) const { return dual{v[i], t ? t[i] : 0.0}; }
};

template <class S>
struct Out;
// Per-cell outputs are written once and read back only by a later kernel, after the L2 has cycled
// through other data: non-temporal stores leave the L2 to the neighbour gathers of the running
// kernel (about 1% on the SST channel step).
template <>
struct Out<double> {
  double* v = nullptr;
  double* t = nullptr;
  AFX_DEV void set(int64_t i, double x) const { __builtin_nontemporal_store(x, &v[i]); }
};
template <>
struct Out<dual> {
  double* v = nullptr;
  double* t = nullptr;
  AFX_DEV void set(int64_t i, dual x) const {
    __builtin_nontemporal_store(x.v, &v[i]);
    __builtin_nontemporal_store(x.d, &t[i]);
  }
};

// Read-only geometry/topology of a mesh, as device pointers (all SoA: vec[j * n + i]).
struct MeshView {
  int32_t C, F;
  const int32_t* owner;
  const int32_t* nb;     // -1 on boundary faces
  const int32_t* cptr;   // [C+1] cell -> face CSR
  const int32_t* cface;  // f (cell owns f) or ~f (cell is f's neighbour)
  const int32_t* cnbr;   // the other cell of each cface entry (-1 on a boundary face)
  const double* area;    // [F]
  const double* nrm;     // [d][F] owner-outward unit normal
  const double* fc;      // [d][F] face centroid
  const double* vol;     // [C]
  const double* cc;      // [d][C] cell centroid
  const int32_t* fu;     // [F] kernel face index -> caller face index (nullptr: the same)
  const int32_t* bsched; // [nbs] block order of the cell kernels (nullptr: memory order)
  int32_t nbs;           // number of 256-cell blocks bsched orders
};

// Per-face geometry seen from a face's owner (P) and neighbour (N); G is double, or `dual` when a
// geometry tangent is carried (the node-coordinate JVP).
template <int D, class G = double>
struct FaceGeo {
  G A, n[D], xf[D], xP[D], xN[D];
  G g;      // interpolation factor (0 on boundary faces)
  G dvec[D];  // xN - xP (0 on boundary faces)
  G dist;   // normal distance: d.n (interior) or (xf - xP).n (boundary)
  G skew[D];  // xf - (xP + g d)
  G dp[D], dn[D];  // xf - xP, xf - xN
  G dpn, dnn;
};

template <int D>
AFX_DEV double vdot(const double* a, const double* b) {
  double s = a[0] * b[0];
#pragma unroll
  for (int j = 1; j < D; ++j) s = s + a[j] * b[j];
  return s;
}

template <int D, class G>
AFX_DEV G vdotg(const G* a, const G* b) {
  G s = a[0] * b[0];
#pragma unroll
  for (int j = 1; j < D; ++j) s = s + a[j] * b[j];
  return s;
}

// The derived quantities of a FaceGeo from its A, n, xf, xP and xN (xN = xP on a boundary face), in
// load_face_geo's order of operations.
template <int D, class G>
AFX_DEV void finish_face_geo(bool interior, FaceGeo<D, G>& q) {
#pragma unroll
  for (int j = 0; j < D; ++j) {
    q.dvec[j] = q.xN[j] - q.xP[j];
    q.dp[j] = q.xf[j] - q.xP[j];
    q.dn[j] = q.xf[j] - q.xN[j];
  }
  q.dpn = vdotg<D, G>(q.dp, q.n);
  q.dnn = vdotg<D, G>(q.dn, q.n);
  if (interior) {
    G dd = vdotg<D, G>(q.dvec, q.dvec);
    q.g = vdotg<D, G>(q.dp, q.dvec) / dd;
    q.dist = vdotg<D, G>(q.dvec, q.n);
  } else {
    q.g = sconst<G>(0.0);
    q.dist = q.dpn;
  }
#pragma unroll
  for (int j = 0; j < D; ++j) q.skew[j] = q.xf[j] - (q.xP[j] + q.dvec[j] * q.g);
}

// Loads the geometry of face f; `P`/`N` are its owner and neighbour (N < 0 on a boundary face).
template <int D>
AFX_DEV void load_face_geo(const MeshView& m, int32_t f, int32_t P, int32_t N, FaceGeo<D>& q) {
  q.A = m.area[f];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    q.n[j] = m.nrm[(int64_t)j * m.F + f];
    q.xf[j] = m.fc[(int64_t)j * m.F + f];
    q.xP[j] = m.cc[(int64_t)j * m.C + P];
    q.xN[j] = N >= 0 ? m.cc[(int64_t)j * m.C + N] : q.xP[j];
    q.dvec[j] = q.xN[j] - q.xP[j];
    q.dp[j] = q.xf[j] - q.xP[j];
    q.dn[j] = q.xf[j] - q.xN[j];
  }
  q.dpn = vdot<D>(q.dp, q.n);
  q.dnn = vdot<D>(q.dn, q.n);
  if (N >= 0) {
    double dd = vdot<D>(q.dvec, q.dvec);
    q.g = vdot<D>(q.dp, q.dvec) / dd;
    q.dist = vdot<D>(q.dvec, q.n);
  } else {
    q.g = 0.0;
    q.dist = q.dpn;
  }
#pragma unroll
  for (int j = 0; j < D; ++j) q.skew[j] = q.xf[j] - (q.xP[j] + q.dvec[j] * q.g);
}

// As load_face_geo, with the owner / neighbour centroids given (e.g. from LDS).
template <int D>
AFX_DEV void load_face_geo_pts(const MeshView& m, int32_t f, bool interior, const double* xP, const double* xN,
                               FaceGeo<D>& q) {
  q.A = m.area[f];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    q.n[j] = m.nrm[(int64_t)j * m.F + f];
    q.xf[j] = m.fc[(int64_t)j * m.F + f];
    q.xP[j] = xP[j];
    q.xN[j] = interior ? xN[j] : q.xP[j];
    q.dvec[j] = q.xN[j] - q.xP[j];
    q.dp[j] = q.xf[j] - q.xP[j];
    q.dn[j] = q.xf[j] - q.xN[j];
  }
  q.dpn = vdot<D>(q.dp, q.n);
  q.dnn = vdot<D>(q.dn, q.n);
  if (interior) {
    double dd = vdot<D>(q.dvec, q.dvec);
    q.g = vdot<D>(q.dp, q.dvec) / dd;
    q.dist = vdot<D>(q.dvec, q.n);
  } else {
    q.g = 0.0;
    q.dist = q.dpn;
  }
#pragma unroll
  for (int j = 0; j < D; ++j) q.skew[j] = q.xf[j] - (q.xP[j] + q.dvec[j] * q.g);
}

// Geometry tangents in the view's face order: face area / normal / centroid [F] / [d][F] and cell
// volume / centroid [C] / [d][C].
struct GeoTan {
  const double *area, *nrm, *fc, *vol, *cc;
};

template <int D>
AFX_DEV void cell_centroid_g(const MeshView& m, const GeoTan& t, int32_t c, dual (&x)[D]) {
#pragma unroll
  for (int j = 0; j < D; ++j) x[j] = dual{m.cc[(int64_t)j * m.C + c], t.cc[(int64_t)j * m.C + c]};
}

template <int D>
AFX_DEV void load_face_geo_g(const MeshView& m, const GeoTan& t, int32_t f, bool interior, const dual* xP,
                             const dual* xN, FaceGeo<D, dual>& q) {
  q.A = dual{m.area[f], t.area[f]};
#pragma unroll
  for (int j = 0; j < D; ++j) {
    q.n[j] = dual{m.nrm[(int64_t)j * m.F + f], t.nrm[(int64_t)j * m.F + f]};
    q.xf[j] = dual{m.fc[(int64_t)j * m.F + f], t.fc[(int64_t)j * m.F + f]};
    q.xP[j] = xP[j];
    q.xN[j] = interior ? xN[j] : xP[j];
  }
  finish_face_geo<D, dual>(interior, q);
}

// Reads the view's geometry as FP64 values (GEO = false) or as (value, tangent) pairs with the
// tangents of `t` (GEO = true): one kernel body serves the residual and its node-coordinate JVP.
template <int D, bool GEO>
struct GeoRead;

template <int D>
struct GeoRead<D, false> {
  using T = double;
  AFX_DEV static double vol(const MeshView& m, const GeoTan&, int32_t c) { return m.vol[c]; }
  AFX_DEV static double cc(const MeshView& m, const GeoTan&, int j, int32_t c) { return m.cc[(int64_t)j * m.C + c]; }
  AFX_DEV static double fc(const MeshView& m, const GeoTan&, int j, int32_t f) { return m.fc[(int64_t)j * m.F + f]; }
  AFX_DEV static double nrm(const MeshView& m, const GeoTan&, int j, int32_t f) { return m.nrm[(int64_t)j * m.F + f]; }
  AFX_DEV static double area(const MeshView& m, const GeoTan&, int32_t f) { return m.area[f]; }
  AFX_DEV static void face(const MeshView& m, const GeoTan&, int32_t f, int32_t P, int32_t N, FaceGeo<D>& q) {
    load_face_geo<D>(m, f, P, N, q);
  }
  AFX_DEV static void face_pts(const MeshView& m, const GeoTan&, int32_t f, bool interior, const double* xP,
                               const double* xN, FaceGeo<D>& q) {
    load_face_geo_pts<D>(m, f, interior, xP, xN, q);
  }
};

template <int D>
struct GeoRead<D, true> {
  using T = dual;
  AFX_DEV static dual vol(const MeshView& m, const GeoTan& t, int32_t c) { return dual{m.vol[c], t.vol[c]}; }
  AFX_DEV static dual cc(const MeshView& m, const GeoTan& t, int j, int32_t c) {
    const int64_t i = (int64_t)j * m.C + c;
    return dual{m.cc[i], t.cc[i]};
  }
  AFX_DEV static dual fc(const MeshView& m, const GeoTan& t, int j, int32_t f) {
    const int64_t i = (int64_t)j * m.F + f;
    return dual{m.fc[i], t.fc[i]};
  }
  AFX_DEV static dual nrm(const MeshView& m, const GeoTan& t, int j, int32_t f) {
    const int64_t i = (int64_t)j * m.F + f;
    return dual{m.nrm[i], t.nrm[i]};
  }
  AFX_DEV static dual area(const MeshView& m, const GeoTan& t, int32_t f) { return dual{m.area[f], t.area[f]}; }
  AFX_DEV static void face(const MeshView& m, const GeoTan& t, int32_t f, int32_t P, int32_t N, FaceGeo<D, dual>& q) {
    dual xP[D], xN[D];
    cell_centroid_g<D>(m, t, P, xP);
    if (N >= 0) cell_centroid_g<D>(m, t, N, xN);
    load_face_geo_g<D>(m, t, f, N >= 0, xP, N >= 0 ? xN : xP, q);
  }
  AFX_DEV static void face_pts(const MeshView& m, const GeoTan& t, int32_t f, bool interior, const dual* xP,
                               const dual* xN, FaceGeo<D, dual>& q) {
    load_face_geo_g<D>(m, t, f, interior, xP, xN, q);
  }
};

// Least-squares fit direction cotangent of one cell's fit (all variables, shared M):
// dbar = sum_v w (r_v y_v - (y_v . d) G_v) - 2 w^2 (y_v . d) r_v d, r_v = dphi_v - G_v . d.
template <int D, int NV>
AFX_DEV void lsq_dir_bar(const double (&d)[D], const double (&dphi)[NV], const double (&G)[NV][D],
                         const double (&y)[NV][D], double (&db)[D]) {
  double dd = 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) dd += d[j] * d[j];
  const double w = 1.0 / dd;
#pragma unroll
  for (int j = 0; j < D; ++j) db[j] = 0.0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double gd = 0.0, yd = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      gd += G[v][j] * d[j];
      yd += y[v][j] * d[j];
    }
    const double r = dphi[v] - gd;
#pragma unroll
    for (int j = 0; j < D; ++j) db[j] += w * (r * y[v][j] - yd * G[v][j]) - 2.0 * w * w * yd * r * d[j];
  }
}

// Decodes one entry of the cell -> face list.
// `f` indexes the view's face arrays (geometry, per-face workspace); `fu` indexes the caller's
// per-face arrays (boundary kinds and values, face fluxes and their cotangents).
struct FaceRef {
  int32_t f, P, N;  // face, owner, neighbour (-1 boundary)
  int32_t fu;       // the caller's index of face f
  bool own;         // the visiting cell is the owner
};
AFX_DEV int32_t user_face(const MeshView& m, int32_t f) { return m.fu ? m.fu[f] : f; }

// Entry k of cell c's face list; the other cell comes from the list's own neighbour array, so no
// owner / neighbour lookup stands between the entry and the neighbour's data.
AFX_DEV FaceRef decode_face(const MeshView& m, int32_t k, int32_t c) {
  FaceRef r;
  const int32_t e = m.cface[k], o = m.cnbr[k];
  r.own = e >= 0;
  r.f = r.own ? e : ~e;
  r.P = r.own ? c : o;
  r.N = r.own ? o : c;
  r.fu = m.fu ? m.fu[r.f] : r.f;
  return r;
}

// Walks cell c's face list with the next entry's (face, other cell) pair loaded one step ahead, so
// that its latency overlaps the current face's work instead of heading the next face's chain of
// dependent loads.  The prefetch is unconditional (the last entry is re-read at the end), so it
// never sits behind a branch.
struct FaceWalk {
  int32_t k, kb, ke, e, o;
  AFX_DEV FaceWalk(const MeshView& m, int32_t c) : k(m.cptr[c]), kb(k), ke(m.cptr[c + 1]) { fetch(m); }
  AFX_DEV void fetch(const MeshView& m) {
    const int32_t kk = k < ke ? k : ke - 1;
    e = m.cface[kk];
    o = m.cnbr[kk];
  }
  AFX_DEV bool more() const { return k < ke; }
  // the current entry (its position in the cell's list is `k - kb` before the call); advances
  AFX_DEV FaceRef next(const MeshView& m, int32_t c) {
    FaceRef r;
    r.own = e >= 0;
    r.f = r.own ? e : ~e;
    r.P = r.own ? c : o;
    r.N = r.own ? o : c;
    r.fu = m.fu ? m.fu[r.f] : r.f;
    ++k;
    fetch(m);
    return r;
  }
};

// Least-squares normal matrix of cell c (geometry only): M = sum_f w dx dx^T, w = 1/|dx|^2.
template <int D>
AFX_DEV void lsq_matrix(const MeshView& m, int32_t c, double M[D][D]) {
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b) M[a][b] = 0.0;
  double ccc[D];
#pragma unroll
  for (int j = 0; j < D; ++j) ccc[j] = m.cc[(int64_t)j * m.C + c];
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    int32_t o = r.own ? r.N : r.P;
    double dx[D];
#pragma unroll
    for (int j = 0; j < D; ++j) dx[j] = (o >= 0 ? m.cc[(int64_t)j * m.C + o] : m.fc[(int64_t)j * m.F + r.f]) - ccc[j];
    double w = 1.0 / vdot<D>(dx, dx);
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) M[a][b] += w * dx[a] * dx[b];
  }
}

// Inverse of a symmetric positive-definite D x D matrix (adjugate / determinant); T is double, or
// `dual` for its tangent.
template <int D, class T = double>
AFX_DEV void sym_inverse(const T M[D][D], T I[D][D]) {
  if (D == 2) {
    T det = M[0][0] * M[1][1] - M[0][1] * M[1][0];
    I[0][0] = M[1][1] / det;
    I[1][1] = M[0][0] / det;
    I[0][1] = -M[0][1] / det;
    I[1][0] = -M[1][0] / det;
  } else {
    T c00 = M[1][1] * M[2][2] - M[1][2] * M[2][1];
    T c01 = M[1][2] * M[2][0] - M[1][0] * M[2][2];
    T c02 = M[1][0] * M[2][1] - M[1][1] * M[2][0];
    T det = M[0][0] * c00 + M[0][1] * c01 + M[0][2] * c02;
    T inv = 1.0 / det;
    I[0][0] = c00 * inv;
    I[1][0] = c01 * inv;
    I[2][0] = c02 * inv;
    I[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) * inv;
    I[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * inv;
    I[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * inv;
    I[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) * inv;
    I[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * inv;
    I[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * inv;
  }
}

#define AFX_PRAGMA_STR(x) #x
#define AFX_UNROLL(n) _Pragma(AFX_PRAGMA_STR(unroll n))
#ifndef AFX_FACE_UNROLL
#define AFX_FACE_UNROLL 1
#endif

constexpr int kBlock = 256;
constexpr int kNumXcd = 8;

// Global thread index with an XCD-aware block order.  Workgroup b runs on XCD b % 8; the remap
// gives XCD x the contiguous range of logical blocks [x q + min(x, r), ...) for n = 8 q + r
// blocks (a bijection onto [0, n)), so the stencil neighbours of a block's cells -- the next cell
// row, the next plane -- are mostly processed on the same XCD and read once into its L2.  On the
// 128 x 64 x 64 hex RANS residual + adjoint this is 1.13x faster with owner-sorted faces; with
// face-family-major faces (consecutive cells' faces 4096 apart) it was 2x slower, because every
// XCD then streams its own scattered face lines instead of sharing them through the fabric.
AFX_DEV uint32_t xcd_logical_block() {
  const uint32_t n = gridDim.x, b = blockIdx.x;
  const uint32_t q = n / kNumXcd, r = n % kNumXcd, x = b % kNumXcd, t = b / kNumXcd;
  return x * q + (x < r ? x : r) + t;
}

AFX_DEV int32_t xcd_thread_index() { return (int32_t)(xcd_logical_block() * blockDim.x + threadIdx.x); }

// Cell index of a cell kernel launched one thread per cell: as xcd_thread_index(), then, when the
// mesh carries a block order for this grid, the logical block is mapped through it.  The order is
// a permutation of the blocks (each XCD keeps its contiguous share and walks it tile by tile along
// the share's thinnest axis), so every cell is still visited exactly once and results do not change;
// only which neighbours are resident in an XCD's L2 at the same time does.
AFX_DEV int32_t xcd_thread_index(const MeshView& m) {
  uint32_t lb = xcd_logical_block();
  if (m.bsched != nullptr && gridDim.x == (uint32_t)m.nbs) lb = (uint32_t)m.bsched[lb];
  return (int32_t)(lb * blockDim.x + threadIdx.x);
}
inline dim3 grid_for(int64_t n, int block = kBlock) { return dim3((unsigned)((n + block - 1) / block)); }

// LDS staging of a workgroup's own cells.  A cell kernel launched one 256-thread workgroup per
// block of 256 consecutive cells (xcd_thread_index(m) == block_base(m) + threadIdx.x) stages the
// data its cells' neighbours read into LDS; a face whose other cell lies in the same block reads it
// there (tile_slot >= 0), the others from global memory.  Values are copied, not recomputed, so
// results are bitwise those of the unstaged gather.
AFX_DEV int32_t block_base(const MeshView& m) {
  uint32_t lb = xcd_logical_block();
  if (m.bsched != nullptr && gridDim.x == (uint32_t)m.nbs) lb = (uint32_t)m.bsched[lb];
  return (int32_t)(lb * kBlock);
}
AFX_DEV int tile_slot(int32_t o, int32_t base) {
  const uint32_t d = (uint32_t)(o - base);
  return d < (uint32_t)kBlock ? (int)d : -1;
}

// Where the body of a staged cell kernel runs.  DevExec: thread t of the 256-thread workgroup that
// owns the block at `base`; the workgroup stages its cells in LDS, synchronises, and reads a
// neighbour inside the block from there.  HostExec: one cell at a time on the host (the OpenMP CPU
// line), with no workgroup to wait for and every neighbour read from memory (slot() is -1), so the
// staging writes are never read back.  Either way a cell computes the same values.
struct DevExec {
  int32_t base;
  int t;
  __device__ explicit DevExec(const MeshView& m) : base(block_base(m)), t((int)threadIdx.x) {}
  __device__ void sync() const { __syncthreads(); }
  __device__ int slot(int32_t o) const { return tile_slot(o, base); }
};
struct HostExec {
  int32_t base;
  int t;
  explicit HostExec(int32_t c) : base(c - c % kBlock), t(c % kBlock) {}
  void sync() const {}
  int slot(int32_t) const { return -1; }
};

// Cells in XCD-contiguous order over a bounded set of resident workgroups: workgroup b on XCD
// x = b % 8 walks that XCD's logical blocks t = b / 8, b / 8 + g_x, ... (g_x = workgroups on XCD
// x), so the cells in flight form a front advancing through each XCD's contiguous range instead of
// the whole range at once.  With gridDim.x equal to the block count this is xcd_thread_index().
template <class Fn>
AFX_DEV void for_each_cell(const MeshView& m, Fn&& fn) {
  const int32_t n = m.C;
  const uint32_t bs = blockDim.x, nb = ((uint32_t)n + bs - 1) / bs;
  const uint32_t G = gridDim.x, b = blockIdx.x;
  const uint32_t q = nb / kNumXcd, r = nb % kNumXcd, x = b % kNumXcd, w = b / kNumXcd;
  const uint32_t start = x * q + (x < r ? x : r), size = q + (x < r ? 1u : 0u);
  const uint32_t gx = G / kNumXcd + (x < G % kNumXcd ? 1u : 0u);
  const bool ordered = m.bsched != nullptr && nb == (uint32_t)m.nbs;
  for (uint32_t t = w; t < size; t += gx) {
    const uint32_t lb = ordered ? (uint32_t)m.bsched[start + t] : start + t;
    const int32_t c = (int32_t)(lb * bs + threadIdx.x);
    if (c < n) fn(c);
  }
}

// Launch size for for_each_cell: every block when that is at most `resident_per_cu` workgroups per
// compute unit, otherwise that many per compute unit (a multiple of the XCD count).
dim3 cell_grid(int64_t n, int resident_per_cu);

}  // namespace afx

// ---------------------------------------------------------------------------------------------
// host-side error plumbing (one message per calling thread)
void afx_set_error(const std::string& msg);
const char* afx_get_error();

#define AFX_CHECK_HIP(expr)                                                          \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess) {                                                          \
      afx_set_error(std::string(#expr) + ": " + hipGetErrorString(_e));              \
      return 1;                                                                      \
    }                                                                                \
  } while (0)

#define AFX_CHECK_LAUNCH()                                                           \
  do {                                                                               \
    hipError_t _e = hipGetLastError();                                               \
    if (_e != hipSuccess) {                                                          \
      afx_set_error(std::string("kernel launch: ") + hipGetErrorString(_e));          \
      return 1;                                                                      \
    }                                                                                \
  } while (0)

#define AFX_REQUIRE(cond, msg)                                                       \
  do {                                                                               \
    if (!(cond)) {                                                                   \
      afx_set_error(msg);                                                            \
      return 2;                                                                      \
    }                                                                                \
  } while (0)
