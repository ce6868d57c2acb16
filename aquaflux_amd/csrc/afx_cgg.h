// Corrected Green-Gauss (skewness-corrected, swept) gradients and a stand-alone Venkatakrishnan
// limiter pass, shared by the flow, RANS and scalar residuals.
//
// Reference: CorrectedGreenGauss.terms / operator / rhs / gradients (schemes/gradient.py:1160-1337)
// with SweptGradientSolve.solve (schemes/gradient.py:618-705) and InverseCellVolume
// (schemes/gradient.py:420-435):  x0 = r V^-1 b (the peeled first sweep), then
// x <- x + r V^-1 (b - A_g x) for sweeps - 1 sweeps, A_g x = V x - sum_f S_f w_f(x) with
// w_f = (1 - g) skew_f . x_P + g skew_f . x_N on interior faces.  b is the compact Green-Gauss sum,
// so the compact gradient kernel (which writes V^-1 b) supplies x0 when r = 1.
#pragma once

#include "afx_common.h"
#include "afx_flow_face.h"

namespace afx {

constexpr int kGradCorrected = 3;

// Gradient rows G[(v * d + j) * C + c] of nv variables: on entry the compact Green-Gauss gradient
// V^-1 b, on exit the swept corrected gradient.  scratch: 2 * nv * d * C doubles.
int cgg_sweeps(const MeshView& m, int dim, int nv, double* G, double* scratch, int sweeps, double relax,
               hipStream_t s);

// Reverse of cgg_sweeps: on entry the cotangent of the swept gradient, on exit the cotangent of the
// Green-Gauss sum b (what the compact gradient's reverse pass takes).  scratch as above.
int cgg_sweeps_adjoint(const MeshView& m, int dim, int nv, double* Gbar, double* scratch, int sweeps, double relax,
                       hipStream_t s);

// Limiter slots: entry l limits state row var[l] with gradient rows var[l] into psi[slot[l] * C + c],
// with ext[slot[l] * C + c] (stencil argmax) and ext[(ext_stride + slot[l]) * C + c] (argmin).
struct LimiterSlots {
  int nl;
  int ext_stride;
  int var[8];
  int slot[8];
  int32_t* kmin;  // nullable: kmin[slot[l] * C + c] = the first face (internal index) attaining the
                  // minimum (the limiter adjoint re-evaluates only it)
};

int venkat_pass(const MeshView& m, int dim, const LimiterSlots& sl, In<double> x, In<double> G, double lim_k,
                Out<double> psi, int32_t* ext, hipStream_t s);
int venkat_pass(const MeshView& m, int dim, const LimiterSlots& sl, In<dual> x, In<dual> G, double lim_k,
                Out<dual> psi, int32_t* ext, hipStream_t s);

// ---------------------------------------------------------------------------------------------
// Node-coordinate derivatives of the swept solve.  With the geometry moving, one sweep
//   x_{k+1} = (1 - r) x_k + r V^-1 (b + corr(x_k)),  corr(x)_P += S_f w_f, corr(x)_N -= S_f w_f,
//   S_f = n_f A_f, w_f = (1 - g_f) skew_f . x_P + g_f skew_f . x_N,  x_0 = r V^-1 b,
// depends on V, S, g and skew as well as on x and b.

// Doubles of iterate storage cgg_sweeps_store needs: x_0 .. x_{sweeps-2}.
inline int64_t cgg_iterate_doubles(int dim, int nv, int64_t C, int sweeps) {
  return sweeps > 1 ? (int64_t)(sweeps - 1) * nv * dim * C : 0;
}

// cgg_sweeps that keeps the iterates x_0 .. x_{sweeps-2} in `iters` (cgg_iterate_doubles), the
// last one in G; the geometry adjoint reads them back.
int cgg_sweeps_store(const MeshView& m, int dim, int nv, double* G, double* iters, int sweeps, double relax,
                     hipStream_t s);

// Geometry cotangents of the swept solve, accumulated over all sweeps.
struct CggGeoBars {
  double* face;  // [2d + 1][F] internal face order: cotangent of S = n A (d), of g (1), of skew (d)
  double* vol;   // [C]
};
inline int64_t cgg_geo_bar_doubles(int dim, int64_t C, int64_t F) { return (int64_t)(2 * dim + 1) * F + C; }

// cgg_sweeps_adjoint that also writes (=) the geometry cotangents of every sweep into `gb`; G and
// iters are what cgg_sweeps_store left.
int cgg_sweeps_adjoint_geometry(const MeshView& m, int dim, int nv, double* Gbar, const double* G,
                                const double* iters, double* scratch, int sweeps, double relax, CggGeoBars gb,
                                hipStream_t s);

// The sweeps on (value, tangent) pairs of the gradient and of the geometry (node-coordinate JVP):
// on entry (G, Gt) hold the compact Green-Gauss gradient and its tangent.  scratch: 4 nv d C doubles.
int cgg_sweeps_tangent(const MeshView& m, const GeoTan& gt, int dim, int nv, double* G, double* Gt,
                       double* scratch, int sweeps, double relax, hipStream_t s);

// The Venkatakrishnan limiter with the geometry (face / cell centroids, volumes) carrying tangents.
int venkat_pass_tangent(const MeshView& m, const GeoTan& gt, int dim, const LimiterSlots& sl, In<dual> x,
                        In<dual> G, double lim_k, Out<dual> psi, hipStream_t s);

// Folds one interior face's accumulated swept-solve cotangents into the cotangents of its area,
// normal, centroid and the two cell centroids (adds).
template <int D>
AFX_DEV void cgg_face_geo_bar(const CggGeoBars& cb, int64_t F, int32_t f, const FaceGeo<D>& q, double& Ab,
                              double (&nb)[D], double (&xfb)[D], double (&xPb)[D], double (&xNb)[D]) {
  FaceGeoBarF<D> qf;
  qf.zero();
  qf.g = cb.face[(int64_t)D * F + f];
#pragma unroll
  for (int j = 0; j < D; ++j) qf.skew[j] = cb.face[(int64_t)(D + 1 + j) * F + f];
  fold_face_geo_bar<D>(q, true, qf);
  Ab += qf.A;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const double Sb = cb.face[(int64_t)j * F + f];
    nb[j] += q.A * Sb + qf.n[j];
    Ab += q.n[j] * Sb;
    xfb[j] += qf.xf[j];
    xPb[j] += qf.xP[j];
    xNb[j] += qf.xN[j];
  }
}

}  // namespace afx
