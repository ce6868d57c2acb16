// Cell gradient of a scalar field with Dirichlet / zero-gradient boundary closures (leading-order
// boundary values), by compact Green-Gauss or least squares.
//
// Reference: ResidualAssembler.gradient (discretization/residual.py:253-296) with
// CompactGreenGauss / LeastSquaresGradient (schemes/gradient.py) -- used for the SST wall-distance
// gradient (turbulence/transport.py:86-110).
#include "afx_common.h"
#include "afx_mesh.h"
#include "afx_scalar_face.h"
#include "afx_profile.h"

using namespace afx;

namespace {

AFX_DEV double lead_bval(int8_t kind, double val, double phiP) {
  if (kind == kBcDirichlet) return val;
  if (kind == kBcNone) return 0.0;
  return phiP;
}

template <int D, int GRAD>
__global__ void __launch_bounds__(256) k_scalar_grad(MeshView m, const int8_t* __restrict__ kind,
                                                     const double* __restrict__ bval,
                                                     const double* __restrict__ phi, double* __restrict__ G) {
  const int32_t c = xcd_thread_index();
  if (c >= m.C) return;
  const double xc = phi[c];
  double acc[D];
#pragma unroll
  for (int j = 0; j < D; ++j) acc[j] = 0.0;
  double M[D][D];
#pragma unroll
  for (int p = 0; p < D; ++p)
#pragma unroll
    for (int q = 0; q < D; ++q) M[p][q] = 0.0;
  for (int32_t k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
    FaceRef r = decode_face(m, m.cface[k]);
    const int32_t o = r.own ? r.N : r.P;
    if (o >= 0) {
      const double xo = phi[o];
      if (GRAD == 0) {
        FaceGeo<D> q;
        load_face_geo<D>(m, r.f, r.P, r.N, q);
        const double w = 1.0 - q.g;
        const double pf = w * (r.own ? xc : xo) + q.g * (r.own ? xo : xc);
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const double t = pf * (q.n[j] * q.A);
          acc[j] = r.own ? acc[j] + t : acc[j] - t;
        }
      } else {
        double dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = m.cc[(int64_t)j * m.C + o] - m.cc[(int64_t)j * m.C + c];
        const double wgt = 1.0 / vdot<D>(dx, dx);
#pragma unroll
        for (int p = 0; p < D; ++p)
#pragma unroll
          for (int q2 = 0; q2 < D; ++q2) M[p][q2] += wgt * dx[p] * dx[q2];
        const double dphi = xo - xc;
#pragma unroll
        for (int j = 0; j < D; ++j) acc[j] += (wgt * dx[j]) * dphi;
      }
    } else {
      const double pb = lead_bval(kind[r.f], bval[r.f], xc);
      if (GRAD == 0) {
        const double A = m.area[r.f];
#pragma unroll
        for (int j = 0; j < D; ++j) acc[j] += pb * (m.nrm[(int64_t)j * m.F + r.f] * A);
      } else {
        double dx[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dx[j] = m.fc[(int64_t)j * m.F + r.f] - m.cc[(int64_t)j * m.C + c];
        const double wgt = 1.0 / vdot<D>(dx, dx);
#pragma unroll
        for (int p = 0; p < D; ++p)
#pragma unroll
          for (int q2 = 0; q2 < D; ++q2) M[p][q2] += wgt * dx[p] * dx[q2];
        const double dphi = pb - xc;
#pragma unroll
        for (int j = 0; j < D; ++j) acc[j] += (wgt * dx[j]) * dphi;
      }
    }
  }
  if (GRAD == 0) {
    const double iv = 1.0 / m.vol[c];
#pragma unroll
    for (int j = 0; j < D; ++j) G[(int64_t)j * m.C + c] = acc[j] * iv;
  } else {
    double I[D][D];
    sym_inverse<D>(M, I);
#pragma unroll
    for (int p = 0; p < D; ++p) {
      double s = I[p][0] * acc[0];
#pragma unroll
      for (int q2 = 1; q2 < D; ++q2) s = s + I[p][q2] * acc[q2];
      G[(int64_t)p * m.C + c] = s;
    }
  }
}

}  // namespace

extern "C" int afx_scalar_gradient(const afx_mesh* mh, const afx_geometry* g, int gradient, const int8_t* kind,
                                   const double* value, const double* phi, double* grad, hipStream_t s) {
  AFX_REQUIRE(mh && g && kind && value && phi && grad, "afx_scalar_gradient: null argument");
  AFX_REQUIRE(gradient == 0 || gradient == 1, "afx_scalar_gradient: unknown gradient scheme");
  MeshView m = afx_view(mh, g);
  if (m.C == 0) return 0;
#define AFX_SG(DD, GG)                                                                                  \
  if (mh->dim == DD && gradient == GG) {                                                                \
    AFX_LAUNCH(kKScalarGrad, (k_scalar_grad<DD, GG>), grid_for(m.C), dim3(kBlock), 0, s, m, kind, value, \
               phi, grad);                                                                              \
    AFX_CHECK_LAUNCH();                                                                                 \
    return 0;                                                                                           \
  }
  AFX_SG(2, 0) AFX_SG(2, 1) AFX_SG(3, 0) AFX_SG(3, 1)
#undef AFX_SG
  afx_set_error("afx_scalar_gradient: unsupported dimension");
  return 2;
}
