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
  int32_t* kmin;  // nullable: kmin[slot[l] * C + c] = position in the cell's face list of the
                  // first face attaining the minimum (the limiter adjoint re-evaluates only it)
};

int venkat_pass(const MeshView& m, int dim, const LimiterSlots& sl, In<double> x, In<double> G, double lim_k,
                Out<double> psi, int32_t* ext, hipStream_t s);
int venkat_pass(const MeshView& m, int dim, const LimiterSlots& sl, In<dual> x, In<dual> G, double lim_k,
                Out<dual> psi, int32_t* ext, hipStream_t s);

}  // namespace afx
