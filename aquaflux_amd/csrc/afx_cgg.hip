// Swept corrected Green-Gauss gradients (forward and reverse) and a stand-alone Venkatakrishnan
// limiter pass.  One thread per cell, gather over the cell's faces, no atomics.
#include "afx_cgg.h"
#include "afx_mesh.h"
#include "afx_flow_face.h"
#include "afx_profile.h"

namespace afx {
namespace {

// x_out = x + r V^-1 (b - (V x - corr(x))), with b = V x0 / r recovered from the peeled sweep.
template <int D>
__global__ void __launch_bounds__(256) k_cgg_sweep(MeshView m, int nv, const double* __restrict__ x0,
                                                   const double* __restrict__ x, double* __restrict__ xo,
                                                   double relax) {
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  const double V = m.vol[c];
  const double iv = 1.0 / V;
  for (int v = 0; v < nv; ++v) {
    double corr[D];
#pragma unroll
    for (int j = 0; j < D; ++j) corr[j] = 0.0;
    for (FaceWalk walk(m, c); walk.more();) {
      FaceRef r = walk.next(m, c);
      if (r.N < 0) continue;
      FaceGeo<D> q;
      load_face_geo<D>(m, r.f, r.P, r.N, q);
      double sP = 0.0, sN = 0.0;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        sP += q.skew[j] * x[(int64_t)(v * D + j) * m.C + r.P];
        sN += q.skew[j] * x[(int64_t)(v * D + j) * m.C + r.N];
      }
      const double w = (1.0 - q.g) * sP + q.g * sN;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const double t = (q.n[j] * q.A) * w;
        corr[j] = r.own ? corr[j] + t : corr[j] - t;
      }
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int64_t i = (int64_t)(v * D + j) * m.C + c;
      const double b = x0[i] * V / relax;
      const double Ax = x[i] * V - corr[j];
      xo[i] = x[i] + relax * ((b - Ax) * iv);
    }
  }
}

// Reverse sweep: z = y / V; bbar += r z; y_out = y - r (V z - corr^T(z)).
template <int D>
__global__ void __launch_bounds__(256) k_cgg_sweep_adj(MeshView m, int nv, const double* __restrict__ y,
                                                       double* __restrict__ yo, double* __restrict__ bbar,
                                                       double relax, int first) {
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  const double V = m.vol[c];
  for (int v = 0; v < nv; ++v) {
    double ct[D];
#pragma unroll
    for (int j = 0; j < D; ++j) ct[j] = 0.0;
    for (FaceWalk walk(m, c); walk.more();) {
      FaceRef r = walk.next(m, c);
      if (r.N < 0) continue;
      FaceGeo<D> q;
      load_face_geo<D>(m, r.f, r.P, r.N, q);
      const double iP = 1.0 / m.vol[r.P], iN = 1.0 / m.vol[r.N];
      double sz = 0.0;
#pragma unroll
      for (int j = 0; j < D; ++j)
        sz += (q.n[j] * q.A) * (y[(int64_t)(v * D + j) * m.C + r.P] * iP - y[(int64_t)(v * D + j) * m.C + r.N] * iN);
      const double wgt = r.own ? 1.0 - q.g : q.g;
#pragma unroll
      for (int j = 0; j < D; ++j) ct[j] += wgt * q.skew[j] * sz;
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int64_t i = (int64_t)(v * D + j) * m.C + c;
      const double z = y[i] / V;
      bbar[i] = (first ? 0.0 : bbar[i]) + relax * z;
      yo[i] = y[i] - relax * (V * z - ct[j]);
    }
  }
}

template <int D>
__global__ void __launch_bounds__(256) k_cgg_final_adj(MeshView m, int nv, const double* __restrict__ y,
                                                       double* __restrict__ bbar, double relax, int first) {
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  const double V = m.vol[c];
  for (int v = 0; v < nv; ++v)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int64_t i = (int64_t)(v * D + j) * m.C + c;
      bbar[i] = (first ? 0.0 : bbar[i]) + relax * (y[i] / V);
    }
}

__global__ void __launch_bounds__(256) k_scale(int64_t n, double* __restrict__ x, double a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= a;
}

__global__ void __launch_bounds__(256) k_zero(int64_t n, double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// Venkatakrishnan limiter of the listed slots (schemes/limiter.py:78-117), same face order and tie
// conventions as the fused gradient kernels.
template <int D, class S>
__global__ void __launch_bounds__(256) k_venkat(MeshView m, LimiterSlots sl, In<S> x, In<S> G, double lim_k,
                                                Out<S> psi, int32_t* __restrict__ ext) {
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  const double e2 = m.vol[c] * (lim_k * lim_k * lim_k);
  double xcen[D];
#pragma unroll
  for (int j = 0; j < D; ++j) xcen[j] = m.cc[(int64_t)j * m.C + c];
  for (int l = 0; l < sl.nl; ++l) {
    const int v = sl.var[l];
    const S xc = x[(int64_t)v * m.C + c];
    S mx = xc, mn = xc;
    int32_t amx = c, amn = c;
    for (FaceWalk walk(m, c); walk.more();) {
      FaceRef r = walk.next(m, c);
      const int32_t o = r.own ? r.N : r.P;
      if (o < 0) continue;
      const S xo = x[(int64_t)v * m.C + o];
      if (val(xo) > val(mx)) {
        mx = xo;
        amx = o;
      }
      if (val(xo) < val(mn)) {
        mn = xo;
        amn = o;
      }
    }
    S g[D];
#pragma unroll
    for (int j = 0; j < D; ++j) g[j] = G[(int64_t)(v * D + j) * m.C + c];
    S lo = sconst<S>(0.0);
    bool first = true;
    int32_t ks = 0;
    for (int32_t k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
      const int32_t e = m.cface[k];
      const int32_t f = e >= 0 ? e : ~e;
      S dm = g[0] * (m.fc[f] - xcen[0]);
#pragma unroll
      for (int j = 1; j < D; ++j) dm = dm + g[j] * (m.fc[(int64_t)j * m.F + f] - xcen[j]);
      const S vf = venkat_face<S>(dm, mx - xc, mn - xc, e2);
      if (first || val(vf) < val(lo)) {
        lo = vf;
        ks = k - m.cptr[c];
      }
      first = false;
    }
    const int sj = sl.slot[l];
    if (sl.kmin) sl.kmin[(int64_t)sj * m.C + c] = ks;
    psi.set((int64_t)sj * m.C + c, dmin(dmax(lo, 0.0), 1.0));
    if (ext) {
      ext[(int64_t)sj * m.C + c] = amx;
      ext[(int64_t)(sl.ext_stride + sj) * m.C + c] = amn;
    }
  }
}

template <class S>
int venkat_impl(const MeshView& m, int dim, const LimiterSlots& sl, In<S> x, In<S> G, double lim_k, Out<S> psi,
                int32_t* ext, hipStream_t s) {
  if (sl.nl == 0) return 0;
  if (dim == 2)
    AFX_LAUNCH(kKLimiter, (k_venkat<2, S>), grid_for(m.C), dim3(kBlock), 0, s, m, sl, x, G, lim_k, psi, ext);
  else
    AFX_LAUNCH(kKLimiter, (k_venkat<3, S>), grid_for(m.C), dim3(kBlock), 0, s, m, sl, x, G, lim_k, psi, ext);
  AFX_CHECK_LAUNCH();
  return 0;
}

}  // namespace

int cgg_sweeps(const MeshView& m, int dim, int nv, double* G, double* scratch, int sweeps, double relax,
               hipStream_t s) {
  const int64_t n = (int64_t)nv * dim * m.C;
  if (sweeps <= 0) {
    hipLaunchKernelGGL(k_zero, grid_for(n), dim3(kBlock), 0, s, n, G);
    AFX_CHECK_LAUNCH();
    return 0;
  }
  if (relax != 1.0) {
    hipLaunchKernelGGL(k_scale, grid_for(n), dim3(kBlock), 0, s, n, G, relax);
    AFX_CHECK_LAUNCH();
  }
  if (sweeps == 1) return 0;
  double* x0 = scratch;
  double* tmp = scratch + n;
  AFX_CHECK_HIP(hipMemcpyAsync(x0, G, n * 8, hipMemcpyDeviceToDevice, s));
  double* cur = G;
  double* nxt = tmp;
  for (int k = 1; k < sweeps; ++k) {
    if (dim == 2)
      AFX_LAUNCH(kKCggSweep, k_cgg_sweep<2>, grid_for(m.C), dim3(kBlock), 0, s, m, nv, x0, cur, nxt, relax);
    else
      AFX_LAUNCH(kKCggSweep, k_cgg_sweep<3>, grid_for(m.C), dim3(kBlock), 0, s, m, nv, x0, cur, nxt, relax);
    AFX_CHECK_LAUNCH();
    double* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (cur != G) AFX_CHECK_HIP(hipMemcpyAsync(G, cur, n * 8, hipMemcpyDeviceToDevice, s));
  return 0;
}

int cgg_sweeps_adjoint(const MeshView& m, int dim, int nv, double* Gbar, double* scratch, int sweeps, double relax,
                       hipStream_t s) {
  const int64_t n = (int64_t)nv * dim * m.C;
  if (sweeps <= 0) {
    hipLaunchKernelGGL(k_zero, grid_for(n), dim3(kBlock), 0, s, n, Gbar);
    AFX_CHECK_LAUNCH();
    return 0;
  }
  double* y = Gbar;
  double* ya = scratch;
  double* bbar = scratch + n;
  int first = 1;
  for (int k = sweeps - 1; k >= 1; --k) {
    if (dim == 2)
      AFX_LAUNCH(kKCggSweepAdj, k_cgg_sweep_adj<2>, grid_for(m.C), dim3(kBlock), 0, s, m, nv, y, ya, bbar, relax,
                 first);
    else
      AFX_LAUNCH(kKCggSweepAdj, k_cgg_sweep_adj<3>, grid_for(m.C), dim3(kBlock), 0, s, m, nv, y, ya, bbar, relax,
                 first);
    AFX_CHECK_LAUNCH();
    first = 0;
    double* t = y;
    y = ya;
    ya = t;
  }
  if (dim == 2)
    AFX_LAUNCH(kKCggSweepAdj, k_cgg_final_adj<2>, grid_for(m.C), dim3(kBlock), 0, s, m, nv, y, bbar, relax, first);
  else
    AFX_LAUNCH(kKCggSweepAdj, k_cgg_final_adj<3>, grid_for(m.C), dim3(kBlock), 0, s, m, nv, y, bbar, relax, first);
  AFX_CHECK_LAUNCH();
  AFX_CHECK_HIP(hipMemcpyAsync(Gbar, bbar, n * 8, hipMemcpyDeviceToDevice, s));
  return 0;
}

int venkat_pass(const MeshView& m, int dim, const LimiterSlots& sl, In<double> x, In<double> G, double lim_k,
                Out<double> psi, int32_t* ext, hipStream_t s) {
  return venkat_impl<double>(m, dim, sl, x, G, lim_k, psi, ext, s);
}

int venkat_pass(const MeshView& m, int dim, const LimiterSlots& sl, In<dual> x, In<dual> G, double lim_k,
                Out<dual> psi, int32_t* ext, hipStream_t s) {
  return venkat_impl<dual>(m, dim, sl, x, G, lim_k, psi, ext, s);
}

}  // namespace afx

// Stand-alone Venkatakrishnan limiter of one scalar (and, with tangents, its JVP).
extern "C" int afx_venkat_limiter(const afx_mesh* mh, const afx_geometry* g, double limiter_k, const double* phi,
                                  const double* grad, const double* phi_t, const double* grad_t, double* psi,
                                  double* psi_t, hipStream_t s) {
  AFX_REQUIRE(mh && g && phi && grad && psi, "afx_venkat_limiter: null argument");
  AFX_REQUIRE((phi_t == nullptr && grad_t == nullptr) || psi_t != nullptr,
              "afx_venkat_limiter: tangents need a psi_t buffer");
  afx::MeshView m;
  if (int rv_ = afx_rview(mh, g, s, &m)) return rv_;
  if (m.C == 0) return 0;
  afx::LimiterSlots sl{};
  sl.nl = 1;
  sl.ext_stride = 1;
  if (psi_t == nullptr)
    return afx::venkat_pass(m, mh->dim, sl, afx::In<double>{phi, nullptr}, afx::In<double>{grad, nullptr}, limiter_k,
                            afx::Out<double>{psi, nullptr}, nullptr, s);
  return afx::venkat_pass(m, mh->dim, sl, afx::In<afx::dual>{phi, phi_t}, afx::In<afx::dual>{grad, grad_t}, limiter_k,
                          afx::Out<afx::dual>{psi, psi_t}, nullptr, s);
}
