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
        ks = f;
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

// ---------------------------------------------------------------------------------------------
// node-coordinate derivatives of the sweeps

// One forward sweep with the gradient and the geometry as (value, tangent) pairs.
template <int D>
__global__ void __launch_bounds__(256) k_cgg_sweep_tan(MeshView m, GeoTan gt, int nv, In<dual> x0, In<dual> x,
                                                       Out<dual> xo, double relax) {
  using GR = GeoRead<D, true>;
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  const dual V = GR::vol(m, gt, c);
  const dual iv = 1.0 / V;
  for (int v = 0; v < nv; ++v) {
    dual corr[D];
#pragma unroll
    for (int j = 0; j < D; ++j) corr[j] = dual{0.0, 0.0};
    for (FaceWalk walk(m, c); walk.more();) {
      FaceRef r = walk.next(m, c);
      if (r.N < 0) continue;
      FaceGeo<D, dual> q;
      GR::face(m, gt, r.f, r.P, r.N, q);
      dual sP = dual{0.0, 0.0}, sN = dual{0.0, 0.0};
#pragma unroll
      for (int j = 0; j < D; ++j) {
        sP = sP + q.skew[j] * x[(int64_t)(v * D + j) * m.C + r.P];
        sN = sN + q.skew[j] * x[(int64_t)(v * D + j) * m.C + r.N];
      }
      const dual w = (1.0 - q.g) * sP + q.g * sN;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const dual t = (q.n[j] * q.A) * w;
        corr[j] = r.own ? corr[j] + t : corr[j] - t;
      }
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int64_t i = (int64_t)(v * D + j) * m.C + c;
      const dual xi = x[i];
      const dual b = x0[i] * V / relax;
      const dual Ax = xi * V - corr[j];
      xo.set(i, xi + relax * ((b - Ax) * iv));
    }
  }
}

// Reverse of one sweep x_{k+1} = (1 - r) x_k + r V^-1 (b + corr(x_k)) with its geometry
// cotangents: y is the cotangent of x_{k+1}; yo receives that of x_k, bbar that of b (= on the first
// call, += after).  With z = r V^-1 y: Vbar -= y . (x_{k+1} - (1 - r) x_k) / V, and at face f, with
// dz = z_P - z_N, Sbar += sum_v dz_v w_v, gbar += sum_v (dz_v . S) skew . (x_N - x_P)_v,
// skewbar += sum_v (dz_v . S) ((1 - g) x_P + g x_N)_v (owner writes).
template <int D>
__global__ void __launch_bounds__(256) k_cgg_sweep_adj_geo(MeshView m, int nv, const double* __restrict__ y,
                                                           const double* __restrict__ xk,
                                                           const double* __restrict__ xk1, double* __restrict__ yo,
                                                           double* __restrict__ bbar, double relax, int first,
                                                           CggGeoBars cb) {
  constexpr int NVMAX = D + 3;
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  const double V = m.vol[c];
  double vb = 0.0;
  for (int v = 0; v < nv; ++v)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int64_t i = (int64_t)(v * D + j) * m.C + c;
      vb -= y[i] * (xk1[i] - (1.0 - relax) * xk[i]);
    }
  cb.vol[c] += vb / V;
  double ct[NVMAX][D];
#pragma unroll
  for (int v = 0; v < NVMAX; ++v)
#pragma unroll
    for (int j = 0; j < D; ++j) ct[v][j] = 0.0;
  for (FaceWalk walk(m, c); walk.more();) {
    FaceRef r = walk.next(m, c);
    if (r.N < 0) continue;
    FaceGeo<D> q;
    load_face_geo<D>(m, r.f, r.P, r.N, q);
    const double iP = 1.0 / m.vol[r.P], iN = 1.0 / m.vol[r.N];
    const double wgt = r.own ? 1.0 - q.g : q.g;
    double Sb[D], gbar = 0.0, kb[D];
#pragma unroll
    for (int j = 0; j < D; ++j) Sb[j] = kb[j] = 0.0;
#pragma unroll
    for (int v = 0; v < NVMAX; ++v) {
      if (v >= nv) break;
      double dz[D], sz = 0.0;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const int64_t row = (int64_t)(v * D + j) * m.C;
        dz[j] = y[row + r.P] * iP - y[row + r.N] * iN;
        sz += (q.n[j] * q.A) * dz[j];
      }
#pragma unroll
      for (int j = 0; j < D; ++j) ct[v][j] += wgt * q.skew[j] * sz;
      if (r.own) {
        double sP = 0.0, sN = 0.0, xP[D], xN[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const int64_t row = (int64_t)(v * D + j) * m.C;
          xP[j] = xk[row + r.P];
          xN[j] = xk[row + r.N];
          sP += q.skew[j] * xP[j];
          sN += q.skew[j] * xN[j];
        }
        const double w = (1.0 - q.g) * sP + q.g * sN;
        const double wb = relax * sz;
        gbar += wb * (sN - sP);
#pragma unroll
        for (int j = 0; j < D; ++j) {
          Sb[j] += relax * dz[j] * w;
          kb[j] += wb * ((1.0 - q.g) * xP[j] + q.g * xN[j]);
        }
      }
    }
    if (r.own) {
#pragma unroll
      for (int j = 0; j < D; ++j) {
        cb.face[(int64_t)j * m.F + r.f] += Sb[j];
        cb.face[(int64_t)(D + 1 + j) * m.F + r.f] += kb[j];
      }
      cb.face[(int64_t)D * m.F + r.f] += gbar;
    }
  }
  for (int v = 0; v < nv; ++v)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int64_t i = (int64_t)(v * D + j) * m.C + c;
      const double z = y[i] / V;
      bbar[i] = (first ? 0.0 : bbar[i]) + relax * z;
      yo[i] = y[i] - relax * (V * z - ct[v][j]);
    }
}

// Reverse of x_0 = r V^-1 b: bbar += r y / V, Vbar -= y . x_0 / V.
template <int D>
__global__ void __launch_bounds__(256) k_cgg_final_adj_geo(MeshView m, int nv, const double* __restrict__ y,
                                                           const double* __restrict__ x0, double* __restrict__ bbar,
                                                           double relax, int first, double* __restrict__ volb) {
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  const double V = m.vol[c];
  double vb = 0.0;
  for (int v = 0; v < nv; ++v)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int64_t i = (int64_t)(v * D + j) * m.C + c;
      bbar[i] = (first ? 0.0 : bbar[i]) + relax * (y[i] / V);
      vb -= y[i] * x0[i];
    }
  volb[c] += vb / V;
}

// The limiter of k_venkat with the centroids and the volume (through V K^3) as (value, tangent)
// pairs; the same face order and first-minimum rule.
template <int D>
__global__ void __launch_bounds__(256) k_venkat_tan(MeshView m, GeoTan gt, LimiterSlots sl, In<dual> x, In<dual> G,
                                                    double lim_k, Out<dual> psi) {
  using GR = GeoRead<D, true>;
  const int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  const dual e2 = GR::vol(m, gt, c) * (lim_k * lim_k * lim_k);
  dual xcen[D];
#pragma unroll
  for (int j = 0; j < D; ++j) xcen[j] = GR::cc(m, gt, j, c);
  for (int l = 0; l < sl.nl; ++l) {
    const int v = sl.var[l];
    const dual xc = x[(int64_t)v * m.C + c];
    dual mx = xc, mn = xc;
    for (FaceWalk walk(m, c); walk.more();) {
      FaceRef r = walk.next(m, c);
      const int32_t o = r.own ? r.N : r.P;
      if (o < 0) continue;
      const dual xo = x[(int64_t)v * m.C + o];
      if (val(xo) > val(mx)) mx = xo;
      if (val(xo) < val(mn)) mn = xo;
    }
    dual g[D];
#pragma unroll
    for (int j = 0; j < D; ++j) g[j] = G[(int64_t)(v * D + j) * m.C + c];
    dual lo = dual{0.0, 0.0};
    bool first = true;
    for (int32_t k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
      const int32_t e = m.cface[k];
      const int32_t f = e >= 0 ? e : ~e;
      dual dm = g[0] * (GR::fc(m, gt, 0, f) - xcen[0]);
#pragma unroll
      for (int j = 1; j < D; ++j) dm = dm + g[j] * (GR::fc(m, gt, j, f) - xcen[j]);
      const dual vf = venkat_face<dual, dual>(dm, mx - xc, mn - xc, e2);
      if (first || val(vf) < val(lo)) lo = vf;
      first = false;
    }
    psi.set((int64_t)sl.slot[l] * m.C + c, dmin(dmax(lo, 0.0), 1.0));
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

int cgg_sweeps_store(const MeshView& m, int dim, int nv, double* G, double* iters, int sweeps, double relax,
                     hipStream_t s) {
  const int64_t n = (int64_t)nv * dim * m.C;
  if (sweeps <= 1) return cgg_sweeps(m, dim, nv, G, nullptr, sweeps, relax, s);
  AFX_REQUIRE(iters, "cgg_sweeps_store: null iterate storage");
  if (relax != 1.0) {
    hipLaunchKernelGGL(k_scale, grid_for(n), dim3(kBlock), 0, s, n, G, relax);
    AFX_CHECK_LAUNCH();
  }
  AFX_CHECK_HIP(hipMemcpyAsync(iters, G, n * 8, hipMemcpyDeviceToDevice, s));
  for (int k = 1; k < sweeps; ++k) {
    const double* in = iters + (int64_t)(k - 1) * n;
    double* out = k < sweeps - 1 ? iters + (int64_t)k * n : G;
    if (dim == 2)
      AFX_LAUNCH(kKCggSweep, k_cgg_sweep<2>, grid_for(m.C), dim3(kBlock), 0, s, m, nv, iters, in, out, relax);
    else
      AFX_LAUNCH(kKCggSweep, k_cgg_sweep<3>, grid_for(m.C), dim3(kBlock), 0, s, m, nv, iters, in, out, relax);
    AFX_CHECK_LAUNCH();
  }
  return 0;
}

int cgg_sweeps_adjoint_geometry(const MeshView& m, int dim, int nv, double* Gbar, const double* G,
                                const double* iters, double* scratch, int sweeps, double relax, CggGeoBars gb,
                                hipStream_t s) {
  const int64_t n = (int64_t)nv * dim * m.C;
  AFX_REQUIRE(nv <= dim + 3, "cgg_sweeps_adjoint_geometry: at most d + 3 variables");
  if (int rz = afx_fill_zero(gb.face, (int64_t)(2 * dim + 1) * m.F, s)) return rz;
  if (int rz = afx_fill_zero(gb.vol, m.C, s)) return rz;
  if (sweeps <= 0) return cgg_sweeps_adjoint(m, dim, nv, Gbar, scratch, sweeps, relax, s);
  AFX_REQUIRE(sweeps == 1 || iters, "cgg_sweeps_adjoint_geometry: null iterates");
  double* y = Gbar;
  double* ya = scratch;
  double* bbar = scratch + n;
  int first = 1;
  for (int k = sweeps - 1; k >= 1; --k) {
    const double* xk = iters + (int64_t)(k - 1) * n;
    const double* xk1 = k < sweeps - 1 ? iters + (int64_t)k * n : G;
    if (dim == 2)
      AFX_LAUNCH(kKCggSweepAdj, k_cgg_sweep_adj_geo<2>, grid_for(m.C), dim3(kBlock), 0, s, m, nv, y, xk, xk1, ya,
                 bbar, relax, first, gb);
    else
      AFX_LAUNCH(kKCggSweepAdj, k_cgg_sweep_adj_geo<3>, grid_for(m.C), dim3(kBlock), 0, s, m, nv, y, xk, xk1, ya,
                 bbar, relax, first, gb);
    AFX_CHECK_LAUNCH();
    first = 0;
    double* t = y;
    y = ya;
    ya = t;
  }
  const double* x0 = sweeps > 1 ? iters : G;
  if (dim == 2)
    AFX_LAUNCH(kKCggSweepAdj, k_cgg_final_adj_geo<2>, grid_for(m.C), dim3(kBlock), 0, s, m, nv, y, x0, bbar, relax,
               first, gb.vol);
  else
    AFX_LAUNCH(kKCggSweepAdj, k_cgg_final_adj_geo<3>, grid_for(m.C), dim3(kBlock), 0, s, m, nv, y, x0, bbar, relax,
               first, gb.vol);
  AFX_CHECK_LAUNCH();
  AFX_CHECK_HIP(hipMemcpyAsync(Gbar, bbar, n * 8, hipMemcpyDeviceToDevice, s));
  return 0;
}

int cgg_sweeps_tangent(const MeshView& m, const GeoTan& gt, int dim, int nv, double* G, double* Gt,
                       double* scratch, int sweeps, double relax, hipStream_t s) {
  const int64_t n = (int64_t)nv * dim * m.C;
  if (sweeps <= 0) {
    hipLaunchKernelGGL(k_zero, grid_for(n), dim3(kBlock), 0, s, n, G);
    AFX_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_zero, grid_for(n), dim3(kBlock), 0, s, n, Gt);
    AFX_CHECK_LAUNCH();
    return 0;
  }
  if (relax != 1.0) {
    hipLaunchKernelGGL(k_scale, grid_for(n), dim3(kBlock), 0, s, n, G, relax);
    AFX_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_scale, grid_for(n), dim3(kBlock), 0, s, n, Gt, relax);
    AFX_CHECK_LAUNCH();
  }
  if (sweeps == 1) return 0;
  AFX_REQUIRE(scratch, "cgg_sweeps_tangent: null scratch");
  double* x0v = scratch;
  double* x0t = scratch + n;
  AFX_CHECK_HIP(hipMemcpyAsync(x0v, G, n * 8, hipMemcpyDeviceToDevice, s));
  AFX_CHECK_HIP(hipMemcpyAsync(x0t, Gt, n * 8, hipMemcpyDeviceToDevice, s));
  double *cv = G, *ct = Gt, *nv_ = scratch + 2 * n, *nt = scratch + 3 * n;
  for (int k = 1; k < sweeps; ++k) {
    if (dim == 2)
      AFX_LAUNCH(kKCggSweep, k_cgg_sweep_tan<2>, grid_for(m.C), dim3(kBlock), 0, s, m, gt, nv, In<dual>{x0v, x0t},
                 In<dual>{cv, ct}, Out<dual>{nv_, nt}, relax);
    else
      AFX_LAUNCH(kKCggSweep, k_cgg_sweep_tan<3>, grid_for(m.C), dim3(kBlock), 0, s, m, gt, nv, In<dual>{x0v, x0t},
                 In<dual>{cv, ct}, Out<dual>{nv_, nt}, relax);
    AFX_CHECK_LAUNCH();
    std::swap(cv, nv_);
    std::swap(ct, nt);
  }
  if (cv != G) {
    AFX_CHECK_HIP(hipMemcpyAsync(G, cv, n * 8, hipMemcpyDeviceToDevice, s));
    AFX_CHECK_HIP(hipMemcpyAsync(Gt, ct, n * 8, hipMemcpyDeviceToDevice, s));
  }
  return 0;
}

int venkat_pass_tangent(const MeshView& m, const GeoTan& gt, int dim, const LimiterSlots& sl, In<dual> x,
                        In<dual> G, double lim_k, Out<dual> psi, hipStream_t s) {
  if (sl.nl == 0) return 0;
  if (dim == 2)
    AFX_LAUNCH(kKLimiter, k_venkat_tan<2>, grid_for(m.C), dim3(kBlock), 0, s, m, gt, sl, x, G, lim_k, psi);
  else
    AFX_LAUNCH(kKLimiter, k_venkat_tan<3>, grid_for(m.C), dim3(kBlock), 0, s, m, gt, sl, x, G, lim_k, psi);
  AFX_CHECK_LAUNCH();
  return 0;
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
