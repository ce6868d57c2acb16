// Mesh handle and device geometry (forward + adjoint with respect to node coordinates).
//
// Geometry follows the reference pipeline (aquaflux/mesh/mesh.py:365-387):
//   1. face area / centroid / winding normal -- 2D edge (mesh/face.py:217-230) or 3D centre fan
//      with the signed-area centroid (mesh/face.py:292-349);
//   2. approximate cell centroids = mean of the cell's face centroids (mesh/cell.py:84-112);
//   3. owner-outward orientation of the normal (mesh/face.py:196-211);
//   4. cell volume and centroid by the divergence theorem, recentred on the approximate centroid
//      (mesh/cell.py:123-170).
// Each step is one thread per face or per cell; cell steps walk the cell -> face list, so every
// reduction is a fixed-order register sum (no atomics, bitwise reproducible).
#include "afx_common.h"
#include "afx_mesh.h"
#include "afx_profile.h"

#include <vector>
#include <algorithm>
#include <type_traits>

static thread_local std::string g_err;
void afx_set_error(const std::string& msg) { g_err = msg; }
const char* afx_get_error() { return g_err.c_str(); }

using namespace afx;

extern "C" const char* afx_last_error(void) { return afx_get_error(); }

extern "C" int afx_mesh_create(int dim, int64_t n_cells, int64_t n_faces, int64_t n_nodes,
                               const int32_t* owner, const int32_t* neighbour,
                               const int64_t* face_node_offsets, const int32_t* face_node_indices,
                               afx_mesh** out) {
  AFX_REQUIRE(out != nullptr, "afx_mesh_create: null output handle");
  *out = nullptr;
  AFX_REQUIRE(dim == 2 || dim == 3, "afx_mesh_create: dim must be 2 or 3");
  AFX_REQUIRE(n_cells > 0 && n_faces > 0 && n_nodes > 0, "afx_mesh_create: empty mesh");
  AFX_REQUIRE(n_cells < INT32_MAX && n_faces < INT32_MAX / 2 && n_nodes < INT32_MAX,
              "afx_mesh_create: mesh exceeds 32-bit indexing");
  // host-side build of the cell -> face CSR (owner entries f, neighbour entries ~f), face order
  std::vector<int32_t> count(n_cells + 1, 0);
  for (int64_t f = 0; f < n_faces; ++f) {
    int32_t P = owner[f], N = neighbour[f];
    AFX_REQUIRE(P >= 0 && P < n_cells, "afx_mesh_create: owner index out of range");
    AFX_REQUIRE(N >= -1 && N < n_cells, "afx_mesh_create: neighbour index out of range");
    count[P + 1]++;
    if (N >= 0) count[N + 1]++;
  }
  for (int64_t c = 0; c < n_cells; ++c) count[c + 1] += count[c];
  std::vector<int32_t> cface(count[n_cells]);
  std::vector<int32_t> fill(count.begin(), count.end() - 1);
  int32_t max_faces = 0;
  for (int64_t f = 0; f < n_faces; ++f) {
    cface[fill[owner[f]]++] = (int32_t)f;
    if (neighbour[f] >= 0) cface[fill[neighbour[f]]++] = ~(int32_t)f;
  }
  for (int64_t c = 0; c < n_cells; ++c) max_faces = std::max(max_faces, count[c + 1] - count[c]);
  std::vector<int32_t> cnbr(cface.size());
  for (size_t k = 0; k < cface.size(); ++k) {
    const int32_t e = cface[k], f = e >= 0 ? e : ~e;
    cnbr[k] = e >= 0 ? neighbour[f] : owner[f];
  }
  // Internal face order.  A caller order that is not sorted by owner (the structured generators
  // number faces family by family, k fastest) is replaced by faces sorted by (rank among their
  // owner's faces, owner, caller index): slot-r faces of consecutive cells are then consecutive
  // records, so each face read of a wave -- the owner's, and one cell / row / plane away the
  // neighbour's -- is contiguous.  Owner-sorted caller orders (tetrahedral and OpenFOAM-style
  // meshes) are kept as they are.
  bool sorted = true;
  for (int64_t f = 1; f < n_faces && sorted; ++f) sorted = owner[f] >= owner[f - 1];
  std::vector<int32_t> fu;
  if (!sorted) {
    std::vector<int32_t> rank(n_faces), by_owner(n_faces);
    std::vector<int64_t> ostart(n_cells + 1, 0);
    for (int64_t f = 0; f < n_faces; ++f) ostart[owner[f] + 1]++;
    for (int64_t c = 0; c < n_cells; ++c) ostart[c + 1] += ostart[c];
    int32_t max_rank = 0;
    for (int64_t f = 0; f < n_faces; ++f) {
      const int64_t pos = ostart[owner[f]]++;
      by_owner[pos] = (int32_t)f;
    }
    for (int64_t i = 0; i < n_faces; ++i) {  // rank = position within the owner's run
      const int32_t f = by_owner[i];
      rank[f] = (i > 0 && owner[by_owner[i - 1]] == owner[f]) ? rank[by_owner[i - 1]] + 1 : 0;
      max_rank = std::max(max_rank, rank[f]);
    }
    std::vector<int64_t> rstart(max_rank + 2, 0);  // stable counting sort of by_owner by rank
    for (int64_t f = 0; f < n_faces; ++f) rstart[rank[f] + 1]++;
    for (int32_t r = 0; r <= max_rank; ++r) rstart[r + 1] += rstart[r];
    fu.resize(n_faces);
    for (int64_t i = 0; i < n_faces; ++i) {
      const int32_t f = by_owner[i];
      fu[rstart[rank[f]]++] = f;
    }
  }
  std::vector<int32_t> owner_i, nb_i, cface_i, cnbr_i;
  if (!sorted) {
    owner_i.resize(n_faces);
    nb_i.resize(n_faces);
    for (int64_t i = 0; i < n_faces; ++i) {
      owner_i[i] = owner[fu[i]];
      nb_i[i] = neighbour[fu[i]];
    }
    cface_i.resize(count[n_cells]);
    std::vector<int32_t> fill_i(count.begin(), count.end() - 1);
    for (int64_t i = 0; i < n_faces; ++i) {
      cface_i[fill_i[owner_i[i]]++] = (int32_t)i;
      if (nb_i[i] >= 0) cface_i[fill_i[nb_i[i]]++] = ~(int32_t)i;
    }
    cnbr_i.resize(cface_i.size());
    for (size_t k = 0; k < cface_i.size(); ++k) {
      const int32_t e = cface_i[k], f = e >= 0 ? e : ~e;
      cnbr_i[k] = e >= 0 ? nb_i[f] : owner_i[f];
    }
  }
  // node -> incidence CSR (for the node-coordinate adjoint gather)
  int64_t n_inc = face_node_offsets[n_faces];
  std::vector<int32_t> ncount(n_nodes + 1, 0);
  for (int64_t i = 0; i < n_inc; ++i) {
    int32_t v = face_node_indices[i];
    AFX_REQUIRE(v >= 0 && v < n_nodes, "afx_mesh_create: face_node_indices out of range");
    ncount[v + 1]++;
  }
  for (int64_t v = 0; v < n_nodes; ++v) ncount[v + 1] += ncount[v];
  std::vector<int32_t> ninc(n_inc);
  {
    std::vector<int32_t> nfill(ncount.begin(), ncount.end() - 1);
    for (int64_t i = 0; i < n_inc; ++i) ninc[nfill[face_node_indices[i]]++] = (int32_t)i;
  }
  std::vector<int32_t> kpos(n_inc);
  for (int64_t k = 0; k < n_inc; ++k) kpos[ninc[k]] = (int32_t)k;
  std::vector<int32_t> inc_face(n_inc);
  for (int64_t f = 0; f < n_faces; ++f)
    for (int64_t i = face_node_offsets[f]; i < face_node_offsets[f + 1]; ++i) inc_face[i] = (int32_t)f;
  std::vector<int32_t> foff32(n_faces + 1);
  for (int64_t f = 0; f <= n_faces; ++f) foff32[f] = (int32_t)face_node_offsets[f];

  afx_mesh* m = new afx_mesh();
  m->dim = dim;
  m->C = (int32_t)n_cells;
  m->F = (int32_t)n_faces;
  m->N = (int32_t)n_nodes;
  m->n_inc = (int32_t)n_inc;
  m->n_cface = (int32_t)cface.size();
  m->max_faces = max_faces;
  AFX_CHECK_HIP(hipGetDevice(&m->device));
  auto up = [&](void** dst, const void* src, size_t bytes) -> hipError_t {
    hipError_t e = hipMalloc(dst, bytes > 0 ? bytes : 4);
    if (e != hipSuccess) return e;
    return bytes ? hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
  };
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = up((void**)&m->owner, owner, n_faces * 4);
  if (e == hipSuccess) e = up((void**)&m->nb, neighbour, n_faces * 4);
  if (e == hipSuccess) e = up((void**)&m->cptr, count.data(), (n_cells + 1) * 4);
  if (e == hipSuccess) e = up((void**)&m->cface, cface.data(), cface.size() * 4);
  if (e == hipSuccess) e = up((void**)&m->cnbr, cnbr.data(), cnbr.size() * 4);
  if (e == hipSuccess) e = up((void**)&m->foff, foff32.data(), (n_faces + 1) * 4);
  if (e == hipSuccess) e = up((void**)&m->fidx, face_node_indices, n_inc * 4);
  if (e == hipSuccess) e = up((void**)&m->inc_face, inc_face.data(), n_inc * 4);
  if (e == hipSuccess) e = up((void**)&m->nptr, ncount.data(), (n_nodes + 1) * 4);
  if (e == hipSuccess) e = up((void**)&m->ninc, ninc.data(), n_inc * 4);
  if (e == hipSuccess) e = up((void**)&m->kpos, kpos.data(), n_inc * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&m->approx, (size_t)n_cells * dim * 8);
  if (e == hipSuccess) e = hipMalloc((void**)&m->normal_raw, (size_t)n_faces * dim * 8);
  if (!sorted) {
    if (e == hipSuccess) e = up((void**)&m->fu, fu.data(), n_faces * 4);
    if (e == hipSuccess) e = up((void**)&m->owner_i, owner_i.data(), n_faces * 4);
    if (e == hipSuccess) e = up((void**)&m->nb_i, nb_i.data(), n_faces * 4);
    if (e == hipSuccess) e = up((void**)&m->cface_i, cface_i.data(), cface_i.size() * 4);
    if (e == hipSuccess) e = up((void**)&m->cnbr_i, cnbr_i.data(), cnbr_i.size() * 4);
  }
  if (e != hipSuccess) {
    afx_set_error(std::string("afx_mesh_create: ") + hipGetErrorString(e));
    afx_mesh_destroy(m);
    return 1;
  }
  *out = m;
  return 0;
}

extern "C" int afx_mesh_destroy(afx_mesh* m) {
  if (!m) return 0;
  void* ptrs[] = {m->owner,   m->nb,      m->cptr,       m->cface,   m->foff, m->fidx,   m->inc_face,
                  m->nptr,    m->ninc,    m->approx,     m->normal_raw, m->scratch, m->fu, m->owner_i,
                  m->nb_i,    m->cface_i, m->pack_scratch, m->kpos, m->cnbr, m->cnbr_i, m->bsched};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete m;
  return 0;
}

extern "C" int afx_mesh_info(const afx_mesh* m, int64_t* out) {
  AFX_REQUIRE(m && out, "afx_mesh_info: null argument");
  out[0] = m->C;
  out[1] = m->F;
  out[2] = m->N;
  out[3] = m->dim;
  out[4] = m->n_cface;
  out[5] = m->max_faces;
  out[6] = m->device;
  return 0;
}

extern "C" int afx_mesh_cell_faces(const afx_mesh* m, int32_t* cptr, int32_t* cface, hipStream_t s) {
  AFX_REQUIRE(m && cptr && cface, "afx_mesh_cell_faces: null argument");
  AFX_CHECK_HIP(hipMemcpyAsync(cptr, m->cptr, (size_t)(m->C + 1) * 4, hipMemcpyDeviceToDevice, s));
  AFX_CHECK_HIP(hipMemcpyAsync(cface, m->cface, (size_t)m->n_cface * 4, hipMemcpyDeviceToDevice, s));
  return 0;
}

namespace {

// ---------------------------------------------------------------------------------------------
// face geometry from nodes (unoriented normal)
template <int D>
__global__ void __launch_bounds__(256) k_face_geometry(int32_t F, int32_t Nn, const int32_t* __restrict__ foff,
                                const int32_t* __restrict__ fidx, const double* __restrict__ X,
                                double* __restrict__ area, double* __restrict__ nrm,
                                double* __restrict__ fc) {
  int32_t f = xcd_thread_index();
  if (f >= F) return;
  int32_t b = foff[f], e = foff[f + 1];
  if (D == 2) {
    int32_t a0 = fidx[b], a1 = fidx[b + 1];
    double p0[2] = {X[a0], X[Nn + a0]}, p1[2] = {X[a1], X[Nn + a1]};
    double d0 = p1[0] - p0[0], d1 = p1[1] - p0[1];
    double sq = d0 * d0 + d1 * d1;
    area[f] = sq > 0.0 ? sqrt(sq) : 0.0;
    fc[f] = 0.5 * (p0[0] + p1[0]);
    fc[F + f] = 0.5 * (p0[1] + p1[1]);
    double v0 = d1, v1 = -d0;
    double s2 = v0 * v0 + v1 * v1;
    double inv = s2 > 0.0 ? sqrt(s2) : 1.0;
    nrm[f] = s2 > 0.0 ? v0 / inv : 0.0;
    nrm[F + f] = s2 > 0.0 ? v1 / inv : 0.0;
  } else {
    int32_t k = e - b;
    double ctr[3] = {0.0, 0.0, 0.0};
    for (int32_t i = b; i < e; ++i) {
      int32_t a = fidx[i];
#pragma unroll
      for (int j = 0; j < 3; ++j) ctr[j] += X[(int64_t)j * Nn + a];
    }
    double invk = 1.0 / (double)k;
#pragma unroll
    for (int j = 0; j < 3; ++j) ctr[j] = ctr[j] * invk;
    double S[3] = {0.0, 0.0, 0.0};
    for (int32_t i = b; i < e; ++i) {
      int32_t a = fidx[i], c = fidx[i + 1 < e ? i + 1 : b];
      double u[3], w[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        u[j] = X[(int64_t)j * Nn + a] - ctr[j];
        w[j] = X[(int64_t)j * Nn + c] - ctr[j];
      }
      S[0] += 0.5 * (u[1] * w[2] - u[2] * w[1]);
      S[1] += 0.5 * (u[2] * w[0] - u[0] * w[2]);
      S[2] += 0.5 * (u[0] * w[1] - u[1] * w[0]);
    }
    double sq = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
    double A = sq > 0.0 ? sqrt(sq) : 0.0;
    double nn[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) nn[j] = sq > 0.0 ? S[j] / A : 0.0;
    double cs[3] = {0.0, 0.0, 0.0};
    for (int32_t i = b; i < e; ++i) {
      int32_t a = fidx[i], c = fidx[i + 1 < e ? i + 1 : b];
      double pb[3], pc[3], u[3], w[3], dir[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        pb[j] = X[(int64_t)j * Nn + a];
        pc[j] = X[(int64_t)j * Nn + c];
        u[j] = pb[j] - ctr[j];
        w[j] = pc[j] - ctr[j];
      }
      dir[0] = 0.5 * (u[1] * w[2] - u[2] * w[1]);
      dir[1] = 0.5 * (u[2] * w[0] - u[0] * w[2]);
      dir[2] = 0.5 * (u[0] * w[1] - u[1] * w[0]);
      double sgn = dir[0] * nn[0] + dir[1] * nn[1] + dir[2] * nn[2];
#pragma unroll
      for (int j = 0; j < 3; ++j) cs[j] += sgn * ((ctr[j] + pb[j] + pc[j]) / 3.0);
    }
    double safe = A > 0.0 ? A : 1.0;
    area[f] = A;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      nrm[(int64_t)j * F + f] = nn[j];
      fc[(int64_t)j * F + f] = cs[j] / safe;
    }
  }
}

template <int D>
__global__ void __launch_bounds__(256) k_approx_centroid(MeshView m, double* __restrict__ approx) {
  int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  double s[D];
#pragma unroll
  for (int j = 0; j < D; ++j) s[j] = 0.0;
  int32_t b = m.cptr[c], e = m.cptr[c + 1];
  for (int32_t k = b; k < e; ++k) {
    int32_t ent = m.cface[k];
    int32_t f = ent >= 0 ? ent : ~ent;
#pragma unroll
    for (int j = 0; j < D; ++j) s[j] += m.fc[(int64_t)j * m.F + f];
  }
  double inv = 1.0 / (double)(e - b);
#pragma unroll
  for (int j = 0; j < D; ++j) approx[(int64_t)j * m.C + c] = s[j] * inv;
}

template <int D>
__global__ void __launch_bounds__(256) k_orient(MeshView m, const double* __restrict__ approx, const double* __restrict__ raw,
                         double* __restrict__ nrm) {
  int32_t f = xcd_thread_index();
  if (f >= m.F) return;
  int32_t P = m.owner[f];
  double proj = 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j)
    proj += (m.fc[(int64_t)j * m.F + f] - approx[(int64_t)j * m.C + P]) * raw[(int64_t)j * m.F + f];
  double sgn = proj < 0.0 ? -1.0 : 1.0;
#pragma unroll
  for (int j = 0; j < D; ++j) nrm[(int64_t)j * m.F + f] = raw[(int64_t)j * m.F + f] * sgn;
}

template <int D>
__global__ void __launch_bounds__(256) k_cell_volume(MeshView m, const double* __restrict__ approx, double* __restrict__ vol,
                              double* __restrict__ cc) {
  int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  double ap[D], v = 0.0, cs[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    ap[j] = approx[(int64_t)j * m.C + c];
    cs[j] = 0.0;
  }
  for (int32_t k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
    int32_t ent = m.cface[k];
    int32_t f = ent >= 0 ? ent : ~ent;
    double loc[D];
#pragma unroll
    for (int j = 0; j < D; ++j) loc[j] = m.fc[(int64_t)j * m.F + f] - ap[j];
    double nl[D];
#pragma unroll
    for (int j = 0; j < D; ++j) nl[j] = m.nrm[(int64_t)j * m.F + f];
    double flux = vdot<D>(loc, nl) * m.area[f];
    if (ent >= 0) {
      v += flux;
#pragma unroll
      for (int j = 0; j < D; ++j) cs[j] += loc[j] * flux;
    } else {
      v += -flux;
#pragma unroll
      for (int j = 0; j < D; ++j) cs[j] += -(loc[j] * flux);
    }
  }
  double V = v / (double)D;
  vol[c] = V;
  double inv = 1.0 / ((double)(D + 1) * V);
#pragma unroll
  for (int j = 0; j < D; ++j) cc[(int64_t)j * m.C + c] = cs[j] * inv + ap[j];
}

// ---------------------------------------------------------------------------------------------
// geometry adjoint (reverse of the four passes above), from cotangents of (area, oriented normal,
// face centroid, volume, cell centroid) to node-coordinate cotangents.

// Cell pass: cotangents of the divergence-theorem sums (csum [d], S = d V) and of the approximate
// centroid, per cell: cellbar[(j)*C + c] for csum_j, [d*C + c] for S, [(d+1+j)*C + c] for approx_j.
// FORM: the approximate centroid is re-formed here from the face centroids (the same sum in the
// same order as k_approx_centroid, so the same value) and written to `approx` for the face pass;
// otherwise `approx` holds the forward pass's.
template <int D, bool FORM>
__global__ void __launch_bounds__(256) k_geom_cell_adj(MeshView m, double* __restrict__ approx, const double* __restrict__ vol_b,
                                const double* __restrict__ cc_b, double* __restrict__ cellbar) {
  int32_t c = xcd_thread_index(m);
  if (c >= m.C) return;
  const double V = m.vol[c];
  double ap[D], csb[D], apb[D];
  double Vb = vol_b ? vol_b[c] : 0.0;
  if (!FORM) {
#pragma unroll
    for (int j = 0; j < D; ++j) ap[j] = approx[(int64_t)j * m.C + c];
  } else {
#pragma unroll
    for (int j = 0; j < D; ++j) ap[j] = 0.0;
    const int32_t b = m.cptr[c], e = m.cptr[c + 1];
    for (int32_t k = b; k < e; ++k) {
      const int32_t ent = m.cface[k];
      const int32_t f = ent >= 0 ? ent : ~ent;
#pragma unroll
      for (int j = 0; j < D; ++j) ap[j] += m.fc[(int64_t)j * m.F + f];
    }
    const double inv = 1.0 / (double)(e - b);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      ap[j] = ap[j] * inv;
      approx[(int64_t)j * m.C + c] = ap[j];
    }
  }
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const double ccb = cc_b ? cc_b[(int64_t)j * m.C + c] : 0.0;
    csb[j] = ccb / ((double)(D + 1) * V);
    Vb -= ccb * (m.cc[(int64_t)j * m.C + c] - ap[j]) / V;
    apb[j] = ccb;
  }
  const double Sb = Vb / (double)D;
  for (int32_t k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
    const int32_t ent = m.cface[k];
    const int32_t f = ent >= 0 ? ent : ~ent;
    const double sg = ent >= 0 ? 1.0 : -1.0;
    const double A = m.area[f];
    double l[D], n[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      l[j] = m.fc[(int64_t)j * m.F + f] - ap[j];
      n[j] = m.nrm[(int64_t)j * m.F + f];
    }
    const double phi = vdot<D>(l, n) * A;
    const double phib = sg * (Sb + vdot<D>(csb, l));
#pragma unroll
    for (int j = 0; j < D; ++j) apb[j] -= sg * csb[j] * phi + phib * A * n[j];
  }
#pragma unroll
  for (int j = 0; j < D; ++j) {
    cellbar[(int64_t)j * m.C + c] = csb[j];
    cellbar[(int64_t)(D + 1 + j) * m.C + c] = apb[j];
  }
  cellbar[(int64_t)D * m.C + c] = Sb;
}

AFX_DEV void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// The 3D centre-fan reverse of the face pass below for a face of k <= 4 nodes (triangles, quads):
// the same operations in the same order as the general loop, with the k node positions loaded once
// and kept in registers (the general loop re-reads each node up to 23 times); per-edge products are
// re-formed from them where needed rather than stored.  Indices into the fixed arrays are
// compile-time constants after unrolling; the run-time k only selects values.
AFX_DEV void face_adj_small(int32_t k, const int32_t* __restrict__ fidx, const int32_t* __restrict__ kpos,
                            int32_t Nn, const double* __restrict__ X, const double* n, const double* nb,
                            const double* xf, const double* xfb, double Ab, double* __restrict__ inc_b) {
  double P[4][3];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int32_t a = i < k ? fidx[i] : fidx[0];
#pragma unroll
    for (int j = 0; j < 3; ++j) P[i][j] = X[(int64_t)j * Nn + a];
  }
  const double invk = 1.0 / (double)k;
  double ctr[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) ctr[j] = i < k ? ctr[j] + P[i][j] : ctr[j];
#pragma unroll
  for (int j = 0; j < 3; ++j) ctr[j] = ctr[j] * invk;
  // edge i runs from node i to node (i + 1) mod k; its end point, fan vectors and product
  auto ends = [&](int i, double* u, double* w, double* q) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      q[j] = i + 1 < k ? P[(i + 1) & 3][j] : P[0][j];
      u[j] = P[i & 3][j] - ctr[j];
      w[j] = q[j] - ctr[j];
    }
  };
  double Sv[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double u[3], w[3], q[3], dir[3];
    ends(i, u, w, q);
    cross3(u, w, dir);
#pragma unroll
    for (int j = 0; j < 3; ++j) Sv[j] = i < k ? Sv[j] + 0.5 * dir[j] : Sv[j];
  }
  const double Ar = sqrt(Sv[0] * Sv[0] + Sv[1] * Sv[1] + Sv[2] * Sv[2]);
  double nh[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) nh[j] = Sv[j] / Ar;
  const double s = (nh[0] * n[0] + nh[1] * n[1] + nh[2] * n[2]) < 0.0 ? -1.0 : 1.0;
  double nhb[3], csb[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    nhb[j] = s * nb[j];
    csb[j] = xfb[j] / Ar;
  }
  Ab -= (xfb[0] * xf[0] + xfb[1] * xf[1] + xfb[2] * xf[2]) / Ar;
  double Sg[4], Sn[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double u[3], w[3], q[3], dir[3];
    ends(i, u, w, q);
    cross3(u, w, dir);
    double sgnb = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) sgnb += csb[j] * ((ctr[j] + P[i][j] + q[j]) / 3.0);
    Sg[i] = sgnb;
#pragma unroll
    for (int j = 0; j < 3; ++j) nhb[j] = i < k ? nhb[j] + sgnb * 0.5 * dir[j] : nhb[j];
    double sgn = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) sgn += 0.5 * dir[j] * nh[j];
    Sn[i] = sgn;
  }
  const double pr = nhb[0] * nh[0] + nhb[1] * nh[1] + nhb[2] * nh[2];
  double Svb[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) Svb[j] = Ab * nh[j] + (nhb[j] - pr * nh[j]) / Ar;
  // edge i's cotangents of its first node (ub) and second node (wb), and its centroid share (tcb)
  auto edge = [&](int i, double* ub, double* wb, double* tcb) {
    double u[3], w[3], q[3], dirb[3];
    ends(i, u, w, q);
#pragma unroll
    for (int j = 0; j < 3; ++j) dirb[j] = 0.5 * (Svb[j] + (0.0));
#pragma unroll
    for (int j = 0; j < 3; ++j) dirb[j] += 0.5 * Sg[i] * nh[j];
    cross3(w, dirb, ub);
    cross3(dirb, u, wb);
#pragma unroll
    for (int j = 0; j < 3; ++j) tcb[j] = Sn[i] * csb[j] / 3.0;
  };
  double ctrb[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double ub[3], wb[3], tcb[3];
    edge(i, ub, wb, tcb);
#pragma unroll
    for (int j = 0; j < 3; ++j) ctrb[j] = i < k ? ctrb[j] + (tcb[j] - ub[j] - wb[j]) : ctrb[j];
  }
  // incidence i collects edge i's ub + tcb and edge i-1's wb + tcb (for the first incidence the
  // last edge's, added after its own share); edge i-1's share is carried from the previous step
  double first[3], prev[3], last[3];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double ub[3], wb[3], tcb[3];
    edge(i, ub, wb, tcb);
    if (i < k) {
      const int64_t kk = (int64_t)kpos[i] * 3;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double mine = ub[j] + tcb[j];
        if (i == 0)
          first[j] = mine;
        else
          inc_b[kk + j] = (0.0 + prev[j]) + mine + ctrb[j] * invk;
      }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      prev[j] = wb[j] + tcb[j];
      last[j] = i < k ? prev[j] : last[j];
    }
  }
  const int64_t k0 = (int64_t)kpos[0] * 3;
#pragma unroll
  for (int j = 0; j < 3; ++j) inc_b[k0 + j] = ((0.0 + first[j]) + last[j]) + ctrb[j] * invk;
}

// Face pass: total cotangents of (area, oriented normal, centroid) of each face, then the reverse of
// its node construction into per-incidence node cotangents, each written once, at the incidence's
// node-major position: inc_b[kpos[i] * D + j] (the node pass then reads contiguous ranges).
template <int D>
__global__ void __launch_bounds__(256) k_geom_face_adj(MeshView m, int32_t Nn, const int32_t* __restrict__ foff,
                                const int32_t* __restrict__ fidx, const int32_t* __restrict__ kpos,
                                const double* __restrict__ X, const double* __restrict__ approx,
                                const double* __restrict__ cellbar, const double* __restrict__ area_b,
                                const double* __restrict__ nrm_b, const double* __restrict__ fc_b,
                                double* __restrict__ inc_b) {
  int32_t f = xcd_thread_index();
  if (f >= m.F) return;
  const int32_t P = m.owner[f], N = m.nb[f];
  const double A = m.area[f];
  double n[D], xf[D], nb[D], xfb[D];
  double Ab = area_b ? area_b[f] : 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    n[j] = m.nrm[(int64_t)j * m.F + f];
    xf[j] = m.fc[(int64_t)j * m.F + f];
    nb[j] = nrm_b ? nrm_b[(int64_t)j * m.F + f] : 0.0;
    xfb[j] = fc_b ? fc_b[(int64_t)j * m.F + f] : 0.0;
  }
  for (int side = 0; side < 2; ++side) {
    const int32_t c = side == 0 ? P : N;
    if (c < 0) continue;
    const double sg = side == 0 ? 1.0 : -1.0;
    double l[D], csb[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      l[j] = xf[j] - approx[(int64_t)j * m.C + c];
      csb[j] = cellbar[(int64_t)j * m.C + c];
    }
    const double Sb = cellbar[(int64_t)D * m.C + c];
    const double ln = vdot<D>(l, n);
    const double phi = ln * A;
    const double phib = sg * (Sb + vdot<D>(csb, l));
    const double inv_nf = 1.0 / (double)(m.cptr[c + 1] - m.cptr[c]);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      nb[j] += phib * A * l[j];
      xfb[j] += sg * csb[j] * phi + phib * A * n[j] + cellbar[(int64_t)(D + 1 + j) * m.C + c] * inv_nf;
    }
    Ab += phib * ln;
  }
  const int32_t b = foff[f], e = foff[f + 1];
  // orientation sign of the stored normal relative to the node-winding normal
  if (D == 2) {
    const int32_t a0 = fidx[b], a1 = fidx[b + 1];
    const double d0 = X[a1] - X[a0], d1 = X[Nn + a1] - X[Nn + a0];
    const double v0 = d1, v1 = -d0;
    const double vv = sqrt(v0 * v0 + v1 * v1);
    const double r0 = v0 / vv, r1 = v1 / vv;
    const double s = (r0 * n[0] + r1 * n[1]) < 0.0 ? -1.0 : 1.0;  // n = s * raw
    const double rb0 = s * nb[0], rb1 = s * nb[1];
    const double proj = rb0 * r0 + rb1 * r1;
    const double vb0 = (rb0 - proj * r0) / vv, vb1 = (rb1 - proj * r1) / vv;
    // area = |d|, raw normal = unit(d1, -d0), centroid = (p0 + p1) / 2
    const double db0 = Ab * d0 / A - vb1, db1 = Ab * d1 / A + vb0;
    const int64_t k0 = (int64_t)kpos[b] * 2, k1 = (int64_t)kpos[b + 1] * 2;
    inc_b[k0] = -db0 + 0.5 * xfb[0];
    inc_b[k0 + 1] = -db1 + 0.5 * xfb[1];
    inc_b[k1] = db0 + 0.5 * xfb[0];
    inc_b[k1 + 1] = db1 + 0.5 * xfb[1];
  } else if (e - b <= 4) {
    face_adj_small(e - b, fidx + b, kpos + b, Nn, X, n, nb, xf, xfb, Ab, inc_b);
  } else {
    const int32_t k = e - b;
    const double invk = 1.0 / (double)k;
    double ctr[3] = {0.0, 0.0, 0.0};
    for (int32_t i = b; i < e; ++i) {
      const int32_t a = fidx[i];
#pragma unroll
      for (int j = 0; j < 3; ++j) ctr[j] += X[(int64_t)j * Nn + a];
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) ctr[j] = ctr[j] * invk;
    double Sv[3] = {0.0, 0.0, 0.0};
    for (int32_t i = b; i < e; ++i) {
      const int32_t a = fidx[i], c2 = fidx[i + 1 < e ? i + 1 : b];
      double u[3], w[3], dir[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        u[j] = X[(int64_t)j * Nn + a] - ctr[j];
        w[j] = X[(int64_t)j * Nn + c2] - ctr[j];
      }
      cross3(u, w, dir);
#pragma unroll
      for (int j = 0; j < 3; ++j) Sv[j] += 0.5 * dir[j];
    }
    const double Ar = sqrt(Sv[0] * Sv[0] + Sv[1] * Sv[1] + Sv[2] * Sv[2]);
    double nh[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) nh[j] = Sv[j] / Ar;
    const double s = (nh[0] * n[0] + nh[1] * n[1] + nh[2] * n[2]) < 0.0 ? -1.0 : 1.0;
    double nhb[3], csb[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      nhb[j] = s * nb[j];
      csb[j] = xfb[j] / Ar;
    }
    // centroid = cs / A
    Ab -= (xfb[0] * xf[0] + xfb[1] * xf[1] + xfb[2] * xf[2]) / Ar;
    // sgn_i = dir_i . nh contributes sum_i sgn_b_i dir_i to nh's cotangent
    for (int32_t i = b; i < e; ++i) {
      const int32_t a = fidx[i], c2 = fidx[i + 1 < e ? i + 1 : b];
      double pb_[3], pc_[3], u[3], w[3], dir[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        pb_[j] = X[(int64_t)j * Nn + a];
        pc_[j] = X[(int64_t)j * Nn + c2];
        u[j] = pb_[j] - ctr[j];
        w[j] = pc_[j] - ctr[j];
      }
      cross3(u, w, dir);
      double sgnb = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) sgnb += csb[j] * ((ctr[j] + pb_[j] + pc_[j]) / 3.0);
#pragma unroll
      for (int j = 0; j < 3; ++j) nhb[j] += sgnb * 0.5 * dir[j];
    }
    // nh = Sv / |Sv|, A = |Sv|
    const double pr = nhb[0] * nh[0] + nhb[1] * nh[1] + nhb[2] * nh[2];
    double Svb[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) Svb[j] = Ab * nh[j] + (nhb[j] - pr * nh[j]) / Ar;
    // reverse of edge (i, i+1) of the centre fan: cotangents ub (node i), wb (node i+1) and the
    // centroid share tcb
    auto edge = [&](int32_t i, double* ub, double* wb, double* tcb) {
      const int32_t a = fidx[i], c2 = fidx[i + 1 < e ? i + 1 : b];
      double pb_[3], pc_[3], u[3], w[3], dir[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        pb_[j] = X[(int64_t)j * Nn + a];
        pc_[j] = X[(int64_t)j * Nn + c2];
        u[j] = pb_[j] - ctr[j];
        w[j] = pc_[j] - ctr[j];
      }
      cross3(u, w, dir);
      double sgn = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) sgn += 0.5 * dir[j] * nh[j];
      double dirb[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) dirb[j] = 0.5 * (Svb[j] + (0.0));
      // sgn_i = dir_i . nh (with dir_i = u x w / 2): cotangent sgnb_i adds sgnb_i nh / 2 to (u x w)'s
      double sgnb = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) sgnb += csb[j] * ((ctr[j] + pb_[j] + pc_[j]) / 3.0);
#pragma unroll
      for (int j = 0; j < 3; ++j) dirb[j] += 0.5 * sgnb * nh[j];
      // d(u x w) = du x w + u x dw:  u_b = w x dirb, w_b = dirb x u
      cross3(w, dirb, ub);
      cross3(dirb, u, wb);
#pragma unroll
      for (int j = 0; j < 3; ++j) tcb[j] = sgn * csb[j] / 3.0;
    };
    double ctrb[3] = {0.0, 0.0, 0.0};
    for (int32_t i = b; i < e; ++i) {
      double ub[3], wb[3], tcb[3];
      edge(i, ub, wb, tcb);
#pragma unroll
      for (int j = 0; j < 3; ++j) ctrb[j] += tcb[j] - ub[j] - wb[j];
    }
    // incidence i collects edge i's ub + tcb and edge i-1's wb + tcb, in the order the edges run
    // (edge i-1 first, except for the first incidence, whose previous edge is the last one)
    for (int32_t i = b; i < e; ++i) {
      const int32_t ip = i > b ? i - 1 : e - 1;
      double ub[3], wb[3], tcb[3], ub_p[3], wb_p[3], tcb_p[3];
      edge(i, ub, wb, tcb);
      edge(ip, ub_p, wb_p, tcb_p);
      const int64_t kk = (int64_t)kpos[i] * 3;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double mine = ub[j] + tcb[j], prev = wb_p[j] + tcb_p[j];
        const double t = i == b ? (0.0 + mine) + prev : (0.0 + prev) + mine;
        inc_b[kk + j] = t + ctrb[j] * invk;
      }
    }
  }
}

// Node pass: node cotangent = sum over the node's incidences.
constexpr int kNodeLanes = 8;  // lanes per node in the node pass of the geometry adjoint

template <int D>
__global__ void __launch_bounds__(256) k_geom_node_adj(int32_t Nn, const int32_t* __restrict__ nptr, const double* __restrict__ inc_b,
                                double* __restrict__ node_b) {
  // kNodeLanes consecutive lanes per node: lane l sums incidences l, l + kNodeLanes, ... of the
  // node's contiguous range (so a wave's loads cover contiguous memory), then an xor butterfly adds
  // the partial sums; every lane of the group forms the same total (each step adds the same two
  // values), so the result does not depend on which lane stores it.
  const int32_t tid = xcd_thread_index();
  const int32_t p = tid / kNodeLanes;
  const int l = tid % kNodeLanes;
  double s[D];
#pragma unroll
  for (int j = 0; j < D; ++j) s[j] = 0.0;
  if (p < Nn) {
    for (int32_t k = nptr[p] + l; k < nptr[p + 1]; k += kNodeLanes) {
#pragma unroll
      for (int j = 0; j < D; ++j) s[j] += inc_b[(int64_t)k * D + j];
    }
  }
#pragma unroll
  for (int off = kNodeLanes / 2; off > 0; off >>= 1) {
#pragma unroll
    for (int j = 0; j < D; ++j) s[j] += __shfl_xor(s[j], off, kNodeLanes);
  }
  if (p < Nn && l == 0) {
#pragma unroll
    for (int j = 0; j < D; ++j) node_b[(int64_t)j * Nn + p] = s[j];
  }
}

}  // namespace

double* afx_mesh_scratch(afx_mesh* m, int64_t n) {
  if (m->scratch_n >= n) return m->scratch;
  if (m->scratch) {
    if (hipFree(m->scratch) != hipSuccess) return nullptr;
    m->scratch = nullptr;
    m->scratch_n = 0;
  }
  if (hipMalloc((void**)&m->scratch, n * 8) != hipSuccess) {
    m->scratch = nullptr;
    return nullptr;
  }
  m->scratch_n = n;
  return m->scratch;
}

afx::MeshView afx_view(const afx_mesh* m, const afx_geometry* g) {
  afx::MeshView v;
  v.C = m->C;
  v.F = m->F;
  v.owner = m->owner;
  v.nb = m->nb;
  v.cptr = m->cptr;
  v.cface = m->cface;
  v.cnbr = m->cnbr;
  v.area = g ? g->area : nullptr;
  v.nrm = g ? g->normal : nullptr;
  v.fc = g ? g->face_centroid : nullptr;
  v.vol = g ? g->volume : nullptr;
  v.cc = g ? g->cell_centroid : nullptr;
  v.fu = nullptr;
  v.bsched = m->bsched;
  v.nbs = m->nbs;
  return v;
}

namespace {

// Face geometry in internal order: pack[(0 | 1+j | 1+D+j) F + i] = (area, normal_j, centroid_j)[fu[i]].
template <int D>
__global__ void __launch_bounds__(256) k_pack_faces(int32_t F, const int32_t* __restrict__ fu,
                                                     const double* __restrict__ area, const double* __restrict__ nrm,
                                                     const double* __restrict__ fc, double* __restrict__ pack) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= F) return;
  const int64_t u = fu[i];
  pack[i] = area[u];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    pack[(int64_t)(1 + j) * F + i] = nrm[(int64_t)j * F + u];
    pack[(int64_t)(1 + D + j) * F + i] = fc[(int64_t)j * F + u];
  }
}

int pack_faces(const afx_mesh* m, const afx_geometry* g, double* pack, hipStream_t s) {
  if (m->dim == 2)
    k_pack_faces<2><<<grid_for(m->F), dim3(kBlock), 0, s>>>(m->F, m->fu, g->area, g->normal, g->face_centroid, pack);
  else
    k_pack_faces<3><<<grid_for(m->F), dim3(kBlock), 0, s>>>(m->F, m->fu, g->area, g->normal, g->face_centroid, pack);
  AFX_CHECK_LAUNCH();
  return 0;
}

}  // namespace

namespace {
__global__ void __launch_bounds__(256) k_fill_zero(double* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
}  // namespace

int afx_fill_zero(double* p, int64_t n, hipStream_t s) {
  if (!p || n <= 0) return 0;
  k_fill_zero<<<grid_for(n), dim3(kBlock), 0, s>>>(p, n);
  AFX_CHECK_LAUNCH();
  return 0;
}

int afx_rview(const afx_mesh* m, const afx_geometry* g, hipStream_t s, afx::MeshView* out) {
  AFX_REQUIRE(m && g && out, "internal view: null argument");
  afx::MeshView v = afx_view(m, g);
  if (m->fu) {
    const double* pack = g->face_pack;
    if (!pack) {
      afx_mesh* mm = const_cast<afx_mesh*>(m);
      if (!mm->pack_scratch) AFX_CHECK_HIP(hipMalloc((void**)&mm->pack_scratch, (size_t)(2 * m->dim + 1) * m->F * 8));
      int rc = pack_faces(m, g, mm->pack_scratch, s);
      if (rc) return rc;
      pack = mm->pack_scratch;
    }
    v.owner = m->owner_i;
    v.nb = m->nb_i;
    v.cface = m->cface_i;
    v.cnbr = m->cnbr_i;
    v.fu = m->fu;
    v.area = pack;
    v.nrm = pack + m->F;
    v.fc = pack + (int64_t)(1 + m->dim) * m->F;
  }
  *out = v;
  return 0;
}

extern "C" int afx_geometry_pack(const afx_mesh* m, const afx_geometry* g, double* face_pack, hipStream_t s) {
  AFX_REQUIRE(m && g && face_pack, "afx_geometry_pack: null argument");
  AFX_REQUIRE(g->area && g->normal && g->face_centroid, "afx_geometry_pack: face geometry pointer missing");
  if (!m->fu) return 0;  // the caller's face order is the internal one: nothing to pack
  return pack_faces(m, g, face_pack, s);
}

extern "C" int afx_mesh_face_order(const afx_mesh* m, int32_t* permuted, int32_t* face_of_internal,
                                   hipStream_t s) {
  AFX_REQUIRE(m && permuted, "afx_mesh_face_order: null argument");
  *permuted = m->fu ? 1 : 0;
  if (m->fu && face_of_internal)
    AFX_CHECK_HIP(hipMemcpyAsync(face_of_internal, m->fu, (size_t)m->F * 4, hipMemcpyDeviceToDevice, s));
  return 0;
}

extern "C" int afx_mesh_set_block_order(afx_mesh* m, const int32_t* order, int64_t n_blocks) {
  AFX_REQUIRE(m, "afx_mesh_set_block_order: null mesh");
  const int64_t nb = ((int64_t)m->C + kBlock - 1) / kBlock;
  if (order == nullptr) {
    if (m->bsched) AFX_CHECK_HIP(hipDeviceSynchronize());
    if (m->bsched) (void)hipFree(m->bsched);
    m->bsched = nullptr;
    m->nbs = 0;
    return 0;
  }
  AFX_REQUIRE(n_blocks == nb, "afx_mesh_set_block_order: n_blocks must be ceil(n_cells / 256)");
  std::vector<uint8_t> seen((size_t)nb, 0);
  for (int64_t i = 0; i < nb; ++i) {
    const int32_t b = order[i];
    AFX_REQUIRE(b >= 0 && b < nb && !seen[(size_t)b], "afx_mesh_set_block_order: order is not a permutation");
    seen[(size_t)b] = 1;
  }
  AFX_CHECK_HIP(hipSetDevice(m->device));
  if (m->bsched == nullptr) AFX_CHECK_HIP(hipMalloc(&m->bsched, (size_t)nb * sizeof(int32_t)));
  AFX_CHECK_HIP(hipMemcpy(m->bsched, order, (size_t)nb * sizeof(int32_t), hipMemcpyHostToDevice));
  m->nbs = (int32_t)nb;
  return 0;
}

namespace {

// approx: where the approximate cell centroids go (the mesh's own buffer, or a caller's record).
int geometry_compute(const afx_mesh* m, const double* node_coords, afx_geometry* g, double* approx,
                     hipStream_t s) {
  AFX_REQUIRE(m && node_coords && g, "afx_geometry_compute: null argument");
  AFX_REQUIRE(g->area && g->normal && g->face_centroid && g->volume && g->cell_centroid,
              "afx_geometry_compute: geometry output pointer missing");
  double* area = const_cast<double*>(g->area);
  double* nrm = const_cast<double*>(g->normal);
  double* fc = const_cast<double*>(g->face_centroid);
  double* vol = const_cast<double*>(g->volume);
  double* cc = const_cast<double*>(g->cell_centroid);
  afx::MeshView v = afx_view(m, g);
  if (m->dim == 2) {
    AFX_LAUNCH(kKGeomFace, k_face_geometry<2>, grid_for(m->F), dim3(kBlock), 0, s, m->F, m->N, m->foff,
                       m->fidx, node_coords, area, m->normal_raw, fc);
    AFX_CHECK_LAUNCH();
    AFX_LAUNCH(kKGeomApprox, k_approx_centroid<2>, grid_for(m->C), dim3(kBlock), 0, s, v, approx);
    AFX_CHECK_LAUNCH();
    AFX_LAUNCH(kKGeomOrient, k_orient<2>, grid_for(m->F), dim3(kBlock), 0, s, v, approx, m->normal_raw, nrm);
    AFX_CHECK_LAUNCH();
    AFX_LAUNCH(kKGeomVolume, k_cell_volume<2>, grid_for(m->C), dim3(kBlock), 0, s, v, approx, vol, cc);
    AFX_CHECK_LAUNCH();
  } else {
    AFX_LAUNCH(kKGeomFace, k_face_geometry<3>, grid_for(m->F), dim3(kBlock), 0, s, m->F, m->N, m->foff,
                       m->fidx, node_coords, area, m->normal_raw, fc);
    AFX_CHECK_LAUNCH();
    AFX_LAUNCH(kKGeomApprox, k_approx_centroid<3>, grid_for(m->C), dim3(kBlock), 0, s, v, approx);
    AFX_CHECK_LAUNCH();
    AFX_LAUNCH(kKGeomOrient, k_orient<3>, grid_for(m->F), dim3(kBlock), 0, s, v, approx, m->normal_raw, nrm);
    AFX_CHECK_LAUNCH();
    AFX_LAUNCH(kKGeomVolume, k_cell_volume<3>, grid_for(m->C), dim3(kBlock), 0, s, v, approx, vol, cc);
    AFX_CHECK_LAUNCH();
  }
  if (g->face_pack && m->fu) return pack_faces(m, g, const_cast<double*>(g->face_pack), s);
  return 0;
}

}  // namespace

extern "C" int afx_geometry_compute(const afx_mesh* m, const double* node_coords, afx_geometry* g,
                                    hipStream_t s) {
  AFX_REQUIRE(m, "afx_geometry_compute: null argument");
  return geometry_compute(m, node_coords, g, m->approx, s);
}

extern "C" int afx_geometry_compute_record(const afx_mesh* m, const double* node_coords, afx_geometry* g,
                                           double* approx, hipStream_t s) {
  AFX_REQUIRE(m && approx, "afx_geometry_compute_record: null argument");
  return geometry_compute(m, node_coords, g, approx, s);
}

extern "C" int afx_geometry_vjp_workspace_bytes(const afx_mesh* m, int64_t* bytes) {
  AFX_REQUIRE(m && bytes, "afx_geometry_vjp_workspace_bytes: null argument");
  const int64_t D = m->dim;
  *bytes = ((D + (2 * D + 1)) * (int64_t)m->C + D * (int64_t)m->n_inc) * 8 + 256;
  return 0;
}

namespace {

// record: null (the approximate centroids are re-formed into the workspace) or the forward's.
int geometry_vjp(const afx_mesh* m, const double* node_coords, const afx_geometry* g, const double* record,
                 const afx_geometry_cotangents* gb, double* node_b, void* work, int64_t work_bytes, hipStream_t s) {
  AFX_REQUIRE(m && node_coords && g && gb && node_b && work, "afx_geometry_vjp: null argument");
  AFX_REQUIRE(g->area && g->normal && g->face_centroid && g->volume && g->cell_centroid,
              "afx_geometry_vjp: geometry pointer missing");
  int64_t need = 0;
  afx_geometry_vjp_workspace_bytes(m, &need);
  AFX_REQUIRE(work_bytes >= need - 256, "afx_geometry_vjp: workspace too small");
  AFX_REQUIRE(m->C > 0 && m->F > 0 && m->N > 0, "afx_geometry_vjp: empty mesh");
  AFX_REQUIRE((int64_t)m->N * kNodeLanes < INT32_MAX, "afx_geometry_vjp: too many nodes for the node pass");
  const int D = m->dim;
  double* approx = record ? const_cast<double*>(record) : (double*)work;
  double* cellbar = (double*)work + (int64_t)D * m->C;
  double* inc_b = cellbar + (int64_t)(2 * D + 1) * m->C;
  afx::MeshView v = afx_view(m, g);
  if (D == 2) {
    if (record)
      AFX_LAUNCH(kKGeomAdj, (k_geom_cell_adj<2, false>), grid_for(m->C), dim3(kBlock), 0, s, v, approx, gb->volume,
                 gb->cell_centroid, cellbar);
    else
      AFX_LAUNCH(kKGeomAdj, (k_geom_cell_adj<2, true>), grid_for(m->C), dim3(kBlock), 0, s, v, approx, gb->volume,
                 gb->cell_centroid, cellbar);
    AFX_CHECK_LAUNCH();
    AFX_LAUNCH(kKGeomAdjFace, k_geom_face_adj<2>, grid_for(m->F), dim3(kBlock), 0, s, v, m->N, m->foff, m->fidx,
               m->kpos, node_coords, approx, cellbar, gb->area, gb->normal, gb->face_centroid, inc_b);
    AFX_CHECK_LAUNCH();
    AFX_LAUNCH(kKGeomAdjNode, k_geom_node_adj<2>, grid_for((int64_t)m->N * kNodeLanes), dim3(kBlock), 0, s, m->N, m->nptr, inc_b,
               node_b);
    AFX_CHECK_LAUNCH();
  } else {
    if (record)
      AFX_LAUNCH(kKGeomAdj, (k_geom_cell_adj<3, false>), grid_for(m->C), dim3(kBlock), 0, s, v, approx, gb->volume,
                 gb->cell_centroid, cellbar);
    else
      AFX_LAUNCH(kKGeomAdj, (k_geom_cell_adj<3, true>), grid_for(m->C), dim3(kBlock), 0, s, v, approx, gb->volume,
                 gb->cell_centroid, cellbar);
    AFX_CHECK_LAUNCH();
    AFX_LAUNCH(kKGeomAdjFace, k_geom_face_adj<3>, grid_for(m->F), dim3(kBlock), 0, s, v, m->N, m->foff, m->fidx,
               m->kpos, node_coords, approx, cellbar, gb->area, gb->normal, gb->face_centroid, inc_b);
    AFX_CHECK_LAUNCH();
    AFX_LAUNCH(kKGeomAdjNode, k_geom_node_adj<3>, grid_for((int64_t)m->N * kNodeLanes), dim3(kBlock), 0, s, m->N, m->nptr, inc_b,
               node_b);
    AFX_CHECK_LAUNCH();
  }
  return 0;
}

}  // namespace

extern "C" int afx_geometry_vjp(const afx_mesh* m, const double* node_coords, const afx_geometry* g,
                                const afx_geometry_cotangents* gb, double* node_b, void* work, int64_t work_bytes,
                                hipStream_t s) {
  return geometry_vjp(m, node_coords, g, nullptr, gb, node_b, work, work_bytes, s);
}

extern "C" int afx_geometry_vjp_record(const afx_mesh* m, const double* node_coords, const afx_geometry* g,
                                       const double* approx, const afx_geometry_cotangents* gb, double* node_b,
                                       void* work, int64_t work_bytes, hipStream_t s) {
  AFX_REQUIRE(approx, "afx_geometry_vjp_record: null record");
  return geometry_vjp(m, node_coords, g, approx, gb, node_b, work, work_bytes, s);
}

// ---- face mass flux (u . n) A ------------------------------------------------------------------
// The verification flux of a prescribed velocity (reference: tests/support/fields.py:16-43), one
// pass over the faces instead of a matrix product over the (d, F) normals.

namespace {

template <int D, bool PF>
__global__ void __launch_bounds__(256) k_face_flux(int64_t F, const double* __restrict__ u, const double* __restrict__ n,
                                                   const double* __restrict__ area, double* __restrict__ mdot) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  double un = 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) un += (PF ? u[(int64_t)j * F + f] : u[j]) * n[(int64_t)j * F + f];
  mdot[f] = un * area[f];
}

template <int D, bool PF>
__global__ void __launch_bounds__(256) k_face_flux_vjp(int64_t F, const double* __restrict__ u,
                                                       const double* __restrict__ n, const double* __restrict__ area,
                                                       const double* __restrict__ mb, double* __restrict__ nb,
                                                       double* __restrict__ ab, double* __restrict__ ub) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const double g = mb[f], A = area[f];
  double uf[D], nf[D], un = 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    uf[j] = PF ? u[(int64_t)j * F + f] : u[j];
    nf[j] = n[(int64_t)j * F + f];
    un += uf[j] * nf[j];
  }
  if (nb) {
#pragma unroll
    for (int j = 0; j < D; ++j) nb[(int64_t)j * F + f] = uf[j] * A * g;
  }
  if (ab) ab[f] = un * g;
  if (PF && ub) {
#pragma unroll
    for (int j = 0; j < D; ++j) ub[(int64_t)j * F + f] = nf[j] * A * g;
  }
}

template <class F2>
int by_dim_pf(int dim, int per_face, F2&& f) {
  if (dim == 2) return per_face ? f(std::integral_constant<int, 2>{}, std::true_type{})
                                : f(std::integral_constant<int, 2>{}, std::false_type{});
  return per_face ? f(std::integral_constant<int, 3>{}, std::true_type{})
                  : f(std::integral_constant<int, 3>{}, std::false_type{});
}

}  // namespace

extern "C" int afx_face_flux(int64_t F, int dim, int per_face, const double* u, const double* normal,
                             const double* area, double* mdot, hipStream_t s) {
  AFX_REQUIRE(F > 0 && (dim == 2 || dim == 3) && (per_face == 0 || per_face == 1) && u && normal && area && mdot,
              "afx_face_flux: bad argument");
  return by_dim_pf(dim, per_face, [&](auto Dc, auto Pc) {
    constexpr int D = decltype(Dc)::value;
    constexpr bool PF = decltype(Pc)::value;
    AFX_LAUNCH(kKFaceFlux, (k_face_flux<D, PF>), grid_for(F), dim3(kBlock), 0, s, F, u, normal, area, mdot);
    AFX_CHECK_LAUNCH();
    return 0;
  });
}

extern "C" int afx_face_flux_vjp(int64_t F, int dim, int per_face, const double* u, const double* normal,
                                 const double* area, const double* mdot_bar, double* normal_bar, double* area_bar,
                                 double* u_bar, hipStream_t s) {
  AFX_REQUIRE(F > 0 && (dim == 2 || dim == 3) && (per_face == 0 || per_face == 1) && u && normal && area &&
                  mdot_bar,
              "afx_face_flux_vjp: bad argument");
  AFX_REQUIRE(per_face == 1 || u_bar == nullptr, "afx_face_flux_vjp: u_bar needs a per-face velocity");
  return by_dim_pf(dim, per_face, [&](auto Dc, auto Pc) {
    constexpr int D = decltype(Dc)::value;
    constexpr bool PF = decltype(Pc)::value;
    AFX_LAUNCH(kKFaceFlux, (k_face_flux_vjp<D, PF>), grid_for(F), dim3(kBlock), 0, s, F, u, normal, area, mdot_bar,
               normal_bar, area_bar, u_bar);
    AFX_CHECK_LAUNCH();
    return 0;
  });
}

// ---- geometry tangent: the four geometry passes on dual numbers --------------------------------
// Node coordinates with a tangent (X, Xt) -> the geometry and its tangent, each pass the forward
// pass's operations in the same order on `dual` values (so the values are the forward's bits and
// the tangents are exact derivatives; branch decisions read the values).

namespace {

template <int D>
__global__ void __launch_bounds__(256) k_face_geometry_t(int32_t F, int32_t Nn, const int32_t* __restrict__ foff,
                                                         const int32_t* __restrict__ fidx, In<dual> X, Out<dual> area,
                                                         Out<dual> raw, Out<dual> fc) {
  const int32_t f = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (f >= F) return;
  const int32_t b = foff[f], e = foff[f + 1];
  if (D == 2) {
    const int32_t a0 = fidx[b], a1 = fidx[b + 1];
    const dual p0[2] = {X[a0], X[Nn + a0]}, p1[2] = {X[a1], X[Nn + a1]};
    const dual d0 = p1[0] - p0[0], d1 = p1[1] - p0[1];
    const dual sq = d0 * d0 + d1 * d1;
    area.set(f, sq.v > 0.0 ? dsqrt(sq) : dual{0.0, 0.0});
    fc.set(f, 0.5 * (p0[0] + p1[0]));
    fc.set(F + f, 0.5 * (p0[1] + p1[1]));
    const dual v0 = d1, v1 = -d0;
    const dual s2 = v0 * v0 + v1 * v1;
    const dual inv = s2.v > 0.0 ? dsqrt(s2) : dual{1.0, 0.0};
    raw.set(f, s2.v > 0.0 ? v0 / inv : dual{0.0, 0.0});
    raw.set(F + f, s2.v > 0.0 ? v1 / inv : dual{0.0, 0.0});
  } else {
    const int32_t k = e - b;
    dual ctr[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (int32_t i = b; i < e; ++i) {
      const int32_t a = fidx[i];
#pragma unroll
      for (int j = 0; j < 3; ++j) ctr[j] += X[(int64_t)j * Nn + a];
    }
    const double invk = 1.0 / (double)k;
#pragma unroll
    for (int j = 0; j < 3; ++j) ctr[j] = ctr[j] * invk;
    dual S[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (int32_t i = b; i < e; ++i) {
      const int32_t a = fidx[i], c = fidx[i + 1 < e ? i + 1 : b];
      dual u[3], w[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        u[j] = X[(int64_t)j * Nn + a] - ctr[j];
        w[j] = X[(int64_t)j * Nn + c] - ctr[j];
      }
      S[0] += 0.5 * (u[1] * w[2] - u[2] * w[1]);
      S[1] += 0.5 * (u[2] * w[0] - u[0] * w[2]);
      S[2] += 0.5 * (u[0] * w[1] - u[1] * w[0]);
    }
    const dual sq = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
    const dual A = sq.v > 0.0 ? dsqrt(sq) : dual{0.0, 0.0};
    dual nn[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) nn[j] = sq.v > 0.0 ? S[j] / A : dual{0.0, 0.0};
    dual cs[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (int32_t i = b; i < e; ++i) {
      const int32_t a = fidx[i], c = fidx[i + 1 < e ? i + 1 : b];
      dual pb[3], pc[3], u[3], w[3], dir[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        pb[j] = X[(int64_t)j * Nn + a];
        pc[j] = X[(int64_t)j * Nn + c];
        u[j] = pb[j] - ctr[j];
        w[j] = pc[j] - ctr[j];
      }
      dir[0] = 0.5 * (u[1] * w[2] - u[2] * w[1]);
      dir[1] = 0.5 * (u[2] * w[0] - u[0] * w[2]);
      dir[2] = 0.5 * (u[0] * w[1] - u[1] * w[0]);
      const dual sgn = dir[0] * nn[0] + dir[1] * nn[1] + dir[2] * nn[2];
#pragma unroll
      for (int j = 0; j < 3; ++j) cs[j] += sgn * ((ctr[j] + pb[j] + pc[j]) / 3.0);
    }
    const dual safe = A.v > 0.0 ? A : dual{1.0, 0.0};
    area.set(f, A);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      raw.set((int64_t)j * F + f, nn[j]);
      fc.set((int64_t)j * F + f, cs[j] / safe);
    }
  }
}

template <int D>
__global__ void __launch_bounds__(256) k_approx_centroid_t(MeshView m, In<dual> fc, Out<dual> approx) {
  const int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= m.C) return;
  dual s[D];
#pragma unroll
  for (int j = 0; j < D; ++j) s[j] = dual{0.0, 0.0};
  const int32_t b = m.cptr[c], e = m.cptr[c + 1];
  for (int32_t k = b; k < e; ++k) {
    const int32_t ent = m.cface[k];
    const int32_t f = ent >= 0 ? ent : ~ent;
#pragma unroll
    for (int j = 0; j < D; ++j) s[j] += fc[(int64_t)j * m.F + f];
  }
  const double inv = 1.0 / (double)(e - b);
#pragma unroll
  for (int j = 0; j < D; ++j) approx.set((int64_t)j * m.C + c, s[j] * inv);
}

template <int D>
__global__ void __launch_bounds__(256) k_orient_t(MeshView m, In<dual> fc, In<dual> approx, In<dual> raw,
                                                  Out<dual> nrm) {
  const int32_t f = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (f >= m.F) return;
  const int32_t P = m.owner[f];
  double proj = 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) proj += (fc[(int64_t)j * m.F + f].v - approx[(int64_t)j * m.C + P].v) * raw[(int64_t)j * m.F + f].v;
  const double sgn = proj < 0.0 ? -1.0 : 1.0;
#pragma unroll
  for (int j = 0; j < D; ++j) nrm.set((int64_t)j * m.F + f, raw[(int64_t)j * m.F + f] * sgn);
}

template <int D>
__global__ void __launch_bounds__(256) k_cell_volume_t(MeshView m, In<dual> approx, In<dual> fc, In<dual> nrm,
                                                       In<dual> area, Out<dual> vol, Out<dual> cc) {
  const int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= m.C) return;
  dual ap[D], v = {0.0, 0.0}, cs[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    ap[j] = approx[(int64_t)j * m.C + c];
    cs[j] = dual{0.0, 0.0};
  }
  for (int32_t k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
    const int32_t ent = m.cface[k];
    const int32_t f = ent >= 0 ? ent : ~ent;
    dual loc[D], nl[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      loc[j] = fc[(int64_t)j * m.F + f] - ap[j];
      nl[j] = nrm[(int64_t)j * m.F + f];
    }
    dual dot = loc[0] * nl[0];
#pragma unroll
    for (int j = 1; j < D; ++j) dot = dot + loc[j] * nl[j];
    const dual flux = dot * area[f];
    if (ent >= 0) {
      v += flux;
#pragma unroll
      for (int j = 0; j < D; ++j) cs[j] += loc[j] * flux;
    } else {
      v += -flux;
#pragma unroll
      for (int j = 0; j < D; ++j) cs[j] += -(loc[j] * flux);
    }
  }
  const dual V = v / (double)D;
  vol.set(c, V);
  const dual inv = 1.0 / ((double)(D + 1) * V);
#pragma unroll
  for (int j = 0; j < D; ++j) cc.set((int64_t)j * m.C + c, cs[j] * inv + ap[j]);
}

template <int D>
int geometry_jvp(const afx_mesh* m, const double* X, const double* Xt, afx_geometry* g, afx_geometry* t, double* work,
                 hipStream_t s) {
  const int64_t C = m->C, F = m->F;
  double* approx = work;
  double* approx_t = approx + D * C;
  double* raw = approx_t + D * C;
  double* raw_t = raw + D * F;
  afx::MeshView v = afx_view(m, nullptr);
  const double* fcv = g->face_centroid;
  const double* fct = t->face_centroid;
  AFX_LAUNCH(kKGeomFace, k_face_geometry_t<D>, grid_for(F), dim3(kBlock), 0, s, m->F, m->N, m->foff, m->fidx,
             In<dual>{X, Xt}, Out<dual>{const_cast<double*>(g->area), const_cast<double*>(t->area)},
             Out<dual>{raw, raw_t}, Out<dual>{const_cast<double*>(fcv), const_cast<double*>(fct)});
  AFX_CHECK_LAUNCH();
  AFX_LAUNCH(kKGeomApprox, k_approx_centroid_t<D>, grid_for(C), dim3(kBlock), 0, s, v, In<dual>{fcv, fct},
             Out<dual>{approx, approx_t});
  AFX_CHECK_LAUNCH();
  AFX_LAUNCH(kKGeomOrient, k_orient_t<D>, grid_for(F), dim3(kBlock), 0, s, v, In<dual>{fcv, fct},
             In<dual>{approx, approx_t}, In<dual>{raw, raw_t},
             Out<dual>{const_cast<double*>(g->normal), const_cast<double*>(t->normal)});
  AFX_CHECK_LAUNCH();
  AFX_LAUNCH(kKGeomVolume, k_cell_volume_t<D>, grid_for(C), dim3(kBlock), 0, s, v, In<dual>{approx, approx_t},
             In<dual>{fcv, fct}, In<dual>{g->normal, t->normal}, In<dual>{g->area, t->area},
             Out<dual>{const_cast<double*>(g->volume), const_cast<double*>(t->volume)},
             Out<dual>{const_cast<double*>(g->cell_centroid), const_cast<double*>(t->cell_centroid)});
  AFX_CHECK_LAUNCH();
  return 0;
}

}  // namespace

extern "C" int afx_geometry_jvp_workspace_bytes(const afx_mesh* m, int64_t* bytes) {
  AFX_REQUIRE(m && bytes, "afx_geometry_jvp_workspace_bytes: null argument");
  *bytes = (int64_t)2 * m->dim * ((int64_t)m->C + m->F) * 8 + 256;
  return 0;
}

extern "C" int afx_geometry_jvp(const afx_mesh* m, const double* node_coords, const double* node_tangent,
                                afx_geometry* out, afx_geometry* tangent, void* work, int64_t work_bytes,
                                hipStream_t s) {
  AFX_REQUIRE(m && node_coords && node_tangent && out && tangent && work, "afx_geometry_jvp: null argument");
  AFX_REQUIRE(out->area && out->normal && out->face_centroid && out->volume && out->cell_centroid,
              "afx_geometry_jvp: geometry output pointer missing");
  AFX_REQUIRE(tangent->area && tangent->normal && tangent->face_centroid && tangent->volume && tangent->cell_centroid,
              "afx_geometry_jvp: tangent output pointer missing");
  int64_t need = 0;
  afx_geometry_jvp_workspace_bytes(m, &need);
  AFX_REQUIRE(work_bytes >= need - 256, "afx_geometry_jvp: workspace too small");
  if (m->dim == 2) return geometry_jvp<2>(m, node_coords, node_tangent, out, tangent, (double*)work, s);
  return geometry_jvp<3>(m, node_coords, node_tangent, out, tangent, (double*)work, s);
}
