// Internal definition of the opaque mesh handle.
#pragma once

#include "aquaflux_amd.h"
#include "afx_common.h"

struct afx_mesh {
  int dim = 0;
  int32_t C = 0, F = 0, N = 0, n_inc = 0, n_cface = 0, max_faces = 0;
  int device = 0;
  int32_t* owner = nullptr;
  int32_t* nb = nullptr;
  int32_t* cptr = nullptr;      // [C+1]
  int32_t* cface = nullptr;     // [n_cface]  f (owner side) or ~f (neighbour side)
  int32_t* cnbr = nullptr;      // [n_cface]  the other cell of each entry (-1: boundary face)
  int32_t* foff = nullptr;      // [F+1] face -> node CSR
  int32_t* fidx = nullptr;      // [n_inc]
  int32_t* inc_face = nullptr;  // [n_inc] face of each incidence
  int32_t* nptr = nullptr;      // [N+1] node -> incidence CSR
  int32_t* ninc = nullptr;      // [n_inc]
  int32_t* kpos = nullptr;      // [n_inc] position of each incidence in the node-major list (ninc^-1)
  double* approx = nullptr;     // [d][C] scratch: approximate cell centroids
  double* normal_raw = nullptr; // [d][F] scratch: winding-order normals
  double* scratch = nullptr;    // lazily grown scratch of entry points without a workspace argument
  int64_t scratch_n = 0;
  // Internal face order of the residual kernels (faces by rank among their owner's faces, then
  // owner), so the faces a run of consecutive cells visits are consecutive records.  Null when the
  // caller's faces are already sorted by owner.
  int32_t* fu = nullptr;        // [F] internal -> caller face index
  int32_t* owner_i = nullptr;   // [F] owner / neighbour in internal order
  int32_t* nb_i = nullptr;
  int32_t* cface_i = nullptr;   // [n_cface] cell -> face list over internal face indices
  int32_t* cnbr_i = nullptr;    // [n_cface] its other cells
  double* pack_scratch = nullptr;  // [(2d+1)][F] face geometry gathered for a call without face_pack
  int32_t* bsched = nullptr;    // [nbs] block order of the cell kernels (afx_mesh_set_block_order)
  int32_t nbs = 0;
};

// Device scratch of at least n doubles owned by the handle (grown on demand, freed with it);
// nullptr on allocation failure.  Not for concurrent use from several host threads.
double* afx_mesh_scratch(afx_mesh* m, int64_t n);

// Zero n doubles on `s` with a fill kernel (capturable in a graph, unlike a memset node).
int afx_fill_zero(double* p, int64_t n, hipStream_t s);

// Caller-order view (the geometry kernels).
afx::MeshView afx_view(const afx_mesh* m, const afx_geometry* g);

// Internal-order view of the residual kernels: internal topology and face geometry, from
// g->face_pack or (when that is null) gathered on `s` into the handle's pack scratch.
int afx_rview(const afx_mesh* m, const afx_geometry* g, hipStream_t s, afx::MeshView* out);
