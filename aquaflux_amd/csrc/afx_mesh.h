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
  int32_t* foff = nullptr;      // [F+1] face -> node CSR
  int32_t* fidx = nullptr;      // [n_inc]
  int32_t* inc_face = nullptr;  // [n_inc] face of each incidence
  int32_t* nptr = nullptr;      // [N+1] node -> incidence CSR
  int32_t* ninc = nullptr;      // [n_inc]
  double* approx = nullptr;     // [d][C] scratch: approximate cell centroids
  double* normal_raw = nullptr; // [d][F] scratch: winding-order normals
  double* scratch = nullptr;    // lazily grown scratch of entry points without a workspace argument
  int64_t scratch_n = 0;
};

// Device scratch of at least n doubles owned by the handle (grown on demand, freed with it);
// nullptr on allocation failure.  Not for concurrent use from several host threads.
double* afx_mesh_scratch(afx_mesh* m, int64_t n);

afx::MeshView afx_view(const afx_mesh* m, const afx_geometry* g);
