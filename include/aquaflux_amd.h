/* aquaflux_amd -- C ABI of the MI355X (gfx950) finite-volume residual library.
 *
 * Every entry point returns 0 on success, 1 on a HIP runtime failure and 2 on an invalid argument;
 * `afx_last_error()` returns the calling thread's last message.  Nothing throws across the ABI.
 * All array arguments are DEVICE pointers on the handle's device unless stated otherwise, laid
 * out structure-of-arrays: a per-face vector field is `v[j * n_faces + f]`, a per-cell one
 * `v[j * n_cells + c]`.  The caller owns every buffer; the mesh handle owns only its topology
 * copies and scratch.  Work is enqueued on `stream` and is stream-ordered; calls on one handle
 * from several host threads must use distinct workspaces.
 *
 * The state vector keeps the reference's field-major layout
 * [u-block, v-block, (w-block), p-block, ...], each block n_cells long
 * (reference: aquaflux/flow/state.py:36-69).
 */
#ifndef AQUAFLUX_AMD_H
#define AQUAFLUX_AMD_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------- */
/* errors                                                                                        */

/* Message of the calling thread's last failure ("" if none). */
const char* afx_last_error(void);

/* ------------------------------------------------------------------------------------------- */
/* kernel timing (no reference counterpart: the reference has no profiler hook)                  */

/* Enables per-launch event timing with a pool of max_launches event pairs (<= 0 disables). */
int afx_profile_enable(int64_t max_launches);
/* Per-kernel total milliseconds and launch counts since enabling (arrays of
 * afx_profile_kernel_count() entries); waits for the recorded events and clears them. */
int afx_profile_collect(double* total_ms, int64_t* count);
int afx_profile_kernel_count(void);
const char* afx_profile_kernel_name(int id);

/* ------------------------------------------------------------------------------------------- */
/* mesh                                                                                          */

typedef struct afx_mesh afx_mesh;

/* Builds a device mesh handle on the current device from HOST arrays.
 * Replaces the connectivity construction of Mesh.from_csr / FaceCellConnectivity /
 * FaceNodeConnectivity (reference: aquaflux/mesh/mesh.py:137-206,
 * aquaflux/mesh/connectivity.py:114-136, :399-459): owner[f], neighbour[f] (-1 = boundary face),
 * face -> node CSR (face_node_offsets[n_faces + 1], face_node_indices[...]). */
int afx_mesh_create(int dim, int64_t n_cells, int64_t n_faces, int64_t n_nodes, const int32_t* owner,
                    const int32_t* neighbour, const int64_t* face_node_offsets,
                    const int32_t* face_node_indices, afx_mesh** out);
int afx_mesh_destroy(afx_mesh* mesh);
/* out[7] = {n_cells, n_faces, n_nodes, dim, n_cell_face_entries, max_faces_per_cell, device} */
int afx_mesh_info(const afx_mesh* mesh, int64_t* out);
/* Copies the cell -> face list (entries f when the cell owns f, ~f when it is f's neighbour). */
int afx_mesh_cell_faces(const afx_mesh* mesh, int32_t* cell_ptr, int32_t* cell_faces, hipStream_t stream);

/* Block order of the cell kernels (no reference counterpart: a scheduling choice of this build).
 * order[n_blocks] (host memory, n_blocks = ceil(n_cells / 256)) is a permutation of the 256-cell
 * blocks: logical workgroup i of a one-thread-per-cell launch processes block order[i].  Results
 * are unchanged (every cell is still computed by exactly one thread, with the same arithmetic);
 * what changes is which cells an XCD holds in its L2 at the same time.  order == NULL restores the
 * memory order.  Synchronous; not to be called while kernels on this mesh are in flight. */
int afx_mesh_set_block_order(afx_mesh* mesh, const int32_t* order, int64_t n_blocks);

/* The residual kernels visit faces in an internal order: when the caller's faces are not sorted by
 * owner, by (rank among the owner's faces, owner), so that consecutive cells read consecutive face
 * records.  *permuted = 0 when the caller's order is kept (it is sorted by owner); otherwise 1, and
 * face_of_internal[i] (nullable, [F] device) receives the caller's index of internal face i.  No
 * reference counterpart (a layout choice of this library). */
int afx_mesh_face_order(const afx_mesh* mesh, int32_t* permuted, int32_t* face_of_internal, hipStream_t stream);

/* ------------------------------------------------------------------------------------------- */
/* geometry                                                                                      */

typedef struct {
  const double* area;          /* [F]    */
  const double* normal;        /* [d][F] owner-outward unit normal */
  const double* face_centroid; /* [d][F] */
  const double* volume;        /* [C]    */
  const double* cell_centroid; /* [d][C] */
  const double* face_pack;     /* optional [(2d+1)][F]: area, normal, centroid in the internal face order
                                  (afx_mesh_face_order), filled by afx_geometry_compute or
                                  afx_geometry_pack; NULL makes every residual call gather it, and it
                                  is unused when the mesh reports no permutation */
} afx_geometry;

/* Face/cell geometry from node coordinates node_coords[d][N] into the (writable) buffers of `out`.
 * Replaces Mesh.geometry / _fused_geometry (reference: aquaflux/mesh/mesh.py:341-387,
 * aquaflux/mesh/face.py:196-349, aquaflux/mesh/cell.py:84-170). */
int afx_geometry_compute(const afx_mesh* mesh, const double* node_coords, afx_geometry* out, hipStream_t stream);

/* Forward-mode derivative of afx_geometry_compute: from node coordinates node_coords[d][N] and a
 * tangent node_tangent[d][N], writes the geometry into `out` (afx_geometry_compute's values up to
 * rounding; face_pack untouched) and its tangent into `tangent` (area, normal, face centroid, volume,
 * cell centroid; face_pack unused).  Workspace from afx_geometry_jvp_workspace_bytes.  Replaces
 * jax.jvp through Mesh.geometry (reference: aquaflux/mesh/mesh.py:341-387,
 * aquaflux/mesh/face.py:196-349, aquaflux/mesh/cell.py:84-170). */
int afx_geometry_jvp_workspace_bytes(const afx_mesh* mesh, int64_t* bytes);
int afx_geometry_jvp(const afx_mesh* mesh, const double* node_coords, const double* node_tangent, afx_geometry* out,
                     afx_geometry* tangent, void* work, int64_t work_bytes, hipStream_t stream);

/* Writes geom's face arrays in the internal face order into face_pack[(2d+1)][F] (a no-op when the
 * mesh reports no permutation).  For geometry not produced by afx_geometry_compute.  No reference
 * counterpart. */
int afx_geometry_pack(const afx_mesh* mesh, const afx_geometry* geom, double* face_pack, hipStream_t stream);

/* Cotangents of the five geometry arrays (written, not accumulated; NULL = not requested). */
typedef struct {
  double* area;          /* [F]    */
  double* normal;        /* [d][F] */
  double* face_centroid; /* [d][F] */
  double* volume;        /* [C]    */
  double* cell_centroid; /* [d][C] */
} afx_geometry_cotangents;

/* Device workspace of afx_geometry_vjp. */
int afx_geometry_vjp_workspace_bytes(const afx_mesh* mesh, int64_t* bytes);

/* node_b[d][N] = (d geometry / d node_coords)^T . geometry cotangents, recomputing the face
 * construction (centre-fan triangles / edges), the approximate centroids and the divergence-theorem
 * sums from the nodes and the forward geometry `geom`.  Replaces jax.vjp through Mesh.geometry
 * (reference: aquaflux/mesh/mesh.py:341-387, aquaflux/mesh/face.py:196-349,
 * aquaflux/mesh/cell.py:84-170). */
int afx_geometry_vjp(const afx_mesh* mesh, const double* node_coords, const afx_geometry* geom,
                     const afx_geometry_cotangents* geom_cotangents, double* node_b, void* work,
                     int64_t work_bytes, hipStream_t stream);

/* afx_geometry_compute / afx_geometry_vjp split at a forward record, as jax.vjp through
 * Mesh.geometry keeps its linearisation point (reference: aquaflux/mesh/mesh.py:341-387): the
 * forward also writes the approximate cell centroids approx[d][C] (the mean of each cell's face
 * centroids, aquaflux/mesh/cell.py:84-112) into the caller's buffer, and the reverse pass reads them
 * instead of forming them again.  afx_geometry_vjp_record returns what afx_geometry_vjp returns for
 * the same nodes and geometry; workspace as afx_geometry_vjp. */
int afx_geometry_compute_record(const afx_mesh* mesh, const double* node_coords, afx_geometry* out, double* approx,
                                hipStream_t stream);
int afx_geometry_vjp_record(const afx_mesh* mesh, const double* node_coords, const afx_geometry* geom,
                            const double* approx, const afx_geometry_cotangents* geom_cotangents, double* node_b,
                            void* work, int64_t work_bytes, hipStream_t stream);

/* ------------------------------------------------------------------------------------------- */
/* coupled (u, v[, w], p) flow residual                                                          */

enum { AFX_FLOW_NONE = 0, AFX_FLOW_WALL = 1, AFX_FLOW_INLET = 2, AFX_FLOW_OUTLET = 3 };
enum { AFX_ADV_NONE = 0, AFX_ADV_UPWIND = 1, AFX_ADV_LIMITED = 2 };
/* AFX_GRAD_CORRECTED: CorrectedGreenGauss with SweptGradientSolve (reference:
 * aquaflux/schemes/gradient.py:618-705, :1122-1337). */
enum { AFX_GRAD_GREEN_GAUSS = 0, AFX_GRAD_LEAST_SQUARES = 1, AFX_GRAD_CORRECTED = 3 };

typedef struct {
  int32_t advection;          /* AFX_ADV_*: None (Stokes) / FirstOrderUpwind / LimitedUpwind(Venkatakrishnan) */
  int32_t gradient;           /* AFX_GRAD_*: CompactGreenGauss / least squares / CorrectedGreenGauss */
  int64_t pressure_pin;       /* cell whose continuity row becomes p - value; -1 = none */
  double pressure_pin_value;
  double body_force[3];       /* uniform force per unit volume */
  double limiter_k;           /* Venkatakrishnan K (eps^2 = V K^3) */
  const int8_t* face_kind;    /* [F] AFX_FLOW_* per face (interior faces: AFX_FLOW_NONE) */
  int32_t gradient_sweeps;    /* AFX_GRAD_CORRECTED: swept-solve sweep count (reference default 4) */
  int32_t reserved;
  double gradient_relaxation; /* AFX_GRAD_CORRECTED: Richardson relaxation (reference default 1) */
} afx_flow_config;

typedef struct {
  const double* state;   /* [(d+1)][C] field-major u, v, (w), p */
  const double* mu;      /* [C] dynamic (effective) viscosity */
  const double* rho;     /* [C] density */
  const double* u_b;     /* [d][F] prescribed velocity on wall / inlet faces */
  const double* p_b;     /* [F] prescribed pressure on outlet faces */
  const double* gamma_b; /* [F] optional boundary-face momentum viscosity override (NULL = owner mu) */
} afx_flow_inputs;

typedef struct {
  double* state;   /* [(d+1)][C] */
  double* mu;      /* [C] */
  double* rho;     /* [C] */
  double* u_b;     /* [d][F] */
  double* p_b;     /* [F] */
  double* gamma_b; /* [F] (required iff inputs.gamma_b is set) */
} afx_flow_cotangents;

/* Workspace (device bytes) for mode 0 = residual, 1 = JVP, 2 = VJP, 3 = VJP with geometry
 * cotangents. */
int afx_flow_workspace_bytes(const afx_mesh* mesh, int mode, int64_t* bytes);

/* R(state).  Replaces MomentumContinuity.residual (reference: aquaflux/flow/momentum.py:880-882,
 * with flow_fields :822-857 and residual_from_fields :859-878). */
int afx_flow_residual(const afx_mesh* mesh, const afx_geometry* geom, const afx_flow_config* cfg,
                      const afx_flow_inputs* in, double* residual, void* work, int64_t work_bytes,
                      hipStream_t stream);

/* R and dR = J . tangent (tangent pointers may be NULL = zero).  Replaces jax.jvp of the residual
 * (reference: aquaflux/solve/newton.py:108-121). */
int afx_flow_jvp(const afx_mesh* mesh, const afx_geometry* geom, const afx_flow_config* cfg,
                 const afx_flow_inputs* in, const afx_flow_inputs* tangent, double* residual,
                 double* residual_tangent, void* work, int64_t work_bytes, hipStream_t stream);

/* R and dR along tangents of the inputs and of the geometry together: geom_tangent holds the
 * tangents of face area, normal, face centroid, volume and cell centroid (e.g. from
 * afx_geometry_jvp of a node-coordinate tangent), and face_pack their face arrays in the internal
 * order when the mesh reorders faces (afx_geometry_pack of the tangent arrays).  Compact
 * Green-Gauss or least-squares gradients.  Replaces jax.jvp of MomentumContinuity.residual with
 * Mesh.node_coords as a differentiable leaf (reference: aquaflux/mesh/mesh.py:63-66,
 * aquaflux/flow/momentum.py:822-882). */
int afx_flow_jvp_geometry(const afx_mesh* mesh, const afx_geometry* geom, const afx_geometry* geom_tangent,
                          const afx_flow_config* cfg, const afx_flow_inputs* in, const afx_flow_inputs* tangent,
                          double* residual, double* residual_tangent, void* work, int64_t work_bytes,
                          hipStream_t stream);

/* Cotangents J^T . cotangent with respect to state, mu, rho, u_b, p_b (gamma_b).  Replaces jax.vjp
 * of the residual (reference: aquaflux/solve/implicit.py:733-761). */
int afx_flow_vjp(const afx_mesh* mesh, const afx_geometry* geom, const afx_flow_config* cfg,
                 const afx_flow_inputs* in, const double* cotangent, afx_flow_cotangents* out, void* work,
                 int64_t work_bytes, hipStream_t stream);

/* As afx_flow_vjp, and also the cotangents of the five geometry arrays (written; caller face
 * order), which afx_geometry_vjp carries on to the node coordinates.  Compact Green-Gauss or
 * least-squares gradients; workspace of mode 3.  Replaces jax.vjp of MomentumContinuity.residual
 * with Mesh.node_coords as a differentiable leaf (reference: aquaflux/mesh/mesh.py:63-66,
 * aquaflux/flow/momentum.py:822-882). */
int afx_flow_vjp_geometry(const afx_mesh* mesh, const afx_geometry* geom, const afx_flow_config* cfg,
                          const afx_flow_inputs* in, const double* cotangent, afx_flow_cotangents* out,
                          afx_geometry_cotangents* geom_out, void* work, int64_t work_bytes, hipStream_t stream);

/* Intermediate fields: gradients [(d+1)d][C], a_P [C], psi [d][C] (limited advection only).
 * Replaces MomentumContinuity.flow_fields' gradient / momentum_matrix_diagonal products
 * (reference: aquaflux/flow/momentum.py:822-857). */
int afx_flow_fields(const afx_mesh* mesh, const afx_geometry* geom, const afx_flow_config* cfg,
                    const afx_flow_inputs* in, double* gradients, double* a_p, double* psi, hipStream_t stream);

/* Rhie-Chow face mass flux mdot[F] (caller face order; boundary faces by their patch closure) and,
 * when `tangent` is not NULL, its tangent mdot_tangent[F].  Workspace as afx_flow_workspace_bytes
 * mode 0 (mode 1 with a tangent).  Replaces MomentumContinuity.mass_flux / FlowFields.mdot
 * (reference: aquaflux/flow/momentum.py:531-554, :884-899; aquaflux/flow/rhie_chow.py:274-353). */
int afx_flow_mass_flux(const afx_mesh* mesh, const afx_geometry* geom, const afx_flow_config* cfg,
                       const afx_flow_inputs* in, const afx_flow_inputs* tangent, double* mdot, double* mdot_tangent,
                       void* work, int64_t work_bytes, hipStream_t stream);

/* ------------------------------------------------------------------------------------------- */
/* scalar fields                                                                                 */

/* Cell gradient grad[j][C] of phi by `gradient` (AFX_GRAD_*; `sweeps` and `relaxation` are read for
 * AFX_GRAD_CORRECTED) with leading-order boundary values: kind[f] == AFX_BC_DIRICHLET -> value[f],
 * any other closure -> the owner value.  Replaces ResidualAssembler.gradient for Dirichlet /
 * zero-gradient closures (reference: aquaflux/discretization/residual.py:253-296), e.g. the SST
 * wall-distance gradient (aquaflux/turbulence/transport.py:86-110). */
int afx_scalar_gradient(const afx_mesh* mesh, const afx_geometry* geom, int gradient, int sweeps, double relaxation,
                        const int8_t* kind, const double* value, const double* phi, double* grad,
                        hipStream_t stream);

/* Gradient and its tangent along a tangent of phi and of the geometry together (geom_tangent as
 * for afx_scalar_jvp_geometry; boundary values constant).  Compact Green-Gauss, least squares or
 * corrected Green-Gauss (`sweeps`, `relaxation` as for afx_scalar_gradient).  Replaces jax.jvp of
 * GradientScheme.gradients with Mesh.node_coords as a differentiable leaf (the SST wall-distance
 * gradient of a differentiated build; reference: aquaflux/schemes/gradient.py:98-153, :1160-1337,
 * aquaflux/turbulence/transport.py:259-313). */
int afx_scalar_gradient_jvp_geometry(const afx_mesh* mesh, const afx_geometry* geom, const afx_geometry* geom_tangent,
                                     int gradient, int sweeps, double relaxation, const int8_t* kind,
                                     const double* value, const double* phi, const double* phi_tangent, double* grad,
                                     double* grad_tangent, hipStream_t stream);

/* Cotangents of the five geometry arrays (written; caller face order) of the gradient of phi for a
 * gradient cotangent grad_bar[d][C] (boundary values and phi held fixed; their cotangents come
 * from afx_scalar_gradient_vjp).  Compact Green-Gauss, least squares or corrected Green-Gauss
 * (`sweeps`, `relaxation` as for afx_scalar_gradient).  Replaces jax.vjp of
 * GradientScheme.gradients with Mesh.node_coords as a differentiable leaf (the SST wall-distance
 * gradient of a differentiated build; reference: aquaflux/schemes/gradient.py:98-153, :1160-1337,
 * aquaflux/turbulence/transport.py:259-313). */
int afx_scalar_gradient_vjp_geometry(const afx_mesh* mesh, const afx_geometry* geom, int gradient, int sweeps,
                                     double relaxation, const int8_t* kind, const double* value, const double* phi,
                                     const double* grad_bar, afx_geometry_cotangents* out, hipStream_t stream);

/* Reverse of afx_scalar_gradient (same scheme, closures and geometry): phi_bar[C] and, when not
 * NULL, value_bar[F] (zero except on Dirichlet faces, caller order) from grad_bar[j][C].  The
 * gradient is linear in (phi, value), so this is its transpose.  Replaces jax.vjp through
 * GradientScheme.gradients (reference: aquaflux/schemes/gradient.py:98-153, :1122-1337). */
int afx_scalar_gradient_vjp(const afx_mesh* mesh, const afx_geometry* geom, int gradient, int sweeps,
                            double relaxation, const int8_t* kind, const double* grad_bar, double* phi_bar,
                            double* value_bar, hipStream_t stream);

/* Venkatakrishnan limiter psi[C] of phi[C] with its gradient grad[j][C] (eps^2 = V K^3, stencil
 * extrema over the face neighbours, minimum over the cell's faces, clipped to [0, 1]); with phi_t /
 * grad_t (either may be NULL = zero) also its tangent psi_t.  Replaces
 * VenkatakrishnanLimiter.limit (reference: aquaflux/schemes/limiter.py:78-117). */
int afx_venkat_limiter(const afx_mesh* mesh, const afx_geometry* geom, double limiter_k, const double* phi,
                       const double* grad, const double* phi_t, const double* grad_t, double* psi, double* psi_t,
                       hipStream_t stream);

enum { AFX_GRAD_NONE = 2 };

typedef struct {
  int32_t advection;     /* AFX_ADV_* */
  int32_t gradient;      /* AFX_GRAD_GREEN_GAUSS / _LEAST_SQUARES / _NONE (orthogonal grids) / _CORRECTED */
  double limiter_k;      /* Venkatakrishnan K */
  const int8_t* bc_kind; /* [F] AFX_BC_* on boundary faces */
  int32_t gradient_sweeps;    /* AFX_GRAD_CORRECTED: swept-solve sweep count */
  int32_t transient;          /* 0 steady; 1 backward Euler (BDF1); 2 BDF2 (TransientTerm) */
  double gradient_relaxation; /* AFX_GRAD_CORRECTED: Richardson relaxation */
  double dt;                  /* timestep of the transient term */
} afx_scalar_config;

typedef struct {
  const double* phi;   /* [C] transported scalar */
  const double* gamma; /* [C] effective diffusivity */
  const double* flux;  /* [F] owner-outward volumetric face flux Q_f */
  const double* value; /* [F] Dirichlet value / Neumann flux q / convective coefficient h */
  const double* t_inf; /* [F] convective ambient value */
  const double* phi_old;   /* [C] previous time level (transient only) */
  const double* phi_older; /* [C] second-previous time level (BDF2 only) */
} afx_scalar_inputs;

typedef struct {
  double* phi;   /* [C] */
  double* gamma; /* [C] */
  double* flux;  /* [F] */
  double* value; /* [F] */
  double* t_inf; /* [F] */
  afx_geometry_cotangents geometry; /* all five set, or all NULL (compact Green-Gauss, least squares or no
                                      * gradient; refused for corrected Green-Gauss) */
  double* phi_old;   /* [C] (transient only) */
  double* phi_older; /* [C] (BDF2 only) */
} afx_scalar_cotangents;

/* Workspace (device bytes) for mode 0 = residual, 1 = JVP, 2 = VJP. */
int afx_scalar_workspace_bytes(const afx_mesh* mesh, int mode, int64_t* bytes);

/* R(phi) = sum of (Q phi_f + diffusion) face fluxes.  Replaces ScalarTransport.residual(flux)(phi)
 * (reference: aquaflux/transport/scalar.py:185-223, ResidualAssembler.residual in
 * aquaflux/discretization/residual.py:252-412). */
int afx_scalar_residual(const afx_mesh* mesh, const afx_geometry* geom, const afx_scalar_config* cfg,
                        const afx_scalar_inputs* in, double* residual, void* work, int64_t work_bytes,
                        hipStream_t stream);

/* afx_scalar_residual without a workspace argument, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  The workspace is a file-scope device array of afx_scalar_noalloc_max_bytes() bytes
 * (64 MiB: up to 762,600 cells in 3D, 1,048,576 in 2D), held on each GPU the library runs on until
 * the process ends; a larger mesh is refused (status 2).  A mesh whose faces the library reorders
 * needs geom->face_pack.  Calls on one GPU share the array: calls that overlap on one GPU race. */
int afx_scalar_residual_noalloc(const afx_mesh* mesh, const afx_geometry* geom, const afx_scalar_config* cfg,
                                const afx_scalar_inputs* in, double* residual, hipStream_t stream);
/* Size in bytes of the workspace array of afx_scalar_residual_noalloc. */
int afx_scalar_noalloc_max_bytes(int64_t* bytes);

/* R and J . tangent (tangent pointers may be NULL = zero).  Replaces jax.jvp of the scalar residual. */
int afx_scalar_jvp(const afx_mesh* mesh, const afx_geometry* geom, const afx_scalar_config* cfg,
                   const afx_scalar_inputs* in, const afx_scalar_inputs* tangent, double* residual,
                   double* residual_tangent, void* work, int64_t work_bytes, hipStream_t stream);

/* afx_scalar_jvp with a tangent of the geometry as well: geom_tangent holds the tangents of area,
 * normal, face centroid, volume and cell centroid (e.g. from afx_geometry_jvp of a node-coordinate
 * tangent), and face_pack their face arrays in the internal order when the mesh reorders faces
 * (afx_geometry_pack of the tangent arrays).  Compact Green-Gauss, least-squares or no gradient
 * reconstruction.  With `tangent` also carrying the face-flux tangent, residual_tangent is the
 * derivative of the residual along the node-coordinate direction.  Replaces jax.jvp of
 * ScalarTransport.residual with Mesh.node_coords as a differentiable leaf (reference:
 * aquaflux/mesh/mesh.py:63-66, aquaflux/transport/scalar.py:185-223). */
int afx_scalar_jvp_geometry(const afx_mesh* mesh, const afx_geometry* geom, const afx_geometry* geom_tangent,
                            const afx_scalar_config* cfg, const afx_scalar_inputs* in,
                            const afx_scalar_inputs* tangent, double* residual, double* residual_tangent, void* work,
                            int64_t work_bytes, hipStream_t stream);

/* Cotangents of phi, gamma, the face flux, the boundary values and (optionally) the geometry.
 * Replaces jax.vjp of the scalar residual, including through MeshGeometry. */
int afx_scalar_vjp(const afx_mesh* mesh, const afx_geometry* geom, const afx_scalar_config* cfg,
                   const afx_scalar_inputs* in, const double* cotangent, afx_scalar_cotangents* out, void* work,
                   int64_t work_bytes, hipStream_t stream);

/* The residual and its VJP split at a forward record, as jax.vjp(f, ...) splits into the primal and
 * a pullback that holds the linearisation point (reference call sites: jax.vjp of
 * ScalarTransport.residual(flux)(phi), aquaflux/transport/scalar.py:185-223).  The record keeps the
 * gradient pass's outputs (gradient, limiter value, the limiter's extremum cells and minimum face:
 * afx_scalar_record_bytes bytes), so afx_scalar_vjp_record does not run that pass again.  The
 * record is valid for the inputs and geometry of the call that wrote it; afx_scalar_vjp_record
 * returns what afx_scalar_vjp returns at those inputs.  Not for the corrected Green-Gauss scheme
 * (status 2: its reverse pass needs the sweep iterates; use afx_scalar_vjp). */
int afx_scalar_record_bytes(const afx_mesh* mesh, int64_t* bytes);
int afx_scalar_residual_record(const afx_mesh* mesh, const afx_geometry* geom, const afx_scalar_config* cfg,
                               const afx_scalar_inputs* in, double* residual, void* record, int64_t record_bytes,
                               hipStream_t stream);
int afx_scalar_vjp_record(const afx_mesh* mesh, const afx_geometry* geom, const afx_scalar_config* cfg,
                          const afx_scalar_inputs* in, const double* cotangent, const void* record,
                          int64_t record_bytes, afx_scalar_cotangents* out, void* work, int64_t work_bytes,
                          hipStream_t stream);

/* ------------------------------------------------------------------------------------------- */
/* coupled RANS (u, v[, w], p, k, omega) residual with the k-omega SST closure                   */

/* Parametrisation of a solved turbulence scalar (reference: DirectScalars / LogScalars,
 * aquaflux/turbulence/coupled.py:155-211): state rows hold phi, or w with phi = e^w. */
enum { AFX_SCALAR_DIRECT = 0, AFX_SCALAR_LOG = 1 };

enum { AFX_BC_NONE = 0, AFX_BC_DIRICHLET = 1, AFX_BC_ZERO_GRADIENT = 2, AFX_BC_NEUMANN = 3, AFX_BC_CONVECTIVE = 4 };

typedef struct {
  afx_flow_config flow;               /* momentum / continuity configuration */
  int32_t scalar_advection;           /* AFX_ADV_* of the k and omega equations */
  int32_t reserved;
  double density;                     /* the SST assembler's density (k/omega volume flux mdot / density) */
  double sst[16];                     /* a1, beta*, alpha1, beta1, sigma_k1, sigma_w1, alpha2, beta2,
                                         sigma_k2, sigma_w2, kappa, E, wall-omega exponent,
                                         wall-omega viscous coeff, beta*^0.25, y*_lam */
  const int8_t* k_kind;               /* [F] AFX_BC_DIRICHLET / AFX_BC_ZERO_GRADIENT on boundary faces */
  const int8_t* omega_kind;           /* [F] */
  const int8_t* wall_face;            /* [F] 1 on the SST wall-patch faces */
  const int8_t* wall_cell;            /* [C] 1 on their owner cells (the omega fixation set) */
  const double* wall_distance;        /* [C] */
  const double* wall_distance_gradient; /* [d][C] */
  int32_t k_transform;                /* AFX_SCALAR_DIRECT / AFX_SCALAR_LOG: the solved k unknown */
  int32_t omega_transform;            /* ... and omega (log: the wall fixation row is log(omega / target)) */
} afx_rans_config;

typedef struct {
  const double* state;       /* [(d+3)][C] field-major u, v, (w), p, k, omega */
  const double* mu;          /* [C] molecular dynamic viscosity */
  const double* rho;         /* [C] density */
  const double* nu;          /* [C] molecular kinematic viscosity of the SST assembler */
  const double* u_b;         /* [d][F] */
  const double* p_b;         /* [F] */
  const double* k_value;     /* [F] Dirichlet values of k */
  const double* omega_value; /* [F] Dirichlet values of omega */
} afx_rans_inputs;

typedef struct {
  double* state; /* [(d+3)][C] */
  double* mu;    /* [C] */
  double* rho;   /* [C] */
  double* nu;    /* [C] */
  double* u_b;   /* [d][F] */
  double* p_b;   /* [F] */
} afx_rans_cotangents;

/* Workspace (device bytes) for mode 0 = residual, 1 = JVP, 2 = VJP, 3 = VJP with geometry
 * cotangents. */
int afx_rans_workspace_bytes(const afx_mesh* mesh, int mode, int64_t* bytes);

/* R(u, p, k, omega) at the solved state (k / omega rows e^w under AFX_SCALAR_LOG).  Replaces
 * CoupledRANS.residual with DirectScalars or LogScalars (reference:
 * aquaflux/turbulence/coupled.py:437-473, closure_fields / k_residual / omega_residual in
 * aquaflux/turbulence/transport.py:548-718). */
int afx_rans_residual(const afx_mesh* mesh, const afx_geometry* geom, const afx_rans_config* cfg,
                      const afx_rans_inputs* in, double* residual, void* work, int64_t work_bytes,
                      hipStream_t stream);

/* R and J . tangent (tangent pointers may be NULL = zero).  Replaces jax.jvp of CoupledRANS.residual
 * (reference: aquaflux/solve/newton.py:108-121). */
int afx_rans_jvp(const afx_mesh* mesh, const afx_geometry* geom, const afx_rans_config* cfg,
                 const afx_rans_inputs* in, const afx_rans_inputs* tangent, double* residual,
                 double* residual_tangent, void* work, int64_t work_bytes, hipStream_t stream);

/* R and dR along tangents of the inputs and of the geometry together: geom_tangent as for
 * afx_flow_jvp_geometry, plus the tangents of the configuration's wall distance [C] and wall
 * distance gradient [d][C] (a differentiated build moves them with the nodes: reference
 * aquaflux/mesh/distance.py:87-140).  Compact Green-Gauss or least-squares gradients; workspace of
 * mode 1.  Replaces jax.jvp of CoupledRANS.residual with Mesh.node_coords as a differentiable leaf
 * (reference: aquaflux/mesh/mesh.py:63-66, aquaflux/turbulence/coupled.py:437-473). */
int afx_rans_jvp_geometry(const afx_mesh* mesh, const afx_geometry* geom, const afx_geometry* geom_tangent,
                          const double* wall_distance_tangent, const double* wall_distance_gradient_tangent,
                          const afx_rans_config* cfg, const afx_rans_inputs* in, const afx_rans_inputs* tangent,
                          double* residual, double* residual_tangent, void* work, int64_t work_bytes,
                          hipStream_t stream);

/* J^T . cotangent with respect to state, mu, rho, nu, u_b, p_b.  Replaces jax.vjp of
 * CoupledRANS.residual (reference: aquaflux/solve/implicit.py:733-761). */
int afx_rans_vjp(const afx_mesh* mesh, const afx_geometry* geom, const afx_rans_config* cfg,
                 const afx_rans_inputs* in, const double* cotangent, afx_rans_cotangents* out, void* work,
                 int64_t work_bytes, hipStream_t stream);

/* As afx_rans_vjp, and also the cotangents of the five geometry arrays (written; caller face order)
 * and of the configuration's wall distance [C] and wall-distance gradient [d][C], which a
 * differentiated build carries on to the geometry (nearest wall-face centroid distance,
 * afx_scalar_gradient_vjp / _vjp_geometry) and then afx_geometry_vjp to the nodes.  Compact
 * Green-Gauss or least-squares gradients; workspace of mode 3.  Replaces jax.vjp of
 * CoupledRANS.residual with Mesh.node_coords as a differentiable leaf (reference:
 * aquaflux/mesh/mesh.py:63-66, aquaflux/mesh/distance.py:87-140,
 * aquaflux/turbulence/coupled.py:437-473). */
int afx_rans_vjp_geometry(const afx_mesh* mesh, const afx_geometry* geom, const afx_rans_config* cfg,
                          const afx_rans_inputs* in, const double* cotangent, afx_rans_cotangents* out,
                          const afx_geometry_cotangents* geom_out, double* wall_distance_bar,
                          double* wall_distance_gradient_bar, void* work, int64_t work_bytes, hipStream_t stream);

/* R and J^T . cotangent in one call, the forward passes shared (the pair jax.vjp returns: the
 * primal output and its pullback applied to `cotangent`).  Workspace as for mode 2. */
int afx_rans_residual_vjp(const afx_mesh* mesh, const afx_geometry* geom, const afx_rans_config* cfg,
                          const afx_rans_inputs* in, const double* cotangent, double* residual,
                          afx_rans_cotangents* out, void* work, int64_t work_bytes, hipStream_t stream);

/* The OpenMP CPU line of afx_rans_residual_vjp: the same passes and per-cell code as the HIP
 * kernels, run one cell at a time by `threads` host threads (<= 0: OpenMP's default) on HOST arrays
 * (owner / neighbour, the five geometry arrays and every input and configuration array in the
 * caller's face order); compact Green-Gauss or least-squares gradients.  No counterpart in the
 * reference: it is the best-CPU baseline the GPU figures are reported against (the reference's own
 * CPU path is JAX on XLA:CPU), and nothing in the device path calls it. */
int afx_rans_residual_vjp_host(int dim, int64_t n_cells, int64_t n_faces, const int32_t* owner,
                               const int32_t* neighbour, const afx_geometry* geom, const afx_rans_config* cfg,
                               const afx_rans_inputs* in, const double* cotangent, double* residual,
                               afx_rans_cotangents* out, int threads);

/* Forward intermediates into `fields` (the mode-0 workspace layout): gradients [(d+3)d][C],
 * limiters [d+2][C], a_P [C], closure [7][C] = nu_t, F1, strain rate, cross-diffusion,
 * wall-function nu_t, wall shear rate, mu_eff.  Replaces SSTTurbulence.closure_fields (reference:
 * aquaflux/turbulence/transport.py:548-579). */
int afx_rans_fields(const afx_mesh* mesh, const afx_geometry* geom, const afx_rans_config* cfg,
                    const afx_rans_inputs* in, double* fields, hipStream_t stream);

/* ------------------------------------------------------------------------------------------- */
/* compressed (coloured) Jacobian assembly                                                       */

/* Cell-block sparsity pattern at stencil `reach` on the interior-face cell graph and a probing
 * colouring of it, from HOST arrays (owner / neighbour of the n_edges interior faces).  Replaces
 * block_stencil_colouring and _saturation_colouring (reference:
 * aquaflux/solve/sparse_jacobian.py:338-534).  HOST outputs: pattern_ptr[n_cells + 1] (always) and,
 * when pattern_cols is not NULL, pattern_cols[pattern_ptr[n_cells]] (each row: the cell itself,
 * then the cells within `reach` steps in ascending order) and, when pattern_depth is not NULL, each
 * entry's graph distance from its row cell; when colour is not NULL, colour[n_cells]
 * and *n_colours (two cells share a colour only if their graph distance exceeds 2 reach, so no row
 * couples to both; chosen by saturation degree, ties by conflict degree). */
int afx_block_colouring(int64_t n_cells, int64_t n_edges, const int32_t* owner, const int32_t* neighbour,
                        int32_t reach, int64_t* pattern_ptr, int32_t* pattern_cols, int8_t* pattern_depth,
                        int32_t* colour, int32_t* n_colours);

/* Probe seed of (colour k, field): seed[v * n_cells + c] = 1 where v == field and colour[c] == k,
 * else 0 (field-major, `block` fields).  Replaces ColumnProbePlan.seed_block (reference:
 * aquaflux/solve/sparse_jacobian.py:208-247). */
int afx_probe_seed(int64_t n_cells, int32_t block, const int32_t* colour, int32_t k, int32_t field, double* seed,
                   hipStream_t stream);

/* Writes column `field` of every block whose column cell has colour k from the probe response
 * J . seed (field-major) into block-ELL storage vals[((s*B + a)*B + field)*C + c] (float when
 * precision == 0, double when 1) over cols[s*C + c] (slot 0 the row cell, a later slot equal to
 * the row cell is padding).  With depth (nullable, int8 [nslot*C], the entries' graph distances)
 * entries farther than max_depth are left untouched: a column probed with a shorter-reach
 * colouring than the pattern (ColumnProbePlan, sparse_jacobian.py:88-247) keeps its outer entries
 * at zero.  Replaces ProbeGather.scatter (reference: aquaflux/solve/sparse_jacobian.py:632-720). */
int afx_probe_scatter(int64_t n_cells, int32_t block, int32_t nslot, const int32_t* cols, const int8_t* depth,
                      int32_t max_depth, const int32_t* colour, int32_t k, int32_t field, const double* response,
                      int precision, void* vals, hipStream_t stream);

/* ------------------------------------------------------------------------------------------- */
/* Newton-Krylov inner operators: compact block Jacobian, SpMV, point-block Jacobi              */

/* Block-ELL slot count per row (1 + max faces per cell). */
int afx_bell_slots(const afx_mesh* mesh, int32_t* nslot);

/* Reach-1 block Jacobian of the flow residual at `in` (gradients, a_P and limiter frozen), stored
 * block-ELL cell-major: vals[(slot*B*B + r*B + k)*C + c] (float if precision == 0, double if 1),
 * cols[slot*C + c]; B = d + 1.  `tslot` (nullable, [nslot*C]) receives, per slot, the slot of the
 * same face in the column cell's row (for afx_bell_spmv_t).  `fields` is scratch of
 * ((d+1)d + 1 + d) * C doubles.
 * Replaces materialize_block_jacobian's coloured-JVP probing (reference:
 * aquaflux/solve/sparse_jacobian.py:722-927) with direct dual-number assembly. */
int afx_flow_jacobian(const afx_mesh* mesh, const afx_geometry* geom, const afx_flow_config* cfg,
                      const afx_flow_inputs* in, int precision, void* vals, int32_t* cols, int32_t* tslot,
                      double* fields, hipStream_t stream);

/* y = A x for a block-ELL matrix with cell-major vectors x[c*B + v] (B = 1 to 6); a slot > 0 whose
 * column is the row itself is padding. */
int afx_bell_spmv(int64_t n_rows, int32_t block, int32_t nslot, int precision, const void* vals,
                  const int32_t* cols, const void* x, void* y, hipStream_t stream);

/* y = A^T x using the transpose slot map of afx_flow_jacobian. */
int afx_bell_spmv_t(int64_t n_rows, int32_t block, int32_t nslot, int precision, const void* vals,
                    const int32_t* cols, const int32_t* tslot, const void* x, void* y, hipStream_t stream);

/* Inverses of the diagonal blocks (slot 0) into inv[(r*B + k)*C + c]; sets *singular = 1 (device
 * int) when a block is singular (that block is replaced by the identity). */
int afx_pbj_setup(int64_t n_rows, int32_t block, int precision, const void* vals, void* inv, int32_t* singular,
                  hipStream_t stream);
/* y = D^-1 x. */
int afx_pbj_apply(int64_t n_rows, int32_t block, int precision, const void* inv, const void* x, void* y,
                  hipStream_t stream);

/* ---- aggregation multigrid on a scalar ELL operator (the block preconditioner's pressure Schur) --
 * Replaces the host smoothed-aggregation V-cycle of the reference's Schur solve
 * (aquaflux/flow/block_preconditioner.py:266-343, aquaflux/solve/multigrid.py:1596-1700).
 * ELL as afx_bell_spmv with block 1 (slot 0 the diagonal). */
/* y = x + omega (b - A x) / diag(A); x may be NULL (a sweep from zero); x and y must differ. */
int afx_ell_jacobi(int64_t n, int32_t nslot, int precision, const void* vals, const int32_t* cols, const void* b,
                   const void* x, void* y, double omega, hipStream_t stream);
/* r = b - A x (x and r must differ). */
int afx_ell_residual(int64_t n, int32_t nslot, int precision, const void* vals, const int32_t* cols, const void* b,
                     const void* x, void* r, hipStream_t stream);
/* rc[I] = sum of r over the members of aggregate I: members[k * n_coarse + I], padded with -1. */
int afx_aggregate_restrict(int64_t n_coarse, int32_t max_members, int precision, const int32_t* members,
                           const void* r, void* rc, hipStream_t stream);
/* x[i] += alpha xc[agg[i]]. */
int afx_aggregate_prolong(int64_t n, int precision, const int32_t* agg, const void* xc, double alpha, void* x,
                          hipStream_t stream);
/* out[0] = a . b (FP64 accumulation): 256 per-block partial sums into partial[256], then their sum;
 * no atomics and no cleared memory, so it may be recorded in a HIP graph. */
int afx_dot(int64_t n, int precision, const void* a, const void* b, double* partial, void* out, hipStream_t stream);
/* y = A x for a dense row-major n x n matrix (the coarsest level's pseudo-inverse); x and y differ. */
int afx_dense_matvec(int64_t n, int precision, const void* A, const void* x, void* y, hipStream_t stream);

/* Face mass flux mdot[f] = (u_f . n_f) A_f from the owner-outward unit normals normal[d][F] and areas
 * area[F]; u is one velocity (per_face = 0, u[d]) or one per face (per_face = 1, u[d][F]).  Replaces
 * the verification flux builder of the reference's tests/support/fields.py:16-43 (the K4 flux). */
int afx_face_flux(int64_t n_faces, int dim, int per_face, const double* u, const double* normal, const double* area,
                  double* mdot, hipStream_t stream);
/* Its transpose for a cotangent mdot_bar[F]: normal_bar[d][F] = u A mdot_bar, area_bar[F] = (u . n)
 * mdot_bar (either may be NULL, each written, not accumulated), and with a per-face u, u_bar[d][F] =
 * n A mdot_bar (NULL to skip).  Replaces jax.vjp through that flux. */
int afx_face_flux_vjp(int64_t n_faces, int dim, int per_face, const double* u, const double* normal,
                      const double* area, const double* mdot_bar, double* normal_bar, double* area_bar, double* u_bar,
                      hipStream_t stream);

/* ---- Krylov basis ------------------------------------------------------------------------------ */
/* x[i] += sum_{j<k} y[j] basis[j * ld + i] with basis FP32 (precision 0) or FP64 (1), y and x FP64,
 * the k terms summed in FP64 in index order.  Replaces the update x + Z y of the restarted GMRES in
 * solve_linear (reference: aquaflux/solve/linear.py:199-298), where the basis is FP64. */
int afx_basis_combine(int64_t n, int32_t k, int precision, const void* basis, int64_t ld, const double* y, double* x,
                      hipStream_t stream);

/* Restarted GMRES with the Arnoldi recurrence on the device (replaces the Arnoldi / Givens / update
 * loop of lineax's GMRES in solve_linear, reference: aquaflux/solve/linear.py:199-298).  The host
 * enqueues, per cycle: afx_gmres_start, then for j = 0, 1, ... the caller's product
 * w = A M^-1 basis[j] followed by afx_gmres_arnoldi(j), then afx_gmres_update; it reads the status
 * doubles of `state` whenever it chooses to synchronise.
 *   state (FP64, afx_gmres_sizes' state_doubles): [0] k = steps accepted in this cycle, [1] done
 *   (the residual estimate reached tol, or breakdown), [2] breakdown, [4] tol, [5] beta = ||r||,
 *   [6] the current residual estimate |g_k|, [8 + j] the estimate after step j.  Once done is set,
 *   further afx_gmres_arnoldi calls of the cycle do nothing.
 *   work (FP64, work_doubles): partial sums.  basis: m + 1 vectors of `ld` entries, FP32
 *   (precision 0) or FP64 (1); every dot product, norm and update is accumulated in FP64 in a fixed
 *   order, so results are bitwise reproducible. */
int afx_gmres_sizes(int64_t n, int32_t m, int64_t* state_doubles, int64_t* work_doubles);
/* beta = ||r||, basis[0] = r / beta, g = beta e_1; when `first`, tol = atol + rtol beta (pass rtol = 0
 * and atol = the tolerance when the initial guess is not zero).  done is set when beta <= tol. */
int afx_gmres_start(int64_t n, int32_t m, int precision, int r_precision, const void* r, double rtol, double atol,
                    int first, void* basis, int64_t ld, double* state, double* work, hipStream_t stream);
/* Step j: orthogonalises w (FP32 or FP64) against basis[0..j] by classical Gram-Schmidt, a second
 * pass when the first removed more than 1 - 1/sqrt(2) of its norm (reorthogonalize = 0) or always
 * (1), stores basis[j + 1] = w / h_{j+1,j}, applies the Givens rotations to the new Hessenberg
 * column and updates the residual estimate. */
int afx_gmres_arnoldi(int64_t n, int32_t m, int32_t j, int precision, int w_precision, const void* w, void* basis,
                      int64_t ld, int reorthogonalize, double* state, double* work, hipStream_t stream);
/* y = R^-1 g over the k accepted steps (k read from state), then x += basis[:k]^T y (x FP32 or FP64,
 * accumulated in FP64). */
int afx_gmres_update(int64_t n, int32_t m, int precision, const void* basis, int64_t ld, double* state,
                     int x_precision, void* x, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* AQUAFLUX_AMD_H */
