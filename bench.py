"""Benchmark: residual + VJP throughput (Gfaces/s) of the face-loop residual on one MI355X.

One step = one residual evaluation R(state) plus one adjoint J^T lambda (cotangents of the state,
viscosity, density and boundary values), all inputs resident in HBM.  The default workload is the
coupled RANS (u, v, w, p, k, omega with the k-omega SST closure) on the 3D hex channel of
128 x 64 x 64 cells (524,288 cells, 1,593,344 faces) with a seeded turbulent-like state;
``--workload flow`` runs the laminar (u, v, w, p) residual on the same mesh.

Prints ONE JSON line.  ``roofline`` is the dominant kernel's algorithmic bytes per launch over its
average launch time, timed with events recorded by the library on the launch stream during the
timed region.  ``cpu_baseline`` is the CPU oracle (torch float64, the reference's pass structure)
timed on this host on a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
S8 = 8  # bytes per FP64 value

# Whole-step algorithmic bytes per face of residual + VJP (pass-structured compulsory traffic of the
# recompute design: 3 forward face passes, 3 reverse passes) for the two 3D hex workloads.
STEP_BYTES_PER_FACE = {"sst": 1386.0, "tet": 890.0}


def _traffic(name, kernel):
    """HBM bytes per launch of ``kernel`` from profiles/<name>_traffic.json (rocprofv3 FETCH_SIZE /
    WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        tk = json.load(fh).get("kernels", {}).get(kernel)
    return tk["hbm_bytes_per_launch"] if tk else None


def _face_pass(D, permuted=True):
    """One face pass: owner + neighbour (2 x int32), area, unit normal and centroid (FP64), the
    cell -> face list entries that reach the face from both sides (2 x int32) and, when the mesh's
    faces are visited in the internal owner-sorted order, the caller index of the face (int32)."""
    return 2 * 4 + S8 * (1 + 2 * D) + 2 * 4 + (4 if permuted else 0)


def _flow_kernel_bytes(D=3, limited=True):
    """Compulsory bytes of each laminar-flow kernel, as per-cell (c), per-face (f) and
    per-boundary-face (b) coefficients: every array a kernel reads or writes counted once."""
    NV = D + 1
    G = NV * D * S8
    psi = D * S8 if limited else 0
    state = NV * S8
    cgeo = (1 + D) * S8
    fp = _face_pass(D)
    bnd = 1 + D * S8 + S8
    return {
        "flow_grad": dict(c=4 + state + cgeo + G + psi, f=fp, b=bnd),
        "flow_ap": dict(c=4 + D * S8 + D * D * S8 + 2 * S8 + cgeo + S8, f=fp, b=1),
        "flow_res": dict(c=4 + state + G + 3 * S8 + psi + cgeo + state, f=fp, b=bnd),
        "flow_res_adj": dict(c=4 + 2 * state + G + 3 * S8 + psi + cgeo + state + G + 3 * S8 + psi,
                             f=fp + (D + 1) * S8, b=bnd),
        "flow_ap_adj": dict(c=4 + D * S8 + D * D * S8 + 2 * S8 + cgeo + S8 + 2 * (D * S8 + 2 * S8)
                            + 2 * D * D * S8 + psi + 2 * D * S8 + 2 * D * 4, f=fp, b=1),
        "flow_grad_adj": dict(c=4 + cgeo + G + 2 * state + 2 * D * S8 + 2 * D * 4, f=fp, b=bnd + (D + 1) * S8),
    }


def _rans_kernel_bytes(D=3, limited=True):
    """Compulsory bytes of each coupled-RANS kernel (same convention as the flow model; ``w`` is per
    wall-adjacent cell, ``wf`` per wall face).  Limited advection of u_i, k, omega (NL = D + 2)."""
    NA, NL = D + 3, D + 2
    state, G = NA * S8, NA * D * S8
    psi = NL * S8 if limited else 0
    ext = 2 * NL * 4 if limited else 0
    cgeo = (1 + D) * S8
    fp = _face_pass(D)
    bnd = 3 + (D + 1) * S8 + 2 * S8 + 1  # kinds, u_b, p_b, k/omega values, wall flag
    closure_in = 2 * S8 + (D * D + 2 * D) * S8 + 4 * S8 + 1  # k, w, grad u/k/w, nu mu rho d, wall flag
    cell_sst = 5 * S8  # nu_t, F1, strain, cross-diffusion, wall shear rate
    res_in = state + G + psi + S8 + 3 * S8 + cell_sst + S8 + 1 + cgeo  # ... + a_P, mu_eff rho nu, d, flag
    bars10 = 10 * S8
    NV = D + 1
    flow_cell = NV * S8 + (D * D + D) * S8 + 2 * S8  # u, p, grad u, grad p, mu_eff, rho
    return {
        # launched once per step by the fused residual + VJP call, which also has it write the
        # extremum owners and the limiter's minimum faces.  Unless corrected-GG sweeps follow, it also
        # runs the cell-local closure from the gradient in registers: reads nu, mu, rho, d and the
        # wall flag (wall cells: the wall-distance gradient and their wall faces), writes the
        # seven closure fields
        "rans_grad": dict(c=4 + state + cgeo + G + psi + ext + 4 * S8 + 1 + 7 * S8, f=fp, b=bnd,
                          w=D * S8 + D * S8, wf=1 + D * S8 + S8 + 4),
        "rans_closure": dict(c=4 + closure_in + 7 * S8, f=0, b=0, w=D * S8 + D * S8, wf=1 + D * S8 + S8 + 4),
        "rans_ap": dict(c=4 + flow_cell + cgeo + S8, f=fp, b=2, w=2 * S8),
        # flow rows of R (momentum, continuity) and the face mass flux (written once per face)
        "rans_flow_res": dict(c=4 + flow_cell + S8 + (D * S8 if limited else 0) + cgeo + NV * S8, f=fp + S8,
                              b=2 + NV * S8, w=2 * S8),
        # k / omega rows: k, omega, their gradients and limiters, the closure (nu_t, F1, strain,
        # cross-diffusion, wall shear rate), nu, d, wall flag, the cotangent rows; writes the k /
        # omega state, gradient and limiter cotangents, six closure cotangents, the two R rows
        # (fused residual + VJP) and the mass-flux cotangent; reads the mass flux
        "rans_scalar_res_adj": dict(c=4 + 2 * S8 + 2 * D * S8 + (2 * S8 if limited else 0) + 7 * S8 + 1 + 2 * S8
                                    + cgeo + 2 * S8 + 2 * D * S8 + (2 * S8 if limited else 0) + 6 * S8 + 2 * S8,
                                    f=fp + 2 * S8, b=3 + 2 * S8),
        # flow rows: the flow cell (+ a_P, limiters), the cotangent rows, the mass-flux cotangent;
        # writes the flow state / gradient / a_P / limiter / mu_eff / rho / mu / wall-nu_t
        # cotangents and the (D + 1) boundary-value cotangent arrays over all faces
        "rans_flow_res_adj": dict(c=4 + flow_cell + S8 + (D * S8 if limited else 0) + NV * S8 + cgeo + NV * S8
                                  + (D * D + D) * S8 + S8 + (D * S8 if limited else 0) + 4 * S8,
                                  f=fp + S8 + NV * S8, b=2 + NV * S8, w=2 * S8),
        "rans_ap_adj": dict(c=4 + D * S8 + D * D * S8 + 4 * S8 + cgeo + S8 + 2 * D * S8 + 2 * D * D * S8
                            + 2 * 4 * S8, f=fp, b=2),
        "rans_closure_adj": dict(c=4 + closure_in + 2 * S8 + bars10 + 2 * G + psi + ext + 2 * state + 3 * S8
                                 + 2 * NL * S8 + 2 * NL * S8 + cgeo, f=2 * 4 + D * S8, b=0,
                                 w=D * S8 + D * S8, wf=1 + 2 * D * S8 + S8),
        "rans_grad_adj": dict(c=4 + G + ext + 2 * NL * S8 + 2 * state + cgeo, f=fp, b=3 + 2 * (D + 1) * S8),
    }


def _scalar_kernel_bytes(D=3):
    """Compulsory bytes of the scalar-transport and geometry kernels (same convention; ``n`` per
    node, ``i`` per face-node incidence).  Limited advection, Green-Gauss gradients."""
    fp = _face_pass(D)
    cgeo = (1 + D) * S8
    fgeo = (1 + 2 * D) * S8
    bnd = 1 + 2 * S8
    return {
        # launched once per step: the residual's forward record (with the stencil-extremum owners and
        # the limiter's minimum face) serves the VJP
        "scalar_grad": dict(c=4 + S8 + cgeo + D * S8 + S8 + 2 * 4 + 4, f=fp, b=bnd),
        "scalar_res": dict(c=4 + 3 * S8 + D * S8 + cgeo + S8, f=fp + S8, b=bnd),
        "scalar_res_adj": dict(c=4 + 4 * S8 + D * S8 + cgeo + 3 * S8 + D * S8 + D * S8 + S8, f=fp + 2 * S8 + fgeo,
                               b=bnd + 2 * S8),
        "scalar_lim_adj": dict(c=4 + 2 * D * S8 + D * S8 + 4 * S8 + S8 + 8 + 2 * S8 + D * S8 + 4 + cgeo
                               + 2 * S8 + 2 * D * S8, f=4 * 2 + D * S8),
        "scalar_grad_adj": dict(c=4 + D * S8 + S8 + 2 * S8 + S8 + 2 * S8 + 8 + 2 * S8 + 4 + D * S8 + cgeo
                                + 2 * D * S8, f=fp + 2 * fgeo, b=bnd + 4 * S8),
        "geom_face": dict(f=4 + fgeo, i=4 + D * S8),
        "geom_approx": dict(c=4 + D * S8, e=4 + D * S8),
        "geom_orient": dict(f=4 + 2 * D * S8 + D * S8, c=D * S8),
        "geom_volume": dict(c=4 + 2 * D * S8 + S8, f=fp),
        "geom_adj_cell": dict(c=4 + cgeo + 2 * D * S8 + S8 + D * S8 + (2 * D + 1) * S8, f=fp),
        "geom_adj_face": dict(f=4 + 4 + fgeo + fgeo, c=(2 * D + 1) * S8 + D * S8 + 4, i=4 + D * S8 + D * S8),
        "geom_adj_node": dict(n=4 + D * S8, i=4 + D * S8),
    }


def algorithmic_bytes(name, counts, workload="sst"):
    """Bytes per launch of kernel ``name`` for mesh counts {C, F, Fb, Cw, Fw, N, I, E}."""
    models = {"sst": _rans_kernel_bytes, "flow": _flow_kernel_bytes, "tet": _scalar_kernel_bytes}
    model = models[workload]()[name]
    return (model.get("c", 0) * counts["C"] + model.get("f", 0) * counts["F"] + model.get("b", 0) * counts["Fb"]
            + model.get("w", 0) * counts.get("Cw", 0) + model.get("wf", 0) * counts.get("Fw", 0)
            + model.get("n", 0) * counts.get("N", 0) + model.get("i", 0) * counts.get("I", 0)
            + model.get("e", 0) * counts.get("E", 0))


CHANNEL_BCS = (("left", 2), ("right", 3), ("bottom", 1), ("top", 1), ("back", 1), ("front", 1))
WALLS = ("bottom", "top", "back", "front")


def _channel_mesh(nx, ny, nz):
    from aquaflux_amd.mesh import structured_grid_3d

    return structured_grid_3d(nx, ny, nz, lx=2.0, named_boundaries=True)


def seeded_state(C, turbulent, seed=2):
    """u = 1 + 0.05 N(0,1), v, w = 0.05 N(0,1), p = N(0, 0.1) [, k ~ LogU(1e-5, 1e-2),
    omega ~ LogU(1, 1e3)], drawn in that order from ``default_rng(seed)``."""
    rng = np.random.default_rng(seed)
    parts = [1.0 + 0.05 * rng.normal(size=C), 0.05 * rng.normal(size=C), 0.05 * rng.normal(size=C),
             rng.normal(0.0, 0.1, C)]
    if turbulent:
        parts += [np.exp(rng.uniform(np.log(1e-5), np.log(1e-2), C)), np.exp(rng.uniform(0.0, np.log(1e3), C))]
    state = np.concatenate(parts)
    return state, rng.normal(size=state.size)


def build_workload(device, workload="sst", nx=128, ny=64, nz=64, nu=1e-5):
    """The 3D hex channel: inlet u = (1, 0, 0) on the left, pressure outlet on the right, no-slip
    walls elsewhere; least-squares gradients and Venkatakrishnan-limited upwind advection.  The SST
    variant adds k = 1e-3, omega = 1 at the inlet, k = 0 and the omega wall fixation on the walls,
    zero gradient at the outlet, nu = 1e-5."""
    import torch

    from aquaflux_amd.boundary import BoundaryConditions, Dirichlet, ZeroGradient
    from aquaflux_amd.discretization import LimitedUpwind
    from aquaflux_amd.flow import MomentumContinuity, NoSlipWall, PressureOutlet, VelocityInlet
    from aquaflux_amd.properties import Constant, PropertyModel
    from aquaflux_amd.schemes import LeastSquaresGradient, VenkatakrishnanLimiter

    mesh = _channel_mesh(nx, ny, nz)
    geom = mesh.geometry(device=device)
    bcs = BoundaryConditions({"left": VelocityInlet((1.0, 0.0, 0.0)), "right": PressureOutlet(0.0),
                              **{w: NoSlipWall() for w in WALLS}})
    props = PropertyModel({"viscosity": Constant(nu), "density": Constant(1.0)})
    adv = LimitedUpwind(VenkatakrishnanLimiter(5.0))
    mc = MomentumContinuity.build(mesh, geom, props, LeastSquaresGradient(), bcs, advection_scheme=adv)
    C = mesh.n_cells
    state, lam = seeded_state(C, workload == "sst")
    x = torch.tensor(state, dtype=torch.float64, device=device)
    lam_t = torch.tensor(lam, dtype=torch.float64, device=device)
    if workload != "sst":
        return mesh, mc, x, lam_t, state, lam
    from aquaflux_amd.turbulence import CoupledRANS, SSTModel, SSTTurbulence

    kb = BoundaryConditions({"left": Dirichlet(1e-3), "right": ZeroGradient(), **{w: Dirichlet(0.0) for w in WALLS}})
    wb = BoundaryConditions({"left": Dirichlet(1.0), "right": ZeroGradient(), **{w: ZeroGradient() for w in WALLS}})
    turb = SSTTurbulence.build(SSTModel(), mesh, geom, LeastSquaresGradient(), adv, 1.0,
                               torch.full((C,), nu, dtype=torch.float64), list(WALLS), kb, wb)
    return mesh, CoupledRANS.build(mc, turb), x, lam_t, state, lam


def _cpu_threads():
    """Threads for the CPU baseline: OMP_NUM_THREADS when set (the host's share of cores given to
    this process), else every core the process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def _time_cpu(fn, warmups, reps):
    """Median wall time of ``reps`` calls of ``fn`` after ``warmups`` untimed ones."""
    for _ in range(warmups):
        fn()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), times


def cpu_baseline(workload, nx, ny, nz, reps=10, warmups=3, nu=1e-5):
    """The CPU oracle (torch float64, reference pass structure) timed on a bounded sample: the same
    channel workload on an nx x ny x nz mesh, residual + VJP (torch autograd), median of ``reps``
    runs after ``warmups`` untimed ones, input construction outside the timed calls."""
    import torch

    from oracle import fv
    from oracle import sst as S

    threads = _cpu_threads()
    torch.set_num_threads(threads)
    mesh = _channel_mesh(nx, ny, nz)
    m = fv.OMesh.from_mesh(mesh)
    g = fv.geometry(m)
    F, d, C = mesh.n_faces, mesh.dim, mesh.n_cells
    kind = torch.zeros(F, dtype=torch.long)
    ub = torch.zeros((F, d), dtype=torch.float64)
    for name, k in CHANNEL_BCS:
        idx = torch.from_numpy(mesh.face_patches.indices(name))
        kind[idx] = k
        if name == "left":
            ub[idx, 0] = 1.0
    s = fv.FlowSetup(kind=kind, u_b=ub, p_b=torch.zeros(F, dtype=torch.float64), advection="limited",
                     gradient="least_squares")
    mu = torch.full((C,), nu, dtype=torch.float64)
    rho = torch.ones(C, dtype=torch.float64)
    state, lam = seeded_state(C, workload == "sst")
    x = torch.tensor(state, dtype=torch.float64, requires_grad=True)
    lam_t = torch.tensor(lam, dtype=torch.float64)
    if workload == "sst":
        wf = np.concatenate([mesh.face_patches.indices(w) for w in WALLS]).astype(np.int64)

        def scalar_bc(inlet, walls_dirichlet):
            kk = torch.zeros(F, dtype=torch.long)
            vv = torch.zeros(F, dtype=torch.float64)
            for name, _ in CHANNEL_BCS:
                idx = torch.from_numpy(mesh.face_patches.indices(name))
                if name == "left":
                    kk[idx], vv[idx] = fv.BC_DIRICHLET, inlet
                elif name in WALLS and walls_dirichlet:
                    kk[idx] = fv.BC_DIRICHLET
                else:
                    kk[idx] = fv.BC_ZERO_GRADIENT
            return fv.ScalarSetup(kind=kk, value=vv, t_inf=torch.zeros(F, dtype=torch.float64))

        t = S.build_sst_setup(m, g, "least_squares", S.SSTModel(), torch.full((C,), nu, dtype=torch.float64), 1.0,
                              wf, scalar_bc(1e-3, True), scalar_bc(1.0, False))

        def fn(z):
            return S.rans_residual(m, g, s, t, z, mu, rho, "limited")
    else:
        def fn(z):
            return fv.flow_residual(m, g, s, z, mu, rho)

    def run():
        torch.autograd.grad(fn(x), (x,), lam_t)

    tm, times = _time_cpu(run, warmups, reps)
    return {
        "value": F / tm / 1e9,
        "unit": "Gfaces/s",
        "cores": threads,
        "host_cpus": os.cpu_count(),
        "kind": "port",
        "sample": f"residual+VJP by the torch-CPU float64 oracle of the same channel workload at {nx}x{ny}x{nz} "
        f"({C} cells, {F} faces): median of {reps} runs after {warmups} warm-ups, {tm:.2f} s each "
        f"(min {min(times):.2f}, max {max(times):.2f}), {threads} threads",
    }


def cpu_openmp_line(op, prim, lam, reps=10, warmups=3):
    """The OpenMP C++ FP64 line: afx_rans_residual_vjp_host runs the device kernels' own cell bodies
    (gradient + limiter + closure, a_P, flow and k / omega face loops and their adjoints) one cell at
    a time over the host threads, on host copies of the full workload made before timing; median of
    ``reps`` runs after ``warmups`` untimed ones."""
    threads = _cpu_threads()
    run = op._raw.host_line(prim, lam)
    tm, times = _time_cpu(lambda: run(threads), warmups, reps)
    F, C = op._raw.mesh.n_faces, op._raw.mesh.n_cells
    return {
        "value": F / tm / 1e9,
        "unit": "Gfaces/s",
        "cores": threads,
        "host_cpus": os.cpu_count(),
        "kind": "port",
        "sample": f"residual+VJP of the full workload ({C} cells, {F} faces) by the OpenMP C++ FP64 build of the "
        f"same cell bodies as the HIP kernels (afx_rans_residual_vjp_host): median of {reps} runs after "
        f"{warmups} warm-ups, {tm * 1e3:.1f} ms each (min {min(times) * 1e3:.1f}, max {max(times) * 1e3:.1f}), "
        f"{threads} threads",
    }


TET_BCS = ("left", "right", "bottom", "top", "back", "front")


def tet_workload(device, n=96, steps=10, warmup=2):
    """Scalar advection-diffusion on the jittered 6-tet split of an n^3 unit cube (Pe = 100, u =
    (1, 0, 0) face flux (u . n) A, Venkatakrishnan-limited upwind, Green-Gauss gradients, Dirichlet 1
    on the left, zero gradient elsewhere).  One step = geometry from the node coordinates, the face
    flux, the residual, and the reverse pass to the node coordinates and phi through the hand-written
    residual and geometry adjoints (torch autograd composes them)."""
    import torch

    from aquaflux_amd import _lib
    from aquaflux_amd.boundary import BoundaryConditions, Dirichlet, ZeroGradient
    from aquaflux_amd.discretization import LimitedUpwind
    from aquaflux_amd.mesh import tet_grid_3d
    from aquaflux_amd.properties import Constant
    from aquaflux_amd.schemes import CompactGreenGauss, VenkatakrishnanLimiter
    from aquaflux_amd.transport import ScalarTransport, face_mass_flux

    t0 = time.perf_counter()
    mesh = tet_grid_3d(n, perturb=0.15, seed=3)
    X = torch.tensor(mesh.node_coords, dtype=torch.float64, device=device, requires_grad=True)
    bcs = BoundaryConditions({p: (Dirichlet(1.0) if p == "left" else ZeroGradient()) for p in TET_BCS})
    geom = mesh.geometry(node_coords=X, device=device)
    st = ScalarTransport.build(mesh, geom, Constant(0.01), bcs, LimitedUpwind(VenkatakrishnanLimiter(5.0)),
                               gradient_scheme=CompactGreenGauss())
    rng = np.random.default_rng(3)
    C, F = mesh.n_cells, mesh.n_faces
    phi = torch.tensor(rng.uniform(0.0, 1.0, C), device=device, requires_grad=True)
    lam = torch.tensor(rng.normal(size=C), device=device)
    print(f"bench: tet workload {n}^3 x 6 built ({C} cells, {F} faces) in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    u = (1.0, 0.0, 0.0)

    def step():  # R and its pullback of lam to (X, phi), as jax.vjp(f, X, phi)[1](lam)
        g = mesh.geometry(node_coords=X, device=device)
        st.geometry = g
        R = st.residual(face_mass_flux(g, u))(phi)
        return torch.autograd.grad(R, (X, phi), lam)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    timer = _lib.KernelTimer(max_launches=32 * (steps + 1))
    with timer:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t1
        kern = timer.collect()
    Fb = int(np.count_nonzero(mesh.neighbour < 0))
    counts = {"C": C, "F": F, "Fb": Fb, "N": mesh.n_nodes, "I": int(mesh.face_node_indices.size),
              "E": 2 * F - Fb}
    dom = max(kern, key=lambda k: kern[k][0])
    avg_ms = kern[dom][0] / kern[dom][1]
    nbytes = algorithmic_bytes(dom, counts, "tet")
    achieved = nbytes / (avg_ms * 1e-3) / 1e9
    ms_step = elapsed / steps * 1e3
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "avg_ms": round(avg_ms, 4),
                "algorithmic_bytes": int(nbytes), "step_bytes_per_face": STEP_BYTES_PER_FACE["tet"],
                "step_frac": round(STEP_BYTES_PER_FACE["tet"] * F / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if n == 96:
        tb = _traffic("scalar_tet6x96", dom)
        if tb is not None:
            roofline["traffic"] = tb
            roofline["traffic_source"] = "rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated (profiles/)"
    return mesh, {
        "metric": "Gfaces/s for residual+VJP",
        "value": round(F * steps / elapsed / 1e9, 4),
        "unit": "Gfaces/s",
        "ms_per_step": round(elapsed / steps * 1e3, 4),
        "steps": steps,
        "dtype": "f64",
        "config": {"workload": f"3D jittered tetrahedra, 6 x {n}^3 ({C:,} cells, {F:,} faces), scalar "
                   "advection-diffusion Pe=100, Venkatakrishnan-limited upwind, Green-Gauss; step = geometry + "
                   "flux + residual + adjoint to node coordinates and phi", "cells": C, "faces": F},
        "kernel_avg_ms": {k: round(v[0] / v[1], 4) for k, v in sorted(kern.items())},
        "roofline": roofline,
    }


def tet_cpu_baseline(n=32, reps=10, warmups=3):
    """The CPU oracle on a bounded sample of the tet workload: geometry, flux, residual and torch
    autograd to node coordinates and phi on the 6 x n^3 mesh, median of ``reps`` runs after
    ``warmups`` untimed ones."""
    import torch

    from aquaflux_amd.mesh import tet_grid_3d
    from oracle import fv

    threads = _cpu_threads()
    torch.set_num_threads(threads)
    mesh = tet_grid_3d(n, perturb=0.15, seed=3)
    F, C = mesh.n_faces, mesh.n_cells
    rng = np.random.default_rng(3)
    phi = torch.tensor(rng.uniform(0.0, 1.0, C), requires_grad=True)
    lam = torch.tensor(rng.normal(size=C))
    kind = torch.where(torch.from_numpy(mesh.neighbour >= 0), 0, fv.BC_ZERO_GRADIENT)
    left = torch.from_numpy(np.asarray(mesh.face_patches.indices("left"), dtype=np.int64))
    kind[left] = fv.BC_DIRICHLET
    value = torch.zeros(F, dtype=torch.float64)
    value[left] = 1.0
    s = fv.ScalarSetup(kind=kind, value=value, t_inf=torch.zeros(F, dtype=torch.float64), advection="limited",
                       gradient="green_gauss")
    gam = torch.full((C,), 0.01, dtype=torch.float64)
    X = torch.tensor(mesh.node_coords, dtype=torch.float64, requires_grad=True)

    def run():
        m = fv.OMesh.from_mesh(mesh, node_coords=X)
        g = fv.geometry(m)
        R = fv.scalar_residual(m, g, s, phi, g.normal[:, 0] * g.area, gam)
        torch.autograd.grad(R, (X, phi), lam)

    tm, times = _time_cpu(run, warmups, reps)
    return {"value": F / tm / 1e9, "unit": "Gfaces/s", "cores": threads, "host_cpus": os.cpu_count(),
            "kind": "port",
            "sample": f"geometry + flux + residual + adjoint by the torch-CPU float64 oracle of the same tet "
            f"workload at 6 x {n}^3 ({C} cells, {F} faces): median of {reps} runs after {warmups} warm-ups, "
            f"{tm:.2f} s each (min {min(times):.2f}, max {max(times):.2f}), {threads} threads"}


def newton_compare(device, nx=200, ny=100, nz=100, mu=1e-3, max_steps=80, cfl0=10.0, preconditioner="block-simpler",
                   block_options=None, restart=60, fp32_outer_basis="fp32", inner_restart=50):
    """Newton to ||R|| <= 1e-12 + 1e-10 ||R0|| on the hex box (u, v, w, p) at Re = 1/mu, from the
    seeded state, twice: FP32 inner solves on the FP32 compact block Jacobian, and the all-FP64 path
    (FP64 Jacobian, FP64 inner solves, FP64 outer Krylov basis).  ``preconditioner``: "inner"
    (GMRES(``inner_restart``) with point-block Jacobi) or "block-<composition>" (the SIMPLE-family
    block preconditioner with that composition, further ``BlockPreconditioner.build`` arguments in
    ``block_options``).  The outer FGMRES restarts every ``restart`` iterations; the FP32 run keeps
    its outer basis in ``fp32_outer_basis`` precision (its residuals, JVPs and update stay FP64)."""
    import torch

    from aquaflux_amd.boundary import BoundaryConditions
    from aquaflux_amd.discretization import FirstOrderUpwind
    from aquaflux_amd.flow import MomentumContinuity, NoSlipWall, PressureOutlet, VelocityInlet
    from aquaflux_amd.mesh import structured_grid_3d
    from aquaflux_amd.properties import Constant, PropertyModel
    from aquaflux_amd.schemes import LeastSquaresGradient
    from aquaflux_amd.block_preconditioner import COMPOSITIONS, BlockPreconditioner
    from aquaflux_amd.solve import DampedNewtonStep, ImplicitNewtonSolver

    if preconditioner != "inner" and preconditioner.removeprefix("block-") not in COMPOSITIONS:
        raise ValueError(f"unknown Newton preconditioner {preconditioner!r}")
    mesh = structured_grid_3d(nx, ny, nz, named_boundaries=True)
    geom = mesh.geometry(device=device)
    bcs = BoundaryConditions({"left": VelocityInlet((1.0, 0.0, 0.0)), "right": PressureOutlet(0.0),
                              "bottom": NoSlipWall(), "top": NoSlipWall(), "back": NoSlipWall(), "front": NoSlipWall()})
    mc = MomentumContinuity.build(mesh, geom, PropertyModel({"viscosity": Constant(mu), "density": Constant(1.0)}),
                                  LeastSquaresGradient(), bcs, advection_scheme=FirstOrderUpwind())
    rng = np.random.default_rng(2)
    C = mesh.n_cells
    x0 = torch.tensor(np.concatenate([1.0 + 0.05 * rng.normal(size=C), 0.05 * rng.normal(size=C),
                                      0.05 * rng.normal(size=C), rng.normal(0.0, 0.1, C)]), dtype=torch.float64,
                      device=device)
    t_setup = time.perf_counter()

    def make_step(prec, **kw):
        # a 60-vector outer FGMRES basis: the block preconditioner's steps take up to ~100 outer
        # iterations, and restarting at 30 cost 7% on the 2M-cell box
        basis = fp32_outer_basis if prec == "fp32" else "fp64"
        if preconditioner == "inner":
            kw.setdefault("inner_iters", inner_restart)
            return DampedNewtonStep(preconditioner="inner", precision=prec, restart=restart, outer_basis=basis,
                                    inner_restart=inner_restart, **kw)
        bp = BlockPreconditioner.build(mc, composition=preconditioner.removeprefix("block-"), **(block_options or {}))
        return DampedNewtonStep(preconditioner="block", precision=prec, block_preconditioner=bp, restart=restart,
                                outer_basis=basis, **kw)

    for prec in ("fp32", "fp64"):  # warm up every kernel and allocation outside the timed solves
        make_step(prec, max_iter=2, inner_iters=2).correction(mc, x0, mc.residual_raw(x0))
    torch.cuda.synchronize()
    print(f"newton: setup + warmup {time.perf_counter() - t_setup:.2f} s", file=sys.stderr, flush=True)
    out = {"workload": f"hex box {nx}x{ny}x{nz} ({C} cells, {4 * C} unknowns), Re={1.0 / mu:g}, first-order upwind, "
                       "least-squares gradients, seeded state; pseudo-transient continuation (SER, CFL0="
                       f"{cfl0:g}) into Newton; stop at ||R|| <= 1e-12 + 1e-10 ||R0||",
           "preconditioner": preconditioner if preconditioner != "inner" else
           f"inner GMRES({inner_restart}) on the block Jacobian with point-block Jacobi",
           "block_options": block_options or {}, "outer_restart": restart,
           "outer_basis": {"fp32": fp32_outer_basis, "fp64": "fp64"}}
    roots = {}
    for prec in ("fp32", "fp64"):
        def progress(step, rel, kits, alpha, secs, prec=prec):
            print(f"newton[{prec}] step {step}: |R|/|R0| = {rel:.3e}, krylov {kits}, alpha {alpha:g}, "
                  f"{secs:.2f} s", file=sys.stderr, flush=True)

        solver = ImplicitNewtonSolver(max_steps=max_steps, progress=progress,
                                      forward_step=make_step(prec, ptc_cfl0=cfl0))
        torch.cuda.synchronize()
        phi, rep = solver.forward(mc, x0)
        roots[prec] = phi
        out[prec] = {"seconds": round(rep.seconds, 3), "newton_steps": rep.steps,
                     "krylov_iterations": rep.krylov_iterations, "final_relative_residual":
                     rep.residual_norms[-1] / rep.residual_norms[0]}
        print(json.dumps({"newton": prec, **out[prec]}), file=sys.stderr, flush=True)
    d = float((roots["fp32"] - roots["fp64"]).abs().max() / roots["fp64"].abs().max())
    out["state_difference_fp32_vs_fp64"] = d
    out["speedup_fp32_inner"] = round(out["fp64"]["seconds"] / out["fp32"]["seconds"], 3)
    return out


def calibrate_copy(mib: int, reps: int = 5) -> int:
    """Copies ``mib`` MiB device to device ``reps`` times: FETCH_SIZE / WRITE_SIZE of these
    dispatches should read ``mib`` MiB each, which calibrates the counters' units."""
    import torch

    n = mib * 1024 * 1024 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.copy_(a)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(json.dumps({"copy_mib": mib, "reps": reps + 1, "copy_GBps": 2 * n * 8 / dt / 1e9}), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=("sst", "flow", "tet"), default="sst")
    ap.add_argument("--no-tet", action="store_true", help="skip the tetrahedral scalar-transport line")
    ap.add_argument("--tet-n", type=int, default=96)
    ap.add_argument("--size", type=int, nargs=3, default=(128, 64, 64), metavar=("NX", "NY", "NZ"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=10, help="timed CPU-baseline runs (after 3 warm-ups)")
    ap.add_argument("--cpu-size", type=int, nargs=3, default=(128, 64, 16), metavar=("NX", "NY", "NZ"))
    ap.add_argument("--no-newton", action="store_true", help="skip the Newton time-to-1e-10 comparison")
    ap.add_argument("--newton-only", action="store_true", help="only the Newton comparison")
    ap.add_argument("--newton-size", type=int, nargs=3, default=(200, 100, 100), metavar=("NX", "NY", "NZ"))
    ap.add_argument("--newton-cfl0", type=float, default=10.0)
    ap.add_argument("--newton-mu", type=float, default=1e-3, help="viscosity of the Newton box (Re = 1/mu)")
    ap.add_argument("--newton-preconditioner", default=None,
                    choices=["inner", "block-triangular", "block-simple", "block-simpler"],
                    help="one Newton comparison with this preconditioner (default: the inner GMRES(30) "
                    "point-block-Jacobi comparison, then the block-SIMPLER one)")
    ap.add_argument("--newton-block-options", type=json.loads, default=None, metavar="JSON",
                    help='further BlockPreconditioner.build arguments, e.g. \'{"schur_iters": 4}\'')
    ap.add_argument("--newton-restart", type=int, default=60, help="outer FGMRES restart of the Newton solves")
    ap.add_argument("--newton-outer-basis", default=None, choices=["fp64", "fp32"],
                    help="precision of the FP32 run's outer FGMRES basis (residuals, JVPs and updates stay FP64; "
                    "default fp64 for the inner GMRES comparison, fp32 for the block ones); the FP64 run is all-FP64")
    ap.add_argument("--newton-inner-restart", type=int, default=30, help="inner GMRES basis of the 'inner' comparison")
    ap.add_argument(
        "--calibrate-copy",
        type=int,
        default=0,
        metavar="MIB",
        help="only run plain device-to-device copies of this many MiB (a known-traffic reference "
        "for hardware byte counters) and exit",
    )
    args = ap.parse_args(argv)
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only: this build's path runs on a single MI355X "
            f"(--gpus {args.gpus} requested)",
            file=sys.stderr,
        )
        return 2

    import torch

    from aquaflux_amd import _lib

    dev = torch.device("cuda", 0)
    def newton_lines():
        precs = [args.newton_preconditioner] if args.newton_preconditioner else ["inner", "block-simpler"]
        out = {}
        for prec in precs:
            basis = args.newton_outer_basis or ("fp64" if prec == "inner" else "fp32")
            key = "newton" if prec == precs[0] else "newton_block"
            out[key] = newton_compare(dev, *args.newton_size, mu=args.newton_mu, cfl0=args.newton_cfl0,
                                      preconditioner=prec, block_options=args.newton_block_options,
                                      restart=args.newton_restart, fp32_outer_basis=basis,
                                      inner_restart=args.newton_inner_restart)
            torch.cuda.empty_cache()
        return out

    if args.newton_only:
        print(json.dumps(newton_lines()), flush=True)
        return 0
    wl = args.workload
    if wl == "tet":
        _, tres = tet_workload(dev, args.tet_n, steps=args.steps, warmup=args.warmup)
        tres.update({"n_gpus": 1, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
                     "vs_baseline": None, "data": "synthetic (jittered mesh seed 3, phi ~ U(0,1), cotangent N(0,1))",
                     "cpu_baseline": None if args.no_cpu_baseline else tet_cpu_baseline(reps=args.cpu_reps)})
        print(json.dumps(tres), flush=True)
        return 0
    nx, ny, nz = args.size
    mesh, op, x, lam, _, _ = build_workload(dev, wl, nx, ny, nz)
    raw = op._raw
    prim = op._detached(x) if wl == "sst" else [x] + [t.detach() for t in op._inputs()[:4]]
    print(f"bench: {wl} workload {nx}x{ny}x{nz} built ({mesh.n_cells} cells, {mesh.n_faces} faces)",
          file=sys.stderr, flush=True)

    if wl == "sst":
        def step():  # R and its pullback from one forward pass, the pair jax.vjp returns
            raw.residual_vjp(prim, lam)
    else:
        def step():
            raw.residual(*prim)
            raw.vjp(prim, lam)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    timer = _lib.KernelTimer(max_launches=16 * (args.steps + 1))
    with timer:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        kern = timer.collect()
    F, C = mesh.n_faces, mesh.n_cells
    Fb = int(np.count_nonzero(mesh.neighbour < 0))
    counts = {"C": C, "F": F, "Fb": Fb}
    if wl == "sst":
        wf = op.turbulence.wall_faces
        counts["Fw"] = int(wf.numel())
        counts["Cw"] = int(op.turbulence.wall_cells.numel())
    ms_per_step = elapsed / args.steps * 1e3
    value = F * args.steps / elapsed / 1e9
    dom = max(kern, key=lambda k: kern[k][0])
    avg_ms = kern[dom][0] / kern[dom][1]
    nbytes = algorithmic_bytes(dom, counts, wl)
    achieved = nbytes / (avg_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "avg_ms": round(avg_ms, 4),
        "algorithmic_bytes": int(nbytes),
    }
    if wl in STEP_BYTES_PER_FACE:
        step_gbs = STEP_BYTES_PER_FACE[wl] * F / (ms_per_step * 1e-3) / 1e9
        roofline["step_bytes_per_face"] = STEP_BYTES_PER_FACE[wl]
        roofline["step_frac"] = round(step_gbs / HBM_PEAK_GBS, 4)
    tag = "rans" if wl == "sst" else "flow"
    tb = _traffic(f"{tag}_hex{nx}x{ny}x{nz}", dom)
    if tb is not None:
        roofline["traffic"] = tb
        roofline["traffic_source"] = "rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated (profiles/)"
    kernels = {k: round(v[0] / v[1], 4) for k, v in sorted(kern.items())}
    desc = {
        "sst": "coupled u,v,w,p,k,omega with the k-omega SST closure (wall functions, omega wall fixation)",
        "flow": "coupled u,v,w,p (laminar)",
    }[wl]
    out = {
        "metric": "Gfaces/s for residual+VJP",
        "value": round(value, 4),
        "unit": "Gfaces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded state u=1+0.05N, v,w=0.05N, p=N(0,0.1)"
        + (", k~LogU(1e-5,1e-2), omega~LogU(1,1e3)" if wl == "sst" else "")
        + "; seeded cotangent N(0,1))",
        "config": {
            "workload": f"3D hex channel {nx}x{ny}x{nz} ({C:,} cells, {F:,} faces), {desc}, least-squares "
            "gradients, Venkatakrishnan-limited upwind, inlet/outlet/walls; residual + VJP w.r.t. state, "
            "viscosity, density, boundary values",
            "cells": C,
            "faces": F,
        },
        "kernel_avg_ms": kernels,
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_tet:
        _, tres = tet_workload(dev, args.tet_n, steps=10, warmup=2)
        out["tet_scalar"] = {k: tres[k] for k in ("value", "unit", "ms_per_step", "config", "kernel_avg_ms",
                                                  "roofline")}
        if not args.no_cpu_baseline:
            out["tet_scalar"]["cpu_baseline"] = tet_cpu_baseline(reps=args.cpu_reps)
        torch.cuda.empty_cache()
    if not args.no_newton:
        out.update(newton_lines())
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, *args.cpu_size, reps=args.cpu_reps)
        if wl == "sst":
            out["cpu_openmp"] = cpu_openmp_line(op, prim, lam, reps=args.cpu_reps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
