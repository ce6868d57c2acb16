"""Benchmark: residual + VJP throughput (Gfaces/s) of the face-loop residual on one MI355X.

One step = one residual evaluation R(state) plus one adjoint J^T lambda (cotangents of the state,
viscosity, density and boundary values), all inputs resident in HBM.  The workload is the 3D hex
channel of 128 x 64 x 64 cells (524,288 cells, 1,593,344 faces) with a seeded turbulent-like state.

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

# Algorithmic (compulsory) bytes per launch of each kernel: every array the kernel must touch,
# counted once.  Units: per cell (c), per face (f), per cell-face entry (e), per boundary face (b).
# D = 3, NV = 4 flow variables, least-squares gradient, limited advection.
def _kernel_bytes(D=3, NV=4, limited=True):
    S = 8
    G = NV * D * S
    psi = D * S if limited else 0
    topo = dict(c=4, e=4, f=8)  # cell_ptr, cell->face list, owner+neighbour
    cgeo = (1 + D) * S  # volume + centroid
    fgeo = (1 + 2 * D) * S  # area + normal + centroid
    bnd = 1 + D * S + S  # kind + u_b + p_b
    state = NV * S
    m = {}
    m["flow_grad"] = dict(c=topo["c"] + state + cgeo + G + psi, e=topo["e"], f=topo["f"] + fgeo, b=bnd)
    m["flow_ap"] = dict(c=topo["c"] + state + G + 2 * S + cgeo + S, e=topo["e"], f=topo["f"] + fgeo, b=1)
    m["flow_res"] = dict(
        c=topo["c"] + state + G + 3 * S + psi + cgeo + state, e=topo["e"], f=topo["f"] + fgeo, b=bnd
    )
    m["flow_res_adj"] = dict(
        c=topo["c"] + state + G + 3 * S + psi + cgeo + state + state + G + 3 * S + psi,
        e=topo["e"],
        f=topo["f"] + fgeo + (D + 1) * S,
        b=bnd,
    )
    m["flow_ap_adj"] = dict(
        c=topo["c"] + state + G + 2 * S + cgeo + S + (state + 2 * S) * 2 + G * 2 + psi + 2 * D * S + 2 * D * 4,
        e=topo["e"],
        f=topo["f"] + fgeo,
        b=1,
    )
    m["flow_grad_adj"] = dict(
        c=topo["c"] + cgeo + G + state * 2 + 2 * D * S + 2 * D * 4, e=topo["e"], f=topo["f"] + fgeo, b=bnd + (D + 1) * S
    )
    return m


def algorithmic_bytes(name, C, F, E, Fb):
    k = _kernel_bytes()[name]
    return k["c"] * C + k["f"] * F + k["e"] * E + k["b"] * Fb


def build_workload(device, nx=128, ny=64, nz=64):
    import torch

    from aquaflux_amd.boundary import BoundaryConditions
    from aquaflux_amd.discretization import LimitedUpwind
    from aquaflux_amd.flow import MomentumContinuity, NoSlipWall, PressureOutlet, VelocityInlet
    from aquaflux_amd.mesh import structured_grid_3d
    from aquaflux_amd.properties import Constant, PropertyModel
    from aquaflux_amd.schemes import LeastSquaresGradient, VenkatakrishnanLimiter

    mesh = structured_grid_3d(nx, ny, nz, named_boundaries=True)
    geom = mesh.geometry(device=device)
    bcs = BoundaryConditions(
        {
            "left": VelocityInlet((1.0, 0.0, 0.0)),
            "right": PressureOutlet(0.0),
            "bottom": NoSlipWall(),
            "top": NoSlipWall(),
            "back": NoSlipWall(),
            "front": NoSlipWall(),
        }
    )
    props = PropertyModel({"viscosity": Constant(1e-5), "density": Constant(1.0)})
    mc = MomentumContinuity.build(
        mesh, geom, props, LeastSquaresGradient(), bcs, advection_scheme=LimitedUpwind(VenkatakrishnanLimiter(5.0))
    )
    rng = np.random.default_rng(2)
    C = mesh.n_cells
    state = np.concatenate(
        [1.0 + 0.05 * rng.normal(size=C), 0.05 * rng.normal(size=C), 0.05 * rng.normal(size=C), rng.normal(0.0, 0.1, C)]
    )
    lam = rng.normal(size=state.size)
    x = torch.tensor(state, dtype=torch.float64, device=device)
    lam_t = torch.tensor(lam, dtype=torch.float64, device=device)
    return mesh, geom, mc, x, lam_t, state, lam


def cpu_baseline(mesh, state, lam, reps=2):
    """The CPU oracle (torch float64, reference pass structure) on the same workload."""
    import torch

    from oracle import fv

    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    m = fv.OMesh.from_mesh(mesh)
    g = fv.geometry(m)
    F, d, C = mesh.n_faces, mesh.dim, mesh.n_cells
    kind = torch.zeros(F, dtype=torch.long)
    ub = torch.zeros((F, d), dtype=torch.float64)
    for name, k in (("left", 2), ("right", 3), ("top", 1), ("bottom", 1), ("back", 1), ("front", 1)):
        idx = torch.from_numpy(mesh.face_patches.indices(name))
        kind[idx] = k
        if name == "left":
            ub[idx, 0] = 1.0
    s = fv.FlowSetup(kind=kind, u_b=ub, p_b=torch.zeros(F, dtype=torch.float64), advection="limited",
                     gradient="least_squares")
    x = torch.tensor(state, dtype=torch.float64, requires_grad=True)
    mu = torch.full((C,), 1e-5, dtype=torch.float64)
    rho = torch.ones(C, dtype=torch.float64)
    lam_t = torch.tensor(lam, dtype=torch.float64)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        R = fv.flow_residual(m, g, s, x, mu, rho)
        torch.autograd.grad(R, (x,), lam_t)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {
        "value": F / t / 1e9,
        "unit": "Gfaces/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{reps} residual+VJP evaluations of the full workload ({F} faces) by the torch-CPU "
        f"float64 oracle, median {t:.2f} s each",
    }


def newton_compare(device, nx=200, ny=100, nz=100, mu=1e-3, max_steps=80, cfl0=10.0):
    """Newton to ||R|| <= 1e-12 + 1e-10 ||R0|| on the Re = 1000 hex box (u, v, w, p), from the seeded
    state, with the inner Krylov solve on the compact block Jacobian in FP32 and in FP64."""
    import torch

    from aquaflux_amd.boundary import BoundaryConditions
    from aquaflux_amd.discretization import FirstOrderUpwind
    from aquaflux_amd.flow import MomentumContinuity, NoSlipWall, PressureOutlet, VelocityInlet
    from aquaflux_amd.mesh import structured_grid_3d
    from aquaflux_amd.properties import Constant, PropertyModel
    from aquaflux_amd.schemes import LeastSquaresGradient
    from aquaflux_amd.solve import DampedNewtonStep, ImplicitNewtonSolver

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
    for prec in ("fp32", "fp64"):  # warm up every kernel and allocation outside the timed solves
        DampedNewtonStep(preconditioner="inner", precision=prec, max_iter=2, inner_iters=2).correction(
            mc, x0, mc.residual_raw(x0))
    torch.cuda.synchronize()
    print(f"newton: setup + warmup {time.perf_counter() - t_setup:.2f} s", file=sys.stderr, flush=True)
    out = {"workload": f"hex box {nx}x{ny}x{nz} ({C} cells, {4 * C} unknowns), Re={1.0 / mu:g}, first-order upwind, "
                       "least-squares gradients, seeded state; pseudo-transient continuation (SER, CFL0="
                       f"{cfl0:g}) into Newton; stop at ||R|| <= 1e-12 + 1e-10 ||R0||"}
    roots = {}
    for prec in ("fp32", "fp64"):
        def progress(step, rel, kits, alpha, secs, prec=prec):
            print(f"newton[{prec}] step {step}: |R|/|R0| = {rel:.3e}, krylov {kits}, alpha {alpha:g}, "
                  f"{secs:.2f} s", file=sys.stderr, flush=True)

        solver = ImplicitNewtonSolver(max_steps=max_steps, progress=progress,
                                      forward_step=DampedNewtonStep(preconditioner="inner", precision=prec,
                                                                    ptc_cfl0=cfl0))
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
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=2)
    ap.add_argument("--no-newton", action="store_true", help="skip the Newton time-to-1e-10 comparison")
    ap.add_argument("--newton-only", action="store_true", help="only the Newton comparison")
    ap.add_argument("--newton-size", type=int, nargs=3, default=(200, 100, 100), metavar=("NX", "NY", "NZ"))
    ap.add_argument("--newton-cfl0", type=float, default=10.0)
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
    if args.newton_only:
        print(json.dumps(newton_compare(dev, *args.newton_size, cfl0=args.newton_cfl0)), flush=True)
        return 0
    mesh, geom, mc, x, lam, state_np, lam_np = build_workload(dev)
    raw = mc._raw
    mu, rho, ub, pb, _ = mc._inputs()
    prim = [x, mu, rho, ub, pb]

    def step():
        raw.residual(x, mu, rho, ub, pb)
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
    E = 2 * (F - Fb) + Fb
    ms_per_step = elapsed / args.steps * 1e3
    value = F * args.steps / elapsed / 1e9
    dom = max(kern, key=lambda k: kern[k][0])
    avg_ms = kern[dom][0] / kern[dom][1]
    nbytes = algorithmic_bytes(dom, C, F, E, Fb)
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
    traffic_file = os.path.join(ROOT, "profiles", "flow_hex128x64x64_traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as fh:
            tk = json.load(fh).get("kernels", {}).get(dom)
        if tk:
            roofline["traffic"] = tk["hbm_bytes_per_launch"]
            roofline["traffic_source"] = "rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated (profiles/)"
    kernels = {k: round(v[0] / v[1], 4) for k, v in sorted(kern.items())}
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
        "data": "synthetic (seeded state u=1+0.05N, v,w=0.05N, p=N(0,0.1); seeded cotangent N(0,1))",
        "config": {
            "workload": "3D hex channel 128x64x64 (524,288 cells, 1,593,344 faces), coupled u,v,w,p, "
            "least-squares gradients, Venkatakrishnan-limited upwind, inlet/outlet/walls; "
            "residual + VJP w.r.t. state, viscosity, density, boundary values (k-omega SST closure not "
            "in this step)",
            "cells": C,
            "faces": F,
        },
        "kernel_avg_ms": kernels,
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_newton:
        out["newton"] = newton_compare(dev)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(mesh, state_np, lam_np, reps=args.cpu_reps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
