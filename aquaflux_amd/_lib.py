"""ctypes binding of ``libaquaflux_amd.so`` (the C ABI declared in ``include/aquaflux_amd.h``).

The library is loaded lazily on first use.  There is no fallback: when the shared library is
missing or a GPU is not available, every compute entry point raises.
"""

from __future__ import annotations

import ctypes
import os
import pathlib

import torch

_LIB = None
LIB_PATH = pathlib.Path(__file__).resolve().parent / "libaquaflux_amd.so"

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_int64_p = ctypes.POINTER(ctypes.c_int64)
c_void_p = ctypes.c_void_p


class AfxGeometry(ctypes.Structure):
    _fields_ = [
        ("area", c_void_p),
        ("normal", c_void_p),
        ("face_centroid", c_void_p),
        ("volume", c_void_p),
        ("cell_centroid", c_void_p),
        ("face_pack", c_void_p),
    ]


class AfxFlowConfig(ctypes.Structure):
    _fields_ = [
        ("advection", ctypes.c_int32),
        ("gradient", ctypes.c_int32),
        ("pressure_pin", ctypes.c_int64),
        ("pressure_pin_value", ctypes.c_double),
        ("body_force", ctypes.c_double * 3),
        ("limiter_k", ctypes.c_double),
        ("face_kind", c_void_p),
        ("gradient_sweeps", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("gradient_relaxation", ctypes.c_double),
    ]


class AfxFlowInputs(ctypes.Structure):
    _fields_ = [
        ("state", c_void_p),
        ("mu", c_void_p),
        ("rho", c_void_p),
        ("u_b", c_void_p),
        ("p_b", c_void_p),
        ("gamma_b", c_void_p),
    ]


class AfxFlowCotangents(ctypes.Structure):
    _fields_ = [
        ("state", c_void_p),
        ("mu", c_void_p),
        ("rho", c_void_p),
        ("u_b", c_void_p),
        ("p_b", c_void_p),
        ("gamma_b", c_void_p),
    ]


class AfxRansConfig(ctypes.Structure):
    _fields_ = [
        ("flow", AfxFlowConfig),
        ("scalar_advection", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("density", ctypes.c_double),
        ("sst", ctypes.c_double * 16),
        ("k_kind", c_void_p),
        ("omega_kind", c_void_p),
        ("wall_face", c_void_p),
        ("wall_cell", c_void_p),
        ("wall_distance", c_void_p),
        ("wall_distance_gradient", c_void_p),
        ("k_transform", ctypes.c_int32),
        ("omega_transform", ctypes.c_int32),
    ]


class AfxRansInputs(ctypes.Structure):
    _fields_ = [
        ("state", c_void_p),
        ("mu", c_void_p),
        ("rho", c_void_p),
        ("nu", c_void_p),
        ("u_b", c_void_p),
        ("p_b", c_void_p),
        ("k_value", c_void_p),
        ("omega_value", c_void_p),
    ]


class AfxRansCotangents(ctypes.Structure):
    _fields_ = [
        ("state", c_void_p),
        ("mu", c_void_p),
        ("rho", c_void_p),
        ("nu", c_void_p),
        ("u_b", c_void_p),
        ("p_b", c_void_p),
    ]


class AfxGeometryCotangents(ctypes.Structure):
    _fields_ = [
        ("area", c_void_p),
        ("normal", c_void_p),
        ("face_centroid", c_void_p),
        ("volume", c_void_p),
        ("cell_centroid", c_void_p),
    ]


class AfxScalarConfig(ctypes.Structure):
    _fields_ = [
        ("advection", ctypes.c_int32),
        ("gradient", ctypes.c_int32),
        ("limiter_k", ctypes.c_double),
        ("bc_kind", c_void_p),
        ("gradient_sweeps", ctypes.c_int32),
        ("transient", ctypes.c_int32),
        ("gradient_relaxation", ctypes.c_double),
        ("dt", ctypes.c_double),
    ]


class AfxScalarInputs(ctypes.Structure):
    _fields_ = [
        ("phi", c_void_p),
        ("gamma", c_void_p),
        ("flux", c_void_p),
        ("value", c_void_p),
        ("t_inf", c_void_p),
        ("phi_old", c_void_p),
        ("phi_older", c_void_p),
    ]


class AfxScalarCotangents(ctypes.Structure):
    _fields_ = [
        ("phi", c_void_p),
        ("gamma", c_void_p),
        ("flux", c_void_p),
        ("value", c_void_p),
        ("t_inf", c_void_p),
        ("geometry", AfxGeometryCotangents),
        ("phi_old", c_void_p),
        ("phi_older", c_void_p),
    ]


# Declared entry points: name -> (argtypes). Every function returns int except afx_last_error.
_SIGNATURES = {
    "afx_mesh_create": [
        ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
        c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_void_p),
    ],
    "afx_mesh_destroy": [c_void_p],
    "afx_mesh_info": [c_void_p, c_void_p],
    "afx_mesh_cell_faces": [c_void_p, c_void_p, c_void_p, c_void_p],
    "afx_mesh_face_order": [c_void_p, c_void_p, c_void_p, c_void_p],
    "afx_mesh_set_block_order": [c_void_p, c_void_p, ctypes.c_int64],
    "afx_block_colouring": [ctypes.c_int64, ctypes.c_int64, c_void_p, c_void_p, ctypes.c_int32, c_void_p, c_void_p,
                            c_void_p, c_void_p, c_void_p],
    "afx_probe_seed": [ctypes.c_int64, ctypes.c_int32, c_void_p, ctypes.c_int32, ctypes.c_int32, c_void_p, c_void_p],
    "afx_probe_scatter": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, c_void_p, c_void_p, ctypes.c_int32,
                          c_void_p, ctypes.c_int32, ctypes.c_int32, c_void_p, ctypes.c_int, c_void_p, c_void_p],
    "afx_geometry_pack": [c_void_p, ctypes.POINTER(AfxGeometry), c_void_p, c_void_p],
    "afx_geometry_compute": [c_void_p, c_void_p, ctypes.POINTER(AfxGeometry), c_void_p],
    "afx_geometry_compute_record": [c_void_p, c_void_p, ctypes.POINTER(AfxGeometry), c_void_p, c_void_p],
    "afx_flow_workspace_bytes": [c_void_p, ctypes.c_int, c_int64_p],
    "afx_flow_residual": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxFlowConfig), ctypes.POINTER(AfxFlowInputs),
        c_void_p, c_void_p, ctypes.c_int64, c_void_p,
    ],
    "afx_flow_jvp": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxFlowConfig), ctypes.POINTER(AfxFlowInputs),
        ctypes.POINTER(AfxFlowInputs), c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p,
    ],
    "afx_flow_vjp": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxFlowConfig), ctypes.POINTER(AfxFlowInputs),
        c_void_p, ctypes.POINTER(AfxFlowCotangents), c_void_p, ctypes.c_int64, c_void_p,
    ],
    "afx_bell_slots": [c_void_p, c_void_p],
    "afx_flow_jacobian": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxFlowConfig), ctypes.POINTER(AfxFlowInputs),
        ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
    ],
    "afx_bell_spmv_t": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "afx_bell_spmv": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "afx_pbj_setup": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "afx_pbj_apply": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "afx_ell_jacobi": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                       ctypes.c_double, c_void_p],
    "afx_ell_residual": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                         c_void_p],
    "afx_aggregate_restrict": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "afx_aggregate_prolong": [ctypes.c_int64, ctypes.c_int, c_void_p, c_void_p, ctypes.c_double, c_void_p, c_void_p],
    "afx_dot": [ctypes.c_int64, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "afx_dense_matvec": [ctypes.c_int64, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "afx_basis_combine": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int, c_void_p, ctypes.c_int64, c_void_p, c_void_p,
                          c_void_p],
    "afx_geometry_jvp_workspace_bytes": [c_void_p, c_void_p],
    "afx_geometry_jvp": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p],
    "afx_face_flux": [ctypes.c_int64, ctypes.c_int, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "afx_face_flux_vjp": [ctypes.c_int64, ctypes.c_int, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                          c_void_p, c_void_p, c_void_p],
    "afx_gmres_sizes": [ctypes.c_int64, ctypes.c_int32, c_void_p, c_void_p],
    "afx_gmres_start": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int, ctypes.c_int, c_void_p, ctypes.c_double,
                        ctypes.c_double, ctypes.c_int, c_void_p, ctypes.c_int64, c_void_p, c_void_p, c_void_p],
    "afx_gmres_arnoldi": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int, ctypes.c_int, c_void_p, c_void_p,
                          ctypes.c_int64, ctypes.c_int, c_void_p, c_void_p, c_void_p],
    "afx_gmres_update": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int, c_void_p, ctypes.c_int64, c_void_p, ctypes.c_int,
                         c_void_p, c_void_p],
    "afx_profile_enable": [ctypes.c_int64],
    "afx_profile_collect": [c_void_p, c_void_p],
    "afx_profile_kernel_count": [],
    "afx_profile_kernel_name": [ctypes.c_int],
    "afx_flow_fields": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxFlowConfig), ctypes.POINTER(AfxFlowInputs),
        c_void_p, c_void_p, c_void_p, c_void_p,
    ],
    "afx_flow_mass_flux": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxFlowConfig), ctypes.POINTER(AfxFlowInputs),
        c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p,
    ],
    "afx_scalar_residual_noalloc": [c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxScalarConfig),
                                    ctypes.POINTER(AfxScalarInputs), c_void_p, c_void_p],
    "afx_scalar_noalloc_max_bytes": [c_int64_p],
    "afx_scalar_gradient_vjp_geometry": [c_void_p, ctypes.POINTER(AfxGeometry), ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, c_void_p, c_void_p, c_void_p, c_void_p,
                                         ctypes.POINTER(AfxGeometryCotangents), c_void_p],
    "afx_rans_vjp_geometry": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxRansConfig), ctypes.POINTER(AfxRansInputs), c_void_p,
        ctypes.POINTER(AfxRansCotangents), ctypes.POINTER(AfxGeometryCotangents), c_void_p, c_void_p, c_void_p,
        ctypes.c_int64, c_void_p,
    ],
    "afx_scalar_gradient_jvp_geometry": [c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxGeometry), ctypes.c_int,
                                         ctypes.c_int, ctypes.c_double, c_void_p, c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p],
    "afx_scalar_gradient_vjp": [c_void_p, ctypes.POINTER(AfxGeometry), ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "afx_venkat_limiter": [c_void_p, ctypes.POINTER(AfxGeometry), ctypes.c_double, c_void_p, c_void_p, c_void_p,
                           c_void_p, c_void_p, c_void_p, c_void_p],
    "afx_rans_residual_vjp_host": [
        ctypes.c_int, ctypes.c_int64, ctypes.c_int64, c_void_p, c_void_p, ctypes.POINTER(AfxGeometry),
        ctypes.POINTER(AfxRansConfig), ctypes.POINTER(AfxRansInputs), c_void_p, c_void_p,
        ctypes.POINTER(AfxRansCotangents), ctypes.c_int,
    ],
    "afx_scalar_gradient": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.c_int, ctypes.c_int, ctypes.c_double, c_void_p, c_void_p,
        c_void_p, c_void_p, c_void_p,
    ],
    "afx_geometry_vjp_workspace_bytes": [c_void_p, c_int64_p],
    "afx_geometry_vjp": [
        c_void_p, c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxGeometryCotangents), c_void_p, c_void_p,
        ctypes.c_int64, c_void_p,
    ],
    "afx_geometry_vjp_record": [
        c_void_p, c_void_p, ctypes.POINTER(AfxGeometry), c_void_p, ctypes.POINTER(AfxGeometryCotangents), c_void_p,
        c_void_p, ctypes.c_int64, c_void_p,
    ],
    "afx_scalar_workspace_bytes": [c_void_p, ctypes.c_int, c_int64_p],
    "afx_scalar_residual": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxScalarConfig), ctypes.POINTER(AfxScalarInputs),
        c_void_p, c_void_p, ctypes.c_int64, c_void_p,
    ],
    "afx_scalar_jvp": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxScalarConfig), ctypes.POINTER(AfxScalarInputs),
        ctypes.POINTER(AfxScalarInputs), c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p,
    ],
    "afx_flow_vjp_geometry": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxFlowConfig), ctypes.POINTER(AfxFlowInputs),
        c_void_p, ctypes.POINTER(AfxFlowCotangents), ctypes.POINTER(AfxGeometryCotangents), c_void_p, ctypes.c_int64,
        c_void_p,
    ],
    "afx_flow_jvp_geometry": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxFlowConfig),
        ctypes.POINTER(AfxFlowInputs), ctypes.POINTER(AfxFlowInputs), c_void_p, c_void_p, c_void_p, ctypes.c_int64,
        c_void_p,
    ],
    "afx_scalar_jvp_geometry": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxScalarConfig),
        ctypes.POINTER(AfxScalarInputs), ctypes.POINTER(AfxScalarInputs), c_void_p, c_void_p, c_void_p, ctypes.c_int64,
        c_void_p,
    ],
    "afx_scalar_vjp": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxScalarConfig), ctypes.POINTER(AfxScalarInputs),
        c_void_p, ctypes.POINTER(AfxScalarCotangents), c_void_p, ctypes.c_int64, c_void_p,
    ],
    "afx_scalar_record_bytes": [c_void_p, c_int64_p],
    "afx_scalar_residual_record": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxScalarConfig), ctypes.POINTER(AfxScalarInputs),
        c_void_p, c_void_p, ctypes.c_int64, c_void_p,
    ],
    "afx_scalar_vjp_record": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxScalarConfig), ctypes.POINTER(AfxScalarInputs),
        c_void_p, c_void_p, ctypes.c_int64, ctypes.POINTER(AfxScalarCotangents), c_void_p, ctypes.c_int64, c_void_p,
    ],
    "afx_rans_workspace_bytes": [c_void_p, ctypes.c_int, c_int64_p],
    "afx_rans_residual": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxRansConfig), ctypes.POINTER(AfxRansInputs),
        c_void_p, c_void_p, ctypes.c_int64, c_void_p,
    ],
    "afx_rans_jvp": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxRansConfig), ctypes.POINTER(AfxRansInputs),
        ctypes.POINTER(AfxRansInputs), c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p,
    ],
    "afx_rans_jvp_geometry": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxGeometry), c_void_p, c_void_p,
        ctypes.POINTER(AfxRansConfig),
        ctypes.POINTER(AfxRansInputs), ctypes.POINTER(AfxRansInputs), c_void_p, c_void_p, c_void_p, ctypes.c_int64,
        c_void_p,
    ],
    "afx_rans_vjp": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxRansConfig), ctypes.POINTER(AfxRansInputs),
        c_void_p, ctypes.POINTER(AfxRansCotangents), c_void_p, ctypes.c_int64, c_void_p,
    ],
    "afx_rans_residual_vjp": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxRansConfig), ctypes.POINTER(AfxRansInputs),
        c_void_p, c_void_p, ctypes.POINTER(AfxRansCotangents), c_void_p, ctypes.c_int64, c_void_p,
    ],
    "afx_rans_fields": [
        c_void_p, ctypes.POINTER(AfxGeometry), ctypes.POINTER(AfxRansConfig), ctypes.POINTER(AfxRansInputs),
        c_void_p, c_void_p,
    ],
}

EXTRA_SIGNATURES: dict = {}


def lib():
    """The loaded library (raises when it has not been built)."""
    global _LIB
    if _LIB is None:
        path = os.environ.get("AFX_LIBRARY", str(LIB_PATH))
        if not os.path.exists(path):
            raise RuntimeError(
                f"aquaflux_amd HIP library not found at {path}; build it with "
                "`python -m aquaflux_amd.build` (hipcc --offload-arch=gfx950)"
            )
        handle = ctypes.CDLL(path)
        handle.afx_last_error.restype = ctypes.c_char_p
        handle.afx_last_error.argtypes = []
        for name, argtypes in {**_SIGNATURES, **EXTRA_SIGNATURES}.items():
            fn = getattr(handle, name)
            fn.restype = ctypes.c_int
            fn.argtypes = argtypes
        handle.afx_profile_kernel_name.restype = ctypes.c_char_p
        _LIB = handle
    return _LIB


def exported_symbols() -> list[str]:
    """Every entry point this binding declares (the checks compare it with the header)."""
    return ["afx_last_error", *_SIGNATURES, *EXTRA_SIGNATURES]


def check(rc: int, what: str) -> None:
    if rc == 0:
        return
    msg = lib().afx_last_error().decode(errors="replace")
    if rc == 2:
        raise ValueError(f"{what}: {msg}")
    raise RuntimeError(f"{what}: {msg}")


def ptr(t) -> int | None:
    """Device (or host) address of a tensor / ndarray, None for None."""
    if t is None:
        return None
    if isinstance(t, torch.Tensor):
        return t.data_ptr()
    return t.ctypes.data


def stream_of(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def require_cuda(t: torch.Tensor, name: str) -> None:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a CUDA tensor")
    if t.dtype != torch.float64:
        raise ValueError(f"{name} must be float64, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


class KernelTimer:
    """Per-kernel durations from event pairs recorded inside the library on each launch stream."""

    def __init__(self, max_launches: int = 100000):
        self.max_launches = max_launches

    def __enter__(self):
        check(lib().afx_profile_enable(self.max_launches), "afx_profile_enable")
        return self

    def collect(self) -> dict:
        """{kernel name: (total ms, launches)} of everything recorded since enabling."""
        L = lib()
        n = L.afx_profile_kernel_count()
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int64 * n)()
        check(L.afx_profile_collect(ctypes.addressof(ms), ctypes.addressof(cnt)), "afx_profile_collect")
        return {L.afx_profile_kernel_name(i).decode(): (ms[i], cnt[i]) for i in range(n) if cnt[i]}

    def __exit__(self, *exc):
        lib().afx_profile_enable(0)
        return False
