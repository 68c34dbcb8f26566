"""ctypes binding of the C ABI declared in ``include/ablate_fv.h``.

The native library is loaded from the package directory.  There is no fallback: if the library is missing the
import of the product path fails loudly (build it with ``python -m ablate_amd.build``).
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int32, c_int64, c_uint64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libablate_fv.so"

# ---- enums (values are part of the ABI) ----
FORM_FP64, FORM_MIXED = 0, 1
FLUX = {"none": 0, "ausm": 1, "average": 2, "off": 3, "rusanov": 4, "ausmpup": 5}
RECON = {"upwind": 0, "none": 0, "leastsquares": 1}
LIMITER = {
    "none": 0, "zero": 1, "minmod": 2, "vanleer": 3, "vanalbada": 4, "superbee": 5, "mc": 6, "sin": 7,
    "barthjespersen": 8,
}
BC_NONE, BC_INFLOW, BC_OUTFLOW, BC_GHOST = 0, 1, 2, 3
ORDER = {"natural": 0, "brick": 1, "rcm": 2, "morton": 3, "column": 4}

_dp = POINTER(c_double)
_ip = POINTER(c_int32)


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("dim", c_int32),
        ("n_cells", c_int64),
        ("n_interior", c_int64),
        ("n_faces", c_int64),
        ("face_cells", _ip),
        ("face_normal", _dp),
        ("face_centroid", _dp),
        ("face_shift", _dp),
        ("cell_centroid", _dp),
        ("cell_volume", _dp),
        ("face_bc", _ip),
    ]


class PhysicsDesc(ctypes.Structure):
    _fields_ = [
        ("gamma", c_double),
        ("Rgas", c_double),
        ("n_species", c_int32),
        ("flux_calculator", c_int32),
        ("reconstruction", c_int32),
        ("limiter", c_int32),
        ("max_lim_grad", c_double),
        ("k", c_double),
        ("mu", c_double),
        ("diff", c_double),
        ("cfl", c_double),
        ("ausmpup_minf", c_double),
        ("compact_species", c_int32),
        ("transport", c_int32),
        ("transport_disabled", c_int32),
        ("conduction_stability_factor", c_double),
        ("viscous_stability_factor", c_double),
    ]


class PgsDesc(ctypes.Structure):
    _fields_ = [
        ("alpha_init", c_double),
        ("domain_length", c_double),
        ("max_alpha_allowed", c_double),
        ("max_delta_pressure_fac", c_double),
    ]


class BcDesc(ctypes.Structure):
    _fields_ = [
        ("kind", c_int32),
        ("rho", c_double),
        ("vel", c_double * 3),
        ("T", c_double),
        ("Yi", _dp),
        ("enforce_at_face", c_int32),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("dim", c_int32),
        ("n_comp", c_int32),
        ("max_faces_per_cell", c_int32),
        ("form", c_int32),
        ("n_cells", c_int64),
        ("n_interior", c_int64),
        ("n_faces", c_int64),
        ("device_bytes", c_int64),
        ("alg_bytes_per_cell", c_double),
    ]


# Every exported symbol with (restype, argtypes); the tests check this list against include/ablate_fv.h.
SIGNATURES = {
    "abl_fv_mesh_box": (c_int32, [c_int32, _ip, _dp, _dp, _ip, c_int32, POINTER(c_void_p)]),
    "abl_fv_mesh_tet_box": (c_int32, [c_int32, c_double, c_uint64, c_int32, POINTER(c_void_p)]),
    "abl_fv_mesh_describe": (c_int32, [c_void_p, POINTER(MeshDesc)]),
    "abl_fv_mesh_free": (None, [c_void_p]),
    "abl_fv_create": (c_int32, [POINTER(MeshDesc), POINTER(PhysicsDesc), c_int32, POINTER(c_void_p)]),
    "abl_fv_destroy": (c_int32, [c_void_p]),
    "abl_fv_last_error": (c_char_p, [c_void_p]),
    "abl_fv_get_info": (c_int32, [c_void_p, POINTER(Info)]),
    "abl_fv_set_boundary": (c_int32, [c_void_p, c_int32, POINTER(BcDesc)]),
    "abl_fv_rhs": (c_int32, [c_void_p, c_double, c_void_p, c_void_p, c_int32, c_void_p]),
    "abl_fv_rhs_vjp": (c_int32, [c_void_p, c_double, c_void_p, c_void_p, c_void_p, _dp, c_void_p]),
    "abl_fv_rhs_vjp_noalloc": (c_int32, [c_void_p, c_double, c_void_p, c_void_p, c_void_p, _dp, c_void_p]),
    "abl_fv_face_interpolate": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "abl_fv_pack_aos": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "abl_fv_unpack_aos": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "abl_fv_check_state": (c_int32, [c_void_p, c_void_p, POINTER(c_int64), c_void_p]),
    "abl_fv_cfl_dt": (c_int32, [c_void_p, c_void_p, _dp, c_void_p]),
    "abl_fv_time_steps": (c_int32, [c_void_p, c_void_p, _dp, POINTER(c_int32), c_void_p]),
    "abl_fv_physics_dt": (c_int32, [c_void_p, c_void_p, _dp, c_void_p]),
    "abl_fv_axpy": (c_int32, [c_void_p, c_double, c_void_p, c_void_p, c_void_p]),
    "abl_fv_ts_euler": (c_int32, [c_void_p, c_void_p, c_void_p, c_double, c_double, c_double, c_double, c_int64, _dp,
                                  POINTER(c_int64), c_void_p]),
    "abl_fv_mesh_gmsh": (c_int32, [c_char_p, c_int32, POINTER(c_void_p)]),
    "abl_fv_mesh_dmplex": (c_int32, [c_int32, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_int32,
                                     POINTER(c_void_p)]),
    "abl_fv_mesh_source_cells": (c_int32, [c_void_p, c_void_p]),
    "abl_fv_set_pgs": (c_int32, [c_void_p, POINTER(PgsDesc)]),
    "abl_fv_pgs_update": (c_int32, [c_void_p, c_void_p, c_double, _dp, _dp, c_void_p]),
    "abl_fv_get_pgs_alpha": (c_int32, [c_void_p, _dp]),
    "abl_fv_set_pgs_alpha": (c_int32, [c_void_p, c_double]),
    "abl_fv_set_profiling": (c_int32, [c_void_p, c_int32]),
    "abl_fv_set_strategy": (c_int32, [c_void_p, c_int32]),
    "abl_fv_get_strategy": (c_int32, [c_void_p, _ip]),
    "abl_fv_set_tile_size": (c_int32, [c_void_p, c_int32]),
    "abl_fv_get_phase_times": (c_int32, [c_void_p, _dp, POINTER(c_int64), c_int32]),
    "abl_hbm_probe": (c_int32, [c_int32, c_int64, c_int32, _dp, _dp]),
}
PHASES = ["ghost", "gradients", "species_gradients", "flux", "vjp", "gather"]
STRATEGY = {"auto": 0, "recompute": 1, "face_flux": 2, "tile": 3, "fused": 4, "march": 5}

_lib = None


def lib() -> ctypes.CDLL:
    """Loads the native library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"native library {LIB_PATH} is missing: build it with `python -m ablate_amd.build` "
                "(there is no CPU fallback)"
            )
        try:
            # One HIP runtime per process: torch ships its own libamdhip64 (same soname). Loading it first makes
            # the dynamic linker bind this library to that runtime instead of starting a second one.
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class NativeError(RuntimeError):
    pass


def check(status: int, handle=None, what: str = "") -> None:
    if status != 0:
        msg = lib().abl_fv_last_error(handle)
        text = msg.decode() if msg else "unknown error"
        if not text.startswith(what + ":"):  # the library's messages mostly name their entry point already
            text = f"{what}: {text}"
        raise NativeError(f"{text} (status {status})")
