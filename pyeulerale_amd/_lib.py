"""ctypes binding of ``libpea_hip.so`` (the C ABI in ``include/pea.h``).

The library is built in-tree (``pyeulerale_amd/libpea_hip.so``) by ``__graft_entry__.build()``
or ``make -C pyeulerale_amd/csrc``.  There is no fallback: if the library is missing or cannot be
loaded every operation raises, so a GPU run never silently takes another code path.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libpea_hip.so"

_c = ctypes
_P = _c.c_void_p
_D = _c.c_double
_I = _c.c_int


class PeaParams(_c.Structure):
    """Mirror of ``pea_params``."""
    _fields_ = [("aoa_deg", _D), ("k2", _D), ("k4", _D), ("c4", _D)]


class PeaGrid(_c.Structure):
    """Mirror of ``pea_grid`` (one grid of a multi-grid launch)."""
    _fields_ = [("M", _c.c_int), ("N", _c.c_int), ("mach", _D), ("prm", PeaParams),
                ("X", _P), ("V", _P), ("U", _P), ("dU", _P), ("R", _P), ("dR", _P)]


# name -> argument types (restype int unless listed in _RESTYPES)
_SIGS = {
    "pea_last_error": [],
    "pea_version": [],
    "pea_set_tile": [_I, _I, _I],
    "pea_get_tile": [_I, _I],
    "pea_num_tiles": [],
    "pea_num_face_tiles": [],
    "pea_set_face_tile": [_I, _I, _I, _I],
    "pea_get_face_tile": [_I, _I, _I, _P, _P],
    "pea_set_free_stream_state": [_P, _D, _I, _I, _P, _P],
    "pea_compute_odes": [_P, _D, _I, _I, _P, _P, _P, _P, _P],
    "pea_compute_odes_batch": [_P, _D, _I, _I, _I, _P, _P, _P, _P, _P],
    "pea_compute_forces": [_I, _I, _P, _P, _P, _D, _D, _P, _P],
    "pea_check_states": [_I, _I, _P, _P, _P, _P],
    "pea_jvp_odes": [_P, _D, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "pea_jvp_odes_multi": [_P, _D, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "pea_jvp_odes_states_dirs": [_P, _D, _I, _I, _I, _P, _P, _P, _P, _P, _I, _P],
    "pea_vjp_odes_states_multi": [_P, _D, _I, _I, _I, _P, _P, _P, _P, _P, _P],
    "pea_jvp_forces": [_I, _I, _P, _P, _P, _P, _P, _P],
    "pea_vjp_odes": [_P, _D, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "pea_vjp_workspace_bytes": [_I, _I],
    "pea_vjp_odes_noalloc": [_P, _D, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "pea_vjp_forces": [_I, _I, _P, _P, _P, _P, _P, _P],
    "pea_hvp_forces": [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P],
    "pea_hvp_odes": [_P, _D, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "pea_cs_jvp_odes": [_P, _D, _I, _I, _P, _P, _P, _P, _P, _P, _P, _D, _P, _P],
    "pea_cs_jvp_odes_multi": [_P, _D, _I, _I, _I, _P, _P, _P, _P, _D, _P, _P],
    "pea_cs_linearize_odes": [_P, _D, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P],
    "pea_apply_odes_wrt_states_fwd": [_I, _I, _P, _P, _P, _P],
    "pea_apply_odes_wrt_states_rev": [_I, _I, _P, _P, _P, _P],
    "pea_apply_odes_wrt_vertices_fwd": [_I, _I, _P, _P, _P, _P],
    "pea_apply_odes_wrt_vertices_rev": [_I, _I, _P, _P, _P, _P],
    "pea_export_bsr": [_I, _I, _P, _D, _D, _I, _P, _P, _P, _P],
    "pea_multi_workspace_bytes": [_I],
    "pea_compute_odes_multi": [_I, _P, _P, _P],
    "pea_multi_prepare": [_I, _P, _I, _P, _P, _P],
    "pea_multi_launch": [_I, _I, _I, _P, _P],
    "pea_jvp_odes_multigrid": [_I, _P, _P, _P],
    "pea_resolvent_last_error": [],
    "pea_block_multi": [_I, _I, _I, _P, _P, _I, _P, _P, _P],
    "pea_line_factor_bytes": [_I, _I, _I],
    "pea_line_factor": [_I, _I, _I, _P, _P, _I, _P, _P],
    "pea_line_solve": [_I, _I, _I, _P, _P, _I, _P, _P, _P],
    "pea_line_solve_lanes": [_I],
    "pea_jacobi_factor": [_I, _I, _I, _P, _P, _I, _P, _P],
    "pea_jacobi_solve": [_I, _I, _I, _P, _P, _P, _P],
    "pea_ring_band": [_I, _I, _I, _P, _P, _I, _I, _I, _P, _P],
    "pea_ring_couple": [_I, _I, _I, _P, _I, _I, _I, _I, _I, _D, _P, _D, _P, _P],
    "pea_dense_lu_factor": [_I, _I, _P, _P, _P, _P],
    "pea_gmres_rotate": [_I, _I, _I, _P, _P, _P, _P, _P, _P],
    "pea_multidot": [_I, _I, _I, _P, _c.c_longlong, _c.c_longlong, _P, _c.c_longlong, _I, _P,
                     _P, _c.c_longlong, _P],
    "pea_gs_update": [_I, _I, _I, _P, _c.c_longlong, _c.c_longlong, _P, _c.c_longlong, _P,
                      _c.c_longlong, _I, _I, _P, _P, _c.c_longlong, _P],
    "pea_gs_orthogonalize": [_I, _I, _I, _P, _c.c_longlong, _c.c_longlong, _P, _c.c_longlong,
                             _I, _P, _P, _P, _P, _c.c_longlong, _P, _P],
    "pea_dense_lu_solve": [_I, _I, _I, _P, _P, _P, _P],
    "pea_nd_last_error": [],
    "pea_nd_assemble": [_I, _I, _I, _P, _P, _P, _I, _P, _P, _P],
    "pea_nd_blocks": [_I, _P, _I, _I, _I, _P, _D, _D, _I, _I, _I, _P, _P],
    "pea_nd_factor": [_I, _I, _I, _P, _P, _P, _I, _P, _P],
    "pea_nd_gather": [_I, _I, _I, _I, _P, _I, _I, _I, _P, _c.c_longlong, _P, _P],
    "pea_nd_scatter": [_I, _I, _I, _I, _P, _I, _I, _I, _P, _P, _c.c_longlong, _P],
    "pea_nd_accumulate": [_I, _I, _P, _P, _P, _P, _I, _P, _P, _c.c_longlong, _P],
}
_RESTYPES = {"pea_last_error": _c.c_char_p, "pea_vjp_workspace_bytes": _c.c_longlong,
             "pea_multi_workspace_bytes": _c.c_longlong,
             "pea_resolvent_last_error": _c.c_char_p, "pea_line_factor_bytes": _c.c_longlong,
             "pea_nd_last_error": _c.c_char_p}

_RESOLVENT = {"pea_block_multi", "pea_line_factor", "pea_line_solve", "pea_line_solve_lanes", "pea_jacobi_factor",
              "pea_jacobi_solve", "pea_ring_band", "pea_ring_couple", "pea_dense_lu_factor",
              "pea_dense_lu_solve", "pea_gmres_rotate",
              "pea_multidot", "pea_gs_update", "pea_gs_orthogonalize"}

_lib = None


class PeaError(RuntimeError):
    """A libpea_hip call returned a non-zero status."""


def load() -> ctypes.CDLL:
    """Loads and binds the library once; raises if it is missing."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise PeaError(f"{LIB_PATH} is not built: run __graft_entry__.build() or "
                           f"make -C {HERE / 'csrc'}")
        lib = _c.CDLL(str(LIB_PATH))
        for name, args in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, _c.c_int)
        _lib = lib
    return _lib


def call(name: str, *args) -> int:
    """Calls ``name`` and raises :class:`PeaError` on a negative status."""
    rc = getattr(load(), name)(*args)
    if rc < 0:
        src = (load().pea_resolvent_last_error if name in _RESOLVENT else
               load().pea_nd_last_error if name.startswith("pea_nd_") else load().pea_last_error)
        msg = src().decode(errors="replace")
        raise PeaError(f"{name} failed ({rc}): {msg}")
    return rc
