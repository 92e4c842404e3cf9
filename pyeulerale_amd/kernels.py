"""Tensor-level entry points over the C ABI of ``libpea_hip.so``.

All tensors live on the GPU in the library's canonical storage (radial index fastest):

* cell fields  -- shape ``(4, N, M)`` contiguous (float64 states/odes, complex128 directions);
* vertex fields -- shape ``(2, N+1, M+1)`` contiguous;
* forces -- shape ``(2, N)`` contiguous.

``storage.transpose(-1, -2)`` is the reference's ``(4, M, N)`` / ``(2, M+1, N+1)`` view
(core.py:94-100); :func:`cells_in` / :func:`verts_in` convert any tensor of reference shape into
canonical storage (no copy when it already is a view of canonical storage).

Every function checks shapes, dtypes, devices and contiguity on the host before launching, and
launches on the current torch stream.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib

F64 = torch.float64
C128 = torch.complex128


@dataclass(frozen=True)
class Params:
    """Scheme constants passed by value to every call (no module globals)."""
    aoa_deg: float = 1.25
    k2: float = 1.0
    k4: float = 1.0 / 32
    c4: float = 2.0

    def c(self) -> _lib.PeaParams:
        return _lib.PeaParams(self.aoa_deg, self.k2, self.k4, self.c4)


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor | None) -> ctypes.c_void_p | None:
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _need(t: torch.Tensor, shape: tuple, dtype: torch.dtype, name: str) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch.Tensor, got {type(t).__name__}")
    if tuple(t.shape) != tuple(shape):
        raise RuntimeError(f"{name}: incorrect shape {tuple(t.shape)}, expected {tuple(shape)}")
    if t.dtype != dtype:
        raise RuntimeError(f"{name}: incorrect dtype {t.dtype}, expected {dtype}")
    if t.device.type != "cuda":
        raise RuntimeError(f"{name}: must be a GPU tensor (got {t.device})")
    if not t.is_contiguous():
        raise RuntimeError(f"{name}: must be contiguous canonical storage")
    return t


def _opt(t, shape, dtype, name):
    return None if t is None else _need(t, shape, dtype, name)


def cells_in(t: torch.Tensor, dtype: torch.dtype, device) -> torch.Tensor:
    """Reference-shaped cell field ``(4|k, M, N)`` -> canonical ``(4|k, N, M)`` storage."""
    t = torch.as_tensor(t)
    return t.to(device=device, dtype=dtype).transpose(-1, -2).contiguous()


def verts_in(t: torch.Tensor, dtype: torch.dtype, device) -> torch.Tensor:
    """Reference-shaped vertex field ``(2, M+1, N+1)`` -> canonical ``(2, N+1, M+1)``."""
    return cells_in(t, dtype, device)


class Grid:
    """Dimensions of one O-grid: M radial and N angular cells."""

    def __init__(self, M: int, N: int):
        if M < 4 or N < 5:
            raise RuntimeError(f"need at least 4 radial and 5 angular cells (got {M}, {N})")
        self.M, self.N = int(M), int(N)

    @property
    def C(self):
        return self.M * self.N

    @property
    def cell_shape(self):
        return (4, self.N, self.M)

    @property
    def vert_shape(self):
        return (2, self.N + 1, self.M + 1)

    @property
    def force_shape(self):
        return (2, self.N)

    def check_primal(self, X, V, U):
        _need(X, self.vert_shape, F64, "vertices")
        _need(V, self.vert_shape, F64, "velocities")
        _need(U, self.cell_shape, F64, "states")


# ------------------------------------------------------------------------------------ primal
def set_free_stream_state(p: Params, mach: float, g: Grid, U: torch.Tensor) -> torch.Tensor:
    _need(U, g.cell_shape, F64, "states")
    _lib.call("pea_set_free_stream_state", ctypes.byref(p.c()), float(mach), g.M, g.N, _ptr(U),
              _stream())
    return U


def compute_odes(p: Params, mach: float, g: Grid, X, V, U, R=None) -> torch.Tensor:
    g.check_primal(X, V, U)
    R = torch.empty(g.cell_shape, dtype=F64, device=U.device) if R is None else \
        _need(R, g.cell_shape, F64, "odes")
    _lib.call("pea_compute_odes", ctypes.byref(p.c()), float(mach), g.M, g.N, _ptr(X), _ptr(V),
              _ptr(U), _ptr(R), _stream())
    return R


def compute_odes_batch(p: Params, mach: float, g: Grid, X, V, U, R=None) -> torch.Tensor:
    """Residuals of a batch of states ``U`` of shape ``(B, 4, N, M)`` on one grid."""
    _need(X, g.vert_shape, F64, "vertices")
    _need(V, g.vert_shape, F64, "velocities")
    B = U.shape[0]
    _need(U, (B, *g.cell_shape), F64, "states")
    R = torch.empty_like(U) if R is None else _need(R, (B, *g.cell_shape), F64, "odes")
    _lib.call("pea_compute_odes_batch", ctypes.byref(p.c()), float(mach), g.M, g.N, B, _ptr(X),
              _ptr(V), _ptr(U), _ptr(R), _stream())
    return R


class MultiGrid:
    """A prepared multi-grid launch over many grids of uneven size.

    ``items`` holds (params, mach, X, V, U) per grid in storage layout, plus dU for
    ``tangent=True`` (the states-direction JVP) or the seed w for ``adjoint=True`` (the states
    adjoint, pea_vjp_odes to the states).  The descriptor table is uploaded once; :meth:`run` is
    then a single launch that allocates and copies nothing (it may be captured in a CUDA graph)
    and returns the outputs: the list of R, of (R, dR), or of (dR/dU)^T w.  The input tensors
    must stay alive and in place while the object is used."""

    def __init__(self, items, tangent: bool = False, adjoint: bool = False):
        if tangent and adjoint:
            raise ValueError("tangent or adjoint, not both")
        self.tangent = bool(tangent)
        self.adjoint = bool(adjoint)
        self.mode = 2 if adjoint else int(self.tangent)
        self.items = list(items)
        G = len(self.items)
        if G < 1:
            raise RuntimeError("no grids")
        grids = (_lib.PeaGrid * G)()
        self.outputs = []
        for i, it in enumerate(self.items):
            p, mach, X, V, U = it[:5]
            g = Grid(U.shape[2], U.shape[1])
            g.check_primal(X, V, U)
            R = None if self.adjoint else torch.empty(g.cell_shape, dtype=F64, device=U.device)
            dU = dR = None
            if self.tangent or self.adjoint:
                dU = _need(it[5], g.cell_shape, C128, "d_odes" if self.adjoint else "d_states")
                dR = torch.empty(g.cell_shape, dtype=C128, device=U.device)
            grids[i] = _lib.PeaGrid(g.M, g.N, float(mach), p.c(), X.data_ptr(), V.data_ptr(),
                                    U.data_ptr(), dU.data_ptr() if dU is not None else None,
                                    R.data_ptr() if R is not None else None,
                                    dR.data_ptr() if dR is not None else None)
            self.outputs.append(dR if self.adjoint else (R, dR) if self.tangent else R)
        nbytes = int(_lib.load().pea_multi_workspace_bytes(G))
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device=self.items[0][4].device)
        total = ctypes.c_int(0)
        _lib.call("pea_multi_prepare", G, grids, self.mode, _ptr(self.ws),
                  ctypes.byref(total), _stream())
        self.G, self.total_tiles = G, total.value

    def run(self):
        _lib.call("pea_multi_launch", self.G, self.total_tiles, self.mode, _ptr(self.ws),
                  _stream())
        return self.outputs


def compute_odes_multi(items):
    """Residuals of many grids of uneven size in one launch.  ``items`` is a list of
    (params, mach, X, V, U) in storage layout; returns the list of R."""
    return MultiGrid(items).run()


def jvp_odes_multigrid(items):
    """States-direction tangents of many grids in one launch: items (params, mach, X, V, U, dU);
    returns the list of (R, dR)."""
    return MultiGrid(items, tangent=True).run()


def vjp_odes_multigrid(items):
    """States adjoints of many grids in one launch: items (params, mach, X, V, U, w); returns the
    list of (dR/dU)^T w.  Every grid needs N >= 18 angular cells."""
    return MultiGrid(items, adjoint=True).run()


def check_states(g: Grid, U, R=None, status=None) -> torch.Tensor:
    """Device int32 tensor (2,): cells with a non-finite or non-physical state (rho <= 0 or
    p <= 0), and cells with a non-finite residual R (when given).  Asynchronous."""
    _need(U, g.cell_shape, F64, "states")
    _opt(R, g.cell_shape, F64, "odes")
    status = torch.empty(2, dtype=torch.int32, device=U.device) if status is None else status
    _lib.call("pea_check_states", g.M, g.N, _ptr(U), _ptr(R), _ptr(status), _stream())
    return status


def compute_forces(g: Grid, X, U, F=None, center=(0.0, 0.0), coef=None):
    _need(X, g.vert_shape, F64, "vertices")
    _need(U, g.cell_shape, F64, "states")
    F = torch.empty(g.force_shape, dtype=F64, device=U.device) if F is None else \
        _need(F, g.force_shape, F64, "forces")
    _opt(coef, (3,), F64, "coef")
    _lib.call("pea_compute_forces", g.M, g.N, _ptr(X), _ptr(U), _ptr(F), float(center[0]),
              float(center[1]), _ptr(coef), _stream())
    return F


# ------------------------------------------------------------------------------------ tangents
def _mach_dir(dmach, device):
    if dmach is None:
        return None
    t = torch.as_tensor(dmach, dtype=C128).reshape(-1).to(device)
    return t.contiguous()


def jvp_odes(p: Params, mach: float, g: Grid, X, V, U, dmach=None, dX=None, dV=None, dU=None,
             dR=None, R=None) -> torch.Tensor:
    """Hand-written forward mode: dR = R_Ma dMa + R_X dX + R_V dV + R_U dU (complex)."""
    g.check_primal(X, V, U)
    _opt(dX, g.vert_shape, C128, "d_vertices")
    _opt(dV, g.vert_shape, C128, "d_velocities")
    _opt(dU, g.cell_shape, C128, "d_states")
    dm = _mach_dir(dmach, U.device)
    if dm is not None and dm.numel() != 1:
        raise RuntimeError("d_mach must hold one value")
    dR = torch.empty(g.cell_shape, dtype=C128, device=U.device) if dR is None else \
        _need(dR, g.cell_shape, C128, "d_odes")
    _opt(R, g.cell_shape, F64, "odes")
    _lib.call("pea_jvp_odes", ctypes.byref(p.c()), float(mach), g.M, g.N, _ptr(X), _ptr(V),
              _ptr(U), _ptr(dm), _ptr(dX), _ptr(dV), _ptr(dU), _ptr(dR), _ptr(R), _stream())
    return dR


def jvp_odes_multi(p: Params, mach: float, g: Grid, X, V, U, dU=None, dX=None, dV=None,
                   dmach=None, dR=None) -> torch.Tensor:
    """k directions at once; direction tensors have a leading axis k.  States-only directions
    (dU alone) take one batched launch of the face kernel and any k; otherwise k in 1, 2, 4, 8."""
    g.check_primal(X, V, U)
    k = next(t.shape[0] for t in (dU, dX, dV) if t is not None)
    _opt(dU, (k, *g.cell_shape), C128, "d_states")
    _opt(dX, (k, *g.vert_shape), C128, "d_vertices")
    _opt(dV, (k, *g.vert_shape), C128, "d_velocities")
    dm = _mach_dir(dmach, U.device)
    if dm is not None and dm.numel() != k:
        raise RuntimeError("d_mach must hold k values")
    dR = torch.empty((k, *g.cell_shape), dtype=C128, device=U.device) if dR is None else \
        _need(dR, (k, *g.cell_shape), C128, "d_odes")
    _lib.call("pea_jvp_odes_multi", ctypes.byref(p.c()), float(mach), g.M, g.N, k, _ptr(X),
              _ptr(V), _ptr(U), _ptr(dm), _ptr(dX), _ptr(dV), _ptr(dU), _ptr(dR), _stream())
    return dR


def jvp_odes_dirs(p: Params, mach: float, g: Grid, X, V, U, dU, dR=None, per: int = 0
                  ) -> torch.Tensor:
    """(dR/dU) dU[j] for k complex directions (dU of shape (k, 4, N, M)) at one state: each
    workgroup evaluates its tile's primal part once and applies the linearised face fluxes to
    `per` of the directions (0: chosen from the grid size).  The same products as
    jvp_odes_multi(..., dU=dU), to round-off."""
    g.check_primal(X, V, U)
    k = dU.shape[0]
    _need(dU, (k, *g.cell_shape), C128, "d_states")
    dR = torch.empty((k, *g.cell_shape), dtype=C128, device=U.device) if dR is None else \
        _need(dR, (k, *g.cell_shape), C128, "d_odes")
    _lib.call("pea_jvp_odes_states_dirs", ctypes.byref(p.c()), float(mach), g.M, g.N, k,
              _ptr(X), _ptr(V), _ptr(U), _ptr(dU), _ptr(dR), int(per), _stream())
    return dR


def vjp_odes_states_multi(p: Params, mach: float, g: Grid, X, V, U, w, gU=None) -> torch.Tensor:
    """(dR/dU)^T w for k seeds at once (w of shape (k, 4, N, M), complex): one batched launch of
    the face adjoint."""
    g.check_primal(X, V, U)
    k = w.shape[0]
    _need(w, (k, *g.cell_shape), C128, "seeds")
    gU = torch.empty_like(w) if gU is None else _need(gU, (k, *g.cell_shape), C128, "d_states")
    _lib.call("pea_vjp_odes_states_multi", ctypes.byref(p.c()), float(mach), g.M, g.N, k,
              _ptr(X), _ptr(V), _ptr(U), _ptr(w), _ptr(gU), _stream())
    return gU


def jvp_forces(g: Grid, X, U, dX=None, dU=None, dF=None) -> torch.Tensor:
    _need(X, g.vert_shape, F64, "vertices")
    _need(U, g.cell_shape, F64, "states")
    _opt(dX, g.vert_shape, C128, "d_vertices")
    _opt(dU, g.cell_shape, C128, "d_states")
    dF = torch.empty(g.force_shape, dtype=C128, device=U.device) if dF is None else \
        _need(dF, g.force_shape, C128, "d_forces")
    _lib.call("pea_jvp_forces", g.M, g.N, _ptr(X), _ptr(U), _ptr(dX), _ptr(dU), _ptr(dF),
              _stream())
    return dF


# ------------------------------------------------------------------------------------ adjoints
def _workspace(g: Grid, device) -> torch.Tensor:
    nbytes = int(_lib.load().pea_vjp_workspace_bytes(g.M, g.N))
    return torch.empty((max(nbytes, 16) + 7) // 8, dtype=F64, device=device)


_TUNED: dict = {}


def autotune_vjp(p: Params, mach: float, g: Grid, X, V, U, reps: int = 5) -> int:
    """Times every tile shape of the cell-based adjoint (the products to vertices, velocities
    and Mach, and to the states of grids narrower than the face-based kernel's tile) that fits
    this grid and registers the fastest (cached per grid size).  Returns the chosen tile
    index."""
    key = (g.M, g.N)
    if key in _TUNED:
        return _TUNED[key]
    lib = _lib.load()
    gen = torch.Generator(device=U.device).manual_seed(0)
    w = torch.randn(g.cell_shape, dtype=C128, device=U.device, generator=gen)
    out = {"states": torch.empty(g.cell_shape, dtype=C128, device=U.device),
           "vertices": torch.empty(g.vert_shape, dtype=C128, device=U.device),
           "velocities": torch.empty(g.vert_shape, dtype=C128, device=U.device),
           "mach": torch.empty(1, dtype=C128, device=U.device)}
    times = {}
    for t in range(lib.pea_num_tiles()):
        if lib.pea_set_tile(g.M, g.N, t) != 0:
            continue
        vjp_odes(p, mach, g, X, V, U, w, want=tuple(out), out=out)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            vjp_odes(p, mach, g, X, V, U, w, want=tuple(out), out=out)
        b.record()
        b.synchronize()
        times[t] = a.elapsed_time(b) / reps
    best = min(times, key=times.get)
    _lib.call("pea_set_tile", g.M, g.N, best)
    _TUNED[key] = best
    return best


# ------------------------------------------------------------------------ face-kernel autotuner
FACE_KINDS = {"residual": 0, "jvp": 1, "vjp": 2}
_FACE_TUNED: dict = {}       # (kind, M, N) -> tile, also (kind + 3, 0, 0) for multi-grid launches


def face_tile(kind: int, M: int = 0, N: int = 0) -> tuple[int, int, int]:
    """(tile index, TM, TN) the face kernel of ``kind`` uses for an M x N grid."""
    tm, tn = ctypes.c_int(0), ctypes.c_int(0)
    t = _lib.load().pea_get_face_tile(kind, M, N, ctypes.byref(tm), ctypes.byref(tn))
    return t, tm.value, tn.value


def set_face_tile(kind: int, M: int, N: int, tile: int) -> None:
    _lib.call("pea_set_face_tile", kind, M, N, tile)
    if tile < 0:
        _FACE_TUNED.pop((kind, M, N), None)
    else:
        _FACE_TUNED[(kind, M, N)] = tile


def _time_events(fn, reps: int, warmup: int = 1) -> float:
    for _ in range(max(1, warmup)):
        fn()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    b.synchronize()
    return a.elapsed_time(b) / reps


def autotune_face(p: Params, mach: float, g: Grid, X, V, U, kinds=("residual", "jvp", "vjp"),
                  reps: int = 10, batch: int = 1, warmup: int = 1) -> dict:
    """Times every tile shape of the face kernels (residual, states tangent, states adjoint) on
    this grid, ``batch`` states / directions / seeds per launch, and registers the fastest for
    grids of this size (``pea_set_face_tile``).  Each shape runs ``warmup`` untimed launches
    before its ``reps`` timed ones (the GPU's clock rises over the first tens of milliseconds of
    load).  The choices are cached per (kind, M, N) for the process; :func:`save_tuning` /
    :func:`load_tuning` keep them across processes.  Returns {kind: (tile, {tile: ms})}."""
    lib = _lib.load()
    gen = torch.Generator(device=U.device).manual_seed(0)
    Ub = U.expand(batch, *U.shape).contiguous()
    Rb = torch.empty_like(Ub)
    d = torch.randn((batch, *g.cell_shape), dtype=C128, device=U.device, generator=gen)
    out = torch.empty_like(d)
    runs = {
        "residual": lambda: compute_odes_batch(p, mach, g, X, V, Ub, Rb),
        "jvp": lambda: jvp_odes_multi(p, mach, g, X, V, U, dU=d, dR=out),
        "vjp": lambda: vjp_odes_states_multi(p, mach, g, X, V, U, d, gU=out),
    }
    result = {}
    for name in kinds:
        kind = FACE_KINDS[name]
        times = {}
        for t in range(lib.pea_num_face_tiles()):
            if lib.pea_set_face_tile(kind, g.M, g.N, t) != 0:
                continue
            times[t] = _time_events(runs[name], reps, warmup)
        best = min(times, key=times.get)
        set_face_tile(kind, g.M, g.N, best)
        result[name] = (best, times)
    return result


def autotune_multigrid(items, reps: int = 5) -> dict:
    """Times every tile shape of the multi-grid residual, states tangent and states adjoint
    launches over ``items`` (params, mach, X, V, U, dU) and registers the fastest process-wide
    (pea_set_face_tile kinds 3-5).  Returns {mode: (tile, {tile: ms})}."""
    lib = _lib.load()
    base = [it[:5] for it in items]
    result = {}
    for mode, name in ((0, "residual"), (1, "jvp"), (2, "vjp")):
        times = {}
        for t in range(lib.pea_num_face_tiles()):
            _lib.call("pea_set_face_tile", 3 + mode, 0, 0, t)
            try:
                mg = MultiGrid(items if mode else base, tangent=mode == 1, adjoint=mode == 2)
            except _lib.PeaError:          # an adjoint tile too wide for one of the grids
                continue
            times[t] = _time_events(mg.run, reps)
        best = min(times, key=times.get)
        set_face_tile(3 + mode, 0, 0, best)
        result[name] = (best, times)
    return result


def save_tuning(path) -> None:
    """Writes the registered face-tile choices as JSON."""
    import json
    rows = [{"kind": k, "M": M, "N": N, "tile": t} for (k, M, N), t in sorted(_FACE_TUNED.items())]
    with open(path, "w") as fh:
        json.dump({"face_tiles": rows}, fh, indent=1)


def load_tuning(path) -> int:
    """Registers the choices of a JSON file written by :func:`save_tuning`; returns how many."""
    import json
    with open(path) as fh:
        rows = json.load(fh)["face_tiles"]
    for r in rows:
        set_face_tile(int(r["kind"]), int(r["M"]), int(r["N"]), int(r["tile"]))
    return len(rows)


def vjp_odes(p: Params, mach: float, g: Grid, X, V, U, w, want=("states",), out=None):
    """Transposed products with complex seed w (no conjugation).  Returns a dict with the
    requested subset of 'states', 'vertices', 'velocities', 'mach'."""
    g.check_primal(X, V, U)
    _need(w, g.cell_shape, C128, "d_odes")
    out = dict(out or {})
    shapes = {"states": g.cell_shape, "vertices": g.vert_shape, "velocities": g.vert_shape,
              "mach": (1,)}
    for k in want:
        if k not in out:
            out[k] = torch.empty(shapes[k], dtype=C128, device=U.device)
        _need(out[k], shapes[k], C128, k)
    ws = _workspace(g, U.device) if "mach" in out else None
    _lib.call("pea_vjp_odes", ctypes.byref(p.c()), float(mach), g.M, g.N, _ptr(X), _ptr(V),
              _ptr(U), _ptr(w), _ptr(out.get("states")), _ptr(out.get("vertices")),
              _ptr(out.get("velocities")), _ptr(out.get("mach")), _ptr(ws), _stream())
    return out


def hvp_odes(p: Params, mach: float, g: Grid, X, V, U, w, dmach=None, dX=None, dV=None, dU=None,
             want=("states",)):
    """Tangent-over-adjoint: directional derivative of the adjoint products along
    (dMa, dX, dV, dU); bilinear in (w, direction)."""
    g.check_primal(X, V, U)
    _need(w, g.cell_shape, C128, "d_odes")
    _opt(dX, g.vert_shape, C128, "d_vertices")
    _opt(dV, g.vert_shape, C128, "d_velocities")
    _opt(dU, g.cell_shape, C128, "d_states")
    dm = _mach_dir(dmach, U.device)
    shapes = {"states": g.cell_shape, "vertices": g.vert_shape, "velocities": g.vert_shape,
              "mach": (1,)}
    out = {k: torch.empty(shapes[k], dtype=C128, device=U.device) for k in want}
    ws = _workspace(g, U.device) if "mach" in out else None
    _lib.call("pea_hvp_odes", ctypes.byref(p.c()), float(mach), g.M, g.N, _ptr(X), _ptr(V),
              _ptr(U), _ptr(w), _ptr(dm), _ptr(dX), _ptr(dV), _ptr(dU),
              _ptr(out.get("states")), _ptr(out.get("vertices")), _ptr(out.get("velocities")),
              _ptr(out.get("mach")), _ptr(ws), _stream())
    return out


def vjp_forces(g: Grid, X, U, wF, want=("states", "vertices")):
    _need(X, g.vert_shape, F64, "vertices")
    _need(U, g.cell_shape, F64, "states")
    _need(wF, g.force_shape, C128, "d_forces")
    out = {}
    if "states" in want:
        out["states"] = torch.empty(g.cell_shape, dtype=C128, device=U.device)
    if "vertices" in want:
        out["vertices"] = torch.empty(g.vert_shape, dtype=C128, device=U.device)
    _lib.call("pea_vjp_forces", g.M, g.N, _ptr(X), _ptr(U), _ptr(wF), _ptr(out.get("states")),
              _ptr(out.get("vertices")), _stream())
    return out


def hvp_forces(g: Grid, X, U, wF, dX=None, dU=None, want=("states", "vertices")):
    """Tangent-over-adjoint of the force adjoints along (dX, dU)."""
    _need(X, g.vert_shape, F64, "vertices")
    _need(U, g.cell_shape, F64, "states")
    _need(wF, g.force_shape, C128, "d_forces")
    _opt(dX, g.vert_shape, C128, "d_vertices")
    _opt(dU, g.cell_shape, C128, "d_states")
    out = {}
    if "states" in want:
        out["states"] = torch.empty(g.cell_shape, dtype=C128, device=U.device)
    if "vertices" in want:
        out["vertices"] = torch.empty(g.vert_shape, dtype=C128, device=U.device)
    _lib.call("pea_hvp_forces", g.M, g.N, _ptr(X), _ptr(U), _ptr(wF), _ptr(dX), _ptr(dU),
              _ptr(out.get("states")), _ptr(out.get("vertices")), _stream())
    return out


# ------------------------------------------------------------------------------------ complex step
def cs_jvp_odes(p: Params, mach: float, g: Grid, X, V, U, dmach=None, dX=None, dV=None, dU=None,
                h: float = 0.0, dR=None) -> torch.Tensor:
    """Matrix-free directional complex step (two complex residuals per complex direction)."""
    g.check_primal(X, V, U)
    _opt(dX, g.vert_shape, C128, "d_vertices")
    _opt(dV, g.vert_shape, C128, "d_velocities")
    _opt(dU, g.cell_shape, C128, "d_states")
    dm = _mach_dir(dmach, U.device)
    dR = torch.empty(g.cell_shape, dtype=C128, device=U.device) if dR is None else \
        _need(dR, g.cell_shape, C128, "d_odes")
    _lib.call("pea_cs_jvp_odes", ctypes.byref(p.c()), float(mach), g.M, g.N, _ptr(X), _ptr(V),
              _ptr(U), _ptr(dm), _ptr(dX), _ptr(dV), _ptr(dU), float(h), _ptr(dR), _stream())
    return dR


def cs_jvp_odes_multi(p: Params, mach: float, g: Grid, X, V, U, dU, h: float = 0.0,
                      dR=None) -> torch.Tensor:
    """Directional complex step along k states directions at once (dU of shape (k, 4, N, M))."""
    g.check_primal(X, V, U)
    k = dU.shape[0]
    _need(dU, (k, *g.cell_shape), C128, "d_states")
    dR = torch.empty_like(dU) if dR is None else _need(dR, (k, *g.cell_shape), C128, "d_odes")
    _lib.call("pea_cs_jvp_odes_multi", ctypes.byref(p.c()), float(mach), g.M, g.N, k, _ptr(X),
              _ptr(V), _ptr(U), _ptr(dU), float(h), _ptr(dR), _stream())
    return dR


def cs_linearize_odes(p: Params, mach: float, g: Grid, X, V, U, JU=None, JX=None, JV=None,
                      JMa=None, want=("states", "vertices", "velocities", "mach")):
    """Complex-step Jacobian blocks (the reference's linearize).  Storage:
    JU (9, 4, 4, N, M), JX/JV (4, 4, 2, N, M), JMa (4, N, M) -- float64."""
    g.check_primal(X, V, U)
    dev = U.device
    if "states" in want and JU is None:
        JU = torch.empty((9, 4, 4, g.N, g.M), dtype=F64, device=dev)
    if "vertices" in want and JX is None:
        JX = torch.empty((4, 4, 2, g.N, g.M), dtype=F64, device=dev)
    if "velocities" in want and JV is None:
        JV = torch.empty((4, 4, 2, g.N, g.M), dtype=F64, device=dev)
    if "mach" in want and JMa is None:
        JMa = torch.empty((4, g.N, g.M), dtype=F64, device=dev)
    _opt(JU, (9, 4, 4, g.N, g.M), F64, "JU")
    _opt(JX, (4, 4, 2, g.N, g.M), F64, "JX")
    _opt(JV, (4, 4, 2, g.N, g.M), F64, "JV")
    _opt(JMa, (4, g.N, g.M), F64, "JMa")
    _lib.call("pea_cs_linearize_odes", ctypes.byref(p.c()), float(mach), g.M, g.N, _ptr(X),
              _ptr(V), _ptr(U), _ptr(JU), _ptr(JX), _ptr(JV), _ptr(JMa), _stream())
    return JU, JX, JV, JMa


def apply_states_fwd(g: Grid, JU, dU, out=None):
    _need(JU, (9, 4, 4, g.N, g.M), F64, "JU")
    _need(dU, g.cell_shape, C128, "d_states")
    out = torch.empty(g.cell_shape, dtype=C128, device=dU.device) if out is None else \
        _need(out, g.cell_shape, C128, "d_odes")
    _lib.call("pea_apply_odes_wrt_states_fwd", g.M, g.N, _ptr(JU), _ptr(dU), _ptr(out), _stream())
    return out


def apply_states_rev(g: Grid, JU, w, out=None):
    _need(JU, (9, 4, 4, g.N, g.M), F64, "JU")
    _need(w, g.cell_shape, C128, "d_odes")
    out = torch.empty(g.cell_shape, dtype=C128, device=w.device) if out is None else \
        _need(out, g.cell_shape, C128, "d_states")
    _lib.call("pea_apply_odes_wrt_states_rev", g.M, g.N, _ptr(JU), _ptr(w), _ptr(out), _stream())
    return out


def apply_vertices_fwd(g: Grid, JX, dX, out=None):
    _need(JX, (4, 4, 2, g.N, g.M), F64, "JX")
    _need(dX, g.vert_shape, C128, "d_vertices")
    out = torch.empty(g.cell_shape, dtype=C128, device=dX.device) if out is None else \
        _need(out, g.cell_shape, C128, "d_odes")
    _lib.call("pea_apply_odes_wrt_vertices_fwd", g.M, g.N, _ptr(JX), _ptr(dX), _ptr(out),
              _stream())
    return out


def apply_vertices_rev(g: Grid, JX, w, out=None):
    _need(JX, (4, 4, 2, g.N, g.M), F64, "JX")
    _need(w, g.cell_shape, C128, "d_odes")
    out = torch.empty(g.vert_shape, dtype=C128, device=w.device) if out is None else \
        _need(out, g.vert_shape, C128, "d_vertices")
    _lib.call("pea_apply_odes_wrt_vertices_rev", g.M, g.N, _ptr(JX), _ptr(w), _ptr(out),
              _stream())
    return out


def export_bsr(g: Grid, JU, shift: complex = 0.0, complex_out: bool = False):
    """0-based BSR of dR/dU - shift*I: (data (nnzb,4,4), indices int32, indptr int32)."""
    _need(JU, (9, 4, 4, g.N, g.M), F64, "JU")
    shift = complex(shift)
    cplx = complex_out or shift.imag != 0.0
    nnzb = (9 * g.M - 6) * g.N
    data = torch.empty((nnzb, 4, 4), dtype=C128 if cplx else F64, device=JU.device)
    idx = torch.empty(nnzb, dtype=torch.int32, device=JU.device)
    ptr = torch.empty(g.M * g.N + 1, dtype=torch.int32, device=JU.device)
    _lib.call("pea_export_bsr", g.M, g.N, _ptr(JU), shift.real, shift.imag, int(cplx),
              _ptr(data), _ptr(idx), _ptr(ptr), _stream())
    return data, idx, ptr
