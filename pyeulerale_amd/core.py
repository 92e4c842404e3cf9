"""Drop-in ``SpatialDiscretization`` over PyTorch-ROCm tensors.

Same constructor keywords, properties and method names as the reference class
(src/py_euler_ale/core.py:24-1488).  Differences by design:

* arrays are float64 / complex128 **GPU tensors**; every property returns a view in the
  reference's shape ``(NUM_VAR, M, N)`` / ``(NUM_DIM, M+1, N+1)`` / ``(NUM_DIM, N)`` of the
  library's canonical storage (radial index fastest), so in-place edits such as
  ``solver.states[:] += update`` work as in the reference (core.py:186-193);
* arguments may be torch tensors or NumPy arrays of the reference shape; outputs passed in
  (``d_odes=...``) must be complex128 views of canonical storage (e.g. an earlier result),
  the analogue of the reference's Fortran-contiguity rule (core.py:229-256);
* ``linearize()`` freezes the linearization point (states, vertices, velocities, Mach, angle of
  attack, JST constants).  The ``apply_*`` products then use hand-written tangent / adjoint
  kernels at that point (``jacobian_mode="matrix_free"``, default) or complex-step assembled
  blocks and block products, the reference's method (``jacobian_mode="complex_step"``);
* the resolvent is solved on the GPU, exactly by block LU over radial rings (``solver=
  "direct"``, many shifts batched) or by batched matrix-free GMRES (``solver="gmres"``); the
  reference's host sparse LU on the exported block matrix stays available (``"superlu"``);
* second-order gradients are exact tangent-over-adjoint products instead of central finite
  differences of complex-step Jacobians (core.py:718-1275).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from . import HEAT_RATIO, NUM_DIM, NUM_VAR
from . import kernels as K
from .grid import read_plot3d

F64, C128 = torch.float64, torch.complex128


def _direct_kind(g) -> str:
    """The exact solver for this grid: the ring block LU while one shift's factors take at most
    16 GB (up to about 257x129 vertices), nested dissection beyond."""
    from .ringlu import bytes_per_shift
    return "ring" if bytes_per_shift(g.M, g.N) <= 16 * 10 ** 9 else "nested"


def _as_tensor(a, dtype, device) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype)
    return torch.as_tensor(np.asarray(a), device=device).to(dtype)


class SpatialDiscretization:
    """Spatial discretization of the compressible 2-D Euler equations around a moving airfoil.

    Residual ``dU/dt = R(Ma, U, X, V)`` and section forces ``F(U, X)`` of a cell-centred JST
    finite-volume scheme on a closed O-grid, with their linearization, resolvent and
    second-order gradients.
    """

    def __init__(
        self,
        grid_file: str | Path | None = None,
        angle_of_attack: float = 1.25,
        mach_number: float = 0.5,
        coefficient_length: float = 1.0,
        coefficient_center: tuple[float, float] = (0.0, 0.0),
        jst_k2: float = 1,
        jst_k4: float = 1 / 32,
        jst_c4: float = 2,
        *,
        vertices=None,
        device: str | torch.device = "cuda",
        jacobian_mode: str = "matrix_free",
    ) -> None:
        """Initialize the discretization (core.py:60-113).

        Either ``grid_file`` (single-block 2-D PLOT3D, read as core.py:213-227) or ``vertices``
        (array of shape ``(NUM_DIM, M+1, N+1)``) gives the grid.
        """
        if (grid_file is None) == (vertices is None):
            raise ValueError("give exactly one of grid_file and vertices")
        if jacobian_mode not in ("matrix_free", "complex_step"):
            raise ValueError(f"unknown jacobian_mode {jacobian_mode!r}")
        self.device = torch.device(device)
        self.jacobian_mode = jacobian_mode
        self.mach_number = mach_number
        self.angle_of_attack = angle_of_attack
        self._coefficient_length = coefficient_length
        self._coefficient_center = coefficient_center
        self._k2, self._k4, self._c4 = jst_k2, jst_k4, jst_c4
        vtx = read_plot3d(grid_file) if grid_file is not None else np.asarray(vertices)
        if vtx.ndim != 3 or vtx.shape[0] != NUM_DIM:
            raise RuntimeError(f"vertices must have shape (2, M+1, N+1), got {vtx.shape}")
        self._num_radial = vtx.shape[1] - 1
        self._num_angular = vtx.shape[2] - 1
        self._grid = K.Grid(self._num_radial, self._num_angular)
        self._X = K.verts_in(_as_tensor(vtx, F64, "cpu"), F64, self.device)
        self._V = torch.zeros_like(self._X)
        self._U = torch.empty(self._grid.cell_shape, dtype=F64, device=self.device)
        self._R = torch.zeros(self._grid.cell_shape, dtype=F64, device=self.device)
        self._F = torch.zeros(self._grid.force_shape, dtype=F64, device=self.device)
        self._lin = None
        self._blocks = None
        self._factors = None
        self.set_free_stream_state()

    # ------------------------------------------------------------------ properties
    @property
    def mach_number(self) -> float:
        """Free-stream Mach number"""
        return self._mach

    @mach_number.setter
    def mach_number(self, mach_number: float) -> None:
        self._mach = float(np.real(mach_number))

    @property
    def angle_of_attack(self) -> float:
        """Far-field angle of attack in degree"""
        return self._aoa

    @angle_of_attack.setter
    def angle_of_attack(self, angle_of_attack: float) -> None:
        self._aoa = float(angle_of_attack)

    @property
    def coefficient_length(self) -> float:
        return self._coefficient_length

    @property
    def coefficient_center(self) -> tuple[float, float]:
        return self._coefficient_center

    @property
    def _dynamic_pressure(self) -> float:
        return self.mach_number ** 2 * HEAT_RATIO / 2

    @property
    def num_radial(self) -> int:
        return self._num_radial

    @property
    def num_angular(self) -> int:
        return self._num_angular

    @property
    def vertices(self) -> torch.Tensor:
        """Grid vertices, view of shape ``(NUM_DIM, num_radial+1, num_angular+1)``."""
        return self._X.transpose(1, 2)

    @property
    def velocities(self) -> torch.Tensor:
        """Grid vertex velocities, view of shape ``(NUM_DIM, num_radial+1, num_angular+1)``."""
        return self._V.transpose(1, 2)

    @property
    def states(self) -> torch.Tensor:
        """Cell states, view of shape ``(NUM_VAR, num_radial, num_angular)``."""
        return self._U.transpose(1, 2)

    @property
    def odes(self) -> torch.Tensor:
        """Residual from the last ``compute_odes``, view ``(NUM_VAR, num_radial, num_angular)``."""
        return self._R.transpose(1, 2)

    @property
    def forces(self) -> torch.Tensor:
        """Section forces from the last ``compute_forces``, shape ``(NUM_DIM, num_angular)``."""
        return self._F

    @property
    def surface_points(self) -> torch.Tensor:
        """Midpoints of consecutive surface vertices, shape ``(NUM_DIM, num_angular)``."""
        v = self.vertices[:, 0, :]
        return (v[:, :-1] + v[:, 1:]) / 2

    def _params(self) -> K.Params:
        return K.Params(aoa_deg=self._aoa, k2=self._k2, k4=self._k4, c4=self._c4)

    # ------------------------------------------------------------------ argument handling
    def _cells(self, a, name):
        t = a if isinstance(a, torch.Tensor) else _as_tensor(a, C128, self.device)
        if tuple(t.shape) != (NUM_VAR, self.num_radial, self.num_angular):
            raise RuntimeError(f"{name}: incorrect shape {tuple(t.shape)}, expected "
                               f"{(NUM_VAR, self.num_radial, self.num_angular)}")
        return K.cells_in(t, C128, self.device)

    def _verts(self, a, name):
        t = a if isinstance(a, torch.Tensor) else _as_tensor(a, C128, self.device)
        if tuple(t.shape) != (NUM_DIM, self.num_radial + 1, self.num_angular + 1):
            raise RuntimeError(f"{name}: incorrect shape {tuple(t.shape)}, expected "
                               f"{(NUM_DIM, self.num_radial + 1, self.num_angular + 1)}")
        return K.verts_in(t, C128, self.device)

    def _forces_in(self, a, name):
        t = _as_tensor(a, C128, self.device)
        if tuple(t.shape) != (NUM_DIM, self.num_angular):
            raise RuntimeError(f"{name}: incorrect shape {tuple(t.shape)}, expected "
                               f"{(NUM_DIM, self.num_angular)}")
        return t.contiguous()

    @staticmethod
    def _out_storage(out, shape_ref, transposed=True):
        """Canonical storage behind a user-provided output, or None."""
        if out is None:
            return None
        if not isinstance(out, torch.Tensor) or out.dtype != C128:
            raise RuntimeError("output must be a complex128 tensor")
        if tuple(out.shape) != tuple(shape_ref):
            raise RuntimeError(f"Incorrect shape. Got {tuple(out.shape)}, expected {shape_ref}")
        st = out.transpose(-1, -2) if transposed else out
        if not st.is_contiguous():
            raise RuntimeError("output is not a view of canonical storage")
        return st

    def _cell_out(self, t):
        return t.transpose(1, 2)

    # ------------------------------------------------------------------ primal
    def set_free_stream_state(self) -> None:
        """Sets ``states`` to the free stream (core.py:265-268)."""
        K.set_free_stream_state(self._params(), self._mach, self._grid, self._U)

    def compute_odes(self) -> None:
        """Computes ``odes`` from ``states``, ``vertices`` and ``velocities`` (core.py:270-276)."""
        K.compute_odes(self._params(), self._mach, self._grid, self._X, self._V, self._U,
                       self._R)

    def check_states(self) -> dict:
        """Counts the cells with a non-finite or non-physical state (rho <= 0 or p <= 0) and the
        cells whose last computed residual is not finite.  The reference evaluates such states
        in complex arithmetic (an imaginary wave speed, spatial_discretization.f90:142-148); in
        FP64 they make the residual NaN wherever their faces enter, and this reports it."""
        st = K.check_states(self._grid, self._U, self._R).cpu().numpy()
        return {"nonphysical_states": int(st[0]), "nonfinite_odes": int(st[1])}

    def compute_forces(self) -> None:
        """Computes ``forces`` from ``states`` and ``vertices`` (core.py:601-607)."""
        K.compute_forces(self._grid, self._X, self._U, self._F)

    # ------------------------------------------------------------------ linearization
    def linearize(self) -> None:
        """Freezes the linearization point used by the ``apply_*`` products (core.py:278-292).

        In ``complex_step`` mode the complex-step Jacobian blocks are assembled as well.
        """
        self._lin = dict(p=self._params(), mach=self._mach, X=self._X.clone(),
                         V=self._V.clone(), U=self._U.clone())
        self._blocks = None
        self._factors = None
        if self.jacobian_mode == "complex_step":
            self._assemble_blocks()

    def _point(self):
        if self._lin is None:
            raise RuntimeError("call linearize() first")
        L = self._lin
        return L["p"], L["mach"], self._grid, L["X"], L["V"], L["U"]

    def _assemble_blocks(self):
        p, mach, g, X, V, U = self._point()
        JU, JX, JV, JMa = K.cs_linearize_odes(p, mach, g, X, V, U)
        self._blocks = dict(U=JU, X=JX, V=JV, Ma=JMa)
        return self._blocks

    def _get_blocks(self):
        return self._blocks if self._blocks is not None else self._assemble_blocks()

    # reference-layout views of the complex-step Jacobians (core.py:101-112)
    @property
    def _odes_wrt_states(self) -> torch.Tensor:
        return self._get_blocks()["U"].permute(1, 2, 4, 3, 0)

    @property
    def _odes_wrt_vertices(self) -> torch.Tensor:
        return self._get_blocks()["X"].permute(1, 2, 4, 3, 0)

    @property
    def _odes_wrt_velocities(self) -> torch.Tensor:
        return self._get_blocks()["V"].permute(1, 2, 4, 3, 0)

    @property
    def _odes_wrt_mach(self) -> torch.Tensor:
        return self._get_blocks()["Ma"].transpose(1, 2)[:, None]

    # ------------------------------------------------------------------ products on R
    def apply_odes_wrt_mach_fwd(self, d_mach, d_odes=None) -> torch.Tensor:
        """``dR/dMa . dMa`` (core.py:294-319)."""
        dm = _as_tensor(d_mach, C128, self.device).reshape(-1)
        if dm.numel() != 1:
            raise RuntimeError("Incorrect shape of d_mach, expected (1,)")
        out = self._out_storage(d_odes, (NUM_VAR, self.num_radial, self.num_angular))
        p, mach, g, X, V, U = self._point()
        if self.jacobian_mode == "complex_step":
            res = self._get_blocks()["Ma"].to(C128) * dm[0]
            if out is not None:
                out.copy_(res)
                res = out
        else:
            res = K.jvp_odes(p, mach, g, X, V, U, dmach=dm, dR=out)
        return self._cell_out(res)

    def apply_odes_wrt_mach_rev(self, d_odes, d_mach=None) -> torch.Tensor:
        """``(dR/dMa)^T . dR`` (core.py:321-346); returns shape ``(1,)``."""
        w = self._cells(d_odes, "d_odes")
        p, mach, g, X, V, U = self._point()
        if self.jacobian_mode == "complex_step":
            res = (self._get_blocks()["Ma"].to(C128) * w).sum().reshape(1)
        else:
            res = K.vjp_odes(p, mach, g, X, V, U, w, want=("mach",))["mach"]
        if d_mach is not None:
            d_mach.copy_(res)
            return d_mach
        return res

    def apply_odes_wrt_states_fwd(self, d_states, d_odes=None) -> torch.Tensor:
        """``dR/dU . dU`` (core.py:348-373)."""
        d = self._cells(d_states, "d_states")
        out = self._out_storage(d_odes, (NUM_VAR, self.num_radial, self.num_angular))
        p, mach, g, X, V, U = self._point()
        if self.jacobian_mode == "complex_step":
            res = K.apply_states_fwd(g, self._get_blocks()["U"], d, out)
        else:
            res = K.jvp_odes(p, mach, g, X, V, U, dU=d, dR=out)
        return self._cell_out(res)

    def apply_odes_wrt_states_rev(self, d_odes, d_states=None) -> torch.Tensor:
        """``(dR/dU)^T . dR`` (core.py:375-400)."""
        w = self._cells(d_odes, "d_odes")
        out = self._out_storage(d_states, (NUM_VAR, self.num_radial, self.num_angular))
        p, mach, g, X, V, U = self._point()
        if self.jacobian_mode == "complex_step":
            res = K.apply_states_rev(g, self._get_blocks()["U"], w, out)
        else:
            res = K.vjp_odes(p, mach, g, X, V, U, w, want=("states",),
                             out=None if out is None else {"states": out})["states"]
        return self._cell_out(res)

    def _grid_fwd(self, which, d, d_odes):
        dd = self._verts(d, f"d_{which}")
        out = self._out_storage(d_odes, (NUM_VAR, self.num_radial, self.num_angular))
        p, mach, g, X, V, U = self._point()
        if self.jacobian_mode == "complex_step":
            res = K.apply_vertices_fwd(g, self._get_blocks()["X" if which == "vertices" else "V"],
                                       dd, out)
        else:
            kw = {"dX": dd} if which == "vertices" else {"dV": dd}
            res = K.jvp_odes(p, mach, g, X, V, U, dR=out, **kw)
        return self._cell_out(res)

    def _grid_rev(self, which, d_odes, out_arr):
        w = self._cells(d_odes, "d_odes")
        out = self._out_storage(out_arr, (NUM_DIM, self.num_radial + 1, self.num_angular + 1))
        p, mach, g, X, V, U = self._point()
        if self.jacobian_mode == "complex_step":
            res = K.apply_vertices_rev(g, self._get_blocks()["X" if which == "vertices" else "V"],
                                       w, out)
        else:
            res = K.vjp_odes(p, mach, g, X, V, U, w, want=(which,),
                             out=None if out is None else {which: out})[which]
        return res.transpose(1, 2)

    def apply_odes_wrt_vertices_fwd(self, d_vertices, d_odes=None) -> torch.Tensor:
        """``dR/dX . dX`` (core.py:402-427)."""
        return self._grid_fwd("vertices", d_vertices, d_odes)

    def apply_odes_wrt_vertices_rev(self, d_odes, d_vertices=None) -> torch.Tensor:
        """``(dR/dX)^T . dR`` (core.py:429-454)."""
        return self._grid_rev("vertices", d_odes, d_vertices)

    def apply_odes_wrt_velocities_fwd(self, d_velocities, d_odes=None) -> torch.Tensor:
        """``dR/dV . dV`` (core.py:456-482)."""
        return self._grid_fwd("velocities", d_velocities, d_odes)

    def apply_odes_wrt_velocities_rev(self, d_odes, d_velocities=None) -> torch.Tensor:
        """``(dR/dV)^T . dR`` (core.py:484-510)."""
        return self._grid_rev("velocities", d_odes, d_velocities)

    # ------------------------------------------------------------------ products on F
    def apply_forces_wrt_states_fwd(self, d_states, d_forces=None) -> torch.Tensor:
        """``dF/dU . dU`` (core.py:609-632)."""
        d = self._cells(d_states, "d_states")
        out = self._out_storage(d_forces, (NUM_DIM, self.num_angular), transposed=False)
        p, mach, g, X, V, U = self._point()
        return K.jvp_forces(g, X, U, dU=d, dF=out)

    def apply_forces_wrt_states_rev(self, d_forces, d_states=None) -> torch.Tensor:
        """``(dF/dU)^T . dF`` (core.py:634-657)."""
        w = self._forces_in(d_forces, "d_forces")
        p, mach, g, X, V, U = self._point()
        res = K.vjp_forces(g, X, U, w, want=("states",))["states"]
        if d_states is not None:
            d_states.copy_(res.transpose(1, 2))
            return d_states
        return self._cell_out(res)

    def apply_forces_wrt_vertices_fwd(self, d_vertices, d_forces=None) -> torch.Tensor:
        """``dF/dX . dX`` (core.py:659-682)."""
        d = self._verts(d_vertices, "d_vertices")
        out = self._out_storage(d_forces, (NUM_DIM, self.num_angular), transposed=False)
        p, mach, g, X, V, U = self._point()
        return K.jvp_forces(g, X, U, dX=d, dF=out)

    def apply_forces_wrt_vertices_rev(self, d_forces, d_vertices=None) -> torch.Tensor:
        """``(dF/dX)^T . dF`` (core.py:684-716)."""
        w = self._forces_in(d_forces, "d_forces")
        p, mach, g, X, V, U = self._point()
        res = K.vjp_forces(g, X, U, w, want=("vertices",))["vertices"]
        if d_vertices is not None:
            d_vertices.copy_(res.transpose(1, 2))
            return d_vertices
        return res.transpose(1, 2)

    # ------------------------------------------------------------------ resolvent
    def solve_odes_wrt_states_fwd(self, d_odes, shift: complex = 0.0, d_states=None,
                                  solver: str = "direct", **opts) -> torch.Tensor:
        """``(dR/dU - shift I)^-1 . d_odes`` at the linearization point (core.py:533-565).

        ``solver``: ``"direct"`` (exact GPU factorisation, kept for further solves with the same
        shift until the next ``linearize``: block LU over radial rings, ``"ring"``, on grids
        where one shift's factors take at most 16 GB, nested dissection, ``"nested"``, beyond),
        ``"superlu"`` (the reference's sparse LU on the host) or ``"gmres"``."""
        return self._solve(d_odes, complex(shift), False, d_states, solver, opts)

    def solve_odes_wrt_states_adj(self, d_states, shift: complex = 0.0, d_odes=None,
                                  solver: str = "direct", **opts) -> torch.Tensor:
        """``(dR/dU - shift I)^-H . d_states`` (core.py:567-599): solves with
        ``(dR/dU)^T - conj(shift) I``."""
        return self._solve(d_states, complex(shift), True, d_odes, solver, opts)

    def solve_odes_wrt_states_batched(self, d_odes, shifts, adjoint: bool = False,
                                      solver: str = "direct", fallback: str | None = "direct",
                                      **opts):
        """All shifts at once: x_k = (dR/dU - s_k I)^-1 d_odes_k (or the adjoint).
        d_odes of shape (S, NUM_VAR, M, N); returns the same shape (views).

        ``solver="direct"``: exact solves, the ring LU factorising all shifts together (in
        chunks that fit ``max_bytes`` of device memory) where it fits, else one nested-dissection
        factorisation per shift (``"ring"`` / ``"nested"`` choose).  ``solver="gmres"`` runs
        batched GMRES; with ``precond="nested"`` the shifts are split into ``groups`` groups by
        |s|, each preconditioned by the exact factorisation at its middle shift
        (``last_groups``); ``groups="adaptive"`` forms the groups from the lowest unsolved |s|
        up as convergence allows and solves the shifts of the lightly damped band one
        factorisation each (``nested.solve_shifts_adaptive``; ``last_exact``).  Shifts whose residual is still above ``rtol`` after ``maxiter`` are
        solved again with the direct method when ``fallback="direct"`` (the default), or left as
        they are (``fallback=None``; ``last_gmres.converged`` tells which)."""
        t = d_odes if isinstance(d_odes, torch.Tensor) else _as_tensor(d_odes, C128, self.device)
        shifts = np.asarray(shifts, dtype=complex).reshape(-1)
        S = len(shifts)
        if tuple(t.shape) != (S, NUM_VAR, self.num_radial, self.num_angular):
            raise RuntimeError(f"d_odes: incorrect shape {tuple(t.shape)}")
        if fallback not in ("direct", None):
            raise ValueError(f"unknown fallback {fallback!r}")
        b = K.cells_in(t, C128, self.device)
        p, mach, g, X, V, U = self._point()
        JU = self._get_blocks()["U"]
        if solver == "direct":
            solver = _direct_kind(g)
        if solver == "ring":
            from .ringlu import solve_shifts
            x, self.last_residual = solve_shifts(g, JU, b, shifts, adjoint=adjoint, **opts)
        elif solver == "nested":
            from .nested import solve_shifts_direct
            x, self.last_residual = solve_shifts_direct(g, JU, b, shifts, adjoint=adjoint, **opts)
        elif solver == "gmres" and opts.get("precond") == "nested" and \
                opts.get("groups") == "adaptive":
            from .gmres import GMRESInfo
            from .nested import solve_shifts_adaptive
            opts = dict(opts)
            opts.pop("precond")
            opts.pop("groups")
            x, groups, exact = solve_shifts_adaptive(g, JU, b, shifts, adjoint=adjoint,
                                                     point=(p, mach, g, X, V, U), **opts)
            it = np.zeros(S, int)
            res = np.zeros(S)
            conv = np.zeros(S, bool)
            for idx, _, inf in groups:
                ok = inf.converged
                it[idx[ok]], res[idx[ok]], conv[idx[ok]] = (inf.iterations[ok],
                                                            inf.residuals[ok], True)
            conv[exact] = True
            self.last_gmres = GMRESInfo(iterations=it, residuals=res, converged=conv,
                                        restarts=sum(inf.restarts for _, _, inf in groups),
                                        seconds=sum(inf.seconds for _, _, inf in groups))
            self.last_groups = groups
            self.last_exact = exact
        elif solver == "gmres" and opts.get("precond") == "nested":
            from .gmres import GMRESInfo
            from .nested import solve_shifts_anchored
            opts = dict(opts)
            opts.pop("precond")
            x, groups = solve_shifts_anchored(g, JU, b, shifts, adjoint=adjoint,
                                              point=(p, mach, g, X, V, U), **opts)
            it = np.zeros(S, int)
            res = np.zeros(S)
            conv = np.zeros(S, bool)
            for idx, _, inf in groups:
                it[idx], res[idx], conv[idx] = inf.iterations, inf.residuals, inf.converged
            self.last_gmres = GMRESInfo(iterations=it, residuals=res, converged=conv,
                                        restarts=sum(inf.restarts for _, _, inf in groups),
                                        seconds=sum(inf.seconds for _, _, inf in groups))
            self.last_groups = groups
            todo = np.flatnonzero(~conv)
            if fallback == "direct" and todo.size:
                x[todo], _ = self._direct_batch(g, JU, b[todo], shifts[todo], adjoint)
        elif solver == "gmres":
            from .gmres import BatchedGMRES
            gm = BatchedGMRES.from_discretization(self, shifts, adjoint=adjoint, **opts)
            x = gm.solve(b)
            self.last_gmres = gm.info
            todo = np.flatnonzero(~gm.info.converged)
            if fallback == "direct" and todo.size:
                x[todo], _ = self._direct_batch(g, JU, b[todo], shifts[todo], adjoint)
        else:
            raise ValueError(f"unknown solver {solver!r}")
        return x.transpose(-1, -2)

    @staticmethod
    def _direct_batch(g, JU, b, shifts, adjoint):
        if _direct_kind(g) == "ring":
            from .ringlu import solve_shifts
            return solve_shifts(g, JU, b, shifts, adjoint=adjoint)
        from .nested import solve_shifts_direct
        return solve_shifts_direct(g, JU, b, shifts, adjoint=adjoint)

    def _direct_factors(self, shift, adjoint, kind):
        key = (complex(shift), bool(adjoint), kind)
        if self._factors is None or self._factors[0] != key:
            self._factors = None        # release the previous factorisation first
            p, mach, g, X, V, U = self._point()
            JU = self._get_blocks()["U"]
            if kind == "ring":
                from .ringlu import RingLU
                fac = RingLU(g, JU, [shift], adjoint=adjoint)
            else:
                from .nested import Factorization
                fac = Factorization(g, JU, shift, adjoint=adjoint)
            self._factors = (key, fac)
        return self._factors[1]

    def _solve(self, rhs, shift, adjoint, out, solver, opts):
        b = self._cells(rhs, "rhs")
        if solver in ("direct", "ring", "nested"):
            kind = _direct_kind(self._grid) if solver == "direct" else solver
            lu = self._direct_factors(shift, adjoint, kind)
            x = lu.solve(b[None])[0]
            self.last_residual = lu.residual
        elif solver == "superlu":
            from .resolvent import direct_solve
            x = direct_solve(self, b, shift, adjoint)
        elif solver == "gmres":
            from .gmres import BatchedGMRES
            gm = BatchedGMRES.from_discretization(self, [shift], adjoint=adjoint, **opts)
            x = gm.solve(b[None])[0]
            self.last_gmres = gm.info
        else:
            raise ValueError(f"unknown solver {solver!r}")
        res = self._cell_out(x)
        return self._copy_out(res, out)

    # ------------------------------------------------------------------ second order
    # grad_C <dA, dA/dB . dB> with <a, b> = vdot(a, b) = sum conj(a) b, evaluated at the current
    # (not the frozen) point as the reference does (core.py:718-1275): one tangent-over-adjoint
    # pass with seed conj(dA) and tangent direction dB.
    def _odes_second(self, d_odes, direction: dict, wrt: str, out):
        w = self._cells(d_odes, "d_odes").conj().resolve_conj()
        g = self._grid
        H = K.hvp_odes(self._params(), self._mach, g, self._X, self._V, self._U, w,
                       want=(wrt,), **direction)[wrt]
        if wrt == "mach":
            res = H
        else:
            res = H.transpose(1, 2)
        return self._copy_out(res, out)

    def _forces_second(self, d_forces, direction: dict, wrt: str, out):
        w = self._forces_in(d_forces, "d_forces").conj().resolve_conj()
        H = K.hvp_forces(self._grid, self._X, self._U, w, want=(wrt,), **direction)[wrt]
        return self._copy_out(H.transpose(1, 2), out)

    def compute_odes_wrt_states_inner_product_wrt_mach(self, d_odes, d_states,
                                                       mach_gradient=None):
        return self._odes_second(d_odes, {"dU": self._cells(d_states, "d_states")}, "mach",
                                 mach_gradient)

    def compute_odes_wrt_states_inner_product_wrt_states(self, d_odes, d_states,
                                                         states_gradient=None):
        return self._odes_second(d_odes, {"dU": self._cells(d_states, "d_states")}, "states",
                                 states_gradient)

    def compute_odes_wrt_states_inner_product_wrt_vertices(self, d_odes, d_states,
                                                           vertices_gradient=None):
        return self._odes_second(d_odes, {"dU": self._cells(d_states, "d_states")}, "vertices",
                                 vertices_gradient)

    def compute_odes_wrt_vertices_inner_product_wrt_mach(self, d_odes, d_vertices,
                                                         mach_gradient=None):
        return self._odes_second(d_odes, {"dX": self._verts(d_vertices, "d_vertices")}, "mach",
                                 mach_gradient)

    def compute_odes_wrt_vertices_inner_product_wrt_states(self, d_odes, d_vertices,
                                                           states_gradient=None):
        return self._odes_second(d_odes, {"dX": self._verts(d_vertices, "d_vertices")},
                                 "states", states_gradient)

    def compute_odes_wrt_vertices_inner_product_wrt_vertices(self, d_odes, d_vertices,
                                                             vertices_gradient=None):
        return self._odes_second(d_odes, {"dX": self._verts(d_vertices, "d_vertices")},
                                 "vertices", vertices_gradient)

    def compute_odes_wrt_velocities_inner_product_wrt_mach(self, d_odes, d_velocities,
                                                           mach_gradient=None):
        return self._odes_second(d_odes, {"dV": self._verts(d_velocities, "d_velocities")},
                                 "mach", mach_gradient)

    def compute_odes_wrt_velocities_inner_product_wrt_states(self, d_odes, d_velocities,
                                                             states_gradient=None):
        return self._odes_second(d_odes, {"dV": self._verts(d_velocities, "d_velocities")},
                                 "states", states_gradient)

    def compute_odes_wrt_velocities_inner_product_wrt_vertices(self, d_odes, d_velocities,
                                                               vertices_gradient=None):
        return self._odes_second(d_odes, {"dV": self._verts(d_velocities, "d_velocities")},
                                 "vertices", vertices_gradient)

    def compute_forces_wrt_states_inner_product_wrt_states(self, d_forces, d_states,
                                                           states_gradient=None):
        return self._forces_second(d_forces, {"dU": self._cells(d_states, "d_states")},
                                   "states", states_gradient)

    def compute_forces_wrt_states_inner_product_wrt_vertices(self, d_forces, d_states,
                                                             vertices_gradient=None):
        return self._forces_second(d_forces, {"dU": self._cells(d_states, "d_states")},
                                   "vertices", vertices_gradient)

    def compute_forces_wrt_vertices_inner_product_wrt_states(self, d_forces, d_vertices,
                                                             states_gradient=None):
        return self._forces_second(d_forces, {"dX": self._verts(d_vertices, "d_vertices")},
                                   "states", states_gradient)

    def compute_forces_wrt_vertices_inner_product_wrt_vertices(self, d_forces, d_vertices,
                                                               vertices_gradient=None):
        return self._forces_second(d_forces, {"dX": self._verts(d_vertices, "d_vertices")},
                                   "vertices", vertices_gradient)

    # ------------------------------------------------------------------ coefficients
    def compute_drag_coefficient(self) -> float:
        """Sum of F_x over the dynamic pressure, without 1/L as in core.py:1277-1287."""
        return float(self._F[0].sum()) / self._dynamic_pressure

    def compute_drag_coefficient_wrt_forces(self, forces_gradient=None) -> torch.Tensor:
        """Gradient of c_d wrt ``forces``; carries 1/(q L) as core.py:1289-1311."""
        g = torch.zeros((NUM_DIM, self.num_angular), dtype=C128, device=self.device)
        g[0] = 1.0 / (self._dynamic_pressure * self.coefficient_length)
        return self._copy_out(g, forces_gradient)

    def compute_lift_coefficient(self) -> float:
        """c_l = sum F_z / (q L) (core.py:1313-1323)."""
        return float(self._F[1].sum()) / (self._dynamic_pressure * self.coefficient_length)

    def compute_lift_coefficient_wrt_forces(self, forces_gradient=None) -> torch.Tensor:
        g = torch.zeros((NUM_DIM, self.num_angular), dtype=C128, device=self.device)
        g[1] = 1.0 / (self._dynamic_pressure * self.coefficient_length)
        return self._copy_out(g, forces_gradient)

    def _moment_arms(self):
        pts = self.surface_points
        cx, cz = self.coefficient_center
        return pts[0] - cx, pts[1] - cz

    def compute_moment_coefficient(self) -> float:
        """c_m about ``coefficient_center`` (core.py:1347-1359)."""
        px, pz = self._moment_arms()
        m = (pz * self._F[0] - px * self._F[1]).sum()
        return float(m) / (self._dynamic_pressure * self.coefficient_length ** 2)

    def compute_moment_coefficient_wrt_forces(self, forces_gradient=None) -> torch.Tensor:
        px, pz = self._moment_arms()
        s = self._dynamic_pressure * self.coefficient_length ** 2
        g = torch.stack([pz / s, -px / s]).to(C128)
        return self._copy_out(g, forces_gradient)

    def compute_moment_coefficient_wrt_vertices(self, vertices_gradient=None) -> torch.Tensor:
        """Gradient of c_m wrt ``vertices`` holding ``forces`` (core.py:1383-1421)."""
        s = self._dynamic_pressure * self.coefficient_length ** 2
        pg = torch.stack([-self._F[1] / s, self._F[0] / s]).to(C128)
        g = torch.zeros((NUM_DIM, self.num_radial + 1, self.num_angular + 1), dtype=C128,
                        device=self.device)
        g[:, 0, :-1] += pg / 2.0
        g[:, 0, 1:] += pg / 2.0
        return self._copy_out(g, vertices_gradient)

    def compute_pressure_coefficients(self, pressure_coefficients=None) -> torch.Tensor:
        """c_p = (p - 1)/q per cell, shape ``(num_radial, num_angular)`` (core.py:1423-1452)."""
        rho, jx, jz, e = self.states
        cp = ((HEAT_RATIO - 1.0) * (e - (jx ** 2 + jz ** 2) / 2.0 / rho) - 1.0) \
            / self._dynamic_pressure
        return self._copy_out(cp, pressure_coefficients)

    def apply_pressure_coefficients_wrt_states_fwd(self, d_states,
                                                   d_pressure_coefficients=None):
        """Tangent of c_p along ``d_states`` (core.py:1454-1488)."""
        rho, jx, jz, _ = self.states
        d = _as_tensor(d_states, C128, self.device)
        drho, djx, djz, de = d
        dcp = (HEAT_RATIO - 1.0) * (de + (jx ** 2 + jz ** 2) / rho ** 2 / 2.0 * drho
                                    - (jx * djx + jz * djz) / rho) / self._dynamic_pressure
        return self._copy_out(dcp, d_pressure_coefficients)

    @staticmethod
    def _copy_out(res, out):
        if out is None:
            return res
        out.copy_(res)
        return out
