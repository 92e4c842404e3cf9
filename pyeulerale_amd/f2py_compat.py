"""The reference's FFI module re-created over the GPU library: ``spatial_discretization`` has
the 20 procedures and the 4 module variables of ``src/spatial_discretization.f90:63-87`` with
f2py's calling convention (hidden dimensions taken from the array shapes, ``intent(inout)``
arrays filled in place), so that the reference's own class can run on it unchanged: replace

    from .euler_ale import spatial_discretization as disc        (core.py:17)

by ``from pyeulerale_amd.f2py_compat import spatial_discretization as disc``.

Arrays stay NumPy, in the reference's layout and dtypes (core.py:94-112): complex128
``vertices``/``velocities`` (2, M+1, N+1), ``states``/``odes`` (4, M, N), ``forces`` (2, N);
float64 Jacobians (4,1,M,N), (4,4,M,N,9), (4,2,M,N,4), (2,4,N), (2,2,N,2).  Every procedure
copies its inputs to the GPU, runs the HIP kernels of ``libpea_hip.so`` and writes its outputs
back.  Jacobians are the complex-step blocks of the reference's method (pea_cs_linearize_odes;
the force Jacobians from the hand-written force tangents, which are linear in each input);
block products run the block-product kernels on the arrays they are given.

Differences: the primal inputs (Mach number, vertices, velocities, states) must be real-valued
(the reference's class only passes real values; complex arithmetic on them is not offered),
and a grid must have M >= 4 and N >= 5 (the reference reads out of bounds below that).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from . import kernels as K

F64, C128 = torch.float64, torch.complex128
# the procedures of spatial_discretization.f90:68-87, in the module's order
PROCEDURES = (
    "set_free_stream_state", "compute_odes", "compute_odes_wrt_mach", "apply_odes_wrt_mach_fwd",
    "apply_odes_wrt_mach_rev", "compute_odes_wrt_states", "convert_odes_wrt_states",
    "apply_odes_wrt_states_fwd", "apply_odes_wrt_states_rev", "compute_odes_wrt_vertices",
    "apply_odes_wrt_vertices_fwd", "apply_odes_wrt_vertices_rev", "compute_odes_wrt_velocities",
    "compute_forces", "compute_forces_wrt_states", "apply_forces_wrt_states_fwd",
    "apply_forces_wrt_states_rev", "compute_forces_wrt_vertices",
    "apply_forces_wrt_vertices_fwd", "apply_forces_wrt_vertices_rev",
)


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError("spatial_discretization runs on the GPU (libpea_hip.so)")
    return torch.device("cuda")


def _real(a, name):
    a = np.asarray(a)
    if np.iscomplexobj(a) and np.any(a.imag != 0):
        raise ValueError(f"{name}: complex values are not supported (real inputs only)")
    return np.real(a).astype(np.float64)


def _cells_real(a, name):
    """(k, M, N) NumPy -> canonical (k, N, M) float64 on the GPU."""
    return torch.as_tensor(np.ascontiguousarray(np.swapaxes(_real(a, name), -1, -2)),
                           device=_dev())


def _cells_cplx(a):
    a = np.asarray(a, dtype=complex)
    return torch.as_tensor(np.ascontiguousarray(np.swapaxes(a, -1, -2)), device=_dev())


def _store(out, t, name):
    """Writes canonical (k, N, M) GPU data into the caller's (k, M, N) array in place."""
    if not isinstance(out, np.ndarray) or not out.flags.writeable:
        raise RuntimeError(f"{name}: must be a writable NumPy array")
    v = np.swapaxes(t.detach().cpu().numpy(), -1, -2)
    if out.shape != v.shape:
        raise RuntimeError(f"{name}: incorrect shape {out.shape}, expected {v.shape}")
    out[...] = v


def _check(a, shape, name):
    if tuple(np.shape(a)) != tuple(shape):
        raise RuntimeError(f"{name}: incorrect shape {np.shape(a)}, expected {tuple(shape)}")


def _mach(mach):
    m = complex(np.asarray(mach).reshape(-1)[0])
    if m.imag != 0.0:
        raise ValueError("mach: complex values are not supported (real inputs only)")
    return m.real


class _Module:
    """The module object: procedures plus the writable 0-d module variables."""

    def __init__(self):
        # spatial_discretization.f90:63-66; the reference writes them with `disc.aoa[...] = x`
        self.aoa = np.zeros((), dtype=np.float64)
        self.jst_k2 = np.array(1.0)
        self.jst_k4 = np.array(1.0 / 32)
        self.jst_c4 = np.array(2.0)

    def _params(self) -> K.Params:
        return K.Params(aoa_deg=float(self.aoa), k2=float(self.jst_k2), k4=float(self.jst_k4),
                        c4=float(self.jst_c4))

    def _point(self, vertices, velocities, states):
        M, N = np.shape(states)[1:]
        _check(vertices, (2, M + 1, N + 1), "vertices")
        _check(velocities, (2, M + 1, N + 1), "velocities")
        g = K.Grid(M, N)
        return g, _cells_real(vertices, "vertices"), _cells_real(velocities, "velocities"), \
            _cells_real(states, "states")

    # ------------------------------------------------------------------ residual (f90:471-537)
    def set_free_stream_state(self, mach, states):
        M, N = np.shape(states)[1:]
        g = K.Grid(M, N)
        U = torch.empty(g.cell_shape, dtype=F64, device=_dev())
        K.set_free_stream_state(self._params(), _mach(mach), g, U)
        _store(states, U, "states")

    def compute_odes(self, mach, vertices, velocities, states, odes):
        g, X, V, U = self._point(vertices, velocities, states)
        _store(odes, K.compute_odes(self._params(), _mach(mach), g, X, V, U), "odes")

    # ------------------------------------------------------------------ Jacobians of R
    def _blocks(self, mach, vertices, velocities, states, want):
        g, X, V, U = self._point(vertices, velocities, states)
        return g, K.cs_linearize_odes(self._params(), _mach(mach), g, X, V, U, want=(want,))

    def compute_odes_wrt_mach(self, mach, vertices, velocities, states, jacs):
        g, (_, _, _, JMa) = self._blocks(mach, vertices, velocities, states, "mach")
        _check(jacs, (4, 1, g.M, g.N), "jacs")
        jacs[...] = np.swapaxes(JMa.cpu().numpy(), -1, -2)[:, None]

    def compute_odes_wrt_states(self, mach, vertices, velocities, states, jacs):
        """Blocks of absent radial neighbours stay untouched, as in f90:727-746."""
        g, (JU, _, _, _) = self._blocks(mach, vertices, velocities, states, "states")
        _check(jacs, (4, 4, g.M, g.N, 9), "jacs")
        blk = JU.permute(1, 2, 4, 3, 0).cpu().numpy()           # (4, 4, M, N, 9)
        m = np.arange(g.M)
        present = np.ones((g.M, 9), bool)
        present[:, 5], present[:, 6] = m >= 1, m >= 2            # i, ii
        present[:, 7], present[:, 8] = m <= g.M - 2, m <= g.M - 3   # o, oo
        for s in range(9):
            rows = present[:, s]
            jacs[:, :, rows, :, s] = blk[:, :, rows, :, s]

    def _grid_jacs(self, mach, vertices, velocities, states, jacs, want):
        g, J = self._blocks(mach, vertices, velocities, states, want)
        _check(jacs, (4, 2, g.M, g.N, 4), "jacs")
        jacs[...] = J[1 if want == "vertices" else 2].permute(1, 2, 4, 3, 0).cpu().numpy()

    def compute_odes_wrt_vertices(self, mach, vertices, velocities, states, jacs):
        self._grid_jacs(mach, vertices, velocities, states, jacs, "vertices")

    def compute_odes_wrt_velocities(self, mach, vertices, velocities, states, jacs):
        self._grid_jacs(mach, vertices, velocities, states, jacs, "velocities")

    # ------------------------------------------------------------------ products on R
    def apply_odes_wrt_mach_fwd(self, jacs, d_mach, d_odes):
        J = _cells_real(np.asarray(jacs)[:, 0], "jacs")           # (4, N, M)
        dm = complex(np.asarray(d_mach).reshape(-1)[0])
        _store(d_odes, J.to(C128) * dm, "d_odes")

    def apply_odes_wrt_mach_rev(self, jacs, d_odes, d_mach):
        J = _cells_real(np.asarray(jacs)[:, 0], "jacs")
        w = _cells_cplx(d_odes)
        d_mach[...] = complex((J.to(C128) * w).sum().cpu())

    @staticmethod
    def _JU(jacs):
        jacs = np.asarray(jacs, dtype=np.float64)
        M, N = jacs.shape[2:4]
        return K.Grid(M, N), torch.as_tensor(
            np.ascontiguousarray(np.transpose(jacs, (4, 0, 1, 3, 2))), device=_dev())

    @staticmethod
    def _JX(jacs):
        jacs = np.asarray(jacs, dtype=np.float64)
        M, N = jacs.shape[2:4]
        return K.Grid(M, N), torch.as_tensor(
            np.ascontiguousarray(np.transpose(jacs, (4, 0, 1, 3, 2))), device=_dev())

    def apply_odes_wrt_states_fwd(self, jacs, d_states, d_odes):
        g, JU = self._JU(jacs)
        _store(d_odes, K.apply_states_fwd(g, JU, _cells_cplx(d_states)), "d_odes")

    def apply_odes_wrt_states_rev(self, jacs, d_odes, d_states):
        g, JU = self._JU(jacs)
        _store(d_states, K.apply_states_rev(g, JU, _cells_cplx(d_odes)), "d_states")

    def apply_odes_wrt_vertices_fwd(self, jacs, d_vertices, d_odes):
        """Also the velocity products (core.py:480-481 passes the velocity blocks here)."""
        g, JX = self._JX(jacs)
        _store(d_odes, K.apply_vertices_fwd(g, JX, _cells_cplx(d_vertices)), "d_odes")

    def apply_odes_wrt_vertices_rev(self, jacs, d_odes, d_vertices):
        g, JX = self._JX(jacs)
        _store(d_vertices, K.apply_vertices_rev(g, JX, _cells_cplx(d_odes)), "d_vertices")

    def convert_odes_wrt_states(self, jacs, data, indices, index_pointers):
        """1-based BSR of the blocks as given (f90:765-814): data (4, 4, nnzb) with
        data[i, j, k] = dR_i/dU_j of block k, ascending 1-based columns per block row."""
        g, JU = self._JU(jacs)
        d, idx, ptr = K.export_bsr(g, JU)
        nnzb = (9 * g.M - 6) * g.N
        _check(data, (4, 4, nnzb), "data")
        _check(indices, (nnzb,), "indices")
        _check(index_pointers, (g.M * g.N + 1,), "index_pointers")
        data[...] = np.moveaxis(d.cpu().numpy(), 0, -1)
        indices[...] = idx.cpu().numpy() + 1
        index_pointers[...] = ptr.cpu().numpy() + 1

    # ------------------------------------------------------------------ forces (f90:1156-1375)
    def compute_forces(self, vertices, states, forces):
        M, N = np.shape(states)[1:]
        g = K.Grid(M, N)
        F = K.compute_forces(g, _cells_real(vertices, "vertices"), _cells_real(states, "states"))
        _check(forces, (2, N), "forces")
        forces[...] = F.cpu().numpy()

    def compute_forces_wrt_states(self, vertices, states, jacs):
        """(2, 4, N): F_n depends on the wall cell (0, n) only, so one tangent per variable
        (along that variable on the whole wall ring) gives every block."""
        M, N = np.shape(states)[1:]
        g = K.Grid(M, N)
        X, U = _cells_real(vertices, "vertices"), _cells_real(states, "states")
        _check(jacs, (2, 4, N), "jacs")
        for v in range(4):
            dU = torch.zeros(g.cell_shape, dtype=C128, device=U.device)
            dU[v, :, 0] = 1.0
            jacs[:, v, :] = K.jvp_forces(g, X, U, dU=dU).real.cpu().numpy()

    def compute_forces_wrt_vertices(self, vertices, states, jacs):
        """(2, 2, N, 2): F_n depends on wall vertices n (block 0) and n + 1 (block 1); tangents
        along one coordinate on the even or on the odd wall vertices separate the two."""
        M, N = np.shape(states)[1:]
        g = K.Grid(M, N)
        X, U = _cells_real(vertices, "vertices"), _cells_real(states, "states")
        _check(jacs, (2, 2, N, 2), "jacs")
        n = np.arange(N)
        for e in range(2):
            for parity in (0, 1):
                dX = torch.zeros(g.vert_shape, dtype=C128, device=X.device)
                dX[e, parity::2, 0] = 1.0
                dF = K.jvp_forces(g, X, U, dX=dX).real.cpu().numpy()   # (2, N)
                first = (n % 2) == parity           # vertex n carries the direction
                jacs[:, e, first, 0] = dF[:, first]
                jacs[:, e, ~first, 1] = dF[:, ~first]

    def apply_forces_wrt_states_fwd(self, jacs, d_states, d_forces):
        J = torch.as_tensor(np.asarray(jacs, dtype=np.float64), device=_dev()).to(C128)
        dU0 = torch.as_tensor(np.asarray(d_states, dtype=complex)[:, 0, :], device=_dev())
        d_forces[...] = torch.einsum("dvn,vn->dn", J, dU0).cpu().numpy()

    def apply_forces_wrt_states_rev(self, jacs, d_forces, d_states):
        J = torch.as_tensor(np.asarray(jacs, dtype=np.float64), device=_dev()).to(C128)
        w = torch.as_tensor(np.asarray(d_forces, dtype=complex), device=_dev())
        d_states[...] = 0.0
        d_states[:, 0, :] = torch.einsum("dvn,dn->vn", J, w).cpu().numpy()

    def apply_forces_wrt_vertices_fwd(self, jacs, d_vertices, d_forces):
        J = torch.as_tensor(np.asarray(jacs, dtype=np.float64), device=_dev()).to(C128)
        dX = torch.as_tensor(np.asarray(d_vertices, dtype=complex)[:, 0, :], device=_dev())
        d_forces[...] = (torch.einsum("den,en->dn", J[..., 0], dX[:, :-1]) +
                         torch.einsum("den,en->dn", J[..., 1], dX[:, 1:])).cpu().numpy()

    def apply_forces_wrt_vertices_rev(self, jacs, d_forces, d_vertices):
        J = torch.as_tensor(np.asarray(jacs, dtype=np.float64), device=_dev()).to(C128)
        w = torch.as_tensor(np.asarray(d_forces, dtype=complex), device=_dev())
        g0 = torch.einsum("den,dn->en", J[..., 0], w).cpu().numpy()
        g1 = torch.einsum("den,dn->en", J[..., 1], w).cpu().numpy()
        d_vertices[...] = 0.0
        d_vertices[:, 0, :-1] += g0
        d_vertices[:, 0, 1:] += g1


spatial_discretization = _Module()
