"""CPU oracle for the spatial discretization -- TEST INFRASTRUCTURE ONLY.

Two interchangeable CPU backends with one NumPy interface, used by ``tests/``, the smoke check
and the CPU-baseline leg of ``bench.py`` as the checker (never by the product path):

* ``"port"`` -- ``liboracle.so``, the C restatement in ``pea_oracle.c``;
* ``"ref"``  -- ``_ref/libpea_ref.so``, the reference's own Fortran module
  (``src/spatial_discretization.f90``) compiled unchanged by ``oracle/Makefile``.  Its module
  procedures are called directly through their compiled symbols: every Fortran argument is
  passed by reference, array dimensions are explicit (f90:496-506) and the four module globals
  ``aoa, jst_k2, jst_k4, jst_c4`` (f90:63-66) are written before each call, as the reference's
  ``_configure_disc`` does (core.py:258-263).

All arrays use the reference's layout: complex128 / float64, Fortran order, shapes
``vertices (2,M+1,N+1)``, ``states (4,M,N)``, ``forces (2,N)``, Jacobians as in core.py:101-112.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_c = ctypes
_P = _c.c_void_p


def _ptr(a: np.ndarray) -> _c.c_void_p:
    return a.ctypes.data_as(_P)


@dataclass
class Params:
    """Scheme constants; defaults of core.py:60-70."""
    aoa: float = 1.25
    k2: float = 1.0
    k4: float = 1.0 / 32
    c4: float = 2.0


class _OrcParams(_c.Structure):
    _fields_ = [("aoa", _c.c_double), ("k2", _c.c_double), ("k4", _c.c_double),
                ("c4", _c.c_double)]


def _fzeros(shape, dtype=complex):
    return np.zeros(shape, dtype=dtype, order="F")


def _fcopy(a, dtype=complex):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


class Oracle:
    """One CPU backend (``"port"`` or ``"ref"``) behind the reference's procedure list."""

    def __init__(self, kind: str = "port"):
        self.kind = kind
        if kind == "port":
            path = HERE / "liboracle.so"
        elif kind == "ref":
            path = HERE / "_ref" / "libpea_ref.so"
        else:
            raise ValueError(kind)
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (run make -C oracle)")
        self.lib = _c.CDLL(str(path))
        if kind == "ref":
            pre = "_QMspatial_discretization"
            self._fn = lambda name: getattr(self.lib, f"{pre}P{name}")
            self._glob = {g: _c.c_double.in_dll(self.lib, f"{pre}E{g}")
                          for g in ("aoa", "jst_k2", "jst_k4", "jst_c4")}
        else:
            self._fn = lambda name: getattr(self.lib, f"orc_{name}")

    @staticmethod
    def available(kind: str) -> bool:
        p = HERE / ("liboracle.so" if kind == "port" else "_ref/libpea_ref.so")
        return p.exists()

    # ----------------------------------------------------------------- plumbing
    def _cfg(self, prm: Params):
        if self.kind == "ref":
            self._glob["aoa"].value = prm.aoa
            self._glob["jst_k2"].value = prm.k2
            self._glob["jst_k4"].value = prm.k4
            self._glob["jst_c4"].value = prm.c4
            return None
        return _c.byref(_OrcParams(prm.aoa, prm.k2, prm.k4, prm.c4))

    def _call_state_fn(self, name, prm, mach, X, V, U, out):
        M, N = U.shape[1], U.shape[2]
        p = self._cfg(prm)
        if self.kind == "ref":
            mc = np.array([mach], dtype=complex)
            self._fn(name)(_ptr(mc), _c.byref(_c.c_int(M)), _c.byref(_c.c_int(N)),
                           _ptr(X), _ptr(V), _ptr(U), _ptr(out))
        else:
            f = self._fn(name)
            f.argtypes = [_P, _c.c_double, _c.c_double, _c.c_int, _c.c_int, _P, _P, _P, _P]
            f(p, complex(mach).real, complex(mach).imag, M, N, _ptr(X), _ptr(V), _ptr(U),
              _ptr(out))
        return out

    # ----------------------------------------------------------------- procedures
    def set_free_stream_state(self, prm: Params, mach: complex, M: int, N: int) -> np.ndarray:
        U = _fzeros((4, M, N))
        p = self._cfg(prm)
        if self.kind == "ref":
            mc = np.array([mach], dtype=complex)
            self._fn("set_free_stream_state")(_ptr(mc), _c.byref(_c.c_int(M)),
                                              _c.byref(_c.c_int(N)), _ptr(U))
        else:
            f = self._fn("set_free_stream_state")
            f.argtypes = [_P, _c.c_double, _c.c_double, _c.c_int, _c.c_int, _P]
            f(p, complex(mach).real, complex(mach).imag, M, N, _ptr(U))
        return U

    def compute_odes(self, prm, mach, X, V, U):
        X, V, U = _fcopy(X), _fcopy(V), _fcopy(U)
        return self._call_state_fn("compute_odes", prm, mach, X, V, U, _fzeros(U.shape))

    def compute_odes_wrt_mach(self, prm, mach, X, V, U):
        X, V, U = _fcopy(X), _fcopy(V), _fcopy(U)
        M, N = U.shape[1:]
        return self._call_state_fn("compute_odes_wrt_mach", prm, mach, X, V, U,
                                   _fzeros((4, 1, M, N), float))

    def _masked(self, name, prm, mach, X, V, U, mask, out, *extra):
        """Port-only: the Jacobian blocks of the cells where ``mask`` (M, N) is true; the blocks
        of other cells stay zero."""
        if self.kind != "port":
            raise NotImplementedError("masked Jacobians exist in the C restatement only")
        M, N = U.shape[1:]
        mk = np.asfortranarray(np.asarray(mask, dtype=np.uint8).reshape(M, N))
        f = self._fn(name)
        f.argtypes = [_P, _c.c_double, _c.c_double, _c.c_int, _c.c_int, _P, _P, _P, _P] + \
            [_c.c_int] * len(extra) + [_P]
        f(self._cfg(prm), complex(mach).real, complex(mach).imag, M, N, _ptr(X), _ptr(V),
          _ptr(U), _ptr(mk), *extra, _ptr(out))
        return out

    def compute_odes_wrt_states(self, prm, mach, X, V, U, mask=None):
        X, V, U = _fcopy(X), _fcopy(V), _fcopy(U)
        M, N = U.shape[1:]
        out = _fzeros((4, 4, M, N, 9), float)
        if mask is not None:
            return self._masked("compute_odes_wrt_states_masked", prm, mach, X, V, U, mask, out)
        return self._call_state_fn("compute_odes_wrt_states", prm, mach, X, V, U, out)

    def compute_odes_wrt_vertices(self, prm, mach, X, V, U, mask=None):
        X, V, U = _fcopy(X), _fcopy(V), _fcopy(U)
        M, N = U.shape[1:]
        out = _fzeros((4, 2, M, N, 4), float)
        if mask is not None:
            return self._masked("compute_odes_wrt_grid_masked", prm, mach, X, V, U, mask, out, 0)
        return self._call_state_fn("compute_odes_wrt_vertices", prm, mach, X, V, U, out)

    def compute_odes_wrt_velocities(self, prm, mach, X, V, U, mask=None):
        X, V, U = _fcopy(X), _fcopy(V), _fcopy(U)
        M, N = U.shape[1:]
        out = _fzeros((4, 2, M, N, 4), float)
        if mask is not None:
            return self._masked("compute_odes_wrt_grid_masked", prm, mach, X, V, U, mask, out, 1)
        return self._call_state_fn("compute_odes_wrt_velocities", prm, mach, X, V, U, out)

    def _apply(self, name, M, N, J, din, out_shape):
        J = _fcopy(J, float)
        din = _fcopy(din)
        out = _fzeros(out_shape)
        if self.kind == "ref":
            self._fn(name)(_c.byref(_c.c_int(M)), _c.byref(_c.c_int(N)), _ptr(J), _ptr(din),
                           _ptr(out))
        else:
            f = self._fn(name)
            f.argtypes = [_c.c_int, _c.c_int, _P, _P, _P]
            f(M, N, _ptr(J), _ptr(din), _ptr(out))
        return out

    def apply_odes_wrt_mach_fwd(self, J, dmach):
        M, N = J.shape[2:]
        return self._apply("apply_odes_wrt_mach_fwd", M, N, J,
                           np.atleast_1d(dmach), (4, M, N))

    def apply_odes_wrt_mach_rev(self, J, dR):
        M, N = J.shape[2:]
        return self._apply("apply_odes_wrt_mach_rev", M, N, J, dR, (1,))

    def apply_odes_wrt_states_fwd(self, J, dU):
        M, N = J.shape[2:4]
        return self._apply("apply_odes_wrt_states_fwd", M, N, J, dU, (4, M, N))

    def apply_odes_wrt_states_rev(self, J, dR):
        M, N = J.shape[2:4]
        return self._apply("apply_odes_wrt_states_rev", M, N, J, dR, (4, M, N))

    def apply_odes_wrt_vertices_fwd(self, J, dX):
        M, N = J.shape[2:4]
        return self._apply("apply_odes_wrt_vertices_fwd", M, N, J, dX, (4, M, N))

    def apply_odes_wrt_vertices_rev(self, J, dR):
        M, N = J.shape[2:4]
        return self._apply("apply_odes_wrt_vertices_rev", M, N, J, dR, (2, M + 1, N + 1))

    def compute_forces(self, prm, X, U):
        X, U = _fcopy(X), _fcopy(U)
        M, N = U.shape[1:]
        F = _fzeros((2, N))
        self._cfg(prm)
        if self.kind == "ref":
            self._fn("compute_forces")(_c.byref(_c.c_int(M)), _c.byref(_c.c_int(N)), _ptr(X),
                                       _ptr(U), _ptr(F))
        else:
            f = self._fn("compute_forces")
            f.argtypes = [_c.c_int, _c.c_int, _P, _P, _P]
            f(M, N, _ptr(X), _ptr(U), _ptr(F))
        return F

    def _forces_jac(self, name, X, U, shape):
        X, U = _fcopy(X), _fcopy(U)
        M, N = U.shape[1:]
        J = _fzeros(shape, float)
        if self.kind == "ref":
            self._fn(name)(_c.byref(_c.c_int(M)), _c.byref(_c.c_int(N)), _ptr(X), _ptr(U),
                           _ptr(J))
        else:
            f = self._fn(name)
            f.argtypes = [_c.c_int, _c.c_int, _P, _P, _P]
            f(M, N, _ptr(X), _ptr(U), _ptr(J))
        return J

    def compute_forces_wrt_states(self, X, U):
        N = U.shape[2]
        return self._forces_jac("compute_forces_wrt_states", X, U, (2, 4, N))

    def compute_forces_wrt_vertices(self, X, U):
        N = U.shape[2]
        return self._forces_jac("compute_forces_wrt_vertices", X, U, (2, 2, N, 2))

    def apply_forces_wrt_states_fwd(self, J, dU):
        N = J.shape[2]
        M = dU.shape[1]
        return self._apply("apply_forces_wrt_states_fwd", M, N, J, dU, (2, N))

    def apply_forces_wrt_states_rev(self, J, dF, M):
        N = J.shape[2]
        return self._apply("apply_forces_wrt_states_rev", M, N, J, dF, (4, M, N))

    def apply_forces_wrt_vertices_fwd(self, J, dX):
        N = J.shape[2]
        M = dX.shape[1] - 1
        return self._apply("apply_forces_wrt_vertices_fwd", M, N, J, dX, (2, N))

    def apply_forces_wrt_vertices_rev(self, J, dF, M):
        N = J.shape[2]
        return self._apply("apply_forces_wrt_vertices_rev", M, N, J, dF, (2, M + 1, N + 1))

    # ----------------------------------------------------------------- port-only helpers
    def cs_jvp(self, prm, mach, X, V, U, dmach=None, dX=None, dV=None, dU=None, h=1e-30):
        """Matrix-free complex-step derivative of R along REAL directions (port only)."""
        if self.kind != "port":
            raise NotImplementedError
        X, V, U = (_fcopy(a, float) for a in (X, V, U))
        M, N = U.shape[1:]
        f = self.lib.orc_cs_jvp
        f.argtypes = [_P, _c.c_double, _c.c_int, _c.c_int, _P, _P, _P, _P, _P, _P, _P,
                      _c.c_double, _P]
        keep = []

        def opt(a, shape):
            if a is None:
                return None
            a = _fcopy(np.broadcast_to(np.asarray(a, float), shape), float)
            keep.append(a)
            return _ptr(a)

        out = _fzeros((4, M, N), float)
        f(self._cfg(prm), float(np.real(mach)), M, N, _ptr(X), _ptr(V), _ptr(U),
          opt(dmach, (1,)), opt(dX, X.shape), opt(dV, V.shape), opt(dU, U.shape), h, _ptr(out))
        return out

    def convert_odes_wrt_states(self, J):
        """0-based BSR (data (nnzb,4,4), indices, indptr) of the dR/dU blocks (port only)."""
        if self.kind != "port":
            raise NotImplementedError
        J = _fcopy(J, float)
        M, N = J.shape[2:4]
        nnzb = (9 * M - 6) * N
        data = np.zeros((nnzb, 4, 4))
        idx = np.zeros(nnzb, dtype=np.intc)
        ptr = np.zeros(M * N + 1, dtype=np.intc)
        f = self.lib.orc_convert_odes_wrt_states
        f.argtypes = [_c.c_int, _c.c_int, _P, _P, _P, _P]
        f.restype = _c.c_long
        got = f(M, N, _ptr(J), _ptr(data), _ptr(idx), _ptr(ptr))
        assert got == nnzb
        return data, idx, ptr

    # ----------------------------------------------------------------- second order
    def inner_product_gradient(self, prm, mach, X, V, U, of: str, along: str, wrt: str,
                               d_of, d_along, mask=None, step_scale: float = 1.0):
        """The reference's second-order gradients grad_wrt <d_of, d(of)/d(along) . d_along>
        (core.py:718-1275), by its own rule (core.py:749-758): central differences along
        d_along of the complex-step Jacobian of ``of`` wrt ``wrt``, transposed onto conj(d_of),
        four passes (sign +-1, real and imaginary part of d_along with factors 1 and 1j) and
        step = 1e-6 sqrt(1 + |along|) / |d_along| (Frobenius norms).

        of: "odes" (along states/vertices/velocities, wrt mach/states/vertices) or "forces"
        (along and wrt states/vertices).  Arrays in the reference's layout.  ``mask`` (M, N)
        restricts the complex-step Jacobians of the odes to those cells (port only); the result
        is then exact only where every contributing row was evaluated.  ``step_scale`` != 1
        scales the reference's step (a probe for cells where the differences straddle a switch
        of the biased maximum, f90:111-120; the reference always uses 1)."""
        X, V, U = _fcopy(X), _fcopy(V), _fcopy(U)
        M, N = U.shape[1:]
        base = {"states": U, "vertices": X, "velocities": V}[along]
        d_along = np.asarray(d_along, dtype=complex)
        w = np.conj(np.asarray(d_of, dtype=complex))
        step = step_scale * 1e-6 * (1.0 + np.linalg.norm(base)) ** 0.5 / np.linalg.norm(d_along)
        total = None
        for sign in (1.0, -1.0):
            for part, factor in ((np.real, 1.0), (np.imag, 1.0j)):
                pert = base + sign * step * part(d_along)
                Xp, Vp, Up = (pert if along == "vertices" else X,
                              pert if along == "velocities" else V,
                              pert if along == "states" else U)
                if of == "odes":
                    if wrt == "mach":
                        J = self.compute_odes_wrt_mach(prm, mach, Xp, Vp, Up)
                        g = self.apply_odes_wrt_mach_rev(J, w)
                    elif wrt == "states":
                        J = self.compute_odes_wrt_states(prm, mach, Xp, Vp, Up, mask=mask)
                        g = self.apply_odes_wrt_states_rev(J, w)
                    elif wrt == "vertices":
                        J = self.compute_odes_wrt_vertices(prm, mach, Xp, Vp, Up, mask=mask)
                        g = self.apply_odes_wrt_vertices_rev(J, w)
                    else:
                        raise ValueError(wrt)
                elif of == "forces":
                    if along == "velocities":
                        raise ValueError("forces do not depend on velocities")
                    if wrt == "states":
                        g = self.apply_forces_wrt_states_rev(
                            self.compute_forces_wrt_states(Xp, Up), w, M)
                    elif wrt == "vertices":
                        g = self.apply_forces_wrt_vertices_rev(
                            self.compute_forces_wrt_vertices(Xp, Up), w, M)
                    else:
                        raise ValueError(wrt)
                else:
                    raise ValueError(of)
                total = sign * factor * g if total is None else total + sign * factor * g
        return total / 2.0 / step
