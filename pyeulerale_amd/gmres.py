"""Batched restarted GMRES for (dR/dU - s_k I) x_k = b_k over many complex shifts at once.

The reference solves one shift at a time with SuperLU on a freshly assembled CSR matrix
(core.py:512-599).  Here all S shifts share one linearization point and advance together:

* operator: ``"blocks"`` -- the assembled complex-step blocks applied to all S vectors in one
  launch (``pea_block_multi``, blocks read once per batch); or ``"jvp"`` -- the hand-written
  tangent (adjoint for the transposed system) per vector, no assembled matrix;
* right preconditioning per shift: ``"line"`` (block LU of every radial line,
  ``pea_line_factor``/``pea_line_solve``), ``"jacobi"`` (4x4 diagonal blocks) or ``None``;
* orthogonalisation: classical Gram-Schmidt applied twice (CGS2) as batched matrix products;
* the small Hessenberg problems (Givens rotations, triangular solves) are done on the host.

The adjoint system (dR/dU - s I)^H y = c is (dR/dU)^T - conj(s) I, as core.py:594-598.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from . import kernels as K

C128 = torch.complex128


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


@dataclass
class GMRESInfo:
    iterations: np.ndarray          # Arnoldi steps per shift until convergence
    residuals: np.ndarray           # final relative residual estimate per shift
    converged: np.ndarray
    restarts: int = 0
    seconds: float = 0.0
    history: list = field(default_factory=list)


class BatchedGMRES:
    """Solver for S shifted systems sharing one linearization point."""

    def __init__(self, g: K.Grid, JU: torch.Tensor | None, shifts, *, adjoint: bool = False,
                 point=None, precond: str | None = "line", operator: str = "blocks",
                 restart: int = 30, maxiter: int = 600, rtol: float = 1e-10):
        self.g = g
        self.adjoint = bool(adjoint)
        self.shifts = np.asarray(shifts, dtype=complex).reshape(-1)
        self.S = len(self.shifts)
        self.JU = JU
        self.point = point
        self.operator = operator
        self.restart = int(restart)
        self.maxiter = int(maxiter)
        self.rtol = float(rtol)
        dev = JU.device if JU is not None else point[5].device
        self.device = dev
        self.sh = torch.as_tensor(self.shifts, dtype=C128, device=dev).contiguous()
        if operator == "blocks" and JU is None:
            raise ValueError("operator 'blocks' needs the assembled blocks")
        if operator == "jvp" and point is None:
            raise ValueError("operator 'jvp' needs the linearization point")
        if precond not in ("line", "jacobi", None):
            raise ValueError(f"unknown preconditioner {precond!r}")
        if precond is not None and JU is None:
            raise ValueError("preconditioning needs the assembled blocks")
        self.precond = precond
        self.F = None
        M, N, S = g.M, g.N, self.S
        st = K._stream()
        if precond == "line":
            nbytes = int(_lib.load().pea_line_factor_bytes(M, N, S))
            self.F = torch.empty(nbytes // 8, dtype=torch.float64, device=dev)
            _lib.call("pea_line_factor", M, N, S, _p(JU), _p(self.sh), int(self.adjoint),
                      _p(self.F), st)
        elif precond == "jacobi":
            self.F = torch.empty(S * 16 * M * N, dtype=C128, device=dev)
            _lib.call("pea_jacobi_factor", M, N, S, _p(JU), _p(self.sh), int(self.adjoint),
                      _p(self.F), st)

    @classmethod
    def from_discretization(cls, sd, shifts, adjoint=False, **opts):
        p, mach, g, X, V, U = sd._point()
        JU = sd._get_blocks()["U"]
        return cls(g, JU, shifts, adjoint=adjoint, point=(p, mach, g, X, V, U), **opts)

    # ------------------------------------------------------------------ operators
    def apply(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        """y_k = (J - s_k I) x_k (or the adjoint operator); x, y of shape (S, 4, N, M)."""
        g = self.g
        if self.operator == "blocks":
            _lib.call("pea_block_multi", g.M, g.N, self.S, _p(self.JU), _p(self.sh),
                      int(self.adjoint), _p(x), _p(y), K._stream())
            return y
        p, mach, gg, X, V, U = self.point
        for k in range(self.S):
            if self.adjoint:
                K.vjp_odes(p, mach, g, X, V, U, x[k], out={"states": y[k]})
                y[k].sub_(x[k] * complex(np.conj(self.shifts[k])))
            else:
                K.jvp_odes(p, mach, g, X, V, U, dU=x[k], dR=y[k])
                y[k].sub_(x[k] * complex(self.shifts[k]))
        return y

    def psolve(self, r: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
        g = self.g
        if self.precond == "line":
            _lib.call("pea_line_solve", g.M, g.N, self.S, _p(self.JU), _p(self.F),
                      int(self.adjoint), _p(r), _p(z), K._stream())
        elif self.precond == "jacobi":
            _lib.call("pea_jacobi_solve", g.M, g.N, self.S, _p(self.F), _p(r), _p(z), K._stream())
        else:
            z.copy_(r)
        return z

    # ------------------------------------------------------------------ solve
    def solve(self, b: torch.Tensor, x0: torch.Tensor | None = None) -> torch.Tensor:
        """Solves for all shifts; b of shape (S, 4, N, M) complex (canonical storage)."""
        t0 = time.perf_counter()
        S, R = self.S, self.restart
        shape = (S, *self.g.cell_shape)
        b = b.to(C128).reshape(shape).contiguous()
        n = b[0].numel()
        x = torch.zeros_like(b) if x0 is None else x0.to(C128).reshape(shape).clone()
        V = torch.empty((S, R + 1, n), dtype=C128, device=self.device)
        w = torch.empty(shape, dtype=C128, device=self.device)
        z = torch.empty(shape, dtype=C128, device=self.device)
        bnorm = torch.linalg.vector_norm(b.reshape(S, n), dim=1).cpu().numpy()
        bnorm[bnorm == 0] = 1.0
        iters = np.zeros(S, dtype=int)
        resid = np.ones(S)
        converged = np.zeros(S, dtype=bool)
        history = []
        total = 0
        restarts = 0
        while True:
            # r = b - A x; the true residual decides convergence, also after the last cycle
            self.apply(x, w)
            r = b - w
            beta = torch.linalg.vector_norm(r.reshape(S, n), dim=1)
            beta_h = beta.cpu().numpy()
            resid = beta_h / bnorm
            converged = resid <= self.rtol
            if converged.all() or total >= self.maxiter:
                break
            V[:, 0] = (r.reshape(S, n) / beta.clamp_min(1e-300)[:, None])
            H = np.zeros((S, R + 1, R), dtype=complex)
            cs = np.zeros((S, R), dtype=complex)
            sn = np.zeros((S, R), dtype=complex)
            gvec = np.zeros((S, R + 1), dtype=complex)
            gvec[:, 0] = beta_h
            active = ~converged
            j_done = 0
            for j in range(R):
                w.copy_(V[:, j].reshape(shape))
                self.psolve(w, z)
                self.apply(z, w)
                wf = w.reshape(S, n, 1)
                Vj = V[:, :j + 1]
                # CGS2: two passes of h = V^H w, w -= V h (the basis itself is never
                # conjugated or copied: h = conj(V conj(w)))
                h = torch.bmm(Vj, wf.conj().resolve_conj()).conj().resolve_conj()
                wf -= torch.bmm(h.transpose(1, 2).contiguous(), Vj).transpose(1, 2)
                h2 = torch.bmm(Vj, wf.conj().resolve_conj()).conj().resolve_conj()
                wf -= torch.bmm(h2.transpose(1, 2).contiguous(), Vj).transpose(1, 2)
                h = h + h2
                hn = torch.linalg.vector_norm(wf[:, :, 0], dim=1)
                V[:, j + 1] = wf[:, :, 0] / hn.clamp_min(1e-300)[:, None]
                hc = h[:, :, 0].cpu().numpy()
                hnh = hn.cpu().numpy()
                H[:, :j + 1, j] = hc
                H[:, j + 1, j] = hnh
                # complex Givens rotations G = [[conj(c), conj(s)], [-s, c]], per shift
                for s in range(S):
                    for i in range(j):
                        hi, hi1 = H[s, i, j], H[s, i + 1, j]
                        H[s, i, j] = np.conj(cs[s, i]) * hi + np.conj(sn[s, i]) * hi1
                        H[s, i + 1, j] = -sn[s, i] * hi + cs[s, i] * hi1
                    a, bb = H[s, j, j], H[s, j + 1, j]
                    den = np.sqrt(abs(a) ** 2 + abs(bb) ** 2)
                    if den == 0.0:
                        cs[s, j], sn[s, j] = 1.0, 0.0
                    else:
                        cs[s, j] = a / den
                        sn[s, j] = bb / den
                    H[s, j, j] = den
                    H[s, j + 1, j] = 0.0
                    gj = gvec[s, j]
                    gvec[s, j] = np.conj(cs[s, j]) * gj
                    gvec[s, j + 1] = -sn[s, j] * gj
                j_done = j + 1
                total += 1
                est = np.abs(gvec[:, j + 1]) / bnorm
                iters[active] += 1
                active &= est > self.rtol
                history.append(est.copy())
                if not active.any() or total >= self.maxiter:
                    break
            # x += P^-1 V y
            Y = np.zeros((S, j_done), dtype=complex)
            for s in range(S):
                if converged[s]:
                    continue
                Hs = H[s, :j_done, :j_done]
                Y[s] = np.linalg.solve(np.triu(Hs), gvec[s, :j_done]) if j_done else Y[s]
            Yd = torch.as_tensor(Y, dtype=C128, device=self.device)
            upd = torch.bmm(Yd[:, None, :], V[:, :j_done]).reshape(shape)
            self.psolve(upd.contiguous(), z)
            x += z
            restarts += 1
        self.info = GMRESInfo(iterations=iters, residuals=resid, converged=converged,
                              restarts=restarts, seconds=time.perf_counter() - t0,
                              history=history)
        return x
