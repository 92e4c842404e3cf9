"""Batched restarted GMRES for (dR/dU - s_k I) x_k = b_k over many complex shifts at once.

The reference solves one shift at a time with SuperLU on a freshly assembled CSR matrix
(core.py:512-599).  Here all S shifts share one linearization point and advance together:

* operator: ``"blocks"`` -- the assembled complex-step blocks applied to all S vectors in one
  launch (``pea_block_multi``, blocks read once per batch); or ``"jvp"`` -- matrix-free, the
  hand-written tangent (adjoint for the transposed system) of all S vectors in one batched
  launch (``pea_jvp_odes_states_dirs`` / ``pea_vjp_odes_states_multi``), no assembled matrix;
* right preconditioning per shift: ``"line"`` (block LU of every radial line,
  ``pea_line_factor``/``pea_line_solve``), ``"jacobi"`` (4x4 diagonal blocks), ``"nested"`` (the
  exact nested-dissection factorisation of one anchor shift for every shift of the batch, all
  shifts in one multi-right-hand-side solve) or ``None``;
* orthogonalisation: classical Gram-Schmidt with selective reorthogonalisation (the projection is
  repeated only where it removed more than half of the vector's norm squared), in two or three
  fused passes over the basis;
* the Hessenberg matrices stay on the device: ``pea_gmres_rotate`` applies the Givens rotations
  of each step, one thread per shift, and the restart's triangular solves are batched.  An
  Arnoldi step is thus device work only; the step for each index j is captured once as a CUDA
  graph and replayed in later restart cycles (``graphs=True``).  The host reads the residual
  estimates every ``check_every`` steps and the true residual once per cycle.

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
                 restart: int = 30, maxiter: int = 600, rtol: float = 1e-10,
                 graphs: bool = True, check_every: int = 10, factor=None,
                 give_up: bool = False):
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
        self.graphs = bool(graphs)
        self.check_every = max(1, int(check_every))
        # stop early once every unconverged shift's observed rate of residual reduction, over
        # the last check_every steps, predicts more than maxiter steps in all
        self.give_up = bool(give_up)
        dev = JU.device if JU is not None else point[5].device
        self.device = dev
        self.sh = torch.as_tensor(self.shifts, dtype=C128, device=dev).contiguous()
        # -s (adjoint: -conj(s)) per vector, broadcast over a cell field
        self._neg_shift = -(self.sh.conj() if self.adjoint else self.sh).view(-1, 1, 1, 1)
        if operator == "blocks" and JU is None:
            raise ValueError("operator 'blocks' needs the assembled blocks")
        if operator == "jvp" and point is None:
            raise ValueError("operator 'jvp' needs the linearization point")
        if precond not in ("line", "jacobi", "nested", None):
            raise ValueError(f"unknown preconditioner {precond!r}")
        if precond is not None and JU is None:
            raise ValueError("preconditioning needs the assembled blocks")
        self.precond = precond
        self.F = None
        M, N, S = g.M, g.N, self.S
        st = K._stream()
        if precond == "nested":
            # one exact factorisation at an anchor shift preconditions every shift of the batch
            # (their systems differ from the anchor's by (s_anchor - s_k) I); its solves run
            # library calls that are not captured, so the Arnoldi steps run eagerly
            if factor is None:
                from .nested import Factorization
                anchor = self.shifts[len(self.shifts) // 2]
                factor = Factorization(g, JU, anchor, adjoint=self.adjoint)
            if factor.adjoint != self.adjoint:
                raise ValueError("the factorisation is of the other system")
            self.F = factor
            self.graphs = False
        elif precond == "line":
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
        # matrix-free: all S products in one launch, then y -= s x with s per vector.  The
        # tangent evaluates each tile's primal part once for all S directions
        # (pea_jvp_odes_states_dirs); the adjoint runs one seed per grid z.
        p, mach, gg, X, V, U = self.point
        if self.adjoint:
            K.vjp_odes_states_multi(p, mach, g, X, V, U, x, gU=y)
        else:
            K.jvp_odes_dirs(p, mach, g, X, V, U, x, dR=y)
        y.addcmul_(x, self._neg_shift)
        return y

    def psolve(self, r: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
        g = self.g
        if self.precond == "line":
            _lib.call("pea_line_solve", g.M, g.N, self.S, _p(self.JU), _p(self.F),
                      int(self.adjoint), _p(r), _p(z), K._stream())
        elif self.precond == "jacobi":
            _lib.call("pea_jacobi_solve", g.M, g.N, self.S, _p(self.F), _p(r), _p(z), K._stream())
        elif self.precond == "nested":
            self.F.precondition(r, z)
        else:
            z.copy_(r)
        return z

    # ------------------------------------------------------------------ solve
    def _step(self, j, st):
        """Arnoldi step j for all shifts, entirely on the device: w = A P^-1 v_j, Gram-Schmidt
        against v_0..v_j, v_{j+1}, column j of H, and its Givens update (``pea_gmres_rotate``)."""
        S, n, R = self.S, st["n"], self.restart
        V, w, z = st["V"], st["w"], st["z"]
        self.psolve(V[j].view(w.shape), z)      # basis vector j of every shift: one block
        self.apply(z, w)
        q = j + 1
        H = st["H"]
        if q <= 32:
            # classical Gram-Schmidt with selective reorthogonalisation (pea_gs_orthogonalize):
            # two passes over the basis rows, a third only for shifts whose first projection
            # removed more than half of |w|^2
            h = st["h1"]
            _lib.call("pea_gs_orthogonalize", S, n, q, _p(V), S * n, n, _p(w), n,
                      st["nb"], _p(st["P"]), _p(st["hw"]), _p(st["flags"]), _p(h), h.shape[1],
                      _p(st["nrm2"]), K._stream())
            hn = st["nrm2"].sqrt()
            H[:, :q, j] = h[:, :q]
        else:
            # CGS2 in three passes over the basis rows: h1 = V^H w (pea_multidot: K split into
            # chunks, which a library GEMM of this shape does not do); w -= V h1 together with
            # h2 = V^H w; w -= V h2 together with |w|^2 (pea_gs_update)
            h1, h2 = st["h1"], st["h2"]
            self._dots(st, V, q, w, h1)
            self._gs_update(st, V, q, h1, w, 1, h2)
            self._gs_update(st, V, q, h2, w, 2, st["nrm"])
            hn = st["nrm"][:, 0].real.sqrt()
            torch.add(h1[:, :q], h2[:, :q], out=H[:, :q, j])
        torch.div(w.reshape(S, n), hn.clamp_min(1e-300)[:, None], out=V[j + 1])
        H[:, j + 1, j] = hn.to(C128)
        _lib.call("pea_gmres_rotate", S, R, j, _p(H), _p(st["cs"]), _p(st["sn"]), _p(st["g"]),
                  _p(st["est"]), K._stream())
        st["hist"][j].copy_(st["est"])

    def _hopeless(self, est, prev, active, total) -> bool:
        """True when every active shift above rtol, reducing its residual estimate at the rate of
        the last check_every steps, would need more than maxiter steps in all."""
        todo = active & (est > self.rtol)
        if not todo.any():
            return False
        with np.errstate(divide="ignore", invalid="ignore"):
            rate = np.log(np.maximum(est[todo], 1e-300) / np.maximum(prev[todo], 1e-300)) \
                / self.check_every
            need = np.where(rate < 0, np.log(self.rtol / est[todo]) / rate, np.inf)
        return bool((total + need > self.maxiter).all())

    def _dots(self, st, V, q, w, out):
        """out[s, i] = sum_k conj(V[i, s, k]) w[s, k] for i < q (basis V of shape (rows, S, n))."""
        n = st["n"]
        _lib.call("pea_multidot", self.S, n, q, _p(V), self.S * n, n, _p(w), n, st["nb"],
                  _p(st["P"]), _p(out), out.shape[1], K._stream())

    def _gs_update(self, st, V, q, h, w, mode, out):
        """w[s] -= sum_{i<q} V[i, s] h[s, i]; mode 1 also gives out[s, i] = V[i, s]^H w_new,
        mode 2 out[s, 0] = |w_new[s]|^2 (basis V of shape (rows, S, n))."""
        n = st["n"]
        _lib.call("pea_gs_update", self.S, n, q, _p(V), self.S * n, n, _p(h), h.shape[1], _p(w), n,
                  mode, st["nb"], _p(st["P"]), _p(out), 0 if out is None else out.shape[1],
                  K._stream())

    def _capture(self, st):
        """One CUDA graph per step index j, captured once (after an eager cycle has warmed the
        libraries up) and replayed in every later restart cycle."""
        graphs = []
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        for j in range(self.restart):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=side):
                self._step(j, st)
            graphs.append(g)
        torch.cuda.current_stream().wait_stream(side)
        return graphs

    def solve(self, b: torch.Tensor, x0: torch.Tensor | None = None) -> torch.Tensor:
        """Solves for all shifts; b of shape (S, 4, N, M) complex (canonical storage)."""
        t0 = time.perf_counter()
        S, R = self.S, self.restart
        shape = (S, *self.g.cell_shape)
        dev = self.device
        b = b.to(C128).reshape(shape).contiguous()
        n = b[0].numel()
        x = torch.zeros_like(b) if x0 is None else x0.to(C128).reshape(shape).clone()
        # basis (rows, shifts, n): vector j of every shift is one contiguous block
        st = dict(n=n, V=torch.empty((R + 1, S, n), dtype=C128, device=dev),
                  w=torch.empty(shape, dtype=C128, device=dev),
                  z=torch.empty(shape, dtype=C128, device=dev),
                  H=torch.zeros((S, R + 1, R), dtype=C128, device=dev),
                  cs=torch.zeros((S, R), dtype=C128, device=dev),
                  sn=torch.zeros((S, R), dtype=C128, device=dev),
                  g=torch.zeros((S, R + 1), dtype=C128, device=dev),
                  est=torch.zeros(S, dtype=torch.float64, device=dev),
                  hist=torch.zeros((R, S), dtype=torch.float64, device=dev),
                  h1=torch.zeros((S, R + 1), dtype=C128, device=dev),
                  h2=torch.zeros((S, R + 1), dtype=C128, device=dev),
                  nrm=torch.zeros((S, 1), dtype=C128, device=dev),
                  nb=max(8, min(1024, 2048 // S)))
        st["P"] = torch.empty((S, st["nb"], R + 1), dtype=C128, device=dev)
        st["hw"] = torch.empty((2, S, R + 1), dtype=C128, device=dev)
        st["flags"] = torch.zeros(S, dtype=torch.int32, device=dev)
        st["nrm2"] = torch.zeros(S, dtype=torch.float64, device=dev)
        w = st["w"]
        bnorm = torch.linalg.vector_norm(b.reshape(S, n), dim=1).cpu().numpy()
        bnorm[bnorm == 0] = 1.0
        bnorm_d = torch.as_tensor(bnorm, device=dev)
        iters = np.zeros(S, dtype=int)
        history = []
        total = restarts = 0
        graphs = None
        gave_up = False
        prev_est = None
        while True:
            # r = b - A x; the true residual decides convergence, also after the last cycle
            self.apply(x, w)
            r = b - w
            beta = torch.linalg.vector_norm(r.reshape(S, n), dim=1)
            resid = beta.cpu().numpy() / bnorm
            converged = resid <= self.rtol
            if converged.all() or total >= self.maxiter or gave_up:
                break
            st["V"][0] = r.reshape(S, n) / beta.clamp_min(1e-300)[:, None]
            for key in ("H", "cs", "sn", "g"):
                st[key].zero_()     # fill kernels: these buffers may also be graph inputs
            st["g"][:, 0] = beta.to(C128)
            active = ~converged
            j_done = 0
            use_graphs = self.graphs and restarts > 0
            if use_graphs and graphs is None:
                graphs = self._capture(st)
            for j in range(R):
                if use_graphs:
                    graphs[j].replay()
                else:
                    self._step(j, st)
                j_done = j + 1
                total += 1
                if j_done % self.check_every == 0 or j_done == R or total >= self.maxiter:
                    est = (st["est"] / bnorm_d).cpu().numpy()
                    if not (active & (est > self.rtol)).any() or total >= self.maxiter:
                        break
                    if self.give_up and prev_est is not None and self._hopeless(
                            est, prev_est, active, total):
                        gave_up = True
                        break
                    prev_est = est
            hist = (st["hist"][:j_done] / bnorm_d).cpu().numpy()
            for jj in range(j_done):
                iters[active] += 1
                active &= hist[jj] > self.rtol
                history.append(hist[jj].copy())
            # x += P^-1 V y, y = H[:j, :j]^-1 g[:j] (upper triangular), zero for shifts that had
            # converged before the cycle
            Hj = torch.triu(st["H"][:, :j_done, :j_done])
            Y = torch.linalg.solve_triangular(Hj, st["g"][:, :j_done, None], upper=True)
            Y[torch.as_tensor(converged, device=dev)] = 0
            # V y as 0 - V (-y): one pass over the basis rows
            upd = torch.zeros(shape, dtype=C128, device=dev)
            negY = (-Y[:, :, 0]).contiguous()
            self._gs_update(st, st["V"], j_done, negY, upd, 0, None)
            self.psolve(upd, st["z"])
            x += st["z"]
            restarts += 1
        self.info = GMRESInfo(iterations=iters, residuals=resid, converged=converged,
                              restarts=restarts, seconds=time.perf_counter() - t0,
                              history=history)
        return x
