"""Exact resolvent (dR/dU - s I)^-1 on the GPU by block LU over radial rings, many shifts at once.

The reference factorises ``dR/dU - s I`` with SuperLU on the host, once per shift
(core.py:533-599).  Here the structure of the O-grid does the ordering instead:

* a *ring* is every cell at one radial index m; its 4N unknowns, ordered (n, v), form one dense
  block (the angular stencil is periodic, so the ring block is a periodic band, stored dense);
* over rings the matrix is block pentadiagonal: ring m couples to rings m-2..m+2 only, and those
  couplings are block diagonal in n (blocks ii, i, o, oo of the stencil).

Block Gaussian elimination from the wall ring outwards, with partial pivoting inside every
dense ring block, needs per ring k::

    B_k = A[k,k-1] - A[k,k-2] W1[k-2]                  (row k, column k-1 after step k-2)
    P_k = A[k,k] - B_k W1[k-1] - A[k,k-2] W2[k-2]      (pivot block)
    W1[k] = P_k^-1 (A[k,k+1] - B_k W2[k-1]),  W2[k] = P_k^-1 A[k,k+2]

and the solve is y_k = P_k^-1 (b_k - B_k y_{k-1} - A[k,k-2] y_{k-2}) followed by
x_k = y_k - W1[k] x_{k+1} - W2[k] x_{k+2}.  The dense products and LU factorisations are batched
over the shifts (complex128 GEMMs, and rocSOLVER's strided-batched getrf/getrs through
``pea_dense_lu_factor`` / ``pea_dense_lu_solve``, on column-major storage); the ring blocks and
couplings come from the assembled complex-step blocks through ``pea_ring_band`` /
``pea_ring_couple``.  Memory per
shift is 3 dense (4N)^2 complex blocks per ring (LU of P_k, W1, W2): 1.6 GB per shift at
129x129 vertices, 6.4 GB at 257x129 (and 412 GB at 1025x513, beyond one GPU).  One step of iterative refinement with the exact block
operator (``pea_block_multi``) follows each solve.

The adjoint system (dR/dU)^T - conj(s) I (core.py:594-598) is factorised the same way with the
transposed ring blocks and couplings.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from . import kernels as K

C128 = torch.complex128


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def bytes_per_shift(M: int, N: int) -> int:
    """Device memory held by the factorisation of one shift."""
    return 3 * M * (4 * N) ** 2 * 16


def check_fits(M: int, N: int, shifts: int, free_bytes: int) -> None:
    """Raises :class:`PeaError` before any allocation when the factorisation of ``shifts``
    shifts needs more than ``free_bytes`` of device memory (a ring LU of 1025x513 vertices needs
    412 GB for one shift)."""
    need = bytes_per_shift(M, N) * max(1, int(shifts))
    if need > free_bytes:
        raise _lib.PeaError(
            f"ring LU of {M}x{N} cells needs {need / 1e9:.1f} GB for {shifts} shift(s), "
            f"{free_bytes / 1e9:.1f} GB of device memory are free")


def _free_bytes() -> int:
    free, _ = torch.cuda.mem_get_info()
    return int(free + torch.cuda.memory_reserved() - torch.cuda.memory_allocated())


class RingLU:
    """Factorisation of (dR/dU - s_k I), or of the adjoint system, for the shifts s_k."""

    def __init__(self, g: K.Grid, JU: torch.Tensor, shifts, *, adjoint: bool = False,
                 refine: int = 1, dense: str = "auto"):
        if JU.device.type != "cuda":
            raise RuntimeError("RingLU runs on the GPU (pea_ring_band / pea_ring_couple)")
        check_fits(g.M, g.N, len(np.atleast_1d(shifts)), _free_bytes())
        K._need(JU, (9, 4, 4, g.N, g.M), torch.float64, "JU")
        self.g = g
        self.JU = JU
        self.adjoint = bool(adjoint)
        self.shifts = np.asarray(shifts, dtype=complex).reshape(-1)
        self.S = len(self.shifts)
        self.refine = int(refine)
        # dense LU of the pivot blocks: rocSOLVER strided-batched for several shifts; for one
        # shift torch's single-matrix LU (faster than a batch of one), same factors
        if dense not in ("auto", "rocsolver", "torch"):
            raise ValueError(f"unknown dense LU backend {dense!r}")
        self.dense = ("torch" if self.S == 1 else "rocsolver") if dense == "auto" else dense
        self.device = JU.device
        self.sh = torch.as_tensor(self.shifts, dtype=C128, device=self.device).contiguous()
        self._factor()

    # ------------------------------------------------------------------ kernels
    # Dense matrices live in column-major storage (the layout of the rocSOLVER factorisation):
    # a storage tensor T of shape (S, q, n) holds the logical (S, n, q) matrix T.mT.
    def _couple(self, k, dk, Y=None, alpha=1.0, Z=None, beta=0.0):
        """Z = beta Z + alpha C_{k,k+dk} Y on column-major storage (Y None: C itself, dense)."""
        g, S, n = self.g, self.S, 4 * self.g.N
        q = n if Y is None else Y.shape[1]
        if Z is None:
            Z = torch.empty((S, q, n), dtype=C128, device=self.device)
            beta = 0.0
        for t in (Y, Z):
            if t is not None and not t.is_contiguous():
                raise RuntimeError("ring operands must be contiguous storage")
        _lib.call("pea_ring_couple", g.M, g.N, S, _p(self.JU), int(self.adjoint), k, dk, 1, q,
                  float(alpha), _p(Y), float(beta), _p(Z), K._stream())
        return Z

    def _band(self, k):
        g, S, n = self.g, self.S, 4 * self.g.N
        D = torch.empty((S, n, n), dtype=C128, device=self.device)
        _lib.call("pea_ring_band", g.M, g.N, S, _p(self.JU), _p(self.sh), int(self.adjoint), k,
                  1, _p(D), K._stream())
        return D

    def _lu_solve(self, k, B):
        """P_k^-1 B for B in column-major storage (S, q, n); returns the same storage (B itself
        with rocSOLVER)."""
        if self.dense == "torch":
            return torch.linalg.lu_solve(self.LU[k], self.piv[k], B.mT).mT.contiguous()
        n = 4 * self.g.N
        _lib.call("pea_dense_lu_solve", self.S, n, B.shape[1], _p(self.LU[k]), _p(self.piv[k]),
                  _p(B), K._stream())
        return B

    # ------------------------------------------------------------------ factorisation
    def _factor(self):
        M, S, n = self.g.M, self.S, 4 * self.g.N
        self.LU, self.piv, self.W1, self.W2 = [None] * M, [None] * M, [None] * M, [None] * M
        info = torch.zeros((M, S), dtype=torch.int32, device=self.device)
        for k in range(M):
            P = self._band(k)
            R1 = self._couple(k, 1) if k + 1 < M else None
            if k >= 1:
                B = self._couple(k, -1)
                if k >= 2:
                    self._couple(k, -2, self.W1[k - 2], -1.0, B, 1.0)
                P.mT.baddbmm_(B.mT, self.W1[k - 1].mT, alpha=-1)
                if k >= 2:
                    self._couple(k, -2, self.W2[k - 2], -1.0, P, 1.0)
                if R1 is not None:
                    R1.mT.baddbmm_(B.mT, self.W2[k - 1].mT, alpha=-1)
                del B
            if self.dense == "torch":
                LU, piv, inf = torch.linalg.lu_factor_ex(P.mT)
                info[k].copy_(inf)
                del P
            else:
                LU = P
                piv = torch.empty((S, n), dtype=torch.int32, device=self.device)
                _lib.call("pea_dense_lu_factor", S, n, _p(P), _p(piv), _p(info[k]), K._stream())
            self.LU[k], self.piv[k] = LU, piv
            if R1 is not None:
                self.W1[k] = self._lu_solve(k, R1)
            if k + 2 < M:
                self.W2[k] = self._lu_solve(k, self._couple(k, 2))
        bad = info.ne(0).any(0).cpu().numpy()
        if bad.any():
            raise _lib.PeaError(f"singular ring block for shifts {self.shifts[bad]}")

    # ------------------------------------------------------------------ solve
    def _sweep(self, b):
        """Forward and back substitution; b of shape (S, M, 1, 4N): per ring the storage of a
        column vector."""
        M = self.g.M
        y = [None] * M
        for k in range(M):
            r = b[:, k].clone()
            if k >= 1:
                self._couple(k, -1, y[k - 1], -1.0, r, 1.0)
                if k >= 2:
                    t = torch.bmm(self.W1[k - 2].mT, y[k - 1].mT).mT.contiguous()
                    self._couple(k, -2, t, 1.0, r, 1.0)
                    self._couple(k, -2, y[k - 2], -1.0, r, 1.0)
            y[k] = self._lu_solve(k, r)
        x = [None] * M
        for k in range(M - 1, -1, -1):
            xk = y[k].mT
            if k + 1 < M:
                xk = torch.baddbmm(xk, self.W1[k].mT, x[k + 1].mT, alpha=-1)
            if k + 2 < M:
                xk = xk.baddbmm_(self.W2[k].mT, x[k + 2].mT, alpha=-1)
            x[k] = xk.mT
        return torch.stack(x, 1)

    def _to_rings(self, v):
        """(S, 4, N, M) canonical storage -> (S, M, 1, 4N), ring unknowns ordered (n, v)."""
        S, g = self.S, self.g
        return v.permute(0, 3, 2, 1).reshape(S, g.M, 1, 4 * g.N).contiguous()

    def _from_rings(self, r):
        S, g = self.S, self.g
        return r.reshape(S, g.M, g.N, 4).permute(0, 3, 2, 1).contiguous()

    def apply(self, x, y):
        """y = A x with the exact blocks, A the (possibly adjoint) shifted system."""
        g = self.g
        _lib.call("pea_block_multi", g.M, g.N, self.S, _p(self.JU), _p(self.sh),
                  int(self.adjoint), _p(x), _p(y), K._stream())
        return y

    def solve(self, b: torch.Tensor) -> torch.Tensor:
        """x_k = A_k^-1 b_k; b of shape (S, 4, N, M) complex (canonical storage)."""
        g = self.g
        b = b.to(C128).reshape(self.S, *g.cell_shape).contiguous()
        x = self._from_rings(self._sweep(self._to_rings(b)))
        w = torch.empty_like(b)
        for _ in range(self.refine):
            r = b - self.apply(x, w)
            x += self._from_rings(self._sweep(self._to_rings(r)))
        self.residual = (torch.linalg.vector_norm((b - self.apply(x, w)).reshape(self.S, -1), dim=1)
                         / torch.linalg.vector_norm(b.reshape(self.S, -1), dim=1).clamp_min(1e-300))
        return x


def solve_shifts(g: K.Grid, JU: torch.Tensor, b: torch.Tensor, shifts, *, adjoint: bool = False,
                 max_bytes: float = 64 * 2 ** 30, refine: int = 1):
    """Solves all shifts, in chunks whose factorisations fit in ``max_bytes`` of device memory
    (and in the free device memory).  Raises :class:`PeaError` before factorising anything when
    not even one shift fits.  Returns (x, relative residual per shift)."""
    shifts = np.asarray(shifts, dtype=complex).reshape(-1)
    per = bytes_per_shift(g.M, g.N)
    free = 0.9 * _free_bytes()
    check_fits(g.M, g.N, 1, int(free))
    chunk = max(1, int(min(float(max_bytes), free) // per))
    xs, res = [], []
    for i in range(0, len(shifts), chunk):
        lu = RingLU(g, JU, shifts[i:i + chunk], adjoint=adjoint, refine=refine)
        xs.append(lu.solve(b[i:i + chunk]))
        res.append(lu.residual)
        del lu
    return torch.cat(xs), torch.cat(res)
