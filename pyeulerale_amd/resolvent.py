"""Resolvent (dR/dU - s I)^-1 of the linearized discretization.

Two solvers:

* :func:`direct_solve` -- the reference's method (core.py:512-599): the dR/dU blocks are
  assembled by complex step on the GPU, exported in O(nnz) to BSR with the shift on the
  diagonal (the reference's ``convert_odes_wrt_states`` scans all cell pairs, f90:765-814) and
  factorised by SuperLU (``scipy.sparse.linalg.spsolve``, core.py:562, 596-598).  Exact; meant
  for grids whose LU fits in host memory.
* :func:`gmres_solve` -- matrix-free restarted GMRES on the GPU for many shifts at once, with
  the hand-written tangent (or adjoint) kernel as the operator and radial-line block
  preconditioning (see :mod:`pyeulerale_amd.precond`).
"""

from __future__ import annotations

import numpy as np
import torch

from . import kernels as K

C128 = torch.complex128


def direct_solve(sd, b: torch.Tensor, shift: complex, adjoint: bool) -> torch.Tensor:
    """Solves (J - s I) x = b, or (J - s I)^H x = b, by sparse LU; b, x in canonical storage."""
    from scipy.sparse import bsr_array
    from scipy.sparse.linalg import spsolve

    p, mach, g, X, V, U = sd._point()
    JU = sd._get_blocks()["U"]
    s = complex(shift).conjugate() if adjoint else complex(shift)
    data, idx, ptr = K.export_bsr(g, JU, s, complex_out=True)
    n = 4 * g.C
    A = bsr_array((data.cpu().numpy(), idx.cpu().numpy(), ptr.cpu().numpy()), shape=(n, n))
    A = A.tocsr()
    if adjoint:
        A = A.T.tocsr()
    # canonical storage (4, N, M) -> BSR unknown order (cell c = n*M + m, var v) = index 4c + v
    rhs = b.permute(1, 2, 0).reshape(-1).cpu().numpy()
    sol = spsolve(A, rhs)
    x = torch.as_tensor(np.asarray(sol, dtype=complex)).reshape(g.N, g.M, 4).permute(2, 0, 1)
    return x.contiguous().to(b.device)


def gmres_solve(sd, b: torch.Tensor, shifts, adjoint: bool = False, **opts) -> torch.Tensor:
    from .gmres import BatchedGMRES
    return BatchedGMRES.from_discretization(sd, shifts, adjoint=adjoint, **opts).solve(b)
