"""The reference's resolvent method, kept for comparison: SuperLU on the exported block matrix.

:func:`direct_solve` follows core.py:512-599.  The dR/dU blocks are assembled by complex step
on the GPU, exported in O(nnz) to BSR with the shift on the diagonal (the reference's
``convert_odes_wrt_states`` scans all cell pairs, f90:765-814), and factorised by SuperLU on
the host (``scipy.sparse.linalg.spsolve``, core.py:562, 596-598).  Exact, but host-bound: about
16 s per shift at 129x129 vertices.  ``solver="superlu"`` of :class:`SpatialDiscretization`
selects it.  The GPU solvers are :mod:`pyeulerale_amd.ringlu` (exact) and
:mod:`pyeulerale_amd.gmres` (iterative).
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
