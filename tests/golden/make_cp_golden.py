"""Writes tests/golden/examples/response_Ma0p5_cp.npz: the steady pressure coefficients of the
response example (NACA0004 129x129, Ma 0.5, chord 10, axis 0; examples/response/run.py) and
their transfer from pitch angle at reduced frequency 1, computed with the CPU oracle: PTC as
run.py:89-123 with SuperLU steps on the oracle's complex-step blocks, then
dU = -(J - sI)^-1 (J_X + s J_V) dX/dalpha (run.py:125-176) and c_p, dc_p (core.py:1423-1488).

The reference's own examples/response/Ma0p5.gz is not used: its c_p does not follow from the
reference's current code (this script, whose PTC history reproduces Ma0p5.out digit for digit,
differs from it by up to 0.05 at the wall), and the reference's CI does not check that file.
Takes about 5 minutes on one core.
"""

import sys
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from oracle.oracle import Oracle, Params  # noqa: E402
from pyeulerale_amd.grid import read_plot3d  # noqa: E402

GAMMA = 1.4


def main() -> None:
    orc = Oracle("port")
    X = read_plot3d(HERE / "examples" / "naca0004_129x129.plot3d.gz").astype(complex)
    V = np.zeros_like(X)
    M, N = X.shape[1] - 1, X.shape[2] - 1
    prm, mach, chord, axis = Params(aoa=0.0), 0.5, 10.0, 0.0
    U = orc.set_free_stream_state(prm, mach, M, N)
    n0 = np.abs(orc.compute_odes(prm, mach, X, V, U)).max()
    eye = sp.eye(4 * M * N)
    hist = []
    for _ in range(100):
        R = orc.compute_odes(prm, mach, X, V, U)
        rel = np.abs(R).max() / n0
        hist.append(rel)
        print(f"{len(hist) - 1:>6} {np.abs(R).max():>15.1e} {rel:>15.1e}", flush=True)
        JU = orc.compute_odes_wrt_states(prm, mach, X, V, U)
        if rel <= 1e-9:
            break
        data, idx, ptr = orc.convert_odes_wrt_states(JU)
        A = sp.bsr_array((data, idx, ptr), shape=(4 * M * N,) * 2).tocsc() - (10.0 * rel) * eye
        dU = spla.spsolve(A.astype(complex), -R.reshape(-1, order="F"))
        U = U + dU.reshape((4, M, N), order="F").real
    dX = np.empty_like(X)
    dX[1] = -(X[0] - axis)
    dX[0] = X[1]
    s = 1j * 1.0 * mach * GAMMA ** 0.5 / (chord / 2)
    JX = orc.compute_odes_wrt_vertices(prm, mach, X, V, U)
    JV = orc.compute_odes_wrt_velocities(prm, mach, X, V, U)
    rhs = orc.apply_odes_wrt_vertices_fwd(JX, dX) + s * orc.apply_odes_wrt_vertices_fwd(JV, dX)
    data, idx, ptr = orc.convert_odes_wrt_states(JU)
    A = sp.bsr_array((data, idx, ptr), shape=(4 * M * N,) * 2).tocsc().astype(complex) - s * eye
    dU = -spla.spsolve(A, rhs.reshape(-1, order="F")).reshape((4, M, N), order="F")
    rho, jx, jz, e = U.real
    q = GAMMA * mach ** 2 / 2
    cp = ((GAMMA - 1) * (e - (jx ** 2 + jz ** 2) / 2 / rho) - 1) / q
    dcp = (GAMMA - 1) * (dU[3] + (jx ** 2 + jz ** 2) / rho ** 2 / 2 * dU[0]
                         - (jx * dU[1] + jz * dU[2]) / rho) / q
    np.savez_compressed(HERE / "examples" / "response_Ma0p5_cp.npz", cp=cp, dcp=dcp,
                        ptc_rel=np.array(hist))


if __name__ == "__main__":
    main()
