#!/usr/bin/env python
"""Writes the golden fixtures of the test suite from the reference's own Fortran module.

The reference module (src/spatial_discretization.f90) is compiled unchanged into
oracle/_ref/libpea_ref.so by oracle/Makefile (amdflang, this container only) and called through
oracle/oracle.py.  Inputs are our own synthetic grids and states (pyeulerale_amd.grid); the
fixtures store inputs and the reference's outputs:

  tiny.npz  -- NACA 0012, 9 x 8 vertices (M=7, N=8), rigidly pitching, Ma 0.5, alpha 1.25 deg:
               R, F, every complex-step Jacobian and fwd/rev products along seeded directions.
  c1.npz    -- NACA 0012, 129 x 65 vertices (M=64, N=128), static, Ma 0.5, alpha 1.25 deg,
               U = U_inf (1 + 1e-3 N(0,1)) from seed 0: R, F and the dR/dU products along the
               16 seeded real directions of the configuration (complex step, the reference's
               block method).  The directions are default_rng(16).standard_normal((16,4,M,N));
               they are regenerated by the tests, which check them against the stored
               checksum (dirs_head, dirs_sum) first.
  zeros.npz -- NACA 0004, 17 x 9 vertices, static, Ma 0.5, alpha 0: rho*w == 0 exactly, so
               |n.j/rho| is evaluated at signed zeros on the z = 0 grid line (derivative
               convention of abs_c, f90:97-102).

Run: python tests/golden/make_golden.py   (needs oracle/_ref built)
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from helpers import case  # noqa: E402
from oracle.oracle import Oracle, Params  # noqa: E402


def c1_directions(M, N):
    """The 16 seeded real directions of the C1 configuration."""
    return np.random.default_rng(16).standard_normal((16, 4, M, N))


def full(orc, prm, mach, X, V, U, rng):
    M, N = U.shape[1:]
    out = dict(X=X, V=V, U=U, mach=mach, aoa=prm.aoa)
    out["R"] = orc.compute_odes(prm, mach, X, V, U).real
    out["F"] = orc.compute_forces(prm, X, U).real
    out["JU"] = orc.compute_odes_wrt_states(prm, mach, X, V, U)
    out["JX"] = orc.compute_odes_wrt_vertices(prm, mach, X, V, U)
    out["JV"] = orc.compute_odes_wrt_velocities(prm, mach, X, V, U)
    out["JMa"] = orc.compute_odes_wrt_mach(prm, mach, X, V, U)
    out["JFU"] = orc.compute_forces_wrt_states(X, U)
    out["JFX"] = orc.compute_forces_wrt_vertices(X, U)
    cr = lambda s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    dU, dX, dV, w, wF = cr(U.shape), cr(X.shape), cr(X.shape), cr(U.shape), cr((2, N))
    dMa = np.array([complex(rng.standard_normal(), rng.standard_normal())])
    out.update(dU=dU, dX=dX, dV=dV, w=w, wF=wF, dMa=dMa)
    out["jvp_U"] = orc.apply_odes_wrt_states_fwd(out["JU"], dU)
    out["jvp_X"] = orc.apply_odes_wrt_vertices_fwd(out["JX"], dX)
    out["jvp_V"] = orc.apply_odes_wrt_vertices_fwd(out["JV"], dV)
    out["jvp_Ma"] = orc.apply_odes_wrt_mach_fwd(out["JMa"], dMa)
    out["vjp_U"] = orc.apply_odes_wrt_states_rev(out["JU"], w)
    out["vjp_X"] = orc.apply_odes_wrt_vertices_rev(out["JX"], w)
    out["vjp_V"] = orc.apply_odes_wrt_vertices_rev(out["JV"], w)
    out["vjp_Ma"] = orc.apply_odes_wrt_mach_rev(out["JMa"], w)
    out["fjvp_U"] = orc.apply_forces_wrt_states_fwd(out["JFU"], dU)
    out["fjvp_X"] = orc.apply_forces_wrt_vertices_fwd(out["JFX"], dX)
    out["fvjp_U"] = orc.apply_forces_wrt_states_rev(out["JFU"], wF, M)
    out["fvjp_X"] = orc.apply_forces_wrt_vertices_rev(out["JFX"], wF, M)
    return out


def main():
    ref = Oracle("ref")
    rng = np.random.default_rng(2025)
    prm = Params(aoa=1.25)
    X, V, U = case(9, 8, thickness=0.12, mach=0.5, aoa=1.25, alpha=0.02, alpha_dot=0.01,
                   seed=1, noise=5e-3)
    np.savez_compressed(HERE / "tiny.npz", **full(ref, prm, 0.5, X, V, U, rng))

    X, V, U = case(129, 65, thickness=0.12, mach=0.5, aoa=1.25, seed=0)
    M, N = U.shape[1:]
    JU = ref.compute_odes_wrt_states(prm, 0.5, X, V, U)
    dirs = c1_directions(M, N)
    prods = np.stack([ref.apply_odes_wrt_states_fwd(JU, d).real for d in dirs])
    np.savez_compressed(HERE / "c1.npz", X=X, V=V, U=U, mach=0.5, aoa=1.25,
                        R=ref.compute_odes(prm, 0.5, X, V, U).real,
                        F=ref.compute_forces(prm, X, U).real, dirs_head=dirs[:, 0, 0, :8],
                        dirs_sum=dirs.sum(axis=(1, 2, 3)), prods=prods)

    prm0 = Params(aoa=0.0)
    X, V, U = case(17, 9, thickness=0.04, mach=0.5, aoa=0.0, seed=3)
    assert np.all(U[2] == 0.0)
    np.savez_compressed(HERE / "zeros.npz", **full(ref, prm0, 0.5, X, V, U,
                                                   np.random.default_rng(9)))
    for f in ("tiny.npz", "c1.npz", "zeros.npz"):
        print(f, (HERE / f).stat().st_size, "bytes")


if __name__ == "__main__":
    main()
