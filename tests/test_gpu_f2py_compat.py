"""The f2py-compatible module (pyeulerale_amd/f2py_compat.py) driven with NumPy arrays exactly
as the reference's class drives its compiled module (core.py:258-716), every procedure against
the oracle's restatement of the same procedure; then the reference's own resolvent round trip
(tests/test_linearization.py:332-356: BSR export, scipy's spsolve, forward product) on it."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import case, crandn, rel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def fz(shape, dtype=complex):
    return np.zeros(shape, dtype=dtype, order="F")


@pytest.fixture(scope="module")
def setup():
    from oracle.oracle import Oracle, Params
    from pyeulerale_amd.f2py_compat import spatial_discretization as disc
    X, V, U = case(33, 17, alpha=0.03, alpha_dot=0.02, seed=8)
    X, V, U = (np.asfortranarray(a.astype(complex)) for a in (X, V, U))
    prm = Params(aoa=2.5, k2=0.8, k4=1 / 40, c4=2.5)
    disc.aoa[...] = prm.aoa
    disc.jst_k2[...] = prm.k2
    disc.jst_k4[...] = prm.k4
    disc.jst_c4[...] = prm.c4
    return disc, Oracle("port"), prm, complex(0.55), X, V, U


def test_residual_forces_and_free_stream(setup):
    disc, orc, prm, mach, X, V, U = setup
    M, N = U.shape[1:]
    odes = fz((4, M, N))
    disc.compute_odes(mach, X, V, U, odes)
    assert rel(odes, orc.compute_odes(prm, mach, X, V, U)) <= 1e-12
    F = fz((2, N))
    disc.compute_forces(X, U, F)
    assert rel(F, orc.compute_forces(prm, X, U)) <= 1e-12
    fs = fz((4, M, N))
    disc.set_free_stream_state(mach, fs)
    assert rel(fs, orc.set_free_stream_state(prm, mach, M, N)) <= 1e-15


def test_jacobians_and_their_products(setup):
    disc, orc, prm, mach, X, V, U = setup
    M, N = U.shape[1:]
    rng = np.random.default_rng(1)
    J = {"mach": fz((4, 1, M, N), float), "states": fz((4, 4, M, N, 9), float),
         "vertices": fz((4, 2, M, N, 4), float), "velocities": fz((4, 2, M, N, 4), float),
         "fstates": fz((2, 4, N), float), "fvertices": fz((2, 2, N, 2), float)}
    disc.compute_odes_wrt_mach(mach, X, V, U, J["mach"])
    disc.compute_odes_wrt_states(mach, X, V, U, J["states"])
    disc.compute_odes_wrt_vertices(mach, X, V, U, J["vertices"])
    disc.compute_odes_wrt_velocities(mach, X, V, U, J["velocities"])
    disc.compute_forces_wrt_states(X, U, J["fstates"])
    disc.compute_forces_wrt_vertices(X, U, J["fvertices"])
    ref = {"mach": orc.compute_odes_wrt_mach(prm, mach, X, V, U),
           "states": orc.compute_odes_wrt_states(prm, mach, X, V, U),
           "vertices": orc.compute_odes_wrt_vertices(prm, mach, X, V, U),
           "velocities": orc.compute_odes_wrt_velocities(prm, mach, X, V, U),
           "fstates": orc.compute_forces_wrt_states(X, U),
           "fvertices": orc.compute_forces_wrt_vertices(X, U)}
    for k in J:
        assert rel(J[k], ref[k]) <= 1e-10, k
    dU, dX, w, wF = (np.asfortranarray(crandn(rng, s)) for s in
                     ((4, M, N), (2, M + 1, N + 1), (4, M, N), (2, N)))
    dMa = np.array([complex(0.3, -0.7)])
    out = fz((4, M, N))
    disc.apply_odes_wrt_states_fwd(J["states"], dU, out)
    assert rel(out, orc.apply_odes_wrt_states_fwd(ref["states"], dU)) <= 1e-12
    disc.apply_odes_wrt_states_rev(J["states"], w, out)
    assert rel(out, orc.apply_odes_wrt_states_rev(ref["states"], w)) <= 1e-12
    disc.apply_odes_wrt_vertices_fwd(J["velocities"], dX, out)
    assert rel(out, orc.apply_odes_wrt_vertices_fwd(ref["velocities"], dX)) <= 1e-12
    disc.apply_odes_wrt_mach_fwd(J["mach"], dMa, out)
    assert rel(out, orc.apply_odes_wrt_mach_fwd(ref["mach"], dMa)) <= 1e-12
    gm = fz((1,))
    disc.apply_odes_wrt_mach_rev(J["mach"], w, gm)
    assert rel(gm, orc.apply_odes_wrt_mach_rev(ref["mach"], w)) <= 1e-12
    gX = fz((2, M + 1, N + 1))
    disc.apply_odes_wrt_vertices_rev(J["vertices"], w, gX)
    assert rel(gX, orc.apply_odes_wrt_vertices_rev(ref["vertices"], w)) <= 1e-12
    dF = fz((2, N))
    disc.apply_forces_wrt_states_fwd(J["fstates"], dU, dF)
    assert rel(dF, orc.apply_forces_wrt_states_fwd(ref["fstates"], dU)) <= 1e-12
    disc.apply_forces_wrt_vertices_fwd(J["fvertices"], dX, dF)
    assert rel(dF, orc.apply_forces_wrt_vertices_fwd(ref["fvertices"], dX)) <= 1e-12
    gU = fz((4, M, N))
    disc.apply_forces_wrt_states_rev(J["fstates"], wF, gU)
    assert rel(gU, orc.apply_forces_wrt_states_rev(ref["fstates"], wF, M)) <= 1e-12
    disc.apply_forces_wrt_vertices_rev(J["fvertices"], wF, gX)
    assert rel(gX, orc.apply_forces_wrt_vertices_rev(ref["fvertices"], wF, M)) <= 1e-12
    # the BSR export, 1-based as the reference's (f90:765-814)
    nnzb = (9 * M - 6) * N
    data, idx, ptr = fz((4, 4, nnzb), float), np.zeros(nnzb, np.int32), np.zeros(M * N + 1, np.int32)
    disc.convert_odes_wrt_states(J["states"], data, idx, ptr)
    d0, i0, p0 = orc.convert_odes_wrt_states(ref["states"])
    assert np.array_equal(idx, i0 + 1) and np.array_equal(ptr, p0 + 1)
    assert rel(np.moveaxis(data, -1, 0), d0) <= 1e-10


def test_boundary_blocks_left_untouched(setup):
    """compute_odes_wrt_states writes no block of a missing radial neighbour (f90:727-746)."""
    disc, orc, prm, mach, X, V, U = setup
    M, N = U.shape[1:]
    J = np.full((4, 4, M, N, 9), 7.0, order="F")
    disc.compute_odes_wrt_states(mach, X, V, U, J)
    assert (J[:, :, 0, :, 5] == 7.0).all() and (J[:, :, :2, :, 6] == 7.0).all()
    assert (J[:, :, M - 1, :, 7] == 7.0).all() and (J[:, :, M - 2:, :, 8] == 7.0).all()
    assert not (J[:, :, 1:, :, 5] == 7.0).any()


def test_reference_round_trip_through_spsolve(setup):
    """tests/test_linearization.py:332-356 on this module: assemble the CSR matrix as
    core.py:512-531 does from convert_odes_wrt_states, solve with scipy, multiply back."""
    from scipy.sparse import bsr_array
    from scipy.sparse.linalg import spsolve
    disc, orc, prm, mach, X, V, U = setup
    M, N = U.shape[1:]
    J = fz((4, 4, M, N, 9), float)
    disc.compute_odes_wrt_states(mach, X, V, U, J)
    nnzb = (9 * M - 6) * N
    data, idx, ptr = fz((4, 4, nnzb), float), np.zeros(nnzb, np.int32), np.zeros(M * N + 1, np.int32)
    disc.convert_odes_wrt_states(J, data, idx, ptr)
    shift = complex(0.1, 0.4)
    A = bsr_array((np.moveaxis(data, -1, 0), idx - 1, ptr - 1), dtype=complex).tocsr()
    A.setdiag(A.diagonal() - shift)
    rng = np.random.default_rng(3)
    b = np.asfortranarray(crandn(rng, (4, M, N)))
    x = spsolve(A, b.reshape(-1, order="F")).reshape((4, M, N), order="F")
    back = fz((4, M, N))
    disc.apply_odes_wrt_states_fwd(J, np.asfortranarray(x), back)
    assert rel(back - shift * x, b) <= 1e-9
