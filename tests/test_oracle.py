"""CPU tests that pin the oracle (oracle/pea_oracle.c) before it is trusted as the checker.

* golden fixtures written by the reference's own Fortran compiled unchanged
  (tests/golden/make_golden.py);
* the compiled reference itself on fresh random cases (when oracle/_ref is built);
* properties the mathematics fixes (free-stream preservation on static and rigidly rotating
  grids, seam, translation and scale invariance, closed-surface force balance), each chosen so
  that a dropped term, a wrong sign or a transposed index fails it;
* the reference's own test method: finite differences against the complex-step Jacobians
  (tests/test_linearization.py:48-330) and the adjoint identity.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from helpers import c1_directions, case, crandn, free_stream, rel
from oracle.oracle import Oracle, Params

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def port():
    return Oracle("port")


def _load(name):
    z = np.load(GOLD / name, allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", ["tiny.npz", "zeros.npz"])
def test_port_reproduces_reference_fixtures(port, name):
    d = _load(name)
    prm = Params(aoa=float(d["aoa"]))
    mach = float(d["mach"])
    X, V, U = d["X"], d["V"], d["U"]
    M = U.shape[1]
    assert rel(port.compute_odes(prm, mach, X, V, U).real, d["R"]) <= 1e-13
    assert rel(port.compute_forces(prm, X, U).real, d["F"]) <= 1e-14
    J = {"JU": port.compute_odes_wrt_states(prm, mach, X, V, U),
         "JX": port.compute_odes_wrt_vertices(prm, mach, X, V, U),
         "JV": port.compute_odes_wrt_velocities(prm, mach, X, V, U),
         "JMa": port.compute_odes_wrt_mach(prm, mach, X, V, U),
         "JFU": port.compute_forces_wrt_states(X, U),
         "JFX": port.compute_forces_wrt_vertices(X, U)}
    for k, v in J.items():
        assert rel(v, d[k]) <= 1e-12, k
    # products with the reference's own blocks
    assert rel(port.apply_odes_wrt_states_fwd(d["JU"], d["dU"]), d["jvp_U"]) <= 1e-14
    assert rel(port.apply_odes_wrt_states_rev(d["JU"], d["w"]), d["vjp_U"]) <= 1e-14
    assert rel(port.apply_odes_wrt_vertices_fwd(d["JX"], d["dX"]), d["jvp_X"]) <= 1e-14
    assert rel(port.apply_odes_wrt_vertices_rev(d["JX"], d["w"]), d["vjp_X"]) <= 1e-14
    assert rel(port.apply_odes_wrt_mach_fwd(d["JMa"], d["dMa"]), d["jvp_Ma"]) <= 1e-14
    assert rel(port.apply_odes_wrt_mach_rev(d["JMa"], d["w"]), d["vjp_Ma"]) <= 1e-14
    assert rel(port.apply_forces_wrt_states_fwd(d["JFU"], d["dU"]), d["fjvp_U"]) <= 1e-14
    assert rel(port.apply_forces_wrt_vertices_fwd(d["JFX"], d["dX"]), d["fjvp_X"]) <= 1e-14
    assert rel(port.apply_forces_wrt_states_rev(d["JFU"], d["wF"], M), d["fvjp_U"]) <= 1e-14
    assert rel(port.apply_forces_wrt_vertices_rev(d["JFX"], d["wF"], M), d["fvjp_X"]) <= 1e-14


def test_port_reproduces_c1_fixture(port):
    d = _load("c1.npz")
    prm = Params(aoa=float(d["aoa"]))
    X, V, U = d["X"], d["V"], d["U"]
    assert rel(port.compute_odes(prm, 0.5, X, V, U).real, d["R"]) <= 1e-13
    assert rel(port.compute_forces(prm, X, U).real, d["F"]) <= 1e-14
    dirs = c1_directions(d)
    for k in (0, 7, 15):
        got = port.cs_jvp(prm, 0.5, X, V, U, dU=dirs[k], h=1e-12)
        assert rel(got, d["prods"][k]) <= 1e-10


@pytest.mark.skipif(not Oracle.available("ref"), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_port_matches_compiled_reference(port, seed):
    ref = Oracle("ref")
    rng = np.random.default_rng(100 + seed)
    A = int(rng.integers(8, 40))
    Rv = int(rng.integers(6, 20))
    mach = float(rng.uniform(0.2, 0.7))
    aoa = float(rng.uniform(-3, 3))
    prm = Params(aoa=aoa, k2=float(rng.uniform(0.5, 1.5)), k4=float(rng.uniform(0.01, 0.05)),
                 c4=float(rng.uniform(1.0, 3.0)))
    X, V, U = case(A, Rv, thickness=float(rng.uniform(0.04, 0.18)), mach=mach, aoa=aoa,
                   alpha=float(rng.uniform(-0.05, 0.05)), alpha_dot=float(rng.uniform(-0.02, 0.02)),
                   seed=seed, noise=1e-2)
    assert rel(port.compute_odes(prm, mach, X, V, U), ref.compute_odes(prm, mach, X, V, U)) \
        <= 1e-13
    assert rel(port.compute_odes_wrt_states(prm, mach, X, V, U),
               ref.compute_odes_wrt_states(prm, mach, X, V, U)) <= 1e-12
    assert rel(port.compute_odes_wrt_vertices(prm, mach, X, V, U),
               ref.compute_odes_wrt_vertices(prm, mach, X, V, U)) <= 1e-12


# ------------------------------------------------------------------ mathematical properties
def _uinf_case(A=33, Rv=17, alpha=0.0, alpha_dot=0.0, aoa=2.0, mach=0.5):
    X, V, _ = case(A, Rv, mach=mach, aoa=aoa, alpha=alpha, alpha_dot=alpha_dot)
    M, N = Rv - 1, A - 1
    U = np.empty((4, M, N))
    U[:] = free_stream(mach, aoa)[:, None, None]
    return X, V, U


def _flux_scale(X, U):
    # |Phi| |n| / area, the magnitude of one face's contribution
    d1 = X[:, 1:, 1:] - X[:, :-1, :-1]
    d2 = X[:, 1:, :-1] - X[:, :-1, 1:]
    area = 0.5 * np.abs(d1[0] * d2[1] - d1[1] * d2[0])
    edge = np.hypot(*(X[:, :-1, 1:] - X[:, :-1, :-1]))
    return np.abs(U[3]).max() * edge / area


@pytest.mark.parametrize("alpha,alpha_dot", [(0.0, 0.0), (0.03, 0.2)])
def test_free_stream_preserved_off_the_wall(port, alpha, alpha_dot):
    """R(U_inf) = 0 except in the wall ring, on static and rigidly rotating grids."""
    X, V, U = _uinf_case(alpha=alpha, alpha_dot=alpha_dot)
    R = port.compute_odes(Params(aoa=2.0), 0.5, X, V, U).real
    scale = _flux_scale(X, U)
    assert np.all(np.abs(R[:, 1:, :]) <= 1e-12 * scale[None, 1:, :])
    assert np.abs(R[:, 0, :]).max() > 1e-6   # the wall ghost does act


def test_seam_roll_and_translation(port):
    X, V, U = case(33, 17, alpha=0.02, alpha_dot=0.05, noise=1e-2)
    prm = Params(aoa=1.0)
    R = port.compute_odes(prm, 0.5, X, V, U).real
    k = 5
    Xr = np.concatenate([X[:, :, k:-1], X[:, :, :k + 1]], axis=2)
    Vr = np.concatenate([V[:, :, k:-1], V[:, :, :k + 1]], axis=2)
    Rr = port.compute_odes(prm, 0.5, Xr, Vr, np.roll(U, -k, axis=2)).real
    assert rel(Rr, np.roll(R, -k, axis=2)) <= 1e-12
    Xt = X + np.array([3.0, -7.0])[:, None, None]
    assert rel(port.compute_odes(prm, 0.5, Xt, V, U).real, R) <= 1e-9


def test_scaling(port):
    """X -> lam X (static) gives R -> R / lam: normals scale with lam, areas with lam^2."""
    X, _, U = case(33, 17, noise=1e-2)
    V = np.zeros_like(X)
    prm = Params(aoa=1.0)
    R = port.compute_odes(prm, 0.5, X, V, U).real
    lam = 4.0
    assert rel(port.compute_odes(prm, 0.5, lam * X, V, U).real, R / lam) <= 1e-13


def test_forces_balance_at_free_stream(port):
    X, V, U = _uinf_case()
    F = port.compute_forces(Params(aoa=2.0), X, U).real
    p = 1.0  # free-stream pressure
    assert np.abs(F.sum(axis=1)).max() <= 1e-13 * p * np.abs(X[:, 0]).max()
    # F = p * (dz, -dx) per surface edge
    dz = X[1, 0, 1:] - X[1, 0, :-1]
    np.testing.assert_allclose(F[0], p * dz, rtol=1e-12, atol=1e-16)


# ------------------------------------------------------------------ the reference's test method
def test_fd_vs_complex_step_blocks(port):
    """Finite differences against the complex-step blocks, every block, with the tolerances
    of tests/test_linearization.py:78-155 (atol 1e-6, rtol 1e-5).  Central differences: our
    grids resolve the wall (first spacing 1e-3 chord), where forward differences with
    h = 1e-7 carry O(h) errors above that tolerance."""
    X, V, U = case(9, 8, alpha=0.02, alpha_dot=0.01, seed=1, noise=5e-3)
    prm = Params(aoa=1.25)
    M, N = U.shape[1:]
    R0 = port.compute_odes(prm, 0.5, X, V, U).real
    JU = port.compute_odes_wrt_states(prm, 0.5, X, V, U)
    nb = lambda m, n, blk: [(m, n), (m, (n - 1) % N), (m, (n - 2) % N), (m, (n + 1) % N),
                            (m, (n + 2) % N), (m - 1, n), (m - 2, n), (m + 1, n),
                            (m + 2, n)][blk]
    for j in range(4):
        for mm in range(M):
            for nn in range(N):
                Up, Um = U.copy(), U.copy()
                Up[j, mm, nn] += 1e-6
                Um[j, mm, nn] -= 1e-6
                col = (port.compute_odes(prm, 0.5, X, V, Up).real -
                       port.compute_odes(prm, 0.5, X, V, Um).real) / 2e-6
                for m in range(M):
                    for n in range(N):
                        for blk in range(9):
                            if nb(m, n, blk) != (mm, nn):
                                continue
                            if blk >= 5 and not (0 <= nb(m, n, blk)[0] < M):
                                continue
                            np.testing.assert_allclose(col[:, m, n], JU[:, j, m, n, blk],
                                                       atol=1e-6, rtol=1e-5)
    # mach derivative
    JMa = port.compute_odes_wrt_mach(prm, 0.5, X, V, U)
    fd = (port.compute_odes(prm, 0.5 + 1e-6, X, V, U).real -
          port.compute_odes(prm, 0.5 - 1e-6, X, V, U).real) / 2e-6
    np.testing.assert_allclose(fd, JMa[:, 0], atol=1e-6, rtol=1e-5)


def test_adjoint_identity(port):
    X, V, U = case(17, 9, alpha=0.02, alpha_dot=0.01, noise=5e-3)
    prm = Params(aoa=1.25)
    JU = port.compute_odes_wrt_states(prm, 0.5, X, V, U)
    JX = port.compute_odes_wrt_vertices(prm, 0.5, X, V, U)
    rng = np.random.default_rng(0)
    du, w, dx = crandn(rng, U.shape), crandn(rng, U.shape), crandn(rng, X.shape)
    lhs = np.sum(w * port.apply_odes_wrt_states_fwd(JU, du))
    rhs = np.sum(port.apply_odes_wrt_states_rev(JU, w) * du)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    lhs = np.sum(w * port.apply_odes_wrt_vertices_fwd(JX, dx))
    rhs = np.sum(port.apply_odes_wrt_vertices_rev(JX, w) * dx)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_bsr_export_matches_dense(port):
    X, V, U = case(9, 8, noise=5e-3)
    JU = port.compute_odes_wrt_states(Params(), 0.5, X, V, U)
    data, idx, ptr = port.convert_odes_wrt_states(JU)
    M, N = U.shape[1:]
    assert len(idx) == (9 * M - 6) * N
    dense = np.zeros((4 * M * N, 4 * M * N))
    for row in range(M * N):
        for e in range(ptr[row], ptr[row + 1]):
            dense[4 * row:4 * row + 4, 4 * idx[e]:4 * idx[e] + 4] = data[e]
        assert np.all(np.diff(idx[ptr[row]:ptr[row + 1]]) > 0)
    rng = np.random.default_rng(1)
    d = crandn(rng, U.shape)
    y = dense @ d.reshape(-1, order="F")
    assert rel(y, port.apply_odes_wrt_states_fwd(JU, d).reshape(-1, order="F")) <= 1e-13


def test_directional_complex_step_kink_at_the_wake_cut(port):
    """On an unpitched symmetric grid at zero incidence, rho w is exactly 0 and the radial faces of
    the trailing-edge wake cut lie on the x axis, so the normal velocity q = n.j/rho of those
    faces is exactly 0: a kink of |q| (f90:146).  A complex step along one input component keeps
    Re q exactly 0 there, but one along a direction in every component gives Re q of order h^2
    and either sign, so the directional complex step differs from the per-column one (the
    reference's linearize), whatever h, and only in the cells beside the cut.  Pitching the grid
    removes the exact zero and the two agree.  The directional comparator of the benchmark at
    its unpitched 1025x513 grid shows this difference."""
    rng = np.random.default_rng(1)
    for alpha, expect_kink in ((0.0, True), (0.01, False)):
        X, V, U = case(65, 33, thickness=0.04, mach=0.5, aoa=0.0, alpha=alpha)
        prm = Params(aoa=0.0)
        d = rng.standard_normal(U.shape)
        J = port.compute_odes_wrt_states(prm, 0.5, X, V, U)
        blk = np.real(port.apply_odes_wrt_states_fwd(J, d))
        for h in (1e-12, 1e-30):
            cs = port.cs_jvp(prm, 0.5, X, V, U, dU=d, h=h)
            diff = np.abs(cs - blk).max(axis=0)                       # (M, N)
            bad = set(np.argwhere(diff > 1e-10 * np.abs(blk).max())[:, 1].tolist())
            if expect_kink:
                assert bad and bad <= {0, U.shape[2] - 1}, bad
            else:
                assert not bad and rel(cs, blk) <= 1e-13
