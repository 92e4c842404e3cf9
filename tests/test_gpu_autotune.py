"""GPU tests of the face-kernel autotuner (pyeulerale_amd.kernels.autotune_face /
autotune_multigrid): every tile shape of the residual, states-tangent and states-adjoint kernels
gives the results of the default shape to round-off (a face on a tile's edge is evaluated by
other instructions than an interior face, so the shapes are not bitwise alike); a multi-grid
launch is bitwise equal to single-grid calls with the same shape; and the tuner registers a
shape per grid size that the launches then use."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def storage(a):
    return torch.as_tensor(np.ascontiguousarray(np.swapaxes(np.asarray(a), -1, -2))).cuda()


def say(msg):
    print(f"[autotune] {msg}", flush=True)


def _outputs(K, p, mach, g, X, V, U, d):
    Ub = torch.stack([U, 1.001 * U])
    return (K.compute_odes_batch(p, mach, g, X, V, Ub),
            K.jvp_odes_multi(p, mach, g, X, V, U, dU=d),
            K.vjp_odes_states_multi(p, mach, g, X, V, U, d))


@pytest.mark.parametrize("A,Rv", [(129, 65), (31, 21), (17, 9)])
def test_every_face_tile_gives_the_same_results(A, Rv):
    from pyeulerale_amd import _lib
    from pyeulerale_amd import kernels as K
    X, V, U = case(A, Rv, alpha=0.02, alpha_dot=0.01, seed=5)
    M, N = U.shape[1:]
    g, p = K.Grid(M, N), K.Params(aoa_deg=1.25)
    Xd, Vd, Ud = storage(X), storage(V), storage(U)
    gen = torch.Generator(device="cuda").manual_seed(1)
    d = torch.randn((2, *g.cell_shape), dtype=torch.complex128, device="cuda", generator=gen)
    ref = _outputs(K, p, 0.5, g, Xd, Vd, Ud, d)
    n = _lib.load().pea_num_face_tiles()
    try:
        for kind in range(3):
            for t in range(n):
                if _lib.load().pea_set_face_tile(kind, M, N, t) != 0:
                    say(f"{A}x{Rv}: kind {kind} tile {t} does not fit")
                    continue
                assert K.face_tile(kind, M, N)[0] == t
                got = _outputs(K, p, 0.5, g, Xd, Vd, Ud, d)
                e = float(torch.linalg.vector_norm(got[kind] - ref[kind])
                          / torch.linalg.vector_norm(ref[kind]))
                say(f"{A}x{Rv}: kind {kind} tile {t}: rel {e:.1e}")
                assert e <= 1e-13, (kind, t, e)
            K.set_face_tile(kind, M, N, -1)
    finally:
        for kind in range(3):
            K.set_face_tile(kind, M, N, -1)


def test_autotune_face_registers_a_shape_per_grid():
    from pyeulerale_amd import kernels as K
    from pyeulerale_amd import workloads as W
    c = W.config("C2")
    g, p = K.Grid(c.M, c.N), K.Params(aoa_deg=c.aoa)
    X, V, U = storage(c.X), storage(c.V), storage(c.U[0])
    gen = torch.Generator(device="cuda").manual_seed(2)
    d = torch.randn((2, *g.cell_shape), dtype=torch.complex128, device="cuda", generator=gen)
    ref = _outputs(K, p, c.mach, g, X, V, U, d)
    try:
        res = K.autotune_face(p, c.mach, g, X, V, U, reps=3, batch=2)
        for name, (best, times) in res.items():
            say(f"257x129 {name}: tile {best}, " +
                ", ".join(f"{t}: {ms * 1e3:.1f} us" for t, ms in times.items()))
            kind = K.FACE_KINDS[name]
            assert K.face_tile(kind, c.M, c.N)[0] == best
            assert times[best] == min(times.values())
        got = _outputs(K, p, c.mach, g, X, V, U, d)
        for a, b in zip(got, ref):
            assert float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)) <= 1e-13
    finally:
        for kind in range(3):
            K.set_face_tile(kind, c.M, c.N, -1)


def test_multigrid_tiles_bitwise_and_autotune():
    from pyeulerale_amd import _lib
    from pyeulerale_amd import kernels as K
    from pyeulerale_amd import workloads as W
    fam = W.grid_family(16, seed=3)
    gen = torch.Generator(device="cuda").manual_seed(3)
    items = []
    for m in fam:
        U = storage(m.case.U[0])
        items.append((K.Params(aoa_deg=m.case.aoa), m.case.mach, storage(m.case.X),
                      storage(m.case.V), U,
                      torch.randn(U.shape, dtype=torch.complex128, device="cuda", generator=gen)))
    try:
        for t in range(_lib.load().pea_num_face_tiles()):
            for mode in range(3):
                K.set_face_tile(3 + mode, 0, 0, t)
            # single-grid calls with the same shape
            singles = []
            for p, mach, X, V, U, d in items:
                g = K.Grid(U.shape[2], U.shape[1])
                for kind in range(3):
                    K.set_face_tile(kind, g.M, g.N, t)
                R2 = torch.empty_like(U)
                dR = K.jvp_odes(p, mach, g, X, V, U, dU=d, R=R2)
                singles.append((K.compute_odes(p, mach, g, X, V, U), R2, dR,
                                K.vjp_odes(p, mach, g, X, V, U, d)["states"]))
                for kind in range(3):
                    K.set_face_tile(kind, g.M, g.N, -1)
            say(f"multi-grid tile {t}")
            R = K.MultiGrid([it[:5] for it in items]).run()
            T = K.MultiGrid(items, tangent=True).run()
            A = K.MultiGrid(items, adjoint=True).run()
            for i, s in enumerate(singles):
                assert torch.equal(R[i], s[0]) and torch.equal(T[i][0], s[1])
                assert torch.equal(T[i][1], s[2]) and torch.equal(A[i], s[3]), (t, i)
        for mode in range(3):
            K.set_face_tile(3 + mode, 0, 0, -1)
        res = K.autotune_multigrid(items, reps=2)
        for name, (best, times) in res.items():
            say(f"multi-grid {name}: tile {best}, " +
                ", ".join(f"{t}: {ms * 1e3:.1f} us" for t, ms in times.items()))
            assert K.face_tile(3 + {"residual": 0, "jvp": 1, "vjp": 2}[name])[0] == best
    finally:
        for mode in range(3):
            K.set_face_tile(3 + mode, 0, 0, -1)
