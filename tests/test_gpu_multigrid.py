"""GPU tests of the multi-grid launch: many O-grids of uneven size (different thickness, size,
Mach, angle of attack, JST constants and pitch motion) in one launch give bitwise the
results of one single-grid call per grid, for the residual, the states tangent and the states
adjoint."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def storage(a):
    return torch.as_tensor(np.ascontiguousarray(np.swapaxes(a, -1, -2))).cuda()


@pytest.fixture(scope="module")
def grids():
    from pyeulerale_amd import kernels as K
    rng = np.random.default_rng(7)
    items = []
    sizes = [(9, 6), (33, 17), (65, 33), (129, 65), (41, 9), (17, 129), (257, 33), (97, 49),
             (11, 5), (65, 65), (129, 17), (201, 101)]
    for A, Rv in sizes:
        mach = float(rng.uniform(0.3, 0.6))
        aoa = float(rng.uniform(-2.0, 3.0))
        amp = np.deg2rad(float(rng.uniform(0.0, 1.0)))
        X, V, U = case(A, Rv, thickness=float(rng.uniform(0.04, 0.18)), mach=mach, aoa=aoa,
                       alpha=amp * 0.3, alpha_dot=amp * 0.2, seed=int(rng.integers(1000)))
        p = K.Params(aoa_deg=aoa, k2=float(rng.uniform(0.5, 1.0)), k4=float(rng.uniform(0.02, 0.04)),
                     c4=2.0)
        items.append((p, mach, storage(X), storage(V), storage(U)))
    return items


def test_multigrid_residual_is_bitwise_single(grids):
    from pyeulerale_amd import kernels as K
    print(f"[multigrid] residual, {len(grids)} grids", flush=True)
    Rs = K.compute_odes_multi(grids)
    for (p, mach, X, V, U), R in zip(grids, Rs):
        g = K.Grid(U.shape[2], U.shape[1])
        ref = K.compute_odes(p, mach, g, X, V, U)
        assert torch.equal(R, ref), (g.M, g.N, float((R - ref).abs().max()))
    # a second call reuses nothing stale
    Rs2 = K.compute_odes_multi(grids[::-1])
    for R, R2 in zip(Rs, Rs2[::-1]):
        assert torch.equal(R, R2)


def test_multigrid_tangent_is_bitwise_single(grids):
    from pyeulerale_amd import kernels as K
    print(f"[multigrid] tangent, {len(grids)} grids", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(3)
    items = []
    for (p, mach, X, V, U) in grids:
        dU = torch.randn(U.shape, dtype=torch.complex128, device="cuda", generator=gen)
        items.append((p, mach, X, V, U, dU))
    outs = K.jvp_odes_multigrid(items)
    for (p, mach, X, V, U, dU), (R, dR) in zip(items, outs):
        g = K.Grid(U.shape[2], U.shape[1])
        R1 = torch.empty_like(U)
        ref = K.jvp_odes(p, mach, g, X, V, U, dU=dU, R=R1)
        assert torch.equal(dR, ref), (g.M, g.N)
        assert torch.equal(R, R1)


def test_multigrid_adjoint_is_bitwise_single(grids):
    from pyeulerale_amd import kernels as K
    from pyeulerale_amd._lib import PeaError
    wide = [it for it in grids if it[4].shape[1] >= 18]   # the adjoint tile needs N >= 18
    print(f"[multigrid] adjoint, {len(wide)} grids", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(5)
    items = [(p, mach, X, V, U, torch.randn(U.shape, dtype=torch.complex128, device="cuda",
                                            generator=gen)) for (p, mach, X, V, U) in wide]
    outs = K.vjp_odes_multigrid(items)
    for (p, mach, X, V, U, w), gU in zip(items, outs):
        g = K.Grid(U.shape[2], U.shape[1])
        ref = K.vjp_odes(p, mach, g, X, V, U, w)["states"]
        assert torch.equal(gU, ref), (g.M, g.N)
    narrow = next(it for it in grids if it[4].shape[1] < 18)
    w = torch.zeros(narrow[4].shape, dtype=torch.complex128, device="cuda")
    with pytest.raises(PeaError):
        K.vjp_odes_multigrid([narrow + (w,)])


def test_multigrid_rejects_bad_descriptors(grids):
    from pyeulerale_amd import kernels as K
    from pyeulerale_amd._lib import PeaError
    p, mach, X, V, U = grids[0]
    bad = torch.empty((4, 4, 3), dtype=torch.float64, device="cuda")   # M = 3 < 4
    with pytest.raises((PeaError, RuntimeError)):
        K.compute_odes_multi([(p, mach, X[:, :5, :4].contiguous(), V[:, :5, :4].contiguous(),
                               bad)])
