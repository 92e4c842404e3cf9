"""GPU tests of the many-direction states tangent (pea_jvp_odes_states_dirs, kernels.jvp_odes_dirs):
the primal part of each tile is computed once and the linearised face fluxes are applied to every
direction.  Checked against the Dual-number tangent (pea_jvp_odes_multi, one direction per grid z)
to round-off, against the oracle's complex-step blocks (the reference's Jacobian,
spatial_discretization.f90:552-1143, applied as apply_odes_wrt_states_fwd, f90:825-864), across
grids with partial tiles, wall and far-field rows inside one tile, the seam, and direction counts
that split over workgroups and over launches."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import case, crandn, dev_cells, dev_verts, host, rel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL_DUAL = 1e-12     # against the Dual-number tangent: the same products, another rounding
TOL_JAC = 1e-10      # against the oracle's complex-step blocks

CASES = {
    # name: (angular vertices, radial vertices, kwargs); M = Rv - 1, N = A - 1
    "tiny": (9, 8, dict()),                                       # 7 x 8: one partial tile
    "narrow": (19, 9, dict(alpha=0.02, alpha_dot=0.01)),          # 8 x 18: seam, wall + far rows
    "ragged": (31, 22, dict(alpha=0.03, alpha_dot=0.02, noise=1e-2)),   # 21 x 30: partial tiles
    "small_pitch": (33, 17, dict(aoa=0.0, alpha=0.01, alpha_dot=0.004, thickness=0.04)),
    "c1": (129, 65, dict()),
}


def say(msg):
    print(f"[dirs] {msg}", flush=True)


def setup(name, k, seed=3):
    from pyeulerale_amd import kernels as K
    A, Rv, kw = CASES[name]
    kw = dict(kw)
    aoa = kw.get("aoa", 1.25)
    X, V, U = case(A, Rv, **kw)
    g, p = K.Grid(U.shape[1], U.shape[2]), K.Params(aoa_deg=aoa)
    rng = np.random.default_rng(seed)
    dU = [crandn(rng, U.shape) for _ in range(k)]
    say(f"{name}: M={g.M} N={g.N} k={k}")
    return K, g, p, X, V, U, dev_verts(X), dev_verts(V), dev_cells(U), dU


@pytest.mark.parametrize("name", list(CASES))
def test_against_dual_tangent_and_oracle(name):
    from oracle.oracle import Oracle, Params as OParams
    k = 3
    K, g, p, X, V, U, Xd, Vd, Ud, dU = setup(name, k)
    dUd = torch.stack([dev_cells(d) for d in dU])
    got = K.jvp_odes_dirs(p, 0.5, g, Xd, Vd, Ud, dUd)
    ref = K.jvp_odes_multi(p, 0.5, g, Xd, Vd, Ud, dU=dUd)
    for i in range(k):
        e = float(torch.linalg.vector_norm(got[i] - ref[i]) / torch.linalg.vector_norm(ref[i]))
        say(f"{name} direction {i}: rel to the Dual tangent {e:.1e}")
        assert e <= TOL_DUAL, (i, e)
    if g.M * g.N <= 1024:
        orc = Oracle("port")
        aoa = CASES[name][2].get("aoa", 1.25)
        J = orc.compute_odes_wrt_states(OParams(aoa=aoa), 0.5, X, V, U)
        for i in range(k):
            assert rel(host(got[i]), orc.apply_odes_wrt_states_fwd(J, dU[i])) <= TOL_JAC


def test_direction_chunks_are_bitwise_alike():
    """Every split of the directions over workgroups gives the same bits: a direction's
    arithmetic does not depend on which workgroup applies it."""
    k = 5
    K, g, p, X, V, U, Xd, Vd, Ud, dU = setup("ragged", k)
    dUd = torch.stack([dev_cells(d) for d in dU])
    base = K.jvp_odes_dirs(p, 0.5, g, Xd, Vd, Ud, dUd, per=k)
    for per in (0, 1, 2, 3, 4, 6):
        assert torch.equal(K.jvp_odes_dirs(p, 0.5, g, Xd, Vd, Ud, dUd, per=per), base), per
    # one direction alone equals the same direction within a batch
    one = K.jvp_odes_dirs(p, 0.5, g, Xd, Vd, Ud, dUd[1:2].contiguous())
    assert torch.equal(one[0], base[1])


def test_257x129_against_dual_tangent():
    from pyeulerale_amd import kernels as K
    from pyeulerale_amd import workloads as W
    c = W.config("C2")
    g, p = K.Grid(c.M, c.N), K.Params(aoa_deg=c.aoa)
    X, V, U = dev_verts(c.X), dev_verts(c.V), dev_cells(c.U[0])
    gen = torch.Generator(device="cuda").manual_seed(4)
    d = torch.randn((8, *g.cell_shape), dtype=torch.complex128, device="cuda", generator=gen)
    got = K.jvp_odes_dirs(p, c.mach, g, X, V, U, d)
    ref = K.jvp_odes_multi(p, c.mach, g, X, V, U, dU=d)
    e = float(torch.linalg.vector_norm(got - ref) / torch.linalg.vector_norm(ref))
    say(f"257x129 pitching, 8 directions: rel to the Dual tangent {e:.1e}")
    assert e <= TOL_DUAL


def test_1025x513_over_two_launches():
    """65 directions at 1025x513: the descriptors' 2^31-byte bound allows 63 per launch, so the
    call makes two launches; directions on both sides of the split match the Dual tangent."""
    from pyeulerale_amd import kernels as K
    from pyeulerale_amd import workloads as W
    c = W.config("C5")
    g, p = K.Grid(c.M, c.N), K.Params(aoa_deg=c.aoa)
    X, V, U = dev_verts(c.X), dev_verts(c.V), dev_cells(c.U[0])
    gen = torch.Generator(device="cuda").manual_seed(5)
    d = torch.randn((65, *g.cell_shape), dtype=torch.complex128, device="cuda", generator=gen)
    got = K.jvp_odes_dirs(p, c.mach, g, X, V, U, d)
    pick = [0, 31, 62, 63, 64]
    ref = K.jvp_odes_multi(p, c.mach, g, X, V, U, dU=d[pick].contiguous())
    for i, j in enumerate(pick):
        e = float(torch.linalg.vector_norm(got[j] - ref[i]) / torch.linalg.vector_norm(ref[i]))
        say(f"1025x513 direction {j}: rel to the Dual tangent {e:.1e}")
        assert e <= TOL_DUAL, (j, e)
