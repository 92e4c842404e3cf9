"""CPU tests that pin the oracle's sampled Jacobians and its restatement of the reference's
second-order rule (core.py:718-1275) before either is used as the checker.

* Masked complex-step Jacobians equal the full ones on the sampled cells and are zero elsewhere.
* The second-order rule (central differences along d_along of complex-step Jacobians, transposed
  onto conj(d_of), core.py:749-758) is checked against a different route to the same number: by
  symmetry of second derivatives, <e, grad_wrt <d_of, dA/d(along) d_along>> equals the central
  difference along e (in the ``wrt`` input) of <conj(d_of), dA/d(along)(wrt + eps e) d_along>,
  which perturbs the other input and applies the forward block products.  A dropped factor 1j,
  a wrong sign, a missing conjugation or a swapped pair of inputs fails it.
* With the compiled reference available, the rule on the reference's own Jacobians agrees with
  the rule on the restatement's.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import case, crandn, rel
from oracle.oracle import Oracle, Params

ODES = [(b, c) for b in ("states", "vertices", "velocities") for c in ("mach", "states", "vertices")]
FORCES = [(b, c) for b in ("states", "vertices") for c in ("states", "vertices")]


@pytest.fixture(scope="module")
def port():
    return Oracle("port")


@pytest.fixture(scope="module")
def small():
    # a state 1e-2 away from the free stream: the dissipation switches (the biased maximum with
    # its 1e-8 margin, f90:111-120) then sit far from their ties, so steps of 1e-7 do not cross
    # them and both difference quotients measure derivatives
    X, V, U = case(14, 11, alpha=0.02, alpha_dot=0.01, seed=3, noise=1e-2)
    return Params(aoa=1.25), 0.5, X, V, U


def test_masked_jacobians_equal_full_on_sampled_cells(port, small):
    prm, mach, X, V, U = small
    M, N = U.shape[1:]
    mask = np.random.default_rng(0).random((M, N)) < 0.3
    for name in ("compute_odes_wrt_states", "compute_odes_wrt_vertices",
                 "compute_odes_wrt_velocities"):
        full = getattr(port, name)(prm, mach, X, V, U)
        part = getattr(port, name)(prm, mach, X, V, U, mask=mask)
        assert np.array_equal(part[:, :, mask], full[:, :, mask]), name
        assert not part[:, :, ~mask].any(), name


def _fwd(port, prm, mach, X, V, U, of, along, d_along):
    """The forward product d(of)/d(along) . d_along with complex-step blocks."""
    M = U.shape[1]
    if of == "odes":
        if along == "states":
            return port.apply_odes_wrt_states_fwd(
                port.compute_odes_wrt_states(prm, mach, X, V, U), d_along)
        J = (port.compute_odes_wrt_vertices if along == "vertices"
             else port.compute_odes_wrt_velocities)(prm, mach, X, V, U)
        return port.apply_odes_wrt_vertices_fwd(J, d_along)
    if along == "states":
        return port.apply_forces_wrt_states_fwd(port.compute_forces_wrt_states(X, U), d_along)
    return port.apply_forces_wrt_vertices_fwd(port.compute_forces_wrt_vertices(X, U), d_along)


def _other_route(port, prm, mach, X, V, U, of, along, wrt, d_of, d_along, e, h=1e-7):
    """Central difference along e in ``wrt`` of <conj(d_of), d(of)/d(along) d_along>."""
    prods = []
    for sgn in (1.0, -1.0):
        m, Xp, Up = mach, X, U
        if wrt == "mach":
            m = mach + sgn * h * e
        elif wrt == "states":
            Up = U + sgn * h * e
        else:
            Xp = X + sgn * h * e
        prods.append(_fwd(port, prm, m, Xp, V, Up, of, along, d_along))
    # difference the products before the inner product: entries that do not depend on the
    # perturbed input (all but the far-field ring for the Mach number) then cancel exactly
    return np.vdot(d_of, prods[0] - prods[1]) / (2.0 * h)


@pytest.mark.parametrize("of,along,wrt", [("odes", b, c) for b, c in ODES] +
                         [("forces", b, c) for b, c in FORCES])
def test_second_order_rule_by_symmetry(port, small, of, along, wrt):
    prm, mach, X, V, U = small
    M, N = U.shape[1:]
    rng = np.random.default_rng(sum(map(ord, of + along + wrt)))
    shp = {"states": (4, M, N), "vertices": (2, M + 1, N + 1), "velocities": (2, M + 1, N + 1)}
    d_of = crandn(rng, (4, M, N) if of == "odes" else (2, N))
    d_along = crandn(rng, shp[along])
    g = port.inner_product_gradient(prm, mach, X, V, U, of, along, wrt, d_of, d_along)
    if wrt == "mach":
        e = 1.0
        lhs = complex(np.asarray(g).reshape(-1)[0])
    else:
        e = rng.standard_normal(shp[wrt])
        lhs = complex(np.sum(g * e))
    rhs = _other_route(port, prm, mach, X, V, U, of, along, wrt, d_of, d_along, e)
    # The rule's own step moves the input by 1e-6 sqrt(1 + |input|) in norm; its truncation
    # error is then ~1e-9 relative along the states but ~4e-6 along the vertices, whose
    # derivatives scale with inverse powers of the 1e-3 wall spacing (the reference's result
    # carries that error too).  Exact zero (forces are linear in the vertices) is compared
    # against the size of the first-order terms.
    tol = 1e-4 if along == "vertices" else 1e-6
    scale = max(abs(rhs), 1e-3 * np.linalg.norm(g) * np.linalg.norm(e),
                1e-4 * np.linalg.norm(d_of) * np.linalg.norm(d_along) * np.linalg.norm(e))
    assert abs(lhs - rhs) <= tol * scale, (lhs, rhs)
    # linear in d_along: the imaginary part enters with the factor 1j
    g_re = port.inner_product_gradient(prm, mach, X, V, U, of, along, wrt, d_of, d_along.real)
    g_im = port.inner_product_gradient(prm, mach, X, V, U, of, along, wrt, d_of, d_along.imag)
    # each call picks its own step from |d_along|, so the parts agree to the rule's O(h^2)
    if of == "forces" and along == wrt == "vertices":
        assert np.linalg.norm(g) <= 1e-6 * np.linalg.norm(d_of) * np.linalg.norm(d_along)
    else:
        assert rel(g_re + 1j * g_im, g) <= tol


@pytest.mark.skipif(not Oracle.available("ref"), reason="oracle/_ref not built")
@pytest.mark.parametrize("of,along,wrt", [("odes", "states", "states"),
                                          ("odes", "vertices", "mach"),
                                          ("odes", "velocities", "vertices"),
                                          ("forces", "vertices", "states")])
def test_second_order_rule_on_reference_jacobians(port, small, of, along, wrt):
    prm, mach, X, V, U = small
    M, N = U.shape[1:]
    rng = np.random.default_rng(7)
    shp = {"states": (4, M, N), "vertices": (2, M + 1, N + 1), "velocities": (2, M + 1, N + 1)}
    d_of = crandn(rng, (4, M, N) if of == "odes" else (2, N))
    d_along = crandn(rng, shp[along])
    a = port.inner_product_gradient(prm, mach, X, V, U, of, along, wrt, d_of, d_along)
    b = Oracle("ref").inner_product_gradient(prm, mach, X, V, U, of, along, wrt, d_of, d_along)
    assert rel(a, b) <= 1e-8


def test_masked_second_order_exact_inside_the_sample(port, small):
    """With a mask that covers the stencil of the checked cells the sampled rule equals the
    full rule there (what the large-grid checks rely on)."""
    prm, mach, X, V, U = small
    M, N = U.shape[1:]
    rng = np.random.default_rng(11)
    d_of, d_along = crandn(rng, (4, M, N)), crandn(rng, (4, M, N))
    mask = np.zeros((M, N), bool)
    mask[:5, :] = True                       # rings 0..4 cover the stencils of rings 0..2
    full = port.inner_product_gradient(prm, mach, X, V, U, "odes", "states", "states",
                                       d_of, d_along)
    part = port.inner_product_gradient(prm, mach, X, V, U, "odes", "states", "states",
                                       d_of, d_along, mask=mask)
    assert np.array_equal(part[:, :3, :], full[:, :3, :])
