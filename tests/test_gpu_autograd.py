"""GPU tests of the torch.autograd layer: first derivatives against the explicit adjoint and
tangent kernels and against torch.autograd.gradcheck (central differences), second
derivatives (double backward = tangent-over-adjoint) against gradgradcheck and against central
differences of the first derivatives."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def storage(a):
    return torch.as_tensor(np.ascontiguousarray(np.swapaxes(a, -1, -2))).cuda()


@pytest.fixture(scope="module")
def tiny():
    """9 x 6 vertices: 5 radial x 8 angular cells with a wall spacing of 0.05 chord, so that the
    derivatives stay O(1-100) and torch's central differences (eps 1e-6) resolve them; pitching,
    perturbed free stream."""
    from helpers import free_stream
    from pyeulerale_amd.grid import naca_ogrid, pitch
    X0 = naca_ogrid(9, 6, 0.12, far_field=4.0, first_spacing=0.05)
    X, V = pitch(X0, 0.02, 0.01)
    rng = np.random.default_rng(0)
    U = np.empty((4, 5, 8))
    U[:] = free_stream(0.5, 1.25)[:, None, None]
    U *= 1.0 + 1e-2 * rng.standard_normal(U.shape)
    return storage(X), storage(V), storage(U)


@pytest.fixture(scope="module")
def small():
    X, V, U = case(33, 17, thickness=0.12, mach=0.5, aoa=1.25, alpha=0.01, alpha_dot=0.004,
                   noise=1e-3)
    return storage(X), storage(V), storage(U)


def test_backward_matches_adjoint_kernel(small):
    from pyeulerale_amd import autograd as A
    from pyeulerale_amd import kernels as K
    X, V, U = small
    print("[autograd] first order vs vjp kernel", flush=True)
    g = K.Grid(U.shape[2], U.shape[1])
    p = K.Params(aoa_deg=1.25)
    Ur, Xr, Vr = (t.clone().requires_grad_() for t in (U, X, V))
    mach = torch.tensor(0.5, dtype=torch.float64, device="cuda", requires_grad=True)
    R = A.odes(Ur, Xr, Vr, mach, p)
    torch.testing.assert_close(R, K.compute_odes(p, 0.5, g, X, V, U), rtol=0, atol=0)
    w = torch.randn_like(R)
    gm, gU, gX, gV = torch.autograd.grad(R, (mach, Ur, Xr, Vr), w)
    ref = K.vjp_odes(p, 0.5, g, X, V, U, w.to(torch.complex128),
                     want=("mach", "states", "vertices", "velocities"))
    torch.testing.assert_close(gU, ref["states"].real, rtol=1e-14, atol=0)
    torch.testing.assert_close(gX, ref["vertices"].real, rtol=1e-14, atol=0)
    torch.testing.assert_close(gV, ref["velocities"].real, rtol=1e-14, atol=0)
    assert abs(float(gm) - float(ref["mach"].real[0])) <= 1e-12 * abs(float(gm))
    F = A.forces(Ur, Xr)
    wF = torch.randn_like(F)
    fU, fX = torch.autograd.grad(F, (Ur, Xr), wF)
    refF = K.vjp_forces(g, X, U, wF.to(torch.complex128))
    torch.testing.assert_close(fU, refF["states"].real, rtol=1e-14, atol=0)
    torch.testing.assert_close(fX, refF["vertices"].real, rtol=1e-14, atol=0)


def test_gradcheck_first_and_second_order(tiny):
    """torch's own central-difference checks (gradcheck / gradgradcheck) on a tiny grid."""
    from pyeulerale_amd import autograd as A
    from pyeulerale_amd import kernels as K
    X, V, U = tiny
    p = K.Params(aoa_deg=1.25)
    print("[autograd] gradcheck / gradgradcheck 9x6", flush=True)
    Ur, Xr, Vr = (t.clone().requires_grad_() for t in (U, X, V))
    mach = torch.tensor(0.5, dtype=torch.float64, device="cuda", requires_grad=True)
    fn = lambda m, u, x, v: A.odes(u, x, v, m, p)
    assert torch.autograd.gradcheck(fn, (mach, Ur, Xr, Vr), eps=1e-6, atol=1e-6, rtol=1e-5)
    assert torch.autograd.gradgradcheck(fn, (mach, Ur, Xr, Vr), eps=1e-6, atol=1e-5, rtol=1e-4)
    fnF = lambda u, x: A.forces(u, x)
    assert torch.autograd.gradcheck(fnF, (Ur, Xr), eps=1e-6, atol=1e-6, rtol=1e-5)
    assert torch.autograd.gradgradcheck(fnF, (Ur, Xr), eps=1e-6, atol=1e-5, rtol=1e-4)


def test_double_backward_is_second_order_gradient(small):
    """grad_z <a, grad_U <w, R>> from double backward equals the central difference of the
    first-order gradient along a.  The step moves U by ~1e-9 relative: the reference's biased
    maximum switches across a band of width 1e-8 (f90:111-120), so larger steps can straddle a
    switch of a near-tied face and measure the jump instead of the derivative."""
    from pyeulerale_amd import autograd as A
    from pyeulerale_amd import kernels as K
    X, V, U = small
    p = K.Params(aoa_deg=1.25)
    print("[autograd] double backward vs FD", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(0)
    w = torch.randn(U.shape, dtype=torch.float64, device="cuda", generator=gen)
    a = torch.randn(U.shape, dtype=torch.float64, device="cuda", generator=gen) * 1e-3 * U.abs()
    dX = torch.randn(X.shape, dtype=torch.float64, device="cuda", generator=gen) * 1e-3
    Ur, Xr = U.clone().requires_grad_(), X.clone().requires_grad_()
    R = A.odes(Ur, Xr, V, 0.5, p)
    (gU,) = torch.autograd.grad(R, (Ur,), w, create_graph=True)
    hU, hX = torch.autograd.grad(gU, (Ur, Xr), a)

    def first(u, x):
        uu = u.clone().requires_grad_()
        (g,) = torch.autograd.grad(A.odes(uu, x, V, 0.5, p), (uu,), w)
        return g
    h = 1e-6
    # d/de <a, grad_U(U + e dU)> for dU = a is <a, H a>; compare the full vectors instead:
    # H_UU a = d/de grad_U(U + e a), H_XU^T a = grad_X <a, grad_U>  checked through dX
    fd_U = (first(U + h * a, X) - first(U - h * a, X)) / (2 * h)
    rel = float(torch.linalg.norm(hU - fd_U) / torch.linalg.norm(fd_U))
    assert rel <= 1e-6, rel
    fd_X = float(((first(U, X + h * dX) - first(U, X - h * dX)) / (2 * h) * a).sum())
    an_X = float((hX * dX).sum())
    assert abs(an_X - fd_X) <= 1e-6 * abs(fd_X), (an_X, fd_X)
