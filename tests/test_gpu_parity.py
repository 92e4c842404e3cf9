"""GPU parity: the HIP library (through its C ABI) against the CPU oracle on identical inputs.

Bars (from the path's specification): R and F <= 1e-12 relative, Jacobian products <= 1e-10
relative, norm-wise.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import case, crandn, dev_cells, dev_verts, host, rel
from oracle.oracle import Oracle, Params as OParams

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL_RF = 1e-12
TOL_JAC = 1e-10

CASES = {
    # name: (angular vertices, radial vertices, kwargs)
    "tiny": (9, 8, dict()),
    "small_pitch": (33, 17, dict(aoa=0.0, alpha=0.01, alpha_dot=0.004, thickness=0.04)),
    "c1": (129, 65, dict()),
    # the face kernels at their limits: N = 18 angular cells (the adjoint tile's width + 3, so
    # its grown tile wraps onto itself) and M = 8 < a tile's rows, so one tile holds the wall
    # and far-field rows and the last tile column is partial and wraps the seam
    "narrow": (19, 9, dict(alpha=0.02, alpha_dot=0.01)),
}


@pytest.fixture(scope="module")
def orc():
    return Oracle("port")


@pytest.fixture(scope="module")
def K():
    from pyeulerale_amd import kernels
    return kernels


def setup_case(name):
    A, Rv, kw = CASES[name]
    kw = dict(kw)
    mach = kw.pop("mach", 0.5)
    aoa = kw.get("aoa", 1.25)
    X, V, U = case(A, Rv, mach=mach, **kw)
    print(f"[case {name}] A={A} Rv={Rv} mach={mach} aoa={aoa}", flush=True)
    return X, V, U, mach, aoa


def _dev(K, X, V, U, aoa):
    from pyeulerale_amd.kernels import Grid, Params
    g = Grid(U.shape[1], U.shape[2])
    return g, Params(aoa_deg=aoa), dev_verts(X), dev_verts(V), dev_cells(U)


@pytest.mark.parametrize("name", list(CASES))
def test_residual_and_forces(name, orc, K):
    X, V, U, mach, aoa = setup_case(name)
    g, p, Xd, Vd, Ud = _dev(K, X, V, U, aoa)
    R = host(K.compute_odes(p, mach, g, Xd, Vd, Ud))
    Rref = orc.compute_odes(OParams(aoa=aoa), mach, X, V, U).real
    assert rel(R, Rref) <= TOL_RF, rel(R, Rref)
    coef = torch.empty(3, dtype=torch.float64, device="cuda")
    F = K.compute_forces(g, Xd, Ud, center=(0.25, 0.0), coef=coef).cpu().numpy()
    Fref = orc.compute_forces(OParams(aoa=aoa), X, U).real
    assert rel(F, Fref) <= TOL_RF
    # reductions behind c_d, c_l, c_m (core.py:1277-1363)
    xp = 0.5 * (X[:, 0, :-1] + X[:, 0, 1:])
    ref = [Fref[0].sum(), Fref[1].sum(),
           ((xp[1] - 0.0) * Fref[0] - (xp[0] - 0.25) * Fref[1]).sum()]
    np.testing.assert_allclose(coef.cpu().numpy(), ref, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("name", list(CASES))
def test_free_stream(name, K):
    X, V, U, mach, aoa = setup_case(name)
    g, p, Xd, Vd, Ud = _dev(K, X, V, U, aoa)
    K.set_free_stream_state(p, mach, g, Ud)
    from helpers import free_stream
    np.testing.assert_array_equal(host(Ud)[:, 3, 4], free_stream(mach, aoa))


def _linearization(orc, X, V, U, mach, aoa):
    op = OParams(aoa=aoa)
    return dict(
        U=orc.compute_odes_wrt_states(op, mach, X, V, U),
        X=orc.compute_odes_wrt_vertices(op, mach, X, V, U),
        V=orc.compute_odes_wrt_velocities(op, mach, X, V, U),
        Ma=orc.compute_odes_wrt_mach(op, mach, X, V, U),
    )


@pytest.mark.parametrize("name", ["tiny", "small_pitch", "c1", "narrow"])
def test_jvp_vjp_against_complex_step_blocks(name, orc, K):
    X, V, U, mach, aoa = setup_case(name)
    g, p, Xd, Vd, Ud = _dev(K, X, V, U, aoa)
    J = _linearization(orc, X, V, U, mach, aoa)
    rng = np.random.default_rng(7)
    dU, dX, dV = crandn(rng, U.shape), crandn(rng, X.shape), crandn(rng, X.shape)
    dMa = complex(rng.standard_normal(), rng.standard_normal())
    # tangents one input at a time and all together
    parts = {
        "U": orc.apply_odes_wrt_states_fwd(J["U"], dU),
        "X": orc.apply_odes_wrt_vertices_fwd(J["X"], dX),
        "V": orc.apply_odes_wrt_vertices_fwd(J["V"], dV),
        "Ma": orc.apply_odes_wrt_mach_fwd(J["Ma"], np.array([dMa])),
    }
    dev = {"U": dict(dU=dev_cells(dU)), "X": dict(dX=dev_verts(dX)),
           "V": dict(dV=dev_verts(dV)), "Ma": dict(dmach=dMa)}
    for key in parts:
        got = host(K.jvp_odes(p, mach, g, Xd, Vd, Ud, **dev[key]))
        assert rel(got, parts[key]) <= TOL_JAC, (key, rel(got, parts[key]))
    allin = {k: v for d in dev.values() for k, v in d.items()}
    Rout = torch.empty_like(Ud)
    got = host(K.jvp_odes(p, mach, g, Xd, Vd, Ud, R=Rout, **allin))
    assert rel(got, sum(parts.values())) <= TOL_JAC
    Rref = orc.compute_odes(OParams(aoa=aoa), mach, X, V, U).real
    assert rel(host(Rout), Rref) <= TOL_RF

    # adjoints
    w = crandn(rng, U.shape)
    out = K.vjp_odes(p, mach, g, Xd, Vd, Ud, dev_cells(w),
                     want=("states", "vertices", "velocities", "mach"))
    ref = {
        "states": orc.apply_odes_wrt_states_rev(J["U"], w),
        "vertices": orc.apply_odes_wrt_vertices_rev(J["X"], w),
        "velocities": orc.apply_odes_wrt_vertices_rev(J["V"], w),
        "mach": orc.apply_odes_wrt_mach_rev(J["Ma"], w),
    }
    for key in ref:
        got = out[key].cpu().numpy() if key == "mach" else host(out[key])
        assert rel(got, ref[key]) <= TOL_JAC, (key, rel(got, ref[key]))
    # the states-only adjoint (face-based kernel) agrees with the all-outputs kernel and is
    # bitwise reproducible (fixed accumulation order, no atomics)
    only = K.vjp_odes(p, mach, g, Xd, Vd, Ud, dev_cells(w), want=("states",))
    assert rel(host(only["states"]), ref["states"]) <= TOL_JAC
    assert rel(host(only["states"]), host(out["states"])) <= 1e-13
    again = K.vjp_odes(p, mach, g, Xd, Vd, Ud, dev_cells(w), want=("states",))
    assert torch.equal(only["states"], again["states"])


@pytest.mark.parametrize("name", ["tiny", "small_pitch", "narrow"])
def test_jvp_multi_directions(name, orc, K):
    """k directions per call.  States-only directions (one batched launch of the face kernel,
    k = 5) are bitwise equal to one pea_jvp_odes per direction and match the oracle's
    complex-step blocks; mixed-input directions (k = 2, the cell-based path) match single
    products to round-off; k adjoint seeds in one launch equal one pea_vjp_odes per seed."""
    X, V, U, mach, aoa = setup_case(name)
    g, p, Xd, Vd, Ud = _dev(K, X, V, U, aoa)
    J = _linearization(orc, X, V, U, mach, aoa)
    rng = np.random.default_rng(11)
    k = 5
    dUs = [crandn(rng, U.shape) for _ in range(k)]
    dUd = torch.stack([dev_cells(d) for d in dUs])
    out = K.jvp_odes_multi(p, mach, g, Xd, Vd, Ud, dU=dUd)
    for i in range(k):
        one = K.jvp_odes(p, mach, g, Xd, Vd, Ud, dU=dUd[i].contiguous())
        assert torch.equal(out[i], one), i
        ref = orc.apply_odes_wrt_states_fwd(J["U"], dUs[i])
        assert rel(host(out[i]), ref) <= TOL_JAC
    dXs = torch.stack([dev_verts(crandn(rng, X.shape)) for _ in range(2)])
    out2 = K.jvp_odes_multi(p, mach, g, Xd, Vd, Ud, dU=dUd[:2].contiguous(), dX=dXs)
    for i in range(2):
        one = K.jvp_odes(p, mach, g, Xd, Vd, Ud, dU=dUd[i].contiguous(), dX=dXs[i].contiguous())
        assert rel(host(out2[i]), host(one)) <= 1e-13
    # the directional complex step, k directions per launch, equals one launch per direction
    cs = K.cs_jvp_odes_multi(p, mach, g, Xd, Vd, Ud, dUd)
    for i in range(k):
        assert torch.equal(cs[i], K.cs_jvp_odes(p, mach, g, Xd, Vd, Ud, dU=dUd[i].contiguous()))
    # k seeds of the states adjoint: one launch of the face adjoint for N >= 18, one adjoint
    # per seed below (tiny: N = 8)
    gs = K.vjp_odes_states_multi(p, mach, g, Xd, Vd, Ud, dUd)
    for i in range(k):
        one = K.vjp_odes(p, mach, g, Xd, Vd, Ud, dUd[i].contiguous(), want=("states",))["states"]
        assert torch.equal(gs[i], one), i
        ref = orc.apply_odes_wrt_states_rev(J["U"], dUs[i])
        assert rel(host(gs[i]), ref) <= TOL_JAC


@pytest.mark.parametrize("name", ["tiny", "small_pitch"])
def test_complex_step_path(name, orc, K):
    X, V, U, mach, aoa = setup_case(name)
    g, p, Xd, Vd, Ud = _dev(K, X, V, U, aoa)
    J = _linearization(orc, X, V, U, mach, aoa)
    JU, JX, JV, JMa = K.cs_linearize_odes(p, mach, g, Xd, Vd, Ud)
    # storage (9,4,4,N,M) -> reference (4,4,M,N,9)
    JUr = JU.permute(1, 2, 4, 3, 0).cpu().numpy()
    JXr = JX.permute(1, 2, 4, 3, 0).cpu().numpy()
    JVr = JV.permute(1, 2, 4, 3, 0).cpu().numpy()
    JMar = host(JMa)[:, None]
    # absent radial blocks are zero in both (the oracle's arrays start at zero)
    for a, b, key in ((JUr, J["U"], "U"), (JXr, J["X"], "X"), (JVr, J["V"], "V"),
                      (JMar, J["Ma"], "Ma")):
        assert rel(a, b) <= TOL_JAC, (key, rel(a, b))
    rng = np.random.default_rng(3)
    dU, dX = crandn(rng, U.shape), crandn(rng, X.shape)
    got = host(K.apply_states_fwd(g, JU, dev_cells(dU)))
    assert rel(got, orc.apply_odes_wrt_states_fwd(J["U"], dU)) <= TOL_JAC
    got = host(K.apply_states_rev(g, JU, dev_cells(dU)))
    assert rel(got, orc.apply_odes_wrt_states_rev(J["U"], dU)) <= TOL_JAC
    got = host(K.apply_vertices_fwd(g, JX, dev_verts(dX)))
    assert rel(got, orc.apply_odes_wrt_vertices_fwd(J["X"], dX)) <= TOL_JAC
    got = host(K.apply_vertices_rev(g, JX, dev_cells(dU)))
    assert rel(got, orc.apply_odes_wrt_vertices_rev(J["X"], dU)) <= TOL_JAC
    # matrix-free directional complex step
    got = host(K.cs_jvp_odes(p, mach, g, Xd, Vd, Ud, dU=dev_cells(dU), dX=dev_verts(dX)))
    ref = orc.apply_odes_wrt_states_fwd(J["U"], dU) + orc.apply_odes_wrt_vertices_fwd(J["X"], dX)
    assert rel(got, ref) <= TOL_JAC
    # O(nnz) BSR export
    data, idx, ptr = K.export_bsr(g, JU)
    d0, i0, p0 = orc.convert_odes_wrt_states(J["U"])
    np.testing.assert_array_equal(idx.cpu().numpy(), i0)
    np.testing.assert_array_equal(ptr.cpu().numpy(), p0)
    assert rel(data.cpu().numpy(), d0) <= TOL_JAC


@pytest.mark.parametrize("name", ["tiny", "small_pitch"])
def test_forces_products(name, orc, K):
    X, V, U, mach, aoa = setup_case(name)
    g, p, Xd, Vd, Ud = _dev(K, X, V, U, aoa)
    JFU = orc.compute_forces_wrt_states(X, U)
    JFX = orc.compute_forces_wrt_vertices(X, U)
    rng = np.random.default_rng(5)
    dU, dX = crandn(rng, U.shape), crandn(rng, X.shape)
    got = K.jvp_forces(g, Xd, Ud, dX=dev_verts(dX), dU=dev_cells(dU)).cpu().numpy()
    ref = orc.apply_forces_wrt_states_fwd(JFU, dU) + orc.apply_forces_wrt_vertices_fwd(JFX, dX)
    assert rel(got, ref) <= TOL_JAC
    wF = crandn(rng, (2, U.shape[2]))
    out = K.vjp_forces(g, Xd, Ud, torch.as_tensor(wF).cuda())
    M = U.shape[1]
    assert rel(host(out["states"]), orc.apply_forces_wrt_states_rev(JFU, wF, M)) <= TOL_JAC
    assert rel(host(out["vertices"]), orc.apply_forces_wrt_vertices_rev(JFX, wF, M)) <= TOL_JAC


@pytest.mark.parametrize("name", ["tiny", "small_pitch"])
def test_hvp_against_central_differences_of_vjp(name, K):
    """Tangent-over-adjoint vs central differences of the adjoint along the direction."""
    X, V, U, mach, aoa = setup_case(name)
    g, p, Xd, Vd, Ud = _dev(K, X, V, U, aoa)
    rng = np.random.default_rng(11)
    w = dev_cells(crandn(rng, U.shape))
    d = dict(dU=crandn(rng, U.shape) * 1e-3 * np.abs(U), dX=crandn(rng, X.shape) * 1e-3,
             dV=crandn(rng, X.shape) * 1e-3, dmach=complex(0.3, -0.2))
    want = ("states", "vertices", "velocities", "mach")
    H = K.hvp_odes(p, mach, g, Xd, Vd, Ud, w, dmach=d["dmach"], dX=dev_verts(d["dX"]),
                   dV=dev_verts(d["dV"]), dU=dev_cells(d["dU"]), want=want)
    # bilinear: H(w, a + ib) = H(w, a) + i H(w, b); difference each real part separately
    h = 1e-6
    fd = {k: 0 for k in want}
    for part, fac in ((np.real, 1.0), (np.imag, 1j)):
        outs = []
        for sgn in (1.0, -1.0):
            Xs = dev_verts(X + sgn * h * part(d["dX"]))
            Vs = dev_verts(V + sgn * h * part(d["dV"]))
            Us = dev_cells(U + sgn * h * part(d["dU"]))
            ms = mach + sgn * h * part(d["dmach"])
            outs.append(K.vjp_odes(p, ms, g, Xs, Vs, Us, w, want=want))
        for k in want:
            diff = (outs[0][k] - outs[1][k]) / (2 * h)
            fd[k] = fd[k] + fac * (diff.cpu().numpy() if k == "mach" else host(diff))
    for k in want:
        got = H[k].cpu().numpy() if k == "mach" else host(H[k])
        assert rel(got, fd[k]) <= 1e-6, (k, rel(got, fd[k]))


def test_coefficients_and_pressure_tangent():
    """The coefficient operators (core.py:1277-1488) through the identities their definitions fix.
    c_l, c_m and c_d are linear in the forces, so grad . F reproduces them, except that c_d's
    gradient carries the 1/L its value lacks (the reference's quirk, core.py:1286 and 1289-1311).
    With the forces held, c_m is linear in the surface vertices, so its vertex gradient predicts a
    vertex move exactly.  The c_p tangent matches central differences of c_p."""
    from pyeulerale_amd.core import SpatialDiscretization
    X, V, U, mach, aoa = setup_case("small_pitch")
    L = 2.5
    sd = SpatialDiscretization(vertices=X, angle_of_attack=aoa, mach_number=mach,
                               coefficient_length=L, coefficient_center=(0.25, 0.05))
    sd.states[:] = torch.as_tensor(U).cuda()
    sd.compute_forces()
    F = sd.forces.to(torch.complex128)
    cl, cd, cm = (sd.compute_lift_coefficient(), sd.compute_drag_coefficient(),
                  sd.compute_moment_coefficient())
    def check(grad, value):   # grad . F == value, to round-off of the summed terms
        terms = grad * F
        assert abs(float(terms.sum().real) - value) <= 1e-13 * float(terms.abs().sum())
    check(sd.compute_lift_coefficient_wrt_forces(), cl)
    check(sd.compute_moment_coefficient_wrt_forces(), cm)
    check(sd.compute_drag_coefficient_wrt_forces(), cd / L)
    rng = np.random.default_rng(3)
    dX = np.zeros_like(X)
    dX[:, 0, :] = 1e-3 * rng.standard_normal(dX[:, 0, :].shape)
    dX[:, 0, -1] = dX[:, 0, 0]                     # the seam vertex moves with its twin
    pred = (sd.compute_moment_coefficient_wrt_vertices() * torch.as_tensor(dX).cuda()).sum().real
    sd.vertices[:] = torch.as_tensor(X + dX).cuda()
    moved = sd.compute_moment_coefficient()        # forces held: not recomputed
    assert abs(moved - cm - float(pred)) <= 1e-12 * abs(float(pred))
    sd.vertices[:] = torch.as_tensor(X).cuda()
    dU = rng.standard_normal(U.shape)
    h = 1e-6
    cp0 = lambda W: (sd.states.copy_(torch.as_tensor(W).cuda()),
                     sd.compute_pressure_coefficients().clone())[1]
    fd = (cp0(U + h * dU) - cp0(U - h * dU)) / (2 * h)
    sd.states[:] = torch.as_tensor(U).cuda()
    tan = sd.apply_pressure_coefficients_wrt_states_fwd(torch.as_tensor(dU).cuda())
    assert rel(tan.real.cpu().numpy(), fd.cpu().numpy()) <= 1e-8
