"""GPU tests of the resolvent: multi-vector block products, the exact ring block-LU solver,
radial-line and block-Jacobi preconditioners, batched GMRES, all against the reference's direct
method (SuperLU on the exported block matrix, core.py:512-599), and the reference's own
round-trip check (test:332-356)."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import case, crandn, dev_cells, host, rel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sd():
    from pyeulerale_amd.core import SpatialDiscretization
    X, V, U = case(33, 17, thickness=0.12, mach=0.5, aoa=1.25, alpha=0.01, alpha_dot=0.004,
                   noise=1e-3)
    s = SpatialDiscretization(vertices=X, angle_of_attack=1.25, mach_number=0.5)
    s.velocities[:] = torch.as_tensor(V).cuda()
    s.states[:] = torch.as_tensor(U).cuda()
    s.linearize()
    print("[resolvent] 33x17 linearized", flush=True)
    return s


@pytest.mark.parametrize("S", [5, 11, 32])
@pytest.mark.parametrize("grid", ["33x17", "43x22"])
def test_block_multi_matches_single_products(request, grid, S):
    """The kernel's four waves take vectors w, w + 4, ...: S = 5 leaves three waves one vector,
    S = 11 gives some waves three and others two, S = 32 (the batched GMRES case) eight each.
    43x22 has a cell count that is not a multiple of the kernel's 64-cell workgroups."""
    from pyeulerale_amd import _lib
    from pyeulerale_amd import kernels as K
    from pyeulerale_amd.gmres import _p
    sd = request.getfixturevalue("sd" if grid == "33x17" else "sd_ragged")
    g = sd._grid
    JU = sd._get_blocks()["U"]
    rng = np.random.default_rng(S)
    x = torch.as_tensor(crandn(rng, (S, *g.cell_shape))).cuda()
    shifts = torch.as_tensor(crandn(rng, S)).cuda()
    for tr in (0, 1):
        y = torch.empty_like(x)
        _lib.call("pea_block_multi", g.M, g.N, S, _p(JU), _p(shifts), tr, _p(x), _p(y),
                  K._stream())
        for k in range(S):
            s = complex(shifts[k])
            ref = (K.apply_states_rev(g, JU, x[k]) - np.conj(s) * x[k]) if tr else \
                (K.apply_states_fwd(g, JU, x[k]) - s * x[k])
            assert rel(y[k].cpu().numpy(), ref.cpu().numpy()) <= 1e-13


@pytest.fixture(scope="module")
def sd_ragged():
    """A grid whose line count (N = 42) is not a multiple of the solve's 16-line groups and whose
    line length (M = 21) is not a multiple of its 8-step blocks."""
    from pyeulerale_amd.core import SpatialDiscretization
    X, V, U = case(43, 22, thickness=0.12, mach=0.5, aoa=1.25, alpha=0.01, alpha_dot=0.004,
                   noise=1e-3)
    s = SpatialDiscretization(vertices=X, angle_of_attack=1.25, mach_number=0.5)
    s.velocities[:] = torch.as_tensor(V).cuda()
    s.states[:] = torch.as_tensor(U).cuda()
    s.linearize()
    print("[resolvent] 43x22 linearized", flush=True)
    return s


@pytest.mark.parametrize("lanes", [4, 16])
@pytest.mark.parametrize("grid", ["33x17", "43x22"])
@pytest.mark.parametrize("adjoint", [False, True])
def test_line_preconditioner_solves_each_radial_line(request, grid, adjoint, lanes):
    """pea_line_factor/pea_line_solve against dense solves of every radial line's system in numpy:
    the line's blocks ii, i, c, o, oo of dR/dU minus s I (transposed, minus conj(s) I, for the
    adjoint), for several shifts at once."""
    from pyeulerale_amd import _lib
    from pyeulerale_amd import kernels as K
    from pyeulerale_amd.gmres import _p
    sd = request.getfixturevalue("sd" if grid == "33x17" else "sd_ragged")
    print(f"[resolvent] line preconditioner {grid} adjoint={adjoint} lanes={lanes}", flush=True)
    g = sd._grid
    M, N = g.M, g.N
    JU = sd._get_blocks()["U"]
    J = JU.cpu().numpy()                      # (slot, i, j, n, m)
    shifts = np.array([complex(0.5, 1.0), complex(0.0, 3.0), complex(-0.2, 0.05)])
    S = len(shifts)
    rng = np.random.default_rng(1)
    b = crandn(rng, (S, 4, N, M))
    slot = {-2: 6, -1: 5, 0: 0, 1: 7, 2: 8}
    ref = np.empty_like(b)
    for n in range(N):
        A = np.zeros((4 * M, 4 * M))
        for m in range(M):
            for off, sl in slot.items():
                if 0 <= m + off < M:
                    A[4 * m:4 * m + 4, 4 * (m + off):4 * (m + off) + 4] = J[sl, :, :, n, m]
        if adjoint:
            A = A.T
        for k, s in enumerate(shifts):
            sh = np.conj(s) if adjoint else s
            rhs = b[k, :, n, :].T.reshape(-1)            # line vector, index 4m + v
            ref[k, :, n, :] = np.linalg.solve(A - sh * np.eye(4 * M), rhs).reshape(M, 4).T
    sh_d = torch.as_tensor(shifts).cuda()
    F = torch.empty(int(_lib.load().pea_line_factor_bytes(M, N, S)) // 8, dtype=torch.float64,
                    device="cuda")
    _lib.call("pea_line_factor", M, N, S, _p(JU), _p(sh_d), int(adjoint), _p(F), K._stream())
    bd = torch.as_tensor(b).cuda()
    x = torch.empty_like(bd)
    _lib.call("pea_line_solve_lanes", lanes)
    try:
        _lib.call("pea_line_solve", M, N, S, _p(JU), _p(F), int(adjoint), _p(bd), _p(x),
                  K._stream())
        assert rel(x.cpu().numpy(), ref) <= 1e-11
        # in place (x aliasing r): the same result
        _lib.call("pea_line_solve", M, N, S, _p(JU), _p(F), int(adjoint), _p(bd), _p(bd),
                  K._stream())
        assert torch.equal(bd, x)
    finally:
        _lib.call("pea_line_solve_lanes", 0)


@pytest.mark.parametrize("q", [1, 5, 16, 23, 40])
def test_gs_update_matches_torch(q):
    """The projections V^H w, and the fused Gram-Schmidt pass, against the same sums in torch:
    the update w -= V h (mode 0), with
    the next projections V^H w (mode 1, and past 32 rows through the separate multi-dot) or the
    squared norm (mode 2); rows beyond q and other shifts' data untouched, repeat calls bitwise
    equal."""
    from pyeulerale_amd import _lib
    from pyeulerale_amd import kernels as K
    from pyeulerale_amd.gmres import _p
    print(f"[resolvent] gs_update q={q}", flush=True)
    c128 = torch.complex128
    S, n, rows, ld = 3, 10007, 41, 45
    rng = np.random.default_rng(q)
    V = torch.as_tensor(crandn(rng, (S, rows, n))).cuda()
    w0 = torch.as_tensor(crandn(rng, (S, n))).cuda()
    h = torch.as_tensor(crandn(rng, (S, ld))).cuda()
    nb = 7
    P = torch.empty((S, nb, rows), dtype=c128, device="cuda")
    ref = w0 - (V[:, :q] * h[:, :q, None]).sum(1)
    dots = (V[:, :q].conj() * ref[:, None]).sum(-1)
    nrm = (ref.abs() ** 2).sum(-1)
    proj = torch.full((S, ld), 7.0, dtype=c128, device="cuda")
    _lib.call("pea_multidot", S, n, q, _p(V), n, rows * n, _p(w0), n, nb, _p(P), _p(proj), ld,
              K._stream())
    ref_proj = (V[:, :q].conj() * w0[:, None]).sum(-1)
    assert rel(proj[:, :q].cpu().numpy(), ref_proj.cpu().numpy()) <= 1e-13
    assert (proj[:, q:] == 7.0).all()
    for mode in (0, 1, 2):
        outs = []
        for _ in range(2):
            w = w0.clone()
            out = torch.full((S, ld), 7.0, dtype=c128, device="cuda")
            _lib.call("pea_gs_update", S, n, q, _p(V), n, rows * n, _p(h), ld, _p(w), n, mode,
                      nb, _p(P), _p(out), ld, K._stream())
            outs.append((w.clone(), out.clone()))
        assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
        w, out = outs[0]
        assert rel(w.cpu().numpy(), ref.cpu().numpy()) <= 1e-14
        if mode == 0:
            assert (out == 7.0).all()
        elif mode == 1:
            assert rel(out[:, :q].cpu().numpy(), dots.cpu().numpy()) <= 1e-12
            assert (out[:, q:] == 7.0).all()
        else:
            assert rel(out[:, 0].real.cpu().numpy(), nrm.cpu().numpy()) <= 1e-13
            assert (out[:, 0].imag == 0).all() and (out[:, 1:] == 7.0).all()


@pytest.mark.parametrize("q", [3, 13, 32])
def test_gs_orthogonalize_selective(q):
    """pea_gs_orthogonalize against the same steps in torch: h1 = V^H w, w -= V h1, and, only for
    shifts where |h1|^2 > |w|^2/2, h2 = V^H w, w -= V h2; outputs h1 + h2 and |w|^2.  Basis rows
    orthonormal per shift; two shifts with w mostly in their span (projection repeated), two
    with random w (not repeated)."""
    from pyeulerale_amd import _lib
    from pyeulerale_amd import kernels as K
    from pyeulerale_amd._lib import PeaError
    from pyeulerale_amd.gmres import _p
    print(f"[resolvent] gs_orthogonalize q={q}", flush=True)
    c128 = torch.complex128
    S, n, rows, nb = 4, 5003, 33, 9
    rng = np.random.default_rng(q)
    Q = [np.linalg.qr(crandn(rng, (n, rows)))[0].T for _ in range(S)]     # orthonormal rows
    V = torch.as_tensor(np.stack(Q)).cuda().contiguous()
    w0 = np.empty((S, n), dtype=complex)
    for k in range(S):
        if k < 2:
            w0[k] = crandn(rng, q) @ Q[k][:q] + 1e-3 * crandn(rng, n)
        else:
            w0[k] = crandn(rng, n)
    wd = torch.as_tensor(w0).cuda()
    P = torch.empty((S, nb, rows + 1), dtype=c128, device="cuda")
    hw = torch.empty((2, S, rows + 1), dtype=c128, device="cuda")
    flags = torch.full((S,), -1, dtype=torch.int32, device="cuda")
    h = torch.full((S, rows + 2), 7.0, dtype=c128, device="cuda")
    nrm2 = torch.zeros(S, dtype=torch.float64, device="cuda")
    _lib.call("pea_gs_orthogonalize", S, n, q, _p(V), n, rows * n, _p(wd), n, nb, _p(P), _p(hw),
              _p(flags), _p(h), rows + 2, _p(nrm2), K._stream())
    Vq = np.stack(Q)[:, :q]
    for k in range(S):
        h1 = Vq[k].conj() @ w0[k]
        w1 = w0[k] - Vq[k].T @ h1
        reorth = np.vdot(h1, h1).real > 0.5 * np.vdot(w0[k], w0[k]).real
        assert int(flags[k]) == int(reorth) == int(k < 2)
        if reorth:
            h2 = Vq[k].conj() @ w1
            w1 = w1 - Vq[k].T @ h2
            h1 = h1 + h2
        assert rel(wd[k].cpu().numpy(), w1) <= 1e-12
        assert rel(h[k, :q].cpu().numpy(), h1) <= 1e-12
        assert abs(float(nrm2[k]) - np.vdot(w1, w1).real) <= 1e-12 * np.vdot(w1, w1).real
    assert (h[:, q:] == 7.0).all()
    with pytest.raises(PeaError):
        _lib.call("pea_gs_orthogonalize", S, n, 33, _p(V), n, rows * n, _p(wd), n, nb, _p(P),
                  _p(hw), _p(flags), _p(h), rows + 2, _p(nrm2), K._stream())


@pytest.mark.parametrize("precond,shifts", [
    ("line", [complex(0.5, 1.0), complex(0.0, 3.0), complex(2.0, 0.0)]),
    # block-Jacobi only copes when the shift dominates the radial coupling of thin cells
    ("jacobi", [complex(3e3, 1.0), complex(1e4, 0.0)]),
])
@pytest.mark.parametrize("adjoint", [False, True])
def test_gmres_matches_direct_solve(sd, precond, shifts, adjoint):
    rng = np.random.default_rng(1)
    b = crandn(rng, (len(shifts), 4, sd.num_radial, sd.num_angular))
    x = sd.solve_odes_wrt_states_batched(b, shifts, adjoint=adjoint, solver="gmres",
                                         fallback=None, precond=precond,
                                         rtol=1e-12).cpu().numpy()
    assert sd.last_gmres.converged.all(), sd.last_gmres.residuals
    for k, s in enumerate(shifts):
        solve = sd.solve_odes_wrt_states_adj if adjoint else sd.solve_odes_wrt_states_fwd
        xd = solve(b[k], shift=s, solver="superlu").cpu().numpy()
        assert rel(x[k], xd) <= 1e-9, (k, rel(x[k], xd))


@pytest.mark.parametrize("adjoint", [False, True])
def test_ring_lu_matches_superlu(sd, adjoint):
    """Exact batched ring block LU against SuperLU, including shifts near zero where GMRES
    stalls (the JST spectrum hugs the imaginary axis)."""
    rng = np.random.default_rng(4)
    shifts = [0.0, complex(0, 0.05), complex(0, 1.0), complex(3.0, 0.5)]
    b = crandn(rng, (len(shifts), 4, sd.num_radial, sd.num_angular))
    print(f"[resolvent] ring LU adjoint={adjoint}", flush=True)
    x = sd.solve_odes_wrt_states_batched(b, shifts, adjoint=adjoint).cpu().numpy()
    # Backward stability: the residual of a partial-pivoting LU with one step of refinement is
    # bounded by a small multiple of eps |A| |x| (norm-wise; |A| <= |J|_F + |s| sqrt(n) in
    # Frobenius norm).  The multiple 16 leaves room for the order of the batched arithmetic.
    JF = float(torch.linalg.vector_norm(sd._get_blocks()["U"]))
    n = 4 * sd.num_radial * sd.num_angular
    eps = np.finfo(float).eps
    for k, s in enumerate(shifts):
        bound = 16 * eps * (JF + abs(s) * np.sqrt(n)) * np.linalg.norm(x[k]) / np.linalg.norm(b[k])
        print(f"[resolvent] shift {s}: residual {float(sd.last_residual[k]):.2e}, "
              f"backward-error bound {bound:.2e}", flush=True)
        assert float(sd.last_residual[k]) <= bound, (k, float(sd.last_residual[k]), bound)
    solve = sd.solve_odes_wrt_states_adj if adjoint else sd.solve_odes_wrt_states_fwd
    for k, s in enumerate(shifts):
        xd = solve(b[k], shift=s, solver="superlu").cpu().numpy()
        assert rel(x[k], xd) <= 1e-10, (k, rel(x[k], xd))
        xs = solve(b[k], shift=s).cpu().numpy()          # single shift, cached factorisation
        assert rel(xs, x[k]) <= 1e-12, (k, rel(xs, x[k]))
    # chunked factorisation (one shift per chunk) gives the same answer
    x1 = sd.solve_odes_wrt_states_batched(b, shifts, adjoint=adjoint, max_bytes=1).cpu().numpy()
    assert rel(x1, x) <= 1e-12


def test_solve_round_trip_like_reference(sd):
    """solve(J x - s x, s) == x and the adjoint analogue, as tests/test_linearization.py:
    332-356, through both the direct and the GMRES solver.  Checked norm-wise: this grid
    resolves the wall (first spacing 1e-3 chord), so the system is far stiffer than the
    reference's 8x9 test grid and single entries lose more digits to conditioning."""
    rng = np.random.default_rng(2)
    s = complex(rng.random(), rng.random())
    x = torch.as_tensor(crandn(rng, (4, sd.num_radial, sd.num_angular))).cuda()
    Jx = sd.apply_odes_wrt_states_fwd(x) - s * x
    for solver in ("direct", "superlu", "gmres"):
        got = sd.solve_odes_wrt_states_fwd(Jx, shift=s, solver=solver, **(
            {"rtol": 1e-13} if solver == "gmres" else {}))
        assert rel(got.cpu().numpy(), x.cpu().numpy()) <= 1e-8, (solver, rel(got.cpu().numpy(), x.cpu().numpy()))
    y = sd.apply_odes_wrt_states_rev(x) - np.conj(s) * x
    for solver in ("direct", "superlu", "gmres"):
        got = sd.solve_odes_wrt_states_adj(y, shift=s, solver=solver, **(
            {"rtol": 1e-13} if solver == "gmres" else {}))
        assert rel(got.cpu().numpy(), x.cpu().numpy()) <= 1e-8, (solver, rel(got.cpu().numpy(), x.cpu().numpy()))


def test_matrix_free_operator_and_batched_equals_sequential(sd):
    rng = np.random.default_rng(3)
    # shifts GMRES converges on (|s| well away from the imaginary-axis spectrum near 0)
    shifts = [complex(0, 10.0), complex(0, 20.0), complex(3.0, 0.0), complex(5.0, 5.0)]
    b = crandn(rng, (4, 4, sd.num_radial, sd.num_angular))
    xb = sd.solve_odes_wrt_states_batched(b, shifts, solver="gmres", fallback=None,
                                          rtol=1e-12).cpu().numpy()
    assert sd.last_gmres.converged.all(), sd.last_gmres.residuals
    xm = sd.solve_odes_wrt_states_batched(b, shifts, solver="gmres", fallback=None, rtol=1e-12,
                                          operator="jvp").cpu().numpy()
    assert rel(xm, xb) <= 1e-9
    for k, s in enumerate(shifts):
        xs = sd.solve_odes_wrt_states_fwd(b[k], shift=s, solver="gmres", rtol=1e-12)
        assert rel(xs.cpu().numpy(), xb[k]) <= 1e-9


def test_gmres_graph_replays_match_eager(sd):
    """The captured Arnoldi steps, replayed over many restart cycles (restart 8, with eager
    residual and update work and host reads between replays), reproduce the uncaptured solver."""
    from pyeulerale_amd.gmres import BatchedGMRES
    rng = np.random.default_rng(5)
    shifts = [complex(0, 10.0), complex(3.0, 0.0), complex(5.0, 5.0)]
    b = torch.as_tensor(crandn(rng, (3, 4, sd.num_angular, sd.num_radial))).cuda()
    out = {}
    for graphs in (False, True):
        gm = BatchedGMRES.from_discretization(sd, shifts, restart=8, maxiter=400, rtol=1e-12,
                                              graphs=graphs, check_every=4)
        print(f"[resolvent] gmres graphs={graphs}", flush=True)
        out[graphs] = (gm.solve(b).cpu().numpy(), gm.info)
    xe, ie = out[False]
    xg, ig = out[True]
    assert ie.converged.all() and ig.converged.all(), (ie.residuals, ig.residuals)
    assert ig.restarts >= 3
    assert (ie.iterations == ig.iterations).all(), (ie.iterations, ig.iterations)
    assert rel(xg, xe) <= 1e-13, rel(xg, xe)
