"""GPU tests of the nested-dissection direct resolvent (pyeulerale_amd/nested.py), through the C
ABI (pea_nd_*):

* on random blocks with the stencil's sparsity, against a dense solve (forward and adjoint);
* on the complex-step Jacobian of a 33x17 grid, against the reference's SuperLU solve
  (core.py:533-599) at shifts from 0 to 3 + 0.5i, forward and adjoint;
* on the 257x129 grid, against the ring block LU;
* at 1025x513 (where the ring LU would need 412 GB), by the residual of the solve with the
  exact block operator, relative to the backward-error scale eps |A| |x| / |b|.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import case, crandn, rel
from test_nested_symbolic import dense, random_blocks

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[nested] {msg}", flush=True)


def storage(a):
    return torch.as_tensor(np.ascontiguousarray(np.swapaxes(np.asarray(a), -1, -2))).cuda()


@pytest.mark.parametrize("M,N,leaf", [(8, 16, 16), (12, 26, 24), (16, 32, 64)])
@pytest.mark.parametrize("adjoint", [False, True])
@pytest.mark.parametrize("inverse", [True, False])
def test_random_blocks_against_dense_solve(M, N, leaf, adjoint, inverse):
    from pyeulerale_amd import kernels as K
    from pyeulerale_amd.nested import Factorization
    say(f"random blocks M={M} N={N} leaf={leaf} adjoint={adjoint}")
    JU = random_blocks(M, N, seed=M * N)
    shift = complex(0.3, 1.7)
    A = dense(JU, M, N) - shift * np.eye(4 * M * N)
    if adjoint:
        A = A.conj().T
    rng = np.random.default_rng(2)
    b = crandn(rng, (3, 4 * M * N))
    ref = np.linalg.solve(A, b.T).T
    f = Factorization(K.Grid(M, N), torch.as_tensor(JU).cuda(), shift, adjoint=adjoint,
                      leaf=leaf, inverse=inverse)
    assert f.singular_pivots == 0
    x = f.solve(torch.as_tensor(b.reshape(3, 4, N, M)).cuda(), refine=0)
    x = x.cpu().numpy().reshape(3, -1)
    assert rel(x, ref) <= 1e-10, rel(x, ref)


@pytest.fixture(scope="module")
def sd33():
    from pyeulerale_amd.core import SpatialDiscretization
    X, V, U = case(33, 17, alpha=0.02, alpha_dot=0.01, seed=4)
    sd = SpatialDiscretization(vertices=X, angle_of_attack=1.25, mach_number=0.5)
    sd.velocities[:] = torch.as_tensor(V).cuda()
    sd.states[:] = torch.as_tensor(U).cuda()
    sd.linearize()
    return sd


@pytest.mark.parametrize("adjoint", [False, True])
def test_jacobian_against_superlu(sd33, adjoint):
    from pyeulerale_amd.nested import Factorization
    rng = np.random.default_rng(3)
    shifts = [0.0, 0.05j, 1j, 3 + 0.5j]
    g = sd33._grid
    JU = sd33._get_blocks()["U"]
    for s in shifts:
        say(f"33x17 shift {s} adjoint={adjoint}")
        b = crandn(rng, (4, sd33.num_radial, sd33.num_angular))
        f = Factorization(g, JU, s, adjoint=adjoint, leaf=16)
        x = f.solve(storage(b)[None])[0].transpose(-1, -2).cpu().numpy()
        solve = sd33.solve_odes_wrt_states_adj if adjoint else sd33.solve_odes_wrt_states_fwd
        ref = solve(b, shift=s, solver="superlu").cpu().numpy()
        assert rel(x, ref) <= 1e-10, (s, rel(x, ref))


@pytest.mark.parametrize("adjoint", [False, True])
def test_adaptive_groups_against_one_factorisation_per_shift(sd33, adjoint):
    """The adaptive batched solver (nested.solve_shifts_adaptive) at 12 shifts i omega,
    omega = logspace(-2, 1, 12), on the 33x17 Jacobian: every shift's solution equal to its own
    exact factorisation's to 1e-8 (GMRES stops at a relative residual of 1e-10), and the groups
    and exact solves together cover every shift once."""
    from pyeulerale_amd.nested import solve_shifts_adaptive, solve_shifts_direct
    g = sd33._grid
    JU = sd33._get_blocks()["U"]
    shifts = 1j * np.logspace(-2, 1, 12)
    rng = np.random.default_rng(8)
    b = storage(crandn(rng, (12, 4, sd33.num_radial, sd33.num_angular)))
    say(f"33x17 adaptive groups, adjoint={adjoint}")
    x, infos, exact = solve_shifts_adaptive(g, JU, b, shifts, adjoint=adjoint)
    ref, _ = solve_shifts_direct(g, JU, b, shifts, adjoint=adjoint)
    covered = sorted([int(k) for idx, _, inf in infos for k in idx[inf.converged]] + exact)
    say(f"33x17 adaptive: groups {[(len(i), inf.converged.sum()) for i, _, inf in infos]}, "
        f"exact {exact}")
    assert covered == list(range(12))
    for k in range(12):
        e = rel(x[k].cpu().numpy(), ref[k].cpu().numpy())
        assert e <= 1e-8, (k, e)


def test_257x129_against_ring_lu():
    from pyeulerale_amd import kernels as K
    from pyeulerale_amd import workloads as W
    from pyeulerale_amd.nested import Factorization
    from pyeulerale_amd.ringlu import RingLU
    c = W.config("C2")
    g, p = K.Grid(c.M, c.N), K.Params(aoa_deg=c.aoa)
    X, V, U = storage(c.X), storage(c.V), storage(c.U[0])
    JU = K.cs_linearize_odes(p, c.mach, g, X, V, U, want=("states",))[0]
    rng = np.random.default_rng(5)
    b = torch.as_tensor(crandn(rng, (2, 4, c.N, c.M))).cuda()
    for s in (0.05j, 2j):
        say(f"257x129 shift {s}")
        f = Factorization(g, JU, s)
        x = f.solve(b)
        xr = RingLU(g, JU, [s, s]).solve(b)
        assert rel(x.cpu().numpy(), xr.cpu().numpy()) <= 1e-10


def test_1025x513_residual():
    """The size the ring LU cannot hold: factorise at one shift and check the residual of the
    solve with the exact block operator against eps |A|_F |x| / |b| (backward stability)."""
    from pyeulerale_amd import _lib
    from pyeulerale_amd import kernels as K
    from pyeulerale_amd import workloads as W
    from pyeulerale_amd.nested import Factorization
    import ctypes
    c = W.config("C5")
    g, p = K.Grid(c.M, c.N), K.Params(aoa_deg=c.aoa)
    X, V, U = storage(c.X), storage(c.V), storage(c.U[0])
    JU = K.cs_linearize_odes(p, c.mach, g, X, V, U, want=("states",))[0]
    s = 0.1j
    say("1025x513 factorisation")
    torch.cuda.synchronize()
    import time
    t0 = time.perf_counter()
    f = Factorization(g, JU, s)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    rng = np.random.default_rng(6)
    b = torch.as_tensor(crandn(rng, (2, 4, c.N, c.M))).cuda()
    x = f.solve(b)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    say(f"1025x513: factor {t1 - t0:.2f} s, solve {t2 - t1:.3f} s, "
        f"{f.singular_pivots} singular pivots")
    sh = torch.full((2,), s, dtype=torch.complex128, device="cuda")
    Ax = torch.empty_like(x)
    _lib.call("pea_block_multi", g.M, g.N, 2, ctypes.c_void_p(JU.data_ptr()),
              ctypes.c_void_p(sh.data_ptr()), 0, ctypes.c_void_p(x.data_ptr()),
              ctypes.c_void_p(Ax.data_ptr()), K._stream())
    res = torch.linalg.vector_norm((Ax - b).reshape(2, -1), dim=1)
    bn = torch.linalg.vector_norm(b.reshape(2, -1), dim=1)
    xn = torch.linalg.vector_norm(x.reshape(2, -1), dim=1)
    JF = float(torch.linalg.vector_norm(JU)) + abs(s) * np.sqrt(4 * c.M * c.N)
    for k in range(2):
        r = float(res[k] / bn[k])
        bound = 64 * np.finfo(float).eps * JF * float(xn[k] / bn[k])
        say(f"1025x513 rhs {k}: residual {r:.2e}, backward-error scale {bound:.2e}")
        assert r <= bound


def test_1025x513_frequency_response_batched_equals_sequential():
    """The 1025x513 configuration end to end: PTC steady state with exact nested-dissection
    solves (examples/response/run.py:89-123), then the pitch transfer functions at 8 of the 32
    frequencies (every fourth of omega = logspace(-2, 1, 32)) solved batched (adaptive groups
    around anchor factorisations, one factorisation per shift in the lightly damped band) and
    one exact factorisation per shift; every shift's residual with the exact operator <= 1e-10
    and c_l, c_m equal to 1e-8.  The coefficients are ~1e4 times more sensitive than the
    residual here (a shift GMRES leaves at 7e-12 moves c_l by 4e-8), so the batched solve runs
    GMRES to 1e-12."""
    from pyeulerale_amd import workloads as W
    from pyeulerale_amd.core import SpatialDiscretization
    from pyeulerale_amd.drivers import frequency_response, pseudo_transient_continuation
    c = W.config("C5")
    sd = SpatialDiscretization(vertices=c.X, angle_of_attack=0.0, mach_number=0.5,
                               coefficient_length=1.0, coefficient_center=(0.25, 0.0))
    hist = pseudo_transient_continuation(sd, rtol=1e-9, max_iter=30, solver="nested",
                                         log=lambda m: say(m.strip()))
    assert hist[-1][2] <= 1e-9
    sd.linearize()
    omega = np.logspace(-2, 1, 32)[::4]
    kred = omega * 0.5 / (0.5 * 1.4 ** 0.5)
    _, cl_s, cm_s = frequency_response(sd, kred, 1.0, 0.25, batched=False, solver="nested",
                                       log=None)
    _, cl_b, cm_b, dU, rhs = frequency_response(sd, kred, 1.0, 0.25, batched=True,
                                                solver="gmres", precond="nested",
                                                groups="adaptive", rtol=1e-12, log=None,
                                                return_states=True)
    say(f"1025x513 response: GMRES steps {sd.last_gmres.iterations.tolist()}, "
        f"converged {sd.last_gmres.converged.tolist()}")
    # residual of every shift with the exact block operator: (J - s I) dU = -rhs
    import ctypes
    from pyeulerale_amd import _lib
    from pyeulerale_amd import kernels as K
    g = sd._grid
    shifts = torch.as_tensor(1j * kred * (0.5 * 1.4 ** 0.5) / 0.5, dtype=torch.complex128,
                             device="cuda")
    x = dU.transpose(-1, -2).contiguous()
    Ax = torch.empty_like(x)
    _lib.call("pea_block_multi", g.M, g.N, len(kred), ctypes.c_void_p(sd._get_blocks()["U"].data_ptr()),
              ctypes.c_void_p(shifts.data_ptr()), 0, ctypes.c_void_p(x.data_ptr()),
              ctypes.c_void_p(Ax.data_ptr()), K._stream())
    b = rhs.transpose(-1, -2).reshape(len(kred), -1)
    res = (torch.linalg.vector_norm((Ax.reshape(len(kred), -1) + b), dim=1)
           / torch.linalg.vector_norm(b, dim=1)).cpu().numpy()
    say(f"1025x513 response: residuals {np.array2string(res, precision=1)}")
    assert (res <= 1e-10).all()
    e_l = np.max(np.abs(cl_b - cl_s) / np.abs(cl_s))
    e_m = np.max(np.abs(cm_b - cm_s) / np.abs(cm_s))
    say(f"1025x513 response: batched vs sequential c_l {e_l:.1e}, c_m {e_m:.1e}")
    assert e_l <= 1e-8 and e_m <= 1e-8
