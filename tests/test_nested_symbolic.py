"""CPU tests of the nested-dissection ordering (pyeulerale_amd/nested.py): the symbolic structure
run through the same multifrontal steps as the GPU (padded fronts, children's updates gathered
at their ring positions, original blocks placed once, partial-pivoting LU of F11, Z = F11^-1 F12,
update F22 - F21 Z, then the two sweeps of the solve) in NumPy reproduces a dense solve of
J - s I and of the adjoint J^T - conj(s) I.  Any cell missing from a ring, a block placed twice
or in the wrong front, or a wrong child position fails it."""

from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg as sla

from pyeulerale_amd.nested import OFFSETS, symbolic


def random_blocks(M, N, seed=0):
    """Blocks (9, 4, 4, N, M) with the stencil's sparsity (radial neighbours beyond the grid
    zero), as the assembled Jacobian's storage."""
    rng = np.random.default_rng(seed)
    JU = rng.standard_normal((9, 4, 4, N, M))
    m = np.arange(M)
    for s, (dm, dn) in enumerate(OFFSETS):
        bad = (m + dm < 0) | (m + dm >= M)
        JU[s][..., bad] = 0.0
    JU[0] += 8.0 * np.eye(4)[:, :, None, None]      # diagonally heavy enough to be well posed
    return JU


def dense(JU, M, N):
    C = M * N
    A = np.zeros((4 * C, 4 * C))
    for n in range(N):
        for m in range(M):
            r = n * M + m
            for s, (dm, dn) in enumerate(OFFSETS):
                mm = m + dm
                if not 0 <= mm < M:
                    continue
                c = ((n + dn) % N) * M + mm
                for i in range(4):
                    A[i * C + r, np.arange(4) * C + c] += JU[s, i, :, n, m]
    return A


def multifrontal(st, JU, shift, b, adjoint=False):
    """The GPU algorithm in NumPy: b (4C, nrhs) in the canonical layout (v * C + cell)."""
    M, N = st.M, st.N
    C = M * N
    sh = np.conj(shift) if adjoint else shift
    upd, fac = None, []
    for lv in st.levels:
        P, Bn, L = lv.P, lv.B, lv.L
        n = len(lv.nodes)
        F = np.zeros((n, L, L), complex)
        # children's update matrices at their ring positions
        for k in range(n):
            for c in range(2):
                kid = lv.child[k, c]
                if kid < 0:
                    continue
                pos = lv.child_pos[k, c]
                idx = np.flatnonzero(pos >= 0)
                rows = (4 * idx[:, None] + np.arange(4)).reshape(-1)
                crow = (4 * pos[idx][:, None] + np.arange(4)).reshape(-1)
                F[k][np.ix_(rows, rows)] += upd[kid][np.ix_(crow, crow)]
            # identity on the padding of S
            for j in range(lv.Pc):
                if lv.cells[k, j] < 0:
                    F[k, 4 * j:4 * j + 4, 4 * j:4 * j + 4] += np.eye(4)
        for k, fr, fc, cell, s in lv.blocks:
            m, nn = cell % M, cell // M
            blk = JU[s, :, :, nn, m]
            if s == 0:
                blk = blk - sh * np.eye(4) if not adjoint else blk - sh * np.eye(4)
            if adjoint:
                F[k, 4 * fc:4 * fc + 4, 4 * fr:4 * fr + 4] += blk.T
            else:
                F[k, 4 * fr:4 * fr + 4, 4 * fc:4 * fc + 4] += blk
        lus, Zs, F21s = [], [], []
        new_upd = np.zeros((n, Bn, Bn), complex)
        for k in range(n):
            lu = sla.lu_factor(F[k, :P, :P])
            Z = sla.lu_solve(lu, F[k, :P, P:])
            new_upd[k] = F[k, P:, P:] - F[k, P:, :P] @ Z
            lus.append(lu)
            Zs.append(Z)
            F21s.append(F[k, P:, :P].copy())
        fac.append((lus, Zs, F21s))
        upd = new_upd
    # solve
    y = np.array(b, complex)

    def gidx(cells):
        ok = cells >= 0
        g = (np.arange(4)[None, :] * C + np.where(ok, cells, 0)[:, None]).reshape(-1)
        return g, np.repeat(ok, 4)

    for lv, (lus, Zs, F21s) in zip(st.levels, fac):
        for k in range(len(lv.nodes)):
            gs, oks = gidx(lv.cells[k, :lv.Pc])
            gb, okb = gidx(lv.cells[k, lv.Pc:])
            t = np.where(oks[:, None], y[gs], 0)
            t = sla.lu_solve(lus[k], t)
            y[gs[oks]] = t[oks]
            w = F21s[k] @ t
            np.subtract.at(y, gb[okb], w[okb])
    for lv, (lus, Zs, F21s) in zip(reversed(st.levels), reversed(fac)):
        for k in range(len(lv.nodes)):
            gs, oks = gidx(lv.cells[k, :lv.Pc])
            gb, okb = gidx(lv.cells[k, lv.Pc:])
            xb = np.where(okb[:, None], y[gb], 0)
            xs = y[gs] - Zs[k] @ xb
            y[gs[oks]] = xs[oks]
    return y


@pytest.mark.parametrize("M,N,leaf", [(6, 8, 16), (8, 16, 16), (12, 26, 24), (16, 32, 64),
                                      (9, 40, 20)])
@pytest.mark.parametrize("adjoint", [False, True])
def test_multifrontal_reproduces_dense_solve(M, N, leaf, adjoint):
    st = symbolic(M, N, leaf)
    cells = np.concatenate([lv.cells[lv.cells >= 0] for lv in st.levels])
    S_cells = np.concatenate([lv.cells[:, :lv.Pc][lv.cells[:, :lv.Pc] >= 0] for lv in st.levels])
    assert np.array_equal(np.sort(S_cells), np.arange(M * N))   # every cell eliminated once
    assert cells.size >= M * N
    JU = random_blocks(M, N, seed=M + N)
    shift = complex(0.3, 1.7)
    A = dense(JU, M, N) - shift * np.eye(4 * M * N)
    if adjoint:
        A = A.conj().T
    rng = np.random.default_rng(1)
    b = rng.standard_normal((4 * M * N, 3)) + 1j * rng.standard_normal((4 * M * N, 3))
    x = multifrontal(st, JU, shift, b, adjoint)
    ref = np.linalg.solve(A, b)
    assert np.linalg.norm(x - ref) <= 1e-10 * np.linalg.norm(ref)


def test_factor_memory_at_1025x513_fits_one_gpu():
    st = symbolic(512, 1024)
    assert 40e9 < st.factor_bytes() < 70e9        # vs 412 GB for the ring LU
    assert 30e12 < st.flops() < 55e12


def test_flop_count_of_one_front():
    """A grid that is a single leaf has one front (its 4 C unknowns, no ring): the count is then
    the textbook one of a complex LU, 8/3 n^3 real operations (a complex multiply-add is 8)."""
    st = symbolic(4, 8, leaf=64)
    assert len(st.levels) == 1 and len(st.levels[0].nodes) == 1 and st.levels[0].B == 0
    n = st.levels[0].P
    assert n == 4 * 4 * 8
    assert st.flops() == 8.0 / 3.0 * n ** 3
