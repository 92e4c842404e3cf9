"""Exact resolvent (dR/dU - s I)^-1 by nested dissection: a multifrontal block LU for grids of
any size that fits the GPU (the ring LU of ``ringlu.py`` needs 3 M (4N)^2 complex values per
shift, 412 GB at 1025x513; this needs ~56 GB there).

The reference factorises ``dR/dU - s I`` with SuperLU on the host, once per shift
(core.py:533-599).  Here the structure of the O-grid gives the ordering:

* the stencil of a cell reaches two cells along each grid line (blocks c, b, bb, f, ff, i, ii,
  o, oo), so a separator two cells wide splits a region in two;
* the periodic grid is cut by two separators of two columns each (n = 0, 1 and N/2, N/2 + 1)
  into two rectangles, each then bisected recursively across its longer side; rectangles of at
  most ``leaf`` cells are eliminated whole;
* eliminating a rectangle couples the cells of the two-cell ring around it (parts of the
  separators of its ancestors), so the front of a node is its separator S followed by that
  ring B.  Front sizes, the index maps and the placement of every original block follow from the
  rectangles alone (:func:`symbolic`, computed once per grid size on the host).

Factorisation, one tree level at a time from the leaves up, every node of a level in one batch
(``Factorization``):

* ``pea_nd_assemble`` writes the fronts: the update matrices of the two children gathered at the
  positions of their rings, identity on the padding of S, then ``pea_nd_blocks`` adds the
  original 4x4 blocks whose first-eliminated cell lies in S, with the shift on the diagonal;
* ``pea_nd_factor``: partial-pivoting LU of F11 (S x S), Z = F11^-1 F12, update
  F22 - F21 Z (rocSOLVER getrf / getrs and a rocBLAS ZGEMM, strided batched over the level);
* LU(F11), its pivots, F21 and Z are kept; F22 is the update matrix passed to the parent.

Solve (any number of right-hand sides): leaves to root, t_S = F11^-1 y_S and y_B -= F21 t_S;
root to leaves, x_S = t_S - Z x_B.  The contributions of the nodes of one level to shared ring
cells are summed in a fixed order (``pea_nd_accumulate``), so solves are deterministic.

The adjoint system (dR/dU)^T - conj(s) I (core.py:594-598) is factorised the same way with the
transposed blocks: the front structure is symmetric.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from functools import lru_cache

import numpy as np
import torch

from . import _lib
from . import kernels as K

C128 = torch.complex128
# stencil offsets of the blocks c, b, bb, f, ff, i, ii, o, oo (spatial_discretization.f90:699-707)
OFFSETS = ((0, 0), (0, -1), (0, -2), (0, 1), (0, 2), (-1, 0), (-2, 0), (1, 0), (2, 0))
OPPOSITE = (0, 3, 4, 1, 2, 7, 8, 5, 6)        # the block of the neighbour that points back


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


@dataclass
class Node:
    depth: int
    rect: tuple | None           # (m0, m1, n0, n1) of the region; None for the whole cylinder
    S: np.ndarray                # separator (or leaf) cells, cell id n*M + m
    B: np.ndarray = None         # ring cells (sorted), eliminated by ancestors
    children: list = field(default_factory=list)
    index: int = -1              # position within its level


@dataclass
class Level:
    """The nodes of one depth, padded to common sizes (in cells: Pc for S, Bc for B)."""
    depth: int
    nodes: list
    Pc: int
    Bc: int
    cells: np.ndarray            # (n, Pc + Bc) front cells, -1 = padding
    child: np.ndarray            # (n, 2) index of each child in the level below, -1 = none
    child_pos: np.ndarray        # (n, 2, Pc + Bc) position in the child's ring, -1 = none
    blocks: np.ndarray           # (e, 5) node, front row cell, front col cell, row cell, block

    @property
    def P(self):
        return 4 * self.Pc

    @property
    def B(self):
        return 4 * self.Bc

    @property
    def L(self):
        return 4 * (self.Pc + self.Bc)


@dataclass
class Structure:
    M: int
    N: int
    leaf: int
    levels: list                 # deepest first

    def factor_bytes(self) -> int:
        """Device memory of one factorisation (LU of F11 with pivots, F21 and Z)."""
        return sum(len(lv.nodes) * (lv.P * lv.P * 16 + lv.P * 4 + 2 * lv.P * lv.B * 16)
                   for lv in self.levels)

    def flops(self) -> float:
        """Real floating-point operations of one factorisation (padded sizes; a complex
        multiply-add is 8 real operations): per front the LU of F11 (8/3 P^3), Z = F11^-1 F12
        (8 P^2 B) and the update F22 - F21 Z (8 P B^2)."""
        tot = 0.0
        for lv in self.levels:
            P, B = lv.P, lv.B
            tot += len(lv.nodes) * (8.0 / 3.0 * P ** 3 + 8.0 * P * P * B + 8.0 * P * B * B)
        return tot


def _rect_cells(M, m0, m1, n0, n1):
    m = np.arange(m0, m1)
    n = np.arange(n0, n1)
    return (n[:, None] * M + m[None, :]).reshape(-1)


def _ring(M, N, rect):
    m0, m1, n0, n1 = rect
    out = []
    for mm in (m0 - 2, m0 - 1, m1, m1 + 1):
        if 0 <= mm < M:
            out.append(np.arange(n0, n1) % N * M + mm)
    m = np.arange(m0, m1)
    for nn in (n0 - 2, n0 - 1, n1, n1 + 1):
        out.append((nn % N) * M + m)
    return np.unique(np.concatenate(out)) if out else np.zeros(0, np.int64)


@lru_cache(maxsize=8)
def symbolic(M: int, N: int, leaf: int = 64) -> Structure:
    """Elimination tree, front sizes and index maps of the nested-dissection ordering."""
    nodes: list[Node] = []

    def build(rect, depth):
        m0, m1, n0, n1 = rect
        h, w = m1 - m0, n1 - n0
        if h * w <= leaf or (h < 5 and w < 5):
            nodes.append(Node(depth, rect, _rect_cells(M, m0, m1, n0, n1)))
            return len(nodes) - 1
        if w >= h and w >= 5 or h < 5:
            c = n0 + (w - 2) // 2
            kids = [build((m0, m1, n0, c), depth + 1), build((m0, m1, c + 2, n1), depth + 1)]
            S = _rect_cells(M, m0, m1, c, c + 2)
        else:
            c = m0 + (h - 2) // 2
            kids = [build((m0, c, n0, n1), depth + 1), build((c + 2, m1, n0, n1), depth + 1)]
            S = _rect_cells(M, c, c + 2, n0, n1)
        nodes.append(Node(depth, rect, S, children=kids))
        return len(nodes) - 1

    if M * N <= leaf or N < 8:
        nodes.append(Node(0, None, np.arange(M * N)))
    else:
        h = N // 2
        kids = [build((0, M, 2, h), 1), build((0, M, h + 2, N), 1)]
        S = np.concatenate([_rect_cells(M, 0, M, 0, 2), _rect_cells(M, 0, M, h, h + 2)])
        nodes.append(Node(0, None, S, children=kids))
    owner = np.full(M * N, -1, np.int64)
    for k, nd in enumerate(nodes):
        if (owner[nd.S] >= 0).any():
            raise AssertionError("separators overlap")
        owner[nd.S] = k
        nd.B = np.zeros(0, np.int64) if nd.rect is None else _ring(M, N, nd.rect)
    if (owner < 0).any():
        raise AssertionError("cells left out of the ordering")
    # region membership: the node's rectangle (the whole grid for the root)
    def in_region(nd, cells):
        if nd.rect is None:
            return np.ones(cells.shape, bool)
        m0, m1, n0, n1 = nd.rect
        m, n = cells % M, cells // M
        return (m >= m0) & (m < m1) & (n >= n0) & (n < n1)

    depth_max = max(nd.depth for nd in nodes)
    by_depth = [[k for k, nd in enumerate(nodes) if nd.depth == d] for d in range(depth_max + 1)]
    for lst in by_depth:
        for i, k in enumerate(lst):
            nodes[k].index = i
    levels = []
    for d in range(depth_max, -1, -1):
        ids = by_depth[d]
        Pc = max(len(nodes[k].S) for k in ids)
        Bc = max(len(nodes[k].B) for k in ids)
        n = len(ids)
        cells = np.full((n, Pc + Bc), -1, np.int64)
        child = np.full((n, 2), -1, np.int64)
        child_pos = np.full((n, 2, Pc + Bc), -1, np.int64)
        ent = []
        for i, k in enumerate(ids):
            nd = nodes[k]
            S, B = nd.S, nd.B
            cells[i, :len(S)] = S
            cells[i, Pc:Pc + len(B)] = B
            # front position of a cell: S (sorted by construction) and B by searchsorted
            if not (np.diff(S) > 0).all():
                raise AssertionError("separator cells not sorted")
            for ci, kid in enumerate(nd.children):
                kn = nodes[kid]
                child[i, ci] = kn.index
                kb = kn.B                                   # sorted
                fc = cells[i]
                valid = fc >= 0
                pos = np.searchsorted(kb, fc[valid])
                pos_c = np.minimum(pos, len(kb) - 1)
                hit = kb[pos_c] == fc[valid]
                cp = np.full(valid.sum(), -1, np.int64)
                cp[hit] = pos_c[hit]
                child_pos[i, ci, valid] = cp
                if hit.sum() != len(kb):
                    raise AssertionError("a child's ring is not inside its parent's front")
            # original blocks first eliminated here
            m, nn = S % M, S // M
            for blk, (dm, dn) in enumerate(OFFSETS):
                mm, nb = m + dm, (nn + dn) % N
                ok = (mm >= 0) & (mm < M)
                src = S[ok]
                nbr = (nb * M + mm)[ok]
                srow = np.searchsorted(S, src)
                own = owner[nbr]
                in_s = own == k
                reg = in_region(nd, nbr)
                in_b = ~reg
                # neighbour in S
                if in_s.any():
                    scol = np.searchsorted(S, nbr[in_s])
                    ent.append(np.stack([np.full(in_s.sum(), i), srow[in_s], scol,
                                         src[in_s], np.full(in_s.sum(), blk)], 1))
                if in_b.any():
                    bpos = Pc + np.searchsorted(B, nbr[in_b])
                    if not np.array_equal(B[bpos - Pc], nbr[in_b]):
                        raise AssertionError("a neighbour outside the region is not in the ring")
                    # row in S, column in the ring, and the ring row's block pointing back
                    ent.append(np.stack([np.full(in_b.sum(), i), srow[in_b], bpos,
                                         src[in_b], np.full(in_b.sum(), blk)], 1))
                    ent.append(np.stack([np.full(in_b.sum(), i), bpos, srow[in_b],
                                         nbr[in_b], np.full(in_b.sum(), OPPOSITE[blk])], 1))
        blocks = np.concatenate(ent) if ent else np.zeros((0, 5), np.int64)
        levels.append(Level(d, ids, Pc, Bc, cells, child, child_pos, blocks))
    return Structure(M, N, leaf, levels)


def _accumulation_lists(lv: Level, C: int):
    """Per target global index (v*C + cell) the (node, front row) contributions of the level's
    rings, sorted by target and then node: the fixed summation order of the solve."""
    k, j = np.nonzero(lv.cells[:, lv.Pc:] >= 0)
    cell = lv.cells[:, lv.Pc:][k, j]
    kk = np.repeat(k, 4)
    jj = (4 * np.repeat(j, 4) + np.tile(np.arange(4), len(j)))
    tg = (np.tile(np.arange(4), len(j)) * C + np.repeat(cell, 4))
    order = np.lexsort((kk, tg))
    kk, jj, tg = kk[order], jj[order], tg[order]
    targets, start = np.unique(tg, return_index=True)
    ptr = np.append(start, len(tg))
    return ptr, kk, jj, targets


class _DeviceLevel:
    pass


_DEVICE_CACHE: dict = {}


def _i32(a, dev):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=dev)


def device_structure(st: Structure, device) -> list:
    """The structure's index maps on the device (int32), checked on the host first: every index
    a kernel will use lies inside the array it addresses."""
    key = (st.M, st.N, st.leaf, str(device))
    if key in _DEVICE_CACHE:
        return _DEVICE_CACHE[key]
    C = st.M * st.N
    out = []
    prev = None
    for lv in st.levels:
        n, Lc = len(lv.nodes), lv.Pc + lv.Bc
        if lv.cells.shape != (n, Lc) or lv.cells.min() < -1 or lv.cells.max() >= C:
            raise AssertionError("front cells out of range")
        has = lv.child >= 0
        if prev is None:
            if has.any():
                raise AssertionError("deepest level with children")
        elif lv.child.max() >= len(prev.nodes) or \
                (lv.child_pos.max(initial=-1) >= prev.Bc):
            raise AssertionError("child index or ring position out of range")
        if (lv.child_pos[~has] >= 0).any():
            raise AssertionError("ring positions given for a missing child")
        b = lv.blocks
        if len(b) and (b[:, 0].min() < 0 or b[:, 0].max() >= n or b[:, 1:3].min() < 0
                       or b[:, 1:3].max() >= Lc or b[:, 3].min() < 0 or b[:, 3].max() >= C
                       or b[:, 4].min() < 0 or b[:, 4].max() > 8):
            raise AssertionError("block entries out of range")
        if len(b) and (np.diff(b[:, 0]) < 0).any():
            raise AssertionError("block entries not sorted by node")
        ptr, kk, jj, tg = _accumulation_lists(lv, C)
        d = _DeviceLevel()
        d.cells = _i32(lv.cells, device)
        d.child = _i32(lv.child, device)
        d.child_pos = _i32(lv.child_pos, device)
        d.blocks = _i32(b.reshape(-1, 5), device)
        d.blk_ptr = np.searchsorted(b[:, 0], np.arange(n + 1)) if len(b) else np.zeros(n + 1, int)
        d.acc_ptr, d.acc_k, d.acc_j, d.acc_target = (_i32(a, device) for a in (ptr, kk, jj, tg))
        d.ntarget = len(tg)
        out.append(d)
        prev = lv
    _DEVICE_CACHE[key] = out
    return out


class Factorization:
    """Nested-dissection factorisation of dR/dU - s I (or of the adjoint system) for one shift.

    ``inverse=True`` keeps the inverses of the pivot blocks F11 (rocSOLVER getri), so that Z, the
    updates and every step of a solve are GEMMs; ``inverse=False`` keeps their LU factors and
    solves by substitution; ``"auto"`` (default) inverts the pivot blocks of the levels with
    fewer than 4096 unknowns per separator (measured at 1025x513: the inverse halves the time of
    the 8192 leaf fronts, LU is 2.7x faster for the single 8192-unknown root).  ``solve``
    follows the elimination with ``refine`` steps of iterative refinement on the exact block
    operator."""

    def __init__(self, g: K.Grid, JU: torch.Tensor, shift: complex = 0.0, *,
                 adjoint: bool = False, leaf: int = 64, inverse="auto",
                 front_bytes: float = 4 * 2 ** 30):
        K._need(JU, (9, 4, 4, g.N, g.M), torch.float64, "JU")
        self.g, self.JU, self.adjoint = g, JU, bool(adjoint)
        self.shift = complex(shift)
        self.st = symbolic(g.M, g.N, leaf)
        self.inverse = [(lv.P < 4096) if inverse == "auto" else bool(inverse)
                        for lv in self.st.levels]
        dev = JU.device
        self.dev = dev
        need = self.st.factor_bytes() + self._transient_bytes(front_bytes)
        free, _ = torch.cuda.mem_get_info(dev)
        free += torch.cuda.memory_reserved(dev) - torch.cuda.memory_allocated(dev)
        if need > free:
            raise _lib.PeaError(f"nested-dissection factorisation of {g.M}x{g.N} cells needs "
                                f"{need / 1e9:.1f} GB, {free / 1e9:.1f} GB are free")
        self.dl = device_structure(self.st, dev)
        self.factors = []
        self._factor(front_bytes)
        self._sh = None

    def _transient_bytes(self, front_bytes):
        upd = [len(lv.nodes) * lv.B * lv.B * 16 for lv in self.st.levels]
        most = max(a + b for a, b in zip(upd, upd[1:] + [0]))
        front = max(min(len(lv.nodes) * lv.L * lv.L * 16, max(front_bytes, lv.L * lv.L * 16))
                    for lv in self.st.levels)
        return most + front

    def _factor(self, front_bytes):
        g, st = self.g, self.st
        s = self.shift.conjugate() if self.adjoint else self.shift
        stream = K._stream()
        upd_prev, Bc_prev = None, 0
        singular = 0
        for lv, d, inv in zip(st.levels, self.dl, self.inverse):
            n, P, B, L = len(lv.nodes), lv.P, lv.B, lv.L
            LU = torch.empty((n, P, P), dtype=C128, device=self.dev)
            piv = torch.empty((n, P), dtype=torch.int32, device=self.dev)
            info = torch.empty(n, dtype=torch.int32, device=self.dev)
            F21t = torch.empty((n, P, B), dtype=C128, device=self.dev)
            Zt = torch.empty((n, B, P), dtype=C128, device=self.dev)
            upd = torch.empty((n, B, B), dtype=C128, device=self.dev) if B else None
            chunk = max(1, int(front_bytes // (L * L * 16)))
            for c0 in range(0, n, chunk):
                nk = min(chunk, n - c0)
                F = torch.empty((nk, L, L), dtype=C128, device=self.dev)
                _lib.call("pea_nd_assemble", nk, lv.Pc, lv.Bc, _p(d.cells[c0:]), _p(d.child[c0:]),
                          _p(d.child_pos[c0:]), Bc_prev, _p(upd_prev), _p(F), stream)
                e0, e1 = int(d.blk_ptr[c0]), int(d.blk_ptr[c0 + nk])
                _lib.call("pea_nd_blocks", e1 - e0, _p(d.blocks[e0:]), c0, g.M, g.N, _p(self.JU),
                          s.real, s.imag, int(self.adjoint), lv.Pc, lv.Bc, _p(F), stream)
                # with the inverse, Z = F11^-1 F12 is written straight into its store
                _lib.call("pea_nd_factor", nk, P, B, _p(F), _p(piv[c0:]), _p(info[c0:]),
                          int(inv), _p(Zt[c0:]) if B else None, stream)
                # F[k] is column-major: tensor index [k, column, row]
                LU[c0:c0 + nk] = F[:, :P, :P]
                if B:
                    F21t[c0:c0 + nk] = F[:, :P, P:]
                    if not inv:
                        Zt[c0:c0 + nk] = F[:, P:, :P]
                    upd[c0:c0 + nk] = F[:, P:, P:]
                del F
            singular += int((info > 0).sum())
            self.factors.append((LU, piv, F21t, Zt, inv))
            upd_prev, Bc_prev = upd, lv.Bc
        self.singular_pivots = singular

    def _eliminate(self, y: torch.Tensor) -> torch.Tensor:
        """y <- A^-1 y in place by the two sweeps; y of shape (nrhs, 4 C)."""
        g, st = self.g, self.st
        nrhs, C4 = y.shape
        stream = K._stream()
        for lv, d, (LU, piv, F21t, Zt, inv) in zip(st.levels, self.dl, self.factors):
            n, P, B, Lc = len(lv.nodes), lv.P, lv.B, lv.Pc + lv.Bc
            T = torch.empty((n, nrhs, P), dtype=C128, device=self.dev)
            _lib.call("pea_nd_gather", n, lv.Pc, 0, Lc, _p(d.cells), nrhs, g.M, g.N, _p(y), C4,
                      _p(T), stream)
            if inv:
                T = torch.bmm(T, LU)          # rows of T are y_S^T; LU[k] holds inv(F11)^T
            else:
                _lib.call("pea_dense_lu_solve", n, P, nrhs, _p(LU), _p(piv), _p(T), stream)
            _lib.call("pea_nd_scatter", n, lv.Pc, 0, Lc, _p(d.cells), nrhs, g.M, g.N, _p(T),
                      _p(y), C4, stream)
            if B:
                W = torch.bmm(T, F21t)
                _lib.call("pea_nd_accumulate", d.ntarget, nrhs, _p(d.acc_ptr), _p(d.acc_k),
                          _p(d.acc_j), _p(d.acc_target), B, _p(W), _p(y), C4, stream)
        for lv, d, (LU, piv, F21t, Zt, inv) in zip(reversed(st.levels), reversed(self.dl),
                                                  reversed(self.factors)):
            n, P, B, Lc = len(lv.nodes), lv.P, lv.B, lv.Pc + lv.Bc
            if not B:
                continue
            XB = torch.empty((n, nrhs, B), dtype=C128, device=self.dev)
            T = torch.empty((n, nrhs, P), dtype=C128, device=self.dev)
            _lib.call("pea_nd_gather", n, lv.Bc, lv.Pc, Lc, _p(d.cells), nrhs, g.M, g.N, _p(y),
                      C4, _p(XB), stream)
            _lib.call("pea_nd_gather", n, lv.Pc, 0, Lc, _p(d.cells), nrhs, g.M, g.N, _p(y), C4,
                      _p(T), stream)
            T.baddbmm_(XB, Zt, alpha=-1)
            _lib.call("pea_nd_scatter", n, lv.Pc, 0, Lc, _p(d.cells), nrhs, g.M, g.N, _p(T),
                      _p(y), C4, stream)
        return y

    def apply(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        """y = A x with the exact blocks (x, y of shape (nrhs, 4, N, M))."""
        g = self.g
        nrhs = x.shape[0]
        if self._sh is None or self._sh.numel() != nrhs:
            self._sh = torch.full((nrhs,), self.shift, dtype=C128, device=self.dev)
        _lib.call("pea_block_multi", g.M, g.N, nrhs, _p(self.JU), _p(self._sh),
                  int(self.adjoint), _p(x), _p(y), K._stream())
        return y

    def solve(self, b: torch.Tensor, refine: int = 1) -> torch.Tensor:
        """x = A^-1 b for right-hand sides b of shape (nrhs, 4, N, M) (canonical storage),
        with ``refine`` steps of iterative refinement; ``self.residual`` holds the relative
        residual |b - A x| / |b| per right-hand side when refine > 0."""
        nrhs = b.shape[0]
        b = b.to(C128).reshape(nrhs, *self.g.cell_shape).contiguous()
        x = self._eliminate(b.reshape(nrhs, -1).clone()).reshape(b.shape)
        w = torch.empty_like(b)
        for it in range(refine):
            r = b - self.apply(x, w)
            x += self._eliminate(r.reshape(nrhs, -1)).reshape(b.shape)
        if refine:
            r = (b - self.apply(x, w)).reshape(nrhs, -1)
            self.residual = (torch.linalg.vector_norm(r, dim=1) /
                             torch.linalg.vector_norm(b.reshape(nrhs, -1), dim=1).clamp_min(1e-300))
        return x

    def precondition(self, r: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
        """z = A^-1 r without refinement (a preconditioner step); r, z of shape
        (nrhs, 4, N, M)."""
        z.copy_(r)
        self._eliminate(z.reshape(r.shape[0], -1))
        return z


def anchor_groups(shifts, groups: int):
    """Splits the shifts into ``groups`` groups of consecutive |s| (as nearly equal in size as
    possible); each group's anchor is its middle shift.  Returns [(indices, anchor)]."""
    shifts = np.asarray(shifts, dtype=complex).reshape(-1)
    order = np.argsort(np.abs(shifts), kind="stable")
    out = []
    for part in np.array_split(order, max(1, min(int(groups), len(shifts)))):
        if len(part):
            out.append((np.sort(part), complex(shifts[part[len(part) // 2]])))
    return out


def solve_shifts_direct(g: K.Grid, JU: torch.Tensor, b: torch.Tensor, shifts, *,
                        adjoint: bool = False, refine: int = 1):
    """Exact solves, one factorisation per shift (the reference's order of work: one sparse
    LU per shift, core.py:533-599).  b of shape (S, 4, N, M); returns (x, relative residual per
    shift)."""
    shifts = np.asarray(shifts, dtype=complex).reshape(-1)
    x = torch.empty_like(b, dtype=C128)
    res = torch.empty(len(shifts), dtype=torch.float64, device=b.device)
    for k, s in enumerate(shifts):
        f = Factorization(g, JU, s, adjoint=adjoint)
        x[k] = f.solve(b[k:k + 1], refine=refine)[0]
        res[k] = f.residual[0]
        del f
    return x, res


def solve_shifts_anchored(g: K.Grid, JU: torch.Tensor, b: torch.Tensor, shifts, *,
                          adjoint: bool = False, groups: int = 4, point=None,
                          operator: str = "blocks", rtol: float = 1e-10, restart: int = 30,
                          maxiter: int = 300):
    """All shifts by batched GMRES, ``groups`` factorisations in all: the shifts are grouped by
    |s| and each group is solved together, preconditioned by the exact factorisation of its
    anchor shift (one multi-right-hand-side elimination per Arnoldi step for the whole group).
    Returns (x, [(indices, anchor, GMRESInfo)])."""
    from .gmres import BatchedGMRES
    shifts = np.asarray(shifts, dtype=complex).reshape(-1)
    x = torch.empty_like(b, dtype=C128)
    infos = []
    for idx, anchor in anchor_groups(shifts, groups):
        f = Factorization(g, JU, anchor, adjoint=adjoint)
        gm = BatchedGMRES(g, JU, shifts[idx], adjoint=adjoint, point=point, precond="nested",
                          operator=operator, rtol=rtol, restart=restart, maxiter=maxiter,
                          factor=f)
        ii = torch.as_tensor(idx, device=b.device)
        x[ii] = gm.solve(b[ii])
        infos.append((idx, anchor, gm.info))
        del gm, f
    return x, infos


def solve_shifts_adaptive(g: K.Grid, JU: torch.Tensor, b: torch.Tensor, shifts, *,
                          adjoint: bool = False, point=None, operator: str = "blocks",
                          rtol: float = 1e-10, budget: int = 40, ratio: float = 8.0,
                          shrink: float = 3.0, min_ratio: float = 1.5, check_every: int = 5,
                          refine: int = 1):
    """All shifts, grouped as their convergence allows (one factorisation per group or shift).

    Going up in |s|: the next group is every unsolved shift with |s| at most ``ratio`` times the
    smallest unsolved |s|, anchored at its middle shift (in log |s|).  Batched GMRES with the
    anchor's exact factorisation as preconditioner solves the group in one multi-right-hand-side
    elimination per Arnoldi step; it stops once the observed rate of every unconverged shift
    predicts more than ``budget`` steps (``BatchedGMRES(give_up=True)``), so a group whose shifts
    sit among lightly damped eigenvalues costs a few steps, not the budget.  Unconverged shifts
    stay unsolved and start the next group.  When a group converges fewer than half of its
    shifts besides the anchor (which its own factorisation always solves), the ratio shrinks by
    ``shrink``; below ``min_ratio`` every remaining shift is solved by its own factorisation (the
    lightly damped band, where a neighbouring shift's factorisation does not precondition).
    Returns (x, infos, exact) with infos [(indices, anchor, GMRESInfo)] and exact the indices
    solved one factorisation each."""
    from .gmres import BatchedGMRES
    shifts = np.asarray(shifts, dtype=complex).reshape(-1)
    x = torch.empty_like(b, dtype=C128)
    mag = np.abs(shifts)
    pending = list(np.argsort(mag, kind="stable"))
    infos, exact = [], []
    direct = False
    while pending:
        if direct:
            k = pending.pop(0)
            f = Factorization(g, JU, shifts[k], adjoint=adjoint)
            x[k] = f.solve(b[k:k + 1], refine=refine)[0]
            exact.append(int(k))
            del f
            continue
        lo = mag[pending[0]]
        group = [k for k in pending if mag[k] <= ratio * max(lo, 1e-300)] or pending[:1]
        anchor_k = group[len(group) // 2]
        f = Factorization(g, JU, shifts[anchor_k], adjoint=adjoint)
        idx = np.array(sorted(group))
        ii = torch.as_tensor(idx, device=b.device)
        gm = BatchedGMRES(g, JU, shifts[idx], adjoint=adjoint, point=point, precond="nested",
                          operator=operator, rtol=rtol, restart=budget, maxiter=budget,
                          factor=f, give_up=True, check_every=check_every)
        x[ii] = gm.solve(b[ii])
        conv = gm.info.converged.copy()
        pos = int(np.flatnonzero(idx == anchor_k)[0])
        if not conv[pos]:   # the anchor's own system: one exact solve with its factorisation
            x[anchor_k] = f.solve(b[anchor_k:anchor_k + 1], refine=refine)[0]
            conv[pos] = True
            gm.info.converged[pos] = True
        infos.append((idx, complex(shifts[anchor_k]), gm.info))
        solved = set(int(k) for k in idx[conv])
        pending = [k for k in pending if int(k) not in solved]
        if len(idx) > 1 and conv.sum() - 1 < (len(idx) - 1) / 2:
            ratio /= shrink
            direct = ratio < min_ratio
        del gm, f
    return x, infos, exact
