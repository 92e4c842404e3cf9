"""The allocation-free adjoint entry point (pea_vjp_odes_noalloc) against pea_vjp_odes with a
caller-given workspace, through the C ABI: two calls in a row on the same GPU, the second on a
smaller grid (the static partial-sum array keeps the first call's contents), every output
bitwise equal."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from helpers import case, crandn, dev_cells, dev_verts

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _outputs(g, dev):
    C128 = torch.complex128
    return {"states": torch.full(g.cell_shape, 7.0, dtype=C128, device=dev),
            "vertices": torch.full(g.vert_shape, 7.0, dtype=C128, device=dev),
            "velocities": torch.full(g.vert_shape, 7.0, dtype=C128, device=dev),
            "mach": torch.full((1,), 7.0, dtype=C128, device=dev)}


def _run(K, name, p, mach, g, X, V, U, w, out, ws=None):
    args = [ctypes.byref(p.c()), float(mach), g.M, g.N, K._ptr(X), K._ptr(V), K._ptr(U),
            K._ptr(w), K._ptr(out["states"]), K._ptr(out["vertices"]),
            K._ptr(out["velocities"]), K._ptr(out["mach"])]
    if name == "pea_vjp_odes":
        args.append(K._ptr(ws))
    K._lib.call(name, *args, K._stream())


@pytest.mark.parametrize("sizes", [((257, 129), (33, 17))])
def test_noalloc_adjoint_matches_workspace_adjoint(sizes):
    from pyeulerale_amd import kernels as K
    for i, (A, Rv) in enumerate(sizes):
        print(f"[no-alloc] call {i}: {A}x{Rv} vertices", flush=True)
        X, V, U = case(A, Rv, alpha=0.02, alpha_dot=0.01, seed=10 + i)
        M, N = U.shape[1:]
        g, p = K.Grid(M, N), K.Params(aoa_deg=1.25)
        Xd, Vd, Ud = dev_verts(X), dev_verts(V), dev_cells(U)
        w = dev_cells(crandn(np.random.default_rng(i), (4, M, N)))
        ref, got = _outputs(g, Ud.device), _outputs(g, Ud.device)
        ws = K._workspace(g, Ud.device)
        _run(K, "pea_vjp_odes", p, 0.5, g, Xd, Vd, Ud, w, ref, ws)
        _run(K, "pea_vjp_odes_noalloc", p, 0.5, g, Xd, Vd, Ud, w, got)
        torch.cuda.synchronize()
        for k in ref:
            assert torch.equal(got[k], ref[k]), (i, k)
        assert float(ref["mach"].abs()[0]) > 0.0
        # states only: the face-based kernel, no partial sums
        so_ref = {"states": torch.empty_like(ref["states"])}
        so_got = {"states": torch.empty_like(ref["states"])}
        for name, o in (("pea_vjp_odes", so_ref), ("pea_vjp_odes_noalloc", so_got)):
            K._lib.call(name, ctypes.byref(p.c()), 0.5, g.M, g.N, K._ptr(Xd), K._ptr(Vd),
                        K._ptr(Ud), K._ptr(w), K._ptr(o["states"]), None, None, None,
                        *(([None] if name == "pea_vjp_odes" else [])), K._stream())
        assert torch.equal(so_got["states"], so_ref["states"])
