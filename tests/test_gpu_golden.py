"""GPU kernels against the golden fixtures written by the reference's compiled Fortran
(tests/golden/make_golden.py).  Bars: R, F <= 1e-12; products <= 1e-10 (relative, norm-wise)."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from helpers import c1_directions, dev_cells, dev_verts, host, rel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def _load(name):
    z = np.load(GOLD / name, allow_pickle=False)
    return {k: z[k] for k in z.files}


def _setup(d):
    from pyeulerale_amd import kernels as K
    U = d["U"]
    g = K.Grid(U.shape[1], U.shape[2])
    p = K.Params(aoa_deg=float(d["aoa"]))
    return K, g, p, float(d["mach"]), dev_verts(d["X"]), dev_verts(d["V"]), dev_cells(U)


@pytest.mark.parametrize("name", ["tiny.npz", "zeros.npz"])
def test_all_products_against_reference_fixture(name):
    d = _load(name)
    print(f"[golden {name}] M,N={d['U'].shape[1:]}", flush=True)
    K, g, p, mach, X, V, U = _setup(d)
    assert rel(host(K.compute_odes(p, mach, g, X, V, U)), d["R"]) <= 1e-12
    assert rel(K.compute_forces(g, X, U).cpu().numpy(), d["F"]) <= 1e-12
    jv = {"jvp_U": dict(dU=dev_cells(d["dU"])), "jvp_X": dict(dX=dev_verts(d["dX"])),
          "jvp_V": dict(dV=dev_verts(d["dV"])), "jvp_Ma": dict(dmach=complex(d["dMa"][0]))}
    for key, kw in jv.items():
        got = host(K.jvp_odes(p, mach, g, X, V, U, **kw))
        assert rel(got, d[key]) <= 1e-10, (key, rel(got, d[key]))
    out = K.vjp_odes(p, mach, g, X, V, U, dev_cells(d["w"]),
                     want=("states", "vertices", "velocities", "mach"))
    for key, k2 in (("vjp_U", "states"), ("vjp_X", "vertices"), ("vjp_V", "velocities")):
        assert rel(host(out[k2]), d[key]) <= 1e-10, (key, rel(host(out[k2]), d[key]))
    assert rel(out["mach"].cpu().numpy(), d["vjp_Ma"]) <= 1e-10
    fj = K.jvp_forces(g, X, U, dX=dev_verts(d["dX"])).cpu().numpy()
    assert rel(fj, d["fjvp_X"]) <= 1e-10
    fj = K.jvp_forces(g, X, U, dU=dev_cells(d["dU"])).cpu().numpy()
    assert rel(fj, d["fjvp_U"]) <= 1e-10
    fv = K.vjp_forces(g, X, U, torch.as_tensor(d["wF"]).cuda())
    assert rel(host(fv["states"]), d["fvjp_U"]) <= 1e-10
    assert rel(host(fv["vertices"]), d["fvjp_X"]) <= 1e-10
    # complex-step blocks assembled on the GPU
    JU, JX, JV, JMa = K.cs_linearize_odes(p, mach, g, X, V, U)
    assert rel(JU.permute(1, 2, 4, 3, 0).cpu().numpy(), d["JU"]) <= 1e-10
    assert rel(JX.permute(1, 2, 4, 3, 0).cpu().numpy(), d["JX"]) <= 1e-10
    assert rel(JV.permute(1, 2, 4, 3, 0).cpu().numpy(), d["JV"]) <= 1e-10
    assert rel(host(JMa)[:, None], d["JMa"]) <= 1e-10


def test_c1_fixture():
    d = _load("c1.npz")
    print("[golden c1.npz]", flush=True)
    K, g, p, mach, X, V, U = _setup(d)
    assert rel(host(K.compute_odes(p, mach, g, X, V, U)), d["R"]) <= 1e-12
    assert rel(K.compute_forces(g, X, U).cpu().numpy(), d["F"]) <= 1e-12
    dirs = c1_directions(d)
    for k in range(dirs.shape[0]):
        dU = dev_cells(dirs[k].astype(complex))
        got = host(K.jvp_odes(p, mach, g, X, V, U, dU=dU)).real
        assert rel(got, d["prods"][k]) <= 1e-10
        got = host(K.cs_jvp_odes(p, mach, g, X, V, U, dU=dU)).real
        assert rel(got, d["prods"][k]) <= 1e-10
