"""CPU tests of the host-side pieces: grid generation, PLOT3D I/O, pitching kinematics, the
C ABI library (loads and exports every symbol include/pea.h declares) and argument checks."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from pyeulerale_amd.grid import naca_ogrid, pitch, read_plot3d, rotation_gain, write_plot3d

ROOT = Path(__file__).resolve().parents[1]


def _areas(X):
    d1 = X[:, 1:, 1:] - X[:, :-1, :-1]
    d2 = X[:, 1:, :-1] - X[:, :-1, 1:]
    return 0.5 * (d1[0] * d2[1] - d1[1] * d2[0])


@pytest.mark.parametrize("A,Rv,t", [(129, 65, 0.12), (257, 129, 0.04), (33, 9, 0.18)])
def test_ogrid_properties(A, Rv, t):
    X = naca_ogrid(A, Rv, t, far_field=20.0)
    assert X.shape == (2, Rv, A)
    np.testing.assert_array_equal(X[:, :, 0], X[:, :, -1])           # closed O-grid
    assert np.all(_areas(X) > 0)                                       # CCW cells
    np.testing.assert_allclose(X[:, 0, 0], [1.0, 0.0], atol=1e-15)     # starts at the TE
    assert X[1, 0, 1] < 0                                              # then the lower side
    r_far = np.hypot(X[0, -1] - 0.5, X[1, -1])
    np.testing.assert_allclose(r_far, 20.0, rtol=1e-12)
    first = np.hypot(*(X[:, 1] - X[:, 0]))
    assert 0.5e-3 < np.median(first) < 2e-3
    # thickness of the NACA section at the max-thickness station ~ t
    assert abs(2 * np.abs(X[1, 0]).max() - t) < 0.02 * t + 1e-3


def test_plot3d_round_trip(tmp_path):
    X = naca_ogrid(17, 9, 0.12)
    f = tmp_path / "g.plot3d"
    write_plot3d(f, X)
    np.testing.assert_array_equal(read_plot3d(f), X)
    head = f.read_text().splitlines()[:2]
    assert head == ["1", "9 17"]


def test_pitch_is_rigid_and_velocity_is_its_derivative():
    X0 = naca_ogrid(33, 9, 0.12)
    a, h = 0.03, 1e-6
    X, V = pitch(X0, a, 1.0)
    Xp, _ = pitch(X0, a + h, 1.0)
    Xm, _ = pitch(X0, a - h, 1.0)
    np.testing.assert_allclose(V, (Xp - Xm) / (2 * h), atol=1e-8)
    np.testing.assert_allclose(_areas(X), _areas(X0), rtol=1e-12)
    np.testing.assert_allclose(rotation_gain(X), V)


# ------------------------------------------------------------------ the C ABI
def _declared_symbols():
    txt = (ROOT / "include" / "pea.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pea_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from pyeulerale_amd import _lib
    lib = _lib.load()
    syms = _declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding declares a signature for every exported entry point
    assert set(syms) <= set(_lib._SIGS), set(syms) - set(_lib._SIGS)
    assert lib.pea_version() >= 1


def test_library_rejects_bad_dimensions_without_gpu_work():
    import ctypes
    from pyeulerale_amd import _lib
    lib = _lib.load()
    p = _lib.PeaParams(1.25, 1.0, 1 / 32, 2.0)
    rc = lib.pea_compute_odes(ctypes.byref(p), 0.5, 3, 8, None, None, None, None, None)
    assert rc == -1 and b"M >= 4" in lib.pea_last_error()
    rc = lib.pea_compute_odes(ctypes.byref(p), 0.5, 8, 4, None, None, None, None, None)
    assert rc == -1
    rc = lib.pea_compute_odes(ctypes.byref(p), 0.5, 8, 8, None, None, None, None, None)
    assert rc == -2
    with pytest.raises(_lib.PeaError):
        _lib.call("pea_jvp_odes_multi", ctypes.byref(p), 0.5, 8, 8, 3, *([None] * 9))
    # the block product's 32-bit lane offsets cover grids below 2^24 cells; the pointers are
    # never dereferenced because the check comes first
    dummy = ctypes.c_void_p(256)
    rc = lib.pea_block_multi(4096, 4096, 1, dummy, dummy, 0, dummy, dummy, None)
    assert rc == -1 and b"2^24" in lib.pea_resolvent_last_error()
    # the face kernels' buffer loads use 32-bit byte offsets: 2^25 cells and more are refused
    # before any launch (a complex cell field of 2^25 cells spans 2 GiB)
    d = [dummy] * 6
    for M, N in ((8192, 4096), (4096, 8193)):
        rc = lib.pea_compute_odes(ctypes.byref(p), 0.5, M, N, *d[:4], None)
        assert rc == -1 and b"32-bit buffer offsets" in lib.pea_last_error()
        rc = lib.pea_jvp_odes(ctypes.byref(p), 0.5, M, N, *d[:3], None, None, None, dummy,
                              dummy, None, None)
        assert rc == -1
        rc = lib.pea_vjp_odes(ctypes.byref(p), 0.5, M, N, *d[:4], dummy, None, None, None,
                              None, None)
        assert rc == -1
    assert lib.pea_compute_odes(ctypes.byref(p), 0.5, 8191, 4096, None, None, None, None,
                                None) == -2            # just below the limit: dims accepted


def test_vjp_workspace_size_is_positive():
    from pyeulerale_amd import _lib
    assert _lib.load().pea_vjp_workspace_bytes(128, 256) >= 16
    assert _lib.load().pea_vjp_workspace_bytes(3, 256) == 0


def test_python_layer_checks_shapes_before_launch():
    import torch
    from pyeulerale_amd import kernels as K
    g = K.Grid(8, 16)
    with pytest.raises(RuntimeError):
        K.Grid(3, 16)
    bad = torch.zeros((4, 8, 16), dtype=torch.float64)   # reference shape, not storage
    with pytest.raises(RuntimeError):
        K.compute_odes(K.Params(), 0.5, g, torch.zeros(g.vert_shape, dtype=torch.float64),
                       torch.zeros(g.vert_shape, dtype=torch.float64), bad)


def test_example_fixtures_parse_and_grids_read():
    """The reference's example grids (gzipped PLOT3D) and CI-pinned outputs copied into
    tests/golden/examples/ are readable: the grid shapes of the 129x129 files, and the PTC /
    frequency-response / gain lines of every .out file."""
    from pathlib import Path
    from pyeulerale_amd.grid import read_plot3d
    from test_gpu_examples import parse_out
    gold = Path(__file__).parent / "golden" / "examples"
    for name in ("naca0004_129x129.plot3d.gz", "naca0012_129x129.plot3d.gz"):
        X = read_plot3d(gold / name)
        assert X.shape == (2, 129, 129)
        np.testing.assert_array_equal(X[:, :, 0], X[:, :, -1])   # closed O-grid (App. C)
    ptc, freq, gains = parse_out(gold / "response_Ma0p5.out")
    assert len(ptc) == 12 and len(freq) == 11 and not gains
    assert freq[0][1] == complex("+7.386e+00-3.482e-01j")
    ptc, freq, gains = parse_out(gold / "gain_Ma0p50.out")
    assert len(ptc) == 11 and not freq and gains == [8.2518, 7.1469]


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with status 2 and says so in its JSON line."""
    import json
    import subprocess
    import sys
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_ring_lu_refuses_what_does_not_fit():
    from pyeulerale_amd import _lib
    from pyeulerale_amd.ringlu import bytes_per_shift, check_fits
    assert bytes_per_shift(512, 1024) > 400e9          # 1025x513 vertices: 412 GB per shift
    with pytest.raises(_lib.PeaError, match="412"):
        check_fits(512, 1024, 1, 288 * 10 ** 9)
    check_fits(128, 128, 32, 288 * 10 ** 9)            # 129x129, 32 shifts: 52 GB
    with pytest.raises(_lib.PeaError):
        check_fits(128, 128, 32, 40 * 10 ** 9)


def test_noalloc_adjoint_refuses_grids_beyond_its_scratch():
    """pea_vjp_odes_noalloc keeps the Mach partial sums in a static 2 MiB array: a grid whose
    partials need more is refused before anything is launched."""
    from pyeulerale_amd import _lib
    lib = _lib.load()
    p = _lib.PeaParams(1.25, 1.0, 1 / 32, 2.0)
    d = ctypes.c_void_p(256)
    assert lib.pea_vjp_workspace_bytes(2048, 1024) > 2 * 2 ** 20
    rc = lib.pea_vjp_odes_noalloc(ctypes.byref(p), 0.5, 2048, 1024, d, d, d, d, None, None,
                                  None, d, None)
    assert rc == -1 and b"static scratch" in lib.pea_last_error()
    rc = lib.pea_vjp_odes_noalloc(ctypes.byref(p), 0.5, 8, 8, None, d, d, d, None, None, None,
                                  None, None)
    assert rc == -2


def test_f2py_compatible_module_has_the_reference_procedures():
    """The 20 procedures and 4 module variables of spatial_discretization.f90:63-87, each with
    f2py's visible argument count (the array dimensions are hidden, e.g. f90:498-500)."""
    import inspect
    from pyeulerale_amd.f2py_compat import PROCEDURES, spatial_discretization as disc
    arity = {"set_free_stream_state": 2, "convert_odes_wrt_states": 4, "compute_forces": 3,
             "compute_forces_wrt_states": 3, "compute_forces_wrt_vertices": 3}
    assert len(PROCEDURES) == 20 == len(set(PROCEDURES))
    for name in PROCEDURES:
        fn = getattr(disc, name)
        want = arity.get(name, 5 if name.startswith("compute_odes") else 3)
        assert len(inspect.signature(fn).parameters) == want, name
    for var, default in (("aoa", 0.0), ("jst_k2", 1.0), ("jst_k4", 1 / 32), ("jst_c4", 2.0)):
        a = getattr(disc, var)
        assert a.shape == () and float(a) == default
    disc.aoa[...] = 1.5                    # as core.py:258-263 writes them
    assert disc._params().aoa_deg == 1.5
    disc.aoa[...] = 0.0


def test_face_tile_registry_and_json_cache(tmp_path):
    """The face-kernel tile registry (host side only: no launch) and the tuner's JSON cache."""
    from pyeulerale_amd import _lib
    from pyeulerale_amd import kernels as K
    lib = _lib.load()
    n = lib.pea_num_face_tiles()
    assert n >= 2
    assert K.face_tile(0, 512, 1024) == (0, 16, 16)           # defaults
    assert K.face_tile(2, 512, 1024) == (0, 14, 15)
    assert K.face_tile(2, 8, 10)[0] == -1                      # no adjoint tile fits N = 10
    assert K.face_tile(2, 8, 14)[0] == 2                       # the 18 x 11 tile needs N >= 14
    with pytest.raises(_lib.PeaError):
        K.set_face_tile(2, 8, 14, 0)                           # 14 x 15 needs N >= 18
    with pytest.raises(_lib.PeaError):
        K.set_face_tile(1, 64, 128, n)
    K.set_face_tile(1, 64, 128, 2)
    K.set_face_tile(4, 0, 0, 1)
    path = tmp_path / "tiles.json"
    K.save_tuning(path)
    K.set_face_tile(1, 64, 128, -1)
    K.set_face_tile(4, 0, 0, -1)
    assert K.face_tile(1, 64, 128)[0] == 0 and K.face_tile(4)[0] == 0
    assert K.load_tuning(path) == 2
    assert K.face_tile(1, 64, 128) == (2, 32, 8) and K.face_tile(4)[0] == 1
    K.set_face_tile(1, 64, 128, -1)
    K.set_face_tile(4, 0, 0, -1)


def test_gmres_gives_up_only_when_every_shift_would_miss_the_budget():
    """BatchedGMRES(give_up=True) stops a group once every unconverged shift, reducing its
    residual estimate at the rate of the last check_every steps, would need more than maxiter
    steps in all; converged and inactive shifts do not count."""
    from types import SimpleNamespace
    from pyeulerale_amd.gmres import BatchedGMRES
    ns = SimpleNamespace(rtol=1e-10, check_every=5, maxiter=40)
    hopeless = lambda est, prev, active, total: BatchedGMRES._hopeless(
        ns, np.asarray(est), np.asarray(prev), np.asarray(active), total)
    on = [True, True]
    # 10x per 5 steps from 1e-3 at step 10: 35 more steps, 45 > 40
    assert hopeless([1e-3, 1e-11], [1e-2, 1e-9], on, 10)
    # 100x per 5 steps: 17.5 more steps, 27.5 <= 40
    assert not hopeless([1e-3, 1e-11], [1e-1, 1e-9], on, 10)
    # one shift in reach keeps the group going
    assert not hopeless([1e-3, 1e-3], [1e-2, 1e-1], on, 10)
    # stagnation is hopeless; a shift below rtol or inactive is ignored
    assert hopeless([1e-3, 1e-3], [1e-3, 1e-1], [True, False], 10)
    assert not hopeless([1e-11, 1e-12], [1e-3, 1e-3], on, 10)
