"""End-to-end runs of the reference's example computations on the reference's own grids, checked
against the outputs its CI pins byte for byte (.github/workflows/tests.yml:47-61):

* examples/response/run.py (Ma 0.05 and 0.5, NACA0004 129x129, chord 10, axis 0): pseudo-
  transient continuation history and the frequency response pitch -> (c_l, c_m) at 11 reduced
  frequencies;
* examples/gain/run.py (Ma 0.1 and 0.5, NACA0012 129x129, chord 10): PTC history, dc_l/dalpha
  and its Mach derivative (second-order products).

Printed figures are compared as printed (2 significant digits for residuals, 4 for the
coefficients).  Residual lines below 1e-10 relative are round-off and only checked against the
tolerance.
"""

from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GOLD = Path(__file__).parent / "golden" / "examples"
_NUM = r"[-+]?\d\.\d+e[-+]\d+"


def parse_out(path: Path):
    ptc, freq, gains = [], [], []
    for line in path.read_text().splitlines():
        t = line.split()
        if len(t) == 3 and t[0].isdigit():
            ptc.append((int(t[0]), t[1], t[2]))
        elif re.fullmatch(rf"\s*{_NUM}\s+{_NUM}{_NUM}j\s+{_NUM}{_NUM}j\s*", line):
            k, cl, cm = t
            freq.append((float(k), complex(cl), complex(cm)))
        elif line.startswith("wrt pitch angle"):
            gains.append(float(t[-2]))
    return ptc, freq, gains


def check_ptc(got_lines, ref_ptc, rtol=1e-9):
    got = [(int(t[0]), t[1], t[2]) for t in (ln.split() for ln in got_lines)
           if len(t) == 3 and t[0].isdigit()]
    assert len(got) == len(ref_ptc), (got, ref_ptc)
    for g, r in zip(got, ref_ptc):
        if float(r[2]) >= 1e-10:
            assert g == r, (g, r)
        else:
            assert float(g[2]) <= rtol, g


def close_printed(a: complex, b: complex, digits: int = 4) -> bool:
    """Equal to within one unit of the last printed digit (real and imaginary parts)."""
    def ok(x, y):
        ulp = 10.0 ** (np.floor(np.log10(max(abs(y), 1e-300))) - (digits - 1))
        return abs(x - y) <= 1.01 * ulp
    return ok(a.real, b.real) and ok(a.imag, b.imag)


def _read_cp(path):
    import gzip
    with gzip.open(path, "rt") as fh:
        lines = [ln for ln in fh.read().splitlines() if ln.strip()]
    return lines, np.array([[float(t) for t in ln.split()] for ln in lines])


def _response(mach: float, cp_file=None):
    from pyeulerale_amd.core import SpatialDiscretization
    from pyeulerale_amd.drivers import frequency_response, pseudo_transient_continuation
    sd = SpatialDiscretization(grid_file=GOLD / "naca0004_129x129.plot3d.gz",
                               mach_number=mach, angle_of_attack=0.0, coefficient_length=10.0,
                               coefficient_center=(0.0, 0.0))
    lines = []
    log = lambda s: (lines.extend(s.splitlines()), print(s, flush=True))
    pseudo_transient_continuation(sd, rtol=1e-9, max_iter=100, log=log)
    k, cl, cm = frequency_response(sd, np.logspace(-2, 0, 11), chord=10.0, axis_location=0.0,
                                   log=log, cp_file=cp_file)
    return sd, lines, k, cl, cm


@pytest.mark.parametrize("case,mach", [("Ma0p05", 0.05), ("Ma0p5", 0.5)])
def test_response_example(case, mach, tmp_path):
    print(f"[examples] response {case}", flush=True)
    ref_ptc, ref_freq, _ = parse_out(GOLD / f"response_{case}.out")
    cp_file = tmp_path / "cp.gz" if mach == 0.5 else None
    sd, lines, k, cl, cm = _response(mach, cp_file)
    check_ptc(lines, ref_ptc)
    assert len(ref_freq) == len(k) == 11
    exact = 0
    for i, (kr, clr, cmr) in enumerate(ref_freq):
        assert abs(k[i] - kr) <= 1e-3 * kr
        assert close_printed(cl[i], clr) and close_printed(cm[i], cmr), (i, cl[i], clr, cm[i], cmr)
        exact += f"{cl[i]:+.3e}" == f"{clr:+.3e}" and f"{cm[i]:+.3e}" == f"{cmr:+.3e}"
    print(f"[examples] {case}: {exact}/11 frequency lines identical as printed", flush=True)
    if cp_file is not None:
        # steady c_p and its transfer at k = 1, against the oracle's (tests/golden/
        # make_cp_golden.py): in memory to 1e-10, and the exported file as printed
        gold = np.load(GOLD / "response_Ma0p5_cp.npz")
        cp = sd.compute_pressure_coefficients().cpu().numpy()
        assert np.abs(cp - gold["cp"]).max() <= 1e-10 * np.abs(gold["cp"]).max()
        _, got = _read_cp(cp_file)
        assert got.shape == (4 * 128 * 128, 5)
        ref = np.stack([np.repeat(gold["cp"].reshape(-1), 4),
                        np.repeat(gold["dcp"].real.reshape(-1), 4),
                        np.repeat(gold["dcp"].imag.reshape(-1), 4)], axis=1)
        ulp = 10.0 ** (np.floor(np.log10(np.maximum(np.abs(ref), 1e-300))) - 5)
        err = np.abs(got[:, 2:] - ref)
        assert (err <= 1.01 * ulp / 2 + 1e-12).all(), float((err / ulp).max())
    if mach == 0.5:
        # batched GMRES with direct fallback and the sequential direct path agree
        from pyeulerale_amd.drivers import frequency_response
        kk = np.array([0.5, 1.0, 4.0])
        _, cl_g, cm_g = frequency_response(sd, kk, chord=10.0, axis_location=0.0, log=None,
                                           solver="gmres", rtol=1e-11)
        _, cl_d, cm_d = frequency_response(sd, kk, chord=10.0, axis_location=0.0, log=None,
                                           batched=False)
        np.testing.assert_allclose(cl_g, cl_d, rtol=1e-7)
        np.testing.assert_allclose(cm_g, cm_d, rtol=1e-7)


@pytest.mark.parametrize("case,mach", [("Ma0p10", 0.1), ("Ma0p50", 0.5)])
def test_gain_example(case, mach):
    from pyeulerale_amd.core import SpatialDiscretization
    from pyeulerale_amd.drivers import pseudo_transient_continuation, static_gain
    print(f"[examples] gain {case}", flush=True)
    ref_ptc, _, ref_gain = parse_out(GOLD / f"gain_{case}.out")
    sd = SpatialDiscretization(grid_file=GOLD / "naca0012_129x129.plot3d.gz",
                               mach_number=mach, angle_of_attack=0.0, coefficient_length=10.0)
    lines = []
    log = lambda s: (lines.extend(s.splitlines()), print(s, flush=True))
    pseudo_transient_continuation(sd, rtol=1e-9, max_iter=100, log=log)
    check_ptc(lines, ref_ptc)
    cl_a, cl_a_m = static_gain(sd, log=log)
    # the reference prints 4 decimals; its Mach derivative comes from finite differences of
    # complex-step Jacobians (core.py:718-1275), this one from exact second-order products
    assert abs(cl_a.real - ref_gain[0]) <= 1.01e-4, (cl_a, ref_gain)
    assert abs(cl_a_m.real - ref_gain[1]) <= 2e-4, (cl_a_m, ref_gain)
