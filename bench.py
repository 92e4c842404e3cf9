#!/usr/bin/env python
"""Benchmark of the spatial-discretization hot path on one MI355X.

Workload (the metric's configuration): NACA 0004 O-grid, 1025 angular x 513 radial vertices
(M = 512 radial, N = 1024 angular cells), chord 1, far field 20 chords, static grid, Ma 0.5,
alpha 0 deg, FP64.  One step = one launch evaluating the residual R(Ma, U, X, V) of a batch of 8
seeded states U_inf (1 + 1e-3 N(0,1)) on that grid (134 MB of states, 134 MB of residuals: more
than the 256 MB Infinity Cache holds with the grid); ``value`` = residual evaluations per second.

Also reported (same JSON line):

* the 64-state batched residual on the 257x129 pitching grid, also beyond the Infinity Cache;

* the roofline of the residual kernel, timed with HIP events on its stream, with its measured
  HBM traffic and FP64 rate from the committed counter summary;
* per-product times of the hand-written tangent (JVP) and adjoint (VJP) against the GPU
  complex-step comparators (directional complex step, and the reference's assembled-block
  products), with their errors, at 257x129 and 1025x513;
* the products along every input (Ma, U, X, V; tangents and adjoints) on the 257x129 pitching
  workload against the complex-step paths;
* second-order gradients (tangent-over-adjoint) at 513x257 against the reference's method,
  central differences of complex-step Jacobian products;
* the 1025x513 steady state (PTC with exact nested-dissection solves) and its 32-shift
  frequency response, batched against one exact solve per shift (time to solution, every
  shift's residual checked);
* 64 O-grids of uneven size in one multi-grid launch against one launch per grid, for the
  residual and for the whole step (residual, JVP and VJP of every grid, one graph);
* the response example's steady state and 32-frequency response on the reference's 129x129 grid
  (batched exact resolvent vs one shift at a time vs the reference's SuperLU);
* the reference's own Fortran (compiled from its source into oracle/_ref) timed on this host's
  cores as the CPU baseline.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
BATCH = 64
HEAD_BATCH = 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-large", action="store_true", help="skip the 1025x513 products")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-freq", action="store_true", help="skip the frequency-response solve")
    ap.add_argument("--no-multigrid", action="store_true", help="skip the 64-grid launch")
    ap.add_argument("--gmres-steps", action="store_true",
                    help="also time 30 unconverged Arnoldi steps of line-preconditioned batched "
                         "GMRES at 1025x513 (step throughput only, not a solve)")
    ap.add_argument("--no-superlu", action="store_true",
                    help="skip timing the reference's host SuperLU solve")
    ap.add_argument("--profile", action="store_true",
                    help="plain launches only (no graphs) for counter collection: a 512 MiB "
                         "copy for calibration, the batched residual, the JVP / VJP / "
                         "complex-step products at 1025x513 and the GMRES line solve, each "
                         "repeated for about 60 ms")
    ap.add_argument("--only-gmres", action="store_true",
                    help="only the 1025x513 batched GMRES section (its JSON on one line), for "
                         "a kernel trace of the Arnoldi steps")
    ap.add_argument("--only-multigrid", action="store_true",
                    help="only the 64-grid section (its JSON on one line), for a kernel trace")
    ap.add_argument("--no-c5-freq", action="store_true",
                    help="skip the 1025x513 steady state and 32-shift frequency response")
    ap.add_argument("--only-c5-freq", action="store_true",
                    help="only the 1025x513 frequency-response section (its JSON on one line)")
    ap.add_argument("--only-head", action="store_true",
                    help="only the headline residual launches (its JSON on one line), for a "
                         "kernel trace of the timed kernel")
    ap.add_argument("--only-freq", action="store_true",
                    help="only the 129x129 frequency-response section without the host SuperLU "
                         "solve (its JSON on one line), for a kernel trace of the ring LU")
    return ap.parse_args()


def make_case(A, Rv, thickness, mach, aoa, alpha, alpha_dot, far, batch, seed=0):
    from pyeulerale_amd.grid import naca_ogrid, pitch
    X0 = naca_ogrid(A, Rv, thickness, far_field=far)
    X, V = pitch(X0, alpha, alpha_dot)
    M, N = Rv - 1, A - 1
    a = np.deg2rad(aoa)
    s = mach * np.sqrt(1.4)
    uinf = np.array([1.0, s * np.cos(a), s * np.sin(a), 1.0 / 0.4 + s * s / 2.0])
    rng = np.random.default_rng(seed)
    U = uinf[None, :, None, None] * (1.0 + 1e-3 * rng.standard_normal((batch, 4, M, N)))
    return X, V, U, M, N


def graph_time_ms(torch, fn, reps=20, warm_ms=50.0):
    """Average GPU time of fn() over `reps` launches captured in one graph (no host gaps), timed
    over 3 replays after the graph has replayed for about warm_ms: the GPU's clock falls while it
    idles (as during capture) and rises over the first ~30 ms of load."""
    fn()
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(2):
            fn()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(reps):
            fn()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    g.replay()
    b.record()
    torch.cuda.synchronize()
    for _ in range(int(np.ceil(warm_ms / max(a.elapsed_time(b), 1e-3)))):
        g.replay()
    a.record()
    for _ in range(3):
        g.replay()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / (3 * reps)


PMC_SUMMARY = ROOT / "profiles" / "pmc_face_kernels_1025x513.md"


def pmc_figures(tm: int, tn: int) -> dict:
    """HBM traffic and FP64 flops per launch of the 8-state 1025x513 residual with tm x tn tiles,
    as measured by the rocprofv3 counter passes summarised in profiles/ (FETCH_SIZE and
    WRITE_SIZE calibrated against a 512 MiB copy; SQ_INSTS_VALU_*_F64).  Empty when the summary
    or the shape's line is absent."""
    out = {}
    try:
        for line in PMC_SUMMARY.read_text().splitlines():
            for key in ("traffic_bytes_per_launch", "fp64_flop_per_launch"):
                if line.startswith(f"{key} k_face_fwd batched residual 1025x513 x8 tile {tm}x{tn}:"):
                    out[key] = float(line.rsplit(":", 1)[1])
    except OSError:
        pass
    return out


def progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rel(a, b):
    import torch
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def products(torch, K, A, Rv, thickness, mach, aoa, alpha, alpha_dot, far, reps):
    """Hand-written JVP/VJP vs complex-step comparators on one grid."""
    X, V, U, M, N = make_case(A, Rv, thickness, mach, aoa, alpha, alpha_dot, far, 1)
    g, p = K.Grid(M, N), K.Params(aoa_deg=aoa)
    dev = lambda a, dt: torch.as_tensor(np.ascontiguousarray(np.swapaxes(a, -1, -2))).to(dt).cuda()
    Xd, Vd, Ud = dev(X, torch.float64), dev(V, torch.float64), dev(U[0], torch.float64)
    rng = np.random.default_rng(1)
    cshape = (4, M, N)
    dU = dev(rng.standard_normal(cshape) + 1j * rng.standard_normal(cshape), torch.complex128)
    w = dev(rng.standard_normal(cshape) + 1j * rng.standard_normal(cshape), torch.complex128)
    out = torch.empty_like(dU)
    ws = K._workspace(g, Ud.device)
    gU = torch.empty_like(dU)
    JU, _, _, _ = K.cs_linearize_odes(p, mach, g, Xd, Vd, Ud, want=("states",))
    tile = K.autotune_vjp(p, mach, g, Xd, Vd, Ud)
    tuned1 = K.autotune_face(p, mach, g, Xd, Vd, Ud, kinds=("jvp", "vjp"), reps=5, batch=1)
    t = {}
    t["jvp"] = graph_time_ms(torch, lambda: K.jvp_odes(p, mach, g, Xd, Vd, Ud, dU=dU, dR=out),
                             reps)
    t["vjp"] = graph_time_ms(torch, lambda: K._lib.call(
        "pea_vjp_odes", K.ctypes.byref(p.c()), float(mach), g.M, g.N, K._ptr(Xd), K._ptr(Vd),
        K._ptr(Ud), K._ptr(w), K._ptr(gU), None, None, None, K._ptr(ws), K._stream()), reps)
    t["cs_dir"] = graph_time_ms(torch, lambda: K.cs_jvp_odes(p, mach, g, Xd, Vd, Ud, dU=dU,
                                                             dR=out), reps)
    t["cs_asm_apply_fwd"] = graph_time_ms(torch, lambda: K.apply_states_fwd(g, JU, dU, out), reps)
    t["cs_asm_apply_rev"] = graph_time_ms(torch, lambda: K.apply_states_rev(g, JU, w, out), reps)
    t["cs_linearize"] = graph_time_ms(torch, lambda: K.cs_linearize_odes(
        p, mach, g, Xd, Vd, Ud, JU=JU, want=("states",)), max(2, reps // 10))
    t["residual"] = graph_time_ms(torch, lambda: K.compute_odes(p, mach, g, Xd, Vd, Ud), reps)
    # 32 products per launch (one per grid z): time per product
    kb = 32
    tuned32 = K.autotune_face(p, mach, g, Xd, Vd, Ud, kinds=("jvp", "vjp"), reps=2, batch=kb)
    db = torch.randn((kb, *g.cell_shape), dtype=torch.complex128, device=Ud.device,
                     generator=torch.Generator(device=Ud.device).manual_seed(2))
    ob = torch.empty_like(db)
    t["jvp_per_product_batch32"] = graph_time_ms(torch, lambda: K.jvp_odes_multi(
        p, mach, g, Xd, Vd, Ud, dU=db, dR=ob), max(2, reps // 5)) / kb
    t["vjp_per_product_batch32"] = graph_time_ms(torch, lambda: K.vjp_odes_states_multi(
        p, mach, g, Xd, Vd, Ud, db, gU=ob), max(2, reps // 5)) / kb
    t["cs_dir_per_product_batch32"] = graph_time_ms(torch, lambda: K.cs_jvp_odes_multi(
        p, mach, g, Xd, Vd, Ud, db, dR=ob), max(2, reps // 5)) / kb
    # the same 32 directions with each tile's primal part evaluated once for all of them
    t["jvp_dirs_per_product_batch32"] = graph_time_ms(torch, lambda: K.jvp_odes_dirs(
        p, mach, g, Xd, Vd, Ud, db, dR=ob), max(2, reps // 5)) / kb
    err_dirs = rel(K.jvp_odes_dirs(p, mach, g, Xd, Vd, Ud, db),
                   K.jvp_odes_multi(p, mach, g, Xd, Vd, Ud, dU=db))
    del db, ob
    # accuracy of the hand-written products against complex step (h = 1e-12)
    jvp = K.jvp_odes(p, mach, g, Xd, Vd, Ud, dU=dU)
    csd = K.apply_states_fwd(g, JU, dU)
    vjp = K.vjp_odes(p, mach, g, Xd, Vd, Ud, w)["states"]
    asm = K.apply_states_rev(g, JU, w)
    C, P = M * N, (M + 1) * (N + 1)
    res = {
        "grid": f"{A}x{Rv}", "cells": C,
        "ms": {k: round(v, 5) for k, v in t.items()},
        "rel_err_jvp_vs_cs_blocks": rel(jvp, csd), "rel_err_vjp_vs_cs_blocks": rel(vjp, asm),
        "rel_err_jvp_vs_cs_dir": rel(jvp, K.cs_jvp_odes(p, mach, g, Xd, Vd, Ud, dU=dU)),
        "vjp_tile": tile,
        "face_tiles_autotuned": {"one_per_launch": {k: v[0] for k, v in tuned1.items()},
                                 "32_per_launch": {k: v[0] for k, v in tuned32.items()}},
        "speedup_vjp_vs_cs_dir": t["cs_dir"] / t["vjp"],
        "speedup_jvp_vs_cs_dir": t["cs_dir"] / t["jvp"],
        "speedup_vjp_vs_cs_asm_rev": t["cs_asm_apply_rev"] / t["vjp"],
        "speedup_vjp_batch32_vs_cs_dir_batch32": t["cs_dir_per_product_batch32"] / t["vjp_per_product_batch32"],
        "speedup_jvp_batch32_vs_cs_dir_batch32": t["cs_dir_per_product_batch32"] / t["jvp_per_product_batch32"],
        "speedup_jvp_dirs_batch32_vs_cs_dir_batch32":
            t["cs_dir_per_product_batch32"] / t["jvp_dirs_per_product_batch32"],
        "rel_err_jvp_dirs_vs_dual_batch32": err_dirs,
        # bytes per product with 32 directions per launch: the direction read and the product
        # written (2 x 64 B per cell), the state and geometry (32 B per cell and vertex) once
        "jvp_dirs_hbm_frac": (128 * C + (32 * C + 32 * P) / kb) /
                             (t["jvp_dirs_per_product_batch32"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "speedup_vjp_vs_cs_linearize_plus_rev": (t["cs_linearize"] + t["cs_asm_apply_rev"]) / t["vjp"],
        "jvp_hbm_frac": (160 * C + 32 * P) / (t["jvp"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "vjp_hbm_frac": (160 * C + 32 * P) / (t["vjp"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    return res


def freq_response(torch, superlu: bool, shifts: int = 32, sequential: bool = True):
    """Steady state and 32-frequency pitch response on the reference's own 129x129 NACA0004 grid
    (examples/response/run.py, Ma 0.5, chord 10): all shifts in one batched ring-LU solve
    against one shift at a time (unless `sequential` is False), plus the reference's host
    SuperLU solve of one shift."""
    from pyeulerale_amd.core import SpatialDiscretization
    from pyeulerale_amd.drivers import frequency_response, pseudo_transient_continuation
    grid = ROOT / "tests" / "golden" / "examples" / "naca0004_129x129.plot3d.gz"
    sd = SpatialDiscretization(grid_file=grid, mach_number=0.5, angle_of_attack=0.0,
                               coefficient_length=10.0, coefficient_center=(0.0, 0.0))
    sync = torch.cuda.synchronize
    t0 = time.perf_counter()
    hist = pseudo_transient_continuation(sd, rtol=1e-9, max_iter=100, log=None)
    sync()
    t_ptc = time.perf_counter() - t0
    k = np.logspace(-2, 1, shifts)
    frequency_response(sd, k[:2], chord=10.0, axis_location=0.0, log=None)   # library warm-up
    sync()
    t0 = time.perf_counter()
    _, cl_b, cm_b = frequency_response(sd, k, chord=10.0, axis_location=0.0, log=None)
    sync()
    t_b = time.perf_counter() - t0
    if not sequential:
        return {"grid": "naca0004_129x129 (reference examples/grids), M=128, N=128",
                "shifts": shifts, "ptc_iterations": len(hist), "ptc_s": t_ptc, "batched_s": t_b,
                "cl_k0.01": [float(cl_b[0].real), float(cl_b[0].imag)]}
    t0 = time.perf_counter()
    _, cl_s, cm_s = frequency_response(sd, k, chord=10.0, axis_location=0.0, log=None,
                                       batched=False)
    sync()
    t_s = time.perf_counter() - t0
    res = {
        "grid": "naca0004_129x129 (reference examples/grids), M=128, N=128", "shifts": shifts,
        "reduced_frequencies": "logspace(-2, 1, 32)",
        "ptc_iterations": len(hist), "ptc_final_rel_residual": hist[-1][2], "ptc_s": t_ptc,
        "batched_s": t_b, "sequential_s": t_s, "speedup_batched_vs_sequential": t_s / t_b,
        "max_rel_diff_batched_vs_sequential": float(max(
            np.max(np.abs(cl_b - cl_s) / np.abs(cl_s)), np.max(np.abs(cm_b - cm_s) / np.abs(cm_s)))),
        "cl_k0.01": [float(cl_b[0].real), float(cl_b[0].imag)],
    }
    if superlu:
        rhs = sd.apply_odes_wrt_vertices_fwd(sd.vertices.new_zeros(sd.vertices.shape,
                                                                   dtype=torch.complex128) + 1.0)
        t0 = time.perf_counter()
        sd.solve_odes_wrt_states_fwd(rhs, shift=0.3j, solver="superlu")
        res["superlu_one_shift_s"] = time.perf_counter() - t0
        res["speedup_batched_per_shift_vs_superlu"] = res["superlu_one_shift_s"] / (t_b / shifts)
    return res


def residual_beyond_cache(torch, K, p, mach, g, Xd, Vd, Ud, Rd, alpha, alpha_dot, reps=20):
    """The batched residual where consecutive launches cannot be served by the 256 MB Infinity
    Cache: (a) 256 states per launch (537 MB of states and residuals), (b) the 64-state launch
    with a 1 GiB buffer written between launches, outside the timed events.  The 64- and
    256-state launches are timed back to back from graphs; the flushed launches need a write
    between launches, so they are timed one by one with HIP events, and so is the unflushed
    64-state launch beside them (per-launch events add a few microseconds to each).  The tile
    shape is the one the project's autotuner picks for the 64-state launch."""
    tuned = K.autotune_face(p, mach, g, Xd, Vd, Ud[0].contiguous(), kinds=("residual",), reps=5,
                            batch=Ud.shape[0], warmup=100)["residual"]
    def timed(fn, flush=None):
        evs = []
        for _ in range(reps):
            if flush is not None:
                flush()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            evs.append((a, b))
        torch.cuda.synchronize()
        return float(np.mean([a.elapsed_time(b) for a, b in evs]))
    big = 256
    _, _, U, M, N = make_case(257, 129, 0.04, mach, 0.0, alpha, alpha_dot, 20.0, big, seed=1)
    Ub = torch.as_tensor(np.ascontiguousarray(np.swapaxes(U, -1, -2))).cuda()
    Rb = torch.empty_like(Ub)
    fn64 = lambda: K.compute_odes_batch(p, mach, g, Xd, Vd, Ud, Rd)
    fn256 = lambda: K.compute_odes_batch(p, mach, g, Xd, Vd, Ub, Rb)
    junk = torch.empty(2 ** 27, dtype=torch.float64, device="cuda")
    flush = lambda: junk.fill_(1.0)
    for _ in range(3):
        fn64(), fn256()
    ms64 = graph_time_ms(torch, fn64, reps)
    ms256 = graph_time_ms(torch, fn256, reps // 2)
    ms64e = timed(fn64)
    ms64f = timed(fn64, flush)
    b64 = Ud.shape[0]
    return {"grid": "257x129", "states_per_launch": [b64, big],
            "tile": K.face_tile(0, g.M, g.N)[1:], "tile_autotuned_ms": tuned[1],
            "ms_per_launch_64": ms64, "ms_per_launch_256": ms256,
            "ms_per_launch_64_events": ms64e, "ms_per_launch_64_cache_flushed": ms64f,
            "states_per_s_64": b64 / (ms64 * 1e-3),
            "states_per_s_256": big / (ms256 * 1e-3),
            "cache_flush_slowdown": ms64f / ms64e}


def profile_launches(torch, K):
    """Kernel launches for hardware-counter passes (rocprofv3 --pmc): every dispatch is a plain
    launch on torch's current stream.  Device-to-device copies of 512 MiB come first; their
    FETCH_SIZE / WRITE_SIZE give the calibration of the byte counters.  Then the face kernels
    (batched residual at 257x129 and at 1025x513 with 8 states; residual, JVP, VJP and the
    complex step at 1025x513; the tangent and adjoint of 32 directions per launch, and the
    tangent with the primal part once per tile) and the radial-line preconditioner solve of
    batched GMRES (32 shifts, 1025x513).  The 8-state 1025x513 residual runs with every tile
    shape of the residual, the others with the default shape.  The inputs are built first; each
    group of launches is repeated for about 60 ms before its last 5 (3 for the 32-direction
    launches, 1 for the block assembly), so that the GPU's clock, which falls while the GPU
    idles and rises over the first ~30 ms of load, has risen; the counts per launch do not
    depend on the clock."""
    src = torch.empty(64 * 2 ** 20, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    del src, dst

    dev = lambda a: torch.as_tensor(np.ascontiguousarray(np.swapaxes(a, -1, -2))).cuda()
    mach2, amp = 0.6, np.deg2rad(0.5)
    alpha, alpha_dot = amp * np.sin(0.4), amp * 0.2 * np.cos(0.4)
    X, V, U, M2, N2 = make_case(257, 129, 0.04, mach2, 0.0, alpha, alpha_dot, 20.0, BATCH)
    g2, p2 = K.Grid(M2, N2), K.Params(aoa_deg=0.0)
    Xd2, Vd2, Ud2 = dev(X), dev(V), dev(U)
    Rd2 = torch.empty_like(Ud2)
    X, V, U, M, N = make_case(1025, 513, 0.04, 0.5, 0.0, 0.0, 0.0, 20.0, HEAD_BATCH)
    g, p = K.Grid(M, N), K.Params(aoa_deg=0.0)
    Xd, Vd, Ud8 = dev(X), dev(V), dev(U)
    Rd = torch.empty_like(Ud8)
    Ud = Ud8[0].contiguous()
    rng = np.random.default_rng(1)
    cshape = (4, M, N)
    dU = dev(rng.standard_normal(cshape) + 1j * rng.standard_normal(cshape))
    w = dev(rng.standard_normal(cshape) + 1j * rng.standard_normal(cshape))
    out = torch.empty_like(dU)
    R1 = torch.empty(g.cell_shape, dtype=torch.float64, device="cuda")
    gen = torch.Generator(device="cuda").manual_seed(3)
    d32 = torch.randn((32, *g.cell_shape), dtype=torch.complex128, device="cuda", generator=gen)
    o32 = torch.empty_like(d32)
    torch.cuda.synchronize()

    for _ in range(1000):
        K.compute_odes_batch(p2, mach2, g2, Xd2, Vd2, Ud2, Rd2)
    for t in range(K._lib.load().pea_num_face_tiles()):   # every residual tile shape
        K.set_face_tile(0, g.M, g.N, t)
        for _ in range(500):
            K.compute_odes_batch(p, 0.5, g, Xd, Vd, Ud8, Rd)
    K.set_face_tile(0, g.M, g.N, -1)
    for _ in range(150):
        K.compute_odes(p, 0.5, g, Xd, Vd, Ud, R1)
        K.jvp_odes(p, 0.5, g, Xd, Vd, Ud, dU=dU, dR=out)
        K.vjp_odes(p, 0.5, g, Xd, Vd, Ud, w)
        K.cs_jvp_odes(p, 0.5, g, Xd, Vd, Ud, dU=dU, dR=out)
    # 32 directions per launch: one per grid z (Dual tangent, face adjoint), and all of them per
    # workgroup with the primal part once per tile
    for _ in range(12):
        K.jvp_odes_multi(p, 0.5, g, Xd, Vd, Ud, dU=d32, dR=o32)
        K.vjp_odes_states_multi(p, 0.5, g, Xd, Vd, Ud, d32, gU=o32)
        K.jvp_odes_dirs(p, 0.5, g, Xd, Vd, Ud, d32, dR=o32)
    del d32, o32, Rd, Ud8, Rd2, Ud2
    # radial-line preconditioner of batched GMRES: factorisation and solves for 32 shifts
    from pyeulerale_amd.gmres import BatchedGMRES
    JU = K.cs_linearize_odes(p, 0.5, g, Xd, Vd, Ud, want=("states",))[0]
    gm = BatchedGMRES(g, JU, 1j * np.logspace(1, 2, 32), restart=2)
    gen = torch.Generator(device="cuda").manual_seed(0)
    rr = torch.randn((32, *g.cell_shape), dtype=torch.complex128, device="cuda", generator=gen)
    zz = torch.empty_like(rr)
    torch.cuda.synchronize()
    for _ in range(15):
        gm.psolve(rr, zz)
    torch.cuda.synchronize()


def all_input_products(torch, K, reps: int = 20):
    """Products along every input of R(Ma, U, X, V) on the 257x129 pitching workload (one
    state): the hand-written tangent along (Ma, U, X, V) together and along X alone, the adjoint
    to all four, against the directional complex step and the complex-step block assembly (the
    reference's linearize, f90:552-1143); errors against the complex-step blocks."""
    mach, aoa = 0.6, 0.0
    amp = np.deg2rad(0.5)
    alpha, alpha_dot = amp * np.sin(0.4), amp * 0.2 * np.cos(0.4)
    X, V, U, M, N = make_case(257, 129, 0.04, mach, aoa, alpha, alpha_dot, 20.0, 1)
    g, p = K.Grid(M, N), K.Params(aoa_deg=aoa)
    dev = lambda a, dt: torch.as_tensor(np.ascontiguousarray(np.swapaxes(a, -1, -2))).to(dt).cuda()
    Xd, Vd, Ud = dev(X, torch.float64), dev(V, torch.float64), dev(U[0], torch.float64)
    rng = np.random.default_rng(2)
    cr = lambda shape: dev(rng.standard_normal(shape) + 1j * rng.standard_normal(shape),
                           torch.complex128)
    dU, dX, dV, w = cr((4, M, N)), cr((2, M + 1, N + 1)), cr((2, M + 1, N + 1)), cr((4, M, N))
    dMa = complex(rng.standard_normal(), rng.standard_normal())
    dMa_d = torch.tensor([dMa], dtype=torch.complex128, device="cuda")   # resident: graphs
    JU, JX, JV, JMa = K.cs_linearize_odes(p, mach, g, Xd, Vd, Ud)
    out = torch.empty_like(dU)
    vout = {"states": torch.empty_like(dU), "vertices": torch.empty_like(dX),
            "velocities": torch.empty_like(dV),
            "mach": torch.empty(1, dtype=torch.complex128, device="cuda")}
    allin = dict(dmach=dMa_d, dX=dX, dV=dV, dU=dU)
    want = ("states", "vertices", "velocities", "mach")
    t = {
        "jvp_all_inputs": graph_time_ms(torch, lambda: K.jvp_odes(p, mach, g, Xd, Vd, Ud,
                                                                  dR=out, **allin), reps),
        "jvp_vertices": graph_time_ms(torch, lambda: K.jvp_odes(p, mach, g, Xd, Vd, Ud, dX=dX,
                                                                dR=out), reps),
        "vjp_all_outputs": graph_time_ms(torch, lambda: K.vjp_odes(p, mach, g, Xd, Vd, Ud, w,
                                                                   want=want, out=vout), reps),
        "cs_dir_all_inputs": graph_time_ms(torch, lambda: K.cs_jvp_odes(
            p, mach, g, Xd, Vd, Ud, dR=out, **allin), reps),
        "cs_dir_vertices": graph_time_ms(torch, lambda: K.cs_jvp_odes(p, mach, g, Xd, Vd, Ud,
                                                                      dX=dX, dR=out), reps),
        "cs_linearize_all": graph_time_ms(torch, lambda: K.cs_linearize_odes(
            p, mach, g, Xd, Vd, Ud, JU=JU, JX=JX, JV=JV, JMa=JMa), max(2, reps // 10)),
    }
    ref_fwd = (K.apply_states_fwd(g, JU, dU) + K.apply_vertices_fwd(g, JX, dX)
               + K.apply_vertices_fwd(g, JV, dV) + JMa.to(torch.complex128) * dMa)
    got = K.vjp_odes(p, mach, g, Xd, Vd, Ud, w, want=want)
    ref_mach = complex((JMa.to(torch.complex128) * w).sum())
    return {
        "grid": "257x129, pitching", "cells": M * N,
        "ms": {k: round(v, 5) for k, v in t.items()},
        "rel_err_jvp_all_vs_cs_blocks": rel(K.jvp_odes(p, mach, g, Xd, Vd, Ud, **allin), ref_fwd),
        "rel_err_vjp_states": rel(got["states"], K.apply_states_rev(g, JU, w)),
        "rel_err_vjp_vertices": rel(got["vertices"], K.apply_vertices_rev(g, JX, w)),
        "rel_err_vjp_velocities": rel(got["velocities"], K.apply_vertices_rev(g, JV, w)),
        "rel_err_vjp_mach": abs(complex(got["mach"][0]) - ref_mach) / abs(ref_mach),
        "speedup_jvp_all_vs_cs_dir": t["cs_dir_all_inputs"] / t["jvp_all_inputs"],
        "speedup_jvp_vertices_vs_cs_dir": t["cs_dir_vertices"] / t["jvp_vertices"],
        "speedup_vjp_all_vs_cs_linearize": t["cs_linearize_all"] / t["vjp_all_outputs"],
    }


def second_order(torch, K, reps: int = 10):
    """Second-order gradients grad_{Ma,U,X} <dR, dR/dU dU> on the NACA0004 513x257 O-grid (far
    field 30 chords, Ma 0.5, alpha 1 deg): the exact tangent-over-adjoint product against the
    reference's method, central differences of complex-step Jacobian products (4 passes of
    linearize + transposed products per call, core.py:718-760), timed on this GPU; the error
    is measured against central differences of the hand-written adjoint."""
    X, V, U, M, N = make_case(513, 257, 0.04, 0.5, 1.0, 0.0, 0.0, 30.0, 1)
    g, p = K.Grid(M, N), K.Params(aoa_deg=1.0)
    dev = lambda a, dt: torch.as_tensor(np.ascontiguousarray(np.swapaxes(a, -1, -2))).to(dt).cuda()
    Xd, Vd, Ud = dev(X, torch.float64), dev(V, torch.float64), dev(U[0], torch.float64)
    rng = np.random.default_rng(3)
    cr = lambda shape: rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    w = dev(cr((4, M, N)), torch.complex128)
    dU = dev(cr((4, M, N)), torch.complex128)
    want = ("states", "vertices", "mach")
    t_hvp = graph_time_ms(torch, lambda: K.hvp_odes(p, 0.5, g, Xd, Vd, Ud, w, dU=dU, want=want),
                          reps)
    JU, JX, _, JMa = K.cs_linearize_odes(p, 0.5, g, Xd, Vd, Ud, want=want)
    t_lin = graph_time_ms(torch, lambda: K.cs_linearize_odes(
        p, 0.5, g, Xd, Vd, Ud, JU=JU, JX=JX, JMa=JMa, want=want), 2)
    t_rev = graph_time_ms(torch, lambda: (K.apply_states_rev(g, JU, w),
                                          K.apply_vertices_rev(g, JX, w)), reps)
    t_ref = 4 * (t_lin + t_rev)
    H = K.hvp_odes(p, 0.5, g, Xd, Vd, Ud, w, dU=dU, want=want)
    # central differences of the adjoint along Re dU and Im dU (step ~1e-9 relative: the biased
    # maximum switches within 1e-8, so larger steps can straddle a switch)
    hstep = 1e-9 * float(torch.linalg.norm(Ud)) / float(torch.linalg.norm(dU.real))
    fd = {}
    for part, d in (("re", dU.real), ("im", dU.imag)):
        a = K.vjp_odes(p, 0.5, g, Xd, Vd, Ud + hstep * d, w, want=want)
        b = K.vjp_odes(p, 0.5, g, Xd, Vd, Ud - hstep * d, w, want=want)
        fd[part] = {k: (a[k] - b[k]) / (2 * hstep) for k in want}
    err = {k: rel(H[k], fd["re"][k] + 1j * fd["im"][k]) for k in ("states", "vertices")}
    return {
        "grid": "513x257", "cells": M * N,
        "ms": {"tangent_over_adjoint": round(t_hvp, 5), "cs_linearize": round(t_lin, 5),
               "cs_block_rev_products": round(t_rev, 5),
               "reference_method_fd_of_cs": round(t_ref, 5)},
        "speedup_vs_reference_method": t_ref / t_hvp,
        "rel_err_vs_central_differences_of_adjoint": err,
    }


def multigrid(torch, K, G: int = 64, seed: int = 0, reps: int = 10):
    """64 O-grids of uneven size in one launch against one launch per grid (both replayed from
    a graph, so no host gaps): NACA 00tt with tt in 04..18, A in 65..513 (odd) angular and R in
    33..257 radial vertices, Ma in U[0.3, 0.6], alpha in U[-2, 3] deg, rigid pitch of amplitude
    <= 1 deg, U = U_inf (1 + 1e-3 N(0,1))."""
    rng = np.random.default_rng(seed)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(np.swapaxes(a, -1, -2))).cuda()
    items, cells = [], 0
    for _ in range(G):
        tt = int(rng.integers(4, 19))
        A = 2 * int(rng.integers(32, 257)) + 1
        Rv = int(rng.integers(33, 258))
        mach, aoa = float(rng.uniform(0.3, 0.6)), float(rng.uniform(-2.0, 3.0))
        amp = np.deg2rad(float(rng.uniform(0.0, 1.0)))
        X, V, U, M, N = make_case(A, Rv, tt / 100.0, mach, aoa, amp * np.sin(0.4),
                                  amp * 0.2 * np.cos(0.4), 20.0, 1, seed=int(rng.integers(1 << 30)))
        items.append((K.Params(aoa_deg=aoa), mach, dev(X), dev(V), dev(U[0])))
        cells += M * N
    # the project's autotuner picks the multi-grid launches' tile shapes for this family
    crand0 = lambda U: torch.randn(U.shape, dtype=torch.complex128, device="cuda")
    tuned = K.autotune_multigrid([it + (crand0(it[4]),) for it in items], reps=3)
    # the one-launch-per-grid comparison uses the same shapes (bitwise equal results)
    for kind, name in enumerate(("residual", "jvp", "vjp")):
        for it in items:
            K.set_face_tile(kind, it[4].shape[2], it[4].shape[1], tuned[name][0])
    mg = K.MultiGrid(items)
    singles = [(p, mach, K.Grid(U.shape[2], U.shape[1]), X, V, U, torch.empty_like(U))
               for p, mach, X, V, U in items]

    def one_by_one():
        for p, mach, g, X, V, U, R in singles:
            K.compute_odes(p, mach, g, X, V, U, R)
    t_multi = graph_time_ms(torch, mg.run, reps)
    t_single = graph_time_ms(torch, one_by_one, reps)
    same = all(torch.equal(R, s[6]) for R, s in zip(mg.outputs, singles))

    # the whole step of every grid -- R, the JVP along a seeded complex dU and the VJP of a
    # seeded complex w -- captured in one graph: one multi-grid launch for R and the JVP and one
    # for the adjoints, against one JVP (with R) and one VJP launch per grid
    crand = lambda U: (torch.randn(U.shape, dtype=torch.float64, device="cuda")
                       + 1j * torch.randn(U.shape, dtype=torch.float64, device="cuda"))
    torch.manual_seed(seed)
    dUs = [crand(U) for _, _, _, _, U in items]
    ws = [crand(U) for _, _, _, _, U in items]
    mgt = K.MultiGrid([it + (dU,) for it, dU in zip(items, dUs)], tangent=True)
    mga = K.MultiGrid([it + (w,) for it, w in zip(items, ws)], adjoint=True)
    adj = [(p, mach, g, X, V, U, w, torch.empty_like(w))
           for (p, mach, g, X, V, U, _), w in zip(singles, ws)]
    jv = [(p, mach, g, X, V, U, dU, torch.empty_like(U), torch.empty_like(dU))
          for (p, mach, g, X, V, U, _), dU in zip(singles, dUs)]

    def vjps():
        for p, mach, g, X, V, U, w, gU in adj:
            K.vjp_odes(p, mach, g, X, V, U, w, out={"states": gU})

    def step_multi():
        mgt.run()
        mga.run()

    def step_single():
        for p, mach, g, X, V, U, dU, R, dR in jv:
            K.jvp_odes(p, mach, g, X, V, U, dU=dU, dR=dR, R=R)
        vjps()
    t_step_multi = graph_time_ms(torch, step_multi, reps)
    t_step_single = graph_time_ms(torch, step_single, reps)
    same_t = all(torch.equal(R, j[7]) and torch.equal(dR, j[8])
                 for (R, dR), j in zip(mgt.outputs, jv))
    same_a = all(torch.equal(gU, a[7]) for gU, a in zip(mga.outputs, adj))
    return {"grids": G, "cells": cells, "tiles": mg.total_tiles,
            "autotuned_tiles": {k: {"tile": v[0], "ms": {str(t): ms for t, ms in v[1].items()}}
                                for k, v in tuned.items()},
            "ms_multi_launch": t_multi, "ms_one_launch_per_grid": t_single,
            "speedup": t_single / t_multi, "cells_per_s_multi": cells / (t_multi * 1e-3),
            "bitwise_equal": bool(same),
            "step_residual_jvp_vjp": {
                "ms_two_multi_grid_launches": t_step_multi,
                "ms_two_launches_per_grid": t_step_single,
                "speedup": t_step_single / t_step_multi,
                "bitwise_equal": bool(same_t and same_a)}}


def gmres_steps(torch, K, S: int = 32, steps: int = 30):
    """Batched GMRES throughput at 1025x513 (where the exact ring LU does not fit one GPU):
    30 Arnoldi steps (graph-replayed after the first cycle, radial-line preconditioner, block
    operator) for 32 shifts s = i*logspace(1, 2, 32) at once, against the same steps for one
    shift at a time; and the batched steps with the matrix-free operator (all 32 tangents in
    one launch).  Time per shift-step; convergence is not the point here."""
    from pyeulerale_amd.gmres import BatchedGMRES
    X, V, U, M, N = make_case(1025, 513, 0.04, 0.5, 0.0, 0.0, 0.0, 30.0, 1)
    g, p = K.Grid(M, N), K.Params(aoa_deg=0.0)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(np.swapaxes(a, -1, -2))).cuda()
    Xd, Vd, Ud = dev(X), dev(V), dev(U[0])
    JU, _, _, _ = K.cs_linearize_odes(p, 0.5, g, Xd, Vd, Ud, want=("states",))
    shifts = 1j * np.logspace(1, 2, S)
    gen = torch.Generator(device="cuda").manual_seed(0)
    b = torch.randn((S, *g.cell_shape), dtype=torch.complex128, device="cuda", generator=gen)

    def run(sh, rhs, operator="blocks"):
        gm = BatchedGMRES(g, JU, sh, restart=steps // 2, maxiter=steps, rtol=0.0,
                          operator=operator, point=(p, 0.5, g, Xd, Vd, Ud))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gm.solve(rhs)
        torch.cuda.synchronize()
        return time.perf_counter() - t0
    run(shifts[:1], b[:1])                         # library warm-up
    t_b = run(shifts, b)
    t_1 = run(shifts[:1], b[:1])
    run(shifts[:1], b[:1], "jvp")
    t_mf = run(shifts, b, "jvp")
    return {"grid": "1025x513", "shifts": S, "arnoldi_steps": steps,
            "ms_per_shift_step_batched": t_b / (S * steps) * 1e3,
            "ms_per_shift_step_single": t_1 / steps * 1e3,
            "speedup_batched_vs_one_shift_at_a_time": (t_1 * S) / t_b,
            "ms_per_shift_step_batched_matrix_free": t_mf / (S * steps) * 1e3}


def freq_response_c5(torch, K, shifts: int = 32, ratio: float = 8.0, budget: int = 40):
    """The batched frequency response at 1025x513: PTC steady state (exact nested-dissection
    solves, examples/response/run.py:89-123), then the pitch transfer functions at 32 shifts
    s = i omega, omega = logspace(-2, 1, 32), right-hand side -(dR/dX + s dR/dV) dX/dalpha
    (run.py:125-176).  Time to solution of all shifts:

    * sequential: one exact nested-dissection factorisation and solve per shift (the order of
      work of the reference's spsolve per shift, core.py:533-565);
    * batched (``nested.solve_shifts_adaptive``): going up in |s|, groups of the unsolved shifts
      within ``ratio`` of the smallest, each solved together by GMRES preconditioned by the
      exact factorisation of its middle shift (stopped once the observed rate predicts more
      than ``budget`` Arnoldi steps); the ratio shrinks after a group that mostly fails, and the
      lightly damped band above is solved one factorisation per shift.

    Every shift's residual is checked with the exact block operator."""
    import ctypes
    from pyeulerale_amd import _lib, workloads as W
    from pyeulerale_amd.core import SpatialDiscretization
    from pyeulerale_amd.drivers import frequency_response, pseudo_transient_continuation
    from pyeulerale_amd.nested import symbolic
    c = W.config("C5")
    t = time.perf_counter()
    st = symbolic(c.M, c.N)
    t_sym = time.perf_counter() - t
    sd = SpatialDiscretization(vertices=c.X, angle_of_attack=0.0, mach_number=0.5,
                               coefficient_length=1.0, coefficient_center=(0.25, 0.0))
    torch.cuda.synchronize()
    t = time.perf_counter()
    hist = pseudo_transient_continuation(sd, rtol=1e-9, max_iter=30, solver="nested", log=None)
    torch.cuda.synchronize()
    t_ptc = time.perf_counter() - t
    progress(f"1025x513 PTC: {len(hist)} iterations, {t_ptc:.1f} s")
    sd.linearize()
    omega = np.logspace(-2, 1, shifts)
    speed = 0.5 * 1.4 ** 0.5
    kred = omega * 0.5 / speed
    g = sd._grid
    JU = sd._get_blocks()["U"]
    sh = torch.as_tensor(1j * omega, dtype=torch.complex128, device="cuda")

    def residuals(dU, rhs):
        x = dU.transpose(-1, -2).contiguous()
        Ax = torch.empty_like(x)
        _lib.call("pea_block_multi", g.M, g.N, len(omega), ctypes.c_void_p(JU.data_ptr()),
                  ctypes.c_void_p(sh.data_ptr()), 0, ctypes.c_void_p(x.data_ptr()),
                  ctypes.c_void_p(Ax.data_ptr()), K._stream())
        b = rhs.transpose(-1, -2).reshape(len(omega), -1)
        return (torch.linalg.vector_norm(Ax.reshape(len(omega), -1) + b, dim=1)
                / torch.linalg.vector_norm(b, dim=1)).cpu().numpy()

    torch.cuda.synchronize()
    t = time.perf_counter()
    _, cl_s, cm_s, dU_s, rhs = frequency_response(sd, kred, 1.0, 0.25, batched=False,
                                                  solver="nested", log=None,
                                                  return_states=True)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t
    res_s = residuals(dU_s, rhs)
    del dU_s
    progress(f"1025x513 sequential: {t_seq:.1f} s")
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, cl_b, cm_b, dU_b, rhs = frequency_response(sd, kred, 1.0, 0.25, batched=True,
                                                  solver="gmres", precond="nested",
                                                  groups="adaptive", ratio=ratio,
                                                  budget=budget, log=None, return_states=True)
    torch.cuda.synchronize()
    t_bat = time.perf_counter() - t
    res_b = residuals(dU_b, rhs)
    inf = sd.last_gmres
    progress(f"1025x513 batched: {t_bat:.1f} s")
    return {
        "grid": "1025x513", "shifts": len(omega), "omega": [float(omega[0]), float(omega[-1])],
        "symbolic_s": t_sym, "factor_bytes": st.factor_bytes(), "factor_flop": st.flops(),
        "ptc": {"iterations": len(hist), "seconds": t_ptc, "rel_residual": hist[-1][2]},
        "seconds_sequential_exact": t_seq, "seconds_batched": t_bat,
        "speedup_batched_vs_sequential": t_seq / t_bat,
        "group_ratio": ratio, "gmres_budget": budget,
        "groups": [{"shifts": len(idx), "anchor_omega": abs(a),
                    "converged": int(gi.converged.sum()), "steps": int(gi.iterations.max())}
                   for idx, a, gi in sd.last_groups],
        "factorisations": len(sd.last_groups) + len(sd.last_exact),
        "solved_one_factorisation_each": len(sd.last_exact),
        "gmres_steps": inf.iterations.tolist(),
        "max_residual_sequential": float(res_s.max()), "max_residual_batched": float(res_b.max()),
        "max_rel_diff_cl": float(np.max(np.abs(cl_b - cl_s) / np.abs(cl_s))),
        "max_rel_diff_cm": float(np.max(np.abs(cm_b - cm_s) / np.abs(cm_s))),
        "cl": [[float(omega[i]), cl_s[i].real, cl_s[i].imag] for i in range(0, len(omega), 4)],
        "cm": [[float(omega[i]), cm_s[i].real, cm_s[i].imag] for i in range(0, len(omega), 4)],
    }


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_worker(args):
    """One host process: residual evaluations of the reference for `seconds` (bounded sample)."""
    kind, seconds, X, V, U, mach, aoa = args
    from oracle.oracle import Oracle, Params
    orc = Oracle("ref" if kind == "reference" else "port")
    prm = Params(aoa=aoa)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds or n == 0:
        orc.compute_odes(prm, mach, X, V, U)
        n += 1
    return n, time.perf_counter() - t0


def cpu_baseline(seconds, X, V, U, mach, aoa, grid):
    """The reference's compiled Fortran compute_odes (oracle/_ref) on the host: one core, like
    the reference's serial build, and one process per core on every core this process may use
    (at most 16), each evaluating the single-state residual for a bounded time.  Run before the
    GPU is initialised, so the worker processes are plain forks of a CPU-only process."""
    import multiprocessing as mp
    from oracle.oracle import Oracle
    kind = "reference" if Oracle.available("ref") else "port"
    Xc, Vc, Uc = (np.asfortranarray(a.astype(complex)) for a in (X, V, U))
    n1, t1 = _cpu_worker((kind, seconds, Xc, Vc, Uc, mach, aoa))
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_worker, [(kind, seconds, Xc, Vc, Uc, mach, aoa)] * cores)
    rate_all = sum(n / t for n, t in res)
    src = ("compiled reference Fortran, f90 compute_odes" if kind == "reference"
           else "C restatement")
    return {"value": rate_all, "unit": "residual evals/s", "cores": cores, "kind": kind,
            "cpu_model": _cpu_model(),
            "sample": f"single-state residual evaluations of the {grid} workload ({src}): "
                      f"{cores} processes, one per core, {sum(n for n, _ in res)} evaluations "
                      f"in {seconds:.0f} s each",
            "one_core": {"value": n1 / t1, "cores": 1,
                         "sample": f"{n1} evaluations on 1 core in {t1:.1f} s"}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus N>1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from pyeulerale_amd import kernels as K
    if args.profile:
        profile_launches(torch, K)
        progress("profile launches done")
        return
    if args.only_gmres:
        print(json.dumps(gmres_steps(torch, K)), flush=True)
        return
    if args.only_multigrid:
        print(json.dumps(multigrid(torch, K)), flush=True)
        return
    if args.only_freq:
        print(json.dumps(freq_response(torch, False, sequential=False)), flush=True)
        return
    if args.only_c5_freq:
        print(json.dumps(freq_response_c5(torch, K)), flush=True)
        return
    # headline workload: NACA0004 1025x513, static grid, Ma 0.5, alpha 0, 8 states per launch
    mach, aoa = 0.5, 0.0
    X, V, U, M, N = make_case(1025, 513, 0.04, mach, aoa, 0.0, 0.0, 20.0, HEAD_BATCH)
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_seconds, X, V, U[0], mach, aoa,
                                                "1025x513")
    if cpu is not None:
        progress("cpu baseline done")
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    g, p = K.Grid(M, N), K.Params(aoa_deg=aoa)
    Xd = torch.as_tensor(np.ascontiguousarray(np.swapaxes(X, -1, -2))).cuda()
    Vd = torch.as_tensor(np.ascontiguousarray(np.swapaxes(V, -1, -2))).cuda()
    Ud = torch.as_tensor(np.ascontiguousarray(np.swapaxes(U, -1, -2))).cuda()
    Rd = torch.empty_like(Ud)
    step = lambda: K.compute_odes_batch(p, mach, g, Xd, Vd, Ud, Rd)
    # The K timed launches are replayed from one HIP graph, and so are the warm-up launches after
    # the first, which runs eagerly so that nothing is loaded during capture.  The events bracket
    # the timed replay on the stream the kernels run on.  The graphs of every residual tile shape
    # are captured first (capture runs nothing); then the project's autotuner times every shape
    # on this grid and batch (300 untimed launches of each before its 5 timed ones) and
    # registers the fastest, and the graphs of that shape replay right after it.  The GPU's clock
    # rises over the first ~30 ms of sustained load (a launch takes ~140 us cold, ~110 us once
    # the clock has risen): the tuner's launches raise it, and the default warm-up of 1000
    # launches (~0.11 s) keeps the timed rate the steady one.  The counter summary in profiles/
    # holds every shape of this launch.
    graph_sets = {}
    for t in range(K._lib.load().pea_num_face_tiles()):
        K.set_face_tile(0, g.M, g.N, t)
        step()
        torch.cuda.synchronize()
        gs = []
        for n in (max(0, args.warmup - 1), args.steps):
            gr = torch.cuda.CUDAGraph()
            if n > 0:
                with torch.cuda.graph(gr):
                    for _ in range(n):
                        step()
            gs.append(gr if n > 0 else None)
        graph_sets[t] = gs
    K.set_face_tile(0, g.M, g.N, -1)
    tuned_head = K.autotune_face(p, mach, g, Xd, Vd, Ud[0].contiguous(), kinds=("residual",),
                                 reps=5, batch=HEAD_BATCH, warmup=300)["residual"]
    head_tile = K.face_tile(0, g.M, g.N)
    assert head_tile[0] == tuned_head[0]
    graphs = graph_sets[head_tile[0]]
    torch.cuda.synchronize()
    if graphs[0] is not None:
        graphs[0].replay()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    graphs[1].replay()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ev0.elapsed_time(ev1) / args.steps
    C, P = M * N, (M + 1) * (N + 1)
    bytes_launch = HEAD_BATCH * 64 * C + 32 * P
    achieved = bytes_launch / (kern_ms * 1e-3) / 1e9
    value = HEAD_BATCH * args.steps / wall
    del Ud, Rd, graphs, graph_sets
    if args.only_head:
        print(json.dumps({"value": value, "ms_per_step": wall / args.steps * 1e3,
                          "kernel_ms": kern_ms, "steps": args.steps,
                          "achieved_GBs": achieved}), flush=True)
        return

    # the 257x129 pitching workload of the products (Ma 0.6, rigid pitch about the quarter chord)
    mach, aoa = 0.6, 0.0
    omega, t_ = 0.2, 2.0
    amp = np.deg2rad(0.5)
    alpha, alpha_dot = amp * np.sin(omega * t_), amp * omega * np.cos(omega * t_)
    X, V, U, M, N = make_case(257, 129, 0.04, mach, aoa, alpha, alpha_dot, 20.0, BATCH)
    g, p = K.Grid(M, N), K.Params(aoa_deg=aoa)
    Xd = torch.as_tensor(np.ascontiguousarray(np.swapaxes(X, -1, -2))).cuda()
    Vd = torch.as_tensor(np.ascontiguousarray(np.swapaxes(V, -1, -2))).cuda()
    Ud = torch.as_tensor(np.ascontiguousarray(np.swapaxes(U, -1, -2))).cuda()
    Rd = torch.empty_like(Ud)
    progress(f"residual 1025x513 x{HEAD_BATCH}: {kern_ms:.4f} ms per launch")
    cache = residual_beyond_cache(torch, K, p, mach, g, Xd, Vd, Ud, Rd, alpha, alpha_dot)
    progress("residual beyond the Infinity Cache done")
    extra = {"residual_beyond_infinity_cache": cache,
             "products_257x129": products(torch, K, 257, 129, 0.04, mach, aoa, alpha,
                                          alpha_dot, 20.0, 50)}
    progress("products 257x129 done")
    if not args.no_large:
        extra["products_1025x513"] = products(torch, K, 1025, 513, 0.04, 0.5, 0.0, 0.0, 0.0,
                                              20.0, 10)
        progress("products 1025x513 done")
    extra["products_all_inputs_257x129"] = all_input_products(torch, K)
    progress("all-input products done")
    if not args.no_large:
        extra["second_order_513x257"] = second_order(torch, K)
        progress("second-order products done")
    if args.gmres_steps:
        extra["gmres_step_throughput_1025x513"] = gmres_steps(torch, K)
        progress("gmres done")
    if not args.no_multigrid:
        extra["multigrid_64_grids"] = multigrid(torch, K)
        progress("multi-grid done")
    if not args.no_freq:
        extra["frequency_response_129x129"] = freq_response(torch, not args.no_superlu)
        progress("frequency response done")
    if not args.no_c5_freq:
        extra["frequency_response_1025x513"] = freq_response_c5(torch, K)
        progress("1025x513 frequency response done")
    pmc = pmc_figures(head_tile[1], head_tile[2])
    line = {
        "metric": "residual evals/s + hand-written VJP vs complex-step JVP speedup (FP64, 1 GPU)",
        "value": value, "unit": "residual evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "NACA0004 O-grid 1025x513 vertices (M=512, N=1024), far field "
                               "20 chords, static, Ma 0.5, alpha 0, batch of 8 seeded states "
                               "per launch (402 MB of states and residuals)",
                   "batch": HEAD_BATCH, "grid": "1025x513"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": pmc.get("traffic_bytes_per_launch"),
                     "kernel": f"k_face_fwd<double, double, {head_tile[1]}, {head_tile[2]}, "
                               f"LdReal, StoreRealB>",
                     "tile_autotuned_ms": {str(k): v for k, v in tuned_head[1].items()},
                     "tile_autotuned_fastest": tuned_head[0],
                     "kernel_ms": kern_ms, "algorithmic_bytes_per_launch": bytes_launch,
                     "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE of the 8-state "
                                       "1025x513 launch with this tile shape, "
                                       "profiles/pmc_face_kernels_1025x513.md",
                     "fp64_tflops": (pmc["fp64_flop_per_launch"] / (kern_ms * 1e-3) / 1e12
                                     if "fp64_flop_per_launch" in pmc else None),
                     "fp64_peak_tflops": FP64_PEAK_TFLOPS},
        # the target is stated at 1025x513; 257x129 when that grid is skipped
        "speedup_vjp_vs_cs_jvp": extra.get("products_1025x513",
                                           extra["products_257x129"])["speedup_vjp_vs_cs_dir"],
        "speedup_grid": "1025x513" if "products_1025x513" in extra else "257x129",
        "extra": extra,
    }
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line))


if __name__ == "__main__":
    main()
