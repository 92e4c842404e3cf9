"""Application drivers over :class:`SpatialDiscretization`, the computations of the reference's
example scripts with all frequencies solved in one batch.

* :func:`pseudo_transient_continuation` -- steady state by switched evolution relaxation,
  examples/response/run.py:89-123 (shift 1/dt = 10 * ||R||_inf / ||R_0||_inf);
* :func:`frequency_response` -- transfer functions pitch angle -> (c_l, c_m),
  examples/response/run.py:125-176, every reduced frequency in one batched resolvent solve;
* :func:`static_gain` -- dc_l/dalpha and its Mach derivative through adjoint and second-order
  products, examples/gain/run.py:80-140.
"""

from __future__ import annotations

import gzip
import math
from typing import Callable

import numpy as np
import torch

from . import HEAT_RATIO
from .grid import rotation_gain

C128 = torch.complex128


def pseudo_transient_continuation(sd, rtol: float = 1e-9, max_iter: int = 100,
                                  solver: str = "direct", log: Callable | None = print,
                                  **solve_opts) -> list[tuple[int, float, float]]:
    """Drives ``sd.states`` to a steady state; returns (iteration, ||R||_inf, relative).

    The shift falls with the residual, so the last steps are Newton steps (shift -> 0).  The
    default ``solver="direct"`` (an exact GPU factorisation, as the reference's sparse LU) converges
    quadratically there.
    ``solver="gmres"`` is exact only while GMRES converges, which for the nearly imaginary
    spectrum of the JST Jacobian needs large shifts (|s| of order 10 on a 65x33 grid); use it
    for the early, strongly shifted steps or with a loose ``rtol`` as an inexact Newton method.
    """
    sd.compute_odes()
    norm0 = float(sd.odes.abs().max())
    if log:
        log(f"\nPseudo-transient continuation:\n{'it':>6} {'residual':>15} {'rel residual':>15}")
    hist = []
    for nt in range(max_iter):
        sd.compute_odes()
        norm = float(sd.odes.abs().max())
        rel = norm / norm0
        hist.append((nt, norm, rel))
        if log:
            log(f"{nt:>6} {norm:>15.1e} {rel:>15.1e}")
        sd.linearize()
        sd.compute_forces()
        if rel <= rtol or math.isnan(rel):
            break
        time_step = 1e-1 * rel ** -1.0
        update = sd.solve_odes_wrt_states_fwd(-sd.odes, shift=time_step ** -1.0, solver=solver,
                                              **solve_opts)
        sd.states[:] += update.real
    return hist


def pitch_gain(sd, axis_location: float) -> torch.Tensor:
    """dX/dalpha of a nose-up pitch about (axis_location, 0) (examples/response/run.py:126-128)."""
    X = sd.vertices.cpu().numpy()
    return torch.as_tensor(rotation_gain(X, (axis_location, 0.0))).to(C128).to(sd.device)


def export_pressure_coefficients(file_name, vertices, pressure_coef, pressure_coef_wrt_aoa) -> None:
    """Writes c_p and its pitch transfer for plotting, gzipped, as the response example does
    (examples/response/run.py:32-64).  Each cell (m, n), in row-major order of (M, N), gets four
    lines ``x z c_p Re(L c_p/L alpha) Im(L c_p/L alpha)`` (its vertices bi, fi, fo, bo) and an
    empty line.  ``vertices`` (2, M+1, N+1), the coefficients (M, N); arrays or tensors."""
    as_np = lambda a: a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    X = as_np(vertices).real
    cp = as_np(pressure_coef).real
    dcp = as_np(pressure_coef_wrt_aoa)
    M, N = cp.shape
    corners = np.stack([X[:, :-1, :-1], X[:, :-1, 1:], X[:, 1:, 1:], X[:, 1:, :-1]], axis=-1)
    rows = np.empty((M, N, 4, 5))
    rows[..., 0] = corners[0]
    rows[..., 1] = corners[1]
    rows[..., 2] = cp[..., None]
    rows[..., 3] = dcp.real[..., None]
    rows[..., 4] = dcp.imag[..., None]
    lines = []
    for cell in rows.reshape(M * N, 4, 5):
        lines.extend(" ".join(f"{v:+.5e}" for v in row) for row in cell)
        lines.append("")
    with gzip.open(file_name, "wt") as fh:
        fh.write("\n".join(lines) + "\n")


def frequency_response(sd, reduced_frequencies, chord: float, axis_location: float,
                       batched: bool = True, solver: str = "direct",
                       log: Callable | None = print, cp_file=None, return_states: bool = False,
                       **solve_opts):
    """Transfer functions from pitch angle to c_l and c_m at the linearization point.

    ``sd`` must be linearized at a steady state and constructed with
    ``coefficient_length=chord`` and ``coefficient_center=(axis_location, 0)``.
    Returns (reduced frequencies, L c_l / L alpha, L c_m / L alpha) as complex arrays, and with
    ``return_states`` also the state responses (S, 4, M, N) and the right-hand sides solved for.
    With ``cp_file`` the steady c_p and its transfer at the last frequency are exported
    (:func:`export_pressure_coefficients`), as the example does.

    ``batched=True`` solves all shifts together: ``solver="direct"`` factorises them in one
    batched ring LU, ``solver="gmres"`` runs batched GMRES and solves the shifts it does not
    converge by the direct method.  ``batched=False`` solves the shifts one at a time, as the
    reference does, with ``solver`` (``"direct"``, ``"superlu"`` or ``"gmres"``).
    """
    k = np.asarray(reduced_frequencies, dtype=float)
    speed = sd.mach_number * HEAT_RATIO ** 0.5
    shifts = 1j * k * speed / (chord / 2)
    dX = pitch_gain(sd, axis_location)
    rX = sd.apply_odes_wrt_vertices_fwd(dX)
    rV = sd.apply_odes_wrt_velocities_fwd(dX)
    rhs = torch.stack([rX + complex(s) * rV for s in shifts])
    if batched:
        dU = -sd.solve_odes_wrt_states_batched(rhs, shifts, solver=solver, **solve_opts)
    else:
        dU = torch.stack([-sd.solve_odes_wrt_states_fwd(rhs[i], shift=complex(s), solver=solver,
                                                        **solve_opts)
                          for i, s in enumerate(shifts)])
    gl = sd.compute_lift_coefficient_wrt_forces().reshape(-1)
    gm = sd.compute_moment_coefficient_wrt_forces().reshape(-1)
    gmx = sd.compute_moment_coefficient_wrt_vertices().reshape(-1)
    fX = sd.apply_forces_wrt_vertices_fwd(dX)
    cl, cm = [], []
    for i in range(len(k)):
        dF = (fX + sd.apply_forces_wrt_states_fwd(dU[i])).reshape(-1)
        cl.append(complex(torch.vdot(gl, dF)))
        cm.append(complex(torch.vdot(gm, dF) + torch.vdot(gmx, dX.reshape(-1))))
    cl, cm = np.array(cl), np.array(cm)
    if cp_file is not None:
        export_pressure_coefficients(
            cp_file, sd.vertices, sd.compute_pressure_coefficients(),
            sd.apply_pressure_coefficients_wrt_states_fwd(dU[-1]))
    if log:
        log(f"\nFrequency response:\n"
            f"{'red frequency':>15} {'lift coefficient':>25} {'moment coefficient':>25}")
        for i in range(len(k)):
            log(f"{k[i]:>15.3e} {cl[i]:>+25.3e} {cm[i]:>+25.3e}")
    if return_states:
        return k, cl, cm, dU, rhs
    return k, cl, cm


def static_gain(sd, solver: str = "direct", log: Callable | None = print, **solve_opts):
    """dc_l/dalpha and d(dc_l/dalpha)/dMa at the linearization point (pitch about the origin),
    examples/gain/run.py:80-140."""
    X = sd.vertices
    dX = torch.empty((2, *X.shape[1:]), dtype=C128, device=sd.device)
    dX[1] = -X[0]
    dX[0] = X[1]
    gF = sd.compute_lift_coefficient_wrt_forces()
    dU_a = -sd.solve_odes_wrt_states_fwd(sd.apply_odes_wrt_vertices_fwd(dX), solver=solver,
                                         **solve_opts)
    cl_a = torch.vdot(gF.reshape(-1), (sd.apply_forces_wrt_vertices_fwd(dX) +
                                       sd.apply_forces_wrt_states_fwd(dU_a)).reshape(-1))
    dU_m = -sd.solve_odes_wrt_states_fwd(
        sd.apply_odes_wrt_mach_fwd(np.asarray([1.0 + 0j])), solver=solver, **solve_opts)
    lam = -sd.solve_odes_wrt_states_adj(sd.apply_forces_wrt_states_rev(gF), solver=solver,
                                        **solve_opts)
    v = lambda a, b: torch.vdot(a.reshape(-1), b.reshape(-1))
    cl_a_m = (
        -2 / sd.mach_number * cl_a
        + v(sd.compute_forces_wrt_states_inner_product_wrt_states(gF, dU_a), dU_m)
        + v(sd.compute_forces_wrt_vertices_inner_product_wrt_states(gF, dX), dU_m)
        + sd.compute_odes_wrt_states_inner_product_wrt_mach(lam, dU_a).reshape(())
        + v(sd.compute_odes_wrt_states_inner_product_wrt_states(lam, dU_m), dU_a)
        + sd.compute_odes_wrt_vertices_inner_product_wrt_mach(lam, dX).reshape(())
        + v(sd.compute_odes_wrt_states_inner_product_wrt_vertices(lam, dU_m), dX)
    )
    cl_a, cl_a_m = complex(cl_a), complex(cl_a_m)
    if log:
        log("\nLift coefficient derivatives:")
        log(f"{'wrt pitch angle:':<36} {cl_a.real:.4f} (1/rad)")
        log(f"{'wrt pitch angle wrt Mach number:':<36} {cl_a_m.real:.4f} (1/rad)")
    return cl_a, cl_a_m
