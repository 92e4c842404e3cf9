"""O-grids around NACA 00tt airfoils, PLOT3D I/O and rigid pitching kinematics.

Array convention (the reference's, core.py:94-100): vertices have shape ``(2, M+1, N+1)``,
index ``[d, m, n]`` with ``d`` = (x, z), ``m`` the radial vertex index (0 = airfoil surface)
and ``n`` the angular vertex index; the closing column ``n = N`` duplicates ``n = 0``.
"""

from __future__ import annotations

import gzip
from pathlib import Path

import numpy as np


def naca_half_thickness(x: np.ndarray, thickness: float) -> np.ndarray:
    """Symmetric NACA 4-digit half thickness (closed trailing edge, chord 1)."""
    return 5.0 * thickness * (0.2969 * np.sqrt(np.clip(x, 0.0, None)) - 0.1260 * x
                              - 0.3516 * x ** 2 + 0.2843 * x ** 3 - 0.1036 * x ** 4)


def _geometric_ratio(n_int: int, first: float, length: float) -> float:
    """Ratio q > 1 with first*(q**n - 1)/(q - 1) == length (bisection)."""
    if first * n_int >= length:
        return 1.0
    lo, hi = 1.0 + 1e-15, 2.0
    while first * (hi ** n_int - 1.0) / (hi - 1.0) < length:
        hi *= 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if first * (mid ** n_int - 1.0) / (mid - 1.0) < length:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def naca_ogrid(num_angular_pts: int, num_radial_pts: int, thickness: float = 0.12,
               chord: float = 1.0, far_field: float = 20.0, first_spacing: float = 1e-3
               ) -> np.ndarray:
    """Closed O-grid around a NACA 00tt airfoil.

    The surface is cosine-clustered from the trailing edge over the lower side to the leading
    edge and back over the upper side (first angular vertex == last), so every cell
    ``(bi, fi, fo, bo)`` is counter-clockwise.  Radial lines run straight from each surface point
    to a far-field circle of radius ``far_field`` chords about mid-chord, with geometric
    stretching whose first spacing is ``first_spacing`` chords.

    Returns vertices of shape ``(2, num_radial_pts, num_angular_pts)`` (float64).
    """
    A, Rv = int(num_angular_pts), int(num_radial_pts)
    if A < 6 or Rv < 5:
        raise ValueError("need at least 6 angular and 5 radial vertices")
    theta = np.linspace(0.0, 2.0 * np.pi, A)
    xs = 0.5 * (1.0 + np.cos(theta))
    zs = naca_half_thickness(xs, thickness) * np.where(theta <= np.pi, -1.0, 1.0)
    zs[0] = zs[-1] = 0.0
    xf = 0.5 + far_field * np.cos(theta)
    zf = -far_field * np.sin(theta)
    xf[-1], zf[-1] = xf[0], zf[0]
    length = np.hypot(xf - xs, zf - zs)
    q = _geometric_ratio(Rv - 1, first_spacing, float(np.mean(length)))
    k = np.arange(Rv, dtype=float)
    s = k / (Rv - 1) if q == 1.0 else (q ** k - 1.0) / (q ** (Rv - 1) - 1.0)
    X = np.empty((2, Rv, A))
    X[0] = xs[None, :] + s[:, None] * (xf - xs)[None, :]
    X[1] = zs[None, :] + s[:, None] * (zf - zs)[None, :]
    X[:, :, -1] = X[:, :, 0]
    return X * chord


def rotation_gain(vertices: np.ndarray, axis: tuple[float, float] = (0.25, 0.0)) -> np.ndarray:
    """dX/dalpha (per radian) of a nose-up rigid pitch about ``axis``: (z - za, -(x - xa)).

    Same convention as the reference's response driver (examples/response/run.py:126-128).
    """
    g = np.empty_like(vertices)
    g[0] = vertices[1] - axis[1]
    g[1] = -(vertices[0] - axis[0])
    return g


def pitch(vertices0: np.ndarray, alpha: float, alpha_dot: float,
          axis: tuple[float, float] = (0.25, 0.0)) -> tuple[np.ndarray, np.ndarray]:
    """Rigidly pitch a grid nose-up by ``alpha`` (rad) about ``axis``.

    Returns ``(X, V)``: rotated vertices and the exact rigid-rotation vertex velocities
    ``alpha_dot * dX/dalpha`` evaluated at the rotated position (area preserving, so the
    geometric conservation law holds exactly for the continuous motion).
    """
    ca, sa = np.cos(alpha), np.sin(alpha)
    dx = vertices0[0] - axis[0]
    dz = vertices0[1] - axis[1]
    X = np.empty_like(vertices0)
    X[0] = axis[0] + ca * dx + sa * dz
    X[1] = axis[1] - sa * dx + ca * dz
    V = alpha_dot * rotation_gain(X, axis)
    return X, V


def read_plot3d(path: str | Path) -> np.ndarray:
    """Reads a single-block 2-D PLOT3D grid as the reference does (core.py:213-227).

    Format: block count, then ``Mpts Npts``, then all x followed by all z with the angular index
    fastest.  Returns float64 vertices of shape ``(2, Mpts, Npts)``.  A ``.gz`` file is read
    through gzip.
    """
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rt") as fh:
        fh.readline()
        mpts, npts = (int(t) for t in fh.readline().split()[:2])
        vals = np.array(fh.read().split(), dtype=float)
    if vals.size != 2 * mpts * npts:
        raise ValueError(f"{path}: expected {2 * mpts * npts} coordinates, got {vals.size}")
    return vals.reshape((2, mpts, npts))


def write_plot3d(path: str | Path, vertices: np.ndarray) -> None:
    """Writes vertices ``(2, Mpts, Npts)`` in the format read by :func:`read_plot3d`."""
    _, mpts, npts = vertices.shape
    with open(path, "w") as fh:
        fh.write(f"1\n{mpts} {npts}\n")
        np.savetxt(fh, np.asarray(vertices, dtype=float).reshape(-1), fmt="%.17e")
