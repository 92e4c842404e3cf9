"""Synthetic workloads of the spatial-discretization path: seeded O-grid cases at the five
benchmark configurations and the family of uneven grids run in one multi-grid launch.

Every case is built from :func:`pyeulerale_amd.grid.naca_ogrid` (cosine-clustered NACA 00tt
O-grid, geometric radial stretching with a first spacing of 1e-3 chord) and
:func:`pyeulerale_amd.grid.pitch` (rigid pitch about the quarter chord with GCL-exact vertex
velocities).  States are the free stream times (1 + noise N(0,1)) from ``default_rng(seed)``.
Arrays are NumPy, in the reference's layout (core.py:94-100): vertices and velocities
``(2, M+1, N+1)``, states ``(batch, 4, M, N)``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import HEAT_RATIO
from .grid import naca_ogrid, pitch


def free_stream(mach: float, aoa_deg: float) -> np.ndarray:
    """U_inf = (1, Ma sqrt(g) cos a, Ma sqrt(g) sin a, 1/(g-1) + (Ma sqrt(g))^2/2)
    (spatial_discretization.f90:250-259)."""
    a = np.deg2rad(aoa_deg)
    s = mach * np.sqrt(HEAT_RATIO)
    return np.array([1.0, s * np.cos(a), s * np.sin(a), 1.0 / (HEAT_RATIO - 1.0) + s * s / 2.0])


@dataclass
class Case:
    """One grid with its flow parameters and a batch of seeded states."""
    X: np.ndarray
    V: np.ndarray
    U: np.ndarray          # (batch, 4, M, N)
    mach: float
    aoa: float
    thickness: float
    far_field: float

    @property
    def M(self) -> int:
        return self.U.shape[-2]

    @property
    def N(self) -> int:
        return self.U.shape[-1]


def case(A: int, Rv: int, thickness: float, mach: float, aoa: float, alpha: float = 0.0,
         alpha_dot: float = 0.0, far_field: float = 20.0, batch: int = 1, seed: int = 0,
         noise: float = 1e-3) -> Case:
    """NACA 00tt O-grid of A angular x Rv radial vertices pitched by ``alpha`` (rad) with rate
    ``alpha_dot`` about the quarter chord, and ``batch`` states U_inf (1 + noise N(0,1))."""
    X0 = naca_ogrid(A, Rv, thickness, far_field=far_field)
    X, V = pitch(X0, alpha, alpha_dot)
    M, N = Rv - 1, A - 1
    rng = np.random.default_rng(seed)
    U = free_stream(mach, aoa)[None, :, None, None] * \
        (1.0 + noise * rng.standard_normal((batch, 4, M, N)))
    return Case(X, V, U, mach, aoa, thickness, far_field)


def pitch_phase(amplitude_deg: float = 0.5, omega: float = 0.2, t: float = 2.0):
    """alpha(t) = amplitude sin(omega t) and its rate, in radians, at a time where both are
    non-zero."""
    amp = np.deg2rad(amplitude_deg)
    return amp * np.sin(omega * t), amp * omega * np.cos(omega * t)


def config(name: str, batch: int = 1, seed: int = 0, noise: float = 1e-3) -> Case:
    """The benchmark configurations by name:

    * ``C1`` NACA0012 129x65, far field 20c, static, Ma 0.5, alpha 1.25 deg;
    * ``C2`` NACA0004 257x129, far field 20c, pitching 0.5 deg sin(0.2 t) at t = 2, Ma 0.6,
      alpha 0;
    * ``C3`` NACA0004 513x257, far field 30c, static, Ma 0.5, alpha 1 deg;
    * ``C5`` NACA0004 1025x513, far field 20c, static, Ma 0.5, alpha 0 (pitching enters the
      frequency response through dX/dalpha only).
    """
    if name == "C1":
        return case(129, 65, 0.12, 0.5, 1.25, batch=batch, seed=seed, noise=noise)
    if name == "C2":
        a, ad = pitch_phase()
        return case(257, 129, 0.04, 0.6, 0.0, a, ad, 20.0, batch, seed, noise)
    if name == "C3":
        return case(513, 257, 0.04, 0.5, 1.0, far_field=30.0, batch=batch, seed=seed,
                    noise=noise)
    if name == "C5":
        return case(1025, 513, 0.04, 0.5, 0.0, batch=batch, seed=seed, noise=noise)
    raise ValueError(f"unknown configuration {name!r}")


@dataclass
class FamilyMember:
    """One grid of the uneven family with its own flow parameters."""
    case: Case
    angular: int
    radial: int


def grid_family(G: int = 64, seed: int = 0) -> list[FamilyMember]:
    """G O-grids of uneven size for one multi-grid launch: NACA 00tt with tt uniform in
    {04..18}, A uniform odd in {65..513} angular and R uniform in {33..257} radial vertices,
    Ma uniform in [0.3, 0.6], alpha uniform in [-2, 3] deg, each rigidly pitching with its own
    amplitude <= 1 deg (phase 0.4 rad, rate 0.2), U = U_inf (1 + 1e-3 N(0,1))."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(G):
        tt = int(rng.integers(4, 19))
        A = 2 * int(rng.integers(32, 257)) + 1
        Rv = int(rng.integers(33, 258))
        mach, aoa = float(rng.uniform(0.3, 0.6)), float(rng.uniform(-2.0, 3.0))
        amp = np.deg2rad(float(rng.uniform(0.0, 1.0)))
        c = case(A, Rv, tt / 100.0, mach, aoa, amp * np.sin(0.4), amp * 0.2 * np.cos(0.4),
                 20.0, 1, seed=int(rng.integers(1 << 30)))
        out.append(FamilyMember(c, A, Rv))
    return out
