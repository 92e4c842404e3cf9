"""Shared fixtures of the test suite: synthetic grids/states in the reference's layout and
conversions to the library's canonical GPU storage."""

from __future__ import annotations

import numpy as np

from pyeulerale_amd.grid import naca_ogrid, pitch

GAMMA = 1.4


def free_stream(mach: float, aoa_deg: float) -> np.ndarray:
    a = np.deg2rad(aoa_deg)
    s = mach * np.sqrt(GAMMA)
    return np.array([1.0, s * np.cos(a), s * np.sin(a), 1.0 / (GAMMA - 1.0) + s * s / 2.0])


def case(A: int, Rv: int, thickness=0.12, mach=0.5, aoa=1.25, alpha=0.0, alpha_dot=0.0,
         seed=0, noise=1e-3, far_field=20.0):
    """Grid (A angular x Rv radial vertices), vertex velocities of a rigid pitch and a seeded
    state U_inf * (1 + noise N(0,1)).  Arrays in reference layout, float64 (2|4, M(+1), N(+1))."""
    X0 = naca_ogrid(A, Rv, thickness, far_field=far_field)
    X, V = pitch(X0, alpha, alpha_dot)
    M, N = Rv - 1, A - 1
    rng = np.random.default_rng(seed)
    U = np.empty((4, M, N))
    U[:] = free_stream(mach, aoa)[:, None, None]
    U *= 1.0 + noise * rng.standard_normal(U.shape)
    return X, V, U


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rel(a, b) -> float:
    """Norm-wise relative difference ||a - b|| / ||b||."""
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


# ---------------------------------------------------------------- canonical GPU storage
def dev_cells(a, dtype=None):
    import torch
    t = torch.as_tensor(np.ascontiguousarray(np.swapaxes(np.asarray(a), -1, -2)))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


dev_verts = dev_cells


def host(t) -> np.ndarray:
    """Canonical storage -> reference layout numpy array."""
    return np.swapaxes(t.detach().cpu().numpy(), -1, -2)
