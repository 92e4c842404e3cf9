"""torch.autograd over the hand-written products: residual R(Ma, U, X, V) and forces F(U, X) as
differentiable functions with exact first and second derivatives.

* ``odes(U, X, V, mach)`` -> R.  Backward is the adjoint kernel (``pea_vjp_odes``): the
  gradients with respect to U, X, V and Ma in one launch.  That backward is itself
  differentiable: for cotangents a = (a_U, a_X, a_V, a_Ma) of the gradients,
  * the seed w receives the tangent J a (``pea_jvp_odes``);
  * the inputs receive the tangent-over-adjoint d/de [J(z + e a)^T w] (``pea_hvp_odes``),
    which by symmetry of the Hessian of <w, R> is grad_z <a, J^T w>.

  Double backward therefore gives the README's second-order gradients
  grad <dR, dR/dU dU> (README.md:181-215; core.py:718-1275 obtains them by finite differences
  of complex-step Jacobians).
* ``forces(U, X)`` -> F, the same way over ``pea_vjp_forces`` / ``pea_jvp_forces`` /
  ``pea_hvp_forces``.

Tensors are FP64 on the GPU in the library's storage layout: U of shape (4, N, M), X and V of
shape (2, N+1, M+1), R like U, F of shape (2, N), with M radial and N angular cells (the
reference's arrays transposed in their last two axes).  ``mach`` is a 0-d float64 tensor (it
may require grad) or a float.  Derivatives of third order are not provided.
"""

from __future__ import annotations

import torch

from . import kernels as K

F64, C128 = torch.float64, torch.complex128


def _grid(U: torch.Tensor) -> K.Grid:
    if U.dim() != 3 or U.shape[0] != 4:
        raise RuntimeError(f"states must have shape (4, N, M), got {tuple(U.shape)}")
    return K.Grid(U.shape[2], U.shape[1])


def _c(t):
    return None if t is None else t.detach().to(C128).contiguous()


def _re(t, like):
    return None if t is None else t.real.to(like.dtype).reshape(like.shape).contiguous()


class _Odes(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mach, U, X, V, params):
        g = _grid(U)
        ctx.params, ctx.g = params, g
        ctx.save_for_backward(mach, U, X, V)
        return K.compute_odes(params, float(mach), g, X.detach(), V.detach(), U.detach())

    @staticmethod
    def backward(ctx, gR):
        mach, U, X, V = ctx.saved_tensors
        need = ctx.needs_input_grad
        gm, gU, gX, gV = _OdesVjp.apply(gR, mach, U, X, V, ctx.params,
                                        (need[0], need[1], need[2], need[3]))
        return gm, gU, gX, gV, None


class _OdesVjp(torch.autograd.Function):
    """(w, z) -> J(z)^T w for z = (Ma, U, X, V), real w."""

    @staticmethod
    def forward(ctx, w, mach, U, X, V, params, need):
        ctx.set_materialize_grads(False)
        g = _grid(U)
        ctx.params, ctx.g, ctx.need = params, g, need
        ctx.save_for_backward(w, mach, U, X, V)
        want = tuple(k for k, n in zip(("mach", "states", "vertices", "velocities"), need) if n)
        out = K.vjp_odes(params, float(mach), g, X.detach().contiguous(),
                         V.detach().contiguous(), U.detach().contiguous(), _c(w), want=want)
        gm = out["mach"].real.reshape(()).to(F64) if need[0] else None
        return (gm, _re(out.get("states"), U), _re(out.get("vertices"), X),
                _re(out.get("velocities"), V))

    @staticmethod
    def backward(ctx, am, aU, aX, aV):
        w, mach, U, X, V = ctx.saved_tensors
        g, p = ctx.g, ctx.params
        # unused outputs arrive as None; an all-zero direction contributes nothing
        dm = None if am is None else am.detach().reshape(1).to(C128)
        dU, dX, dV = _c(aU), _c(aX), _c(aV)
        if dm is None and dU is None and dX is None and dV is None:
            return None, None, None, None, None, None, None
        Xd, Vd, Ud = X.detach().contiguous(), V.detach().contiguous(), U.detach().contiguous()
        gw = None
        if ctx.needs_input_grad[0]:
            gw = K.jvp_odes(p, float(mach), g, Xd, Vd, Ud, dmach=dm, dX=dX, dV=dV, dU=dU).real
        need = ctx.needs_input_grad[1:5]
        want = tuple(k for k, n in zip(("mach", "states", "vertices", "velocities"), need) if n)
        out = {}
        if want:
            out = K.hvp_odes(p, float(mach), g, Xd, Vd, Ud, _c(w), dmach=dm, dX=dX, dV=dV,
                             dU=dU, want=want)
        gm = out["mach"].real.reshape(()).to(F64) if "mach" in out else None
        return (gw, gm, _re(out.get("states"), U), _re(out.get("vertices"), X),
                _re(out.get("velocities"), V), None, None)


class _Forces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, X):
        g = _grid(U)
        ctx.g = g
        ctx.save_for_backward(U, X)
        return K.compute_forces(g, X.detach().contiguous(), U.detach().contiguous())

    @staticmethod
    def backward(ctx, gF):
        U, X = ctx.saved_tensors
        return _ForcesVjp.apply(gF, U, X, ctx.needs_input_grad[:2])


class _ForcesVjp(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, U, X, need):
        ctx.set_materialize_grads(False)
        g = _grid(U)
        ctx.g, ctx.need = g, need
        ctx.save_for_backward(w, U, X)
        want = tuple(k for k, n in zip(("states", "vertices"), need) if n)
        out = K.vjp_forces(g, X.detach().contiguous(), U.detach().contiguous(), _c(w), want=want)
        return _re(out.get("states"), U), _re(out.get("vertices"), X)

    @staticmethod
    def backward(ctx, aU, aX):
        w, U, X = ctx.saved_tensors
        g = ctx.g
        dU, dX = _c(aU), _c(aX)
        if dU is None and dX is None:
            return None, None, None, None
        Xd, Ud = X.detach().contiguous(), U.detach().contiguous()
        gw = None
        if ctx.needs_input_grad[0]:
            gw = K.jvp_forces(g, Xd, Ud, dX=dX, dU=dU).real
        want = tuple(k for k, n in zip(("states", "vertices"), ctx.needs_input_grad[1:3]) if n)
        out = K.hvp_forces(g, Xd, Ud, _c(w), dX=dX, dU=dU, want=want) if want else {}
        return gw, _re(out.get("states"), U), _re(out.get("vertices"), X), None


def odes(U: torch.Tensor, X: torch.Tensor, V: torch.Tensor, mach, params: K.Params | None = None):
    """Differentiable residual R(Ma, U, X, V) (f90:496-537), storage layout."""
    if not isinstance(mach, torch.Tensor):
        mach = torch.tensor(float(mach), dtype=F64, device=U.device)
    for name, t in (("states", U), ("vertices", X), ("velocities", V)):
        if t.dtype != F64 or t.device.type != "cuda":
            raise RuntimeError(f"{name} must be a float64 GPU tensor")
    return _Odes.apply(mach.to(F64).reshape(()), U.contiguous(), X.contiguous(), V.contiguous(),
                       params or K.Params())


def forces(U: torch.Tensor, X: torch.Tensor):
    """Differentiable surface forces F(U, X) (f90:1156-1174), storage layout."""
    for name, t in (("states", U), ("vertices", X)):
        if t.dtype != F64 or t.device.type != "cuda":
            raise RuntimeError(f"{name} must be a float64 GPU tensor")
    return _Forces.apply(U.contiguous(), X.contiguous())
