"""torch.autograd wrapper of the HIP evaluator and the Newton polish built on it.

The differentiable input is an FP64 offset ``delta`` (per candidate) from exact 64.64 base
points ``V0``: v = V0 + delta.  A float64 tensor cannot hold v ~ 3.5e18 to sub-unit precision,
so the exact part stays in the integer words and only the small offset flows through autograd.
``V0 + delta`` is formed exactly on the device (delta rounded to 2^-64).

forward:  K2(v) -- the kernel computes K2 and, when a gradient is needed, K2' (and K2'' when the
          caller announces a second derivative, ``second_order=True``) in the same pass (the
          adjoint sums share the phase and sin/cos of the value sum); a second derivative taken
          without that announcement costs one more pass for K2'';
backward: grad_delta = grad_out * K2'; that product is itself differentiable, with
          d/d delta (grad_out * K2') = grad_out * K2'', so torch.autograd.grad(...,
          create_graph=True) gives the diagonal second derivative used by Newton.

Newton polish (:func:`newton_polish`) follows the step rule of the reference CPU evaluator
(search/path1_eval_fast.cpp:103-141): step = K'/K'' clamped to +-clamp, stop once
|step| < 1e-13, require |K''| > 1e-12, keep the start point when it was larger in |K2|.
"""
from __future__ import annotations

import torch

from . import evaluator as ev


def offset_points(V0: torch.Tensor, delta: torch.Tensor) -> torch.Tensor:
    """Exact V0 + rint(delta 2^64) on 64.64 words held as int64 pairs (lo, hi)."""
    d = torch.round(delta.detach().to(torch.float64) * 2.0 ** 64)
    if torch.any(d.abs() >= 2.0 ** 63):
        raise ValueError("offset too large for the exact 64.64 update (|delta| < 0.5 required)")
    di = d.to(torch.int64)
    lo = V0[:, 0]
    hi = V0[:, 1]
    new_lo = lo + di                       # wraps mod 2^64 (two's complement int64)
    # carry/borrow out of the low word, compared as unsigned
    bias = torch.tensor(-(1 << 63), dtype=torch.int64, device=V0.device)
    ulo = lo ^ bias
    unew = new_lo ^ bias
    carry = ((di > 0) & (unew < ulo)).to(torch.int64) - ((di < 0) & (unew > ulo)).to(torch.int64)
    new_hi = hi + carry
    return torch.stack([new_lo, new_hi], dim=1).contiguous()


class _Deriv(torch.autograd.Function):
    """K2'(delta) as a function of delta (values from the forward pass); its derivative is
    K2'' -- from the forward pass when it computed it, else from one more pass at the same exact
    points (third derivatives are not provided)."""

    @staticmethod
    def forward(ctx, delta, dK, d2K, ev_args):
        ctx.have_d2 = d2K is not None
        ctx.ev_args = ev_args
        ctx.save_for_backward(d2K if d2K is not None else dK)
        return dK.clone()

    @staticmethod
    def backward(ctx, g):
        (d2K,) = ctx.saved_tensors
        if not ctx.have_d2:
            table, V, k, n_zeros, mixed = ctx.ev_args
            flags = ev.VALUE | ev.D1 | ev.D2 | (ev.MIXED if mixed else 0)
            _, _, d2K = table.eval_device(V, k, n_zeros=n_zeros, flags=flags)
        return g * d2K, None, None, None


class K2Function(torch.autograd.Function):
    @staticmethod
    def forward(ctx, delta, table, V0, k, n_zeros, mixed, second_order):
        V = offset_points(V0, delta)
        need_d1 = ctx.needs_input_grad[0]
        flags = ev.VALUE | (ev.D1 if need_d1 else 0) | (ev.D2 if need_d1 and second_order else 0)
        flags |= ev.MIXED if mixed else 0
        K, dK, d2K = table.eval_device(V, k, n_zeros=n_zeros, flags=flags)
        ctx.ev_args = (table, V, k, n_zeros, mixed)
        ctx.have_d2 = bool(flags & ev.D2)
        if need_d1:
            ctx.save_for_backward(delta, dK, d2K if ctx.have_d2 else dK)
        return K

    @staticmethod
    def backward(ctx, grad_out):
        delta, dK, d2K = ctx.saved_tensors
        d2 = d2K if ctx.have_d2 else None
        return (grad_out * _Deriv.apply(delta, dK, d2, ctx.ev_args),
                None, None, None, None, None, None)


def k2(delta: torch.Tensor, table: ev.K2Table, V0: torch.Tensor, k: float, n_zeros=None,
       mixed: bool = False, second_order: bool = False) -> torch.Tensor:
    """Differentiable K2 at v = V0 + delta.  second_order=True computes K2'' in the forward pass
    as well (for callers that differentiate twice, e.g. Newton through autograd)."""
    return K2Function.apply(delta, table, V0, float(k), n_zeros, mixed, second_order)


def _second_derivative_graph(delta, table, V0, k, n_zeros, mixed):
    """K2' and K2'' via autograd (first derivative with create_graph, then its derivative)."""
    d = delta.detach().requires_grad_(True)
    K = k2(d, table, V0, k, n_zeros, mixed, second_order=True)
    (g,) = torch.autograd.grad(K.sum(), d, create_graph=True)
    (h,) = torch.autograd.grad(g.sum(), d)
    return K.detach(), g.detach(), h.detach()


def newton_polish(table: ev.K2Table, V0: torch.Tensor, k: float, n_steps: int = 8,
                  clamp: float = 1e-4, n_zeros=None, mixed: bool = False, tol: float = 1e-13,
                  keep_start: bool = True):
    """FP64 Newton on K2' = 0 driven through autograd: step = K'/K'' clamped to +-clamp while
    |K''| > 1e-12, stopping once |step| < tol (path1_eval_fast.cpp:105-129).  Returns
    (V, K, K', steps); with keep_start the start point is kept where it was larger in |K2|."""
    n = V0.shape[0]
    delta = torch.zeros(n, dtype=torch.float64, device=V0.device)
    active = torch.ones(n, dtype=torch.bool, device=V0.device)
    steps = torch.zeros(n, dtype=torch.int32, device=V0.device)
    K0 = None
    K = g = None
    for it in range(n_steps + 1):
        K, g, h = _second_derivative_graph(delta, table, V0, k, n_zeros, mixed)
        if K0 is None:
            K0 = K.clone()
        if it == n_steps:
            break
        ok = active & (h.abs() > 1e-12)
        step = torch.clamp(g / torch.where(ok, h, torch.ones_like(h)), -clamp, clamp)
        step = torch.where(ok, step, torch.zeros_like(step))
        delta = delta - step
        steps += ok.to(torch.int32)
        active = ok & (step.abs() >= tol)
        if not bool(active.any()):
            K, g, h = _second_derivative_graph(delta, table, V0, k, n_zeros, mixed)
            break
    V = offset_points(V0, delta)
    if not keep_start:
        return V, K, g, steps
    keep = K0.abs() > K.abs()
    V = torch.where(keep[:, None], V0, V)
    Kf = torch.where(keep, K0, K)
    return V, Kf, g, steps


def polish(table: ev.K2Table, V0: torch.Tensor, k: float, n_steps: int = 8, clamp: float = 1e-4,
           n_zeros=None, mixed: bool = False):
    """Device Newton polish (k2_polish_device: one fused K/K'/K'' pass + an exact 64.64 update
    per step on the device; the still-moving candidates are read back between steps to compact
    the next pass).  Same rule and return values as :func:`newton_polish`; with
    mixed=True the sums run in the precision form and a candidate stops once |step| < 1e-11, the
    noise floor that form's FP32 rounding leaves in v."""
    V = V0.clone()
    # K0 (the unpolished value) comes from the polish's first pass, which is at V0
    K, dK, d2K, steps, K0 = table.polish_device(V, k, n_steps=n_steps, clamp=clamp,
                                                n_zeros=n_zeros, flags=ev.MIXED if mixed else 0,
                                                start=True)
    keep = K0.abs() > K.abs()
    V = torch.where(keep[:, None], V0, V)
    Kf = torch.where(keep, K0, K)
    return V, Kf, dK, steps
