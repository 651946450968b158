"""Candidate pipelines built on the evaluator: pre-screen, grid scans and Newton polish.

* :func:`screen` -- staged screen on the first zeros with shifted copies (k2_screen_device; the
  reference's GPU leaf screen, search/gpu_mertens_stream.cu:186-250), survivors moved to their best
  shift.
* :func:`grid_then_newton` -- the reference's per-peak polish (search/path1_eval_fast.cpp:58-143):
  h0 over all zeros, a (2*steps+1)-point grid of +-half_width on the first grid_zeros zeros (a grid
  point replaces the start only when strictly larger in |h| than the best so far, h0 included,
  scanned in order of m), Newton on all zeros from the best grid point, keep the larger of |h0| and
  |h_final|.  Here every point is an exact 64.64 word and each Newton step is a fresh exact
  evaluation instead of a rotation of precomputed A/B rows.
* :func:`cascade` -- the hierarchical grid polish of search/gpu_mertens_stream.cu:381-505: stages
  of (zeros, half-steps, step, threshold, centred on the previous best?); a candidate stops at the
  first stage whose best |h| is below its threshold.
"""
from __future__ import annotations

import torch

from . import autograd as ag
from . import evaluator as ev

REFERENCE_SCREEN = ((64, 0.78), (128, 0.855), (256, 0.892))
# (zeros, half-steps M, step, threshold, centred on the previous stage's best)
REFERENCE_CASCADE = ((256, 38, 1.5e-4, 0.895, False), (512, 228, 2.5e-5, 0.918, False),
                     (2048, 60, 1.0e-5, 0.948, True))


def shifted(V: torch.Tensor, offsets: torch.Tensor) -> torch.Tensor:
    """V (n, 2) x offsets (m,) -> (n*m, 2) exact 64.64 words V + rint(offset 2^64), row-major."""
    n, m = V.shape[0], offsets.shape[0]
    Vr = V.repeat_interleave(m, dim=0)
    d = offsets.to(torch.float64).repeat(n)
    return ag.offset_points(Vr, d)


def screen(table: ev.K2Table, V: torch.Tensor, k: float = 3e-9, shift_step: float = 0.0014,
           n_shift: int = 3, stages=REFERENCE_SCREEN):
    """Screen V; returns (indices of survivors, survivors moved to their best shift, stage values)."""
    val, best, ok = table.screen_device(V, k, shift_step, n_shift, stages)
    idx = torch.nonzero(ok, as_tuple=False).flatten()
    Vs = ag.offset_points(V[idx], best[idx].to(torch.float64) * shift_step)
    return idx, Vs, val


def grid_then_newton(table: ev.K2Table, V: torch.Tensor, k: float, half_width: float = 2.5e-4,
                     steps: int = 40, grid_zeros: int = 4096, newton_steps: int = 6,
                     clamp: float = 1e-4, n_zeros=None):
    """Returns (h_unpol, h_pol, V_pol, ddy) like polish_peaks_with_weights, with V_pol exact."""
    n = V.shape[0]
    nz_all = table.n if n_zeros is None else int(n_zeros)
    h0, _, _ = table.eval_device(V, k, n_zeros=nz_all, flags=ev.VALUE)
    h0 = h0.clone()
    best_abs = h0.abs()
    best_d = torch.zeros(n, dtype=torch.float64, device=V.device)
    if steps > 0:
        m = torch.arange(-steps, steps + 1, device=V.device, dtype=torch.float64)
        dd = (half_width * m) / steps
        Vg = shifted(V, dd)
        hs, _, _ = table.eval_device(Vg, k, n_zeros=min(grid_zeros, nz_all), flags=ev.VALUE)
        hs = hs.view(n, 2 * steps + 1).abs()
        # first strictly larger value in scan order, starting from |h0|
        for j in range(2 * steps + 1):
            better = hs[:, j] > best_abs
            best_abs = torch.where(better, hs[:, j], best_abs)
            best_d = torch.where(better, dd[j].expand(n), best_d)
    Vb = ag.offset_points(V, best_d)
    Vp = Vb.clone()
    K, dK, d2K, st = table.polish_device(Vp, k, n_steps=newton_steps, clamp=clamp, n_zeros=nz_all)
    keep = h0.abs() > K.abs()
    V_pol = torch.where(keep[:, None], V, Vp)
    h_pol = torch.where(keep, h0, K)
    ddy = _delta(V, V_pol)
    return h0, h_pol, V_pol, ddy


def _delta(V0: torch.Tensor, V1: torch.Tensor) -> torch.Tensor:
    """(V1 - V0) / 2^64 as float64 for nearby points."""
    lo = V1[:, 0] - V0[:, 0]
    bias = torch.tensor(-(1 << 63), dtype=torch.int64, device=V0.device)
    borrow = ((V1[:, 0] ^ bias) < (V0[:, 0] ^ bias)).to(torch.int64)
    hi = V1[:, 1] - V0[:, 1] - borrow
    lo_f = lo.to(torch.float64)
    lo_f = torch.where(lo < 0, lo_f + 2.0 ** 64, lo_f)
    return hi.to(torch.float64) + lo_f * 2.0 ** -64


def cascade(table: ev.K2Table, V: torch.Tensor, k: float, stages=REFERENCE_CASCADE):
    """Hierarchical grid scan.  Returns (best |h| of the last stage reached, signed best h,
    offset of the best point, number of stages passed) per candidate."""
    n = V.shape[0]
    dev = V.device
    center = torch.zeros(n, dtype=torch.float64, device=dev)
    best_h = torch.zeros(n, dtype=torch.float64, device=dev)
    best_d = torch.zeros(n, dtype=torch.float64, device=dev)
    passed = torch.zeros(n, dtype=torch.int32, device=dev)
    alive = torch.arange(n, device=dev)
    for s, (nz, M, step, thr, centred) in enumerate(stages):
        if alive.numel() == 0:
            break
        m = torch.arange(-M, M + 1, device=dev, dtype=torch.float64) * step
        base = ag.offset_points(V[alive], center[alive]) if centred else V[alive]
        Vg = shifted(base, m)
        h, _, _ = table.eval_device(Vg, k, n_zeros=min(nz, table.n), flags=ev.VALUE)
        h = h.view(alive.numel(), 2 * M + 1)
        j = torch.argmax(h.abs(), dim=1)          # first maximum in scan order
        hb = h.gather(1, j[:, None]).flatten()
        db = m[j] + (center[alive] if centred else 0.0)
        best_h[alive] = hb
        best_d[alive] = db
        center[alive] = db
        ok = hb.abs() >= thr
        passed[alive[ok]] = s + 1
        alive = alive[ok]
    return best_h.abs(), best_h, best_d, passed
