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
* :func:`cascade` -- the hierarchical grid polish ``ZeroData::polish`` of
  search/gpu_mertens_stream.cu:381-505: five grid stages (256/512/2,048 zeros with thresholds
  0.895/0.918/0.948, then 641 points of 2.5e-6 on the first 15,000 zeros and 241 points of 5e-7
  on the whole table); a candidate stops at the first thresholded stage whose best |h| is below
  its threshold.
* :func:`polish_single_peak` -- search/path1_kernel_opt.py:94-147: a grid on all zeros, Newton
  from 0 and from the best grid point, the largest |h| kept.
* :func:`fast_polish_batch` -- search/path1_kernel_opt.py:150-221: a 121-point grid (evaluated
  exactly here instead of through degree-16 Taylor moments), then 4 Newton steps for the 25
  largest peaks.
* :func:`scan` -- search/polish.py:51-73: the best |h| on the grid y0 + m step, |m| <= half/step.
* :func:`polish_with_weights` -- polish_peaks_with_weights (search/path1_eval_fast.cpp:58-143) for
  several real weight vectors at once, as search/eval_tech1_600k.py:252-271 runs it per vector.
"""
from __future__ import annotations

import numpy as np
import torch

from . import autograd as ag
from . import evaluator as ev

REFERENCE_SCREEN = ((64, 0.78), (128, 0.855), (256, 0.892))
# Stages of ZeroData::polish (search/gpu_mertens_stream.cu:21-29,381-505):
# (zeros (0: the whole table), half-steps M, step, threshold (None: none), centred on the previous
#  stage's best offset, sets the returned h).  The 641-point stage on the first 15,000 zeros moves
# the point only; the 241-point full-table stage sets both the point and h.
REFERENCE_CASCADE = ((256, 38, 1.5e-4, 0.895, False, True),
                     (512, 228, 2.5e-5, 0.918, False, True),
                     (2048, 60, 1.0e-5, 0.948, True, True),
                     (15000, 320, 2.5e-6, None, True, False),
                     (0, 120, 5e-7, None, True, True))


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
    """Hierarchical grid polish (search/gpu_mertens_stream.cu:381-505) for a batch of candidates.
    Stage s evaluates K2 over its first zeros at the points d = c + m step, |m| <= M (c = the
    previous stage's best offset when centred, else 0; d in FP64 as the reference forms dy),
    each the exact word V + rint(d 2^64); the first point of largest |h| in scan order becomes
    the best (when |h| > 0, as the reference's strict comparison from 0).  A thresholded stage
    whose best |h| is below its threshold ends the candidate there.  Stages that do not set h
    move the point only.  Returns (|h|, h, offset d of the returned point, thresholded stages
    passed) per candidate."""
    n = V.shape[0]
    dev = V.device
    best_h = torch.zeros(n, dtype=torch.float64, device=dev)
    best_d = torch.zeros(n, dtype=torch.float64, device=dev)
    passed = torch.zeros(n, dtype=torch.int32, device=dev)
    alive = torch.arange(n, device=dev)
    for nz, M, step, thr, centred, sets_h in stages:
        if alive.numel() == 0:
            break
        na = alive.numel()
        m = torch.arange(-M, M + 1, device=dev, dtype=torch.float64) * step
        d = m[None, :].expand(na, -1)
        if centred:
            d = best_d[alive][:, None] + m[None, :]
        Vg = ag.offset_points(V[alive].repeat_interleave(2 * M + 1, dim=0), d.reshape(-1))
        nzs = table.n if nz == 0 else min(nz, table.n)
        h, _, _ = table.eval_device(Vg, k, n_zeros=nzs, flags=ev.VALUE)
        h = h.view(na, 2 * M + 1)
        j = torch.argmax(h.abs(), dim=1)          # first maximum in scan order
        hb = h.gather(1, j[:, None]).flatten()
        db = d.gather(1, j[:, None]).flatten()
        upd = hb.abs() > 0
        best_d[alive] = torch.where(upd, db, best_d[alive])
        if sets_h:
            best_h[alive] = torch.where(upd, hb, best_h[alive])
        if thr is not None:
            ok = hb.abs() >= thr
            passed[alive[ok]] += 1
            alive = alive[ok]
    return best_h.abs(), best_h, best_d, passed


def polish_single_peak(table: ev.K2Table, V: torch.Tensor, k: float, n_zeros=None,
                       grid_hw: float = 4e-5, grid_steps: int = 40, newton_steps: int = 8,
                       clamp: float = 2e-5, tol: float = 1e-14):
    """search/path1_kernel_opt.py:94-147 on exact points, for a batch of peaks: h0 over the first
    n_zeros zeros; a (2 grid_steps + 1)-point grid of +-grid_hw over the same zeros (first
    maximum of |h|); Newton (clamp, stop once |step| < tol or |K''| <= 1e-12) from offset 0 and
    from the best grid point; a Newton end point replaces the best so far (h0 first) only when
    strictly larger in |h|.  Returns (h0, h_pol, V_pol, ddy)."""
    n = V.shape[0]
    nz = table.n if n_zeros is None else int(n_zeros)
    h0, _, _ = table.eval_device(V, k, n_zeros=nz, flags=ev.VALUE)
    h0 = h0.clone()
    starts = [torch.zeros(n, dtype=torch.float64, device=V.device)]
    if grid_steps > 0:
        dd = torch.from_numpy(np.linspace(-grid_hw, grid_hw, 2 * grid_steps + 1)).to(V.device)
        hs, _, _ = table.eval_device(shifted(V, dd), k, n_zeros=nz, flags=ev.VALUE)
        j = torch.argmax(hs.view(n, -1).abs(), dim=1)
        starts.append(dd[j])
    best_h, best_V = h0, V
    for d0 in starts:
        Vs = ag.offset_points(V, d0)
        Vf, Kf, _, _ = ag.newton_polish(table, Vs, k, n_steps=newton_steps, clamp=clamp,
                                        n_zeros=nz, tol=tol, keep_start=False)
        better = Kf.abs() > best_h.abs()
        best_h = torch.where(better, Kf, best_h)
        best_V = torch.where(better[:, None], Vf, best_V)
    return h0, best_h, best_V, _delta(V, best_V)


def fast_polish_batch(table: ev.K2Table, V: torch.Tensor, k: float, n_zeros=None,
                      grid_hw: float = 3e-5, grid_points: int = 121, top: int = 25,
                      newton_steps: int = 4, clamp: float = 1e-5, tol: float = 1e-13):
    """search/path1_kernel_opt.py:150-221: h_unpol at offset 0; the grid value of largest |h| per
    peak (every grid point evaluated exactly); for the `top` peaks of largest grid |h|, Newton
    from their grid point, kept when larger in |h| than h_unpol, else h_unpol at offset 0.
    Returns (h_unpol, h_pol, V_pol, ddy)."""
    n = V.shape[0]
    nz = table.n if n_zeros is None else int(n_zeros)
    h_unpol, _, _ = table.eval_device(V, k, n_zeros=nz, flags=ev.VALUE)
    h_unpol = h_unpol.clone()
    grid = torch.from_numpy(np.linspace(-grid_hw, grid_hw, grid_points)).to(V.device)
    vals, _, _ = table.eval_device(shifted(V, grid), k, n_zeros=nz, flags=ev.VALUE)
    vals = vals.view(n, grid_points)
    j = torch.argmax(vals.abs(), dim=1)
    h_pol = vals.gather(1, j[:, None]).flatten()
    ddy = grid[j].clone()
    V_pol = ag.offset_points(V, ddy)
    sel = torch.argsort(-h_pol.abs(), stable=True)[:top]
    if sel.numel():
        Vf, Kf, _, _ = ag.newton_polish(table, V_pol[sel], k, n_steps=newton_steps, clamp=clamp,
                                        n_zeros=nz, tol=tol, keep_start=False)
        take = Kf.abs() > h_unpol[sel].abs()
        h_pol[sel] = torch.where(take, Kf, h_unpol[sel])
        V_pol[sel] = torch.where(take[:, None], Vf, V[sel])
        ddy = _delta(V, V_pol)
    return h_unpol, h_pol, V_pol, ddy


def scan(table: ev.K2Table, V0: torch.Tensor, k: float, half_width: float = 0.25,
         step: float = 2e-5, n_zeros=None, chunk: int = 4000):
    """search/polish.py:51-73 for a batch of seeds: K at V0 + rint(m step 2^64) / 2^64 for
    |m| <= round(half_width / step), scanned in chunks of m; the first point of largest |h|.
    Returns (best h, best V, best m) per seed."""
    n = V0.shape[0]
    nz = table.n if n_zeros is None else int(n_zeros)
    m_half = int(round(half_width / step))
    if m_half * step >= 0.5:
        raise ValueError("scan: the grid must stay within +-0.5 of the seed")
    best_h = torch.zeros(n, dtype=torch.float64, device=V0.device)
    best_m = torch.zeros(n, dtype=torch.int64, device=V0.device)
    have = torch.zeros(n, dtype=torch.bool, device=V0.device)
    for m0 in range(-m_half, m_half + 1, chunk):
        m1 = min(m0 + chunk, m_half + 1)
        m = torch.arange(m0, m1, dtype=torch.int64, device=V0.device)
        h, _, _ = table.eval_device(shifted(V0, m.to(torch.float64) * step), k, n_zeros=nz,
                                    flags=ev.VALUE)
        h = h.view(n, m1 - m0)
        j = torch.argmax(h.abs(), dim=1)
        v = h.gather(1, j[:, None]).flatten()
        better = ~have | (v.abs() > best_h.abs())
        best_h = torch.where(better, v, best_h)
        best_m = torch.where(better, m[j], best_m)
        have = torch.ones_like(have)
    best_V = ag.offset_points(V0, best_m.to(torch.float64) * step)
    return best_h, best_V, best_m


def reference_weight_vectors(gamma: np.ndarray, n_trunc: int = 10_000):
    """The weight vectors of search/eval_tech1_600k.py:243-250, as rows of an array: the JP
    kernel (1 - u) cos(pi u) + sin(pi u) / pi with u = gamma / gamma_max, the Gaussians at
    k = 1.5e-10 and 1e-10, raw (all ones) and the first n_trunc zeros only.  Returns
    (names, array (5, n))."""
    g = np.asarray(gamma, dtype=np.float64)
    u = g / g[-1]
    trunc = np.zeros_like(g)
    trunc[:n_trunc] = 1.0
    rows = [(1.0 - u) * np.cos(np.pi * u) + np.sin(np.pi * u) / np.pi,
            np.exp(-1.5e-10 * g * g), np.exp(-1.0e-10 * g * g), np.ones_like(g), trunc]
    return ["jp", "g15", "g10", "raw", "trunc"], np.stack(rows)


def polish_with_weights(table: ev.K2Table, V: torch.Tensor, wt, half_width: float = 2.5e-4,
                        steps: int = 40, grid_zeros: int = 4096, newton_steps: int = 6,
                        clamp: float = 1e-4, n_zeros=None):
    """search/path1_eval_fast.cpp:58-143 for each row of wt (weight vectors over the zeros,
    n_w <= 8) on exact points: h0 over the first n_zeros zeros and the (2 steps + 1)-point grid
    over the first grid_zeros, both for all rows in one pass each (k2_eval_weights_device); a
    grid point replaces the start only when strictly larger in |h| (h0 included, scan order);
    then per row Newton on K'/K'' (clamp, stop once |step| < 1e-13 or |K''| <= 1e-12) with the
    offset accumulated in FP64 like the reference's ddy, and the larger of |h0| and |h_final|
    kept.  Returns (h_unpol, h_pol, V_pol, ddy): (n_w, n), (n_w, n), (n_w, n, 2), (n_w, n)."""
    n = V.shape[0]
    nz = table.n if n_zeros is None else int(n_zeros)
    wt = torch.as_tensor(np.asarray(wt, dtype=np.float64)).to(V.device)
    if wt.dim() == 1:
        wt = wt[None, :]
    n_w = wt.shape[0]
    h0 = table.eval_weights_device(V, wt, n_zeros=nz, flags=ev.VALUE)[0].clone()
    best_abs = h0.abs()
    best_d = torch.zeros((n_w, n), dtype=torch.float64, device=V.device)
    if steps > 0:
        mm = torch.arange(-steps, steps + 1, device=V.device, dtype=torch.float64)
        dd = (half_width * mm) / steps
        hs = table.eval_weights_device(shifted(V, dd), wt, n_zeros=min(grid_zeros, nz),
                                       flags=ev.VALUE)[0].view(n_w, n, 2 * steps + 1).abs()
        for j in range(2 * steps + 1):
            better = hs[:, :, j] > best_abs
            best_abs = torch.where(better, hs[:, :, j], best_abs)
            best_d = torch.where(better, dd[j].expand(n_w, n), best_d)
    h_pol = torch.empty_like(h0)
    ddy = torch.empty_like(h0)
    V_pol = torch.empty((n_w, n, 2), dtype=V.dtype, device=V.device)
    for r in range(n_w):
        d = best_d[r].clone()
        active = torch.ones(n, dtype=torch.bool, device=V.device)
        row = wt[r:r + 1]
        for _ in range(newton_steps):
            if not bool(active.any()):
                break
            _, fp, fpp = table.eval_weights_device(ag.offset_points(V, d), row, n_zeros=nz,
                                                   flags=ev.VALUE | ev.D1 | ev.D2)
            fp, fpp = fp[0], fpp[0]
            ok = active & (fpp.abs() > 1e-12)
            step = torch.clamp(fp / torch.where(ok, fpp, torch.ones_like(fpp)), -clamp, clamp)
            step = torch.where(ok, step, torch.zeros_like(step))
            d = d - step
            active = ok & (step.abs() >= 1e-13)
        Vf = ag.offset_points(V, d)
        hf = table.eval_weights_device(Vf, row, n_zeros=nz, flags=ev.VALUE)[0][0]
        keep = h0[r].abs() > hf.abs()
        h_pol[r] = torch.where(keep, h0[r], hf)
        ddy[r] = torch.where(keep, torch.zeros_like(d), d)
        V_pol[r] = torch.where(keep[:, None], V, Vf)
    return h0, h_pol, V_pol, ddy
