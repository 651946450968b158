"""Restatements of the reference's Python polish and scan control flow, on exact 64.64 points and
evaluated with the oracle's FP64 tier -- TEST INFRASTRUCTURE ONLY (the checker for
mertens_bound_amd.pipeline; see this package's __init__ for who may import it).

* :func:`polish_single_peak`  search/path1_kernel_opt.py:94-147
* :func:`fast_polish_batch`   search/path1_kernel_opt.py:150-221 (grid values evaluated exactly
                              instead of through the degree-16 Taylor moments: the moment
                              truncation itself is parity unpinned)
* :func:`scan`                search/polish.py:51-73
* :func:`cascade`             search/gpu_mertens_stream.cu:381-505 (ZeroData::polish, five grid
                              stages; stage constants at :21-29 and :474-503)

Every point is the exact 64.64 word V + rint(offset * 2^64); offsets accumulate in FP64 like the
reference's ddy.  Pinned by closed forms on a one-zero table (tests/test_oracle.py), where the
peak and the Newton limit are known exactly.
"""
from __future__ import annotations

import numpy as np

from . import D1, D2, VALUE, eval_f64


def offset(V: np.ndarray, d) -> np.ndarray:
    """Exact V + rint(d 2^64) per row (|d| < 0.5)."""
    d = np.broadcast_to(np.asarray(d, dtype=np.float64), (V.shape[0],))
    out = np.empty_like(V)
    for i, ((lo, hi), x) in enumerate(zip(V, d)):
        v = ((int(hi) << 64) | int(lo)) + int(round(float(x) * 2.0 ** 64))
        v &= (1 << 128) - 1
        out[i, 0] = v & ((1 << 64) - 1)
        out[i, 1] = v >> 64
    return out


def _value(table, V, k, nz):
    return eval_f64(table, V, k, flags=VALUE, n_zeros=nz)[0]


def newton(table, V, k, nz, steps, clamp, tol):
    """The reference loop (path1_kernel_opt.py:117-134): while |f''| > 1e-12, ddy -= clamp(f'/f''),
    stop once |step| < tol; returns (ddy, h at the end point)."""
    n = V.shape[0]
    ddy = np.zeros(n)
    active = np.ones(n, dtype=bool)
    for _ in range(steps):
        if not active.any():
            break
        idx = np.nonzero(active)[0]
        _, fp, fpp = eval_f64(table, offset(V[idx], ddy[idx]), k, flags=VALUE | D1 | D2,
                              n_zeros=nz)
        for j, i in enumerate(idx):
            if abs(fpp[j]) > 1e-12:
                step = max(-clamp, min(clamp, fp[j] / fpp[j]))
                ddy[i] -= step
                if abs(step) < tol:
                    active[i] = False
            else:
                active[i] = False
    return ddy, _value(table, offset(V, ddy), k, nz)


def polish_single_peak(table, V, k, n_zeros=None, grid_hw=4e-5, grid_steps=40, newton_steps=8,
                       clamp=2e-5, tol=1e-14):
    nz = table.n if n_zeros is None else int(n_zeros)
    n = V.shape[0]
    h0 = _value(table, V, k, nz)
    starts = [np.zeros(n)]
    if grid_steps > 0:
        dd = np.linspace(-grid_hw, grid_hw, 2 * grid_steps + 1)
        hs = np.stack([_value(table, offset(V, x), k, nz) for x in dd], axis=1)
        starts.append(dd[np.argmax(np.abs(hs), axis=1)])
    best_h, best_d = h0.copy(), np.zeros(n)
    for d0 in starts:
        Vs = offset(V, d0)
        dd_n, h = newton(table, Vs, k, nz, newton_steps, clamp, tol)
        better = np.abs(h) > np.abs(best_h)
        best_h = np.where(better, h, best_h)
        best_d = np.where(better, d0 + dd_n, best_d)
    return h0, best_h, best_d


def fast_polish_batch(table, V, k, n_zeros=None, grid_hw=3e-5, grid_points=121, top=25,
                      newton_steps=4, clamp=1e-5, tol=1e-13):
    nz = table.n if n_zeros is None else int(n_zeros)
    h_unpol = _value(table, V, k, nz)
    grid = np.linspace(-grid_hw, grid_hw, grid_points)
    vals = np.stack([_value(table, offset(V, x), k, nz) for x in grid], axis=1)
    j = np.argmax(np.abs(vals), axis=1)
    h_pol = vals[np.arange(V.shape[0]), j].copy()
    ddy = grid[j].copy()
    sel = np.argsort(-np.abs(h_pol), kind="stable")[:top]
    if sel.size:
        dd_n, h = newton(table, offset(V[sel], ddy[sel]), k, nz, newton_steps, clamp, tol)
        take = np.abs(h) > np.abs(h_unpol[sel])
        h_pol[sel] = np.where(take, h, h_unpol[sel])
        ddy[sel] = np.where(take, ddy[sel] + dd_n, 0.0)
    return h_unpol, h_pol, ddy


def scan(table, V0, k, half_width=0.25, step=2e-5, n_zeros=None, chunk=4000):
    nz = table.n if n_zeros is None else int(n_zeros)
    m_half = int(round(half_width / step))
    n = V0.shape[0]
    best_h = np.zeros(n)
    best_m = np.zeros(n, dtype=np.int64)
    have = np.zeros(n, dtype=bool)
    for m0 in range(-m_half, m_half + 1, chunk):
        m1 = min(m0 + chunk, m_half + 1)
        m = np.arange(m0, m1)
        h = np.stack([_value(table, offset(V0, float(x) * step), k, nz) for x in m], axis=1)
        j = np.argmax(np.abs(h), axis=1)
        v = h[np.arange(n), j]
        better = ~have | (np.abs(v) > np.abs(best_h))
        best_h = np.where(better, v, best_h)
        best_m = np.where(better, m[j], best_m)
        have[:] = True
    return best_h, best_m


# ZeroData::polish stages (search/gpu_mertens_stream.cu:21-29,474-503): zeros (0 = all), M, step,
# threshold (None = none), centred on the previous best, sets the returned h
CASCADE_STAGES = ((256, 38, 1.5e-4, 0.895, False, True), (512, 228, 2.5e-5, 0.918, False, True),
                  (2048, 60, 1.0e-5, 0.948, True, True), (15000, 320, 2.5e-6, None, True, False),
                  (0, 120, 5e-7, None, True, True))


def cascade(table, V, k, stages=CASCADE_STAGES):
    """ZeroData::polish per candidate, as the reference writes it: in each stage a running best
    |s| starting from 0 is replaced by every strictly larger grid value in scan order
    (:398-409, :425-433, :451-459, :484-490, :494-502); best_dy = center + m step in FP64; a
    thresholded stage with best below its threshold returns (best_h, best_dy) (:410, :435, :461).
    Every grid value is K2 over the stage's zeros at the exact word V + rint(dy 2^64).
    Returns (h, dy, thresholded stages passed) arrays."""
    n = V.shape[0]
    H = np.zeros(n); D = np.zeros(n); P = np.zeros(n, dtype=np.int32)
    for c in range(n):
        best_h = 0.0
        best_dy = 0.0
        for nz, M, step, thr, centred, sets_h in stages:
            nzs = table.n if nz == 0 else min(nz, table.n)
            center = best_dy if centred else 0.0
            dys = np.array([center + m * step for m in range(-M, M + 1)])
            vals = _value(table, offset(np.repeat(V[c:c + 1], dys.size, axis=0), dys), k, nzs)
            best_abs = 0.0
            for dy, s in zip(dys, vals):
                if abs(s) > best_abs:
                    best_abs = abs(s)
                    best_dy = float(dy)
                    if sets_h:
                        best_h = float(s)
            if thr is not None:
                if best_abs < thr:
                    break
                P[c] += 1
        H[c] = best_h
        D[c] = best_dy
    return H, D, P
