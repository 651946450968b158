"""CPU oracle for the K2 evaluator -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may import this
package, and only as the checker (or as the timed CPU baseline).  The product path in
``mertens_bound_amd`` never imports it.

Pieces:

* ``libk2oracle.so`` (k2_oracle.c): exact integer phase words, the FP64/libm restatement of the
  reference CPU evaluator (search/path1_eval_fast.cpp:15-53,77-81), a binary128 accuracy oracle,
  the Newton polish rule (path1_eval_fast.cpp:105-141) and reference-format restatements of
  ``compute_peak_AB`` / ``polish_peaks_with_weights``.
* :func:`mp_k2`: pure-Python exact-integer phase + mpmath at >= 40 digits, for small cases; it
  pins the binary128 tier.
* ``_ref/libpath1_eval.so``: the reference evaluator compiled from its own source by the
  Makefile here (only where the reference tree is mounted); :func:`load_ref` returns it or None.

Parity pins (each exercised in tests/test_oracle.py):
  phase words     -- Python big-integer brute force; closed forms (G = 2^110, V = n + 1/4).
  FP64 / binary128-- mpmath (mp_k2) on small tables; closed forms (single zero, v = 0,
                     linearity in w, w -> i w); the compiled reference (reference format).
  derivatives     -- binary128 central differences at h = 2^-40 (error << 1e-12 relative).
  polish          -- compiled reference (reference format); K' = 0 at the returned point.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libk2oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libpath1_eval.so")

VALUE, D1, D2 = 1, 2, 4

_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int)
_lib = None


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)


def _p(a, t):
    return a.ctypes.data_as(t)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.k2o_phase_word.restype = ctypes.c_uint64
        L.k2o_phase_word.argtypes = [ctypes.c_uint64] * 4
        L.k2o_phase_words.argtypes = [ctypes.c_int, _u64p, ctypes.c_int, _u64p, _u64p]
        L.k2o_eval_f64.argtypes = [ctypes.c_int, _u64p, _f64p, _f64p, ctypes.c_int, _u64p,
                                   ctypes.c_double, ctypes.c_uint, _f64p, _f64p, _f64p, ctypes.c_int]
        L.k2o_eval_quad.argtypes = [ctypes.c_int, _u64p, _f64p, ctypes.c_int, _u64p,
                                    ctypes.c_double, ctypes.c_uint, _f64p, _f64p, _f64p, _f64p,
                                    ctypes.c_int]
        L.k2o_eval_quad_pair.argtypes = [ctypes.c_int, _u64p, _f64p, ctypes.c_int, _u64p,
                                         ctypes.c_double, _f64p, _f64p]
        L.k2o_polish_f64.argtypes = [ctypes.c_int, _u64p, _f64p, _f64p, ctypes.c_int, _u64p,
                                     ctypes.c_double, ctypes.c_int, ctypes.c_double, _u64p,
                                     _f64p, _f64p, _f64p, _i32p]
        L.k2o_ref_compute_peak_AB.argtypes = [ctypes.c_int, _u64p, _u64p, _f64p, ctypes.c_int,
                                              _f64p, _f64p] + [_u64p] * 6 + [_f64p, _f64p]
        L.k2o_ref_polish_peaks_with_weights.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, _f64p,
                                                        _f64p, _f64p, ctypes.c_double, ctypes.c_int,
                                                        _f64p, _f64p, _f64p]
        L.k2o_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _w2(w):
    return _c(np.stack([np.real(w), np.imag(w)], axis=1), np.float64)


def phase_words(G, V) -> np.ndarray:
    """(nc, nz) uint64 phase words floor(G V / 2^110) mod 2^64."""
    G = _c(G, np.uint64); V = _c(V, np.uint64)
    out = np.empty((V.shape[0], G.shape[0]), dtype=np.uint64)
    lib().k2o_phase_words(V.shape[0], _p(V, _u64p), G.shape[0], _p(G, _u64p), _p(out, _u64p))
    return out


def eval_f64(table, V, k, flags=VALUE | D1 | D2, nthreads=0, n_zeros=None):
    nz = table.n if n_zeros is None else int(n_zeros)
    G = _c(table.G[:nz], np.uint64); g = _c(table.gamma[:nz], np.float64); w = _w2(table.w[:nz])
    V = _c(V, np.uint64); nc = V.shape[0]
    K = np.zeros(nc); dK = np.zeros(nc); d2K = np.zeros(nc)
    lib().k2o_eval_f64(nz, _p(G, _u64p), _p(g, _f64p), _p(w, _f64p), nc, _p(V, _u64p),
                       float(k), flags, _p(K, _f64p), _p(dK, _f64p), _p(d2K, _f64p), int(nthreads))
    return K, dK, d2K


def eval_quad(table, V, k, flags=VALUE | D1 | D2, nthreads=0, n_zeros=None):
    """binary128 oracle: returns (K, K', K'', sum of |terms|) as float64 arrays."""
    nz = table.n if n_zeros is None else int(n_zeros)
    G = _c(table.G[:nz], np.uint64); w = _w2(table.w[:nz])
    V = _c(V, np.uint64); nc = V.shape[0]
    K = np.zeros(nc); dK = np.zeros(nc); d2K = np.zeros(nc); aS = np.zeros(nc)
    lib().k2o_eval_quad(nz, _p(G, _u64p), _p(w, _f64p), nc, _p(V, _u64p), float(k), flags,
                        _p(K, _f64p), _p(dK, _f64p), _p(d2K, _f64p), _p(aS, _f64p), int(nthreads))
    return K, dK, d2K, aS


def eval_quad_pair(table, V, k):
    G = _c(table.G, np.uint64); w = _w2(table.w); V = _c(V, np.uint64); nc = V.shape[0]
    hi = np.zeros(nc); lo = np.zeros(nc)
    lib().k2o_eval_quad_pair(table.n, _p(G, _u64p), _p(w, _f64p), nc, _p(V, _u64p), float(k),
                             _p(hi, _f64p), _p(lo, _f64p))
    return hi, lo


def polish_f64(table, V0, k, n_steps=8, clamp=1e-4):
    G = _c(table.G, np.uint64); g = _c(table.gamma, np.float64); w = _w2(table.w)
    V0 = _c(V0, np.uint64); nc = V0.shape[0]
    Vout = np.zeros_like(V0); K0 = np.zeros(nc); Kf = np.zeros(nc); dKf = np.zeros(nc)
    st = np.zeros(nc, dtype=np.int32)
    lib().k2o_polish_f64(table.n, _p(G, _u64p), _p(g, _f64p), _p(w, _f64p), nc, _p(V0, _u64p),
                         float(k), int(n_steps), float(clamp), _p(Vout, _u64p), _p(K0, _f64p),
                         _p(Kf, _f64p), _p(dKf, _f64p), _p(st, _i32p))
    return Vout, K0, Kf, dKf, st


def max_threads() -> int:
    return int(lib().k2o_max_threads())


# ---- reference format ---------------------------------------------------------------------

def _ref_args(x_lo, x_hi, dy0, alpha, gamma, ghi_lo, ghi_hi, glo_lo, glo_hi, psi_lo, psi_hi):
    arrs = [_c(x_lo, np.uint64), _c(x_hi, np.uint64), _c(dy0, np.float64), _c(alpha, np.float64),
            _c(gamma, np.float64)] + [_c(a, np.uint64) for a in (ghi_lo, ghi_hi, glo_lo, glo_hi,
                                                                    psi_lo, psi_hi)]
    return arrs


def _call_cpab(fn, *args):
    x_lo, x_hi, dy0, alpha, gamma, ghl, ghh, gll, glh, psl, psh = _ref_args(*args)
    npk, nz = x_lo.shape[0], gamma.shape[0]
    A = np.empty((npk, nz)); B = np.empty((npk, nz))
    fn(npk, _p(x_lo, _u64p), _p(x_hi, _u64p), _p(dy0, _f64p), nz, _p(alpha, _f64p),
       _p(gamma, _f64p), _p(ghl, _u64p), _p(ghh, _u64p), _p(gll, _u64p), _p(glh, _u64p),
       _p(psl, _u64p), _p(psh, _u64p), _p(A, _f64p), _p(B, _f64p))
    return A, B


def _call_polish(fn, A, B, gamma, w, half, steps):
    A = _c(A, np.float64); B = _c(B, np.float64); gamma = _c(gamma, np.float64); w = _c(w, np.float64)
    npk, nz = A.shape
    hu = np.empty(npk); hp = np.empty(npk); dd = np.empty(npk)
    fn(npk, nz, _p(A, _f64p), _p(B, _f64p), _p(gamma, _f64p), _p(w, _f64p), float(half), int(steps),
       _p(hu, _f64p), _p(hp, _f64p), _p(dd, _f64p))
    return hu, hp, dd


def ref_compute_peak_AB(*args):
    return _call_cpab(lib().k2o_ref_compute_peak_AB, *args)


def ref_polish_peaks_with_weights(A, B, gamma, w, half=2.5e-4, steps=40):
    return _call_polish(lib().k2o_ref_polish_peaks_with_weights, A, B, gamma, w, half, steps)


def load_ref():
    """The compiled reference evaluator (oracle/_ref), or None where it was not built."""
    if not os.path.exists(REF_PATH):
        return None
    L = ctypes.CDLL(REF_PATH)
    L.compute_peak_AB.argtypes = [ctypes.c_int, _u64p, _u64p, _f64p, ctypes.c_int, _f64p, _f64p] \
        + [_u64p] * 6 + [_f64p, _f64p]
    L.polish_peaks_with_weights.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, _f64p, _f64p, _f64p,
                                            ctypes.c_double, ctypes.c_int, _f64p, _f64p, _f64p]
    return L


def ref_lib_compute_peak_AB(L, *args):
    return _call_cpab(L.compute_peak_AB, *args)


def ref_lib_polish(L, A, B, gamma, w, half=2.5e-4, steps=40):
    return _call_polish(L.polish_peaks_with_weights, A, B, gamma, w, half, steps)


# ---- mpmath oracle (small cases) ----------------------------------------------------------

def mp_k2(G, w, V, k, dps=40, derivs=True):
    """Exact-integer phase + mpmath at ``dps`` digits; returns lists (K, K', K'')."""
    import mpmath
    with mpmath.workdps(dps):
        twopi = 2 * mpmath.pi
        kk = mpmath.mpf(float(k))
        out = ([], [], [])
        gs = [twopi * mpmath.mpf((int(h) << 64) | int(l)) / mpmath.mpf(2) ** 110 for l, h in G]
        for vl, vh in V:
            Vi = (int(vh) << 64) | int(vl)
            s0 = s1 = s2 = mpmath.mpf(0)
            for (gl, gh), g, wj in zip(G, gs, w):
                Gi = (int(gh) << 64) | int(gl)
                num = (Gi * Vi) % (1 << 174)
                th = twopi * mpmath.mpf(num) / mpmath.mpf(2) ** 174
                a = 2 * mpmath.exp(-kk * g * g)
                c, s = mpmath.cos(th), mpmath.sin(th)
                re, im = mpmath.mpf(float(np.real(wj))), mpmath.mpf(float(np.imag(wj)))
                val = a * (re * c - im * s)
                s0 += val
                if derivs:
                    s1 += a * g * (-re * s - im * c)
                    s2 -= g * g * val
            out[0].append(s0); out[1].append(s1); out[2].append(s2)
        return out
