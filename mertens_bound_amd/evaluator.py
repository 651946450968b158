"""Python host layer over libk2eval.so (C ABI in include/k2eval.h), through ctypes.

It keeps the reference evaluator's call shape -- a zero table plus a candidate list in,
K2 values out -- as exposed to the reference's Python drivers by ctypes
(search/eval_tech1_600k.py:201-267): ``K2Table(table).eval(V, k)`` replaces
``compute_peak_AB`` + the weighted sum ``h0 = sum_i w_i A_i`` (search/path1_eval_fast.cpp:15-53,
77-81), and :func:`compute_peak_AB` / :func:`polish_peaks_with_weights` keep the reference's
exact signatures for replaying its callers.

There is no CPU fallback: every entry point runs the HIP library, and loading fails loudly when
the library is not built or no GPU is visible.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import synth

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libk2eval.so")

VALUE, D1, D2, MIXED = 1, 2, 4, 8
E_MESSAGES = {1: "invalid argument", 2: "HIP runtime error", 3: "device allocation failed"}

_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


class K2Error(RuntimeError):
    pass


class k2_request(ctypes.Structure):
    _fields_ = [("n_zeros", ctypes.c_uint32), ("n_cand", ctypes.c_uint32),
                ("cand_offset", ctypes.c_uint64), ("k", ctypes.c_double)]


class k2_tiling(ctypes.Structure):
    _fields_ = [("nc", ctypes.c_int32), ("target_blocks", ctypes.c_int32),
                ("mixed_head", ctypes.c_int32), ("mixed_tail", ctypes.c_double),
                ("z_max_cand", ctypes.c_int32), ("wg_threads", ctypes.c_int32)]


class k2_batch_tiling(ctypes.Structure):
    _fields_ = [("nc", ctypes.c_int32), ("cz", ctypes.c_int32), ("chunk", ctypes.c_int32),
                ("c_min_fill", ctypes.c_int32), ("wg_per_cu", ctypes.c_int32)]


class k2_batch_info(ctypes.Structure):
    _fields_ = [("n_items", ctypes.c_int64), ("grid", ctypes.c_int32), ("n_k", ctypes.c_int32),
                ("useful_evals", ctypes.c_double), ("padded_evals", ctypes.c_double),
                ("partial_bytes", ctypes.c_int64)]


_lib = None


def _declare(L):
    L.k2_last_error.restype = ctypes.c_char_p
    L.k2_version.restype = ctypes.c_int
    L.k2_table_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_uint32, _u64p, _f64p, _f64p]
    L.k2_table_destroy.argtypes = [_vp]
    L.k2_table_size.argtypes = [_vp]
    L.k2_table_size.restype = ctypes.c_uint32
    L.k2_set_tiling.argtypes = [_vp, ctypes.POINTER(k2_tiling)]
    L.k2_get_tiling.argtypes = [_vp, ctypes.POINTER(k2_tiling)]
    L.k2_eval.argtypes = [_vp, ctypes.c_uint32, _u64p, ctypes.c_double, ctypes.c_uint32,
                          ctypes.c_uint32, _f64p, _f64p, _f64p]
    L.k2_eval_device.argtypes = [_vp, ctypes.c_uint32, _vp, ctypes.c_double, ctypes.c_uint32,
                                 ctypes.c_uint32, _vp, _vp, _vp, _vp]
    L.k2_eval_batched_device.argtypes = [_vp, ctypes.c_uint32, ctypes.POINTER(k2_request), _vp,
                                         ctypes.c_uint64, ctypes.c_uint32, _vp, _vp, _vp, _vp]
    L.k2_batch_create.argtypes = [_vp, ctypes.c_uint32, ctypes.POINTER(k2_request), _vp,
                                  ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(k2_batch_tiling),
                                  _vp, _vp, _vp, ctypes.POINTER(_vp)]
    L.k2_batch_run.argtypes = [_vp, _vp]
    L.k2_batch_run_direct.argtypes = [_vp, _vp]
    L.k2_batch_get_info.argtypes = [_vp, ctypes.POINTER(k2_batch_info)]
    L.k2_batch_destroy.argtypes = [_vp]
    L.k2_screen_device.argtypes = [_vp, ctypes.c_uint32, _vp, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                                   ctypes.POINTER(ctypes.c_float), _vp, _vp, _vp, _vp]
    L.k2_enclose_device.argtypes = [_vp, ctypes.c_uint32, _vp, ctypes.c_double, ctypes.c_uint32,
                                    ctypes.c_uint32, _vp, _vp, _vp, _vp, _vp, _vp]
    L.k2_exp_neg_host.argtypes = [ctypes.c_uint32, _f64p, _f64p]
    L.k2_split_plan.argtypes = [ctypes.c_int32] * 4 + [ctypes.POINTER(ctypes.c_int32)]
    L.k2_eval_weights_device.argtypes = [_vp, ctypes.c_uint32, _vp, ctypes.c_uint32, _vp,
                                         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _vp,
                                         _vp, _vp, _vp]
    L.k2_polish_stats.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64)]
    L.k2_mixed_head.argtypes = [_vp, ctypes.c_double, ctypes.c_uint32,
                                ctypes.POINTER(ctypes.c_uint32)]
    L.k2_polish_device.argtypes = [_vp, ctypes.c_uint32, _vp, ctypes.c_double, ctypes.c_uint32,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_uint32, _vp, _vp, _vp,
                                   _vp, _vp, _vp]
    L.k2_phase_words.argtypes = [_vp, ctypes.c_uint32, _u64p, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_int, _u64p]
    L.k2_sincos_words.argtypes = [ctypes.c_uint32, _u64p, ctypes.c_int, _f64p]
    L.k2_valu_rates.argtypes = [_f64p]
    L.k2_profile.argtypes = [_vp, ctypes.c_int]
    L.k2_kernel_time.argtypes = [_vp, _f64p, ctypes.POINTER(ctypes.c_int64)]
    L.compute_peak_AB.argtypes = [ctypes.c_int, _u64p, _u64p, _f64p, ctypes.c_int, _f64p, _f64p] \
        + [_u64p] * 6 + [_f64p, _f64p]
    L.compute_peak_AB.restype = None
    L.polish_peaks_with_weights.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, _f64p, _f64p,
                                            _f64p, ctypes.c_double, ctypes.c_int, _f64p, _f64p,
                                            _f64p]
    L.polish_peaks_with_weights.restype = None
    return L


def load_library():
    """Load libk2eval.so.  torch (when installed) is imported first so that both share one HIP
    runtime in the process."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise K2Error(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                          "python -m mertens_bound_amd._build")
        _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        msg = load_library().k2_last_error().decode(errors="replace")
        raise K2Error(f"{what} failed ({E_MESSAGES.get(rc, rc)}): {msg}")


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _p(a, t):
    return a.ctypes.data_as(t)


class K2Table:
    """A zero table resident on the GPU (canonical encoding, see synth.ZeroTable)."""

    def __init__(self, table: synth.ZeroTable):
        L = load_library()
        self._G = _c(table.G, np.uint64)
        self._gamma = _c(table.gamma, np.float64)
        self._w = _c(np.stack([table.w.real, table.w.imag], axis=1), np.float64)
        self.n = table.n
        self.gamma = self._gamma
        h = _vp()
        _check(L.k2_table_create(ctypes.byref(h), self.n, _p(self._G, _u64p),
                                 _p(self._gamma, _f64p), _p(self._w, _f64p)), "k2_table_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load_library().k2_table_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # tiling -----------------------------------------------------------------------------
    def set_tiling(self, nc=None, target_blocks=None, mixed_head=None, mixed_tail=None,
                   z_max_cand=None, wg_threads=None):
        """mixed_head 0 chooses the FP64 head from the weight magnitudes (tail <= mixed_tail);
        candidate counts <= z_max_cand run as Z tiles (-1: never); wg_threads 256 or 512."""
        cur = self.get_tiling()
        t = k2_tiling(nc if nc is not None else cur["nc"],
                      target_blocks if target_blocks is not None else cur["target_blocks"],
                      mixed_head if mixed_head is not None else cur["mixed_head"],
                      mixed_tail if mixed_tail is not None else cur["mixed_tail"],
                      z_max_cand if z_max_cand is not None else cur["z_max_cand"],
                      wg_threads if wg_threads is not None else cur["wg_threads"])
        _check(load_library().k2_set_tiling(self._h, ctypes.byref(t)), "k2_set_tiling")

    def get_tiling(self):
        t = k2_tiling()
        _check(load_library().k2_get_tiling(self._h, ctypes.byref(t)), "k2_get_tiling")
        return {"nc": t.nc, "target_blocks": t.target_blocks, "mixed_head": t.mixed_head,
                "mixed_tail": t.mixed_tail, "z_max_cand": t.z_max_cand,
                "wg_threads": t.wg_threads}

    def mixed_head(self, k, n_zeros=None):
        """FP64 head (zeros) the mixed form uses for k over the first n_zeros zeros."""
        h = ctypes.c_uint32()
        _check(load_library().k2_mixed_head(self._h, float(k), int(n_zeros or 0), ctypes.byref(h)),
               "k2_mixed_head")
        return h.value

    # host-buffer entry points -------------------------------------------------------------
    def eval(self, V, k, n_zeros=None, flags=VALUE | D1 | D2):
        """K2, K2', K2'' (numpy float64) of the 64.64 candidates V over the first n_zeros zeros."""
        V = _c(V, np.uint64).reshape(-1, 2)
        n = V.shape[0]
        K = np.zeros(n); dK = np.zeros(n); d2K = np.zeros(n)
        _check(load_library().k2_eval(self._h, n, _p(V, _u64p), float(k), int(n_zeros or 0),
                                      int(flags), _p(K, _f64p), _p(dK, _f64p), _p(d2K, _f64p)),
               "k2_eval")
        return K, dK, d2K

    def phase_words(self, V, z0=0, nz=None, mode=0):
        V = _c(V, np.uint64).reshape(-1, 2)
        nz = self.n - z0 if nz is None else nz
        out = np.zeros((V.shape[0], nz), dtype=np.uint64)
        _check(load_library().k2_phase_words(self._h, V.shape[0], _p(V, _u64p), z0, nz, mode,
                                             _p(out, _u64p)), "k2_phase_words")
        return out

    # device-buffer entry points (torch tensors on the GPU) ----------------------------------
    def eval_device(self, V, k, n_zeros=None, flags=VALUE | D1 | D2, out=None, stream=None):
        """V: (n, 2) int64/uint64 CUDA tensor of 64.64 words.  Returns (K, dK, d2K) tensors."""
        import torch
        n = V.shape[0]
        if out is None:
            out = tuple(torch.empty(n, dtype=torch.float64, device=V.device) for _ in range(3))
        K, dK, d2K = out
        st = stream if stream is not None else torch.cuda.current_stream(V.device).cuda_stream
        _check(load_library().k2_eval_device(
            self._h, n, V.data_ptr(), float(k), int(n_zeros or 0), int(flags), K.data_ptr(),
            dK.data_ptr() if dK is not None else None, d2K.data_ptr() if d2K is not None else None,
            st), "k2_eval_device")
        return K, dK, d2K

    def eval_batched_device(self, requests, V, flags=VALUE | D1, out=None, stream=None):
        """requests: iterable of (n_zeros, n_cand, cand_offset, k); V: (N, 2) tensor.  One
        persistent launch; waits for it to finish."""
        import torch
        arr, n = _requests(requests)
        N = V.shape[0]
        if out is None:
            out = tuple(torch.zeros(N, dtype=torch.float64, device=V.device) for _ in range(3))
        K, dK, d2K = out
        st = stream if stream is not None else torch.cuda.current_stream(V.device).cuda_stream
        _check(load_library().k2_eval_batched_device(
            self._h, n, arr, V.data_ptr(), N, int(flags), K.data_ptr(),
            dK.data_ptr() if dK is not None else None, d2K.data_ptr() if d2K is not None else None,
            st), "k2_eval_batched_device")
        return K, dK, d2K

    def polish_device(self, V, k, n_steps=8, clamp=1e-4, n_zeros=None, flags=0, stream=None,
                      start=False):
        """Newton polish in place on V (tensor).  Returns (K, dK, d2K, steps) at the final point,
        and with start=True also K at the starting points (from the first pass)."""
        import torch
        n = V.shape[0]
        K, dK, d2K = (torch.empty(n, dtype=torch.float64, device=V.device) for _ in range(3))
        K0 = torch.empty(n, dtype=torch.float64, device=V.device) if start else None
        steps = torch.zeros(n, dtype=torch.int32, device=V.device)
        st = stream if stream is not None else torch.cuda.current_stream(V.device).cuda_stream
        _check(load_library().k2_polish_device(
            self._h, n, V.data_ptr(), float(k), int(n_zeros or 0), int(n_steps), float(clamp),
            int(flags), K.data_ptr(), dK.data_ptr(), d2K.data_ptr(), steps.data_ptr(),
            K0.data_ptr() if start else None, st), "k2_polish_device")
        return (K, dK, d2K, steps, K0) if start else (K, dK, d2K, steps)

    def polish_stats(self):
        """Candidate evaluations of the last polish_device call (converged candidates drop out)."""
        n = ctypes.c_int64()
        _check(load_library().k2_polish_stats(self._h, ctypes.byref(n)), "k2_polish_stats")
        return n.value

    def screen_device(self, V, k, shift_step=0.0014, n_shift=3,
                      stages=((64, 0.78), (128, 0.855), (256, 0.892)), stream=None):
        """Staged pre-screen (k2_screen_device).  Returns (values [n, 4] float32 with -1 where a
        stage was not reached, best shift index [n] int32, passed-all-stages [n] int32)."""
        import torch
        n = V.shape[0]
        nz = (ctypes.c_int32 * len(stages))(*[int(a) for a, _ in stages])
        th = (ctypes.c_float * len(stages))(*[float(b) for _, b in stages])
        val = torch.empty((n, 4), dtype=torch.float32, device=V.device)
        best = torch.empty(n, dtype=torch.int32, device=V.device)
        ok = torch.empty(n, dtype=torch.int32, device=V.device)
        st = stream if stream is not None else torch.cuda.current_stream(V.device).cuda_stream
        _check(load_library().k2_screen_device(self._h, n, V.data_ptr(), float(k),
                                               float(shift_step), int(n_shift), len(stages), nz,
                                               th, val.data_ptr(), best.data_ptr(),
                                               ok.data_ptr(), st), "k2_screen_device")
        return val, best, ok

    def eval_weights_device(self, V, wt, n_zeros=None, flags=VALUE, stream=None):
        """K_m = sum_j wt[m, j] 2 Re(w_j e^{i g_j v}) for the rows of wt ((n_w, >= n_zeros)
        float64 CUDA tensor, n_w <= 8), sharing the phase and sin/cos.  Returns (K, dK, d2K),
        each (n_w, n) or None where not requested."""
        import torch
        n = V.shape[0]
        wt = wt.to(dtype=torch.float64).contiguous()
        n_w, ld = wt.shape
        mk = lambda f: (torch.empty((n_w, n), dtype=torch.float64, device=V.device)
                        if flags & f else None)
        K, dK, d2K = mk(VALUE | D1 | D2), mk(D1), mk(D2)
        st = stream if stream is not None else torch.cuda.current_stream(V.device).cuda_stream
        _check(load_library().k2_eval_weights_device(
            self._h, n, V.data_ptr(), n_w, wt.data_ptr(), ld, int(n_zeros or 0),
            int(flags) | VALUE, K.data_ptr(), dK.data_ptr() if dK is not None else None,
            d2K.data_ptr() if d2K is not None else None, st), "k2_eval_weights_device")
        return K, dK, d2K

    def enclose_device(self, V, k, n_zeros=None, d1=True, stream=None):
        """Enclosure mode (k2_enclose_device): the exact K2 (and K2') over the first n_zeros
        zeros lies within K +- K_rad (dK +- dK_rad).  Returns (K, K_rad, dK, dK_rad, tail) with
        tail = [bound on sum_{j >= n_zeros} 2 e^{-k g^2} |w_j|, the same with the factor g_j]
        (float64 CUDA tensors; dK, dK_rad None when d1 is False)."""
        import torch
        n = V.shape[0]
        mk = lambda: torch.empty(n, dtype=torch.float64, device=V.device)
        K, K_rad = mk(), mk()
        dK, dK_rad = (mk(), mk()) if d1 else (None, None)
        tail = torch.empty(2, dtype=torch.float64, device=V.device)
        st = stream if stream is not None else torch.cuda.current_stream(V.device).cuda_stream
        _check(load_library().k2_enclose_device(
            self._h, n, V.data_ptr(), float(k), int(n_zeros or 0), VALUE | (D1 if d1 else 0),
            K.data_ptr(), K_rad.data_ptr(), dK.data_ptr() if d1 else None,
            dK_rad.data_ptr() if d1 else None, tail.data_ptr(), st), "k2_enclose_device")
        return K, K_rad, dK, dK_rad, tail

    def enclose(self, V, k, n_zeros=None, d1=True):
        """Host-array form of enclose_device: V (n, 2) uint64 words; returns numpy arrays."""
        import torch
        Vt = v_tensor(np.ascontiguousarray(V, dtype=np.uint64))
        res = self.enclose_device(Vt, k, n_zeros, d1)
        torch.cuda.synchronize()
        return tuple(None if r is None else r.cpu().numpy() for r in res)

    # profiling ----------------------------------------------------------------------------
    def profile(self, enable=True):
        _check(load_library().k2_profile(self._h, 1 if enable else 0), "k2_profile")

    def kernel_time(self):
        ms = ctypes.c_double(); n = ctypes.c_int64()
        _check(load_library().k2_kernel_time(self._h, ctypes.byref(ms), ctypes.byref(n)),
               "k2_kernel_time")
        return ms.value, n.value


def valu_rates():
    """Measured issue rates (lane operations/s) of v_mad_u64_u32, v_fma_f64, v_fma_f32 and
    v_sin_f32 on the current GPU (k2_valu_rates)."""
    out = np.zeros(4)
    _check(load_library().k2_valu_rates(_p(out, _f64p)), "k2_valu_rates")
    return {"int32_mad": out[0], "fp64_fma": out[1], "fp32_fma": out[2], "trans": out[3]}


def exp_neg_host(x):
    """The enclosure mode's e^{-x} evaluated on the host (same operations as on the device)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    _check(load_library().k2_exp_neg_host(x.size, x.ctypes.data_as(_f64p),
                                          out.ctypes.data_as(_f64p)), "k2_exp_neg_host")
    return out


def split_plan(nz, n_ctiles, target, factor):
    """The zero split plan of a C-tile launch (host arithmetic only): dict with the split count,
    zeros per short split, long splits and zeros per long split, and the launch's workgroups."""
    out = (ctypes.c_int32 * 5)()
    _check(load_library().k2_split_plan(nz, n_ctiles, target, factor, out), "k2_split_plan")
    return {"n": out[0], "zps": out[1], "n_long": out[2], "zps_long": out[3], "blocks": out[4]}


def _requests(requests):
    reqs = list(requests)
    arr = (k2_request * max(1, len(reqs)))()
    for i, (nz, nc, off, k) in enumerate(reqs):
        arr[i] = k2_request(int(nz), int(nc), int(off), float(k))
    return arr, len(reqs)


class K2Batch:
    """A batched plan: uneven requests (n_zeros, n_cand, cand_offset, k) over one table, bound to
    the candidate tensor V and output tensors, run as one persistent launch; ``run`` replays a
    HIP graph captured on first use, ``run_direct`` issues the launches without the graph."""

    def __init__(self, table: "K2Table", requests, V, flags=VALUE | D1, tiling=None, out=None):
        import torch
        L = load_library()
        self.table = table          # keeps the table alive
        self.V = V
        N = V.shape[0]
        if out is None:
            out = tuple(torch.zeros(N, dtype=torch.float64, device=V.device) for _ in range(3))
        self.K, self.dK, self.d2K = out
        arr, n = _requests(requests)
        ti = None
        if tiling is not None:
            ti = k2_batch_tiling(int(tiling.get("nc", 0)), int(tiling.get("cz", 0)),
                                 int(tiling.get("chunk", 0)), int(tiling.get("c_min_fill", 0)),
                                 int(tiling.get("wg_per_cu", 0)))
        h = _vp()
        _check(L.k2_batch_create(table.handle, n, arr, V.data_ptr(), N, int(flags),
                                 ctypes.byref(ti) if ti is not None else None, self.K.data_ptr(),
                                 self.dK.data_ptr(), self.d2K.data_ptr(), ctypes.byref(h)),
               "k2_batch_create")
        self._h = h

    def run(self, stream=None):
        import torch
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _check(load_library().k2_batch_run(self._h, st), "k2_batch_run")
        return self.K, self.dK, self.d2K

    def run_direct(self, stream=None):
        import torch
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _check(load_library().k2_batch_run_direct(self._h, st), "k2_batch_run_direct")
        return self.K, self.dK, self.d2K

    def info(self):
        i = k2_batch_info()
        _check(load_library().k2_batch_get_info(self._h, ctypes.byref(i)), "k2_batch_get_info")
        return {"n_items": i.n_items, "grid": i.grid, "n_k": i.n_k,
                "useful_evals": i.useful_evals, "padded_evals": i.padded_evals,
                "partial_bytes": i.partial_bytes}

    def close(self):
        if getattr(self, "_h", None):
            load_library().k2_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sincos_words(W, mode=0):
    W = _c(W, np.uint64).ravel()
    out = np.zeros((W.shape[0], 2))
    _check(load_library().k2_sincos_words(W.shape[0], _p(W, _u64p), mode, _p(out, _f64p)),
           "k2_sincos_words")
    return out[:, 0], out[:, 1]


def v_tensor(V, device="cuda"):
    """(n, 2) uint64 64.64 words -> int64 tensor on the GPU (same bytes)."""
    import torch
    return torch.from_numpy(_c(V, np.uint64).view(np.int64)).to(device)


def v_from_tensor(Vt) -> np.ndarray:
    return Vt.cpu().numpy().view(np.uint64).reshape(-1, 2)


# ---- reference signatures (search/path1_eval_fast.cpp) --------------------------------------

def compute_peak_AB(n_peaks, x_lo, x_hi, dy0, n_zeros, alpha, gamma, ghi_lo, ghi_hi, glo_lo,
                    glo_hi, psi_lo, psi_hi, out_A, out_B):
    """search/path1_eval_fast.cpp:15-31, evaluated on the GPU; outputs filled in place."""
    L = load_library()
    args = [_c(a, np.uint64) for a in (x_lo, x_hi)] + [_c(dy0, np.float64)]
    zs = [_c(alpha, np.float64), _c(gamma, np.float64)] + \
        [_c(a, np.uint64) for a in (ghi_lo, ghi_hi, glo_lo, glo_hi, psi_lo, psi_hi)]
    if not (out_A.flags.c_contiguous and out_B.flags.c_contiguous):
        raise K2Error("compute_peak_AB: outputs must be C-contiguous float64")
    L.compute_peak_AB(int(n_peaks), _p(args[0], _u64p), _p(args[1], _u64p), _p(args[2], _f64p),
                      int(n_zeros), _p(zs[0], _f64p), _p(zs[1], _f64p),
                      *[_p(a, _u64p) for a in zs[2:]], _p(out_A, _f64p), _p(out_B, _f64p))
    if np.isnan(out_A.flat[0]) if out_A.size else False:
        raise K2Error("compute_peak_AB: " + L.k2_last_error().decode(errors="replace"))


def polish_peaks_with_weights(n_peaks, n_zeros, A, B, gamma, w, grid_half_width, grid_steps,
                              out_h_unpol, out_h_pol, out_ddy_pol):
    """search/path1_eval_fast.cpp:58-70, evaluated on the GPU; outputs filled in place."""
    L = load_library()
    A = _c(A, np.float64); B = _c(B, np.float64)
    gamma = _c(gamma, np.float64); w = _c(w, np.float64)
    L.polish_peaks_with_weights(int(n_peaks), int(n_zeros), _p(A, _f64p), _p(B, _f64p),
                                _p(gamma, _f64p), _p(w, _f64p), float(grid_half_width),
                                int(grid_steps), _p(out_h_unpol, _f64p), _p(out_h_pol, _f64p),
                                _p(out_ddy_pol, _f64p))
    if n_peaks and np.isnan(out_h_pol[0]):
        raise K2Error("polish_peaks_with_weights: " + L.k2_last_error().decode(errors="replace"))
