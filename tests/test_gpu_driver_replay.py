"""The reference driver's call sequence (search/eval_tech1_600k.py:201-271: read zeros_600k.bin,
take CLUSTER candidates, compute_peak_AB, then polish_peaks_with_weights once per weight vector)
replayed through plain ctypes -- the binding a maintainer keeps, changing only the library path --
against the reference library built from its own source (oracle/_ref) and libk2eval.so."""
import ctypes
import os

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from mertens_bound_amd import formats, pipeline, synth

pytestmark = pytest.mark.gpu

u64_p = ctypes.POINTER(ctypes.c_uint64)
f64_p = ctypes.POINTER(ctypes.c_double)


def _bind(path):
    lib = ctypes.CDLL(path)
    lib.compute_peak_AB.argtypes = [ctypes.c_int, u64_p, u64_p, f64_p, ctypes.c_int, f64_p, f64_p,
                                    u64_p, u64_p, u64_p, u64_p, u64_p, u64_p, f64_p, f64_p]
    lib.compute_peak_AB.restype = None
    lib.polish_peaks_with_weights.argtypes = [ctypes.c_int, ctypes.c_int, f64_p, f64_p, f64_p,
                                              f64_p, ctypes.c_double, ctypes.c_int, f64_p, f64_p,
                                              f64_p]
    lib.polish_peaks_with_weights.restype = None
    return lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _driver(lib, zfile, cands):
    z = formats.read_zeros_600k_bin(zfile)
    n = len(z["gamma"])
    x_lo, x_hi = formats.x_words([c["x"] for c in cands])
    dy0 = np.ascontiguousarray([c["dy"] for c in cands], dtype=np.float64)
    n_p = len(cands)
    A = np.empty((n_p, n)); B = np.empty((n_p, n))
    lib.compute_peak_AB(n_p, _p(x_lo, u64_p), _p(x_hi, u64_p), _p(dy0, f64_p), n,
                        _p(z["alpha"], f64_p), _p(z["gamma"], f64_p),
                        *[_p(z[k], u64_p) for k in ("ghi_lo", "ghi_hi", "glo_lo", "glo_hi",
                                                    "psi_lo", "psi_hi")],
                        _p(A, f64_p), _p(B, f64_p))
    names, wts = pipeline.reference_weight_vectors(z["gamma"], n_trunc=1000)
    out = {}
    for name, w in zip(names, wts):
        w = np.ascontiguousarray(w)
        hu = np.empty(n_p); hp = np.empty(n_p); dd = np.empty(n_p)
        lib.polish_peaks_with_weights(n_p, n, _p(A, f64_p), _p(B, f64_p), _p(z["gamma"], f64_p),
                                      _p(w, f64_p), 2.5e-4, 40, _p(hu, f64_p), _p(hp, f64_p),
                                      _p(dd, f64_p))
        out[name] = (hu, hp, dd)
    return A, B, out


def test_eval_tech1_call_sequence_matches_reference(tmp_path):
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libpath1_eval.so")
    if not os.path.exists(ref_path):
        pytest.skip("reference library not built (no reference tree at build time)")
    import torch  # noqa: F401  (shares the HIP runtime with libk2eval.so)
    tb = synth.make_zero_table(2000, 23)
    zfile = str(tmp_path / "zeros_600k.bin")
    formats.write_zeros_600k_bin(zfile, formats.reference_words(tb))
    lines = open(os.path.join(GOLDEN, "cluster_sample.txt")).read().splitlines()
    cands = formats.select_candidates(formats.parse_cluster_lines(lines),
                                      **formats.SUB35_SELECTION)[:20]
    assert len(cands) >= 5
    A0, B0, ref = _driver(_bind(ref_path), zfile, cands)
    A1, B1, ours = _driver(_bind(os.path.join(ROOT, "mertens_bound_amd", "libk2eval.so")),
                           zfile, cands)
    z = formats.read_zeros_600k_bin(zfile)
    th_max = 2 * np.pi + z["gamma"][None, :] * np.max(np.abs([c["dy"] for c in cands]))
    tol = 2 * z["alpha"][None, :] * (th_max * 2.3e-16 + 4.5e-16)
    assert np.all(np.abs(A1 - A0) <= tol) and np.all(np.abs(B1 - B0) <= tol)
    for name in ref:
        for a, b, atol in zip(ours[name], ref[name], (1e-12, 1e-10, 1e-9)):
            np.testing.assert_allclose(a, b, rtol=0, atol=atol, err_msg=name)
