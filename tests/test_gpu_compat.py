"""The reference-signature entry points (compute_peak_AB, polish_peaks_with_weights) on the
GPU against the reference evaluator compiled from its own source (oracle/_ref) -- or, where that
library was not built, against the oracle's restatement, which the CPU tests pin bit for bit to
it -- and against the committed fixture."""
import numpy as np
import pytest

from conftest import golden
from golden.make_golden import ref_inputs
from mertens_bound_amd import evaluator as ev

pytestmark = pytest.mark.gpu


def _reference(oracle_lib, args):
    L = oracle_lib.load_ref()
    if L is not None:
        return oracle_lib.ref_lib_compute_peak_AB(L, *args)
    return oracle_lib.ref_compute_peak_AB(*args)


def test_compute_peak_AB(oracle_lib):
    tb, rw, args = ref_inputs()
    A0, B0 = _reference(oracle_lib, args)
    npk, nz = A0.shape
    A = np.empty_like(A0); B = np.empty_like(B0)
    ev.compute_peak_AB(npk, *args[:3], nz, *args[3:], A, B)
    # the argument th = top64 2pi/2^64 + g dy0 is rounded once (|th| <= 2pi + g |dy0|), so the
    # two builds may differ by one ulp of th on top of the sin/cos rounding
    th_max = 2 * np.pi + rw["gamma"][None, :] * np.max(np.abs(args[2]))
    tol = 2 * rw["alpha"][None, :] * (th_max * 2.3e-16 + 4.5e-16)
    assert np.all(np.abs(A - A0) <= tol) and np.all(np.abs(B - B0) <= tol)
    g = golden("ref_compat.npz")
    np.testing.assert_allclose(A[:, :8], g["A_head"], rtol=0, atol=3e-16)
    np.testing.assert_allclose(A.sum(axis=1), g["A_rowsum"], rtol=0, atol=1e-13)


def test_polish_peaks_with_weights(oracle_lib):
    tb, rw, args = ref_inputs()
    A0, B0 = _reference(oracle_lib, args)
    w = np.exp(-3e-9 * rw["gamma"] ** 2)
    npk, nz = A0.shape
    hu = np.empty(npk); hp = np.empty(npk); dd = np.empty(npk)
    ev.polish_peaks_with_weights(npk, nz, A0, B0, rw["gamma"], w, 2.5e-4, 40, hu, hp, dd)
    g = golden("ref_compat.npz")
    np.testing.assert_allclose(hu, g["h_unpol"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(hp, g["h_pol"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(dd, g["ddy_pol"], rtol=0, atol=1e-10)
