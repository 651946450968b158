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


def test_compute_peak_AB_many_peaks_and_large_arguments(oracle_lib):
    """More peaks than one grid dimension holds (70,000 > 65,535: two chunks of rows) with the
    reference's own call, and sub-grid offsets large enough that the angle leaves the range of
    the three-part pi/2 reduction (|g dy0| up to 2.6e6): the reference function accepts any int
    and any double (search/path1_eval_fast.cpp:15-53)."""
    tb, rw, args = ref_inputs()
    nz = 64
    npk = 70_000
    rng = np.random.default_rng(17)
    sel = rng.integers(0, args[0].shape[0], npk)
    dy0 = rng.uniform(-4e-3, 4e-3, npk)
    dy0[:500] = rng.uniform(-2500.0, 2500.0, 500)            # |theta| far beyond 1.6e6
    a = [args[0][sel], args[1][sel], dy0] + [x[:nz] for x in args[3:]]
    A0, B0 = oracle_lib.ref_compute_peak_AB(*a)
    A = np.empty_like(A0); B = np.empty_like(B0)
    ev.compute_peak_AB(npk, *a[:3], nz, *a[3:], A, B)
    th_max = 2 * np.pi + rw["gamma"][None, :nz] * np.abs(dy0)[:, None]
    tol = 2 * rw["alpha"][None, :nz] * (th_max * 2.3e-16 + 4.5e-16)
    assert np.all(np.isfinite(A)) and np.all(np.isfinite(B))
    assert np.all(np.abs(A - A0) <= tol) and np.all(np.abs(B - B0) <= tol)
    print(f"70,000 peaks: max |dA| {np.abs(A - A0).max():.2e}, large-argument rows "
          f"{np.abs(A - A0)[:500].max():.2e}")


def test_polish_peaks_with_weights_repeated_calls(oracle_lib):
    """The driver's loop over weight vectors (search/eval_tech1_600k.py:252-267): repeated calls
    with different weights on the same A/B give what each gives alone (the library keeps its
    device buffers and stream across calls)."""
    tb, rw, args = ref_inputs()
    A0, B0 = _reference(oracle_lib, args)
    npk, nz = A0.shape
    outs = []
    for kk in (3e-9, 1.5e-10, 1e-10, 3e-9):
        w = np.exp(-kk * rw["gamma"] ** 2)
        hu = np.empty(npk); hp = np.empty(npk); dd = np.empty(npk)
        ev.polish_peaks_with_weights(npk, nz, A0, B0, rw["gamma"], w, 2.5e-4, 40, hu, hp, dd)
        outs.append((hu, hp, dd))
    for x, y in zip(outs[0], outs[3]):
        assert np.array_equal(x, y)
    hu_r, hp_r, dd_r = oracle_lib.ref_polish_peaks_with_weights(
        A0, B0, rw["gamma"], np.exp(-1.5e-10 * rw["gamma"] ** 2), 2.5e-4, 40)
    np.testing.assert_allclose(outs[1][0], hu_r, rtol=0, atol=1e-12)
    np.testing.assert_allclose(outs[1][1], hp_r, rtol=0, atol=1e-10)


def test_shims_in_many_chunks_and_pieces(oracle_lib):
    """Arrays larger than the device staging: 2,000 peaks x 20,000 zeros are three device chunks
    of 838 rows (A and B of a chunk fill the 256 MiB buffer), each array chunk 134 MB moving
    through the two 64 MiB pinned buffers in three pieces, and the polish uploads the chunks
    into alternating device halves on its second stream.  Rows on both sides of every chunk
    boundary and a random sample against the oracle's restatement of the reference."""
    from mertens_bound_amd import formats, synth
    tb = synth.make_zero_table(20_000, 9)
    rw = formats.reference_words(tb)
    npk, nz = 2000, tb.n
    V = synth.make_candidates(npk, 3.4e18, 3.6e18, 9)
    xs, dy = formats.candidates_to_x_dy(V)
    rng = np.random.default_rng(9)
    dy = dy + rng.uniform(-4e-3, 4e-3, npk)
    xlo, xhi = formats.x_words(xs)
    zw = (rw["alpha"], rw["gamma"], rw["ghi_lo"], rw["ghi_hi"], rw["glo_lo"], rw["glo_hi"],
          rw["psi_lo"], rw["psi_hi"])
    A = np.empty((npk, nz)); B = np.empty((npk, nz))
    ev.compute_peak_AB(npk, xlo, xhi, dy, nz, *zw, A, B)
    sel = np.unique(np.concatenate([[0, 1, 837, 838, 839, 1675, 1676, 1677, 1999],
                                    rng.choice(npk, 24, replace=False)]))
    A0, B0 = oracle_lib.ref_compute_peak_AB(xlo[sel], xhi[sel], dy[sel], *zw)
    th_max = 2 * np.pi + rw["gamma"][None, :] * np.abs(dy[sel])[:, None]
    tol = 2 * rw["alpha"][None, :] * (th_max * 2.3e-16 + 4.5e-16)
    assert np.all(np.abs(A[sel] - A0) <= tol) and np.all(np.abs(B[sel] - B0) <= tol)
    w = np.exp(-3e-9 * rw["gamma"] ** 2)
    hu = np.empty(npk); hp = np.empty(npk); dd = np.empty(npk)
    ev.polish_peaks_with_weights(npk, nz, A, B, rw["gamma"], w, 2.5e-4, 5, hu, hp, dd)
    hu0, hp0, dd0 = oracle_lib.ref_polish_peaks_with_weights(A[sel], B[sel], rw["gamma"], w,
                                                             2.5e-4, 5)
    print(f"2,000 x 20,000 shims: max |dA| {np.abs(A[sel] - A0).max():.2e}, h_unpol "
          f"{np.abs(hu[sel] - hu0).max():.2e}, h_pol {np.abs(hp[sel] - hp0).max():.2e}, ddy "
          f"{np.abs(dd[sel] - dd0).max():.2e}")
    np.testing.assert_allclose(hu[sel], hu0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(hp[sel], hp0, rtol=0, atol=1e-11)
    np.testing.assert_allclose(dd[sel], dd0, rtol=0, atol=1e-9)
