"""Pins the CPU oracle (oracle/) before anything is checked against it.

Each oracle tier is checked against something that does not retype its own formula:
Python big-integer arithmetic, mpmath, closed forms, structural identities between the value
and derivative formulas, finite differences in binary128, and the reference evaluator compiled
from its own source (oracle/_ref, where built).
"""
import numpy as np
import pytest

from conftest import golden
from mertens_bound_amd import formats, synth

M64 = (1 << 64) - 1
M128 = (1 << 128) - 1


def _table(n, seed=1):
    return synth.make_zero_table(n, seed)


def _one_zero(G: int, w: complex):
    Gw = np.array([[G & M64, G >> 64]], dtype=np.uint64)
    g = synth.gamma_from_words(Gw)
    return synth.ZeroTable(G=Gw, gamma=g, w=np.array([w], dtype=np.complex128))


def test_phase_words_match_big_integers(oracle_lib):
    rng = np.random.default_rng(7)
    G = rng.integers(0, 2**64, size=(64, 2), dtype=np.uint64)
    V = rng.integers(0, 2**64, size=(16, 2), dtype=np.uint64)
    edge = [0, 1, M128, 1 << 110, (1 << 127) + 12345, M64, 1 << 64]
    G = np.concatenate([G, synth.v_from_int(edge)])
    V = np.concatenate([V, synth.v_from_int(edge)])
    W = oracle_lib.phase_words(G, V)
    for c, v in enumerate(synth.v_to_int(V)):
        for j, g in enumerate(synth.v_to_int(G)):
            assert int(W[c, j]) == ((g * v) >> 110) & M64


def test_phase_word_quarter_turn_closed_form(oracle_lib):
    # G = 2^110 is gamma = 2 pi; v = n + 1/4 turns the phase by exactly a quarter
    for n in (0, 1, 3 * 10**18, (1 << 63) - 1):
        V = synth.v_from_int([(n << 64) + (1 << 62)])
        W = oracle_lib.phase_words(synth.v_from_int([1 << 110]), V)
        assert int(W[0, 0]) == 1 << 62


@pytest.mark.parametrize("tier", ["f64", "quad"])
def test_single_zero_closed_forms(oracle_lib, tier):
    k = 1.5e-10
    w = complex(0.03, -0.07)
    tb = _one_zero(1 << 110, w)                  # gamma = 2 pi exactly
    n = 3 * 10**18
    V = synth.v_from_int([n << 64, (n << 64) + (1 << 62), (n << 64) + (1 << 63)])
    if tier == "f64":
        K, dK, d2K = oracle_lib.eval_f64(tb, V, k)
    else:
        K, dK, d2K, _ = oracle_lib.eval_quad(tb, V, k)
    a = 2 * np.exp(-k * (2 * np.pi) ** 2)
    g = 2 * np.pi
    # e^{i theta} = 1, i, -1 at v = n, n + 1/4, n + 1/2
    np.testing.assert_allclose(K, [a * w.real, -a * w.imag, -a * w.real], rtol=0, atol=1e-15)
    # K' = a g Re(i w e^{i theta}), K'' = -g^2 K
    np.testing.assert_allclose(dK, [-a * g * w.imag, -a * g * w.real, a * g * w.imag],
                               rtol=0, atol=1e-14)
    np.testing.assert_allclose(d2K, -g * g * K, rtol=1e-14, atol=1e-14)


def test_v_zero_is_weight_sum(oracle_lib):
    tb = _table(300)
    k = 3e-9
    V = synth.v_from_int([0])
    K, _, _, _ = oracle_lib.eval_quad(tb, V, k)
    ref = 2 * np.sum(np.exp(-k * tb.gamma ** 2) * tb.w.real)
    assert abs(K[0] - ref) < 1e-15


def test_linearity_and_derivative_identities(oracle_lib):
    """K is linear in w; K'[w] = K[i g w] and K''[w] = -K[g^2 w] tie the derivative formulas
    to the value formula."""
    tb = _table(400, seed=3)
    V = synth.make_candidates(6, 1e18, 4e18, 3)
    k = 1e-9
    K1, dK1, d2K1, _ = oracle_lib.eval_quad(tb, V, k)
    w2 = np.random.default_rng(1).normal(size=tb.n) + 1j * np.random.default_rng(2).normal(size=tb.n)
    tb2 = synth.ZeroTable(tb.G, tb.gamma, w2 * 1e-2)
    tbs = synth.ZeroTable(tb.G, tb.gamma, tb.w + w2 * 1e-2)
    K2, _, _, _ = oracle_lib.eval_quad(tb2, V, k)
    Ks, _, _, _ = oracle_lib.eval_quad(tbs, V, k)
    np.testing.assert_allclose(Ks, K1 + K2, rtol=0, atol=1e-14)
    tbi = synth.ZeroTable(tb.G, tb.gamma, 1j * tb.gamma * tb.w)
    Ki, _, _, _ = oracle_lib.eval_quad(tbi, V, k)
    np.testing.assert_allclose(dK1, Ki, rtol=1e-12, atol=1e-12)
    tbg = synth.ZeroTable(tb.G, tb.gamma, tb.gamma ** 2 * tb.w)
    Kg, _, _, _ = oracle_lib.eval_quad(tbg, V, k)
    np.testing.assert_allclose(d2K1, -Kg, rtol=1e-12, atol=1e-9)


def test_quad_matches_mpmath(oracle_lib):
    tb = _table(60, seed=4)
    V = synth.make_candidates(3, 3.4e18, 3.6e18, 4)
    k = 1.5e-10
    K, dK, d2K, _ = oracle_lib.eval_quad(tb, V, k)
    mK, mdK, md2K = oracle_lib.mp_k2(tb.G, tb.w, V, k, dps=40)
    np.testing.assert_allclose(K, [float(x) for x in mK], rtol=0, atol=1e-16)
    np.testing.assert_allclose(dK, [float(x) for x in mdK], rtol=1e-15, atol=1e-14)
    np.testing.assert_allclose(d2K, [float(x) for x in md2K], rtol=1e-15, atol=1e-10)


def test_derivatives_match_binary128_differences(oracle_lib):
    tb = _table(200, seed=5)
    k = 1.5e-10
    v0 = synth.v_to_int(synth.make_candidates(4, 3.4e18, 3.6e18, 5))
    _, dK, d2K, _ = oracle_lib.eval_quad(tb, synth.v_from_int(v0), k)
    # first derivative: h = 2^-30 (truncation (g h)^2/6 < 1e-13 relative)
    h = 1 << 34
    hp, lp = oracle_lib.eval_quad_pair(tb, synth.v_from_int([v + h for v in v0]), k)
    hm, lm = oracle_lib.eval_quad_pair(tb, synth.v_from_int([v - h for v in v0]), k)
    fd1 = ((hp - hm) + (lp - lm)) / (2 * 2.0 ** -30)
    np.testing.assert_allclose(fd1, dK, rtol=1e-10, atol=1e-9)
    # second derivative: h = 2^-16
    h2 = 1 << 48
    h0, l0 = oracle_lib.eval_quad_pair(tb, synth.v_from_int(v0), k)
    hp, lp = oracle_lib.eval_quad_pair(tb, synth.v_from_int([v + h2 for v in v0]), k)
    hm, lm = oracle_lib.eval_quad_pair(tb, synth.v_from_int([v - h2 for v in v0]), k)
    fd2 = ((hp - 2 * h0 + hm) + (lp - 2 * l0 + lm)) / (2.0 ** -16) ** 2
    np.testing.assert_allclose(fd2, d2K, rtol=1e-6, atol=1e-6 * np.max(np.abs(d2K)))


def test_f64_tier_tracks_quad(oracle_lib):
    tb = synth.make_zero_table(10_000, 0)
    V = synth.make_candidates(32, 1e18, 4e18, 0)
    K, dK, d2K = oracle_lib.eval_f64(tb, V, 1.5e-10)
    Kq, dKq, d2Kq, aS = oracle_lib.eval_quad(tb, V, 1.5e-10)
    assert np.max(np.abs(K - Kq)) < 1e-13
    assert np.max(np.abs(dK - dKq)) < 1e-9


def test_golden_w1024x10000_reproduces(oracle_lib):
    g = golden("w1024x10000_oracle.npz")
    c = synth.W1024x10000
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    assert str(g["table_sha256"]) == tb.sha256()
    V = synth.make_candidates(c["n_cand"], c["v_lo"], c["v_hi"], c["seed"])
    assert np.array_equal(V, g["V"])
    K, dK, d2K, _ = oracle_lib.eval_quad(tb, V[:16], c["k"])
    assert np.array_equal(K, g["K"][:16]) and np.array_equal(dK, g["dK"][:16])


def test_polish_oracle_reaches_stationary_point(oracle_lib):
    g = golden("polish_small.npz")
    tb = synth.make_zero_table(3000, 11)
    assert str(g["table_sha256"]) == tb.sha256()
    Vout, K0, Kf, dKf, st = oracle_lib.polish_f64(tb, g["V0"], float(g["k"]))
    assert np.array_equal(Vout, g["V"]) and np.array_equal(Kf, g["K"])
    assert np.all(np.abs(Kf) >= np.abs(K0))
    # scale-aware stationarity: |K'| small against sum g |a|
    scale = np.sum(2 * np.exp(-float(g["k"]) * tb.gamma ** 2) * tb.gamma * np.abs(tb.w))
    # candidates whose Newton loop stopped on |step| < 1e-13 sit on a stationary point
    done = (st < 8) & np.any(Vout != g["V0"], axis=1)
    assert done.sum() >= 4
    assert np.all(np.abs(dKf[done]) < 1e-9 * scale)


def test_restated_reference_matches_compiled_reference(oracle_lib):
    L = oracle_lib.load_ref()
    if L is None:
        pytest.skip("oracle/_ref not built (reference source not mounted where it was built)")
    from golden.make_golden import ref_inputs
    tb, rw, args = ref_inputs()
    A0, B0 = oracle_lib.ref_lib_compute_peak_AB(L, *args)
    A1, B1 = oracle_lib.ref_compute_peak_AB(*args)
    assert np.array_equal(A0, A1) and np.array_equal(B0, B1)
    w = np.exp(-3e-9 * rw["gamma"] ** 2)
    r0 = oracle_lib.ref_lib_polish(L, A0, B0, rw["gamma"], w)
    r1 = oracle_lib.ref_polish_peaks_with_weights(A0, B0, rw["gamma"], w)
    for a, b in zip(r0, r1):
        assert np.array_equal(a, b)
    gold = golden("ref_compat.npz")
    assert np.array_equal(gold["h_pol"], r0[1]) and np.array_equal(gold["A_head"], A0[:, :8])


def test_reference_format_matches_canonical(oracle_lib):
    """The reference's two-word phase (ghi/glo split, psi) on canonical-derived words gives the
    canonical K2: 2 Re(w e^{i g y}) = A with y = x/2^30 + dy0 (path1_eval_fast.cpp:7-14)."""
    tb = synth.make_zero_table(500, 9)
    rw = formats.reference_words(tb)
    V = synth.make_candidates(8, 3.4e18, 3.6e18, 9)
    xs, dy = formats.candidates_to_x_dy(V)
    xlo, xhi = formats.x_words(xs)
    A, B = oracle_lib.ref_compute_peak_AB(xlo, xhi, dy, rw["alpha"], rw["gamma"], rw["ghi_lo"],
                                          rw["ghi_hi"], rw["glo_lo"], rw["glo_hi"], rw["psi_lo"],
                                          rw["psi_hi"])
    k = 1.5e-10
    h0 = A @ np.exp(-k * tb.gamma ** 2)
    K, _, _, _ = oracle_lib.eval_quad(tb, V, k)
    # dy0 < 2^-30 is applied in FP64 as g * dy0 (<= 1e-5 rad), psi carries 53 bits
    np.testing.assert_allclose(h0, K, rtol=0, atol=5e-12)


def _one_zero_peak():
    """One zero with gamma = 2 pi G / 2^110 ~ 1000.3 and w = 0.6 e^{0.4 i}: K2(v) =
    2 e^{-k g^2} 0.6 cos(g v + 0.4), whose extrema sit exactly at g v + 0.4 = m pi."""
    import mpmath
    from mertens_bound_amd import synth
    G = int(round(1000.3 / (2 * np.pi) * 2.0 ** 110))
    with mpmath.workdps(50):
        g = 2 * mpmath.pi * mpmath.mpf(G) / mpmath.mpf(2) ** 110
        m = int(mpmath.nint((5 * g + mpmath.mpf("0.4")) / mpmath.pi))
        vstar = (m * mpmath.pi - mpmath.mpf("0.4")) / g
    w = np.array([0.6 * np.exp(0.4j)])
    tb = synth.ZeroTable(np.array([[G & (2 ** 64 - 1), G >> 64]], dtype=np.uint64),
                         np.array([float(g)]), w)
    return tb, vstar, float(g)


def test_pipeline_restatements_reach_the_one_zero_peak(oracle_lib):
    import mpmath
    from oracle import pipelines as P
    from mertens_bound_amd import synth
    tb, vstar, g = _one_zero_peak()
    k = 1e-9
    amp = 2 * np.exp(-k * g * g) * 0.6
    offs = np.array([3e-6, -7e-6, 1.2e-5])                 # start points near the peak
    V0 = synth.v_from_int([int(mpmath.nint((vstar + mpmath.mpf(float(o))) * 2 ** 64))
                           for o in offs])
    want_d = np.array([float(vstar - mpmath.mpf(v) / mpmath.mpf(2) ** 64)
                       for v in synth.v_to_int(V0)])
    h0, h, d = P.polish_single_peak(tb, V0, k)
    assert np.allclose(np.abs(h), amp, rtol=0, atol=1e-15)
    assert np.max(np.abs(d - want_d)) < 1e-12
    assert np.all(np.abs(h) >= np.abs(h0))
    hu, hp, dp = P.fast_polish_batch(tb, V0, k)
    assert np.allclose(np.abs(hp), amp, rtol=0, atol=1e-15)
    assert np.max(np.abs(dp - want_d)) < 1e-12
    # scan: the grid point nearest the peak (cos is even about it) is the first maximum
    step = 1e-4
    hb, mb = P.scan(tb, V0, k, half_width=2e-3, step=step)
    assert np.array_equal(mb, np.rint(want_d / step).astype(np.int64))
    assert np.all(np.abs(hb) <= amp + 1e-15)
    # the two Newton variants differ in clamp and tolerance only; a start too far for the clamp
    # budget (8 x 2e-5) stays short of the peak
    far = synth.v_from_int([int(mpmath.nint((vstar + mpmath.mpf("4e-4")) * 2 ** 64))])
    d_far, _ = P.newton(tb, far, k, 1, 8, 2e-5, 1e-14)
    assert abs(d_far[0] + 8 * 2e-5) < 1e-15


def _one_zero_table(amp, k):
    """One zero (gamma ~ 1000.3) with |w| chosen so that 2 e^{-k g^2} |w| = amp, arg w = 0.4."""
    import mpmath
    from mertens_bound_amd import synth
    G = int(round(1000.3 / (2 * np.pi) * 2.0 ** 110))
    with mpmath.workdps(50):
        g = 2 * mpmath.pi * mpmath.mpf(G) / mpmath.mpf(2) ** 110
    r = amp / (2 * np.exp(-k * float(g) ** 2))
    tb = synth.ZeroTable(np.array([[G & (2 ** 64 - 1), G >> 64]], dtype=np.uint64),
                         np.array([float(g)]), np.array([r * np.exp(0.4j)]))
    return tb, g


def _nearest_peak_cascade(g, amp, v0, stages):
    """Closed form of ZeroData::polish on one zero: K2(v) = amp cos(g v + 0.4), so in every grid
    stage the first point of largest |K2| is the first point whose phase g v + 0.4 lies nearest
    a multiple of pi (computed in mpmath at the exact word of each point)."""
    import mpmath
    best_h, best_dy, passed = 0.0, 0.0, 0
    with mpmath.workdps(50):
        for nz, M, step, thr, centred, sets_h in stages:
            center = best_dy if centred else 0.0
            best = None
            for m in range(-M, M + 1):
                dy = center + m * step
                v = mpmath.mpf(v0 + int(round(dy * 2.0 ** 64))) / mpmath.mpf(2) ** 64
                t = (g * v + mpmath.mpf("0.4")) / mpmath.pi
                dist = abs(t - mpmath.nint(t))
                if best is None or dist < best[0]:
                    best = (dist, dy, amp * mpmath.cos(mpmath.pi * t))
            best_dy = best[1]
            if sets_h:
                best_h = float(best[2])
            if thr is not None:
                if abs(float(best[2])) < thr:
                    break
                passed += 1
    return best_h, best_dy, passed


def test_cascade_restatement_on_one_zero(oracle_lib):
    """The oracle's five-stage ZeroData::polish (reference stage table) against the closed form:
    amplitudes 0.99 (every stage passes: the full-table stage sets the returned point and h),
    0.93 (stops at the 2,048-zero threshold 0.948), 0.9 (stops at 0.918) and 0.5 (stops at the
    first stage's 0.895)."""
    from oracle import pipelines as P
    from mertens_bound_amd import synth
    k = 3e-9
    rng = np.random.default_rng(5)
    for amp, want_passed in ((0.99, 3), (0.93, 2), (0.9, 1), (0.5, 0)):
        tb, g = _one_zero_table(amp, k)
        v0s = [int(x) for x in rng.integers(1 << 60, 1 << 61, size=2, dtype=np.uint64)]
        V = synth.v_from_int(v0s)
        H, D, Pa = P.cascade(tb, V, k)
        for i, v0 in enumerate(v0s):
            h, dy, passed = _nearest_peak_cascade(g, amp, v0, P.CASCADE_STAGES)
            assert Pa[i] == passed == want_passed, (amp, Pa[i], passed)
            assert D[i] == dy, (amp, D[i], dy)
            assert abs(H[i] - h) <= 1e-15
        if want_passed == 3:
            # the returned point lies within half a final grid step (2.5e-7) of a peak
            assert np.all(np.abs(np.abs(H) - amp) <= amp * (1 - np.cos(float(g) * 2.5e-7)) + 1e-15)
