"""Bit-exact phase words and the sin/cos of the HIP path against the CPU oracle."""
import numpy as np
import pytest

from mertens_bound_amd import evaluator as ev
from mertens_bound_amd import synth

pytestmark = pytest.mark.gpu
M64 = (1 << 64) - 1


@pytest.fixture(scope="module")
def small():
    tb = synth.make_zero_table(2000, 1)
    return tb, ev.K2Table(tb)


@pytest.mark.parametrize("mode", [0, 1])
def test_phase_words_bit_exact(small, oracle_lib, mode):
    tb, t = small
    V = synth.make_candidates(256, 1e18, 4e18, 1)
    W = t.phase_words(V, mode=mode)
    Wo = oracle_lib.phase_words(tb.G, V)
    assert np.array_equal(W, Wo)


def test_phase_fixup_cases_hit_and_exact(small, oracle_lib):
    """Pairs whose partial product has bits [98,110) all ones take the exact recomputation;
    they occur at rate 2^-12, so 512 x 2000 pairs contain ~250 of them."""
    tb, t = small
    V = synth.make_candidates(512, 3.4e18, 3.6e18, 2)
    W = t.phase_words(V, mode=0)
    Wo = oracle_lib.phase_words(tb.G, V)
    assert np.array_equal(W, Wo)
    hits = 0
    for c, v in enumerate(synth.v_to_int(V[:64])):
        for j, g in enumerate(synth.v_to_int(tb.G)):
            gl = [(g >> (32 * i)) & 0xFFFFFFFF for i in range(4)]
            vl = [(v >> (32 * i)) & 0xFFFFFFFF for i in range(4)]
            U = sum(gl[i] * vl[j2] << (32 * (i + j2)) for i in range(4) for j2 in range(4)
                    if i + j2 >= 2)
            hits += ((U >> 98) & 0xFFF) == 0xFFF
    assert hits > 0


def test_split_table_words_bit_exact(small, oracle_lib):
    """The one-launch kernel's remainder words (packed column 5, carry case -- product bits
    [98,111) all ones, rate 2^-13 -- fixed up from the columns) give W bits [1,64) exactly, on
    the pairs of the fix-up test (about 125 carry cases) and on edge words."""
    tb, t = small
    V = synth.make_candidates(512, 3.4e18, 3.6e18, 2)
    Wo = oracle_lib.phase_words(tb.G, V)
    assert np.array_equal(t.phase_words(V, mode=4), Wo & np.uint64(M64 - 1))
    hits = 0
    for v in synth.v_to_int(V[:128]):
        vl = [(v >> (32 * i)) & 0xFFFFFFFF for i in range(4)]
        for g in synth.v_to_int(tb.G):
            gl = [(g >> (32 * i)) & 0xFFFFFFFF for i in range(4)]
            U = sum(gl[i] * vl[j2] << (32 * (i + j2)) for i in range(4) for j2 in range(4)
                    if i + j2 >= 2)
            hits += ((U >> 98) & 0x1FFF) == 0x1FFF
    assert hits > 0   # the fix-up path ran (about 30 of these 256,000 pairs)
    edge = [0, 1, (1 << 128) - 1, 1 << 110, (1 << 127) + 12345, M64, 1 << 64, (1 << 98) - 1]
    G = synth.v_from_int(edge)
    te = ev.K2Table(synth.ZeroTable(G=G, gamma=np.ones(len(edge)), w=np.zeros(len(edge), complex)))
    Ve = synth.v_from_int(edge + [(3 << 64) + (1 << 62)])
    assert np.array_equal(te.phase_words(Ve, mode=4),
                          oracle_lib.phase_words(G, Ve) & np.uint64(M64 - 1))


def test_phase_words_edges(oracle_lib):
    edge = [0, 1, (1 << 128) - 1, 1 << 110, (1 << 127) + 12345, M64, 1 << 64, (1 << 98) - 1]
    G = synth.v_from_int(edge)
    tb = synth.ZeroTable(G=G, gamma=np.ones(len(edge)), w=np.zeros(len(edge), complex))
    t = ev.K2Table(tb)
    V = synth.v_from_int(edge + [(3 << 64) + (1 << 62)])
    for mode in (0, 1):
        assert np.array_equal(t.phase_words(V, mode=mode), oracle_lib.phase_words(G, V))
    _check_window(t.phase_words(V, mode=2), oracle_lib.phase_words(G, V))
    _check_window_f(t.phase_words(V, mode=3), oracle_lib.phase_words(G, V), check_mean=False)


def _check_window(win, Wo):
    """The FP32 form's window is bits [128,174) of the product without columns 0..2, which sum
    to less than 2^130 = 4 units of bit 128: floor(W / 2^18) minus 0..4 (mod 2^46)."""
    m46 = np.uint64((1 << 46) - 1)
    assert np.all(win <= m46)
    d = ((Wo >> np.uint64(18)) - win) & m46
    assert d.max() <= 4
    return d


def test_phase_window_within_4_units(small, oracle_lib):
    tb, t = small
    V = synth.make_candidates(512, 3.4e18, 3.6e18, 3)
    d = _check_window(t.phase_words(V, mode=2), oracle_lib.phase_words(tb.G, V))
    # the dropped columns (about 1.5 units of bit 128 on average) do move the window for most
    # pairs, so the bound is exercised, not vacuous
    assert 0 < np.count_nonzero(d) < d.size


def test_sincos_of_words():
    rng = np.random.default_rng(0)
    W = rng.integers(0, 2**64, size=100_000, dtype=np.uint64)
    W[:8] = np.array([0, 1 << 62, 1 << 63, 3 << 62, (1 << 61), (1 << 61) - 1, M64, 1 << 60],
                     dtype=np.uint64)
    s, c = ev.sincos_words(W, mode=0)
    th = 2 * np.pi * (W.astype(np.float64) / 2.0 ** 64)
    import mpmath
    # exact reference on a sample (40 digits)
    with mpmath.workdps(40):
        for i in list(range(8)) + list(range(8, 100_000, 997)):
            t = 2 * mpmath.pi * mpmath.mpf(int(W[i])) / mpmath.mpf(2) ** 64
            assert abs(s[i] - float(mpmath.sin(t))) <= 2.3e-16
            assert abs(c[i] - float(mpmath.cos(t))) <= 2.3e-16
    # numpy evaluates at th rounded to double (|th| up to 2pi: ulp 8.9e-16)
    assert np.max(np.abs(s - np.sin(th))) < 2e-15
    s32, c32 = ev.sincos_words(W, mode=1)
    assert np.max(np.abs(s32 - np.sin(th))) < 3e-7 and np.max(np.abs(c32 - np.cos(th))) < 3e-7


def test_sincos_split_table_of_words():
    """The one-launch kernel's sin/cos (2,048-point split table, sin x = x - x^3/6): against long
    double over random words and both edges of every interval (the largest remainders, where the
    dropped x^5/120 is 0.64 u), and against mpmath on a sample."""
    U = 2.0 ** -53
    rng = np.random.default_rng(5)
    W = rng.integers(0, 2**64, size=1 << 20, dtype=np.uint64)
    k = np.arange(2048, dtype=np.uint64)
    step = np.uint64(1 << 53)
    W[:2048] = k << np.uint64(53)                               # largest negative remainder
    W[2048:4096] = (k << np.uint64(53)) + step - np.uint64(1)   # largest positive remainder
    W[4096:4104] = np.array([0, 1, 1 << 62, 1 << 63, 3 << 62, M64, 1 << 53, (1 << 53) - 1],
                            dtype=np.uint64)
    s, c = ev.sincos_words(W, mode=2)
    two_pi = np.longdouble("6.283185307179586476925286766559005768")
    th = two_pi * (W.astype(np.longdouble) / np.longdouble(2.0) ** 64)
    es = np.abs(s.astype(np.longdouble) - np.sin(th)) / U
    ec = np.abs(c.astype(np.longdouble) - np.cos(th)) / U
    print(f"sincos_tabw max error: sin {float(es.max()):.3f} u, cos {float(ec.max()):.3f} u")
    assert float(es.max()) <= 3.0 and float(ec.max()) <= 3.0
    # a wrong interval, half-step or sign misses by 1e-3 or more
    import mpmath
    with mpmath.workdps(40):
        for i in list(range(0, 4104, 257)) + list(range(4104, 1 << 20, 9973)):
            t = 2 * mpmath.pi * mpmath.mpf(int(W[i])) / mpmath.mpf(2) ** 64
            assert abs(s[i] - float(mpmath.sin(t))) <= 3.5e-16
            assert abs(c[i] - float(mpmath.cos(t))) <= 3.5e-16


def _check_window_f(win, Wo, check_mean=True):
    """The FP32 kernels' window sums column 3 in FP32, in radians with the table offset folded
    in (phase_window_f): the zero's limbs times 2 pi / 2^44 and the candidate's times 2^-34,
    each rounded once to float, four FMAs starting from minus 3/4 of a step.  Worst case
    2 2^-24 S3 with S3 < 2^66 (units of 2^96), 2048 units of bit 128, plus the FMAs' four
    roundings at |t| < 2^-8 rad, 2^-33 rad each = 5218 units, plus 4 for columns 0..2 and 1 for
    the conversion to units here: |window - floor(W / 2^18)| <= 7271 (mod 2^46).  The roundings
    are to nearest and each limb's own, so the mean error stays within tens of units over random
    words (not over a few dozen extreme ones)."""
    m46 = (1 << 46) - 1
    assert np.all(win <= np.uint64(m46))
    d = ((win.astype(object) - (Wo >> np.uint64(18)).astype(object)) & m46)
    d = np.array([x - (1 << 46) if x >= (1 << 45) else x for x in d.ravel()], dtype=np.int64)
    assert np.abs(d).max() <= 7271
    if check_mean:
        assert abs(d.mean()) <= 64
    return d


def test_phase_window_f_within_bound(small, oracle_lib):
    tb, t = small
    V = synth.make_candidates(512, 3.4e18, 3.6e18, 3)
    d = _check_window_f(t.phase_words(V, mode=3), oracle_lib.phase_words(tb.G, V))
    # FP32 rounding moves the window by tens to thousands of units, so the bound is exercised
    assert np.abs(d).max() > 64


def test_phase_words_full_table(oracle_lib):
    """All 2,325,572 zeros of the full-size table against candidates spread over
    [3.4e18, 3.6e18]: every 64-bit word bit-exact (the one-launch kernel's words in W bits
    [1,64)), every FP32 window within 4 units."""
    c = synth.W65536x2325572
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    t = ev.K2Table(tb)
    V = np.concatenate([synth.v_from_int([int(3.4e18) << 64, (int(3.5e18) << 64) + 12345,
                                          (int(3.6e18) << 64) - 1]),
                        synth.make_candidates(1, c["v_lo"], c["v_hi"], 77)])
    Wo = oracle_lib.phase_words(tb.G, V)
    assert np.array_equal(t.phase_words(V, mode=0), Wo)
    assert np.array_equal(t.phase_words(V, mode=4), Wo & np.uint64(M64 - 1))
    _check_window(t.phase_words(V, mode=2), Wo)
    _check_window_f(t.phase_words(V, mode=3), Wo)
