"""The FP32 form's phase window and table sin/cos, restated in exact integers and numpy float32
on the CPU and checked against the exact phase: the bounds the kernels rely on (the window is
floor(phase 2^46) minus 0..4; sin/cos within FP32 rounding of the exact values)."""
import numpy as np

from mertens_bound_amd import synth

M64 = (1 << 64) - 1


def _limbs(x):
    return [(x >> (32 * i)) & 0xFFFFFFFF for i in range(4)]


def _window(G, V, tb=11):
    """Columns 3..5 of the limb product plus half a table step, bits [128, 192) (as the kernels
    form them: columns 0..2 are skipped)."""
    g, v = _limbs(G), _limbs(V)
    P = sum(g[i] * v[j] << (32 * (i + j)) for i in range(4) for j in range(4) if i + j >= 3)
    P += 1 << (173 - tb)
    return (P >> 128) & M64


def _cases(n, seed):
    rng = np.random.default_rng(seed)
    Gs = [int(x) for x in synth.v_to_int(synth.make_zero_table(n, seed).G)]
    Vs = [int(x) for x in synth.v_to_int(synth.make_candidates(n, 3.4e18, 3.6e18, seed))]
    edge = [(1 << 128) - 1, (1 << 127) | ((1 << 96) - 1), 1 << 110, (1 << 96) - 1, 0xFFFFFFFF << 64]
    pairs = list(zip(Gs, Vs)) + [(a, b) for a in edge for b in edge]
    pairs += [(int(rng.integers(0, 2**63)) << 65 | int(rng.integers(0, 2**63)),
               int(rng.integers(0, 2**63)) << 65 | int(rng.integers(0, 2**63))) for _ in range(n)]
    return pairs


def test_window_is_floor_minus_at_most_4():
    m46 = (1 << 46) - 1
    diffs = []
    for G, V in _cases(2000, 1):
        exact = ((G * V) >> 128) & m46
        win = (_window(G, V) - (1 << 34)) & m46
        diffs.append((exact - win) & m46)
    diffs = np.array(diffs)
    assert diffs.max() <= 4
    assert diffs.min() == 0 and diffs.max() >= 2   # the skipped columns do carry


def test_window_sincos_float32_within_rounding():
    """Index = top 11 bits (rounded), remainder R -> x = 2 pi (R - 2^31) 2^-43 in FP32,
    cos(t + x) = c - x (s + c x/2), sin(t + x) = s + x (c - s x/2) in FP32; compared with
    sin/cos of the exact phase (to 2^-64 turn, evaluated in float64)."""
    f32 = np.float32
    two_pi = 2 * np.pi
    k = np.arange(2048)
    tab_c = np.cos(two_pi * k / 2048).astype(f32)
    tab_s = np.sin(two_pi * k / 2048).astype(f32)
    step = f32(two_pi) / f32(2.0 ** 43)
    half = step * f32(2.0 ** 31)
    errs = []
    for G, V in _cases(3000, 2):
        w = _window(G, V)
        hi, lo = w >> 32, w & 0xFFFFFFFF
        idx = (hi >> 3) & 0x7FF
        R = ((hi << 32 | lo) >> 3) & 0xFFFFFFFF
        u = f32(R)
        x = f32(u * step - half)        # the kernels fuse these two roundings into one FMA
        h = f32(u * f32(0.5 * step) - f32(0.5 * half))
        c, s = tab_c[idx], tab_s[idx]
        C = f32(c - x * f32(s + h * c))
        S = f32(s + x * f32(c - h * s))
        phase = ((G * V) >> 110) & M64          # exact to 2^-64 turn
        th = two_pi * (phase / 2.0 ** 64)
        errs.append(max(abs(float(C) - np.cos(th)), abs(float(S) - np.sin(th))))
    errs = np.array(errs)
    # table rounding (<= 2^-25 each), two FP32 roundings of the rotation, |x|^3/6 <= 6e-10:
    # well inside 2.5e-7; a wrong sign, index or half step misses by ~1e-3 or more
    assert errs.max() < 2.5e-7, errs.max()


def _fma32(a, b, c):
    """float32 fused multiply-add: the product of two float32 is exact in float64, the sum
    rounds once in float64 and once more to float32 (a double rounding that can move the last
    bit, which the bound below does not rely on)."""
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


TWO_PI = 6.283185307179586476925286766559005768


def _window_consts(tb=11):
    """The kernels' FP32 constants: the remainder step 2 pi 2^-(32+tb) and three quarters of a
    table step, as the device code forms them in float32."""
    f32 = np.float32
    step = f32(f32(TWO_PI) / f32(2.0 ** (32 + tb)))
    return step, f32(step * f32(3221225472.0))


def _window_f(G, V, tb=11):
    """The FP32 kernels' window (phase_window_f): columns 4..5 exact plus three quarters of a
    table step (column 5 mod 2^32, of which the kernels use the low 14 bits), and beside it
    column 3 summed in float32, in radians, from minus three quarters of a step: the zero's limbs
    times 2 pi / 2^44 and the candidate's times 2^-34 as floats, t3 = 2 pi S3 / 2^78 - 3/4 step."""
    g, v = _limbs(G), _limbs(V)
    _, tq = _window_consts(tb)
    gscale = TWO_PI / 2.0 ** 44
    gf = [np.float32(float(x) * gscale) for x in g]
    vf = [np.float32(np.float32(x) * np.float32(2.0 ** -34)) for x in v]
    t = _fma32(gf[3], vf[0], -tq)
    t = _fma32(gf[2], vf[1], t)
    t = _fma32(gf[1], vf[2], t)
    t = _fma32(gf[0], vf[3], t)
    s4 = (g[1] * v[3] + g[2] * v[2] + g[3] * v[1] + (3 << (44 - tb))) & M64
    s5 = ((s4 >> 32) + g[2] * v[3] + g[3] * v[2]) & 0xFFFFFFFF
    return (s5 << 32) | (s4 & 0xFFFFFFFF), t


def test_window_f_column3_in_float32_within_bound():
    """The phase the FP32 kernels use, window - offset + (t3 + 3/4 step) / (2 pi 2^-46) (units
    of bit 128): column 3 in FP32 errs by at most 2 2^-24 S3 with S3 < 2^66 (units of 2^96),
    2048 units of bit 128, plus four FMA roundings at |t| < 2^-8 rad (2^-33 rad = 1304.5 units
    each) and 4 for columns 0..2: 7270 units.  Rounding to nearest keeps the mean within a few
    units.  A dropped or mis-scaled column-3 term misses by 2^32 units or more, a wrong offset by
    2^33."""
    from fractions import Fraction
    m46 = 1 << 46
    _, tq = _window_consts()
    unit = TWO_PI / 2.0 ** 46
    d = []
    for G, V in _cases(3000, 4):
        exact = ((G * V) >> 128) % m46
        w, t3 = _window_f(G, V)
        eff = Fraction(w - (3 << 33)) + Fraction((float(t3) + float(tq)) / unit)
        x = (eff - exact) % m46
        d.append(float(x - m46 if x >= m46 // 2 else x))
    d = np.array(d)
    assert np.abs(d).max() <= 7270, np.abs(d).max()
    assert abs(d.mean()) <= 64, d.mean()
    assert np.abs(d).max() > 16          # FP32 rounding does move the phase


def test_window_f_sincos_float32_within_rounding():
    """sincos_window_f: index = top 11 bits of the window (rounded at three quarters of a step),
    x = R 2 pi 2^-43 + t3 in FP32 (one FMA; t3 carries column 3 and minus 3/4 of a step), h = x/2,
    and the rotation; compared with sin/cos of the exact phase.  |x| stays within 3/4 of a step,
    so sin x = x errs by at most 2e-9: the total stays inside 2.5e-7, where a wrong offset, scale
    or index misses by 1e-3 or more."""
    f32 = np.float32
    two_pi = 2 * np.pi
    k = np.arange(2048)
    tab_c = np.cos(two_pi * k / 2048).astype(f32)
    tab_s = np.sin(two_pi * k / 2048).astype(f32)
    step, _ = _window_consts()
    errs = []
    for G, V in _cases(3000, 5):
        w, t3 = _window_f(G, V)
        hi, lo = w >> 32, w & 0xFFFFFFFF
        idx = (hi >> 3) & 0x7FF
        R = ((hi << 32 | lo) >> 3) & 0xFFFFFFFF
        x = _fma32(f32(R), step, t3)
        h = f32(f32(0.5) * x)
        c, s = tab_c[idx], tab_s[idx]
        C = _fma32(-x, _fma32(h, c, s), c)
        S = _fma32(x, _fma32(-h, s, c), s)
        phase = ((G * V) >> 110) & M64
        th = two_pi * (phase / 2.0 ** 64)
        errs.append(max(abs(float(C) - np.cos(th)), abs(float(S) - np.sin(th))))
    errs = np.array(errs)
    assert errs.max() < 2.5e-7, errs.max()
