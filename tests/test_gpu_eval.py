"""K2 / K2' / K2'' of the HIP evaluator against the oracle (golden fixtures and live CPU runs).

Tolerances: K2 within 1e-9 absolute (the accuracy bar of the metric; FP64 lands near 1e-13);
derivatives in a scale-aware relative form |dK'| / sum_j g_j |a_j| <= 1e-12 (FP64) and
|dK''| / sum_j g_j^2 |a_j| <= 1e-12, since K' ~ 0 near peaks makes plain relative error
meaningless.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden
from mertens_bound_amd import evaluator as ev
from mertens_bound_amd import synth

pytestmark = pytest.mark.gpu
ALL = ev.VALUE | ev.D1 | ev.D2


@pytest.fixture(scope="module")
def w10k():
    c = synth.W1024x10000
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    V = synth.make_candidates(c["n_cand"], c["v_lo"], c["v_hi"], c["seed"])
    return tb, ev.K2Table(tb), V, c["k"]


def _scales(tb, k, n=None):
    n = tb.n if n is None else n
    a = 2 * np.exp(-k * tb.gamma[:n] ** 2) * np.abs(tb.w[:n])
    return np.sum(a), np.sum(tb.gamma[:n] * a), np.sum(tb.gamma[:n] ** 2 * a)


def test_w1024x10000_plain_vs_golden(w10k):
    tb, t, V, k = w10k
    g = golden("w1024x10000_oracle.npz")
    K, dK, d2K = t.eval(V, k, flags=ALL)
    s0, s1, s2 = _scales(tb, k)
    errK = np.max(np.abs(K - g["K"]))
    print(f"w1024x10000 plain max|dK2|={errK:.3e}")
    assert errK <= 1e-12
    assert np.max(np.abs(dK - g["dK"])) / s1 <= 1e-12
    assert np.max(np.abs(d2K - g["d2K"])) / s2 <= 1e-12


def test_w1024x10000_mixed_vs_golden(w10k):
    tb, t, V, k = w10k
    g = golden("w1024x10000_oracle.npz")
    K, dK, d2K = t.eval(V, k, flags=ALL | ev.MIXED)
    errK = np.max(np.abs(K - g["K"]))
    s0, s1, s2 = _scales(tb, k)
    print(f"w1024x10000 mixed max|dK2|={errK:.3e} max|dK2'|/scale={np.max(np.abs(dK - g['dK'])) / s1:.3e}")
    assert errK <= 1e-9
    assert np.max(np.abs(dK - g["dK"])) / s1 <= 1e-6


def test_deterministic(w10k):
    tb, t, V, k = w10k
    a = t.eval(V, k, flags=ALL)
    b = t.eval(V, k, flags=ALL)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("nc,wg", [(1, 256), (2, 256), (4, 256), (8, 256), (1, 512), (2, 512)])
@pytest.mark.parametrize("target", [37, 4096])
def test_tilings_agree(w10k, nc, wg, target):
    tb, t, V, k = w10k
    g = golden("w1024x10000_oracle.npz")
    t.set_tiling(nc=nc, target_blocks=target, wg_threads=wg)
    try:
        K, dK, _ = t.eval(V, k, flags=ev.VALUE | ev.D1)
        Km, _, _ = t.eval(V, k, flags=ev.VALUE | ev.D1 | ev.MIXED)
    finally:
        t.set_tiling(nc=2, target_blocks=3072, wg_threads=256)
    assert np.max(np.abs(K - g["K"])) <= 1e-12
    assert np.max(np.abs(Km - g["K"])) <= 1e-9


@pytest.mark.parametrize("n_cand,n_zeros", [(1, 1), (1, 10_000), (255, 257), (257, 3333),
                                            (1000, 4096), (300, 9999)])
def test_ragged_sizes_vs_port(w10k, oracle_lib, n_cand, n_zeros):
    tb, t, V, k = w10k
    Vs = V[:n_cand]
    K, dK, d2K = t.eval(Vs, k, n_zeros=n_zeros, flags=ALL)
    Ko, dKo, d2Ko = oracle_lib.eval_f64(tb, Vs, k, n_zeros=n_zeros)
    s0, s1, s2 = _scales(tb, k, n_zeros)
    assert np.max(np.abs(K - Ko)) <= 1e-12
    assert np.max(np.abs(dK - dKo)) / s1 <= 1e-12
    assert np.max(np.abs(d2K - d2Ko)) / s2 <= 1e-12


def test_value_only_and_device_path(w10k):
    import torch
    tb, t, V, k = w10k
    K, _, _ = t.eval(V, k, flags=ev.VALUE)
    Kd, dKd, _ = t.eval_device(ev.v_tensor(V), k, flags=ev.VALUE | ev.D1)
    torch.cuda.synchronize()
    Kh, dKh, _ = t.eval(V, k, flags=ev.VALUE | ev.D1)
    assert np.array_equal(Kd.cpu().numpy(), Kh) and np.array_equal(dKd.cpu().numpy(), dKh)
    assert np.array_equal(K, Kh)


def test_argument_errors(w10k):
    tb, t, V, k = w10k
    with pytest.raises(ev.K2Error):
        t.eval(V[:4], k, n_zeros=tb.n + 1)
    with pytest.raises(ev.K2Error):
        t.eval(V[:4], -1.0)
    with pytest.raises(ev.K2Error):
        t.set_tiling(nc=3)
    with pytest.raises(ev.K2Error):
        t.set_tiling(nc=4, wg_threads=512)
    with pytest.raises(ev.K2Error):
        t.set_tiling(wg_threads=384)
    K, _, _ = t.eval(V[:0], k)
    assert K.shape == (0,)


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "w65536x2325572_sub256.npz")),
                    reason="65,536 x 2,325,572 fixture not generated")
def test_65536x2325572_subsample_vs_golden():
    """256 candidates of the full-size workload against all 2,325,572 zeros."""
    c = synth.W65536x2325572
    g = golden("w65536x2325572_sub256.npz")
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    assert tb.sha256() == str(g["table_sha256"])
    t = ev.K2Table(tb)
    K, dK, d2K = t.eval(g["V"], c["k"], flags=ALL)
    s0, s1, s2 = _scales(tb, c["k"])
    e = np.max(np.abs(K - g["K"]))
    print(f"w65536x2325572 sub256 plain max|dK2|={e:.3e}")
    assert e <= 1e-9
    assert np.max(np.abs(dK - g["dK"])) / s1 <= 1e-11
    Km, dKm, _ = t.eval(g["V"], c["k"], flags=ev.VALUE | ev.D1 | ev.MIXED)
    em = np.max(np.abs(Km - g["K"]))
    print(f"w65536x2325572 sub256 mixed max|dK2|={em:.3e}")
    assert em <= 1e-9


def _head_rule(tb, k, nz, tail=0.16):
    """The mixed form's FP64 head: shortest multiple of 4096 zeros after which the weight
    magnitudes 2 e^{-k g^2} |w| of the whole 4096-zero blocks up to nz sum to <= tail."""
    a = 2 * np.exp(-k * tb.gamma ** 2) * np.abs(tb.w)
    npad = (tb.n + 4095) // 4096 * 4096
    ap = np.zeros(npad)
    ap[:tb.n] = a
    blocks = ap.reshape(-1, 4096).sum(axis=1)
    nb = min(len(blocks), (nz + 4095) // 4096)
    s, hb = 0.0, nb
    while hb > 0 and s + blocks[hb - 1] <= tail:
        s += blocks[hb - 1]
        hb -= 1
    return min(nz, hb * 4096)


@pytest.fixture(scope="module")
def mid_table():
    tb = synth.make_zero_table(200_000, 5)
    return tb, ev.K2Table(tb)


@pytest.mark.parametrize("k", list(synth.BATCH_KS))
def test_adaptive_mixed_head(mid_table, k):
    tb, t = mid_table
    for nz in (tb.n, 50_000, 3000):
        h = t.mixed_head(k, nz)
        want = _head_rule(tb, k, nz)
        print(f"k={k} n_zeros={nz}: FP64 head {h} (rule {want})")
        assert abs(h - want) <= 4096 and h % 4096 in (0, nz % 4096)
    V = synth.make_candidates(512, 3.4e18, 3.6e18, 5)
    Kp, dKp, _ = t.eval(V, k, flags=ev.VALUE | ev.D1)
    Km, dKm, _ = t.eval(V, k, flags=ev.VALUE | ev.D1 | ev.MIXED)
    err = np.max(np.abs(Km - Kp))
    print(f"k={k}: mixed vs FP64 max|dK2|={err:.3e}")
    assert err <= 1e-9
    t.set_tiling(mixed_head=8192)
    try:
        assert t.mixed_head(k) == 8192
    finally:
        t.set_tiling(mixed_head=0)
    assert t.get_tiling()["mixed_tail"] == pytest.approx(0.16)


@pytest.mark.parametrize("n_cand", [1, 5, 8, 17, 64, 65, 200])
def test_z_tile_path_matches_c_tiles(w10k, n_cand):
    """Few candidates run as Z tiles (lanes over zeros); forced C tiles must agree, and both
    must match the oracle's golden values."""
    tb, t, V, k = w10k
    g = golden("w1024x10000_oracle.npz")
    Vs = V[:n_cand]
    t.set_tiling(z_max_cand=1000)
    try:
        Kz, dKz, d2Kz = t.eval(Vs, k, flags=ALL)
        Kzm, _, _ = t.eval(Vs, k, flags=ev.VALUE | ev.D1 | ev.MIXED)
        t.set_tiling(z_max_cand=-1)
        Kc, dKc, d2Kc = t.eval(Vs, k, flags=ALL)
    finally:
        t.set_tiling(z_max_cand=0)
    s0, s1, s2 = _scales(tb, k)
    assert np.max(np.abs(Kz - g["K"][:n_cand])) <= 1e-12
    assert np.max(np.abs(Kz - Kc)) <= 1e-13
    assert np.max(np.abs(dKz - dKc)) / s1 <= 1e-13
    assert np.max(np.abs(d2Kz - d2Kc)) / s2 <= 1e-13
    assert np.max(np.abs(Kzm - g["K"][:n_cand])) <= 1e-9


def test_mixed_one_z_group_runs_fp64(w10k):
    """With at most 8 candidates (one Z group) the mixed form runs in FP64 throughout: its
    results equal the plain form's bit for bit; 9 candidates take the FP32 tail again."""
    tb, t, V, k = w10k
    for n in (1, 8):
        Kp, dKp, _ = t.eval(V[:n], k, flags=ev.VALUE | ev.D1)
        Km, dKm, _ = t.eval(V[:n], k, flags=ev.VALUE | ev.D1 | ev.MIXED)
        assert np.array_equal(Kp, Km) and np.array_equal(dKp, dKm)
    Kp, _, _ = t.eval(V[:9], k, flags=ev.VALUE | ev.D1)
    Km, _, _ = t.eval(V[:9], k, flags=ev.VALUE | ev.D1 | ev.MIXED)
    assert not np.array_equal(Kp, Km) and np.max(np.abs(Kp - Km)) <= 1e-9


def test_far_candidates_rotated_table():
    """Candidates above y = 2^64 evaluate exactly through formats.rotate_table: the GPU at
    v' = v - 2^64 q on the rotated table against exact big-integer phases with mpmath."""
    import mpmath
    from mertens_bound_amd import formats
    tb = synth.make_zero_table(120, 22)
    vals = [(int(1.8788774099540933515e19) << 64) + 7, int(1.957e19) << 64,
            (int(3.5e18) << 64) + 99]
    k = 3e-9
    want = []
    with mpmath.workdps(40):
        for V in vals:
            s = mpmath.mpf(0)
            for (lo, hi), wj in zip(tb.G, tb.w):
                Gi = (int(hi) << 64) | int(lo)
                th = 2 * mpmath.pi * mpmath.mpf((Gi * V) % (1 << 174)) / mpmath.mpf(2) ** 174
                gg = 2 * mpmath.pi * mpmath.mpf(Gi) / mpmath.mpf(2) ** 110
                s += 2 * mpmath.exp(-k * gg * gg) * (mpmath.mpf(wj.real) * mpmath.cos(th)
                                                     - mpmath.mpf(wj.imag) * mpmath.sin(th))
            want.append(float(s))
    got = np.zeros(len(vals))
    for q, (ix, words) in formats.split_far_candidates(vals).items():
        got[ix] = ev.K2Table(formats.rotate_table(tb, q)).eval(words, k, flags=ev.VALUE)[0]
    assert np.max(np.abs(got - np.array(want))) < 1e-13


def test_handles_from_threads_and_caller_stream(w10k):
    """Distinct table handles used from two threads at once give the sequential results, and
    the device entry points run on the caller's stream (a side stream here)."""
    import threading

    import torch
    tb, t, V, k = w10k
    want = t.eval(V, k, flags=ALL)
    tables = [ev.K2Table(tb), ev.K2Table(tb)]
    got = [None, None]

    def work(i):
        for _ in range(3):
            got[i] = tables[i].eval(V, k, flags=ALL)

    th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    for g in got:
        for a, b in zip(g, want):
            assert np.array_equal(a, b)
    side = torch.cuda.Stream()
    Vd = ev.v_tensor(V)
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        K, dK, d2K = t.eval_device(Vd, k, flags=ALL, stream=side.cuda_stream)
    side.synchronize()
    assert np.array_equal(K.cpu().numpy(), want[0])
