"""The seeded generator: determinism, the Riemann-von Mangoldt count, encodings."""
import numpy as np

from conftest import golden
from mertens_bound_amd import synth


def test_deterministic_and_pinned():
    a = synth.make_zero_table(10_000, 0)
    b = synth.make_zero_table(10_000, 0)
    assert a.sha256() == b.sha256()
    assert a.sha256() == str(golden("w1024x10000_oracle.npz")["table_sha256"])
    V1 = synth.make_candidates(1024, 1e18, 4e18, 0)
    assert np.array_equal(V1, golden("w1024x10000_oracle.npz")["V"])


def test_rvm_count_and_jitter():
    n = 20_000
    tb = synth.make_zero_table(n, 2)
    t = tb.gamma / (2 * np.pi)
    N = t * np.log(t / np.e) + 7 / 8
    j = np.arange(1, n + 1)
    # jitter is at most 0.1 mean spacings, i.e. N moves by at most ~0.1
    assert np.max(np.abs(N - (j - 0.5))) < 0.11
    assert np.all(np.diff(tb.gamma) > 0)


def test_words_define_gamma():
    tb = synth.make_zero_table(1000, 4)
    exact = np.array([((int(h) << 64) | int(l)) / 2.0 ** 110 * 2 * np.pi for l, h in tb.G])
    np.testing.assert_allclose(tb.gamma, exact, rtol=4e-16)
    # the words carry bits below the double mantissa (full 128-bit entropy)
    assert np.any(tb.G[:, 0] & np.uint64(0xFFFF))


def test_weights_distribution():
    tb = synth.make_zero_table(50_000, 6)
    rho = 0.5 + 1j * tb.gamma
    zp = 1.0 / (tb.w * rho)
    z = np.log(np.abs(zp)) - np.log(np.log(tb.gamma / (2 * np.pi)))
    assert abs(np.mean(z)) < 0.02 and abs(np.std(z) - 0.8) < 0.02
    ang = np.angle(zp)
    assert abs(np.mean(np.cos(ang))) < 0.02


def test_full_size_count_below_cutoff():
    g = synth.rvm_gammas(2_325_573)
    assert g[2_325_571] < 1.3e6 < g[2_325_572]


def test_candidates_range_and_batches():
    V = synth.make_candidates(5000, 3.4e18, 3.6e18, 1)
    v = synth.v_to_float(V)
    assert v.min() >= 3.4e18 and v.max() < 3.6e18
    tb = synth.make_zero_table(200_000, 0)
    nz, ks, cnt, offs, Vb = synth.make_batched_requests(tb, n_req=64, seed=3)
    assert len(nz) == 64 and cnt.min() >= 1 and cnt.max() <= 4096
    assert set(np.unique(ks)) <= set(synth.BATCH_KS)
    assert Vb.shape[0] == cnt.sum() and offs[-1] + cnt[-1] == Vb.shape[0]
