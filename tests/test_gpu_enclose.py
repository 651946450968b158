"""Enclosure mode (k2_enclose_device): the binary128 oracle's K2 and K2' must lie inside the
returned balls K +- K_rad for every candidate, the radii must be small, a closed form must be
enclosed, the tail bound must cover the zeros left out, and the sin/cos the kernels use must
meet the per-evaluation bound (3 * 2^-53) the radius is built on.

A golden value rounded to FP64 is itself off by up to 2^-53 |K|; that is added where fixtures
are compared."""
import numpy as np
import pytest

from conftest import golden
from mertens_bound_amd import evaluator as ev
from mertens_bound_amd import synth

pytestmark = pytest.mark.gpu
U = 2.0 ** -53


@pytest.fixture(scope="module")
def w10k():
    c = synth.W1024x10000
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    V = synth.make_candidates(c["n_cand"], c["v_lo"], c["v_hi"], c["seed"])
    return tb, ev.K2Table(tb), V, c["k"]


def test_w1024x10000_encloses_binary128(w10k, oracle_lib):
    tb, t, V, k = w10k
    K, Kr, dK, dKr, tail = t.enclose(V, k)
    hi, lo = oracle_lib.eval_quad_pair(tb, V, k)          # K2 in binary128, as hi + lo
    _, dKq, _, _ = oracle_lib.eval_quad(tb, V, k, flags=oracle_lib.VALUE | oracle_lib.D1)
    dev = np.abs((K - hi) - lo)
    print(f"w1024x10000 enclosure: max|K - K_q| {dev.max():.3e}, max rad {Kr.max():.3e}, "
          f"max|dK - dK_q| {np.abs(dK - dKq).max():.3e}, max dK rad {dKr.max():.3e}")
    assert np.all(dev <= Kr)
    assert np.all(np.abs(dK - dKq) <= dKr + U * np.abs(dKq))
    assert Kr.max() < 1e-13 and dKr.max() < 1e-9
    assert tail[0] == pytest.approx(1e-300, rel=1e-6) and tail[1] == pytest.approx(1e-300, rel=1e-6)
    # the midpoint is the plain evaluator's value to within the radius
    Kp, _, _ = t.eval(V, k, flags=ev.VALUE)
    assert np.all(np.abs(Kp - K) <= Kr + 1e-14)


def test_w1024x10000_d1_off_and_repeatable(w10k):
    tb, t, V, k = w10k
    a = t.enclose(V, k, d1=False)
    b = t.enclose(V, k, d1=True)
    assert a[2] is None and a[3] is None
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])   # bit-identical reruns


def test_tail_bound_covers_rest_of_table(w10k):
    tb, t, V, k = w10k
    g = golden("w1024x10000_oracle.npz")
    assert np.array_equal(g["V"], V)
    N = 5_000
    K, Kr, dK, dKr, tail = t.enclose(V, k, n_zeros=N)
    a = 2 * np.exp(-k * tb.gamma[N:] ** 2) * np.abs(tb.w[N:])
    assert tail[0] >= a.sum() * (1 - 1e-12) and tail[0] <= a.sum() * 1.02
    assert tail[1] >= (tb.gamma[N:] * a).sum() * (1 - 1e-12)
    # full-table K2 (oracle) lies within the N-zero ball widened by the tail bound
    assert np.all(np.abs(g["K"] - K) <= Kr + tail[0] + U * np.abs(g["K"]))
    assert np.all(np.abs(g["dK"] - dK) <= dKr + tail[1] + U * np.abs(g["dK"]))


def test_closed_form_quarter_turn():
    """gamma = 2 pi (G = 2^110) and v = n + 1/4: the phase is a quarter turn, so
    K2 = -2 e^{-k (2 pi)^2} Im w and K2' = -2 pi 2 e^{-k (2 pi)^2} Re w exactly."""
    import mpmath
    w = np.array([0.3 - 0.7j, 0.0 + 0.0j])
    G = np.array([[0, 1 << 46], [0, 1 << 46]], dtype=np.uint64)
    tb = synth.ZeroTable(G, np.full(2, 2 * np.pi), w)
    t = ev.K2Table(tb)
    V = np.array([[1 << 62, n] for n in (0, 1, 12345, 3_500_000_000_000_000_000)],
                 dtype=np.uint64)
    k = 1.5e-3
    K, Kr, dK, dKr, _ = t.enclose(V, k, n_zeros=1)
    with mpmath.workdps(40):
        f = 2 * mpmath.exp(-mpmath.mpf(k) * (2 * mpmath.pi) ** 2)
        Kx = -f * mpmath.mpf(w[0].imag)
        dKx = -2 * mpmath.pi * f * mpmath.mpf(w[0].real)
        for i in range(V.shape[0]):
            assert abs(mpmath.mpf(float(K[i])) - Kx) <= Kr[i]
            assert abs(mpmath.mpf(float(dK[i])) - dKx) <= dKr[i]
    assert Kr.max() < 1e-14


@pytest.mark.skipif(not __import__("os").path.exists(
    __import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                               "w65536x2325572_sub256.npz")), reason="65,536 x 2,325,572 fixture not generated")
def test_65536x2325572_subsample_enclosed():
    """256 candidates of the full-size workload against all 2,325,572 zeros."""
    c = synth.W65536x2325572
    g = golden("w65536x2325572_sub256.npz")
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    assert tb.sha256() == str(g["table_sha256"])
    t = ev.K2Table(tb)
    K, Kr, dK, dKr, tail = t.enclose(g["V"], c["k"])
    dev = np.abs(K - g["K"])
    print(f"w65536x2325572 sub256 enclosure: max|K - K_q| {dev.max():.3e}, max rad {Kr.max():.3e}, "
          f"dK rad {dKr.max():.3e}")
    assert np.all(dev <= Kr + U * np.abs(g["K"]))
    assert np.all(np.abs(dK - g["dK"]) <= dKr + U * np.abs(g["dK"]))
    assert Kr.max() < 1e-12


def test_sincos_meets_enclosure_bound():
    """Per-evaluation sin/cos error of the FP64 table path against long double, over random
    words and the worst remainders: the two edges of each of the 1,024 table intervals, where
    the remainder from the interval's midpoint is largest."""
    rng = np.random.default_rng(3)
    W = rng.integers(0, 2**64, size=1 << 20, dtype=np.uint64)
    k = rng.integers(0, 1024, size=4096).astype(np.uint64)
    step = np.uint64(1 << 54)
    W[:4096] = k << np.uint64(54)                                # largest negative remainder
    W[4096:8192] = (k << np.uint64(54)) + step - np.uint64(1)    # largest positive remainder
    W[8192:8200] = np.array([0, 1, 1 << 62, 1 << 63, 3 << 62, (1 << 64) - 1, 1 << 55,
                             (1 << 55) - 1], dtype=np.uint64)
    s, c = ev.sincos_words(W, mode=0)
    two_pi = np.longdouble("6.283185307179586476925286766559005768")
    th = two_pi * (W.astype(np.longdouble) / np.longdouble(2.0) ** 64)
    es = np.abs(s.astype(np.longdouble) - np.sin(th)) / U
    ec = np.abs(c.astype(np.longdouble) - np.cos(th)) / U
    print(f"sincos_tab64 max error: sin {float(es.max()):.3f} u, cos {float(ec.max()):.3f} u")
    assert float(es.max()) <= 3.0 and float(ec.max()) <= 3.0


def test_enclose_argument_errors(w10k):
    import torch
    tb, t, V, k = w10k
    Vt = ev.v_tensor(V)
    out = [torch.empty(V.shape[0], dtype=torch.float64, device="cuda") for _ in range(4)]
    L = ev.load_library()
    st = torch.cuda.current_stream().cuda_stream
    rc = L.k2_enclose_device(t._h, V.shape[0], Vt.data_ptr(), k, 0, ev.VALUE | ev.D2,
                             *[o.data_ptr() for o in out], None, st)
    assert rc != 0 and b"K2 and K2'" in L.k2_last_error()
    rc = L.k2_enclose_device(t._h, V.shape[0], Vt.data_ptr(), k, tb.n + 1, ev.VALUE,
                             *[o.data_ptr() for o in out], None, st)
    assert rc != 0
    rc = L.k2_enclose_device(t._h, V.shape[0], Vt.data_ptr(), -1.0, 0, ev.VALUE,
                             *[o.data_ptr() for o in out], None, st)
    assert rc != 0
