"""Host-side pins of the enclosure mode's error budget (no GPU): the bounded e^{-x} that the
enclosure kernels use (same operations on host and device) stays within its stated
2.5 * 2^-53 relative error, checked against long double and mpmath."""
import numpy as np

from mertens_bound_amd import evaluator as ev

U = 2.0 ** -53


def _rel_err(x, y):
    ref = np.exp(-x.astype(np.longdouble))
    return np.abs((y.astype(np.longdouble) - ref) / ref) / U


def test_exp_neg_bound_long_double():
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(0.0, 1.0, 300_000), rng.uniform(0.0, 708.0, 300_000),
                        np.linspace(0.0, 708.0, 100_001), rng.uniform(0.0, 1e-6, 50_000),
                        np.log(2.0) * (np.arange(1, 1021) + 0.5)])   # reduction boundaries <= 708
    y = ev.exp_neg_host(x)
    rel = _rel_err(x, y)
    print(f"exp_neg_bounded max rel err {float(rel.max()):.3f} u")
    assert float(rel.max()) <= 2.5


def test_exp_neg_bound_mpmath_worst():
    import mpmath
    rng = np.random.default_rng(8)
    x = rng.uniform(0.0, 708.0, 200_000)
    y = ev.exp_neg_host(x)
    worst = np.argsort(-_rel_err(x, y))[:50]
    with mpmath.workprec(200):
        for i in worst:
            r = mpmath.exp(-mpmath.mpf(float(x[i])))
            assert abs((mpmath.mpf(float(y[i])) - r) / r) <= 2.5 * mpmath.mpf(2) ** -53


def test_exp_neg_edges():
    y = ev.exp_neg_host(np.array([0.0, 1e-300, 708.0, 708.0000001, 1e6, np.inf, np.nan]))
    assert y[0] == 1.0 and y[1] == 1.0
    assert 0.0 < y[2] < 1e-300 and y[2] >= np.finfo(np.float64).tiny
    assert y[3] == 0.0 and y[4] == 0.0 and y[5] == 0.0 and y[6] == 0.0
