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


def test_weight_lines_within_stated_bounds():
    """The enclosure weights kernel's arithmetic replayed on the host (IEEE FP64, the same
    operations: g from the two words of G, x = (k g) g, e^{-x} by the library's own host copy,
    a = 2 e^{-x} w per component, b = g a) against mpmath at 60 digits, line by line, for
    gamma over the table's range and x over [0, 708]: |g - gamma| <= 3.5 u gamma,
    |x - X| <= 9 u X, |a - A| <= (3.5 + 9 X) u |A|, |b - B| <= (8 + 9 X) u |B| (the per-term
    budget of k2_enclose.hip adds the sin/cos's 3 u and the group sums' roundings)."""
    import mpmath
    rng = np.random.default_rng(11)
    n = 3000
    gam = np.exp(rng.uniform(np.log(14.0), np.log(1.3e6), n))
    G = [int(gm / (2 * np.pi) * 2.0 ** 110) + int(rng.integers(0, 1 << 62)) for gm in gam]
    xt = np.concatenate([rng.uniform(0.0, 708.0, n // 2),
                         np.exp(rng.uniform(np.log(1e-12), np.log(708.0), n - n // 2))])
    w = rng.normal(size=n) * 0.05 + 1j * rng.normal(size=n) * 0.05
    ghi = np.array([g >> 64 for g in G], dtype=np.uint64)
    glo = np.array([g & (2 ** 64 - 1) for g in G], dtype=np.uint64)
    s = ghi.astype(np.float64) + glo.astype(np.float64) * 5.421010862427522e-20
    g = s * 8.928943354902097e-14
    k = xt / (g * g)                        # a double k per sample
    x = (k * g) * g
    E = ev.exp_neg_host(x)
    f = 2.0 * E
    a = f * w.real + 1j * (f * w.imag)
    b = g * a.real + 1j * (g * a.imag)
    worst = np.zeros(4)
    with mpmath.workdps(60):
        twopi = 2 * mpmath.pi
        for i in range(n):
            gm = twopi * mpmath.mpf(G[i]) / mpmath.mpf(2) ** 110
            X = mpmath.mpf(float(k[i])) * gm * gm
            e = 2 * mpmath.exp(-X)
            Ar, Ai = e * mpmath.mpf(float(w[i].real)), e * mpmath.mpf(float(w[i].imag))
            Am = mpmath.sqrt(Ar * Ar + Ai * Ai)
            Xf = float(X)
            r = [abs(mpmath.mpf(float(g[i])) - gm) / gm / U,
                 abs(mpmath.mpf(float(x[i])) - X) / X / U if X > 0 else 0,
                 mpmath.sqrt((mpmath.mpf(float(a[i].real)) - Ar) ** 2 +
                             (mpmath.mpf(float(a[i].imag)) - Ai) ** 2) / Am / U / (3.5 + 9 * Xf),
                 mpmath.sqrt((mpmath.mpf(float(b[i].real)) - gm * Ar) ** 2 +
                             (mpmath.mpf(float(b[i].imag)) - gm * Ai) ** 2) / (gm * Am) / U
                 / (8 + 9 * Xf)]
            worst = np.maximum(worst, [float(v) for v in r])
    print(f"worst: g {worst[0]:.2f} u (bound 3.5), x {worst[1]:.2f} u (bound 9), a {worst[2]:.3f} "
          f"of its bound, b {worst[3]:.3f} of its bound")
    assert worst[0] <= 3.5 and worst[1] <= 9.0 and worst[2] <= 1.0 and worst[3] <= 1.0
