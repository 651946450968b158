"""Seeded synthetic zero tables and candidate lists in the canonical fixed-point encoding.

Canonical encoding (shared by the HIP kernels, the C-ABI and the CPU oracle):

* zero j: ``G_j`` is an unsigned 128-bit word holding gamma_j / 2pi in Q18.110 fixed point,
  stored as two little-endian uint64 words ``(lo, hi)``.  gamma_j is *defined* as
  ``2 pi G_j / 2^110`` so the phase gamma_j * v / 2pi is an exact integer product.
  This plays the role of the reference's ``glo``/``ghi`` words
  (search/path1_build_zeros_100k.py:56-58,102-104), collapsed to one exactly-defined word.
* ``gamma`` (float64) is the nearest double to that exact value, used for the Gaussian weight
  and the derivative factors.
* ``w`` (complex128) is w_j = 1/(rho_j zeta'(rho_j)); the reference stores |w| and the phase
  word ``psi`` separately (search/path1_build_zeros_100k.py:93-104); here it is kept complex.
* candidate c: ``V_c`` is an unsigned 128-bit 64.64 fixed-point word ``(lo=frac, hi=int)``,
  v_c = hi + lo / 2^64.  The reference's (x, dy0) pair (search/eval_tech1_600k.py:216-218)
  maps to it through :func:`candidates_from_x_dy`.

Zeros follow the Riemann-von Mangoldt count (gamma/2pi) log(gamma/(2pi e)) + 7/8 = j - 1/2,
jittered by U(-0.1, 0.1) mean spacings; |zeta'| is log-normal with median log(gamma/2pi) and
sigma 0.8, and arg zeta' is uniform.
"""
from __future__ import annotations

import dataclasses
import hashlib

import numpy as np

TWO_PI = 2.0 * np.pi
M64 = (1 << 64) - 1

# stream tags so the zero table and the candidate list of one seed are independent draws
_TAG_ZEROS = 0x5A45524F
_TAG_CANDS = 0x43414E44


@dataclasses.dataclass
class ZeroTable:
    G: np.ndarray        # (Z, 2) uint64, [lo, hi] of gamma/2pi in Q18.110
    gamma: np.ndarray    # (Z,) float64
    w: np.ndarray        # (Z,) complex128, 1/(rho zeta'(rho))

    @property
    def n(self) -> int:
        return int(self.G.shape[0])

    def sha256(self) -> str:
        h = hashlib.sha256()
        for a in (self.G, self.gamma, self.w):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()

    def head(self, n: int) -> "ZeroTable":
        return ZeroTable(self.G[:n], self.gamma[:n], self.w[:n])

    def count_below(self, T: float) -> int:
        """Number of zeros with gamma < T (the table is sorted by gamma)."""
        return int(np.searchsorted(self.gamma, T, side="left"))


def rvm_gammas(n: int) -> np.ndarray:
    """gamma_j for j = 1..n solving (t) log(t/e) + 7/8 = j - 1/2 with t = gamma/2pi (t > 1)."""
    j = np.arange(1, n + 1, dtype=np.float64)
    rhs = j - 0.5 - 0.875
    # start to the right of the root on the convex increasing branch t > 1; Newton then
    # decreases monotonically onto it
    t = 2.0 * (j + 1.0) / np.log(j + 2.0) + np.e
    for _ in range(60):
        f = t * (np.log(t) - 1.0) - rhs
        step = f / np.log(t)
        t = t - step
        if np.max(np.abs(step) / t) < 1e-15:
            break
    return TWO_PI * t


def _g_to_words(g: np.ndarray, rng: np.random.Generator):
    """floor(g * 2^110) with the bits below g's double mantissa filled uniformly at random."""
    m, e = np.frexp(g)                                    # g = m 2^e, m in [0.5, 1)
    M = np.ldexp(m, 53).astype(np.uint64)                 # 53-bit integer mantissa
    s = (e + 57).astype(np.int64)                         # G0 = M << s, s in [59, 75]
    if np.any(s < 1) or np.any(s > 75):
        raise ValueError("gamma/2pi outside the Q18.110 range")
    r_lo = rng.integers(0, 2**64, size=g.shape, dtype=np.uint64)
    r_hi = rng.integers(0, 2**64, size=g.shape, dtype=np.uint64)
    big = s >= 64
    sh_small = np.where(big, 0, s).astype(np.uint64)            # s < 64 branch
    sh_big = np.where(big, s - 64, 0).astype(np.uint64)         # s >= 64 branch
    mask_small = np.where(big, np.uint64(M64), (np.uint64(1) << sh_small) - np.uint64(1))
    lo_small = (M << sh_small) | (r_lo & mask_small)
    hi_small = M >> (np.uint64(64) - np.where(big, 1, sh_small).astype(np.uint64))
    mask_big = (np.uint64(1) << sh_big) - np.uint64(1)
    lo_big = r_lo
    hi_big = (M << sh_big) | (r_hi & mask_big)
    lo = np.where(big, lo_big, lo_small)
    hi = np.where(big, hi_big, hi_small)
    return np.stack([lo, hi], axis=1).astype(np.uint64)


def gamma_from_words(G: np.ndarray) -> np.ndarray:
    """float64 gamma = 2 pi G / 2^110 (hi word carries the leading bits)."""
    g = G[:, 1].astype(np.float64) * 2.0 ** -46 + G[:, 0].astype(np.float64) * 2.0 ** -110
    return g * TWO_PI


def make_zero_table(n: int, seed: int = 0) -> ZeroTable:
    rng = np.random.default_rng([seed, _TAG_ZEROS])
    gam = rvm_gammas(n)
    spacing = TWO_PI / np.log(gam / TWO_PI)
    gam = gam + rng.uniform(-0.1, 0.1, size=n) * spacing
    G = _g_to_words(gam / TWO_PI, rng)
    gamma = gamma_from_words(G)
    if not np.all(np.diff(gamma) > 0):
        raise AssertionError("synthetic gammas are not strictly increasing")
    log_med = np.log(np.log(gamma / TWO_PI))
    zmag = np.exp(rng.normal(log_med, 0.8))
    zarg = rng.uniform(0.0, TWO_PI, size=n)
    zp = zmag * np.exp(1j * zarg)
    rho = 0.5 + 1j * gamma
    w = 1.0 / (rho * zp)
    return ZeroTable(G=G, gamma=gamma, w=w.astype(np.complex128))


def make_candidates(n: int, v_lo: float, v_hi: float, seed: int = 0) -> np.ndarray:
    """(n, 2) uint64 64.64 words [frac, int], v uniform in [v_lo, v_hi)."""
    rng = np.random.default_rng([seed, _TAG_CANDS])
    vi = rng.integers(int(v_lo), int(v_hi), size=n, dtype=np.uint64)
    vf = rng.integers(0, 2**64, size=n, dtype=np.uint64)
    return np.stack([vf, vi], axis=1).astype(np.uint64)


def v_to_float(V: np.ndarray) -> np.ndarray:
    return V[:, 1].astype(np.float64) + V[:, 0].astype(np.float64) * 2.0 ** -64


def v_to_int(V: np.ndarray) -> list:
    """Exact 64.64 words as Python ints (value * 2^64)."""
    return [(int(h) << 64) | int(l) for l, h in V]


def v_from_int(vals) -> np.ndarray:
    arr = np.empty((len(vals), 2), dtype=np.uint64)
    for i, x in enumerate(vals):
        x &= (1 << 128) - 1
        arr[i, 0] = x & M64
        arr[i, 1] = x >> 64
    return arr


def candidates_from_x_dy(x, dy0) -> np.ndarray:
    """Reference candidate (x in Q98.30, dy0 float) -> 64.64 words, V = x 2^34 + rint(dy0 2^64).

    The reference evaluates at y = x / 2^30 + dy0 (search/eval_tech1_600k.py:216-218,
    legacy/superseded/certify_peak2_600k.py:175-177).  A 64.64 word holds 0 <= y < 2^64
    (1.8447e19); candidates outside raise ValueError rather than wrap (formats.split_far_candidates
    and formats.rotate_table evaluate larger y).
    """
    out = []
    for xi, d in zip(x, dy0):
        v = (int(xi) << 34) + int(round(float(d) * 2.0 ** 64))
        if not 0 <= v < (1 << 128):
            raise ValueError(f"candidate y = {int(xi) / 2.0 ** 30 + float(d):.6e} is outside the "
                             "64.64 range [0, 2^64)")
        out.append(v)
    return v_from_int(out)


# the five workloads of this evaluator -----------------------------------------------------

W1024x10000 = dict(n_zeros=10_000, n_cand=1_024, v_lo=1e18, v_hi=4e18, k=1.5e-10, seed=0)
W65536x2325572 = dict(n_zeros=2_325_572, n_cand=65_536, v_lo=3.4e18, v_hi=3.6e18, k=1.5e-10, seed=0)
BATCH_KS = (3e-9, 1e-9, 1.5e-10)


def make_batched_requests(table: ZeroTable, n_req: int = 512, seed: int = 0,
                          t_lo: float = 1e5, t_hi: float = 1.3e6, c_max: int = 4096,
                          v_lo: float = 3.4e18, v_hi: float = 3.6e18):
    """Uneven requests: cutoff T log-uniform, k from BATCH_KS, candidate count log-uniform in [1, c_max].

    Returns (n_zeros per request, k per request, counts, offsets, V for all requests concatenated).
    """
    rng = np.random.default_rng([seed, 0x42415443])
    T = np.exp(rng.uniform(np.log(t_lo), np.log(t_hi), size=n_req))
    ks = np.asarray(BATCH_KS)[rng.integers(0, len(BATCH_KS), size=n_req)]
    cnt = np.floor(np.exp(rng.uniform(0.0, np.log(c_max + 1.0), size=n_req))).astype(np.int64)
    cnt = np.clip(cnt, 1, c_max)
    nz = np.searchsorted(table.gamma, T, side="left").astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)
    V = make_candidates(int(cnt.sum()), v_lo, v_hi, seed=seed + 7)
    return nz, ks, cnt, offs, V
