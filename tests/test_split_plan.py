"""The zero split plans of the C-tile launches (host arithmetic, no GPU): every plan covers the
zero range exactly once in contiguous splits, guided plans put a multiple of 8 long splits first
and leave enough short splits to end the launch evenly, and the block dealing of the kernels
(block_split: XCD x of the 8 takes every 8th block and walks splits x, x + 8, ...) reaches every
(split, candidate tile) pair once."""
import numpy as np
import pytest

from mertens_bound_amd import evaluator as ev

SLOTS = 1024   # resident C-tile workgroups assumed by the plan (256 CUs x 4)


def _ranges(p, nz):
    out = []
    for s in range(p["n"]):
        if s < p["n_long"]:
            z0 = s * p["zps_long"]
            z1 = z0 + p["zps_long"]
        else:
            z0 = p["n_long"] * p["zps_long"] + (s - p["n_long"]) * p["zps"]
            z1 = min(nz, z0 + p["zps"])
        out.append((z0, z1))
    return out


def _check_plan(nz, n_ctiles, target, factor):
    p = ev.split_plan(nz, n_ctiles, target, factor)
    if nz == 0:
        assert p["n"] == 0
        return p
    r = _ranges(p, nz)
    assert r[0][0] == 0 and r[-1][1] == nz
    assert all(a[1] == b[0] for a, b in zip(r, r[1:]))
    assert all(z1 > z0 for z0, z1 in r)
    if p["zps"] >= 256:
        assert p["zps"] % 256 == 0
    if p["n_long"]:
        assert factor > 1 and p["n_long"] % 8 == 0
        assert p["zps_long"] == factor * p["zps"]
        n_short = p["n"] - p["n_long"]
        # the short splits fill every slot for the last factor rounds
        assert n_short * n_ctiles >= factor * SLOTS - n_ctiles
        assert p["blocks"] == n_ctiles * ((p["n"] + 7) // 8 * 8)
    else:
        assert p["blocks"] == n_ctiles * p["n"]
    return p


def test_metric_plan():
    # 65,536 candidates at 4 per lane (64 tiles) over the full table at 49,152 target workgroups
    p = _check_plan(2_325_572, 64, 49_152, 8)
    assert p == {"n": 254, "zps": 3072, "n_long": 72, "zps_long": 24576, "blocks": 16384}
    u = _check_plan(2_325_572, 64, 49_152, 1)
    assert u["n_long"] == 0 and u["n"] == 758


@pytest.mark.parametrize("seed", range(4))
def test_random_plans(seed):
    rng = np.random.default_rng(seed)
    for _ in range(2000):
        nz = int(rng.choice([rng.integers(0, 3000), rng.integers(1, 300_000),
                             rng.integers(1, 3_000_000)]))
        n_ctiles = int(rng.integers(1, 300))
        target = int(rng.choice([3072, 12_288, 49_152, 98_304, rng.integers(1, 200_000)]))
        _check_plan(nz, n_ctiles, target, int(rng.choice([1, 2, 4, 8, 16])))


def test_block_dealing_reaches_every_pair_once():
    for nz, n_ctiles, target in [(2_325_572, 64, 49_152), (2_260_036, 64, 98_304),
                                 (1_000_003, 7, 24_576), (3_000_001, 32, 98_304)]:
        p = _check_plan(nz, n_ctiles, target, 8)
        assert p["n_long"] > 0
        b = np.arange(p["blocks"])
        k = b >> 3
        j = k // n_ctiles
        ct = k - j * n_ctiles
        split = 8 * j + (b & 7)
        real = split < p["n"]
        pairs = split[real] * n_ctiles + ct[real]
        assert np.array_equal(np.sort(pairs), np.arange(p["n"] * n_ctiles))
        # padding blocks start at or past the end of the zeros (they return at once)
        starts = p["n_long"] * p["zps_long"] + (split[~real] - p["n_long"]) * p["zps"]
        assert np.all(starts >= nz)
        # every XCD gets the same number of long splits
        xcd_long = np.bincount((split[real & (split < p["n_long"])]) % 8, minlength=8)
        assert len(set(xcd_long.tolist())) == 1


def test_rejects_bad_arguments():
    with pytest.raises(RuntimeError):
        ev.split_plan(100, 0, 10, 1)
