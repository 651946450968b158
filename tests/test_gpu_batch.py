"""Batched uneven requests in one persistent launch (and its HIP graph replay) against
per-request one-launch evaluations and the oracle."""
import numpy as np
import pytest

from mertens_bound_amd import evaluator as ev
from mertens_bound_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import torch
    tb = synth.make_zero_table(30_000, 4)
    t = ev.K2Table(tb)
    rng = np.random.default_rng(9)
    n_req = 40
    nz = np.exp(rng.uniform(np.log(500), np.log(30_000), n_req)).astype(int)
    ks = np.array(synth.BATCH_KS)[rng.integers(0, 3, n_req)]
    cnt = np.clip(np.exp(rng.uniform(0, np.log(3000), n_req)).astype(int), 1, 3000)
    cnt[:3] = [1, 1024, 2048]                       # single candidate, exact C tiles
    offs = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    V = synth.make_candidates(int(cnt.sum()), 3.4e18, 3.6e18, 9)
    reqs = [(int(a), int(b), int(c), float(d)) for a, b, c, d in zip(nz, cnt, offs, ks)]
    return tb, t, reqs, V, torch


def _per_request(t, reqs, V, flags):
    K = np.zeros(V.shape[0]); dK = np.zeros(V.shape[0])
    for nz, n, off, k in reqs:
        a, b, _ = t.eval(V[off:off + n], k, n_zeros=nz, flags=flags)
        K[off:off + n] = a; dK[off:off + n] = b
    return K, dK


@pytest.mark.parametrize("tiling", [None, {"nc": 2, "cz": 8, "chunk": 4096},
                                    {"nc": 4, "cz": 16, "chunk": 2048, "c_min_fill": 64}])
def test_batch_matches_per_request(setup, tiling):
    tb, t, reqs, V, torch = setup
    Vt = ev.v_tensor(V)
    b = ev.K2Batch(t, reqs, Vt, flags=ev.VALUE | ev.D1, tiling=tiling)
    K, dK, _ = b.run_direct()
    torch.cuda.synchronize()
    K = K.cpu().numpy(); dK = dK.cpu().numpy()
    K0, dK0 = _per_request(t, reqs, V, ev.VALUE | ev.D1)
    info = b.info()
    print(f"batch items={info['n_items']} grid={info['grid']} "
          f"fill={info['useful_evals'] / info['padded_evals']:.3f}")
    assert np.max(np.abs(K - K0)) <= 1e-12
    assert np.max(np.abs(dK - dK0)) <= 1e-12 * np.max(np.abs(dK0))
    b.close()


def test_batch_vs_oracle_sample(setup, oracle_lib):
    tb, t, reqs, V, torch = setup
    Vt = ev.v_tensor(V)
    K, dK, _ = t.eval_batched_device(reqs, Vt, flags=ev.VALUE | ev.D1)
    K = K.cpu().numpy()
    for nz, n, off, k in reqs[:6]:
        m = min(n, 8)
        Ko, _, _ = oracle_lib.eval_f64(tb, V[off:off + m], k, n_zeros=nz)
        assert np.max(np.abs(K[off:off + m] - Ko)) <= 1e-12


def test_graph_replays_match_direct(setup):
    tb, t, reqs, V, torch = setup
    Vt = ev.v_tensor(V)
    b = ev.K2Batch(t, reqs, Vt, flags=ev.VALUE | ev.D1 | ev.D2)
    ref = [x.clone() for x in b.run_direct()]
    torch.cuda.synchronize()
    for rep in range(3):
        for x in (b.K, b.dK, b.d2K):
            x.fill_(float("nan"))
        # eager work between replays
        junk = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        out = b.run()
        torch.cuda.synchronize()
        for x, y in zip(out, ref):
            assert torch.equal(x, y), f"replay {rep} differs"
        assert torch.isfinite(junk)
    b.close()


def test_batch_mixed_accuracy(setup):
    tb, t, reqs, V, torch = setup
    Vt = ev.v_tensor(V)
    Kp, _, _ = t.eval_batched_device(reqs, Vt, flags=ev.VALUE | ev.D1)
    Km, _, _ = t.eval_batched_device(reqs, Vt, flags=ev.VALUE | ev.D1 | ev.MIXED)
    d = (Km - Kp).abs().max().item()
    print(f"batch mixed vs plain max|dK2|={d:.3e}")
    assert d <= 1e-9


def test_batch_argument_errors(setup):
    tb, t, reqs, V, torch = setup
    Vt = ev.v_tensor(V)
    with pytest.raises(ev.K2Error):
        ev.K2Batch(t, [(100, 10, 0, 1e-9), (100, 10, 5, 1e-9)], Vt)       # overlap
    with pytest.raises(ev.K2Error):
        ev.K2Batch(t, [(100, 1, i, 1e-9 * (i + 1)) for i in range(5)], Vt)  # 5 distinct k
    with pytest.raises(ev.K2Error):
        ev.K2Batch(t, [(tb.n + 1, 1, 0, 1e-9)], Vt)
