"""Calls on one table handle queued on two streams without host synchronisation give the results
each gives alone: the table's scratch (partial sums, screen records, enclosure buffers) and its
cache of per-k weights (4 sets, least recently used evicted) are shared, and every device entry
point orders itself after the previous call's work when that ran on another stream."""
import numpy as np
import pytest

from mertens_bound_amd import evaluator as ev
from mertens_bound_amd import synth

pytestmark = pytest.mark.gpu


def test_two_streams_one_table_match_sequential():
    import torch
    tb = synth.make_zero_table(200_000, 43)
    t = ev.K2Table(tb)
    Va = ev.v_tensor(synth.make_candidates(65_536, 3.4e18, 3.6e18, 71))
    Vb = ev.v_tensor(synth.make_candidates(3_000, 3.4e18, 3.6e18, 72))
    Vs = ev.v_tensor(synth.make_candidates(8_192, 3.4e18, 3.6e18, 73))
    # six distinct k: the weight cache (4 sets) evicts sets that calls queued earlier read
    ks = [1.5e-10, 1e-9, 3e-9, 2e-10, 5e-10, 7e-10]
    fl = ev.VALUE | ev.D1
    calls = [("eval", Va, ks[0], fl), ("eval", Vb, ks[1], fl | ev.MIXED),
             ("screen", Vs, ks[2], None), ("eval", Va, ks[3], fl),
             ("enclose", Vb, ks[4], None), ("eval", Vb, ks[5], fl),
             ("eval", Va, ks[0], fl | ev.MIXED), ("screen", Vs, ks[1], None)]

    def run(c, stream):
        kind, V, k, f = c
        if kind == "eval":
            return t.eval_device(V, k, flags=f, stream=stream)
        if kind == "screen":
            return t.screen_device(V, k, 0.0014, 3, ((64, 0.3), (128, 0.4)), stream=stream)
        return t.enclose_device(V, k, stream=stream)[:2]

    def copy(r):
        return [x.clone() for x in r if x is not None]

    alone = []
    for c in calls:
        r = run(c, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        alone.append(copy(r))
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    got = []
    for i, c in enumerate(calls):
        s = sa if i % 2 == 0 else sb
        with torch.cuda.stream(s):
            r = run(c, s.cuda_stream)
            got.append(r)
            # each result is copied on the stream that made it, before a later call can reuse
            # the tensors' memory
            got[-1] = copy(r)
    torch.cuda.synchronize()
    for i, (g, w) in enumerate(zip(got, alone)):
        for x, y in zip(g, w):
            assert torch.equal(x, y), (i, calls[i][0], calls[i][2])


def test_batch_plan_runs_on_two_streams():
    """One batched plan replayed on two streams in turn (the plan's partial sums and work counter
    are shared by its runs): every replay gives the direct launch's results."""
    import torch
    tb = synth.make_zero_table(60_000, 44)
    t = ev.K2Table(tb)
    V = synth.make_candidates(5_000, 3.4e18, 3.6e18, 74)
    Vt = ev.v_tensor(V)
    reqs = [(60_000, 2_000, 0, 1.5e-10), (30_000, 2_500, 2_000, 1e-9), (45_000, 500, 4_500, 3e-9)]
    b = ev.K2Batch(t, reqs, Vt, flags=ev.VALUE | ev.D1)
    b.run_direct()
    torch.cuda.synchronize()
    want = (b.K.clone(), b.dK.clone())
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    outs = []
    for i in range(6):
        s = sa if i % 2 == 0 else sb
        with torch.cuda.stream(s):
            b.run(stream=s.cuda_stream)
            outs.append((b.K.clone(), b.dK.clone()))
    torch.cuda.synchronize()
    for K, dK in outs:
        assert torch.equal(K, want[0]) and torch.equal(dK, want[1])
    assert np.all(np.isfinite(want[0].cpu().numpy()))
