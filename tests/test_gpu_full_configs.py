"""The batched and polish workloads at their full size on the 2,325,572-zero table, against the
oracle.

* 512 uneven requests (cutoffs T log-uniform in [1e5, 1.3e6], k in {3e-9, 1e-9, 1.5e-10},
  1..4,096 candidates each) in one persistent launch replayed from its HIP graph, plain and
  mixed: the sampled candidates against binary128 (batch512_sample.npz), every candidate against
  the per-request loop of one-launch calls that the reference's drivers make
  (search/eval_tech1_600k.py:252-267).
* 4,096 candidates near synthetic peaks, 8 Newton steps (the step rule of
  search/path1_eval_fast.cpp:103-141), plain and mixed: the sampled candidates' end points and
  values against the oracle's own polish (polish4096_sample.npz, binary128 values), every sampled
  end point of both forms against the oracle evaluated at that same word, and K2' against
  central differences at h = 1e-6.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden
from mertens_bound_amd import autotune
from mertens_bound_amd import evaluator as ev
from mertens_bound_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full():
    import torch
    c = synth.W65536x2325572
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    return tb, ev.K2Table(tb), c, torch


def _scales(tb, k, nz, power):
    g = tb.gamma[:nz]
    return float(np.sum(2 * np.exp(-k * g * g) * np.abs(tb.w[:nz]) * g ** power))


def test_batch512_full_table_vs_oracle(full):
    tb, t, c, torch = full
    g = golden("batch512_sample.npz")
    assert str(g["table_sha256"]) == tb.sha256()
    nz, ks, cnt, offs, V = synth.make_batched_requests(tb, n_req=512, seed=0)
    idx = g["idx"]
    assert np.array_equal(V[idx], g["V"])
    assert set(np.round(g["k"] * 1e12).astype(int)) == {150, 1000, 3000}   # every k sampled
    reqs = [(int(a), int(b), int(o), float(k)) for a, b, o, k in zip(nz, cnt, offs, ks)]
    Vt = ev.v_tensor(V)
    res = {}
    for form, flags, bar_k, bar_d in (("plain", ev.VALUE | ev.D1, 1e-12, 1e-12),
                                      ("mixed", ev.VALUE | ev.D1 | ev.MIXED, 1e-9, 1e-6)):
        b = ev.K2Batch(t, reqs, Vt, flags=flags, tiling=autotune.batch_tiling(form))
        b.run()
        K, dK, _ = b.run()                      # the second graph replay
        torch.cuda.synchronize()
        K = K.cpu().numpy(); dK = dK.cpu().numpy()
        res[form] = (K, dK)
        eK = np.abs(K[idx] - g["K"])
        eD = np.abs(dK[idx] - g["dK"]) / g["gabsS"]
        print(f"batch512 {form}: {len(idx)} sampled candidates, max |dK2| {eK.max():.2e}, "
              f"max |dK2'|/scale {eD.max():.2e}")
        assert eK.max() <= bar_k
        assert eD.max() <= bar_d
        b.close()
    # every candidate of every request: the batch against one launch per request
    Kl = torch.zeros(V.shape[0], dtype=torch.float64, device="cuda")
    dKl = torch.zeros_like(Kl)
    for a, n, o, k in reqs:
        t.eval_device(Vt[o:o + n], k, n_zeros=a, flags=ev.VALUE | ev.D1,
                      out=(Kl[o:o + n], dKl[o:o + n], None))
    torch.cuda.synchronize()
    Kl = Kl.cpu().numpy(); dKl = dKl.cpu().numpy()
    Kp, dKp = res["plain"]
    Km, _ = res["mixed"]
    print(f"batch512 all {V.shape[0]} candidates: plain vs per-request max |dK2| "
          f"{np.abs(Kp - Kl).max():.2e}, mixed vs plain {np.abs(Km - Kp).max():.2e}")
    assert np.abs(Kp - Kl).max() <= 1e-12
    assert np.abs(dKp - dKl).max() <= 1e-12 * np.abs(dKl).max()
    assert np.abs(Km - Kp).max() <= 1e-9


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "polish4096_sample.npz")),
                    reason="polish4096_sample.npz not generated")
def test_polish4096_full_table_vs_oracle(full, oracle_lib):
    from mertens_bound_amd import workloads
    tb, t, c, torch = full
    g = golden("polish4096_sample.npz")
    assert str(g["table_sha256"]) == tb.sha256()
    k = float(g["k"])
    V0 = workloads.polish_candidates(t, 4096, seed=0)
    idx = g["idx"]
    assert np.array_equal(V0[idx], g["V0"])
    s0, s1, s2 = (_scales(tb, k, tb.n, p) for p in (0, 1, 2))
    conv = g["steps"] < 8                      # the oracle's polish stopped on the step rule
    V0t = ev.v_tensor(V0)
    ends = {}
    for form, mixed in (("plain", False), ("mixed", True)):
        Vt = V0t.clone()
        K, dK, d2K, steps = t.polish_device(Vt, k, n_steps=8, clamp=1e-4,
                                            flags=ev.MIXED if mixed else 0)
        K0, _, _ = t.eval_device(V0t, k, flags=ev.VALUE)
        torch.cuda.synchronize()
        K, K0, steps = K.cpu().numpy(), K0.cpu().numpy(), steps.cpu().numpy()
        Vf = ev.v_from_tensor(Vt)
        keep = np.abs(K0) > np.abs(K)           # path1_eval_fast.cpp:135-141
        Vsel = np.where(keep[:, None], V0, Vf)[idx]
        ends[form] = Vsel
        # the GPU's values at its own end points against the oracle at the same words
        Ko, dKo, d2Ko = oracle_lib.eval_f64(tb, Vsel, k)
        Kg, dKg, d2Kg = t.eval(Vsel, k, flags=ev.VALUE | ev.D1 | ev.D2 |
                               (ev.MIXED if mixed else 0))
        eK = np.abs(Kg - Ko).max(); e1 = np.abs(dKg - dKo).max() / s1
        e2 = np.abs(d2Kg - d2Ko).max() / s2
        # and the value the polish itself returned (the start point's where that was larger)
        eR = np.abs(np.where(keep, K0, K)[idx] - Ko).max()
        assert eR <= (1e-9 if mixed else 1e-12), eR
        dv = np.array([abs(a - b) / 2.0 ** 64 for a, b in
                       zip(synth.v_to_int(Vsel), synth.v_to_int(g["V"]))])
        print(f"polish4096 {form}: steps {np.bincount(steps, minlength=9)}; at its end points vs "
              f"the oracle: |dK2| {eK:.2e}, |dK2'|/scale {e1:.2e}, |dK2''|/scale {e2:.2e}; end "
              f"points vs the oracle's polish: max |dv| {dv.max():.2e} (converged "
              f"{dv[conv].max() if conv.any() else 0:.2e}), |dK2| vs binary128 "
              f"{np.abs(np.where(keep[idx], K0[idx], K[idx]) - g['K']).max():.2e}")
        if not mixed:
            assert eK <= 1e-12 and e1 <= 1e-12 and e2 <= 1e-12
            # the same Newton trajectories as the oracle: the same step counts, the same end
            # points to well within one FP64 step rounding, binary128 values within the bar
            assert np.array_equal(steps[idx], g["steps"])
            assert dv.max() <= 1e-12
            assert np.abs(np.where(keep[idx], K0[idx], K[idx]) - g["K"]).max() <= 1e-9
            assert np.abs(dKg - g["dK"]).max() / s1 <= 1e-6
        else:
            # the precision form at its own end points, every sampled candidate
            assert eK <= 1e-9 and e1 <= 1e-6
            # where the oracle's polish converged, the precision form lands on the same peak
            # (its stop rule 1e-11 and FP32 rounding in K2' leave about 5e-12 in v)
            assert dv[conv].max() <= 1e-10
    # K2' against central differences at h = 1e-6 at the plain end points: the Taylor remainder
    # h^2/6 sum g^3 |a| bounds the difference (plus the two values' rounding / 2h)
    h = 1e-6
    Vsel = ends["plain"]
    Vp = synth.v_from_int([v + int(round(h * 2.0 ** 64)) for v in synth.v_to_int(Vsel)])
    Vm = synth.v_from_int([v - int(round(h * 2.0 ** 64)) for v in synth.v_to_int(Vsel)])
    Kp, _, _ = t.eval(Vp, k, flags=ev.VALUE)
    Km, _, _ = t.eval(Vm, k, flags=ev.VALUE)
    _, dK, _ = t.eval(Vsel, k, flags=ev.VALUE | ev.D1)
    fd = (Kp - Km) / (2 * h)
    bound = h * h / 6 * _scales(tb, k, tb.n, 3) + 1e-12 * s0 / h
    print(f"central FD at h=1e-6: max |FD - K2'| {np.abs(fd - dK).max():.3e} (bound {bound:.3e}, "
          f"scale {s1:.3e})")
    assert np.abs(fd - dK).max() <= bound
