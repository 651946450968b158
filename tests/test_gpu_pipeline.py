"""Pre-screen, grid + Newton polish and the hierarchical grid cascade against the oracle and the
reference's polish semantics."""
import numpy as np
import pytest

from mertens_bound_amd import evaluator as ev
from mertens_bound_amd import formats, pipeline, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import torch
    tb = synth.make_zero_table(5_000, 21)
    return tb, ev.K2Table(tb), torch


def _shift_int(V, m, d):
    return synth.v_from_int([v + int(round(m * d * 2.0 ** 64)) for v in synth.v_to_int(V)])


def test_screen_values_vs_oracle(setup, oracle_lib):
    tb, t, torch = setup
    k, d = 3e-9, 0.0014
    V = synth.make_candidates(512, 3.4e18, 3.6e18, 21)
    stages = ((64, 0.0), (128, 0.0), (256, 0.0))           # thresholds 0: every stage runs
    val, best, ok = t.screen_device(ev.v_tensor(V), k, d, 3, stages)
    val = val.cpu().numpy()
    assert np.all(ok.cpu().numpy() == 1)
    for s, (nz, _) in enumerate(stages):
        hs = np.stack([oracle_lib.eval_f64(tb, _shift_int(V, m, d), k, flags=1, n_zeros=nz)[0]
                       for m in range(-3, 4)], axis=1)
        want = np.max(np.abs(hs), axis=1)
        scale = np.sum(2 * np.exp(-k * tb.gamma[:nz] ** 2) * np.abs(tb.w[:nz]))
        err = np.max(np.abs(val[:, s] - want)) / scale
        print(f"screen stage {nz}: max error {err:.2e} of sum|a|")
        assert err <= 2e-6
        # the best shift (that of the last stage) is the oracle's argmax wherever the two
        # largest |h_m| differ by more than the error bar
        if s == len(stages) - 1:
            srt = np.sort(np.abs(hs), axis=1)
            clear = srt[:, -1] - srt[:, -2] > 4e-6 * scale
            want_best = np.argmax(np.abs(hs), axis=1) - 3
            assert clear.sum() > 400
            np.testing.assert_array_equal(best.cpu().numpy()[clear], want_best[clear])
    np.testing.assert_array_equal(val[:, 3], -1.0)


@pytest.mark.parametrize("n", [1, 63, 65, 257, 3000])
def test_screen_ragged_sizes(setup, oracle_lib, n):
    """Candidate counts that fill no whole wave or workgroup: every stage value against the
    oracle, and with thresholds that stop some candidates, the staged outputs against the run in
    which every candidate takes every stage (bit for bit)."""
    tb, t, torch = setup
    k, d = 3e-9, 0.0014
    V = synth.make_candidates(n, 3.4e18, 3.6e18, 100 + n)
    Vt = ev.v_tensor(V)
    allv, allb, allok = t.screen_device(Vt, k, d, 3, ((64, 0.0), (128, 0.0), (256, 0.0)))
    allv = allv.cpu().numpy()
    assert np.all(allok.cpu().numpy() == 1)
    for s, nz in enumerate((64, 128, 256)):
        hs = np.stack([oracle_lib.eval_f64(tb, _shift_int(V, m, d), k, flags=1, n_zeros=nz)[0]
                       for m in range(-3, 4)], axis=1)
        scale = np.sum(2 * np.exp(-k * tb.gamma[:nz] ** 2) * np.abs(tb.w[:nz]))
        assert np.max(np.abs(allv[:, s] - np.max(np.abs(hs), axis=1))) <= 2e-6 * scale
    thr = np.float32(np.median(allv[:, 0])) if n > 1 else np.float32(0.0)
    stages = ((64, float(thr)), (128, 0.0), (256, 0.0))
    val, best, ok = (x.cpu().numpy() for x in t.screen_device(Vt, k, d, 3, stages))
    reach = allv[:, 0] >= thr
    assert np.array_equal(ok == 1, reach)
    assert np.array_equal(val[reach, :3], allv[reach, :3])
    assert np.all(val[~reach, 1:] == -1)
    assert np.array_equal(best[reach], allb.cpu().numpy()[reach])


def test_screen_no_candidates(setup):
    tb, t, torch = setup
    Vt = ev.v_tensor(synth.make_candidates(1, 3.4e18, 3.6e18, 1))[:0]
    val, best, ok = t.screen_device(Vt, 3e-9, 0.0014, 3, ((64, 0.0), (128, 0.0), (256, 0.0)))
    assert val.shape[0] == 0 and ok.shape[0] == 0


def test_screen_staging_and_survivors(setup):
    tb, t, torch = setup
    V = synth.make_candidates(4096, 3.4e18, 3.6e18, 22)
    Vt = ev.v_tensor(V)
    val, best, ok = t.screen_device(Vt, 3e-9, 0.0014, 3, ((64, 0.3), (128, 0.4), (256, 0.5)))
    val = val.cpu().numpy(); ok = ok.cpu().numpy()
    # a candidate passes iff every stage value reaches its threshold; failed stages stop the rest
    thr = np.array([0.3, 0.4, 0.5])
    for i in range(len(ok)):
        reached = val[i, :3] >= 0
        n = int(reached.sum())
        assert np.all(val[i, :n - 1] >= thr[:n - 1])
        assert ok[i] == int(n == 3 and val[i, 2] >= thr[2])
        assert np.all(val[i, n:] == -1)
    idx, Vs, _ = pipeline.screen(t, Vt, 3e-9, 0.0014, 3, ((64, 0.3), (128, 0.4), (256, 0.5)))
    assert idx.numel() == int(ok.sum())
    # survivors sit at their best shift: K over 256 zeros there equals the stage-3 value
    K, _, _ = t.eval_device(Vs, 3e-9, n_zeros=256, flags=ev.VALUE)
    np.testing.assert_allclose(np.abs(K.cpu().numpy()), val[idx.cpu().numpy(), 2], atol=2e-5)


def test_screen_calls_queue_without_host_sync(setup):
    """Two screens with different k, shifts and stages queued on one stream (the call does not
    synchronise; the second rewrites the table's zero records after the first has read them)
    give what each gives alone."""
    tb, t, torch = setup
    Vt = ev.v_tensor(synth.make_candidates(2048, 3.4e18, 3.6e18, 23))
    a = (3e-9, 0.0014, 3, ((64, 0.2), (128, 0.3), (256, 0.4)))
    b = (1.5e-10, 0.002, 5, ((100, 0.25), (400, 0.35)))
    alone = []
    for p in (a, b):
        r = t.screen_device(Vt, *p)
        torch.cuda.synchronize()
        alone.append([x.clone() for x in r])
    ra = t.screen_device(Vt, *a)
    rb = t.screen_device(Vt, *b)
    torch.cuda.synchronize()
    for got, want in zip((ra, rb), alone):
        for x, y in zip(got, want):
            assert torch.equal(x, y)
    assert not torch.equal(ra[0], rb[0])


def test_screen_sparse_survivors_staged(setup):
    """Later stages run over compacted survivor lists (appended with atomics, in any order): with
    thresholds that stop most candidates, every output equals the single-stage-at-a-time
    definition -- a candidate's stage values are those of a run where every candidate takes every
    stage, it reaches stage s exactly when the values before reach their thresholds, and two
    calls give the same bits."""
    tb, t, torch = setup
    V = synth.make_candidates(20_000, 3.4e18, 3.6e18, 25)
    Vt = ev.v_tensor(V)
    k, d = 3e-9, 0.0014
    allv, allb, _ = t.screen_device(Vt, k, d, 3, ((64, 0.0), (128, 0.0), (256, 0.0)))
    allv = allv.cpu().numpy()
    # stage 1 stops 90% of the candidates; each later threshold is the median of the values of
    # the candidates that reach it
    thr = np.zeros(3, dtype=np.float32)
    alive = np.ones(len(V), dtype=bool)
    for s, q in enumerate((0.9, 0.5, 0.5)):
        thr[s] = np.quantile(allv[alive, s], q)
        alive &= allv[:, s] >= thr[s]
    stages = tuple((n, float(x)) for n, x in zip((64, 128, 256), thr))
    r1 = t.screen_device(Vt, k, d, 3, stages)
    r2 = t.screen_device(Vt, k, d, 3, stages)
    torch.cuda.synchronize()
    for x, y in zip(r1, r2):
        assert torch.equal(x, y)
    val, best, ok = (x.cpu().numpy() for x in r1)
    reach = np.ones(len(V), dtype=bool)
    for s in range(3):
        assert np.array_equal(val[reach, s], allv[reach, s])
        assert np.all(val[~reach, s] == -1)
        reach = reach & (allv[:, s] >= thr[s])
    assert np.array_equal(ok == 1, reach)
    print(f"staged screen: {np.sum(val[:, 1] >= 0)} of {len(V)} reach stage 2, "
          f"{np.sum(val[:, 2] >= 0)} stage 3, {reach.sum()} pass")
    assert 0 < reach.sum() < np.sum(val[:, 1] >= 0) < len(V)
    # the best shift of a candidate that completes every stage is that of its last stage
    assert np.array_equal(best[reach], allb.cpu().numpy()[reach])


def test_grid_then_newton_matches_reference_polish(setup, oracle_lib):
    """Canonical grid + Newton vs the reference's polish_peaks_with_weights on A/B rows of the
    same table (reference-format words, dy0 = 0)."""
    tb, t, torch = setup
    k = 3e-9
    V = synth.make_candidates(24, 3.4e18, 3.6e18, 23)
    h0, hp, Vp, ddy = pipeline.grid_then_newton(t, ev.v_tensor(V), k)
    rw = formats.reference_words(tb)
    xs, dy = formats.candidates_to_x_dy(V)
    xlo, xhi = formats.x_words(xs)
    A, B = oracle_lib.ref_compute_peak_AB(xlo, xhi, dy, rw["alpha"], rw["gamma"], rw["ghi_lo"],
                                          rw["ghi_hi"], rw["glo_lo"], rw["glo_hi"], rw["psi_lo"],
                                          rw["psi_hi"])
    w = np.exp(-k * tb.gamma ** 2)
    hu_r, hp_r, dd_r = oracle_lib.ref_polish_peaks_with_weights(A, B, tb.gamma, w, 2.5e-4, 40)
    np.testing.assert_allclose(h0.cpu().numpy(), hu_r, atol=1e-11)
    np.testing.assert_allclose(hp.cpu().numpy(), hp_r, atol=1e-9)
    np.testing.assert_allclose(ddy.cpu().numpy(), dd_r, atol=1e-9)


def test_cascade_reference_stages_vs_restatement(oracle_lib):
    """The five-stage ZeroData::polish with the reference's stage table (REFERENCE_CASCADE)
    against the oracle's restatement: same stages passed, the same grid point (offset) and h to
    1e-12.  A 20,000-zero table with weights scaled 2.5x, so that the 40 largest of 20,000
    random candidates on the first 256 zeros spread over every stage (plus 24 random ones, which
    stop at the first), including the 641-point grid on the first 15,000 zeros and the
    241-point grid over the whole table."""
    import torch
    from oracle import pipelines as P
    tb0 = synth.make_zero_table(20_000, 21)
    tb = synth.ZeroTable(tb0.G, tb0.gamma, tb0.w * 2.5)
    k = 3e-9
    pool = synth.make_candidates(20_000, 3.4e18, 3.6e18, 24)
    K256 = oracle_lib.eval_f64(tb, pool, k, flags=1, n_zeros=256)[0]
    top = np.argsort(-np.abs(K256), kind="stable")[:40]
    V = np.concatenate([pool[top], pool[-24:]])
    assert pipeline.REFERENCE_CASCADE == P.CASCADE_STAGES
    t = ev.K2Table(tb)
    babs, bh, bd, passed = pipeline.cascade(t, ev.v_tensor(V), k)
    torch.cuda.synchronize()
    H, D, Pa = P.cascade(tb, V, k)
    passed = passed.cpu().numpy(); bd = bd.cpu().numpy(); bh = bh.cpu().numpy()
    print("stages passed:", np.bincount(Pa, minlength=4))
    assert np.bincount(Pa, minlength=4).min() >= 1        # every outcome is exercised
    np.testing.assert_array_equal(passed, Pa)
    np.testing.assert_array_equal(bd, D)
    np.testing.assert_allclose(bh, H, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(babs.cpu().numpy(), np.abs(bh))


def test_cascade_custom_stages(setup, oracle_lib):
    """Custom stage tables: h at the returned offset over the last stage's zeros."""
    tb, t, torch = setup
    k = 3e-9
    V = synth.make_candidates(64, 3.4e18, 3.6e18, 24)
    stages = ((256, 10, 1.5e-4, 0.2, False, True), (512, 20, 2.5e-5, 0.3, False, True),
              (2048, 8, 1e-5, 0.4, True, True))
    babs, bh, bd, passed = pipeline.cascade(t, ev.v_tensor(V), k, stages)
    bd = bd.cpu().numpy(); passed = passed.cpu().numpy(); bh = bh.cpu().numpy()
    for i in range(0, 64, 9):
        s = passed[i]
        last = min(s, len(stages) - 1)
        nz = stages[last][0]
        want = oracle_lib.eval_f64(tb, _shift_int(V[i:i + 1], 1, bd[i]), k, flags=1,
                                   n_zeros=nz)[0][0]
        assert abs(want - bh[i]) <= 1e-12


def _near_peaks(t, tb, k, n, seed, torch):
    """Candidates a few 1e-5 from local maxima of |K| (found by a fine GPU scan)."""
    V = synth.make_candidates(n, 3.4e18, 3.6e18, seed)
    hb, Vb, mb = pipeline.scan(t, ev.v_tensor(V), k, half_width=2e-3, step=2e-5, n_zeros=512)
    rng = np.random.default_rng(seed)
    d = rng.uniform(-3e-5, 3e-5, n)
    return ev.v_from_tensor(pipeline.ag.offset_points(Vb, torch.tensor(d, device=Vb.device)))


def test_polish_single_peak_vs_restatement(setup, oracle_lib):
    from oracle import pipelines as P
    tb, t, torch = setup
    k = 3e-9
    V = _near_peaks(t, tb, k, 12, 31, torch)
    h0, h, Vp, d = pipeline.polish_single_peak(t, ev.v_tensor(V), k)
    h0r, hr, dr = P.polish_single_peak(tb, V, k)
    np.testing.assert_allclose(h0.cpu().numpy(), h0r, atol=1e-12)
    np.testing.assert_allclose(h.cpu().numpy(), hr, atol=1e-11)
    np.testing.assert_allclose(d.cpu().numpy(), dr, atol=1e-10)


def test_fast_polish_batch_vs_restatement(setup, oracle_lib):
    from oracle import pipelines as P
    tb, t, torch = setup
    k = 3e-9
    V = _near_peaks(t, tb, k, 40, 32, torch)
    hu, hp, Vp, d = pipeline.fast_polish_batch(t, ev.v_tensor(V), k, top=10)
    hur, hpr, dr = P.fast_polish_batch(tb, V, k, top=10)
    np.testing.assert_allclose(hu.cpu().numpy(), hur, atol=1e-12)
    np.testing.assert_allclose(hp.cpu().numpy(), hpr, atol=1e-11)
    np.testing.assert_allclose(d.cpu().numpy(), dr, atol=1e-10)


def test_scan_vs_restatement(setup, oracle_lib):
    from oracle import pipelines as P
    tb, t, torch = setup
    k = 3e-9
    V = synth.make_candidates(6, 3.4e18, 3.6e18, 33)
    hb, Vb, mb = pipeline.scan(t, ev.v_tensor(V), k, half_width=1e-3, step=2e-5, n_zeros=2000,
                               chunk=37)
    hr, mr = P.scan(tb, V, k, half_width=1e-3, step=2e-5, n_zeros=2000, chunk=37)
    np.testing.assert_allclose(hb.cpu().numpy(), hr, atol=1e-12)
    assert np.array_equal(mb.cpu().numpy(), mr)
    want = P.offset(V, mr * 2e-5)
    assert np.array_equal(ev.v_from_tensor(Vb), want)
