"""Several weight vectors in one pass (k2_eval_weights_device) against the oracle: weight vector
m folded into the table's w (w_j -> wt_m[j] w_j) and evaluated with k = 0 gives the same K_m,
K_m', K_m'' by definition; Gaussian weights must reproduce the one-vector evaluator."""
import numpy as np
import pytest

from mertens_bound_amd import evaluator as ev
from mertens_bound_amd import pipeline, synth

pytestmark = pytest.mark.gpu
ALL = ev.VALUE | ev.D1 | ev.D2


@pytest.fixture(scope="module")
def setup():
    import torch
    tb = synth.make_zero_table(3000, 12)
    return tb, ev.K2Table(tb), torch


def _oracle(oracle_lib, tb, V, wt, nz):
    out = []
    for row in wt:
        tm = synth.ZeroTable(tb.G, tb.gamma, tb.w * row)
        out.append(oracle_lib.eval_f64(tm, V, 0.0, n_zeros=nz))
    return [np.stack([o[q] for o in out]) for q in range(3)]


@pytest.mark.parametrize("nz", [3000, 1700])
def test_reference_vectors_vs_oracle(setup, oracle_lib, nz):
    tb, t, torch = setup
    names, wt = pipeline.reference_weight_vectors(tb.gamma, n_trunc=1000)
    V = synth.make_candidates(300, 3.4e18, 3.6e18, 12)
    K, dK, d2K = t.eval_weights_device(ev.v_tensor(V), torch.tensor(wt, device="cuda"),
                                       n_zeros=nz, flags=ALL)
    Ko, dKo, d2Ko = _oracle(oracle_lib, tb, V, wt, nz)
    a = 2 * np.abs(wt[:, :nz] * tb.w[None, :nz])
    for q, (x, y, p) in enumerate(((K, Ko, 0), (dK, dKo, 1), (d2K, d2Ko, 2))):
        scale = (a * tb.gamma[None, :nz] ** p).sum(axis=1)[:, None]
        err = np.max(np.abs(x.cpu().numpy() - y) / scale)
        print(f"n_zeros={nz} q={q}: max scaled error {err:.2e}")
        assert err <= 1e-13


def test_gaussian_rows_match_plain_evaluator(setup):
    tb, t, torch = setup
    ks = [3e-9, 1e-9, 1.5e-10]
    wt = np.stack([np.exp(-k * tb.gamma ** 2) for k in ks])
    V = synth.make_candidates(700, 3.4e18, 3.6e18, 13)
    K, dK, _ = t.eval_weights_device(ev.v_tensor(V), torch.tensor(wt, device="cuda"),
                                     flags=ev.VALUE | ev.D1)
    for m, k in enumerate(ks):
        Kp, dKp, _ = t.eval(V, k, flags=ev.VALUE | ev.D1)
        assert np.max(np.abs(K[m].cpu().numpy() - Kp)) <= 1e-13
        assert np.max(np.abs(dK[m].cpu().numpy() - dKp)) <= 1e-13 * np.abs(dKp).max()


def test_weight_argument_errors(setup):
    tb, t, torch = setup
    V = ev.v_tensor(synth.make_candidates(8, 3.4e18, 3.6e18, 14))
    with pytest.raises(ev.K2Error):
        t.eval_weights_device(V, torch.ones((9, tb.n), dtype=torch.float64, device="cuda"))
    with pytest.raises(ev.K2Error):
        t.eval_weights_device(V, torch.ones((2, 100), dtype=torch.float64, device="cuda"))
    K, _, _ = t.eval_weights_device(V, torch.ones((2, tb.n), dtype=torch.float64, device="cuda"),
                                    n_zeros=100)
    assert K.shape == (2, 8)


def test_polish_with_weights_vs_reference_polish(setup, oracle_lib):
    """The reference's per-vector polish (restated polish_peaks_with_weights on compute_peak_AB
    rows of the same table, reference-format words, dy0 = 0) for the JP-kernel and raw vectors,
    against the fused-grid GPU pipeline on exact points."""
    from mertens_bound_amd import formats
    tb, t, torch = setup
    names, wt = pipeline.reference_weight_vectors(tb.gamma, n_trunc=1000)
    wt = wt[[0, 3]]
    V = synth.make_candidates(16, 3.4e18, 3.6e18, 15)
    h0, hp, Vp, ddy = pipeline.polish_with_weights(t, ev.v_tensor(V), wt)
    rw = formats.reference_words(tb)
    xs, dy = formats.candidates_to_x_dy(V)
    xlo, xhi = formats.x_words(xs)
    A, B = oracle_lib.ref_compute_peak_AB(xlo, xhi, dy, rw["alpha"], rw["gamma"], rw["ghi_lo"],
                                          rw["ghi_hi"], rw["glo_lo"], rw["glo_hi"], rw["psi_lo"],
                                          rw["psi_hi"])
    for r in range(2):
        hu_r, hp_r, dd_r = oracle_lib.ref_polish_peaks_with_weights(A, B, tb.gamma, wt[r],
                                                                    2.5e-4, 40)
        np.testing.assert_allclose(h0[r].cpu().numpy(), hu_r, atol=1e-11)
        np.testing.assert_allclose(hp[r].cpu().numpy(), hp_r, atol=1e-9)
        np.testing.assert_allclose(ddy[r].cpu().numpy(), dd_r, atol=1e-9)
