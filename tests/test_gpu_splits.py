"""Guided zero splits on the GPU: a plan with long splits first (and padding splits past the
end) against a plan of uniform splits of the same launch, against the oracle's FP64 port, run
to run, and in the precision form."""
import numpy as np
import pytest

from mertens_bound_amd import evaluator as ev
from mertens_bound_amd import synth

pytestmark = pytest.mark.gpu
N_ZEROS, N_CAND, K = 600_000, 65_536, 1.5e-10
GUIDED, UNIFORM = 49_152, 4_096     # target workgroups: long splits first / uniform splits


@pytest.fixture(scope="module")
def setup():
    tb = synth.make_zero_table(N_ZEROS, 11)
    t = ev.K2Table(tb)
    V = synth.make_candidates(N_CAND, 3.4e18, 3.6e18, 12)
    a = 2 * np.exp(-K * tb.gamma ** 2) * np.abs(tb.w)
    return tb, t, V, (np.sum(a), np.sum(tb.gamma * a))


def _run(t, V, target, flags=ev.VALUE | ev.D1):
    t.set_tiling(nc=4, target_blocks=target, wg_threads=256)
    try:
        return t.eval(V, K, flags=flags)
    finally:
        t.set_tiling(nc=2, target_blocks=3072, wg_threads=256)


def test_plans_are_what_the_test_needs():
    g = ev.split_plan(N_ZEROS, N_CAND // 1024, GUIDED, 8)
    u = ev.split_plan(N_ZEROS, N_CAND // 1024, UNIFORM, 8)
    assert g["n_long"] > 0 and g["n"] % 8 != 0      # long splits and padding splits
    assert u["n_long"] == 0


def test_guided_vs_uniform_and_port(setup, oracle_lib):
    tb, t, V, (s0, s1) = setup
    Kg, dKg, _ = _run(t, V, GUIDED)
    Ku, dKu, _ = _run(t, V, UNIFORM)
    # the same terms grouped into other splits: FP64 rounding only
    assert np.max(np.abs(Kg - Ku)) / s0 <= 1e-13
    assert np.max(np.abs(dKg - dKu)) / s1 <= 1e-13
    sel = np.arange(0, N_CAND, 1021)
    Ko, dKo, _ = oracle_lib.eval_f64(tb, V[sel], K)
    assert np.max(np.abs(Kg[sel] - Ko)) <= 1e-12
    assert np.max(np.abs(dKg[sel] - dKo)) / s1 <= 1e-12
    # run to run: the same plan, bit for bit
    Kg2, dKg2, _ = _run(t, V, GUIDED)
    assert np.array_equal(Kg, Kg2) and np.array_equal(dKg, dKg2)


def test_guided_precision_form(setup):
    tb, t, V, (s0, s1) = setup
    Kf, _, _ = _run(t, V, GUIDED)
    Km, _, _ = _run(t, V, GUIDED, ev.VALUE | ev.D1 | ev.MIXED)
    assert np.max(np.abs(Km - Kf)) <= 1e-9
