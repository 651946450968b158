"""Device Newton polish (8 FP64 steps on the exact 64.64 point) against the oracle's polish."""
import numpy as np
import pytest

from conftest import golden
from mertens_bound_amd import evaluator as ev
from mertens_bound_amd import synth

pytestmark = pytest.mark.gpu


def test_polish_matches_oracle():
    g = golden("polish_small.npz")
    tb = synth.make_zero_table(3000, 11)
    t = ev.K2Table(tb)
    k = float(g["k"])
    V0 = g["V0"]
    K0, _, _ = t.eval(V0, k, flags=ev.VALUE)
    Vt = ev.v_tensor(V0)
    K, dK, d2K, steps = t.polish_device(Vt, k, n_steps=8, clamp=1e-4)
    Vf = ev.v_from_tensor(Vt)
    K = K.cpu().numpy()
    keep = np.abs(K0) > np.abs(K)          # path1_eval_fast.cpp:135-141
    Vsel = np.where(keep[:, None], V0, Vf)
    Ksel = np.where(keep, K0, K)
    np.testing.assert_allclose(Ksel, g["K"], rtol=0, atol=1e-12)
    dv = np.array([abs(a - b) for a, b in zip(synth.v_to_int(Vsel), synth.v_to_int(g["V"]))])
    assert np.all(dv <= 2 ** 64 * 1e-9)
    assert np.array_equal(steps.cpu().numpy(), g["steps"])


def test_polish_compaction_counts_and_matches_full_passes():
    """After each step only the candidates still moving are evaluated again; the result must not
    depend on that (a per-candidate sum does not depend on which other candidates share a pass
    beyond rounding), and the evaluation count must reflect the early stops."""
    g = golden("polish_small.npz")
    tb = synth.make_zero_table(3000, 11)
    t = ev.K2Table(tb)
    k = float(g["k"])
    V0 = g["V0"]
    n = V0.shape[0]
    Vt = ev.v_tensor(V0)
    K, dK, d2K, steps = t.polish_device(Vt, k, n_steps=8, clamp=1e-4)
    evals = t.polish_stats()
    st = steps.cpu().numpy()
    print(f"polish: {n} candidates, mean steps {st.mean():.2f}, candidate evaluations {evals} "
          f"(full passes would be {n * 9})")
    assert 2 * n <= evals <= 9 * n
    if st.max() < 8:
        assert evals < 9 * n
    # every candidate is evaluated once per step it took, and once more when it stopped because
    # its step fell below 1e-13, plus the first and the final pass
    assert evals >= n + int(st.sum())
