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


@pytest.mark.parametrize("flags, tol", [(0, 1e-13), (ev.MIXED, 1e-9)])
def test_polish_start_values_from_first_pass(flags, tol):
    """K0 of k2_polish_device (K at the starting points, taken from its first pass) against a
    separate evaluation at those points in the same form; with no Newton step the final pass is
    at the start and K0 equals K."""
    g = golden("polish_small.npz")
    tb = synth.make_zero_table(3000, 11)
    t = ev.K2Table(tb)
    k = float(g["k"])
    V0 = g["V0"]
    Kref, _, _ = t.eval(V0, k, flags=ev.VALUE | flags)
    K, dK, d2K, steps, K0 = t.polish_device(ev.v_tensor(V0), k, n_steps=8, clamp=1e-4,
                                            flags=flags, start=True)
    np.testing.assert_allclose(K0.cpu().numpy(), Kref, rtol=0, atol=tol)
    K, dK, d2K, steps, K0 = t.polish_device(ev.v_tensor(V0), k, n_steps=0, flags=flags,
                                            start=True)
    assert np.array_equal(K0.cpu().numpy(), K.cpu().numpy())
    np.testing.assert_allclose(K.cpu().numpy(), Kref, rtol=0, atol=tol)


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


def test_mixed_polish_matches_fp64_polish(oracle_lib):
    """The polish in the precision form (FP32 tail after the integer phase reduction, FP64 head
    and Newton step, stop once |step| < 1e-11) against the FP64 polish and the oracle.  Where
    both forms stop on their step rules they land on the same peak: the same point to within
    1e-10 (the form's K2' error over |K2''|) and K within 1e-9.  Every candidate's end point,
    converged or not, is checked against the oracle evaluated at that same word: K within 1e-9,
    K2' within 1e-6 of its scale.  Step counts differ (the two stop rules differ), so they are
    not compared."""
    import torch
    from mertens_bound_amd import autograd as ag
    from mertens_bound_amd import workloads
    tb = synth.make_zero_table(200_000, 5)
    t = ev.K2Table(tb)
    k = 1.5e-10
    V0 = workloads.polish_candidates(t, 256, seed=3, n_peaks=8, n_screen=4096, k=k)
    V0t = ev.v_tensor(V0)
    Va, Ka, dKa, sa = ag.polish(t, V0t, k, n_steps=8)
    Vb, Kb, dKb, sb = ag.polish(t, V0t, k, n_steps=8, mixed=True)
    torch.cuda.synchronize()
    Ka, Kb = Ka.cpu().numpy(), Kb.cpu().numpy()
    sa, sb = sa.cpu().numpy(), sb.cpu().numpy()
    Vb_np = ev.v_from_tensor(Vb)
    dv = np.array([abs(a - b) / 2.0 ** 64 for a, b in
                   zip(synth.v_to_int(ev.v_from_tensor(Va)), synth.v_to_int(Vb_np))])
    conv = (sa < 8) & (sb < 8)   # both stopped on their step rules
    dK = np.abs(Ka - Kb)
    Ko, dKo, _ = oracle_lib.eval_f64(tb, Vb_np, k, flags=oracle_lib.VALUE | oracle_lib.D1)
    s1 = np.sum(2 * np.exp(-k * tb.gamma ** 2) * np.abs(tb.w) * tb.gamma)
    eK = np.abs(Kb - Ko)
    print(f"mixed vs FP64 polish: {conv.sum()} of {len(sa)} converged in both; converged: max "
          f"|dK| {dK[conv].max():.3e}, max |dv| {dv[conv].max():.3e}; all vs the oracle at the "
          f"mixed end points: max |dK| {eK.max():.3e}; mixed steps {np.bincount(sb, minlength=9)}"
          f"; FP64 steps {np.bincount(sa, minlength=9)}")
    assert conv.sum() >= len(sa) // 2
    assert dK[conv].max() <= 1e-9
    assert dv[conv].max() <= 1e-10
    assert eK.max() <= 1e-9
    # K2' the polish returned against the oracle at the same words (where the start point was
    # kept, the returned K2' is that of the Newton end point, so compare only the others)
    moved = np.array([a != b for a, b in zip(synth.v_to_int(Vb_np), synth.v_to_int(V0))])
    assert np.max(np.abs(dKb.cpu().numpy() - dKo)[moved]) / s1 <= 1e-6
