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
