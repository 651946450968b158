"""torch.autograd wrapper: gradients against the oracle's analytic derivative and central
differences, second derivative through create_graph, and the autograd-driven Newton polish
against the device polish."""
import numpy as np
import pytest

from mertens_bound_amd import autograd as ag
from mertens_bound_amd import evaluator as ev
from mertens_bound_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import torch
    tb = synth.make_zero_table(20_000, 12)
    return tb, ev.K2Table(tb), torch


@pytest.mark.parametrize("second_order", [False, True])
def test_gradient_vs_oracle_and_fd(setup, oracle_lib, second_order):
    """K2'' through create_graph: from the forward pass (second_order=True) or from one more pass
    at the same points (the default forward computes K2 and K2' only)."""
    tb, t, torch = setup
    k = 1.5e-10
    V0 = synth.make_candidates(64, 3.4e18, 3.6e18, 12)
    V0t = ev.v_tensor(V0)
    delta = torch.zeros(64, dtype=torch.float64, device="cuda", requires_grad=True)
    K = ag.k2(delta, t, V0t, k, second_order=second_order)
    (g,) = torch.autograd.grad(K.sum(), delta, create_graph=True)
    (h,) = torch.autograd.grad(g.sum(), delta)
    Ko, dKo, d2Ko, _ = oracle_lib.eval_quad(tb, V0, k)
    a = 2 * np.exp(-k * tb.gamma ** 2) * np.abs(tb.w)
    s1, s2 = np.sum(tb.gamma * a), np.sum(tb.gamma ** 2 * a)
    rel1 = np.max(np.abs(g.detach().cpu().numpy() - dKo)) / s1
    rel2 = np.max(np.abs(h.cpu().numpy() - d2Ko)) / s2
    print(f"scale-relative gradient error {rel1:.2e}, second derivative {rel2:.2e}")
    assert rel1 <= 1e-6 and rel2 <= 1e-6
    # central differences at h = 1e-6 (truncation ~ (g_eff h)^2/6: about 1e-3 relative here)
    hstep = 1e-6
    with torch.no_grad():
        Kp = ag.k2(torch.full((64,), hstep, dtype=torch.float64, device="cuda"), t, V0t, k)
        Km = ag.k2(torch.full((64,), -hstep, dtype=torch.float64, device="cuda"), t, V0t, k)
    fd = ((Kp - Km) / (2 * hstep)).cpu().numpy()
    assert np.max(np.abs(fd - dKo)) / s1 <= 2e-3


def test_offset_points_exact(setup):
    tb, t, torch = setup
    V0 = synth.v_from_int([(5 << 64) + (3 << 62), (7 << 64), (1 << 64) - 1])
    V0t = ev.v_tensor(V0)
    d = torch.tensor([0.375, -0.25, 2.0 ** -64], dtype=torch.float64, device="cuda")
    out = ev.v_from_tensor(ag.offset_points(V0t, d))
    # carry into the integer word, borrow out of it, carry from the last fraction bit
    want = [(6 << 64) + (1 << 61), (7 << 64) - (1 << 62), 1 << 64]
    assert synth.v_to_int(out) == want


def test_newton_autograd_matches_device_polish(setup):
    tb, t, torch = setup
    k = 3e-9
    V0 = synth.make_candidates(32, 3.4e18, 3.6e18, 13)
    V0t = ev.v_tensor(V0)
    Va, Ka, ga, sa = ag.newton_polish(t, V0t, k, n_steps=8)
    Vb, Kb, gb, sb = ag.polish(t, V0t, k, n_steps=8)
    torch.cuda.synchronize()
    assert torch.max(torch.abs(Ka - Kb)).item() <= 1e-10
    dv = [abs(a - b) for a, b in zip(synth.v_to_int(ev.v_from_tensor(Va)),
                                     synth.v_to_int(ev.v_from_tensor(Vb)))]
    assert max(dv) <= 2 ** 64 * 1e-10
