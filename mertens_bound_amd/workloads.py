"""Workload builders that need the GPU evaluator (inputs of the polish workload)."""
from __future__ import annotations

import numpy as np

from . import evaluator as ev
from . import synth


def polish_candidates(table: ev.K2Table, n: int = 4096, seed: int = 0, n_peaks: int = 16,
                      n_screen: int = 65_536, k: float = 1.5e-10, sigma: float = 1e-3):
    """Candidates near synthetic peaks: screen n_screen uniform candidates in [3.4e18, 3.6e18]
    with K2, polish the n_peaks largest |K2| with 8 Newton steps, and draw n / n_peaks points
    v0 + N(0, sigma) around each polished peak (64.64 words).  Deterministic for a seed: the
    GPU sums are evaluated in a fixed order."""
    import torch
    from . import autograd as ag
    V = synth.make_candidates(n_screen, 3.4e18, 3.6e18, seed + 101)
    K, _, _ = table.eval_device(ev.v_tensor(V), k, flags=ev.VALUE)
    top = np.argsort(-np.abs(K.cpu().numpy()), kind="stable")[:n_peaks]
    Vp, _, _, _ = ag.polish(table, ev.v_tensor(V[top]), k, n_steps=8)
    peaks = synth.v_to_int(ev.v_from_tensor(Vp))
    torch.cuda.synchronize()
    rng = np.random.default_rng([seed, 0x504F4C49])
    per = n // n_peaks
    out = []
    for p in peaks:
        d = rng.normal(0.0, sigma, size=per)
        out.extend(p + int(round(x * 2.0 ** 64)) for x in d)
    return synth.v_from_int(out)
