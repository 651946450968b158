"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

Inputs are regenerated from the seeded generator (mertens_bound_amd.synth); the fixtures hold
the input sha256 values and the oracle outputs only.

  w1024x10000_oracle.npz   all 1,024 candidates of the 10,000-zero workload (k = 1.5e-10, seed 0):
                       binary128 K2, K2', K2'', sum |terms|
  w65536x2325572_sub256.npz   the first 256 candidates of the 65,536-candidate workload over all
                       2,325,572 zeros: binary128 K2, K2', K2''  (about 5 min on 8 cores)
  polish_small.npz     FP64 Newton polish (8 steps) of 16 candidates on a 3,000-zero table
  ref_compat.npz       the reference evaluator (oracle/_ref, compiled from
                       search/path1_eval_fast.cpp) on reference-format words of a small table:
                       row sums and leading entries of compute_peak_AB's A/B, and the
                       polish_peaks_with_weights outputs

Usage: python tests/golden/make_golden.py [w1024x10000|w65536x2325572|polish|ref|all]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from mertens_bound_amd import formats, synth  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def w1024x10000():
    c = synth.W1024x10000
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    V = synth.make_candidates(c["n_cand"], c["v_lo"], c["v_hi"], c["seed"])
    t = time.time()
    K, dK, d2K, aS = oracle.eval_quad(tb, V, c["k"])
    print(f"w1024x10000 oracle {time.time() - t:.1f}s")
    np.savez_compressed(os.path.join(OUT, "w1024x10000_oracle.npz"), K=K, dK=dK, d2K=d2K, absS=aS,
                        table_sha256=tb.sha256(), V=V)


def w65536x2325572():
    c = synth.W65536x2325572
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    V = synth.make_candidates(c["n_cand"], c["v_lo"], c["v_hi"], c["seed"])[:256]
    t = time.time()
    K, dK, d2K, aS = oracle.eval_quad(tb, V, c["k"])
    print(f"w65536x2325572 subsample oracle {time.time() - t:.1f}s")
    np.savez_compressed(os.path.join(OUT, "w65536x2325572_sub256.npz"), K=K, dK=dK, d2K=d2K, absS=aS,
                        table_sha256=tb.sha256(), V=V)


def polish():
    tb = synth.make_zero_table(3000, 11)
    V = synth.make_candidates(16, 3.4e18, 3.6e18, 11)
    Vout, K0, Kf, dKf, st = oracle.polish_f64(tb, V, 3e-9, n_steps=8, clamp=1e-4)
    np.savez_compressed(os.path.join(OUT, "polish_small.npz"), V0=V, V=Vout, K0=K0, K=Kf, dK=dKf,
                        steps=st, table_sha256=tb.sha256(), k=3e-9)


def ref_inputs():
    """Reference-format inputs of the ref_compat fixture (also used by the tests)."""
    tb = synth.make_zero_table(2000, 5)
    rw = formats.reference_words(tb)
    V = synth.make_candidates(24, 3.4e18, 3.6e18, 5)
    xs, dy = formats.candidates_to_x_dy(V)
    dy = ref_dy0(dy)
    xlo, xhi = formats.x_words(xs)
    return tb, rw, (xlo, xhi, dy, rw["alpha"], rw["gamma"], rw["ghi_lo"], rw["ghi_hi"],
                    rw["glo_lo"], rw["glo_hi"], rw["psi_lo"], rw["psi_hi"])


def ref_dy0(dy):
    # sub-grid offsets of the size the reference's candidates carry (|dy0| <~ 4e-3)
    rng = np.random.default_rng(5)
    return dy + rng.uniform(-4e-3, 4e-3, size=len(dy))


def ref():
    L = oracle.load_ref()
    if L is None:
        raise SystemExit("oracle/_ref is not built (needs the reference source)")
    tb, rw, args = ref_inputs()
    A, B = oracle.ref_lib_compute_peak_AB(L, *args)
    w = np.exp(-3e-9 * rw["gamma"] ** 2)
    hu, hp, dd = oracle.ref_lib_polish(L, A, B, rw["gamma"], w, 2.5e-4, 40)
    np.savez_compressed(os.path.join(OUT, "ref_compat.npz"), dy0=args[2], A_rowsum=A.sum(axis=1),
                        B_rowsum=B.sum(axis=1), A_head=A[:, :8], B_head=B[:, :8], h_unpol=hu,
                        h_pol=hp, ddy_pol=dd, table_sha256=tb.sha256())


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    oracle.build()
    for name, fn in (("w1024x10000", w1024x10000), ("polish", polish), ("ref", ref),
                     ("w65536x2325572", w65536x2325572)):
        if what in (name, "all"):
            fn()
