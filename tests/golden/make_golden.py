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
  batch512_sample.npz  the 512 uneven requests on the 2,325,572-zero table (seed 0): two
                       candidates (first and last) of each of 32 requests (every 16th), each at
                       its own cutoff and k: binary128 K2, K2' and sum |terms|, sum g |terms|
                       (about 1 min on 8 cores)
  polish4096_sample.npz the 4,096-candidate polish workload on the 2,325,572-zero table: the
                       start words of 16 candidates (every 256th), the oracle's FP64 Newton polish
                       of them (8 steps, clamp 1e-4, the reference's step rule) and binary128 K2,
                       K2', K2'' at its end points.  The start points come from a GPU screen
                       (mertens_bound_amd.workloads.polish_candidates), so this one runs where a
                       GPU is visible: OMP_NUM_THREADS=1 make_golden.py polish4096 --out <dir>
                       (one single-threaded oracle polish per sampled candidate, side by side)

  screen_sample.npz    the leaf-screen workload (2^24 candidates, k = 3e-9, 7 shifts of 0.0014):
                       for the first 512 candidates, max over the shifts of |K2| over the first
                       64, 128 and 256 zeros (the three stage values) in the oracle's FP64 tier,
                       and sum |a| per stage (the values' scale)

Usage: python tests/golden/make_golden.py [w1024x10000|w65536x2325572|polish|ref|batch512|
                                           polish4096|screen|all] [--out DIR]
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


def batch512_sample_index(nz, ks, cnt, offs, every=16):
    """(request, candidate) pairs of the batch fixture: the first and last candidate of every
    `every`-th request (one where a request has a single candidate)."""
    pairs = []
    for r in range(0, len(nz), every):
        pairs.append((r, int(offs[r])))
        if cnt[r] > 1:
            pairs.append((r, int(offs[r] + cnt[r] - 1)))
    return pairs


def batch512():
    c = synth.W65536x2325572
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    nz, ks, cnt, offs, V = synth.make_batched_requests(tb, n_req=512, seed=0)
    pairs = batch512_sample_index(nz, ks, cnt, offs)
    t = time.time()
    K = []; dK = []; aS = []; gS = []
    for r, i in pairs:
        sub = synth.ZeroTable(tb.G[:nz[r]], tb.gamma[:nz[r]], tb.w[:nz[r]])
        k0, k1, _, s = oracle.eval_quad(sub, V[i:i + 1], ks[r], flags=oracle.VALUE | oracle.D1)
        K.append(k0[0]); dK.append(k1[0]); aS.append(s[0])
        gS.append(float(np.sum(2 * np.exp(-ks[r] * sub.gamma ** 2) * np.abs(sub.w) * sub.gamma)))
    print(f"batch512 sample ({len(pairs)} candidates) oracle {time.time() - t:.1f}s")
    req = np.array([r for r, _ in pairs]); idx = np.array([i for _, i in pairs])
    np.savez_compressed(os.path.join(OUT, "batch512_sample.npz"), req=req, idx=idx, V=V[idx],
                        n_zeros=nz[req], k=ks[req], K=np.array(K), dK=np.array(dK),
                        absS=np.array(aS), gabsS=np.array(gS), table_sha256=tb.sha256())


def polish4096():
    import concurrent.futures
    import torch
    from mertens_bound_amd import evaluator as ev
    from mertens_bound_amd import workloads
    c = synth.W65536x2325572
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    torch.cuda.set_device(0)
    V0 = workloads.polish_candidates(ev.K2Table(tb), 4096, seed=0)
    idx = np.arange(0, 4096, 256)
    t = time.time()
    # one oracle polish per candidate, in parallel (the library releases the GIL)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(idx)) as ex:
        res = list(ex.map(lambda i: oracle.polish_f64(tb, V0[i:i + 1], c["k"], n_steps=8,
                                                      clamp=1e-4), idx))
    Vp = np.concatenate([r[0] for r in res]); steps = np.concatenate([r[4] for r in res])
    K0 = np.concatenate([r[1] for r in res])
    print(f"polish4096 oracle polish of {len(idx)} candidates {time.time() - t:.1f}s")
    t = time.time()
    K, dK, d2K, aS = oracle.eval_quad(tb, Vp, c["k"], nthreads=len(idx))
    print(f"polish4096 binary128 at the end points {time.time() - t:.1f}s")
    np.savez_compressed(os.path.join(OUT, "polish4096_sample.npz"), idx=idx, V0=V0[idx], V=Vp,
                        steps=steps, K0=K0, K=K, dK=dK, d2K=d2K, absS=aS, k=c["k"],
                        table_sha256=tb.sha256())


def screen():
    c = synth.W65536x2325572
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    V = synth.make_candidates(1 << 24, c["v_lo"], c["v_hi"], c["seed"])[:512]
    k, d = 3e-9, 0.0014
    vals = np.zeros((V.shape[0], 3))
    scale = np.zeros(3)
    for s, nz in enumerate((64, 128, 256)):
        hs = []
        for m in range(-3, 4):
            Vm = synth.v_from_int([v + int(round(m * d * 2.0 ** 64)) for v in synth.v_to_int(V)])
            hs.append(oracle.eval_f64(tb, Vm, k, flags=oracle.VALUE, n_zeros=nz)[0])
        vals[:, s] = np.max(np.abs(np.stack(hs, axis=1)), axis=1)
        scale[s] = np.sum(2 * np.exp(-k * tb.gamma[:nz] ** 2) * np.abs(tb.w[:nz]))
    np.savez_compressed(os.path.join(OUT, "screen_sample.npz"), V=V, val=vals, scale=scale, k=k,
                        shift=d, table_sha256=tb.sha256())


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
    args = sys.argv[1:]
    if "--out" in args:
        i = args.index("--out")
        OUT = os.path.abspath(args[i + 1])
        os.makedirs(OUT, exist_ok=True)
        del args[i:i + 2]
    what = args[0] if args else "all"
    oracle.build()
    for name, fn in (("w1024x10000", w1024x10000), ("polish", polish), ("ref", ref),
                     ("w65536x2325572", w65536x2325572), ("batch512", batch512),
                     ("polish4096", polish4096), ("screen", screen)):
        if what in (name, "all"):
            fn()
