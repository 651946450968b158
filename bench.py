"""Throughput of the K2 evaluator on one MI355X.

Workload (one "step"): K2 and dK2/dv of 65,536 candidates v in [3.4e18, 3.6e18] (64.64 fixed
point) against all 2,325,572 synthetic zeros (gamma < 1.3e6, Q18.110 words, complex128 weights),
k = 1.5e-10, seed 0 -- 1.524e11 candidate-zero evaluations.  Inputs are resident in HBM before
the timed region.  Metric: candidate-zero evaluations per second, valid only with |dK2| <= 1e-9
against the multiprecision oracle (checked on the committed 256-candidate fixture).

Prints one JSON line.  --form plain (FP64, default) or mixed (FP32 sin/cos after the exact phase).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic work per candidate-zero evaluation of K2 + K2' (see DESIGN.md):
FP64_FLOPS_PER_EVAL = 42      # sin/cos of the reduced phase + weighting + two accumulations
INT32_MAD_PER_EVAL = 15       # exact 128x128-bit phase product window
FP64_PEAK_TFLOPS = 78.6       # MI355X vector FP64, published
CDNA_CLOCK_HZ = 2.4e9


def cpu_baseline(tb, V, k, budget_s=20.0, n_c=64, chunk=1 << 18):
    """The reference evaluator (oracle/_ref, compiled from search/path1_eval_fast.cpp; 16 OpenMP
    threads, hard-coded there) on a bounded sample: compute_peak_AB over zero chunks plus the
    weighted sums K = sum e^{-k g^2} A and K' = -sum g e^{-k g^2} B.  Falls back to the oracle's
    fused FP64 port where the reference library was not built."""
    import numpy as np
    import oracle
    from mertens_bound_amd import formats
    L = oracle.load_ref()
    Vs = V[:n_c]
    if L is not None:
        xs, dy = formats.candidates_to_x_dy(Vs)
        xlo, xhi = formats.x_words(xs)
        chunks = []
        for z0 in range(0, tb.n, chunk):            # format conversion is not timed
            z1 = min(tb.n, z0 + chunk)
            sub = type(tb)(tb.G[z0:z1], tb.gamma[z0:z1], tb.w[z0:z1])
            wk = np.exp(-k * sub.gamma ** 2)
            chunks.append((z1 - z0, ref_words_fast(sub), wk, sub.gamma * wk))
        K = np.zeros(n_c); dK = np.zeros(n_c)
        evals = 0
        t0 = time.perf_counter()
        for n, rw, wk, gwk in chunks:
            A, B = oracle._call_cpab(L.compute_peak_AB, xlo, xhi, dy, rw["alpha"], rw["gamma"],
                                     rw["ghi_lo"], rw["ghi_hi"], rw["glo_lo"], rw["glo_hi"],
                                     rw["psi_lo"], rw["psi_hi"])
            K += A @ wk
            dK -= B @ gwk
            evals += n_c * n
            if time.perf_counter() - t0 > budget_s:
                break
        dt = time.perf_counter() - t0
        return {"value": evals / dt, "unit": "candidate-zero evals/s", "cores": 16,
                "kind": "reference",
                "sample": f"compute_peak_AB (reference, 16 threads) + weighted sums on {n_c} "
                          f"candidates x {evals // n_c} zeros in chunks of {chunk}, {dt:.1f} s"}
    nt = os.cpu_count() or 1
    t0 = time.perf_counter()
    oracle.eval_f64(tb, Vs, k, flags=oracle.VALUE | oracle.D1, nthreads=nt)
    dt = time.perf_counter() - t0
    return {"value": n_c * tb.n / dt, "unit": "candidate-zero evals/s", "cores": nt,
            "kind": "port", "sample": f"oracle FP64 port on {n_c} candidates x {tb.n} zeros, "
                                      f"{dt:.1f} s"}


def ref_words_fast(tb):
    """Reference-format words (vectorised; same values as formats.reference_words)."""
    import numpy as np
    lo = tb.G[:, 0].astype(np.uint64); hi = tb.G[:, 1].astype(np.uint64)
    s36 = np.uint64(36); s28 = np.uint64(28); s12 = np.uint64(12); s52 = np.uint64(52)
    ghi_lo = lo << s36
    ghi_hi = (hi << s36) | (lo >> s28)
    # glo = round(G / 2^12)
    add = np.uint64(1 << 11)
    lo2 = lo + add
    carry = (lo2 < lo).astype(np.uint64)
    hi2 = hi + carry
    glo_lo = (lo2 >> s12) | (hi2 << s52)
    glo_hi = hi2 >> s12
    frac = np.mod(-np.angle(tb.w) / (2 * np.pi), 1.0)
    m = np.round(frac * 2.0 ** 53).astype(np.uint64)            # 53-bit fraction
    psi_hi = m << np.uint64(11)                                  # m << 75 over 128 bits
    psi_lo = np.zeros_like(m)
    return {"gamma": tb.gamma, "alpha": np.abs(tb.w), "ghi_lo": ghi_lo, "ghi_hi": ghi_hi,
            "glo_lo": glo_lo, "glo_hi": glo_hi, "psi_lo": psi_lo, "psi_hi": psi_hi}


def cpu_info():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        d = dict(l.split(":", 1) for l in out.splitlines() if ":" in l)
        return {"model": d.get("Model name", "").strip(), "cpus": d.get("CPU(s)", "").strip()}
    except Exception:
        return {}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--form", choices=["plain", "mixed"], default="plain")
    ap.add_argument("--n-cand", type=int, default=65_536)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nc", type=int, default=0)
    ap.add_argument("--target-blocks", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the evaluator shards by candidates with no "
              "communication, and multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from mertens_bound_amd import evaluator as ev
    from mertens_bound_amd import synth

    c = synth.CONFIG2
    t_gen = time.perf_counter()
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    V = synth.make_candidates(args.n_cand, c["v_lo"], c["v_hi"], c["seed"])
    print(f"[bench] inputs generated in {time.perf_counter() - t_gen:.1f}s", file=sys.stderr,
          flush=True)
    torch.cuda.set_device(0)
    t = ev.K2Table(tb)
    if args.nc or args.target_blocks:
        t.set_tiling(nc=args.nc or None, target_blocks=args.target_blocks or None)
    Vd = ev.v_tensor(V)
    flags = ev.VALUE | ev.D1 | (ev.MIXED if args.form == "mixed" else 0)
    n = Vd.shape[0]
    out = tuple(torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3))
    stream = torch.cuda.current_stream()

    # accuracy gate: the first 256 candidates against the multiprecision oracle fixture
    acc = None
    gpath = os.path.join(ROOT, "tests", "golden", "config2_sub256.npz")
    if os.path.exists(gpath):
        g = np.load(gpath, allow_pickle=False)
        if np.array_equal(g["V"], V[:256]):
            K, dK, _ = t.eval(V[:256], c["k"], flags=flags)
            acc = float(np.max(np.abs(K - g["K"])))

    for _ in range(args.warmup):
        t.eval_device(Vd, c["k"], flags=flags, out=out, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    t.profile(True)
    e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        t.eval_device(Vd, c["k"], flags=flags, out=out, stream=stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = e0.elapsed_time(e1) / args.steps
    k_ms, k_n = t.kernel_time()
    t.profile(False)
    kern_ms = k_ms / max(1, k_n)
    evals = float(n) * tb.n
    value = evals / (ms * 1e-3)
    kern_rate = evals / (kern_ms * 1e-3)
    fp64_tflops = kern_rate * FP64_FLOPS_PER_EVAL / 1e12
    int_tops = kern_rate * INT32_MAD_PER_EVAL / 1e12
    # HBM bytes that must move per launch: the zero table records read by the kernel
    # (16 B phase word + 16 B weight + 16 B derivative weight per zero) plus the candidates and
    # the outputs
    alg_bytes = tb.n * 48 + n * (16 + 16)
    res = {
        "metric": "candidate-zero evals/sec (K2 + dK2/dv, 2.33M zeros) at |dK2|<=1e-9 vs oracle",
        "value": value,
        "unit": "candidate-zero evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "plain" else "f32+f64",
        "data": "synthetic (seeded RvM zeros, log-normal |zeta'|, uniform candidates)",
        "config": {"workload": "K2+dK2/dv, 65,536 candidates x 2,325,572 zeros, k=1.5e-10",
                   "form": args.form, "n_cand": n, "n_zeros": tb.n, "tiling": t.get_tiling()},
        "accuracy": {"max_abs_dK2_vs_oracle_256": acc, "bar": 1e-9,
                     "ok": (acc is not None and acc <= 1e-9)},
        "roofline": {"bound": "valu", "achieved": fp64_tflops, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": fp64_tflops / FP64_PEAK_TFLOPS,
                     "traffic": None,
                     "kernel_ms": kern_ms, "kernel_evals_per_s": kern_rate,
                     "int32_mad_tops": int_tops,
                     "algorithmic_bytes": alg_bytes,
                     "hbm_GBps_algorithmic": alg_bytes / (kern_ms * 1e-3) / 1e9,
                     "work_per_eval": {"fp64_flops": FP64_FLOPS_PER_EVAL,
                                       "int32_mad": INT32_MAD_PER_EVAL}},
        "wall_s": wall,
    }
    if not args.no_cpu:
        try:
            cb = cpu_baseline(tb, V, c["k"])
            cb["host"] = cpu_info()
            res["cpu_baseline"] = cb
        except Exception as exc:  # reported, never silently dropped
            res["cpu_baseline"] = {"error": repr(exc)}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
