"""Throughput of the K2 evaluator on one MI355X.

Default workload (one "step"): K2 and dK2/dv of 65,536 candidates v in [3.4e18, 3.6e18] (64.64
fixed point) against all 2,325,572 synthetic zeros (gamma < 1.3e6, Q18.110 words, complex128
weights), k = 1.5e-10, seed 0 -- 1.524e11 candidate-zero evaluations, plain FP64 form.  Inputs are
resident in HBM before the timed region.  The metric (candidate-zero evaluations per second)
counts only with |dK2| <= 1e-9 against the multiprecision oracle, checked on the committed
256-candidate fixture.

Other workloads (--workload): mixed (FP32 form of the same), batched (512 uneven requests in one
persistent launch, graph replay), polish (4,096 candidates, 8 Newton steps), small (10,000 zeros x
1,024 candidates), enclose (the default workload with rigorous radii), weights (five weight
vectors in one pass), screen (the staged 64/128/256-zero leaf screen on 2^24 candidates), copy (a
2 GiB device copy, the byte-counter calibration for rocprofv3 --pmc).

Prints one JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic work per candidate-zero evaluation, by instruction class (the frozen per-evaluation
# constants of the path's roofline; flops count an FMA as 2):
#   plain K2 + K2':        42 FP64 flops (sin/cos of the reduced phase, weighting, two sums)
#                          + 15 INT32 multiply-adds (the exact 128 x 128-bit phase window)
#   plain K2 + K2' + K2'': 46 FP64 flops + 15 INT32 multiply-adds
#   precision form:        2 transcendentals (sin, cos) + 8 FP32 flops (weight, compensated sum)
#                          + 15 INT32 multiply-adds per FP32-tail evaluation; its FP64 head at the
#                          plain constants
#   enclosure K2 + K2':    the plain work + 1.5 FP64 additions (TwoSum of the K2 and K2' group
#                          sums, one per 8 zeros)
#   leaf screen (one zero, 7 shifted sums): 15 INT32 multiply-adds + 2 transcendentals + 19 FP32
#                          flops (weight rotation 6, unshifted sum 1, per shift pair the two FMAs
#                          of Re q cos(m g d) and Im q sin(m g d), whose sum and difference are the
#                          pair's terms)
# roofline.frac is the largest class fraction (achieved class rate / class peak).  Class peaks:
# FP64 78.6 TFLOP/s and FP32 157.3 TFLOP/s (MI355X published); the INT32 multiply-add
# (v_mad_u64_u32) and transcendental (v_sin_f32) rates are measured on the same GPU in the same
# run by the library's microbenchmark (k2_valu_rates).
ALG = {"plain": {"fp64_flops": 42, "int32_mad": 15},
       "plain_d2": {"fp64_flops": 46, "int32_mad": 15},
       "mixed": {"trans": 2, "fp32_flops": 8, "int32_mad": 15},
       "enclose": {"fp64_flops": 43.5, "int32_mad": 15},
       "screen": {"int32_mad": 15, "trans": 2, "fp32_flops": 19}}
PUBLISHED_PEAKS = {"fp64_flops": 78.6e12, "fp32_flops": 157.3e12}
# measured once on gfx950 (k2_valu_rates, lane operations/s); used only where the live
# measurement fails
RECORDED_RATES = {"int32_mad": 33.0e12, "trans": 9.8e12}
# As-executed VALU issue model (not the roofline): SQ_INSTS_VALU per evaluation from the
# committed counter summaries (wave instructions x 64 / evaluations), each wave64 VALU
# instruction taking one 4-cycle issue slot of a SIMD: 39.3 T lane-instructions/s on 256 CUs at
# 2.4 GHz.  The VALU dual-issues some pairs (SQ_ACTIVE_INST_VALU2), so a loop can pass 1.
VALU_SLOT_PEAK = 39.3e12
VALU_INSTS = {"plain": 39.0, "mixed": 21.7, "enclose": 41.5, "screen": 29.7}
_RATES = {}


def class_peaks():
    """Published FP64/FP32 peaks and the INT32 multiply-add / transcendental rates measured on
    this GPU now (recorded values if the measurement fails)."""
    if not _RATES:
        peaks = dict(PUBLISHED_PEAKS)
        src = "measured (k2_valu_rates)"
        try:
            from mertens_bound_amd import evaluator as ev
            r = ev.valu_rates()
            peaks["int32_mad"] = r["int32_mad"]
            peaks["trans"] = r["trans"]
            peaks["measured_fp64_fma_lane_ops"] = r["fp64_fma"]
            peaks["measured_fp32_fma_lane_ops"] = r["fp32_fma"]
        except Exception as exc:   # reported in the line
            peaks.update(RECORDED_RATES)
            src = f"recorded (measurement failed: {exc!r})"
        peaks["source_int32_trans"] = src
        _RATES.update(peaks)
    return _RATES


def roofline(form, evals, kern_ms, alg_bytes, workload, head_frac=0.0):
    """Roofline of the dominant kernel: per-class achieved rates against the class peaks,
    frac = the largest.  head_frac: for the mixed form, the share of the evaluations in its
    FP64 head (charged at the plain constants)."""
    rate = evals / (kern_ms * 1e-3)
    peaks = class_peaks()
    if form == "mixed":
        work = {"fp64_flops": head_frac * ALG["plain"]["fp64_flops"],
                "int32_mad": ALG["mixed"]["int32_mad"],
                "trans": (1.0 - head_frac) * ALG["mixed"]["trans"],
                "fp32_flops": (1.0 - head_frac) * ALG["mixed"]["fp32_flops"]}
    else:
        work = dict(ALG[form])
    units = {"fp64_flops": "TFLOP/s", "fp32_flops": "TFLOP/s", "int32_mad": "T MAD/s",
             "trans": "T op/s"}
    classes = {}
    for cl, per in work.items():
        ach = rate * per
        classes[cl] = {"per_eval": per, "achieved": ach / 1e12, "peak": peaks[cl] / 1e12,
                       "unit": units[cl], "frac": ach / peaks[cl]}
    bind = max(classes, key=lambda c: classes[c]["frac"])
    b = classes[bind]
    insts = VALU_INSTS.get(form) if form != "mixed" else (
        head_frac * VALU_INSTS["plain"] + (1.0 - head_frac) * VALU_INSTS["mixed"])
    r = {"bound": "valu", "binding_class": bind, "achieved": b["achieved"], "peak": b["peak"],
         "unit": b["unit"], "frac": b["frac"], "traffic": traffic_summary(workload),
         "classes": classes,
         "peaks_source": {"fp64_flops": "published", "fp32_flops": "published",
                          "int32_mad": peaks["source_int32_trans"],
                          "trans": peaks["source_int32_trans"]},
         "kernel_ms": kern_ms, "kernel_evals_per_s": rate,
         "algorithmic_bytes": alg_bytes,
         "hbm_GBps_algorithmic": alg_bytes / (kern_ms * 1e-3) / 1e9,
         "valu_issue": {"sq_insts_valu_per_eval": insts,
                        "valu_issue_frac": (rate * insts / VALU_SLOT_PEAK
                                            if insts is not None else None),
                        "note": "issue model, not the roofline: VALU instructions per "
                                "evaluation (committed counters) against one wave64 "
                                "instruction per 4 cycles per SIMD, 39.3 T lane-instr/s"}}
    for k in ("measured_fp64_fma_lane_ops", "measured_fp32_fma_lane_ops"):
        if k in peaks:
            r.setdefault("measured_rates", {})[k] = peaks[k]
    return r


def cpu_baseline(tb, V, k, budget_s=15.0, n_c=256, chunk=1 << 17):
    """The reference evaluator (oracle/_ref, compiled from search/path1_eval_fast.cpp; 16 OpenMP
    threads, hard-coded there) on a bounded sample: batches of n_c candidates against the whole
    table until about budget_s seconds have run -- compute_peak_AB over zero chunks plus the
    weighted sums K = sum e^{-k g^2} A and K' = -sum g e^{-k g^2} B.  Beside it: the oracle's
    fused FP64 port of the same arithmetic on n_c candidates x the whole table at all physical
    cores the process may use (lscpu, capped at its CPU share) and at 16 threads, and both
    evaluators on the 1,024 x 10,000 workload in full.  Falls back to the port as the headline
    where the reference library was not built."""
    import numpy as np
    import oracle
    from mertens_bound_amd import formats, synth
    info = cpu_info()
    threads = info.get("threads_usable") or 16
    out = {"host": info}
    L = oracle.load_ref()
    runs = []
    # the fused port (libm sin/cos, exact phase, fused sums), n_c candidates x all zeros
    for nt in sorted({threads, 16}):
        t0 = time.perf_counter()
        oracle.eval_f64(tb, V[:n_c], k, flags=oracle.VALUE | oracle.D1, nthreads=nt)
        dt = time.perf_counter() - t0
        runs.append({"kind": "port", "threads": nt, "workload": f"{n_c} x {tb.n}",
                     "value": n_c * tb.n / dt, "s": dt})
    # the 1,024 x 10,000 workload in full
    c1 = synth.W1024x10000
    tb1 = synth.make_zero_table(c1["n_zeros"], c1["seed"])
    V1 = synth.make_candidates(c1["n_cand"], c1["v_lo"], c1["v_hi"], c1["seed"])
    for nt in sorted({threads, 16}):
        t0 = time.perf_counter()
        oracle.eval_f64(tb1, V1, c1["k"], flags=oracle.VALUE | oracle.D1, nthreads=nt)
        dt = time.perf_counter() - t0
        runs.append({"kind": "port", "threads": nt, "workload": f"{tb1.n} zeros x {V1.shape[0]}",
                     "value": V1.shape[0] * tb1.n / dt, "s": dt})
    if L is not None:
        xs, dy = formats.candidates_to_x_dy(V1)
        xlo, xhi = formats.x_words(xs)
        rw = ref_words_fast(tb1)
        wk = np.exp(-c1["k"] * tb1.gamma ** 2)
        t0 = time.perf_counter()
        A, B = oracle._call_cpab(L.compute_peak_AB, xlo, xhi, dy, rw["alpha"], rw["gamma"],
                                 rw["ghi_lo"], rw["ghi_hi"], rw["glo_lo"], rw["glo_hi"],
                                 rw["psi_lo"], rw["psi_hi"])
        _ = A @ wk, B @ (tb1.gamma * wk)
        dt = time.perf_counter() - t0
        runs.append({"kind": "reference", "threads": 16,
                     "workload": f"{tb1.n} zeros x {V1.shape[0]}",
                     "value": V1.shape[0] * tb1.n / dt, "s": dt})
        # compute_peak_AB alone at its published size (path1_report.txt:34: 0.737 s)
        sub, a, _ = compat_inputs(tb)
        t0 = time.perf_counter()
        A, B = oracle._call_cpab(L.compute_peak_AB, *a)
        dt = time.perf_counter() - t0
        del A, B
        runs.append({"kind": "reference", "threads": 16,
                     "workload": f"compute_peak_AB {a[0].shape[0]} peaks x {sub.n} zeros "
                                 f"(published 0.737 s)",
                     "value": a[0].shape[0] * sub.n / dt, "s": dt})
    out["runs"] = runs
    if L is not None:
        batches = []                                 # candidate words: not timed
        for b0 in range(0, min(V.shape[0], 64 * n_c), n_c):
            xs, dy = formats.candidates_to_x_dy(V[b0:b0 + n_c])
            xlo, xhi = formats.x_words(xs)
            batches.append((xlo, xhi, dy))
        chunks = []
        for z0 in range(0, tb.n, chunk):            # format conversion is not timed
            z1 = min(tb.n, z0 + chunk)
            sub = type(tb)(tb.G[z0:z1], tb.gamma[z0:z1], tb.w[z0:z1])
            wk = np.exp(-k * sub.gamma ** 2)
            chunks.append((z1 - z0, ref_words_fast(sub), wk, sub.gamma * wk))
        evals = 0
        done = 0
        t0 = time.perf_counter()
        for xlo, xhi, dy in batches:
            nb = xlo.shape[0]
            K = np.zeros(nb); dK = np.zeros(nb)
            for n, rw, wk, gwk in chunks:
                A, B = oracle._call_cpab(L.compute_peak_AB, xlo, xhi, dy, rw["alpha"],
                                         rw["gamma"], rw["ghi_lo"], rw["ghi_hi"], rw["glo_lo"],
                                         rw["glo_hi"], rw["psi_lo"], rw["psi_hi"])
                K += A @ wk
                dK -= B @ gwk
                evals += nb * n
            done += nb
            if time.perf_counter() - t0 > budget_s:
                break
        dt = time.perf_counter() - t0
        out.update({"value": evals / dt, "unit": "candidate-zero evals/s", "cores": 16,
                    "kind": "reference",
                    "sample": f"compute_peak_AB (reference, 16 threads) + weighted sums: {done} "
                              f"candidates x all {tb.n} zeros in zero chunks of {chunk}, "
                              f"{dt:.1f} s"})
        return out
    best = max((r for r in runs if r["workload"] == f"{n_c} x {tb.n}"), key=lambda r: r["value"])
    out.update({"value": best["value"], "unit": "candidate-zero evals/s",
                "cores": best["threads"], "kind": "port",
                "sample": f"oracle FP64 port on {n_c} candidates x {tb.n} zeros, "
                          f"{best['s']:.1f} s"})
    return out


def ref_words_fast(tb):
    """Reference-format words (vectorised; same values as formats.reference_words)."""
    import numpy as np
    lo = tb.G[:, 0].astype(np.uint64); hi = tb.G[:, 1].astype(np.uint64)
    s36 = np.uint64(36); s28 = np.uint64(28); s12 = np.uint64(12); s52 = np.uint64(52)
    ghi_lo = lo << s36
    ghi_hi = (hi << s36) | (lo >> s28)
    add = np.uint64(1 << 11)                 # glo = round(G / 2^12)
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


def cpu_share():
    """CPUs this process may use: its affinity mask, capped by a cgroup CPU quota if one is set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            n = min(n, max(1, int(int(q) // int(per))))
    except Exception:
        pass
    return n


def cpu_info():
    """CPU model and physical core count (lscpu) and the threads this process can use."""
    d = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        d = dict((l.split(":", 1)[0].strip(), l.split(":", 1)[1].strip())
                 for l in out.splitlines() if ":" in l)
    except Exception:
        pass
    phys = None
    try:
        phys = int(d.get("Core(s) per socket", "0")) * int(d.get("Socket(s)", "0")) or None
    except ValueError:
        pass
    share = cpu_share()
    return {"model": d.get("Model name", ""), "logical_cpus": d.get("CPU(s)", ""),
            "physical_cores": phys, "cpu_share": share,
            "threads_usable": min(phys, share) if phys else share}


def traffic_summary(workload):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE scaled by a copy kernel of known size), or None."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        d = json.load(open(p))
        return d.get(workload, {}).get("hbm_bytes_per_launch")
    except Exception:
        return None


def timed(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    e0.record(st)
    for _ in range(steps):
        fn()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps, time.perf_counter() - w0


def base_record(args, metric, value, unit, ms, dtype, config):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic (seeded RvM zeros, log-normal |zeta'|, uniform candidates)",
            "config": config}


def run_eval(args, form):
    import numpy as np
    import torch
    from mertens_bound_amd import autotune
    from mertens_bound_amd import evaluator as ev
    from mertens_bound_amd import synth
    c = synth.W1024x10000 if args.workload == "small" else synth.W65536x2325572
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    n_cand = args.n_cand or c["n_cand"]
    V = synth.make_candidates(n_cand, c["v_lo"], c["v_hi"], c["seed"])
    torch.cuda.set_device(0)
    t = ev.K2Table(tb)
    if args.nc or args.target_blocks:
        t.set_tiling(nc=args.nc or None, target_blocks=args.target_blocks or None)
    else:
        autotune.apply_tuned(t, n_cand, tb.n, "mixed" if form == "mixed" else "plain")
    Vd = ev.v_tensor(V)
    flags = ev.VALUE | ev.D1 | (ev.MIXED if form == "mixed" else 0)
    out = tuple(torch.empty(n_cand, dtype=torch.float64, device="cuda") for _ in range(3))
    st = torch.cuda.current_stream().cuda_stream

    acc = None
    gname = "w1024x10000_oracle.npz" if args.workload == "small" else "w65536x2325572_sub256.npz"
    gpath = os.path.join(ROOT, "tests", "golden", gname)
    if os.path.exists(gpath) and not args.no_check:
        g = np.load(gpath, allow_pickle=False)
        m = min(len(g["K"]), n_cand)
        if np.array_equal(g["V"][:m], V[:m]):
            K, _, _ = t.eval(V[:m], c["k"], flags=flags)
            acc = float(np.max(np.abs(K - g["K"][:m])))

    def step():
        t.eval_device(Vd, c["k"], flags=flags, out=out, stream=st)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ms, wall = timed(step, args.steps, 0, torch)
    # the kernels' own average from a second, profiled pass (per-launch events cost microseconds,
    # which would show in the timed region of the small workloads)
    t.profile(True)
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    k_ms, k_n = t.kernel_time()
    t.profile(False)
    evals = float(n_cand) * tb.n
    alg_bytes = tb.n * 48 + n_cand * (16 + 16)     # zero records read + candidates + outputs
    res = base_record(
        args, "candidate-zero evals/sec (K2 + dK2/dv, 2.33M zeros) at |dK2|<=1e-9 vs oracle",
        evals / (ms * 1e-3), "candidate-zero evals/s", ms,
        "f64" if form == "plain" else "f32+f64",
        {"workload": f"K2+dK2/dv, {n_cand:,} candidates x {tb.n:,} zeros, k={c['k']}",
         "form": form, "n_cand": n_cand, "n_zeros": tb.n, "tiling": t.get_tiling(),
         "fp64_head_zeros": t.mixed_head(c["k"]) if form == "mixed" else tb.n})
    res["accuracy"] = {"max_abs_dK2_vs_oracle": acc, "oracle_sample": gname, "bar": 1e-9,
                       "ok": (acc is not None and acc <= 1e-9)}
    head = res["config"]["fp64_head_zeros"] / tb.n if form == "mixed" else 0.0
    res["roofline"] = roofline(form, evals, k_ms / max(1, k_n), alg_bytes,
                               f"{args.workload}_{form}", head)
    res["wall_s"] = wall
    if form == "mixed" and not args.no_check:
        # the FP64 kernel on the same inputs and GPU: speed-up and |dK2| for every candidate
        out64 = tuple(torch.empty(n_cand, dtype=torch.float64, device="cuda") for _ in range(3))
        t64 = ev.K2Table(tb)
        if not (args.nc or args.target_blocks):
            autotune.apply_tuned(t64, n_cand, tb.n, "plain")

        def step64():
            t64.eval_device(Vd, c["k"], flags=ev.VALUE | ev.D1, out=out64, stream=st)

        ms64, _ = timed(step64, args.steps, args.warmup, torch)
        step()
        torch.cuda.synchronize()
        d = (out[0] - out64[0]).abs()
        res["vs_plain"] = {"plain_ms_per_step": ms64, "speedup": ms64 / ms,
                           "max_abs_dK2_vs_fp64_kernel_all_candidates": float(d.max().item()),
                           "mean_abs_dK2_vs_fp64_kernel": float(d.mean().item())}
    if form == "plain" and args.workload == "metric" and not args.no_check:
        # the precision form on the same resident inputs, timed the same way: speed-up over the
        # FP64 line above and its accuracy (oracle sample and every candidate vs FP64)
        tm = ev.K2Table(tb)
        autotune.apply_tuned(tm, n_cand, tb.n, "mixed")
        outm = tuple(torch.empty(n_cand, dtype=torch.float64, device="cuda") for _ in range(3))
        flm = ev.VALUE | ev.D1 | ev.MIXED

        def stepm():
            tm.eval_device(Vd, c["k"], flags=flm, out=outm, stream=st)

        msm, _ = timed(stepm, args.steps, args.warmup, torch)
        step()
        stepm()
        torch.cuda.synchronize()
        d = (outm[0] - out[0]).abs()
        accm = None
        if os.path.exists(gpath):
            g = np.load(gpath, allow_pickle=False)
            m = min(len(g["K"]), n_cand)
            if np.array_equal(g["V"][:m], V[:m]):
                accm = float(np.max(np.abs(outm[0][:m].cpu().numpy() - g["K"][:m])))
        res["mixed_form"] = {"value": evals / (msm * 1e-3), "ms_per_step": msm,
                             "speedup_vs_fp64": ms / msm, "dtype": "f32+f64",
                             "fp64_head_zeros": tm.mixed_head(c["k"]),
                             "max_abs_dK2_vs_oracle": accm,
                             "max_abs_dK2_vs_fp64_all_candidates": float(d.max().item())}
    if form == "plain" and args.workload == "metric" and not args.no_check:
        # the other full-size workloads on the same table, each timed the same way with its own
        # accuracy check and roofline: 512 batched requests and the 4,096-candidate polish in
        # both forms, and the leaf screen
        ctx = (tb, ev.K2Table(tb))          # the library's default tiling, as standalone
        bp = run_batched(args, "plain", ctx)
        bm = run_batched(args, "mixed", ctx)
        res["batched"] = {"plain": sub_record(bp), "mixed": sub_record(bm),
                          "mixed_speedup_vs_batched_plain": bp["ms_per_step"] / bm["ms_per_step"],
                          "mixed_evals_per_s_vs_plain_fp64_kernel": bm["value"] / res["value"]}
        pp = run_polish(args, "plain", ctx)
        pm = run_polish(args, "mixed", ctx)
        res["polish"] = {"plain": sub_record(pp), "mixed": sub_record(pm),
                         "mixed_speedup": pp["ms_per_step"] / pm["ms_per_step"]}
        res["screen"] = sub_record(run_screen(args, ctx))
        res["screen_survivors"] = sub_record(run_screen(args, ctx, peaks=True))
        res["compat"] = run_compat(args, ctx)
        res["weights"] = sub_record(run_weights(args, ctx),
                                    keep=("value", "unit", "ms_per_step", "config", "accuracy",
                                          "vs_separate_passes", "wall_s"))
        res["enclose"] = sub_record(run_enclose(args, ctx),
                                    keep=("value", "unit", "ms_per_step", "config", "accuracy",
                                          "radius_max_all", "roofline", "wall_s"))
        # the CPU-runnable case (1,024 candidates x 10,000 zeros) in both forms: every
        # candidate against binary128 (plain) and against the FP64 kernel (precision form)
        a1 = argparse.Namespace(**vars(args))
        a1.workload, a1.n_cand, a1.steps, a1.warmup = "small", 0, max(args.steps, 50), 5
        keep1 = ("value", "unit", "ms_per_step", "dtype", "config", "accuracy", "roofline",
                 "vs_plain", "wall_s")
        s1p = sub_record(run_eval(a1, "plain"), keep=keep1)
        s1m = sub_record(run_eval(a1, "mixed"), keep=keep1)
        res["config1"] = {"plain": s1p, "mixed": s1m,
                          "mixed_speedup": s1p["ms_per_step"] / s1m["ms_per_step"]}
    if not args.no_cpu and args.workload == "metric":
        try:
            cb = cpu_baseline(tb, V, c["k"])
            res["cpu_baseline"] = cb
        except Exception as exc:  # reported, never silently dropped
            res["cpu_baseline"] = {"error": repr(exc)}
    return res


def full_table(ctx=None):
    """(table, K2Table) of the 2,325,572-zero workload, reused when the caller built them."""
    import torch
    from mertens_bound_amd import evaluator as ev
    from mertens_bound_amd import synth
    if ctx is not None:
        return ctx
    c = synth.W65536x2325572
    tb = synth.make_zero_table(c["n_zeros"], c["seed"])
    torch.cuda.set_device(0)
    return tb, ev.K2Table(tb)


def golden_npz(name):
    """A committed fixture under tests/golden (data only), or None."""
    import numpy as np
    p = os.path.join(ROOT, "tests", "golden", name)
    return np.load(p, allow_pickle=False) if os.path.exists(p) else None


def run_batched(args, form=None, ctx=None):
    """512 uneven requests (cutoffs, k, counts) in one persistent launch replayed from its HIP
    graph; accuracy on the committed binary128 sample (batch512_sample.npz) and against the
    per-request loop of one-launch calls (search/eval_tech1_600k.py:252-267)."""
    import numpy as np
    import torch
    from mertens_bound_amd import autotune
    from mertens_bound_amd import evaluator as ev
    from mertens_bound_amd import synth
    tb, t = full_table(ctx)
    nz, ks, cnt, offs, V = synth.make_batched_requests(tb, n_req=512, seed=0)
    reqs = [(int(a), int(b), int(o), float(k)) for a, b, o, k in zip(nz, cnt, offs, ks)]
    Vt = ev.v_tensor(V)
    form = form or ("mixed" if args.form == "mixed" else "plain")
    flags = ev.VALUE | ev.D1 | (ev.MIXED if form == "mixed" else 0)
    b = ev.K2Batch(t, reqs, Vt, flags=flags, tiling=autotune.batch_tiling(form))
    useful = float(np.sum(cnt.astype(float) * nz))
    ms, wall = timed(b.run, args.steps, args.warmup, torch)
    # the same requests one launch each, for comparison
    out = tuple(torch.zeros(V.shape[0], dtype=torch.float64, device="cuda") for _ in range(3))

    def loop():
        for a, n, o, k in reqs:
            t.eval_device(Vt[o:o + n], k, n_zeros=a, flags=flags,
                          out=(out[0][o:o + n], out[1][o:o + n], out[2][o:o + n]))

    ms_loop, _ = timed(loop, 1, 1, torch)
    b.run()
    torch.cuda.synchronize()
    d = (out[0] - b.K).abs().max().item()
    info = b.info()
    res = base_record(args, "candidate-zero evals/sec, 512 uneven requests (one graph launch)",
                      useful / (ms * 1e-3), "candidate-zero evals/s", ms,
                      "f64" if form == "plain" else "f32+f64",
                      {"workload": "512 requests: T log-uniform [1e5,1.3e6], k in "
                                   "{3e-9,1e-9,1.5e-10}, counts log-uniform [1,4096], 2.33M-zero "
                                   "table", "form": form, "batch": info})
    res["per_request_loop"] = {"ms": ms_loop, "evals_per_s": useful / (ms_loop * 1e-3),
                               "max_abs_diff_vs_batch": d}
    g = None if args.no_check else golden_npz("batch512_sample.npz")
    if g is not None and np.array_equal(V[g["idx"]], g["V"]):
        K = b.K.cpu().numpy()[g["idx"]]
        dK = b.dK.cpu().numpy()[g["idx"]]
        eK = float(np.max(np.abs(K - g["K"])))
        res["accuracy"] = {"max_abs_dK2_vs_oracle": eK,
                           "max_dK2prime_vs_oracle_over_scale":
                               float(np.max(np.abs(dK - g["dK"]) / g["gabsS"])),
                           "oracle_sample": f"batch512_sample.npz ({len(g['idx'])} candidates of "
                                            f"32 requests, binary128)",
                           "bar": 1e-9, "ok": eK <= 1e-9}
    head = 0.0
    if form == "mixed":   # share of the useful evaluations in the requests' FP64 heads
        head = sum(float(n) * t.mixed_head(k, a) for a, n, _, k in reqs) / useful
    res["roofline"] = roofline(form, useful, ms, tb.n * 48 * 3 + V.shape[0] * 32,
                               f"batched_{form}", head)
    res["wall_s"] = wall
    b.close()
    return res


def run_polish(args, form=None, ctx=None):
    """4,096 candidates near synthetic peaks, 8 Newton steps on the exact 64.64 points (the step
    rule of search/path1_eval_fast.cpp:103-141), K2/K2'/K2'' passes on the device; accuracy on the
    committed sample of the oracle's own polish (polish4096_sample.npz, binary128 values)."""
    import numpy as np
    import torch
    from mertens_bound_amd import autograd as ag
    from mertens_bound_amd import evaluator as ev
    from mertens_bound_amd import synth, workloads
    tb, t = full_table(ctx)
    c = synth.W65536x2325572
    V0 = workloads.polish_candidates(t, 4096, seed=0)
    V0t = ev.v_tensor(V0)
    k = c["k"]
    mixed = (form or args.form) == "mixed"
    # the tuned tiling of the polish's passes (the table may be shared with other sub-records)
    from mertens_bound_amd import autotune
    prev_tiling = t.get_tiling()
    ti = autotune.polish_tiling("mixed" if mixed else "plain")
    if ti:
        t.set_tiling(nc=ti["nc"], target_blocks=ti["target_blocks"],
                     wg_threads=ti.get("wg_threads", 256))

    def step():
        ag.polish(t, V0t, k, n_steps=8, mixed=mixed)

    ms, wall = timed(step, args.steps, args.warmup, torch)
    V, K, dK, steps = ag.polish(t, V0t, k, n_steps=8, mixed=mixed)
    torch.cuda.synchronize()
    # candidate evaluations per polish: the Newton passes (the first, over all candidates, also
    # gives the start values; converged candidates drop out after each step) and the final value
    cand_evals = t.polish_stats()
    evals = float(cand_evals) * tb.n
    res = base_record(args, "candidate-zero evals/sec, Newton polish (K2, K2', K2'')",
                      evals / (ms * 1e-3), "candidate-zero evals/s", ms,
                      "f32+f64" if mixed else "f64",
                      {"workload": "4,096 candidates near synthetic peaks (v0 + N(0,1e-3)) x "
                                   "2,325,572 zeros, 8 Newton steps",
                       "form": "mixed" if mixed else "plain",
                       "candidate_evals_per_polish": cand_evals,
                       "passes_equivalent": cand_evals / V0.shape[0]})
    st = steps.cpu().numpy()
    res["polish"] = {"steps_histogram": np.bincount(st, minlength=9).tolist(),
                     "max_abs_dK_final": float(dK.abs().max().item()),
                     "mean_abs_K": float(K.abs().mean().item())}
    g = None if args.no_check else golden_npz("polish4096_sample.npz")
    if g is not None and np.array_equal(V0[g["idx"]], g["V0"]):
        idx = g["idx"]
        Vn = ev.v_from_tensor(V)[idx]
        Kn = K.cpu().numpy()[idx]
        dv = np.array([abs(a - b) / 2.0 ** 64 for a, b in
                       zip(synth.v_to_int(Vn), synth.v_to_int(g["V"]))])
        conv = (g["steps"] < 8) & (st[idx] < 8)
        eK = np.abs(Kn - g["K"])
        acc = {"oracle_sample": f"polish4096_sample.npz ({len(idx)} candidates: the oracle's FP64 "
                                f"polish, binary128 values at its end points)",
               "converged_in_both": int(conv.sum()),
               "max_abs_dK2_converged": float(eK[conv].max()) if conv.any() else None,
               "max_dv_converged": float(dv[conv].max()) if conv.any() else None,
               "max_abs_dK2_all": float(eK.max()), "max_dv_all": float(dv.max()), "bar": 1e-9}
        if not mixed:
            acc["same_steps_as_oracle"] = bool(np.array_equal(st[idx], g["steps"]))
            acc["ok"] = bool(eK.max() <= 1e-9 and acc["same_steps_as_oracle"])
        else:
            acc["ok"] = bool(conv.any() and eK[conv].max() <= 1e-9)
            acc["note"] = ("candidates still moving after 8 steps end at other points in the "
                           "precision form (its steps follow its own K2'), so only candidates "
                           "converged in both forms are compared with the oracle's end points")
        res["accuracy"] = acc
    res["roofline"] = roofline("plain_d2" if not mixed else "mixed", evals, ms,
                               tb.n * 64 + V0.shape[0] * 40, "polish_" + ("mixed" if mixed else
                                                                          "plain"),
                               t.mixed_head(k) / tb.n if mixed else 0.0)
    res["config"]["tiling"] = t.get_tiling()
    res["wall_s"] = wall
    t.set_tiling(nc=prev_tiling["nc"], target_blocks=prev_tiling["target_blocks"],
                 wg_threads=prev_tiling["wg_threads"])
    return res


def compat_inputs(tb, n_peaks=1919, n_zeros=100_000, seed=0):
    """Reference-format inputs of the published compute_peak_AB size
    (search/results/path1_report.txt:34): the first n_zeros zeros of the synthetic table as
    (alpha, gamma, ghi, glo, psi) words and n_peaks candidates as (x, dy0) with sub-grid offsets
    of the size the reference's CLUSTER lines carry (|dy0| <= 4e-3).  Returns (sub-table,
    argument tuple after n_peaks/n_zeros, V of the points x / 2^30 + dy0 as 64.64 words)."""
    import numpy as np
    from mertens_bound_amd import formats, synth
    c = synth.W65536x2325572
    sub = type(tb)(tb.G[:n_zeros], tb.gamma[:n_zeros], tb.w[:n_zeros])
    rw = ref_words_fast(sub)
    V = synth.make_candidates(n_peaks, c["v_lo"], c["v_hi"], seed + 1919)
    xs, dy = formats.candidates_to_x_dy(V)
    dy = dy + np.random.default_rng(seed).uniform(-4e-3, 4e-3, n_peaks)
    xlo, xhi = formats.x_words(xs)
    Vd = synth.candidates_from_x_dy(xs, dy)
    return sub, (xlo, xhi, dy, rw["alpha"], rw["gamma"], rw["ghi_lo"], rw["ghi_hi"],
                 rw["glo_lo"], rw["glo_hi"], rw["psi_lo"], rw["psi_hi"]), Vd


def run_compat(args, ctx=None):
    """The reference drivers' ctypes sequence on libk2eval.so at the published size: one
    compute_peak_AB of 1,919 peaks x 100,000 zeros (A and B, 3.07 GB, back in host memory), then
    polish_peaks_with_weights per weight vector for the five vectors of
    search/eval_tech1_600k.py:243-267.  The reference took 0.737 s for that compute_peak_AB
    (search/results/path1_report.txt:34, 16 threads)."""
    import numpy as np
    import torch
    from mertens_bound_amd import evaluator as ev
    from mertens_bound_amd import pipeline
    tb, t = full_table(ctx)
    sub, a, Vd = compat_inputs(tb)
    npk, nz = a[0].shape[0], sub.n
    A = np.empty((npk, nz)); B = np.empty((npk, nz))
    reps = max(1, min(args.steps, 3))
    ev.compute_peak_AB(npk, *a[:3], nz, *a[3:], A, B)          # first call: buffers and stream
    t0 = time.perf_counter()
    for _ in range(reps):
        ev.compute_peak_AB(npk, *a[:3], nz, *a[3:], A, B)
    cpab_s = (time.perf_counter() - t0) / reps
    names, wt = pipeline.reference_weight_vectors(sub.gamma)
    hu = np.empty(npk); hp = np.empty(npk); dd = np.empty(npk)
    t0 = time.perf_counter()
    for m in range(len(names)):
        ev.polish_peaks_with_weights(npk, nz, A, B, sub.gamma, wt[m], 2.5e-4, 40, hu, hp, dd)
    pol_s = time.perf_counter() - t0
    # A/B rows weighted by the Gaussian against the fused evaluator at the same points (the
    # reference words round gamma/2pi and arg w differently: agreement to about 1e-11)
    k = 1.5e-10
    wk = np.exp(-k * sub.gamma ** 2)
    t2 = ev.K2Table(sub)
    K, dK, _ = t2.eval(Vd[:64], k, flags=ev.VALUE | ev.D1)
    err = float(np.max(np.abs(A[:64] @ wk - K)))
    res = {"value": npk * nz / cpab_s, "unit": "candidate-zero evals/s (A, B in host memory)",
           "compute_peak_AB_s": cpab_s, "reference_published_s": 0.737,
           "speedup_vs_published": 0.737 / cpab_s,
           "polish_five_vectors_s": pol_s,
           "config": {"workload": f"compute_peak_AB {npk:,} peaks x {nz:,} zeros + "
                                  f"polish_peaks_with_weights x {len(names)} weight vectors",
                      "reference": "search/results/path1_report.txt:34"},
           "accuracy": {"max_abs_dK2_rowsum_vs_fused_evaluator": err, "sample": 64,
                        "bar": 1e-9, "ok": err <= 1e-9},
           "note": "A and B (16 B per evaluation) cross PCIe into pageable host memory: the "
                   "time is the transfer's, not the kernel's"}
    del A, B
    return res


def sub_record(res, keep=("value", "unit", "ms_per_step", "dtype", "config", "accuracy",
                          "roofline", "per_request_loop", "polish", "screen", "wall_s")):
    """The fields of a workload's own line that go into the default line as a sub-record."""
    return {k: res[k] for k in keep if k in res}


def run_enclose(args, ctx=None):
    """Enclosure mode (midpoint + rigorous radius of K2 and K2') on the metric workload."""
    import numpy as np
    import torch
    from mertens_bound_amd import autotune
    from mertens_bound_amd import evaluator as ev
    from mertens_bound_amd import synth
    c = synth.W65536x2325572
    tb, _ = full_table(ctx)
    n_cand = args.n_cand if (args.n_cand and ctx is None) else c["n_cand"]
    V = synth.make_candidates(n_cand, c["v_lo"], c["v_hi"], c["seed"])
    t = ev.K2Table(tb)
    autotune.apply_tuned(t, n_cand, tb.n, "plain")
    Vd = ev.v_tensor(V)
    acc = None
    gpath = os.path.join(ROOT, "tests", "golden", "w65536x2325572_sub256.npz")
    if os.path.exists(gpath) and not args.no_check:
        g = np.load(gpath, allow_pickle=False)
        K, Kr, dK, dKr, _ = t.enclose(g["V"], c["k"])
        inside = np.abs(K - g["K"]) <= Kr + 2.0 ** -53 * np.abs(g["K"])
        acc = {"max_abs_dK2_vs_oracle": float(np.max(np.abs(K - g["K"]))),
               "max_radius": float(Kr.max()), "max_dK_radius": float(dKr.max()),
               "oracle_inside_all": bool(inside.all()), "oracle_sample": "w65536x2325572_sub256.npz",
               "bar": 1e-9, "ok": bool(inside.all() and np.max(np.abs(K - g["K"])) <= 1e-9)}
    state = {}

    def step():
        state["r"] = t.enclose_device(Vd, c["k"])

    ms, wall = timed(step, args.steps, args.warmup, torch)
    K, Kr = state["r"][0], state["r"][1]
    evals = float(n_cand) * tb.n
    alg_bytes = tb.n * (16 + 32) + n_cand * (16 + 32)
    res = base_record(
        args, "candidate-zero evals/sec (K2 + dK2/dv enclosure, 2.33M zeros)",
        evals / (ms * 1e-3), "candidate-zero evals/s", ms, "f64",
        {"workload": f"K2+dK2/dv with radii, {n_cand:,} candidates x {tb.n:,} zeros, "
                     f"k={c['k']}", "form": "enclose", "n_cand": n_cand, "n_zeros": tb.n,
         "tiling": t.get_tiling()})
    res["accuracy"] = acc
    res["radius_max_all"] = float(Kr.max().item())
    res["roofline"] = roofline("enclose", evals, ms, alg_bytes, "enclose")
    res["wall_s"] = wall
    return res


def run_weights(args, ctx=None):
    """The reference's five weight vectors (JP kernel, Gaussians at 1.5e-10 and 1e-10, raw,
    first 10,000 zeros) for the metric's candidates and table in one pass
    (k2_eval_weights_device), against five one-vector passes of the same kernel.  Accuracy: the
    Gaussian row at k = 1.5e-10 is K2 itself, checked on the committed binary128 sample of the
    same candidates (w65536x2325572_sub256.npz)."""
    import numpy as np
    import torch
    from mertens_bound_amd import evaluator as ev
    from mertens_bound_amd import pipeline, synth
    c = synth.W65536x2325572
    tb, _ = full_table(ctx)
    n_cand = args.n_cand if (args.n_cand and ctx is None) else c["n_cand"]
    V = synth.make_candidates(n_cand, c["v_lo"], c["v_hi"], c["seed"])
    t = ev.K2Table(tb)
    t.set_tiling(target_blocks=49152)
    Vd = ev.v_tensor(V)
    names, wt = pipeline.reference_weight_vectors(tb.gamma)
    wtd = torch.tensor(wt, device="cuda")
    ms, wall = timed(lambda: t.eval_weights_device(Vd, wtd, flags=ev.VALUE), args.steps,
                     args.warmup, torch)
    ms1, _ = timed(lambda: [t.eval_weights_device(Vd, wtd[m:m + 1], flags=ev.VALUE)
                            for m in range(len(names))], max(1, args.steps // 2), 1, torch)
    evals = float(n_cand) * tb.n * len(names)
    res = base_record(args, "candidate-zero-weight evals/sec (5 weight vectors, 2.33M zeros)",
                      evals / (ms * 1e-3), "candidate-zero-vector evals/s", ms, "f64",
                      {"workload": f"K2 for {len(names)} weight vectors ({', '.join(names)}), "
                                   f"{n_cand:,} candidates x {tb.n:,} zeros", "form": "weights"})
    res["vs_separate_passes"] = {"ms": ms1, "speedup": ms1 / ms}
    g = None if args.no_check else golden_npz("w65536x2325572_sub256.npz")
    if g is not None and np.array_equal(V[:g["V"].shape[0]], g["V"]):
        m = g["V"].shape[0]
        K = t.eval_weights_device(Vd[:m], wtd, flags=ev.VALUE)[0][names.index("g15")]
        e = float(np.max(np.abs(K.cpu().numpy() - g["K"])))
        res["accuracy"] = {"max_abs_dK2_gaussian_row_vs_oracle": e,
                           "oracle_sample": "w65536x2325572_sub256.npz (binary128, k = 1.5e-10)",
                           "bar": 1e-9, "ok": e <= 1e-9}
    res["wall_s"] = wall
    return res


def screen_peak_candidates(t, n_cand, seed=0, every=16, sigma=3e-4):
    """2^24-style screen input with survivors: uniform candidates in [3.4e18, 3.6e18], every
    `every`-th replaced by a point near one of the 1,024 best of 2^18 uniform candidates on the
    256-zero stage value (moved to its best shift, then v + N(0, sigma))."""
    import numpy as np
    import torch
    from mertens_bound_amd import autograd as ag
    from mertens_bound_amd import evaluator as ev
    from mertens_bound_amd import synth
    c = synth.W65536x2325572
    pool = ev.v_tensor(synth.make_candidates(1 << 18, c["v_lo"], c["v_hi"], seed + 5))
    val, best, _ = t.screen_device(pool, 3e-9, 0.0014, 3, ((64, 0.0), (128, 0.0), (256, 0.0)))
    top = torch.argsort(-val[:, 2], stable=True)[:1024]
    peaks = ag.offset_points(pool[top], best[top].to(torch.float64) * 0.0014)
    V = ev.v_tensor(synth.make_candidates(n_cand, c["v_lo"], c["v_hi"], c["seed"]))
    pos = torch.arange(0, n_cand, every, device=V.device)
    rng = np.random.default_rng(seed)
    off = torch.tensor(rng.normal(0.0, sigma, pos.numel()), device=V.device)
    V[pos] = ag.offset_points(peaks[torch.arange(pos.numel(), device=V.device) % 1024], off)
    torch.cuda.synchronize()
    return V


def run_screen(args, ctx=None, peaks=False):
    """The staged leaf screen (k2_screen_device) of search/gpu_mertens_stream.cu:186-250 with its
    constants: 64/128/256 zeros, thresholds 0.78/0.855/0.892, 7 shifts of 0.0014, on 2^24
    candidates v in [3.4e18, 3.6e18] against the first zeros of the full table (k = 3e-9).  A
    zero evaluation is one candidate against one zero with all 7 shifted terms; a candidate that
    fails a stage stops there.  peaks=True: the survivor case -- every 16th candidate near a peak
    of this table's 256-zero sum and the thresholds halved (0.39/0.4275/0.446), so that a known
    share of the candidates reaches the later stages; per-stage times from calls with the first
    one, two and three stages, and the survivors' stage values checked bit for bit against a run
    in which every candidate takes every stage."""
    import numpy as np
    import torch
    from mertens_bound_amd import evaluator as ev
    from mertens_bound_amd import pipeline, synth
    c = synth.W65536x2325572
    tb, t = full_table(ctx)
    n_cand = args.n_cand if (args.n_cand and ctx is None) else (1 << 24)
    k, d, m = 3e-9, 0.0014, 3
    if peaks:
        Vd = screen_peak_candidates(t, n_cand)
        V = None
        stages = tuple((a, b / 2) for a, b in pipeline.REFERENCE_SCREEN)
    else:
        V = synth.make_candidates(n_cand, c["v_lo"], c["v_hi"], c["seed"])
        Vd = ev.v_tensor(V)
        stages = pipeline.REFERENCE_SCREEN
    state = {}

    def step():
        state["r"] = t.screen_device(Vd, k, d, m, stages)

    ms, wall = timed(step, args.steps, args.warmup, torch)
    val = state["r"][0].cpu().numpy()
    passed = int(state["r"][2].sum().item())
    nz = np.array([a for a, _ in stages])
    reached = (val[:, :len(stages)] >= 0).sum(axis=0)            # candidates reaching each stage
    zero_evals = float(np.sum(reached * np.diff(np.concatenate([[0], nz]))))
    name = "screen_peaks" if peaks else "screen"
    res = base_record(args, "screened candidates/sec (64/128/256-zero staged screen, 7 shifts)",
                      n_cand / (ms * 1e-3), "candidates/s", ms, "u32+f32",
                      {"workload": f"leaf screen, {n_cand:,} candidates"
                                   f"{' (every 16th near a peak)' if peaks else ''}, stages "
                                   f"{'/'.join(str(a) for a, _ in stages)} zeros, thresholds "
                                   f"{'/'.join(str(b) for _, b in stages)}, {2 * m + 1} shifts "
                                   f"of {d}, k={k}", "form": name, "n_cand": n_cand})
    res["screen"] = {"reached_per_stage": [int(x) for x in reached], "passed": passed,
                     "zero_evals_per_step": zero_evals,
                     "zero_evals_per_s": zero_evals / (ms * 1e-3)}
    if peaks:
        # per-stage times: calls with the first one, two and all stages, interleaved over rounds,
        # each call between its own HIP events; the medians' differences are the stages' times
        # (the later stages recompute the earlier zeros of their survivors)
        st = torch.cuda.current_stream()
        ev_pairs = []
        for _ in range(max(10, 2 * args.steps)):
            for ns in range(1, len(stages) + 1):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(st)
                t.screen_device(Vd, k, d, m, stages[:ns])
                e1.record(st)
                ev_pairs.append((ns, e0, e1))
        torch.cuda.synchronize()
        ms_pre = [float(np.median([a.elapsed_time(b) for n_, a, b in ev_pairs if n_ == ns]))
                  for ns in range(1, len(stages) + 1)]
        per = []
        for s_ in range(len(stages)):
            dt = ms_pre[s_] - (ms_pre[s_ - 1] if s_ else 0.0)
            ev_s = float(reached[s_]) * nz[s_]             # zeros the stage's launch sums
            per.append({"stage": s_ + 1, "zeros": int(nz[s_]), "candidates": int(reached[s_]),
                        "ms": dt, "zero_evals_per_s": ev_s / (dt * 1e-3) if dt > 0 else None,
                        "median_call_ms_through_this_stage": ms_pre[s_]})
        res["screen"]["per_stage"] = per
        # every candidate through every stage: the survivors' values must be the same bits
        allv, _, _ = t.screen_device(Vd, k, d, m, tuple((a, 0.0) for a, _ in stages))
        allv = allv.cpu().numpy()
        got = val[:, :len(stages)]
        mask = got >= 0
        thr = np.array([b for _, b in stages], dtype=np.float32)
        want_pass = np.all(allv[:, :len(stages)] >= thr[None, :], axis=1)
        res["accuracy"] = {"survivor_values_equal_all_stage_run":
                               bool(np.array_equal(got[mask], allv[:, :len(stages)][mask])),
                           "pass_flags_match_thresholds":
                               bool(np.array_equal(state["r"][2].cpu().numpy() == 1, want_pass)),
                           "ok": None}
        res["accuracy"]["ok"] = bool(res["accuracy"]["survivor_values_equal_all_stage_run"] and
                                     res["accuracy"]["pass_flags_match_thresholds"])
    g = None if (args.no_check or peaks) else golden_npz("screen_sample.npz")
    if g is not None and np.array_equal(V[:g["V"].shape[0]], g["V"]):
        n = g["V"].shape[0]
        got = val[:n, :3]
        err = np.where(got >= 0, np.abs(got - g["val"]) / g["scale"][None, :], 0.0)
        res["accuracy"] = {"max_stage_value_error_over_sum_abs_a": float(err.max()),
                           "values_compared": int((got >= 0).sum()),
                           "oracle_sample": "screen_sample.npz (first 512 candidates, FP64 "
                                            "oracle, every stage reached)",
                           "bar": 2e-6, "ok": bool(err.max() <= 2e-6)}
    res["roofline"] = roofline("screen", zero_evals, ms, n_cand * 40, name)
    res["wall_s"] = wall
    return res


def run_copy(args):
    import torch
    n = 1 << 28                                  # 2 GiB of float64
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    ms, wall = timed(lambda: y.copy_(x), args.steps, args.warmup, torch)
    return {"metric": "device copy (byte-counter calibration)", "value": 2 * 8 * n / (ms * 1e-3),
            "unit": "B/s", "bytes_read": 8 * n, "bytes_written": 8 * n, "ms": ms}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["metric", "mixed", "batched", "polish", "small",
                                           "enclose", "weights", "screen", "screen_peaks", "compat",
                                           "copy"],
                    default="metric")
    ap.add_argument("--form", choices=["plain", "mixed"], default="plain")
    ap.add_argument("--n-cand", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the accuracy checks and comparison runs (profiling)")
    ap.add_argument("--nc", type=int, default=0)
    ap.add_argument("--target-blocks", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the evaluator shards by candidates with no "
              "communication, and multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.workload in ("metric", "small"):
        res = run_eval(args, args.form)
    elif args.workload == "mixed":
        res = run_eval(args, "mixed")
    elif args.workload == "batched":
        res = run_batched(args)
    elif args.workload == "polish":
        res = run_polish(args)
    elif args.workload == "enclose":
        res = run_enclose(args)
    elif args.workload == "weights":
        res = run_weights(args)
    elif args.workload == "screen":
        res = run_screen(args)
    elif args.workload == "screen_peaks":
        res = run_screen(args, peaks=True)
    elif args.workload == "compat":
        res = run_compat(args)
    else:
        res = run_copy(args)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
