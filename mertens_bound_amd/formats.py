"""The reference's table and candidate formats, and conversion to the canonical encoding.

* ``zeros.bin`` (search/path1_build_zeros_100k.py:34-47): int32 n, then planar arrays
  two_astar[n] (2 |w| e^{-3e-9 g^2}), gamma[n], ghi_lo, ghi_hi, glo_lo, glo_hi, psi_lo, psi_hi.
* ``zeros_600k.bin`` (legacy/superseded/certify_peak2_600k.py:119-143, read by
  search/eval_tech1_600k.py:38-50): uint64 n, gamma[n], alpha[n] (= |w|), then n x 6 uint64
  rows (ghi_lo, ghi_hi, glo_lo, glo_hi, psi_lo, psi_hi).
* words: glo = rint(g 2^98), ghi = rint(g 2^146) mod 2^128, psi = rint(frac(-arg w / 2pi) 2^128)
  with g = gamma / 2pi (search/path1_build_zeros_100k.py:56-58,102-104).
* ``path1_zeros_100k.npz`` (search/path1_build_zeros_100k.py:145-159): the same words plus
  gamma, re_c, im_c (w = re_c + i im_c), alpha and two_astar.
* zero chunk text files, one zero per line ``gamma re im`` in decimal (read by
  search/polish.py:23-38 and search/path1_build_zeros_100k.py:63-90).
* candidates: ``CLUSTER`` log lines carry x (Q98.30 integer) and dy, y = x / 2^30 + dy
  (search/eval_sub35_candidates.py:26-46, search/eval_tech1_600k.py:27-29,216-218;
  search/path1_kernel_opt.py:224-264); the peak summary CSV of
  search/results/subrecord_peaks_summary.csv carries x_int, dy and y_exact_base.

Canonical words: G = floor(g 2^110) (Q18.110) and V = v 2^64 (64.64).  A canonical table maps
exactly onto reference words (ghi = G 2^36 mod 2^128, glo = round(G / 2^12)); a reference table
maps back with :func:`table_from_reference`.
"""
from __future__ import annotations

import re
import struct

import numpy as np

from . import synth

M64 = (1 << 64) - 1
M128 = (1 << 128) - 1


def _words(x: int):
    return x & M64, (x >> 64) & M64


def _u128(lo, hi) -> int:
    return (int(hi) << 64) | int(lo)


def reference_words(table: synth.ZeroTable, weight_k: float | None = None):
    """Reference-format arrays for a canonical table.

    Returns dict(gamma, alpha, ghi_lo, ghi_hi, glo_lo, glo_hi, psi_lo, psi_hi) where alpha = |w|
    (zeros_600k.bin convention) or, with weight_k, two_astar = 2|w| e^{-k g^2} (zeros.bin).
    """
    n = table.n
    out = {k: np.zeros(n, dtype=np.uint64) for k in
           ("ghi_lo", "ghi_hi", "glo_lo", "glo_hi", "psi_lo", "psi_hi")}
    arg = np.angle(table.w)
    frac = np.mod(-arg / (2 * np.pi), 1.0)
    for j in range(n):
        G = _u128(*table.G[j])
        ghi = (G << 36) & M128
        glo = (G + (1 << 11)) >> 12
        # psi: 53 significant bits of the phase fraction, scaled to 2^128
        psi = int(round(float(frac[j]) * 2.0 ** 53)) << 75 & M128
        out["ghi_lo"][j], out["ghi_hi"][j] = _words(ghi)
        out["glo_lo"][j], out["glo_hi"][j] = _words(glo)
        out["psi_lo"][j], out["psi_hi"][j] = _words(psi)
    alpha = np.abs(table.w)
    out["gamma"] = table.gamma.copy()
    if weight_k is None:
        out["alpha"] = alpha
    else:
        out["two_astar"] = 2.0 * alpha * np.exp(-weight_k * table.gamma ** 2)
    return out


def table_from_reference(gamma, alpha, ghi_lo, ghi_hi, glo_lo, glo_hi, psi_lo, psi_hi):
    """Canonical table from reference words.

    ghi = rint(g 2^146) mod 2^128 keeps only g mod 2^-18 (the rest never changes the
    reference's phase, path1_eval_fast.cpp:45); glo = rint(g 2^98) carries the leading bits.
    g 2^146 = K 2^128 + ghi with K = round((glo 2^48 - ghi) / 2^128), then G = round(g 2^146 / 2^36).
    w = alpha e^{-2 pi i psi / 2^128}.
    """
    n = len(gamma)
    G = np.zeros((n, 2), dtype=np.uint64)
    for j in range(n):
        ghi = _u128(ghi_lo[j], ghi_hi[j])
        glo = _u128(glo_lo[j], glo_hi[j])
        K = ((glo << 48) - ghi + (1 << 127)) >> 128
        full = (K << 128) + ghi
        G[j, 0], G[j, 1] = _words((full + (1 << 35)) >> 36)
    ang = np.array([-2 * np.pi * (_u128(l, h) / 2.0 ** 128) for l, h in zip(psi_lo, psi_hi)])
    w = np.asarray(alpha, dtype=np.float64) * np.exp(1j * ang)
    return synth.ZeroTable(G=G, gamma=np.asarray(gamma, dtype=np.float64), w=w)


def candidates_to_x_dy(V):
    """64.64 words -> reference (x in Q98.30, dy0 float) with V = x 2^34 + dy0 2^64 exactly
    representable remainder (dy0 in [0, 2^-30))."""
    xs, dys = [], []
    for lo, hi in V:
        v = _u128(lo, hi)
        x = v >> 34
        r = v - (x << 34)
        xs.append(x)
        dys.append(r / 2.0 ** 64)
    return xs, np.array(dys)


def x_words(xs):
    lo = np.array([x & M64 for x in xs], dtype=np.uint64)
    hi = np.array([(x >> 64) & M64 for x in xs], dtype=np.uint64)
    return lo, hi


# ---- binary table files --------------------------------------------------------------------

def write_zeros_bin(path, ref: dict):
    """search/path1_build_zeros_100k.py:34-47 layout (needs ref['two_astar'])."""
    n = len(ref["gamma"])
    with open(path, "wb") as f:
        f.write(struct.pack("<i", n))
        for key in ("two_astar", "gamma", "ghi_lo", "ghi_hi", "glo_lo", "glo_hi", "psi_lo",
                    "psi_hi"):
            f.write(np.ascontiguousarray(ref[key][:n]).tobytes())


def read_zeros_bin(path) -> dict:
    with open(path, "rb") as f:
        n = struct.unpack("<i", f.read(4))[0]
        out = {"two_astar": np.frombuffer(f.read(8 * n), dtype=np.float64).copy(),
               "gamma": np.frombuffer(f.read(8 * n), dtype=np.float64).copy()}
        for key in ("ghi_lo", "ghi_hi", "glo_lo", "glo_hi", "psi_lo", "psi_hi"):
            out[key] = np.frombuffer(f.read(8 * n), dtype=np.uint64).copy()
    return out


def write_zeros_600k_bin(path, ref: dict):
    """legacy/superseded/certify_peak2_600k.py:119-143 layout (needs ref['alpha'])."""
    n = len(ref["gamma"])
    rows = np.stack([ref[k] for k in ("ghi_lo", "ghi_hi", "glo_lo", "glo_hi", "psi_lo",
                                      "psi_hi")], axis=1).astype(np.uint64)
    with open(path, "wb") as f:
        f.write(struct.pack("<Q", n))
        f.write(np.ascontiguousarray(ref["gamma"], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(ref["alpha"], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(rows).tobytes())


def read_zeros_600k_bin(path) -> dict:
    """search/eval_tech1_600k.py:38-50."""
    with open(path, "rb") as f:
        n = struct.unpack("<Q", f.read(8))[0]
        gamma = np.frombuffer(f.read(n * 8), dtype=np.float64).copy()
        alpha = np.frombuffer(f.read(n * 8), dtype=np.float64).copy()
        raw = np.frombuffer(f.read(n * 48), dtype=np.uint64).reshape(n, 6)
    out = {"gamma": gamma, "alpha": alpha}
    for i, key in enumerate(("ghi_lo", "ghi_hi", "glo_lo", "glo_hi", "psi_lo", "psi_hi")):
        out[key] = np.ascontiguousarray(raw[:, i])
    return out


# ---- CLUSTER log lines ------------------------------------------------------------------

# the fields the reference's parsers read: |h|, h, y and the (x, dy) pair
# (search/eval_sub35_candidates.py:26-46; search/eval_tech1_600k.py:27-29)
_CLUSTER = re.compile(
    r"CLUSTER #(?P<n>\d+)(?:\s+\[coset=(?P<coset>[^\]]+)\])?\s+\|h\|=(?P<absh>[0-9.]+)\s+"
    r"h=(?P<h>[+\-0-9.]+).*?\(x=(?P<x>\d+),\s*dy=(?P<dy>[+\-0-9.eE]+)\)")


_Y = re.compile(r"\sy=([0-9.]+)\s")


def parse_cluster_lines(lines):
    """Candidates from CLUSTER lines: list of dict(n, coset, absh, h, x, dy, y) (y the logged
    value, None where the line has none)."""
    out = []
    for line in lines:
        m = _CLUSTER.search(line)
        if not m:
            continue
        my = _Y.search(line)
        out.append({"n": int(m["n"]), "coset": m["coset"], "absh": float(m["absh"]),
                    "h": float(m["h"]), "x": int(m["x"]), "dy": float(m["dy"]),
                    "y": float(my.group(1)) if my else None})
    return out


# the two reference selections of CLUSTER candidates
SUB35_SELECTION = {"y_max": 3.4849e18, "shift": 32, "min_abs_h": None, "y_from": "log"}
KERNEL_OPT_SELECTION = {"y_max": 1.95718788505622e19, "shift": 18, "min_abs_h": 0.96,
                        "y_from": "x_dy"}


def select_candidates(cands, y_max, shift, min_abs_h=None, y_from="log"):
    """Filter and deduplicate CLUSTER candidates like the reference drivers: keep
    0 < y < y_max (y as logged, or x / 2^30 + dy) and |h| >= min_abs_h, then the largest |h| per
    x >> shift, sorted by |h| (search/eval_sub35_candidates.py:26-46 with SUB35_SELECTION;
    search/path1_kernel_opt.py:224-264 with KERNEL_OPT_SELECTION)."""
    keep = []
    for c in cands:
        y = c.get("y") if y_from == "log" else c["x"] / 2.0 ** 30 + c["dy"]
        if y is None or not (0 < y < y_max):
            continue
        if min_abs_h is not None and c["absh"] < min_abs_h:
            continue
        keep.append(c)
    return dedupe_candidates(keep, shift=shift)


def read_peaks_csv(path):
    """Rows of search/results/subrecord_peaks_summary.csv as dicts; x (int), dy (float) and
    y_exact (the decimal string of x / 2^30 + dy) are parsed, the other columns kept as text."""
    import csv
    out = []
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            d = dict(row)
            d["x"] = int(row["x_int"])
            d["dy"] = float(row["dy"])
            d["y_exact"] = row["y_exact_base"]
            out.append(d)
    return out


# ---- candidates beyond the 64.64 range ----------------------------------------------------

def split_far_candidates(vals):
    """Exact values V = v 2^64 (Python ints, v may exceed 2^64) -> {q: (indices, words)} with
    v = 2^64 q + v', 0 <= v' < 2^64, words the 64.64 words of v'.  Evaluate group q against
    rotate_table(table, q)."""
    groups = {}
    for i, x in enumerate(vals):
        q, r = divmod(int(x), 1 << 128)
        if q < 0:
            raise ValueError("negative candidate")
        groups.setdefault(q, ([], []))
        groups[q][0].append(i)
        groups[q][1].append(r)
    return {q: (np.array(ix), synth.v_from_int(rs)) for q, (ix, rs) in groups.items()}


def rotate_table(table: synth.ZeroTable, q: int) -> synth.ZeroTable:
    """The table for candidates shifted down by 2^64 q: e^{i g (2^64 q + v')} = e^{i g 2^64 q}
    e^{i g v'}, and g 2^64 q / 2pi = G q / 2^46 exactly, so w_j becomes w_j e^{2 pi i f_j} with
    f_j = ((G_j q) mod 2^46) / 2^46 (a 46-bit fraction, exact in FP64).  K2 and its derivatives
    at v are those of the rotated table at v', up to the FP64 rounding of the rotated weights."""
    if q == 0:
        return table
    f = np.array([((_u128(lo, hi) * q) & ((1 << 46) - 1)) / 2.0 ** 46 for lo, hi in table.G])
    rot = np.exp(2j * np.pi * f)
    return synth.ZeroTable(G=table.G, gamma=table.gamma, w=table.w * rot)


# ---- zero chunk text files and the .npz table ---------------------------------------------

def read_zero_chunks(paths, gamma_max=None):
    """Zeros from text files of ``gamma re im`` lines (decimal).  G is the 64.64-style word
    round(gamma / 2pi 2^110) of the decimal gamma at 60 significant digits, so the table's gamma
    (2 pi G / 2^110) is the decimal value to within 2^-110 relative, independent of any double
    rounding; w = re + i im.  Zeros with gamma >= gamma_max are skipped (search/polish.py:31-33)."""
    import mpmath
    Gs, gam, w = [], [], []
    with mpmath.workdps(60):
        scale = mpmath.mpf(2) ** 110 / (2 * mpmath.pi)
        for p in paths:
            with open(p) as f:
                for line in f:
                    parts = line.split()
                    if len(parts) < 3:
                        continue
                    g = mpmath.mpf(parts[0])
                    if gamma_max is not None and g >= gamma_max:
                        continue
                    Gs.append(int(mpmath.nint(g * scale)))
                    gam.append(float(g))
                    w.append(complex(float(parts[1]), float(parts[2])))
    G = np.zeros((len(Gs), 2), dtype=np.uint64)
    for j, x in enumerate(Gs):
        G[j, 0], G[j, 1] = _words(x)
    return synth.ZeroTable(G=G, gamma=np.array(gam, dtype=np.float64),
                           w=np.array(w, dtype=np.complex128))


def write_zero_chunk(path, gamma_text, w):
    """One text file of ``gamma re im`` lines (gamma as given decimal strings)."""
    with open(path, "w") as f:
        for g, c in zip(gamma_text, w):
            f.write(f"{g} {c.real!r} {c.imag!r}\n")


_NPZ_KEYS = ("gamma", "re_c", "im_c", "alpha", "two_astar", "ghi_lo", "ghi_hi", "glo_lo",
             "glo_hi", "psi_lo", "psi_hi")


def write_zeros_npz(path, table: synth.ZeroTable, weight_k: float = 3e-9):
    """search/path1_build_zeros_100k.py:145-159 layout from a canonical table."""
    ref = reference_words(table, weight_k=weight_k)
    ref_a = reference_words(table)
    np.savez(path, gamma=table.gamma, re_c=table.w.real, im_c=table.w.imag,
             alpha=ref_a["alpha"], two_astar=ref["two_astar"],
             **{k: ref[k] for k in ("ghi_lo", "ghi_hi", "glo_lo", "glo_hi", "psi_lo", "psi_hi")})


def read_zeros_npz(path) -> dict:
    with np.load(path, allow_pickle=False) as d:
        return {k: np.array(d[k]) for k in _NPZ_KEYS if k in d.files}


def table_from_npz(d: dict) -> synth.ZeroTable:
    """Canonical table from .npz fields: G from the ghi/glo words (as table_from_reference),
    w = re_c + i im_c (the stored weights, not their alpha/psi rounding)."""
    t = table_from_reference(d["gamma"], d["alpha"], d["ghi_lo"], d["ghi_hi"], d["glo_lo"],
                             d["glo_hi"], d["psi_lo"], d["psi_hi"])
    return synth.ZeroTable(G=t.G, gamma=t.gamma, w=d["re_c"] + 1j * d["im_c"])


def dedupe_candidates(cands, shift=32):
    """Keep the largest |h| per bucket x >> shift (search/eval_sub35_candidates.py dedupes on
    x >> 32; search/path1_kernel_opt.py:224-264 on x >> 18)."""
    best = {}
    for c in cands:
        key = c["x"] >> shift
        if key not in best or c["absh"] > best[key]["absh"]:
            best[key] = c
    return sorted(best.values(), key=lambda c: -c["absh"])
