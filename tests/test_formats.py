"""Reference table/candidate formats and their conversion to the canonical encoding."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from mertens_bound_amd import formats, synth


def test_reference_words_round_trip():
    tb = synth.make_zero_table(400, 3)
    rw = formats.reference_words(tb)
    back = formats.table_from_reference(rw["gamma"], rw["alpha"], rw["ghi_lo"], rw["ghi_hi"],
                                        rw["glo_lo"], rw["glo_hi"], rw["psi_lo"], rw["psi_hi"])
    assert np.array_equal(back.G, tb.G)
    assert np.max(np.abs(back.w - tb.w)) < 1e-16


def test_reference_words_relation():
    tb = synth.make_zero_table(50, 8)
    rw = formats.reference_words(tb)
    for j in range(50):
        G = (int(tb.G[j, 1]) << 64) | int(tb.G[j, 0])
        ghi = (int(rw["ghi_hi"][j]) << 64) | int(rw["ghi_lo"][j])
        glo = (int(rw["glo_hi"][j]) << 64) | int(rw["glo_lo"][j])
        assert ghi == (G << 36) % (1 << 128)
        assert abs(glo * 4096 - G) <= 2048


def test_candidate_round_trip():
    V = synth.make_candidates(64, 1e18, 4e18, 2)
    xs, dy = formats.candidates_to_x_dy(V)
    assert np.all((dy >= 0) & (dy < 2.0 ** -30))
    assert np.array_equal(synth.candidates_from_x_dy(xs, dy), V)


def test_zero_files_round_trip(tmp_path):
    tb = synth.make_zero_table(300, 1)
    rw = formats.reference_words(tb, weight_k=3e-9)
    p = tmp_path / "zeros.bin"
    formats.write_zeros_bin(p, rw)
    back = formats.read_zeros_bin(p)
    for k in ("two_astar", "gamma", "ghi_lo", "ghi_hi", "glo_lo", "glo_hi", "psi_lo", "psi_hi"):
        assert np.array_equal(back[k], rw[k])
    rw2 = formats.reference_words(tb)
    p2 = tmp_path / "zeros_600k.bin"
    formats.write_zeros_600k_bin(p2, rw2)
    back2 = formats.read_zeros_600k_bin(p2)
    for k in ("alpha", "gamma", "ghi_lo", "psi_hi"):
        assert np.array_equal(back2[k], rw2[k])
    assert p2.stat().st_size == 8 + 300 * 64


def test_cluster_lines_from_reference_logs():
    lines = open(f"{GOLDEN}/cluster_sample.txt").read().splitlines()
    cands = formats.parse_cluster_lines(lines)
    assert len(cands) == len(lines)
    first = cands[0]
    assert first["x"] == 434177590594559139596739939 and first["dy"] == -0.00045
    # y = x / 2^30 + dy reproduces the logged y to the printed precision
    assert abs(first["x"] / 2 ** 30 + first["dy"] - 404359391512869987.260789382) < 1e2
    d = formats.dedupe_candidates(cands, shift=32)
    assert len(d) <= len(cands) and d[0]["absh"] == max(c["absh"] for c in cands)
    V = synth.candidates_from_x_dy([c["x"] for c in cands if c["dy"] >= 0],
                                   [c["dy"] for c in cands if c["dy"] >= 0])
    assert V.shape[1] == 2


def test_peaks_csv_candidate_encoding():
    """The reference's own peak list (search/results/subrecord_peaks_summary.csv): each row's
    (x_int, dy) maps to a 64.64 word whose value matches the row's y_exact_base, printed by the
    reference to 10 decimals, within that rounding."""
    from decimal import Decimal, getcontext
    getcontext().prec = 60
    rows = formats.read_peaks_csv(os.path.join(GOLDEN, "subrecord_peaks_summary.csv"))
    assert len(rows) == 25
    assert [int(r["rank_at_k3e9"]) for r in rows] == list(range(1, 26))
    inside = [r for r in rows if Decimal(r["y_exact"]) < Decimal(2 ** 64)]
    outside = [r for r in rows if Decimal(r["y_exact"]) >= Decimal(2 ** 64)]
    assert len(inside) == 24 and len(outside) == 1       # one peak sits above y = 2^64
    V = synth.candidates_from_x_dy([r["x"] for r in inside], [r["dy"] for r in inside])
    for r, v in zip(inside, synth.v_to_int(V)):
        val = Decimal(v) / Decimal(2 ** 64)
        assert abs(val - Decimal(r["y_exact"])) <= Decimal("1e-10"), (r["rank_at_k3e9"], val)
    with pytest.raises(ValueError):
        synth.candidates_from_x_dy([outside[0]["x"]], [outside[0]["dy"]])


def test_cluster_selection_presets():
    lines = open(os.path.join(GOLDEN, "cluster_sample.txt")).read().splitlines()
    cands = formats.parse_cluster_lines(lines)
    assert cands and all(c["y"] is not None for c in cands)
    for preset in (formats.SUB35_SELECTION, formats.KERNEL_OPT_SELECTION):
        sel = formats.select_candidates(cands, **preset)
        keys = [c["x"] >> preset["shift"] for c in sel]
        assert len(keys) == len(set(keys))                      # one per bucket
        assert all(a["absh"] >= b["absh"] for a, b in zip(sel, sel[1:]))
        for c in sel:
            y = c["y"] if preset["y_from"] == "log" else c["x"] / 2.0 ** 30 + c["dy"]
            assert 0 < y < preset["y_max"]
            if preset["min_abs_h"] is not None:
                assert c["absh"] >= preset["min_abs_h"]
            # the kept one is the largest |h| among the admissible candidates of its bucket
            same = [d for d in cands if d["x"] >> preset["shift"] == c["x"] >> preset["shift"]]
            adm = [d for d in same
                   if 0 < (d["y"] if preset["y_from"] == "log" else d["x"] / 2.0 ** 30 + d["dy"])
                   < preset["y_max"]
                   and (preset["min_abs_h"] is None or d["absh"] >= preset["min_abs_h"])]
            assert c["absh"] == max(d["absh"] for d in adm)


def test_zero_chunk_text_exact_words(tmp_path):
    """gamma from decimal text at 60 digits: G = round(gamma / 2pi 2^110) exactly (checked with
    rational arithmetic at 100 digits), independent of the double rounding of gamma."""
    import mpmath
    gammas = ["14.134725141734693790457251983562470270784257115699",
              "21.022039638771554992628479593896902777334340524903",
              "1299999.9876543210987654321098765432109876543210"]
    w = [0.01 - 0.02j, -0.03 + 0.001j, 1e-7 + 2e-7j]
    p = tmp_path / "chunk_000.txt"
    formats.write_zero_chunk(p, gammas, w)
    tb = formats.read_zero_chunks([p])
    with mpmath.workdps(100):
        for j, g in enumerate(gammas):
            want = int(mpmath.nint(mpmath.mpf(g) / (2 * mpmath.pi) * mpmath.mpf(2) ** 110))
            assert synth.v_to_int(tb.G[j:j + 1])[0] == want
            assert tb.gamma[j] == float(mpmath.mpf(g))
    np.testing.assert_array_equal(tb.w, np.array(w))
    assert formats.read_zero_chunks([p], gamma_max=100.0).n == 2


def test_npz_round_trip(tmp_path):
    tb = synth.make_zero_table(2000, 5)
    p = tmp_path / "path1_zeros_100k.npz"
    formats.write_zeros_npz(p, tb)
    d = formats.read_zeros_npz(p)
    assert set(d) >= {"gamma", "re_c", "im_c", "ghi_lo", "glo_lo", "psi_hi", "two_astar"}
    back = formats.table_from_npz(d)
    np.testing.assert_array_equal(back.G, tb.G)
    np.testing.assert_array_equal(back.w, tb.w)
    np.testing.assert_array_equal(back.gamma, tb.gamma)


def _mp_k2(G, w, gamma, Vint, k):
    """K2 from exact big-integer phases (frac(G V / 2^174) turns) and mpmath at 40 digits."""
    import mpmath
    out = []
    with mpmath.workdps(40):
        for V in Vint:
            s = mpmath.mpf(0)
            for (lo, hi), wj, g in zip(G, w, gamma):
                Gi = (int(hi) << 64) | int(lo)
                fr = mpmath.mpf((Gi * V) % (1 << 174)) / mpmath.mpf(2) ** 174
                th = 2 * mpmath.pi * fr
                gg = 2 * mpmath.pi * mpmath.mpf(Gi) / mpmath.mpf(2) ** 110
                a = 2 * mpmath.exp(-mpmath.mpf(k) * gg * gg)
                s += a * (mpmath.mpf(wj.real) * mpmath.cos(th) - mpmath.mpf(wj.imag) * mpmath.sin(th))
            out.append(float(s))
    return np.array(out)


def test_far_candidates_by_rotated_table(oracle_lib):
    """y beyond 2^64 (the reference's peak at 1.8789e19 and its search bound 1.957e19): the
    rotated table at v' = v - 2^64 q reproduces K2 from exact big-integer phases."""
    tb = synth.make_zero_table(150, 21)
    rows = formats.read_peaks_csv(os.path.join(GOLDEN, "subrecord_peaks_summary.csv"))
    far = [r for r in rows if r["x"] >= 1 << 94][0]
    vals = [(far["x"] << 34) + int(round(far["dy"] * 2.0 ** 64)),
            int(1.957e19) << 64, (int(1.2e19) << 64) + 12345]
    k = 3e-9
    want = _mp_k2(tb.G, tb.w, tb.gamma, vals, k)
    got = np.zeros(len(vals))
    groups = formats.split_far_candidates(vals)
    assert set(groups) == {0, 1}
    for q, (ix, words) in groups.items():
        got[ix] = oracle_lib.eval_quad(formats.rotate_table(tb, q), words, k)[0]
    assert np.max(np.abs(got - want)) < 1e-14
