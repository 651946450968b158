"""Reference table/candidate formats and their conversion to the canonical encoding."""
import numpy as np

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
