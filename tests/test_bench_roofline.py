"""bench.py's roofline arithmetic (CPU): the fraction is the largest class fraction of the frozen
per-evaluation work at the achieved rate, the mixed form charges its FP64 head at the plain
constants, the issue figure uses the counted VALU instructions, and the traffic comes from the
committed counter summary."""
import json
import os

import pytest

import bench

PEAKS = {"fp64_flops": 78.6e12, "fp32_flops": 157.3e12, "int32_mad": 32.0e12, "trans": 18.0e12,
         "source_int32_trans": "test"}


@pytest.fixture(autouse=True)
def fixed_peaks(monkeypatch):
    monkeypatch.setattr(bench, "_RATES", dict(PEAKS))


def test_plain_binds_on_fp64():
    evals, ms = 65_536 * 2_325_572, 150.0
    r = bench.roofline("plain", evals, ms, 113.7e6, "metric_plain")
    rate = evals / (ms * 1e-3)
    assert r["binding_class"] == "fp64_flops"
    assert r["frac"] == pytest.approx(42 * rate / 78.6e12)
    assert r["classes"]["int32_mad"]["frac"] == pytest.approx(15 * rate / 32.0e12)
    assert r["valu_issue"]["valu_issue_frac"] == pytest.approx(
        bench.VALU_INSTS["plain"] * rate / bench.VALU_SLOT_PEAK)
    assert r["frac"] == max(c["frac"] for c in r["classes"].values())


def test_mixed_head_charged_at_plain_constants():
    evals, ms, h = 1.0e11, 50.0, 0.03
    r = bench.roofline("mixed", evals, ms, 0.0, "mixed_mixed", head_frac=h)
    rate = evals / (ms * 1e-3)
    c = r["classes"]
    assert c["fp64_flops"]["per_eval"] == pytest.approx(h * 42)
    assert c["trans"]["per_eval"] == pytest.approx((1 - h) * 2)
    assert c["fp32_flops"]["per_eval"] == pytest.approx((1 - h) * 8)
    assert c["int32_mad"]["per_eval"] == 15
    assert r["binding_class"] == "int32_mad"
    assert r["frac"] == pytest.approx(15 * rate / 32.0e12)
    insts = h * bench.VALU_INSTS["plain"] + (1 - h) * bench.VALU_INSTS["mixed"]
    assert r["valu_issue"]["sq_insts_valu_per_eval"] == pytest.approx(insts)


def test_traffic_from_committed_summary():
    d = json.load(open(os.path.join(bench.ROOT, "profiles", "traffic.json")))
    r = bench.roofline("plain", 1e11, 100.0, 0.0, "metric_plain")
    assert r["traffic"] == d["metric_plain"]["hbm_bytes_per_launch"]
    assert bench.traffic_summary("no_such_workload") is None
