"""The bench line's roofline objects keep the fields the measurement
contract names, with the FP64 and HBM peaks of MI355X, on synthetic run
records (no GPU: the helpers only do arithmetic on a finished run)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def run_record(spmv_bytes):
    # a 1M-particle impact run: 103.8 M live bonds, 4.9 products per step
    return dict(n=1_005_000, live=103_809_666, iters=[5, 5, 5, 4, 5], elapsed=0.0828,
                steps=20, spmv_ms=98 * 0.5114, spmv_n=98, spmv_bytes=spmv_bytes)


def check_roofline(rf):
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in rf, key
    assert rf["bound"] in ("hbm", "mfma")
    assert rf["unit"] in ("GB/s", "TFLOP/s")
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-12


def test_matrix_free_roofline_is_hbm_bound():
    r = run_record(1_374_355_992)
    rf = bench.mf_roofline(r)
    check_roofline(rf)
    assert rf["bound"] == "hbm" and rf["peak"] == 8000.0 and rf["unit"] == "GB/s"
    # 1.374 GB per 0.5114 ms product
    assert abs(rf["achieved"] - 1.374355992 / 0.5114e-3) < 1e-6 * rf["achieved"]
    # its FP64 operations per algorithmic byte lie below the FP64 / HBM ridge
    assert abs(rf["ridge_flops_per_byte"] - 78.6e3 / 8000.0) < 1e-12
    assert rf["flops_per_byte"] < rf["ridge_flops_per_byte"]
    assert abs(rf["fp64_frac"] - rf["fp64_tflops"] / 78.6) < 1e-12
    # the traffic comes from the committed counter summary
    assert rf["traffic"] is not None and rf["traffic"] > r["spmv_bytes"]
    # and so do the utilisation counters that say what limits it
    c = rf["counters"]
    assert c is not None and 0.0 < c["valu_busy"] < 1.0 and 0.0 < c["td_busy"] <= 1.0
    assert 50 < c["valu_instr_per_launch"] * 64 / r["live"] < 300
    # the time its VALU instructions alone need (four cycles per wave64
    # instruction, sixteen per FP64 reciprocal or square root, 1,024 SIMDs
    # at 2.4 GHz) is below the time it takes
    want = (4 * c["valu_instr_per_launch"] + 12 * c["fp64_transcendental_per_launch"]) \
        / (1024 * 2.4e9) * 1e3
    assert abs(rf["valu_issue_floor_ms"] - want) < 1e-12
    assert abs(rf["valu_issue_frac"] - want / 0.5114) < 1e-12
    assert 0.3 < rf["valu_issue_frac"] < 1.0


def test_step_rooflines():
    r = run_record(1_374_355_992)
    for rf in (bench.mf_step_roofline(r), bench.step_roofline(dict(r, spmv_bytes=5.49e9))):
        assert rf["bound"] == "hbm" and rf["unit"] == "GB/s"
        assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-12
        assert abs(rf["ms_per_step"] - 1e3 * r["elapsed"] / r["steps"]) < 1e-9
        assert rf["cg_iterations_mean"] == float(np.mean(r["iters"]))


def test_committed_traffic_summaries_parse():
    for key in bench.TRAFFIC_FILES:
        assert bench.read_traffic(key) > 1e9, key
