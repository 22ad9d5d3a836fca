"""Measurement tooling on the CPU: the PMC traffic summary and the bench's
lookup of a committed traffic measurement."""
import csv
import json

import pytest

from opencaeporox_amd import pmcsum


def _csv(path, rows, counter):
    cols = ["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols)
        w.writeheader()
        for d, (name, v) in enumerate(rows, 1):
            w.writerow({"Correlation_Id": d, "Dispatch_Id": d, "Kernel_Name": name, "Counter_Name": counter,
                        "Counter_Value": v})


def test_pmcsum_calibrates_against_the_copy(tmp_path):
    # a 1 GiB copy whose FETCH_SIZE reads half its bytes and WRITE_SIZE all of them
    _csv(tmp_path / "cf.csv", [("__amd_rocclr_copyBuffer", 524288.0)], "FETCH_SIZE")
    _csv(tmp_path / "cw.csv", [("__amd_rocclr_copyBuffer", 1048576.0)], "WRITE_SIZE")
    # a kernel dispatched twice; --last 1 keeps the second dispatch
    _csv(tmp_path / "f.csv", [("k<1>", 10.0), ("k<1>", 100.0)], "FETCH_SIZE")
    _csv(tmp_path / "w.csv", [("k<1>", 20.0), ("k<1>", 50.0)], "WRITE_SIZE")
    out = tmp_path / "t.json"
    pmcsum.main(["--fetch", str(tmp_path / "f.csv"), "--write", str(tmp_path / "w.csv"),
                 "--copy-fetch", str(tmp_path / "cf.csv"), "--copy-write", str(tmp_path / "cw.csv"),
                 "--copy-mib", "1024", "--last", "1", "--workload", "key", "--json", str(out)])
    d = json.loads(out.read_text())
    assert d["fetch_factor"] == pytest.approx(2.0) and d["write_factor"] == pytest.approx(1.0)
    assert d["workload"] == "key"
    k = d["kernels"]["k<1>"]
    assert k["dispatches"] == 1 and k["fetch_kib"] == 100.0 and k["write_kib"] == 50.0
    assert k["hbm_bytes"] == pytest.approx((100.0 * 2 + 50.0) * 1024)


def test_pmcsum_counters_and_derived_figures(tmp_path):
    """FP64 operations from the instruction counters (64 lanes, FMA twice),
    lane utilisation, L2 hit rate and GFLOP/s over the traced mean duration;
    the last dispatch of each kernel."""
    a, b = tmp_path / "a.csv", tmp_path / "b.csv"
    cols = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    with open(a, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols)
        w.writeheader()
        for d, (c, v) in enumerate([("SQ_INSTS_VALU_FMA_F64", 1.0), ("SQ_INSTS_VALU_ADD_F64", 1.0),
                                    ("SQ_INSTS_VALU_MUL_F64", 1.0), ("SQ_INSTS_VALU_TRANS_F64", 1.0),
                                    ("SQ_ACTIVE_INST_VALU", 1.0), ("SQ_THREAD_CYCLES_VALU", 1.0)], 1):
            w.writerow({"Dispatch_Id": 1, "Kernel_Name": "k<9>", "Counter_Name": c, "Counter_Value": v})
        for c, v in [("SQ_INSTS_VALU_FMA_F64", 10.0), ("SQ_INSTS_VALU_ADD_F64", 4.0),
                     ("SQ_INSTS_VALU_MUL_F64", 3.0), ("SQ_INSTS_VALU_TRANS_F64", 1.0),
                     ("SQ_ACTIVE_INST_VALU", 100.0), ("SQ_THREAD_CYCLES_VALU", 3200.0)]:
            w.writerow({"Dispatch_Id": 2, "Kernel_Name": "k<9>", "Counter_Name": c, "Counter_Value": v})
    _csv(b, [("k<9>", 3.0), ("k<9>", 30.0)], "TCC_HIT_sum")
    with open(b, "a", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name",
                                          "Counter_Value"])
        w.writerow({"Correlation_Id": 2, "Dispatch_Id": 2, "Kernel_Name": "k<9>", "Counter_Name": "TCC_MISS_sum",
                    "Counter_Value": 10.0})
    tr = tmp_path / "stats.csv"
    tr.write_text("Name,Calls,TotalDurationNs,AverageNs\nk<9>,2,2000,1000\n")
    out = tmp_path / "c.json"
    pmcsum.main(["--counters", str(a), str(b), "--last", "1", "--trace", str(tr), "--json", str(out)])
    k = json.loads(out.read_text())["k<9>"]
    ops = 64 * (2 * 10 + 4 + 3 + 1)
    assert k["fp64_ops_upper"] == ops and k["fp64_gflops_upper"] == pytest.approx(ops / 1000)
    assert k["valu_lane_util"] == pytest.approx(0.5)
    assert k["l2_hit_rate"] == pytest.approx(30 / 40)
    assert k["mean_ms"] == pytest.approx(1e-3)


def test_bench_reports_traffic_only_for_its_own_configuration(tmp_path, monkeypatch):
    import bench
    f = tmp_path / "traffic.json"
    f.write_text(json.dumps({"workload": "cells=8 Nc=9 seed=2 assembly=4x4x4",
                             "kernels": {"void ocpx::res_asm<7>(...)": {"hbm_bytes": 123.0}}}))
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(f))
    assert bench.measured_traffic("res_asm<7>", "cells=8 Nc=9 seed=2 assembly=4x4x4") == 123.0
    assert bench.measured_traffic("res_asm<7>", "cells=16 Nc=9 seed=2 assembly=4x4x4") is None
    assert bench.measured_traffic("res_asm<6>", "cells=8 Nc=9 seed=2 assembly=4x4x4") is None
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(tmp_path / "missing.json"))
    assert bench.measured_traffic("res_asm<7>", "cells=8 Nc=9 seed=2 assembly=4x4x4") is None
