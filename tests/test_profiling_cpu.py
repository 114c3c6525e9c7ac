"""The rocprofv3 summary tool: kernel names, counter calibration and the
role split of the loop kernels, on small synthetic counter files."""
import csv
import json

from bubbles_amd import profiling

HEAD = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id",
        "Grid_Size", "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size",
        "VGPR_Count", "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value",
        "Start_Timestamp", "End_Timestamp"]


def _write(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(HEAD)
        for d, (name, kib) in enumerate(rows, 1):
            w.writerow([d, d, "Agent 1", 0, 1, 1, 256, 1, name, 256, 0, 0, 8, 0, 16, counter, kib, 0, 1])


def test_short_names():
    assert profiling.short_name("void bbx::k_buckets_density<false, 0>(bbx::SimConst, int)") == "buckets_density"
    assert profiling.short_name("void bbx::k_buckets_density<false, 2>(bbx::SimConst, int)") == "buckets_density_warm"
    assert profiling.short_name("bbx::k_pressure_force(bbx::SimConst)") == "pressure_force"
    assert profiling.short_name("bbx::k_nonpressure<2>(x)") == "nonpressure_warm"


def test_traffic_calibration_and_roles(tmp_path):
    n = 1 << 20
    copy = "bbx::k_stream_copy(a)"
    pf = "bbx::k_pressure_force(a)"
    pp = "bbx::k_predict_pressure(a)"
    # FETCH_SIZE reads half the copied KiB (scale 2), WRITE_SIZE all of it (scale 1)
    fetch = [(copy, 2 ** 19), (pp, 60_000), (pf, 80_000), (pf, 85_000), (pp, 10), (pf, 9)]
    write = [(copy, 2 ** 20), (pp, 12_288), (pf, 16_384), (pf, 65_536), (pp, 0), (pf, 0)]
    _write(tmp_path / "f.csv", "FETCH_SIZE", fetch)
    _write(tmp_path / "w.csv", "WRITE_SIZE", write)
    rec = profiling.traffic(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), n)
    assert rec["fetch_scale"] == 2.0 and rec["write_scale"] == 1.0
    k = rec["kernels"]
    assert k["stream_copy"]["bytes_per_launch"] == 2 * 2 ** 30
    assert k["predict_pressure"]["fetch_bytes"] == 60_000 * 1024 * 2
    assert k["pressure_force_predict"]["write_bytes"] == 16_384 * 1024
    assert k["pressure_force_integrate"]["write_bytes"] == 65_536 * 1024
    assert "pressure_force_skipped" in k and "predict_pressure_skipped" in k
    md = profiling.traffic_markdown(rec, "cmd", {"predict_pressure": 28})
    assert "| predict_pressure | 1 |" in md and "| 28 |" in md
    json.dumps(rec)
