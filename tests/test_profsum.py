"""profiles/ summaries (pyrit_amd/profsum.py) from rocprofv3 CSV output:
kernel statistics with the working-launch column, and HBM traffic scaled by
the copy probe with skipped launches left out."""
import csv
import json

from pyrit_amd import profsum


def _write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_stats_with_trace(tmp_path):
    stats = tmp_path / "s.csv"
    trace = tmp_path / "t.csv"
    _write(stats, ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs", "StdDev"],
           [["void pyr::k_sort_pass<8, 20, true, true>(pyr::SortArgs)", 3, 6020, 2006.7, 60.0, 20, 3000, 0],
            ["pyr::(anonymous namespace)::k_emit(int)", 1, 4000, 4000, 40.0, 4000, 4000, 0]])
    _write(trace, ["Kernel_Name", "Start_Timestamp", "End_Timestamp", "VGPR_Count", "LDS_Block_Size"],
           [["void pyr::k_sort_pass<8, 20, true, true>(pyr::SortArgs)", 0, 3000, 72, 49664],
            ["void pyr::k_sort_pass<8, 20, true, true>(pyr::SortArgs)", 5000, 8000, 72, 49664],
            ["void pyr::k_sort_pass<8, 20, true, true>(pyr::SortArgs)", 9000, 9020, 72, 49664],
            ["pyr::(anonymous namespace)::k_emit(int)", 0, 4000, 36, 0]])
    out = tmp_path / "o.md"
    profsum.stats(str(stats), str(out), ["cmd a"], trace=str(trace))
    text = out.read_text()
    assert "    cmd a" in text
    row = [l for l in text.splitlines() if l.startswith("| k_sort_pass<8,20,true,true>")][0].split("|")
    # calls 3, working calls 2 averaging 3.0 us, VGPRs and LDS from the trace
    assert row[2].strip() == "3" and row[10].strip() == "2" and row[11].strip() == "3.0"
    assert row[8].strip() == "72" and row[9].strip() == "49664"
    assert any(l.startswith("| k_emit |") for l in text.splitlines())


def test_traffic_scaling_and_working_launches(tmp_path):
    hdr = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    fetch = tmp_path / "f.csv"
    write = tmp_path / "w.csv"
    # probe: 4096 KiB really read, the counter reads half of it
    _write(fetch, hdr, [["k_copy_probe(unsigned long const*)", "FETCH_SIZE", 2048],
                        ["void pyr::k_sort_pass<8, 20, true, true>(pyr::SortArgs)", "FETCH_SIZE", 1000],
                        ["void pyr::k_sort_pass<8, 20, true, true>(pyr::SortArgs)", "FETCH_SIZE", 1200],
                        ["void pyr::k_sort_pass<8, 20, true, true>(pyr::SortArgs)", "FETCH_SIZE", 1]])
    _write(write, hdr, [["k_copy_probe(unsigned long const*)", "WRITE_SIZE", 4096],
                        ["void pyr::k_sort_pass<8, 20, true, true>(pyr::SortArgs)", "WRITE_SIZE", 800],
                        ["void pyr::k_sort_pass<8, 20, true, true>(pyr::SortArgs)", "WRITE_SIZE", 1000],
                        ["void pyr::k_sort_pass<8, 20, true, true>(pyr::SortArgs)", "WRITE_SIZE", 0]])
    out = tmp_path / "t.json"
    profsum.traffic(str(fetch), str(write), 123, 4096 * 1024, str(out), ["c1", "c2"])
    d = json.loads(out.read_text())
    assert abs(d["calibration"]["fetch_scale"] - 2.0) < 1e-12 and abs(d["calibration"]["write_scale"] - 1.0) < 1e-12
    k = [e for e in d["kernels"] if e["kernel"] == "k_sort_pass<8,20,true,true>"][0]
    assert k["launches"] == 2 and k["voxels"] == 123
    assert abs(k["fetch_bytes"] - 1100 * 1024 * 2.0) < 1e-6
    assert abs(k["write_bytes"] - 900 * 1024 * 1.0) < 1e-6


def test_counters_l2_hit_rate(tmp_path):
    hdr = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    src = tmp_path / "c.csv"
    _write(src, hdr, [["void pyr::k_trace_camera(pyr::DagView)", "TCC_HIT_sum", 300],
                      ["void pyr::k_trace_camera(pyr::DagView)", "TCC_MISS_sum", 100],
                      ["void pyr::k_trace_camera(pyr::DagView)", "TCC_HIT_sum", 500],
                      ["void pyr::k_trace_camera(pyr::DagView)", "TCC_MISS_sum", 100],
                      ["k_other()", "TCC_HIT_sum", 1]])
    out = tmp_path / "o.json"
    profsum.counters(str(src), ["k_trace_camera"], str(out), ["cmd"])
    d = json.loads(out.read_text())["kernels"]
    assert list(d) == ["k_trace_camera"]
    k = d["k_trace_camera"]
    assert k["TCC_HIT_sum"] == 400 and k["TCC_MISS_sum"] == 100 and abs(k["l2_hit_rate"] - 0.8) < 1e-12
