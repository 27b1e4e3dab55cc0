"""The rocprofv3 summarizer of profiles/ on small synthetic CSV files (no GPU)."""
import csv
import json

from gatt_dns_amd import profile_summary as ps


def _write(path, rows, fields):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_hbm_bytes_from_calibrated_counters(tmp_path):
    fields = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    # the calibration copy reads and writes 3 x 256 MiB: FETCH counts half the read KiB, WRITE all
    kib = 3 * (1 << 25) * 8 / 1024
    fetch = [{"Dispatch_Id": 1, "Kernel_Name": "__amd_rocclr_copyBuffer", "Counter_Name": "FETCH_SIZE",
              "Counter_Value": kib / 2}]
    write = [{"Dispatch_Id": 1, "Kernel_Name": "__amd_rocclr_copyBuffer", "Counter_Name": "WRITE_SIZE",
              "Counter_Value": kib}]
    # one library kernel, two dispatches over 2 steps: 3 arrays of 8 points read, 1 written per dispatch
    points = 8
    for d in (2, 3):
        fetch.append({"Dispatch_Id": d, "Kernel_Name": "gatt::k(gatt::A)", "Counter_Name": "FETCH_SIZE",
                      "Counter_Value": 3 * 8 * points / 1024 / 2})
        write.append({"Dispatch_Id": d, "Kernel_Name": "gatt::k(gatt::A)", "Counter_Name": "WRITE_SIZE",
                      "Counter_Value": 1 * 8 * points / 1024})
    _write(tmp_path / "f.csv", fetch, fields)
    _write(tmp_path / "w.csv", write, fields)
    res, md = ps.hbm(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), 2, points, "cmd")
    assert abs(res["calibration"]["read_bytes_per_fetch_kib"] - 2048.0) < 1e-9
    assert abs(res["calibration"]["write_bytes_per_write_kib"] - 1024.0) < 1e-9
    assert abs(res["array_touches_per_point_step"] - 4.0) < 1e-12
    k = res["kernels"]["gatt::k(gatt::A)"]
    assert k["dispatches_per_step"] == 1.0
    assert "3.00 / 1.00" in md
    json.dumps(res)


def test_kernel_stats_busy_union_and_start_grid(tmp_path):
    fields = ["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"]
    rows = [
        # a small set-up kernel before the large case starts: left out with start_grid
        {"Kernel_Name": "gatt::setup(int)", "Start_Timestamp": 0, "End_Timestamp": 500,
         "Grid_Size_X": 64, "Grid_Size_Y": 1, "Grid_Size_Z": 1},
        {"Kernel_Name": "gatt::a(int)", "Start_Timestamp": 1000, "End_Timestamp": 3000,
         "Grid_Size_X": 4096, "Grid_Size_Y": 1, "Grid_Size_Z": 1},
        # overlaps the first: the union is 1000..4000
        {"Kernel_Name": "gatt::b(int)", "Start_Timestamp": 2000, "End_Timestamp": 4000,
         "Grid_Size_X": 64, "Grid_Size_Y": 1, "Grid_Size_Z": 1},
        {"Kernel_Name": "gatt::a(int)", "Start_Timestamp": 5000, "End_Timestamp": 6000,
         "Grid_Size_X": 4096, "Grid_Size_Y": 1, "Grid_Size_Z": 1},
    ]
    _write(tmp_path / "t.csv", rows, fields)
    md = ps.stats(str(tmp_path / "t.csv"), 1, "cmd", start_grid=1000)
    assert "busy union 0.0040 ms" in md          # (4000 - 1000) + (6000 - 5000) ns
    assert "kernel time sum 0.0050 ms" in md     # 2000 + 2000 + 1000 ns
    assert "gatt::setup" not in md
    assert "| `gatt::a(int)` | 2 | 1.5 | 3.0 |" in md
