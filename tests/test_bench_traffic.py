"""The traffic tables of bench.py (--traffic): HBM bytes per launch from two rocprofv3 counter
runs, each counter scaled by what it reads for the 1 GiB copy of the same run."""
import csv
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

COLS = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _write(d, counter, rows):
    os.makedirs(d / "host", exist_ok=True)
    with open(d / "host" / "run_counter_collection.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=COLS)
        w.writeheader()
        for i, (name, val) in enumerate(rows):
            w.writerow({"Dispatch_Id": i, "Kernel_Name": name, "Counter_Name": counter, "Counter_Value": val})


def test_traffic_tables_scale_by_the_copy(tmp_path, capsys):
    gib_kib = float(1 << 20)
    # the copy reads half of what it moves by FETCH_SIZE (scale 2) and all of it by WRITE_SIZE
    fetch = [("__amd_rocclr_copyBuffer", gib_kib / 2), ("void reef::bk::kb_bicg_t<double, true>(reef::TankSet, int, int)",
                                                        100000.0),
             ("void reef::bk::kb_bicg_t<double, true>(reef::TankSet, int, int)", 120000.0),
             ("void reef::bk::kb_bicg_x<double>(reef::TankSet, int, int)", 50000.0)]
    write = [("__amd_rocclr_copyBuffer", gib_kib), ("void reef::bk::kb_bicg_t<double, true>(reef::TankSet, int, int)",
                                                    32768.0),
             ("void reef::bk::kb_bicg_t<double, true>(reef::TankSet, int, int)", 32768.0),
             ("void reef::bk::kb_bicg_x<double>(reef::TankSet, int, int)", 65536.0)]
    _write(tmp_path / "f", "FETCH_SIZE", fetch)
    _write(tmp_path / "w", "WRITE_SIZE", write)
    out = tmp_path / "out"
    out.mkdir()
    bench.traffic_tables(str(tmp_path / "f"), str(tmp_path / "w"), out_dir=str(out))
    t = json.load(open(out / "traffic_bicg_t.json"))
    assert abs(t["fetch_scale"] - 2.0) < 1e-12 and abs(t["write_scale"] - 1.0) < 1e-12
    # mean over the two dispatches: 110000 KiB fetched (x2), 32768 KiB written
    assert t["bytes_per_launch"] == int(1024.0 * (2.0 * 110000.0 + 32768.0))
    x = json.load(open(out / "traffic_bicg_x.json"))
    assert x["bytes_per_launch"] == int(1024.0 * (2.0 * 50000.0 + 65536.0))
    assert "| bicg_t |" in capsys.readouterr().out
