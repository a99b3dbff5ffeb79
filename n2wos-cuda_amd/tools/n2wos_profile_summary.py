#!/usr/bin/env python3
"""Summarise rocprofv3 counter runs of the persistent walker into the JSON
records that bench.py reports (profiles/wos_walker_mesh2_traffic.json and
profiles/wos_walker_mesh2_valu.json), stamped with the walker sources' hash.

Inputs are rocprofv3 output directories (``-d DIR ... --output-format csv``);
the *_counter_collection.csv and *_kernel_trace.csv files under them are
found recursively.

  traffic: --fetch-dir / --write-dir of `bench.py --traffic-probe` runs with
           FETCH_SIZE and WRITE_SIZE.  The 1 GiB device copy in that run
           calibrates each counter (bytes the copy must read or write over
           the counter's reading); the walker dispatch's readings are scaled
           by those factors.
  valu:    --raw-dir (SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_WAVES
           GRBM_GUI_ACTIVE), --valu-dir (VALUBusy VALUUtilization) and
           --occ-dir (OccupancyPercent MeanOccupancyPerCU) of
           `bench.py --section walker` runs, plus --walk-steps (the section's
           printed step total) and --kernel-ms (the walker launch's duration
           from a kernel trace or the bench line).
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
WALKER = "k_wos_persistent"
COPY = "copyBuffer"


def counters(d: str) -> list[dict]:
    """Per-dispatch counter values: [{"kernel", "dispatch", name: value, ...}]
    in dispatch order (values summed over the rows of one dispatch)."""
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no *counter_collection.csv under {d}")
    rows: dict[tuple, dict] = {}
    for f in files:
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                key = (f, int(r["Dispatch_Id"]))
                e = rows.setdefault(key, {"kernel": r["Kernel_Name"], "dispatch": int(r["Dispatch_Id"])})
                e[r["Counter_Name"]] = e.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return [rows[k] for k in sorted(rows)]


def walker_dispatch(ds: list[dict]) -> dict:
    w = [d for d in ds if WALKER in d["kernel"]]
    if not w:
        raise SystemExit("no walker dispatch in the counter file")
    return w[-1]


def source_hash() -> str:
    sys.path.insert(0, ROOT)
    import bench  # noqa: E402
    return bench.walker_source_hash()


def traffic(a) -> dict:
    f, w = counters(a.fetch_dir), counters(a.write_dir)
    copies_f = [d["FETCH_SIZE"] for d in f if COPY in d["kernel"]]
    copies_w = [d["WRITE_SIZE"] for d in w if COPY in d["kernel"]]
    if not copies_f or not copies_w:
        raise SystemExit("no calibration copy dispatch")
    kib = a.copy_bytes / 1024.0
    fs = kib / (sum(copies_f) / len(copies_f))
    ws = kib / (sum(copies_w) / len(copies_w))
    wf, ww = walker_dispatch(f), walker_dispatch(w)
    rd = wf["FETCH_SIZE"] * 1024.0 * fs
    wr = ww["WRITE_SIZE"] * 1024.0 * ws
    return {"kernel": wf["kernel"], "workload": a.workload, "fetch_size_kib": wf["FETCH_SIZE"],
            "write_size_kib": ww["WRITE_SIZE"], "copy_fetch_kib": copies_f, "copy_write_kib": copies_w,
            "fetch_scale": fs, "write_scale": ws, "hbm_read_bytes_per_launch": rd,
            "hbm_write_bytes_per_launch": wr, "hbm_bytes_per_launch": rd + wr, "source_sha256": source_hash(),
            "commands": a.command}


def valu(a) -> dict:
    raw = walker_dispatch(counters(a.raw_dir))
    v = walker_dispatch(counters(a.valu_dir))
    o = walker_dispatch(counters(a.occ_dir))
    insts = raw["SQ_INSTS_VALU"]
    return {"kernel": raw["kernel"], "workload": a.workload, "walk_steps": a.walk_steps,
            "valu_insts": insts, "valu_insts_per_step": insts / a.walk_steps,
            "active_inst_valu": raw.get("SQ_ACTIVE_INST_VALU"), "waves": raw.get("SQ_WAVES"),
            "gui_active_cycles_sum_xcds": raw.get("GRBM_GUI_ACTIVE"),
            "valu_busy": v["VALUBusy"] / 100.0, "valu_utilization": v["VALUUtilization"] / 100.0,
            "waves_per_cu": o.get("MeanOccupancyPerCU"), "occupancy_percent": o.get("OccupancyPercent"),
            "kernel_ms": a.kernel_ms, "source": "rocprofv3 --pmc on bench.py --section walker",
            "source_sha256": source_hash(), "commands": a.command}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("kind", choices=["traffic", "valu"])
    ap.add_argument("--fetch-dir")
    ap.add_argument("--write-dir")
    ap.add_argument("--copy-bytes", type=float, default=float(1 << 30))
    ap.add_argument("--raw-dir")
    ap.add_argument("--valu-dir")
    ap.add_argument("--occ-dir")
    ap.add_argument("--walk-steps", type=float)
    ap.add_argument("--kernel-ms", type=float)
    ap.add_argument("--workload", default="benchmark problem 2, SAH BVH leaf 4, 170,132 points x 256 walks")
    ap.add_argument("--command", action="append", default=[], help="a profiling command, recorded verbatim")
    ap.add_argument("--output", required=True)
    a = ap.parse_args(argv)
    doc = traffic(a) if a.kind == "traffic" else valu(a)
    with open(a.output, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    print(json.dumps(doc))


if __name__ == "__main__":
    main()
