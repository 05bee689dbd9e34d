"""HBM traffic of the counting phase from rocprofv3 counter runs.

Two ``rocprofv3 --pmc`` runs of the same command (one with FETCH_SIZE, one with WRITE_SIZE,
``--output-format csv``) give per-dispatch KiB figures for every kernel.  Each counter is
scaled by what it reads for ``pbf_copy_device`` of a known size (kernel ``k_copy``, which
reads and writes exactly that many bytes), run under the same two counters.  The traffic of
one counting phase is the sum over its kernels, averaged over the phases in the run (one per
``k_count2<.., 2>`` dispatch).

    python -m parbutterfly_amd.profiling --fetch DIR --write DIR --cal-fetch DIR --cal-write DIR \\
        --cal-bytes N [--out JSON]
"""
import argparse
import csv
import glob
import json
import os
from collections import defaultdict

# kernels launched between the counting phase's start and end events (count.hip)
COUNT_PHASE = ("k_rinfo", "k_slotinfo", "k_fill_u32", "k_fill_u64", "k_classify", "k_classify_big", "k_wedge_chunks",
               "k_scan_tiles", "k_tile_sums", "k_scan_sums", "k_light_hist", "k_light_plan", "k_light_scatter",
               "k_split_pass1", "k_split_credit", "k_split_pass2", "k_count2", "k_count_hbm", "k_hot_reduce",
               "k_convert_vert", "k_oidx", "k_edge_out_s", "k_edge_out_o")


def _short(name):
    base = name.replace("(anonymous namespace)::", "").replace("void ", "")
    return base.split("(")[0].split("<")[0].strip()


def per_dispatch(run_dir, counter):
    """{dispatch_id: (kernel short name, full name, value)} for one counter."""
    files = glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {run_dir}")
    vals = defaultdict(float)
    names = {}
    for fn in files:
        with open(fn) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                d = int(r["Dispatch_Id"])
                vals[d] += float(r["Counter_Value"])
                names[d] = r["Kernel_Name"]
    return {d: (_short(names[d]), names[d], v) for d, v in vals.items()}


def calibration(run_dir, counter, nbytes):
    """Bytes per counter unit from the copy kernel's dispatches."""
    ds = [v for (k, _, v) in per_dispatch(run_dir, counter).values() if k == "k_copy"]
    if not ds:
        raise ValueError(f"no k_copy dispatch in {run_dir}")
    return nbytes / (sum(ds) / len(ds))


def count_phase_traffic(fetch_dir, write_dir, cal_fetch, cal_write):
    out = {}
    for label, d, scale, cn in (("fetch", fetch_dir, cal_fetch, "FETCH_SIZE"), ("write", write_dir, cal_write, "WRITE_SIZE")):
        disp = per_dispatch(d, cn)
        phases = sum(1 for (k, full, _) in disp.values() if k == "k_count2" and ", 2>" in full)
        if phases == 0:
            raise ValueError(f"no light-group k_count2 dispatch in {d}")
        per_kernel = defaultdict(float)
        for (k, _, v) in disp.values():
            if k in COUNT_PHASE:
                per_kernel[k] += v * scale / phases
        out[label] = {"bytes_per_phase": sum(per_kernel.values()), "per_kernel": dict(per_kernel), "phases": phases}
    out["bytes_per_launch"] = out["fetch"]["bytes_per_phase"] + out["write"]["bytes_per_phase"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--cal-fetch", required=True)
    ap.add_argument("--cal-write", required=True)
    ap.add_argument("--cal-bytes", type=int, required=True)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    cf = calibration(a.cal_fetch, "FETCH_SIZE", a.cal_bytes)
    cw = calibration(a.cal_write, "WRITE_SIZE", a.cal_bytes)
    res = count_phase_traffic(a.fetch, a.write, cf, cw)
    res["calibration"] = {"fetch_bytes_per_unit": cf, "write_bytes_per_unit": cw, "copy_bytes": a.cal_bytes}
    txt = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt + "\n")
    print(txt)


if __name__ == "__main__":
    main()
