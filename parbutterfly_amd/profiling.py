"""Profiling helpers: workloads to run under rocprofv3, and HBM traffic from its counter runs.

    python -m parbutterfly_amd.profiling workload NAME [--reps N]
    python -m parbutterfly_amd.profiling stats DIR [--top N]   # rocprofv3 --stats table
    python -m parbutterfly_amd.profiling counters DIR [DIR ...] --kernels "K1;K2"   # --pmc table

runs one workload on device 0 (graph generated and uploaded first, one untimed warm-up call,
then N calls), so that a kernel trace or a counter run covers that workload alone.  NAME:
  vert-chunglu   per-vertex counts, Chung-Lu nv = 2^20, nu = 2^18, m = 2^24, gamma 2.1, seed 2, ADEG
  edge-chunglu   per-edge counts of the same graph (ADEG)
  edge-rmat      per-edge counts, R-MAT 2^22 x 2^22, m = 1.2e8, (.5, .1, .1, .3), seed 3, ACOCORE
  tip-chunglu    the tip peel of the Chung-Lu graph from its per-vertex counts (counts untimed)
  wing-rmat      the wing peel of the R-MAT graph from its per-edge counts (counts untimed)

HBM traffic: two ``rocprofv3 --pmc`` runs of the same command (one with FETCH_SIZE, one with WRITE_SIZE,
``--output-format csv``) give per-dispatch KiB figures for every kernel.  Each counter is
scaled by what it reads for ``pbf_copy_device`` of a known size (kernel ``k_copy``, which
reads and writes exactly that many bytes), run under the same two counters.  The traffic of
one counting phase is the sum over its kernels, averaged over the phases in the run (one per
``k_count2<.., 2>`` dispatch).

    python -m parbutterfly_amd.profiling traffic --fetch DIR --write DIR --cal-fetch DIR --cal-write DIR \\
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
               "k_convert_vert", "k_oidx", "k_edge_out_s", "k_edge_out_o", "k_hlog_reduce", "k_endpoint_wedges")


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


def kernel_stats(run_dir, top=30):
    """Markdown table of rocprofv3 --stats (kernel_stats.csv): calls, average and total time."""
    files = glob.glob(os.path.join(run_dir, "**", "*kernel_stats.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no kernel_stats.csv under {run_dir}")
    rows = defaultdict(lambda: [0, 0.0])
    for fn in files:
        with open(fn) as f:
            for r in csv.DictReader(f):
                name = r["Name"].replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0].strip()
                rows[name][0] += int(r["Calls"])
                rows[name][1] += float(r["TotalDurationNs"])
    total = sum(v[1] for v in rows.values())
    out = ["| kernel | calls | avg us | total ms | % |", "|---|---:|---:|---:|---:|"]
    for name, (calls, ns) in sorted(rows.items(), key=lambda kv: -kv[1][1])[:top]:
        out.append(f"| `{name}` | {calls} | {ns / calls / 1e3:.1f} | {ns / 1e6:.2f} | {100 * ns / total:.1f} |")
    out.append(f"\nAll kernels: {total / 1e6:.2f} ms.")
    return "\n".join(out)


def counter_table(run_dirs, kernels):
    """Markdown table of rocprofv3 --pmc counters (one or more runs): each counter's value per
    dispatch (summed over the device), averaged over the dispatches of each kernel (matched on
    the full name's prefix, e.g. "k_count2<0, 1>")."""
    vals = defaultdict(lambda: defaultdict(list))
    for d in run_dirs:
        for fn in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            per = defaultdict(float)
            name = {}
            with open(fn) as f:
                for r in csv.DictReader(f):
                    key = (int(r["Dispatch_Id"]), r["Counter_Name"])
                    per[key] += float(r["Counter_Value"])
                    name[int(r["Dispatch_Id"])] = r["Kernel_Name"].replace("(anonymous namespace)::", "").replace("void ", "")
            for (disp, cn), v in per.items():
                full = name[disp]
                for k in kernels:
                    if full.startswith(k):
                        vals[cn][k].append(v)
    out = ["| counter (per dispatch) | " + " | ".join(f"`{k}`" for k in kernels) + " |",
           "|---|" + "---:|" * len(kernels)]
    for cn in sorted(vals):
        row = []
        for k in kernels:
            xs = vals[cn].get(k, [])
            row.append(f"{sum(xs) / len(xs):.3g}" if xs else "")
        out.append(f"| {cn} | " + " | ".join(row) + " |")
    return "\n".join(out)


WORKLOADS = ("vert-chunglu", "edge-chunglu", "edge-rmat", "tip-chunglu", "wing-rmat")


def run_workload(name, reps):
    import sys
    import time

    import parbutterfly_amd as pb
    if name not in WORKLOADS:
        raise SystemExit(f"unknown workload {name}; one of {WORKLOADS}")
    if name.endswith("chunglu"):
        g = pb.Graph.chung_lu(1 << 20, 1 << 18, 1 << 24, 2.1, 2.1, 2)
        rank = "ADEG"
    else:
        g = pb.Graph.rmat(22, 22, 120_000_000, 0.5, 0.1, 0.1, 0.3, 3)
        rank = "ACOCORE"
    ctx = pb.Context(0)
    dg = g.to_device(ctx)
    use_v, _ = ctx.side(dg)
    per = "VERT" if name.startswith(("vert", "tip")) else "EDGE"
    counts = ctx.count(dg, use_v, per, "GPU", rank)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        if name.startswith("tip"):
            ctx.peel(dg, use_v, counts)
        elif name.startswith("wing"):
            ctx.peel_edge(dg, use_v, counts)
        else:
            ctx.count(dg, use_v, per, "GPU", rank, out=counts)
    ctx.synchronize()
    print(json.dumps({"workload": name, "reps": reps, "ms_per_call": (time.perf_counter() - t0) * 1e3 / reps,
                      "timing": ctx.timing()}), flush=True)
    sys.stdout.flush()


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    w = sub.add_parser("workload")
    w.add_argument("name")
    w.add_argument("--reps", type=int, default=2)
    st = sub.add_parser("stats")
    st.add_argument("dir")
    st.add_argument("--top", type=int, default=30)
    cnt = sub.add_parser("counters")
    cnt.add_argument("dirs", nargs="+")
    cnt.add_argument("--kernels", required=True, help="kernel name prefixes separated by semicolons")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--cal-fetch", required=True)
    t.add_argument("--cal-write", required=True)
    t.add_argument("--cal-bytes", type=int, required=True)
    t.add_argument("--out", default=None)
    a = ap.parse_args()
    if a.cmd == "workload":
        run_workload(a.name, a.reps)
        return
    if a.cmd == "stats":
        print(kernel_stats(a.dir, a.top))
        return
    if a.cmd == "counters":
        print(counter_table(a.dirs, [k.strip() for k in a.kernels.split(";")]))
        return
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
