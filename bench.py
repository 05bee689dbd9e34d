"""Benchmark: per-vertex butterfly counting throughput (Gwedges/s) on one MI355X.

Workload (the configuration the headline metric is quoted on): Chung-Lu power-law bipartite
graph, nv = 2^20, nu = 2^18, m = 2^24 distinct edges, degree exponent 2.1 on both sides, seed 2,
uint32 ids, ADEG ranking, per-vertex uint64 counts.  One step = the reference's timed region
for Count (apps/hyper/BipartiteButterflies.C:219-222): ranking + relabel + wedge counting +
rank-to-id conversion, from the CSR resident in HBM to the counts resident in HBM.

value = W / (seconds per step) / 1e9, W = rank-filtered wedges (rankWedgeF) under the ranking.
roofline: the counting phase, algorithmic bytes B_alg(VERT) = 2 * (4 W + 32 m) + 4 (n + 1) + 8 n_S
over its duration measured with HIP events on the library's stream, against 8 TB/s.
cpu_baseline: the reference's own BATCHS path (oracle/_ref, compiled from the reference's
headers) on the same graph, all host cores, median of its timed rounds.

traffic: HBM bytes of one counting phase from rocprofv3 FETCH_SIZE / WRITE_SIZE runs, each
scaled by the same counter's reading for pbf_copy_device of a known size (--calibrate-copy),
computed by parbutterfly_amd/profiling.py into profiles/traffic_count_chunglu_2p24.json.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--traffic JSON]
    python bench.py --calibrate-copy BYTES     # one device copy of BYTES (counter calibration)
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CFG = dict(nv=1 << 20, nu=1 << 18, m=1 << 24, gamma=2.1, seed=2)


def b_alg_vert(W, m, n, ns):
    return 2 * (4 * W + 32 * m) + 4 * (n + 1) + 8 * ns


def cpu_baseline(graph, use_v, rounds=2):
    """Times the compiled reference (BATCHS, ADEG, per vertex) on the same graph."""
    from oracle import oracle as O
    if not O.ref_available():
        return None
    cores = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
        cores = min(cores, aff)
    except AttributeError:
        pass
    cores = min(cores, 16)
    with tempfile.TemporaryDirectory() as td:
        gp = os.path.join(td, "g.bcsr")
        graph.write_binary(gp)
        t0 = time.time()
        uv, _, counts, _, info = O.run_ref("vert", gp, os.path.join(td, "o.bin"), "BATCHS", "ADEG",
                                           rounds=rounds, threads=cores, timeout=900)
        wall = time.time() - t0
    secs = sorted(info["seconds"])
    return dict(seconds=secs[len(secs) // 2], all_seconds=info["seconds"], cores=cores, wall=wall,
                use_v=uv, counts=counts)


def gen_batch(seeds, workers=16):
    """Graphs of the mixed-batch distribution (seed per graph), generated in parallel threads
    (the generator is C code that releases the interpreter lock)."""
    import parbutterfly_amd as pb
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(pb.Graph.batch_graph, seeds))


def batched_bench(pb, ctx, count=4096, first_seed=1000):
    """Batched workload: `count` graphs (seed first_seed + g; nv, nu log-uniform in [64, 65536],
    average degree log-uniform in [2, 32], exponent uniform in [2, 3]), all CSRs resident in HBM,
    per-graph VERT and TOTAL counts into HBM.  graphs/s of the plain per-graph loop, the
    union pass, the captured union pass, and the captured pass after autotuning; every output
    compared bit-exactly with the plain loop's."""
    t0 = time.time()
    hg = gen_batch([first_seed + g for g in range(count)], workers=min(16, os.cpu_count() or 1))
    gen_s = time.time() - t0
    dgs = [g.to_device(ctx) for g in hg]
    edges = int(sum(g.m for g in hg))
    print(f"# batch: {count} graphs, {edges} edges, generated in {gen_s:.1f}s", file=sys.stderr, flush=True)
    del hg
    res = {"graphs": count, "edges": edges}

    def run(mode, capture=False, outputs=None, reps=1):
        best = None
        for _ in range(reps):
            ctx.synchronize()
            t = time.perf_counter()
            out = ctx.count_batch(dgs, "ADEG", mode=mode, capture=capture, outputs=outputs)
            dt = time.perf_counter() - t
            outputs = out
            best = dt if best is None else min(best, dt)
        return best, outputs

    def same(a, b):
        va, oa, ta = (x.numpy() for x in a)
        vb, ob, tb = (x.numpy() for x in b)
        return bool(np.array_equal(oa, ob) and np.array_equal(ta[:count], tb[:count]) and
                    np.array_equal(va[:int(oa[count])], vb[:int(ob[count])]))

    _, ref = run("union")  # warms the scratch
    t_union, out_u = run("union", reps=2)
    t_plain, out_p = run("plain")
    res["plain_graphs_s"] = round(count / t_plain, 1)
    res["union_graphs_s"] = round(count / t_union, 1)
    res["bit_exact_union_vs_plain"] = same(out_u, out_p)
    _, out_c = run("union", capture=True, outputs=out_u)  # eager pass + capture
    t_cap, out_c = run("union", capture=True, outputs=out_c, reps=3)  # replays
    res["captured_graphs_s"] = round(count / t_cap, 1)
    res["bit_exact_captured_vs_plain"] = same(out_c, out_p)
    best, best_ms = ctx.batch_autotune(dgs)
    _, out_t = run("union", capture=True, outputs=out_c)
    t_tuned, out_t = run("union", capture=True, outputs=out_t, reps=3)
    res["autotuned_config"] = best
    res["autotuned_graphs_s"] = round(count / t_tuned, 1)
    res["bit_exact_autotuned_vs_plain"] = same(out_t, out_p)
    res["speedup_vs_plain"] = round(t_plain / t_tuned, 2)
    ctx.set_tuning()
    # batched + learned (pbf_count_batch_learned): features, MLP selection, one union pass per
    # chosen ranking
    from parbutterfly_amd import selector as S
    if os.path.exists(S.MODEL_PATH):
        model = S.load()
        out_l, pick = S.run_learned(ctx, model, dgs)  # warms every part's scratch
        best_t = None
        for _ in range(2):
            ctx.synchronize()
            t = time.perf_counter()
            out_l, pick = S.run_learned(ctx, model, dgs, outputs=out_l)
            dt = time.perf_counter() - t
            best_t = dt if best_t is None else min(best_t, dt)
        verts, tots = S.gather(out_l, count)
        vp, op, tp = (x.numpy() for x in out_p)
        exact = bool(np.array_equal(tots, tp[:count]) and
                     all(np.array_equal(verts[k], vp[int(op[k]):int(op[k + 1])]) for k in range(count)))
        res["learned_graphs_s"] = round(count / best_t, 1)
        res["learned_speedup_vs_plain"] = round(t_plain / best_t, 2)
        res["bit_exact_learned_vs_plain"] = exact
        res["learned_variant_histogram"] = np.bincount(pick, minlength=S.NVAR).tolist()
    ctx.trim()
    return res


def edge_bench(pb, ctx, dg, use_v, W, m, n, vert_sum, steps=5):
    """Per-edge counts of the headline graph: throughput, roofline fraction of the counting phase,
    and sum E = 4 TOTAL = 2 sum_S B (the per-vertex run's sum)."""
    eout = pb.DeviceArray(ctx, m, np.uint64)
    ctx.count(dg, use_v, "EDGE", "GPU", "ADEG", out=eout)
    ctx.synchronize()
    eph = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.count(dg, use_v, "EDGE", "GPU", "ADEG", out=eout)
        eph.append(ctx.timing())
    ctx.synchronize()
    edge_ms = (time.perf_counter() - t0) * 1e3 / steps
    edge_count_ms = float(np.mean([p["count_ms"] for p in eph]))
    balg = 2 * (4 * W + 32 * m) + 4 * (n + 1) + 8 * m
    esum = int(pb.as_u64(eout).sum(dtype=np.uint64))
    del eout
    return {"gwedges_s": round(W / (edge_ms * 1e-3) / 1e9, 3), "ms_per_step": round(edge_ms, 4),
            "count_ms": round(edge_count_ms, 4), "b_alg_bytes": balg,
            "roofline_frac": round(balg / (edge_count_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "sum_is_4_total": esum == 2 * vert_sum}


def tip_bench(pb, ctx, dg, use_v, counts):
    """Tip numbers of the headline graph from its per-vertex counts (Peel, the CLI's default
    after VERT counts): one call, timed with the library's events and the host clock."""
    t0 = time.perf_counter()
    tip = ctx.peel(dg, use_v, counts)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    tm = ctx.timing()
    T = tip.numpy()
    B = pb.as_u64(counts)
    return {"peel_s": round(wall, 3), "peel_ms_events": round(tm["peel_ms"], 1), "rounds": int(tm["rounds"]),
            "levels": int(len(np.unique(T))), "max_tip": int(T.max()), "tip_le_count": bool(np.all(T <= B))}


def sparse_bench(pb, ctx, dg, exact_total, denom=25, seed=0, reps=5):
    """Approximate TOTAL by sparsification (the driver's -sparseType EDGE / COLOR, -d 25, seed 0)
    on the headline graph in HBM: sparsify + side + rank + relabel + TOTAL count per call.  The
    sparsification passes are HBM streams; their algorithmic bytes are one read of both input
    CSRs and one write of both sparsified CSRs."""
    out = {}
    for st in ("EDGE", "COLOR"):
        h = ctx.sparsify(dg, st, denom, seed)
        nv2, nu2, m2 = h.nv, h.nu, h.m
        del h
        ctx.count_total_sparse(dg, st, denom, seed)
        ctx.synchronize()
        ph = []
        t0 = time.perf_counter()
        for _ in range(reps):
            approx, exact = ctx.count_total_sparse(dg, st, denom, seed)
            ph.append(ctx.timing())
        call_ms = (time.perf_counter() - t0) * 1e3 / reps
        sp_ms = float(np.mean([p["sparsify_ms"] for p in ph]))
        balg = 4 * ((dg.nv + 1) + 2 * dg.m + (dg.nu + 1)) + 4 * ((nv2 + 1) + 2 * m2 + (nu2 + 1))
        out[st] = {"denom": denom, "seed": seed, "sparsified": {"nv": nv2, "nu": nu2, "m": m2},
                   "approx_total": approx, "sparse_total": exact, "exact_total": exact_total,
                   "rel_error": round((approx - exact_total) / exact_total, 5) if exact_total else None,
                   "ms_per_call": round(call_ms, 4), "sparsify_ms": round(sp_ms, 4),
                   "count_ms": round(float(np.mean([p["count_ms"] for p in ph])), 4),
                   "sparsify_b_alg_bytes": balg,
                   "sparsify_roofline_frac": round(balg / (sp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    return out


def read_bench(pb, ctx, g, reps=3):
    """readBipartite of the headline graph from a file in the page cache: the host reader (parse
    on one core, then the CSR copied to HBM by the first count) against the device reader (file
    bytes copied to HBM, tokenised, parsed and validated there), for the AdjacencyHypergraph text
    and the binary CSR.  Medians of `reps` reads; the device reads are checked against the host
    read."""
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for kind, path, write in (("text", os.path.join(td, "g.adj"), g.write),
                                  ("binary", os.path.join(td, "g.bcsr"), g.write_binary)):
            write(path)
            size = os.path.getsize(path)
            th, td_ = [], []
            for _ in range(reps):
                t0 = time.perf_counter()
                h = pb.Graph.read(path)
                th.append(time.perf_counter() - t0)
                del h
                t0 = time.perf_counter()
                d = pb.DeviceGraph.read(ctx, path)
                td_.append(time.perf_counter() - t0)
            same = all(np.array_equal(getattr(g, k), getattr(d, k).numpy())
                       for k in ("offsetsV", "edgesV", "offsetsU", "edgesU"))
            del d
            hs, ds = sorted(th)[reps // 2], sorted(td_)[reps // 2]
            out[kind] = {"bytes": size, "host_read_s": round(hs, 4), "device_read_s": round(ds, 4),
                         "device_read_gb_s": round(size / ds / 1e9, 2), "speedup": round(hs / ds, 1),
                         "same_as_host": bool(same)}
    return out


def rmat_bench(pb, ctx, log_n=22, m=120_000_000, sub_edges=4_000_000):
    """Count + peel workload: R-MAT bipartite graph (a, b, c, d) = (0.5, 0.1, 0.1, 0.3), 2^22 x 2^22
    vertices, 1.2e8 distinct edges, seed 3, ACOCORE ranking, per-edge counts, then the full wing
    peel.  Properties checked at full size: identical per-edge counts under ADEG and ACOCORE,
    sum E = 4 TOTAL, wing <= E, and for the top wing level k (and the largest level whose
    subgraph has at most sub_edges edges) the subgraph H_k = {e : wing(e) >= k} has at least k
    butterflies on every edge, exactly k on some edge at the top level."""
    t0 = time.time()
    g = pb.Graph.rmat(log_n, log_n, m, 0.5, 0.1, 0.1, 0.3, 3)
    print(f"# rmat: nv={g.nv} nu={g.nu} m={g.m} generated in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    dg = g.to_device(ctx)
    use_v, _ = ctx.side(dg)
    res = {"workload": f"rmat bipartite 2^{log_n} x 2^{log_n}, m={m}, (a,b,c,d)=(.5,.1,.1,.3) seed=3, ACOCORE, "
                       "per-edge counts + wing peel",
           "m": int(g.m), "use_v": bool(use_v)}
    e = pb.DeviceArray(ctx, g.m, np.uint64)
    ctx.count(dg, use_v, "EDGE", "GPU", "ACOCORE", out=e)  # warms the scratch
    ph = []
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(3):
        ctx.count(dg, use_v, "EDGE", "GPU", "ACOCORE", out=e)
        ph.append(ctx.timing())
    ctx.synchronize()
    step_ms = (time.perf_counter() - t) * 1e3 / 3
    W = ph[-1]["wedges"]
    count_ms = float(np.mean([p["count_ms"] for p in ph]))
    n = g.nv + g.nu
    balg = 2 * (4 * W + 32 * g.m) + 4 * (n + 1) + 8 * g.m
    res.update({"wedges": int(W), "edge_count_ms_per_step": round(step_ms, 3),
                "edge_gwedges_s": round(W / (step_ms * 1e-3) / 1e9, 3),
                "edge_phases_ms": {k: round(float(np.mean([p[k] for p in ph])), 3)
                                   for k in ("rank_ms", "relabel_ms", "count_ms", "convert_ms")},
                "edge_count_roofline_frac": round(balg / (count_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)})
    E = e.numpy()
    Ead = ctx.count(dg, use_v, "EDGE", "GPU", "ADEG").numpy()
    total = ctx.count(dg, use_v, "TOTAL", "GPU", "ACOCORE")
    res["butterflies"] = int(total)
    res["edge_counts_equal_across_rankings"] = bool(np.array_equal(E, Ead))
    res["edge_sum_is_4_total"] = int(E.sum(dtype=np.uint64)) == 4 * int(total)
    del Ead
    ctx.synchronize()
    t = time.perf_counter()
    wing = ctx.peel_edge(dg, use_v, e)
    ctx.synchronize()
    res["peel_s"] = round(time.perf_counter() - t, 3)
    res["peel_rounds"] = int(ctx.timing()["rounds"])
    wn = wing.numpy()
    res["max_wing"] = int(wn.max())
    res["wing_le_count"] = bool(np.all(wn <= E))
    print(f"# rmat: count {step_ms:.1f} ms/step, peel {res['peel_s']} s, {res['peel_rounds']} rounds",
          file=sys.stderr, flush=True)
    # O-CSR edge list: edge p is (owner, neighbour) of the O side's CSR
    offO, edgO = (g.offsetsV, g.edgesV) if use_v else (g.offsetsU, g.edgesU)
    owner = np.repeat(np.arange(len(offO) - 1, dtype=np.uint32), np.diff(offO.astype(np.int64)))
    vv, uu = (owner, edgO) if use_v else (edgO, owner)
    levels = np.unique(wn)[::-1]
    sizes = np.cumsum([int(np.count_nonzero(wn == k)) for k in levels])
    ks = {int(levels[0])}
    fit = [int(levels[j]) for j in range(len(levels)) if sizes[j] <= sub_edges]
    if fit:
        ks.add(fit[-1])
    checks = []
    for k in sorted(ks, reverse=True):
        sel = wn >= k
        h = pb.Graph.from_edges(g.nv, g.nu, np.ascontiguousarray(vv[sel]), np.ascontiguousarray(uu[sel]))
        huv, _ = ctx.side(h)
        hc = ctx.count(h, huv, "EDGE", "GPU", "ACOCORE")
        mn = int(hc.min()) if len(hc) else 0
        ok = mn >= k and (k != int(levels[0]) or mn == k)
        checks.append({"k": k, "edges": int(sel.sum()), "min_butterflies": mn, "ok": bool(ok)})
    res["wing_subgraph_checks"] = checks
    del e, wing, dg
    ctx.trim()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic_count_chunglu_2p24.json"),
                    help="JSON with measured HBM bytes per counting phase")
    ap.add_argument("--calibrate-copy", type=int, default=0, help="run one device copy of this many bytes and exit")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 4096-graph workload")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--no-rmat", action="store_true", help="skip the R-MAT count + wing-peel workload")
    ap.add_argument("--no-edge", action="store_true", help="skip the per-edge counts on the Chung-Lu graph")
    ap.add_argument("--no-sparse", action="store_true", help="skip the sparsified (approximate) totals")
    ap.add_argument("--no-tip", action="store_true", help="skip the tip peel of the Chung-Lu graph")
    ap.add_argument("--no-read", action="store_true", help="skip the file-reader timings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import parbutterfly_amd as pb

    if args.calibrate_copy:
        ctx = pb.Context(0)
        nb = args.calibrate_copy
        a = pb.DeviceArray(ctx, nb, np.uint8)
        b = pb.DeviceArray(ctx, nb, np.uint8)
        b.copy_from(a)
        print(json.dumps({"calibrate_copy_bytes": nb}), flush=True)
        return

    t0 = time.time()
    g = pb.Graph.chung_lu(CFG["nv"], CFG["nu"], CFG["m"], CFG["gamma"], CFG["gamma"], CFG["seed"])
    gen_s = time.time() - t0
    print(f"# generated graph nv={g.nv} nu={g.nu} m={g.m} in {gen_s:.1f}s", file=sys.stderr, flush=True)
    ctx = pb.Context(0)
    dg = g.to_device(ctx)
    use_v, side_w = ctx.side(dg)
    ns = g.nu if use_v else g.nv
    n = g.nv + g.nu
    out = pb.DeviceArray(ctx, ns, np.uint64)

    # H2D-inclusive rate (host graph in, host counts out), reported separately: the first call
    # also allocates the context's buffers, so the median of the next three is reported
    host_counts = ctx.count(g, use_v, "VERT", "GPU", "ADEG")
    hts = []
    for _ in range(3):
        t_h = time.perf_counter()
        ctx.count(g, use_v, "VERT", "GPU", "ADEG", out=host_counts)
        hts.append(time.perf_counter() - t_h)
    h2d_s = sorted(hts)[1]

    for _ in range(args.warmup):
        ctx.count(dg, use_v, "VERT", "GPU", "ADEG", out=out)
    ctx.synchronize()
    phases = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.count(dg, use_v, "VERT", "GPU", "ADEG", out=out)
        phases.append(ctx.timing())
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    W = phases[-1]["wedges"]
    assert all(p["wedges"] == W for p in phases)
    gwedges = W / (ms_per_step * 1e-3) / 1e9
    count_ms = float(np.mean([p["count_ms"] for p in phases]))
    total_ms = float(np.mean([p["total_ms"] for p in phases]))
    balg = b_alg_vert(W, g.m, n, ns)
    achieved = balg / (count_ms * 1e-3) / 1e9
    gpu_counts = pb.as_u64(out)
    assert np.array_equal(gpu_counts, pb.as_u64(host_counts)), "device and host paths disagree"

    # per-edge counts on the same graph (same ranking, same W): B_alg(EDGE) = 2 B_pass + 4 (n + 1) + 8 m
    edge = None
    if not args.no_edge:
        edge = edge_bench(pb, ctx, dg, use_v, W, g.m, n, int(gpu_counts.sum(dtype=np.uint64)))

    tipres = None
    if not args.no_tip:
        tipres = tip_bench(pb, ctx, dg, use_v, out)
        print(f"# tip peel {tipres['peel_s']} s, {tipres['rounds']} rounds", file=sys.stderr, flush=True)

    sparse = None
    if not args.no_sparse:
        sparse = sparse_bench(pb, ctx, dg, int(gpu_counts.sum(dtype=np.uint64)) // 2)

    reads = None
    if not args.no_read:
        reads = read_bench(pb, ctx, g)

    traffic = None
    if args.traffic and os.path.exists(args.traffic):
        with open(args.traffic) as f:
            traffic = json.load(f).get("bytes_per_launch")

    rec = {
        "metric": "Gwedges/s per-vertex count",
        "value": round(gwedges, 3),
        "unit": "Gwedges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 ids / u64 counts",
        "data": "synthetic (seeded Chung-Lu generator)",
        "config": {"workload": "chung-lu bipartite nv=2^20 nu=2^18 m=2^24 gamma=2.1 seed=2, ADEG, per-vertex counts",
                   "wedges": W, "use_v": use_v, "n_S": ns},
        "phases_ms": {k: round(float(np.mean([p[k] for p in phases])), 4)
                      for k in ("rank_ms", "relabel_ms", "count_ms", "total_ms")},
        "roofline": {"bound": "hbm", "kernel": "counting phase (k_slotinfo ... k_count2 ... k_convert_vert)", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "b_alg_bytes": balg, "count_ms": round(count_ms, 4)},
        "host_inclusive_gwedges_s": round(W / h2d_s / 1e9, 3),
        "per_edge": edge,
        "tip_peel": tipres,
        "sparse_total": sparse,
        "read_path": reads,
    }
    if not args.no_batch:
        del dg, out
        ctx.trim()
        rec["batched"] = batched_bench(pb, ctx, count=args.batch_count)
    if not args.no_rmat:
        rec["rmat_count_peel"] = rmat_bench(pb, ctx)
    if not args.no_cpu_baseline:
        cb = cpu_baseline(g, use_v)
        if cb is not None:
            ok = cb["use_v"] == use_v and np.array_equal(cb["counts"].astype(np.uint64), gpu_counts)
            rec["cpu_baseline"] = {"value": round(W / cb["seconds"] / 1e9, 4), "unit": "Gwedges/s",
                                   "cores": cb["cores"], "kind": "reference",
                                   "sample": "full workload (same graph), reference BATCHS/ADEG Count(), "
                                             f"median of {len(cb['all_seconds'])} rounds",
                                   "seconds": round(cb["seconds"], 3), "bit_exact_vs_gpu": bool(ok)}
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
