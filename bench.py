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
headers) on the same graph, all host cores (at most 16: a one-GPU job's share on the GPU hosts),
median of 3 timed rounds after a warm-up, compared bit-exactly.  The other sections carry their
own baselines beside the GPU figures: per_edge (CountE, same graph), rmat_count_peel (CountE on
the same graph with its peak memory; PeelE on a bounded R-MAT sample), tip_peel (Peel on a
bounded Chung-Lu sample), batched (the per-graph loop over the first 512 graphs) and config1
(the CPU-runnable case: SERIAL on one core and BATCHS).  Peels report their algorithmic bytes
(tip: 4 sum_O deg^2 + 32 m; wing: 8 sum_e deg(u) deg(v)) against the HBM peak.

traffic: HBM bytes of one counting phase from rocprofv3 FETCH_SIZE / WRITE_SIZE runs, each
scaled by the same counter's reading for pbf_copy_device of a known size (--calibrate-copy),
computed by parbutterfly_amd/profiling.py into profiles/traffic_count_chunglu_2p24.json (per-edge:
profiles/traffic_count_edge_chunglu_2p24.json, reported as per_edge.traffic).

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


def host_cores():
    """Host threads for the CPU baselines: the cores this process may use, at most 16, which is
    the CPU share a one-GPU job gets on the GPU hosts (their nproc can show many more)."""
    cores = os.cpu_count() or 1
    try:
        cores = min(cores, len(os.sched_getaffinity(0)))
    except AttributeError:
        pass
    return max(1, min(cores, 16))


def ref_run(op, graph, count_type, rank_type, rounds=1, threads=None, peel_type="BATCHS", timeout=900):
    """Runs the compiled reference (oracle/_ref) on a graph (or, for op "vertlist", a list of
    graphs); returns (use_v, counts, peel, info, seconds), seconds = the median of the timed
    rounds after the first (the warm-up) when there are several.  None when not built."""
    from oracle import oracle as O
    if not O.ref_available():
        return None
    import resource
    with tempfile.TemporaryDirectory() as td:
        if op == "vertlist":
            paths = []
            for k, g in enumerate(graph):
                paths.append(os.path.join(td, f"g{k}.bcsr"))
                g.write_binary(paths[-1])
            gp = os.path.join(td, "list.txt")
            with open(gp, "w") as f:
                f.write("\n".join(paths) + "\n")
        else:
            gp = os.path.join(td, "g.bcsr")
            graph.write_binary(gp)
        uv, _, counts, peel, info = O.run_ref(op, gp, os.path.join(td, "o.bin"), count_type, rank_type,
                                              peel_type=peel_type, rounds=rounds, threads=threads or host_cores(),
                                              timeout=timeout)
    info["max_rss_gb"] = round(resource.getrusage(resource.RUSAGE_CHILDREN).ru_maxrss / 2 ** 20, 2)
    secs = info["peel_seconds"] if op in ("peelv", "peele") else info["seconds"]
    timed = secs[1:] if len(secs) > 1 else secs
    return uv, counts, peel, info, sorted(timed)[len(timed) // 2]


def cpu_baseline(graph, use_v, rounds=4):
    """The reference's BATCHS / ADEG Count on the headline graph, all host cores."""
    r = ref_run("vert", graph, "BATCHS", "ADEG", rounds=rounds, timeout=900)
    if r is None:
        return None
    uv, counts, _, info, sec = r
    return dict(seconds=sec, all_seconds=info["seconds"], cores=host_cores(), use_v=uv, counts=counts)


def gen_batch(seeds, workers=16):
    """Graphs of the mixed-batch distribution (seed per graph), generated in parallel threads
    (the generator is C code that releases the interpreter lock)."""
    import parbutterfly_amd as pb
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(pb.Graph.batch_graph, seeds))


def batched_bench(pb, ctx, count=4096, first_seed=1000, cpu_sample=512):
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
    sample = hg[:cpu_sample] if cpu_sample else []
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
    # the whole set as one union pass under each selector variant (ranking x knob setting):
    # what a batch-level choice of variant could gain over the fixed one (variant 3)
    var_ms, var_exact = [], True
    for v in range(9):
        rank, lm, lw, bpc = pb.variant_tuning(v)
        ctx.set_tuning(light_max=lm, light_work=lw, blocks_per_cu=bpc)
        ctx.count_batch(dgs, rank, mode="union", outputs=out_u)  # warms this variant's scratch
        bt = None
        for _ in range(2):
            ctx.synchronize()
            t = time.perf_counter()
            out_u = ctx.count_batch(dgs, rank, mode="union", outputs=out_u)
            dt = time.perf_counter() - t
            bt = dt if bt is None else min(bt, dt)
        var_ms.append(round(bt * 1e3, 2))
        var_exact = var_exact and same(out_u, out_p)
    res["union_ms_by_variant"] = var_ms
    res["bit_exact_union_variants_vs_plain"] = var_exact
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
    if sample:
        # the reference's per-graph loop (Compute's side choice + BATCHS / ADEG Count, graph after
        # graph) over the first graphs of the set, all host cores
        r = ref_run("vertlist", sample, "BATCHS", "ADEG", rounds=3)
        if r is not None:
            _, tots, _, info, sec = r
            tp = out_p[2].numpy()
            res["cpu_baseline"] = {"value": round(len(sample) / sec, 1), "unit": "graphs/s", "cores": host_cores(),
                                   "kind": "reference",
                                   "sample": f"the first {len(sample)} graphs of the set, reference BATCHS/ADEG Count() "
                                             "per graph, median of 2 rounds after a warm-up",
                                   "seconds": round(sec, 3),
                                   "bit_exact_vs_gpu": bool(np.array_equal(tots.astype(np.uint64), tp[:len(sample)]))}
    ctx.trim()
    return res


def edge_bench(pb, ctx, dg, use_v, W, m, n, vert_sum, steps=5, host_graph=None):
    """Per-edge counts of the headline graph: throughput, roofline fraction of the counting phase,
    and sum E = 4 TOTAL = 2 sum_S B (the per-vertex run's sum); with host_graph, the reference's
    BATCHS / ADEG CountE on the same graph as the CPU baseline, compared bit-exactly."""
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
    E = pb.as_u64(eout)
    esum = int(E.sum(dtype=np.uint64))
    del eout
    res = {"gwedges_s": round(W / (edge_ms * 1e-3) / 1e9, 3), "ms_per_step": round(edge_ms, 4),
           "count_ms": round(edge_count_ms, 4), "b_alg_bytes": balg,
           "roofline_frac": round(balg / (edge_count_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "sum_is_4_total": esum == 2 * vert_sum}
    if host_graph is not None:
        r = ref_run("edge", host_graph, "BATCHS", "ADEG", rounds=3)
        if r is not None:
            uv, counts, _, info, sec = r
            res["cpu_baseline"] = {"value": round(W / sec / 1e9, 4), "unit": "Gwedges/s", "cores": host_cores(),
                                   "kind": "reference",
                                   "sample": "full workload (same graph), reference BATCHS/ADEG CountE(), median of 2 "
                                             "rounds after a warm-up",
                                   "seconds": round(sec, 3), "max_rss_gb": info["max_rss_gb"],
                                   "bit_exact_vs_gpu": bool(uv == use_v and np.array_equal(counts.astype(np.uint64), E))}
    return res


def tip_b_alg(g, use_v):
    """Algorithmic bytes of a tip peel: every S vertex u is a frontier item once, and its walk
    reads the lists of its centres, 4 * sum_{v in N(u)} deg(v) + 32 * deg(u) bytes (the
    two-hop walk of PeelOrigParallel, butterfly_peel.h:670-702), i.e. 4 sum_{v in O} deg(v)^2
    + 32 m in all."""
    dO = np.diff((g.offsetsV if use_v else g.offsetsU).astype(np.int64))
    return int(4 * (dO * dO).sum() + 32 * g.m)


def wing_b_alg(g, use_v):
    """Algorithmic bytes of a wing peel: every edge e = (u, v) (u on S) is a frontier item once,
    and for each u2 in N(v) its walk merges N(u) with N(u2), 4 (deg u + deg u2) bytes
    (PeelEOrigParallel, butterfly_epeel.h:364-429): 8 sum_e deg(u) deg(v) in all."""
    offS = (g.offsetsU if use_v else g.offsetsV).astype(np.int64)
    edgS = (g.edgesU if use_v else g.edgesV).astype(np.int64)
    dS = np.diff(offS)
    dO = np.diff((g.offsetsV if use_v else g.offsetsU).astype(np.int64))
    return int(8 * (np.repeat(dS, dS) * dO[edgS]).sum())


def tip_bench(pb, ctx, dg, use_v, counts, host_graph, cpu=True):
    """Tip numbers of the headline graph from its per-vertex counts (Peel, the CLI's default
    after VERT counts): one call, timed with the library's events and the host clock, its
    algorithmic bytes against the HBM peak, and the frontiers (items peeled per round and per
    tip level).  The reference's Peel takes hours at this size, so the CPU baseline is a bounded
    sample: Chung-Lu nv = 2^16, nu = 2^14, m = 2^20, gamma 2.1, seed 2, peeled by the reference's
    PBATCHS Peel and by the GPU, compared bit-exactly."""
    t0 = time.perf_counter()
    tip = ctx.peel(dg, use_v, counts)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    tm = ctx.timing()
    T = tip.numpy()
    B = pb.as_u64(counts)
    _, per_level = np.unique(T, return_counts=True)
    balg = tip_b_alg(host_graph, use_v)
    res = {"peel_s": round(wall, 3), "peel_ms_events": round(tm["peel_ms"], 1), "rounds": int(tm["rounds"]),
           "levels": int(len(per_level)), "max_tip": int(T.max()), "tip_le_count": bool(np.all(T <= B)),
           "frontier": {"items": int(len(T)), "mean_per_round": round(len(T) / max(tm["rounds"], 1), 3),
                        "max_per_level": int(per_level.max())},
           "b_alg_bytes": balg, "roofline_frac": round(balg / (tm["peel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}
    if cpu:
        sg = pb.Graph.chung_lu(1 << 16, 1 << 14, 1 << 20, 2.1, 2.1, 2)
        sdg = sg.to_device(ctx)
        su, _ = ctx.side(sdg)
        sc = ctx.count(sdg, su, "VERT")
        ctx.peel(sdg, su, sc)  # warm
        ctx.synchronize()
        t0 = time.perf_counter()
        stip = ctx.peel(sdg, su, sc)
        ctx.synchronize()
        gsec = time.perf_counter() - t0
        rounds = int(ctx.timing()["rounds"])
        r = ref_run("peelv", sg, "BATCHS", "ADEG", rounds=1, peel_type="BATCHS")
        if r is not None:
            uv, _, rtip, info, sec = r
            res["cpu_baseline"] = {"value": round(sec, 3), "unit": "s", "higher_is_better": False,
                                   "cores": host_cores(), "kind": "reference",
                                   "sample": "Chung-Lu nv=2^16 nu=2^14 m=2^20 gamma=2.1 seed=2, reference PBATCHS Peel() "
                                             "of the reference's own counts, one run",
                                   "gpu_s": round(gsec, 4), "rounds": rounds,
                                   "bit_exact_vs_gpu": bool(uv == su and np.array_equal(rtip.astype(np.uint64),
                                                                                        stip.numpy()))}
        del sdg, sc, stip
    return res


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


def rmat_bench(pb, ctx, log_n=22, m=120_000_000, sub_edges=4_000_000, cpu=True):
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
    res["peel_b_alg_bytes"] = wing_b_alg(g, use_v)
    res["peel_roofline_frac"] = round(res["peel_b_alg_bytes"] / res["peel_s"] / 1e9 / HBM_PEAK_GBS, 5)
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
        ok = mn == k  # at every level the least count among the edges with wing >= k is k
        checks.append({"k": k, "edges": int(sel.sum()), "min_butterflies": mn, "ok": bool(ok)})
    res["wing_subgraph_checks"] = checks
    del wing, dg
    if cpu:
        # the reference's BATCHS / ACOCORE CountE on the same graph (dense arrays: its peak
        # resident memory is reported), then a bounded wing-peel sample on both sides
        r = ref_run("edge", g, "BATCHS", "ACOCORE", rounds=1, timeout=1200)
        if r is not None:
            uv, counts, _, info, sec = r
            res["cpu_baseline"] = {"value": round(W / sec / 1e9, 4), "unit": "Gwedges/s", "cores": host_cores(),
                                   "kind": "reference",
                                   "sample": "full workload (same graph), reference BATCHS/ACOCORE CountE(), one run",
                                   "seconds": round(sec, 3), "max_rss_gb": info["max_rss_gb"],
                                   "bit_exact_vs_gpu": bool(uv == use_v and np.array_equal(counts.astype(np.uint64), E))}
            del counts
        sg = pb.Graph.rmat(14, 14, 1_000_000, 0.5, 0.1, 0.1, 0.3, 3)
        sdg = sg.to_device(ctx)
        su, _ = ctx.side(sdg)
        se = ctx.count(sdg, su, "EDGE", "GPU", "ACOCORE")
        ctx.peel_edge(sdg, su, se)  # warm
        ctx.synchronize()
        t0 = time.perf_counter()
        sw = ctx.peel_edge(sdg, su, se)
        ctx.synchronize()
        gsec = time.perf_counter() - t0
        srounds = int(ctx.timing()["rounds"])
        r = ref_run("peele", sg, "BATCHS", "ACOCORE", rounds=1, peel_type="BATCHS")
        if r is not None:
            uv, _, rw, info, sec = r
            res["peel_cpu_baseline"] = {"value": round(sec, 3), "unit": "s", "higher_is_better": False,
                                        "cores": host_cores(), "kind": "reference",
                                        "sample": "R-MAT 2^14 x 2^14 m=1e6 (.5,.1,.1,.3) seed=3, reference PBATCHS PeelE() "
                                                  "of the reference's own ACOCORE counts, one run",
                                        "gpu_s": round(gsec, 4), "rounds": srounds,
                                        "b_alg_bytes": wing_b_alg(sg, su),
                                        "bit_exact_vs_gpu": bool(uv == su and np.array_equal(rw.astype(np.uint64),
                                                                                             sw.numpy()))}
        del sdg, se, sw
    del e, E
    ctx.trim()
    return res


def config1_bench(pb, ctx, steps=50, cpu=True):
    """The reference's CPU-runnable case: uniform random bipartite graph, nv = nu = 1000, 10,000
    distinct edges, seed 1, ADEG, per-vertex counts.  GPU: the graph resident in HBM, one step
    = ranking + relabel + counting + conversion, as one call and as a captured one-graph batch
    (replayed as a HIP graph); CPU: the reference's SERIAL Count on one core
    and its BATCHS Count on all host cores (medians of 3 rounds after a warm-up), compared
    bit-exactly."""
    g = pb.Graph.uniform(1000, 1000, 10000, 1)
    dg = g.to_device(ctx)
    use_v, _ = ctx.side(dg)
    out = pb.DeviceArray(ctx, g.nu if use_v else g.nv, np.uint64)
    for _ in range(5):
        ctx.count(dg, use_v, "VERT", "GPU", "ADEG", out=out)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.count(dg, use_v, "VERT", "GPU", "ADEG", out=out)
    ctx.synchronize()
    sec = (time.perf_counter() - t0) / steps
    W = int(ctx.timing()["wedges"])
    gc = pb.as_u64(out)
    res = {"workload": "uniform bipartite nv=nu=1000 m=10000 seed=1, ADEG, per-vertex counts", "wedges": W,
           "gpu_ms_per_step": round(sec * 1e3, 4), "gpu_gwedges_s": round(W / sec / 1e9, 4)}
    # the same graph as a one-graph captured batch: the whole step replayed as one HIP graph, with
    # no host synchronisation inside it (a graph this small is bound by launches and syncs)
    outs = ctx.count_batch([dg], "ADEG", mode="union", capture=True)
    outs = ctx.count_batch([dg], "ADEG", mode="union", capture=True, outputs=outs)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        outs = ctx.count_batch([dg], "ADEG", mode="union", capture=True, outputs=outs)
    ctx.synchronize()
    csec = (time.perf_counter() - t0) / steps
    vb = outs[0].numpy()[:len(gc)]
    res["gpu_captured_ms_per_step"] = round(csec * 1e3, 4)
    res["gpu_captured_gwedges_s"] = round(W / csec / 1e9, 4)
    res["captured_bit_exact"] = bool(np.array_equal(vb, gc))
    if cpu:
        for ct, thr in (("SERIAL", 1), ("BATCHS", host_cores())):
            r = ref_run("vert", g, ct, "ADEG", rounds=4, threads=thr)
            if r is None:
                continue
            uv, counts, _, info, csec = r
            res["cpu_" + ct.lower()] = {"ms": round(csec * 1e3, 4), "gwedges_s": round(W / csec / 1e9, 4),
                                        "cores": thr, "kind": "reference",
                                        "bit_exact_vs_gpu": bool(uv == use_v and
                                                                 np.array_equal(counts.astype(np.uint64), gc))}
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
        edge = edge_bench(pb, ctx, dg, use_v, W, g.m, n, int(gpu_counts.sum(dtype=np.uint64)),
                          host_graph=None if args.no_cpu_baseline else g)

    tipres = None
    if not args.no_tip:
        tipres = tip_bench(pb, ctx, dg, use_v, out, g, cpu=not args.no_cpu_baseline)
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
    # the per-edge phase's HBM traffic from its own counter runs
    etraffic = os.path.join(ROOT, "profiles", "traffic_count_edge_chunglu_2p24.json")
    if edge is not None and os.path.exists(etraffic):
        with open(etraffic) as f:
            edge["traffic"] = json.load(f).get("bytes_per_launch")

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
        rec["batched"] = batched_bench(pb, ctx, count=args.batch_count, cpu_sample=0 if args.no_cpu_baseline else 512)
    if not args.no_rmat:
        rec["rmat_count_peel"] = rmat_bench(pb, ctx, cpu=not args.no_cpu_baseline)
    rec["config1"] = config1_bench(pb, ctx, cpu=not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        cb = cpu_baseline(g, use_v)
        if cb is not None:
            ok = cb["use_v"] == use_v and np.array_equal(cb["counts"].astype(np.uint64), gpu_counts)
            rec["cpu_baseline"] = {"value": round(W / cb["seconds"] / 1e9, 4), "unit": "Gwedges/s",
                                   "cores": cb["cores"], "kind": "reference",
                                   "sample": "full workload (same graph), reference BATCHS/ADEG Count(), "
                                             f"median of {len(cb['all_seconds']) - 1} rounds after a warm-up",
                                   "seconds": round(cb["seconds"], 3), "bit_exact_vs_gpu": bool(ok)}
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
