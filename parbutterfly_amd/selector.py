"""Learned variant selector for batches of graphs (no counterpart in the reference).

A small MLP (16 features -> 64 -> 64 -> 9, BF16 on the matrix cores, trained by the library's
own kernels) predicts the log2 runtime of each of the 9 counting variants (ranking SIDE /
ADEG / ACOCORE x 3 settings of the counting knobs, see pbf_variant_tuning) from cheap graph
statistics (pbf_batch_features).  A batch is then split by the ranking of the predicted best
variant and every part runs as one disjoint-union pass (pbf_count_batch).  Results are bit-exact whatever
the choice: the variants only change the order of work.

    python -m parbutterfly_amd.selector collect --first-seed 1000000 --count 20000 --out DIR
    python -m parbutterfly_amd.selector train --data DIR --out MODEL.npz [--epochs 300]
    python -m parbutterfly_amd.selector evaluate --model MODEL.npz --data DIR2 [--out JSON]
    python -m parbutterfly_amd.selector evaluate-fixture [--fixture NPZ] [--out JSON]
"""
import argparse
import glob
import json
import os
import sys
import time

import numpy as np

import parbutterfly_amd as pb

NVAR = 9
FIXED_VARIANT = 3  # ADEG with the default knobs: the method without the selector
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_model.npz")
HELDOUT_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden",
                            "selector_heldout512.npz")


def variant_settings(v):
    rank, lmax, lwork, bpc = pb.variant_tuning(v)
    return rank, dict(light_max=lmax, light_work=lwork, blocks_per_cu=bpc)


def measure(ctx, dg, reps=1):
    """log2 milliseconds of each variant's per-graph pipeline (ranking + relabel + counting,
    device time between the library's events) on one device graph."""
    use_v, _ = ctx.side(dg)
    out = np.zeros(NVAR)
    ctx.count(dg, use_v, "VERT", "GPU", "ADEG")  # warms the scratch for this graph's size
    for v in range(NVAR):
        rank, knobs = variant_settings(v)
        ctx.set_tuning(**knobs)
        best = None
        for _ in range(reps):
            ctx.count(dg, use_v, "VERT", "GPU", rank)
            ms = ctx.timing()["total_ms"]
            best = ms if best is None else min(best, ms)
        out[v] = np.log2(max(best, 1e-3))
    ctx.set_tuning()
    return out


def collect(ctx, first_seed, count, out_dir, chunk=512, workers=16):
    """Features and measured variant runtimes of graphs first_seed .. first_seed + count - 1,
    written per chunk as DIR/chunk_<start>.npz (F float32 [k, 16], T float32 [k, 9] log2 ms)."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(out_dir, exist_ok=True)
    for start in range(first_seed, first_seed + count, chunk):
        fn = os.path.join(out_dir, f"chunk_{start}.npz")
        if os.path.exists(fn):
            continue
        seeds = list(range(start, min(start + chunk, first_seed + count)))
        t0 = time.time()
        with ThreadPoolExecutor(max_workers=workers) as ex:
            hg = list(ex.map(pb.Graph.batch_graph, seeds))
        dgs = [g.to_device(ctx) for g in hg]
        F = ctx.batch_features(dgs)
        T = np.stack([measure(ctx, d) for d in dgs]).astype(np.float32)
        np.savez(fn, F=F, T=T, seeds=np.array(seeds, np.int64))
        del dgs, hg
        ctx.trim()
        print(f"collect: seeds {seeds[0]}..{seeds[-1]} in {time.time() - t0:.1f}s", flush=True)


def load_fixture(path=HELDOUT_PATH):
    """(F, T, seeds) of a held-out fixture written by tests/golden/make_selector_heldout.py."""
    z = np.load(path)
    return z["F"], z["T"], z["seeds"]


def load_dir(d):
    F, T = [], []
    for fn in sorted(glob.glob(os.path.join(d, "chunk_*.npz"))):
        z = np.load(fn)
        F.append(z["F"])
        T.append(z["T"])
    return np.concatenate(F), np.concatenate(T)


def train(ctx, F, T, epochs=300, lr=2e-3, seed=1):
    fmu, fsd = F.mean(0), F.std(0) + 1e-6
    tmu, tsd = T.mean(0), T.std(0) + 1e-6
    X = ((F - fmu) / fsd).astype(np.float32)
    Y = ((T - tmu) / tsd).astype(np.float32)
    p = pb.mlp_init(seed)
    loss = ctx.mlp_train(p, X, Y, epochs=epochs, lr=lr, seed=seed + 1)
    return dict(params=p, fmu=fmu.astype(np.float32), fsd=fsd.astype(np.float32), tmu=tmu.astype(np.float32),
                tsd=tsd.astype(np.float32), loss=np.float32(loss))


def save(model, path):
    np.savez(path, **model)


def load(path=MODEL_PATH):
    z = np.load(path)
    return {k: z[k] for k in z.files}


def predict(ctx, model, F):
    """Predicted log2 ms [n, 9] from feature rows."""
    X = ((F - model["fmu"]) / model["fsd"]).astype(np.float32)
    return ctx.mlp_predict(model["params"], X) * model["tsd"] + model["tmu"]


def evaluate(ctx, model, F, T):
    """Regret against exhaustive search on measured runtimes, and the gain over the method
    without the selector (the fixed variant)."""
    P = predict(ctx, model, F)
    pick = P.argmin(1)
    t = 2.0 ** T
    best = t.min(1)
    tp = t[np.arange(len(t)), pick]
    tf = t[:, FIXED_VARIANT]
    return {"graphs": int(len(t)),
            "mean_regret_selected": float((tp / best - 1).mean()),
            "mean_regret_fixed": float((tf / best - 1).mean()),
            "total_ms_exhaustive_best": float(best.sum()), "total_ms_selected": float(tp.sum()),
            "total_ms_fixed": float(tf.sum()),
            "gain_selected_vs_fixed": float(tf.sum() / tp.sum()),
            "picked_best_fraction": float((pick == t.argmin(1)).mean()),
            "variant_histogram": np.bincount(pick, minlength=NVAR).tolist()}


def run_learned(ctx, model, graphs, outputs=None):
    """Batched + learned through pbf_count_batch_learned: features, one batched inference, then
    one union pass per chosen ranking with the knob setting most chosen among its graphs; the
    outputs land in pbf_count_batch's layout.  Returns (outputs as Context.count_batch gives
    them, the chosen variant of every graph)."""
    return ctx.count_batch_learned(graphs, model, outputs)


def gather(res, n):
    """Per-graph VERT arrays and totals (host) from run_learned's outputs."""
    if len(res) == 3:  # device outputs
        vout, voff, tot = (x.numpy() for x in res)
        return [vout[int(voff[k]):int(voff[k + 1])] for k in range(n)], tot[:n]
    return res


def count_learned(ctx, model, graphs):
    """run_learned + gather: (per-graph VERT arrays, totals, chosen variants)."""
    res, pick = run_learned(ctx, model, graphs)
    verts, tots = gather(res, len(graphs))
    return verts, tots, pick


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--first-seed", type=int, required=True)
    c.add_argument("--count", type=int, required=True)
    c.add_argument("--out", required=True)
    c.add_argument("--chunk", type=int, default=512)
    t = sub.add_parser("train")
    t.add_argument("--data", required=True, nargs="+")
    t.add_argument("--out", default=MODEL_PATH)
    t.add_argument("--epochs", type=int, default=300)
    t.add_argument("--lr", type=float, default=2e-3)
    e = sub.add_parser("evaluate")
    e.add_argument("--model", default=MODEL_PATH)
    e.add_argument("--data", required=True)
    e.add_argument("--out", default=None)
    x = sub.add_parser("evaluate-fixture")
    x.add_argument("--model", default=MODEL_PATH)
    x.add_argument("--fixture", default=HELDOUT_PATH)
    x.add_argument("--out", default=None)
    a = ap.parse_args()
    ctx = pb.Context(0)
    if a.cmd == "collect":
        collect(ctx, a.first_seed, a.count, a.out, a.chunk)
    elif a.cmd == "train":
        Fs, Ts = zip(*(load_dir(d) for d in a.data))
        F, T = np.concatenate(Fs), np.concatenate(Ts)
        t0 = time.time()
        m = train(ctx, F, T, epochs=a.epochs, lr=a.lr)
        save(m, a.out)
        print(json.dumps({"samples": int(len(F)), "epochs": a.epochs, "final_loss": float(m["loss"]),
                          "train_s": round(time.time() - t0, 2)}), flush=True)
    else:
        if a.cmd == "evaluate-fixture":
            F, T, _ = load_fixture(a.fixture)
        else:
            F, T = load_dir(a.data)
        r = evaluate(ctx, load(a.model), F, T)
        txt = json.dumps(r)
        if a.out:
            with open(a.out, "w") as f:
                f.write(txt + "\n")
        print(txt, flush=True)


if __name__ == "__main__":
    sys.exit(main())
