"""Writes the golden vectors in tests/golden/ from the compiled reference (oracle/_ref).

Run in a container that has the reference tree (oracle/Makefile builds oracle/_ref/ref_harness
from the reference's own headers):
    python tests/golden/make_golden.py [--sparse-only | --side-only | --konect-only]
For every graph below it stores the AdjacencyHypergraph input (<name>.adj) and the reference's
outputs (<name>.npz): cmpWedgeCounts, per-vertex counts (SERIAL and BATCHS), per-edge counts,
totals, tip numbers (PBATCHS), wing numbers (PBATCHS), and the reference's ADEG / ACOCORE orders
with the wedge count W each gives; sparse_cases.npz holds the graphs that the reference's
sparsifiers make of them (make_sparse); side_cases.npz holds cmpWedgeCounts (and the counts)
of graphs on which its 32-bit C(d,2) wraps (make_side); konect_*.tsv / .adj are KONECT edge
lists and what the reference's utils/KONECTtoHyperAdj.C writes for them (make_konect).  The reference prints only totals
(BipartiteButterflies.C:215, 234, 262); the harness dumps the arrays that Compute discards.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import parbutterfly_amd as pb  # noqa: E402
from oracle import oracle as O  # noqa: E402


def graphs():
    kab = np.repeat(np.arange(4, dtype=np.uint32), 5), np.tile(np.arange(5, dtype=np.uint32), 4)
    return [
        ("uniform_1000x1000_m10000_s1", pb.Graph.uniform(1000, 1000, 10000, 1)),
        ("chunglu_300x200_m2000_g21_s7", pb.Graph.chung_lu(300, 200, 2000, 2.1, 2.1, 7)),
        ("chunglu_2000x500_m12000_g21_s2", pb.Graph.chung_lu(2000, 500, 12000, 2.1, 2.1, 2)),
        ("rmat_8x8_m3000_s3", pb.Graph.rmat(8, 8, 3000, 0.5, 0.1, 0.1, 0.3, 3)),
        ("k_4_5", pb.Graph.from_edges(4, 5, *kab)),
    ]


SPARSE_CASES = [("EDGE", 2, 5), ("EDGE", 3, 1), ("EDGE", 25, 7), ("COLOR", 2, 5), ("COLOR", 3, 1), ("COLOR", 25, 7),
                ("EDGE", 4, 123456789), ("COLOR", 4, 123456789)]


def make_sparse(tmp):
    """sparse_cases.npz: for every graph and (type, denom, seed) of SPARSE_CASES the graph that
    eSparseBipartite / clrSparseBipartite (with delZeroDeg) make of it, and CountTotal of that
    graph with the driver's scaling."""
    out_bin = os.path.join(tmp, "s.bin")
    rec = {}
    for name, _ in graphs():
        adj = os.path.join(HERE, name + ".adj")
        for st, d, seed in SPARSE_CASES:
            _, _, c, p, _ = O.run_ref("sparse", adj, out_bin, "BATCHS", "ADEG",
                                      extra=("-sparseType", st, "-d", str(d), "-seed", str(seed)))
            nv, nu, m = (int(x) for x in c[:3])
            k = f"{name}__{st}_{d}_{seed}__"
            rec[k + "sizes"] = np.array([nv, nu, m], np.int64)
            rec[k + "csr"] = c[3:].astype(np.uint32)  # offsetsV, edgesV, offsetsU, edgesU
            rec[k + "totals"] = p  # total of the sparsified graph, scaled total
    np.savez_compressed(os.path.join(HERE, "sparse_cases.npz"), **rec)
    print("sparse_cases", len(rec) // 3)


SIDE_CASES = {
    # case -> whether the counts are stored too.  hub_wrap: the per-vertex term wraps and flips
    # the side choice (tests/graphs.py); the headline graph: hub degrees up to 2^18, the
    # wrapped sum is the reference's num_wedges
    "hub_wrap": True,
    "chunglu_2p20x2p18_m2p24_g21_s2": False,
}


def side_case_graph(name):
    if name == "hub_wrap":
        from tests.graphs import hub_wrap_graph
        return hub_wrap_graph()
    return pb.Graph.chung_lu(1 << 20, 1 << 18, 1 << 24, 2.1, 2.1, 2)


def make_side(tmp):
    """side_cases.npz: <case>__use_v, <case>__side_wedges from the reference's cmpWedgeCounts,
    and for the small case its BATCHS / ADEG per-vertex and per-edge counts and total."""
    out_bin = os.path.join(tmp, "s.bin")
    rec = {}
    for name, counts in SIDE_CASES.items():
        g = side_case_graph(name)
        gp = os.path.join(tmp, name + ".bcsr")
        g.write_binary(gp)
        use_v, sw, _, _, _ = O.run_ref("side", gp, out_bin)
        rec[name + "__use_v"] = np.array([use_v])
        rec[name + "__side_wedges"] = np.array([sw], np.int64)
        if counts:
            for op in ("vert", "edge", "total"):
                _, _, c, _, _ = O.run_ref(op, gp, out_bin, "BATCHS", "ADEG")
                rec[name + "__" + op] = c
        print("side case", name, use_v, sw)
    np.savez_compressed(os.path.join(HERE, "side_cases.npz"), **rec)


def konect_inputs():
    """KONECT edge lists (name -> text): comment lines, 2 to 4 columns, duplicate edges, sparse
    and unsorted ids."""
    rng = np.random.default_rng(11)
    out = {}
    rows = ["% bip unweighted", "% 14 6 7"]
    for v, h in [(3, 10), (3, 11), (50, 10), (7, 2), (3, 10), (50, 2), (7, 11), (9, 40), (9, 2), (7, 10), (50, 11),
                 (3, 2), (9, 10), (7, 40)]:
        rows.append(f"{v}\t{h}")
    out["konect_small"] = "\n".join(rows) + "\n"
    rows = ["% bip weighted positive", "%  note: extra header line"]
    vs, hs = rng.integers(1, 400, 3000) * 3 + 1, rng.integers(1, 150, 3000) * 7
    for v, h, w, t in zip(vs, hs, rng.integers(1, 9, 3000), rng.integers(10 ** 9, 2 * 10 ** 9, 3000)):
        rows.append(f"{v} {h} {w} {t}")
    out["konect_weighted"] = "\n".join(rows) + "\n"
    g = pb.Graph.chung_lu(3000, 1000, 20000, 2.1, 2.3, 13)
    owner = np.repeat(np.arange(g.nv), np.diff(g.offsetsV.astype(np.int64)))
    perm = rng.permutation(g.m)
    rows = ["% sym unweighted"] + [f"{owner[i] + 1}   {g.edgesV[i] + 1}  1" for i in perm]
    out["konect_chunglu"] = "\n".join(rows) + "\n"
    return out


def make_konect(tmp):
    """<name>.tsv (the input) and <name>.adj (the reference's KONECTtoHyperAdj output)."""
    conv = os.path.join(ROOT, "oracle", "_ref", "konect_to_hyperadj")
    for name, text in konect_inputs().items():
        src = os.path.join(HERE, name + ".tsv")
        with open(src, "w") as f:
            f.write(text)
        subprocess.run([conv, src, os.path.join(HERE, name + ".adj")], check=True, stdout=subprocess.DEVNULL)
        print("konect", name)


def main():
    O.build()
    if not O.ref_available():
        raise SystemExit("oracle/_ref/ref_harness is not built (needs the reference tree)")
    tmp = tempfile.mkdtemp()
    if "--side-only" in sys.argv:
        make_side(tmp)
        return
    if "--konect-only" in sys.argv:
        make_konect(tmp)
        return
    make_konect(tmp)
    make_sparse(tmp)
    make_side(tmp)
    if "--sparse-only" in sys.argv:
        return
    out_bin = os.path.join(tmp, "o.bin")
    for name, g in graphs():
        adj = os.path.join(HERE, name + ".adj")
        g.write(adj)
        rec = {}
        use_v, sw, c, _, _ = O.run_ref("side", adj, out_bin)
        rec["use_v"] = np.array([use_v])
        rec["side_wedges"] = np.array([sw], np.int64)
        for ct in ("SERIAL", "BATCHS"):
            for rt in ("ADEG", "ACOCORE"):
                _, _, c, _, _ = O.run_ref("vert", adj, out_bin, ct, rt)
                rec[f"vert_{ct}_{rt}"] = c
                _, _, c, _, _ = O.run_ref("edge", adj, out_bin, ct, rt)
                rec[f"edge_{ct}_{rt}"] = c
                _, _, c, _, _ = O.run_ref("total", adj, out_bin, ct, rt)
                rec[f"total_{ct}_{rt}"] = c
        _, _, _, tip, _ = O.run_ref("peelv", adj, out_bin, "BATCHS", "ADEG", "BATCHS")
        rec["tip_PBATCHS"] = tip
        _, _, _, wing, _ = O.run_ref("peele", adj, out_bin, "BATCHS", "ADEG", "BATCHS")
        rec["wing_PBATCHS"] = wing
        for rt in ("ADEG", "ACOCORE"):
            _, _, order, w, _ = O.run_ref("rank", adj, out_bin, "BATCHS", rt)
            rec[f"order_{rt}"] = order.astype(np.uint32)
            rec[f"W_{rt}"] = w
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
        print(name, "m", g.m, "total", int(rec["total_SERIAL_ADEG"][0]))


if __name__ == "__main__":
    main()
