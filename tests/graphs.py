"""Graph fixtures shared by the tests (generated with the library's seeded generators)."""
import numpy as np

import parbutterfly_amd as pb


def complete_bipartite(a, b):
    v = np.repeat(np.arange(a, dtype=np.uint32), b)
    u = np.tile(np.arange(b, dtype=np.uint32), a)
    return pb.Graph.from_edges(a, b, v, u)


def empty_graph(nv, nu):
    return pb.Graph.from_edges(nv, nu, np.zeros(0, np.uint32), np.zeros(0, np.uint32))


def small_graphs():
    """(tag, graph) pairs small enough for every oracle function, brute force included."""
    out = []
    for s in range(4):
        out.append((f"uniform-30x25-s{s}", pb.Graph.uniform(30, 25, 150, s)))
        out.append((f"chunglu-60x40-s{s}", pb.Graph.chung_lu(60, 40, 200, 2.1, 2.1, s)))
    out.append(("K_4_5", complete_bipartite(4, 5)))
    out.append(("K_1_7", complete_bipartite(1, 7)))
    out.append(("rmat-6x6", pb.Graph.rmat(6, 6, 600, 0.5, 0.1, 0.1, 0.3, 3)))
    return out


def hub_wrap_graph():
    """A graph on which cmpWedgeCounts' 32-bit C(d,2) wraps and flips the side choice
    (butterfly_utils.h:1061-1086): V vertex 0 is adjacent to all 65537 U vertices (its term
    65537 * 65536 wraps to 65536 modulo 2^32, i.e. C(d,2) = 32768 instead of 2147516416), and
    V vertices 1..999 are adjacent to U vertices 0 and 1.  Wrapped, wV = 32768 + 999 <= wU =
    2 * C(1000, 2), so the counts live on U; exactly, wV > wU.  Butterflies: every pair of V
    vertices with U vertices 0 and 1, C(1000, 2) in all."""
    nu = 65537
    v = np.concatenate([np.zeros(nu, np.uint32), np.repeat(np.arange(1, 1000, dtype=np.uint32), 2)])
    u = np.concatenate([np.arange(nu, dtype=np.uint32), np.tile(np.array([0, 1], np.uint32), 999)])
    return pb.Graph.from_edges(1000, nu, v, u)


def wide_centre_graph(k=69999):
    """Two adjacent hubs of degree k + 1: V vertex 0 on U vertices 0..k, U vertex 0 on V
    vertices 0..k, plus the edges (j, j) for j = 1..k.  The butterflies are {0, j} x {0, j}, one
    per j: edge (0, 0) lies in k of them, every other edge in one.  Under ADEG the two hubs rank
    first and the U hub is the centre of a wedge from every V vertex j; with k + 1 > 65536,
    deg(hub) - 1 times the largest degree - 1 reaches 2^32, so the per-edge pass credits those
    chunks in 64 bits."""
    j = np.arange(1, k + 1, dtype=np.uint32)
    z = np.zeros(k, np.uint32)
    v = np.concatenate([np.zeros(1, np.uint32), z, j, j])
    u = np.concatenate([np.zeros(1, np.uint32), j, z, j])
    return pb.Graph.from_edges(k + 1, k + 1, v, u)


def hbm_graph(keys, pad):
    """S = U (wV <= wU) under SIDE ranking: U vertex `keys + pad` shares the V hub v0 with U
    vertices 0..keys-1, so as an endpoint it has `keys` distinct keys at a rank above 16384
    (beyond every LDS table: the HBM-histogram path).  A U hub with keys + 1 leaves in V makes
    wU > wV."""
    i = keys + pad
    v = [0] * (keys + 1)
    u = list(range(keys)) + [i]
    leaves = keys + 1
    v += list(range(1, 1 + leaves))
    u += [i + 1] * leaves
    return pb.Graph.from_edges(1 + leaves, i + 2, np.array(v, np.uint32), np.array(u, np.uint32))


def two_hubs_with_leaves(hub, leaves):
    """V vertices 0 and 1 adjacent to U vertices 0..hub-1; U vertex j also adjacent to `leaves`
    private V vertices 2 + leaves * j + t (degree 1).  The butterflies are {0, 1} x {j, k}: a hub
    edge lies in hub - 1 of them, a leaf edge in none."""
    j = np.arange(hub, dtype=np.uint32)
    v = [np.zeros(hub, np.uint32), np.ones(hub, np.uint32)]
    u = [j, j]
    for t in range(leaves):
        v.append(2 + leaves * j + t)
        u.append(j)
    return pb.Graph.from_edges(2 + leaves * hub, hub, np.concatenate(v).astype(np.uint32), np.concatenate(u).astype(np.uint32))
