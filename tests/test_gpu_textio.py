"""The device reader (readBipartite, apps/hyper/butterfly_utils.h:1008-1058, parsed and validated
on the GPU) against the host reader: identical arrays for every golden input and for binary
files, identical error codes for malformed files (one fault per file, and files with several
faults, where both must report the one met first in file order), and counts from the
device-read graph equal to those of the host-read one."""
import os

import numpy as np
import pytest

import parbutterfly_amd as pb

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["uniform_1000x1000_m10000_s1", "chunglu_300x200_m2000_g21_s7", "chunglu_2000x500_m12000_g21_s2",
         "rmat_8x8_m3000_s3", "k_4_5"]


def _same(h, d):
    assert (h.nv, h.nu, h.m) == (d.nv, d.nu, d.m)
    for k in ("offsetsV", "edgesV", "offsetsU", "edgesU"):
        assert np.array_equal(np.asarray(getattr(h, k)), getattr(d, k).numpy()), k


def _status(fn):
    try:
        fn()
    except pb.PbfError as e:
        return e.name
    return "PBF_OK"


def test_golden_files(ctx, tmp_path):
    for name in NAMES:
        print("case", name, flush=True)
        p = os.path.join(GOLDEN, name + ".adj")
        h = pb.Graph.read(p)
        d = pb.DeviceGraph.read(ctx, p)
        _same(h, d)
        use_v, _ = ctx.side(d)
        assert np.array_equal(pb.as_u64(ctx.count(d, use_v, "VERT")), pb.as_u64(ctx.count(h, use_v, "VERT")))
        b = str(tmp_path / (name + ".bcsr"))
        h.write_binary(b)
        _same(h, pb.DeviceGraph.read(ctx, b))


def test_generated_with_odd_spacing(ctx, tmp_path):
    # tokens separated by runs of mixed whitespace (spaces, tabs, CR, NUL), files whose size is
    # not a multiple of the 4096-byte tile, and tokens straddling tile boundaries
    g = pb.Graph.chung_lu(5000, 3000, 60000, 2.1, 2.1, 3)
    p = str(tmp_path / "g.adj")
    g.write(p)
    text = open(p, "rb").read()
    seps = [b" ", b"\t", b"\r\n", b"  \x00 ", b"\n\n\n"]
    toks = text.split()
    rng = np.random.default_rng(1)
    out = bytearray()
    for i, t in enumerate(toks):
        out += t + seps[int(rng.integers(len(seps)))]
    q = str(tmp_path / "odd.adj")
    open(q, "wb").write(bytes(out))
    h = pb.Graph.read(q)
    _same(h, pb.DeviceGraph.read(ctx, q))
    # the full headline-sized text file goes through the same kernels
    big = pb.Graph.chung_lu(1 << 18, 1 << 16, 1 << 21, 2.1, 2.1, 2)
    r = str(tmp_path / "big.adj")
    big.write(r)
    _same(big, pb.DeviceGraph.read(ctx, r))


CASES = {
    "bad header": "AdjacencyGraph\n1\n1\n1\n1\n0\n0\n0\n0\n",
    "mv != mu": "AdjacencyHypergraph\n1\n1\n1\n2\n0\n0\n0\n0\n0\n",
    "missing token": "AdjacencyHypergraph\n1\n1\n1\n1\n0\n0\n0\n",
    "trailing token": "AdjacencyHypergraph\n1\n1\n1\n1\n0\n0\n0\n0\n7\n",
    "non-numeric": "AdjacencyHypergraph\n1\n1\n1\n1\n0\nx\n0\n0\n",
    "range V": "AdjacencyHypergraph\n1\n1\n1\n1\n0\n5\n0\n0\n",
    "range U": "AdjacencyHypergraph\n1\n1\n1\n1\n0\n0\n0\n3\n",
    "unsorted": "AdjacencyHypergraph\n1\n2\n2\n2\n0\n1\n0\n0\n1\n0\n0\n",
    "transpose": "AdjacencyHypergraph\n1\n1\n2\n1\n0\n1\n0\n1\n0\n",
    "offsets decrease": "AdjacencyHypergraph\n3\n2\n2\n2\n0\n2\n1\n0\n1\n0\n1\n0\n0\n",
    "offset start": "AdjacencyHypergraph\n1\n1\n1\n1\n1\n0\n0\n0\n",
    # several faults: a non-numeric token after an out-of-range one, and the reverse
    "range then text": "AdjacencyHypergraph\n1\n2\n2\n2\n0\n9\nq\n0\n1\n0\n0\n",
    "text then range": "AdjacencyHypergraph\n1\n2\n2\n2\n0\nq\n9\n0\n1\n0\n0\n",
    # unsorted V list and a U-side range fault: the reader meets the U range fault first
    "unsorted then range": "AdjacencyHypergraph\n1\n2\n2\n2\n0\n1\n0\n0\n1\n0\n4\n",
    "valid": "AdjacencyHypergraph\n1\n1\n1\n1\n0\n0\n0\n0\n",
    "empty graph": "AdjacencyHypergraph\n3\n0\n2\n0\n0\n0\n0\n0\n0\n",
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_error_codes_match_host_reader(ctx, tmp_path, name):
    p = str(tmp_path / "x.adj")
    open(p, "w").write(CASES[name])
    want = _status(lambda: pb.Graph.read(p))
    got = _status(lambda: pb.DeviceGraph.read(ctx, p))
    assert got == want, (name, got, want)
    if name in ("valid", "empty graph"):
        assert want == "PBF_OK"
        _same(pb.Graph.read(p), pb.DeviceGraph.read(ctx, p))


def test_binary_errors(ctx, tmp_path):
    g = pb.Graph.uniform(300, 200, 3000, 5)
    b = str(tmp_path / "g.bcsr")
    g.write_binary(b)
    raw = bytearray(open(b, "rb").read())
    # an edgesU entry out of range, then a swapped pair in an edgesV list
    off_eU = 32 + 4 * ((g.nv + 1) + g.m + (g.nu + 1))
    bad = bytearray(raw)
    bad[off_eU + 40:off_eU + 44] = np.uint32(g.nv + 3).tobytes()
    q = str(tmp_path / "bad1.bcsr")
    open(q, "wb").write(bytes(bad))
    assert _status(lambda: pb.DeviceGraph.read(ctx, q)) == _status(lambda: pb.Graph.read(q)) == "PBF_ERR_RANGE"
    v = int(np.argmax(np.diff(g.offsetsV.astype(np.int64)) >= 2))
    a = 32 + 4 * ((g.nv + 1) + int(g.offsetsV[v]))
    bad = bytearray(raw)
    bad[a:a + 4], bad[a + 4:a + 8] = raw[a + 4:a + 8], raw[a:a + 4]
    q = str(tmp_path / "bad2.bcsr")
    open(q, "wb").write(bytes(bad))
    assert _status(lambda: pb.DeviceGraph.read(ctx, q)) == _status(lambda: pb.Graph.read(q)) == "PBF_ERR_UNSORTED"
    q = str(tmp_path / "short.bcsr")
    open(q, "wb").write(bytes(raw[:-4]))
    assert _status(lambda: pb.DeviceGraph.read(ctx, q)) == _status(lambda: pb.Graph.read(q)) == "PBF_ERR_FORMAT"
    assert _status(lambda: pb.DeviceGraph.read(ctx, str(tmp_path / "none"))) == "PBF_ERR_IO"
