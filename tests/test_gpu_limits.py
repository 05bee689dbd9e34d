"""Memory bounds of the counting passes (pbf_set_limits, the reference's -m / -a): counting in
rank ranges of bounded wedges, and a capped HBM histogram pool, must not change any count
(bit-exact against the oracle)."""
import numpy as np
import pytest

import parbutterfly_amd as pb
from oracle import oracle as O
from tests.graphs import hbm_graph

pytestmark = pytest.mark.gpu


@pytest.fixture
def limits(ctx):
    yield ctx
    ctx.set_limits(0, 0)


@pytest.mark.parametrize("rank_type", ["ADEG", "ACOCORE"])
def test_wedge_batches(limits, rank_type):
    ctx = limits
    # U hubs of degree up to 8884: split endpoints, workgroup and light endpoints in every range
    g = pb.Graph.chung_lu(20000, 5000, 300000, 2.1, 2.1, 3)
    use_v, _ = ctx.side(g)
    ctx.count(g, use_v, "TOTAL", "GPU", rank_type)
    W = ctx.timing()["wedges"]
    assert ctx.timing()["batches"] == 1
    ov, oe = O.count_vert(g, use_v, rank_type), O.count_edge(g, use_v, rank_type)
    for parts in (3, 11):
        ctx.set_limits(W // parts + 1)
        v = pb.as_u64(ctx.count(g, use_v, "VERT", "GPU", rank_type))
        assert ctx.timing()["batches"] >= parts
        assert np.array_equal(v, ov), parts
        assert np.array_equal(pb.as_u64(ctx.count(g, use_v, "EDGE", "GPU", rank_type)), oe), parts
        assert ctx.count(g, use_v, "TOTAL", "GPU", rank_type) == int(ov.sum()) // 2, parts


def test_wedge_bound_below_one_endpoint(limits):
    # an endpoint with more wedges than the bound on its own cannot be counted
    ctx = limits
    g = pb.Graph.chung_lu(3000, 2000, 30000, 2.2, 2.3, 5)
    use_v, _ = ctx.side(g)
    ctx.set_limits(1)
    with pytest.raises(pb.PbfError) as e:
        ctx.count(g, use_v, "VERT")
    assert e.value.name == "PBF_ERR_CAPACITY"
    ctx.set_limits(0)
    assert np.array_equal(pb.as_u64(ctx.count(g, use_v, "VERT")), O.count_vert(g, use_v))


def test_array_bound(limits):
    # histograms for exactly one endpoint at a time (HBM-histogram path), then batches as well
    ctx = limits
    g = hbm_graph(45000, 5000)
    ref = O.count_vert(g, True)
    ctx.set_limits(0, 2 * (g.nv + g.nu))
    v = pb.as_u64(ctx.count(g, True, "VERT", "GPU", "SIDE"))
    assert ctx.timing()["hub_vertices"] >= 1
    assert np.array_equal(v, ref)
    ctx.count(g, True, "TOTAL", "GPU", "SIDE")
    W = ctx.timing()["wedges"]
    ctx.set_limits(W // 4 + 1, 2 * (g.nv + g.nu))
    v = pb.as_u64(ctx.count(g, True, "VERT", "GPU", "SIDE"))
    assert ctx.timing()["batches"] >= 2
    assert np.array_equal(v, ref)
