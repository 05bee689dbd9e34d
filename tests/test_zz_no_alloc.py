"""pbf_count_vert_noalloc (per-vertex counting in the library's fixed per-device scratch, no
allocation inside the call) against pbf_count_vert, bit-exact, on two calls in a row, the
second with a smaller graph, then a graph too large for the scratch (PBF_ERR_CAPACITY, the
context still working afterwards)."""
import numpy as np
import pytest

import parbutterfly_amd as pb
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _both(ctx, g, rank_type):
    dg = g.to_device(ctx)
    use_v, _ = ctx.side(dg)
    ref = pb.as_u64(ctx.count(dg, use_v, "VERT", "GPU", rank_type))
    got = pb.as_u64(ctx.count_vert_noalloc(dg, use_v, "GPU", rank_type))
    return use_v, ref, got


@pytest.mark.parametrize("rank_type", ["ADEG", "ACOCORE"])
def test_noalloc_matches_count_vert(ctx, rank_type):
    # the largest size the header promises to fit: nv + nu = 2^20, m = 2^21, power law (hubs
    # on split and HBM-fallback paths), then a smaller graph over the same array
    big = pb.Graph.chung_lu(1 << 19, 1 << 19, 1 << 21, 2.1, 2.1, 31)
    print("noalloc big", rank_type, flush=True)
    _, ref, got = _both(ctx, big, rank_type)
    assert np.array_equal(got, ref)
    small = pb.Graph.chung_lu(3000, 2000, 30000, 2.2, 2.3, 32)
    print("noalloc small", rank_type, flush=True)
    use_v, ref, got = _both(ctx, small, rank_type)
    assert np.array_equal(got, ref)
    assert np.array_equal(got, O.count_vert(small, use_v, rank_type))


def test_noalloc_capacity_and_recovery(ctx):
    # 2^24 edges need far more than the fixed array: a capacity error, nothing written
    g = pb.Graph.chung_lu(1 << 20, 1 << 18, 1 << 24, 2.1, 2.1, 2)
    dg = g.to_device(ctx)
    use_v, _ = ctx.side(dg)
    out = pb.DeviceArray(ctx, g.nu if use_v else g.nv, np.uint64, np.full(g.nu if use_v else g.nv, 7, np.uint64))
    print("noalloc capacity", flush=True)
    with pytest.raises(pb.PbfError) as e:
        ctx.count_vert_noalloc(dg, use_v, out=out)
    assert e.value.name == "PBF_ERR_CAPACITY"
    assert np.all(out.numpy() == 7)
    # host graphs are refused; the context's own path and the fixed path both still work
    with pytest.raises(TypeError):
        ctx.count_vert_noalloc(g, use_v)
    small = pb.Graph.uniform(1000, 1000, 10000, 1)
    ds = small.to_device(ctx)
    su, _ = ctx.side(ds)
    exp = O.count_vert(small, su, "ADEG")
    assert np.array_equal(pb.as_u64(ctx.count(ds, su, "VERT")), exp)
    assert np.array_equal(pb.as_u64(ctx.count_vert_noalloc(ds, su)), exp)
