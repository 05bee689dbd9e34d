"""Device buffers and the device-to-device copy kernel (pbf_copy_device)."""
import numpy as np
import pytest

import parbutterfly_amd as pb
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 15, 16, 17, 4096 + 7, (1 << 20) + 3])
def test_copy_device(ctx, n):
    rng = np.random.default_rng(n)
    h = rng.integers(0, 256, n, dtype=np.uint8)
    a = pb.DeviceArray(ctx, n, np.uint8, host=h)
    b = pb.DeviceArray(ctx, n, np.uint8, host=np.zeros(n, np.uint8))
    b.copy_from(a)
    assert np.array_equal(b.numpy(), h)


def test_caller_stream(ctx):
    # pbf_set_stream: the context's work runs on a stream the caller created; NULL returns to
    # the context's own stream
    import ctypes
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    s = ctypes.c_void_p()
    assert hip.hipStreamCreateWithFlags(ctypes.byref(s), ctypes.c_uint(1)) == 0  # hipStreamNonBlocking
    own = ctx.stream
    g = pb.Graph.chung_lu(3000, 2000, 30000, 2.2, 2.3, 12)
    try:
        ctx.set_stream(s.value)
        assert ctx.stream == s.value
        dg = g.to_device(ctx)
        use_v, _ = ctx.side(dg)
        v = pb.as_u64(ctx.count(dg, use_v, "VERT"))
        e = pb.as_u64(ctx.count(g, use_v, "EDGE"))
        tip = pb.as_u64(ctx.peel(dg, use_v, v))
    finally:
        ctx.set_stream(None)
        assert hip.hipStreamDestroy(s) == 0
    assert ctx.stream == own
    assert np.array_equal(v, O.count_vert(g, use_v))
    assert np.array_equal(e, O.count_edge(g, use_v))
    assert np.array_equal(tip, O.peel_vert(g, use_v, v))
