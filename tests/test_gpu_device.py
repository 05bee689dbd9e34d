"""Device buffers and the device-to-device copy kernel (pbf_copy_device)."""
import numpy as np
import pytest

import parbutterfly_amd as pb

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 15, 16, 17, 4096 + 7, (1 << 20) + 3])
def test_copy_device(ctx, n):
    rng = np.random.default_rng(n)
    h = rng.integers(0, 256, n, dtype=np.uint8)
    a = pb.DeviceArray(ctx, n, np.uint8, host=h)
    b = pb.DeviceArray(ctx, n, np.uint8, host=np.zeros(n, np.uint8))
    b.copy_from(a)
    assert np.array_equal(b.numpy(), h)
