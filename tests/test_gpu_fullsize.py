"""Known answers at the full single-beam size (1024 channels, 1500 MHz top,
300 MHz band, 64 us, 2^20 samples, 4096 DM trials of 0.2; 13255-sample
maximum delay), where the oracle cannot run whole: with a channel-only
input x[c][t] = c mod 7 every dedispersed sample is the same exact integer
sum, whatever the delays, so every sampled row must equal it bit for bit in
each storage format.  This exercises the 64-bit offsets, the row-bounded
staging and the tile edges at the real shape."""
import sys

import numpy as np
import pytest

from mdsm_amd import abi

pytestmark = pytest.mark.gpu

ROWS = [0, 1, 17, 1023, 2048, 3000, 4094, 4095]


def _check(dtype, x, expect):
    s = abi.Survey.make(1500.0, -300.0 / 1024, 1024, 64e-6, 0.0, 0.2, 4096, 1 << 20)
    plan = abi.Plan(s, dtype)
    ld = plan.min_ld
    print(f"[full size dtype {dtype}] C=1024 N=2^20 M=4096 maxshift={plan.maxshift} ld={ld}",
          file=sys.stderr, flush=True)
    din = abi.DeviceBuffer.from_numpy(np.ascontiguousarray(np.broadcast_to(x[:, None], (1024, ld))))
    dout = abi.DeviceBuffer(4096 * (1 << 20) * 4)
    plan.dedisperse(din.ptr, ld, dout.ptr)
    abi.synchronize()
    for d in ROWS:
        row = dout.download(np.float32, (1 << 20,), offset=d * (1 << 20) * 4)
        assert (row == expect).all(), (dtype, d, row.min(), row.max(), expect)


def test_full_size_channel_constant_f32():
    x = (np.arange(1024) % 7).astype(np.float32)
    _check(abi.F32, x, np.float32(x.astype(np.float64).sum()))


def test_full_size_channel_constant_f16():
    x = (np.arange(1024) % 7).astype(np.float16).view(np.uint16)
    _check(abi.F16, x, np.float32((np.arange(1024) % 7).sum()))


def test_full_size_channel_constant_u8():
    u = (np.arange(1024) % 7).astype(np.uint8)
    expect = np.float32(float(u.astype(np.int64).sum()) * (8.0 / 255.0) - 4.0 * 1024)
    _check(abi.U8, u, expect)
