"""Mu-law quantiser kernels against the oracle (bit exact), with the
reference's first-buffer maximum kept across calls."""
import ctypes as C

import numpy as np
import pytest

import oracle
from mdsm_amd import abi

pytestmark = pytest.mark.gpu


class MulawState(C.Structure):
    _fields_ = [("bits", C.c_uint32), ("compression", C.c_uint32), ("max_value", C.c_float)]


lib = abi.lib()
lib.mdsm_mulaw_encode_power.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(MulawState), C.c_void_p, C.c_void_p]
lib.mdsm_mulaw_encode_complex.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]


@pytest.mark.parametrize("bits,mu", [(8, 255), (4, 100), (2, 1)])
def test_power_matches_oracle(bits, mu):
    rng = np.random.default_rng(bits)
    first = (rng.standard_normal(100000) ** 2 * 3).astype(np.float32)
    second = (rng.standard_normal(70001) ** 2 * 5).astype(np.float32)  # may exceed the first max x2
    st = MulawState(bits, mu, 0.0)
    outs = []
    for x in (first, second):
        d_x = abi.DeviceBuffer.from_numpy(x)
        d_o = abi.DeviceBuffer(x.size)
        abi.check(lib.mdsm_mulaw_encode_power(d_x.ptr, x.size, C.byref(st), d_o.ptr, None))
        abi.synchronize()
        outs.append(d_o.download(np.uint8, (x.size,)))
    assert st.max_value == np.float32(first.max()) * 2
    np.testing.assert_array_equal(outs[0], oracle.mulaw_power(first, bits, mu, st.max_value))
    np.testing.assert_array_equal(outs[1], oracle.mulaw_power(second, bits, mu, st.max_value))


def test_complex_matches_oracle():
    rng = np.random.default_rng(3)
    pairs = rng.integers(-32768, 32768, (50001, 2)).astype(np.int16)
    pairs[:3] = [[-32768, -32768], [0, 0], [32767, -1]]
    d_x = abi.DeviceBuffer.from_numpy(pairs)
    d_o = abi.DeviceBuffer(len(pairs))
    abi.check(lib.mdsm_mulaw_encode_complex(d_x.ptr, len(pairs), 3, 255, d_o.ptr, None))
    abi.synchronize()
    np.testing.assert_array_equal(d_o.download(np.uint8, (len(pairs),)), oracle.mulaw_complex(pairs, 3, 255))
