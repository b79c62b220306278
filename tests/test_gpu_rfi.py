"""Bandpass / RFI kernels on the GPU against the oracle: bit-exact sums,
flags, clipped data and powers (the kernels keep the reference's
reduction orders), and the whole stage (mdsm_rfi_clean) against the oracle
chain, where only the host fit differs (SVD vs QR, ~1e-12)."""
import numpy as np
import pytest

import oracle
from mdsm_amd import abi

pytestmark = pytest.mark.gpu


def _rows(rng, nchans, ld, burst=True):
    x = (rng.standard_normal((nchans, ld)) * 2 + 20).astype(np.float32)
    x += np.linspace(0, 5, nchans, dtype=np.float32)[:, None]  # sloped bandpass
    if burst:
        x[3, 700:1500] += 40  # narrow-band burst
        x[:, 2000] += 30  # broadband spikes
        x[:, 2500:2503] -= 30
    return x


def _upload(x):
    return abi.DeviceBuffer.from_numpy(np.ascontiguousarray(x))


@pytest.mark.parametrize("nchans,ld,col0,n", [(64, 4200, 100, 4096), (7, 3000, 0, 2999), (130, 5000, 37, 4963)])
def test_bandpass_sum_bit_exact(nchans, ld, col0, n):
    rng = np.random.default_rng(nchans)
    x = _rows(rng, nchans, ld)
    d_x = _upload(x)
    d_bp = abi.DeviceBuffer(8 * nchans)
    print(f"bandpass_sum nchans={nchans} n={n}", flush=True)
    abi.check(abi.lib().mdsm_bandpass_sum(d_x.ptr, ld, nchans, col0, n, d_bp.ptr, None))
    abi.synchronize()
    np.testing.assert_array_equal(d_bp.download(np.float64, (nchans,)), oracle.bandpass_sum(x, col0, n))


@pytest.mark.parametrize("cb", [64, 100, 1024])
def test_channel_flag_and_clip_bit_exact(cb):
    rng = np.random.default_rng(cb)
    nchans, ld, col0, n = 48, 4300, 50, 4200
    x = _rows(rng, nchans, ld)
    fitted = oracle.bandpass_sum(x, col0, n) + rng.standard_normal(nchans) * 0.01
    thresh = np.float32(3.0)
    want_flags = oracle.channel_flagger(x, col0, n, fitted, cb, thresh)
    assert want_flags[3].any()
    d_x, d_fit = _upload(x), _upload(fitted)
    nb = (n + cb - 1) // cb
    d_flags = abi.DeviceBuffer(nchans * nb)
    d_flags.upload(np.full(nchans * nb, 7, np.uint8))  # the call must clear it
    print(f"channel_flag cb={cb}", flush=True)
    lib = abi.lib()
    abi.check(lib.mdsm_rfi_channel_flag(d_x.ptr, ld, nchans, col0, n, d_fit.ptr, cb, thresh, d_flags.ptr, None))
    abi.synchronize()
    flags = d_flags.download(np.uint8, (nchans, nb))
    np.testing.assert_array_equal(flags, want_flags)
    abi.check(lib.mdsm_rfi_channel_clip(d_x.ptr, ld, nchans, col0, n, d_fit.ptr, cb, d_flags.ptr, None))
    abi.synchronize()
    np.testing.assert_array_equal(d_x.download(np.float32, (nchans, ld)),
                                  oracle.channel_clipper(x, col0, n, fitted, cb, want_flags))


def test_spectrum_clip_bit_exact():
    rng = np.random.default_rng(11)
    nchans, ld, col0, n = 96, 3300, 200, 3000
    x = _rows(rng, nchans, ld)
    fitted = oracle.bandpass_sum(x, col0, n)
    hi, lo = np.float32(np.mean(fitted) + 1.0), np.float32(np.mean(fitted) - 1.0)
    want, k = oracle.spectrum_clipper(x, col0, n, fitted, hi, lo)
    assert k >= 4
    d_x, d_fit = _upload(x), _upload(fitted)
    d_cnt = abi.DeviceBuffer(4)
    d_cnt.zero()
    abi.check(abi.lib().mdsm_rfi_spectrum_clip(d_x.ptr, ld, nchans, col0, n, d_fit.ptr, hi, lo, d_cnt.ptr, None))
    abi.synchronize()
    np.testing.assert_array_equal(d_x.download(np.float32, (nchans, ld)), want)
    assert int(d_cnt.download(np.uint32, (1,))[0]) == k


def test_voltage_to_power_bit_exact():
    rng = np.random.default_rng(12)
    v = rng.integers(-32768, 32768, (5, 1000, 2)).astype(np.int16)
    x = v.view(np.float32).reshape(5, 1000).copy()
    d_x = _upload(x)
    abi.check(abi.lib().mdsm_voltage_to_power(d_x.ptr, 1000, 5, 10, 900, None))
    abi.synchronize()
    got = d_x.download(np.uint32, (5, 1000))
    want = oracle.voltage_to_power(x, 10, 900)
    np.testing.assert_array_equal(got, want.view(np.uint32))


@pytest.mark.parametrize("apply_clip", [False, True])
def test_rfi_clean_matches_oracle_chain(apply_clip):
    rng = np.random.default_rng(13)
    nchans, ld, col0, n, cb = 64, 4600, 256, 4096, 512
    x = _rows(rng, nchans, ld)
    masks = [(10, 12)]
    params = abi.RfiParams.make(ncoeffs=8, channel_block=cb, channel_thresh=3.0, spectrum_thresh=4.0,
                                masks=masks, apply_channel_clip=apply_clip)
    fitted, st = oracle.bandpass_fit(x, col0, n, 8, masks, 3.0, 4.0)
    flags = oracle.channel_flagger(x, col0, n, fitted, cb, st["channel_thresh"])
    y = oracle.channel_clipper(x, col0, n, fitted, cb, flags) if apply_clip else x
    want, k = oracle.spectrum_clipper(y, col0, n, fitted, st["spectrum_hi"], st["spectrum_lo"])
    assert k >= 4 and flags.any()
    d_x = _upload(x)
    d_fit = abi.DeviceBuffer(8 * nchans)
    nb = (n + cb - 1) // cb
    d_flags = abi.DeviceBuffer(nchans * nb)
    got_st = abi.rfi_clean(d_x.ptr, ld, nchans, col0, n, params, d_fit.ptr, d_flags.ptr)
    np.testing.assert_allclose(d_fit.download(np.float64, (nchans,)), fitted, rtol=1e-11)
    assert got_st["mean"] == st["mean"] and got_st["std"] == st["std"]
    np.testing.assert_array_equal(d_flags.download(np.uint8, (nchans, nb)), flags)
    got = d_x.download(np.float32, (nchans, ld))
    # replacement values are rz(fitted): the two fits may differ in the last
    # bit of a few channels; everything else is bit-identical
    np.testing.assert_array_max_ulp(got, want, maxulp=1)
    assert ((got == want) | (np.abs(got - want) <= np.spacing(np.abs(want)))).all()
