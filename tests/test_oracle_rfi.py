"""Pins of the oracle's bandpass / RFI stage (the reference's
bandpass_fitting + rfi_clipping, multibeam_dedispersion_thread.cu:224-390,
and its kernels, multibeam_dedispersion_kernel.cu:302-450): exact answers on
integer data, float64 references within rounding, hand-built bursts."""
import numpy as np
import pytest

import oracle


def test_bandpass_sum_exact_on_integers():
    rng = np.random.default_rng(1)
    x = rng.integers(0, 200, (7, 1000)).astype(np.float32)
    for col0, n in [(0, 1000), (13, 987), (100, 64), (5, 1)]:
        got = oracle.bandpass_sum(x, col0, n)
        want = x[:, col0:col0 + n].astype(np.int64).sum(1) / n
        np.testing.assert_array_equal(got, want)


def test_bandpass_sum_matches_float64_mean():
    rng = np.random.default_rng(2)
    x = (rng.standard_normal((5, 4099)) * 3 + 20).astype(np.float32)
    got = oracle.bandpass_sum(x, 3, 4090)
    want = x[:, 3:4093].astype(np.float64).mean(1)
    np.testing.assert_allclose(got, want, rtol=1e-13)


def test_bandpass_mask_neighbours():
    bp = np.arange(10, dtype=np.float64) * 0.1 + 1.0
    got = oracle.bandpass_mask(bp, [(0, 1), (4, 5), (8, 9)])
    f = lambda v: float(np.float32(v))  # stored through a float
    assert got[0] == got[1] == f((bp[2] + bp[3]) * 0.5)
    assert got[4] == got[5] == f((bp[3] + bp[6]) * 0.5)
    assert got[8] == got[9] == f((bp[7] + bp[6]) * 0.5)
    assert got[2] == bp[2] and got[7] == bp[7]


def test_poly_fit_recovers_polynomial():
    n = 512
    xs = np.arange(n) / n
    coeffs = np.array([3.0, -1.0, 0.5, 2.0, -0.25, 0.1, 0.0, 0.05])
    bp = np.polynomial.polynomial.polyval(xs, coeffs)
    c = oracle.poly_fit(bp, 8)
    np.testing.assert_allclose(c, coeffs, atol=1e-6)
    fitted = oracle.poly_eval(c, n)
    np.testing.assert_allclose(fitted, bp, rtol=1e-11)
    st = oracle.bandpass_stats(bp, fitted, 10.0, 6.0)
    assert st["rmse"] < 1e-9


def test_poly_fit_residual_orthogonal():
    rng = np.random.default_rng(3)
    n = 300
    bp = rng.standard_normal(n) + 5
    c = oracle.poly_fit(bp, 6)
    r = bp - oracle.poly_eval(c, n)
    X = np.vander(np.arange(n) / n, 6, increasing=True)
    np.testing.assert_allclose(X.T @ r, 0, atol=1e-9)


def test_bandpass_stats_constant_and_float64():
    bp = np.full(64, 7.25)
    st = oracle.bandpass_stats(bp, bp, 10.0, 6.0)
    assert st["mean"] == 7.25 and st["std"] == 0.0 and st["rmse"] == 0.0
    assert st["spectrum_hi"] == st["spectrum_lo"] == 7.25 and st["channel_thresh"] == 0.0
    rng = np.random.default_rng(4)
    bp = rng.standard_normal(1024) * 2 + 10
    fitted = bp + rng.standard_normal(1024) * 0.1
    st = oracle.bandpass_stats(bp, fitted, 10.0, 6.0)
    assert st["mean"] == pytest.approx(bp.mean(), rel=1e-6)
    assert st["std"] == pytest.approx(bp.std(), rel=1e-5)
    assert st["rmse"] == pytest.approx(np.sqrt(np.mean((bp - fitted) ** 2)), rel=1e-5)
    assert st["spectrum_hi"] == pytest.approx(bp.mean() + 6 * bp.std(), rel=1e-5)
    assert st["spectrum_lo"] == pytest.approx(bp.mean() - 6 * bp.std(), rel=1e-4)


def test_rz_conversion():
    rng = np.random.default_rng(5)
    for d in np.concatenate([rng.standard_normal(200) * 1e3, [1.5, -2.0, 0.1, -0.1, 1e-40]]):
        f = oracle.f64_to_f32_rz(d)
        assert abs(float(f)) <= abs(d)
        assert abs(float(np.nextafter(f, np.float32(np.sign(d) * np.inf)))) > abs(d) or float(f) == d
    assert oracle.f64_to_f32_rz(0.1) == np.nextafter(np.float32(0.1), np.float32(0))  # 0.1f rounds up
    assert oracle.f64_to_f32_rz(3e38 * 10) == np.finfo(np.float32).max


def _flag_case(nb, cb, burst_blocks, nchans=6, chan=3):
    x = np.full((nchans, nb * cb + 10), 5.0, np.float32)
    for b in burst_blocks:
        x[chan, 10 + b * cb:10 + (b + 1) * cb] = 100.0
    fitted = np.full(nchans, 5.0)
    return oracle.channel_flagger(x, 10, nb * cb, fitted, cb, 10.0), x, fitted


def test_channel_flagger_neighbours():
    nb, cb = 8, 128
    flags, _, _ = _flag_case(nb, cb, [4])
    assert np.flatnonzero(flags[3]).tolist() == [3, 4, 5]
    assert flags.sum() == 3
    assert np.flatnonzero(_flag_case(nb, cb, [0])[0][3]).tolist() == [0, 1]
    assert np.flatnonzero(_flag_case(nb, cb, [nb - 1])[0][3]).tolist() == [nb - 2, nb - 1]
    # the reference's `b < num_blocks - 2`: a burst in the last-but-one block
    # does not flag the last block
    assert np.flatnonzero(_flag_case(nb, cb, [nb - 2])[0][3]).tolist() == [nb - 3, nb - 2]


def test_channel_flagger_threshold_is_strict():
    cb = 64
    x = np.full((1, 4 * cb), 15.0, np.float32)  # mean exactly bp + thresh
    assert not oracle.channel_flagger(x, 0, 4 * cb, np.full(1, 5.0), cb, 10.0).any()
    x[0, :cb] = 15.5
    assert oracle.channel_flagger(x, 0, 4 * cb, np.full(1, 5.0), cb, 10.0)[0].tolist() == [1, 1, 0, 0]


def test_channel_clipper():
    nb, cb = 6, 64
    flags, x, fitted = _flag_case(nb, cb, [2])
    fitted = np.full(x.shape[0], 5.3)
    y = oracle.channel_clipper(x, 10, nb * cb, fitted, cb, flags)
    rz = oracle.f64_to_f32_rz(5.3)
    assert (y[3, 10 + cb:10 + 4 * cb] == rz).all()
    np.testing.assert_array_equal(np.delete(y, 3, 0), np.delete(x, 3, 0))
    np.testing.assert_array_equal(y[3, :10 + cb], x[3, :10 + cb])


def test_spectrum_clipper():
    rng = np.random.default_rng(6)
    x = (rng.standard_normal((32, 500)) + 10).astype(np.float32)
    x[:, 100] += 50  # broadband spike
    x[:, 300] -= 50  # broadband dip
    fitted = np.linspace(9, 11, 32)
    y, k = oracle.spectrum_clipper(x, 0, 500, fitted, 13.0, 7.0)
    assert k == 2
    want = np.array([oracle.f64_to_f32_rz(v) for v in fitted], np.float32)
    np.testing.assert_array_equal(y[:, 100], want)
    np.testing.assert_array_equal(y[:, 300], want)
    np.testing.assert_array_equal(np.delete(y, [100, 300], 1), np.delete(x, [100, 300], 1))


def test_voltage_to_power():
    rng = np.random.default_rng(7)
    v = rng.integers(-2000, 2000, (4, 50, 2)).astype(np.int16)
    x = v.view(np.float32).reshape(4, 50).copy()
    y = oracle.voltage_to_power(x, 0, 50)
    want = (v[..., 0].astype(np.int64) ** 2 + v[..., 1].astype(np.int64) ** 2).astype(np.float32)
    np.testing.assert_array_equal(y, want)
