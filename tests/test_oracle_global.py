"""Pins the oracle's reference-semantics global detector (process_brute /
process_brute_clustering, multibeam_dedispersion_output.cpp:148-248), the
deterministic log / sin-cos of the synthetic generators, and the reference's
LDS halo expression.  CPU only.

The global statistics exist in two orders: "device" (the product's fixed
chunked reduction, which the GPU tests compare with bit for bit) and
"sequential" (the reference's one loop).  Both are pinned here by exact sums
on integer-valued data (every order gives the same bits), a closed form, and
the error bound of each order against math.fsum (a correctly rounded sum);
the threshold by a planted case where x - mean == rms * thresh exactly and
by numpy's vectorised comparison."""
import math

import numpy as np
import pytest

import oracle as O

EPS = 2.0 ** -53


def _exact_stats(x):
    """(mean, rms) from exact integer sums of integer-valued data: the sums
    are exact in any order, so only the final divisions and sqrt round."""
    xi = x.astype(np.int64)
    s1, s2, n = int(xi.sum()), int((xi * xi).sum()), float(x.size)
    mean = float(s1) / n
    return mean, math.sqrt(float(s2) / n - mean * mean)


@pytest.mark.parametrize("shape,sc", [((3, 70000), None), ((5, 200), None), ((2, 140000), 131073),
                                      ((7, 65536), 65535), ((1, 1), None)])
@pytest.mark.parametrize("order", ["device", "sequential"])
def test_global_stats_exact_on_integer_data(shape, sc, order):
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    x = rng.integers(-1000, 1001, shape).astype(np.float32)
    got = O.global_stats(x, sc, order)
    want = _exact_stats(x if sc is None else x[:, :sc])
    assert got == want, (got, want)


@pytest.mark.parametrize("order", ["device", "sequential"])
def test_global_stats_closed_form(order):
    # +a and -a in equal numbers, spread over several 65536-sample chunks
    a = 3.25
    x = np.full((4, 100000), a, np.float32)
    x[:, 1::2] = -a
    assert O.global_stats(x, None, order) == (0.0, a)


def test_global_stats_error_bounds_against_fsum():
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((3, 140000)) + 5.0).astype(np.float32)
    v = x.astype(np.float64).ravel()
    n = v.size
    s1, s2 = math.fsum(v), math.fsum(v * v)  # v*v is exact: squares of floats
    a1, a2 = float(np.abs(v).sum()), float((v * v).sum())
    m_dev, r_dev = O.global_stats(x, None, "device")
    m_seq, r_seq = O.global_stats(x, None, "sequential")
    # device order: 256 sequential adds per lane, 8 tree levels, then the
    # partials (6 here: one add each) and 8 more levels
    depth_dev = 256 + 8 + 1 + 8
    assert abs(m_dev * n - s1) <= (depth_dev + 2) * EPS * a1
    assert abs(m_seq * n - s1) <= (n + 2) * EPS * a1
    mean = s1 / n
    var = s2 / n - mean * mean
    for m, r, depth in ((m_dev, r_dev, depth_dev), (m_seq, r_seq, n)):
        assert abs(r * r - var) <= (depth + 8) * EPS * (a2 / n) * 4
    # the two orders agree to rounding but are different computations
    assert abs(m_dev - m_seq) <= 2 * (n + depth_dev) * EPS * a1 / n


def test_global_stats_prefix_and_stride_read_only_their_columns():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((6, 70001)).astype(np.float32)
    sc = 66000
    # the prefix of a strided array reduces in the same order as the same
    # columns copied into a contiguous array
    assert O.global_stats(x, sc) == O.global_stats(np.ascontiguousarray(x[:, :sc]))
    y = x.copy()
    y[:, sc:] = 1e30  # columns past the prefix must not enter the sums
    assert O.global_stats(y, sc) == O.global_stats(x, sc)


def test_threshold_planted_equality():
    # 72 values: one +6, one -6, 70 zeros: mean 0, rms sqrt(72/72) = 1 exactly
    x = np.zeros((8, 9), np.float32)
    x[5, 7], x[2, 1] = 6.0, -6.0
    mean, rms = O.global_stats(x)
    assert (mean, rms) == (0.0, 1.0)
    d, t, snr, n = O.threshold_ref(x, mean, rms, 6.0)
    assert n == 1 and (d[0], t[0], snr[0]) == (5, 7, np.float32(6.0))  # x - mean == rms*thr counts
    above = float(np.nextafter(np.float32(6.0), np.float32(7.0)))
    assert O.threshold_ref(x, mean, rms, above)[3] == 0
    # the reference compares in double against the float threshold promoted
    # to double: a threshold of 0.1 (float) keeps exactly the values >= (double)0.1f
    z = np.array([[0.1, np.float32(0.1), 0.10000001]], np.float32)
    d, t, _, n = O.threshold_ref(z, 0.0, 1.0, 0.1)
    assert n == 3


def test_threshold_matches_numpy_vectorised():
    rng = np.random.default_rng(11)
    x = (rng.standard_normal((40, 3000)) * 2 + 1).astype(np.float32)
    x[17, 2900] = 30.0
    x[0, 0] = 25.0
    mean, rms = O.global_stats(x)
    thr = np.float32(3.5)
    d, t, snr, n = O.threshold_ref(x, mean, rms, float(thr))
    v = x.astype(np.float64)
    hit = (v - mean) >= rms * np.float64(thr)
    wd, wt = np.nonzero(hit)
    assert n == len(wd) and np.array_equal(d, wd) and np.array_equal(t, wt)
    assert np.array_equal(snr, ((v[wd, wt] - mean) / rms).astype(np.float32))


def test_det_log_and_sincos_accuracy():
    rng = np.random.default_rng(2)
    for x in list(rng.uniform(0, 1, 2000)) + [2.0 ** -25, 0.5, 1.0, 0.7071067811865476, 1e-300]:
        assert abs(O.det_log(x) - math.log(x)) <= 4 * EPS * abs(math.log(x)) + 1e-300
    assert O.det_log(1.0) == 0.0 and O.det_log(0.5) == -0.6931471805599453
    for u in list(rng.uniform(0, 1, 2000)) + [0.0, 0.125, 0.25, 0.5, 0.75, 0.875, 1.0]:
        s, c = O.det_sincos_turn(u)
        assert abs(s - math.sin(2 * math.pi * u)) <= 8 * EPS
        assert abs(c - math.cos(2 * math.pi * u)) <= 8 * EPS
    # exact quarter turns
    assert O.det_sincos_turn(0.0) == (0.0, 1.0)
    assert O.det_sincos_turn(0.25) == (1.0, -0.0)
    assert O.det_sincos_turn(0.5) == (-0.0, -1.0)


def test_shared_shift_hand_values():
    # round(k*(lowdm + M*step)/tsamp) - round(k*(lowdm + (M - 32)*step)/tsamp)
    # (multibeam_dedispersion_thread.cu:542-543), float products, round half away
    assert O.shared_shift_ref(0.5, 0.0, 0.25, 64, 0.5, 32) == 8       # 16 - 8, exact
    assert O.shared_shift_ref(0.75, 0.0, 0.5, 10, 1.0, 4) == 2        # round(3.75) - round(2.25)
    assert O.shared_shift_ref(0.5, 0.0, 1.0, 5, 1.0, 2) == 1          # round(2.5) - round(1.5): 3 - 2
    assert O.shared_shift_ref(0.5, 1.0, 1.0, 5, 1.0, 2) == 1          # round(3.0) - round(2.0)
