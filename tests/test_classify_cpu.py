"""Reference cluster classification (Cluster::computeTransientProbability,
dbscan.cpp:25-171): the library against the oracle on random clusters
(identical scores and cluster properties), and the oracle against
hand-derived cases: the zero-DM and width rules, the reference's
asymmetric histogram smoothing, the median of tied peak bins; and the whole
score against a closed form: for a cluster whose smoothed DM-S/N curve is
a multiple of the dispersion-smearing response of Cordes & McLaughlin
(2003, ApJ 596, 1142, eq. 12: sqrt(pi)/2 erf(z)/z with
z = 6.91e-3 dDM dnu_MHz / (W_ms nu_GHz^3)), the score is 0 when only the
peak bin passes the 0.9 cut, and (1/m - 1)^2 sum(f^2)/N when the bins
above the cut have mean m."""
import numpy as np
import pytest

import oracle
from mdsm_amd import abi

BAND = dict(fch1=1500.0, foff=-0.29296875, nchans=1024)


def _both(t, d, s, min_dm=0.0, step=0.2, ndm=600, mpw=1.0):
    a = abi.classify_cluster_ref(t, d, s, min_dm, step, ndm, mpw, **BAND)
    b = oracle.classify_ref(t, d, s, min_dm, step, ndm, mpw, **BAND)
    return a, b


@pytest.mark.parametrize("seed", range(12))
def test_library_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(4, 300))
    dm0 = rng.uniform(0, 100)
    d = np.round((dm0 + rng.normal(0, rng.uniform(0.2, 5), n)).clip(0, 119) / 0.2) * 0.2
    t = 1.0 + rng.uniform(0, rng.uniform(1e-4, 0.05), n)
    s = rng.uniform(6, 30, n)
    (sa, ia), (sb, ib) = _both(t, d, s)
    assert sa == sb
    np.testing.assert_array_equal(ia, ib)


def test_zero_dm_rule():
    d = np.array([0.0, 0.2, 0.4, 0.2, 0.2])
    (score, info), _ = _both(np.linspace(1, 1.01, 5), d, np.full(5, 10.0))
    assert score == 1.0 and info[0] <= 2.0


def test_width_rule():
    # all points at one time: zero width is not above the minimum pulse width
    d = np.array([40.0, 40.2, 40.4, 40.2])
    (score, info), _ = _both(np.full(4, 2.5), d, np.array([8.0, 12.0, 9.0, 11.0]))
    assert score == 1.0 and info[2] == 0.0
    # 10 ms of extent -> width 0.4 * 10 ms = 4 ms (> 1 ms): scored
    (score, info), _ = _both(np.array([2.5, 2.51, 2.505, 2.5]), d, np.array([8.0, 12.0, 9.0, 11.0]))
    assert info[2] == pytest.approx(4.0, rel=1e-5) and score != 1.0


def test_asymmetric_smoothing_picks_the_peak_bin():
    # raw histogram: bin 100 -> 9, bin 101 -> 3.  The reference smooths with
    # (h[i] + h[i] + h[i+1]) / 3: bin 100 -> 7, 101 -> 2, 99 -> 3, so the peak
    # set is {100} alone and the peak DM is 20.0 (a centred 3-tap mean would
    # give two bins, 100 and 101, and a peak DM of 20.1).
    t = np.array([1.0, 1.002, 1.004])
    d = np.array([20.0, 20.0, 20.2])
    s = np.array([4.0, 5.0, 3.0])
    score, info = oracle.classify_ref(t, d, s, 0.0, 0.2, 600, 0.1, **BAND)
    assert info[0] == pytest.approx(20.0) and info[1] == 5.0


def test_even_number_of_peak_bins_takes_the_midpoint():
    # two separated bins of equal S/N: each smooths to (2h + 0)/3, equal peaks
    t = np.array([1.0, 1.01, 1.0, 1.01])
    d = np.array([30.0, 30.0, 31.0, 31.0])
    s = np.array([5.0, 5.0, 5.0, 5.0])
    _, info = oracle.classify_ref(t, d, s, 0.0, 0.2, 600, 0.1, **BAND)
    assert info[0] == pytest.approx(30.5, abs=1e-5)


def _smearing_cluster(width_ms, peak_bin, ndm=600, step=0.2, amp=10.0):
    """Points whose histogram, after the reference's smoothing
    (2 h[i] + h[i+1]) / 3 on bins 1..N-2, equals amp * f with f the
    Cordes-McLaughlin response around peak_bin (f = 1 at the peak), and whose
    time extent at the peak DM is width_ms / 0.4."""
    from scipy.special import erf
    band = abs(BAND["foff"] * BAND["nchans"])
    nu3 = (1e-3 * BAND["fch1"] - 0.5e-3 * band) ** 3
    x = (np.arange(ndm) - peak_bin) * step
    z = 6.91e-3 * x * band / (width_ms * nu3)
    with np.errstate(invalid="ignore", divide="ignore"):
        f = np.where(z == 0, 1.0, np.sqrt(np.pi) / 2 * erf(z) / z)
    S = amp * f
    h = np.empty(ndm)
    h[ndm - 1] = S[ndm - 1]
    for i in range(ndm - 2, 0, -1):     # invert the smoothing from the top bin down
        h[i] = (3 * S[i] - h[i + 1]) / 2
    h[0] = S[0]
    assert (h > 0).all()
    dt = width_ms / 0.4 * 1e-3
    t0 = 1.0
    times, dms, snrs = [t0, t0 + dt], [peak_bin * step] * 2, [h[peak_bin] / 2] * 2
    for i in range(ndm):
        if i != peak_bin:
            times.append(t0 + dt / 2)
            dms.append(i * step)
            snrs.append(h[i])
    return np.array(times), np.array(dms), np.array(snrs), f


@pytest.mark.parametrize("width_ms,peak_bin", [(0.2, 250), (0.25, 77), (2.0, 250), (3.0, 400)])
def test_score_closed_form_on_smearing_curve(width_ms, peak_bin):
    ndm, step = 600, 0.2
    t, d, s, f = _smearing_cluster(width_ms, peak_bin, ndm, step)
    above = f > 0.9
    # keep the cut away from rounding: no bin within 1e-3 of 0.9
    assert np.abs(f - 0.9).min() > 1e-3
    m = f[above].mean()
    want = (1 / m - 1) ** 2 * (f ** 2).sum() / ndm
    for score, info in (oracle.classify_ref(t, d, s, 0.0, step, ndm, 0.1, **BAND),
                        abi.classify_cluster_ref(t, d, s, 0.0, step, ndm, 0.1, **BAND)):
        assert info[0] == pytest.approx(peak_bin * step, abs=1e-4)
        assert info[2] == pytest.approx(width_ms, rel=1e-6)
        if above.sum() == 1:
            assert abs(score) < 1e-10, score
        else:
            assert score == pytest.approx(want, rel=1e-4, abs=1e-10), (score, want)
