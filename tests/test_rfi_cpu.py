"""Host side of the bandpass / RFI stage: the library's least-squares fit
and statistics against the oracle (numpy SVD lstsq, the method of the
reference's gsl_multifit_linear), and argument validation."""
import numpy as np
import pytest

import oracle
from mdsm_amd import abi

MASKS = [(0, 3), (100, 120), (1020, 1023)]


@pytest.mark.parametrize("ncoeffs", [1, 4, 8, 12])
def test_fit_matches_oracle(ncoeffs):
    rng = np.random.default_rng(ncoeffs)
    n = 1024
    bp = np.polynomial.polynomial.polyval(np.arange(n) / n, [30, -10, 5, 2, -2.5, 1]) + rng.standard_normal(n)
    p = abi.RfiParams.make(ncoeffs=ncoeffs, masks=MASKS)
    fitted, st = abi.bandpass_fit(bp, p)
    masked = oracle.bandpass_mask(bp, MASKS)
    want = oracle.poly_eval(oracle.poly_fit(masked, ncoeffs), n)
    # two backward-stable solvers agree to about cond(X) * eps; the
    # Vandermonde condition number grows quickly past 8 terms
    np.testing.assert_allclose(fitted, want, rtol=1e-11 if ncoeffs <= 8 else 1e-9)
    ost = oracle.bandpass_stats(masked, want, 10.0, 6.0)
    # mean and std do not depend on the fit: identical float accumulations
    assert (st["mean"], st["std"], st["spectrum_hi"], st["spectrum_lo"]) == \
        (ost["mean"], ost["std"], ost["spectrum_hi"], ost["spectrum_lo"])
    assert st["rmse"] == pytest.approx(ost["rmse"], rel=1e-6)
    assert st["channel_thresh"] == pytest.approx(ost["channel_thresh"], rel=1e-6)


def test_fit_exact_polynomial():
    n = 4096
    coeffs = [2.0, 0.5, -0.75, 0.125]
    bp = np.polynomial.polynomial.polyval(np.arange(n) / n, coeffs)
    fitted, st = abi.bandpass_fit(bp, abi.RfiParams.make(ncoeffs=4))
    np.testing.assert_allclose(fitted, bp, rtol=1e-13)
    assert st["rmse"] < 1e-12


def test_fit_rejects_bad_masks():
    bp = np.ones(64)
    lib = abi.lib()
    out = np.zeros(64)
    st = abi.BandpassStats()
    for masks in [[(5, 3)], [(0, 62)], [(1, 63)], [(10, 64)]]:
        p = abi.RfiParams.make(masks=masks)
        assert lib.mdsm_bandpass_fit(abi._ptr(bp), 64, abi.C.byref(p), abi._ptr(out), abi.C.byref(st)) \
            == abi.MDSM_ERR_ARG, masks
    p = abi.RfiParams.make(ncoeffs=65)
    assert lib.mdsm_bandpass_fit(abi._ptr(bp), 64, abi.C.byref(p), abi._ptr(out), abi.C.byref(st)) \
        == abi.MDSM_ERR_ARG


def test_rfi_argument_errors():
    lib = abi.lib()
    fake = 0x1000  # never dereferenced: validation fails first
    assert lib.mdsm_bandpass_sum(fake, 100, 4, 50, 51, fake, None) == abi.MDSM_ERR_ARG
    assert lib.mdsm_bandpass_sum(fake, 100, 0, 0, 10, fake, None) == abi.MDSM_ERR_ARG
    assert lib.mdsm_rfi_channel_flag(fake, 100, 4, 0, 100, fake, 0, 1.0, fake, None) == abi.MDSM_ERR_ARG
    assert lib.mdsm_rfi_spectrum_clip(fake, 100, 4, 0, 100, None, 1.0, 0.0, None, None) == abi.MDSM_ERR_ARG
    assert lib.mdsm_voltage_to_power(fake, 100, 70000, 0, 100, None) == abi.MDSM_ERR_ARG
