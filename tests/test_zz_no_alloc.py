"""mdsm_detect_global_noalloc (the global detector on the library's fixed
device workspace) against mdsm_detect_global (per-stream workspace) through
the C ABI: statistics and candidates bit for bit on two calls in a row, the
second on a smaller region (the fixed workspace keeps the first call's
contents), plus the oracle; regions beyond the workspace are rejected before
any device work."""
import sys

import numpy as np
import pytest

import oracle as O
from mdsm_amd import abi

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _run(fn, d_y, nrows, nsamp, ld, sc, thr, cap=1 << 16):
    d_c, d_n = abi.DeviceBuffer(cap * 16), abi.DeviceBuffer(4)
    mean, rms = fn(d_y.ptr, nrows, nsamp, thr, d_c.ptr, cap, d_n.ptr, ld=ld, stats_nsamp=sc)
    return (mean, rms), abi.read_candidates(d_c, d_n, cap)


def test_noalloc_matches_detect_global_on_two_calls():
    rng = np.random.default_rng(404)
    # first call: 6 rows x 200000 samples (24 chunk partials); second call:
    # 3 rows of a strided 70001-sample layout, 69000 thresholded, statistics
    # over 66001 (6 partials)
    big = (rng.standard_normal((6, 200000)) * 2 + 1).astype(np.float32)
    big[2, 12345] += 40.0
    small = (rng.standard_normal((3, 70001)) * 3 - 2).astype(np.float32)
    small[1, 777] += 50.0
    cases = [(big, 200000, 200000, 200000, 4.5), (small, 69000, 70001, 66001, 4.5)]
    for y, nsamp, ld, sc, thr in cases:
        nrows = y.shape[0]
        _say(f"[noalloc] rows={nrows} nsamp={nsamp} ld={ld} stats={sc}")
        d_y = abi.DeviceBuffer.from_numpy(y)
        st_a, c_a = _run(abi.detect_global_noalloc, d_y, nrows, nsamp, ld, sc, thr)
        st_b, c_b = _run(abi.detect_global, d_y, nrows, nsamp, ld, sc, thr)
        assert st_a == st_b
        assert len(c_a) == len(c_b) > 0
        assert np.array_equal(c_a, c_b)
        # and the oracle's restatement of the same order
        om, orms = O.global_stats(np.ascontiguousarray(y[:, :nsamp]), sc)
        assert st_a == (om, orms)
        od, ot, osnr, on = O.threshold_ref(np.ascontiguousarray(y[:, :nsamp]), om, orms, thr)
        assert list(zip(c_a["dm_idx"], c_a["sample"])) == list(zip(od, ot))
        assert np.array_equal(c_a["snr"], osnr)


def test_noalloc_rejects_regions_beyond_its_workspace():
    # 65535 rows x 131073 samples need 65535 * 3 chunk partials (> 131072);
    # the call is rejected by its argument check, before any device work
    d_y = abi.DeviceBuffer(4096)
    d_c, d_n = abi.DeviceBuffer(16), abi.DeviceBuffer(4)
    with pytest.raises(abi.MdsmError, match="static workspace"):
        abi.detect_global_noalloc(d_y.ptr, 65535, 131073, 6.0, d_c.ptr, 1, d_n.ptr)
