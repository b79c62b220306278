"""Pins the detection half of the CPU oracle: block statistics, dyadic boxcar
detection, 2-D DBSCAN (against scikit-learn), the reference 3-D DBSCAN
(hand-derived cases from the reference algorithm), median of five, legacy
detector, detrending.  CPU only."""
import itertools

import numpy as np
import pytest

import oracle as O


def _numpy_detect(x, L, nwidths, thresh):
    """Independent restatement with cumulative sums (different summation
    order from the oracle's dyadic tree)."""
    nrows, n = x.shape
    m, s = O.block_stats(x, L)
    out = []
    cs = np.concatenate([np.zeros((nrows, 1)), np.cumsum(x, axis=1)], axis=1)
    for d in range(nrows):
        best = np.full(n, -np.inf)
        bk = np.full(n, -1)
        for k in range(nwidths):
            w = 1 << k
            t = np.arange(0, n - w + 1)
            sw = cs[d, t + w] - cs[d, t]
            b = t // L
            snr = (sw - w * m[d, b]) / (s[d, b] * np.sqrt(w))
            upd = snr > best[t]
            best[t[upd]] = snr[upd]
            bk[t[upd]] = k
        for t in np.nonzero((bk >= 0) & (best >= thresh))[0]:
            out.append((d, int(t), int(bk[t]), best[t]))
    return out


def test_block_stats_integer_exact_and_numpy():
    rng = np.random.default_rng(0)
    x = rng.integers(-40, 40, (3, 1000)).astype(np.float64)
    m, s = O.block_stats(x, 256)
    assert m.shape == (3, 4)
    for d in range(3):
        for b in range(4):
            seg = x[d, b * 256:(b + 1) * 256]
            assert m[d, b] == seg.sum() / len(seg)
            assert s[d, b] == np.sqrt((seg * seg).sum() / len(seg) - m[d, b] ** 2)
    y = rng.standard_normal((2, 5000))
    m, s = O.block_stats(y, 1024)
    assert np.allclose(m[0, 4], y[0, 4096:].mean(), rtol=1e-12)
    assert np.allclose(s[1, 2], y[1, 2048:3072].std(), rtol=1e-10)


def test_detect_exact_on_integer_data():
    rng = np.random.default_rng(1)
    x = rng.integers(-3, 4, (6, 3000)).astype(np.float64)
    x[2, 1000:1016] += 6          # width-16 pulse
    x[4, 2990:2999] += 9          # near the end: widths beyond the edge skipped
    got = O.detect(x, 1024, 7, 5.0)
    ref = _numpy_detect(x, 1024, 7, 5.0)
    assert got[4] == len(ref) > 0
    assert [(a, b, c) for a, b, c, _ in ref] == list(zip(got[0].tolist(), got[1].tolist(), got[2].tolist()))
    assert np.allclose(got[3], [r[3] for r in ref], rtol=1e-6)


def test_detect_recovers_pulse_width_and_guard_band_vs_cumsum():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 8192))
    x[1, 3000:3032] += 1.5        # SNR 1.5*sqrt(32) = 8.5 at width 32
    d, t, k, snr, n = O.detect(x, 2048, 7, 6.0)
    hit = (d == 1) & (t == 3000)
    assert hit.any() and k[hit][0] == 5
    ref = _numpy_detect(x, 2048, 7, 6.0)
    got = set(zip(d.tolist(), t.tolist(), k.tolist()))
    exp = set((a, b, c) for a, b, c, _ in ref)
    # different summation order: only guard-band candidates may differ
    for (a, b, c, s) in ref:
        if (a, b, c) not in got:
            assert abs(s - 6.0) < 1e-9
    assert len(got ^ exp) <= 2


def test_dbscan2d_matches_sklearn():
    from sklearn.cluster import DBSCAN
    rng = np.random.default_rng(3)
    pts = np.concatenate([
        rng.integers(0, 400, (300, 2)),
        rng.integers(0, 6, (40, 2)) + [100, 200],
        rng.integers(0, 10, (60, 2)) + [300, 50],
        np.stack([np.arange(30) * 2, np.arange(30) * 3 + 10], 1),    # a chain
    ]).astype(np.int32)
    pts = np.unique(pts, axis=0)
    order = np.lexsort((pts[:, 1], pts[:, 0]))
    pts = pts[order]
    lab, n = O.dbscan2d(pts[:, 0], pts[:, 1], 3, 4)
    sk = DBSCAN(eps=3, min_samples=4, metric="chebyshev").fit(pts).labels_
    assert n == sk.max() + 1 > 2
    assert np.array_equal(lab, sk)


def test_dbscan_ref_hand_cases():
    T = lambda *v: np.array(v, np.float64)
    # three points inside one box, minPoints 2 (neighbours exclude self): one cluster
    lab, n = O.dbscan_ref(T(0, .1, .2), T(5, 5, 5), T(7, 7, 7), 1.0, 1.0, 1.0, 2)
    assert n == 1 and lab.tolist() == [1, 1, 1]
    # two close points, minPoints 2: point 0 has one neighbour -> noise; point 1
    # then finds point 0 already visited -> noise too
    lab, n = O.dbscan_ref(T(0, .1), T(5, 5), T(7, 7), 1.0, 1.0, 1.0, 2)
    assert n == 0 and lab.tolist() == [-1, -1]
    # strict '<': a gap equal to the range is not a neighbour, so with
    # minPoints 1 every point has 0 neighbours -> all noise; a gap just below
    # the range chains them into one cluster
    lab, n = O.dbscan_ref(T(0, 1, 2), T(5, 5, 5), T(7, 7, 7), 1.0, 1.0, 1.0, 1)
    assert n == 0 and lab.tolist() == [-1, -1, -1]
    lab, n = O.dbscan_ref(T(0, .99, 1.98), T(5, 5, 5), T(7, 7, 7), 1.0, 1.0, 1.0, 1)
    assert n == 1 and lab.tolist() == [1, 1, 1]
    # SNR is a clustering dimension: equal time/DM, SNR apart -> not neighbours
    lab, n = O.dbscan_ref(T(0, 0), T(5, 5), T(7, 9), 1.0, 1.0, 1.0, 1)
    assert lab.tolist() == [-1, -1]
    lab, n = O.dbscan_ref(T(0, 0), T(5, 5), T(7, 7.5), 1.0, 1.0, 1.0, 1)
    assert lab.tolist() == [1, 1]
    # chain A-B-C (A,C out of range), minPoints 1: expansion through B
    lab, n = O.dbscan_ref(T(0, .6, 1.2), T(5, 5, 5), T(7, 7, 7), 1.0, 1.0, 1.0, 1)
    assert n == 1 and lab.tolist() == [1, 1, 1]
    # visited-exclusion: P0-P2 and P2-P1 in range, P0-P1 not, minPoints 2.
    # P0 and P1 each see one neighbour -> noise; P2 then sees only visited
    # points -> noise.  (Textbook DBSCAN would make one cluster.)
    lab, n = O.dbscan_ref(T(0, 1.0, .5), T(5, 5, 5), T(7, 7, 7), 0.6, 1.0, 1.0, 2)
    assert n == 0 and lab.tolist() == [-1, -1, -1]


@pytest.mark.parametrize("min_points", [1, 2, 3, 4])
def test_dbscan_ref_isolated_boxes_match_sklearn(min_points):
    """Where the neighbour graph is a disjoint union of cliques (blobs that
    fit in one range box, far apart) the reference's exclusion rules do not
    matter and its DBSCAN reduces to the textbook one: a blob of more than
    min_points points is one cluster, a smaller one is noise.  scikit-learn
    on range-scaled coordinates (Chebyshev, eps 1, min_samples counting the
    point itself = min_points + 1) then gives the same partition, with the
    reference's cluster ids = scikit-learn's + 1 (both number clusters in
    order of discovery).  Different ranges per dimension pin which range
    scales which coordinate."""
    from sklearn.cluster import DBSCAN
    rng = np.random.default_rng(40 + min_points)
    ranges = np.array([0.5, 2.0, 1.5])  # time, DM, S/N
    pts = []
    for k in range(40):
        centre = np.array([k * 7.0, (k % 5) * 23.0, (k % 3) * 17.0]) * ranges
        size = int(rng.integers(1, 7))
        pts.append(centre + rng.uniform(0.0, 0.95, (size, 3)) * ranges)
    pts = np.concatenate(pts)[rng.permutation(sum(len(p) for p in pts))]
    lab, n = O.dbscan_ref(pts[:, 0], pts[:, 1], pts[:, 2], *ranges, min_points)
    sk = DBSCAN(eps=1.0, min_samples=min_points + 1, metric="chebyshev").fit(pts / ranges).labels_
    assert n == sk.max() + 1 > 0
    assert np.array_equal(lab, np.where(sk < 0, -1, sk + 1))


def _components_rule(tm, dm, snr, ranges, min_points):
    """The reference's DBSCAN (dbscan.cpp:194-276, neighbours :390-415) in
    closed form: points are taken as seeds in index order, skipping visited
    ones; a seed with fewer than min_points unvisited neighbours is noise
    (visited, and never a neighbour again); otherwise its cluster is the
    whole connected component that contains it in the neighbour graph
    restricted to the points not yet visited (every listed point expands,
    core or not), and all of them become visited.  Neighbours: the float
    distances |a - b| strictly below the float ranges in all three
    dimensions.  Components come from scipy's graph labelling, not from a
    restatement of the reference's marking loop."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    n = len(tm)
    adj = np.ones((n, n), bool)
    for v, r in zip((tm, dm, snr), ranges):
        adj &= np.abs(v[:, None] - v[None, :]).astype(np.float32) < np.float32(r)
    np.fill_diagonal(adj, False)
    lab = np.zeros(n, np.int32)
    visited = np.zeros(n, bool)
    nclusters = 0
    for i in range(n):
        if visited[i]:
            continue
        open_ = ~visited
        if (adj[i] & open_).sum() < min_points:
            visited[i] = True
            lab[i] = -1
            continue
        idx = np.flatnonzero(open_)
        _, comp = connected_components(csr_matrix(adj[np.ix_(idx, idx)]), directed=False)
        members = idx[comp == comp[np.searchsorted(idx, i)]]
        nclusters += 1
        lab[members] = nclusters
        visited[members] = True
    return lab, nclusters


@pytest.mark.parametrize("min_points", [1, 2, 3, 5])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_dbscan_ref_chained_neighbourhoods(min_points, seed):
    """Random walks in (time, DM, S/N) make chained, branching neighbourhoods
    in which border points, noise taken early and seeds that fail their
    count all occur; the oracle equals the component rule on every one."""
    rng = np.random.default_rng(100 * seed + min_points)
    ranges = (0.5, 2.0, 1.5)
    walks = []
    for w in range(12):
        steps = rng.uniform(-1.0, 1.0, (int(rng.integers(1, 25)), 3)) * np.array(ranges) * 0.8
        walks.append(np.array([w * 9.0, (w % 4) * 30.0, 5.0]) + np.cumsum(steps, axis=0))
    isolated = rng.uniform(0, 1, (15, 3)) * np.array([110.0, 120.0, 20.0])
    pts = np.concatenate(walks + [isolated])[rng.permutation(sum(map(len, walks)) + 15)]
    lab, n = O.dbscan_ref(pts[:, 0], pts[:, 1], pts[:, 2], *ranges, min_points)
    want, nw = _components_rule(pts[:, 0], pts[:, 1], pts[:, 2], ranges, min_points)
    assert n == nw and np.array_equal(lab, want)
    # the cases are not trivial: clusters, noise, and clusters of more than
    # one walk step occur
    assert n > 2 and (lab == -1).any() and np.bincount(lab[lab > 0]).max() > 3


def test_median5_network_is_a_median():
    for perm in itertools.permutations([1.0, 2.0, 3.0, 4.0, 5.0]):
        assert O.median5(*perm) == 3.0
    for perm in set(itertools.permutations([1.0, 1.0, 2.0, 3.0, 3.0])):
        assert O.median5(*perm) == 2.0
    rng = np.random.default_rng(4)
    v = rng.standard_normal((2000, 5)).astype(np.float32)
    assert all(O.median5(*map(float, r)) == float(np.float32(np.median(r))) for r in v)


def test_median_filter_vs_scipy():
    from scipy.ndimage import median_filter
    x = np.random.default_rng(5).standard_normal((3, 500)).astype(np.float32)
    assert np.array_equal(O.median_filter5(x), median_filter(x, size=(1, 5), mode="nearest"))


def test_legacy_detector_block_mean_quirk():
    # vals = 3 * 32768: blocks 0 and 1 enter the mean, the last block does not
    x = np.zeros((3, 32768), np.float32)
    x[0] = 1.0
    x[1] = 3.0
    x[2] = 100.0
    _, _, _, n, mean, rms = O.process_brute_legacy(x, 0.0, 2.0, 6.0)
    assert mean == 2.0 and rms == 1.0
    # median of five then threshold; DM must exceed 1 (row 0 has DM 0)
    y = np.zeros((2, 1000), np.float32)
    y[:, 500:503] = 50.0     # 3 of 5 high -> median high at l = 498..500
    y[:, 700] = 80.0         # a single spike is removed by the median
    d, t, s, n, mean, rms = O.process_brute_legacy(y, 0.0, 2.0, 3.0)
    assert d.tolist() == [1, 1, 1] and t.tolist() == [498, 499, 500]


def test_legacy_detector_chunk_statistics():
    # per-chunk statistics: the same 3-sample pulse passes in chunk 0 (mean 0,
    # rms 1) and fails in chunk 1 (mean 40); snr = (median - mean) / rms
    y = np.zeros((3, 400), np.float32)
    y[:, 100:103] = 50.0
    y[:, 300:303] = 50.0
    mean = np.array([0.0, 0.0, 40.0, 40.0], np.float32)
    rms = np.array([1.0, 1.0, 2.0, 2.0], np.float32)
    d, t, s, n = O.process_brute_legacy_chunks(y, 0.0, 0.75, 6.0, mean, rms, 100)
    # rows 0 and 1 have DM 0 and 0.75 (not above 1); row 2 has DM 1.5
    assert n == 3 and d.tolist() == [2, 2, 2] and t.tolist() == [98, 99, 100]
    assert s.tolist() == [50.0, 50.0, 50.0]
    # with threshold 4 the second pulse passes too: (50 - 40) / 2 = 5 >= 4
    d, t, s, n = O.process_brute_legacy_chunks(y, 0.0, 0.75, 4.0, mean, rms, 100)
    assert t.tolist() == [98, 99, 100, 298, 299, 300] and s.tolist()[3:] == [5.0, 5.0, 5.0]
    # the chunk index is l / spc on the median's first sample: l = 199 sits in
    # chunk 1 (mean 0), l = 200 in chunk 2 (mean 40)
    z = np.zeros((2, 400), np.float32)
    z[1, 199:204] = 30.0
    d, t, s, n = O.process_brute_legacy_chunks(z, 0.0, 2.0, 6.0, mean, rms, 100)
    assert t.tolist() == [197, 198, 199]


def test_detrend_vs_polyfit():
    rng = np.random.default_rng(6)
    L = 1024
    t = np.arange(2 * L)
    x = (0.01 * t + 3 + rng.standard_normal(2 * L)).astype(np.float32)
    got = O.detrend_normalise(x, L)
    for s in range(2):
        seg = x[s * L:(s + 1) * L].astype(np.float64)
        xi = -L / 2 + 0.5 + np.arange(L)
        b, a = np.polyfit(xi, seg, 1)
        r = seg - (a + b * xi)
        assert np.allclose(got[s * L:(s + 1) * L], r / np.sqrt((r * r).mean()), atol=1e-5)


def test_detection_division_filter_never_rejects_a_passing_level():
    # The boxcar detector skips the FP64 division of a level unless
    # a >= thresh * (1 - 2^-40) * den; that must hold whenever the correctly
    # rounded a / den reaches thresh, and a rejected level must round below
    # thresh (numpy float64 is IEEE double with round-to-nearest-even).
    rng = np.random.default_rng(11)
    n = 400000
    thresh = rng.uniform(0.5, 20.0, n)
    den = np.exp(rng.uniform(np.log(1e-6), np.log(1e6), n))
    # quotients placed within a few ulps of the threshold, and spread widely
    near = thresh * den * (1.0 + rng.integers(-8, 9, n) * 2.0 ** -52)
    far = thresh * den * rng.uniform(0.5, 1.5, n)
    for a in (near, far):
        q = a / den
        lim = thresh * (1.0 - 2.0 ** -40)
        passes = a >= lim * den
        assert not np.any((q >= thresh) & ~passes)
        assert np.all(q[~passes] < thresh[~passes])
        assert (~passes).any() or a is near
