"""The triage operating rules on hand-built event lists (CPU): recall counts
pulses (chunk, pulse) found by at least one kept event, false positives
count kept events of no pulse, and the rules pick the best point within the
false-positive budget."""
import numpy as np

from mdsm_amd import abi
from mdsm_amd.triage_train import (THRESHOLDS, hybrid_rule, operating_threshold, recall_fp,
                                   recall_fp_rule, recall_fp_wide, wide_rule)


def _events(rows):
    ev = np.zeros(len(rows), abi.EVENT_DTYPE)
    for i, (chunk, pulse, label, snr, k) in enumerate(rows):
        ev[i]["chunk"], ev[i]["pulse"], ev[i]["label"], ev[i]["snr"], ev[i]["width_log2"] = chunk, pulse, label, snr, k
    return ev


def test_recall_and_false_positives():
    # pulses (0,0), (0,1), (1,0); pulse (0,0) has two events; three noise events
    ev = _events([(0, 0, 1, 9.0, 0), (0, 0, 1, 7.0, 1), (0, 1, 1, 6.5, 6), (1, 0, 1, 20.0, 2),
                  (0, -1, 0, 6.2, 6), (1, -1, 0, 8.0, 0), (1, -1, 0, 30.0, 6)])
    stats = {"injected": 4, "false_clusters": 3}
    prob = np.array([0.9, 0.1, 0.01, 0.95, 0.02, 0.3, 0.001], np.float32)
    assert recall_fp(prob, ev, stats, 0.5) == (2 / 4, 0)
    assert recall_fp(prob, ev, stats, 0.05) == (2 / 4, 1)        # the 0.3 noise event
    assert recall_fp(prob, ev, stats, 0.005) == (3 / 4, 2)
    # S/N rescue at 25: the 30-sigma noise event joins
    assert recall_fp_rule(prob, ev, stats, 0.5, 25.0) == (2 / 4, 1)
    # widest-boxcar rescue at 6.4: pulse (0,1) (6.5, k = 6) is kept, the
    # 6.2-sigma noise event at k = 6 is not, the 30-sigma one is
    assert recall_fp_wide(prob, ev, stats, 0.5, 6.4) == (3 / 4, 1)
    assert recall_fp_wide(prob, ev, stats, 0.5, 6.0) == (3 / 4, 2)


def test_rules_pick_within_budget():
    rng = np.random.default_rng(3)
    n = 4000
    label = (rng.random(n) < 0.3).astype(int)
    pulse = np.where(label == 1, np.arange(n), -1)
    k = rng.integers(0, 7, n)
    snr = 6 + rng.exponential(3.0, n)
    prob = np.clip(np.where(label == 1, rng.beta(5, 1, n), rng.beta(1, 5, n)), 0, 1).astype(np.float32)
    # faint wide pulses the network scores low
    faint = (label == 1) & (k == 6) & (snr < 7.5)
    prob[faint] = 1e-4
    ev = _events(list(zip(np.zeros(n, int), pulse, label, snr, k)))
    stats = {"injected": int(label.sum()), "false_clusters": int((label == 0).sum())}
    budget = 0.5 * stats["false_clusters"]
    t = operating_threshold(prob, ev, stats, fp_fraction=0.5)
    assert t in THRESHOLDS and recall_fp(prob, ev, stats, t)[1] <= budget
    s = hybrid_rule(prob, ev, stats, t)
    assert recall_fp_rule(prob, ev, stats, t, s)[1] <= budget
    tw, w = wide_rule(prob, ev, stats)
    rw, fw = recall_fp_wide(prob, ev, stats, tw, w)
    assert fw <= budget
    # it is the best point of the grid within the budget
    for tt in THRESHOLDS:
        for ww in (1e9, 12.0, 10.0, 9.0, 8.0, 7.5, 7.0, 6.5, 6.0):
            r, f = recall_fp_wide(prob, ev, stats, tt, ww)
            assert f > budget or r <= rw
