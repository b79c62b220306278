"""Train and evaluate the learned triage on synthetic search streams.

The plain search (bandpass normalisation, brute-force dedispersion,
per-block FP64 statistics, boxcar detection at 6 sigma, 2-D DBSCAN with
eps=3 / minPts=4, cluster peaks) runs over seeded chunks with injected
dispersed pulses (DM 10-800, width 1-64 samples, S/N 6-30) and synthetic
RFI (zero-DM broadband spikes, narrowband bursts, 50 Hz-modulated and
on/off channels).  Every cluster peak becomes an event with two BF16
snippets (64 DM trials x 256 samples; 64 subbands x 256 samples
dedispersed at the event DM) and a label (1 when it matches an injected
pulse).  The network trains on the training stream with the library's own
MFMA forward/backward kernels and is evaluated on held-out streams against
the plain path: recall on injected pulses and false-positive clusters.

    python -m mdsm_amd.triage_train --events 200000 --epochs 6 --json out.json
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import abi

SNIP = (64, 256)
ROW = 2 * SNIP[0] * SNIP[1]

# Training stream: the full single-beam survey (1024 channels over 300 MHz
# from 1500 MHz, 64 us, 4096 DM trials of 0.2), 2^17-sample chunks.
TRAIN_SURVEY = dict(fch1=1500.0, foff=-0.29296875, nchans=1024, tsamp=64e-6, lowdm=0.0, dmstep=0.2,
                    ndms=4096, nsamp=1 << 17)
# Held-out streams: beams shaped like the uneven multi-beam batch
# (512/1024 channels, DM steps 0.1-0.5, 1024-4096 trials).
EVAL_SURVEYS = [
    dict(fch1=1500.0, foff=-300.0 / 512, nchans=512, tsamp=64e-6, lowdm=0.0, dmstep=0.4, ndms=2048,
         nsamp=1 << 17),
    dict(fch1=1500.0, foff=-0.29296875, nchans=1024, tsamp=64e-6, lowdm=0.0, dmstep=0.2, ndms=4096,
         nsamp=1 << 17),
    dict(fch1=1500.0, foff=-0.29296875, nchans=1024, tsamp=64e-6, lowdm=0.0, dmstep=0.5, ndms=1700,
         nsamp=1 << 17),
]


def make_config(sv, npulses=40, n_spikes=15, n_bursts=10, n_mains=3, n_telegraph=3):
    s = abi.Survey.make(sv["fch1"], sv["foff"], sv["nchans"], sv["tsamp"], sv["lowdm"], sv["dmstep"],
                        sv["ndms"], sv["nsamp"])
    dm_hi = min(800.0, sv["lowdm"] + sv["dmstep"] * (sv["ndms"] - 1))
    return abi.TriageConfig.make(s, npulses=npulses, dm=(10.0, dm_hi), width=(1.0, 64.0), snr=(6.0, 30.0),
                                 n_spikes=n_spikes, n_bursts=n_bursts, n_mains=n_mains,
                                 n_telegraph=n_telegraph, rfi=(8.0, 40.0), snip=SNIP)


def build_stream(cfg, seed, max_events, max_chunks=100000, log=print):
    """Events + device snippets from chunks of one stream until max_events."""
    sim = abi.TriageSim(cfg, seed)
    d_snip = abi.DeviceBuffer(max_events * ROW * 2)
    events = np.zeros(max_events, abi.EVENT_DTYPE)
    n, t0 = 0, time.time()
    while n < max_events and sim.stats.chunks < max_chunks:
        _, n = sim.run(min(50, max_chunks - sim.stats.chunks), d_snip.ptr, max_events, events, n)
        log(f"  stream seed {seed}: {sim.stats.chunks} chunks, {n} events, {time.time() - t0:.0f} s")
    stats = sim.stats.as_dict()
    sim.close()
    return d_snip, events[:n], stats, time.time() - t0


THRESHOLDS = (0.9, 0.5, 0.3, 0.2, 0.1, 0.05, 0.02, 0.01, 0.005, 0.002, 0.001, 0.0005, 0.0002, 0.0001)


def recall_fp(prob, events, stats, t):
    """Recall on injected pulses and false-positive count of the events kept at prob > t."""
    keep = prob > t
    hits = keep & (events["label"] == 1)
    found = len(set(zip(events["chunk"][hits].tolist(), events["pulse"][hits].tolist())))
    return found / max(stats["injected"], 1), int(((events["label"] == 0) & keep).sum())


def operating_threshold(prob, events, stats, fp_fraction=0.25):
    """Lowest threshold whose false positives stay under fp_fraction of the
    plain path's on a validation stream (lower thresholds keep more pulses)."""
    budget = fp_fraction * stats["false_clusters"]
    ok = [t for t in THRESHOLDS if recall_fp(prob, events, stats, t)[1] <= budget]
    return min(ok) if ok else max(THRESHOLDS)


def build_mixed(cfgs, seeds, max_events, log=print):
    """One training stream drawn from several surveys: equal shares of
    max_events from each (cfg, seed), appended into one snippet buffer."""
    d_snip = abi.DeviceBuffer(max_events * ROW * 2)
    events = np.zeros(max_events, abi.EVENT_DTYPE)
    n, t0, stats = 0, time.time(), []
    for i, (cfg, seed) in enumerate(zip(cfgs, seeds)):
        quota = max_events * (i + 1) // len(cfgs)
        sim = abi.TriageSim(cfg, seed)
        while n < quota:
            _, n = sim.run(50, d_snip.ptr, quota, events, n)
            log(f"  stream seed {seed}: {sim.stats.chunks} chunks, {n} events, {time.time() - t0:.0f} s")
        stats.append(sim.stats.as_dict())
        sim.close()
    total = {k: sum(st[k] for st in stats) for k in stats[0]}
    return d_snip, events[:n], total, time.time() - t0


def score(prob, events, stats, thresholds):
    """Plain vs learned triage on one held-out stream."""
    out = {"plain": {"recall": stats["injected_found"] / stats["injected"], "false_positives":
                     stats["false_clusters"], "clusters": stats["clusters"]}}
    for t in thresholds:
        recall, fp = recall_fp(prob, events, stats, t)
        out[f"cnn@{t}"] = {"recall": recall, "false_positives": fp, "kept": int((prob > t).sum())}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--events", type=int, default=200000)
    ap.add_argument("--eval-chunks", type=int, default=40)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--threshold", type=float, default=None,
                    help="operating threshold (default: picked on a validation stream)")
    ap.add_argument("--single-survey", action="store_true",
                    help="train on the single-beam survey only (default: three beam shapes)")
    ap.add_argument("--save", default=None, help="write trained parameters (.npy, canonical layout)")
    ap.add_argument("--json", default=None, help="write the results as JSON")
    args = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    res = {"config": vars(args)}
    if args.single_survey:
        log(f"building the training stream ({args.events} events, one survey)")
        d_tr, ev_tr, st_tr, t_sim = build_stream(make_config(TRAIN_SURVEY), args.seed, args.events, log=log)
    else:
        # equal shares from the single-beam survey and the two other beam
        # shapes of the held-out streams (other seeds than the held-out ones)
        shapes = [TRAIN_SURVEY, EVAL_SURVEYS[0], EVAL_SURVEYS[2]]
        log(f"building the training stream ({args.events} events from {len(shapes)} beam shapes)")
        d_tr, ev_tr, st_tr, t_sim = build_mixed([make_config(sv) for sv in shapes],
                                                [args.seed + k for k in range(len(shapes))],
                                                args.events, log=log)
    res["train"] = {"events": len(ev_tr), "pulses": int(ev_tr["label"].sum()), "stats": st_tr,
                    "sim_seconds": t_sim}
    lab = abi.DeviceBuffer.from_numpy(ev_tr["label"].astype(np.int32))
    net = abi.TriageNet(SNIP[0], SNIP[1], args.batch)
    net.init(args.seed)
    res["epochs"] = []
    for ep in range(args.epochs):
        t = time.time()
        lr = args.lr * (0.5 ** (ep // 2))
        loss = abi.train_epoch(net, d_tr.ptr, lab.ptr, len(ev_tr), args.batch, lr, args.seed * 1000 + ep)
        abi.synchronize()
        dt = time.time() - t
        res["epochs"].append({"loss": loss, "seconds": dt, "events_per_s": len(ev_tr) / dt})
        log(f"epoch {ep}: loss {loss:.4f}, {len(ev_tr) / dt:.0f} events/s")
    del d_tr
    # operating point from a validation stream of the single-beam survey
    d_va, ev_va, st_va, _ = build_stream(make_config(TRAIN_SURVEY), args.seed + 50, 2000 * args.eval_chunks,
                                         max_chunks=max(args.eval_chunks // 2, 1), log=log)
    if len(ev_va) >= 2000 * args.eval_chunks:
        raise RuntimeError("validation stream filled its event buffer; raise the capacity")
    d_prob = abi.DeviceBuffer(4 * max(len(ev_va), 1))
    abi.predict(net, d_va.ptr, len(ev_va), d_prob.ptr)
    abi.synchronize()
    pv = d_prob.download(np.float32, (len(ev_va),))
    thr = args.threshold if args.threshold is not None else operating_threshold(pv, ev_va, st_va)
    # a second operating point at the target bound itself (half the
    # plain path's false positives on the validation stream)
    thr_half = operating_threshold(pv, ev_va, st_va, fp_fraction=0.5)
    res["validation"] = score(pv, ev_va, st_va, THRESHOLDS)
    res["threshold"] = thr
    res["threshold_half"] = thr_half
    log(f"operating thresholds {thr} (quarter of the plain false positives), {thr_half} (half)")
    del d_va, d_prob
    thresholds = THRESHOLDS
    res["eval"] = []
    for i, sv in enumerate(EVAL_SURVEYS):
        # room for every event of every chunk: a stream cut inside a chunk
        # would count that chunk's pulses for the plain path but not for the
        # network
        cap = 2000 * args.eval_chunks
        d_te, ev_te, st_te, _ = build_stream(make_config(sv), args.seed + 100 + i, cap,
                                             max_chunks=args.eval_chunks, log=log)
        if len(ev_te) >= cap:
            raise RuntimeError("held-out stream filled its event buffer; raise the capacity")
        n = len(ev_te)
        d_prob = abi.DeviceBuffer(4 * max(n, 1))
        t = time.time()
        abi.predict(net, d_te.ptr, n, d_prob.ptr)
        abi.synchronize()
        t_inf = time.time() - t
        prob = d_prob.download(np.float32, (n,))
        r = score(prob, ev_te, st_te, thresholds)
        r["survey"] = sv
        r["operating_point"] = r[f"cnn@{thr}"]
        r["operating_point_half"] = r[f"cnn@{thr_half}"]
        r["inference_events_per_s"] = n / max(t_inf, 1e-9)
        res["eval"].append(r)
        log(json.dumps(r))
    if args.save:
        np.save(args.save, net.params())
    if args.json:
        with open(args.json, "w") as f:
            json.dump(res, f, indent=1)
    net.close()
    return res


if __name__ == "__main__":
    main()
