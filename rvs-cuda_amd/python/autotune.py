"""Autotuner of the batched and served paths.

Picks, by measurement on the device it runs on:
  * the raster tile edge (16 or 32 pixels) of a launch set: both are timed on
    a sample batch, the faster is kept (results are the same bits either way);
  * the binned-entry list capacity per view (entries of (quad, tile) pairs)
    from what the binning pass reports for a sample batch, with headroom -- it
    sizes every workspace and hence how many views fit a memory budget;
  * the serving batch size (views per launch set) and the number of batches in
    flight (slots, one stream each) from short closed-loop trials, maximizing
    views/s subject to a p99 latency bound.
"""
import itertools
import time

import rvs_amd as R
import serving as S


def yuv_bytes(t):
    """Bytes of a target's encoded YUV420 planes."""
    return t.cam.w * t.cam.h * 3 // 2 * (1 if t.bitdepth <= 8 else 2)


def tune_tile(sources, targets, tiles=(16, 32), reps=3):
    """Raster tile edge for batches like `targets`: each candidate's launch
    set (encoded planes only) is run once to warm up, then timed `reps` times
    (wall clock of the synchronous call); returns (fastest tile, {tile: median
    ms}).  The targets' own tile fields are left as they are."""
    import statistics
    ms = {}
    for t in tiles:
        bs = R.BatchSynthesizer(sources, targets, want_color=False, tile=t)
        bs.run()
        times = []
        for _ in range(reps):
            a = time.perf_counter()
            bs.run()
            times.append(time.perf_counter() - a)
        ms[t] = round(1e3 * statistics.median(times), 3)
        del bs
    return min(ms, key=ms.get), ms


def tune_entry_cap(sources, targets, headroom=1.25):
    """Binned entries per view for batches like `targets`: the sample's
    entries per view times `headroom` (rounded up to a multiple of 2^16)."""
    bs = R.BatchSynthesizer(sources, targets)
    bs.run()
    R.profile_enable(True)
    try:
        bs.run()
        c = R.profile_counters()
    finally:
        R.profile_enable(False)
    per_view = int(c["bin_entries"] * headroom / len(targets)) + 1
    return ((per_view + 0xFFFF) >> 16) << 16, c


def tune_serving(scenes, requests, memory_budget, entry_cap, p99_bound_ms=50.0, batches=(4, 8, 16, 32, 64),
                 slots=(1, 2, 3), trial_batches=4, log=None, tie=0.02):
    """Closed-loop trials over (max_batch, slots); each trial serves
    trial_batches * max_batch * slots requests (cycling `requests`) at
    concurrency max_batch * slots.  Among the trials within the p99 bound,
    those within `tie` (relative) of the best rate are equal within the short
    trials' noise, and the one with the lowest p99 is chosen; with none within
    the bound, the lowest p99.  Returns (best, trials)."""
    trials = []
    for b, s in itertools.product(batches, slots):
        n = trial_batches * b * s
        reqs = [requests[k % len(requests)] for k in range(n)]
        srv = S.Server(scenes, memory_budget=memory_budget, max_batch=b, slots=s, entry_cap=entry_cap,
                       host_staging_bytes=b * max(yuv_bytes(t) for _, t in requests))
        try:
            srv.run_closed(reqs[: b * s], concurrency=b * s)   # warm-up
            t0 = time.perf_counter()
            st = srv.run_closed(reqs, concurrency=b * s)
            wall = time.perf_counter() - t0
        finally:
            srv.close()
        summ = S.summarize(st, wall_s=wall)
        summ.update(max_batch=b, slots=s)
        trials.append(summ)
        if log:
            log(summ)
    return pick_serving(trials, p99_bound_ms, tie), trials


def pick_serving(trials, p99_bound_ms=50.0, tie=0.02):
    """The chosen trial: lowest p99 among those within the bound whose rate is
    within `tie` of the best such rate; the lowest p99 overall if none is."""
    ok = [t for t in trials if t["p99_ms"] <= p99_bound_ms]
    if not ok:
        return min(trials, key=lambda t: t["p99_ms"])
    top = max(t["views_s"] for t in ok)
    return min((t for t in ok if t["views_s"] >= (1.0 - tie) * top), key=lambda t: t["p99_ms"])
