"""Autotuner: the bucket size of each batched cloud and the neighbor search's workgroup shape, chosen by measurement.

The reference fixes its tree parameters per call (bucketSize) and its GPU traversal tiling at compile time
(traversal/find_neighbors.cuh:29-58 TravConfig). Here the bucket size of every cloud of a batch is picked from
{16, 32, 64, 128, 256} by timing: clouds are grouped into classes (distribution kind x size decade); for each class a
representative batch of clouds of that class is built and stepped (drift + keys + sort + one warm update + link +
centers + layout + neighbors, captured in a HIP graph) at every candidate bucket size, and the class takes the bucket
size with the least time per particle. The choices are persisted in tuned.json next to this file, which the batched
benchmark reads; `python bench.py --workload autotune` reruns the sweep and rewrites it.

A cloud's bucket size changes its tree and therefore the work of its neighbor search; the batched result of a cloud is
bit-identical to the plain pipeline run with the same bucket size, which is what the parity tests compare.

The neighbor search's waves per workgroup (2, 4 or 8; hipcstone_set_neighbor_waves) is chosen on a plain warm step
of a Plummer sphere (the headline workload's shape at a quarter of its size), timing the neighbor stage with the
pipeline's own HIP events. It never changes a result.
"""

from __future__ import annotations

import json
import os
import time

import numpy as np

CANDIDATES = (16, 32, 64, 128, 256)
KINDS = ("uniform", "plummer")
# size bands of the classes: [1e3, 1e4), [1e4, 1e5), [1e5, 1e6]
BANDS = ((1_000, 10_000), (10_000, 100_000), (100_000, 1_000_001))
TUNED_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")
DEFAULT_BUCKET = 64


def band_of(size: int) -> int:
    for i, (lo, hi) in enumerate(BANDS):
        if size < hi:
            return i
    return len(BANDS) - 1


def load(path: str = TUNED_PATH):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cloud_buckets(sizes, kinds, tuned=None):
    """The bucket size of every cloud from the persisted choices (DEFAULT_BUCKET where nothing was tuned)."""
    tuned = load() if tuned is None else tuned
    table = (tuned or {}).get("bucket", {})
    out = np.empty(len(sizes), dtype=np.uint32)
    for c, (n, kind) in enumerate(zip(sizes, kinds)):
        row = table.get(str(kind))
        out[c] = row[band_of(int(n))] if row else DEFAULT_BUCKET
    return out


def _class_batch(kind, band, clouds, seed, target):
    from . import inputs

    lo, hi = BANDS[band]
    u = inputs.uniform(seed, 40 + band, clouds)
    sizes = np.exp(np.log(lo) + (np.log(min(hi, 1_000_000)) - np.log(lo)) * u).astype(np.int64)
    xs, ys, zs, hs, boxes = [], [], [], [], []
    for c, n in enumerate(sizes):
        n = int(n)
        if kind == "uniform":
            x, y, z = inputs.uniform_cube(n, seed=seed * 7919 + c, dtype=np.float32)
            h = np.full(n, inputs.uniform_h(n, target), dtype=np.float32)
            boxes.append((0.0, 1.0))
        else:
            x, y, z, r = inputs.plummer(n, 0.1, seed=seed * 7919 + c, dtype=np.float32)
            h = inputs.plummer_h(r, n, 0.1, target, dtype=np.float32)
            boxes.append((-1.0, 1.0))
        xs.append(x), ys.append(y), zs.append(z), hs.append(h)
    return sizes, boxes, np.concatenate(xs), np.concatenate(ys), np.concatenate(zs), np.concatenate(hs)


def time_batch(hip, sizes, boxes, buckets, x, y, z, h, curve, ngmax, drift, steps=4, reps=3):
    """Milliseconds per warm step of a batch (graph of `steps` steps, best of `reps` replays after one warm-up)."""
    s = hip.Stream()
    b = hip.Batched(sizes, [hip.make_box(lo, hi, np.float32) for lo, hi in boxes], buckets, curve=curve, ngmax=ngmax)
    b.set_particles(x, y, z, h, stream=s)
    b.build(stream=s)
    b.capture(steps, drift, 5, 0, s)
    b.replay(s)
    s.sync()
    best = float("inf")
    for _ in range(reps):
        e0, e1 = hip.Event(), hip.Event()
        e0.record(s)
        b.replay(s)
        e1.record(s)
        s.sync()
        best = min(best, e0.elapsed_ms(e1) / steps)
    del b
    return best


def tune(hip, clouds_per_class=8, seed=4242, target=64.0, ngmax=128, curve=None, drift=1e-3, log=print):
    """Runs the sweep on the current GPU; returns the tuned table (and the measured times)."""
    curve = hip.HILBERT if curve is None else curve
    table, sweep = {}, {}
    for kind in KINDS:
        row = []
        for band in range(len(BANDS)):
            sizes, boxes, x, y, z, h = _class_batch(kind, band, clouds_per_class, seed + 17 * band, target)
            n = int(sizes.sum())
            times = {}
            for bucket in CANDIDATES:
                ms = time_batch(hip, sizes, boxes, np.full(len(sizes), bucket, np.uint32), x, y, z, h, curve, ngmax,
                                drift)
                times[bucket] = ms
                log(f"autotune {kind} band {band} ({n} particles): bucket {bucket}: {ms:.3f} ms/step "
                    f"({n / ms / 1e3:.0f} Mparticles/s)")
            best = min(times, key=times.get)
            row.append(int(best))
            sweep[f"{kind}/{band}"] = {"particles": n, "ms_per_step": {str(k): round(v, 4) for k, v in times.items()},
                                       "chosen": int(best)}
        table[kind] = row
    return {"bucket": table, "bands": [list(b) for b in BANDS], "candidates": list(CANDIDATES),
            "sweep": sweep, "clouds_per_class": clouds_per_class, "target_neighbors": target, "ngmax": ngmax,
            "drift": drift, "measured": time.strftime("%Y-%m-%d")}


WAVES = (2, 4, 8)


def tune_neighbor_waves(hip, n=4_194_304, steps=4, log=print):
    """Neighbor-stage milliseconds per warm step of a plain pipeline (fp32 Plummer a = 0.05, Hilbert-64, bucket 64,
    ngmax 150, h for ~100 neighbors) for each workgroup shape; returns (best, {waves: ms})."""
    from . import inputs

    x, y, z, r = inputs.plummer(n, 0.05, seed=77, dtype=np.float32)
    h = inputs.plummer_h(r, n, 0.05, 100.0, dtype=np.float32)
    box = hip.make_box(-1.0, 1.0, np.float32)
    pipe = hip.Pipeline(curve=hip.HILBERT, key_bits=64, real_bits=32, bucket_size=64, ngmax=150)
    d = [hip.DeviceArray.from_numpy(a) for a in (x, y, z, h)]
    nb = hip.DeviceArray(n * 150, np.uint32)
    nc = hip.DeviceArray(n, np.uint32)
    pipe.sync(*d, n, box, full_build=True, neighbors=nb, counts=nc)
    before = hip.get_neighbor_waves()
    times = {}
    try:
        for w in WAVES:
            hip.set_neighbor_waves(w)
            v = pipe.view()
            pipe.sync(v.x, v.y, v.z, v.h, n, box, neighbors=nb, counts=nc)  # warm-up at this shape
            pipe.set_timing(True)
            for k in range(steps):
                v = pipe.view()
                hip.displace(v.x, v.y, v.z, v.h, n, 1e-3, 3, k, box)
                pipe.sync(v.x, v.y, v.z, v.h, n, box, neighbors=nb, counts=nc)
            ms, cnt = pipe.stage_times()["neighbors"]
            pipe.set_timing(False)
            times[w] = ms / max(cnt, 1)
            log(f"autotune neighbor waves {w}: {times[w]:.3f} ms per search ({n} particles)")
    finally:
        hip.set_neighbor_waves(before)
    return min(times, key=times.get), times


def apply(hip, tuned=None):
    """Applies the persisted workgroup shape of the neighbor search (if any); returns it."""
    tuned = load() if tuned is None else tuned
    w = int((tuned or {}).get("neighbor_waves", 0) or 0)
    if w:
        hip.set_neighbor_waves(w)
    return w or hip.get_neighbor_waves()


def save(tuned, path: str = TUNED_PATH):
    with open(path, "w") as f:
        json.dump(tuned, f, indent=1, sort_keys=True)
        f.write("\n")
