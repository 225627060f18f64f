#!/usr/bin/env python3
"""Benchmark: the plain single-domain pipeline keys -> sort -> tree -> neighbors on one MI355X.

Workload (16,777,216 fp32 particles): Plummer sphere (a = 0.05, truncated at r = 1) in the box [-1,1]^3, h from
the analytic Plummer density refined by 3 iterations of h <- h (100/n_i)^(1/3) on GPU neighbor counts
(100 +- 10 neighbors), 64-bit Hilbert keys, bucketSize 64, ngmax 150. One step = displace every particle by
1e-3 h (uniform in [-1,1)^3) + keys + stable sort/gather + one warm-start updateOctree + link + centers + layout +
findNeighbors. Inputs stay resident in HBM; W warm-up steps, then K timed steps.

Prints ONE JSON line (rank 0). Extra fields: `roofline` for the dominant kernel (the neighbor search), `stages`
(per-stage HIP-event times over the timed steps), `latency_ms` (p50/p99 per step) and `cpu_baseline` (the
reference's CPU path, compiled from its own headers into oracle/_ref, timed on this host on a bounded sample).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cornerstone-octree_amd"))

METRIC = "Mparticles/s keys→sort→tree→neighbors (16M fp32), batched vs plain; p99 ms"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name (lscpu), stated next to every CPU baseline."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def stage_report(stages, bytes_per_launch):
    """Per-stage HIP-event times with each stage's algorithmic bytes per launch, achieved GB/s and HBM fraction."""
    out = {}
    for name, (ms, c) in stages.items():
        if c <= 0:
            continue
        avg = ms / c
        rec = {"avg_ms": round(avg, 4), "launches": c}
        b = bytes_per_launch.get(name)
        if b is not None and avg > 0:
            gbs = b / (avg / 1e3) / 1e9
            rec.update({"algorithmic_bytes": int(b), "GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
        out[name] = rec
    return out


def pipeline_stage_bytes(n, leaves, nodes, kbar, real_bytes=4, key_bytes=8):
    """Algorithmic bytes per launch of each plain-pipeline stage: every input read once, every output written once.
    keys: x,y,z in, key out; sort+gather: key read, key + u32 index written, index + x,y,z,h read and written;
    tree update: ~32 B per leaf (leaves and counts in and out, one pass over the keys by searches); link: prefix in,
    child offsets, internal/leaf maps and parents out per node; centers: prefix in, two Vec3 out per node; layout:
    counts in, offsets out; neighbors: x,y,z,h in, count and the stored neighbor indices out."""
    R, K = real_bytes, key_bytes
    return {
        "keys": n * (3 * R + K),
        "sort_gather": n * (K + K + 4) + n * (4 + 4 * R + 4 * R),
        "tree_update": 32.0 * leaves,
        "link": nodes * (K + 4 + 4 + 4 + 0.5),
        "centers": nodes * (K + 6 * R),
        "layout": 8.0 * leaves,
        "neighbors": n * (4 * R + 4) + 4.0 * kbar * n,
    }


def prepare_h(hip, pipe, x, y, z, h, n, box, target=100.0, iters=3):
    """Analytic initial h, then `iters` updates h <- h (target/n_i)^(1/3) from GPU neighbor counts.
    Returns host arrays in SFC order."""
    nc = hip.DeviceArray(n, np.uint32)
    d = [hip.DeviceArray.from_numpy(a) for a in (x, y, z, h)]
    for it in range(iters):
        pipe.sync(*d, n, box, full_build=True, neighbors=None, counts=nc)
        a = pipe.arrays()
        cnt = nc.download().astype(np.float64)
        hs = a["h"].astype(np.float64) * np.cbrt(target / np.maximum(cnt, 1.0))
        x, y, z, h = a["x"], a["y"], a["z"], hs.astype(np.float32)
        log(f"  h iteration {it}: neighbors mean {cnt.mean():.1f} p1 {np.percentile(cnt, 1):.0f} "
            f"p99 {np.percentile(cnt, 99):.0f}")
        for dv, hv in zip(d, (x, y, z, h)):
            dv.upload(hv)
    pipe.sync(*d, n, box, full_build=True, neighbors=None, counts=nc)
    cnt = nc.download()
    a = pipe.arrays()
    return a["x"], a["y"], a["z"], a["h"], cnt


def cpu_baseline(x, y, z, h, stride, seed):
    """Reference CPU path (oracle/_ref) on the same particles: one full build, then one displaced warm step timed,
    neighbors of every `stride`-th particle; the neighbor time is scaled by `stride`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_py import REF, make_box

    if REF is None:
        return None
    box = make_box(-1.0, 1.0, np.float32)
    x, y, z, h = (np.array(a, dtype=np.float32, copy=True) for a in (x, y, z, h))
    tree, counts, t0, _ = REF.pipeline(x, y, z, h, box, 64, 150, nb_stride=stride)
    rng = np.random.default_rng(seed)
    disp = (rng.random((3, x.size)) * 2 - 1).astype(np.float32) * (np.float32(1e-3) * h)
    x = np.clip(x + disp[0], -1, 1).astype(np.float32)
    y = np.clip(y + disp[1], -1, 1).astype(np.float32)
    z = np.clip(z + disp[2], -1, 1).astype(np.float32)
    tree, counts, t, nc = REF.pipeline(x, y, z, h, box, 64, 150, nb_stride=stride, tree=tree, counts=counts)
    total = t[0] + t[1] + t[2] + t[3] + t[4] + t[5] * stride
    return {
        "value": x.size / (total / 1e3) / 1e6,
        "unit": "Mparticles/s",
        "cores": REF.omp_threads(),
        "cpu_model": cpu_model(),
        "kind": "reference",
        "sample": (f"one warm step of the same {x.size}-particle workload through the reference CPU OpenMP path "
                   f"(oracle/_ref); keys/sort/tree/link/centers on all particles ({sum(t[:5]):.0f} ms), neighbors for "
                   f"every {stride}th particle ({t[5]:.0f} ms, x{stride}); mean neighbors {nc.mean():.1f}"),
        "stage_ms": {"keys": t[0], "sort_gather": t[1], "tree_update": t[2], "link": t[3], "centers_layout": t[4],
                     "neighbors_extrapolated": t[5] * stride},
    }


def c1_bench(args):
    """Plain CPU-runnable pipeline at its own size: 1,000,000 fp64 particles uniform in [0,1)^3, h for about 100
    neighbors, Hilbert-64 keys, bucketSize 64; one step = a full build from the unsorted input (keys, stable sort +
    gather, the converged leaf array from a single root, link, centers, layout) + findNeighbors with ngmax 150."""
    import hipcstone as hip
    from hipcstone import inputs

    n, K, W, ngmax = 1000000, args.steps, args.warmup, 150
    T = np.float64
    x, y, z = inputs.uniform_cube(n, seed=42, dtype=T)
    h = np.full(n, inputs.uniform_h(n, 100.0), dtype=T)
    box = hip.make_box(0.0, 1.0, T)
    pipe = hip.Pipeline(curve=hip.HILBERT, key_bits=64, real_bits=64, bucket_size=64, ngmax=ngmax)
    d = [hip.DeviceArray.from_numpy(a) for a in (x, y, z, h)]
    nb = hip.DeviceArray(n * ngmax, np.uint32)
    nc = hip.DeviceArray(n, np.uint32)

    def step():
        pipe.sync(*d, n, box, full_build=True, neighbors=nb, counts=nc)

    for _ in range(W):
        step()
    hip.device_sync()
    events = [hip.Event() for _ in range(K + 1)]
    pipe.set_timing(True)
    hip.device_sync()
    events[0].record()
    for k in range(K):
        step()
        events[k + 1].record()
    hip.device_sync()
    per_step = np.array([events[k].elapsed_ms(events[k + 1]) for k in range(K)])
    total_ms = events[0].elapsed_ms(events[K])
    stages = pipe.stage_times()
    pipe.set_timing(False)
    counts = nc.download()
    kbar = float(np.minimum(counts, ngmax).mean())
    v = pipe.view()
    nb_ms, nb_launch = stages["neighbors"]
    nb_avg = nb_ms / max(nb_launch, 1)
    sb = pipeline_stage_bytes(n, int(v.num_leaves), int(v.num_nodes), kbar, real_bytes=8)
    achieved = sb["neighbors"] / (nb_avg / 1e3) / 1e9
    step_bytes = float(sum(sb.values()))
    ms_per_step = total_ms / K
    result = {
        "metric": "Mparticles/s keys→sort→tree→neighbors (1M fp64 uniform, full build)",
        "value": round(n * K / (total_ms / 1e3) / 1e6, 2),
        "unit": "Mparticles/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": ("plain CPU-runnable pipeline: 1,000,000 fp64 particles uniform in [0,1)^3, h = "
                         "0.5 (100 / (4/3 pi N))^(1/3), Hilbert-64 keys, bucketSize 64, full build from a single "
                         "root per step, findNeighbors ngmax 150"),
            "particles": n, "leaves": int(v.num_leaves), "nodes": int(v.num_nodes),
            "mean_neighbors": round(float(counts.mean()), 2), "max_neighbors": int(counts.max()),
        },
        "latency_ms": {"p50": round(float(np.percentile(per_step, 50)), 3),
                       "p99": round(float(np.percentile(per_step, 99)), 3)},
        "roofline": {"kernel": "findNeighborsKernel", "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "avg_launch_ms": round(nb_avg, 3),
                     "algorithmic_bytes_per_launch": sb["neighbors"]},
        "pipeline_roofline": {"algorithmic_bytes_per_step": step_bytes,
                              "achieved_GBs": round(step_bytes / (ms_per_step / 1e3) / 1e9, 1),
                              "frac": round(step_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "stages": stage_report(stages, sb),
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle_py import REF, make_box

        if REF is not None:
            stride = args.cpu_stride
            t0 = time.time()
            _, _, t, ncs = REF.pipeline(x.copy(), y.copy(), z.copy(), h.copy(), make_box(0.0, 1.0, T), 64, ngmax,
                                        nb_stride=stride)
            log(f"cpu baseline (c1): {time.time() - t0:.1f} s")
            total = t[0] + t[1] + t[2] + t[3] + t[4] + t[5] * stride
            result["cpu_baseline"] = {
                "value": round(n / (total / 1e3) / 1e6, 4), "unit": "Mparticles/s", "cores": REF.omp_threads(),
                "cpu_model": cpu_model(), "kind": "reference",
                "sample": (f"one full build of the same 1M-particle input through the reference CPU OpenMP path "
                           f"(oracle/_ref): keys/sort/tree/link/centers on all particles ({sum(t[:5]):.0f} ms), "
                           f"neighbors for every {stride}th particle ({t[5]:.0f} ms, x{stride}); "
                           f"mean neighbors {ncs.mean():.1f}"),
            }
    print(json.dumps(result), flush=True)


def load_profile(name):
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def load_traffic():
    return load_profile("neighbors_c2_traffic.json")


def valu_busy(avg_launch_ms):
    """Fraction of the device's VALU issue capacity the neighbor kernel uses: its wave64 VALU instruction count per
    launch (SQ_INSTS_VALU, committed counter summary) x 4 cycles each / (live launch time x clock x SIMDs)."""
    mix = load_profile("neighbors_c2_instruction_mix.json")
    if not mix or avg_launch_ms <= 0:
        return None
    cycles = mix["wave_instructions_per_launch"]["valu"] * mix["valu_cycles_per_wave_instruction"]
    capacity = avg_launch_ms / 1e3 * mix["engine_clock_ghz"] * 1e9 * mix["simds"]
    return round(cycles / capacity, 3)


def halo_bench(args):
    """Rank-share workload: 134,217,728 fp64 Gaussian-mixture particles (64 clusters, 20 % background, h for ~100
    neighbors from the local component density), Morton-64 keys, bucketSize 128; one step = keys + sort/gather +
    computeOctree from the root + link + centers + layout + makeSfcAssignment(8) + discoverHalos of rank 2
    (search boxes 2h) + node layout + halo gather of x,y,z,h. Inputs are generated on the device and stay resident."""
    import hipcstone as hip

    n, K, W = args.n_halo, args.steps, args.warmup
    ranks, rank = 8, 2
    T = np.float64
    d = [hip.DeviceArray(n, T) for _ in range(4)]
    hip.gaussian_mixture(*d, n, seed=7, clusters=64, background=0.2, target_neighbors=100.0)
    box = hip.make_box(0.0, 1.0, T)
    pipe = hip.Pipeline(curve=hip.MORTON, key_bits=64, real_bits=64, bucket_size=128, ngmax=0)
    pipe.sync(*d, n, box, full_build=True)
    m = pipe.halos(ranks, rank)
    cap = int(m * 1.05) + 1024
    outs = [hip.DeviceArray(cap, T) for _ in range(4)]
    log(f"input: {n} mixture particles; local share of rank {rank}: {m} particles")

    def step():
        pipe.sync(*d, n, box, full_build=True)
        return pipe.halos(ranks, rank, outs, cap)

    for _ in range(W):
        step()
    hip.device_sync()
    events = [hip.Event() for _ in range(K + 1)]
    pipe.set_timing(True)
    hip.device_sync()
    events[0].record()
    locals_ = []
    for k in range(K):
        locals_.append(step())
        events[k + 1].record()
    hip.device_sync()
    per_step = np.array([events[k].elapsed_ms(events[k + 1]) for k in range(K)])
    total_ms = events[0].elapsed_ms(events[K])
    stages = pipe.stage_times()
    pipe.set_timing(False)
    v = pipe.view()
    assert len(set(locals_)) == 1, "rank share changed between identical steps"
    ms_per_step = total_ms / K
    assigned = int(hip.download_ptr(v.layout, v.num_leaves + 1, np.uint32)[v.assigned_last]) - int(
        hip.download_ptr(v.layout, v.num_leaves + 1, np.uint32)[v.assigned_first])
    # compulsory bytes per step: keys 24 read + 8 write, sort 8 read + 12 write, gather 4 + 32 + 32, leaf array and
    # link/centers ~2 per particle, search boxes over the assigned particles 32, halo gather 64 per local particle
    step_bytes = n * (32 + 20 + 68 + 2) + 32 * assigned + 64 * m
    result = {
        "metric": "Mparticles/s rank-share chain keys->sort->tree->assignment->findHalos->halo gather (134M fp64)",
        "value": round(n * K / (total_ms / 1e3) / 1e6, 2),
        "unit": "Mparticles/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": (f"rank {rank} of {ranks} share at real size: {n:,} fp64 Gaussian-mixture particles "
                         "(64 clusters, sigma log-uniform [0.005,0.05], 20% background, [0,1)^3), Morton-64 keys, "
                         "bucketSize 128, full build each step, discoverHalos with 2h search boxes, halo gather"),
            "particles": n, "bucket_size": 128, "keys": "morton-64", "ranks": ranks, "rank": rank,
            "leaves": int(v.num_leaves), "nodes": int(v.num_nodes), "assigned_leaves": v.assigned_last - v.assigned_first,
            "assigned_particles": assigned, "local_particles": int(m),
        },
        "latency_ms": {"p50": round(float(np.percentile(per_step, 50)), 3),
                       "p99": round(float(np.percentile(per_step, 99)), 3), "max": round(float(per_step.max()), 3)},
        "pipeline_roofline": {"algorithmic_bytes_per_step": step_bytes,
                              "achieved_GBs": round(step_bytes / (ms_per_step / 1e3) / 1e9, 1),
                              "frac": round(step_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "stages": {name: {"avg_ms": round(ms / max(c, 1), 4), "launches": c} for name, (ms, c) in stages.items()
                   if c > 0},
    }
    result["cpu_baseline"] = None if args.no_cpu_baseline else halo_cpu_baseline(hip, args.halo_cpu_sample)
    print(json.dumps(result), flush=True)


def halo_cpu_baseline(hip, n):
    """The reference's CPU rank-share chain (oracle/_ref) on a bounded sample: the same mixture workload generated
    at n particles (h for ~100 neighbors at that count), the whole chain run once."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_py import MORTON, REF, make_box

    if REF is None:
        return None
    d = [hip.DeviceArray(n, np.float64) for _ in range(4)]
    hip.gaussian_mixture(*d, n, seed=7, clusters=64, background=0.2, target_neighbors=100.0)
    x, y, z, h = (a.download() for a in d)
    t0 = time.time()
    m, t = REF.halo_chain(x, y, z, h, make_box(0.0, 1.0, np.float64), 128, 8, 2, curve=MORTON)
    log(f"cpu baseline (halo chain, {n} particles): {time.time() - t0:.1f} s")
    total = float(t.sum())
    return {
        "value": round(n / (total / 1e3) / 1e6, 3),
        "unit": "Mparticles/s",
        "cores": REF.omp_threads(),
        "cpu_model": cpu_model(),
        "kind": "reference",
        "sample": (f"the same mixture workload generated at {n} particles, one pass of the whole rank-share chain "
                   f"through the reference CPU path (oracle/_ref): {total:.0f} ms, {m} local particles"),
        "stage_ms": {k: round(float(v), 1) for k, v in zip(
            ("keys", "sort_gather", "tree", "link", "centers_layout", "assignment", "halos", "layout_gather"), t)},
    }


def domain_bench(args):
    """One-rank Domain::sync + findNeighbors on the focus tree: the same 16,777,216 fp32 Plummer particles (h from
    the analytic density), Hilbert-64 keys, bucketSize 64 for the global leaf array and 16 for the focus tree, box
    fitted to the particles; one step = displacement by 1e-3 h + Domain::sync (box, keys, sort, global tree, focus
    tree update and counts, centers, layout, particles back in SFC order) + findNeighbors (ngmax 150)."""
    import hipcstone as hip
    from hipcstone import inputs

    n, K, W, ngmax = args.n, args.steps, args.warmup, args.ngmax
    T = np.float32
    a = 0.05
    x, y, z, r = inputs.plummer(n, a, seed=2024, dtype=T)
    h = inputs.plummer_h(r, n, a, 100.0, dtype=T)
    box = hip.make_box(-1.0, 1.0, T)
    dom = hip.Domain(box, key_bits=64, bucket_size=64, bucket_size_focus=16)
    d = [hip.DeviceArray.from_numpy(v) for v in (x, y, z, h)]
    nb = hip.DeviceArray(n * ngmax, np.uint32)
    nc = hip.DeviceArray(n, np.uint32)
    nb_stream_events = [hip.Event(), hip.Event()]

    def step(k, timing=None):
        hip.displace(d[0], d[1], d[2], d[3], n, 1e-3, 7, k, box)
        dom.sync(*d, n)
        v = dom.view()
        gb = hip.make_box(v.box_lo[:], v.box_hi[:], T)
        if timing is not None:
            timing[0].record()
        hip.find_neighbors(d[0], d[1], d[2], d[3], 0, n, gb, v.child_offsets, v.parents, v.internal_to_leaf,
                           v.layout, v.centers, v.sizes, v.num_nodes, ngmax, nb, nc)
        if timing is not None:
            timing[1].record()

    step(0)
    for k in range(W):
        step(1 + k)
    hip.device_sync()
    events = [hip.Event() for _ in range(K + 1)]
    dom.set_timing(True)
    nb_ms = 0.0
    events[0].record()
    for k in range(K):
        step(1 + W + k, nb_stream_events)
        events[k + 1].record()
        hip.device_sync()
        nb_ms += nb_stream_events[0].elapsed_ms(nb_stream_events[1])
    per_step = np.array([events[k].elapsed_ms(events[k + 1]) for k in range(K)])
    total_ms = float(per_step.sum())
    stages = dom.stage_times()
    dom.set_timing(False)
    v = dom.view()
    counts = nc.download()
    result = {
        "metric": "Mparticles/s one-rank Domain::sync + findNeighbors (16M fp32)",
        "value": round(n * K / (total_ms / 1e3) / 1e6, 2),
        "unit": "Mparticles/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": round(total_ms / K, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {
            "workload": ("one-rank Domain::sync + findNeighbors: 16,777,216 fp32 Plummer particles (a=0.05), fitted box, "
                         "Hilbert-64 keys, bucketSize 64 / focus 16, 1e-3 h displacement per step, ngmax 150"),
            "particles": n, "global_leaves": int(v.num_global_leaves), "focus_leaves": int(v.num_leaves),
            "mean_neighbors": round(float(counts.mean()), 2),
        },
        "latency_ms": {"p50": round(float(np.percentile(per_step, 50)), 3),
                       "p99": round(float(np.percentile(per_step, 99)), 3)},
        "stages": dict({name: {"avg_ms": round(ms / max(c, 1), 4), "launches": c} for name, (ms, c) in stages.items()
                        if c > 0}, neighbors={"avg_ms": round(nb_ms / K, 4), "launches": K}),
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)


def grav_bench(args):
    """One-rank Domain::syncGrav: the same 16,777,216 fp32 Plummer particles with equal masses 1/n, Hilbert-64 keys,
    bucketSize 64 for the global leaf array and 16 for the focus tree, theta 0.5, box fitted to the particles; one
    step = displacement by 1e-3 h + Domain::syncGrav (sync, masses reordered with the particles, focus-tree expansion
    centers: leaf mass centers, upsweep, vector-MAC radii)."""
    import hipcstone as hip
    from hipcstone import inputs

    n, K, W = args.n, args.steps, args.warmup
    T = np.float32
    a = 0.05
    x, y, z, r = inputs.plummer(n, a, seed=2024, dtype=T)
    h = inputs.plummer_h(r, n, a, 100.0, dtype=T)
    m = np.full(n, 1.0 / n, dtype=T)
    box = hip.make_box(-1.0, 1.0, T)
    dom = hip.Domain(box, key_bits=64, bucket_size=64, bucket_size_focus=16, theta=0.5)
    d = [hip.DeviceArray.from_numpy(v) for v in (x, y, z, h, m)]

    def step(k):
        hip.displace(d[0], d[1], d[2], d[3], n, 1e-3, 7, k, box)
        dom.sync_grav(*d, n)

    step(0)
    for k in range(W):
        step(1 + k)
    hip.device_sync()
    events = [hip.Event() for _ in range(K + 1)]
    dom.set_timing(True)
    events[0].record()
    for k in range(K):
        step(1 + W + k)
        events[k + 1].record()
    hip.device_sync()
    per_step = np.array([events[k].elapsed_ms(events[k + 1]) for k in range(K)])
    total_ms = float(per_step.sum())
    stages = dom.stage_times()
    dom.set_timing(False)
    v = dom.view()
    result = {
        "metric": "Mparticles/s one-rank Domain::syncGrav (16M fp32)",
        "value": round(n * K / (total_ms / 1e3) / 1e6, 2),
        "unit": "Mparticles/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": round(total_ms / K, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {
            "workload": ("one-rank Domain::syncGrav: 16,777,216 fp32 Plummer particles (a=0.05), masses 1/n, fitted "
                         "box, Hilbert-64 keys, bucketSize 64 / focus 16, theta 0.5, 1e-3 h displacement per step"),
            "particles": n, "global_leaves": int(v.num_global_leaves), "focus_leaves": int(v.num_leaves),
            "focus_nodes": int(v.num_nodes),
        },
        "latency_ms": {"p50": round(float(np.percentile(per_step, 50)), 3),
                       "p99": round(float(np.percentile(per_step, 99)), 3)},
        "stages": {name: {"avg_ms": round(ms / max(c, 1), 4), "launches": c} for name, (ms, c) in stages.items()
                   if c > 0},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        result["cpu_stage_baseline"] = grav_cpu_stage_baseline(dom, d)
    print(json.dumps(result), flush=True)


def grav_cpu_stage_baseline(dom, d):
    """The reference's CPU expansion-center functions (computeLeafMassCenter, upsweep with CombineSourceCenter,
    setMac, through oracle/_ref) timed on the GPU domain's last focus tree and particles, beside the GPU stage."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_py import REF, make_box

    if REF is None:
        return None
    g = dom.arrays()
    x, y, z, m = (a.download() for a in (d[0], d[1], d[2], d[4]))
    N, L = g["prefixes"].size, g["layout"].size - 1
    l2i = g["leafToInternal"][N - L:]
    box = make_box(g["box_lo"], g["box_hi"], np.float32)
    t0 = time.perf_counter()
    c = REF.leaf_mass_centers(x, y, z, m, l2i, g["layout"], N)
    c = REF.upsweep_centers(g["levelRange"], g["childOffsets"], c, g["levelRange"].size - 2)
    c = REF.set_mac(g["prefixes"], c, float(np.float32(1.0 / np.float64(np.float32(0.5)))), box)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    ok = bool(np.array_equal(c.view(np.uint8), dom.expansion_centers().view(np.uint8)))
    return {"stage": "expansion_centers", "cpu_ms": round(cpu_ms, 2), "cores": REF.omp_threads(),
            "cpu_model": cpu_model(), "kind": "reference", "matches_gpu": ok}


def ijloop_bench(args):
    """ijloop over the plain pipeline's tree: the same 16,777,216 fp32 Plummer particles (h from the analytic
    density), Hilbert-64 keys, bucketSize 64, a converged tree; one step = one ijLoop of the example interaction
    (hipcstone_ijloop.hpp: one fused kernel, each wave's neighbor walk followed by its bodies' interaction fold over
    their first ngmax = 150 neighbors: kernel sum, count, min / max distance, odd-symmetric x moment, postamble).
    The line also carries the neighbor search alone over the same tree, with and without its lists."""
    import hipcstone as hip
    from hipcstone import inputs

    n, K, W, ngmax = args.n, args.steps, args.warmup, args.ngmax
    T = np.float32
    a = 0.05
    x, y, z, r = inputs.plummer(n, a, seed=2024, dtype=T)
    h = inputs.plummer_h(r, n, a, 100.0, dtype=T)
    m = (np.random.default_rng(3).random(n) + 0.5).astype(T)
    box = hip.make_box(-1.0, 1.0, T)
    pipe = hip.Pipeline(curve=hip.HILBERT, key_bits=64, real_bits=32, bucket_size=64, ngmax=0)
    d = [hip.DeviceArray.from_numpy(v) for v in (x, y, z, h)]
    pipe.sync(*d, n, box, full_build=True)
    v = pipe.view()
    perm = hip.download_ptr(v.perm, n, np.uint32)
    dm = hip.DeviceArray.from_numpy(m[perm])
    outs = [hip.DeviceArray(n, T), hip.DeviceArray(n, np.uint32), hip.DeviceArray(n, T), hip.DeviceArray(n, T),
            hip.DeviceArray(n, T)]

    def step():
        hip.ijloop_example(v.x, v.y, v.z, v.h, dm, 0, n, box, v.child_offsets, v.parents, v.internal_to_leaf,
                           v.layout, v.centers, v.sizes, v.num_nodes, ngmax, *outs, total_bodies=n)

    for _ in range(W):
        step()
    hip.device_sync()
    events = [hip.Event() for _ in range(K + 1)]
    events[0].record()
    for k in range(K):
        step()
        events[k + 1].record()
    hip.device_sync()
    per_step = np.array([events[k].elapsed_ms(events[k + 1]) for k in range(K)])
    total_ms = float(per_step.sum())
    cnt = outs[1].download()

    # the neighbor search alone over the same tree and bodies: with its lists (ngmax 150, row-major, every list
    # written to device memory) and count-only (ngmax 0)
    def search_ms(ng):
        nb = hip.DeviceArray(max(n * ng, 1), np.uint32)
        nc = hip.DeviceArray(n, np.uint32)

        def run():
            hip.find_neighbors(v.x, v.y, v.z, v.h, 0, n, box, v.child_offsets, v.parents, v.internal_to_leaf,
                               v.layout, v.centers, v.sizes, v.num_nodes, ng, nb, nc)

        run()
        hip.device_sync()
        e0, e1 = hip.Event(), hip.Event()
        e0.record()
        for _ in range(K):
            run()
        e1.record()
        hip.device_sync()
        return e0.elapsed_ms(e1) / K

    search = {"with_lists_ms": round(search_ms(ngmax), 3), "count_only_ms": round(search_ms(0), 3)}
    search["ijloop_over_count_only"] = round(total_ms / K / search["count_only_ms"], 3)
    result = {
        "metric": "Mparticles/s ijLoop (neighbor search + interaction fold, 16M fp32)",
        "value": round(n * K / (total_ms / 1e3) / 1e6, 2),
        "unit": "Mparticles/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": round(total_ms / K, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {
            "workload": ("ijLoop of the example interaction over 16,777,216 fp32 Plummer particles (a=0.05), "
                         "Hilbert-64 keys, bucketSize 64, converged tree, ngmax 150"),
            "particles": n, "leaves": int(v.num_leaves), "mean_terms": round(float(cnt.mean()), 2),
        },
        "latency_ms": {"p50": round(float(np.percentile(per_step, 50)), 3),
                       "p99": round(float(np.percentile(per_step, 99)), 3)},
        "search_alone": search,
        "cpu_baseline": None if args.no_cpu_baseline else ijloop_cpu_baseline(v, n, m[perm], ngmax),
    }
    print(json.dumps(result), flush=True)


def ijloop_cpu_baseline(v, n, m, ngmax, stride=32):
    """The reference's CPU ijloop (oracle/_ref, CpuAlwaysTraverseNeighborhood) with the same interaction over the
    first n/stride bodies of the same sorted particles and tree, scaled to n."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_py import REF, make_box
    import hipcstone as hip

    if REF is None:
        return None
    T = np.float32
    L, N = int(v.num_leaves), int(v.num_nodes)
    arrs = {k: hip.download_ptr(getattr(v, k), n, T) for k in ("x", "y", "z", "h")}
    tree = {"childOffsets": hip.download_ptr(v.child_offsets, N, np.int32),
            "parents": hip.download_ptr(v.parents, max((N - 1) // 8, 1), np.int32),
            "internalToLeaf": hip.download_ptr(v.internal_to_leaf, N, np.int32)}
    layout = hip.download_ptr(v.layout, L + 1, np.uint32)
    centers = hip.download_ptr(v.centers, 3 * N, T)
    sizes = hip.download_ptr(v.sizes, 3 * N, T)
    last = n // stride
    t0 = time.time()
    REF.ijloop_example(arrs["x"], arrs["y"], arrs["z"], arrs["h"], m, 0, last, make_box(-1.0, 1.0, T), tree, layout,
                       centers, sizes, ngmax)
    ms = (time.time() - t0) * 1e3
    return {"value": round(last / (ms / 1e3) / 1e6, 3), "unit": "Mparticles/s", "cores": REF.omp_threads(),
            "cpu_model": cpu_model(), "kind": "reference",
            "sample": (f"the reference CPU ijloop (CpuAlwaysTraverseNeighborhood, oracle/_ref) with the same "
                       f"interaction over bodies [0, {last}) of the same sorted particles and tree: {ms:.0f} ms")}


class PlainClouds:
    """One plain pipeline per cloud (32-bit keys, the cloud's box and bucket size): the per-cloud baseline of the
    batched workload, stepped the way a batched step moves every cloud (the same drift, cloud by cloud)."""

    def __init__(self, hip, sizes, boxes, buckets, curve, ngmax, x, y, z, h):
        self.hip = hip
        self.pipes, self.boxes, self.nb, self.nc = [], [], [], []
        off = np.concatenate([[0], np.cumsum(sizes)])
        self.inputs = []
        for c, n in enumerate(sizes):
            lo, hi = int(off[c]), int(off[c + 1])
            self.boxes.append(hip.make_box(boxes[c][0], boxes[c][1], np.float32))
            self.pipes.append(hip.Pipeline(curve=curve, key_bits=32, real_bits=32, bucket_size=int(buckets[c]),
                                           ngmax=ngmax))
            self.inputs.append([hip.DeviceArray.from_numpy(np.ascontiguousarray(a[lo:hi])) for a in (x, y, z, h)])
            self.nb.append(hip.DeviceArray(int(n) * ngmax, np.uint32))
            self.nc.append(hip.DeviceArray(int(n), np.uint32))

    def build(self):
        for c, p in enumerate(self.pipes):
            d = self.inputs[c]
            p.sync(*d, d[0].count, self.boxes[c], full_build=True, neighbors=self.nb[c], counts=self.nc[c])
        self.inputs = None

    def step(self, scale, seed, step):
        for c, p in enumerate(self.pipes):
            v = p.view()
            self.hip.displace(v.x, v.y, v.z, v.h, v.num_particles, scale, seed + c, step, self.boxes[c])
            p.sync(v.x, v.y, v.z, v.h, v.num_particles, self.boxes[c], full_build=False, neighbors=self.nb[c],
                   counts=self.nc[c])


def batched_bench(args):
    """The batched workload: 512 fp32 clouds (sizes log-uniform in [1e3, 1e6], uniform or Plummer a = 0.1, h for ~64
    neighbors), Morton-32 keys, per-cloud bucket sizes from the autotuner, ngmax 128; a step is a drift of 1e-3 h plus
    keys, sort, one warm update, link, centers, layout and neighbors of every cloud. The batched form runs 20 steps as
    one captured HIP graph; the baseline runs the same steps as one plain pipeline call per cloud."""
    import hipcstone as hip
    from hipcstone import autotune, inputs

    t0 = time.time()
    sizes, boxes, kinds, x, y, z, h = inputs.cloud_mix(args.clouds)
    n = int(sizes.sum())
    tuned = autotune.load()
    buckets = autotune.cloud_buckets(sizes, kinds, tuned)
    log(f"input: {args.clouds} clouds, {n} particles ({time.time() - t0:.1f} s); buckets "
        f"{dict(zip(*np.unique(buckets, return_counts=True)))}")
    ngmax, drift, G = args.ngmax_batched, 1e-3, args.graph_steps
    s = hip.Stream()
    curve = hip.HILBERT if args.batched_curve == "hilbert" else hip.MORTON
    b = hip.Batched(sizes, [hip.make_box(lo, hi, np.float32) for lo, hi in boxes], buckets, curve=curve,
                    ngmax=ngmax)
    b.set_particles(x, y, z, h, stream=s)
    t1 = time.time()
    b.build(stream=s)
    s.sync()
    log(f"batched build: {time.time() - t1:.2f} s, {b.view(stream=s).num_leaves} leaves")
    b.capture(G, drift, 99, 0, s)
    for _ in range(max(1, args.warmup)):
        b.replay(s)
    s.sync()
    K = max(1, args.steps // G)
    ev = [hip.Event() for _ in range(K + 1)]
    ev[0].record(s)
    for k in range(K):
        b.replay(s)
        ev[k + 1].record(s)
    s.sync()
    faults = hip.device_faults()
    per_graph = np.array([ev[k].elapsed_ms(ev[k + 1]) for k in range(K)])
    total_ms = ev[0].elapsed_ms(ev[K])
    steps = K * G
    value = n * steps / (total_ms / 1e3) / 1e6
    v = b.view(stream=s)
    counts = hip.download_ptr(v.neighbor_counts, n, np.uint32)
    kbar = float(np.minimum(counts, ngmax).mean())
    step_bytes = float(sum(pipeline_stage_bytes(n, int(v.num_leaves), int(v.num_nodes), kbar).values()))
    ms_per_step = total_ms / steps
    log(f"batched: {ms_per_step:.3f} ms/step, {value:.0f} Mparticles/s, mean neighbors {counts.mean():.1f}")
    dev_bytes = b.device_bytes()
    del b

    # the plain per-cloud baseline: the same steps through one pipeline per cloud
    plain = PlainClouds(hip, sizes, boxes, buckets, curve, ngmax, x, y, z, h)
    t2 = time.time()
    plain.build()
    hip.device_sync()
    log(f"plain build of {args.clouds} clouds: {time.time() - t2:.2f} s")
    plain.step(drift, 99, 0)
    hip.device_sync()
    P = args.plain_steps
    e0, e1 = hip.Event(), hip.Event()
    e0.record()
    for k in range(P):
        plain.step(drift, 99, 1 + k)
    e1.record()
    hip.device_sync()
    plain_ms = e0.elapsed_ms(e1) / P
    plain_value = n / (plain_ms / 1e3) / 1e6
    log(f"plain per cloud: {plain_ms:.2f} ms/step, {plain_value:.0f} Mparticles/s")
    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mparticles/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": max(1, args.warmup) * G,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"batched: {args.clouds} fp32 clouds, sizes log-uniform in [1e3, 1e6] ({n} particles), uniform "
                         f"or Plummer a=0.1 50/50, h for ~64 neighbors, {args.batched_curve}-32 keys, autotuned bucket "
                         f"per cloud, ngmax {ngmax}, {G} drift steps (1e-3 h) captured as one HIP graph"),
            "clouds": args.clouds, "particles": n, "ngmax": ngmax, "keys": f"{args.batched_curve}-32",
            "buckets": {str(int(k)): int(c) for k, c in zip(*np.unique(buckets, return_counts=True))},
            "leaves": int(v.num_leaves), "mean_neighbors": round(float(counts.mean()), 2),
            "device_bytes": dev_bytes,
        },
        "batched_vs_plain": round(value / plain_value, 3),
        "plain_per_cloud": {"value": round(plain_value, 2), "unit": "Mparticles/s", "ms_per_step": round(plain_ms, 3),
                            "steps": P, "launches": "one plain pipeline sync (keys..neighbors) per cloud per step"},
        "latency_ms": {"graph_of_steps_p50": round(float(np.percentile(per_graph, 50)), 3),
                       "graph_of_steps_p99": round(float(np.percentile(per_graph, 99)), 3)},
        "pipeline_roofline": {
            "algorithmic_bytes_per_step": step_bytes,
            "achieved_GBs": round(step_bytes / (ms_per_step / 1e3) / 1e9, 1),
            "frac": round(step_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "device_faults": faults,
        "autotune": {"source": "hipcstone/tuned.json" if tuned else None,
                     "bucket": tuned.get("bucket") if tuned else None},
        "cpu_baseline": None,
    }
    del plain
    return result


def served_bench(args):
    """The served workload: Poisson arrivals at 200 requests/s (sizes log-uniform in [1e4, 4e6], fp32 Plummer a = 0.05,
    h for ~100 neighbors, Hilbert-64 keys, bucket 64, ngmax 128), 30 % full builds of a new cloud and 70 % updates of a
    resident cloud (displacement N(0, (5e-4 h)^2), one warm step), admitted by a 200 GB HBM budget into 4 worker
    slots (hipcstone_serve_run). Reports per-request latency (arrival to completion) p50/p99 and throughput; then the
    same requests released all at once through 4 slots and through 1 slot (one request at a time, the plain per-call
    form), whose throughput ratio is what continuous batching gains."""
    import hipcstone as hip
    from hipcstone import inputs

    t0 = time.time()
    reqs = inputs.served_requests(duration_s=args.serve_seconds, rate=200.0, pool=args.serve_pool)
    n_total = sum(int(r["n"]) for r in reqs)
    builds = sum(r["kind"] == 0 for r in reqs)
    log(f"served: {len(reqs)} requests ({builds} builds), {n_total} particles ({time.time() - t0:.1f} s)")

    def summary(res, st, tag):
        ok = [r for r in res if r["status"] != hip.SERVE_REJECTED]
        # from release (the arrival time when requests are released at their arrival times, else the start)
        lat = np.array([(r["end_s"] - r["arrival_s"]) * 1e3 for r in ok])
        wait = np.array([(r["admit_s"] - r["arrival_s"]) * 1e3 for r in ok])
        svc = np.array([(r["end_s"] - r["start_s"]) * 1e3 for r in ok])
        done = sum(int(q["n"]) for q, r in zip(reqs, res) if r["status"] != hip.SERVE_REJECTED)
        out = {"requests": len(res), "completed": len(ok), "particles": done,
               "makespan_s": round(st["makespan_s"], 3),
               "throughput_Mparticles_s": round(done / st["makespan_s"] / 1e6, 2),
               "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                              "max": round(float(lat.max()), 3)},
               "queue_ms": {"p50": round(float(np.percentile(wait, 50)), 3), "p99": round(float(np.percentile(wait, 99)), 3)},
               "service_ms": {"p50": round(float(np.percentile(svc, 50)), 3), "p99": round(float(np.percentile(svc, 99)), 3)},
               "evictions": st["evictions"], "rebuilt": sum(r["status"] == hip.SERVE_REBUILT for r in res),
               "rejected": sum(r["status"] == hip.SERVE_REJECTED for r in res),
               "peak_resident_GB": round(max(r["resident_bytes"] for r in res) / 1e9, 2)}
        log(f"{tag}: {out['throughput_Mparticles_s']} Mparticles/s, latency p50 {out['latency_ms']['p50']} ms, "
            f"p99 {out['latency_ms']['p99']} ms")
        return out

    res, st = hip.serve_run(reqs, hbm_budget=200 * 10**9, slots=args.serve_slots, realtime=True)
    live = summary(res, st, "served at 200 req/s")
    res, st = hip.serve_run(reqs, hbm_budget=200 * 10**9, slots=args.serve_slots, realtime=False)
    sat = summary(res, st, f"saturated, {args.serve_slots} slots")
    res, st = hip.serve_run(reqs, hbm_budget=200 * 10**9, slots=1, realtime=False)
    one = summary(res, st, "saturated, 1 slot")
    result = {
        "metric": METRIC,
        "value": live["throughput_Mparticles_s"],
        "unit": "Mparticles/s",
        "n_gpus": 1,
        "steps": len(reqs),
        "warmup": 0,
        "ms_per_step": round(live["makespan_s"] * 1e3 / max(1, len(reqs)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"served: Poisson arrivals at 200 req/s for {args.serve_seconds} s, sizes log-uniform in "
                         "[1e4, 4e6], fp32 Plummer a=0.05, Hilbert-64, bucket 64, ngmax 128, 30% builds / 70% updates "
                         f"(N(0,(5e-4 h)^2)), 200 GB HBM admission, {args.serve_slots} worker slots"),
            "requests": len(reqs), "builds": int(builds), "particles": n_total},
        "latency_ms": live["latency_ms"],
        "served": live,
        "saturated": sat,
        "saturated_one_slot": one,
        "slots_vs_one_slot": round(sat["throughput_Mparticles_s"] / one["throughput_Mparticles_s"], 3),
        "cpu_baseline": None,
    }
    return result


def autotune_run(args):
    import hipcstone as hip
    from hipcstone import autotune

    curve = hip.HILBERT if args.batched_curve == "hilbert" else hip.MORTON
    waves, wtimes = autotune.tune_neighbor_waves(hip, log=log)
    hip.set_neighbor_waves(waves)
    tuned = autotune.tune(hip, clouds_per_class=args.tune_clouds, curve=curve, log=log)
    tuned["curve"] = args.batched_curve
    tuned["neighbor_waves"] = int(waves)
    tuned["neighbor_waves_ms"] = {str(k): round(v, 4) for k, v in wtimes.items()}
    out = args.tune_out or os.path.join(ROOT, "gpurun_out", "tuned.json")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    autotune.save(tuned, out)
    print(json.dumps(tuned), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16777216)
    ap.add_argument("--ngmax", type=int, default=150)
    ap.add_argument("--cpu-stride", type=int, default=32, help="neighbor sample stride of the CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="plain workload only: skip the batched and served summaries added to its line")
    ap.add_argument("--workload", choices=["plain", "c1", "halo", "domain", "grav", "ijloop", "batched", "autotune",
                                               "served"],
                    default="plain",
                    help="plain: the 16.8M neighbor pipeline (default); c1: the 1M fp64 full-build pipeline; "
                         "halo: the 134M rank-share chain; "
                         "domain: one-rank Domain::sync + neighbors at 16.8M; grav: one-rank Domain::syncGrav at "
                         "16.8M; ijloop: the ijLoop example at 16.8M; batched: 512 clouds in one graph vs one "
                         "plain pipeline per cloud; autotune: the bucket-size sweep of the batched clouds")
    ap.add_argument("--clouds", type=int, default=512)
    ap.add_argument("--ngmax-batched", type=int, default=128)
    ap.add_argument("--graph-steps", type=int, default=20)
    ap.add_argument("--batched-curve", choices=["hilbert", "morton"], default="hilbert",
                    help="32-bit key curve of the batched clouds (node boxes are decoded as Hilbert keys, as the "
                         "reference's nodeFpCenters does, so Morton trees give truncated neighbor sets there and here)")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--tune-clouds", type=int, default=8)
    ap.add_argument("--tune-out", default=None)
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-pool", type=int, default=6)
    ap.add_argument("--n-halo", type=int, default=134217728)
    ap.add_argument("--halo-cpu-sample", type=int, default=8388608,
                    help="particles of the CPU baseline sample of the halo workload")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device 3 times (reference reading for HBM traffic counters)")
    args = ap.parse_args()

    if args.copy_calibration:
        import hipcstone as hip

        nbytes = 1 << 30
        src = hip.DeviceArray(nbytes, np.uint8)
        dst = hip.DeviceArray(nbytes, np.uint8)
        src.fill(1)
        ev = [hip.Event() for _ in range(4)]
        ev[0].record()
        for k in range(3):
            hip._check(hip._lib.hipcstone_memcpy_d2d(hip._p(dst), hip._p(src), hip.C.c_size_t(nbytes), None), "copy")
            ev[k + 1].record()
        hip.device_sync()
        ms = [ev[k].elapsed_ms(ev[k + 1]) for k in range(3)]
        print(json.dumps({"copy_bytes": nbytes, "ms": ms, "GBs": [2 * nbytes / (t / 1e3) / 1e9 for t in ms]}))
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding "
              "of the pipeline is a design in DESIGN.md, not a runnable path", file=sys.stderr)
        sys.exit(2)

    if args.workload != "autotune":
        import hipcstone as hip
        from hipcstone import autotune

        waves = autotune.apply(hip)
        log(f"neighbor search: {waves} waves per workgroup")

    if args.workload == "halo":
        halo_bench(args)
        return
    if args.workload == "c1":
        c1_bench(args)
        return
    if args.workload == "ijloop":
        return ijloop_bench(args)
    if args.workload == "domain":
        domain_bench(args)
        return
    if args.workload == "grav":
        grav_bench(args)
        return
    if args.workload == "batched":
        print(json.dumps(batched_bench(args)), flush=True)
        return
    if args.workload == "autotune":
        autotune_run(args)
        return
    if args.workload == "served":
        print(json.dumps(served_bench(args)), flush=True)
        return

    import hipcstone as hip
    from hipcstone import inputs

    n, K, W, ngmax = args.n, args.steps, args.warmup, args.ngmax
    T = np.float32
    a = 0.05
    t0 = time.time()
    x, y, z, r = inputs.plummer(n, a, seed=2024, dtype=T)
    h = inputs.plummer_h(r, n, a, 100.0, dtype=T)
    log(f"input: {n} Plummer particles ({time.time() - t0:.1f} s)")
    box = hip.make_box(-1.0, 1.0, T)

    pipe = hip.Pipeline(curve=hip.HILBERT, key_bits=64, real_bits=32, bucket_size=64, ngmax=ngmax)
    x, y, z, h, cnt0 = prepare_h(hip, pipe, x, y, z, h, n, box)
    host_particles = (x.copy(), y.copy(), z.copy(), h.copy())
    log(f"h ready: neighbors mean {cnt0.mean():.1f}, within 100+-10: {np.mean(np.abs(cnt0 - 100.0) <= 10):.3f}")

    nb = hip.DeviceArray(n * ngmax, np.uint32)
    nc = hip.DeviceArray(n, np.uint32)
    d = [hip.DeviceArray.from_numpy(v) for v in (x, y, z, h)]
    pipe.sync(*d, n, box, full_build=True, neighbors=nb, counts=nc)  # converged start tree

    def step(k):
        v = pipe.view()
        hip.displace(v.x, v.y, v.z, v.h, n, 1e-3, 7, k, box)
        pipe.sync(v.x, v.y, v.z, v.h, n, box, full_build=False, neighbors=nb, counts=nc)

    for k in range(W):
        step(k)
    hip.device_sync()
    log(f"warm-up done ({W} steps)")

    events = [hip.Event() for _ in range(K + 1)]
    pipe.set_timing(True)
    hip.device_sync()
    events[0].record()
    for k in range(K):
        step(W + k)
        events[k + 1].record()
    hip.device_sync()
    per_step = np.array([events[k].elapsed_ms(events[k + 1]) for k in range(K)])
    total_ms = events[0].elapsed_ms(events[K])
    stages = pipe.stage_times()
    pipe.set_timing(False)

    counts = nc.download()
    kbar = float(np.minimum(counts, ngmax).mean())
    v = pipe.view()
    ms_per_step = total_ms / K
    value = n * K / (total_ms / 1e3) / 1e6

    nb_ms, nb_launch = stages["neighbors"]
    nb_avg = nb_ms / max(nb_launch, 1)
    nb_bytes = n * (16 + 4) + 4.0 * float(np.minimum(counts, ngmax).sum())
    achieved = nb_bytes / (nb_avg / 1e3) / 1e9
    traffic = load_traffic()

    # whole-step algorithmic bytes of the pipeline's stages (the synthetic displacement that moves the input between
    # steps is inside the timed step but not counted): keys 12+8, sort 8 read + 8+4 write, gather 4+16+16, tree,
    # link, centers and layout per leaf / node, neighbors 16+4+4k per particle
    step_bytes = float(sum(pipeline_stage_bytes(n, int(v.num_leaves), int(v.num_nodes), kbar).values()))
    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mparticles/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("plain single-GPU pipeline: 16,777,216 fp32 Plummer particles (a=0.05, r<=1, box [-1,1]^3), "
                         "Hilbert-64 keys, bucketSize 64, warm-start updateOctree per step after a 1e-3 h "
                         "displacement, findNeighbors ngmax 150"),
            "particles": n, "bucket_size": 64, "ngmax": ngmax, "keys": "hilbert-64",
            "leaves": int(v.num_leaves), "nodes": int(v.num_nodes), "mean_neighbors": round(float(counts.mean()), 2),
            "neighbor_waves_per_workgroup": hip.get_neighbor_waves(),
        },
        "latency_ms": {"p50": round(float(np.percentile(per_step, 50)), 3),
                       "p99": round(float(np.percentile(per_step, 99)), 3),
                       "max": round(float(per_step.max()), 3)},
        "roofline": {
            "kernel": "findNeighborsKernel",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic.get("bytes_per_launch") if traffic else None,
            "avg_launch_ms": round(nb_avg, 3),
            "algorithmic_bytes_per_launch": nb_bytes,
            # the kernel is issue-bound on the VALU, not on HBM: share of the VALU issue capacity it keeps busy
            "valu_busy": valu_busy(nb_avg),
        },
        "pipeline_roofline": {
            "algorithmic_bytes_per_step": step_bytes,
            "achieved_GBs": round(step_bytes / (ms_per_step / 1e3) / 1e9, 1),
            "frac": round(step_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "stages": stage_report(stages, pipeline_stage_bytes(n, int(v.num_leaves), int(v.num_nodes), kbar)),
    }
    if not args.no_cpu_baseline:
        t1 = time.time()
        cb = cpu_baseline(*host_particles, args.cpu_stride, seed=11)
        if cb is not None:
            log(f"cpu baseline: {time.time() - t1:.1f} s")
        result["cpu_baseline"] = cb
    else:
        result["cpu_baseline"] = None
    if not args.no_extras:
        # the other two forms of the metric, run after the timed region of the plain pipeline (their own timed
        # regions): the 512-cloud batched mix against one plain pipeline per cloud, and the served request loop
        del nb, nc, d, pipe
        result["batched"] = extra_summary(batched_bench, args, ("value", "unit", "ms_per_step", "batched_vs_plain",
                                                                "plain_per_cloud", "pipeline_roofline", "steps"),
                                          lambda r: {"clouds": r["config"]["clouds"], "particles": r["config"]["particles"],
                                                     "keys": r["config"]["keys"], "buckets": r["config"]["buckets"],
                                                     "mean_neighbors": r["config"]["mean_neighbors"]})
        result["served"] = extra_summary(served_bench, args, ("value", "unit", "latency_ms", "slots_vs_one_slot"),
                                         lambda r: {"requests": r["config"]["requests"],
                                                    "builds": r["config"]["builds"],
                                                    "queue_ms": r["served"]["queue_ms"],
                                                    "peak_resident_GB": r["served"]["peak_resident_GB"],
                                                    "saturated_Mparticles_s": r["saturated"]["throughput_Mparticles_s"],
                                                    "saturated_one_slot_Mparticles_s":
                                                        r["saturated_one_slot"]["throughput_Mparticles_s"]})
    print(json.dumps(result), flush=True)


def extra_summary(fn, args, keys, more):
    """Runs one of the other workloads and keeps its headline fields; a failure is reported in the line instead of
    losing the plain pipeline's result."""
    try:
        t0 = time.time()
        r = fn(args)
        out = {k: r[k] for k in keys if k in r}
        out.update(more(r))
        out["workload"] = r["config"]["workload"]
        log(f"{fn.__name__}: {time.time() - t0:.1f} s")
        return out
    except Exception as e:  # noqa: BLE001 - reported, not hidden
        log(f"{fn.__name__} failed: {e!r}")
        return {"error": repr(e)}


if __name__ == "__main__":
    main()
