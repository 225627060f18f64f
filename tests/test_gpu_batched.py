"""Batched clouds (hipcstone_batched_*) against one plain pipeline per cloud (hipcstone_pipeline_* with 32-bit keys,
the cloud's box, bucket size and ngmax), cloud by cloud, bit for bit: keys, particle order, x/y/z/h, leaf array,
counts, layout, neighbor counts and rows, after the full build and after warm drift steps; the same for a sequence
of warm steps captured in one HIP graph and replayed twice (with eager work and synchronizations in between). The
plain pipeline itself is pinned to the oracle and the reference (test_gpu_pipeline.py, test_gpu_configs.py); one
cloud per case is also checked against the oracle here directly."""

import numpy as np
import pytest

from helpers import neighbor_rows_equal
from oracle_py import MORTON as O_MORTON, HILBERT as O_HILBERT
from oracle_py import make_box as orc_box

pytestmark = pytest.mark.gpu

BUCKETS = (16, 32, 64, 128, 256)


def _mix(seed, num, nmin, nmax, target=40.0):
    from hipcstone import inputs

    return inputs.cloud_mix(num_clouds=num, nmin=nmin, nmax=nmax, seed=seed, target=target)


class PlainClouds:
    """One plain pipeline per cloud, stepped the way a batched step moves every cloud."""

    def __init__(self, hip, sizes, boxes, buckets, curve, ngmax, x, y, z, h):
        self.hip = hip
        self.ngmax = ngmax
        self.pipes, self.boxes, self.nb, self.nc, self.inputs = [], [], [], [], []
        off = np.concatenate([[0], np.cumsum(sizes)])
        for c, n in enumerate(sizes):
            lo, hi = off[c], off[c + 1]
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

    def step(self, scale, seed, step):
        for c, p in enumerate(self.pipes):
            v = p.view()
            n = v.num_particles
            if scale != 0:
                self.hip.displace(v.x, v.y, v.z, v.h, n, scale, seed + c, step, self.boxes[c])
            p.sync(v.x, v.y, v.z, v.h, n, self.boxes[c], full_build=False, neighbors=self.nb[c], counts=self.nc[c])

    def arrays(self, c):
        a = self.pipes[c].arrays()
        n = a["keys"].size
        a["neighbor_counts"] = self.nc[c].download()
        a["neighbors"] = self.nb[c].download().reshape(n, self.ngmax)
        return a


def _compare(b, plain, tag):
    a = b.arrays()
    for c in range(len(plain.pipes)):
        g = b.cloud(a, c)
        r = plain.arrays(c)
        for k in ("keys", "perm", "x", "y", "z", "h", "leaves", "counts"):
            assert np.array_equal(g[k], r[k]), (tag, c, k)
        assert np.array_equal(g["layout"], r["layout"].astype(np.int64)), (tag, c, "layout")
        assert neighbor_rows_equal(g["neighbors"], g["neighbor_counts"], r["neighbors"].astype(np.int64),
                                   r["neighbor_counts"], plain.ngmax), (tag, c, "neighbors")
    return a


@pytest.mark.parametrize("curve_name", ["morton", "hilbert"])
def test_batched_equals_plain_per_cloud(hip, oracle, curve_name):
    curve = hip.MORTON if curve_name == "morton" else hip.HILBERT
    sizes, boxes, kinds, x, y, z, h = _mix(seed=77 if curve == hip.MORTON else 78, num=24, nmin=300, nmax=30000)
    buckets = np.array([BUCKETS[c % len(BUCKETS)] for c in range(len(sizes))], dtype=np.uint32)
    ngmax = 48
    print(f"batched vs plain, {curve_name}: {len(sizes)} clouds, {int(sizes.sum())} particles", flush=True)
    b = hip.Batched(sizes, [hip.make_box(lo, hi, np.float32) for lo, hi in boxes], buckets, curve=curve, ngmax=ngmax)
    b.set_particles(x, y, z, h)
    b.build()
    plain = PlainClouds(hip, sizes, boxes, buckets, curve, ngmax, x, y, z, h)
    plain.build()
    a = _compare(b, plain, "build")

    # one cloud against the oracle directly (keys, order, converged tree and neighbor rows)
    c = int(np.argmax(sizes))
    off = np.concatenate([[0], np.cumsum(sizes)])
    lo, hi = off[c], off[c + 1]
    obox = orc_box(boxes[c][0], boxes[c][1], np.float32)
    keys = oracle.compute_sfc_keys(x[lo:hi], y[lo:hi], z[lo:hi], obox, np.uint32,
                                   O_MORTON if curve == hip.MORTON else O_HILBERT)
    ks, perm = oracle.sort_by_key(keys)
    tree, counts = oracle.compute_octree(ks, int(buckets[c]))
    g = b.cloud(a, c)
    assert np.array_equal(g["keys"], ks) and np.array_equal(g["perm"], perm)
    assert np.array_equal(g["leaves"], tree) and np.array_equal(g["counts"], counts)
    t = oracle.build_octree(tree)
    cen, siz = oracle.node_fp_centers(t["prefixes"], obox, np.float32)
    layout = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    ids = np.arange(0, hi - lo, 3, dtype=np.uint32)
    nb_o, nc_o = oracle.find_neighbors_subset(x[lo:hi][perm], y[lo:hi][perm], z[lo:hi][perm], h[lo:hi][perm], ids,
                                              obox, t, layout, cen, siz, ngmax)
    assert neighbor_rows_equal(g["neighbors"][ids], g["neighbor_counts"][ids], nb_o.astype(np.int64), nc_o, ngmax)

    # warm drift steps: larger moves than the benchmark's so that leaves split and merge
    for step in range(3):
        print(f"batched warm step {step}", flush=True)
        b.step(0.3, 900, step)
        plain.step(0.3, 900, step)
        _compare(b, plain, f"step {step}")


def test_batched_graph_replays_equal_plain_steps(hip):
    sizes, boxes, kinds, x, y, z, h = _mix(seed=79, num=17, nmin=1000, nmax=20000)
    buckets = np.array([BUCKETS[(3 * c + 1) % len(BUCKETS)] for c in range(len(sizes))], dtype=np.uint32)
    ngmax = 40
    print(f"batched graph: {len(sizes)} clouds, {int(sizes.sum())} particles", flush=True)
    s = hip.Stream()
    b = hip.Batched(sizes, [hip.make_box(lo, hi, np.float32) for lo, hi in boxes], buckets, curve=hip.MORTON,
                    ngmax=ngmax)
    b.set_particles(x, y, z, h, stream=s)
    b.build(stream=s)
    b.capture(4, 0.2, 1234, 10, s)
    plain = PlainClouds(hip, sizes, boxes, buckets, hip.MORTON, ngmax, x, y, z, h)
    plain.build()
    for rep in range(2):
        print(f"graph replay {rep}", flush=True)
        b.replay(s)
        s.sync()
        for k in range(4):
            plain.step(0.2, 1234, 10 + k)
        hip.device_sync()
        _compare(b, plain, f"replay {rep}")
        # eager work between the replays: a view read (device-to-host copies and a synchronization)
        assert b.view(stream=s).num_leaves > 0


def test_batched_rejects_bad_arguments(hip):
    box = hip.make_box(0.0, 1.0, np.float32)
    with pytest.raises(hip.HipcstoneError):
        hip.Batched(np.array([100], np.uint64), [box], np.array([0], np.uint32))
    with pytest.raises(hip.HipcstoneError):
        hip.Batched(np.zeros(513, np.uint64) + 10, [box] * 513, np.full(513, 16, np.uint32))
    b = hip.Batched(np.array([100, 200], np.uint64), [box, box], np.array([16, 16], np.uint32))
    with pytest.raises(hip.HipcstoneError):
        b.step(0.0, 0, 0)  # before the build
    s = hip.Stream()
    b.set_particles(*(np.random.default_rng(3).random(300).astype(np.float32) for _ in range(3)),
                    np.full(300, 0.05, np.float32), stream=s)
    b.build(stream=s)
    with pytest.raises(hip.HipcstoneError):
        b.capture(3, 0.1, 1, 0, s)  # odd step count
