"""GPU parity of the ijloop API (hipcstone_ijloop.hpp, through the example interaction's C entry) with the oracle's
restatement of the reference CPU ijloop, bit for bit."""

import numpy as np
import pytest

from helpers import cpu_pipeline
from oracle_py import HILBERT
from oracle_py import make_box as orc_box

pytestmark = pytest.mark.gpu


def dev(hip, a):
    return hip.DeviceArray.from_numpy(np.ascontiguousarray(a))


def _case(oracle, T, n, bnd, seed):
    from hipcstone import inputs

    x, y, z, r = inputs.plummer(n, 0.3, seed=seed, dtype=T)
    h = inputs.plummer_h(r, n, 0.3, 40, dtype=T)
    obox = orc_box((-1.0, -1.05, -1.0), (1.0, 1.0, 1.1), T, bnd)
    p = cpu_pipeline(oracle, x, y, z, h, obox, np.uint64, 32, HILBERT)
    m = (np.random.default_rng(seed).random(n) + 0.5).astype(T)
    return p, obox, m


def _gpu(hip, p, T, bnd, m, first, last, ngmax, ext=1.0, groups=None):
    gbox = hip.make_box((-1.0, -1.05, -1.0), (1.0, 1.0, 1.1), T, bnd)
    d = {k: dev(hip, p[k]) for k in ("x", "y", "z", "h", "childOffsets", "parents", "internalToLeaf", "layout")}
    dc, ds = dev(hip, p["centers"].reshape(-1)), dev(hip, p["sizes"].reshape(-1))
    n = p["x"].size
    outs = [hip.DeviceArray(n, T), hip.DeviceArray(n, np.uint32), hip.DeviceArray(n, T), hip.DeviceArray(n, T),
            hip.DeviceArray(n, T)]
    for o in outs:
        o.fill(0)
    hip.ijloop_example(d["x"], d["y"], d["z"], d["h"], dev(hip, m), first, last, gbox, d["childOffsets"],
                       d["parents"], d["internalToLeaf"], d["layout"], dc, ds, p["prefixes"].size, ngmax, *outs,
                       ext=ext, groups=groups)
    return [o.download() for o in outs]


@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("bnd", [(0, 0, 0), (1, 1, 1), (1, 0, 2)])
@pytest.mark.parametrize("ngmax,first,last,ext", [(64, 0, 20000, 1.0), (7, 333, 17000, 1.2)])
def test_ijloop_example_gpu(hip, oracle, T, bnd, ngmax, first, last, ext):
    print(f"ijloop {np.dtype(T).name} bnd {bnd} ngmax {ngmax} [{first},{last}) ext {ext}", flush=True)
    p, obox, m = _case(oracle, T, 20000, bnd, seed=12)
    tree = {k: p[k] for k in ("childOffsets", "parents", "internalToLeaf")}
    want = oracle.ijloop_example(p["x"], p["y"], p["z"], p["h"], m, first, last, obox, tree, p["layout"],
                                 p["centers"], p["sizes"], ngmax, ext)
    got = _gpu(hip, p, T, bnd, m, first, last, ngmax, ext)
    for name, g, w in zip(("rho", "count", "dmin", "dmax", "fx"), got, want):
        assert np.array_equal(g, w), name
    assert want[1][first:last].min() >= 1 and want[1][first:last].max() >= min(ngmax + 1, 9)


def test_ijloop_example_gpu_chunked(hip, oracle):
    """ngmax = 65536: a wave's scratch slice is 16 MiB, so the 2 GiB scratch cap leaves 128 waves, and each wave
    walks and folds several of the 469 groups of 64 bodies."""
    T = np.float32
    print("ijloop chunked", flush=True)
    p, obox, m = _case(oracle, T, 30000, (0, 0, 0), seed=13)
    tree = {k: p[k] for k in ("childOffsets", "parents", "internalToLeaf")}
    want = oracle.ijloop_example(p["x"], p["y"], p["z"], p["h"], m, 0, 30000, obox, tree, p["layout"], p["centers"],
                                 p["sizes"], 65536)
    got = _gpu(hip, p, T, (0, 0, 0), m, 0, 30000, 65536)
    for name, g, w in zip(("rho", "count", "dmin", "dmax", "fx"), got, want):
        assert np.array_equal(g, w), name


def test_ijloop_example_gpu_subgroups(hip, oracle):
    """subgroup(): only the bodies of the given ranges are computed, with the same values as the full loop."""
    T = np.float64
    print("ijloop subgroups", flush=True)
    p, obox, m = _case(oracle, T, 20000, (1, 1, 1), seed=14)
    tree = {k: p[k] for k in ("childOffsets", "parents", "internalToLeaf")}
    first, last = 200, 19000
    want = oracle.ijloop_example(p["x"], p["y"], p["z"], p["h"], m, first, last, obox, tree, p["layout"],
                                 p["centers"], p["sizes"], 32)
    starts = np.array([210, 500, 4001, 12000], dtype=np.uint32)
    ends = np.array([300, 3000, 9000, 18999], dtype=np.uint32)
    got = _gpu(hip, p, T, (1, 1, 1), m, first, last, 32, groups=(4, dev(hip, starts), dev(hip, ends)))
    inside = np.zeros(20000, bool)
    for b, e in zip(starts, ends):
        inside[b:e] = True
    for name, g, w in zip(("rho", "count", "dmin", "dmax", "fx"), got, want):
        assert np.array_equal(g[inside], w[inside]), name
        assert not g[~inside].any(), name
