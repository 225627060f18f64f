"""GPU parity of the plain pipeline's sync steps (full build, then warm-start steps) against the oracle."""

import numpy as np
import pytest

from helpers import neighbor_rows_equal
from oracle_py import HILBERT, MORTON
from oracle_py import make_box as orc_box

pytestmark = pytest.mark.gpu


def _oracle_step(oracle, x, y, z, h, box, kd, bucket, curve, tree, counts, ngmax):
    keys = oracle.compute_sfc_keys(x, y, z, box, kd, curve)
    ks, perm = oracle.sort_by_key(keys)
    xs, ys, zs, hs = x[perm], y[perm], z[perm], h[perm]
    if tree is None:
        tree, counts = oracle.compute_octree(ks, bucket)
    else:
        tree, counts, _ = oracle.update_octree(ks, bucket, tree, counts)
    o = oracle.build_octree(tree)
    c, s = oracle.node_fp_centers(o["prefixes"], box, x.dtype)
    layout = np.concatenate([[0], np.cumsum(counts, dtype=np.uint64)]).astype(np.uint32)
    nb, nc = oracle.find_neighbors(xs, ys, zs, hs, 0, x.size, box, o, layout, c, s, ngmax)
    return dict(keys=ks, perm=perm, x=xs, y=ys, z=zs, h=hs, leaves=tree, counts=counts, centers=c, sizes=s,
                layout=layout, nb=nb, nc=nc, **o)


@pytest.mark.parametrize("kd,T,curve,bucket,ngmax", [
    (np.uint64, np.float32, HILBERT, 64, 150),
    (np.uint64, np.float64, HILBERT, 64, 150),
    (np.uint32, np.float32, MORTON, 32, 64),
])
def test_pipeline_steps_match_oracle(hip, oracle, kd, T, curve, bucket, ngmax):
    from hipcstone import inputs

    n = 120000
    x, y, z, r = inputs.plummer(n, 0.05, seed=77, dtype=T)
    h = inputs.plummer_h(r, n, 0.05, 100, dtype=T)
    gbox = hip.make_box(-1.0, 1.0, T)
    obox = orc_box(-1.0, 1.0, T)
    pipe = hip.Pipeline(curve=curve, key_bits=64 if kd == np.uint64 else 32,
                        real_bits=32 if T == np.float32 else 64, bucket_size=bucket, ngmax=ngmax)
    nb = hip.DeviceArray(n * ngmax, np.uint32)
    nc = hip.DeviceArray(n, np.uint32)
    tree = counts = None
    rng = np.random.default_rng(5)
    for step in range(4):
        d = [hip.DeviceArray.from_numpy(a) for a in (x, y, z, h)]
        pipe.sync(*d, n, gbox, full_build=(step == 0), neighbors=nb, counts=nc)
        ref = _oracle_step(oracle, x, y, z, h, obox, kd, bucket, curve, tree, counts, ngmax)
        got = pipe.arrays()
        for k in ("keys", "perm", "x", "y", "z", "h", "leaves", "counts", "prefixes", "childOffsets",
                  "internalToLeaf", "leafToInternal", "levelRange", "centers", "sizes", "layout"):
            assert np.array_equal(got[k], ref[k]), (step, k)
        if ref["prefixes"].size > 1:
            assert np.array_equal(got["parents"], ref["parents"]), step
        assert neighbor_rows_equal(nb.download().reshape(n, ngmax), nc.download(), ref["nb"], ref["nc"], ngmax), step
        tree, counts = ref["leaves"], ref["counts"]
        # next step: displace the sorted particles by 1e-3 h (clamped into the box)
        x, y, z, h = ref["x"], ref["y"], ref["z"], ref["h"]
        disp = (rng.random((3, n)) * 2 - 1).astype(T) * (T(1e-3) * h)
        x = np.clip(x + disp[0], T(-1), T(1)).astype(T)
        y = np.clip(y + disp[1], T(-1), T(1)).astype(T)
        z = np.clip(z + disp[2], T(-1), T(1)).astype(T)


def test_pipeline_reuses_its_own_sorted_arrays(hip, oracle):
    """The view's sorted arrays can be fed back as the next input (double-buffered inside)."""
    from hipcstone import inputs

    T = np.float32
    n = 50000
    x, y, z, r = inputs.plummer(n, 0.1, seed=78, dtype=T)
    h = inputs.plummer_h(r, n, 0.1, 60, dtype=T)
    gbox = hip.make_box(-1.0, 1.0, T)
    pipe = hip.Pipeline(bucket_size=64, ngmax=0)
    d = [hip.DeviceArray.from_numpy(a) for a in (x, y, z, h)]
    pipe.sync(*d, n, gbox, full_build=True)
    first = pipe.arrays()
    v = pipe.view()
    pipe.sync(v.x, v.y, v.z, v.h, n, gbox)  # already sorted: keys and order unchanged
    second = pipe.arrays()
    assert np.array_equal(second["keys"], first["keys"])
    assert np.array_equal(second["perm"], np.arange(n, dtype=np.uint32))
    assert np.array_equal(second["x"], first["x"])
    assert np.array_equal(second["leaves"], first["leaves"])


@pytest.mark.parametrize("kd,T,curve", [(np.uint64, np.float32, HILBERT), (np.uint32, np.float64, MORTON)])
def test_pipeline_warm_sort_far_from_previous_order(hip, oracle, kd, T, curve):
    """Warm steps sort through the buckets of the previous leaf array; the result must be the stable sort's for
    inputs far from the previous order: a shuffled input, coincident particles (equal keys, kept in input order),
    1500 particles crowded into one previous leaf (a bucket of many chunks), a changed particle count, and 6000
    crowded particles (a bucket above the limit: the radix sort takes over)."""
    from hipcstone import inputs

    n = 60000
    x, y, z, r = inputs.plummer(n, 0.05, seed=91, dtype=T)
    h = inputs.plummer_h(r, n, 0.05, 60, dtype=T)
    gbox, obox = hip.make_box(-1.0, 1.0, T), orc_box(-1.0, 1.0, T)
    bucket = 16
    pipe = hip.Pipeline(curve=curve, key_bits=64 if kd == np.uint64 else 32,
                        real_bits=32 if T == np.float32 else 64, bucket_size=bucket, ngmax=0)
    rng = np.random.default_rng(8)
    tree = counts = None
    for step in range(6):
        d = [hip.DeviceArray.from_numpy(np.ascontiguousarray(a)) for a in (x, y, z, h)]
        pipe.sync(*d, x.size, gbox, full_build=(step == 0))
        keys = oracle.compute_sfc_keys(x, y, z, obox, kd, curve)
        ks, perm = oracle.sort_by_key(keys)
        if tree is None:
            tree, counts = oracle.compute_octree(ks, bucket)
        else:
            tree, counts, _ = oracle.update_octree(ks, bucket, tree, counts)
        got = pipe.arrays()
        assert np.array_equal(got["keys"], ks), step
        assert np.array_equal(got["perm"], perm), step
        for k, a in (("x", x), ("y", y), ("z", z), ("h", h)):
            assert np.array_equal(got[k], a[perm]), (step, k)
        assert np.array_equal(got["leaves"], tree) and np.array_equal(got["counts"], counts), step
        x, y, z, h = (a[perm].copy() for a in (x, y, z, h))
        m = x.size
        if step == 0:  # shuffled
            p = rng.permutation(m)
            x, y, z, h = x[p], y[p], z[p], h[p]
        elif step == 1:  # every third particle onto its predecessor
            for a in (x, y, z):
                a[1::3] = a[0:m - 1:3][: a[1::3].size]
        elif step in (2, 4):  # particles crowded into one tiny cube
            c = 1500 if step == 2 else 6000
            sel = rng.choice(m, c, replace=False)
            for a in (x, y, z):
                a[sel] = (T(0.3 if step == 2 else -0.6) + rng.random(c).astype(T) * T(1e-6)).astype(T)
        elif step == 3:  # fewer particles
            x, y, z, h = x[:-777], y[:-777], z[:-777], h[:-777]
