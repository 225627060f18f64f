"""The benchmark configurations at their own sizes, checked against the oracle.

* C1, the reference's CPU-runnable case: 1,000,000 fp64 particles uniform in [0,1)^3, h for about 100 neighbors,
  Hilbert-64 keys, bucketSize 64, one full build + findNeighbors with ngmax 150 (the pattern of
  test/unit/neighbors/findneighbors.cpp:43-93).
* C2, the headline case: 16,777,216 fp32 Plummer particles (a = 0.05), Hilbert-64, bucketSize 64, one full build
  and warm steps after a 1e-3 h displacement; one step first crowds 3000 particles into one point, so the next warm
  sort meets a bucket above 2048 keys and takes the radix sort.

Keys, order, reordered coordinates, leaf array, counts, linked tree, centers and layout are compared in full;
neighbor rows and counts on a stride of targets.
"""

import numpy as np
import pytest

from helpers import neighbor_rows_equal
from oracle_py import HILBERT
from oracle_py import make_box as orc_box

pytestmark = pytest.mark.gpu

TREE_KEYS = ("leaves", "counts", "prefixes", "childOffsets", "internalToLeaf", "leafToInternal", "levelRange",
             "centers", "sizes", "layout")


def oracle_tree(oracle, ks, box, T, bucket, tree=None, counts=None):
    if tree is None:
        tree, counts = oracle.compute_octree(ks, bucket)
    else:
        tree, counts, _ = oracle.update_octree(ks, bucket, tree, counts)
    o = oracle.build_octree(tree)
    c, s = oracle.node_fp_centers(o["prefixes"], box, T)
    layout = np.concatenate([[0], np.cumsum(counts, dtype=np.uint64)]).astype(np.uint32)
    return dict(leaves=tree, counts=counts, centers=c, sizes=s, layout=layout, **o)


def check_step(hip, oracle, pipe, got, ref, nb, nc, n, ngmax, box, stride, tag):
    for k in ("keys", "perm", "x", "y", "z", "h") + TREE_KEYS:
        assert np.array_equal(got[k], ref[k]), (tag, k)
    if ref["prefixes"].size > 1:
        assert np.array_equal(got["parents"], ref["parents"]), tag
    ids = np.arange(0, n, stride, dtype=np.uint32)
    nb_o, nc_o = oracle.find_neighbors_subset(ref["x"], ref["y"], ref["z"], ref["h"], ids, box, ref, ref["layout"],
                                              ref["centers"], ref["sizes"], ngmax)
    g_nb = nb.download().reshape(n, ngmax)[ids]
    g_nc = nc.download()[ids]
    assert neighbor_rows_equal(g_nb, g_nc, nb_o, nc_o, ngmax), tag
    return g_nc


def test_c1_full_build_and_neighbors(hip, oracle):
    from hipcstone import inputs

    n, ngmax, bucket, T, kd = 1000000, 150, 64, np.float64, np.uint64
    x, y, z = inputs.uniform_cube(n, seed=42, dtype=T)
    h = np.full(n, inputs.uniform_h(n, 100.0), dtype=T)
    print("C1: 1M fp64 uniform, full build + neighbors", flush=True)
    pipe = hip.Pipeline(curve=hip.HILBERT, key_bits=64, real_bits=64, bucket_size=bucket, ngmax=ngmax)
    nb = hip.DeviceArray(n * ngmax, np.uint32)
    nc = hip.DeviceArray(n, np.uint32)
    pipe.sync(*(hip.DeviceArray.from_numpy(a) for a in (x, y, z, h)), n, hip.make_box(0.0, 1.0, T), full_build=True,
              neighbors=nb, counts=nc)
    got = pipe.arrays()
    box = orc_box(0.0, 1.0, T)
    keys = oracle.compute_sfc_keys(x, y, z, box, kd, HILBERT)
    ks, perm = oracle.sort_by_key(keys)
    ref = dict(keys=ks, perm=perm, x=x[perm], y=y[perm], z=z[perm], h=h[perm],
               **oracle_tree(oracle, ks, box, T, bucket))
    g_nc = check_step(hip, oracle, pipe, got, ref, nb, nc, n, ngmax, box, 16, "c1")
    # the configuration's shape: ~100 neighbors on average, a converged tree of 32768 leaves
    assert 90 < g_nc.mean() < 110 and ref["leaves"].size - 1 == 32768


def test_c2_full_size_steps(hip, oracle):
    from hipcstone import inputs

    n, ngmax, bucket, T, kd = 16777216, 150, 64, np.float32, np.uint64
    x, y, z, r = inputs.plummer(n, 0.05, seed=2024, dtype=T)
    h = inputs.plummer_h(r, n, 0.05, 100.0, dtype=T)
    del r
    gbox = hip.make_box(-1.0, 1.0, T)
    box = orc_box(-1.0, 1.0, T)
    pipe = hip.Pipeline(curve=hip.HILBERT, key_bits=64, real_bits=32, bucket_size=bucket, ngmax=ngmax)
    nb = hip.DeviceArray(n * ngmax, np.uint32)
    nc = hip.DeviceArray(n, np.uint32)
    d = [hip.DeviceArray.from_numpy(a) for a in (x, y, z, h)]
    tree = counts = None
    rng = np.random.default_rng(2025)
    for step in range(4):
        print(f"C2 full size: step {step}", flush=True)
        pipe.sync(*d, n, gbox, full_build=(step == 0), neighbors=nb, counts=nc)
        got = pipe.arrays()
        keys = oracle.compute_sfc_keys(x, y, z, box, kd, HILBERT)
        ks, perm = oracle.sort_by_key(keys)
        ref = dict(keys=ks, perm=perm, x=x[perm], y=y[perm], z=z[perm], h=h[perm],
                   **oracle_tree(oracle, ks, box, T, bucket, tree, counts))
        check_step(hip, oracle, pipe, got, ref, nb, nc, n, ngmax, box, 64, f"c2 step {step}")
        tree, counts = ref["leaves"], ref["counts"]
        # next input: the sorted particles displaced by 1e-3 h; after step 1, 3000 particles of one region are
        # moved onto a single point (more than 2048 keys in one bucket of the previous leaf array)
        x, y, z, h = ref["x"], ref["y"], ref["z"], ref["h"]
        disp = (rng.random((3, n), dtype=np.float32) * 2 - 1) * (T(1e-3) * h)
        x = np.clip(x + disp[0], T(-1), T(1)).astype(T)
        y = np.clip(y + disp[1], T(-1), T(1)).astype(T)
        z = np.clip(z + disp[2], T(-1), T(1)).astype(T)
        del disp
        if step == 1:
            sel = np.arange(n // 3, n // 3 + 3000)
            x[sel], y[sel], z[sel] = T(0.01), T(0.02), T(-0.03)
        for dv, hv in zip(d, (x, y, z, h)):
            dv.upload(hv)
