"""GPU parity: every C-ABI entry point against the oracle on the same seeded inputs (bit-exact)."""

import numpy as np
import pytest

from helpers import OctreeMaker, cpu_pipeline, neighbor_rows_equal, uniform_n_level_tree
from oracle_py import HILBERT, MORTON
from oracle_py import make_box as orc_box

pytestmark = pytest.mark.gpu

KDS = [np.uint32, np.uint64]
TS = [np.float32, np.float64]


def dev(hip, a):
    return hip.DeviceArray.from_numpy(np.ascontiguousarray(a))


def boxes(hip, T, lo, hi, bnd=(0, 0, 0)):
    return hip.make_box(lo, hi, T, bnd), orc_box(lo, hi, T, bnd)


def plummer_input(n, T, a=0.1, seed=21, target=60):
    from hipcstone import inputs

    x, y, z, r = inputs.plummer(n, a, seed=seed, dtype=T)
    h = inputs.plummer_h(r, n, a, target, dtype=T)
    return x, y, z, h


# ---------------------------------------------------------------- keys


@pytest.mark.parametrize("kd", KDS)
@pytest.mark.parametrize("curve", [HILBERT, MORTON])
@pytest.mark.parametrize("T", TS)
def test_keys(hip, oracle, kd, curve, T):
    x, y, z, _ = plummer_input(100003, T)
    for lo, hi in ((-1.0, 1.0), ((-1.0, -1.1, -0.95), (1.05, 1.3, 1.0))):
        gbox, obox = boxes(hip, T, lo, hi)
        dx, dy, dz = dev(hip, x), dev(hip, y), dev(hip, z)
        dk = hip.DeviceArray(x.size, kd)
        hip.compute_sfc_keys(dx, dy, dz, dk, gbox, curve)
        assert np.array_equal(dk.download(), oracle.compute_sfc_keys(x, y, z, obox, kd, curve))


@pytest.mark.parametrize("kd", KDS)
@pytest.mark.parametrize("T", TS)
def test_keys_uniform_large(hip, oracle, kd, T):
    """Every grid cell pattern of the Hilbert state machine: uniform points over the whole box at 4.5M particles
    (several passes of the 4096-block grid-stride launch), with the box corners, faces and points past the
    upper bound included, and every 5th key a removal marker that must be kept."""
    n = 4_500_007
    rng = np.random.default_rng(1234)
    pts = rng.uniform(-0.5, 1.5, size=(3, n)).astype(T)
    pts[:, :8] = np.array([[-0.5, 1.5][(c >> s) & 1] for c in range(8) for s in (2, 1, 0)], dtype=T).reshape(8, 3).T
    pts[:, 8:16] = np.nextafter(T(1.5), T(0))
    pts[0, 16:32] = T(1.7)
    x, y, z = (np.ascontiguousarray(p) for p in pts)
    gbox, obox = boxes(hip, T, -0.5, 1.5)
    rm = kd(1) << kd(3 * (10 if kd == np.uint32 else 21))
    keys_in = np.zeros(n, dtype=kd)
    keys_in[::5] = rm
    for curve in (HILBERT, MORTON):
        dk = dev(hip, keys_in)
        hip.compute_sfc_keys(dev(hip, x), dev(hip, y), dev(hip, z), dk, gbox, curve, honor_remove_key=True)
        expect = oracle.compute_sfc_keys(x, y, z, obox, kd, curve, keys_in=keys_in)
        assert np.array_equal(dk.download(), expect)


def test_keys_remove_key(hip, oracle):
    T, kd = np.float64, np.uint64
    x = np.linspace(0.01, 0.99, 1000)
    keys_in = np.zeros(1000, dtype=kd)
    keys_in[::7] = 1 << 63
    gbox, obox = boxes(hip, T, 0.0, 1.0)
    dk = dev(hip, keys_in)
    hip.compute_sfc_keys(dev(hip, x), dev(hip, x), dev(hip, x), dk, gbox, HILBERT, honor_remove_key=True)
    assert np.array_equal(dk.download(), oracle.compute_sfc_keys(x, x, x, obox, kd, HILBERT, keys_in=keys_in))


# ---------------------------------------------------------------- sort


@pytest.mark.parametrize("kd", KDS)
@pytest.mark.parametrize("n", [1, 2, 63, 4095, 4096, 4097, 100000, 1 << 20])
def test_sort_by_key_stable(hip, kd, n):
    rng = np.random.default_rng(n)
    hi = 1 << (30 if kd == np.uint32 else 63)
    keys = rng.integers(0, hi, size=n, dtype=np.uint64).astype(kd)
    keys[::3] = keys[0]  # duplicates pin the tie order
    vals = rng.permutation(n).astype(np.uint32)
    dk, dv = dev(hip, keys), dev(hip, vals)
    hip.sort_by_key(dk, dv)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(dk.download(), keys[order])
    assert np.array_equal(dv.download(), vals[order])


@pytest.mark.parametrize("end_bit", [8, 17, 30])
def test_sort_partial_bits_and_keys_only(hip, end_bit):
    rng = np.random.default_rng(end_bit)
    keys = rng.integers(0, 1 << end_bit, size=50000, dtype=np.uint64).astype(np.uint32)
    dk = dev(hip, keys)
    hip.sort_by_key(dk, None, end_bit=end_bit)
    assert np.array_equal(dk.download(), np.sort(keys))


@pytest.mark.parametrize("kd", KDS)
@pytest.mark.parametrize("T", TS)
@pytest.mark.parametrize("order", ["random", "near_sorted"])
def test_sort_gather(hip, oracle, kd, T, order):
    """Random input order gathers through packed x,y,z,h records, input close to key order gathers the arrays
    directly; both must give the oracle's order and values."""
    x, y, z, h = plummer_input(300001, T)
    gbox, obox = boxes(hip, T, -1.0, 1.0)
    if order == "near_sorted":
        _, p0 = oracle.sort_by_key(oracle.compute_sfc_keys(x, y, z, obox, kd, HILBERT))
        p0 = p0.copy()
        sw = np.random.default_rng(3).integers(0, x.size - 1, size=2000)
        p0[sw], p0[sw + 1] = p0[sw + 1], p0[sw].copy()
        x, y, z, h = (a[p0] for a in (x, y, z, h))
    keys = oracle.compute_sfc_keys(x, y, z, obox, kd, HILBERT)
    ks, perm = oracle.sort_by_key(keys)
    for hh in (h, None):
        dk = dev(hip, keys)
        dp = hip.DeviceArray(x.size, np.uint32)
        outs = [hip.DeviceArray(x.size, T) for _ in range(4)]
        outs[3].fill(0)
        hip.sort_gather(dk, dp, dev(hip, x), dev(hip, y), dev(hip, z), None if hh is None else dev(hip, hh), *outs)
        assert np.array_equal(dk.download(), ks)
        assert np.array_equal(dp.download(), perm)
        for o, a in zip(outs[:3], (x, y, z)):
            assert np.array_equal(o.download(), a[perm])
        if hh is None:
            assert not outs[3].download().any()  # h_out is not written without h
        else:
            assert np.array_equal(outs[3].download(), h[perm])


# ---------------------------------------------------------------- leaf array


@pytest.mark.parametrize("kd", KDS)
def test_node_counts(hip, oracle, kd):
    tree = OctreeMaker(kd).divide().divide(0).divide(3).divide(0, 4).make_tree()
    rng = np.random.default_rng(7)
    keys = np.sort(rng.integers(0, int(tree[-1]), size=100000, dtype=np.uint64).astype(kd))
    dc = hip.DeviceArray(tree.size - 1, np.uint32)
    hip.compute_node_counts(dev(hip, tree), dc, tree.size - 1, dev(hip, keys))
    assert np.array_equal(dc.download(), oracle.compute_node_counts(tree, keys))
    hip.compute_node_counts(dev(hip, tree), dc, tree.size - 1, dev(hip, keys), max_count=100)
    assert np.array_equal(dc.download(), oracle.compute_node_counts(tree, keys, max_count=100))


@pytest.mark.parametrize("kd", KDS)
def test_node_ops_and_rebalance_kat(hip, oracle, kd):
    # csarray.cpp:108-122 / GPU csarray.cu:85-129: scanned ops {0,1,1,...}
    tree = OctreeMaker(kd).divide().divide(0).make_tree()
    counts = np.array([1, 1, 1, 0, 0, 0, 0, 0, 2, 3, 4, 5, 6, 7, 8], dtype=np.uint32)
    ops = hip.DeviceArray(16, np.int32)
    new_n, conv = hip.compute_node_ops(dev(hip, tree), 15, dev(hip, counts), 4, ops)
    ref_ops, ref_conv = oracle.rebalance_decision(tree, counts, 4)
    scanned = np.concatenate([[0], np.cumsum(ref_ops)]).astype(np.int32)
    assert np.array_equal(ops.download(), scanned) and new_n == scanned[-1] and conv == ref_conv
    new_tree = hip.DeviceArray(new_n + 1, kd)
    hip.rebalance_tree(dev(hip, tree), 15, new_n, ops, new_tree)
    t_ref, _, _ = oracle.update_octree(np.zeros(0, dtype=kd), 4, tree, counts)
    assert np.array_equal(new_tree.download(), t_ref)


@pytest.mark.parametrize("kd", KDS)
@pytest.mark.parametrize("bucket", [1, 16, 64, 128])
def test_compute_octree(hip, oracle, kd, bucket):
    x, y, z, _ = plummer_input(200000, np.float64, a=0.05)
    _, obox = boxes(hip, np.float64, -1.0, 1.0)
    keys = np.sort(oracle.compute_sfc_keys(x, y, z, obox, kd, HILBERT))
    t_gpu, c_gpu = hip.compute_octree(dev(hip, keys), bucket)
    t_ref, c_ref = oracle.compute_octree(keys, bucket)
    assert np.array_equal(t_gpu, t_ref) and np.array_equal(c_gpu, c_ref)


@pytest.mark.parametrize("kd", KDS)
def test_update_octree_warm_steps(hip, oracle, kd):
    """Exactly one updateOctree per step from the previous leaves and counts (csarray.hpp:408-426)."""
    x, y, z, _ = plummer_input(150000, np.float64, a=0.05)
    _, obox = boxes(hip, np.float64, -1.0, 1.0)
    keys = np.sort(oracle.compute_sfc_keys(x, y, z, obox, kd, HILBERT))
    tree, counts = oracle.compute_octree(keys, 64)
    rng = np.random.default_rng(3)
    for step in range(4):
        x = np.clip(x + rng.normal(0, 0.01, x.size), -1, 1)
        keys = np.sort(oracle.compute_sfc_keys(x, y, z, obox, kd, HILBERT))
        cap = tree.size * 64
        dt, dc = dev(hip, tree), dev(hip, counts)
        nt, nc = hip.DeviceArray(cap + 1, kd), hip.DeviceArray(cap, np.uint32)
        nn, conv = hip.update_octree(dev(hip, keys), 64, dt, dc, tree.size - 1, nt, nc, cap)
        rt, rc, rconv = oracle.update_octree(keys, 64, tree, counts)
        assert np.array_equal(nt.download(nn + 1), rt) and np.array_equal(nc.download(nn), rc) and conv == rconv
        tree, counts = rt, rc


# ---------------------------------------------------------------- link, centers, layout


@pytest.mark.parametrize("kd", KDS)
def test_build_octree(hip, oracle, kd):
    cases = [uniform_n_level_tree(kd, 2), OctreeMaker(kd).divide().divide(0).make_tree(),
             OctreeMaker(kd).make_tree()]
    x, y, z, _ = plummer_input(300000, np.float64, a=0.05)
    _, obox = boxes(hip, np.float64, -1.0, 1.0)
    keys = np.sort(oracle.compute_sfc_keys(x, y, z, obox, kd, HILBERT))
    cases.append(oracle.compute_octree(keys, 16)[0])
    for tree in cases:
        g = hip.build_octree(tree)
        r = oracle.build_octree(tree)
        N = r["prefixes"].size
        for k in ("prefixes", "childOffsets", "levelRange", "internalToLeaf", "leafToInternal"):
            assert np.array_equal(g[k], r[k]), k
        if N > 1:
            assert np.array_equal(g["parents"], r["parents"])


@pytest.mark.parametrize("kd", KDS)
@pytest.mark.parametrize("T", TS)
def test_geo_centers(hip, oracle, kd, T):
    x, y, z, _ = plummer_input(200000, np.float64, a=0.05)
    _, obox64 = boxes(hip, np.float64, -1.0, 1.0)
    keys = np.sort(oracle.compute_sfc_keys(x, y, z, obox64, kd, HILBERT))
    pf = oracle.build_octree(oracle.compute_octree(keys, 8)[0])["prefixes"]
    for lo, hi in ((-1.0, 1.0), ((-1.0, -1.1, -0.95), (1.05, 1.3, 1.0))):
        gbox, obox = boxes(hip, T, lo, hi)
        c, s = hip.compute_geo_centers(pf, gbox)
        rc, rs = oracle.node_fp_centers(pf, obox, T)
        assert np.array_equal(c, rc) and np.array_equal(s, rs)


@pytest.mark.parametrize("L", [1, 100, 4096, 70001])
def test_leaf_layout(hip, L):
    counts = np.random.default_rng(L).integers(0, 200, size=L).astype(np.uint32)
    lay = hip.leaf_layout(dev(hip, counts)).download()
    assert lay[0] == 0 and np.array_equal(lay[1:], np.cumsum(counts, dtype=np.uint64).astype(np.uint32))


# ---------------------------------------------------------------- neighbors


def _gpu_neighbors(hip, p, gbox, first, last, ngmax, ext=1.0):
    T = p["x"].dtype
    d = {k: dev(hip, p[k]) for k in ("x", "y", "z", "h", "childOffsets", "parents", "internalToLeaf", "layout")}
    dc, ds = dev(hip, p["centers"].reshape(-1)), dev(hip, p["sizes"].reshape(-1))
    n = last - first
    nb = hip.DeviceArray(max(1, n * ngmax), np.uint32) if ngmax else None
    nc = hip.DeviceArray(max(1, n), np.uint32)
    hip.find_neighbors(d["x"], d["y"], d["z"], d["h"], first, last, gbox, d["childOffsets"], d["parents"],
                       d["internalToLeaf"], d["layout"], dc, ds, p["prefixes"].size, ngmax, nb, nc, ext)
    counts = nc.download(n)
    rows = nb.download(n * ngmax).reshape(n, ngmax) if ngmax else np.zeros((n, 0), np.uint32)
    return rows, counts


@pytest.mark.parametrize("T", TS)
@pytest.mark.parametrize("kd", KDS)
def test_find_neighbors(hip, oracle, T, kd):
    n = 60000
    x, y, z, h = plummer_input(n, T, a=0.2, target=80)
    gbox, obox = boxes(hip, T, -1.0, 1.0)
    p = cpu_pipeline(oracle, x, y, z, h, obox, kd, 64, HILBERT)
    tree = {k: p[k] for k in ("childOffsets", "parents", "internalToLeaf")}
    for ngmax in (150, 32):
        ref_nb, ref_nc = oracle.find_neighbors(p["x"], p["y"], p["z"], p["h"], 0, n, obox, tree, p["layout"],
                                               p["centers"], p["sizes"], ngmax)
        nb, nc = _gpu_neighbors(hip, p, gbox, 0, n, ngmax)
        assert neighbor_rows_equal(nb, nc, ref_nb, ref_nc, ngmax)


def test_find_neighbors_mixed_child_geometry(hip, oracle):
    """fp32 open-box searches read each block of 8 children in a compact form (two center values and one size per
    dimension) and fall back to the children's own centers and sizes for blocks where that form does not hold. Here
    some blocks are made irregular by one-ulp changes of a child's center or size, so both forms run in one walk;
    rows must still equal the oracle's on the same (changed) geometry, with and without a search extension."""
    T = np.float32
    n = 50000
    x, y, z, h = plummer_input(n, T, a=0.2, target=80)
    gbox, obox = boxes(hip, T, -1.0, 1.0)
    p = cpu_pipeline(oracle, x, y, z, h, obox, np.uint64, 32, HILBERT)
    N = p["prefixes"].size
    centers = p["centers"].reshape(-1, 3).copy()
    sizes = p["sizes"].reshape(-1, 3).copy()
    rng = np.random.default_rng(7)
    blocks = (N - 1) // 8
    assert blocks > 50
    picked = rng.choice(blocks, size=blocks // 5, replace=False)
    for i, b in enumerate(picked):
        node = 1 + 8 * int(b) + int(rng.integers(0, 8))
        dim = int(rng.integers(0, 3))
        if i % 2 == 0:
            centers[node, dim] = np.nextafter(centers[node, dim], T(np.inf))
        else:
            sizes[node, dim] = np.nextafter(sizes[node, dim], T(np.inf))
    q = dict(p, centers=centers.reshape(-1), sizes=sizes.reshape(-1))
    tree = {k: p[k] for k in ("childOffsets", "parents", "internalToLeaf")}
    for ext in (1.0, 1.2):
        ref_nb, ref_nc = oracle.find_neighbors(p["x"], p["y"], p["z"], p["h"], 0, n, obox, tree, p["layout"],
                                               q["centers"], q["sizes"], 96, ext)
        nb, nc = _gpu_neighbors(hip, q, gbox, 0, n, 96, ext)
        assert neighbor_rows_equal(nb, nc, ref_nb, ref_nc, 96), ext


@pytest.mark.parametrize("boundary", [(1, 1, 1), (1, 0, 0)])
def test_find_neighbors_pbc_ext_subrange(hip, oracle, boundary):
    T = np.float64
    n = 40000
    x, y, z, h = plummer_input(n, T, a=0.3, target=50)
    gbox, obox = boxes(hip, T, (-1.0, -1.05, -1.0), (1.0, 1.0, 1.1), boundary)
    p = cpu_pipeline(oracle, x, y, z, h, obox, np.uint64, 32, HILBERT)
    tree = {k: p[k] for k in ("childOffsets", "parents", "internalToLeaf")}
    first, last = 1234, 31000
    ref_nb, ref_nc = oracle.find_neighbors(p["x"], p["y"], p["z"], p["h"], first, last, obox, tree, p["layout"],
                                           p["centers"], p["sizes"], 64, 1.25)
    nb, nc = _gpu_neighbors(hip, p, gbox, first, last, 64, 1.25)
    assert neighbor_rows_equal(nb, nc, ref_nb, ref_nc, 64)
    # counts only
    _, nc0 = _gpu_neighbors(hip, p, gbox, first, last, 0, 1.25)
    assert np.array_equal(nc0, ref_nc)


@pytest.mark.parametrize("boundary", [(0, 0, 0), (1, 1, 1)])
def test_find_neighbors_slot_major(hip, oracle, boundary):
    """The slot-major layout holds the same rows: neighbor k of relative body r at ((r//64)*ngmax + k)*64 + r%64."""
    T = np.float32
    n = 30011
    x, y, z, h = plummer_input(n, T, a=0.3, target=50)
    gbox, obox = boxes(hip, T, -1.0, 1.0, boundary)
    p = cpu_pipeline(oracle, x, y, z, h, obox, np.uint64, 32, HILBERT)
    first, last, ngmax = 77, 29000, 48
    rows, counts = _gpu_neighbors(hip, p, gbox, first, last, ngmax)
    d = {k: dev(hip, p[k]) for k in ("x", "y", "z", "h", "childOffsets", "parents", "internalToLeaf", "layout")}
    m = last - first
    groups = (m + 63) // 64
    nb = hip.DeviceArray(groups * 64 * ngmax, np.uint32)
    nb.fill(0xFF)
    nc = hip.DeviceArray(m, np.uint32)
    hip.find_neighbors(d["x"], d["y"], d["z"], d["h"], first, last, gbox, d["childOffsets"], d["parents"],
                       d["internalToLeaf"], d["layout"], dev(hip, p["centers"].reshape(-1)),
                       dev(hip, p["sizes"].reshape(-1)), p["prefixes"].size, ngmax, nb, nc, slot_major=True)
    assert np.array_equal(nc.download(), counts)
    sm = nb.download().reshape(groups, ngmax, 64).transpose(0, 2, 1).reshape(groups * 64, ngmax)[:m]
    assert neighbor_rows_equal(sm, counts, rows, counts, ngmax) and counts.max() > ngmax


@pytest.mark.parametrize("T", TS)
@pytest.mark.parametrize("kd", KDS)
def test_find_neighbors_dense_leaves(hip, oracle, T, kd):
    """Leaves far above the bucket size: a 400-particle cluster narrower than a max-depth cell and 100 coincident
    particles end up in max-depth leaves (csarray.cpp:159-190), so a leaf is drained in many particle chunks, its own
    targets sit in several of them, and rows fill before the leaf ends (ngmax 150 and 32 against ~400 neighbors)."""
    rng = np.random.default_rng(5)
    x0, y0, z0, h0 = plummer_input(6000, T, a=0.2, target=60)
    cx = (0.3 + rng.random(400) * 2e-7).astype(T)
    cy = (-0.2 + rng.random(400) * 2e-7).astype(T)
    cz = (0.1 + rng.random(400) * 2e-7).astype(T)
    x = np.concatenate([x0, cx, np.full(100, -0.4, T)])
    y = np.concatenate([y0, cy, np.full(100, 0.4, T)])
    z = np.concatenate([z0, cz, np.full(100, 0.0, T)])
    h = np.concatenate([h0, np.full(400, 1e-3, T), np.full(100, 2e-3, T)])
    order = rng.permutation(x.size)
    x, y, z, h = x[order], y[order], z[order], h[order]
    n = x.size
    gbox, obox = boxes(hip, T, -1.0, 1.0)
    p = cpu_pipeline(oracle, x, y, z, h, obox, kd, 64, HILBERT)
    assert np.diff(p["layout"]).max() >= 100
    tree = {k: p[k] for k in ("childOffsets", "parents", "internalToLeaf")}
    for ngmax in (150, 32):
        ref_nb, ref_nc = oracle.find_neighbors(p["x"], p["y"], p["z"], p["h"], 0, n, obox, tree, p["layout"],
                                               p["centers"], p["sizes"], ngmax)
        nb, nc = _gpu_neighbors(hip, p, gbox, 0, n, ngmax)
        assert neighbor_rows_equal(nb, nc, ref_nb, ref_nc, ngmax)
        assert ref_nc.max() > 150


@pytest.mark.parametrize("n", [1, 50, 64])
def test_find_neighbors_root_leaf(hip, oracle, n):
    """n <= bucket: the tree is the root alone, a leaf (traversal.hpp:35-45 visits it directly)."""
    T = np.float32
    x, y, z, h = plummer_input(n, T, a=0.3, target=20)
    h = np.maximum(h, T(0.05))
    gbox, obox = boxes(hip, T, -1.0, 1.0)
    p = cpu_pipeline(oracle, x, y, z, h, obox, np.uint64, 64, HILBERT, ngmax=16)
    assert p["prefixes"].size == 1
    nb, nc = _gpu_neighbors(hip, p, gbox, 0, n, 16)
    assert neighbor_rows_equal(nb, nc, p["nb"], p["nc"], 16)


def test_find_neighbors_coincident_and_tiny(hip, oracle):
    """Coincident particles are neighbors (only j == i is excluded, findneighbors.hpp:141); n < one wave."""
    T = np.float32
    x = np.array([0.5, 0.5, 0.5, 0.25, 0.75, 0.5], dtype=T)
    y = np.array([0.5, 0.5, 0.5, 0.25, 0.75, 0.51], dtype=T)
    z = np.array([0.5, 0.5, 0.5, 0.25, 0.75, 0.5], dtype=T)
    h = np.full(6, 0.05, dtype=T)
    gbox, obox = boxes(hip, T, 0.0, 1.0)
    p = cpu_pipeline(oracle, x, y, z, h, obox, np.uint64, 2, HILBERT, ngmax=8)
    nb, nc = _gpu_neighbors(hip, p, gbox, 0, 6, 8)
    assert neighbor_rows_equal(nb, nc, p["nb"], p["nc"], 8)
    assert p["nc"].max() == 3


# ---------------------------------------------------------------- search boxes and halos


@pytest.mark.parametrize("kd", KDS)
@pytest.mark.parametrize("T", TS)
def test_bbox_and_halos(hip, oracle, kd, T):
    from hipcstone import inputs

    n = 120000
    x, y, z = inputs.gaussian_mixture(n, clusters=16, seed=31, dtype=T)
    h = np.full(n, 0.004, dtype=T)
    gbox, obox = boxes(hip, T, 0.0, 1.0)
    p = cpu_pipeline(oracle, x, y, z, h, obox, kd, 64, MORTON)
    L = p["counts"].size
    NI = p["prefixes"].size - L
    init = p["centers"][p["leafToInternal"][NI:]]
    first, last = L // 4, L // 2
    rc, rs = oracle.compute_bounding_box(p["x"], p["y"], p["z"], p["h"], p["layout"], first, last, T(2), init)
    d = {k: dev(hip, p[k]) for k in ("x", "y", "z", "h", "layout", "prefixes", "childOffsets", "parents", "leaves")}
    dsc = dev(hip, init.reshape(-1))
    dss = hip.DeviceArray(3 * L, T)
    dss.fill(0)
    hip.compute_bounding_box(d["x"], d["y"], d["z"], d["h"], d["layout"], first, last, T(2), dsc, dss)
    gc = dsc.download().reshape(L, 3)
    gs = dss.download().reshape(L, 3)
    assert np.array_equal(gc[first:last], rc[first:last]) and np.array_equal(gs[first:last], rs[first:last])
    ref = oracle.find_halos(p["prefixes"], p["childOffsets"], p["parents"], p["centers"], p["sizes"], p["leaves"], rc,
                            rs, obox, first, last)
    assert ref.sum() > 0
    num_nodes = p["prefixes"].size
    wsb = hip.find_halos_workspace_size(num_nodes, last - first, T)
    dc, dsz = dev(hip, p["centers"].reshape(-1)), dev(hip, p["sizes"].reshape(-1))
    # None: allocated in the call; otherwise a caller workspace, reused (the walk resets its own counters)
    ws = hip.DeviceArray(wsb, np.uint8)
    ws.fill(0xFF)
    for w in (None, ws, ws):
        flags = hip.DeviceArray(num_nodes, np.uint8)
        flags.fill(0)
        hip.find_halos(d["prefixes"], d["childOffsets"], d["parents"], dc, dsz, d["leaves"], dsc, dss, gbox, first,
                       last, num_nodes, flags, workspace=w)
        assert np.array_equal(flags.download(), ref)
    with pytest.raises(hip.HipcstoneError):
        hip.find_halos(d["prefixes"], d["childOffsets"], d["parents"], dc, dsz, d["leaves"], dsc, dss, gbox, first,
                       last, num_nodes, flags, workspace=ws, workspace_bytes=wsb - 1)


# ---------------------------------------------------------------- findHalos edge cases


def _halo_case(hip, oracle, tree, T, boxlim, first, last, sizes_scale, seed):
    """findHalos on a given leaf array with random search boxes, GPU against the oracle."""
    kd = tree.dtype
    o = oracle.build_octree(tree)
    obox = orc_box(boxlim[0], boxlim[1], T)
    gbox = hip.make_box(boxlim[0], boxlim[1], T)
    c, s = oracle.node_fp_centers(o["prefixes"], obox, T)
    L = tree.size - 1
    N = o["prefixes"].size
    rng = np.random.default_rng(seed)
    NI = N - L
    sc = c[o["leafToInternal"][NI:]].copy()
    ss = (s[o["leafToInternal"][NI:]] * (1.0 + sizes_scale * rng.random((L, 3)))).astype(T)
    want = oracle.find_halos(o["prefixes"], o["childOffsets"], o["parents"], c, s, tree, sc, ss, obox, first, last)
    flags = hip.DeviceArray(N, np.uint8)
    flags.fill(0)
    hip.find_halos(dev(hip, o["prefixes"]), dev(hip, o["childOffsets"]), dev(hip, o["parents"]),
                   dev(hip, c.reshape(-1)), dev(hip, s.reshape(-1)), dev(hip, tree), dev(hip, sc.reshape(-1)),
                   dev(hip, ss.reshape(-1)), gbox, first, last, N, flags)
    got = flags.download()
    assert np.array_equal(got, want)
    return got


@pytest.mark.parametrize("kd", KDS)
def test_find_halos_edge_cases(hip, oracle, kd):
    T = np.float64
    # root-only tree: the single leaf is local, nothing is outside the local range
    root = np.array([0, int(1 << (63 if kd == np.uint64 else 30))], dtype=kd)
    assert not _halo_case(hip, oracle, root, T, (0.0, 1.0), 0, 1, 0.5, 1).any()
    # one level: every leaf local (one rank) flags nothing; an empty local range flags nothing
    t1 = uniform_n_level_tree(kd, 1)
    assert not _halo_case(hip, oracle, t1, T, (0.0, 1.0), 0, 8, 2.0, 2).any()
    assert not _halo_case(hip, oracle, t1, T, (0.0, 1.0), 3, 3, 2.0, 3).any()
    # one local leaf of a one-level tree with boxes reaching its neighbours
    assert _halo_case(hip, oracle, t1, T, (0.0, 1.0), 3, 4, 2.0, 4).sum() > 1
    # an irregular tree, local range in the middle, boxes of up to 3x the leaf size
    tree = OctreeMaker(kd).divide().divide(0).divide(3).divide(0, 4).divide(7).divide(7, 2).make_tree()
    L = tree.size - 1
    for first, last in ((0, L), (1, L - 1), (L // 3, L // 2), (L - 1, L)):
        _halo_case(hip, oracle, tree, T, (-1.0, 2.0), first, last, 3.0, first + 7)
