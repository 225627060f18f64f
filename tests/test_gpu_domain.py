"""GPU parity of the one-rank Domain::sync (box, keys, order, global leaf array, focus tree, its octree, centers and
layout) against the oracle's one-rank domain sequence (itself pinned to the reference's own Domain) over several
steps, and the neighbor search on the focus tree."""

import numpy as np
import pytest

from helpers import domain_move, neighbor_rows_equal
from oracle_py import CpuDomain

pytestmark = pytest.mark.gpu


def dev(hip, a):
    return hip.DeviceArray.from_numpy(np.ascontiguousarray(a))


def plummer_input(n, T, seed=31):
    from hipcstone import inputs

    x, y, z, r = inputs.plummer(n, 0.1, seed=seed, dtype=T)
    h = inputs.plummer_h(r, n, 0.1, 60, dtype=T)
    return x, y, z, h


@pytest.mark.parametrize("kd,T,bucket,bucket_focus", [(np.uint64, np.float32, 64, 16), (np.uint64, np.float64, 128, 32),
                                                      (np.uint32, np.float32, 64, 64), (np.uint32, np.float64, 32, 8)])
@pytest.mark.parametrize("boundary", [(0, 0, 0), (1, 2, 3)])
def test_domain_sync_steps(hip, oracle, kd, T, bucket, bucket_focus, boundary):
    n = 120000
    x, y, z, h = plummer_input(n, T)
    lo, hi = -1.0, 1.0
    gbox = hip.make_box(lo, hi, T, boundary)
    dom = hip.Domain(gbox, key_bits=64 if kd == np.uint64 else 32, bucket_size=bucket,
                     bucket_size_focus=bucket_focus)
    ref = CpuDomain(oracle, kd, T, bucket, bucket_focus, lo, hi, boundary)
    rng = np.random.default_rng(32)
    d = [dev(hip, a) for a in (x, y, z, h)]
    keys_out = hip.DeviceArray(n, kd)
    for step in range(6):
        print(f"domain step {step}: {np.dtype(kd).name} {np.dtype(T).name} bucket {bucket}/{bucket_focus} "
              f"boundary {boundary}", flush=True)
        xs, ys, zs, hs = (a.download() for a in d)
        o = ref.sync(xs, ys, zs, hs)
        dom.sync(*d, n, keys_out)
        g = dom.arrays()
        v = dom.view()
        assert g["box_lo"] == [float(t) for t in o["lo"]] and g["box_hi"] == [float(t) for t in o["hi"]]
        assert np.array_equal(keys_out.download(), o["keys"]) and np.array_equal(g["keys"], o["keys"])
        for name, a in zip("xyzh", d):
            assert np.array_equal(a.download(), o[name]), name
        assert np.array_equal(g["global_leaves"], o["gtree"]) and np.array_equal(g["global_counts"], o["gcounts"])
        assert v.global_steps == o["gsteps"] and v.focus_steps == o["fsteps"]
        assert bool(v.focus_converged) == o["converged"]
        assert np.array_equal(g["leaves"], o["fleaves"])
        assert np.array_equal(g["counts"], o["fcounts"]) and np.array_equal(g["node_counts"], o["fnodecounts"])
        t = oracle.build_octree(o["fleaves"])
        for k in ("prefixes", "childOffsets", "internalToLeaf", "leafToInternal"):
            assert np.array_equal(g[k], t[k]), k
        c, s = oracle.node_fp_centers(t["prefixes"], ref.box(), T)
        assert np.array_equal(g["centers"], c) and np.array_equal(g["sizes"], s)
        assert np.array_equal(g["layout"], np.concatenate([[0], np.cumsum(o["fcounts"])]).astype(np.uint32))
        assert (v.start_index, v.end_index) == (0, n)
        # the next input (host-side, identical for both): displaced; shuffled after step 1 (the later sorts bucket
        # by the previous leaves), rearranged after step 2, contracted by 20 % after step 3 (the box shrink limit)
        for a, m in zip(d, domain_move(rng, o, step)):
            a.upload(np.ascontiguousarray(m))


def test_domain_neighbors_on_focus_tree(hip, oracle):
    """findNeighbors over the domain's focus tree (Domain::octreeProperties) equals the oracle's search over the
    oracle domain's tree."""
    n, T, kd = 80000, np.float32, np.uint64
    x, y, z, h = plummer_input(n, T, seed=33)
    gbox = hip.make_box(-1.0, 1.0, T)
    dom = hip.Domain(gbox, key_bits=64, bucket_size=64, bucket_size_focus=16)
    ref = CpuDomain(oracle, kd, T, 64, 16, -1.0, 1.0)
    d = [dev(hip, a) for a in (x, y, z, h)]
    o = ref.sync(x, y, z, h)
    dom.sync(*d, n)
    v = dom.view()
    gb = hip.make_box(v.box_lo[:], v.box_hi[:], T)
    ngmax = 64
    nb = hip.DeviceArray(n * ngmax, np.uint32)
    nc = hip.DeviceArray(n, np.uint32)
    hip.find_neighbors(d[0], d[1], d[2], d[3], 0, n, gb, v.child_offsets, v.parents, v.internal_to_leaf, v.layout,
                       v.centers, v.sizes, v.num_nodes, ngmax, nb, nc)
    t = oracle.build_octree(o["fleaves"])
    c, s = oracle.node_fp_centers(t["prefixes"], ref.box(), T)
    layout = np.concatenate([[0], np.cumsum(o["fcounts"])]).astype(np.uint32)
    ids = np.arange(0, n, 7, dtype=np.uint32)
    nb_o, nc_o = oracle.find_neighbors_subset(o["x"], o["y"], o["z"], o["h"], ids, ref.box(), t, layout, c, s, ngmax)
    g_nb = nb.download().reshape(n, ngmax)[ids]
    g_nc = nc.download()[ids]
    assert neighbor_rows_equal(g_nb, g_nc, nb_o, nc_o, ngmax)
    assert g_nc.mean() > 20


def test_domain_rejects_bad_arguments(hip):
    gbox = hip.make_box(0.0, 1.0, np.float32)
    with pytest.raises(hip.HipcstoneError):
        hip.Domain(gbox, bucket_size=16, bucket_size_focus=64)
    dom = hip.Domain(gbox)
    x = hip.DeviceArray(10, np.float32)
    with pytest.raises(hip.HipcstoneError):
        dom.sync(x, x, x, None, 10)


def test_domain_gathers_properties_in_sync_order(hip, oracle):
    """Properties follow the sync's SFC order (Domain::sync particleProperties / reapplySync at one rank)."""
    n, T, kd = 50000, np.float64, np.uint64
    x, y, z, h = plummer_input(n, T, seed=34)
    dom = hip.Domain(hip.make_box(-1.0, 1.0, T), key_bits=64, bucket_size=32, bucket_size_focus=8)
    ref = CpuDomain(oracle, kd, T, 32, 8, -1.0, 1.0)
    d = [dev(hip, a) for a in (x, y, z, h)]
    ids = np.arange(n, dtype=np.uint32)
    props = [dev(hip, ids), dev(hip, (ids * 3).astype(np.uint64)), dev(hip, ids.astype(np.uint8)),
             dev(hip, ids.astype(np.uint16))]
    o = ref.sync(x, y, z, h)
    dom.sync(*d, n)
    dom.gather_properties(props)
    perm = o["perm"]
    assert np.array_equal(props[0].download(), ids[perm])
    assert np.array_equal(props[1].download(), (ids * 3).astype(np.uint64)[perm])
    assert np.array_equal(props[2].download(), ids.astype(np.uint8)[perm])
    assert np.array_equal(props[3].download(), ids.astype(np.uint16)[perm])


@pytest.mark.parametrize("kd,T,bucket,bucket_focus,theta", [(np.uint64, np.float32, 64, 16, 0.5),
                                                            (np.uint64, np.float64, 32, 8, 0.7),
                                                            (np.uint32, np.float32, 64, 32, 0.6)])
def test_domain_sync_grav_steps(hip, oracle, kd, T, bucket, bucket_focus, theta):
    """Domain::syncGrav at one rank: sync's order and trees, m reordered with x, y, z, h, and the focus tree's
    expansion centers (mass centers + squared vector-MAC radius of 1/theta), bit for bit, over several steps; then
    updateExpansionCenters on moved-in-place coordinates."""
    n = 90000
    x, y, z, h = plummer_input(n, T, seed=35)
    rng = np.random.default_rng(36)
    m = (rng.random(n) * 2 - 0.4).astype(T)
    m[rng.random(n) < 0.03] = 0
    gbox = hip.make_box(-1.0, 1.0, T)
    dom = hip.Domain(gbox, key_bits=64 if kd == np.uint64 else 32, bucket_size=bucket,
                     bucket_size_focus=bucket_focus, theta=theta)
    ref = CpuDomain(oracle, kd, T, bucket, bucket_focus, -1.0, 1.0)
    d = [dev(hip, a) for a in (x, y, z, h, m)]
    for step in range(3):
        print(f"syncGrav step {step}: {np.dtype(kd).name} {np.dtype(T).name} bucket {bucket}/{bucket_focus}",
              flush=True)
        host = [a.download() for a in d]
        o = ref.sync_grav(*host, theta)
        dom.sync_grav(*d)
        for name, a in zip("xyzhm", d):
            assert np.array_equal(a.download(), o[name]), name
        g = dom.arrays()
        assert np.array_equal(g["leaves"], o["fleaves"]) and np.array_equal(g["prefixes"], o["octree"]["prefixes"])
        ec = dom.expansion_centers()
        assert ec is not None and np.array_equal(ec.view(np.uint8), o["expansion_centers"].view(np.uint8))
        assert ec[0, 3] > 0 and (ec[:, 3] == 0).any()
        moved = [(a.download() + rng.normal(0, 0.004, n)).astype(T) for a in d[:3]]
        for a, mv in zip(d[:3], moved):
            a.upload(np.clip(mv, -1, 1).astype(T))
    # plain sync clears the centers; updateExpansionCenters recomputes them for the arrays given
    host = [a.download() for a in d]
    o = ref.sync(*host[:4])
    dom.sync(*d[:4])
    assert dom.expansion_centers() is None
    dom.gather_properties([d[4]])
    ms = host[4][o["perm"]]
    assert np.array_equal(d[4].download(), ms)
    dom.update_expansion_centers(d[0], d[1], d[2], d[4])
    ref_c, _ = ref.expansion_centers(o["x"], o["y"], o["z"], ms, theta)
    assert np.array_equal(dom.expansion_centers().view(np.uint8), ref_c.view(np.uint8))


def test_domain_sync_crowded_bucket_takes_the_radix_sort(hip, oracle):
    """A later sync whose input crowds more than 2048 particles into one leaf of the previous global leaf array hands
    that step's sort from the buckets to the radix sort; the result equals the oracle's domain either way."""
    n, T, kd = 60000, np.float32, np.uint64
    x, y, z, h = plummer_input(n, T, seed=37)
    dom = hip.Domain(hip.make_box(-1.0, 1.0, T), key_bits=64, bucket_size=64, bucket_size_focus=16)
    ref = CpuDomain(oracle, kd, T, 64, 16, -1.0, 1.0)
    d = [dev(hip, a) for a in (x, y, z, h)]
    for step in range(3):
        print(f"domain crowded bucket: step {step}", flush=True)
        xs, ys, zs, hs = (a.download() for a in d)
        o = ref.sync(xs, ys, zs, hs)
        dom.sync(*d, n)
        g = dom.arrays()
        assert np.array_equal(g["keys"], o["keys"])
        for name, a in zip("xyzh", d):
            assert np.array_equal(a.download(), o[name]), name
        assert np.array_equal(g["global_leaves"], o["gtree"]) and np.array_equal(g["leaves"], o["fleaves"])
        nxt = [np.array(o[k], copy=True) for k in "xyzh"]
        if step == 0:
            # 3000 particles of one region onto one point inside the box
            sel = np.arange(n // 2, n // 2 + 3000)
            nxt[0][sel], nxt[1][sel], nxt[2][sel] = T(0.013), T(-0.021), T(0.007)
        for a, m in zip(d, nxt):
            a.upload(np.ascontiguousarray(m))
