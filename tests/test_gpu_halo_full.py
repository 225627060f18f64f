"""The rank-share chain at its real size: 134,217,728 fp64 Gaussian-mixture particles, Morton-64 keys,
bucketSize 128, rank 2 of 8. Every stage is checked against the oracle on the GPU's own intermediate arrays
(sorted keys, leaf array, octree, centers, search boxes, halo flags, local layout, gathered particles), plus the
size-independent properties of the sort (sortedness, permutation, gather by the permutation)."""

import os

import numpy as np
import pytest

from oracle_py import MORTON
from oracle_py import make_box as orc_box

pytestmark = pytest.mark.gpu

N = int(os.environ.get("HIPCSTONE_HALO_N", 134217728))


def test_halo_chain_full_size(hip, oracle):
    T, kd, bucket, ranks, rank = np.float64, np.uint64, 128, 8, 2
    print(f"halo chain: n={N} fp64 morton-64 bucket {bucket} rank {rank}/{ranks}", flush=True)
    d = [hip.DeviceArray(N, T) for _ in range(4)]
    hip.gaussian_mixture(*d, N, seed=7, clusters=64, background=0.2, target_neighbors=100.0)
    pipe = hip.Pipeline(curve=MORTON, key_bits=64, real_bits=64, bucket_size=bucket, ngmax=0)
    pipe.sync(*d, N, hip.make_box(0.0, 1.0, T), full_build=True)
    m = pipe.halos(ranks, rank)
    outs = [hip.DeviceArray(m, T) for _ in range(4)]
    assert pipe.halos(ranks, rank, outs, m) == m
    hip.device_sync()
    a = pipe.arrays()
    v = pipe.view()
    L, NN = v.num_leaves, v.num_nodes
    obox = orc_box(0.0, 1.0, T)
    print(f"  leaves {L}, nodes {NN}, local {m}", flush=True)

    # sort + gather: sorted keys, a permutation, particles gathered by it, keys of the sorted particles
    keys, perm = a["keys"], a["perm"]
    assert np.all(keys[1:] >= keys[:-1])
    seen = np.zeros(N, dtype=np.bool_)
    seen[perm] = True
    assert seen.all()
    del seen
    for name, dv in zip("xyzh", d):
        assert np.array_equal(dv.download()[perm], a[name]), name
    assert np.array_equal(oracle.compute_sfc_keys(a["x"], a["y"], a["z"], obox, kd, MORTON), keys)
    print("  keys/sort/gather ok", flush=True)

    # leaf array, octree, centers
    tree, counts = oracle.compute_octree(keys, bucket)
    assert np.array_equal(tree, a["leaves"]) and np.array_equal(counts, a["counts"])
    o = oracle.build_octree(tree)
    for k in ("prefixes", "childOffsets", "internalToLeaf", "leafToInternal"):
        assert np.array_equal(o[k], a[k]), k
    c, s = oracle.node_fp_centers(o["prefixes"], obox, T)
    assert np.array_equal(c, a["centers"]) and np.array_equal(s, a["sizes"])
    print("  tree/link/centers ok", flush=True)

    # assignment, halos, layout, gather
    off, _, _ = oracle.make_sfc_assignment(tree, counts, ranks)
    first, last = int(off[rank]), int(off[rank + 1])
    assert (v.assigned_first, v.assigned_last) == (first, last)
    t = {k: a[k] for k in ("prefixes", "childOffsets", "parents", "leafToInternal")}
    sc, ss, flags = oracle.discover_halos(t, a["centers"], a["sizes"], a["leaves"], a["x"], a["y"], a["z"], a["h"],
                                          a["layout"], first, last, obox)
    assert np.array_equal(hip.download_ptr(v.halo_flags, NN, np.uint8), flags)
    gsc = hip.download_ptr(v.search_centers, 3 * L, T).reshape(L, 3)
    gss = hip.download_ptr(v.search_sizes, 3 * L, T).reshape(L, 3)
    assert np.array_equal(gsc[first:last], sc[first:last]) and np.array_equal(gss[first:last], ss[first:last])
    local = oracle.compute_node_layout(counts, flags, a["leafToInternal"][NN - L:], first, last)
    assert np.array_equal(hip.download_ptr(v.local_layout, L + 1, np.uint32), local)
    assert int(local[-1]) == m
    for name, ov in zip("xyzh", outs):
        assert np.array_equal(ov.download(), oracle.gather_leaf_ranges(a["layout"], local, a[name])), name
    print(f"  halos ok: {int(flags.sum())} flagged nodes, {m - int(a['layout'][last] - a['layout'][first])} halo "
          f"particles", flush=True)
