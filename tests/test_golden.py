"""Golden fixtures written by the reference's own CPU code (tests/golden/make_golden.py): the oracle (CPU) and the
HIP path (GPU) must reproduce every one bit for bit."""

import numpy as np
import pytest

from conftest import load_golden
from helpers import neighbor_rows_equal
from oracle_py import HILBERT, MORTON
from oracle_py import make_box as orc_box

VARIANTS = [(c, cs, kd, ks, T, ts) for c, cs in ((HILBERT, "hilbert"), (MORTON, "morton"))
            for kd, ks in ((np.uint32, "u32"), (np.uint64, "u64")) for T, ts in ((np.float32, "f32"),
                                                                                  (np.float64, "f64"))]


@pytest.mark.parametrize("curve,cs,kd,ks,T,ts", VARIANTS)
def test_oracle_keys_golden(oracle, curve, cs, kd, ks, T, ts):
    g = load_golden("keys.npz")
    box = orc_box(g["box_lo"], g["box_hi"], T)
    keys = oracle.compute_sfc_keys(g[f"x_{ts}"], g[f"y_{ts}"], g[f"z_{ts}"], box, kd, curve)
    assert np.array_equal(keys, g[f"keys_{cs}_{ks}_{ts}"])


def test_oracle_tree_golden(oracle):
    g = load_golden("tree.npz")
    box = orc_box(-1.0, 1.0, np.float32)
    keys = oracle.compute_sfc_keys(g["x"], g["y"], g["z"], box, np.uint64, HILBERT)
    assert np.array_equal(keys, g["keys"])
    ks, perm = oracle.sort_by_key(keys)
    assert np.array_equal(ks, g["sorted_keys"]) and np.array_equal(perm, g["perm"])
    tree, counts = oracle.compute_octree(ks, 16)
    assert np.array_equal(tree, g["leaves"]) and np.array_equal(counts, g["counts"])
    o = oracle.build_octree(tree)
    for k in ("prefixes", "childOffsets", "parents", "levelRange", "internalToLeaf", "leafToInternal"):
        assert np.array_equal(o[k], g[k]), k
    c, s = oracle.node_fp_centers(o["prefixes"], box, np.float32)
    assert np.array_equal(c, g["centers"]) and np.array_equal(s, g["sizes"])
    p = g["perm"]
    nb, nc = oracle.find_neighbors(g["x"][p], g["y"][p], g["z"][p], g["h"][p], 0, p.size, box, o,
                                   np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32), c, s, 48)
    assert neighbor_rows_equal(nb, nc, g["neighbors"], g["neighbor_counts"], 48)
    keys2 = oracle.compute_sfc_keys(g["x2"], g["y"][p], g["z"][p], box, np.uint64, HILBERT)
    assert np.array_equal(keys2, g["keys2"])
    ks2, perm2 = oracle.sort_by_key(keys2)
    assert np.array_equal(perm2, g["perm2"])
    t2, c2, conv = oracle.update_octree(ks2, 16, tree, counts)
    assert np.array_equal(t2, g["leaves2"]) and np.array_equal(c2, g["counts2"]) and conv == bool(g["converged2"])


def test_oracle_halos_golden(oracle):
    g = load_golden("halos.npz")
    T = np.float64
    box = orc_box(0.0, 1.0, T)
    first, last = int(g["first"]), int(g["last"])
    sc, ss = oracle.compute_bounding_box(g["x"], g["y"], g["z"], g["h"], g["layout"], first, last, T(2),
                                         g["search_init"])
    assert np.array_equal(sc, g["search_centers"]) and np.array_equal(ss, g["search_sizes"])
    flags = oracle.find_halos(g["prefixes"], g["childOffsets"], g["parents"], g["centers"], g["sizes"], g["leaves"],
                              sc, ss, box, first, last)
    assert np.array_equal(flags, g["flags"])


# ---------------------------------------------------------------- GPU against the same fixtures


def _dev(hip, a):
    return hip.DeviceArray.from_numpy(np.ascontiguousarray(a))


@pytest.mark.gpu
@pytest.mark.parametrize("curve,cs,kd,ks,T,ts", VARIANTS)
def test_gpu_keys_golden(hip, curve, cs, kd, ks, T, ts):
    g = load_golden("keys.npz")
    box = hip.make_box(g["box_lo"], g["box_hi"], T)
    dk = hip.DeviceArray(g[f"x_{ts}"].size, kd)
    hip.compute_sfc_keys(_dev(hip, g[f"x_{ts}"]), _dev(hip, g[f"y_{ts}"]), _dev(hip, g[f"z_{ts}"]), dk, box, curve)
    assert np.array_equal(dk.download(), g[f"keys_{cs}_{ks}_{ts}"])


@pytest.mark.gpu
def test_gpu_pipeline_golden(hip):
    g = load_golden("tree.npz")
    n = g["x"].size
    box = hip.make_box(-1.0, 1.0, np.float32)
    pipe = hip.Pipeline(curve=hip.HILBERT, key_bits=64, real_bits=32, bucket_size=16, ngmax=48)
    nb = hip.DeviceArray(n * 48, np.uint32)
    nc = hip.DeviceArray(n, np.uint32)
    pipe.sync(*(_dev(hip, g[k]) for k in ("x", "y", "z", "h")), n, box, full_build=True, neighbors=nb, counts=nc)
    a = pipe.arrays()
    assert np.array_equal(a["keys"], g["sorted_keys"]) and np.array_equal(a["perm"], g["perm"])
    for k in ("leaves", "counts", "prefixes", "childOffsets", "parents", "levelRange", "internalToLeaf",
              "leafToInternal", "centers", "sizes"):
        assert np.array_equal(a[k], g[k]), k
    assert neighbor_rows_equal(nb.download().reshape(n, 48), nc.download(), g["neighbors"], g["neighbor_counts"], 48)
    # the warm step of the fixture: x displaced, y/z/h as sorted
    p = g["perm"]
    pipe.sync(_dev(hip, g["x2"]), _dev(hip, g["y"][p]), _dev(hip, g["z"][p]), _dev(hip, g["h"][p]), n, box)
    a = pipe.arrays()
    assert np.array_equal(a["perm"], g["perm2"])
    assert np.array_equal(a["leaves"], g["leaves2"]) and np.array_equal(a["counts"], g["counts2"])


@pytest.mark.gpu
def test_gpu_halos_golden(hip):
    g = load_golden("halos.npz")
    T = np.float64
    box = hip.make_box(0.0, 1.0, T)
    first, last = int(g["first"]), int(g["last"])
    L = g["counts"].size
    N = g["prefixes"].size
    sc = _dev(hip, g["search_init"].reshape(-1))
    ss = hip.DeviceArray(3 * L, T)
    ss.fill(0)
    hip.compute_bounding_box(*(_dev(hip, g[k]) for k in ("x", "y", "z", "h", "layout")), first, last, T(2), sc, ss)
    gsc, gss = sc.download().reshape(L, 3), ss.download().reshape(L, 3)
    assert np.array_equal(gsc[first:last], g["search_centers"][first:last])
    assert np.array_equal(gss[first:last], g["search_sizes"][first:last])
    flags = hip.DeviceArray(N, np.uint8)
    flags.fill(0)
    hip.find_halos(_dev(hip, g["prefixes"]), _dev(hip, g["childOffsets"]), _dev(hip, g["parents"]),
                   _dev(hip, g["centers"].reshape(-1)), _dev(hip, g["sizes"].reshape(-1)), _dev(hip, g["leaves"]), sc,
                   ss, box, first, last, N, flags)
    assert np.array_equal(flags.download(), g["flags"])
