"""Golden fixtures written by the reference's own CPU code (tests/golden/make_golden.py): the oracle (CPU) and the
HIP path (GPU) must reproduce every one bit for bit."""

import numpy as np
import pytest

from conftest import load_golden
from helpers import DOMAIN_GOLDEN_CASES, domain_input, domain_move, neighbor_rows_equal
from oracle_py import HILBERT, MORTON, CpuDomain
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


def _domain_check(g, name, step, got, T):
    """One call's outputs against domain.npz: `got` holds keys, x/y/z/h (reordered), box, trees and layout."""
    p = f"{name}{step}_"
    assert np.array_equal(got["keys"], g[p + "keys"])
    box = np.array([got["lo"][0], got["hi"][0], got["lo"][1], got["hi"][1], got["lo"][2], got["hi"][2]], dtype=T)
    assert np.array_equal(box, g[p + "box"])
    for k in ("gtree", "gcounts", "fleaves", "fcounts", "fnodecounts", "layout"):
        assert np.array_equal(got[k], g[p + k]), (name, step, k)


@pytest.mark.parametrize("case", DOMAIN_GOLDEN_CASES, ids=[c[0] for c in DOMAIN_GOLDEN_CASES])
def test_oracle_domain_golden(oracle, case):
    """The oracle's one-rank Domain sequence reproduces the reference Domain's calls stored in domain.npz."""
    g = load_golden("domain.npz")
    name, kd, T, bucket, bucket_focus, boundary, theta, n, calls, grav, seed = case
    arrays = domain_input(n, T, seed, with_mass=grav)
    assert np.array_equal([float(np.sum(a, dtype=np.float64)) for a in arrays], g[f"{name}_input_sum"]), \
        "the input generator drifted from the one that wrote the fixture"
    dom = CpuDomain(oracle, kd, T, bucket, bucket_focus, -1.0, 1.0, boundary)
    rng = np.random.default_rng(seed + 1000)
    for step in range(calls):
        r = dom.sync_grav(*arrays, theta) if grav else dom.sync(*arrays)
        assert np.array_equal(r["perm"], g[f"{name}{step}_perm"])
        r["layout"] = np.concatenate([[0], np.cumsum(r["fcounts"])]).astype(np.uint32)
        _domain_check(g, name, step, r, T)
        if grav:
            assert np.array_equal(r["expansion_centers"].view(np.uint8).ravel(),
                                  g[f"{name}{step}_expansion_centers"].view(np.uint8).ravel())
        arrays = domain_move(rng, r, step, carried=("h", "m") if grav else ("h",))


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


@pytest.mark.gpu
@pytest.mark.parametrize("case", DOMAIN_GOLDEN_CASES, ids=[c[0] for c in DOMAIN_GOLDEN_CASES])
def test_gpu_domain_golden(hip, case):
    """hipcstone_domain_sync / _sync_grav reproduce the reference Domain's calls stored in domain.npz, bit for bit:
    box, keys, particle order, global and focus trees, layout and expansion centers."""
    g = load_golden("domain.npz")
    name, kd, T, bucket, bucket_focus, boundary, theta, n, calls, grav, seed = case
    arrays = domain_input(n, T, seed, with_mass=grav)
    assert np.array_equal([float(np.sum(a, dtype=np.float64)) for a in arrays], g[f"{name}_input_sum"]), \
        "the input generator drifted from the one that wrote the fixture"
    dom = hip.Domain(hip.make_box(-1.0, 1.0, T, boundary), key_bits=64 if kd == np.uint64 else 32,
                     bucket_size=bucket, bucket_size_focus=bucket_focus, theta=theta)
    rng = np.random.default_rng(seed + 1000)
    keys_out = hip.DeviceArray(n, kd)
    for step in range(calls):
        print(f"domain golden case {name} call {step}", flush=True)
        d = [_dev(hip, a) for a in arrays]
        if grav:
            dom.sync_grav(*d, keys_out=keys_out)
        else:
            dom.sync(*d, n, keys_out)
        a = dom.arrays()
        perm = g[f"{name}{step}_perm"]
        got = dict(keys=keys_out.download(), lo=[T(v) for v in a["box_lo"]], hi=[T(v) for v in a["box_hi"]],
                   gtree=a["global_leaves"], gcounts=a["global_counts"], fleaves=a["leaves"], fcounts=a["counts"],
                   fnodecounts=a["node_counts"], layout=a["layout"])
        _domain_check(g, name, step, got, T)
        out = {}
        for k, dv, hv in zip("xyzhm", d, arrays):
            out[k] = dv.download()
            assert np.array_equal(out[k], hv[perm]), (name, step, k)
        if grav:
            ec = dom.expansion_centers()
            assert np.array_equal(ec.view(np.uint8).ravel(),
                                  g[f"{name}{step}_expansion_centers"].view(np.uint8).ravel())
        arrays = domain_move(rng, out, step, carried=("h", "m") if grav else ("h",))
