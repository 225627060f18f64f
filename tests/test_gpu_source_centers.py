"""GPU parity of the focus-tree source centers (leaf mass centers, upsweep, setMac, geoMacSpheres) with the oracle's
restatement of the reference CPU path, bit for bit."""

import numpy as np
import pytest

from test_oracle_source_centers import _tree

pytestmark = pytest.mark.gpu


def dev(hip, a):
    return hip.DeviceArray.from_numpy(np.ascontiguousarray(a))


def _gpu_centers(hip, p, T):
    N = p["prefixes"].size
    d = {k: dev(hip, p[k]) for k in ("x", "y", "z", "m", "l2i", "layout", "childOffsets")}
    c = hip.DeviceArray(4 * N, T)
    c.fill(0)
    hip.leaf_mass_centers(d["x"], d["y"], d["z"], d["m"], d["l2i"], d["layout"], c)
    leaf_only = c.download().reshape(N, 4)
    hip.upsweep_centers(p["levelRange"], d["childOffsets"], c)
    return c, leaf_only


@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("kd", [np.uint32, np.uint64])
def test_source_centers_and_macs(hip, oracle, T, kd):
    from oracle_py import make_box

    p = _tree(oracle, T, kd, n=60000)
    N = p["prefixes"].size
    c, leaf_only = _gpu_centers(hip, p, T)
    ref_leaf = oracle.leaf_mass_centers(p["x"], p["y"], p["z"], p["m"], p["l2i"], p["layout"], N)
    assert np.array_equal(leaf_only.view(np.uint8), ref_leaf.view(np.uint8))
    ref_up = oracle.upsweep_centers(p["levelRange"], p["childOffsets"], ref_leaf, p["levelRange"].size - 2)
    assert np.array_equal(c.download().reshape(N, 4).view(np.uint8), ref_up.view(np.uint8))
    gbox = hip.make_box(-1.0, 1.0, T)
    dp = dev(hip, p["prefixes"])
    for inv_theta in (2.0, 1.0 / 0.7):
        cm = dev(hip, ref_up)
        hip.set_mac(dp, cm, inv_theta, gbox)
        ref_mac = oracle.set_mac(p["prefixes"], ref_up, inv_theta, p["box"])
        assert np.array_equal(cm.download().reshape(N, 4).view(np.uint8), ref_mac.view(np.uint8))
    g = hip.DeviceArray(4 * N, T)
    hip.geo_mac_spheres(dp, g, 2.5, gbox)
    ref_g = oracle.geo_mac_spheres(p["prefixes"], T, 2.5, p["box"])
    assert np.array_equal(g.download().reshape(N, 4).view(np.uint8), ref_g.view(np.uint8))


def test_source_centers_exact_sums_and_tiny_trees(hip, oracle):
    """Integer masses on a dyadic grid (exact leaf sums, total mass at the root) and a root-only tree."""
    T = np.float64
    p = _tree(oracle, T, np.uint64, n=30000, dyadic=True)
    c, _ = _gpu_centers(hip, p, T)
    got = c.download().reshape(-1, 4)
    assert got[0, 3] == np.abs(p["m"]).sum()
    ref = oracle.upsweep_centers(p["levelRange"], p["childOffsets"],
                                 oracle.leaf_mass_centers(p["x"], p["y"], p["z"], p["m"], p["l2i"], p["layout"],
                                                          p["prefixes"].size), p["levelRange"].size - 2)
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
    q = _tree(oracle, np.float32, np.uint64, n=10, bucket=16)
    assert q["prefixes"].size == 1
    c1, _ = _gpu_centers(hip, q, np.float32)
    r1 = oracle.leaf_mass_centers(q["x"], q["y"], q["z"], q["m"], q["l2i"], q["layout"], 1)
    assert np.array_equal(c1.download().reshape(1, 4).view(np.uint8), r1.view(np.uint8))
