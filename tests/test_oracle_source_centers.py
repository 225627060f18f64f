"""The oracle's restatement of the focus-tree source centers (focus/source_center.hpp: computeLeafMassCenter,
CombineSourceCenter upsweep, setMac, geoMacSpheres) against the reference's own CPU functions (oracle/_ref), and
against what the arithmetic fixes: exact sums for integer masses on a dyadic grid, the root's mass, the total mass
center, and the min-MAC radius a vector MAC reduces to when the mass center is the geometric center."""

import numpy as np
import pytest

from helpers import cpu_pipeline
from oracle_py import HILBERT, MORTON
from oracle_py import make_box


def _tree(lib, T, kd, n=20000, seed=3, bucket=16, curve=HILBERT, dyadic=False):
    from hipcstone import inputs

    rng = np.random.default_rng(seed)
    if dyadic:
        # coordinates on a 2^-12 grid and integer masses: every weighted sum of a leaf is exact in f64
        x, y, z = ((rng.integers(-4096, 4096, size=(3, n)) / 4096.0).astype(T))
        h = np.full(n, 0.01, T)
        m = rng.integers(-3, 4, size=n).astype(T)
    else:
        x, y, z, r = inputs.plummer(n, 0.2, seed=seed, dtype=T)
        h = inputs.plummer_h(r, n, 0.2, 60, dtype=T)
        m = (rng.random(n) * 2 - 0.5).astype(T)
        m[rng.random(n) < 0.05] = 0
    box = make_box(-1.0, 1.0, T)
    p = cpu_pipeline(lib, x, y, z, h, box, kd, bucket, curve)
    p["m"] = m[p["perm"]]
    n_int = p["prefixes"].size - (p["layout"].size - 1)
    p["l2i"] = p["leafToInternal"][n_int:]
    p["box"] = box
    return p


def _centers(lib, p):
    c = lib.leaf_mass_centers(p["x"], p["y"], p["z"], p["m"], p["l2i"], p["layout"], p["prefixes"].size)
    return lib.upsweep_centers(p["levelRange"], p["childOffsets"], c, p["levelRange"].size - 2)


@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("kd", [np.uint32, np.uint64])
@pytest.mark.parametrize("curve", [HILBERT, MORTON])
def test_source_centers_match_reference(oracle, ref, T, kd, curve):
    p = _tree(oracle, T, kd, curve=curve)
    N = p["prefixes"].size
    lo = oracle.leaf_mass_centers(p["x"], p["y"], p["z"], p["m"], p["l2i"], p["layout"], N)
    lr = ref.leaf_mass_centers(p["x"], p["y"], p["z"], p["m"], p["l2i"], p["layout"], N)
    assert np.array_equal(lo.view(np.uint8), lr.view(np.uint8))
    L = p["levelRange"].size - 2
    uo = oracle.upsweep_centers(p["levelRange"], p["childOffsets"], lo, L)
    ur = ref.upsweep_centers(p["levelRange"], p["childOffsets"], lr, L)
    assert np.array_equal(uo.view(np.uint8), ur.view(np.uint8))
    for inv_theta in (1.0 / 0.5, 1.0 / 0.8):
        so = oracle.set_mac(p["prefixes"], uo, inv_theta, p["box"])
        sr = ref.set_mac(p["prefixes"], ur, inv_theta, p["box"])
        assert np.array_equal(so.view(np.uint8), sr.view(np.uint8))
    go = oracle.geo_mac_spheres(p["prefixes"], T, 2.5, p["box"])
    gr = ref.geo_mac_spheres(p["prefixes"], T, 2.5, p["box"])
    assert np.array_equal(go.view(np.uint8), gr.view(np.uint8))
    # zero-mass leaves exist and keep (0, 0, 0, 0); setMac leaves their radius at 0
    zero = uo[:, 3] == 0
    assert zero.any() and not uo[zero, :3].any() and not so[zero, 3].any()


def test_leaf_centers_exact_sums(oracle):
    """f64, integer masses, dyadic coordinates: a leaf's sums are exact, so its center is each exact weighted sum
    times the rounded 1/M, as normalizeMass computes it (source_center.hpp:59-68)."""
    p = _tree(oracle, np.float64, np.uint64, n=30000, dyadic=True)
    N = p["prefixes"].size
    c = oracle.leaf_mass_centers(p["x"], p["y"], p["z"], p["m"], p["l2i"], p["layout"], N)
    lay = p["layout"].astype(np.int64)
    w = np.abs(p["m"])
    for leaf in range(0, lay.size - 1, 7):
        a, b = lay[leaf], lay[leaf + 1]
        M = w[a:b].sum()
        node = p["l2i"][leaf]
        assert c[node, 3] == M
        inv = 1.0 / M if M != 0 else 1.0
        for d, coord in enumerate(("x", "y", "z")):
            assert c[node, d] == (w[a:b] * p[coord][a:b]).sum() * inv


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_root_is_total_mass_center(oracle, T):
    p = _tree(oracle, T, np.uint64, n=20000, dyadic=True)
    c = _centers(oracle, p)
    w = np.abs(p["m"]).astype(np.float64)
    assert c[0, 3] == w.sum()  # integer masses: every partial sum is exact
    for d, coord in enumerate(("x", "y", "z")):
        exact = (w * p[coord].astype(np.float64)).sum() / w.sum()
        assert abs(c[0, d] - exact) < (1e-5 if T == np.float32 else 1e-12)
    # every internal node's mass is the sum of its children's
    co = p["childOffsets"]
    internal = np.nonzero(co)[0]
    kids = co[internal][:, None] + np.arange(8)[None, :]
    assert np.array_equal(c[internal, 3], c[kids, 3].sum(axis=1))


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_vector_mac_at_geometric_center_is_min_mac(oracle, T):
    """With the expansion center at the node's geometric center, |center - geoCenter| = 0 and the vector MAC radius
    (l * invTheta + 0)^2 equals the min-MAC radius of geoMacSpheres with the same invTheta (macs.hpp)."""
    p = _tree(oracle, T, np.uint64, n=8000)
    g = oracle.geo_mac_spheres(p["prefixes"], T, 2.0, p["box"])
    c = g.copy()
    c[:, 3] = 1
    s = oracle.set_mac(p["prefixes"], c, 2.0, p["box"])
    assert np.array_equal(s[:, 3], g[:, 3])
    # geometric centers agree with nodeFpCenters, and l = 2 * max(size)
    gc, gs = oracle.node_fp_centers(p["prefixes"], p["box"], T)
    assert np.array_equal(g[:, :3], gc)
    l = (T(2) * gs.max(axis=1)).astype(T)
    assert np.array_equal(g[:, 3], (l * T(2.0)) * (l * T(2.0)))
