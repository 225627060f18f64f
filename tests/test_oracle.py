"""Pins the oracle (the C restatement in oracle/) before anything is checked against it.

Each oracle function is checked against (a) known-answer tests taken from the reference's own unit tests,
(b) brute force / closed forms, and (c) the reference's headers compiled into oracle/_ref when present.
"""

import ctypes as C

import numpy as np
import pytest

from helpers import OctreeMaker, brute_neighbors, cpu_pipeline, neighbor_rows_equal, node_range, uniform_n_level_tree
from oracle_py import HILBERT, MORTON, make_box

KDS = [np.uint32, np.uint64]


def _lib_fn(oracle, name, restype, argtypes):
    f = getattr(oracle.lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


def _ks(kd):
    return "u32" if kd == np.uint32 else "u64"


def _ctype(kd):
    return C.c_uint32 if kd == np.uint32 else C.c_uint64


# ---------------------------------------------------------------- keys


@pytest.mark.parametrize("kd", KDS)
def test_morton_kat(oracle, kd):
    # test/unit/sfc/morton.cpp:26-40: iMorton(5,3,6) at tree level 3 == pad(0b101011110, 9)
    imorton = _lib_fn(oracle, f"orc_imorton_{_ks(kd)}", _ctype(kd), [C.c_uint, C.c_uint, C.c_uint])
    L = 10 if kd == np.uint32 else 21
    shifts = L - 3
    code = imorton(5 << shifts, 3 << shifts, 6 << shifts)
    assert code == (0b101011110 << (3 * L - 9))


def test_morton_decode_kat(oracle):
    # morton.cpp:42-73
    dec32 = _lib_fn(oracle, "orc_decode_morton_u32", None, [C.c_uint32] + [C.POINTER(C.c_uint)] * 3)
    dec64 = _lib_fn(oracle, "orc_decode_morton_u64", None, [C.c_uint64] + [C.POINTER(C.c_uint)] * 3)
    a, b, c = C.c_uint(), C.c_uint(), C.c_uint()
    dec32(340, C.byref(a), C.byref(b), C.byref(c))
    assert (a.value, b.value, c.value) == (5, 2, 4)
    dec64(0x7FFFFFFFFFFFFFFF, C.byref(a), C.byref(b), C.byref(c))
    assert (a.value, b.value, c.value) == ((1 << 21) - 1,) * 3
    dec64(0x1249249241249249, C.byref(a), C.byref(b), C.byref(c))
    assert c.value == (1 << 21) - 512 - 1
    dec64(0b0111 << 60, C.byref(a), C.byref(b), C.byref(c))
    assert (a.value, b.value, c.value) == (1 << 20, 1 << 20, 1 << 20)
    dec64(0b0011 << 60, C.byref(a), C.byref(b), C.byref(c))
    assert (a.value, b.value, c.value) == (0, 1 << 20, 1 << 20)


@pytest.mark.parametrize("kd", KDS)
def test_hilbert_first_order_curve(oracle, kd):
    # test/unit/sfc/hilbert.cpp:83-118: the level-1 octant visit order
    ihil = _lib_fn(oracle, f"orc_ihilbert_{_ks(kd)}", _ctype(kd), [C.c_uint, C.c_uint, C.c_uint])
    L = 10 if kd == np.uint32 else 21
    hilbert_to_morton = [0, 1, 3, 2, 6, 7, 5, 4]
    half = (1 << L) // 2
    for xi in (0, 1):
        for yi in (0, 1):
            for zi in (0, 1):
                morton_octant = 4 * xi + 2 * yi + zi
                for off in (0, half - 1):
                    key = ihil(half * xi + off, half * yi + off, half * zi + off)
                    octant = (key >> (3 * (L - 1))) & 7
                    assert hilbert_to_morton[octant] == morton_octant


@pytest.mark.parametrize("kd", KDS)
def test_hilbert_continuity_and_inversion(oracle, kd):
    # hilbert.cpp:120-188: consecutive keys are face neighbours; decode inverts encode
    L = 10 if kd == np.uint32 else 21
    dec = _lib_fn(oracle, f"orc_decode_hilbert_{_ks(kd)}", None, [_ctype(kd)] + [C.POINTER(C.c_uint)] * 3)
    ihil = _lib_fn(oracle, f"orc_ihilbert_{_ks(kd)}", _ctype(kd), [C.c_uint, C.c_uint, C.c_uint])

    def decode(k):
        a, b, c = C.c_uint(), C.c_uint(), C.c_uint()
        dec(k, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    for level in range(1, L):
        for octant in range(8 if level > 1 else 7):
            last = (octant + 1) * int(node_range(kd, level)) - 1
            p, q = decode(last), decode(last + 1)
            assert sum(abs(int(u) - int(v)) for u, v in zip(p, q)) == 1
    rng = np.random.default_rng(0)
    pts = rng.integers(0, 1 << L, size=(1000, 3))
    for px, py, pz in pts:
        assert decode(ihil(int(px), int(py), int(pz))) == (px, py, pz)


@pytest.mark.parametrize("kd", KDS)
@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_sfc3d_closed_form_unit_box(oracle, kd, T):
    # on the unit box sfc3D reduces to ix = min(floor(x * 2^L), 2^L - 1) (sfc.hpp:142-179)
    L = 10 if kd == np.uint32 else 21
    rng = np.random.default_rng(1)
    x, y, z = (rng.random(2000).astype(T) for _ in range(3))
    keys = oracle.compute_sfc_keys(x, y, z, make_box(0, 1, T), kd, MORTON)
    imorton = _lib_fn(oracle, f"orc_imorton_{_ks(kd)}", _ctype(kd), [C.c_uint, C.c_uint, C.c_uint])
    grid = lambda a: np.minimum(np.floor(a.astype(np.float64) * (1 << L)), (1 << L) - 1).astype(np.int64)
    ix, iy, iz = grid(x), grid(y), grid(z)
    for k in range(0, 2000, 7):
        assert keys[k] == imorton(int(ix[k]), int(iy[k]), int(iz[k]))


def test_remove_key_is_kept(oracle):
    # sfc.hpp:268-276: keys equal to removeKey = 2^63 are not recomputed
    x = np.array([0.1, 0.2, 0.3])
    keys_in = np.array([0, 1 << 63, 5], dtype=np.uint64)
    k = oracle.compute_sfc_keys(x, x, x, make_box(0, 1), np.uint64, HILBERT, keys_in=keys_in)
    assert k[1] == 1 << 63 and k[0] != 0


# ---------------------------------------------------------------- sort


@pytest.mark.parametrize("kd", KDS)
def test_sort_is_stable_sort(oracle, kd):
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 50, size=5000).astype(kd)  # many duplicates: pins tie order
    ks, perm = oracle.sort_by_key(keys)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(perm, ref) and np.array_equal(ks, keys[ref])
    # gather.cpp:24-37 sortInvert KAT: {2,1,5,4} -> {1,0,3,2}
    ks, perm = oracle.sort_by_key(np.array([2, 1, 5, 4], dtype=kd))
    assert list(perm) == [1, 0, 3, 2]


# ---------------------------------------------------------------- leaf array


@pytest.mark.parametrize("kd", KDS)
def test_node_counts_kat(oracle, kd):
    # test/unit/tree/csarray.cpp:56-80
    tree = OctreeMaker(kd).divide().divide(0).make_tree()
    t = [int(v) for v in tree]
    keys = np.array([t[1], t[1], t[1] + 10, t[1] + 100, t[2] - 1, t[2] + 1, t[11], t[11] + 2, t[12], t[12] + 1000,
                     t[12] + 2000, t[13] - 10, t[13], t[13] + 1], dtype=kd)
    ref = [0, 5, 1, 0, 0, 0, 0, 0, 0, 0, 0, 2, 4, 2, 0]
    assert list(oracle.compute_node_counts(tree, keys)) == ref
    # with guesses (csarray.hpp:212-223) the result is the same
    assert list(oracle.compute_node_counts(tree, keys, guess=np.ones(15, np.uint32))) == ref


@pytest.mark.parametrize("kd", KDS)
def test_node_counts_brute_force(oracle, kd):
    rng = np.random.default_rng(3)
    tree = OctreeMaker(kd).divide().divide(0).divide(3).divide(0, 4).make_tree()
    keys = np.sort(rng.integers(0, int(tree[-1]), size=20000, dtype=np.uint64).astype(kd))
    lb = np.searchsorted(keys, tree, side="left")
    assert np.array_equal(oracle.compute_node_counts(tree, keys), np.diff(lb).astype(np.uint32))
    # max count caps
    assert oracle.compute_node_counts(tree, keys, max_count=7).max() <= 7


@pytest.mark.parametrize("kd", KDS)
def test_rebalance_decision_kat(oracle, kd):
    # csarray.cpp:108-122
    tree = OctreeMaker(kd).divide().divide(0).make_tree()
    counts = np.array([1, 1, 1, 0, 0, 0, 0, 0, 2, 3, 4, 5, 6, 7, 8], dtype=np.uint32)
    ops, conv = oracle.rebalance_decision(tree, counts, 4)
    assert list(ops) == [1, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 8, 8, 8, 8] and not conv


@pytest.mark.parametrize("kd", KDS)
def test_rebalance_tree_kat(oracle, kd):
    # csarray.cpp:193-212
    tree = OctreeMaker(kd).divide().divide(0).make_tree()
    ops = np.array([1, 0, 0, 0, 0, 0, 0, 0, 1, 8, 1, 1, 1, 1, 8, 0], dtype=np.int32)
    f = getattr(oracle.lib, f"orc_rebalance_tree_{_ks(kd)}")
    f.restype = C.c_int
    new_n = f(tree.ctypes.data_as(C.c_void_p), C.c_int(15), ops.ctypes.data_as(C.c_void_p), C.c_void_p(0))
    ops = np.array([1, 0, 0, 0, 0, 0, 0, 0, 1, 8, 1, 1, 1, 1, 8, 0], dtype=np.int32)
    new_tree = np.zeros(new_n + 1, dtype=kd)
    f(tree.ctypes.data_as(C.c_void_p), C.c_int(15), ops.ctypes.data_as(C.c_void_p),
      new_tree.ctypes.data_as(C.c_void_p))
    ref = OctreeMaker(kd).divide().divide(2).divide(7).make_tree()
    assert np.array_equal(new_tree, ref)


def _check_converged_tree(tree, counts, keys, bucket, kd):
    # cs_util.hpp:42-62 invariants + the unique fixpoint of the rebalance rule
    L = 10 if kd == np.uint32 else 21
    assert tree[0] == 0 and tree[-1] == node_range(kd, 0)
    rng = np.diff(tree.astype(np.uint64))
    assert np.all(rng > 0)
    levels = np.array([(L - (int(r).bit_length() - 1) // 3) for r in rng])
    assert np.all(rng == np.array([1 << (3 * (L - l)) for l in levels], dtype=np.uint64))
    assert np.all(tree[:-1].astype(np.uint64) % rng == 0)  # aligned
    assert np.array_equal(np.diff(np.searchsorted(keys, tree)).astype(np.uint32), counts)
    assert np.all((counts <= bucket) | (levels == L))


@pytest.mark.parametrize("kd", KDS)
@pytest.mark.parametrize("bucket", [1, 16, 64, 1024])
def test_compute_octree_invariants(oracle, kd, bucket):
    # ComputeOctreeTester pattern (csarray.cpp:303-348) on Gaussian-clustered keys
    rng = np.random.default_rng(4)
    L = 10 if kd == np.uint32 else 21
    pts = np.clip(rng.normal(0.5, 0.1, size=(30000, 3)), 0, 1 - 1e-9)
    x, y, z = pts.T.copy()
    keys = np.sort(oracle.compute_sfc_keys(x, y, z, make_box(0, 1), kd, HILBERT))
    tree, counts = oracle.compute_octree(keys, bucket)
    _check_converged_tree(tree, counts, keys, bucket, kd)
    # every internal node of the converged tree holds more than bucket particles (merge rule, csarray.hpp:273-281)
    o = oracle.build_octree(tree)
    pf = o["prefixes"]
    internal = o["childOffsets"] != 0
    for p in pf[internal][:200]:
        plen = int(p).bit_length() - 1
        start = (int(p) ^ (1 << plen)) << (3 * L - plen)
        end = start + (1 << (3 * L - plen))
        cnt = np.searchsorted(keys, end) - np.searchsorted(keys, start)
        assert cnt > bucket


# ---------------------------------------------------------------- link


def _check_connectivity(o, leaves, kd):
    # octree.cpp:27-74 checkConnectivity
    L = 10 if kd == np.uint32 else 21
    pf = o["prefixes"]
    co = o["childOffsets"]
    par = o["parents"]
    N = pf.size
    for i in range(N):
        if co[i]:
            plen = int(pf[i]).bit_length() - 1
            for c in range(8):
                child = co[i] + c
                assert int(pf[child]).bit_length() - 1 == plen + 3
                assert (int(pf[child]) >> 3) == int(pf[i])
                assert par[(child - 1) // 8] == i
    # leaves in leaf order map to their prefixes
    NI = N - (leaves.size - 1)
    for li in range(leaves.size - 1):
        node = o["leafToInternal"][NI + li]
        assert co[node] == 0 and o["internalToLeaf"][node] == li


@pytest.mark.parametrize("kd", KDS)
def test_link_uniform_kat(oracle, kd):
    # octree.cpp:99-124: 4x4x4 grid -> 73 nodes, 1/8/64 per level
    tree = uniform_n_level_tree(kd, 2)
    o = oracle.build_octree(tree)
    lr = o["levelRange"]
    assert lr[1] - lr[0] == 1 and lr[2] - lr[1] == 8 and lr[3] - lr[2] == 64 and lr[-1] == 73
    _check_connectivity(o, tree, kd)


@pytest.mark.parametrize("kd", KDS)
def test_link_irregular_kat(oracle, kd):
    # octree.cpp:134-182: root divided, then octant 0 divided -> 15 leaves, 2 internal nodes
    tree = OctreeMaker(kd).divide().divide(0).make_tree()
    o = oracle.build_octree(tree)
    assert o["prefixes"].size == 17
    assert int(np.count_nonzero(o["childOffsets"])) == 2
    _check_connectivity(o, tree, kd)


# ---------------------------------------------------------------- centers


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_centers_closed_form(oracle, T):
    # box.hpp:318-335 on a cube: center = min + (imax+imin) * 0.5 * 2^-L * len, size = (imax-imin) * ...
    kd = np.uint64
    tree = uniform_n_level_tree(kd, 1)
    o = oracle.build_octree(tree)
    box = make_box(-1, 1, T)
    c, s = oracle.node_fp_centers(o["prefixes"], box, T)
    assert np.allclose(c[0], 0) and np.allclose(s[0], 1)
    for i in range(1, 9):
        assert np.allclose(s[i], 0.5)
        assert np.all(np.abs(np.abs(c[i]) - 0.5) < 1e-7)
    # the 8 level-1 centers are the 8 distinct octant centers
    assert len({tuple(np.sign(v)) for v in c[1:9]}) == 8


# ---------------------------------------------------------------- neighbors


@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("kd,curve", [(np.uint64, HILBERT), (np.uint32, HILBERT)])
def test_neighbors_vs_brute_force(oracle, T, kd, curve):
    # findneighbors.cpp:43-136 pattern: the tree search equals the all-pairs search
    from hipcstone import inputs

    n = 3000
    x, y, z = inputs.uniform_cube(n, seed=5, dtype=T)
    h = np.full(n, inputs.uniform_h(n, 60), dtype=T)
    box = make_box(0, 1, T)
    p = cpu_pipeline(oracle, x, y, z, h, box, kd, 16, curve, ngmax=200)
    nb, nc = brute_neighbors(p["x"], p["y"], p["z"], p["h"], 200)
    assert neighbor_rows_equal(p["nb"], p["nc"], nb, nc, 200)


def test_neighbors_truncation_keeps_smallest(oracle):
    from hipcstone import inputs

    n = 4000
    x, y, z = inputs.uniform_cube(n, seed=6)
    h = np.full(n, inputs.uniform_h(n, 100))
    box = make_box(0, 1)
    full = cpu_pipeline(oracle, x, y, z, h, box, np.uint64, 32, ngmax=400)
    p = cpu_pipeline(oracle, x, y, z, h, box, np.uint64, 32, ngmax=20)
    assert np.array_equal(full["nc"], p["nc"]) and p["nc"].max() > 20
    assert np.array_equal(full["nb"][:, :20], p["nb"])


# ---------------------------------------------------------------- halos


@pytest.mark.parametrize("kd", KDS)
def test_find_halos_kat(oracle, kd):
    # test/unit/traversal/discovery.cpp:52-109: 4x4x4 grid, radius 0.1, local halves -> 21 flags each
    T = np.float64
    box = make_box(0, 1, T)
    tree = uniform_n_level_tree(kd, 2)
    o = oracle.build_octree(tree)
    c, s = oracle.node_fp_centers(o["prefixes"], box, T)
    NI = o["prefixes"].size - 64
    l2i = o["leafToInternal"][NI:]
    tC = c[l2i]
    tS = s[l2i] + 0.1
    for first, last in ((0, 32), (32, 64)):
        flags = oracle.find_halos(o["prefixes"], o["childOffsets"], o["parents"], c, s, tree, tC, tS, box, first, last)
        assert int(flags.sum()) == 21


# ---------------------------------------------------------------- against the compiled reference


@pytest.mark.parametrize("kd", KDS)
@pytest.mark.parametrize("curve", [HILBERT, MORTON])
@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_keys_vs_ref(oracle, ref, kd, curve, T):
    from hipcstone import inputs

    x, y, z, _ = inputs.plummer(20000, 0.1, seed=8, dtype=T)
    for box in (make_box(-1, 1, T), make_box((-1.0, -1.1, -0.95), (1.05, 1.3, 1.0), T)):
        assert np.array_equal(oracle.compute_sfc_keys(x, y, z, box, kd, curve),
                              ref.compute_sfc_keys(x, y, z, box, kd, curve))


@pytest.mark.parametrize("kd", KDS)
def test_tree_steps_vs_ref(oracle, ref, kd):
    """computeOctree, then warm updateOctree steps on displaced particles: identical tree/count sequences."""
    from hipcstone import inputs

    T = np.float64
    x, y, z, _ = inputs.plummer(30000, 0.05, seed=9, dtype=T)
    box = make_box(-1, 1, T)
    keys = np.sort(oracle.compute_sfc_keys(x, y, z, box, kd, HILBERT))
    t1, c1 = oracle.compute_octree(keys, 32)
    t2, c2 = ref.compute_octree(keys, 32)
    assert np.array_equal(t1, t2) and np.array_equal(c1, c2)
    rng = np.random.default_rng(10)
    for step in range(3):
        x = np.clip(x + rng.normal(0, 0.02, x.size), -1, 1)
        keys = np.sort(oracle.compute_sfc_keys(x, y, z, box, kd, HILBERT))
        t1, c1, v1 = oracle.update_octree(keys, 32, t1, c1)
        t2, c2, v2 = ref.update_octree(keys, 32, t2, c2)
        assert np.array_equal(t1, t2) and np.array_equal(c1, c2) and v1 == v2
        assert np.array_equal(oracle.build_octree(t1)["prefixes"], ref.build_octree(t2)["prefixes"])


@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("boundary", [(0, 0, 0), (1, 1, 1), (1, 0, 1)])
def test_neighbors_vs_ref(oracle, ref, T, boundary):
    from hipcstone import inputs

    n = 20000
    x, y, z, r = inputs.plummer(n, 0.2, seed=11, dtype=T)
    h = inputs.plummer_h(r, n, 0.2, 60, dtype=T)
    box = make_box((-1.0, -1.05, -1.0), (1.0, 1.0, 1.1), T, boundary)
    p = cpu_pipeline(oracle, x, y, z, h, box, np.uint64, 32, HILBERT)
    tree = {k: p[k] for k in ("childOffsets", "parents", "internalToLeaf")}
    for ext in (1.0, 1.2):
        nb1, nc1 = oracle.find_neighbors(p["x"], p["y"], p["z"], p["h"], 0, n, box, tree, p["layout"], p["centers"],
                                         p["sizes"], 64, ext)
        nb2, nc2 = ref.find_neighbors(p["x"], p["y"], p["z"], p["h"], 0, n, box, tree, p["layout"], p["centers"],
                                      p["sizes"], 64, ext)
        assert neighbor_rows_equal(nb1, nc1, nb2, nc2, 64)
    c1, s1 = oracle.node_fp_centers(p["prefixes"], box, T)
    c2, s2 = ref.node_fp_centers(p["prefixes"], box, T)
    assert np.array_equal(c1, c2) and np.array_equal(s1, s2)


@pytest.mark.parametrize("kd", KDS)
@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_bbox_and_halos_vs_ref(oracle, ref, kd, T):
    from hipcstone import inputs

    n = 40000
    x, y, z = inputs.gaussian_mixture(n, clusters=8, seed=12, dtype=T)
    h = np.full(n, 0.01, dtype=T)
    box = make_box(0, 1, T)
    p = cpu_pipeline(oracle, x, y, z, h, box, kd, 64, MORTON)
    L = p["counts"].size
    NI = p["prefixes"].size - L
    init = p["centers"][p["leafToInternal"][NI:]]
    first, last = L // 4, L // 2
    b1 = oracle.compute_bounding_box(p["x"], p["y"], p["z"], p["h"], p["layout"], first, last, T(2), init)
    b2 = ref.compute_bounding_box(p["x"], p["y"], p["z"], p["h"], p["layout"], first, last, T(2), init)
    assert np.array_equal(b1[0], b2[0]) and np.array_equal(b1[1], b2[1])
    f1 = oracle.find_halos(p["prefixes"], p["childOffsets"], p["parents"], p["centers"], p["sizes"], p["leaves"],
                           b1[0], b1[1], box, first, last)
    f2 = ref.find_halos(p["prefixes"], p["childOffsets"], p["parents"], p["centers"], p["sizes"], p["leaves"],
                        b2[0], b2[1], box, first, last)
    assert np.array_equal(f1, f2) and f1.sum() > 0
