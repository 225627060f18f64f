"""Pins the oracle's one-rank domain step: spanning trees, the focus-tree update (mergeCountAndMacOp,
enforceKeys, protectAncestors, injectKeys), focus counts, the global-tree rule and the box rules.

Known answers come from the reference's unit tests (test/unit/sfc/common.cpp, test/unit/focus/octree_focus.cpp);
random cases are checked against the reference's own functions compiled into oracle/_ref (CombinedUpdate::updateFocus,
computeSpanningTree, computeNodeCounts + upsweep, rebalanceDecision / rebalanceTree, limitBoxShrinking). The
one-rank sequence that composes them (oracle_py.CpuDomain) is run here over both libraries, and pinned against the
reference's own Domain in test_oracle_refdomain.py and by the golden fixture domain.npz (test_golden.py).
"""

import ctypes as C

import numpy as np
import pytest

from helpers import OctreeMaker, node_range
from oracle_py import CpuDomain, fitting_box, limit_box_shrinking

KDS = [np.uint32, np.uint64]


def _ks(kd):
    return "u32" if np.dtype(kd) == np.uint32 else "u64"


def _kt(kd):
    return C.c_uint32 if np.dtype(kd) == np.uint32 else C.c_uint64


def pad(kd, v, length):
    """pad(prefix, length): the prefix of `length` bits moved to the front of the key (sfc/common.hpp)."""
    maxl = 10 if np.dtype(kd) == np.uint32 else 21
    return kd(int(v) << (3 * maxl - length))


# ---------------------------------------------------------------- sfc/common.hpp KATs


@pytest.mark.parametrize("kd", KDS)
def test_make_prefix_and_span_kat(oracle, kd):
    mp = getattr(oracle.lib, f"orc_make_prefix_{_ks(kd)}")
    mp.restype = _kt(kd)
    if kd == np.uint32:
        # test/unit/sfc/common.cpp:265-270
        assert mp(_kt(kd)(0)) == 1 and mp(_kt(kd)(1)) == 0o10000000001 and mp(_kt(kd)(0o1000000000)) == 0o11
    span = getattr(oracle.lib, f"orc_span_sfc_range_{_ks(kd)}")

    def run(a, b):
        out = np.zeros(64, dtype=kd)
        n = span(_kt(kd)(int(a)), _kt(kd)(int(b)), out.ctypes.data_as(C.c_void_p))
        return out[:n].tolist()

    # test/unit/sfc/common.cpp:376-419
    assert run(0, 0o7777) == [0, 0o1000, 0o2000, 0o3000, 0o4000, 0o5000, 0o6000, 0o7000, 0o7100, 0o7200, 0o7300,
                              0o7400, 0o7500, 0o7600, 0o7700, 0o7710, 0o7720, 0o7730, 0o7740, 0o7750, 0o7760, 0o7770,
                              0o7771, 0o7772, 0o7773, 0o7774, 0o7775, 0o7776]
    assert run(1, 0o742) == [1, 2, 3, 4, 5, 6, 7, 0o10, 0o20, 0o30, 0o40, 0o50, 0o60, 0o70, 0o100, 0o200, 0o300,
                             0o400, 0o500, 0o600, 0o700, 0o710, 0o720, 0o730, 0o740, 0o741]
    assert run(0o41305, 0o46000) == [0o41305, 0o41306, 0o41307, 0o41310, 0o41320, 0o41330, 0o41340, 0o41350,
                                     0o41360, 0o41370, 0o41400, 0o41500, 0o41600, 0o41700, 0o42000, 0o43000,
                                     0o44000, 0o45000]
    assert len(run(0o40000, 0o50000)) == 1 and len(run(0o40000, 0o60000)) == 2
    assert run(pad(kd, 1, 3), node_range(kd, 0)) == [int(pad(kd, d, 3)) for d in range(1, 8)]
    assert run(0, 0) == [] and run(1, 1) == []


@pytest.mark.parametrize("kd", KDS)
def test_spanning_tree_vs_ref(oracle, ref, kd):
    rng = np.random.default_rng(51)
    top = int(node_range(kd, 0))
    for trial in range(30):
        k = rng.integers(0, top, int(rng.integers(0, 40)), dtype=np.uint64).astype(kd)
        keys = np.unique(np.concatenate([[0], k, [top]]).astype(kd))
        a = oracle.compute_spanning_tree(keys)
        b = ref.compute_spanning_tree(keys)
        assert np.array_equal(a, b)
        assert np.all(np.isin(keys, a)) and np.all(a[1:] > a[:-1])


# ---------------------------------------------------------------- focus decisions (octree_focus.cpp KATs)


def _node_inputs(oracle, kd, tree, leaf_counts, leaf_macs, internal_macs):
    """counts upswept from leaves and macs scattered from leaves plus internal overrides, as the reference test's
    computeNodeOps builds them (test/unit/focus/octree_focus.cpp:25-57)."""
    o = oracle.build_octree(tree)
    N = o["prefixes"].size
    L = tree.size - 1
    NI = N - L
    l2i = o["leafToInternal"][NI:]
    counts = np.zeros(N, dtype=np.uint64)
    counts[l2i] = leaf_counts
    lr, co = o["levelRange"], o["childOffsets"]
    for level in range(lr.size - 2, -1, -1):
        for i in range(lr[level], lr[level + 1]):
            if co[i]:
                counts[i] = counts[co[i]:co[i] + 8].sum()
    macs = np.zeros(N, dtype=np.uint8)
    macs[l2i] = leaf_macs
    for key, val in internal_macs:
        macs[np.nonzero(o["prefixes"] == kd(key))[0][0]] = val
    return o, counts.astype(np.uint32), macs, l2i


def _leaf_ops(oracle, kd, tree, leaf_counts, leaf_macs, internal_macs, fs, fe, bucket):
    o, counts, macs, l2i = _node_inputs(oracle, kd, tree, leaf_counts, leaf_macs, internal_macs)
    N = o["prefixes"].size
    ops = np.zeros(N, dtype=np.int32)
    f = getattr(oracle.lib, f"orc_focus_node_ops_{_ks(kd)}")
    conv = f(o["prefixes"].ctypes.data_as(C.c_void_p), o["childOffsets"].ctypes.data_as(C.c_void_p),
             o["parents"].ctypes.data_as(C.c_void_p), counts.ctypes.data_as(C.c_void_p),
             macs.ctypes.data_as(C.c_void_p), C.c_int(N), _kt(kd)(int(fs)), _kt(kd)(int(fe)), C.c_uint(bucket),
             ops.ctypes.data_as(C.c_void_p))
    return ops[l2i].tolist(), bool(conv)


@pytest.mark.parametrize("kd", KDS)
def test_focus_rebalance_decision_kat(oracle, kd):
    # test/unit/focus/octree_focus.cpp:62-186
    t = OctreeMaker(kd).divide().divide(0).divide(7).make_tree()
    im = [(1, 1), (0o10, 1), (0o17, 1)]
    ops, conv = _leaf_ops(oracle, kd, t, [1, 1, 1, 2, 1, 1, 1, 1, 1, 1, 2, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0],
                          [0, 0, 1, 0, 1, 1, 1, 1, 1, 0, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0], im, t[0], t[8], 1)
    assert ops == [1, 1, 1, 8, 1, 1, 1, 1, 1, 1, 8, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0] and not conv
    ops, conv = _leaf_ops(oracle, kd, t, [1] * 14 + [0, 0, 2, 1, 0, 0, 0, 0],
                          [0, 0, 1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0], im, t[0], t[8], 1)
    assert ops == [1] * 22 and conv
    ops, conv = _leaf_ops(oracle, kd, t, [1] * 14 + [0, 0, 2, 1, 0, 0, 0, 0],
                          [0, 0, 1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0],
                          [(1, 1), (0o10, 1), (0o17, 0)], t[0], t[8], 1)
    assert ops == [1] * 15 + [0] * 7 and not conv
    t = OctreeMaker(kd).divide().divide(0).divide(1).make_tree()
    ops, conv = _leaf_ops(oracle, kd, t, [1, 2, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 1, 1, 2, 1, 2, 1, 1, 2, 1, 1],
                          [0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0],
                          [(1, 1), (0o10, 1), (0o11, 0)], t[2], t[10], 1)
    assert ops == [1, 8] + [1] * 14 + [8] + [1] * 5 and not conv
    t = OctreeMaker(kd).divide().divide(6).divide(7).make_tree()
    ops, conv = _leaf_ops(oracle, kd, t, [1] * 22, [1] * 6 + [0] * 16, [(1, 1), (0o16, 0), (0o17, 0)], t[14], t[22],
                          1)
    assert ops == [1] * 7 + [0] * 7 + [1] * 8 and not conv


def test_key_enforcement_kat(oracle):
    # test/unit/focus/octree_focus.cpp:228-293 (32-bit keys)
    kd = np.uint32
    t = OctreeMaker(kd).divide().divide(1).make_tree()
    o = oracle.build_octree(t)
    N = o["prefixes"].size
    enf = getattr(oracle.lib, "orc_enforce_key_single_u32")
    prot = getattr(oracle.lib, "orc_protect_ancestors_u32")
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    # decodePlaceholderBit (sfc/common.hpp:206-213) for 32-bit keys
    dec = lambda v: C.c_uint32((int(v) ^ (1 << (int(v).bit_length() - 1))) << (30 - int(v).bit_length() + 1))  # noqa

    def run(key_placeholder, ops0, protect=False):
        ops = np.array(ops0, dtype=np.int32)
        st = enf(dec(key_placeholder), ptr(o["prefixes"]), ptr(o["childOffsets"]), ptr(o["parents"]), ptr(ops))
        if protect:
            prot(ptr(o["prefixes"]), ptr(o["parents"]), C.c_int(N), ptr(ops))
        return st, ops.tolist()

    base = [1, 1] + [0] * 15
    assert run(0o111, base) == (1, [1] * 17)
    assert run(0o1112, base) == (2, [1] * 10 + [8] + [1] * 6)
    assert run(0o101, base, protect=True) == (2, [1, 8] + [1] * 8 + [0] * 7)
    ops = np.array([1] * 10 + [0] * 7, dtype=np.int32)
    st1 = enf(dec(0o101), ptr(o["prefixes"]), ptr(o["childOffsets"]), ptr(o["parents"]), ptr(ops))
    st2 = enf(dec(0o1011), ptr(o["prefixes"]), ptr(o["childOffsets"]), ptr(o["parents"]), ptr(ops))
    assert (st1, st2) == (2, 3) and ops.tolist() == [1, 8] + [1] * 8 + [0] * 7


# ---------------------------------------------------------------- focus update and counts against the reference


def _random_tree(oracle, kd, n, bucket, seed):
    from hipcstone import inputs

    x, y, z, _ = inputs.plummer(n, 0.1, seed=seed, dtype=np.float64)
    from oracle_py import HILBERT, make_box

    keys = np.sort(oracle.compute_sfc_keys(x, y, z, make_box(-1, 1, np.float64), kd, HILBERT))
    tree, counts = oracle.compute_octree(keys, bucket)
    return keys, tree, counts


def _monotone_macs(parents, N, rng):
    """Random MAC flags with the monotonicity of a real MAC (a node that passes has passing descendants):
    flag[child] may be 1 only if flag[parent] is 1. Nodes are level-major, so parents come first."""
    macs = np.zeros(N, dtype=np.uint8)
    macs[0] = 1
    draw = rng.random(N) < 0.7
    for i in range(1, N):
        macs[i] = macs[parents[(i - 1) // 8]] & draw[i]
    return macs


@pytest.mark.parametrize("kd", KDS)
def test_update_focus_and_counts_vs_ref(oracle, ref, kd):
    top = int(node_range(kd, 0))
    for seed, bf, bg in ((61, 8, 64), (62, 16, 16), (63, 32, 256)):
        keys, gtree, _ = _random_tree(oracle, kd, 40000, bg, seed)
        # focus tree: start from a tree of another resolution, so splits, merges and key injection all occur
        _, leaves, _ = _random_tree(oracle, kd, 20000, 4 * bf, seed + 100)
        for step in range(4):
            lc1, nc1 = oracle.focus_counts(leaves, keys)
            lc2, nc2 = ref.focus_counts(leaves, keys)
            assert np.array_equal(lc1, lc2) and np.array_equal(nc1, nc2)
            o = oracle.build_octree(leaves)
            macs = _monotone_macs(o["parents"], nc1.size, np.random.default_rng(seed + step))
            for fs, fe in ((0, top), (int(gtree[gtree.size // 3]), int(gtree[2 * gtree.size // 3]))):
                a, ca = oracle.update_focus(leaves, bf, fs, fe, gtree, nc1, macs)
                b, cb = ref.update_focus(leaves, bf, fs, fe, gtree, nc1, macs)
                assert np.array_equal(a, b) and ca == cb, (seed, step, fs, fe)
            leaves, _ = oracle.update_focus(leaves, bf, 0, top, gtree, nc1)


@pytest.mark.parametrize("kd", KDS)
def test_domain_global_rule_vs_ref(oracle, ref, kd):
    keys, _, _ = _random_tree(oracle, kd, 60000, 64, 71)
    top = int(node_range(kd, 0))
    for bucket in (16, 64, 200):
        t1 = t2 = np.array([0, top], dtype=kd)
        c1 = c2 = np.array([bucket - 1], dtype=np.uint32)
        for first in (True, False, False):
            t1, c1, s1 = oracle.domain_global_update(keys, bucket, t1, c1, first)
            t2, c2, s2 = ref.domain_global_update(keys, bucket, t2, c2, first)
            assert np.array_equal(t1, t2) and np.array_equal(c1, c2) and s1 == s2


def test_box_rules_vs_ref(ref):
    from oracle_py import _ptr

    rng = np.random.default_rng(81)
    for T in (np.float32, np.float64):
        for trial in range(50):
            prev_lo = rng.uniform(-2, 0, 3).astype(T)
            prev_hi = (prev_lo + rng.uniform(0.1, 3, 3)).astype(T)
            fit_lo = (prev_lo + rng.uniform(-0.2, 0.2, 3)).astype(T)
            fit_hi = (prev_hi + rng.uniform(-0.2, 0.2, 3)).astype(T)
            lo, hi = limit_box_shrinking(list(fit_lo), list(fit_hi), list(prev_lo), list(prev_hi))
            out = np.zeros(6, dtype=T)
            fit = np.array([fit_lo[0], fit_hi[0], fit_lo[1], fit_hi[1], fit_lo[2], fit_hi[2]], dtype=T)
            prev = np.array([prev_lo[0], prev_hi[0], prev_lo[1], prev_hi[1], prev_lo[2], prev_hi[2]], dtype=T)
            ref._f(f"limit_box_shrinking_{'f32' if T == np.float32 else 'f64'}")(
                _ptr(fit), _ptr(prev), _ptr(np.zeros(3, np.int32)), _ptr(out))
            assert np.array_equal(np.array(lo, T), out[0::2]) and np.array_equal(np.array(hi, T), out[1::2])
    # fitting box: periodic / fixed dimensions keep their limits, cubic_open widens to the largest side
    x = np.array([0.1, 0.5, 0.3])
    y = np.array([0.2, 0.9, 0.4])
    z = np.array([0.0, 0.2, 0.1])
    lo, hi = fitting_box(x, y, z, [-1, -1, -1], [1, 1, 1], (1, 0, 3))
    assert lo == [-1, 0.2, 0.0] and hi == [1, 0.9, 2.0]  # the largest side is the periodic x side (2)
    lo, hi = fitting_box(x, y, z, [-1, -1, -1], [1, 1, 1], (0, 0, 3))
    assert lo == [0.1, 0.2, 0.0] and hi == [0.5, 0.9, 0.7]


# ---------------------------------------------------------------- the one-rank domain sequence


@pytest.mark.parametrize("kd", KDS)
@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_domain_sequence_oracle_vs_ref(oracle, ref, kd, T):
    from hipcstone import inputs

    n = 60000
    x, y, z, r = inputs.plummer(n, 0.1, seed=91, dtype=T)
    h = inputs.plummer_h(r, n, 0.1, 60, dtype=T)
    for boundary, lo, hi in (((0, 0, 0), -1, 1), ((1, 0, 2), -1.0, 1.0)):
        do = CpuDomain(oracle, kd, T, 64, 16, lo, hi, boundary)
        dr = CpuDomain(ref, kd, T, 64, 16, lo, hi, boundary)
        xx, yy, zz, hh = x.copy(), y.copy(), z.copy(), h.copy()
        rng = np.random.default_rng(92)
        for step in range(4):
            a = do.sync(xx, yy, zz, hh)
            b = dr.sync(xx, yy, zz, hh)
            for k in ("keys", "perm", "gtree", "gcounts", "fleaves", "fcounts", "fnodecounts"):
                assert np.array_equal(a[k], b[k]), (boundary, step, k)
            assert a["lo"] == b["lo"] and a["hi"] == b["hi"] and a["fsteps"] == b["fsteps"]
            assert int(a["fcounts"].sum()) == n and int(a["gcounts"].sum()) == n
            if step == 0:  # the first call converges the focus tree (later calls make one update each)
                assert a["fcounts"].max() <= 16 and a["gcounts"].max() <= 64
            # the focus tree holds every global leaf key
            assert np.all(np.isin(a["gtree"], a["fleaves"]))
            xx = np.clip(a["x"] + rng.normal(0, 0.01, n), -1, 1).astype(T)
            yy, zz, hh = a["y"], a["z"], a["h"]
