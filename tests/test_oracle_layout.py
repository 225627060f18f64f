"""Pins the oracle's SFC assignment, halo discovery driver, node layout, marked-range extraction and range gathers.

Known-answer tests come from the reference's unit tests (test/unit/domain/domaindecomp.cpp, layout.cpp,
test/unit_cuda/halos/gather_halos_gpu.cu); random cases are checked against the reference headers compiled into
oracle/_ref and against direct numpy restatements.
"""

import numpy as np
import pytest

from helpers import cpu_pipeline
from oracle_py import MORTON, make_box

KDS = [np.uint32, np.uint64]
UMAX = 0xFFFFFFFF


def _bins(lib, counts, num_bins, kd=np.uint64):
    counts = np.asarray(counts, dtype=np.uint32)
    tree = np.arange(counts.size + 1, dtype=kd)
    return lib.make_sfc_assignment(tree, counts, num_bins)


# ---------------------------------------------------------------- uniformBins / makeSfcAssignment


@pytest.mark.parametrize("kd", KDS)
def test_uniform_bins_kat(oracle, kd):
    # test/unit/domain/domaindecomp.cpp:23-71
    off, _, cnt = _bins(oracle, [UMAX - 10, 5, 5, UMAX - 11, 5, 6], 2, kd)
    assert off.tolist() == [0, 3, 6] and cnt.tolist() == [UMAX, UMAX]
    off, _, _ = _bins(oracle, [5, 5, 5, 15, 1, 0], 2, kd)
    assert off.tolist() == [0, 3, 6]
    _, _, cnt = _bins(oracle, [15, 0, 1, 5, 5, 5], 2, kd)
    assert cnt.min() == 15 and cnt.max() == 16
    _, _, cnt = _bins(oracle, [4, 3, 4, 3, 4, 3, 4, 3, 4, 3], 7, kd)
    assert cnt.min() == 3 and cnt.max() == 7


@pytest.mark.parametrize("kd", KDS)
def test_make_sfc_assignment_kat(oracle, kd):
    # test/unit/domain/domaindecomp.cpp:74-90
    tree = np.array([0, 10, 20, 30, 40], dtype=kd)
    off, bnd, cnt = oracle.make_sfc_assignment(tree, np.array([5, 5, 5, 5], np.uint32), 2)
    assert bnd.tolist() == [0, 20, 40] and off.tolist() == [0, 2, 4] and np.diff(off).tolist() == [2, 2]
    assert cnt.tolist() == [10, 10]


def _uniform_bins_numpy(counts, num_bins):
    """Direct restatement with Python integers: bin i starts at the first leaf whose count scan reaches
    floor(i * (total / num_bins)) (the double quotient, as the reference computes it)."""
    scan = np.concatenate([[0], np.cumsum(counts.astype(np.uint64))]).astype(np.uint64)
    total = int(scan[-1])
    bc = float(total) / num_bins
    off = [0] + [int(np.searchsorted(scan, np.uint64(int(i * bc)), side="left")) for i in range(1, num_bins)]
    off.append(counts.size)
    cnt = [int(scan[off[i + 1]] - scan[off[i]]) & UMAX for i in range(num_bins)]
    return np.array(off, np.int32), np.array(cnt, np.uint32)


@pytest.mark.parametrize("kd", KDS)
def test_make_sfc_assignment_random(oracle, ref, kd):
    rng = np.random.default_rng(31)
    for trial in range(40):
        L = int(rng.integers(1, 3000))
        counts = rng.integers(0, 300, L).astype(np.uint32)
        counts[rng.random(L) < 0.2] = 0
        tree = np.sort(rng.choice(np.iinfo(kd).max // 2, L + 1, replace=False).astype(kd))
        for ranks in (1, 2, 8, int(rng.integers(1, 40)), L + 3):
            o1 = oracle.make_sfc_assignment(tree, counts, ranks)
            o2 = ref.make_sfc_assignment(tree, counts, ranks)
            for a, b in zip(o1, o2):
                assert np.array_equal(a, b), (trial, ranks)
            off, cnt = _uniform_bins_numpy(counts, ranks)
            assert np.array_equal(o1[0], off) and np.array_equal(o1[2], cnt)
            assert np.array_equal(o1[1], tree[off])


# ---------------------------------------------------------------- extractMarkedElements


@pytest.mark.parametrize("kd", KDS)
def test_extract_marked_kat(oracle, kd):
    # test/unit/domain/layout.cpp:61-101
    leaves = np.arange(11, dtype=kd)
    flags = np.array([0, 0, 0, 0, 1, 2, 3, 3, 4, 4, 5], dtype=np.uint32)
    cases = [((0, 0), []), ((0, 3), []), ((0, 4), [3, 4]), ((0, 5), [3, 5]), ((0, 7), [3, 6]),
             ((0, 10), [3, 6, 7, 8, 9, 10]), ((9, 10), [9, 10])]
    for (a, b), want in cases:
        assert oracle.extract_marked_elements(leaves, flags, a, b).ravel().tolist() == want, (a, b)


def _runs_numpy(source, flags, first, second):
    marked = [bool(flags[i + 1] > flags[i]) for i in range(first, second)]
    out, i = [], 0
    while i < len(marked):
        if marked[i]:
            j = i
            while j < len(marked) and marked[j]:
                j += 1
            out.append((source[first + i], source[first + j]))
            i = j
        else:
            i += 1
    return np.array(out, dtype=source.dtype).reshape(-1, 2)


@pytest.mark.parametrize("kd", KDS)
def test_extract_marked_random(oracle, ref, kd):
    rng = np.random.default_rng(32)
    for trial in range(60):
        L = int(rng.integers(1, 500))
        inc = rng.integers(0, 4, L) * (rng.random(L) < rng.random())
        flags = np.concatenate([[0], np.cumsum(inc)]).astype(np.uint32)
        source = np.sort(rng.integers(0, 1 << 30, L + 1)).astype(kd)
        a = int(rng.integers(0, L + 1))
        b = int(rng.integers(a, L + 1))
        p1 = oracle.extract_marked_elements(source, flags, a, b)
        p2 = ref.extract_marked_elements(source, flags, a, b)
        assert np.array_equal(p1, p2), trial
        assert np.array_equal(p1, _runs_numpy(source, flags, a, b)), trial


# ---------------------------------------------------------------- computeNodeLayout


def test_node_layout_random(oracle, ref):
    rng = np.random.default_rng(33)
    for trial in range(30):
        L = int(rng.integers(1, 4000))
        N = L + (L - 1) // 7
        counts = rng.integers(0, 200, L).astype(np.uint32)
        flags = (rng.random(N) < 0.1).astype(np.uint8)
        l2i = rng.permutation(N)[:L].astype(np.int32)
        a = int(rng.integers(0, L + 1))
        b = int(rng.integers(a, L + 1))
        l1 = oracle.compute_node_layout(counts, flags, l2i, a, b)
        l2 = ref.compute_node_layout(counts, flags, l2i, a, b)
        assert np.array_equal(l1, l2), trial
        i = np.arange(L)
        keep = ((i >= a) & (i < b)) | (flags[l2i] != 0)
        want = np.concatenate([[0], np.cumsum(np.where(keep, counts, 0))]).astype(np.uint32)
        assert np.array_equal(l1, want)


# ---------------------------------------------------------------- range gathers


def test_gather_ranges_kat(oracle):
    # test/unit_cuda/halos/gather_halos_gpu.cu:30-51
    src = np.arange(30, dtype=np.int32)
    out = oracle.gather_ranges([0, 4, 7], [4, 12, 22], src, 10)
    assert out.tolist() == [4, 5, 6, 7, 12, 13, 14, 22, 23, 24]


def test_gather_ranges_random(oracle):
    rng = np.random.default_rng(34)
    for trial in range(20):
        R = int(rng.integers(1, 300))
        lens = rng.integers(0, 50, R)
        scan = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint32)
        total = int(lens.sum())
        src = rng.random((2000, 3)).astype(np.float32)
        offs = rng.integers(0, 2000 - 50, R).astype(np.uint32)
        out = oracle.gather_ranges(scan, offs, src, total)
        want = np.concatenate([src[o:o + n] for o, n in zip(offs, lens)] + [src[:0]])
        assert np.array_equal(out, want)


def test_gather_leaf_ranges(oracle):
    rng = np.random.default_rng(35)
    L = 500
    counts = rng.integers(0, 40, L).astype(np.uint32)
    src_layout = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    keep = rng.random(L) < 0.3
    dst_layout = np.concatenate([[0], np.cumsum(np.where(keep, counts, 0))]).astype(np.uint32)
    src = rng.random(int(src_layout[-1]))
    out = oracle.gather_leaf_ranges(src_layout, dst_layout, src)
    want = np.concatenate([src[src_layout[i]:src_layout[i + 1]] for i in range(L) if keep[i]])
    assert np.array_equal(out, want)


# ---------------------------------------------------------------- discoverHalos


@pytest.mark.parametrize("kd", KDS)
@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_discover_halos_vs_ref_parts(oracle, ref, kd, T):
    """The oracle's discoverHalos equals the reference's computeBoundingBox + findHalos composed as
    octree_focus_mpi.hpp:538-567 composes them, on rank 2 of an 8-way assignment."""
    from hipcstone import inputs

    n = 60000
    x, y, z = inputs.gaussian_mixture(n, clusters=16, seed=13, dtype=T)
    h = np.full(n, 0.006, dtype=T)
    box = make_box(0, 1, T)
    p = cpu_pipeline(oracle, x, y, z, h, box, kd, 32, MORTON)
    off, _, _ = oracle.make_sfc_assignment(p["leaves"], p["counts"], 8)
    first, last = int(off[2]), int(off[3])
    tree = {k: p[k] for k in ("prefixes", "childOffsets", "parents", "leafToInternal")}
    sc, ss, flags = oracle.discover_halos(tree, p["centers"], p["sizes"], p["leaves"], p["x"], p["y"], p["z"], p["h"],
                                          p["layout"], first, last, box)
    L = p["counts"].size
    NI = p["prefixes"].size - L
    seed = p["centers"][p["leafToInternal"][NI:]]
    rc, rs = ref.compute_bounding_box(p["x"], p["y"], p["z"], p["h"], p["layout"], first, last, T(2), seed)
    rf = ref.find_halos(p["prefixes"], p["childOffsets"], p["parents"], p["centers"], p["sizes"], p["leaves"], rc, rs,
                        box, first, last)
    assert np.array_equal(sc[first:last], rc[first:last]) and np.array_equal(ss[first:last], rs[first:last])
    assert np.array_equal(flags, rf) and flags.sum() > 0
    # accumulate keeps earlier flags
    marked = np.zeros_like(flags)
    marked[::97] = 1
    _, _, acc = oracle.discover_halos(tree, p["centers"], p["sizes"], p["leaves"], p["x"], p["y"], p["z"], p["h"],
                                      p["layout"], first, last, box, accumulate_flags=marked)
    assert np.array_equal(acc, flags | marked)
    # the halo-gather chain: kept layout, then packed arrays
    local = ref.compute_node_layout(p["counts"], flags, p["leafToInternal"][NI:], first, last)
    assert np.array_equal(local, oracle.compute_node_layout(p["counts"], flags, p["leafToInternal"][NI:], first, last))
    xs = oracle.gather_leaf_ranges(p["layout"], local, p["x"])
    lo, hi = int(p["layout"][first]), int(p["layout"][last])
    a0 = int(local[first])
    assert np.array_equal(xs[a0:a0 + hi - lo], p["x"][lo:hi])
    assert xs.size > hi - lo  # there are halos
    # the same particles through extractMarkedElements + gatherRanges (the reference's own composition)
    pairs = ref.extract_marked_elements(p["layout"], local, 0, L)
    lens = (pairs[:, 1] - pairs[:, 0]).astype(np.uint32)
    scan = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint32)
    xs2 = oracle.gather_ranges(scan, pairs[:, 0], p["x"], int(lens.sum()))
    assert np.array_equal(xs, xs2)


def test_reference_halo_chain_matches_oracle_chain(oracle, ref):
    """The reference's CPU rank-share chain (used as the halo benchmark's CPU baseline) and the oracle's chain agree
    on the number of local particles."""
    from hipcstone import inputs

    T = np.float64
    n = 200000
    x, y, z = inputs.gaussian_mixture(n, clusters=16, seed=14, dtype=T)
    h = np.full(n, 0.004, dtype=T)
    box = make_box(0, 1, T)
    p = cpu_pipeline(oracle, x, y, z, h, box, np.uint64, 128, MORTON)
    off, _, _ = oracle.make_sfc_assignment(p["leaves"], p["counts"], 8)
    first, last = int(off[2]), int(off[3])
    t = {k: p[k] for k in ("prefixes", "childOffsets", "parents", "leafToInternal")}
    _, _, flags = oracle.discover_halos(t, p["centers"], p["sizes"], p["leaves"], p["x"], p["y"], p["z"], p["h"],
                                        p["layout"], first, last, box)
    L = p["counts"].size
    local = oracle.compute_node_layout(p["counts"], flags, p["leafToInternal"][p["prefixes"].size - L:], first, last)
    xs, ys, zs, hs = (a.copy() for a in (x, y, z, h))
    m, times = ref.halo_chain(xs, ys, zs, hs, box, 128, 8, 2)
    assert m == int(local[-1]) and m > int(p["layout"][last] - p["layout"][first])
    assert np.array_equal(xs, p["x"]) and times.sum() > 0
