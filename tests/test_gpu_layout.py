"""GPU parity of the rank-share path: SFC assignment, halo discovery, node layout, marked ranges and the halo
gathers, against the oracle (bit-exact), and the pipeline's rank-share step on Gaussian-mixture inputs."""

import numpy as np
import pytest

from oracle_py import MORTON
from oracle_py import make_box as orc_box

pytestmark = pytest.mark.gpu

KDS = [np.uint32, np.uint64]
UMAX = 0xFFFFFFFF


def dev(hip, a):
    return hip.DeviceArray.from_numpy(np.ascontiguousarray(a))


def gpu_assign(hip, tree, counts, ranks):
    tree = np.ascontiguousarray(tree)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    return hip.make_sfc_assignment(dev(hip, tree), dev(hip, counts), counts.size, ranks)


# ---------------------------------------------------------------- assignment


@pytest.mark.parametrize("kd", KDS)
def test_assignment_kat(hip, oracle, kd):
    # test/unit/domain/domaindecomp.cpp:23-90, including per-leaf counts near 2^32 (64-bit count scan)
    cases = [([UMAX - 10, 5, 5, UMAX - 11, 5, 6], 2), ([5, 5, 5, 15, 1, 0], 2), ([15, 0, 1, 5, 5, 5], 2),
             ([4, 3, 4, 3, 4, 3, 4, 3, 4, 3], 7)]
    for counts, ranks in cases:
        tree = np.arange(len(counts) + 1, dtype=kd)
        g = gpu_assign(hip, tree, counts, ranks)
        o = oracle.make_sfc_assignment(tree, np.array(counts, np.uint32), ranks)
        for a, b in zip(g, o):
            assert np.array_equal(a, b), (counts, ranks)
    g = gpu_assign(hip, np.array([0, 10, 20, 30, 40], kd), [5, 5, 5, 5], 2)
    assert g[1].tolist() == [0, 20, 40] and g[0].tolist() == [0, 2, 4] and g[2].tolist() == [10, 10]


@pytest.mark.parametrize("kd", KDS)
def test_assignment_random(hip, oracle, kd):
    rng = np.random.default_rng(41)
    for L in (1, 7, 4096, 4097, 300001, 2000000):
        counts = rng.integers(0, 300, L).astype(np.uint32)
        counts[rng.random(L) < 0.25] = 0
        tree = np.sort(rng.choice(np.iinfo(kd).max // 2, L + 1, replace=False).astype(kd))
        for ranks in (1, 8, 37, L + 2):
            g = gpu_assign(hip, tree, counts, ranks)
            o = oracle.make_sfc_assignment(tree, counts, ranks)
            for a, b in zip(g, o):
                assert np.array_equal(a, b), (L, ranks)


# ---------------------------------------------------------------- extractMarkedElements


@pytest.mark.parametrize("kd", KDS)
def test_extract_marked_kat(hip, oracle, kd):
    # test/unit/domain/layout.cpp:61-101
    leaves = np.arange(11, dtype=kd)
    flags = np.array([0, 0, 0, 0, 1, 2, 3, 3, 4, 4, 5], dtype=np.uint32)
    cases = [((0, 0), []), ((0, 3), []), ((0, 4), [3, 4]), ((0, 5), [3, 5]), ((0, 7), [3, 6]),
             ((0, 10), [3, 6, 7, 8, 9, 10]), ((9, 10), [9, 10])]
    ds, df = dev(hip, leaves), dev(hip, flags)
    for (a, b), want in cases:
        assert hip.extract_marked_elements(ds, df, a, b).ravel().tolist() == want, (a, b)


@pytest.mark.parametrize("kd", KDS)
def test_extract_marked_random(hip, oracle, kd):
    rng = np.random.default_rng(42)
    for L in (1, 50, 4095, 4096, 70000, 1000003):
        p = rng.random()
        inc = rng.integers(0, 4, L) * (rng.random(L) < p)
        flags = np.concatenate([[0], np.cumsum(inc)]).astype(np.uint32)
        source = np.sort(rng.integers(0, 1 << 40, L + 1)).astype(kd)
        ds, df = dev(hip, source), dev(hip, flags)
        for a, b in ((0, L), (L // 3, L // 3), (L // 5, L - L // 7), (L - 1, L)):
            g = hip.extract_marked_elements(ds, df, a, b)
            o = oracle.extract_marked_elements(source, flags, a, b)
            assert np.array_equal(g, o), (L, a, b)


# ---------------------------------------------------------------- computeNodeLayout


def test_node_layout_random(hip, oracle):
    rng = np.random.default_rng(43)
    for L in (1, 9, 4096, 100001, 1500000):
        N = L + (L - 1) // 7
        counts = rng.integers(0, 200, L).astype(np.uint32)
        flags = (rng.random(N) < 0.05).astype(np.uint8)
        l2i = rng.permutation(N)[:L].astype(np.int32)
        dc, dfl, dl2i = dev(hip, counts), dev(hip, flags), dev(hip, l2i)
        for a, b in ((0, L), (0, 0), (L // 4, L // 2), (L - 1, L)):
            out = hip.DeviceArray(L + 1, np.uint32)
            hip.compute_node_layout(dc, dfl, dl2i, L, a, b, out)
            assert np.array_equal(out.download(), oracle.compute_node_layout(counts, flags, l2i, a, b)), (L, a, b)


# ---------------------------------------------------------------- gathers


def test_gather_ranges_kat(hip):
    # test/unit_cuda/halos/gather_halos_gpu.cu:30-51
    src = dev(hip, np.arange(30, dtype=np.int32))
    buf = hip.DeviceArray(10, np.int32)
    hip.gather_ranges(dev(hip, np.array([0, 4, 7], np.uint32)), dev(hip, np.array([4, 12, 22], np.uint32)), 3, src,
                      buf, 10, 4)
    assert buf.download().tolist() == [4, 5, 6, 7, 12, 13, 14, 22, 23, 24]


@pytest.mark.parametrize("elem", [(np.uint8, 1), (np.uint16, 1), (np.float32, 1), (np.float64, 1), (np.float32, 3),
                                  (np.float32, 4)])
@pytest.mark.parametrize("shape", ["long", "tiny"])
def test_gather_ranges_random(hip, oracle, elem, shape):
    """Long ranges (window search) and many tiny or empty ranges (more than a window per chunk)."""
    dt, width = elem
    rng = np.random.default_rng(44)
    S = 3000000
    src = rng.integers(0, 250, (S, width)).astype(dt)
    if shape == "long":
        R = 400
        lens = rng.integers(0, 20000, R)
    else:
        R = 200000
        lens = rng.integers(0, 3, R)
    offs = rng.integers(0, S - lens.max() - 1, R).astype(np.uint32)
    assert np.all(offs.astype(np.int64) + lens <= S)  # every source range lies inside src
    scan = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint32)
    total = int(lens.sum())
    want = oracle.gather_ranges(scan, offs, src, total)
    buf = hip.DeviceArray(total * width, dt)
    hip.gather_ranges(dev(hip, scan), dev(hip, offs), R, dev(hip, src.ravel()), buf, total,
                      np.dtype(dt).itemsize * width)
    assert np.array_equal(buf.download().reshape(total, width), want)


@pytest.mark.parametrize("dt", [np.float32, np.float64, np.uint32, np.uint64])
def test_gather_leaf_ranges(hip, oracle, dt):
    rng = np.random.default_rng(45)
    L = 200000
    counts = rng.integers(0, 130, L).astype(np.uint32)
    counts[rng.random(L) < 0.01] = 700  # leaves longer than a wave
    src_layout = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    keep = rng.random(L) < 0.3
    keep[L // 3:L // 2] = True
    dst_layout = np.concatenate([[0], np.cumsum(np.where(keep, counts, 0))]).astype(np.uint32)
    n, m = int(src_layout[-1]), int(dst_layout[-1])
    arrays = [rng.integers(0, 1 << 30, n).astype(dt) for _ in range(4)]
    dsl, ddl = dev(hip, src_layout), dev(hip, dst_layout)
    for k in (1, 4):
        outs = [hip.DeviceArray(m, dt) for _ in range(k)]
        hip.gather_leaf_ranges(dsl, ddl, L, [dev(hip, a) for a in arrays[:k]], outs)
        for a, o in zip(arrays[:k], outs):
            assert np.array_equal(o.download(), oracle.gather_leaf_ranges(src_layout, dst_layout, a))


# ---------------------------------------------------------------- the rank-share step on a clustered set


def mixture(hip, n, T, seed=7):
    x, y, z, h = (hip.DeviceArray(n, T) for _ in range(4))
    hip.gaussian_mixture(x, y, z, h, n, seed=seed)
    return x, y, z, h


@pytest.mark.parametrize("kd,T,bucket,ranks,rank", [(np.uint64, np.float64, 128, 8, 2),
                                                    (np.uint64, np.float32, 64, 8, 0),
                                                    (np.uint32, np.float32, 32, 5, 4),
                                                    (np.uint32, np.float64, 16, 3, 1)])
def test_pipeline_halos_vs_oracle(hip, oracle, kd, T, bucket, ranks, rank):
    """keys -> sort -> computeOctree -> link -> centers -> assignment -> discoverHalos -> node layout -> halo gather
    on the GPU; every intermediate against the oracle on the GPU's own sorted particles and tree (the tree itself
    against the oracle's computeOctree)."""
    n = 400000
    d = mixture(hip, n, T)
    gbox = hip.make_box(0.0, 1.0, T)
    obox = orc_box(0.0, 1.0, T)
    pipe = hip.Pipeline(curve=MORTON, key_bits=64 if kd == np.uint64 else 32, real_bits=64 if T == np.float64 else 32,
                        bucket_size=bucket, ngmax=0)
    pipe.sync(*d, n, gbox, full_build=True)
    m = pipe.halos(ranks, rank)
    outs = [hip.DeviceArray(m, T) for _ in range(4)]
    assert pipe.halos(ranks, rank, outs, m) == m
    a = pipe.arrays()
    v = pipe.view()
    L, N = v.num_leaves, v.num_nodes
    tree, counts = oracle.compute_octree(a["keys"], bucket)
    assert np.array_equal(tree, a["leaves"]) and np.array_equal(counts, a["counts"])
    off, bnd, rc = oracle.make_sfc_assignment(a["leaves"], a["counts"], ranks)
    goff = hip.download_ptr(v.tree_offsets, ranks + 1, np.int32)
    assert np.array_equal(goff, off)
    first, last = int(off[rank]), int(off[rank + 1])
    assert (v.assigned_first, v.assigned_last) == (first, last)
    t = {k: a[k] for k in ("prefixes", "childOffsets", "parents", "leafToInternal")}
    sc, ss, flags = oracle.discover_halos(t, a["centers"], a["sizes"], a["leaves"], a["x"], a["y"], a["z"], a["h"],
                                          a["layout"], first, last, obox)
    gflags = hip.download_ptr(v.halo_flags, N, np.uint8)
    gsc = hip.download_ptr(v.search_centers, 3 * L, T).reshape(L, 3)
    gss = hip.download_ptr(v.search_sizes, 3 * L, T).reshape(L, 3)
    assert np.array_equal(gsc[first:last], sc[first:last]) and np.array_equal(gss[first:last], ss[first:last])
    assert np.array_equal(gflags, flags)
    local = oracle.compute_node_layout(a["counts"], flags, a["leafToInternal"][N - L:], first, last)
    assert np.array_equal(hip.download_ptr(v.local_layout, L + 1, np.uint32), local)
    assert m == int(local[-1]) == v.num_local
    for name, o in zip("xyzh", outs):
        assert np.array_equal(o.download(), oracle.gather_leaf_ranges(a["layout"], local, a[name])), name
    # the assigned range arrives whole and in order; halos exist on both sides for an interior rank
    lo, hi = int(a["layout"][first]), int(a["layout"][last])
    assert np.array_equal(outs[0].download()[local[first]:local[first] + hi - lo], a["x"][lo:hi])
    assert m > hi - lo
    # the same packing through extractMarkedElements + gatherRanges on the GPU
    pairs = hip.extract_marked_elements(dev(hip, a["layout"]), dev(hip, local), 0, L)
    lens = (pairs[:, 1] - pairs[:, 0]).astype(np.uint32)
    scan = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint32)
    buf = hip.DeviceArray(m, T)
    hip.gather_ranges(dev(hip, scan), dev(hip, pairs[:, 0].astype(np.uint32)), pairs.shape[0], dev(hip, a["x"]), buf,
                      m, np.dtype(T).itemsize)
    assert np.array_equal(buf.download(), outs[0].download())


def test_pipeline_halos_capacity_and_errors(hip):
    n, T = 50000, np.float32
    d = mixture(hip, n, T)
    pipe = hip.Pipeline(curve=MORTON, key_bits=64, real_bits=32, bucket_size=64, ngmax=0)
    with pytest.raises(hip.HipcstoneError):
        pipe.halos(8, 2)  # no synced tree yet
    pipe.sync(*d, n, hip.make_box(0.0, 1.0, T), full_build=True)
    m = pipe.halos(8, 2)
    small = [hip.DeviceArray(max(1, m - 1), T) for _ in range(4)]
    with pytest.raises(hip.HipcstoneError):
        pipe.halos(8, 2, small, m - 1)
    with pytest.raises(hip.HipcstoneError):
        pipe.halos(8, 8)
    # one rank owns everything: no halos, the local arrays are the sorted arrays
    m1 = pipe.halos(1, 0)
    assert m1 == n


def test_gaussian_mixture_generator(hip):
    """The device generator's uniforms are the host generator's streams: background membership and positions
    agree bit-for-bit with hipcstone.inputs.uniform."""
    from hipcstone import inputs

    n = 100000
    x, y, z, h = mixture(hip, n, np.float64, seed=7)
    xs = x.download()
    bg = inputs.uniform(7, 14, n) < 0.2
    assert np.array_equal(xs[bg], np.minimum(inputs.uniform(7, 20, n)[bg], np.nextafter(1.0, 0.0)))
    hs = h.download()
    assert np.all(hs > 0) and np.all(np.isfinite(hs))
    for a in (xs, y.download(), z.download()):
        assert a.min() >= 0.0 and a.max() < 1.0
