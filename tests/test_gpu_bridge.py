"""The reference-side drop-in on the GPU: cstone::Domain<uint64_t, T, execution::Gpu> (the specialization of
cornerstone-octree_amd/bridge/cstone_hipcstone.hpp, driven through its reference-shaped C++ interface by
tests/bridge/domain_driver.cpp) over several sync steps against the oracle's one-rank domain (itself pinned to the
reference's own Domain run): box, keys, particle order, carried properties, focus leaves and layout; then
cstone::findNeighbors(execution::Gpu, ...) over Domain::octreeProperties() against the oracle's CPU search."""

import ctypes as C
import os

import numpy as np
import pytest

from helpers import domain_move, neighbor_rows_equal
from oracle_py import CpuDomain

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "bridge", "libbridge_driver.so")


@pytest.fixture(scope="module")
def bridge():
    if not os.path.exists(DRIVER):
        pytest.skip("the bridge is built only where the reference headers are present (tests/bridge)")
    lib = C.CDLL(DRIVER)
    lib.bridge_last_error.restype = C.c_char_p
    return lib


def dev(hip, a):
    return hip.DeviceArray.from_numpy(np.ascontiguousarray(a))


def _ok(bridge, rc, what):
    assert rc == 0, f"{what}: {bridge.bridge_last_error().decode()}"


@pytest.mark.parametrize("T,bucket,bucket_focus,bnd", [(np.float32, 64, 16, 0), (np.float64, 32, 8, 1)])
def test_domain_through_the_shim(hip, oracle, bridge, T, bucket, bucket_focus, bnd):
    from hipcstone import inputs

    rs = "f32" if T == np.float32 else "f64"
    n = 60000
    x, y, z, r = inputs.plummer(n, 0.1, seed=41, dtype=T)
    h = inputs.plummer_h(r, n, 0.1, 60, dtype=T)
    lim = (C.c_double * 6)(-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)
    dom = C.c_void_p()
    _ok(bridge, getattr(bridge, f"bridge_domain_create_{rs}")(C.byref(dom), C.c_uint(bucket), C.c_uint(bucket_focus),
                                                               C.c_float(0.5), lim, C.c_int(bnd)), "create")
    ref = CpuDomain(oracle, np.uint64, T, bucket, bucket_focus, -1.0, 1.0, (bnd, bnd, bnd))
    rng = np.random.default_rng(42)
    d = [dev(hip, a) for a in (x, y, z, h)]
    ids = np.arange(n, dtype=np.uint32)
    prop = dev(hip, ids)
    keys = hip.DeviceArray(n, np.uint64)
    try:
        for step in range(4):
            print(f"Domain<u64,{rs},Gpu> through the shim: step {step}", flush=True)
            xs, ys, zs, hs = (a.download() for a in d)
            o = ref.sync(xs, ys, zs, hs)
            prop.upload(ids)  # the carried property: each particle's index in this step's input
            _ok(bridge, getattr(bridge, f"bridge_domain_sync_{rs}")(dom, *(C.c_void_p(a.ptr) for a in d), None,
                                                                     C.c_size_t(n), C.c_void_p(keys.ptr),
                                                                     C.c_void_p(prop.ptr), 0), "sync")
            info = (C.c_int64 * 10)()
            box = (C.c_double * 6)()
            _ok(bridge, getattr(bridge, f"bridge_domain_info_{rs}")(dom, info, box), "info")
            assert list(info[:3]) == [0, n, n]
            assert [box[0], box[2], box[4]] == [float(t) for t in o["lo"]]
            assert [box[1], box[3], box[5]] == [float(t) for t in o["hi"]]
            assert np.array_equal(keys.download(), o["keys"])
            for name, a in zip("xyzh", d):
                assert np.array_equal(a.download(), o[name]), name
            assert np.array_equal(prop.download(), ids[o["perm"]])
            L = int(info[3])
            assert L == o["fleaves"].size - 1 and info[9] == L + 1
            leaves = hip.download_ptr(int(info[5]), L + 1, np.uint64)
            layout = hip.download_ptr(int(info[6]), L + 1, np.uint32)
            assert np.array_equal(leaves, o["fleaves"])
            assert np.array_equal(layout, np.concatenate([[0], np.cumsum(o["fcounts"])]).astype(np.uint32))
            for a, m in zip(d, domain_move(rng, o, step)):
                a.upload(np.ascontiguousarray(m))

        # findNeighbors(execution::Gpu, ...) over octreeProperties() of the last sync (run it once more so that the
        # arrays and the tree belong to the same step)
        xs, ys, zs, hs = (a.download() for a in d)
        o = ref.sync(xs, ys, zs, hs)
        _ok(bridge, getattr(bridge, f"bridge_domain_sync_{rs}")(dom, *(C.c_void_p(a.ptr) for a in d), None,
                                                                 C.c_size_t(n), C.c_void_p(keys.ptr),
                                                                 C.c_void_p(prop.ptr), 0), "sync")
        ngmax = 96
        nb = hip.DeviceArray(n * ngmax, np.uint32)
        nc = hip.DeviceArray(n, np.uint32)
        _ok(bridge, getattr(bridge, f"bridge_find_neighbors_{rs}")(dom, *(C.c_void_p(a.ptr) for a in d), C.c_uint32(0),
                                                                    C.c_uint32(n), C.c_uint(ngmax),
                                                                    C.c_void_p(nb.ptr), C.c_void_p(nc.ptr)),
            "findNeighbors")
        hip.device_sync()
        t = oracle.build_octree(o["fleaves"])
        c, s = oracle.node_fp_centers(t["prefixes"], ref.box(), T)
        lay = np.concatenate([[0], np.cumsum(o["fcounts"])]).astype(np.uint32)
        sel = np.arange(0, n, 5, dtype=np.uint32)
        nb_o, nc_o = oracle.find_neighbors_subset(o["x"], o["y"], o["z"], o["h"], sel, ref.box(), t, lay, c, s, ngmax)
        g_nb = nb.download().reshape(n, ngmax)[sel]
        g_nc = nc.download()[sel]
        assert neighbor_rows_equal(g_nb, g_nc, nb_o, nc_o, ngmax)
        assert g_nc.mean() > 20
    finally:
        getattr(bridge, f"bridge_domain_destroy_{rs}")(dom)

