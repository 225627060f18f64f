"""Python binding of libhipcstone (the MI355X-native single-domain octree pipeline) over its C ABI.

The product is the C ABI in include/hipcstone.h, implemented by lib/libhipcstone.so (HIP kernels for gfx950 plus
C++ host orchestration). This module loads that library with ctypes and exposes:

* ``DeviceArray`` — device memory owned from Python (hipMalloc through the library), with numpy upload/download;
* ``make_box`` — the reference's Box<T> value (sfc/box.hpp:95-175) as the C struct;
* one function per C ABI entry (keys, sort, counts, rebalance, update, link, centers, layout, bbox, halos,
  neighbors) taking DeviceArrays;
* ``Pipeline`` — the plain single-domain sync step (keys -> sort -> tree -> link -> centers -> layout ->
  neighbors).

There is no CPU fallback: if the library is missing, importing this module raises.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libhipcstone.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"libhipcstone.so not found at {LIB_PATH}; build it with `make -C cornerstone-octree_amd` "
        "(or __graft_entry__.build())"
    )

_lib = C.CDLL(LIB_PATH)

HILBERT = 1
MORTON = 0
OPEN, PERIODIC, FIXED, CUBIC_OPEN = 0, 1, 2, 3

STATUS = {0: "OK", 1: "INVALID_ARG", 2: "OUT_OF_MEMORY", 3: "HIP_ERROR", 4: "CAPACITY", 5: "DEVICE_FAULT"}
DEVICE_FAULT = 5
FAULT_SORT_LOOKBACK, FAULT_WALK_BUDGET, FAULT_MALFORMED_TREE = 1, 2, 4

_lib.hipcstone_last_error.restype = C.c_char_p
_lib.hipcstone_build_info.restype = C.c_char_p


class HipcstoneError(RuntimeError):
    def __init__(self, status, where):
        msg = _lib.hipcstone_last_error().decode(errors="replace")
        super().__init__(f"{where}: {STATUS.get(status, status)}: {msg}")
        self.status = status


def _check(status, where):
    if status != 0:
        raise HipcstoneError(status, where)


def build_info() -> str:
    return _lib.hipcstone_build_info().decode()


def device_faults(stream=None):
    """Synchronizes the stream, reads and clears the device's fault word; returns its HIPCSTONE_FAULT_* bits (0 when
    no kernel guard fired)."""
    v = C.c_uint32()
    st = _lib.hipcstone_device_faults(_stream(stream), C.byref(v))
    if st not in (0, DEVICE_FAULT):
        _check(st, "device_faults")
    return int(v.value)


def set_neighbor_waves(waves):
    """Tuning knob: waves per workgroup of later neighbor search launches (2, 4 or 8; 0 = default 4)."""
    _check(_lib.hipcstone_set_neighbor_waves(C.c_int32(int(waves))), "set_neighbor_waves")


def get_neighbor_waves():
    return int(_lib.hipcstone_get_neighbor_waves())


def set_sort_lookback_limit(polls):
    """Test hook: polls a radix sort tile makes of an unpublished predecessor before giving up (0 = default)."""
    _check(_lib.hipcstone_set_sort_lookback_limit(C.c_uint32(int(polls))), "set_sort_lookback_limit")


# ---------------------------------------------------------------- boxes


class BoxF32(C.Structure):
    _fields_ = [("lim", C.c_float * 6), ("len", C.c_float * 3), ("ilen", C.c_float * 3), ("bnd", C.c_uint8 * 3)]


class BoxF64(C.Structure):
    _fields_ = [("lim", C.c_double * 6), ("len", C.c_double * 3), ("ilen", C.c_double * 3), ("bnd", C.c_uint8 * 3)]


def make_box(lo, hi, dtype=np.float32, boundary=(OPEN, OPEN, OPEN)):
    """Box(xmin,xmax,ymin,ymax,zmin,zmax,bx,by,bz) as built by the reference constructor (box.hpp:100-122).

    ``lo``/``hi`` are scalars (cube) or 3-sequences.
    """
    lo = np.broadcast_to(np.asarray(lo, dtype=np.float64), (3,))
    hi = np.broadcast_to(np.asarray(hi, dtype=np.float64), (3,))
    b = [int(v) for v in boundary]
    if np.dtype(dtype) == np.float32:
        box = BoxF32()
        _lib.hipcstone_make_box_f32(*(C.c_float(float(np.float32(v))) for v in (lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])),
                                    *b, C.byref(box))
    else:
        box = BoxF64()
        _lib.hipcstone_make_box_f64(*(C.c_double(float(v)) for v in (lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])),
                                    *b, C.byref(box))
    return box


def box_dtype(box):
    return np.float32 if isinstance(box, BoxF32) else np.float64


# ---------------------------------------------------------------- device memory and streams


class DeviceArray:
    """A device allocation (hipMalloc via the library) with a numpy dtype and element count."""

    def __init__(self, count, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        self._ptr = C.c_void_p()
        _check(_lib.hipcstone_device_alloc(C.byref(self._ptr), C.c_size_t(max(self.nbytes, 1))), "device_alloc")

    @property
    def ptr(self):
        return self._ptr.value

    @classmethod
    def from_numpy(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size <= self.count
        _check(_lib.hipcstone_memcpy_h2d(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes),
                                         _stream(stream)), "memcpy_h2d")

    def download(self, count=None, dtype=None, stream=None):
        dt = np.dtype(dtype) if dtype is not None else self.dtype
        n = self.nbytes // dt.itemsize if count is None else int(count)
        out = np.empty(n, dtype=dt)
        _check(_lib.hipcstone_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), C.c_size_t(out.nbytes),
                                         _stream(stream)), "memcpy_d2h")
        return out

    def fill(self, byte=0, stream=None):
        _check(_lib.hipcstone_fill_bytes(C.c_void_p(self.ptr), C.c_uint8(byte), C.c_size_t(self.nbytes),
                                         _stream(stream)), "fill_bytes")

    def free(self):
        if self._ptr and self._ptr.value:
            _lib.hipcstone_device_free(self._ptr)
            self._ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download_ptr(ptr, count, dtype, stream=None):
    dt = np.dtype(dtype)
    out = np.empty(int(count), dtype=dt)
    if out.nbytes:
        _check(_lib.hipcstone_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(out.nbytes),
                                         _stream(stream)), "memcpy_d2h")
    return out


def _stream(s):
    if s is None:
        return C.c_void_p(0)
    return C.c_void_p(s.handle if isinstance(s, Stream) else s)


def _p(a):
    if a is None:
        return C.c_void_p(0)
    if isinstance(a, DeviceArray):
        return C.c_void_p(a.ptr)
    return C.c_void_p(int(a))


class Stream:
    def __init__(self):
        h = C.c_void_p()
        _check(_lib.hipcstone_stream_create(C.byref(h)), "stream_create")
        self.handle = h.value

    def sync(self):
        _check(_lib.hipcstone_stream_sync(C.c_void_p(self.handle)), "stream_sync")

    def __del__(self):
        try:
            _lib.hipcstone_stream_destroy(C.c_void_p(self.handle))
        except Exception:
            pass


class Event:
    def __init__(self):
        h = C.c_void_p()
        _check(_lib.hipcstone_event_create(C.byref(h)), "event_create")
        self.handle = h.value

    def record(self, stream=None):
        _check(_lib.hipcstone_event_record(C.c_void_p(self.handle), _stream(stream)), "event_record")

    def elapsed_ms(self, later):
        ms = C.c_float()
        _check(_lib.hipcstone_event_elapsed_ms(C.c_void_p(self.handle), C.c_void_p(later.handle), C.byref(ms)),
               "event_elapsed")
        return ms.value

    def __del__(self):
        try:
            _lib.hipcstone_event_destroy(C.c_void_p(self.handle))
        except Exception:
            pass


def device_sync():
    _check(_lib.hipcstone_device_sync(), "device_sync")


# ---------------------------------------------------------------- kernels

_KS = {np.dtype(np.uint32): "u32", np.dtype(np.uint64): "u64"}
_RS = {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64"}


def _fn(name):
    return getattr(_lib, name)


def compute_sfc_keys(x, y, z, keys, box, curve=HILBERT, honor_remove_key=False, stream=None):
    """keys[i] = sfc3D(x[i], y[i], z[i], box) (sfc/sfc.hpp:142-179, sfc_gpu.cu:24-33)."""
    cname = "hilbert" if curve == HILBERT else "morton"
    f = _fn(f"hipcstone_compute_sfc_keys_{cname}_{_KS[keys.dtype]}_{_RS[x.dtype]}")
    _check(f(_stream(stream), _p(x), _p(y), _p(z), _p(keys), C.c_size_t(x.count), C.byref(box),
             C.c_int(int(honor_remove_key))), "compute_sfc_keys")


def sort_by_key_workspace_size(n, key_dtype):
    f = _fn(f"hipcstone_sort_by_key_workspace_size_{_KS[np.dtype(key_dtype)]}")
    f.restype = C.c_size_t
    return f(C.c_size_t(n))


def sort_by_key(keys, values, n=None, end_bit=0, stream=None):
    """Stable ascending sort of keys (in place) carrying values (primitives_gpu.h:119-130)."""
    n = keys.count if n is None else n
    ws = DeviceArray(sort_by_key_workspace_size(n, keys.dtype), np.uint8)
    kb = DeviceArray(n, keys.dtype)
    vb = DeviceArray(n, np.uint32) if values is not None else None
    f = _fn(f"hipcstone_sort_by_key_{_KS[keys.dtype]}")
    _check(f(_stream(stream), _p(keys), _p(values), C.c_size_t(n), _p(kb), _p(vb), _p(ws), C.c_size_t(ws.nbytes),
             C.c_int(end_bit)), "sort_by_key")
    device_sync()


def sort_gather(keys, perm, x, y, z, h, xo, yo, zo, ho, stream=None):
    n = keys.count
    f = _fn(f"hipcstone_sort_gather_workspace_size_{_KS[keys.dtype]}")
    f.restype = C.c_size_t
    ws = DeviceArray(f(C.c_size_t(n)), np.uint8)
    kb = DeviceArray(n, keys.dtype)
    vb = DeviceArray(n, np.uint32)
    g = _fn(f"hipcstone_sort_gather_{_KS[keys.dtype]}_{_RS[x.dtype]}")
    _check(g(_stream(stream), _p(keys), _p(perm), C.c_size_t(n), _p(x), _p(y), _p(z), _p(h), _p(xo), _p(yo), _p(zo),
             _p(ho), _p(kb), _p(vb), _p(ws), C.c_size_t(ws.nbytes)), "sort_gather")
    device_sync()


def gather(perm, src, dst, stream=None):
    f = _fn("hipcstone_gather_u32" if src.dtype.itemsize == 4 else "hipcstone_gather_u64")
    _check(f(_stream(stream), _p(perm), C.c_size_t(perm.count), _p(src), _p(dst)), "gather")


def compute_node_counts(tree, counts, num_nodes, keys, num_keys=None, max_count=0xFFFFFFFF, stream=None):
    num_keys = keys.count if num_keys is None else num_keys
    f = _fn(f"hipcstone_compute_node_counts_{_KS[tree.dtype]}")
    _check(f(_stream(stream), _p(tree), _p(counts), C.c_int32(num_nodes), _p(keys), C.c_size_t(num_keys),
             C.c_uint32(max_count), C.c_int(1)), "compute_node_counts")


def compute_node_ops(tree, num_nodes, counts, bucket_size, node_ops, stream=None):
    """Returns (new_num_nodes, converged); node_ops receives the exclusive scan of the ops (n+1 entries)."""
    _lib.hipcstone_node_ops_workspace_size.restype = C.c_size_t
    ws = DeviceArray(_lib.hipcstone_node_ops_workspace_size(C.c_int32(num_nodes)), np.uint8)
    nn = C.c_int32()
    conv = C.c_int()
    f = _fn(f"hipcstone_compute_node_ops_{_KS[tree.dtype]}")
    _check(f(_stream(stream), _p(tree), C.c_int32(num_nodes), _p(counts), C.c_uint32(bucket_size), _p(node_ops),
             _p(ws), C.c_size_t(ws.nbytes), C.byref(nn), C.byref(conv)), "compute_node_ops")
    return nn.value, bool(conv.value)


def rebalance_tree(tree, num_nodes, new_num_nodes, node_ops, new_tree, stream=None):
    f = _fn(f"hipcstone_rebalance_tree_{_KS[tree.dtype]}")
    _check(f(_stream(stream), _p(tree), C.c_int32(num_nodes), C.c_int32(new_num_nodes), _p(node_ops), _p(new_tree)),
           "rebalance_tree")


def update_octree(keys, bucket_size, tree, counts, num_nodes, new_tree, new_counts, capacity, max_count=0xFFFFFFFF,
                  stream=None):
    """One updateOctree step (update_gpu.cuh:43-68). Returns (new_num_nodes, converged)."""
    _lib.hipcstone_update_octree_workspace_size.restype = C.c_size_t
    ws = DeviceArray(_lib.hipcstone_update_octree_workspace_size(C.c_int32(num_nodes)), np.uint8)
    nn = C.c_int32(num_nodes)
    conv = C.c_int()
    f = _fn(f"hipcstone_update_octree_{_KS[keys.dtype]}")
    st = f(_stream(stream), _p(keys), C.c_size_t(keys.count), C.c_uint32(bucket_size), _p(tree), _p(counts),
           C.byref(nn), _p(new_tree), _p(new_counts), C.c_int32(capacity), C.c_uint32(max_count), _p(ws),
           C.c_size_t(ws.nbytes), C.byref(conv))
    _check(st, "update_octree")
    return nn.value, bool(conv.value)


def compute_octree(keys, bucket_size, max_count=0xFFFFFFFF, stream=None):
    """computeOctree (csarray.hpp:429-440) on the device: returns (leaves, counts) as numpy arrays."""
    kd = keys.dtype
    root = np.array([0, 1 << (63 if kd == np.uint64 else 30)], dtype=kd)
    cap = max(1024, 8 * (keys.count // max(bucket_size, 1) + 8))
    trees = [DeviceArray(cap + 1, kd), DeviceArray(cap + 1, kd)]
    cnts = [DeviceArray(cap, np.uint32), DeviceArray(cap, np.uint32)]
    trees[0].upload(root)
    cnts[0].upload(np.array([keys.count], dtype=np.uint32))
    n, cur = 1, 0
    while True:
        nn, conv = update_octree(keys, bucket_size, trees[cur], cnts[cur], n, trees[1 - cur], cnts[1 - cur], cap,
                                 max_count, stream)
        cur, n = 1 - cur, nn
        if conv:
            break
    return trees[cur].download(n + 1), cnts[cur].download(n)


def build_octree(leaves):
    """Link a leaf array (numpy) on the device; returns a dict of numpy arrays (octree.hpp:285-360 layout)."""
    kd = leaves.dtype
    L = leaves.size - 1
    NI = (L - 1) // 7
    N = L + NI
    L_max = 21 if kd == np.uint64 else 10
    d_leaves = DeviceArray.from_numpy(leaves)
    prefixes = DeviceArray(N, kd)
    co = DeviceArray(N, np.int32)
    par = DeviceArray(max(1, (N - 1) // 8), np.int32)
    lr = DeviceArray(L_max + 2, np.int32)
    i2l = DeviceArray(N, np.int32)
    l2i = DeviceArray(N, np.int32)
    fws = _fn(f"hipcstone_build_octree_workspace_size_{_KS[kd]}")
    fws.restype = C.c_size_t
    ws = DeviceArray(fws(C.c_int32(L)), np.uint8)
    lrh = (C.c_int32 * (L_max + 2))()
    f = _fn(f"hipcstone_build_octree_{_KS[kd]}")
    _check(f(None, _p(d_leaves), C.c_int32(L), _p(prefixes), _p(co), _p(par), _p(lr), _p(i2l), _p(l2i), _p(ws),
             C.c_size_t(ws.nbytes), lrh), "build_octree")
    out = dict(prefixes=prefixes.download(), childOffsets=co.download(), parents=par.download(),
               levelRange=lr.download(), internalToLeaf=i2l.download(), leafToInternal=l2i.download())
    assert np.array_equal(out["levelRange"], np.array(list(lrh), dtype=np.int32))
    return out


def compute_geo_centers(prefixes, box):
    """Geometric node centers/sizes (source_center_gpu.h:93-99) of a numpy prefix array; returns (centers, sizes)."""
    T = box_dtype(box)
    N = prefixes.size
    dp = DeviceArray.from_numpy(prefixes)
    c = DeviceArray(3 * N, T)
    s = DeviceArray(3 * N, T)
    f = _fn(f"hipcstone_compute_geo_centers_{_KS[prefixes.dtype]}_{_RS[np.dtype(T)]}")
    _check(f(None, _p(dp), C.c_int32(N), _p(c), _p(s), C.byref(box)), "compute_geo_centers")
    return c.download().reshape(N, 3), s.download().reshape(N, 3)


def leaf_mass_centers(x, y, z, m, leaf_to_internal, layout, centers, stream=None):
    """Leaf mass centers (source_center_gpu.h:66-74, the CPU path's summation order) into the device array
    `centers` (4 values per node) at leaf_to_internal[i] for every leaf i; all arguments are DeviceArrays."""
    T = np.dtype(x.dtype)
    f = _fn(f"hipcstone_leaf_mass_centers_{_RS[T]}")
    _check(f(stream, _p(x), _p(y), _p(z), _p(m), _p(leaf_to_internal), C.c_int32(leaf_to_internal.count),
             _p(layout), _p(centers)), "leaf_mass_centers")


def upsweep_centers(level_range, child_offsets, centers, stream=None):
    """Mass centers of the internal nodes from their children, level by level (source_center_gpu.h:86-90);
    level_range is a host array of num_levels + 2 entries, child_offsets and centers DeviceArrays."""
    lr = np.ascontiguousarray(level_range, dtype=np.int32)
    T = np.dtype(centers.dtype)
    f = _fn(f"hipcstone_upsweep_centers_{_RS[T]}")
    _check(f(stream, C.c_int32(lr.size - 2), lr.ctypes.data_as(C.c_void_p), _p(child_offsets), _p(centers)),
           "upsweep_centers")


def set_mac(prefixes, centers, inv_theta, box, stream=None):
    """centers[i][3] = squared vector-MAC radius where the mass is nonzero (source_center_gpu.h:110-116);
    prefixes and centers are DeviceArrays."""
    T = box_dtype(box)
    f = _fn(f"hipcstone_set_mac_{_KS[np.dtype(prefixes.dtype)]}_{_RS[np.dtype(T)]}")
    _check(f(stream, _p(prefixes), C.c_int32(prefixes.count), _p(centers), C.c_float(inv_theta), C.byref(box)),
           "set_mac")


def geo_mac_spheres(prefixes, centers, inv_theta, box, stream=None):
    """centers[i] = {geometric center, squared min-MAC radius} (source_center_gpu.h:102-108)."""
    T = box_dtype(box)
    f = _fn(f"hipcstone_geo_mac_spheres_{_KS[np.dtype(prefixes.dtype)]}_{_RS[np.dtype(T)]}")
    _check(f(stream, _p(prefixes), C.c_int32(prefixes.count), _p(centers), C.c_float(inv_theta), C.byref(box)),
           "geo_mac_spheres")


def leaf_layout(counts, stream=None):
    """layout = [0, inclusive_scan(counts)] of a device counts array; returns a DeviceArray of L+1."""
    L = counts.count
    _lib.hipcstone_scan_workspace_size.restype = C.c_size_t
    ws = DeviceArray(_lib.hipcstone_scan_workspace_size(C.c_size_t(L)), np.uint8)
    out = DeviceArray(L + 1, np.uint32)
    _check(_lib.hipcstone_leaf_layout(_stream(stream), _p(counts), C.c_int32(L), _p(out), _p(ws),
                                      C.c_size_t(ws.nbytes)), "leaf_layout")
    return out


def compute_bounding_box(x, y, z, h, layout, first, last, scale, centers, sizes, stream=None):
    T = x.dtype
    ctype = C.c_float if T == np.float32 else C.c_double
    f = _fn(f"hipcstone_compute_bounding_box_{_RS[T]}")
    _check(f(_stream(stream), _p(x), _p(y), _p(z), _p(h), _p(layout), C.c_int32(first), C.c_int32(last),
             ctype(scale), _p(centers), _p(sizes)), "compute_bounding_box")


def find_halos_workspace_size(num_nodes, num_local_leaves, real_dtype):
    _lib.hipcstone_find_halos_workspace_size.restype = C.c_size_t
    return int(_lib.hipcstone_find_halos_workspace_size(C.c_int32(num_nodes), C.c_int32(num_local_leaves),
                                                        C.c_int32(np.dtype(real_dtype).itemsize)))


def find_halos(prefixes, child_offsets, parents, centers, sizes, leaves, search_centers, search_sizes, box, first,
               last, num_nodes, flags, workspace=None, workspace_bytes=None, stream=None):
    """findHalos (collisions.hpp:68-94). workspace: a DeviceArray of at least find_halos_workspace_size() bytes, or
    None for a stream-ordered allocation inside the call."""
    f = _fn(f"hipcstone_find_halos_{_KS[prefixes.dtype]}_{_RS[centers.dtype]}")
    wsb = 0 if workspace is None else (workspace.nbytes if workspace_bytes is None else int(workspace_bytes))
    _check(f(_stream(stream), _p(prefixes), _p(child_offsets), _p(parents), _p(centers), _p(sizes), _p(leaves),
             _p(search_centers), _p(search_sizes), C.byref(box), C.c_int32(first), C.c_int32(last),
             C.c_int32(num_nodes), _p(flags), _p(workspace), C.c_size_t(wsb)), "find_halos")


def ijloop_example(x, y, z, h, m, first, last, box, child_offsets, parents, internal_to_leaf, layout, centers, sizes,
                   num_nodes, ngmax, rho, count, dmin, dmax, fx, ext=1.0, total_bodies=None, groups=None,
                   stream=None):
    """The fixed example interaction through the ijloop API (hipcstone_ijloop.hpp; reference
    traversal/ijloop/gpu_alwaystraverse.cuh:174-204): outputs are body-indexed device arrays. groups: optional
    (num_groups, group_start, group_end) device arrays for the subgroup path."""
    T = np.dtype(box_dtype(box))
    if total_bodies is None:
        if not isinstance(x, DeviceArray):
            raise ValueError("ijloop_example: pass total_bodies (the length of x..m) with raw device pointers")
        total_bodies = x.count
    total = total_bodies
    ng, gs, ge = (0, None, None) if groups is None else groups
    f = _fn(f"hipcstone_ijloop_example_{_RS[T]}")
    _check(f(_stream(stream), _p(child_offsets), _p(parents), _p(internal_to_leaf), _p(layout), _p(centers), _p(sizes),
             C.c_int32(num_nodes), C.byref(box), C.c_uint32(first), C.c_uint32(last), _p(x), _p(y), _p(z), _p(h),
             _p(m), C.c_uint32(ngmax), C.c_float(ext), C.c_uint32(total), C.c_uint32(ng), _p(gs), _p(ge), _p(rho),
             _p(count), _p(dmin), _p(dmax), _p(fx)), "ijloop_example")


def _dtype_of(a, default=None):
    if isinstance(a, DeviceArray):
        return np.dtype(a.dtype)
    if default is None:
        raise ValueError("raw device pointer: pass the element dtype")
    return np.dtype(default)


def make_sfc_assignment(leaves, counts, num_leaves, num_ranks, key_dtype=None, stream=None):
    """makeSfcAssignment (domaindecomp.hpp:103-118): equal-count bins over the leaf array. Returns host arrays
    (tree_offsets[num_ranks+1], boundaries[num_ranks+1], rank_counts[num_ranks])."""
    K = _dtype_of(leaves, key_dtype)
    _lib.hipcstone_sfc_assignment_workspace_size.restype = C.c_size_t
    ws = DeviceArray(_lib.hipcstone_sfc_assignment_workspace_size(C.c_int32(num_leaves)), np.uint8)
    off = DeviceArray(num_ranks + 1, np.int32)
    bnd = DeviceArray(num_ranks + 1, K)
    cnt = DeviceArray(num_ranks, np.uint32)
    f = _fn(f"hipcstone_make_sfc_assignment_{_KS[K]}")
    _check(f(_stream(stream), _p(leaves), _p(counts), C.c_int32(num_leaves), C.c_int32(num_ranks), _p(off), _p(bnd),
             _p(cnt), _p(ws), C.c_size_t(ws.nbytes)), "make_sfc_assignment")
    return off.download(), bnd.download(), cnt.download()


def discover_halos(prefixes, child_offsets, parents, leaf_to_internal, centers, sizes, leaves, num_nodes, x, y, z, h,
                   layout, first, last, box, search_centers, search_sizes, flags, search_ext=1.0, accumulate=False,
                   key_dtype=None, workspace=None, stream=None):
    """FocusedOctree::discoverHalos (octree_focus_mpi.hpp:511-568) on the device; leaf_to_internal is the leaf part
    of the map and layout the absolute leaf layout of x..h."""
    K = _dtype_of(prefixes, key_dtype)
    T = np.dtype(box_dtype(box))
    f = _fn(f"hipcstone_discover_halos_{_KS[K]}_{_RS[T]}")
    _check(f(_stream(stream), _p(prefixes), _p(child_offsets), _p(parents), _p(leaf_to_internal), _p(centers),
             _p(sizes), _p(leaves), C.c_int32(num_nodes), _p(x), _p(y), _p(z), _p(h), _p(layout), C.c_int32(first),
             C.c_int32(last), C.byref(box), C.c_float(search_ext), _p(search_centers), _p(search_sizes), _p(flags),
             C.c_int32(int(accumulate)), _p(workspace), C.c_size_t(0 if workspace is None else workspace.nbytes)),
           "discover_halos")


def compute_node_layout(counts, flags, leaf_to_internal, num_leaves, first, last, layout, stream=None):
    """computeNodeLayout (layout.hpp:153-185): layout[num_leaves+1] of the assigned + halo leaves."""
    _lib.hipcstone_scan_workspace_size.restype = C.c_size_t
    ws = DeviceArray(_lib.hipcstone_scan_workspace_size(C.c_size_t(num_leaves)), np.uint8)
    _check(_lib.hipcstone_compute_node_layout(_stream(stream), _p(counts), _p(flags), _p(leaf_to_internal),
                                              C.c_int32(num_leaves), C.c_int32(first), C.c_int32(last), _p(layout),
                                              _p(ws), C.c_size_t(ws.nbytes)), "compute_node_layout")
    return layout


def extract_marked_elements(source, flags, first, second, source_dtype=None, stream=None):
    """extractMarkedElements (layout.hpp:111-141): an (n, 2) host array of (source[a], source[b]) per marked run."""
    V = _dtype_of(source, source_dtype)
    _lib.hipcstone_extract_marked_workspace_size.restype = C.c_size_t
    ws = DeviceArray(_lib.hipcstone_extract_marked_workspace_size(C.c_int32(second - first)), np.uint8)
    cap = max(1, (second - first + 1) // 2 + 1)
    pairs = DeviceArray(2 * cap, V)
    num = DeviceArray(1, np.uint32)
    f = _fn(f"hipcstone_extract_marked_elements_{_KS[V]}")
    _check(f(_stream(stream), _p(source), _p(flags), C.c_int32(first), C.c_int32(second), _p(pairs),
             C.c_uint32(cap), _p(num), _p(ws), C.c_size_t(ws.nbytes)), "extract_marked_elements")
    n = int(num.download()[0])
    return pairs.download()[: 2 * n].reshape(n, 2)


def gather_ranges(range_scan, range_offsets, num_ranges, src, buffer, buffer_size, elem_bytes, stream=None):
    """gatherRanges (gather_halos_gpu.h:24-31): buffer[t] = src[range_offsets[r] + t - range_scan[r]]."""
    _check(_lib.hipcstone_gather_ranges(_stream(stream), _p(range_scan), _p(range_offsets), C.c_int32(num_ranges),
                                        _p(src), _p(buffer), C.c_size_t(buffer_size), C.c_size_t(elem_bytes)),
           "gather_ranges")


def gather_leaf_ranges(src_layout, dst_layout, num_leaves, srcs, dsts, dtype=None, stream=None):
    """Fused halo gather of up to four arrays from the full leaf layout to the local (kept-leaf) layout."""
    srcs = list(srcs) + [None] * (4 - len(srcs))
    dsts = list(dsts) + [None] * (4 - len(dsts))
    T = _dtype_of(next(a for a in srcs if a is not None), dtype)
    name = {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64", np.dtype(np.uint32): "u32",
            np.dtype(np.uint64): "u64"}[T]
    f = _fn(f"hipcstone_gather_leaf_ranges_{name}")
    _check(f(_stream(stream), _p(src_layout), _p(dst_layout), C.c_int32(num_leaves), *[_p(a) for a in srcs],
             *[_p(a) for a in dsts]), "gather_leaf_ranges")


def find_neighbors(x, y, z, h, first, last, box, child_offsets, parents, internal_to_leaf, layout, centers, sizes,
                   num_nodes, ngmax, neighbors, counts, ext=1.0, stream=None, slot_major=False):
    """CPU-semantics radius search on the device (findneighbors.hpp:156-177). slot_major: rows of each group of 64
    bodies stored slot-major (neighbor k of relative body r at ((r // 64) * ngmax + k) * 64 + r % 64)."""
    T = x.dtype if isinstance(x, DeviceArray) else np.dtype(box_dtype(box))
    f = _fn(f"hipcstone_find_neighbors_{'slot_major_' if slot_major else ''}{_RS[np.dtype(T)]}")
    _check(f(_stream(stream), _p(x), _p(y), _p(z), _p(h), C.c_uint32(first), C.c_uint32(last), C.byref(box),
             _p(child_offsets), _p(parents), _p(internal_to_leaf), _p(layout), _p(centers), _p(sizes),
             C.c_int32(num_nodes), C.c_float(ext), C.c_uint32(ngmax), _p(neighbors), _p(counts)), "find_neighbors")


def displace(x, y, z, h, n, scale, seed, step, box, stream=None):
    """Synthetic multi-step motion: x += scale*h*u, u uniform in [-1,1)^3 (counter-based), clamped into the box."""
    if isinstance(box, BoxF32):
        _check(_lib.hipcstone_synth_displace_f32(_stream(stream), _p(x), _p(y), _p(z), _p(h), C.c_size_t(n),
                                                 C.c_float(scale), C.c_uint64(seed), C.c_uint64(step), C.byref(box)),
               "synth_displace")
    else:
        _check(_lib.hipcstone_synth_displace_f64(_stream(stream), _p(x), _p(y), _p(z), _p(h), C.c_size_t(n),
                                                 C.c_double(scale), C.c_uint64(seed), C.c_uint64(step), C.byref(box)),
               "synth_displace")


def gaussian_mixture(x, y, z, h, n, offset=0, seed=7, clusters=64, background=0.2, target_neighbors=100.0,
                     n_total=None, stream=None):
    """Gaussian-mixture particles with density-matched h, generated on the device (hipcstone_synth_gaussian_mixture)."""
    T = _dtype_of(x)
    f = _fn(f"hipcstone_synth_gaussian_mixture_{_RS[T]}")
    _check(f(_stream(stream), _p(x), _p(y), _p(z), _p(h), C.c_size_t(n), C.c_uint64(offset), C.c_uint64(seed),
             C.c_uint32(clusters), C.c_double(background), C.c_double(target_neighbors),
             C.c_double(float(n if n_total is None else n_total))), "synth_gaussian_mixture")


# ---------------------------------------------------------------- pipeline


class PipelineView(C.Structure):
    _fields_ = [
        ("num_particles", C.c_size_t),
        ("num_leaves", C.c_int32),
        ("num_nodes", C.c_int32),
        ("num_internal", C.c_int32),
        ("converged_calls", C.c_int32),
    ] + [(name, C.c_void_p) for name in (
        "keys", "perm", "x", "y", "z", "h", "leaves", "counts", "prefixes", "child_offsets", "parents",
        "internal_to_leaf", "leaf_to_internal", "level_range", "centers", "sizes", "layout")] + [
        ("level_range_host", C.c_int32 * 24),
        ("assigned_first", C.c_int32),
        ("assigned_last", C.c_int32),
        ("num_local", C.c_size_t),
    ] + [(name, C.c_void_p) for name in (
        "tree_offsets", "halo_flags", "search_centers", "search_sizes", "local_layout")]


class Pipeline:
    """Plain single-domain sync step (hipcstone_pipeline_*): keys, sort+gather, tree update, link, centers,
    layout and neighbors."""

    def __init__(self, curve=HILBERT, key_bits=64, real_bits=32, bucket_size=64, ngmax=150):
        self.handle = C.c_void_p()
        self.key_dtype = np.uint64 if key_bits == 64 else np.uint32
        self.real_dtype = np.float32 if real_bits == 32 else np.float64
        self.ngmax = ngmax
        self.max_level = 21 if key_bits == 64 else 10
        _check(_lib.hipcstone_pipeline_create(C.byref(self.handle), C.c_int(curve), C.c_int(key_bits),
                                              C.c_int(real_bits), C.c_uint32(bucket_size), C.c_uint32(ngmax)),
               "pipeline_create")

    def sync(self, x, y, z, h, n, box, full_build=False, neighbors=None, counts=None, stream=None):
        _check(_lib.hipcstone_pipeline_sync(self.handle, _stream(stream), _p(x), _p(y), _p(z), _p(h), C.c_size_t(n),
                                            C.byref(box), C.c_int(int(full_build)), _p(neighbors), _p(counts)),
               "pipeline_sync")

    STAGES = ("keys", "sort_gather", "tree_update", "link", "centers", "layout", "neighbors", "assignment",
              "halo_discovery", "node_layout", "halo_gather")

    def halos(self, num_ranks, rank, outputs=None, capacity=0, search_ext=1.0, stream=None):
        """One rank's share of the synced tree: assignment, discoverHalos, node layout and (with outputs = four
        device arrays) the halo gather. Returns the local particle count."""
        nl = C.c_size_t()
        outs = list(outputs) if outputs is not None else [None] * 4
        _check(_lib.hipcstone_pipeline_halos(self.handle, _stream(stream), C.c_int32(num_ranks), C.c_int32(rank),
                                             C.c_float(search_ext), *[_p(a) for a in outs], C.c_size_t(capacity),
                                             C.byref(nl)), "pipeline_halos")
        return nl.value

    def set_timing(self, enable=True):
        _check(_lib.hipcstone_pipeline_set_timing(self.handle, C.c_int(int(enable))), "pipeline_set_timing")

    def stage_times(self):
        """{stage: (total_ms, launches)} since timing was enabled (synchronizes)."""
        ms = (C.c_double * len(self.STAGES))()
        cnt = (C.c_int32 * len(self.STAGES))()
        _check(_lib.hipcstone_pipeline_stage_times(self.handle, ms, cnt), "pipeline_stage_times")
        return {name: (ms[i], cnt[i]) for i, name in enumerate(self.STAGES)}

    def view(self) -> PipelineView:
        v = PipelineView()
        _check(_lib.hipcstone_pipeline_get_view(self.handle, C.byref(v)), "pipeline_get_view")
        return v

    def arrays(self):
        """Download the pipeline state as numpy arrays."""
        v = self.view()
        n, L, N = v.num_particles, v.num_leaves, v.num_nodes
        kd, T = self.key_dtype, self.real_dtype
        return dict(
            keys=download_ptr(v.keys, n, kd), perm=download_ptr(v.perm, n, np.uint32),
            x=download_ptr(v.x, n, T), y=download_ptr(v.y, n, T), z=download_ptr(v.z, n, T),
            h=download_ptr(v.h, n, T), leaves=download_ptr(v.leaves, L + 1, kd),
            counts=download_ptr(v.counts, L, np.uint32), prefixes=download_ptr(v.prefixes, N, kd),
            childOffsets=download_ptr(v.child_offsets, N, np.int32),
            parents=download_ptr(v.parents, max(1, (N - 1) // 8), np.int32),
            internalToLeaf=download_ptr(v.internal_to_leaf, N, np.int32),
            leafToInternal=download_ptr(v.leaf_to_internal, N, np.int32),
            levelRange=np.array(list(v.level_range_host)[: self.max_level + 2], dtype=np.int32),
            centers=download_ptr(v.centers, 3 * N, T).reshape(N, 3), sizes=download_ptr(v.sizes, 3 * N, T).reshape(N, 3),
            layout=download_ptr(v.layout, L + 1, np.uint32),
        )

    def __del__(self):
        try:
            _lib.hipcstone_pipeline_destroy(self.handle)
        except Exception:
            pass


# ---------------------------------------------------------------- batched clouds


class BatchedView(C.Structure):
    _fields_ = [
        ("num_clouds", C.c_uint32),
        ("num_waves", C.c_uint32),
        ("num_particles", C.c_size_t),
        ("num_leaves", C.c_int32),
        ("num_nodes", C.c_int32),
        ("leaf_capacity", C.c_int32),
    ] + [(name, C.c_void_p) for name in (
        "keys", "perm", "x", "y", "z", "h", "leaves", "counts", "layout", "prefixes", "child_offsets",
        "internal_to_leaf", "centers", "sizes", "cloud_roots", "neighbors", "neighbor_counts")]


CLOUD_SHIFT, LOCAL_SHIFT = 54, 24


class Batched:
    """Many fp32 clouds through the pipeline at once (hipcstone_batched_*): each stage one launch over all clouds,
    warm steps without host synchronization (capturable in one HIP graph). Cloud c is bit-identical to a plain
    32-bit-key pipeline on that cloud alone."""

    def __init__(self, counts, boxes, buckets, curve=MORTON, ngmax=128, leaf_capacity_factor=0.5):
        counts = np.ascontiguousarray(counts, dtype=np.uint64)
        buckets = np.ascontiguousarray(buckets, dtype=np.uint32)
        C_ = len(counts)
        barr = (BoxF32 * C_)(*boxes)
        self.handle = C.c_void_p()
        self.counts = counts
        self.offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        self.ngmax = ngmax
        _check(_lib.hipcstone_batched_create(C.byref(self.handle), C.c_int(curve), C.c_uint32(C_),
                                             counts.ctypes.data_as(C.c_void_p), barr,
                                             buckets.ctypes.data_as(C.c_void_p), C.c_uint32(ngmax),
                                             C.c_double(leaf_capacity_factor)), "batched_create")

    def set_particles(self, x, y, z, h, stream=None):
        """x, y, z, h of all clouds in cloud order: numpy arrays (host) or DeviceArrays."""
        dev = isinstance(x, DeviceArray)
        if dev:
            ptrs = [_p(a) for a in (x, y, z, h)]
        else:
            arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (x, y, z, h)]
            ptrs = [a.ctypes.data_as(C.c_void_p) for a in arrs]
        _check(_lib.hipcstone_batched_set_particles(self.handle, _stream(stream), *ptrs, C.c_int(int(dev))),
               "batched_set_particles")

    def build(self, stream=None):
        _check(_lib.hipcstone_batched_build(self.handle, _stream(stream)), "batched_build")

    def step(self, drift_scale=0.0, seed=0, step=0, stream=None):
        _check(_lib.hipcstone_batched_step(self.handle, _stream(stream), C.c_float(drift_scale), C.c_uint64(seed),
                                           C.c_uint64(step)), "batched_step")

    def capture(self, steps, drift_scale, seed, first_step, stream):
        _check(_lib.hipcstone_batched_capture(self.handle, _stream(stream), C.c_int32(steps), C.c_float(drift_scale),
                                              C.c_uint64(seed), C.c_uint64(first_step)), "batched_capture")

    def replay(self, stream):
        _check(_lib.hipcstone_batched_replay(self.handle, _stream(stream)), "batched_replay")

    def view(self, stream=None) -> BatchedView:
        v = BatchedView()
        _check(_lib.hipcstone_batched_get_view(self.handle, _stream(stream), C.byref(v)), "batched_get_view")
        return v

    def device_bytes(self):
        _lib.hipcstone_batched_device_bytes.restype = C.c_size_t
        return int(_lib.hipcstone_batched_device_bytes(self.handle))

    def arrays(self, with_neighbors=True):
        v = self.view()
        n, L = v.num_particles, v.num_leaves
        out = dict(keys=download_ptr(v.keys, n, np.uint64), perm=download_ptr(v.perm, n, np.uint32),
                   x=download_ptr(v.x, n, np.float32), y=download_ptr(v.y, n, np.float32),
                   z=download_ptr(v.z, n, np.float32), h=download_ptr(v.h, n, np.float32),
                   leaves=download_ptr(v.leaves, L + 1, np.uint64), counts=download_ptr(v.counts, L, np.uint32),
                   layout=download_ptr(v.layout, L + 1, np.uint32),
                   roots=download_ptr(v.cloud_roots, v.num_clouds, np.int32),
                   counts_nb=download_ptr(v.neighbor_counts, n, np.uint32))
        if with_neighbors and self.ngmax:
            out["neighbors"] = download_ptr(v.neighbors, n * self.ngmax, np.uint32).reshape(n, self.ngmax)
        return out

    def cloud(self, a, c):
        """Cloud c's part of arrays() in the cloud's own terms (what a plain pipeline on the cloud alone returns):
        32-bit keys and leaves, local permutation, layout and neighbor indices."""
        lo, hi = int(self.offsets[c]), int(self.offsets[c + 1])
        sel = (a["leaves"][:-1] >> np.uint64(CLOUD_SHIFT)) == np.uint64(c)
        li = np.nonzero(sel)[0]
        leaves = ((a["leaves"][li] >> np.uint64(LOCAL_SHIFT)) & np.uint64((1 << 30) - 1)).astype(np.uint32)
        leaves = np.concatenate([leaves, [np.uint32(1 << 30)]]).astype(np.uint32)
        out = dict(keys=((a["keys"][lo:hi] >> np.uint64(LOCAL_SHIFT)) & np.uint64((1 << 30) - 1)).astype(np.uint32),
                   perm=(a["perm"][lo:hi].astype(np.int64) - lo).astype(np.uint32),
                   x=a["x"][lo:hi], y=a["y"][lo:hi], z=a["z"][lo:hi], h=a["h"][lo:hi], leaves=leaves,
                   counts=a["counts"][li],
                   layout=np.concatenate([a["layout"][li], [a["layout"][li[-1] + 1]]]).astype(np.int64) - lo,
                   neighbor_counts=a["counts_nb"][lo:hi])
        if "neighbors" in a:
            out["neighbors"] = a["neighbors"][lo:hi].astype(np.int64) - lo
        return out

    def __del__(self):
        try:
            _lib.hipcstone_batched_destroy(self.handle)
        except Exception:
            pass


# ---------------------------------------------------------------- one-rank domain


class DomainView(C.Structure):
    _fields_ = [
        ("num_particles", C.c_size_t),
        ("start_index", C.c_size_t),
        ("end_index", C.c_size_t),
        ("num_global_leaves", C.c_int32),
        ("num_leaves", C.c_int32),
        ("num_nodes", C.c_int32),
        ("num_internal", C.c_int32),
        ("global_steps", C.c_int32),
        ("focus_steps", C.c_int32),
        ("focus_converged", C.c_int32),
        ("box_lo", C.c_double * 3),
        ("box_hi", C.c_double * 3),
    ] + [(name, C.c_void_p) for name in (
        "keys", "global_leaves", "global_counts", "leaves", "counts", "node_counts", "prefixes", "child_offsets",
        "parents", "internal_to_leaf", "leaf_to_internal", "level_range", "centers", "sizes", "layout",
        "expansion_centers")]


class Domain:
    """One-rank Domain::sync (hipcstone_domain_*): fitted box, Hilbert keys, sort, global leaf array, focus tree;
    sync reorders the particle arrays in place into SFC order."""

    STAGES = ("box", "keys", "sort_gather", "global_tree", "focus_update", "focus_counts", "finish",
              "expansion_centers")

    def __init__(self, box, key_bits=64, bucket_size=64, bucket_size_focus=16, theta=0.5):
        self.handle = C.c_void_p()
        self.key_dtype = np.uint64 if key_bits == 64 else np.uint32
        self.real_dtype = np.dtype(box_dtype(box))
        self.max_level = 21 if key_bits == 64 else 10
        real_bits = 32 if self.real_dtype == np.float32 else 64
        _check(_lib.hipcstone_domain_create(C.byref(self.handle), C.c_int(key_bits), C.c_int(real_bits),
                                            C.c_uint32(bucket_size), C.c_uint32(bucket_size_focus), C.c_float(theta),
                                            C.byref(box)), "domain_create")

    def sync(self, x, y, z, h, n=None, keys_out=None, stream=None):
        n = x.count if n is None else n
        _check(_lib.hipcstone_domain_sync(self.handle, _stream(stream), _p(x), _p(y), _p(z), _p(h), C.c_size_t(n),
                                          _p(keys_out)), "domain_sync")

    def sync_grav(self, x, y, z, h, m, n=None, keys_out=None, stream=None):
        """Domain::syncGrav: sync, m reordered with x, y, z, h, and the focus tree's expansion centers."""
        n = x.count if n is None else n
        _check(_lib.hipcstone_domain_sync_grav(self.handle, _stream(stream), _p(x), _p(y), _p(z), _p(h), _p(m),
                                               C.c_size_t(n), _p(keys_out)), "domain_sync_grav")

    def update_expansion_centers(self, x, y, z, m, stream=None):
        """Domain::updateExpansionCenters for arrays in the order of the last sync."""
        _check(_lib.hipcstone_domain_update_expansion_centers(self.handle, _stream(stream), _p(x), _p(y), _p(z),
                                                              _p(m)), "domain_update_expansion_centers")

    def expansion_centers(self):
        """(num_nodes, 4) mass center x, y, z and squared MAC radius per focus node, or None."""
        v = self.view()
        if not v.expansion_centers:
            return None
        return download_ptr(v.expansion_centers, 4 * v.num_nodes, self.real_dtype).reshape(v.num_nodes, 4)

    def view(self) -> DomainView:
        v = DomainView()
        _check(_lib.hipcstone_domain_get_view(self.handle, C.byref(v)), "domain_get_view")
        return v

    def gather_properties(self, arrays, stream=None):
        """Reorder device arrays (in the order the last sync received x, y, z, h) into that sync's SFC order."""
        arrays = list(arrays)
        ptrs = (C.c_void_p * max(1, len(arrays)))(*[a.ptr for a in arrays])
        sizes = (C.c_size_t * max(1, len(arrays)))(*[np.dtype(a.dtype).itemsize for a in arrays])
        _check(_lib.hipcstone_domain_gather_properties(self.handle, _stream(stream), ptrs, sizes, C.c_int(len(arrays))),
               "domain_gather_properties")

    def set_timing(self, enable=True):
        _check(_lib.hipcstone_domain_set_timing(self.handle, C.c_int(int(enable))), "domain_set_timing")

    def stage_times(self):
        ms = (C.c_double * len(self.STAGES))()
        cnt = (C.c_int32 * len(self.STAGES))()
        _check(_lib.hipcstone_domain_stage_times(self.handle, ms, cnt), "domain_stage_times")
        return {name: (ms[i], cnt[i]) for i, name in enumerate(self.STAGES)}

    def arrays(self):
        v = self.view()
        n, G, L, N = v.num_particles, v.num_global_leaves, v.num_leaves, v.num_nodes
        kd, T = self.key_dtype, self.real_dtype
        return dict(
            keys=download_ptr(v.keys, n, kd), global_leaves=download_ptr(v.global_leaves, G + 1, kd),
            global_counts=download_ptr(v.global_counts, G, np.uint32), leaves=download_ptr(v.leaves, L + 1, kd),
            counts=download_ptr(v.counts, L, np.uint32), node_counts=download_ptr(v.node_counts, N, np.uint32),
            prefixes=download_ptr(v.prefixes, N, kd), childOffsets=download_ptr(v.child_offsets, N, np.int32),
            parents=download_ptr(v.parents, max(1, (N - 1) // 8), np.int32),
            internalToLeaf=download_ptr(v.internal_to_leaf, N, np.int32),
            leafToInternal=download_ptr(v.leaf_to_internal, N, np.int32),
            levelRange=download_ptr(v.level_range, self.max_level + 2, np.int32),
            centers=download_ptr(v.centers, 3 * N, T).reshape(N, 3), sizes=download_ptr(v.sizes, 3 * N, T).reshape(N, 3),
            layout=download_ptr(v.layout, L + 1, np.uint32),
            box_lo=list(v.box_lo), box_hi=list(v.box_hi),
        )

    def __del__(self):
        try:
            _lib.hipcstone_domain_destroy(self.handle)
        except Exception:
            pass


# ---------------------------------------------------------------- served clouds

SERVE_BUILD, SERVE_UPDATE = 0, 1
SERVE_BUILT, SERVE_UPDATED, SERVE_REBUILT, SERVE_REJECTED = 0, 1, 2, 3


class ServeRequest(C.Structure):
    _fields_ = [("id", C.c_uint64), ("cloud", C.c_uint32), ("kind", C.c_uint32), ("n", C.c_uint64),
                ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("h", C.c_void_p),
                ("arrival_s", C.c_double), ("seed", C.c_uint64)]


class ServeResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_uint32), ("slot", C.c_uint32), ("arrival_s", C.c_double),
                ("admit_s", C.c_double), ("start_s", C.c_double), ("end_s", C.c_double), ("digest", C.c_uint64),
                ("num_leaves", C.c_int32), ("pad", C.c_uint32), ("resident_bytes", C.c_uint64)]


class ServeConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("slots", C.c_int32), ("bucket_size", C.c_uint32), ("ngmax", C.c_uint32),
                ("curve", C.c_int32), ("sigma", C.c_float), ("box_lo", C.c_float), ("box_hi", C.c_float),
                ("realtime", C.c_int32), ("digests", C.c_int32)]


class ServeStats(C.Structure):
    _fields_ = [("makespan_s", C.c_double), ("evictions", C.c_uint64)]


def serve_run(requests, hbm_budget=200 << 30, slots=4, bucket_size=64, ngmax=128, curve=HILBERT, sigma=5e-4,
              box=(-1.0, 1.0), realtime=True, digests=False):
    """Runs requests (dicts: id, cloud, kind, n, x, y, z, h (float32 host arrays, kept alive by the caller),
    arrival_s, seed) through hipcstone_serve_run; returns (list of result dicts, stats dict)."""
    arr = (ServeRequest * max(1, len(requests)))()
    for i, r in enumerate(requests):
        q = arr[i]
        q.id, q.cloud, q.kind, q.n = int(r["id"]), int(r["cloud"]), int(r["kind"]), int(r["n"])
        for k in "xyzh":
            a = r.get(k)
            setattr(q, k, a.ctypes.data if a is not None else None)
        q.arrival_s, q.seed = float(r["arrival_s"]), int(r.get("seed", 0))
    res = (ServeResult * max(1, len(requests)))()
    cfg = ServeConfig(int(hbm_budget), int(slots), int(bucket_size), int(ngmax), int(curve), float(sigma),
                      float(box[0]), float(box[1]), int(bool(realtime)), int(bool(digests)))
    stats = ServeStats()
    _check(_lib.hipcstone_serve_run(C.byref(cfg), arr, C.c_size_t(len(requests)), res, C.byref(stats)), "serve_run")
    out = [{f: getattr(res[i], f) for f, _ in ServeResult._fields_ if f != "pad"} for i in range(len(requests))]
    return out, {"makespan_s": stats.makespan_s, "evictions": int(stats.evictions)}
