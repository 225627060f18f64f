"""ORACLE — test infrastructure, not product code.

ctypes access to the two CPU checkers:
  * ``ORACLE`` — oracle/liboracle.so, the C restatement of the reference's algorithms (oracle.c);
  * ``REF``    — oracle/_ref/libcstone_ref.so, the reference's own headers compiled from /root/reference by
                 `make -C oracle ref` (absent where the reference is not mounted; then ``REF`` is None).
Both expose the same Python methods, so tests can run one check against either. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libcstone_ref.so")

HILBERT, MORTON = 1, 0


class BoxF32(C.Structure):
    _fields_ = [("lim", C.c_float * 6), ("len", C.c_float * 3), ("ilen", C.c_float * 3), ("bnd", C.c_uint8 * 3)]


class BoxF64(C.Structure):
    _fields_ = [("lim", C.c_double * 6), ("len", C.c_double * 3), ("ilen", C.c_double * 3), ("bnd", C.c_uint8 * 3)]


def make_box(lo, hi, dtype=np.float64, boundary=(0, 0, 0)):
    """Box<T> with lengths and inverse lengths evaluated in T, as sfc/box.hpp:100-122 does."""
    T = np.dtype(dtype).type
    lo = np.broadcast_to(np.asarray(lo, dtype=np.float64), (3,)).astype(T)
    hi = np.broadcast_to(np.asarray(hi, dtype=np.float64), (3,)).astype(T)
    b = BoxF32() if T == np.float32 else BoxF64()
    for d in range(3):
        b.lim[2 * d] = lo[d]
        b.lim[2 * d + 1] = hi[d]
        L = T(hi[d] - lo[d])
        b.len[d] = L
        b.ilen[d] = T(T(1.0) / L)
        b.bnd[d] = int(boundary[d])
    return b


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else C.c_void_p(0)


_KS = {np.dtype(np.uint32): "u32", np.dtype(np.uint64): "u64"}
_RS = {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64"}


class CpuLib:
    """Common interface over liboracle.so ('orc_' prefix) and libcstone_ref.so ('ref_' prefix)."""

    def __init__(self, path, kind):
        self.kind = kind
        self.lib = C.CDLL(path)
        self.pre = "orc_" if kind == "oracle" else "ref_"
        if kind == "oracle":
            self.lib.orc_malloc.restype = C.c_void_p

    def _f(self, name):
        return getattr(self.lib, self.pre + name)

    # ---- keys
    def compute_sfc_keys(self, x, y, z, box, key_dtype, curve=HILBERT, keys_in=None):
        x = np.ascontiguousarray(x)
        kd = np.dtype(key_dtype)
        keys = np.zeros(x.size, dtype=kd) if keys_in is None else np.array(keys_in, dtype=kd)
        if self.kind == "oracle":
            f = self._f(f"compute_sfc_keys_{_KS[kd]}_{_RS[x.dtype]}")
            f(C.c_int(curve), _ptr(x), _ptr(np.ascontiguousarray(y)), _ptr(np.ascontiguousarray(z)), _ptr(keys),
              C.c_size_t(x.size), C.byref(box), C.c_int(1))
        else:
            f = self._f(f"compute_sfc_keys_{_KS[kd]}")
            f(C.c_int(curve), _ptr(x), _ptr(np.ascontiguousarray(y)), _ptr(np.ascontiguousarray(z)), _ptr(keys),
              C.c_size_t(x.size), C.byref(box), C.c_int(int(x.dtype == np.float64)))
        return keys

    # ---- sort
    def sort_by_key(self, keys, values=None):
        k = np.array(keys, copy=True)
        v = np.arange(k.size, dtype=np.uint32) if values is None else np.array(values, dtype=np.uint32)
        self._f(f"sort_by_key_{_KS[k.dtype]}")(_ptr(k), _ptr(v), C.c_size_t(k.size))
        return k, v

    # ---- leaf array
    def compute_node_counts(self, tree, keys, max_count=0xFFFFFFFF, guess=None):
        n = tree.size - 1
        counts = np.zeros(n, dtype=np.uint32) if guess is None else np.array(guess, dtype=np.uint32)
        self._f(f"compute_node_counts_{_KS[tree.dtype]}")(_ptr(tree), _ptr(counts), C.c_int(n), _ptr(keys),
                                                           C.c_size_t(keys.size), C.c_uint(max_count),
                                                           C.c_int(int(guess is not None)))
        return counts

    def rebalance_decision(self, tree, counts, bucket):
        n = tree.size - 1
        ops = np.zeros(n + 1, dtype=np.int32)
        conv = self._f(f"rebalance_decision_{_KS[tree.dtype]}")(_ptr(tree), _ptr(counts), C.c_int(n),
                                                                 C.c_uint(bucket), _ptr(ops))
        return ops[:n], bool(conv)

    def compute_octree(self, keys, bucket, max_count=0xFFFFFFFF):
        if self.kind == "oracle":
            tp = C.c_void_p()
            cp = C.c_void_p()
            nn = C.c_int()
            calls = self._f(f"compute_octree_{_KS[keys.dtype]}")(_ptr(keys), C.c_size_t(keys.size),
                                                                  C.c_uint(bucket), C.byref(tp), C.byref(nn),
                                                                  C.byref(cp), C.c_uint(max_count))
            tree = np.ctypeslib.as_array(C.cast(tp, C.POINTER(C.c_uint64 if keys.dtype == np.uint64 else C.c_uint32)),
                                         shape=(nn.value + 1,)).copy()
            counts = np.ctypeslib.as_array(C.cast(cp, C.POINTER(C.c_uint32)), shape=(nn.value,)).copy()
            self.lib.orc_free(tp)
            self.lib.orc_free(cp)
            return tree, counts
        cap = keys.size + 64
        while True:
            tree = np.zeros(cap + 1, dtype=keys.dtype)
            counts = np.zeros(cap, dtype=np.uint32)
            nn = self._f(f"compute_octree_{_KS[keys.dtype]}")(_ptr(keys), C.c_size_t(keys.size), C.c_uint(bucket),
                                                               _ptr(tree), C.c_int(cap), _ptr(counts))
            if nn <= cap:
                return tree[: nn + 1].copy(), counts[:nn].copy()
            cap = nn

    def update_octree(self, keys, bucket, tree, counts, max_count=0xFFFFFFFF):
        """One updateOctree step; returns (tree, counts, converged)."""
        if self.kind == "oracle":
            n = tree.size - 1
            # the oracle replaces the malloc'd tree/counts it is given (orc_update_octree_*)
            libc = C.CDLL(None)
            libc.malloc.restype = C.c_void_p
            tbuf = libc.malloc(C.c_size_t(tree.nbytes))
            cbuf = libc.malloc(C.c_size_t(max(counts.nbytes, 4)))
            C.memmove(tbuf, tree.ctypes.data, tree.nbytes)
            C.memmove(cbuf, counts.ctypes.data, counts.nbytes)
            tpp = C.c_void_p(tbuf)
            cpp = C.c_void_p(cbuf)
            nn = C.c_int(n)
            conv = self._f(f"update_octree_{_KS[keys.dtype]}")(_ptr(keys), C.c_size_t(keys.size), C.c_uint(bucket),
                                                                C.byref(tpp), C.byref(nn), C.byref(cpp),
                                                                C.c_uint(max_count))
            newt = np.ctypeslib.as_array(C.cast(tpp, C.POINTER(C.c_uint64 if keys.dtype == np.uint64 else C.c_uint32)),
                                         shape=(nn.value + 1,)).copy()
            newc = np.ctypeslib.as_array(C.cast(cpp, C.POINTER(C.c_uint32)), shape=(nn.value,)).copy()
            self.lib.orc_free(tpp)
            self.lib.orc_free(cpp)
            return newt, newc, bool(conv)
        cap = max(tree.size * 4096, 64)
        t = np.zeros(cap + 1, dtype=tree.dtype)
        c = np.zeros(cap, dtype=np.uint32)
        t[: tree.size] = tree
        c[: counts.size] = counts
        nn = C.c_int(tree.size - 1)
        conv = self._f(f"update_octree_{_KS[keys.dtype]}")(_ptr(keys), C.c_size_t(keys.size), C.c_uint(bucket),
                                                            _ptr(t), C.byref(nn), _ptr(c), C.c_int(cap))
        return t[: nn.value + 1].copy(), c[: nn.value].copy(), bool(conv)

    # ---- link
    def build_octree(self, leaves):
        L = leaves.size - 1
        NI = (L - 1) // 7
        N = L + NI
        Lmax = 21 if leaves.dtype == np.uint64 else 10
        out = dict(prefixes=np.zeros(N, dtype=leaves.dtype), childOffsets=np.zeros(N, dtype=np.int32),
                   parents=np.zeros(max(1, (N - 1) // 8), dtype=np.int32), levelRange=np.zeros(Lmax + 2, np.int32),
                   internalToLeaf=np.zeros(N, np.int32), leafToInternal=np.zeros(N, np.int32))
        self._f(f"build_octree_{_KS[leaves.dtype]}")(_ptr(leaves), C.c_int(L), _ptr(out["prefixes"]),
                                                      _ptr(out["childOffsets"]), _ptr(out["parents"]),
                                                      _ptr(out["levelRange"]), _ptr(out["internalToLeaf"]),
                                                      _ptr(out["leafToInternal"]))
        return out

    # ---- geometry
    def node_fp_centers(self, prefixes, box, real_dtype, curve=HILBERT):
        T = np.dtype(real_dtype)
        N = prefixes.size
        c = np.zeros((N, 3), dtype=T)
        s = np.zeros((N, 3), dtype=T)
        f = self._f(f"node_fp_centers_{_KS[prefixes.dtype]}_{_RS[T]}")
        f(C.c_int(curve), _ptr(prefixes), C.c_size_t(N), _ptr(c), _ptr(s), C.byref(box))
        return c, s

    # ---- source (mass) centers and MAC radii of the focus tree
    def leaf_mass_centers(self, x, y, z, m, leaf_to_internal, layout, num_nodes, centers=None):
        """centers[leafToInternal[leaf]] = (mass center, total |m|); other nodes keep `centers` (zeros if None)."""
        T = x.dtype
        c = np.zeros((num_nodes, 4), dtype=T) if centers is None else np.array(centers, dtype=T, copy=True)
        l2i = np.ascontiguousarray(leaf_to_internal, dtype=np.int32)
        self._f(f"leaf_mass_centers_{_RS[T]}")(_ptr(x), _ptr(y), _ptr(z), _ptr(np.ascontiguousarray(m, dtype=T)),
                                               _ptr(l2i), C.c_int(l2i.size),
                                               _ptr(np.ascontiguousarray(layout, dtype=np.uint32)), _ptr(c))
        return c

    def upsweep_centers(self, level_range, child_offsets, centers, num_levels):
        c = np.array(centers, copy=True)
        self._f(f"upsweep_centers_{_RS[c.dtype]}")(_ptr(np.ascontiguousarray(level_range, dtype=np.int32)),
                                                    C.c_int(num_levels),
                                                    _ptr(np.ascontiguousarray(child_offsets, dtype=np.int32)),
                                                    _ptr(c))
        return c

    def set_mac(self, prefixes, centers, inv_theta, box):
        c = np.array(centers, copy=True)
        self._f(f"set_mac_{_KS[prefixes.dtype]}_{_RS[c.dtype]}")(_ptr(prefixes), C.c_size_t(prefixes.size), _ptr(c),
                                                                   C.c_float(inv_theta), C.byref(box))
        return c

    def geo_mac_spheres(self, prefixes, real_dtype, inv_theta, box):
        T = np.dtype(real_dtype)
        c = np.zeros((prefixes.size, 4), dtype=T)
        self._f(f"geo_mac_spheres_{_KS[prefixes.dtype]}_{_RS[T]}")(_ptr(prefixes), C.c_size_t(prefixes.size),
                                                                     _ptr(c), C.c_float(inv_theta), C.byref(box))
        return c

    def compute_bounding_box(self, x, y, z, h, layout, first, last, scale, centers):
        T = x.dtype
        c = np.array(centers, dtype=T, copy=True)
        s = np.zeros_like(c)
        ct = C.c_float if T == np.float32 else C.c_double
        self._f(f"compute_bounding_box_{_RS[T]}")(_ptr(x), _ptr(y), _ptr(z), _ptr(h), _ptr(layout), C.c_int(first),
                                                   C.c_int(last), ct(scale), _ptr(c), _ptr(s))
        return c, s

    def find_halos(self, prefixes, childOffsets, parents, centers, sizes, leaves, sc, ss, box, first, last):
        flags = np.zeros(prefixes.size, dtype=np.uint8)
        f = self._f(f"find_halos_{_KS[prefixes.dtype]}_{_RS[centers.dtype]}")
        f(_ptr(prefixes), _ptr(childOffsets), _ptr(parents), _ptr(np.ascontiguousarray(centers)),
          _ptr(np.ascontiguousarray(sizes)), _ptr(leaves), _ptr(np.ascontiguousarray(sc)),
          _ptr(np.ascontiguousarray(ss)), C.byref(box), C.c_int(first), C.c_int(last), _ptr(flags))
        return flags

    def find_neighbors(self, x, y, z, h, first, last, box, tree, layout, centers, sizes, ngmax, ext=1.0):
        T = x.dtype
        n = last - first
        nb = np.zeros(max(1, n * ngmax), dtype=np.uint32)
        nc = np.zeros(max(1, n), dtype=np.uint32)
        f = self._f(f"find_neighbors_{_RS[T]}")
        f(_ptr(x), _ptr(y), _ptr(z), _ptr(h), C.c_uint32(first), C.c_uint32(last), C.byref(box),
          _ptr(tree["childOffsets"]), _ptr(tree["parents"]), _ptr(tree["internalToLeaf"]), _ptr(layout),
          _ptr(np.ascontiguousarray(centers)), _ptr(np.ascontiguousarray(sizes)), C.c_float(ext), C.c_uint(ngmax),
          _ptr(nb), _ptr(nc))
        return nb[: n * ngmax].reshape(n, ngmax) if ngmax else nb[:0].reshape(n, 0), nc[:n]

    def ijloop_example(self, x, y, z, h, m, first, last, box, tree, layout, centers, sizes, ngmax, ext=1.0):
        """The example ijloop interaction (rho, count, dmin, dmax, fx), body-indexed; entries outside [first, last)
        stay zero."""
        T = x.dtype
        n = x.size
        out = [np.zeros(n, T), np.zeros(n, np.uint32), np.zeros(n, T), np.zeros(n, T), np.zeros(n, T)]
        f = self._f(f"ijloop_example_{_RS[T]}")
        f(_ptr(x), _ptr(y), _ptr(z), _ptr(h), _ptr(np.ascontiguousarray(m, dtype=T)), C.c_uint32(first),
          C.c_uint32(last), C.byref(box), _ptr(tree["childOffsets"]), _ptr(tree["parents"]),
          _ptr(tree["internalToLeaf"]), _ptr(layout), _ptr(np.ascontiguousarray(centers)),
          _ptr(np.ascontiguousarray(sizes)), C.c_float(ext), C.c_uint(ngmax), *[_ptr(a) for a in out])
        return out

    def find_neighbors_subset(self, x, y, z, h, ids, box, tree, layout, centers, sizes, ngmax, ext=1.0):
        """Per-particle search for a list of indices (oracle only)."""
        assert self.kind == "oracle"
        T = x.dtype
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        nb = np.zeros(max(1, ids.size * ngmax), dtype=np.uint32)
        nc = np.zeros(max(1, ids.size), dtype=np.uint32)
        f = self._f(f"find_neighbors_subset_{_RS[T]}")
        f(_ptr(x), _ptr(y), _ptr(z), _ptr(h), _ptr(ids), C.c_size_t(ids.size), C.byref(box),
          _ptr(tree["childOffsets"]), _ptr(tree["parents"]), _ptr(tree["internalToLeaf"]), _ptr(layout),
          _ptr(np.ascontiguousarray(centers)), _ptr(np.ascontiguousarray(sizes)), C.c_float(ext), C.c_uint(ngmax),
          _ptr(nb), _ptr(nc))
        return nb[: ids.size * ngmax].reshape(ids.size, ngmax), nc[: ids.size]

    # ---- reference pipeline timing (REF only)
    def pipeline(self, x, y, z, h, box, bucket, ngmax, nb_stride=1, tree=None, counts=None, curve=HILBERT,
                 key_dtype=np.uint64):
        """One reference CPU pipeline step (keys, stable sort + gather, computeOctree or one warm updateOctree,
        link, layout, centers, neighbors of every nb_stride-th particle); x,y,z,h are reordered in place.
        Returns (tree, counts, stage_times_ms[keys, sort, tree, link, centers, neighbors, total], counts_sampled)."""
        assert self.kind == "ref"
        T = x.dtype
        n = x.size
        cap = max(8 * (n // max(bucket, 1) + 64), 1024)
        t = np.zeros(cap + 1, dtype=key_dtype)
        c = np.zeros(cap, dtype=np.uint32)
        nn = C.c_int(0)
        if tree is not None:
            t[: tree.size] = tree
            c[: counts.size] = counts
            nn = C.c_int(tree.size - 1)
        m = (n + nb_stride - 1) // nb_stride
        nb = np.zeros(max(1, m * ngmax), dtype=np.uint32)
        nc = np.zeros(max(1, m), dtype=np.uint32)
        times = np.zeros(7, dtype=np.float64)
        f = self._f(f"pipeline_{_KS[np.dtype(key_dtype)]}_{_RS[T]}")
        f(C.c_int(curve), _ptr(x), _ptr(y), _ptr(z), _ptr(h), C.c_size_t(n), C.byref(box), C.c_uint(bucket), _ptr(t),
          C.byref(nn), _ptr(c), C.c_int(cap), C.c_uint(ngmax), C.c_uint32(nb_stride), _ptr(nb), _ptr(nc),
          _ptr(times))
        return t[: nn.value + 1].copy(), c[: nn.value].copy(), times, nc[:m]

    # ---- domain decomposition and halo layout
    def make_sfc_assignment(self, tree, counts, num_ranks):
        """makeSfcAssignment (domaindecomp.hpp:103-118): (treeOffsets, boundaries, rank counts)."""
        tree = np.ascontiguousarray(tree)
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        L = counts.size
        off = np.zeros(num_ranks + 1, dtype=np.int32)
        bnd = np.zeros(num_ranks + 1, dtype=tree.dtype)
        rc = np.zeros(num_ranks, dtype=np.uint32)
        self._f(f"make_sfc_assignment_{_KS[tree.dtype]}")(_ptr(tree), _ptr(counts), C.c_int(L), C.c_int(num_ranks),
                                                          _ptr(off), _ptr(bnd), _ptr(rc))
        return off, bnd, rc

    def extract_marked_elements(self, source, flags, first, second):
        """extractMarkedElements (layout.hpp:111-141) as an (n, 2) array of pairs."""
        source = np.ascontiguousarray(source)
        flags = np.ascontiguousarray(flags, dtype=np.uint32)
        cap = max(1, (second - first + 1) // 2 + 1)
        pairs = np.zeros(2 * cap, dtype=source.dtype)
        f = self._f(f"extract_marked_elements_{_KS[source.dtype]}")
        if self.kind == "oracle":
            f.restype = C.c_size_t
            n = f(_ptr(source), _ptr(flags), C.c_int(first), C.c_int(second), _ptr(pairs), C.c_size_t(cap))
        else:
            f.restype = C.c_size_t
            n = f(_ptr(source), C.c_size_t(source.size), _ptr(flags), C.c_size_t(flags.size), C.c_int(first),
                  C.c_int(second), _ptr(pairs), C.c_size_t(cap))
        assert n <= cap
        return pairs[: 2 * n].reshape(n, 2).copy()

    def compute_node_layout(self, counts, flags, leaf_to_internal, first, last):
        """computeNodeLayout (layout.hpp:153-185), CPU branch: layout of numLeaves + 1 entries. leaf_to_internal is
        the leaf part of the map (leafToInternal + numInternalNodes, octree.hpp:372-375)."""
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        flags = np.ascontiguousarray(flags, dtype=np.uint8)
        l2i = np.ascontiguousarray(leaf_to_internal, dtype=np.int32)
        L = counts.size
        layout = np.zeros(L + 1, dtype=np.uint32)
        if self.kind == "oracle":
            self.lib.orc_compute_node_layout(_ptr(counts), _ptr(flags), _ptr(l2i), C.c_int(L), C.c_int(first),
                                             C.c_int(last), _ptr(layout))
        else:
            self.lib.ref_compute_node_layout(_ptr(counts), _ptr(flags), _ptr(l2i), C.c_int(L), C.c_int(flags.size),
                                             C.c_int(first), C.c_int(last), _ptr(layout))
        return layout

    def gather_ranges(self, range_scan, range_offsets, src, buffer_size):
        """gatherRanges (gather_halos_gpu.cu:28-44) restated as a loop over ranges (oracle only)."""
        assert self.kind == "oracle"
        rs = np.ascontiguousarray(range_scan, dtype=np.uint32)
        ro = np.ascontiguousarray(range_offsets, dtype=np.uint32)
        src = np.ascontiguousarray(src)
        out = np.zeros((buffer_size,) + src.shape[1:], dtype=src.dtype)
        elem = src.dtype.itemsize * int(np.prod(src.shape[1:], dtype=np.int64))
        self.lib.orc_gather_ranges(_ptr(rs), _ptr(ro), C.c_int(rs.size), _ptr(src), _ptr(out),
                                   C.c_size_t(buffer_size), C.c_size_t(elem))
        return out

    def gather_leaf_ranges(self, src_layout, dst_layout, src):
        """Kept-leaf gather from the full arrays to the local layout (oracle only)."""
        assert self.kind == "oracle"
        sl = np.ascontiguousarray(src_layout, dtype=np.uint32)
        dl = np.ascontiguousarray(dst_layout, dtype=np.uint32)
        src = np.ascontiguousarray(src)
        out = np.zeros(int(dl[-1]), dtype=src.dtype)
        self.lib.orc_gather_leaf_ranges(_ptr(sl), _ptr(dl), C.c_int(dl.size - 1), _ptr(src), _ptr(out),
                                        C.c_size_t(src.dtype.itemsize))
        return out

    def discover_halos(self, tree, centers, sizes, leaves, x, y, z, h, layout, first, last, box, search_ext=1.0,
                       accumulate_flags=None):
        """discoverHalos (octree_focus_mpi.hpp:511-568) over the leaf layout (oracle only):
        (searchCenters, searchSizes, flags)."""
        assert self.kind == "oracle"
        T = x.dtype
        K = leaves.dtype
        N = tree["prefixes"].size
        L = leaves.size - 1
        sc = np.zeros((L, 3), dtype=T)
        ss = np.zeros((L, 3), dtype=T)
        flags = (np.zeros(N, dtype=np.uint8) if accumulate_flags is None
                 else np.array(accumulate_flags, dtype=np.uint8, copy=True))
        l2i = np.ascontiguousarray(tree["leafToInternal"][N - L:])  # the leaf part, octree.hpp:252
        self._f(f"discover_halos_{_KS[K]}_{_RS[T]}")(
            _ptr(tree["prefixes"]), _ptr(tree["childOffsets"]), _ptr(tree["parents"]), _ptr(l2i),
            _ptr(np.ascontiguousarray(centers)), _ptr(np.ascontiguousarray(sizes)), _ptr(leaves), C.c_int(N),
            _ptr(x), _ptr(y), _ptr(z), _ptr(h), _ptr(np.ascontiguousarray(layout, dtype=np.uint32)),
            C.c_int(first), C.c_int(last), C.byref(box), C.c_float(search_ext), _ptr(sc), _ptr(ss), _ptr(flags),
            C.c_int(0 if accumulate_flags is None else 1))
        return sc, ss, flags

    # ---- focus tree and the domain's global-tree rule
    def _max_key(self, kd):
        return np.array(1 << (63 if np.dtype(kd) == np.uint64 else 30), dtype=kd)

    def update_focus(self, leaves, bucket, focus_start, focus_end, mandatory, counts, macs=None):
        """CombinedUpdate::updateFocus (octree_focus.hpp:64-111): (new leaves, converged)."""
        kd = leaves.dtype
        mandatory = np.ascontiguousarray(mandatory, dtype=kd)
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        macs = None if macs is None else np.ascontiguousarray(macs, dtype=np.uint8)
        L = leaves.size - 1
        kt = C.c_uint32 if kd == np.uint32 else C.c_uint64
        if self.kind == "oracle":
            p = C.c_void_p(self.lib.orc_malloc(C.c_size_t(leaves.nbytes)))
            C.memmove(p, leaves.ctypes.data, leaves.nbytes)
            nl = C.c_int(L)
            f = self._f(f"update_focus_{_KS[kd]}")
            conv = f(C.byref(p), C.byref(nl), C.c_uint(bucket), kt(int(focus_start)), kt(int(focus_end)),
                     _ptr(mandatory), C.c_int(mandatory.size), _ptr(counts), _ptr(macs))
            out = np.ctypeslib.as_array(C.cast(p, C.POINTER(kt)), shape=(nl.value + 1,)).copy()
            self.lib.orc_free(p)
            return out, bool(conv)
        cap = max(64, 16 * (L + 1) + 8 * mandatory.size)
        while True:
            buf = np.zeros(cap + 1, dtype=kd)
            buf[: L + 1] = leaves
            nl = C.c_int(L)
            conv = self._f(f"update_focus_{_KS[kd]}")(_ptr(buf), C.byref(nl), C.c_int(cap), C.c_uint(bucket),
                                                      kt(int(focus_start)), kt(int(focus_end)), _ptr(mandatory),
                                                      C.c_int(mandatory.size), _ptr(counts), _ptr(macs))
            if nl.value <= cap:
                return buf[: nl.value + 1].copy(), bool(conv)
            cap = nl.value + 1

    def focus_counts(self, leaves, keys):
        """Leaf counts and upswept node counts of a focus tree (octree_focus.hpp:264-270)."""
        L = leaves.size - 1
        N = L + (L - 1) // 7
        lc = np.zeros(L, dtype=np.uint32)
        nc = np.zeros(N, dtype=np.uint32)
        keys = np.ascontiguousarray(keys)
        self._f(f"focus_counts_{_KS[leaves.dtype]}")(_ptr(np.ascontiguousarray(leaves)), C.c_int(L), _ptr(keys),
                                                       C.c_size_t(keys.size), _ptr(lc), _ptr(nc))
        return lc, nc

    def compute_spanning_tree(self, keys):
        keys = np.ascontiguousarray(keys)
        kd = keys.dtype
        if self.kind == "oracle":
            f = self.lib[f"orc_compute_spanning_tree_{_KS[kd]}"]
            f.restype = C.c_void_p
            n = C.c_int()
            p = f(_ptr(keys), C.c_int(keys.size), C.byref(n))
            kt = C.c_uint32 if kd == np.uint32 else C.c_uint64
            out = np.ctypeslib.as_array(C.cast(p, C.POINTER(kt)), shape=(n.value,)).copy()
            self.lib.orc_free(C.c_void_p(p))
            return out
        cap = 1024
        while True:
            out = np.zeros(cap, dtype=kd)
            n = self._f(f"compute_spanning_tree_{_KS[kd]}")(_ptr(keys), C.c_int(keys.size), _ptr(out), C.c_int(cap))
            if n <= cap:
                return out[:n].copy()
            cap = n

    def domain_global_update(self, sorted_keys, bucket, tree, counts, first_call):
        """The domain's global-tree rule at one rank (assignment.hpp:115-123): (tree, counts, steps)."""
        kd = tree.dtype
        L = tree.size - 1
        if self.kind == "oracle":
            kt = C.c_uint32 if kd == np.uint32 else C.c_uint64
            tp = C.c_void_p(self.lib.orc_malloc(C.c_size_t(tree.nbytes)))
            C.memmove(tp, tree.ctypes.data, tree.nbytes)
            c = np.ascontiguousarray(counts, dtype=np.uint32)
            cp = C.c_void_p(self.lib.orc_malloc(C.c_size_t(max(4, c.nbytes))))
            C.memmove(cp, c.ctypes.data, c.nbytes)
            nn = C.c_int(L)
            steps = self._f(f"domain_global_update_{_KS[kd]}")(_ptr(sorted_keys), C.c_size_t(sorted_keys.size),
                                                               C.c_uint(bucket), C.byref(tp), C.byref(nn),
                                                               C.byref(cp), C.c_int(int(first_call)))
            t = np.ctypeslib.as_array(C.cast(tp, C.POINTER(kt)), shape=(nn.value + 1,)).copy()
            cc = np.ctypeslib.as_array(C.cast(cp, C.POINTER(C.c_uint32)), shape=(nn.value,)).copy()
            self.lib.orc_free(tp)
            self.lib.orc_free(cp)
            return t, cc, steps
        cap = max(4096, 8 * (L + 1))
        while True:
            t = np.zeros(cap + 1, dtype=kd)
            t[: L + 1] = tree
            cc = np.zeros(cap, dtype=np.uint32)
            cc[:L] = counts
            nn = C.c_int(L)
            steps = self._f(f"domain_global_update_{_KS[kd]}")(_ptr(sorted_keys), C.c_size_t(sorted_keys.size),
                                                               C.c_uint(bucket), _ptr(t), C.byref(nn), _ptr(cc),
                                                               C.c_int(cap), C.c_int(int(first_call)))
            if nn.value <= cap:
                return t[: nn.value + 1].copy(), cc[: nn.value].copy(), steps
            cap = nn.value

    def halo_chain(self, x, y, z, h, box, bucket, num_ranks, rank, curve=MORTON, search_ext=1.0):
        """The rank-share chain through the reference's CPU functions (ref only): (local particles, stage ms[8]).
        x, y, z, h are sorted in place."""
        assert self.kind == "ref"
        T = x.dtype
        times = np.zeros(8, dtype=np.float64)
        f = self._f(f"halo_chain_u64_{_RS[T]}")
        f.restype = C.c_size_t
        m = f(C.c_int(curve), _ptr(x), _ptr(y), _ptr(z), _ptr(h), C.c_size_t(x.size), C.byref(box), C.c_uint(bucket),
              C.c_int(num_ranks), C.c_int(rank), C.c_float(search_ext), _ptr(times))
        return int(m), times

    def omp_threads(self):
        assert self.kind == "ref"
        return int(self.lib.ref_omp_threads())


ORACLE = CpuLib(ORACLE_PATH, "oracle") if os.path.exists(ORACLE_PATH) else None
REF = CpuLib(REF_PATH, "ref") if os.path.exists(REF_PATH) else None


def fitting_box(x, y, z, prev_lo, prev_hi, boundary):
    """makeGlobalBox (sfc/box_mpi.hpp:51-105) at one rank: periodic and fixed dimensions keep the previous limits,
    the others take the coordinates' min/max; cubic_open dimensions are widened to the largest side. Limits in T."""
    T = np.asarray(x).dtype.type
    lo = [T(v) for v in prev_lo]
    hi = [T(v) for v in prev_hi]
    for d, a in enumerate((x, y, z)):
        if boundary[d] not in (1, 2) and a.size:
            lo[d], hi[d] = T(a.min()), T(a.max())
    side = max(T(hi[0] - lo[0]), T(hi[1] - lo[1]), T(hi[2] - lo[2]))
    for d in range(3):
        if boundary[d] == 3:
            hi[d] = max(hi[d], T(lo[d] + side))
    return lo, hi


def limit_box_shrinking(fit_lo, fit_hi, prev_lo, prev_hi):
    """limitBoxShrinking (sfc/box.hpp:398-415) with shrinkLimit 0.05, in T."""
    T = type(fit_lo[0])
    s = T(0.05)
    lo, hi = [], []
    for d in range(3):
        L = T(prev_hi[d] - prev_lo[d])
        lo.append(min(fit_lo[d], T(prev_lo[d] + T(s * L))))
        hi.append(max(fit_hi[d], T(prev_hi[d] - T(s * L))))
    return lo, hi


class CpuDomain:
    """Domain<KeyType, T>::sync at one rank (domain/domain.hpp:169-218), sequenced over a CPU lib's functions:
    fitted / shrink-limited box (assignment.hpp:101-103), Hilbert keys and a stable sort, the global tree rule
    (assignment.hpp:115-123) from the initial root {0, nodeRange(0)} with counts bucketSize-1 (assignment.hpp:62-66,
    the spanning tree of initialDomainSplits at one rank), particles gathered into SFC order, then the focus tree
    (octree_focus_mpi.hpp:61-87: root with count bucketSizeFocus+1): converged on the first call
    (:585-603), one update per call (domain.hpp:200-203). The focus area is the whole key range and every global
    leaf key is mandatory (octree_focus_mpi.hpp:120-121), so MAC flags do not enter the decisions."""

    def __init__(self, lib, key_dtype, real_dtype, bucket, bucket_focus, lo, hi, boundary=(0, 0, 0)):
        self.lib = lib
        self.kd = np.dtype(key_dtype)
        self.T = np.dtype(real_dtype).type
        self.bucket, self.bucket_focus = bucket, bucket_focus
        self.lo = [self.T(v) for v in np.broadcast_to(lo, (3,))]
        self.hi = [self.T(v) for v in np.broadcast_to(hi, (3,))]
        self.boundary = tuple(boundary)
        top = 1 << (63 if self.kd == np.uint64 else 30)
        self.top = top
        self.gtree = np.array([0, top], dtype=self.kd)
        self.gcounts = np.array([bucket - 1], dtype=np.uint32)
        self.fleaves = np.array([0, top], dtype=self.kd)
        self.fnodecounts = np.array([bucket_focus + 1], dtype=np.uint32)
        self.fleafcounts = self.fnodecounts.copy()
        self.first = True

    def box(self):
        return make_box(self.lo, self.hi, self.T, self.boundary)

    def sync(self, x, y, z, h):
        lo, hi = fitting_box(x, y, z, self.lo, self.hi, self.boundary)
        if not self.first:
            if isinstance(self.lib, CpuLib) and self.lib.kind == "ref":
                out = np.zeros(6, dtype=self.T)
                fit = np.array([lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]], dtype=self.T)
                prev = np.array([self.lo[0], self.hi[0], self.lo[1], self.hi[1], self.lo[2], self.hi[2]],
                                dtype=self.T)
                bnd = np.array(self.boundary, dtype=np.int32)
                self.lib._f(f"limit_box_shrinking_{_RS[np.dtype(self.T)]}")(_ptr(fit), _ptr(prev), _ptr(bnd),
                                                                             _ptr(out))
                lo, hi = [out[0], out[2], out[4]], [out[1], out[3], out[5]]
            else:
                lo, hi = limit_box_shrinking(lo, hi, self.lo, self.hi)
        self.lo, self.hi = [self.T(v) for v in lo], [self.T(v) for v in hi]
        box = self.box()
        keys = self.lib.compute_sfc_keys(x, y, z, box, self.kd, HILBERT)
        keys, perm = self.lib.sort_by_key(keys)
        self.gtree, self.gcounts, gsteps = self.lib.domain_global_update(keys, self.bucket, self.gtree, self.gcounts,
                                                                         self.first)
        xs, ys, zs, hs = x[perm], y[perm], z[perm], h[perm]
        fsteps = 0
        if self.first:
            while True:
                fsteps += 1
                self.fleaves, conv = self.lib.update_focus(self.fleaves, self.bucket_focus, 0, self.top, self.gtree,
                                                           self.fnodecounts)
                self.fleafcounts, self.fnodecounts = self.lib.focus_counts(self.fleaves, keys)
                if conv:
                    break
        fsteps += 1
        self.fleaves, conv = self.lib.update_focus(self.fleaves, self.bucket_focus, 0, self.top, self.gtree,
                                                   self.fnodecounts)
        self.fleafcounts, self.fnodecounts = self.lib.focus_counts(self.fleaves, keys)
        self.first = False
        return dict(keys=keys, perm=perm, x=xs, y=ys, z=zs, h=hs, lo=self.lo, hi=self.hi, gtree=self.gtree,
                    gcounts=self.gcounts, gsteps=gsteps, fleaves=self.fleaves, fcounts=self.fleafcounts,
                    fnodecounts=self.fnodecounts, fsteps=fsteps, converged=conv)

    def sync_grav(self, x, y, z, h, m, theta):
        """Domain::syncGrav at one rank (domain.hpp:220-287): the MAC-driven refinement selects nothing (every node
        lies in the focus), so the order and trees are sync's; m is reordered with x, y, z, h and the expansion
        centers are those of expansion_centers() on the final focus tree."""
        r = self.sync(x, y, z, h)
        r["m"] = m[r["perm"]]
        r["expansion_centers"], r["octree"] = self.expansion_centers(r["x"], r["y"], r["z"], r["m"], theta)
        return r

    def expansion_centers(self, x, y, z, m, theta):
        """Domain::updateExpansionCenters (domain.hpp:366-372): the focus tree's leaf mass centers over the layout of
        its leaf counts (updateCenters, octree_focus_mpi.hpp:351-414), the upsweep, and the squared vector-MAC radius
        of 1.0 / theta (a double, passed as float to setMacRadius, :446-453); returns (centers, octree)."""
        o = self.lib.build_octree(self.fleaves)
        L = self.fleaves.size - 1
        n_int = o["prefixes"].size - L
        layout = np.zeros(L + 1, dtype=np.uint32)
        layout[1:] = np.cumsum(self.fleafcounts, dtype=np.uint64).astype(np.uint32)
        c = self.lib.leaf_mass_centers(x, y, z, m, o["leafToInternal"][n_int:], layout, o["prefixes"].size)
        c = self.lib.upsweep_centers(o["levelRange"], o["childOffsets"], c, o["levelRange"].size - 2)
        inv_theta = np.float32(1.0 / np.float64(np.float32(theta)))
        return self.lib.set_mac(o["prefixes"], c, float(inv_theta), self.box()), o


REFDOM_PATH = os.path.join(HERE, "_ref", "libcstone_ref_domain.so")


class RefDomain:
    """The reference's own one-rank Domain<KeyType, T, Cpu> (domain/domain.hpp:38-420), run as an MPI singleton
    through oracle/_ref/libcstone_ref_domain.so (oracle/ref_domain.cpp). sync() / sync_grav() return the same keys
    as CpuDomain.sync, plus what only a run of the reference knows: its focus-tree layout, the linked focus tree with
    its geometric centers, start/end indices and (syncGrav) the expansion centers. Only available where the
    reference sources and MPI were present at build time."""

    _lib = None

    @classmethod
    def available(cls):
        return os.path.exists(REFDOM_PATH)

    @classmethod
    def lib(cls):
        if cls._lib is None:
            lib = C.CDLL(REFDOM_PATH)
            lib.refdom_create.restype = C.c_void_p
            lib.refdom_last_error.restype = C.c_char_p
            cls._lib = lib
        return cls._lib

    def __init__(self, key_dtype, real_dtype, bucket, bucket_focus, lo, hi, boundary=(0, 0, 0), theta=0.5):
        self.kd = np.dtype(key_dtype)
        self.T = np.dtype(real_dtype).type
        self.box0 = make_box(lo, hi, self.T, boundary)
        lib = self.lib()
        self.h = lib.refdom_create(C.c_int(8 * self.kd.itemsize), C.c_int(8 * np.dtype(self.T).itemsize),
                                   C.c_uint(bucket), C.c_uint(bucket_focus), C.c_float(theta), C.byref(self.box0))
        if not self.h:
            raise RuntimeError("refdom_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib().refdom_destroy(C.c_void_p(self.h))
            self.h = None

    def _call(self, name, *args):
        return getattr(self.lib(), "refdom_" + name)(C.c_void_p(self.h), *args)

    def sync(self, x, y, z, h, m=None):
        n = x.size
        xs, ys, zs, hs = (np.array(a, dtype=self.T, copy=True) for a in (x, y, z, h))
        ms = None if m is None else np.array(m, dtype=self.T, copy=True)
        ids = np.zeros(n, np.uint32)
        keys = np.zeros(n, self.kd)
        rc = self._call("sync", _ptr(xs), _ptr(ys), _ptr(zs), _ptr(hs), _ptr(ms), _ptr(ids), C.c_size_t(n),
                        _ptr(keys))
        if rc != 0:
            raise RuntimeError("reference Domain sync failed: " + self.lib().refdom_last_error().decode())
        sz = np.zeros(7, np.int64)
        self._call("sizes", _ptr(sz))
        start, end, nwh, ng, nf, nn, ni = (int(v) for v in sz)
        bx = np.zeros(6)
        self._call("box", _ptr(bx))
        gtree, gcounts = np.zeros(ng + 1, self.kd), np.zeros(ng, np.uint32)
        self._call("global", _ptr(gtree), _ptr(gcounts))
        fl, flc, fnc = np.zeros(nf + 1, self.kd), np.zeros(nf, np.uint32), np.zeros(nn, np.uint32)
        layout = np.zeros(nf + 1, np.uint32)
        self._call("focus", _ptr(fl), _ptr(flc), _ptr(fnc), _ptr(layout))
        L = 11 if self.kd == np.uint32 else 22
        pf, co, i2l = np.zeros(nn, self.kd), np.zeros(nn, np.int32), np.zeros(nn, np.int32)
        lr = np.zeros(L + 1, np.int32)
        cen, siz = np.zeros(3 * nn, self.T), np.zeros(3 * nn, self.T)
        self._call("octree", _ptr(pf), _ptr(co), _ptr(i2l), _ptr(lr), _ptr(cen), _ptr(siz))
        r = dict(keys=keys, perm=ids, x=xs, y=ys, z=zs, h=hs, lo=[self.T(bx[0]), self.T(bx[2]), self.T(bx[4])],
                 hi=[self.T(bx[1]), self.T(bx[3]), self.T(bx[5])], gtree=gtree, gcounts=gcounts, fleaves=fl,
                 fcounts=flc, fnodecounts=fnc, layout=layout, start=start, end=end, with_halos=nwh,
                 prefixes=pf, childOffsets=co, internalToLeaf=i2l, levelRange=lr, centers=cen, sizes=siz,
                 num_internal=ni)
        if ms is not None:
            ec = np.zeros(4 * nn, self.T)
            self._call("expansion_centers", _ptr(ec))
            r["m"] = ms
            r["expansion_centers"] = ec
        return r

    def sync_grav(self, x, y, z, h, m):
        return self.sync(x, y, z, h, m)
