"""The C-ABI library loads and exports every entry point include/hipcstone.h declares (no GPU needed)."""

import ctypes
import os
import re
import subprocess

import numpy as np

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "hipcstone.h")
LIB = os.path.join(ROOT, "cornerstone-octree_amd", "lib", "libhipcstone.so")


def declared_functions():
    """Function names declared by the header after macro expansion (gcc -E)."""
    out = subprocess.run(["gcc", "-E", "-P", "-I", os.path.dirname(HEADER), HEADER], capture_output=True, text=True,
                         check=True).stdout
    names = set(re.findall(r"\b(hipcstone_[a-z0-9_]+)\s*\(", out))
    return sorted(names)


def test_header_declares_expected_families():
    names = declared_functions()
    for prefix in ("hipcstone_compute_sfc_keys_hilbert_u64_f32", "hipcstone_sort_by_key_u64",
                   "hipcstone_sort_gather_u64_f32", "hipcstone_compute_node_counts_u64",
                   "hipcstone_update_octree_u64", "hipcstone_build_octree_u64", "hipcstone_compute_geo_centers_u64_f32",
                   "hipcstone_find_halos_u64_f64", "hipcstone_find_neighbors_f32", "hipcstone_pipeline_sync"):
        assert prefix in names
    assert len(names) >= 60


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build with `make -C cornerstone-octree_amd`"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_box_matches_reference_constructor():
    # hipcstone_make_box_* evaluates lengths and inverse lengths in T like sfc/box.hpp:100-122
    import hipcstone

    for T in (np.float32, np.float64):
        b = hipcstone.make_box((-1.0, 0.1, 0.0), (1.0, 0.7, 3.0), T, (1, 0, 2))
        lo = np.array([-1.0, 0.1, 0.0], dtype=T)
        hi = np.array([1.0, 0.7, 3.0], dtype=T)
        assert np.array_equal(np.array(list(b.len), dtype=T), hi - lo)
        assert np.array_equal(np.array(list(b.ilen), dtype=T), T(1) / (hi - lo))
        assert list(b.bnd) == [1, 0, 2]


def test_errors_are_status_codes_not_exits():
    # the reference exits on GPU errors (cuda/errorcheck.cuh:15-27); the C ABI returns a status instead
    import hipcstone

    lib = ctypes.CDLL(LIB)
    st = lib.hipcstone_compute_sfc_keys_hilbert_u64_f32(None, None, None, None, None, ctypes.c_size_t(10), None, 0)
    assert st == 1
    assert b"null" in hipcstone._lib.hipcstone_last_error()
    st = lib.hipcstone_pipeline_create(None, 1, 64, 32, 64, 150)
    assert st == 1
    h = ctypes.c_void_p()
    assert lib.hipcstone_pipeline_create(ctypes.byref(h), 1, 48, 32, 64, 150) == 1
    assert lib.hipcstone_pipeline_create(ctypes.byref(h), 1, 64, 32, 0, 150) == 1


def test_struct_layouts_match_the_binding(tmp_path):
    """The ctypes mirrors of the header's structs have the C compiler's sizes and field offsets."""
    import hipcstone

    src = tmp_path / "layout.c"
    fields = [f for f, _ in hipcstone.PipelineView._fields_]
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"hipcstone.h\"\nint main(void){\n"
                   + 'printf("%zu %zu %zu\\n", sizeof(hipcstone_pipeline_view), sizeof(hipcstone_box_f32), '
                     "sizeof(hipcstone_box_f64));\n"
                   + "".join(f'printf("%zu\\n", offsetof(hipcstone_pipeline_view, {f}));\n' for f in fields)
                   + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    sizes, offsets = [int(v) for v in out[:3]], [int(v) for v in out[3:]]
    assert sizes == [ctypes.sizeof(hipcstone.PipelineView), ctypes.sizeof(hipcstone.BoxF32),
                     ctypes.sizeof(hipcstone.BoxF64)]
    assert offsets == [getattr(hipcstone.PipelineView, f).offset for f in fields]


def test_domain_view_layout_matches_the_binding(tmp_path):
    import hipcstone

    src = tmp_path / "dlayout.c"
    fields = [f for f, _ in hipcstone.DomainView._fields_]
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"hipcstone.h\"\nint main(void){\n"
                   + 'printf("%zu\\n", sizeof(hipcstone_domain_view));\n'
                   + "".join(f'printf("%zu\\n", offsetof(hipcstone_domain_view, {f}));\n' for f in fields)
                   + "return 0;}\n")
    exe = tmp_path / "dlayout"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(hipcstone.DomainView)
    assert out[1:] == [getattr(hipcstone.DomainView, f).offset for f in fields]
