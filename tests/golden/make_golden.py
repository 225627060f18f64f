#!/usr/bin/env python3
"""Writes the golden fixtures of this directory from the reference's own CPU code (oracle/_ref, built from the
reference headers by `make -C oracle ref`). Run in a container where the reference sources are mounted:

    python tests/golden/make_golden.py

Fixtures are small .npz files of inputs and expected outputs (no code): keys for all curve/key/float variants,
a stable sort, a converged leaf array, its linked octree and geometric centers, a warm update step, neighbor
lists, per-leaf search boxes and halo flags; and (domain.npz) the outputs of the reference's own one-rank Domain
(oracle/_ref/libcstone_ref_domain.so, `make -C oracle refdomain`) over several sync / syncGrav calls.

    python tests/golden/make_golden.py [keys tree halos domain]   (default: all)
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "cornerstone-octree_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle_py import HILBERT, MORTON, REF, RefDomain, make_box  # noqa: E402

from hipcstone import inputs  # noqa: E402
from helpers import DOMAIN_GOLDEN_CASES, domain_input, domain_move  # noqa: E402

def domain_fixture():
    """Runs the reference Domain over each case's moving input (tests/helpers.py domain_input / domain_move, whose
    next input is made from the previous call's output) and stores every call's outputs."""
    if not RefDomain.available():
        sys.exit("oracle/_ref/libcstone_ref_domain.so is missing: run `make -C oracle refdomain`")
    out = {}
    for name, kd, T, bucket, bucket_focus, boundary, theta, n, calls, grav, seed in DOMAIN_GOLDEN_CASES:
        dom = RefDomain(kd, T, bucket, bucket_focus, -1.0, 1.0, boundary, theta=theta)
        arrays = domain_input(n, T, seed, with_mass=grav)
        out[f"{name}_input_sum"] = np.array([float(np.sum(a, dtype=np.float64)) for a in arrays])
        rng = np.random.default_rng(seed + 1000)
        for step in range(calls):
            r = dom.sync(*arrays)
            p = f"{name}{step}_"
            out[p + "keys"], out[p + "perm"] = r["keys"], r["perm"]
            out[p + "box"] = np.array([r["lo"][0], r["hi"][0], r["lo"][1], r["hi"][1], r["lo"][2], r["hi"][2]],
                                      dtype=T)
            for k in ("gtree", "gcounts", "fleaves", "fcounts", "fnodecounts", "layout"):
                out[p + k] = r[k]
            if grav:
                out[p + "expansion_centers"] = r["expansion_centers"]
            arrays = domain_move(rng, r, step, carried=("h", "m") if grav else ("h",))
    np.savez_compressed(os.path.join(HERE, "domain.npz"), **out)


def main():
    which = set(sys.argv[1:]) or {"keys", "tree", "halos", "domain"}
    if which & {"keys", "tree", "halos"}:
        if REF is None:
            sys.exit("oracle/_ref/libcstone_ref.so is missing: run `make -C oracle ref` where the reference is mounted")
        golden_components(which)
    if "domain" in which:
        domain_fixture()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


def golden_components(which):

    # ---- keys: 3000 Plummer points in a non-cubic box, every variant
    out = {}
    for T, ts in ((np.float32, "f32"), (np.float64, "f64")):
        x, y, z, _ = inputs.plummer(3000, 0.1, seed=101, dtype=T)
        lo, hi = (-1.0, -1.1, -0.95), (1.05, 1.3, 1.0)
        box = make_box(lo, hi, T)
        out[f"x_{ts}"], out[f"y_{ts}"], out[f"z_{ts}"] = x, y, z
        for kd, ks in ((np.uint32, "u32"), (np.uint64, "u64")):
            for curve, cs in ((HILBERT, "hilbert"), (MORTON, "morton")):
                out[f"keys_{cs}_{ks}_{ts}"] = REF.compute_sfc_keys(x, y, z, box, kd, curve)
    out["box_lo"] = np.array((-1.0, -1.1, -0.95))
    out["box_hi"] = np.array((1.05, 1.3, 1.0))
    if "keys" in which:
        np.savez_compressed(os.path.join(HERE, "keys.npz"), **out)

    # ---- tree: 20000 Plummer points, Hilbert-64, bucket 16; sort, computeOctree, link, centers, one warm step
    T = np.float32
    n = 20000
    x, y, z, r = inputs.plummer(n, 0.05, seed=102, dtype=T)
    h = inputs.plummer_h(r, n, 0.05, 40, dtype=T)
    box = make_box(-1.0, 1.0, T)
    keys = REF.compute_sfc_keys(x, y, z, box, np.uint64, HILBERT)
    ks, perm = REF.sort_by_key(keys)
    tree, counts = REF.compute_octree(ks, 16)
    o = REF.build_octree(tree)
    c, s = REF.node_fp_centers(o["prefixes"], box, T)
    xs, ys, zs, hs = x[perm], y[perm], z[perm], h[perm]
    layout = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    nb, nc = REF.find_neighbors(xs, ys, zs, hs, 0, n, box, o, layout, c, s, 48)
    rng = np.random.default_rng(103)
    x2 = np.clip(xs + (rng.random(n).astype(T) * 2 - 1) * T(0.02) * hs, -1, 1).astype(T)
    keys2 = REF.compute_sfc_keys(x2, ys, zs, box, np.uint64, HILBERT)
    ks2, perm2 = REF.sort_by_key(keys2)
    tree2, counts2, conv2 = REF.update_octree(ks2, 16, tree, counts)
    if "tree" in which:
        np.savez_compressed(os.path.join(HERE, "tree.npz"), x=x, y=y, z=z, h=h, keys=keys, sorted_keys=ks,
                            perm=perm,
                            leaves=tree, counts=counts, centers=c, sizes=s, neighbors=nb, neighbor_counts=nc,
                            x2=x2, keys2=keys2, perm2=perm2, leaves2=tree2, counts2=counts2,
                            converged2=np.array(conv2), **o)

    # ---- halos: Gaussian mixture, Morton-64 keys (centers decoded as Hilbert, as the reference does), bucket 32
    T = np.float64
    n = 30000
    x, y, z = inputs.gaussian_mixture(n, clusters=6, seed=104, dtype=T)
    h = np.full(n, 0.006, dtype=T)
    box = make_box(0.0, 1.0, T)
    keys = REF.compute_sfc_keys(x, y, z, box, np.uint64, MORTON)
    ks, perm = REF.sort_by_key(keys)
    tree, counts = REF.compute_octree(ks, 32)
    o = REF.build_octree(tree)
    c, s = REF.node_fp_centers(o["prefixes"], box, T)
    xs, ys, zs, hs = x[perm], y[perm], z[perm], h[perm]
    layout = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    L = counts.size
    NI = o["prefixes"].size - L
    init = c[o["leafToInternal"][NI:]]
    first, last = L // 3, (2 * L) // 3
    sc, ss = REF.compute_bounding_box(xs, ys, zs, hs, layout, first, last, T(2), init)
    flags = REF.find_halos(o["prefixes"], o["childOffsets"], o["parents"], c, s, tree, sc, ss, box, first, last)
    if "halos" in which:
        np.savez_compressed(os.path.join(HERE, "halos.npz"), x=xs, y=ys, z=zs, h=hs, leaves=tree, counts=counts,
                            layout=layout, centers=c, sizes=s, search_init=init, search_centers=sc,
                            search_sizes=ss, first=np.array(first), last=np.array(last), flags=flags, **o)


if __name__ == "__main__":
    main()
