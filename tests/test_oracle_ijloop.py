"""The oracle's ijloop restatement (example interaction) against the reference's own CPU ijloop
(traversal/ijloop/cpu_alwaystraverse.hpp through oracle/_ref) and against an independent per-operation fold in numpy
scalars of the same type."""

import numpy as np
import pytest

from helpers import cpu_pipeline
from oracle_py import HILBERT, MORTON
from oracle_py import make_box


def _case(lib, T, n=6000, bnd=(0, 0, 0), seed=5):
    from hipcstone import inputs

    x, y, z, r = inputs.plummer(n, 0.3, seed=seed, dtype=T)
    h = inputs.plummer_h(r, n, 0.3, 40, dtype=T)
    box = make_box((-1.0, -1.05, -1.0), (1.0, 1.0, 1.1), T, bnd)
    p = cpu_pipeline(lib, x, y, z, h, box, np.uint64, 32, HILBERT)
    m = (np.random.default_rng(seed).random(n) + 0.5).astype(T)
    return p, box, m


def _run(lib, p, box, m, first, last, ngmax, ext=1.0):
    tree = {k: p[k] for k in ("childOffsets", "parents", "internalToLeaf")}
    return lib.ijloop_example(p["x"], p["y"], p["z"], p["h"], m, first, last, box, tree, p["layout"], p["centers"],
                              p["sizes"], ngmax, ext)


@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("bnd", [(0, 0, 0), (1, 1, 1), (1, 0, 2)])
@pytest.mark.parametrize("ngmax,first,last", [(64, 0, 6000), (7, 100, 5000)])
def test_ijloop_example_matches_reference(oracle, ref, T, bnd, ngmax, first, last):
    p, box, m = _case(oracle, T, bnd=bnd)
    a = _run(oracle, p, box, m, first, last, ngmax)
    b = _run(ref, p, box, m, first, last, ngmax)
    for name, u, v in zip(("rho", "count", "dmin", "dmax", "fx"), a, b):
        assert np.array_equal(u, v), name
    assert a[1][first:last].max() > 2 and not a[1][:first].any() and not a[1][last:].any()


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_ijloop_example_fold_restated(oracle, T):
    """Per body: the neighbor list of findNeighbors, then the interaction folded step by step in T."""
    p, box, m = _case(oracle, T, n=3000, bnd=(1, 0, 0), seed=9)
    ngmax = 40
    rho, cnt, dmin, dmax, fx = _run(oracle, p, box, m, 0, 3000, ngmax)
    tree = {k: p[k] for k in ("childOffsets", "parents", "internalToLeaf")}
    nb, nc = oracle.find_neighbors(p["x"], p["y"], p["z"], p["h"], 0, 3000, box, tree, p["layout"], p["centers"],
                                   p["sizes"], ngmax)
    x, y, z, h = p["x"], p["y"], p["z"], p["h"]
    L, IL = box.len, box.ilen
    for i in list(range(0, 3000, 97)) + [2999]:
        hi = h[i]
        s2 = T(2) * hi
        inside = all(lo <= c - s2 and c + s2 <= up for c, lo, up in
                     ((x[i], box.lim[0], box.lim[1]), (y[i], box.lim[2], box.lim[3]), (z[i], box.lim[4], box.lim[5])))
        pbc = not inside
        r2 = T(4) * hi * hi
        u = T(1) - T(0) / r2
        srho, sn, smin, smax, sfx = m[i] * (u * u * u), 1, np.finfo(T).max, -np.finfo(T).max, m[i] * T(0)
        for j in nb[i, : min(nc[i], ngmax)]:
            d = [x[i] - x[j], y[i] - y[j], z[i] - z[j]]
            if pbc:
                for k in range(3):
                    d[k] = d[k] - T(box.bnd[k] == 1) * T(L[k]) * T(np.rint(d[k] * T(IL[k])))
            d2 = d[0] * d[0] + (d[1] * d[1] + d[2] * d[2])
            uu = T(1) - d2 / r2
            srho = srho + m[j] * (uu * uu * uu)
            sn += 1
            smin = d2 if d2 < smin else smin
            smax = d2 if smax < d2 else smax
            sfx = sfx + m[j] * d[0]
        assert rho[i] == T(srho / (hi * hi * hi)) and cnt[i] == sn, i
        assert dmin[i] == smin and dmax[i] == smax and fx[i] == sfx, i
        assert cnt[i] == 1 + min(nc[i], ngmax)
