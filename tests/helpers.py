"""Shared test helpers: small tree builders in the style of the reference's test utilities, brute-force checkers,
and a CPU pipeline assembled from the oracle."""

import numpy as np

MAXL = {np.dtype(np.uint32): 10, np.dtype(np.uint64): 21}


def node_range(kd, level):
    return kd(1) << kd(3 * (MAXL[np.dtype(kd)] - level))


class OctreeMaker:
    """Cornerstone leaf arrays built by dividing nodes, like tree/cs_util.hpp:137-197 OctreeMaker."""

    def __init__(self, kd):
        self.kd = kd
        self.nodes = [(0, 0)]  # (key, level) of leaves in order

    def divide(self, *path):
        """Divide the node reached by the octant path (root when empty)."""
        kd = self.kd
        key = 0
        for lvl, d in enumerate(path, start=1):
            key += d * int(node_range(kd, lvl))
        level = len(path)
        idx = self.nodes.index((key, level))
        children = [(key + c * int(node_range(kd, level + 1)), level + 1) for c in range(8)]
        self.nodes[idx:idx + 1] = children
        return self

    def make_tree(self):
        keys = [k for k, _ in self.nodes] + [int(node_range(self.kd, 0))]
        return np.array(keys, dtype=self.kd)


def uniform_n_level_tree(kd, level):
    """All 8^level nodes of one level plus the end key (cs_util.hpp:97-105)."""
    r = int(node_range(kd, level))
    return np.array([i * r for i in range(8 ** level)] + [int(node_range(kd, 0))], dtype=kd)


def brute_neighbors(x, y, z, h, ngmax):
    """all2allNeighbors + sortNeighbors (test/unit/neighbors/all_to_all.hpp:26-65) for an open box."""
    n = x.size
    T = x.dtype.type
    nb = np.zeros((n, ngmax), dtype=np.uint32)
    nc = np.zeros(n, dtype=np.uint32)
    for i in range(n):
        dx = x - x[i]
        dy = y - y[i]
        dz = z - z[i]
        d2 = dx * dx + dy * dy + dz * dz
        r = T(2) * h[i]
        hits = np.nonzero(d2 < r * r)[0]
        hits = hits[hits != i]
        nc[i] = hits.size
        nb[i, : min(ngmax, hits.size)] = hits[:ngmax]
    return nb, nc


def cpu_pipeline(lib, x, y, z, h, box, kd, bucket, curve=1, ngmax=None):
    """keys -> stable sort -> gather -> computeOctree -> link -> centers -> layout (-> neighbors) on a CPU lib."""
    keys = lib.compute_sfc_keys(x, y, z, box, kd, curve)
    ks, perm = lib.sort_by_key(keys)
    xs, ys, zs, hs = x[perm], y[perm], z[perm], h[perm]
    tree, counts = lib.compute_octree(ks, bucket)
    o = lib.build_octree(tree)
    c, s = lib.node_fp_centers(o["prefixes"], box, x.dtype)
    layout = np.zeros(counts.size + 1, dtype=np.uint32)
    layout[1:] = np.cumsum(counts, dtype=np.uint64).astype(np.uint32)
    out = dict(keys=ks, perm=perm, x=xs, y=ys, z=zs, h=hs, leaves=tree, counts=counts, centers=c, sizes=s,
               layout=layout, **o)
    if ngmax is not None:
        nb, nc = lib.find_neighbors(xs, ys, zs, hs, 0, x.size, box, o, layout, c, s, ngmax)
        out["nb"], out["nc"] = nb, nc
    return out


def neighbor_rows_equal(nb1, nc1, nb2, nc2, ngmax):
    """Counts equal and the first min(count, ngmax) entries of every row equal."""
    if not np.array_equal(nc1, nc2):
        return False
    m = np.minimum(nc1, ngmax)
    cols = np.arange(ngmax)[None, :]
    mask = cols < m[:, None]
    return np.array_equal(np.where(mask, nb1, 0), np.where(mask, nb2, 0))


# ---- the moving input of the one-rank Domain checks (reference Domain, CPU oracle, GPU, golden fixtures)

# the one-rank Domain cases of tests/golden/domain.npz: (name, key dtype, real dtype, bucket, bucket focus,
# boundary, theta, particles, calls, syncGrav, input seed)
DOMAIN_GOLDEN_CASES = [
    ("a", np.uint64, np.float32, 64, 16, (0, 0, 0), 0.5, 8000, 5, False, 201),
    ("b", np.uint32, np.float64, 32, 8, (1, 2, 3), 0.5, 6000, 5, False, 202),
    ("g", np.uint64, np.float64, 32, 8, (0, 0, 0), 0.7, 6000, 3, True, 203),
]


def domain_input(n, T, seed, with_mass=False):
    """Plummer particles (a = 0.1) with h for about 60 neighbors, and optionally masses in [-0.4, 1.6) with 3 % zeros."""
    from hipcstone import inputs

    x, y, z, r = inputs.plummer(n, 0.1, seed=seed, dtype=T)
    h = inputs.plummer_h(r, n, 0.1, 60, dtype=T)
    if not with_mass:
        return [x, y, z, h]
    rng = np.random.default_rng(seed + 1)
    m = (rng.random(n) * 2 - 0.4).astype(T)
    m[rng.random(n) < 0.03] = 0
    return [x, y, z, h, m]


def domain_move(rng, out, step, carried=("h",), lo=-1.0, hi=1.0):
    """The next step's input from a sync's output: x, y, z displaced, the `carried` arrays kept; a shuffled order
    after step 1, a large displacement (a rearranged tree) after step 2, and after step 3 a contraction by 20 %,
    which the box may follow by only 5 % per side and call (limitBoxShrinking, sfc/box.hpp:398-415)."""
    T = out["x"].dtype.type
    n = out["x"].size
    sigma = 0.05 if step == 2 else 0.005
    scale = 0.8 if step == 3 else 1.0
    nxt = [np.clip(out[k] * T(scale) + rng.normal(0, sigma, n), lo, hi).astype(T) for k in "xyz"]
    nxt += [np.array(out[k], copy=True) for k in carried]
    if step == 1:
        p = rng.permutation(n)
        nxt = [a[p] for a in nxt]
    return nxt
