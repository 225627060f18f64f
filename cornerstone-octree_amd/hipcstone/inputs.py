"""Deterministic synthetic particle sets for the pipeline's workloads (host side, numpy).

Every random stream is a counter-based splitmix64 sequence: value k of stream s under seed q is
splitmix64(splitmix64(q * 0x9E3779B97F4A7C15 + s) + k), mapped to [0, 1) through its top 53 bits. The same
generator feeds the GPU runs, the CPU oracle and the golden fixtures, so every party sees identical inputs.
"""

from __future__ import annotations

import numpy as np

_M = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n doubles in [0, 1) from stream `stream` of seed `seed`, starting at counter `offset`."""
    with np.errstate(over="ignore"):
        base = splitmix64(np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + np.uint64(stream))
        k = np.arange(offset, offset + n, dtype=np.uint64) + base
    return (splitmix64(k) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform_cube(n: int, seed: int = 42, lo=0.0, hi=1.0, dtype=np.float64):
    """x, y, z uniform in [lo, hi)^3."""
    span = hi - lo
    return tuple((lo + span * uniform(seed, s, n)).astype(dtype) for s in range(3))


def plummer(n: int, a: float, seed: int = 42, rmax: float = 1.0, center=(0.0, 0.0, 0.0), dtype=np.float32):
    """Plummer sphere of scale radius a, truncated at r <= rmax, isotropic directions.

    Radii by inversion of the cumulative mass M(r) = r^3/(r^2+a^2)^(3/2): r = a / sqrt(u^(-2/3) - 1); draws with
    r > rmax are rejected (the next counters of the same streams replace them, in order).
    """
    out_r = np.empty(0)
    out_k = np.empty(0, dtype=np.int64)
    offset = 0
    need = n
    while need > 0:
        m = int(need * 1.05) + 64
        u = uniform(seed, 0, m, offset)
        u = np.maximum(u, 1e-300)
        with np.errstate(divide="ignore", over="ignore"):
            r = a / np.sqrt(np.power(u, -2.0 / 3.0) - 1.0)
        ok = np.isfinite(r) & (r <= rmax)
        idx = np.nonzero(ok)[0][:need]
        out_r = np.concatenate([out_r, r[idx]])
        out_k = np.concatenate([out_k, idx + offset])
        need -= idx.size
        offset += m
    # directions from the counters of the accepted radii
    cz = 2.0 * uniform_at(seed, 1, out_k) - 1.0
    phi = 2.0 * np.pi * uniform_at(seed, 2, out_k)
    sz = np.sqrt(np.maximum(0.0, 1.0 - cz * cz))
    x = center[0] + out_r * sz * np.cos(phi)
    y = center[1] + out_r * sz * np.sin(phi)
    z = center[2] + out_r * cz
    return x.astype(dtype), y.astype(dtype), z.astype(dtype), out_r


def uniform_at(seed: int, stream: int, counters: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        base = splitmix64(np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + np.uint64(stream))
        k = counters.astype(np.uint64) + base
    return (splitmix64(k) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def plummer_h(r: np.ndarray, n_total: int, a: float, target: float = 100.0, dtype=np.float32):
    """Smoothing length from the analytic Plummer density so that the 2h sphere holds `target` particles:
    rho(r) = 3N/(4 pi a^3) (1 + r^2/a^2)^(-5/2), (4/3) pi (2h)^3 rho = target."""
    rho = 3.0 * n_total / (4.0 * np.pi * a**3) * np.power(1.0 + (r * r) / (a * a), -2.5)
    two_h = np.cbrt(target * 3.0 / (4.0 * np.pi * rho))
    return (0.5 * two_h).astype(dtype)


def uniform_h(n: int, target: float = 100.0):
    """h = 0.5 (target / (4/3 pi N))^(1/3) for N uniform points in the unit cube."""
    return 0.5 * (target / (4.0 / 3.0 * np.pi * n)) ** (1.0 / 3.0)


def gaussian_mixture(n: int, clusters: int = 64, background: float = 0.2, seed: int = 7, dtype=np.float64):
    """`clusters` isotropic Gaussians (sigma log-uniform in [0.005, 0.05], centres uniform in [0,1)^3) plus a
    uniform background fraction; points clipped to [0, 1)."""
    cs = np.stack([uniform(seed, 10 + d, clusters) for d in range(3)], axis=1)
    sig = np.exp(np.log(0.005) + (np.log(0.05) - np.log(0.005)) * uniform(seed, 13, clusters))
    u_bg = uniform(seed, 14, n)
    which = np.minimum((uniform(seed, 15, n) * clusters).astype(np.int64), clusters - 1)
    # Box-Muller normals
    u1 = np.maximum(uniform(seed, 16, n), 1e-300)
    u2 = uniform(seed, 17, n)
    u3 = np.maximum(uniform(seed, 18, n), 1e-300)
    u4 = uniform(seed, 19, n)
    r1 = np.sqrt(-2.0 * np.log(u1))
    r2 = np.sqrt(-2.0 * np.log(u3))
    g = np.stack([r1 * np.cos(2 * np.pi * u2), r1 * np.sin(2 * np.pi * u2), r2 * np.cos(2 * np.pi * u4)], axis=1)
    pts = cs[which] + sig[which, None] * g
    bg = u_bg < background
    if bg.any():
        pts[bg] = np.stack([uniform(seed, 20 + d, n)[bg] for d in range(3)], axis=1)
    pts = np.clip(pts, 0.0, np.nextafter(1.0, 0.0))
    return pts[:, 0].astype(dtype), pts[:, 1].astype(dtype), pts[:, 2].astype(dtype)


def cloud_mix(num_clouds: int = 512, nmin: int = 1000, nmax: int = 1_000_000, seed: int = 2026,
              target: float = 64.0, plummer_a: float = 0.1):
    """The batched workload: `num_clouds` independent fp32 clouds with sizes log-uniform in [nmin, nmax]; each is,
    50/50, uniform in the unit cube (box [0, 1]^3) or a Plummer sphere of scale radius `plummer_a` truncated at r <= 1
    (box [-1, 1]^3), with h for about `target` neighbors (uniform_h / plummer_h). Returns the sizes, the per-cloud
    boxes as (lo, hi) and kinds, and the concatenated x, y, z, h (cloud order)."""
    u = uniform(seed, 30, num_clouds)
    sizes = np.exp(np.log(nmin) + (np.log(nmax) - np.log(nmin)) * u).astype(np.int64)
    kinds = np.where(uniform(seed, 31, num_clouds) < 0.5, "uniform", "plummer")
    xs, ys, zs, hs, boxes = [], [], [], [], []
    for c in range(num_clouds):
        n = int(sizes[c])
        if kinds[c] == "uniform":
            x, y, z = uniform_cube(n, seed=seed * 1000 + c, dtype=np.float32)
            h = np.full(n, uniform_h(n, target), dtype=np.float32)
            boxes.append((0.0, 1.0))
        else:
            x, y, z, r = plummer(n, plummer_a, seed=seed * 1000 + c, dtype=np.float32)
            h = plummer_h(r, n, plummer_a, target, dtype=np.float32)
            boxes.append((-1.0, 1.0))
        xs.append(x), ys.append(y), zs.append(z), hs.append(h)
    return sizes, boxes, kinds, np.concatenate(xs), np.concatenate(ys), np.concatenate(zs), np.concatenate(hs)


def served_requests(duration_s: float = 5.0, rate: float = 200.0, seed: int = 5150, nmin: int = 10_000,
                    nmax: int = 4_000_000, build_frac: float = 0.3, a: float = 0.05, target: float = 100.0,
                    pool: int = 6):
    """The served workload: Poisson arrivals at `rate` requests/s over `duration_s`; sizes log-uniform in
    [nmin, nmax]; a fraction `build_frac` are full builds of a new cloud (fp32 Plummer a, truncated at r <= 1, h for
    `target` neighbors), the rest updates of an earlier cloud (chosen uniformly among the clouds built so far). A
    build of n particles takes the first n of one of `pool` Plummer samples of nmax particles (i.i.d. draws, so any
    prefix is itself a Plummer sample) with its own h for n. Returns the request dicts (host arrays kept in them)."""
    gap = -np.log(1.0 - uniform(seed, 50, 1 << 16)) / rate
    arrivals = np.cumsum(gap)
    arrivals = arrivals[arrivals < duration_s]
    m = arrivals.size
    kind_u = uniform(seed, 51, m)
    size_u = uniform(seed, 52, m)
    pick_u = uniform(seed, 53, m)
    seeds = (uniform(seed, 54, m) * 2**52).astype(np.uint64)
    samples = [plummer(nmax, a, seed=seed * 31 + p, dtype=np.float32) for p in range(pool)]
    reqs, clouds = [], []
    for i in range(m):
        build = kind_u[i] < build_frac or not clouds
        if build:
            n = int(np.exp(np.log(nmin) + (np.log(nmax) - np.log(nmin)) * size_u[i]))
            x, y, z, r = samples[len(clouds) % pool]
            h = plummer_h(r[:n], n, a, target, dtype=np.float32)
            cloud = {"id": len(clouds), "n": n, "x": x[:n], "y": y[:n], "z": z[:n], "h": h}
            clouds.append(cloud)
            reqs.append({"id": i, "cloud": cloud["id"], "kind": 0, "n": n, "x": cloud["x"], "y": cloud["y"],
                         "z": cloud["z"], "h": h, "arrival_s": float(arrivals[i]), "seed": int(seeds[i])})
        else:
            cloud = clouds[min(int(pick_u[i] * len(clouds)), len(clouds) - 1)]
            reqs.append({"id": i, "cloud": cloud["id"], "kind": 1, "n": cloud["n"], "x": None, "y": None, "z": None,
                         "h": None, "arrival_s": float(arrivals[i]), "seed": int(seeds[i])})
    return reqs
