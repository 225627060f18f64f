"""Served clouds (hipcstone_serve_run): continuous batching over resident clouds.

* Each request's result (digest of its keys, neighbor counts and stored neighbor rows) is the same when the
  requests run one at a time (one worker) and in the mix (four workers, every request released at once so that
  admitted requests overlap on the device), and equals a plain pipeline run by hand on the same cloud.
* With a small HBM budget the loop evicts idle clouds (their particles are kept on the host) and brings them back:
  every request completes and a cloud's requests keep their order.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(seed, duration=0.25, nmax=120_000, rate=200.0):
    from hipcstone import inputs

    return inputs.served_requests(duration_s=duration, rate=rate, seed=seed, nmin=5_000, nmax=nmax, pool=2)


def test_serve_results_same_alone_and_in_the_mix(hip):
    reqs = _requests(11)
    assert sum(r["kind"] == 1 for r in reqs) > 10 and sum(r["kind"] == 0 for r in reqs) > 5
    print(f"serve: {len(reqs)} requests, alone (1 slot)", flush=True)
    alone, _ = hip.serve_run(reqs, slots=1, realtime=False, digests=True)
    print("serve: mixed (4 slots, all released at once)", flush=True)
    mixed, st = hip.serve_run(reqs, slots=4, realtime=False, digests=True)
    assert all(r["status"] in (hip.SERVE_BUILT, hip.SERVE_UPDATED) for r in mixed)
    assert [r["digest"] for r in alone] == [r["digest"] for r in mixed]
    assert [r["num_leaves"] for r in alone] == [r["num_leaves"] for r in mixed]
    # the mix really overlapped: some request started before an earlier-started one ended, on another worker
    spans = sorted((r["start_s"], r["end_s"], r["slot"]) for r in mixed)
    assert any(spans[i + 1][0] < spans[i][1] and spans[i + 1][2] != spans[i][2] for i in range(len(spans) - 1))
    assert len({r["slot"] for r in mixed}) > 1
    assert st["evictions"] == 0


def test_serve_build_digest_equals_a_plain_pipeline(hip):
    reqs = _requests(12, duration=0.05)
    first = reqs[0]
    assert first["kind"] == 0
    res, _ = hip.serve_run([first], slots=1, realtime=False, digests=True)
    # the same build through a plain pipeline, digested on the host the way the server digests on the device
    n, ngmax = first["n"], 128
    pipe = hip.Pipeline(curve=hip.HILBERT, key_bits=64, real_bits=32, bucket_size=64, ngmax=ngmax)
    d = [hip.DeviceArray.from_numpy(np.ascontiguousarray(first[k])) for k in "xyzh"]
    nb = hip.DeviceArray(n * ngmax, np.uint32)
    nc = hip.DeviceArray(n, np.uint32)
    pipe.sync(*d, n, hip.make_box(-1.0, 1.0, np.float32), full_build=True, neighbors=nb, counts=nc)
    keys = pipe.arrays()["keys"]
    assert res[0]["digest"] == host_digest(keys, nc.download(), nb.download().reshape(n, ngmax), ngmax)


M64 = (1 << 64) - 1


def mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def mix64_np(z):
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def host_digest(keys, counts, nb, ngmax):
    """The server's digest restated: sum over i of mix64(i * phi ^ key_i) + mix64(i << 32 | count_i) + sum over the
    first min(count_i, ngmax) stored neighbors k of mix64(nb << 20 ^ k << 52 ^ i), wrapping at 2^64."""
    with np.errstate(over="ignore"):
        n = keys.size
        i = np.arange(n, dtype=np.uint64)
        d = mix64_np((i * np.uint64(0x9E3779B97F4A7C15)) ^ keys) + mix64_np((i << np.uint64(32)) | counts.astype(np.uint64))
        m = np.minimum(counts, ngmax)
        cols = np.arange(ngmax, dtype=np.uint64)[None, :]
        mask = np.arange(ngmax)[None, :] < m[:, None]
        t = mix64_np((nb.astype(np.uint64) << np.uint64(20)) ^ (cols << np.uint64(52)) ^ i[:, None])
        d = d + np.where(mask, t, np.uint64(0)).sum(axis=1, dtype=np.uint64)
        return int(d.sum(dtype=np.uint64))


def test_serve_evicts_under_a_small_budget(hip):
    reqs = _requests(13, duration=0.3, nmax=60_000)
    budget = 160 << 20
    print(f"serve: {len(reqs)} requests under a {budget >> 20} MiB budget", flush=True)
    # released at their arrival times: an idle cloud with no queued request gets evicted for a later build, and a
    # later update of it brings it back from its spilled particles
    res, st = hip.serve_run(reqs, hbm_budget=budget, slots=3, realtime=True, digests=True)
    assert st["evictions"] > 0
    assert all(r["status"] != hip.SERVE_REJECTED for r in res)
    assert any(r["status"] == hip.SERVE_REBUILT for r in res)
    # per cloud, requests ran in arrival order
    by_cloud = {}
    for q, r in zip(reqs, res):
        by_cloud.setdefault(q["cloud"], []).append((q["arrival_s"], r["start_s"], r["end_s"]))
    for runs in by_cloud.values():
        runs.sort()
        assert all(runs[i][2] <= runs[i + 1][1] for i in range(len(runs) - 1))


def test_serve_rejects_bad_config(hip):
    with pytest.raises(hip.HipcstoneError):
        hip.serve_run([], slots=0)
