"""Pins the oracle's one-rank Domain sequence (oracle_py.CpuDomain) against a run of the reference's own
Domain<KeyType, T, Cpu> (domain/domain.hpp:63-86 ctor, :169-218 sync, :220-287 syncGrav, :340-392 accessors), built
from the reference headers into oracle/_ref/libcstone_ref_domain.so and run as an MPI singleton.

Every step compares the box, the sorted keys, the particle order and the reordered x, y, z, h, the global leaf array
and its counts (GlobalAssignment, assignment.hpp:92-144), the focus leaves, leaf and node counts (FocusedOctree,
octree_focus_mpi.hpp:100-252, :585-603), the particle layout and assigned range, the linked focus tree with its
geometric centers, and for syncGrav the reordered masses and the expansion centers. The inputs move between steps,
one step hands in a shuffled order and one a large displacement (several global update steps).
"""

import numpy as np
import pytest

from helpers import domain_input, domain_move
from oracle_py import CpuDomain, RefDomain

pytestmark = pytest.mark.skipif(not RefDomain.available(),
                                reason="oracle/_ref/libcstone_ref_domain.so not built (needs the reference and MPI)")

CASES = [(np.uint64, np.float32, 64, 16), (np.uint64, np.float64, 128, 32), (np.uint32, np.float32, 64, 64),
         (np.uint32, np.float64, 32, 8)]


def compare(oracle, a, b, T, kd, n, box):
    for k in ("keys", "perm", "x", "y", "z", "h", "gtree", "gcounts", "fleaves", "fcounts", "fnodecounts"):
        assert a[k].dtype == b[k].dtype and np.array_equal(a[k], b[k]), k
    assert a["lo"] == b["lo"] and a["hi"] == b["hi"]
    assert (b["start"], b["end"], b["with_halos"]) == (0, n, n)
    assert np.array_equal(b["layout"], np.concatenate([[0], np.cumsum(a["fcounts"])]).astype(np.uint32))
    t = oracle.build_octree(a["fleaves"])
    for k in ("prefixes", "childOffsets", "internalToLeaf"):
        assert np.array_equal(t[k], b[k]), k
    assert np.array_equal(t["levelRange"], b["levelRange"][: t["levelRange"].size])
    c, s = oracle.node_fp_centers(t["prefixes"], box, T)
    assert np.array_equal(c.ravel(), b["centers"]) and np.array_equal(s.ravel(), b["sizes"])


@pytest.mark.parametrize("kd,T,bucket,bucket_focus", CASES)
@pytest.mark.parametrize("boundary", [(0, 0, 0), (1, 2, 3)])
def test_cpu_domain_matches_reference_domain(oracle, kd, T, bucket, bucket_focus, boundary):
    n = 40000
    x, y, z, h = domain_input(n, T, seed=41)
    cpu = CpuDomain(oracle, kd, T, bucket, bucket_focus, -1.0, 1.0, boundary)
    ref = RefDomain(kd, T, bucket, bucket_focus, -1.0, 1.0, boundary)
    rng = np.random.default_rng(42)
    gsteps = []
    for step in range(6):
        a = cpu.sync(x, y, z, h)
        b = ref.sync(x, y, z, h)
        compare(oracle, a, b, T, kd, n, cpu.box())
        # invariants of every step: all particles counted once, keys sorted and equal to the recomputed keys
        assert int(a["fcounts"].sum()) == n and int(a["gcounts"].sum()) == n
        assert np.all(np.diff(a["keys"].astype(np.float64)) >= 0)
        assert np.array_equal(oracle.compute_sfc_keys(a["x"], a["y"], a["z"], cpu.box(), kd), a["keys"])
        if step == 0:  # the first call converges both trees
            assert a["fcounts"].max() <= bucket_focus and a["gcounts"].max() <= bucket
        gsteps.append(a["gsteps"])
        x, y, z, h = domain_move(rng, a, step)
    assert max(gsteps[3:]) >= 1


@pytest.mark.parametrize("kd,T,bucket,bucket_focus,theta", [(np.uint64, np.float32, 64, 16, 0.5),
                                                            (np.uint64, np.float64, 32, 8, 0.7),
                                                            (np.uint32, np.float32, 64, 32, 0.6)])
def test_cpu_domain_sync_grav_matches_reference_domain(oracle, kd, T, bucket, bucket_focus, theta):
    n = 30000
    x, y, z, h, m = domain_input(n, T, seed=43, with_mass=True)
    rng = np.random.default_rng(44)
    cpu = CpuDomain(oracle, kd, T, bucket, bucket_focus, -1.0, 1.0)
    ref = RefDomain(kd, T, bucket, bucket_focus, -1.0, 1.0, theta=theta)
    for step in range(3):
        a = cpu.sync_grav(x, y, z, h, m, theta)
        b = ref.sync_grav(x, y, z, h, m)
        compare(oracle, a, b, T, kd, n, cpu.box())
        assert np.array_equal(a["m"], b["m"])
        assert np.array_equal(a["expansion_centers"].view(np.uint8).ravel(),
                              b["expansion_centers"].view(np.uint8).ravel())
        x, y, z, h, m = domain_move(rng, a, step, carried=("h", "m"))
