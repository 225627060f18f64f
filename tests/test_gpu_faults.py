"""Kernel guards report through the device's fault word instead of returning silently wrong output with OK.

Each case forces its guard from the host while every kernel still ends normally: the radix sort's look-back is given
a poll limit of one (a tile gives up at its second poll of a predecessor that has not published yet; the wait is
shortened, never lengthened), and the neighbor walk gets trees whose indices all lie inside the arrays but that violate the
linked-octree rules (a child range at or before its parent, or a DAG whose paths multiply past the step budget).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dev(hip, a):
    return hip.DeviceArray.from_numpy(np.ascontiguousarray(a))


def test_sort_lookback_give_up_is_reported(hip):
    n = 1 << 24
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    print("sort with look-back poll limit 1", flush=True)
    hip.set_sort_lookback_limit(1)  # the smallest limit (0 restores the default)
    try:
        dk = dev(hip, keys)
        dv = dev(hip, np.arange(n, dtype=np.uint32))
        hip.sort_by_key(dk, dv)
        faults = hip.device_faults()
    finally:
        hip.set_sort_lookback_limit(0)
    assert faults & hip.FAULT_SORT_LOOKBACK
    # with the default limit the same sort is correct and raises nothing
    dk = dev(hip, keys)
    dv = dev(hip, np.arange(n, dtype=np.uint32))
    hip.sort_by_key(dk, dv)
    assert hip.device_faults() == 0
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(dv.download(), order.astype(np.uint32))


def _flat_tree(T, levels):
    """A 'tree' of 1 + 8 * levels nodes in which all 8 nodes of level k point to the same 8 nodes of level k + 1
    (so the number of root-to-leaf paths is 8^levels), every node box covering the whole unit cube, the last level
    leaves 0..7 of 4 particles each. All indices lie inside the arrays."""
    N = 1 + 8 * levels
    co = np.zeros(N, np.int32)
    co[0] = 1
    for k in range(1, levels):
        co[1 + 8 * (k - 1): 1 + 8 * k] = 1 + 8 * k
    i2l = np.full(N, -1, np.int32)
    i2l[N - 8:] = np.arange(8, dtype=np.int32)
    centers = np.full(3 * N, 0.5, T)
    sizes = np.full(3 * N, 0.5, T)
    layout = (np.arange(9, dtype=np.uint32) * 4)
    return N, co, i2l, centers, sizes, layout


@pytest.mark.parametrize("case", ["budget", "backward_child"])
def test_neighbor_walk_guards_are_reported(hip, case):
    T = np.float32
    levels = 8
    N, co, i2l, centers, sizes, layout = _flat_tree(T, levels)
    expect = hip.FAULT_WALK_BUDGET
    if case == "backward_child":
        co[1] = 1  # node 1's children would start at node 1 itself
        expect = hip.FAULT_MALFORMED_TREE
    n = 32
    rng = np.random.default_rng(6)
    x, y, z = (rng.random(n).astype(T) for _ in range(3))
    h = np.full(n, 0.2, T)
    box = hip.make_box(0.0, 1.0, T)
    ngmax = 16
    nb = hip.DeviceArray(n * ngmax, np.uint32)
    nc = hip.DeviceArray(n, np.uint32)
    parents = np.zeros(N, np.int32)
    print(f"neighbor walk on a malformed tree: {case}", flush=True)
    hip.find_neighbors(dev(hip, x), dev(hip, y), dev(hip, z), dev(hip, h), 0, n, box, dev(hip, co), dev(hip, parents),
                       dev(hip, i2l), dev(hip, layout), dev(hip, centers), dev(hip, sizes), N, ngmax, nb, nc)
    assert hip.device_faults() & expect
    assert hip.device_faults() == 0  # read and cleared
