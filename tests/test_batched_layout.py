"""CPU checks of the two constructions the batched pipeline relies on (csrc/batched.hip), restated here in Python on
real cornerstone leaf arrays from the oracle:

* the slot link: internal node P gets the unsorted-layout slot (tid + binaryKeyWeight(leaves[tid], level)) / 7 of the
  leaf pair that creates it (tree/octree.hpp:86-99), the root's slot is swapped with 0, children of P are nodes
  1 + 8 slot(P) + digit, each node finds its parent's slot by one binary search for the parent's child-4 boundary.
  Checked: the node indices are a bijection onto [0, N), children of every internal node are its 8 children in
  digit order, and a depth-first walk in child order visits the leaves in SFC order (what the neighbor walk needs);
* composite keys: the 64-bit composite leaf array of several clouds, split per cloud and shifted back, is each
  cloud's own 32-bit converged tree, and the per-cloud rebalance rule (bucket per cloud, deepest level 13, no merge
  at or above level 3) on the composite array equals each cloud's own csarray decision.
"""

import bisect

import numpy as np
import pytest

MAXL64 = 21


def node_range(level, maxl=MAXL64):
    return 1 << (3 * (maxl - level))


def tree_level(rng, maxl=MAXL64):
    lv = 0
    while node_range(lv, maxl) > rng:
        lv += 1
    assert node_range(lv, maxl) == rng
    return lv


def octal_digit(key, level, maxl=MAXL64):
    return (key >> (3 * (maxl - level))) & 7


def common_prefix(a, b, bits=63):
    x = a ^ b
    return bits if x == 0 else bits - x.bit_length()


def binary_key_weight(key, level):
    ret = 0
    for l in range(1, level + 2):
        d = octal_digit(key, l)
        ret += (7 - d) if d >= 4 else -d
    return ret


def slot_link(leaves):
    """The batched link restated: returns prefixes {node index: (key, level)}, child offsets and leaf indices."""
    leaves = [int(k) for k in leaves]
    L = len(leaves) - 1
    I = (L - 1) // 7

    def internal_slot(key, level):
        boundary = key + 4 * node_range(level + 1)
        tid = bisect.bisect_left(leaves, boundary) - 1
        return (tid + binary_key_weight(leaves[tid], level)) // 7

    root_slot = internal_slot(0, 0) if L > 1 else 0

    def remap(s):
        return 0 if s == root_slot else (root_slot if s == 0 else s)

    def node_index(key, level):
        if level == 0:
            return 0
        parent = key & ~(node_range(level - 1) - 1)
        return 1 + 8 * remap(internal_slot(parent, level - 1)) + octal_digit(key, level)

    nodes, child, leaf_of = {}, {}, {}
    for t in range(L):
        key = leaves[t]
        lv = tree_level(leaves[t + 1] - key)
        idx = 0 if L == 1 else node_index(key, lv)
        assert idx not in nodes
        nodes[idx] = (key, lv)
        child[idx] = 0
        leaf_of[idx] = t
        if t < L - 1:
            pl = common_prefix(key, leaves[t + 1])
            if pl % 3 == 0:
                lv2 = pl // 3
                o = (t + binary_key_weight(key, lv2)) // 7
                nkey = 0 if lv2 == 0 else key & ~(node_range(lv2) - 1)
                ix = node_index(nkey, lv2)
                assert ix not in nodes
                nodes[ix] = (nkey, lv2)
                child[ix] = 1 + 8 * remap(o)
    return nodes, child, leaf_of, L, I


@pytest.mark.parametrize("seed,bucket,n", [(1, 16, 3000), (2, 64, 20000), (3, 8, 500), (4, 128, 40000)])
def test_slot_link_is_a_consistent_octree(oracle, seed, bucket, n):
    rng = np.random.default_rng(seed)
    # clustered keys (so the tree has leaves on many levels)
    centers = rng.integers(0, 1 << 63, 5, dtype=np.uint64)
    keys = np.concatenate([(c + rng.integers(0, 1 << (63 - 3 * int(rng.integers(3, 12))), n // 5, dtype=np.uint64))
                           for c in centers])
    keys = np.sort(keys % np.uint64(1 << 63))
    tree, counts = oracle.compute_octree(keys, bucket)
    nodes, child, leaf_of, L, I = slot_link(tree)
    N = L + I
    assert sorted(nodes) == list(range(N))
    for idx, (key, lv) in nodes.items():
        if child[idx]:
            co = child[idx]
            assert co + 8 <= N
            for k in range(8):
                ck, cl = nodes[co + k]
                assert cl == lv + 1 and ck == key + k * node_range(lv + 1)
    # depth-first in child order from the root visits the leaves in SFC order
    order, stack = [], [0]
    while stack:
        i = stack.pop()
        if child[i] == 0:
            order.append(leaf_of[i])
        else:
            stack.extend(child[i] + k for k in reversed(range(8)))
    assert order == list(range(L))


def _cloud_tree(oracle, keys32, bucket):
    tree, counts = oracle.compute_octree(keys32, bucket)
    return tree, counts


def test_composite_tree_is_the_clouds_trees(oracle):
    rng = np.random.default_rng(9)
    clouds = [(rng.integers(0, 1 << 30, int(n), dtype=np.uint32), b)
              for n, b in ((5000, 16), (300, 64), (12000, 32), (40, 16), (9000, 128))]
    # composite keys; cloud slots past the last are empty level-3 leaves
    # the converged composite tree built cloud by cloud (what the batched full build constructs level by level)
    leaves = []
    for c in range(512):
        if c < len(clouds):
            k, b = clouds[c]
            t, _ = _cloud_tree(oracle, np.sort(k), b)
            leaves.extend(int((np.uint64(c) << np.uint64(54)) | (np.uint64(v) << np.uint64(24))) for v in t[:-1])
        else:
            leaves.append(c << 54)
    leaves.append(1 << 63)
    leaves = np.array(leaves, dtype=np.uint64)
    # it is a cornerstone array: its slot link is consistent
    nodes, child, leaf_of, L, I = slot_link(leaves)
    assert sorted(nodes) == list(range(L + I))
    # the cloud roots are the level-3 nodes
    for c in range(len(clouds)):
        roots = [i for i, (k, lv) in nodes.items() if lv == 3 and k == c << 54]
        assert len(roots) == 1
    # the per-cloud rebalance decision on the composite array equals each cloud's own (one warm step after moving
    # the keys): compare the counts-based decision through the oracle's update_octree on each cloud
    changed = 0
    for c, (k, b) in enumerate(clouds):
        k2 = np.sort((k.astype(np.int64) + rng.integers(-(1 << 26), 1 << 26, k.size)) .clip(0, (1 << 30) - 1)).astype(np.uint32)
        t, _ = _cloud_tree(oracle, np.sort(k), b)
        # counts of the moved keys on the old tree: the decision then splits and merges
        cnt = np.diff(np.searchsorted(k2, t.astype(np.int64))).astype(np.uint32)
        t2, cnt2, conv = oracle.update_octree(k2, b, t, cnt)
        # composite: shift, apply the per-cloud rule with deepest level 13
        ct = [(c << 54) + (int(v) << 24) for v in t]
        ops = composite_ops(ct, cnt, b)
        new = composite_rebalance(ct, ops)
        back = [((v >> 24) & ((1 << 30) - 1)) for v in new[:-1]] + [1 << 30]
        # the rebalanced shape equals the cloud's own (counts are recomputed from keys afterwards on both sides)
        assert back == [int(v) for v in t2]
        changed += int(len(back) != len(t))
    assert changed >= 2  # the moves made real rebalances


def composite_ops(tree, counts, bucket):
    """nodeOpsKernel of batched.hip restated for one cloud's part of the composite array (last entry: the end key)."""
    L = len(tree) - 1
    ops = []
    for i in range(L):
        key = tree[i]
        lv = tree_level(tree[i + 1] - key)
        op = 1
        merged = False
        if lv > 3:
            s = octal_digit(key, lv)
            if s > 0 and i - s + 8 <= L and tree[i - s + 8] == tree[i - s] + node_range(lv - 1):
                if sum(int(v) for v in counts[i - s:i - s + 8]) <= bucket:
                    op, merged = 0, True
        if not merged:
            cnt = int(counts[i])
            if cnt > bucket * 512 and lv + 3 < 13:
                op = 4096
            elif cnt > bucket * 64 and lv + 2 < 13:
                op = 512
            elif cnt > bucket * 8 and lv + 1 < 13:
                op = 64
            elif cnt > bucket and lv < 13:
                op = 8
        ops.append(op)
    return ops


def composite_rebalance(tree, ops):
    out = []
    for i, op in enumerate(ops):
        key = tree[i]
        lv = tree_level(tree[i + 1] - key)
        if op == 1:
            out.append(key)
        elif op >= 8:
            diff = {8: 1, 64: 2, 512: 3, 4096: 4}[op]
            out.extend(key + s * node_range(lv + diff) for s in range(op))
    out.append(tree[-1])
    return out
