"""The tree update of the syncs after a Domain's first (oracle updateNsTree) -- CPU tests.

Pinned three ways:
* bit for bit against the reference's own single-rank focused octree (cstone::FocusedOctreeSingleNode,
  compiled from the reference headers into oracle/_ref) over a sequence of syncs of a moving flow that
  splits and merges leaves;
* the fixed point: updating repeatedly on fixed keys ends at the converged cornerstone tree of those
  keys (computeOctree, itself pinned against the reference), whatever tree it starts from; an update
  reports no change exactly when its decision (on the previous counts) keeps every node;
* invariants of one update: the leaves stay a cornerstone array (every leaf an aligned octree node),
  only leaves split (in 8) or sibling groups merge, and the counts are the brute-force counts of the
  new keys.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sphexa_amd import abi

BOX = abi.Box.make([-0.5] * 3, [0.5] * 3, abi.PERIODIC)


def sorted_keys(lib, pts):
    return np.sort(O.sfc_keys(lib, BOX, *[np.ascontiguousarray(c) for c in pts]))


def collapse(pts, step):
    """a collapsing, rotating periodic flow: every call pulls particles towards the centre and turns them"""
    r2 = (pts ** 2).sum(0)
    pts = pts * (1 - 0.25 * np.exp(-r2 / 0.05))[None, :]
    c, s = np.cos(0.3 + 0.1 * step), np.sin(0.3 + 0.1 * step)
    pts = np.stack([pts[0] * c - pts[1] * s, pts[0] * s + pts[1] * c, pts[2]])
    return (pts + 0.5) % 1.0 - 0.5


def expand(pts, step):
    """the reverse: particles pushed out of the centre (sibling groups empty and merge)"""
    r2 = (pts ** 2).sum(0)
    pts = pts * (1 + 0.3 * np.exp(-r2 / 0.02))[None, :]
    return (pts + 0.5) % 1.0 - 0.5


def check_cornerstone(leaves):
    d = np.diff(leaves.astype(object))
    for k, r in zip(leaves[:-1].astype(object), d):
        lvl = (r.bit_length() - 1) // 3
        assert r == 1 << (3 * lvl) and k % r == 0, (k, r)
    assert leaves[0] == 0 and int(leaves[-1]) == 1 << 63


def brute_counts(leaves, keys):
    return np.diff(np.searchsorted(keys, leaves, side="left")).astype(np.uint32)


@pytest.mark.parametrize("flow", [collapse, expand])
def test_tree_update_matches_reference_focus_tree(orc, ref, flow):
    rng = np.random.default_rng(3)
    pts = rng.uniform(-0.5, 0.5, (3, 20000))
    if flow is expand:
        for s in range(6):
            pts = collapse(pts, s)  # start clustered, then spread
    ft = O.FocusTree(ref, 64)
    t = None
    changed = 0
    for step in range(8):
        k = sorted_keys(orc, pts)
        lr, cr = ft.sync(k, BOX)
        if t is None:
            t = O.Tree(orc, k, BOX, 64)
        else:
            before = t.leaves.copy()
            t.update(k, BOX, 64)
            changed += int(not np.array_equal(before, t.leaves))
        np.testing.assert_array_equal(t.leaves, lr, err_msg=f"step {step}")
        np.testing.assert_array_equal(t.counts, cr, err_msg=f"step {step}")
        pts = flow(pts, step)
    assert changed >= 5, "the flow should change the tree at most syncs"


def test_tree_update_converges_to_the_cornerstone_tree(orc):
    rng = np.random.default_rng(5)
    a = rng.uniform(-0.5, 0.5, (3, 30000))
    b = rng.normal(0, 0.12, (3, 30000)).clip(-0.49, 0.49)  # another distribution: many splits and merges
    ka, kb = sorted_keys(orc, a), sorted_keys(orc, b)
    t = O.Tree(orc, ka, BOX, 64)
    target = O.Tree(orc, kb, BOX, 64)
    # the first update decides on the counts of ka, on which the tree is converged: it changes nothing
    # and reports so (as the reference's does), then counts kb
    assert t.update(kb, BOX, 64)
    assert not np.array_equal(t.leaves, target.leaves)
    updates = 0
    while True:
        conv = t.update(kb, BOX, 64)
        updates += 1
        same = np.array_equal(t.leaves, target.leaves)
        if conv:
            assert same
            break
        assert updates < 40
    assert updates >= 3  # a leaf of 1328 particles needs two rounds of splits, then one round that changes nothing
    np.testing.assert_array_equal(t.counts, target.counts)
    np.testing.assert_array_equal(t.prefixes, target.prefixes)
    np.testing.assert_array_equal(t.childOffsets, target.childOffsets)
    np.testing.assert_array_equal(t.centers, target.centers)
    # a converged tree with unchanged keys stays as it is
    assert t.update(kb, BOX, 64)
    np.testing.assert_array_equal(t.leaves, target.leaves)


def test_tree_update_invariants(orc):
    rng = np.random.default_rng(8)
    pts = rng.uniform(-0.5, 0.5, (3, 25000))
    k0 = sorted_keys(orc, pts)
    t = O.Tree(orc, k0, BOX, 64)
    for step in range(4):
        pts = collapse(pts, step)
        k = sorted_keys(orc, pts)
        old, old_counts = t.leaves.copy(), t.counts.copy()
        t.update(k, BOX, 64)
        check_cornerstone(t.leaves)
        np.testing.assert_array_equal(t.counts, brute_counts(t.leaves, k))
        np.testing.assert_array_equal(t.layout[1:], np.cumsum(t.counts))
        # every old leaf of more than 64 particles split into its 8 children, and only those
        for i in np.nonzero(old_counts > 64)[0]:
            r = (int(old[i + 1]) - int(old[i])) // 8
            kids = [int(old[i]) + c * r for c in range(8)]
            pos = np.searchsorted(t.leaves, np.uint64(kids[0]))
            assert [int(v) for v in t.leaves[pos:pos + 8]] == kids
        # a leaf of at most 64 either stays or lies inside a merged ancestor that held at most 64 old
        # particles in all
        old_layout = np.concatenate([[0], np.cumsum(old_counts)])
        for i in np.nonzero(old_counts <= 64)[0]:
            s, e = int(old[i]), int(old[i + 1])
            j = int(np.searchsorted(t.leaves, np.uint64(s), side="right")) - 1
            ns, ne = int(t.leaves[j]), int(t.leaves[j + 1])
            assert ns <= s and e <= ne
            if (ns, ne) != (s, e):
                a, b = np.searchsorted(old, [np.uint64(ns), np.uint64(ne)])
                assert old_layout[b] - old_layout[a] <= 64
        assert t.numLeafNodes == len(t.leaves) - 1
