"""The neighbour-list rule of the device octree builder (csrc/octree_gpu.hpp), restated on the
CPU and checked against the reference's own tree and the host builder (no GPU).

The host builder relinks neighbours one subdivision at a time (OctTreeNode::addNeighbors,
OctTreeNode.cpp:46-136); the device rebuilds every list once per level from the previous
level's lists:
  * a leaf keeps its unsubdivided neighbours in order, then, for each neighbour subdivided in
    the level in increasing id, that neighbour's children (in child order) that reach it;
  * a new child has only its sibling on a wall inside the parent, and on a wall on the
    parent's face the parent's neighbours that reach its quadrant, by the same rule.
The lists are then sorted by decreasing overlap with a stable sort, which is what std::sort
does for ranges of at most 16 entries (TreeNode::sortNeighbors, TreeNode.cpp:200-207).  If
this statement of the rule is right, it reproduces the reference tree's sorted lists and the
host builder's exactly; a variant that puts a subdivided neighbour's children in its place in
the list instead of at the end does not."""

import numpy as np

from helpers import PC, child_first, golden, load_pkg


def _span(axis):
    return {0: (1, 2), 1: (0, 2), 2: (0, 1)}[axis]


def _quad(tb, S, bu, bw, u, w):
    tu = tb[u] <= S[u] if bu == 0 else tb[u + 3] >= S[u]
    tw = tb[w] <= S[w] if bw == 0 else tb[w + 3] >= S[w]
    return tu and tw


def _overlap(base, nb, wall):
    u = 1 if wall < 2 else 0
    w = 2 if wall < 4 else 1
    du = min(base[u + 3], nb[u + 3]) - max(base[u], nb[u])
    dw = min(base[w + 3], nb[w + 3]) - max(base[w], nb[w])
    return max(du, 0.0) * max(dw, 0.0)


def level_rule(box, divide, in_place=False):
    """Nodes, first children and per-wall lists of the tree whose node l (creation order) is
    subdivided when divide[l], built level by level with the device's list rule (with
    in_place, a subdivided neighbour's children take its place in the list instead)."""
    B = [np.array(box, float)]
    F = [-1]
    P = [-1]
    lists = {0: [[] for _ in range(6)]}
    lbeg, lend = 0, 1
    while lend > lbeg:
        if not any(divide[i] for i in range(lbeg, lend)):
            break
        nxt = lend
        for i in range(lbeg, lend):
            if not divide[i]:
                continue
            b = B[i]
            xc, yc, zc = 0.5 * (b[0] + b[3]), 0.5 * (b[1] + b[4]), 0.5 * (b[2] + b[5])
            F[i] = nxt
            for k in range(8):
                B.append(np.array([xc if k & 1 else b[0], yc if k & 2 else b[1], zc if k & 4 else b[2],
                                   b[3] if k & 1 else xc, b[4] if k & 2 else yc, b[5] if k & 4 else zc]))
                F.append(-1)
                P.append(i)
            nxt += 8

        def children(tb, e, wall):
            axis, side = wall // 2, wall % 2
            u, w = _span(axis)
            S = B[F[e]][3:6]
            return [F[e] + k for k in range(8)
                    if ((k >> axis) & 1) == 1 - side and _quad(tb, S, (k >> u) & 1, (k >> w) & 1, u, w)]

        new = {}
        for n in range(nxt):
            L = [[] for _ in range(6)]
            if F[n] >= 0:
                new[n] = L
                continue
            for wall in range(6):
                axis, side = wall // 2, wall % 2
                if n < lend:
                    src = lists[n][wall]
                else:
                    p, k = P[n], (n - lend) % 8
                    if ((k >> axis) & 1) != side:
                        L[wall] = [n + ((k ^ (1 << axis)) - k)]
                        continue
                    u, w = _span(axis)
                    Sp = B[F[p]][3:6]
                    src = [e for e in lists[p][wall] if _quad(B[e], Sp, (k >> u) & 1, (k >> w) & 1, u, w)]
                if in_place:
                    L[wall] = [c for e in src for c in ([e] if F[e] < 0 else children(B[n], e, wall))]
                    continue
                L[wall] = [e for e in src if F[e] < 0]
                for e in sorted(e for e in src if F[e] >= 0):
                    L[wall] += children(B[n], e, wall)
            new[n] = L
        lists = new
        lbeg, lend = lend, nxt
    # sorted as TreeNode::sortNeighbors (stable for the list lengths met here)
    for n, L in lists.items():
        for wall in range(6):
            L[wall] = sorted(L[wall], key=lambda e: -_overlap(B[n], B[e], wall))
    return np.array(B), np.array(F), lists


def _check(box, divide, bounds, cf, nb_off, nb_id, in_place=False):
    B, F, lists = level_rule(box, divide, in_place)
    assert np.array_equal(B, bounds) and np.array_equal(F, cf)
    bad = 0
    for n in np.nonzero(cf < 0)[0]:
        for wall in range(6):
            q = 6 * n + wall
            want = list(nb_id[nb_off[q]:nb_off[q + 1]])
            assert len(want) <= 16
            bad += lists[n][wall] != want
    return bad


def test_rule_reproduces_the_reference_tree():
    """The reference's spiral-disk octree (4320 leaves, its own sorted lists)."""
    t = golden("tree_paths.npz")
    cf = child_first(t["child_begin"], t["child_count"], t["child_index"])
    assert _check(t["bounds"][0], cf >= 0, t["bounds"], cf, t["nb_off"], t["nb_id"]) == 0


def test_rule_reproduces_the_host_builder():
    """The host builder's density-refined tree at levels 2..6 (4e4 nodes); putting a
    subdivided neighbour's children in its place instead of at the end changes lists."""
    pkg = load_pkg()
    h = 10 * PC
    box = [-h, -h, -h, h, h, h]
    ex = pkg.Octree.build(box, pkg.disk_model(tau_v=1.0, spiral=True), 2, 6, 1e-5, 50, 9).export()
    args = (box, ex["child_first"] >= 0, ex["bounds"], ex["child_first"], ex["nb_off"], ex["nb_id"])
    assert _check(*args) == 0
    assert _check(*args, in_place=True) > 0
