"""Tree topology of build.py:138-186 (median split, LIFO numbering).

The topology is a pure function of (N, leaf_size).  The oracle's stack
restatement is pinned here by properties the algorithm fixes independently of
how it is written, and checked against the per-node closed form that the HIP
kernel k_topology evaluates (one thread per (depth, position) slot).
"""

import numpy as np
import pytest

from oracle import oracle as O


def leaves_closed_form(c: int, L: int) -> int:
    if c <= L:
        return 1
    return leaves_closed_form(c // 2, L) + leaves_closed_form(c - c // 2, L)


def closed_form_topology(N: int, L: int):
    """Per-slot evaluation used by k_topology: a node is reached from the
    root by its left/right path; its index is 1 + 2*r(parent) + side, where
    r(v) is the number of internal nodes before v in right-first preorder
    (the LIFO pop order).  r(right child) = r(v) + 1 and
    r(left child) = r(v) + 1 + internal(right sibling)."""
    out = {}
    depth = 0
    c = N
    while c > L:  # deepest possible level is bounded by repeated halving
        c = c - c // 2
        depth += 1
    for k in range(depth + 1):
        for j in range(1 << k):
            c, s, r, idx, lrank, parent = N, 0, 0, 0, 0, -1
            ok = True
            for t in range(1, k + 1):
                if c <= L:
                    ok = False
                    break
                half = c // 2
                bit = (j >> (k - t)) & 1
                if bit == 0:
                    nidx = 1 + 2 * r
                    r = r + 1 + (leaves_closed_form(c - half, L) - 1)
                    c = half
                else:
                    nidx = 2 + 2 * r
                    lrank += leaves_closed_form(half, L)
                    r = r + 1
                    s, c = s + half, c - half
                parent, idx = idx, nidx
            if ok:
                out[idx] = (s, c, parent, 1 + 2 * r if c > L else -1, lrank if c <= L else -1)
    M = len(out)
    start = np.array([out[i][0] for i in range(M)])
    count = np.array([out[i][1] for i in range(M)])
    parent = np.array([out[i][2] for i in range(M)])
    first_child = np.array([out[i][3] for i in range(M)])
    lrank = np.array([out[i][4] for i in range(M)])
    return start, count, parent, first_child, lrank


def check_invariants(t, N, L):
    M = t.num_nodes
    assert t.count[0] == N and t.start[0] == 0
    assert M == 2 * int(t.is_leaf.sum()) - 1
    assert M <= O.lib().orc_max_nodes(N, L)
    for k in range(M):
        if t.is_leaf[k]:
            assert t.count[k] <= L and t.first_child[k] == -1 and t.child_count[k] == 0
        else:
            assert t.count[k] > L and t.child_count[k] == 2
            a, b = t.first_child[k], t.first_child[k] + 1
            # children split the parent's range at the count median
            assert t.start[a] == t.start[k] and t.count[a] == t.count[k] // 2
            assert t.start[b] == t.start[k] + t.count[k] // 2
            assert t.count[a] + t.count[b] == t.count[k]
    # leaves tile [0, N) exactly once
    lv = np.nonzero(t.is_leaf)[0]
    order = np.argsort(t.start[lv])
    s, c = t.start[lv][order], t.count[lv][order]
    assert s[0] == 0 and np.all(s[1:] == (s + c)[:-1]) and (s + c)[-1] == N


def subtree(t, k):
    out, st = [], [k]
    while st:
        v = st.pop()
        out.append(v)
        if not t.is_leaf[v]:
            st += [t.first_child[v], t.first_child[v] + 1]
    return out


@pytest.mark.parametrize("N,L", [(1, 32), (2, 1), (3, 1), (10, 4), (10, 1), (33, 32), (64, 32), (100, 7),
                                 (1000, 3), (4097, 16), (10_000, 32), (12_345, 5)])
def test_invariants_and_lifo_numbering(N, L):
    t = O.topology(N, L)
    check_invariants(t, N, L)
    # LIFO numbering: the right child's subtree is numbered before the left
    # child's, and children are allocated as a consecutive pair
    for k in range(min(t.num_nodes, 400)):
        if not t.is_leaf[k]:
            a, b = t.first_child[k], t.first_child[k] + 1
            if not t.is_leaf[b] and not t.is_leaf[a]:
                assert max(subtree(t, b)[1:]) < min(subtree(t, a)[1:])


@pytest.mark.parametrize("N,L,M,leaves", [
    (10_000, 32, 1023, 512),           # config 1
    (1_000_000, 32, 65535, 32768),     # config 2
    (4_000_000, 32, 262143, 131072),   # config 5
    (16_000_000, 16, 2097151, 1048576),
    (16_000_000, 32, 1048575, 524288),
    (16_000_000, 64, 524287, 262144),
])
def test_config_node_counts(N, L, M, leaves):
    assert 2 * leaves_closed_form(N, L) - 1 == M
    if N <= 1_000_000:
        t = O.topology(N, L)
        assert t.num_nodes == M and int(t.is_leaf.sum()) == leaves


@pytest.mark.parametrize("N,L", [(1, 1), (2, 1), (5, 2), (11, 1), (37, 4), (100, 3), (513, 8), (2000, 32),
                                 (3001, 17), (9999, 32)])
def test_closed_form_matches_stack_algorithm(N, L):
    t = O.topology(N, L)
    start, count, parent, first_child, lrank = closed_form_topology(N, L)
    np.testing.assert_array_equal(start, t.start)
    np.testing.assert_array_equal(count, t.count)
    np.testing.assert_array_equal(first_child, t.first_child)
    for k in range(1, t.num_nodes):
        p = parent[k]
        assert k in (t.first_child[p], t.first_child[p] + 1)
    lv = np.nonzero(t.is_leaf)[0]
    np.testing.assert_array_equal(np.sort(lrank[lv]), np.arange(len(lv)))
    assert np.all(np.argsort(lrank[lv]) == np.argsort(t.start[lv]))
