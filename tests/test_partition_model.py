"""The device build replaces each std::partition pass with a rank-and-swap
step.  This checks that characterization against a literal transcription of
libstdc++'s two-ended partition (/usr/include/c++/11/bits/stl_algo.h:1487-1510)
on random inputs, and that the oracle's whole-tree leaf order equals the
sequence of such passes on small trees."""
import numpy as np


def std_partition_bidir(a, lo, hi, pred):
    first, last = lo, hi
    while True:
        while True:
            if first == last:
                return first
            elif pred(a[first]):
                first += 1
            else:
                break
        last -= 1
        while True:
            if first == last:
                return first
            elif not pred(a[last]):
                last -= 1
            else:
                break
        a[first], a[last] = a[last], a[first]
        first += 1


def rank_swap(a, lo, hi, pred):
    flags = np.array([bool(pred(x)) for x in a[lo:hi]], dtype=bool)
    t = int(flags.sum())
    left = [lo + i for i in range(t) if not flags[i]]
    right = [lo + i for i in range(t, hi - lo) if flags[i]][::-1]
    assert len(left) == len(right)
    for i, j in zip(left, right):
        a[i], a[j] = a[j], a[i]
    return lo + t


def test_rank_swap_equals_two_ended_partition():
    rng = np.random.default_rng(0)
    for trial in range(400):
        n = int(rng.integers(0, 60))
        vals = list(rng.integers(0, 5, size=n))
        lo = int(rng.integers(0, n + 1))
        hi = int(rng.integers(lo, n + 1))
        c = int(rng.integers(0, 5))
        a1, a2 = list(vals), list(vals)
        s1 = std_partition_bidir(a1, lo, hi, lambda x: x == c)
        s2 = rank_swap(a2, lo, hi, lambda x: x == c)
        assert s1 == s2 and a1 == a2, (trial, vals, lo, hi, c)


def test_sequential_cell_passes_match_oracle_leaf_order():
    """Root-level passes over all cells reproduce the oracle's first-level
    arrangement for a one-level tree (leaf_size large enough that children are
    leaves)."""
    import oracle

    rng = np.random.default_rng(3)
    n = 200
    pts = rng.random((n, 2))
    lo = np.array([-0.01, -0.01])
    hi = np.array([1.01, 1.01])
    t = oracle.Tree(pts, None, np.ones(n), lo, hi, split_log=2, leaf_size=n - 1)
    order = list(range(n))
    p32 = pts.astype(np.float32)
    lo32, hi32 = lo.astype(np.float32), hi.astype(np.float32)

    def cell(i):
        q = []
        for a in range(2):
            v = np.float32(np.float32(4.0) * (p32[i, a] - lo32[a])) / np.float32(hi32[a] - lo32[a])
            q.append(min(max(int(v), 0), 3))
        return (q[1] << 2) + q[0]

    s = 0
    for c in range(16):
        s = rank_swap(order, s, n, lambda i: cell(i) == c)
    assert order == list(t.leaf_order())
