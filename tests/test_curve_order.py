"""The spatial order of SBH_EVAL_SORT_QUERIES: Hilbert-curve keys of the
queries' grid cells in the tree's root box (21 bits per axis in 3D, 31 in
2D).  hilbert_keys() is the restatement the GPU order is checked against
(tests/test_gpu_parity.py); it is pinned here by what makes a curve a
Hilbert curve: the keys of a 2^b grid are a bijection onto [0, 2^(d b)), and
consecutive keys are face neighbours."""
import numpy as np
import pytest


def hilbert_index(cells, bits):
    """Hilbert index of integer cells (n, d) on a 2^bits grid (Skilling's
    transpose algorithm, J. Skilling, AIP Conf. Proc. 707, 2004)."""
    d = cells.shape[1]
    X = [cells[:, a].astype(np.int64).copy() for a in range(d)]
    M = 1 << (bits - 1)
    Q = M
    while Q > 1:
        P = Q - 1
        for i in range(d):
            hi = (X[i] & Q) != 0
            X[0] = np.where(hi, X[0] ^ P, X[0])
            t = np.where(hi, 0, (X[0] ^ X[i]) & P)
            X[0] ^= t
            X[i] ^= t
        Q >>= 1
    for i in range(1, d):
        X[i] ^= X[i - 1]
    t = np.zeros_like(X[0])
    Q = M
    while Q > 1:
        t = np.where((X[d - 1] & Q) != 0, t ^ (Q - 1), t)
        Q >>= 1
    for i in range(d):
        X[i] ^= t
    key = np.zeros(len(cells), np.uint64)
    for b in range(bits - 1, -1, -1):
        for i in range(d):
            key = (key << np.uint64(1)) | ((X[i] >> b) & 1).astype(np.uint64)
    return key


def hilbert_keys(q, lo, hi):
    """Keys of the queries as the device computes them: cell coordinates from
    the float rounding of queries and root box, in float64."""
    dim = q.shape[1]
    cells = 2097152.0 if dim == 3 else 2147483648.0
    bits = 21 if dim == 3 else 31
    q = q.astype(np.float32).astype(np.float64)
    lo = np.asarray(lo, np.float32).astype(np.float64)
    hi = np.asarray(hi, np.float32).astype(np.float64)
    t = (q - lo) / (hi - lo) * cells
    c = np.clip(t, 0.0, cells - 1.0).astype(np.int64)
    return hilbert_index(c, bits)


@pytest.mark.parametrize("dim,bits", [(2, 1), (2, 3), (2, 5), (3, 1), (3, 2), (3, 4)])
def test_hilbert_index_is_a_hilbert_curve(dim, bits):
    n = 1 << bits
    grid = np.stack(np.meshgrid(*[np.arange(n)] * dim, indexing="ij"), -1).reshape(-1, dim)
    k = hilbert_index(grid, bits)
    assert sorted(k.tolist()) == list(range(n ** dim))
    path = grid[np.argsort(k)]
    steps = np.abs(np.diff(path, axis=0)).sum(1)
    assert (steps == 1).all()
    assert (path[0] == 0).all()  # starts at the origin


def test_hilbert_order_refines_the_coarse_order():
    """A cell's key prefix is its parent cell's key: sorting by the 21-bit
    keys orders the 2^3-cell coarse grid consistently (a property a Morton
    curve shares; a plain reordering of axes would not keep it)."""
    rng = np.random.default_rng(3)
    c = rng.integers(0, 1 << 21, (5000, 3))
    fine = hilbert_index(c, 21)
    coarse = hilbert_index(c >> 18, 3)
    o = np.argsort(fine, kind="stable")
    assert (np.diff(coarse[o].astype(np.int64)) >= 0).all()
