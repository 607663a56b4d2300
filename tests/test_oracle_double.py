"""Pins for the oracle's double build (the reference compiled with USE_DOUBLE,
include/utils.h:13-29: FLOAT = double, literals FP(x) = x, PI = M_PI).

The double build shares the float build's code, with the scalar type as a
template parameter; these checks pin what the type changes: the float64
kernel constants (closed forms with 0.1, 1e-5 and pi as doubles), the
quantisation and cell bounds in float64 (dyadic-grid leaf order, agreement
with the float build where every operation is exact in both), the moments
(float64 sums), the Kahan brute force (math.fsum), Barnes-Hut (infinite and
tiny beta, an independent recursion for the vote groups) and the estimator
(exact expectation and variance by enumeration, sampling with the 32-bit
roulette uniform)."""
import math

import numpy as np
import pytest

import oracle
import workloads as W
from test_oracle_pins import morton_key

KERNELS = ["gravity", "wn", "smoothdist", "helmholtz_x", "helmholtz_y", "gravity_grad_x",
           "gravity_grad_y", "gravity_grad_z"]


def closed_form64(kernel, dim, p, n, m, alpha, q):
    """include/kernel.h:31-83 with double literals (0.1, 1e-5, M_PI)."""
    d = p - q
    r = np.sqrt((d * d).sum(1))
    if kernel == "gravity":
        return -m / (r + 0.1)
    if kernel == "wn":
        dn = (n * d).sum(1)
        if dim == 2:
            return m * dn / (2 * np.pi * (r * r + 1e-5))
        return m * dn / (4 * np.pi * (r ** 3 + 1e-5))
    if kernel == "helmholtz_x":
        return m * np.cos(4 * np.pi * r) / (r + 0.1)
    if kernel == "helmholtz_y":
        return m * np.sin(2 * np.pi * r) / (r + 0.1)
    if kernel.startswith("gravity_grad_"):
        a = "xyz".index(kernel[-1])
        return -m / (r + 0.1) ** 3 * (q[:, a] - p[:, a])
    return np.exp(-alpha * r)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kernel", KERNELS)
def test_double_kernel_closed_forms(kernel, dim):
    if kernel == "gravity_grad_z" and dim == 2:
        pytest.skip("z component needs 3D")
    rng = np.random.default_rng(21)
    n = 400
    p, nv, q = (rng.normal(size=(n, dim)) for _ in range(3))
    m = rng.uniform(0.5, 2, n)
    got = oracle.kernel_pairs_double(dim, kernel, p, nv, m, 3.0, q)
    ref = closed_form64(kernel, dim, p, nv, m, 3.0, q)
    # the trigonometric kernels pass through zero: bound them by their envelope
    env = m / (np.sqrt(((p - q) ** 2).sum(1)) + 0.1) if kernel.startswith("helm") else np.abs(ref)
    assert (np.abs(got - ref) <= 1e-13 * env + 1e-300).all()
    # the float constants would be visible: 0.1 vs float(0.1) shifts gravity by ~1e-9
    if kernel == "gravity":
        f32c = -m / (np.sqrt(((p - q) ** 2).sum(1)) + np.float64(np.float32(0.1)))
        assert np.abs(got - f32c).max() > 1e-10


@pytest.mark.parametrize("dim,k,levels", [(3, 1, 6), (3, 2, 3), (2, 3, 2)])
def test_double_dyadic_grid_leaf_order_is_morton_order(dim, k, levels):
    g = 1 << (k * levels)
    rng = np.random.default_rng(dim * 10 + k + 1)
    n = min(3000, g ** dim)
    flat = rng.choice(g ** dim, size=n, replace=False)
    cells = np.stack([(flat // g ** a) % g for a in range(dim)], 1).astype(np.int64)
    pts = (cells + 0.5) / g
    t = oracle.Tree(pts, None, np.ones(n), np.zeros(dim), np.ones(dim), split_log=k,
                    dtype=np.float64)
    keys = morton_key(cells, k, dim, levels)
    assert t.leaf_order().tolist() == sorted(range(n), key=lambda i: keys[i])


def test_double_build_matches_float_build_where_both_are_exact():
    """Points of 8 significant bits in a power-of-two box: every quantisation
    and cell bound is exact in float and in double, so both builds make the
    same tree; their moments differ only by the float rounding."""
    rng = np.random.default_rng(5)
    pts = np.round(rng.random((4000, 3)) * 256) / 256
    m = rng.integers(1, 4, 4000).astype(np.float64)
    args = (pts, None, m, np.zeros(3), np.full(3, 2.0))
    for k in (1, 2):
        a = oracle.Tree(*args, split_log=k)
        b = oracle.Tree(*args, split_log=k, dtype=np.float64)
        np.testing.assert_array_equal(a.leaf_order(), b.leaf_order())
        na, nb = a.nodes(), b.nodes()
        for key in ("start", "end", "nchild", "depth", "first_child"):
            np.testing.assert_array_equal(na[key], nb[key])
        np.testing.assert_array_equal(na["bmin"].astype(np.float64), nb["bmin"])
        np.testing.assert_allclose(na["centroid"], nb["centroid"], rtol=2e-6, atol=1e-6)
        assert nb["bmin"].dtype == np.float64 and nb["mass"].dtype == np.float64


def test_double_moments_match_float64_sums():
    pts, _, m, _ = W.uniform_cube(2000, 1, seed=3)
    nrm = np.random.default_rng(5).normal(size=pts.shape)
    lo, hi = W.bounds(pts)
    t = oracle.Tree(pts, nrm, m, lo, hi, split_log=2, dtype=np.float64)
    nd = t.nodes()
    order = t.leaf_order()
    p, n, mm = pts[order], nrm[order], m[order]
    for v in range(0, t.num_nodes, 5):
        s, e = nd["start"][v], nd["end"][v]
        M = mm[s:e].sum()
        np.testing.assert_allclose(nd["mass"][v], M, rtol=1e-13)
        np.testing.assert_allclose(nd["centroid"][v], (mm[s:e, None] * p[s:e]).sum(0) / M,
                                   rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(nd["normal"][v], (mm[s:e, None] * n[s:e]).sum(0) / M,
                                   rtol=1e-11, atol=1e-13)
        # cell bounds: the box's diagonal
        np.testing.assert_allclose(nd["diam"][v], np.linalg.norm(nd["bmax"][v] - nd["bmin"][v]),
                                   rtol=1e-15)
    np.testing.assert_array_equal(nd["mass"], nd["mass64"])  # one float64 recurrence


@pytest.fixture(scope="module")
def case64():
    pts, _, m, q = W.uniform_cube(3000, 200, seed=17)
    nrm = np.random.default_rng(18).normal(size=pts.shape)
    lo, hi = W.bounds(pts)
    return pts, nrm, m, q, lo, hi


@pytest.mark.parametrize("kernel", ["gravity", "wn", "smoothdist"])
def test_double_brute_force_is_the_exact_sum(case64, kernel):
    """Kahan summation of float64 terms: within a few ulp of the exactly
    rounded sum (math.fsum) of the same terms."""
    pts, nrm, m, q, lo, hi = case64
    t = oracle.Tree(pts, nrm, m, lo, hi, alpha=2.0, dtype=np.float64)
    got = t.brute_force(kernel, q)
    assert got.dtype == np.float64
    for i in range(0, len(q), 13):
        terms = closed_form64(kernel, 3, pts, nrm, m, 2.0, q[i][None, :])
        exact = math.fsum(terms)
        scale = np.abs(terms).sum()
        assert abs(got[i] - exact) <= 1e-13 * scale


def test_double_barnes_hut_limits(case64):
    pts, nrm, m, q, lo, hi = case64
    t = oracle.Tree(pts, nrm, m, lo, hi, split_log=2, dtype=np.float64)
    ref = t.brute_force("gravity", q)
    got, cnt = t.barnes_hut("gravity", q, np.inf, counts=True)
    np.testing.assert_allclose(got, ref, rtol=1e-13)
    assert (cnt == len(pts)).all()
    nd = t.nodes()
    tiny, cnt = t.barnes_hut("gravity", q, 1e-300, counts=True)
    root = -nd["mass"][0] / (np.linalg.norm(q - nd["centroid"][0], axis=1) + 0.1)
    np.testing.assert_allclose(tiny, root, rtol=1e-14)
    assert (cnt == 1).all()


def test_double_barnes_hut_vote_groups_against_recursion(case64):
    pts, nrm, m, q, lo, hi = case64
    t = oracle.Tree(pts[:500], nrm[:500], m[:500], lo, hi, split_log=1, dtype=np.float64)
    nd = t.nodes()
    order = t.leaf_order()
    p, mm, qq = pts[:500][order], m[:500][order], q[:64]

    def grav(c, mass, qs):
        return -mass / (np.linalg.norm(qs - c, axis=1) + 0.1)

    for group in (1, 32, 64):
        tot = np.zeros(len(qq))

        def visit(v, mask):
            if not mask.any():
                return
            r = (np.full(len(qq), np.inf) if nd["leaf"][v] else
                 np.linalg.norm(qq - nd["centroid"][v], axis=1) / nd["diam"][v])
            acc = (r >= 1.5).reshape(-1, group).all(1).repeat(group)
            use = mask & acc
            if use.any():
                if nd["leaf"][v]:
                    for j in range(nd["start"][v], nd["end"][v]):
                        tot[use] += grav(p[j], mm[j], qq[use])
                else:
                    tot[use] += grav(nd["centroid"][v], nd["mass"][v], qq[use])
            rest = mask & ~acc
            if rest.any() and not nd["leaf"][v]:
                for c in range(nd["first_child"][v], nd["first_child"][v] + nd["nchild"][v]):
                    visit(c, rest)

        visit(0, np.ones(len(qq), bool))
        got = t.barnes_hut("gravity", qq, 1.5, vote_group=group)
        np.testing.assert_allclose(got, tot, rtol=1e-12)


def enumerate_estimator64(t, pts, m, q, spl):
    """Mean and variance of the double-build estimator for one query, by
    enumerating every sampled point and roulette outcome (float64 nodes)."""
    nd = t.nodes()
    order = t.leaf_order()
    p, mm = pts[order], m[order]

    def f_node(v):
        if nd["leaf"][v]:
            s, e = nd["start"][v], nd["end"][v]
            return (-mm[s:e] / (np.linalg.norm(p[s:e] - q, axis=1) + 0.1)).sum()
        return -nd["mass"][v] / (np.linalg.norm(nd["centroid"][v] - q) + 0.1)

    def ratio(v):
        return np.inf if nd["leaf"][v] else np.linalg.norm(q - nd["centroid"][v]) / nd["diam"][v]

    def child_containing(v, idx):
        for c in range(nd["first_child"][v], nd["first_child"][v] + nd["nchild"][v]):
            if nd["start"][c] <= idx < nd["end"][c]:
                return child_containing(c, idx) if nd["nchild"][c] == 1 else c
        raise AssertionError

    mean_total, var_total = 0.0, 0.0
    for si in range(nd["nchild"][0]):
        S = nd["first_child"][0] + si
        mean_total += f_node(S)
        if nd["leaf"][S]:
            continue
        nS = nd["end"][S] - nd["start"][S]
        e1 = e2 = 0.0
        for idx in range(nd["start"][S], nd["end"][S]):
            cur, prob, ev, acc = S, 1.0, 1.0, 0.0
            cur_beta = max(1.0, ratio(S))
            while not nd["leaf"][cur]:
                ffr = ratio(cur)
                tp = min(cur_beta / ffr, 1.0)
                e1 += prob * (1 - tp) * acc / nS
                e2 += prob * (1 - tp) * acc * acc / nS
                prob *= tp
                ev *= tp
                prev, cur = cur, child_containing(cur, idx)
                cur_beta = max(cur_beta, ffr)
                kids = range(nd["first_child"][prev], nd["first_child"][prev] + nd["nchild"][prev])
                diff = sum(f_node(c) for c in kids) - f_node(prev)
                acc += nS / (ev * (nd["end"][prev] - nd["start"][prev])) * diff
            e1 += prob * acc / nS
            e2 += prob * acc * acc / nS
        mean_total += e1
        var_total += (e2 - e1 * e1) / spl
    return mean_total, var_total


def test_double_stochastic_enumerated_expectation_and_sampling():
    rng = np.random.default_rng(31)
    n = 60
    pts = rng.random((n, 3))
    m = rng.uniform(0.5, 1.5, n)
    t = oracle.Tree(pts, None, m, np.full(3, -0.01), np.full(3, 1.01), split_log=1,
                    dtype=np.float64)
    queries = np.array([[0.5, 0.5, 0.5], [1.3, -0.2, 0.4], [0.05, 0.9, 0.5]])
    exact = t.brute_force("gravity", queries)
    spl, seeds = 2, 4000
    for qi, qv in enumerate(queries):
        mu, var = enumerate_estimator64(t, pts, m, qv, spl)
        assert abs(mu - exact[qi]) <= 1e-12 * abs(exact[qi])  # telescoping
        assert var > 0
        ests = np.array([t.stochastic("gravity", qv[None, :], spl, seed=s, group=1)[0]
                         for s in range(seeds)])
        assert ests.dtype == np.float64
        se = math.sqrt(var / seeds)
        assert abs(ests.mean() - mu) <= 4 * se
        assert abs(ests.var(ddof=1) / var - 1.0) <= 0.12


def test_double_stochastic_shares_paths_with_the_float_build():
    """Same seed, same tree shape: both builds draw the same sample points
    (Philox words 0-1), so on a tree whose roulette always continues (queries
    far away: every tp = 1) the two estimates agree to float precision."""
    rng = np.random.default_rng(41)
    pts = np.round(rng.random((3000, 3)) * 256) / 256
    m = np.ones(3000)
    args = (pts, None, m, np.zeros(3), np.full(3, 2.0))
    a = oracle.Tree(*args, split_log=1)
    b = oracle.Tree(*args, split_log=1, dtype=np.float64)
    q = np.array([[40.0, 35.0, -30.0], [-50.0, 20.0, 60.0]])
    for seed in (0, 9):
        fa = a.stochastic("gravity", q, 3, seed=seed)
        fb = b.stochastic("gravity", q, 3, seed=seed)
        np.testing.assert_allclose(fa, fb, rtol=1e-6)
    # and another seed changes the double estimate
    assert (b.stochastic("gravity", q, 3, seed=1) != b.stochastic("gravity", q, 3, seed=2)).any()


def test_double_stochastic_counts_and_groups(case64):
    pts, nrm, m, q, lo, hi = case64
    t = oracle.Tree(pts, nrm, m, lo, hi, split_log=2, dtype=np.float64)
    a, ca = t.stochastic("gravity", q[:128], 2, seed=3, group=64, counts=True)
    assert (ca > 0).all()
    pad = np.repeat(q[70:71], 65, axis=0)  # query 70 alone in group 1
    b = t.stochastic("gravity", pad, 2, seed=3, group=64)
    assert b[-1] == a[70]
