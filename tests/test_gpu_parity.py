"""Parity of the HIP path (through sbh_gpu -> libsbh_hip.so) with the CPU
oracle on the same seeded inputs.

Bars (from the build's contract):
  * tree: leaf order, node ranges / numbering / bounds / diameters, and the
    fp32 node moments are bit-exact;
  * brute force and Barnes-Hut: within 1e-5 relative of the oracle
    (denominator max(|y|, 1) for the winding number, |y| otherwise);
  * stochastic: the same seed gives the same per-query estimate within 1e-5
    relative (shared Philox paths), and the same evaluation counts.
"""
import numpy as np
import pytest

import oracle
import workloads as W
from test_curve_order import hilbert_keys

pytestmark = pytest.mark.gpu

TOL = 1e-5
KERNELS = ["gravity", "wn", "smoothdist"]


def rel_err(got, ref, kernel):
    ref = np.asarray(ref, np.float64)
    den = np.maximum(np.abs(ref), 1.0) if kernel == "wn" else np.abs(ref)
    return np.abs(np.asarray(got, np.float64) - ref) / den


def gpu_class(sbh_gpu, dim, k):
    return getattr(sbh_gpu, ("OctreeGPU%d" if dim == 3 else "QuadtreeGPU%d") % k)


def make_pair(sbh_gpu, pts, nrm, m, lo, hi, k, leaf_size=1, alpha=1.0, moments64=False):
    dim = pts.shape[1]
    print(f"[case] tree dim={dim} k={k} n={len(pts)} leaf={leaf_size}", flush=True)
    g = gpu_class(sbh_gpu, dim, k)(pts, nrm, m, lo, hi, alpha=alpha, leaf_size=leaf_size,
                                   moments64=moments64)
    o = oracle.Tree(pts, nrm, m, lo, hi, alpha=alpha, split_log=k, leaf_size=leaf_size)
    return g, o


def assert_tree_equal(g, o):
    go = g.leaf_order()
    np.testing.assert_array_equal(go, o.leaf_order())
    gn, on = g.nodes(), o.nodes()
    assert len(gn["start"]) == len(on["start"])
    for key in ("start", "end", "nchild", "depth"):
        np.testing.assert_array_equal(gn[key].astype(np.int64), on[key].astype(np.int64), key)
    leaf = on["leaf"].astype(bool)
    np.testing.assert_array_equal(gn["first_child"][~leaf].astype(np.int64),
                                  on["first_child"][~leaf])
    np.testing.assert_array_equal(gn["parent"][1:].astype(np.int64), on["parent"][1:])
    for key in ("bmin", "bmax", "diam", "centroid", "mass", "normal"):
        np.testing.assert_array_equal(gn[key].view(np.uint32), on[key].view(np.uint32), key)
    m64 = gn["moments64"]
    np.testing.assert_allclose(m64[:, 3], on["mass64"], rtol=1e-12)
    np.testing.assert_allclose(m64[:, :o.dim], on["centroid64"], rtol=1e-9, atol=1e-12)


TREE_CASES = [
    ("cube3d_k1", 3, 1, 1, 3000),
    ("cube3d_k2", 3, 2, 1, 6000),
    ("cube3d_k3", 3, 3, 1, 5000),
    ("cube3d_k1_leaf8", 3, 1, 8, 16384),
    ("cube2d_k1", 2, 1, 1, 5000),
    ("cube2d_k2_leaf4", 2, 2, 4, 9000),
    ("plummer_k1_large", 3, 1, 1, 40000),
    ("plummer_k2_large", 3, 2, 1, 40000),
]


@pytest.mark.parametrize("name,dim,k,leaf,n", TREE_CASES, ids=[c[0] for c in TREE_CASES])
def test_tree_bit_exact(sbh_gpu, name, dim, k, leaf, n):
    if name.startswith("plummer"):
        pts, nrm, m, _ = W.plummer(n, 1, rmax=50.0)
        m = np.random.default_rng(5).uniform(0.5, 1.5, n)
    else:
        pts, nrm, m, _ = W.uniform_cube(n, 1, dim=dim, seed=n)
        nrm = np.random.default_rng(9).normal(size=pts.shape)
    lo, hi = W.bounds(pts)
    g, o = make_pair(sbh_gpu, pts, nrm, m, lo, hi, k, leaf_size=leaf)
    assert_tree_equal(g, o)
    assert g.info()["num_nodes"] == o.num_nodes
    assert g.info()["depth"] == o.depth


@pytest.mark.parametrize("dim,k,leaf", [(3, 1, 1), (3, 2, 1), (3, 1, 8), (2, 2, 1)])
def test_tree_partition_size_classes(sbh_gpu, dim, k, leaf):
    """Nodes of exactly 63 ... 1000 points: the sizes around the wavefront
    partition's 64-position slabs and its 512-point limit (above it a
    workgroup partitions the node).  Each cluster fills one cell of the
    4-per-axis grid alone, so some node holds exactly its points."""
    rng = np.random.default_rng(17 + dim + k + leaf)
    sizes = [63, 64, 65, 127, 128, 129, 511, 512, 513, 514, 1000]
    cells = rng.permutation(4 ** dim)[: len(sizes)]
    pts = []
    for s, c in zip(sizes, cells):
        idx = np.array([(c >> (2 * a)) & 3 for a in range(dim)])
        pts.append(0.125 + 0.25 * idx + 0.01 * (rng.random((s, dim)) - 0.5))
    pts = np.concatenate(pts)
    m = rng.uniform(0.5, 1.5, len(pts))
    g, o = make_pair(sbh_gpu, pts, np.zeros_like(pts), m, np.zeros(dim), np.ones(dim), k,
                     leaf_size=leaf)
    counts = o.nodes()["end"] - o.nodes()["start"]
    assert set(sizes) <= set(counts.tolist())
    assert_tree_equal(g, o)


def test_tree_duplicates_and_max_depth(sbh_gpu):
    """Coincident points never separate: leaves hit max_depth=100 and hold
    several points, whose internal order comes from the partition passes."""
    rng = np.random.default_rng(11)
    base = rng.random((300, 3))
    pts = np.concatenate([base, base[:40], base[:40], base[5:9]])
    m = rng.uniform(0.5, 1.5, len(pts))
    lo, hi = W.bounds(pts)
    g, o = make_pair(sbh_gpu, pts, np.zeros_like(pts), m, lo, hi, 1)
    assert o.depth == 100
    assert_tree_equal(g, o)


def test_tree_single_point_and_root_leaf(sbh_gpu):
    pts = np.array([[0.25, 0.5, 0.75]])
    g, o = make_pair(sbh_gpu, pts, np.zeros_like(pts), np.ones(1), [0, 0, 0], [1, 1, 1], 1)
    assert_tree_equal(g, o)
    pts = np.random.default_rng(2).random((6, 3))
    g, o = make_pair(sbh_gpu, pts, np.zeros_like(pts), np.ones(6), [0, 0, 0], [1, 1, 1], 2,
                     leaf_size=8)
    assert_tree_equal(g, o)
    q = np.random.default_rng(3).random((70, 3))
    for kern in KERNELS:
        fn = getattr(sbh_gpu, "multi_level_prefix_control_variate_" + kern)
        got, _ = fn(g, q, 4)
        ref = o.stochastic(kern, q, 4)
        assert rel_err(got, ref, kern).max() <= TOL


def _ring(n, q, seed):
    """2D winding-number workload: dipoles on a jittered circle
    (python/sampler.py:5-20), queries in the bounding square."""
    rng = np.random.default_rng(seed)
    ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    r = 0.4 * (1 + 0.05 * rng.normal(size=n))
    pts = r[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1)
    nxt = np.roll(pts, -1, axis=0)
    e = nxt - pts
    nrm = np.stack([e[:, 1], -e[:, 0]], 1)
    length = np.linalg.norm(nrm, axis=1)
    nrm /= length[:, None]
    queries = rng.uniform(-0.6, 0.6, (q, 2))
    return pts, nrm, length, queries


def _cube(n, q, dim, seed, kernel):
    if kernel == "wn":
        if dim == 3:
            level = 3 if n <= 2000 else 4 if n <= 6000 else 5
            pts, nrm, m, queries = W.noisy_icosphere(level, q, seed=seed)
        else:
            pts, nrm, m, queries = _ring(n, q, seed)
    else:
        pts, nrm, m, queries = W.uniform_cube(n, q, dim=dim, seed=seed)
    lo, hi = W.bounds(pts)
    return pts, nrm, m, queries, lo, hi


@pytest.mark.parametrize("dim", [3, 2])
@pytest.mark.parametrize("kernel", KERNELS)
def test_brute_force(sbh_gpu, dim, kernel):
    pts, nrm, m, q, lo, hi = _cube(7000, 1001, dim, 21, kernel)
    alpha = 10.0 if kernel == "smoothdist" else 1.0
    g, o = make_pair(sbh_gpu, pts, nrm, m, lo, hi, 1, alpha=alpha)
    got, ms = getattr(sbh_gpu, "eval_brute_force_" + kernel)(g, q)
    ref = o.brute_force(kernel, q)
    ref64 = o.brute_force(kernel, q, fp64=True)
    assert got.dtype == np.float32 and got.shape == (len(q),)
    assert ms > 0
    assert rel_err(got, ref, kernel).max() <= TOL
    assert rel_err(got, ref64, kernel).max() <= TOL
    g64, _ = getattr(sbh_gpu, "eval_brute_force_f64_" + kernel)(g, q)
    assert rel_err(g64, ref64, kernel).max() <= 1e-12


@pytest.mark.parametrize("operands", ["f32", "bf16"])
@pytest.mark.parametrize("kernel", ["gravity", "smoothdist"])
@pytest.mark.parametrize("workload", ["cube", "plummer"])
def test_brute_force_mfma(sbh_gpu, kernel, workload, operands):
    """Matrix-core squared distances with the near field recomputed: within
    1e-5 of the oracle and of float64, for every query (ragged count).  The
    untruncated Plummer sphere (coordinates up to ~1e3) covers large offsets
    for gravity; for smooth distance it is truncated at r <= 10, since
    exp(-10 d) underflows to zero even in float64 for its outer queries."""
    if workload == "cube":
        pts, nrm, m, q = W.uniform_cube(9000, 3001, seed=27)
    else:
        pts, nrm, m, q = W.plummer(9000, 3001, rmax=None if kernel == "gravity" else 10.0)
    lo, hi = W.bounds(pts)
    alpha = 10.0 if kernel == "smoothdist" else 1.0
    g, o = make_pair(sbh_gpu, pts, nrm, m, lo, hi, 1, alpha=alpha)
    got, ms = getattr(sbh_gpu, "eval_brute_force_" + kernel)(g, q, mfma=operands)
    ref = o.brute_force(kernel, q)
    ref64 = o.brute_force(kernel, q, fp64=True)
    assert np.isfinite(got).all()
    if kernel == "smoothdist":
        assert (ref64 > 0).all()
    err = rel_err(got, ref64, kernel)
    print(f"[mfma] {operands} {kernel} {workload}: max {err.max():.2e} mean {err.mean():.2e}")
    assert rel_err(got, ref, kernel).max() <= TOL
    assert err.max() <= TOL


def test_brute_force_ill_conditioned_wn(sbh_gpu):
    """Random dipole directions cancel; the error bar is then relative to the
    sum's condition (sum of |terms|), the bound fp32 terms can meet."""
    pts, _, m, q = W.uniform_cube(6000, 500, seed=23)
    nrm = np.random.default_rng(24).normal(size=pts.shape)
    lo, hi = W.bounds(pts)
    g, o = make_pair(sbh_gpu, pts, nrm, m, lo, hi, 1)
    got, _ = sbh_gpu.eval_brute_force_wn(g, q)
    ref64 = o.brute_force("wn", q, fp64=True)
    cond = o.abs_sum("wn", q)
    assert (np.abs(got - ref64) / cond).max() <= TOL


@pytest.mark.parametrize("vote", [1, 32, 64])
@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("dim,k", [(3, 1), (3, 2), (2, 1)])
def test_barnes_hut(sbh_gpu, vote, kernel, dim, k):
    pts, nrm, m, q, lo, hi = _cube(6000, 777, dim, 31 + k, kernel)
    alpha = 10.0 if kernel == "smoothdist" else 1.0
    g, o = make_pair(sbh_gpu, pts, nrm, m, lo, hi, k, alpha=alpha)
    fn = getattr(sbh_gpu, "eval_barnes_hut_" + kernel)
    for beta in (1.5, 2.0, 3.0, np.inf):
        ref, rcnt = o.barnes_hut(kernel, q, beta, vote_group=vote, counts=True)
        for exact in (True, False):
            got, ms, cnt = fn(g, q, beta, vote_group=vote, return_counts=True, exact_terms=exact)
            np.testing.assert_array_equal(cnt, rcnt)
            err = rel_err(got, ref, kernel)
            assert err.max() <= TOL, (beta, vote, exact, err.max())
            if exact and kernel != "smoothdist":
                # same terms, same order, correctly rounded: identical sums
                np.testing.assert_array_equal(got, ref)
            if not exact:
                # the launch without counts computes the same values
                np.testing.assert_array_equal(fn(g, q, beta, vote_group=vote)[0], got)


def test_barnes_hut_sorted_queries_identical(sbh_gpu):
    pts, nrm, m, q, lo, hi = _cube(20000, 5000, 3, 41, "gravity")
    g, o = make_pair(sbh_gpu, pts, nrm, m, lo, hi, 1)
    a, _ = sbh_gpu.eval_barnes_hut_gravity(g, q, 2.0, vote_group=1)
    b, _ = sbh_gpu.eval_barnes_hut_gravity(g, q, 2.0, vote_group=1, sort_queries=True)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("dim,k", [(3, 2), (3, 1), (2, 2)])
def test_stochastic_same_seed(sbh_gpu, kernel, dim, k):
    pts, nrm, m, q, lo, hi = _cube(6000, 300, dim, 51 + k, kernel)
    alpha = 10.0 if kernel == "smoothdist" else 1.0
    g, o = make_pair(sbh_gpu, pts, nrm, m, lo, hi, k, alpha=alpha)
    fn = getattr(sbh_gpu, "multi_level_prefix_control_variate_" + kernel)
    for spl, seed in ((1, 0), (4, 7), (3, 2**40 + 5)):
        ref, rcnt = o.stochastic(kernel, q, spl, seed=seed, group=64, counts=True)
        got, ms, cnt = fn(g, q, spl, seed=seed, return_counts=True, exact_terms=True)
        np.testing.assert_array_equal(cnt, rcnt)
        err = rel_err(got, ref, kernel)
        assert err.max() <= TOL, (spl, seed, err.max())
        if kernel != "smoothdist":
            np.testing.assert_array_equal(got, ref)
        counted = fn(g, q, spl, seed=seed, return_counts=True)
        _check_fast(counted, ref, rcnt, kernel, f"{kernel} d{dim} k{k} spl{spl}")
        # the launch without counts computes the same values
        np.testing.assert_array_equal(fn(g, q, spl, seed=seed)[0], counted[0])


def _check_fast(res, ref, rcnt, kernel, label):
    """The fast mode (hardware sqrt / reciprocal / exp2 for terms, opening
    ratios and roulette probabilities) follows the same paths as the exact
    estimator except where a roulette uniform falls within a few ulp of the
    probability; on every other query the terms agree to fp32 rounding."""
    fast, _, fcnt = res
    same = fcnt == rcnt
    print(f"[fast] {label}: {np.count_nonzero(~same)} of {len(same)} paths differ", flush=True)
    assert np.count_nonzero(~same) <= max(2, len(same) // 100)
    ferr = rel_err(fast[same], ref[same], kernel)
    print(f"[fast] {label}: mean {ferr.mean():.2e} max {ferr.max():.2e}")
    assert ferr.mean() <= TOL and np.percentile(ferr, 99) <= 10 * TOL


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("dim,k,leaf", [(3, 2, 1), (3, 1, 4), (2, 2, 1), (2, 3, 3)])
def test_stochastic_per_query_paths(sbh_gpu, kernel, dim, k, leaf):
    """One sample path per query (the reference CPU sampler's draws): values
    and term counts equal the oracle's 1-query groups, and do not depend on
    the evaluation order.  Leaf sizes above 1 cover multi-point leaves."""
    pts, nrm, m, q, lo, hi = _cube(5000, 257, dim, 91 + k, kernel)
    alpha = 10.0 if kernel == "smoothdist" else 1.0
    g, o = make_pair(sbh_gpu, pts, nrm, m, lo, hi, k, leaf_size=leaf, alpha=alpha)
    fn = getattr(sbh_gpu, "multi_level_prefix_control_variate_" + kernel)
    for spl, seed in ((1, 3), (5, 2**33 + 1)):
        ref, rcnt = o.stochastic(kernel, q, spl, seed=seed, group=1, counts=True)
        got, _, cnt = fn(g, q, spl, seed=seed, return_counts=True, exact_terms=True,
                         per_query_paths=True)
        np.testing.assert_array_equal(cnt, rcnt)
        assert rel_err(got, ref, kernel).max() <= TOL
        if kernel != "smoothdist":
            np.testing.assert_array_equal(got, ref)
        srt, _, scnt = fn(g, q, spl, seed=seed, return_counts=True, exact_terms=True,
                          per_query_paths=True, sort_queries=True)
        np.testing.assert_array_equal(srt, got)
        np.testing.assert_array_equal(scnt, cnt)
        # a query's result does not depend on the other queries of the call
        one, _ = fn(g, q[37:38], spl, seed=seed, exact_terms=True, per_query_paths=True)
        assert rel_err(one, o.stochastic(kernel, q[37:38], spl, seed=seed, group=1),
                       kernel).max() <= TOL
        _check_fast(fn(g, q, spl, seed=seed, per_query_paths=True, return_counts=True), ref,
                    rcnt, kernel, f"per-query {kernel} d{dim} k{k} leaf{leaf} spl{spl}")
    # the shared-path estimator on multi-point leaves
    ref = o.stochastic(kernel, q, 2, seed=9, group=64)
    got, _ = fn(g, q, 2, seed=9, exact_terms=True)
    assert rel_err(got, ref, kernel).max() <= TOL


@pytest.mark.parametrize("cls,dim,k", [("Octree1", 3, 1), ("Octree2", 3, 2), ("Octree3", 3, 3),
                                       ("Quadtree1", 2, 1), ("Quadtree2", 2, 2),
                                       ("Quadtree3", 2, 3)])
def test_cpu_plugin_api(sbh_gpu, cls, dim, k):
    """The sbh module's evaluation functions (src/sbh.cpp:216-270): the
    reference CPU plugin's per-query Barnes-Hut and per-query-path stochastic
    estimator, bit-identical to the oracle with its 1-query groups."""
    import sbh

    for kernel in KERNELS:
        pts, nrm, m, q, lo, hi = _cube(3000, 200, dim, 97 + k, kernel)
        alpha = 10.0 if kernel == "smoothdist" else 1.0
        t = getattr(sbh, cls)(pts, nrm, m, lo, hi, alpha)
        o = oracle.Tree(pts, nrm, m, lo, hi, alpha=alpha, split_log=k)
        bh = getattr(sbh, "eval_barnes_hut_" + kernel)(t, q, 2.0)
        assert isinstance(bh, np.ndarray) and bh.dtype == np.float32 and bh.shape == (200,)
        ref_bh = o.barnes_hut(kernel, q, 2.0, vote_group=1)
        vals, cnt = getattr(sbh, "multi_level_prefix_control_variate_" + kernel)(t, q, 2)
        assert vals.dtype == np.float32 and cnt.dtype == np.int32
        ref_s, ref_c = o.stochastic(kernel, q, 2, seed=0, group=1, counts=True)
        np.testing.assert_array_equal(cnt, ref_c)
        if kernel == "smoothdist":
            assert rel_err(bh, ref_bh, kernel).max() <= TOL
            assert rel_err(vals, ref_s, kernel).max() <= TOL
        else:
            np.testing.assert_array_equal(bh, ref_bh)
            np.testing.assert_array_equal(vals, ref_s)
        exact = getattr(sbh, "eval_barnes_hut_" + kernel)(t, q, np.inf)
        assert rel_err(exact, o.brute_force(kernel, q, fp64=True), kernel).max() <= TOL


def test_stochastic_unbiased_and_variance(sbh_gpu):
    """Mean over 256 seeds within 3 standard errors of brute force (pooled z
    test plus a per-query bound with multiple-testing slack), and per-query
    variance within 10% of the oracle estimator's at the same sample budget."""
    pts, nrm, m, q, lo, hi = _cube(4096, 128, 3, 61, "gravity")
    g, o = make_pair(sbh_gpu, pts, nrm, m, lo, hi, 2)
    exact = o.brute_force("gravity", q, fp64=True)
    seeds = 256
    est = np.stack([sbh_gpu.multi_level_prefix_control_variate_gravity(g, q, 1, seed=s)[0]
                    for s in range(seeds)]).astype(np.float64)
    mean, sd = est.mean(0), est.std(0, ddof=1)
    se = sd / np.sqrt(seeds)
    z = (mean - exact) / np.where(se > 0, se, 1.0)
    assert abs(z.mean()) * np.sqrt(len(z)) <= 3.0 * max(1.0, z.std())
    assert np.abs(z).max() <= 4.5  # 128 queries: per-query 3-SE bound with Bonferroni slack
    # oracle estimator, independent seeds, same budget
    ests_o = np.stack([o.stochastic("gravity", q, 1, seed=10_000 + s) for s in range(seeds)])
    var_g = est.var(0, ddof=1).mean()
    var_o = ests_o.astype(np.float64).var(0, ddof=1).mean()
    assert abs(var_g / var_o - 1.0) <= 0.10


def test_ragged_and_empty_queries(sbh_gpu):
    pts, nrm, m, q, lo, hi = _cube(3000, 130, 3, 71, "gravity")
    g, o = make_pair(sbh_gpu, pts, nrm, m, lo, hi, 2)
    for nq in (0, 1, 63, 65, 130):
        qq = q[:nq]
        a, _ = sbh_gpu.eval_brute_force_gravity(g, qq)
        b, _ = sbh_gpu.eval_barnes_hut_gravity(g, qq, 2.0)
        c, _ = sbh_gpu.multi_level_prefix_control_variate_gravity(g, qq, 2, exact_terms=True)
        assert a.shape == b.shape == c.shape == (nq,)
        if nq:
            assert rel_err(a, o.brute_force("gravity", qq), "gravity").max() <= TOL
            assert rel_err(b, o.barnes_hut("gravity", qq, 2.0, vote_group=32), "gravity").max() <= TOL
            assert rel_err(c, o.stochastic("gravity", qq, 2), "gravity").max() <= TOL


def test_host_tree_constructor(sbh_gpu):
    import sbh

    pts, nrm, m, q, lo, hi = _cube(2000, 100, 3, 81, "gravity")
    cpu = sbh.Octree2(pts, nrm, m, lo, hi)
    g = sbh_gpu.OctreeGPU2(cpu)
    o = oracle.Tree(pts, nrm, m, lo, hi, split_log=2)
    assert_tree_equal(g, o)
    with pytest.raises(TypeError):
        sbh_gpu.OctreeGPU1(cpu)


@pytest.mark.parametrize("dim", [3, 2])
@pytest.mark.parametrize("n", [1, 2, 2047, 2048, 2049, 6001, 100003, (1 << 20) + 5])
def test_query_order_is_a_stable_hilbert_sort(sbh_gpu, dim, n):
    """sbh_query_order (the order SBH_EVAL_SORT_QUERIES evaluates in): a
    permutation, Hilbert keys non-decreasing, equal keys in input order.
    Sizes around the sort's 2048-pair tiles and its merge passes, with
    repeated keys (the same point many times) mixed in."""
    import sbh_capi as C

    rng = np.random.default_rng(40 + dim + n)
    pts = rng.random((5000, dim))
    lo, hi = W.bounds(pts)
    rep = min(30, n // 100)
    q = np.concatenate([lo + (hi - lo) * rng.random((n - 100 * rep, dim)),
                        np.repeat(pts[:rep], 100, axis=0)])
    q = q[rng.permutation(n)] if rep else q
    t = C.Tree(pts, None, np.ones(len(pts)), lo, hi)
    soa = np.ascontiguousarray(q.astype(np.float32).T)
    order = np.empty(n, np.uint32)
    C.check(C.lib().sbh_query_order(t.h, soa.ctypes.data, n, order.ctypes.data, 0))
    np.testing.assert_array_equal(np.sort(order), np.arange(n, dtype=np.uint32))
    qkeys = hilbert_keys(q, lo, hi)
    keys = qkeys[order]
    assert (keys[1:] >= keys[:-1]).all()
    same = keys[1:] == keys[:-1]
    assert (order[1:][same] > order[:-1][same]).all()
    # the order by (key, input index) is unique
    np.testing.assert_array_equal(order, np.lexsort((np.arange(n), qkeys)))
    if rep:
        assert same.sum() >= 99 * rep


@pytest.mark.parametrize("dim,k", [(3, 1), (3, 2), (2, 2)])
def test_small_launches_split(sbh_gpu, dim, k):
    """Launches too small to fill the GPU split Barnes-Hut over a frontier of
    subtrees and the sorted estimator over the root's subdomains.  Queries
    far outside the box accept
    the root itself (its moment, one node term); near queries walk every
    subtree.  Values agree with the oracle to fp32 rounding, counts exactly."""
    pts, nrm, m, q, lo, hi = _cube(3000, 300, dim, 61 + k, "gravity")
    d = q - 0.5
    far = 0.5 + 60.0 * d / np.linalg.norm(d, axis=1, keepdims=True)  # ratio to the root >> beta
    qs = np.concatenate([q, far])
    g, o = make_pair(sbh_gpu, pts, nrm, m, lo, hi, k)
    for beta in (1.5, 3.0):
        got, _, cnt = sbh_gpu.eval_barnes_hut_gravity(g, qs, beta, vote_group=1,
                                                      return_counts=True)
        ref, rcnt = o.barnes_hut("gravity", qs, beta, vote_group=1, counts=True)
        np.testing.assert_array_equal(cnt, rcnt)
        assert (rcnt[len(q):] == 1).all()  # the far queries take the root's moment
        assert rel_err(got, ref, "gravity").max() <= TOL
        # without counts: the same split, the same values
        np.testing.assert_array_equal(
            sbh_gpu.eval_barnes_hut_gravity(g, qs, beta, vote_group=1)[0], got)
    st, _, scnt = sbh_gpu.multi_level_prefix_control_variate_gravity(g, qs, 3, seed=5,
                                                                     return_counts=True)
    np.testing.assert_array_equal(
        sbh_gpu.multi_level_prefix_control_variate_gravity(g, qs, 3, seed=5)[0], st)
    ref, rcnt = o.stochastic("gravity", qs, 3, seed=5, group=64, counts=True)
    same = scnt == rcnt
    assert np.count_nonzero(~same) <= max(2, len(same) // 100)
    assert rel_err(st[same], ref[same], "gravity").max() <= TOL


def _walk_split_tasks(sbh_gpu, tree, kern, q):
    """Tasks a Barnes-Hut launch of q queries splits into, from the serving
    engine's working-set accounting (a split request holds nsplit x q float
    parts beyond the whole-walk request's working set)."""
    def peak(no_split):
        srv = sbh_gpu.Server(256 << 20, 1)
        srv.add_tree(tree)
        srv.submit(0, "barnes_hut", kern, q, beta=2.0, no_split=no_split).wait(60.0)
        b = srv.stats()["peak_bytes"]
        del srv
        return b

    extra = peak(False) - peak(True)
    ns = round(extra / (len(q) * 4))
    assert abs(extra - ns * len(q) * 4) < 512
    return max(ns, 1)


def _spread_estimates(pts, lo, hi, q, order):
    """Per 64-query wavefront of a launch in `order`: the points its queries'
    bounding box would hold at the density of the 32^3-cell grid cells the
    queries lie in (mean point count x (box diagonal / cell diagonal)^3),
    the measure the walk compares with 512 to walk lane by lane."""
    lo32, hi32 = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    g = 32
    width = (hi32 - lo32).astype(np.float64)

    def cells(x):
        c = np.clip((x.astype(np.float32) - lo32) * (np.float32(g) / (hi32 - lo32)), 0, g - 1)
        c = c.astype(np.int64)
        return (c[:, 2] * g + c[:, 1]) * g + c[:, 0]

    counts = np.bincount(cells(pts), minlength=g ** 3)
    diag = np.linalg.norm(width / g)
    est = []
    for w in range(0, len(q), 64):
        qq = q[order[w:w + 64]].astype(np.float32).astype(np.float64)
        dens = counts[cells(qq)].mean()
        est.append(dens * (np.linalg.norm(qq.max(0) - qq.min(0)) / diag) ** 3)
    return np.asarray(est)


def test_barnes_hut_spread_wavefronts(sbh_gpu):
    """Wavefronts whose queries are spread out walk with every lane on its own
    path, the others share one walk (the union of their lanes' walks): either
    way each query takes exactly the oracle's per-query walk
    (include/barnes_hut.h:23-78).  Sparse uniform queries over a concentrated
    Plummer tree (a wavefront's queries span a region holding, at the local
    density, far more than 512 of its points) beside a tight cluster of
    queries (far fewer), in input order and Hilbert-sorted, split into tasks
    and as whole walks."""
    # 2^18 points: the split launches' subtrees hold over 1024 nodes on average,
    # so their wavefronts measure the spread too
    pts, nrm, m, _ = W.plummer(1 << 18, 1, seed_sources=31, seed_queries=32, rmax=10.0)
    lo, hi = W.bounds(pts)
    rng = np.random.default_rng(33)
    sparse = lo + (hi - lo) * rng.random((1500, 3))
    tight = 0.01 * rng.standard_normal((700, 3))
    q = np.concatenate([sparse, tight])
    g, o = make_pair(sbh_gpu, pts, nrm, m, lo, hi, 1)
    ns = _walk_split_tasks(sbh_gpu, g, "gravity", q)
    assert ns > 1 and len(g.nodes()["diam"]) / ns > 1024, ns

    def spread(order):
        return _spread_estimates(pts, lo, hi, q, order)

    for order in (np.arange(len(q)), np.argsort(hilbert_keys(q, lo, hi), kind="stable")):
        est = spread(order)
        assert (est > 1.5 * 512).sum() >= 5 and (est < 512 / 4).sum() >= 5, est
    for beta in (1.5, 3.0):
        ref, rcnt = o.barnes_hut("gravity", q, beta, vote_group=1, counts=True)
        for sort in (False, True):
            for no_split in (False, True):
                got, _, cnt = sbh_gpu.eval_barnes_hut_gravity(
                    g, q, beta, vote_group=1, sort_queries=sort, return_counts=True,
                    no_split=no_split)
                np.testing.assert_array_equal(cnt, rcnt)
                assert rel_err(got, ref, "gravity").max() <= TOL, (beta, sort, no_split)


def _clustered_with_shallow_leaves():
    """An octree whose depth-1 level holds two leaves (clusters of 5 and 7
    points at leaf size 8) beside one populated octant: its split frontier
    has leaves above the frontier's depth."""
    rng = np.random.default_rng(77)
    a = 0.05 + 0.1 * rng.random((5, 3))
    c = np.array([0.85, 0.05, 0.85]) + 0.1 * rng.random((7, 3))
    b = 0.5 + 0.5 * rng.random((3000, 3))
    pts = np.concatenate([a, b, c])
    return pts, np.zeros_like(pts), rng.uniform(0.5, 1.5, len(pts)), np.zeros(3), np.ones(3)


def _shell_queries(q, center, size, rng):
    """Queries at 0.3 to 4 box widths from the box centre in random
    directions: far ones accept the root, middle ones nodes below it."""
    d = rng.normal(size=(q, len(center)))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return center + size * rng.uniform(0.3, 4.0, (q, 1)) * d


@pytest.mark.parametrize("case", ["octree_k1", "quadtree_k1", "plummer_k1_leaf8",
                                  "shallow_leaves_leaf8", "octree_k1_150k"])
def test_barnes_hut_split_frontiers(sbh_gpu, case):
    """The split Barnes-Hut launch over balanced frontiers: from the root's
    children, the heaviest subtree is replaced by its children while the
    frontier stays within 256 tasks and depth 8, and a launch takes the
    smallest kept frontier (4, 8, 16, ... tasks) that makes 16k wavefronts,
    else the largest.  A query that accepts an ancestor above its task's
    subtree takes that ancestor's moment once.  Counts equal the oracle's
    per-query walk exactly, values agree to fp32 rounding, with near queries
    (walks into every task) and shells of queries that accept the root or
    nodes between the root and the frontier.  The task count is checked where
    the rule fixes it: a small launch takes the largest frontier, which stops
    only when no subtree's children fit (within 2^d - 1 of 256 here)."""
    rng = np.random.default_rng(sum(map(ord, case)))
    leaf, k, dim, want = 1, 1, 3, None
    largest = (256 - 7, 256)
    if case == "octree_k1":
        pts, nrm, m, q, lo, hi = _cube(6000, 700, 3, 5, "gravity")
        want = largest
    elif case == "quadtree_k1":
        pts, nrm, m, q, lo, hi = _cube(3000, 700, 2, 6, "gravity")
        dim, want = 2, (256 - 3, 256)
    elif case == "plummer_k1_leaf8":
        pts, nrm, m, q = W.plummer(1 << 13, 700, seed_sources=8, seed_queries=9)
        lo, hi = W.bounds(pts)
        leaf = 8
    elif case == "shallow_leaves_leaf8":
        # the two depth-1 leaves stay tasks of their own beside the
        # populated octant's subtrees
        pts, nrm, m, lo, hi = _clustered_with_shallow_leaves()
        q = rng.random((700, 3))
        leaf, want = 8, largest
    else:
        pts, nrm, m, q, lo, hi = _cube(6000, 150000, 3, 7, "gravity")
        want = (8, 8)  # 150000 queries fill the GPU with the root's 8 children already
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    q = np.concatenate([q, _shell_queries(len(q) // 2, (lo + hi) / 2, (hi - lo).max(), rng)])
    g, o = make_pair(sbh_gpu, pts, nrm, m, lo, hi, k, leaf_size=leaf)
    ns = _walk_split_tasks(sbh_gpu, g, "gravity", q)
    print(f"[frontier] {case}: {ns} tasks for {len(q)} queries", flush=True)
    assert ns > 1
    if want is not None:
        assert want[0] <= ns <= want[1], (ns, want)
    for beta in (1.5, 3.0):
        got, _, cnt = sbh_gpu.eval_barnes_hut_gravity(g, q, beta, vote_group=1,
                                                      return_counts=True)
        ref, rcnt = o.barnes_hut("gravity", q, beta, vote_group=1, counts=True)
        np.testing.assert_array_equal(cnt, rcnt)
        assert rel_err(got, ref, "gravity").max() <= TOL, (case, beta)
        np.testing.assert_array_equal(
            sbh_gpu.eval_barnes_hut_gravity(g, q, beta, vote_group=1)[0], got)
        whole, _ = sbh_gpu.eval_barnes_hut_gravity(g, q, beta, vote_group=1, no_split=True)
        assert rel_err(got, whole, "gravity").max() <= TOL
