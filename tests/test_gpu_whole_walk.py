"""The whole-walk kernels that large launches run (the per-query Barnes-Hut
walk and the sorted-sample estimator, with and without term counts) against
the oracle on the same seeded inputs.

Barnes-Hut launches of fewer than 2^18 queries, and estimator launches of any
size on trees with two or more internal root subdomains, normally split each
walk into tasks; `no_split=True` (SBH_EVAL_NO_SPLIT) makes them run the
whole-walk kernels, and one case of 2^18 + 77 queries reaches the Barnes-Hut
one by size alone.  Bars:

  * Barnes-Hut per query, fast terms: within 1e-5 of
    oracle.barnes_hut(vote_group=1) (include/barnes_hut.h:23-78), counts
    identical, the launch without counts bit-identical to the counted one;
  * the estimator (include/tree_sampler_gpu.h:25-140 with the 64 queries of
    a wavefront sharing their paths): at most 1 % of queries on different
    sample paths than oracle.stochastic(group=64) (a roulette uniform within
    a few ulp of its probability), the rest within 1e-5 (mean) of it;
  * Hilbert-sorted launches (sort_queries=True): the oracle evaluates the
    queries permuted into the launch order (sbh_query_order), and its
    results are permuted back.
"""
import numpy as np
import pytest

import oracle
import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-5
KERNELS = ["gravity", "wn", "smoothdist"]


def rel_err(got, ref, kernel):
    ref = np.asarray(ref, np.float64)
    den = np.maximum(np.abs(ref), 1.0) if kernel == "wn" else np.abs(ref)
    return np.abs(np.asarray(got, np.float64) - ref) / den


def _inputs(n, q, dim, seed, kernel):
    if kernel == "wn" and dim == 3:
        pts, nrm, m, qs = W.noisy_icosphere(4 if n <= 6000 else 5, q, seed=seed)
    elif kernel == "wn":
        rng = np.random.default_rng(seed)
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        r = 0.4 * (1 + 0.05 * rng.normal(size=n))
        pts = r[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1)
        e = np.roll(pts, -1, axis=0) - pts
        nrm = np.stack([e[:, 1], -e[:, 0]], 1)
        m = np.linalg.norm(nrm, axis=1)
        nrm /= m[:, None]
        qs = rng.uniform(-0.6, 0.6, (q, 2))
    else:
        pts, nrm, m, qs = W.uniform_cube(n, q, dim=dim, seed=seed)
    lo, hi = W.bounds(pts)
    return pts, nrm, m, qs, lo, hi


def _trees(sbh_gpu, pts, nrm, m, lo, hi, k, alpha):
    dim = pts.shape[1]
    print(f"[case] dim={dim} k={k} n={len(pts)} alpha={alpha}", flush=True)
    cls = getattr(sbh_gpu, ("OctreeGPU%d" if dim == 3 else "QuadtreeGPU%d") % k)
    g = cls(pts, nrm, m, lo, hi, alpha=alpha)
    o = oracle.Tree(pts, nrm, m, lo, hi, alpha=alpha, split_log=k)
    return g, o


def launch_order(pts, m, lo, hi, queries):
    """order[i] = the query evaluated in slot i of a sorted launch.  The order
    depends only on the tree's box and dimension."""
    import sbh_capi as C

    t = C.Tree(pts, None, m, lo, hi)
    soa = np.ascontiguousarray(np.asarray(queries, np.float32).T)
    order = np.empty(len(queries), np.uint32)
    C.check(C.lib().sbh_query_order(t.h, soa.ctypes.data, len(queries), order.ctypes.data, 0))
    t.close()
    assert np.array_equal(np.sort(order), np.arange(len(queries), dtype=np.uint32))
    return order.astype(np.int64)


def oracle_in_order(fn, queries, order, *a, **kw):
    """The oracle run on the queries in launch order, results in input order."""
    res = fn(queries[order], *a, **kw)
    if isinstance(res, tuple):
        out = []
        for r in res:
            u = np.empty_like(r)
            u[order] = r
            out.append(u)
        return tuple(out)
    u = np.empty_like(res)
    u[order] = res
    return u


def check_same_paths(fast, fcnt, ref, rcnt, kernel, label):
    same = fcnt == rcnt
    ndiff = int(np.count_nonzero(~same))
    ferr = rel_err(fast[same], ref[same], kernel)
    print(f"[paths] {label}: {ndiff} of {len(same)} differ; mean {ferr.mean():.2e} "
          f"p99 {np.percentile(ferr, 99):.2e} max {ferr.max():.2e}", flush=True)
    assert ndiff <= max(2, len(same) // 100), label
    assert ferr.mean() <= TOL and np.percentile(ferr, 99) <= 10 * TOL, label


def check_bh(sbh_gpu, g, o, q, kernel, beta, label, **kw):
    fn = getattr(sbh_gpu, "eval_barnes_hut_" + kernel)
    got, _, cnt = fn(g, q, beta, vote_group=1, return_counts=True, **kw)
    ref, rcnt = o.barnes_hut(kernel, q, beta, vote_group=1, counts=True)
    np.testing.assert_array_equal(cnt, rcnt)
    err = rel_err(got, ref, kernel)
    print(f"[bh] {label} beta={beta}: max {err.max():.2e}", flush=True)
    assert err.max() <= TOL, (label, beta, err.max())
    # the launch without counts (the walk the benchmark times) computes the same values
    plain, _ = fn(g, q, beta, vote_group=1, **kw)
    np.testing.assert_array_equal(plain, got)
    return got


def check_stochastic(sbh_gpu, g, o, q, kernel, spl, seed, label, order=None, **kw):
    fn = getattr(sbh_gpu, "multi_level_prefix_control_variate_" + kernel)
    fast, _, fcnt = fn(g, q, spl, seed=seed, return_counts=True, **kw)
    if order is None:
        ref, rcnt = o.stochastic(kernel, q, spl, seed=seed, group=64, counts=True)
    else:
        ref, rcnt = oracle_in_order(
            lambda qq: o.stochastic(kernel, qq, spl, seed=seed, group=64, counts=True), q, order)
    check_same_paths(fast, fcnt, ref, rcnt, kernel, label)
    plain, _ = fn(g, q, spl, seed=seed, **kw)
    np.testing.assert_array_equal(plain, fast)
    return fast


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("dim,k", [(3, 2), (3, 1), (2, 2), (2, 3)])
def test_whole_walk_kernels_small_launch(sbh_gpu, kernel, dim, k):
    pts, nrm, m, q, lo, hi = _inputs(6000, 777, dim, 131 + k, kernel)
    alpha = 10.0 if kernel == "smoothdist" else 1.0
    g, o = _trees(sbh_gpu, pts, nrm, m, lo, hi, k, alpha)
    label = f"{kernel} d{dim} k{k}"
    for beta in (1.5, 2.0, 3.0):
        a = check_bh(sbh_gpu, g, o, q, kernel, beta, label, no_split=True)
        # the split launch visits the same terms: equal to fp32 rounding
        b, _ = getattr(sbh_gpu, "eval_barnes_hut_" + kernel)(g, q, beta, vote_group=1)
        assert rel_err(b, a, kernel).max() <= TOL
    for spl, seed in ((1, 0), (4, 7), (3, 2**40 + 5)):
        check_stochastic(sbh_gpu, g, o, q, kernel, spl, seed, f"{label} spl{spl}",
                         no_split=True)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("dim,k", [(3, 2), (2, 2)])
def test_whole_walk_kernels_sorted_launch(sbh_gpu, kernel, dim, k):
    """sort_queries=True: the benchmark's launch.  Barnes-Hut per query is
    order independent; the estimator's wavefront groups are the launch
    order's consecutive 64 queries."""
    pts, nrm, m, q, lo, hi = _inputs(6000, 1500, dim, 151 + k, kernel)
    alpha = 10.0 if kernel == "smoothdist" else 1.0
    g, o = _trees(sbh_gpu, pts, nrm, m, lo, hi, k, alpha)
    order = launch_order(pts, m, lo, hi, q)
    label = f"sorted {kernel} d{dim} k{k}"
    a = check_bh(sbh_gpu, g, o, q, kernel, 2.0, label, no_split=True, sort_queries=True)
    b, _ = getattr(sbh_gpu, "eval_barnes_hut_" + kernel)(g, q, 2.0, vote_group=1, no_split=True)
    np.testing.assert_array_equal(a, b)
    for spl, seed in ((2, 3), (5, 11)):
        check_stochastic(sbh_gpu, g, o, q, kernel, spl, seed, f"{label} spl{spl}", order=order,
                         no_split=True, sort_queries=True)


def test_whole_walk_kernels_by_size(sbh_gpu):
    """2^18 + 77 queries on a 4096-point octree (k = 2): large enough that
    the default Barnes-Hut launch takes the whole-walk kernel without any
    flag, in input order and in Hilbert order; the estimator's default
    launch (split over ranges of root subdomains) agrees with its whole-walk
    launch to fp32 rounding."""
    pts, nrm, m, q, lo, hi = _inputs(4096, (1 << 18) + 77, 3, 171, "gravity")
    g, o = _trees(sbh_gpu, pts, nrm, m, lo, hi, 2, 1.0)
    a = check_bh(sbh_gpu, g, o, q, "gravity", 2.0, "2^18+77")
    b = check_bh(sbh_gpu, g, o, q, "gravity", 2.0, "2^18+77 sorted", sort_queries=True)
    np.testing.assert_array_equal(a, b)
    # and equal to the no-split flag's launch, which is the same kernel
    c, _ = sbh_gpu.eval_barnes_hut_gravity(g, q, 2.0, vote_group=1, no_split=True)
    np.testing.assert_array_equal(a, c)
    check_stochastic(sbh_gpu, g, o, q, "gravity", 2, 19, "2^18+77 spl2")
    order = launch_order(pts, m, lo, hi, q)
    s = check_stochastic(sbh_gpu, g, o, q, "gravity", 2, 19, "2^18+77 spl2 sorted", order=order,
                         sort_queries=True)
    t = check_stochastic(sbh_gpu, g, o, q, "gravity", 2, 19, "2^18+77 spl2 sorted whole",
                         order=order, sort_queries=True, no_split=True)
    assert rel_err(s, t, "gravity").max() <= TOL


def test_plummer_headline_shape_small(sbh_gpu):
    """The benchmark's workload shape at reduced size: untruncated Plummer
    sources and queries (a root box hundreds of times wider than the core, so
    a handful of populated subdomains), octree k = 2, Hilbert-sorted launch,
    Barnes-Hut at theta 1.0 and the estimator at 64 samples per subdomain."""
    pts, nrm, m, q = W.plummer(1 << 15, 4096)
    lo, hi = W.bounds(pts)
    g, o = _trees(sbh_gpu, pts, nrm, m, lo, hi, 2, 1.0)
    order = launch_order(pts, m, lo, hi, q)
    check_bh(sbh_gpu, g, o, q, "gravity", 1.0, "plummer", no_split=True, sort_queries=True)
    check_stochastic(sbh_gpu, g, o, q, "gravity", 64, 1, "plummer spl64", order=order,
                     no_split=True, sort_queries=True)


def test_sorted_estimator_past_packed_keys(sbh_gpu):
    """2^23 + 3001 points: the sorted estimator draws 512-sample chunks with
    the sample index packed under the point index while a subdomain's points
    lie below 2^23, and 256-sample chunks with the index in the roulette
    word's low bits past it (the last subdomains here); both against the
    oracle, in the whole-walk and the split launch."""
    pts, nrm, m, q, lo, hi = _inputs((1 << 23) + 3001, 512, 3, 303, "gravity")
    g, o = _trees(sbh_gpu, pts, nrm, m, lo, hi, 2, 1.0)
    for spl in (300, 600):
        check_stochastic(sbh_gpu, g, o, q, "gravity", spl, 23, f"2^23+3001 spl{spl}",
                         no_split=True)
    check_stochastic(sbh_gpu, g, o, q, "gravity", 300, 29, "2^23+3001 spl300 split")


def test_work_counts(sbh_gpu):
    """SBH_EVAL_WORK_COUNTS (the instrumentation behind the benchmark's
    roofline): values unchanged; with one sample per subdomain nothing is
    reused, so the kernel terms are the reference counts plus one own term per
    internal subdomain and per evaluated path level; with many sorted samples
    the node cache evaluates far fewer terms than the reference sums; modes it
    does not cover are refused."""
    import sbh_capi as C

    pts, nrm, m, q, lo, hi = _inputs(6000, 700, 3, 191, "gravity")
    t = C.Tree(pts, None, m, lo, hi, split_log=2)
    nq = len(q)
    nsub_internal = int(np.count_nonzero(
        t_nodes_internal_children(t)))
    for spl in (1, 64):
        ref, _, rc = t.eval_host(C.STOCHASTIC, "gravity", q, samples_per_subdomain=spl, seed=4,
                                 flags=C.EVAL_SPLIT_COUNTS | C.EVAL_NO_SPLIT, counts=True)
        got, _, wc = t.eval_host(C.STOCHASTIC, "gravity", q, samples_per_subdomain=spl, seed=4,
                                 flags=C.EVAL_WORK_COUNTS, counts=True)
        np.testing.assert_array_equal(got, ref)
        ref_terms = rc[:nq] + rc[nq:]
        terms, levels = wc[:nq].astype(np.int64), wc[nq:].astype(np.int64)
        print(f"[work] spl {spl}: reference {ref_terms.sum()} terms, evaluated {terms.sum()}, "
              f"levels {levels.sum()}", flush=True)
        assert (levels >= 0).all() and (terms > 0).all()
        if spl == 1:
            assert (terms >= ref_terms).all()
            assert (terms <= ref_terms + nsub_internal + levels).all()
        else:
            assert terms.sum() * 4 < ref_terms.sum()
    for bad in (C.EVAL_EXACT_TERMS, C.EVAL_PER_QUERY_PATHS):
        with pytest.raises(C.SbhError):
            t.eval_host(C.STOCHASTIC, "gravity", q, samples_per_subdomain=2,
                        flags=C.EVAL_WORK_COUNTS | bad, counts=True)
    with pytest.raises(C.SbhError):
        t.eval_host(C.BARNES_HUT, "gravity", q, beta=2.0, flags=C.EVAL_WORK_COUNTS, counts=True)
    with pytest.raises(C.SbhError):
        t.eval_host(C.STOCHASTIC, "gravity", q, samples_per_subdomain=2, flags=C.EVAL_WORK_COUNTS)
    t.close()


def t_nodes_internal_children(t):
    """Per root child of a sbh_capi tree: 1 if it is an internal node."""
    import ctypes

    import sbh_capi as C

    nn = t.info()["num_nodes"]
    first = np.empty(nn, np.uint32)
    nch = np.empty(nn, np.uint32)
    C.check(C.lib().sbh_tree_nodes(t.h, None, None, first.ctypes.data, nch.ctypes.data,
                                   *([None] * 10)))
    kids = np.arange(first[0], first[0] + nch[0])
    return (nch[kids] > 0).astype(np.int64)
