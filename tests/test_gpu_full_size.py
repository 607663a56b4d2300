"""The headline workload at its full size (untruncated Plummer sphere,
N = Q = 2^20, octree k = 2): properties that do not depend on the size.

* the tree (leaf order, every node range, count, depth and fp32 moment) is
  the oracle's tree, bit for bit, at 2^20 points;
* the stochastic estimator at 512 samples per subdomain and Barnes-Hut at
  theta 1.0 stay under the 1e-2 mean relative error of the metric against
  the float64 all-pairs sum (GPU float64 kernel, pinned against the oracle at
  small sizes) on 16384 queries;
* Barnes-Hut per query is independent of the query order (Hilbert-sorted and
  unsorted launches agree bit for bit);
* the estimator is a function of (seed, query group): the same call twice
  gives the same bits, another seed gives a different sample;
* the estimator's default launch, split over ranges of root subdomains,
  agrees with its whole-walk launch to fp32 rounding.
"""
import numpy as np
import pytest

import oracle
import workloads as W

pytestmark = pytest.mark.gpu

N = 1 << 20
ERR_Q = 16384


@pytest.fixture(scope="module")
def headline(sbh_gpu):
    pts, nrm, m, q = W.plummer(N, N, seed_sources=1, seed_queries=2)
    lo, hi = W.bounds(pts)
    print(f"[case] full size Plummer n={N} q={N} k=2", flush=True)
    g = sbh_gpu.OctreeGPU2(pts, nrm, m, lo, hi)
    ref, _ = sbh_gpu.eval_brute_force_f64_gravity(g, q[:ERR_Q])
    return pts, nrm, m, lo, hi, q, g, np.asarray(ref, np.float64)


def _err(got, ref):
    return float(np.mean(np.abs(np.asarray(got, np.float64) - ref) / np.abs(ref)))


def test_full_size_tree_is_the_oracle_tree(headline):
    pts, nrm, m, lo, hi, _, g, _ = headline
    o = oracle.Tree(pts, nrm, m, lo, hi, split_log=2)
    np.testing.assert_array_equal(g.leaf_order(), o.leaf_order())
    gn, on = g.nodes(), o.nodes()
    assert len(gn["start"]) == len(on["start"])
    for key in ("start", "end", "nchild", "depth"):
        np.testing.assert_array_equal(gn[key].astype(np.int64), on[key].astype(np.int64), key)
    for key in ("centroid", "mass", "diam"):
        np.testing.assert_array_equal(gn[key].view(np.uint32), on[key].view(np.uint32), key)


def test_full_size_errors_within_the_metric(sbh_gpu, headline):
    *_, q, g, ref = headline
    st, _ = sbh_gpu.multi_level_prefix_control_variate_gravity(g, q, 512, seed=1,
                                                                sort_queries=True)
    assert st.shape == (N,) and np.isfinite(st).all()
    e = _err(st[:ERR_Q], ref)
    print(f"[case] stochastic k=2 spl=512 mean rel err {e:.3e}", flush=True)
    assert e <= 1e-2
    bh, _ = sbh_gpu.eval_barnes_hut_gravity(g, q, 1.0, vote_group=1, sort_queries=True)
    e = _err(bh[:ERR_Q], ref)
    print(f"[case] barnes-hut theta=1 mean rel err {e:.3e}", flush=True)
    assert e <= 1e-2


def test_full_size_barnes_hut_independent_of_query_order(sbh_gpu, headline):
    *_, q, g, _ = headline
    a, _ = sbh_gpu.eval_barnes_hut_gravity(g, q, 1.0, vote_group=1, sort_queries=True)
    b, _ = sbh_gpu.eval_barnes_hut_gravity(g, q, 1.0, vote_group=1, sort_queries=False)
    np.testing.assert_array_equal(a, b)


def test_full_size_estimator_is_a_function_of_the_seed(sbh_gpu, headline):
    *_, q, g, _ = headline
    a, _ = sbh_gpu.multi_level_prefix_control_variate_gravity(g, q, 16, seed=7)
    b, _ = sbh_gpu.multi_level_prefix_control_variate_gravity(g, q, 16, seed=7)
    c, _ = sbh_gpu.multi_level_prefix_control_variate_gravity(g, q, 16, seed=8)
    np.testing.assert_array_equal(a, b)
    assert np.mean(a != c) > 0.5


def test_full_size_estimator_split_matches_whole_walk(sbh_gpu, headline):
    """The default sorted launch (tasks per 64-query block and range of root
    subdomains, heaviest ranges first) against the whole-walk launch: the
    same sample paths and terms, the parts added in subdomain order, so
    equal to fp32 rounding."""
    *_, q, g, _ = headline
    a, _ = sbh_gpu.multi_level_prefix_control_variate_gravity(g, q, 32, seed=5, sort_queries=True)
    b, _ = sbh_gpu.multi_level_prefix_control_variate_gravity(g, q, 32, seed=5, sort_queries=True,
                                                              no_split=True)
    err = np.abs(a.astype(np.float64) - b) / np.abs(b.astype(np.float64))
    print(f"[case] split vs whole: mean {err.mean():.2e} max {err.max():.2e}", flush=True)
    assert err.mean() <= 1e-6 and err.max() <= 1e-4
