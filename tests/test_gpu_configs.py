"""The workload configurations the measurements are quoted on, checked
against the oracle on the same seeded inputs:

  * the reference's CPU-runnable case (3D gravity, 16384 uniform sources with
    masses U[0.5, 1.5], 16384 queries in [-0.1, 1.1]^3) at full size:
    Barnes-Hut at beta = 2 with leaf size 8 (include/tree3d.h:20,34-36) and
    leaf size 1, vote groups 1 and 32, exact and fast terms; the estimator at
    256 samples per query (octree k = 2, 4 per subdomain);
  * the 2D variant of the 1M-source GPU case (8-component Gaussian mixture,
    include/tree2d.h:237-365) in fp32 at reduced size: brute force,
    Barnes-Hut and the estimator on QuadtreeGPU1 / QuadtreeGPU2, in input and
    Hilbert launch order.
"""
import numpy as np
import pytest

import oracle
import workloads as W
from test_gpu_whole_walk import (TOL, check_bh, check_same_paths, check_stochastic,
                                 launch_order, rel_err)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cube16k():
    pts, nrm, m, q = W.uniform_cube(16384, 16384, seed=0)
    lo, hi = W.bounds(pts)
    return pts, np.zeros_like(pts), m, q, lo, hi


@pytest.mark.parametrize("leaf", [8, 1])
def test_cpu_case_barnes_hut(sbh_gpu, cube16k, leaf):
    pts, nrm, m, q, lo, hi = cube16k
    print(f"[case] cube 16384 leaf={leaf}", flush=True)
    g = sbh_gpu.OctreeGPU1(pts, nrm, m, lo, hi, leaf_size=leaf)
    o = oracle.Tree(pts, nrm, m, lo, hi, split_log=1, leaf_size=leaf)
    np.testing.assert_array_equal(g.leaf_order(), o.leaf_order())
    exact64 = o.brute_force("gravity", q, fp64=True)
    for vote in (1, 32):
        ref, rcnt = o.barnes_hut("gravity", q, 2.0, vote_group=vote, counts=True)
        for exact in (True, False):
            got, _, cnt = sbh_gpu.eval_barnes_hut_gravity(g, q, 2.0, vote_group=vote,
                                                          return_counts=True, exact_terms=exact)
            np.testing.assert_array_equal(cnt, rcnt)
            err = rel_err(got, ref, "gravity")
            print(f"[bh] leaf {leaf} vote {vote} exact {exact}: max {err.max():.2e}, "
                  f"vs float64 mean {rel_err(got, exact64, 'gravity').mean():.2e}", flush=True)
            assert err.max() <= TOL
            if exact:
                np.testing.assert_array_equal(got, ref)
    # the whole-walk launch and the Hilbert-sorted launch the benchmark uses
    a = check_bh(sbh_gpu, g, o, q, "gravity", 2.0, f"cube16k leaf{leaf}", no_split=True)
    b = check_bh(sbh_gpu, g, o, q, "gravity", 2.0, f"cube16k leaf{leaf} sorted",
                 sort_queries=True)
    assert rel_err(b, a, "gravity").max() <= TOL


def test_cpu_case_stochastic(sbh_gpu, cube16k):
    """256 samples per query: 4 per subdomain over the 64 root cells of the
    k = 2 octree."""
    pts, nrm, m, q, lo, hi = cube16k
    g = sbh_gpu.OctreeGPU2(pts, nrm, m, lo, hi)
    o = oracle.Tree(pts, nrm, m, lo, hi, split_log=2)
    assert int(g.nodes()["nchild"][0]) == 64
    ref, rcnt = o.stochastic("gravity", q, 4, seed=1, group=64, counts=True)
    got, _, cnt = sbh_gpu.multi_level_prefix_control_variate_gravity(
        g, q, 4, seed=1, return_counts=True, exact_terms=True)
    np.testing.assert_array_equal(cnt, rcnt)
    np.testing.assert_array_equal(got, ref)
    check_stochastic(sbh_gpu, g, o, q, "gravity", 4, 1, "cube16k spl4", no_split=True)
    order = launch_order(pts, m, lo, hi, q)
    check_stochastic(sbh_gpu, g, o, q, "gravity", 4, 1, "cube16k spl4 sorted", order=order,
                     sort_queries=True, no_split=True)
    # per-query paths (the reference CPU sampler's draws): bit-identical
    ref1, rc1 = o.stochastic("gravity", q, 4, seed=1, group=1, counts=True)
    got1, _, c1 = sbh_gpu.multi_level_prefix_control_variate_gravity(
        g, q, 4, seed=1, return_counts=True, exact_terms=True, per_query_paths=True)
    np.testing.assert_array_equal(c1, rc1)
    np.testing.assert_array_equal(got1, ref1)


@pytest.fixture(scope="module")
def gmm2d():
    pts, nrm, m, q = W.gmm2d(1 << 16, 4096, seed=3)
    lo, hi = W.bounds(pts)
    return pts, np.zeros_like(pts), m, q, lo, hi


@pytest.mark.parametrize("k", [1, 2])
def test_gmm2d_fp32(sbh_gpu, gmm2d, k):
    pts, nrm, m, q, lo, hi = gmm2d
    print(f"[case] gmm2d 65536 k={k}", flush=True)
    g = getattr(sbh_gpu, "QuadtreeGPU%d" % k)(pts, nrm, m, lo, hi)
    o = oracle.Tree(pts, nrm, m, lo, hi, split_log=k)
    np.testing.assert_array_equal(g.leaf_order(), o.leaf_order())
    bf, _ = sbh_gpu.eval_brute_force_gravity(g, q)
    assert rel_err(bf, o.brute_force("gravity", q), "gravity").max() <= TOL
    assert rel_err(bf, o.brute_force("gravity", q, fp64=True), "gravity").max() <= TOL
    for beta in (1.0, 2.0):
        for vote in (1, 32):
            ref, rcnt = o.barnes_hut("gravity", q, beta, vote_group=vote, counts=True)
            got, _, cnt = sbh_gpu.eval_barnes_hut_gravity(g, q, beta, vote_group=vote,
                                                          return_counts=True)
            np.testing.assert_array_equal(cnt, rcnt)
            assert rel_err(got, ref, "gravity").max() <= TOL
        check_bh(sbh_gpu, g, o, q, "gravity", beta, f"gmm2d k{k}", no_split=True,
                 sort_queries=True)
    order = launch_order(pts, m, lo, hi, q)
    for spl in (1, 4):
        ref, rcnt = o.stochastic("gravity", q, spl, seed=2, group=64, counts=True)
        got, _, cnt = sbh_gpu.multi_level_prefix_control_variate_gravity(
            g, q, spl, seed=2, return_counts=True, exact_terms=True)
        np.testing.assert_array_equal(cnt, rcnt)
        np.testing.assert_array_equal(got, ref)
        fast = sbh_gpu.multi_level_prefix_control_variate_gravity(g, q, spl, seed=2,
                                                                  return_counts=True)
        check_same_paths(fast[0], fast[2], ref, rcnt, "gravity", f"gmm2d k{k} spl{spl} split")
        check_stochastic(sbh_gpu, g, o, q, "gravity", spl, 2, f"gmm2d k{k} spl{spl} sorted",
                         order=order, sort_queries=True, no_split=True)
