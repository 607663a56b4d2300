"""Serving engine (sbh_server_* through sbh_gpu.Server): every request's
values are identical alone and in a mix with other trees, kernels and
methods; admission never exceeds the HBM budget; queued requests can be
cancelled; an oversized request is refused."""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trees(sbh_gpu):
    rng = np.random.default_rng(0)
    out = []
    p, n, m, _ = W.plummer(20000, 1, rmax=10.0)
    lo, hi = W.bounds(p)
    out.append(("gravity", sbh_gpu.OctreeGPU2(p, n, m, lo, hi), lo, hi, 3))
    p, n, m, _ = W.noisy_icosphere(4, 1)
    lo, hi = W.bounds(p)
    out.append(("wn", sbh_gpu.OctreeGPU1(p, n, m, lo, hi), lo, hi, 3))
    p, n, m, _ = W.gmm2d(15000, 1)
    lo, hi = W.bounds(p)
    out.append(("gravity", sbh_gpu.QuadtreeGPU2(p, n, m, lo, hi), lo, hi, 2))
    p, n, m, _ = W.uniform_cube(12000, 1, seed=3)
    lo, hi = W.bounds(p)
    out.append(("smoothdist", sbh_gpu.OctreeGPU2(p, n, m, lo, hi, alpha=10.0), lo, hi, 3))
    return out


def make_requests(trees, count, seed, max_q=4000):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        ti = int(rng.integers(len(trees)))
        kern, tree, lo, hi, dim = trees[ti]
        q = int(rng.integers(1, max_q))
        pts = lo + (hi - lo) * rng.random((q, dim))
        method = "stochastic" if rng.random() < 0.5 else "barnes_hut"
        reqs.append(dict(ti=ti, kernel=kern, method=method, queries=pts,
                         beta=float(rng.choice([1.5, 2.0, 3.0])),
                         spl=int(rng.integers(1, 4)), seed=int(rng.integers(1 << 40))))
    return reqs


def one_shot(sbh_gpu, trees, r):
    kern, tree, *_ = trees[r["ti"]]
    if r["method"] == "barnes_hut":
        fn = getattr(sbh_gpu, "eval_barnes_hut_" + kern)
        return fn(tree, r["queries"], r["beta"], vote_group=1)[0]
    fn = getattr(sbh_gpu, "multi_level_prefix_control_variate_" + kern)
    # the engine evaluates each request in its queries' spatial order
    return fn(tree, r["queries"], r["spl"], seed=r["seed"], sort_queries=True)[0]


def test_results_identical_alone_and_in_mix(sbh_gpu, trees):
    srv = sbh_gpu.Server(64 << 20, 4)
    ids = [srv.add_tree(t[1]) for t in trees]
    reqs = make_requests(trees, 40, seed=1)
    handles = []
    for r in reqs:
        handles.append(srv.submit(ids[r["ti"]], r["method"], r["kernel"], r["queries"],
                                  beta=r["beta"], samples_per_subdomain=r["spl"], seed=r["seed"]))
    solo = sbh_gpu.Server(64 << 20, 1)
    solo_ids = [solo.add_tree(t[1]) for t in trees]
    for r, h in zip(reqs, handles):
        got, lat = h.result(60.0)
        assert lat > 0
        # the same request served alone: identical bits
        a = solo.submit(solo_ids[r["ti"]], r["method"], r["kernel"], r["queries"],
                        beta=r["beta"], samples_per_subdomain=r["spl"], seed=r["seed"])
        np.testing.assert_array_equal(got, a.result(60.0)[0])
        # and the one-shot evaluation: the same terms and sample paths, summed
        # in whole-query order by the engine and per root subtree or
        # subdomain by a one-shot launch this small
        ref = one_shot(sbh_gpu, trees, r)
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["peak_bytes"] <= st["budget_bytes"]
    # several launches overlapped or queued: the mix was batched
    assert st["launches"] >= 1


def test_admission_by_budget_and_cancel(sbh_gpu, trees):
    kern, tree, lo, hi, dim = trees[0]
    q = 3000
    rng = np.random.default_rng(5)
    pts = lo + (hi - lo) * rng.random((q, dim))
    # one request's working set, as the engine accounts it
    probe = sbh_gpu.Server(64 << 20, 1)
    probe.add_tree(tree)
    probe.submit(0, "stochastic", kern, pts, samples_per_subdomain=1).wait(60.0)
    per_request = probe.stats()["peak_bytes"]
    assert q * (dim + 1) * 4 < per_request < q * (dim + 12) * 4
    del probe
    srv = sbh_gpu.Server(2 * per_request, 1)
    tid = srv.add_tree(tree)
    hs = [srv.submit(tid, "stochastic", kern, pts, samples_per_subdomain=8, seed=i)
          for i in range(12)]
    cancelled = srv.cancel(hs[-1])
    for h in hs[:-1]:
        assert h.wait(60.0) == sbh_gpu.REQ_DONE
    assert hs[-1].wait(60.0) == (sbh_gpu.REQ_CANCELLED if cancelled else sbh_gpu.REQ_DONE)
    st = srv.stats()
    assert st["peak_bytes"] <= 2 * per_request
    with pytest.raises(RuntimeError, match="budget"):
        srv.submit(tid, "barnes_hut", kern, lo + (hi - lo) * rng.random((3 * q, dim)))


def test_invalid_requests(sbh_gpu, trees):
    srv = sbh_gpu.Server(1 << 20, 2)
    tid = srv.add_tree(trees[0][1])
    pts = np.random.default_rng(1).random((10, 3))
    with pytest.raises(ValueError):
        srv.submit(tid, "brute_force", "gravity", pts)
    with pytest.raises(ValueError):
        srv.submit(tid + 5, "barnes_hut", "gravity", pts)
    with pytest.raises(ValueError):
        srv.submit(tid, "stochastic", "gravity", pts, samples_per_subdomain=0)
