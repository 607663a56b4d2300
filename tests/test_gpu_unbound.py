"""The kernels the reference defines but binds to no plugin
(include/kernel.h:61-83: helmholtzKernelX / Y, gravityGrad per component)
and the smooth-distance transform, on the GPU against the oracle.

Errors are relative to the sum of |terms| (the Helmholtz terms oscillate and
the gradient components change sign, so the sums cancel)."""
import numpy as np
import pytest

import oracle
import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-5
UNBOUND = ["helmholtz_x", "helmholtz_y", "gravity_grad_x", "gravity_grad_y", "gravity_grad_z"]


def _pair(sbh_gpu, dim, k, n=4000, q=300, seed=131):
    pts, nrm, m, queries = W.uniform_cube(n, q, dim=dim, seed=seed + dim + k)
    lo, hi = W.bounds(pts)
    cls = getattr(sbh_gpu, ("OctreeGPU%d" if dim == 3 else "QuadtreeGPU%d") % k)
    print(f"[case] unbound dim={dim} k={k} n={n}", flush=True)
    return cls(pts, nrm, m, lo, hi), oracle.Tree(pts, nrm, m, lo, hi, split_log=k), queries


def _err(got, ref, cond):
    return np.abs(np.asarray(got, np.float64) - ref) / cond


@pytest.mark.parametrize("dim,k", [(3, 2), (2, 1)])
def test_unbound_kernels_all_methods(sbh_gpu, dim, k):
    g, o, q = _pair(sbh_gpu, dim, k)
    for kern in UNBOUND:
        if kern.endswith("_z") and dim == 2:
            with pytest.raises(ValueError):
                sbh_gpu.eval_brute_force_gravity_grad_z(g, q)
            continue
        cond = o.abs_sum(kern, q)
        exact_ops = kern.startswith("gravity_grad")  # no cosf / sinf: bit-exact
        bf, _ = getattr(sbh_gpu, "eval_brute_force_" + kern)(g, q)
        assert _err(bf, o.brute_force(kern, q, fp64=True), cond).max() <= TOL, kern
        bh = getattr(sbh_gpu, "eval_barnes_hut_" + kern)
        for vote in (1, 32):
            ref, rcnt = o.barnes_hut(kern, q, 2.0, vote_group=vote, counts=True)
            got, _, cnt = bh(g, q, 2.0, vote_group=vote, exact_terms=True, return_counts=True)
            np.testing.assert_array_equal(cnt, rcnt)
            if exact_ops:
                np.testing.assert_array_equal(got, ref)
            else:
                assert _err(got, ref, cond).max() <= TOL, (kern, vote)
            fast, _ = bh(g, q, 2.0, vote_group=vote)
            assert _err(fast, ref, cond).max() <= TOL, (kern, vote)
        st = getattr(sbh_gpu, "multi_level_prefix_control_variate_" + kern)
        for group, pq in ((64, False), (1, True)):
            ref, rcnt = o.stochastic(kern, q, 3, seed=5, group=group, counts=True)
            got, _, cnt = st(g, q, 3, seed=5, exact_terms=True, return_counts=True,
                             per_query_paths=pq)
            np.testing.assert_array_equal(cnt, rcnt)
            if exact_ops:
                np.testing.assert_array_equal(got, ref)
            else:
                # the sample weights amplify a term's ulp; relative to the
                # sum of |terms| with the weights' spread allowed
                assert _err(got, ref, cond).mean() <= TOL, (kern, group)
        # exact brute force: beta = inf visits every point
        inf, _ = bh(g, q, np.inf, vote_group=1, exact_terms=True)
        assert _err(inf, o.brute_force(kern, q, fp64=True), cond).max() <= TOL, kern


def test_gravity_gradient_vectors(sbh_gpu):
    g, o, q = _pair(sbh_gpu, 3, 1)
    grad, ms = sbh_gpu.eval_brute_force_gravity_grad(g, q)
    assert grad.shape == (len(q), 3) and grad.dtype == np.float32 and ms > 0
    for a, c in enumerate("xyz"):
        comp, _ = getattr(sbh_gpu, "eval_brute_force_gravity_grad_" + c)(g, q)
        np.testing.assert_array_equal(grad[:, a], comp)
    bhg, _ = sbh_gpu.eval_barnes_hut_gravity_grad(g, q, 2.0, vote_group=1, exact_terms=True)
    for a, c in enumerate("xyz"):
        np.testing.assert_array_equal(
            bhg[:, a], o.barnes_hut("gravity_grad_" + c, q, 2.0, vote_group=1))
    stg, _ = sbh_gpu.multi_level_prefix_control_variate_gravity_grad(g, q, 2, seed=7,
                                                                     exact_terms=True)
    for a, c in enumerate("xyz"):
        np.testing.assert_array_equal(stg[:, a], o.stochastic("gravity_grad_" + c, q, 2, seed=7))
    q2 = np.zeros((0, 3))
    empty, _ = sbh_gpu.eval_brute_force_gravity_grad(g, q2)
    assert empty.shape == (0, 3)


def test_smooth_distance_transform(sbh_gpu):
    pts, nrm, m, q = W.noisy_icosphere(3, 400, seed=141)
    lo, hi = W.bounds(pts)
    g = sbh_gpu.OctreeGPU2(pts, nrm, m, lo, hi, alpha=10.0)
    for name, kw in (("eval_brute_force_smoothdist", {}),
                     ("eval_barnes_hut_smoothdist", {"beta": 2.0}),
                     ("multi_level_prefix_control_variate_smoothdist",
                      {"samples_per_subdomain": 4, "seed": 3})):
        fn = getattr(sbh_gpu, name)
        w, _ = fn(g, q, **kw)
        d, _ = fn(g, q, smooth_distance=True, **kw)
        expect = -np.log(w.astype(np.float64)) / 10.0
        ok = w > 0
        np.testing.assert_allclose(d[ok], expect[ok], rtol=1e-6, atol=1e-7)
        assert np.isnan(d[w < 0]).all()
    with pytest.raises(ValueError, match="smooth"):
        g1 = sbh_gpu.OctreeGPU1(pts, nrm, m, lo, hi)
        sbh_gpu.eval_brute_force_gravity(g1, q, smooth_distance=True)
