"""Serving engine (sbh_server_* through sbh_gpu.Server): every request's
values equal the oracle's (Barnes-Hut per query within 1e-5; the estimator
on the request's Hilbert launch order with its 64-query path groups, on the
same sample paths, within 1e-5), are identical alone and in a mix with other
trees, kernels and methods, and are bit-identical to a one-shot sorted
launch of the whole-walk kernels; admission never exceeds the HBM budget;
queued requests can be cancelled; an oversized request is refused."""
import numpy as np
import pytest

import oracle
import workloads as W
from test_gpu_whole_walk import check_same_paths, launch_order, oracle_in_order, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def trees(sbh_gpu):
    """(kernel, device tree, lo, hi, dim, oracle tree, points, masses)."""
    out = []

    def add(kern, cls, k, p, n, m, alpha=1.0):
        lo, hi = W.bounds(p)
        g = cls(p, n, m, lo, hi, alpha=alpha)
        o = oracle.Tree(p, n, m, lo, hi, alpha=alpha, split_log=k)
        out.append((kern, g, lo, hi, p.shape[1], o, p, m))

    p, n, m, _ = W.plummer(20000, 1, rmax=10.0)
    add("gravity", sbh_gpu.OctreeGPU2, 2, p, n, m)
    p, n, m, _ = W.noisy_icosphere(4, 1)
    add("wn", sbh_gpu.OctreeGPU1, 1, p, n, m)
    p, n, m, _ = W.gmm2d(15000, 1)
    add("gravity", sbh_gpu.QuadtreeGPU2, 2, p, n, m)
    p, n, m, _ = W.uniform_cube(12000, 1, seed=3)
    add("smoothdist", sbh_gpu.OctreeGPU2, 2, p, n, m, alpha=10.0)
    return out


def make_requests(trees, count, seed, max_q=4000):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        ti = int(rng.integers(len(trees)))
        kern, tree, lo, hi, dim = trees[ti][:5]
        q = int(rng.integers(1, max_q))
        pts = lo + (hi - lo) * rng.random((q, dim))
        method = "stochastic" if rng.random() < 0.5 else "barnes_hut"
        reqs.append(dict(ti=ti, kernel=kern, method=method, queries=pts,
                         beta=float(rng.choice([1.5, 2.0, 3.0])),
                         spl=int(rng.integers(1, 4)), seed=int(rng.integers(1 << 40)),
                         no_split=bool(rng.random() < 0.25)))
    return reqs


def one_shot(sbh_gpu, trees, r, no_split=False):
    """The request as one sorted launch (its walks split into tasks as the
    one-shot launch splits them, unless no_split), with term counts:
    (values, counts)."""
    kern, tree, *_ = trees[r["ti"]]
    if r["method"] == "barnes_hut":
        fn = getattr(sbh_gpu, "eval_barnes_hut_" + kern)
        v, _, c = fn(tree, r["queries"], r["beta"], vote_group=1, sort_queries=True,
                     no_split=no_split, return_counts=True)
        return v, c
    fn = getattr(sbh_gpu, "multi_level_prefix_control_variate_" + kern)
    # the engine evaluates each request in its queries' spatial order
    v, _, c = fn(tree, r["queries"], r["spl"], seed=r["seed"], sort_queries=True,
                 no_split=no_split, return_counts=True)
    return v, c


def check_against_oracle(trees, r, got, cnt, label):
    """Served values against the oracle: Barnes-Hut per query within 1e-5
    with identical counts; the estimator on the request's launch order, on
    the same sample paths (counts of the one-shot twin), within 1e-5."""
    kern, _, lo, hi, _, o, pts, m = trees[r["ti"]]
    q = r["queries"]
    if r["method"] == "barnes_hut":
        ref, rcnt = o.barnes_hut(kern, q, r["beta"], vote_group=1, counts=True)
        np.testing.assert_array_equal(cnt, rcnt)
        err = rel_err(got, ref, kern)
        assert err.max() <= TOL, (label, err.max())
        return
    order = launch_order(pts, m, lo, hi, q)
    ref, rcnt = oracle_in_order(
        lambda qq: o.stochastic(kern, qq, r["spl"], seed=r["seed"], group=64, counts=True),
        q, order)
    check_same_paths(got, cnt, ref, rcnt, kern, label)


def test_results_identical_alone_and_in_mix(sbh_gpu, trees):
    srv = sbh_gpu.Server(64 << 20, 4)
    ids = [srv.add_tree(t[1]) for t in trees]
    reqs = make_requests(trees, 40, seed=1)
    handles = []
    for r in reqs:
        handles.append(srv.submit(ids[r["ti"]], r["method"], r["kernel"], r["queries"],
                                  beta=r["beta"], samples_per_subdomain=r["spl"], seed=r["seed"],
                                  no_split=r["no_split"]))
    solo = sbh_gpu.Server(64 << 20, 1)
    solo_ids = [solo.add_tree(t[1]) for t in trees]
    for i, (r, h) in enumerate(zip(reqs, handles)):
        got, lat = h.result(60.0)
        assert lat > 0
        # the same request served alone: identical bits
        a = solo.submit(solo_ids[r["ti"]], r["method"], r["kernel"], r["queries"],
                        beta=r["beta"], samples_per_subdomain=r["spl"], seed=r["seed"],
                        no_split=r["no_split"])
        np.testing.assert_array_equal(got, a.result(60.0)[0])
        # the one-shot sorted launch (split the same way): identical bits
        ref, cnt = one_shot(sbh_gpu, trees, r, no_split=r["no_split"])
        np.testing.assert_array_equal(got, ref)
        # and the oracle
        check_against_oracle(trees, r, got, cnt, f"request {i} {r['method']} "
                             f"{r['kernel']} q={len(r['queries'])}")
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["peak_bytes"] <= st["budget_bytes"]
    # several launches overlapped or queued: the mix was batched
    assert st["launches"] >= 1


def test_admission_by_budget_and_cancel(sbh_gpu, trees):
    kern, tree, lo, hi, dim = trees[0][:5]
    q = 3000
    rng = np.random.default_rng(5)
    pts = lo + (hi - lo) * rng.random((q, dim))
    # one request's working set, as the engine accounts it
    probe = sbh_gpu.Server(64 << 20, 1)
    probe.add_tree(tree)
    probe.submit(0, "stochastic", kern, pts, samples_per_subdomain=1, no_split=True).wait(60.0)
    whole = probe.stats()["peak_bytes"]
    assert q * (dim + 1) * 4 < whole < q * (dim + 12) * 4
    del probe
    # split over ranges of the root's children (one internal child each), a
    # request also holds its parts: a {Barnes-Hut, control-variate} float
    # pair per query and range
    probe = sbh_gpu.Server(64 << 20, 1)
    probe.add_tree(tree)
    probe.submit(0, "stochastic", kern, pts, samples_per_subdomain=1).wait(60.0)
    per_request = probe.stats()["peak_bytes"]
    nodes = tree.nodes()
    nch = np.asarray(nodes["nchild"])
    subs = np.arange(nodes["first_child"][0], nodes["first_child"][0] + nch[0])
    nranges = int(np.count_nonzero(nch[subs]))
    assert nranges >= 2
    assert 0 <= per_request - whole - nranges * q * 8 < 512
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
        srv.submit(tid, "stochastic", kern, lo + (hi - lo) * rng.random((3 * q, dim)))


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


def test_request_keeps_server_alive(sbh_gpu, trees):
    """A Request holds its Server: dropping the Server object first leaves the
    engine running until the request has been read."""
    import gc

    kern, tree, lo, hi, dim = trees[0][:5]
    srv = sbh_gpu.Server(64 << 20, 2)
    tid = srv.add_tree(tree)
    pts = lo + (hi - lo) * np.random.default_rng(8).random((500, dim))
    r = srv.submit(tid, "barnes_hut", kern, pts)
    del srv
    gc.collect()
    vals, lat = r.result(60.0)
    ref, _ = getattr(sbh_gpu, "eval_barnes_hut_" + kern)(tree, pts, 2.0, vote_group=1,
                                                          sort_queries=True)
    np.testing.assert_array_equal(vals, ref)


class _Req:
    """ctypes twin of sbh_request (include/sbh_hip.h)."""

    @staticmethod
    def make(tree_id, method, kernel, params, q_ptr, nq, out_ptr):
        import ctypes

        import sbh_capi as C

        class R(ctypes.Structure):
            _fields_ = [("tree_id", ctypes.c_int32), ("method", ctypes.c_int32),
                        ("kernel", ctypes.c_int32), ("params", C.EvalParams),
                        ("queries", ctypes.c_void_p), ("q", ctypes.c_int64),
                        ("results", ctypes.c_void_p)]

        return R(tree_id, method, kernel, params, q_ptr, nq, out_ptr)


def test_c_abi_server_flags_and_shutdown(sbh_gpu):
    """Through the C ABI: eval flags the engine does not implement are refused
    (SBH_ERR_INVALID) instead of ignored; destroying a server with requests
    still queued cancels them and returns."""
    import ctypes

    import sbh_capi as C

    L = C.lib()
    L.sbh_server_create.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
    L.sbh_server_destroy.argtypes = [ctypes.c_void_p]
    L.sbh_server_add_tree.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.POINTER(ctypes.c_int32)]
    L.sbh_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.POINTER(ctypes.c_int64)]
    pts, _, m, q = W.uniform_cube(5000, 4000, seed=17)
    lo, hi = W.bounds(pts)
    t = C.Tree(pts, None, m, lo, hi, split_log=2)
    soa = np.ascontiguousarray(q.astype(np.float32).T)
    outs = [np.empty(len(q), np.float32) for _ in range(16)]
    srv = ctypes.c_void_p()
    C.check(L.sbh_server_create(1 << 20, 1, ctypes.byref(srv)))  # room for one request
    tid = ctypes.c_int32()
    C.check(L.sbh_server_add_tree(srv, t.h, ctypes.byref(tid)))
    rid = ctypes.c_int64()
    for bad in (C.EVAL_EXACT_TERMS, C.EVAL_PER_QUERY_PATHS, C.EVAL_SMOOTH_DISTANCE,
                C.EVAL_SPLIT_COUNTS, C.EVAL_DEVICE_POINTERS):
        p = C.EvalParams(2.0, 4, 1, 1, bad)
        r = _Req.make(tid.value, C.STOCHASTIC, C.GRAVITY, p, soa.ctypes.data, len(q),
                      outs[0].ctypes.data)
        assert L.sbh_server_submit(srv, ctypes.addressof(r), ctypes.byref(rid)) == C.ERR_INVALID
    keep = []
    for i in range(16):
        p = C.EvalParams(2.0, 64, i, 1, C.EVAL_SORT_QUERIES | C.EVAL_NO_SPLIT)
        r = _Req.make(tid.value, C.STOCHASTIC, C.GRAVITY, p, soa.ctypes.data, len(q),
                      outs[i].ctypes.data)
        keep.append(r)
        C.check(L.sbh_server_submit(srv, ctypes.addressof(r), ctypes.byref(rid)))
    C.check(L.sbh_server_destroy(srv))  # cancels what is still queued, waits for the rest
    t.close()
