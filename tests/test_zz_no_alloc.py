"""sbh_eval_noalloc, the entry point that makes no device allocation (its
scratch is a fixed per-GPU array), against sbh_eval through the C ABI: every
method, sorted and split launches, counts, on two calls in a row with the
second smaller, bit-identical results; evaluations that need more scratch
than the array holds, and host pointers, are refused."""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def _cases(C):
    bh = dict(beta=2.0)
    st = dict(samples_per_subdomain=3, seed=11)
    return [
        ("brute", C.BRUTE_FORCE, "gravity", 0, {}, False),
        ("brute_counts", C.BRUTE_FORCE, "gravity", C.EVAL_SPLIT_COUNTS, {}, True),
        ("brute_mfma", C.BRUTE_FORCE, "gravity", C.EVAL_MFMA, {}, False),
        ("brute_mfma_bf16", C.BRUTE_FORCE, "smoothdist", C.EVAL_MFMA_BF16, {}, False),
        ("bh_split", C.BARNES_HUT, "gravity", 0, bh, False),
        ("bh_split_counts", C.BARNES_HUT, "gravity", C.EVAL_SPLIT_COUNTS, bh, True),
        ("bh_sorted", C.BARNES_HUT, "wn", C.EVAL_SORT_QUERIES, bh, False),
        ("bh_vote32", C.BARNES_HUT, "gravity", 0, dict(beta=2.0, vote_group=32), True),
        ("bh_exact_sorted", C.BARNES_HUT, "gravity", C.EVAL_SORT_QUERIES | C.EVAL_EXACT_TERMS,
         bh, True),
        ("stoch_split", C.STOCHASTIC, "gravity", 0, st, False),
        ("stoch_split_counts", C.STOCHASTIC, "gravity", C.EVAL_SPLIT_COUNTS, st, True),
        ("stoch_sorted", C.STOCHASTIC, "gravity", C.EVAL_SORT_QUERIES, st, False),
        ("stoch_sorted_nosplit", C.STOCHASTIC, "smoothdist",
         C.EVAL_SORT_QUERIES | C.EVAL_NO_SPLIT | C.EVAL_SMOOTH_DISTANCE, st, False),
        ("stoch_per_query", C.STOCHASTIC, "gravity",
         C.EVAL_PER_QUERY_PATHS | C.EVAL_EXACT_TERMS, st, True),
        ("stoch_work", C.STOCHASTIC, "gravity", C.EVAL_WORK_COUNTS, st, True),
    ]


@pytest.fixture(scope="module")
def setup():
    import torch

    import sbh_capi as C

    pts, _, m, q = W.uniform_cube(20000, 40001, seed=5)
    lo, hi = W.bounds(pts)
    nrm = np.random.default_rng(6).normal(size=pts.shape)
    t = C.Tree(pts, nrm, m, lo, hi, alpha=10.0, split_log=2)
    p2, _, m2, q2 = W.gmm2d(30000, 40000, seed=7)
    lo2, hi2 = W.bounds(p2)
    t2 = C.Tree(p2, None, m2, lo2, hi2, split_log=2)
    yield C, torch, [(t, q), (t2, q2)]
    t.close()
    t2.close()


def _run(C, torch, tree, qs, nq, method, kern, flags, kw, counts, fn):
    dev = torch.device("cuda:0")
    qd = torch.from_numpy(np.ascontiguousarray(qs[:nq].astype(np.float32).T)).to(dev)
    out = torch.full((nq,), float("nan"), dtype=torch.float32, device=dev)
    cnt = torch.full((2 * nq,), -7, dtype=torch.int32, device=dev) if counts else None
    torch.cuda.synchronize()
    fn(tree, method, kern, qd.data_ptr(), nq, out.data_ptr(),
       counts_ptr=cnt.data_ptr() if counts else 0, flags=flags, **kw)
    torch.cuda.synchronize()
    return out.cpu().numpy(), (cnt.cpu().numpy() if counts else None)


def test_noalloc_matches_sbh_eval(setup):
    C, torch, trees = setup
    for ti, (tree, qs) in enumerate(trees):
        for name, method, kern, flags, kw, counts in _cases(C):
            if tree.dim == 2 and (flags & (C.EVAL_MFMA | C.EVAL_MFMA_BF16) or kern == "wn"):
                continue
            for nq in (len(qs), 5001):  # two calls in a row, the second smaller
                print(f"[noalloc] tree {ti} {name} nq={nq}", flush=True)
                a, ca = _run(C, torch, tree, qs, nq, method, kern, flags, kw, counts,
                             C.Tree.eval_device)
                b, cb = _run(C, torch, tree, qs, nq, method, kern, flags, kw, counts,
                             C.Tree.eval_device_noalloc)
                assert np.isfinite(a).all() or kern == "smoothdist"
                np.testing.assert_array_equal(b, a, err_msg=f"{name} nq={nq}")
                if counts:
                    np.testing.assert_array_equal(cb, ca, err_msg=f"{name} nq={nq}")


def test_noalloc_refusals(setup):
    C, torch, trees = setup
    tree, _ = trees[0]
    dev = torch.device("cuda:0")
    # a sorted launch whose order and sort work space exceed the fixed scratch
    nq = 10_500_000
    qd = torch.rand((3, nq), dtype=torch.float32, device=dev)
    out = torch.empty(nq, dtype=torch.float32, device=dev)
    with pytest.raises(C.SbhError, match="scratch") as e:
        tree.eval_device_noalloc(C.BARNES_HUT, "gravity", qd.data_ptr(), nq, out.data_ptr(),
                                 beta=2.0, flags=C.EVAL_SORT_QUERIES)
    assert e.value.code == C.ERR_INVALID
    # a launch split over the 64 root subdomains (all internal) with counts:
    # its partial sums (16 bytes per query and subdomain range) exceed the
    # array
    nq2 = 270001
    q2 = torch.rand((3, nq2), dtype=torch.float32, device=dev)
    o2 = torch.empty(nq2, dtype=torch.float32, device=dev)
    c2 = torch.empty(2 * nq2, dtype=torch.int32, device=dev)
    with pytest.raises(C.SbhError, match="partial sums"):
        tree.eval_device_noalloc(C.STOCHASTIC, "gravity", q2.data_ptr(), nq2, o2.data_ptr(),
                                 counts_ptr=c2.data_ptr(), samples_per_subdomain=2,
                                 flags=C.EVAL_SPLIT_COUNTS)
    # the whole-walk kernels need none
    tree.eval_device_noalloc(C.STOCHASTIC, "gravity", q2.data_ptr(), nq2, o2.data_ptr(),
                             counts_ptr=c2.data_ptr(), samples_per_subdomain=2,
                             flags=C.EVAL_SPLIT_COUNTS | C.EVAL_NO_SPLIT)
    torch.cuda.synchronize()
    assert torch.isfinite(o2).all().item()
    # the same launch in input order needs no scratch at all
    tree.eval_device_noalloc(C.BARNES_HUT, "gravity", qd.data_ptr(), nq, out.data_ptr(),
                             beta=2.0, flags=C.EVAL_NO_SPLIT)
    torch.cuda.synchronize()
    assert torch.isfinite(out).all().item()
    # host pointers would need staging buffers
    import ctypes

    q = np.random.default_rng(1).random((10, 3)).astype(np.float32)
    soa = np.ascontiguousarray(q.T)
    res = np.empty(10, np.float32)
    p = C.EvalParams(2.0, 1, 0, 1, 0)
    ms = ctypes.c_float()
    rc = C.lib().sbh_eval_noalloc(tree.h, C.BARNES_HUT, C.GRAVITY, ctypes.byref(p),
                                  soa.ctypes.data, 10, res.ctypes.data, None, ctypes.byref(ms))
    assert rc == C.ERR_INVALID
