"""The double build on the GPU (sbh_gpu64; the reference's USE_DOUBLE
configuration) against the oracle's double build on the same inputs.

Bit for bit: the tree (leaf order, ranges, float64 bounds, diameters and
moments), brute force, Barnes-Hut (vote groups 1, 32, 64, with counts) and the
stochastic estimator (wavefront-shared and per-query paths, with counts) for
gravity, the winding number and the gravity gradient, whose terms use only
correctly rounded operations on both sides.  The kernels with exp / cos / sin
(smooth distance, Helmholtz) are compared within 1e-13 of the sum's scale:
the device and host libraries round those functions differently."""
import numpy as np
import pytest

import oracle
import workloads as W

pytestmark = pytest.mark.gpu

BITWISE = ["gravity", "wn", "gravity_grad_x", "gravity_grad_z"]
LIBM = ["smoothdist", "helmholtz_x", "helmholtz_y"]


@pytest.fixture(scope="module")
def sbh_gpu64(sbh_gpu):
    import sbh_gpu64 as m
    assert m.fp == 64
    return m


def _cls(mod, dim, k):
    return getattr(mod, ("OctreeGPU" if dim == 3 else "QuadtreeGPU") + str(k))


CASES = [  # (name, dim, k, n, q)
    ("cube3d_k1", 3, 1, 3000, 300),
    ("cube3d_k2", 3, 2, 5000, 257),
    ("gmm2d_k2", 2, 2, 4000, 200),
    ("plummer3d_k3", 3, 3, 4000, 130),
]


def _inputs(name, dim, n, q):
    if name.startswith("gmm"):
        pts, nrm, m, qs = W.gmm2d(n, q, seed=5)
    elif name.startswith("plummer"):
        pts, nrm, m, qs = W.plummer(n, q, seed_sources=6, seed_queries=7, rmax=20.0)
    else:
        pts, nrm, m, qs = W.uniform_cube(n, q, seed=8)
    nrm = np.random.default_rng(10).normal(size=pts.shape)  # every case has normals
    lo, hi = W.bounds(pts)
    return (np.asarray(pts, np.float64), np.asarray(nrm, np.float64), np.asarray(m, np.float64),
            np.asarray(qs, np.float64), np.asarray(lo, np.float64), np.asarray(hi, np.float64))


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request, sbh_gpu64):
    name, dim, k, n, q = request.param
    pts, nrm, m, qs, lo, hi = _inputs(name, dim, n, q)
    print(f"[case] double build {name} n={n} q={q} k={k}", flush=True)
    g = _cls(sbh_gpu64, dim, k)(pts, nrm, m, lo, hi, 2.0)
    o = oracle.Tree(pts, nrm, m, lo, hi, alpha=2.0, split_log=k, dtype=np.float64)
    return dim, k, qs, g, o


def _same_bits(a, b):
    np.testing.assert_array_equal(np.asarray(a, np.float64).view(np.uint64),
                                  np.asarray(b, np.float64).view(np.uint64))


def test_double_tree_is_the_oracle_tree(case):
    dim, k, _, g, o = case
    np.testing.assert_array_equal(g.leaf_order(), o.leaf_order())
    gn, on = g.nodes(), o.nodes()
    assert len(gn["start"]) == len(on["start"])
    for key in ("start", "end", "nchild", "depth"):
        np.testing.assert_array_equal(gn[key].astype(np.int64), on[key].astype(np.int64), key)
    for key in ("bmin", "bmax", "diam", "centroid", "mass", "normal"):
        assert gn[key].dtype == np.float64
        _same_bits(gn[key], on[key])


def _kernels(dim):
    return [kk for kk in BITWISE + LIBM if not (kk == "gravity_grad_z" and dim == 2)]


def _compare(kernel, got, ref, scale):
    if kernel in BITWISE:
        _same_bits(got, ref)
    else:
        assert (np.abs(got - ref) <= 1e-13 * scale).all(), np.abs(got - ref).max()


def test_double_brute_force(sbh_gpu64, case):
    dim, k, qs, g, o = case
    for kern in _kernels(dim):
        got, ms = getattr(sbh_gpu64, "eval_brute_force_" + kern)(g, qs)
        assert got.dtype == np.float64 and ms >= 0
        ref = o.brute_force(kern, qs)
        _compare(kern, got, ref, np.abs(ref).max() + 1.0)


@pytest.mark.parametrize("group", [1, 32, 64])
def test_double_barnes_hut(sbh_gpu64, case, group):
    dim, k, qs, g, o = case
    for beta in (1.0, 2.0, np.inf):
        for kern in _kernels(dim):
            got, _, cnt = getattr(sbh_gpu64, "eval_barnes_hut_" + kern)(
                g, qs, beta, vote_group=group, return_counts=True)
            ref, rc = o.barnes_hut(kern, qs, beta, vote_group=group, counts=True)
            np.testing.assert_array_equal(cnt, rc)
            _compare(kern, got, ref, np.abs(ref).max() + 1.0)


@pytest.mark.parametrize("per_query", [False, True])
def test_double_stochastic(sbh_gpu64, case, per_query):
    dim, k, qs, g, o = case
    for kern in _kernels(dim):
        for spl, seed in ((1, 0), (4, 12345)):
            got, _, cnt = getattr(sbh_gpu64, "multi_level_prefix_control_variate_" + kern)(
                g, qs, spl, seed=seed, return_counts=True, per_query_paths=per_query)
            ref, rc = o.stochastic(kern, qs, spl, seed=seed, group=1 if per_query else 64,
                                   counts=True)
            np.testing.assert_array_equal(cnt, rc)
            _compare(kern, got, ref, np.abs(ref).max() + 1.0)


def _launch_order(g32, qs):
    """The spatial order SBH_EVAL_SORT_QUERIES uses (Hilbert order of the
    queries' float rounding in the tree's float-rounded root cell)."""
    import ctypes

    import sbh_capi as C

    q32 = np.ascontiguousarray(qs.astype(np.float32).T)
    order = np.empty(len(qs), np.uint32)
    C.check(C.lib().sbh_query_order(g32.h, q32.ctypes.data, len(qs), order.ctypes.data, 0))
    return order.astype(np.int64)


def test_double_sorted_queries(sbh_gpu64, case):
    """Sorting the launch changes nothing per query (vote group 1, per-query
    paths) and regroups the shared forms exactly as the oracle does on the
    queries taken in launch order."""
    import sbh_capi as C

    dim, k, qs, g, o = case
    g32 = C.Tree(o.points, None, o.masses, o.min_corner, o.max_corner, split_log=k)
    order = _launch_order(g32, qs)
    assert sorted(order.tolist()) == list(range(len(qs)))
    for kern in ("gravity", "wn"):
        bh = getattr(sbh_gpu64, "eval_barnes_hut_" + kern)
        a, _ = bh(g, qs, 2.0, vote_group=1)
        b, _ = bh(g, qs, 2.0, vote_group=1, sort_queries=True)
        _same_bits(a, b)
        c, _ = bh(g, qs, 2.0, vote_group=32, sort_queries=True)
        ref = o.barnes_hut(kern, qs[order], 2.0, vote_group=32)
        _same_bits(c[order], ref)
        st = getattr(sbh_gpu64, "multi_level_prefix_control_variate_" + kern)
        a, _ = st(g, qs, 2, seed=5, per_query_paths=True)
        b, _ = st(g, qs, 2, seed=5, per_query_paths=True, sort_queries=True)
        _same_bits(a, b)
        c, _ = st(g, qs, 2, seed=5, sort_queries=True)
        ref = o.stochastic(kern, qs[order], 2, seed=5, group=64)
        _same_bits(c[order], ref)
    g32.close()


def test_double_smooth_distance_transform(sbh_gpu64, case):
    dim, k, qs, g, o = case
    w, _ = sbh_gpu64.eval_barnes_hut_smoothdist(g, qs, 2.0, vote_group=1)
    d, _ = sbh_gpu64.eval_barnes_hut_smoothdist(g, qs, 2.0, vote_group=1, smooth_distance=True)
    np.testing.assert_allclose(d, -np.log(w) / 2.0, rtol=1e-14, atol=1e-14)


def test_double_build_is_more_accurate_than_float(sbh_gpu, sbh_gpu64):
    """Barnes-Hut at beta = inf sums every point: the double build's error
    against the exactly rounded sum is far below the float build's."""
    pts, nrm, m, qs, lo, hi = _inputs("cube3d_k2", 3, 20000, 512)
    g64 = sbh_gpu64.OctreeGPU2(pts, nrm, m, lo, hi)
    g32 = sbh_gpu.OctreeGPU2(pts, nrm, m, lo, hi)
    ref, _ = sbh_gpu.eval_brute_force_f64_gravity(g32, qs)  # float64 sum of fp32 inputs
    o = oracle.Tree(pts, nrm, m, lo, hi, split_log=2, dtype=np.float64)
    exact64 = o.brute_force("gravity", qs)
    d64, _ = sbh_gpu64.eval_barnes_hut_gravity(g64, qs, np.inf, vote_group=1)
    d32, _ = sbh_gpu.eval_barnes_hut_gravity(g32, qs, np.inf, vote_group=1, exact_terms=True)
    e64 = np.abs(d64 - exact64) / np.abs(exact64)
    e32 = np.abs(d32 - ref) / np.abs(ref)
    assert e64.max() < 1e-12 and e32.max() > 1e-9


def test_double_tree_refused_by_float_entry_points(sbh_gpu, sbh_gpu64, tmp_path):
    import sbh_capi as C

    pts = np.random.default_rng(3).random((100, 3))
    g = sbh_gpu64.OctreeGPU1(pts, np.zeros_like(pts), np.ones(100), [0, 0, 0], [1, 1, 1])
    info = g.info()
    assert info["num_points"] == 100 and info["device_bytes"] > 100 * 64
    with pytest.raises(ValueError):
        sbh_gpu64.eval_barnes_hut_gravity(g, pts[:, :2], 2.0)
    with pytest.raises(TypeError):  # a double tree is not an sbh_gpu tree
        sbh_gpu.eval_barnes_hut_gravity(g, pts.astype(np.float32), 2.0)
    # the C ABI refuses a double tree in the float entry points and vice versa
    L = C.lib()
    t32 = C.Tree(pts, None, np.ones(100), [0, 0, 0], [1, 1, 1])
    out = np.zeros(100)
    p = C.EvalParamsDouble(beta=2.0, samples_per_subdomain=1, seed=0, vote_group=1, flags=0)
    q = np.ascontiguousarray(pts.T)
    assert L.sbh_eval_double(t32.h, C.BARNES_HUT, 0, p, q.ctypes.data, 100, out.ctypes.data,
                             None, None) == C.ERR_INVALID
    assert "double-build" in L.sbh_last_error().decode()


@pytest.mark.parametrize("shape", ["leaf4", "duplicates", "single_point"])
def test_double_edge_trees(sbh_gpu64, shape):
    """Leaves of several points, duplicates that reach the depth limit, and a
    one-point tree: the double build still equals the oracle bit for bit."""
    rng = np.random.default_rng(77)
    leaf = 1
    if shape == "leaf4":
        pts = rng.random((1500, 3))
        leaf = 4
    elif shape == "duplicates":
        base = rng.random((60, 3))
        pts = np.concatenate([base, base[:12], base[:5]])
    else:
        pts = np.array([[0.3, 0.6, 0.2]])
    n = len(pts)
    nrm = rng.normal(size=pts.shape)
    m = rng.uniform(0.5, 1.5, n)
    lo, hi = np.full(3, -0.01), np.full(3, 1.01)
    qs = rng.random((100, 3)) * 1.4 - 0.2
    print(f"[case] double edge tree {shape} n={n} leaf={leaf}", flush=True)
    g = sbh_gpu64.OctreeGPU1(pts, nrm, m, lo, hi, 1.0, leaf)
    o = oracle.Tree(pts, nrm, m, lo, hi, split_log=1, leaf_size=leaf, dtype=np.float64)
    np.testing.assert_array_equal(g.leaf_order(), o.leaf_order())
    assert g.info()["depth"] == o.depth
    for kern in ("gravity", "wn"):
        _same_bits(getattr(sbh_gpu64, "eval_brute_force_" + kern)(g, qs)[0],
                   o.brute_force(kern, qs))
        for group in (1, 64):
            got, _, cnt = getattr(sbh_gpu64, "eval_barnes_hut_" + kern)(
                g, qs, 2.0, vote_group=group, return_counts=True)
            ref, rc = o.barnes_hut(kern, qs, 2.0, vote_group=group, counts=True)
            _same_bits(got, ref)
            np.testing.assert_array_equal(cnt, rc)
        for per_query in (False, True):
            got, _, cnt = getattr(sbh_gpu64, "multi_level_prefix_control_variate_" + kern)(
                g, qs, 3, seed=4, return_counts=True, per_query_paths=per_query)
            ref, rc = o.stochastic(kern, qs, 3, seed=4, group=1 if per_query else 64,
                                   counts=True)
            _same_bits(got, ref)
            np.testing.assert_array_equal(cnt, rc)
    # no queries: empty results, no launch
    empty, ms = sbh_gpu64.eval_barnes_hut_gravity(g, np.zeros((0, 3)), 2.0)
    assert empty.shape == (0,) and ms == 0.0


def test_double_build_through_ctypes_device_pointers(sbh_gpu64, case):
    """The ctypes binding's double-build handle with device-resident float64
    buffers (SBH_EVAL_DEVICE_POINTERS) gives the module's results."""
    import torch

    import sbh_capi as C

    dim, k, qs, g, o = case
    t = C.TreeDouble(o.points, o.normals, o.masses, o.min_corner, o.max_corner, alpha=2.0,
                     split_log=k)
    dev = torch.device("cuda:0")
    qd = torch.from_numpy(np.ascontiguousarray(qs.T)).to(dev)
    out = torch.empty(len(qs), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    t.eval_device(C.BARNES_HUT, "gravity", qd.data_ptr(), len(qs), out.data_ptr(), beta=2.0,
                  vote_group=32)
    torch.cuda.synchronize()
    ref, _ = sbh_gpu64.eval_barnes_hut_gravity(g, qs, 2.0, vote_group=32)
    _same_bits(out.cpu().numpy(), ref)
    st, _ = t.eval_host(C.STOCHASTIC, "wn", qs, samples_per_subdomain=2, seed=3)
    ref, _ = sbh_gpu64.multi_level_prefix_control_variate_wn(g, qs, 2, seed=3)
    _same_bits(st, ref)
    t.close()
