"""Tree images on the GPU: a saved and reloaded tree is the same tree (every
node array, the leaf order, every evaluation result), and damaged images are
refused before they reach the device."""
import os

import numpy as np
import pytest

import oracle
import tree_image as TI
import workloads as W

pytestmark = pytest.mark.gpu


def _tree(sbh_gpu, case):
    if case == "cube3d_k2":
        pts, nrm, m, q = W.uniform_cube(5000, 300, dim=3, seed=111)
        cls, kw = sbh_gpu.OctreeGPU2, {}
    elif case == "cube2d_k1_leaf4":
        pts, nrm, m, q = W.uniform_cube(4000, 300, dim=2, seed=112)
        cls, kw = sbh_gpu.QuadtreeGPU1, {"leaf_size": 4}
    else:  # icosphere with normals, float64 moments, smooth-distance alpha
        pts, nrm, m, q = W.noisy_icosphere(3, 300, seed=113)
        cls, kw = sbh_gpu.OctreeGPU3, {"moments64": True, "alpha": 7.0}
    lo, hi = W.bounds(pts)
    print(f"[case] tree image {case} n={len(pts)}", flush=True)
    return cls(pts, nrm, m, lo, hi, **kw), q


def _same_nodes(a, b):
    na, nb = a.nodes(), b.nodes()
    assert na.keys() == nb.keys()
    for key in na:
        np.testing.assert_array_equal(na[key], nb[key], key)
    np.testing.assert_array_equal(a.leaf_order(), b.leaf_order())


@pytest.mark.parametrize("case", ["cube3d_k2", "cube2d_k1_leaf4", "ico3d_k3_m64"])
def test_save_load_round_trip(sbh_gpu, tmp_path, case):
    t, q = _tree(sbh_gpu, case)
    path = os.path.join(tmp_path, "t.sbht")
    t.save(path)
    img = open(path, "rb").read()
    assert t.to_bytes() == img
    for u in (sbh_gpu.load_tree(path), sbh_gpu.tree_from_bytes(img)):
        assert type(u) is type(t)
        assert u.info()["num_nodes"] == t.info()["num_nodes"]
        _same_nodes(t, u)
        assert u.to_bytes() == img
        for kern in ("gravity", "wn", "smoothdist"):
            a, _ = getattr(sbh_gpu, "eval_brute_force_" + kern)(t, q)
            b, _ = getattr(sbh_gpu, "eval_brute_force_" + kern)(u, q)
            np.testing.assert_array_equal(a, b)
            a, _ = getattr(sbh_gpu, "eval_barnes_hut_" + kern)(t, q, 2.0, vote_group=1)
            b, _ = getattr(sbh_gpu, "eval_barnes_hut_" + kern)(u, q, 2.0, vote_group=1)
            np.testing.assert_array_equal(a, b)
            fn = getattr(sbh_gpu, "multi_level_prefix_control_variate_" + kern)
            np.testing.assert_array_equal(fn(t, q, 3, seed=4)[0], fn(u, q, 3, seed=4)[0])


def test_image_matches_documented_format_and_oracle(sbh_gpu):
    pts, nrm, m, _ = W.uniform_cube(3000, 1, dim=3, seed=114)
    lo, hi = W.bounds(pts)
    t = sbh_gpu.OctreeGPU1(pts, nrm, m, lo, hi)
    im = TI.parse(t.to_bytes())
    o = oracle.Tree(pts, nrm, m, lo, hi, split_log=1)
    np.testing.assert_array_equal(im["perm"].astype(np.int64), o.leaf_order())
    on = o.nodes()
    np.testing.assert_array_equal(im["n_start"].astype(np.int64), on["start"])
    np.testing.assert_array_equal(im["n_end"].astype(np.int64), on["end"])
    np.testing.assert_array_equal(im["nodeC"][:, 3].view(np.uint32), on["mass"].view(np.uint32))
    assert im["checksum"] == TI.fnv1a(t.to_bytes()[:-8])


def test_damaged_images_are_refused(sbh_gpu):
    t, _ = _tree(sbh_gpu, "cube3d_k2")
    img = t.to_bytes()
    im = TI.parse(img)
    with pytest.raises(ValueError, match="checksum"):
        bad = bytearray(img)
        bad[len(bad) // 2] ^= 0xFF
        sbh_gpu.tree_from_bytes(bytes(bad))
    with pytest.raises(ValueError, match="truncated|checksum"):
        sbh_gpu.tree_from_bytes(img[: len(img) // 3])
    with pytest.raises(ValueError, match="magic"):
        sbh_gpu.tree_from_bytes(b"NOTATREE" + img[8:])
    # structurally wrong but correctly sealed: the root's first child out of range
    bad = bytearray(img)
    off = im["offsets"]["nodeL"] + 4  # root record, field y
    bad[off:off + 4] = np.uint32(im["num_nodes"]).tobytes()
    with pytest.raises(ValueError, match="child"):
        sbh_gpu.tree_from_bytes(bytes(TI.reseal(bad)))
    # a rope pointing backwards (a walk that would not end)
    bad = bytearray(img)
    off = im["offsets"]["dfsL"] + 16 * 5 + 8  # record 5, field z
    bad[off:off + 4] = np.uint32(1).tobytes()
    with pytest.raises(ValueError, match="rope"):
        sbh_gpu.tree_from_bytes(bytes(TI.reseal(bad)))
    # the intact image still loads
    sbh_gpu.tree_from_bytes(img)


def test_loaded_tree_serves(sbh_gpu):
    t, q = _tree(sbh_gpu, "cube3d_k2")
    u = sbh_gpu.tree_from_bytes(t.to_bytes())
    srv = sbh_gpu.Server(1 << 26)
    tid = srv.add_tree(u)
    r = srv.submit(tid, "stochastic", "gravity", q, samples_per_subdomain=2, seed=9)
    got, _ = r.result()
    ref, _ = sbh_gpu.multi_level_prefix_control_variate_gravity(t, q, 2, seed=9,
                                                                 sort_queries=True)
    # same sample paths (seed, 64-query block of the sorted order); the serving
    # kernel is a different compilation of the same fast-math terms
    np.testing.assert_allclose(got, ref, rtol=1e-5)
