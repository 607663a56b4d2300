"""CPU-only checks of the boundary: libsbh_hip.so loads and exports every
function include/sbh_hip.h declares; argument validation happens on the host
before any device work (so it is testable without a GPU); the plugin modules
import and keep the reference's names."""
import ctypes
import os
import re
import sys

import numpy as np
import pytest

from conftest import PKG, ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "sbh_hip.h")
LIB = os.path.join(PKG, "libsbh_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|const char\*)\s+(sbh_\w+)\s*\(", text,
                                 flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 12
    lib = ctypes.CDLL(LIB)
    for n in names:
        assert hasattr(lib, n), n


def test_version_and_last_error():
    lib = ctypes.CDLL(LIB)
    lib.sbh_version.restype = ctypes.c_int32
    assert lib.sbh_version() == 100
    lib.sbh_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.sbh_last_error(), bytes)


def _create(lib, dim=3, k=1, leaf=1, depth=100, pts=None, m=None, lo=None, hi=None, alpha=1.0):
    n = 0 if pts is None else len(pts)
    pts = np.zeros((max(n, 1), dim), np.float32) if pts is None else np.asarray(pts, np.float32)
    m = np.ones(max(n, 1), np.float32) if m is None else np.asarray(m, np.float32)
    lo = np.zeros(dim, np.float32) if lo is None else np.asarray(lo, np.float32)
    hi = np.ones(dim, np.float32) if hi is None else np.asarray(hi, np.float32)
    h = ctypes.c_void_p()
    f = lib.sbh_tree_create
    f.argtypes = [ctypes.c_int32] * 4 + [ctypes.c_int64] + [ctypes.c_void_p] * 5 + [
        ctypes.c_float, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p]
    return f(dim, k, leaf, depth, n, pts.ctypes.data, None, m.ctypes.data, lo.ctypes.data,
             hi.ctypes.data, alpha, 0, ctypes.byref(h), None)


@pytest.mark.parametrize("kw", [
    dict(dim=4), dict(k=0), dict(k=4), dict(leaf=0), dict(depth=-1),
    dict(pts=np.zeros((0, 3))),
    dict(pts=[[0.5, np.nan, 0.5]]),
    dict(pts=[[0.5, 0.5, 0.5]], m=[0.0]),
    dict(pts=[[0.5, 0.5, 0.5]], m=[-1.0]),
    dict(pts=[[0.5, 0.5, 0.5]], lo=[1, 1, 1], hi=[0, 0, 0]),
    dict(pts=[[0.5, 0.5, 0.5]], alpha=float("inf")),
])
def test_invalid_tree_arguments(kw):
    lib = ctypes.CDLL(LIB)
    assert _create(lib, **kw) == 1  # SBH_ERR_INVALID, before touching a device


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_valid_tree_without_gpu_reports_no_device():
    lib = ctypes.CDLL(LIB)
    assert _create(lib, pts=[[0.5, 0.5, 0.5]]) == 4  # SBH_ERR_NODEVICE


def test_plugin_modules_keep_reference_names():
    import sbh
    import sbh_gpu

    for c in ("Quadtree1", "Quadtree2", "Quadtree3", "Octree1", "Octree2", "Octree3"):
        assert hasattr(sbh, c)
        assert hasattr(sbh_gpu, c.replace("tree", "treeGPU"))
    for kern in ("gravity", "wn", "smoothdist"):
        for fn in ("eval_brute_force_", "eval_barnes_hut_", "multi_level_prefix_control_variate_"):
            assert hasattr(sbh_gpu, fn + kern)
        for fn in ("eval_barnes_hut_", "multi_level_prefix_control_variate_"):
            assert hasattr(sbh, fn + kern)
    assert sbh.fp == 32 and sbh_gpu.fp == 32


def test_host_tree_holds_float32_inputs():
    import sbh

    rng = np.random.default_rng(0)
    pts = rng.random((50, 3))
    t = sbh.Octree2(pts, np.zeros_like(pts), np.ones(50), pts.min(0), pts.max(0), 2.5)
    assert t.dim == 3 and t.split_log == 2 and t.num_points == 50
    np.testing.assert_array_equal(t.points, pts.astype(np.float32))
    assert t.alpha == 2.5
    with pytest.raises(ValueError):
        sbh.Octree1(pts[:, :2], np.zeros((50, 2)), np.ones(50), [0, 0], [1, 1])
    # queries of the wrong dimension are rejected before any device work
    with pytest.raises(ValueError):
        sbh.eval_barnes_hut_gravity(t, pts[:, :2], 2.0)
    # a tree of another module is not a host tree
    with pytest.raises(TypeError):
        sbh.eval_barnes_hut_gravity(object(), pts, 2.0)


def test_cpu_plugin_api_needs_a_gpu():
    """The sbh evaluation functions run on the GPU only: without one they
    raise instead of computing on the host."""
    import sbh
    import sbh_gpu

    if sbh_gpu.device_count() > 0:
        pytest.skip("a GPU is visible")
    pts = np.random.default_rng(2).random((40, 3))
    t = sbh.Octree1(pts, np.zeros_like(pts), np.ones(40), pts.min(0), pts.max(0))
    with pytest.raises(RuntimeError):
        sbh.eval_barnes_hut_gravity(t, pts, 2.0)
    with pytest.raises(RuntimeError):
        sbh.multi_level_prefix_control_variate_wn(t, pts, 1)


def test_gpu_tree_validation_errors_on_host():
    import sbh_gpu

    pts = np.random.default_rng(1).random((20, 3))
    with pytest.raises(ValueError):
        sbh_gpu.OctreeGPU1(pts, np.zeros_like(pts), -np.ones(20), [0, 0, 0], [1, 1, 1])
    with pytest.raises(ValueError):
        sbh_gpu.OctreeGPU1(pts[:, :2], np.zeros((20, 2)), np.ones(20), [0, 0, 0], [1, 1, 1])


def test_tree_image_loader_refuses_damage_before_the_device():
    """Damaged images fail with SBH_ERR_INVALID whether or not a GPU exists:
    the loader checks them on the host first."""
    import ctypes

    import sbh_capi as C

    L = C.lib()
    h = ctypes.c_void_p()
    assert L.sbh_tree_deserialize(b"garbage!" * 4, 32, ctypes.byref(h)) == C.ERR_INVALID
    assert "magic" in L.sbh_last_error().decode()
    assert L.sbh_tree_deserialize(b"SBHTREE\0" + b"\0" * 8, 16, ctypes.byref(h)) == C.ERR_INVALID
    assert L.sbh_tree_deserialize(None, 0, ctypes.byref(h)) == C.ERR_INVALID
    assert L.sbh_tree_load(b"/nonexistent/tree.sbht", ctypes.byref(h)) == C.ERR_INVALID
    assert not h.value


GOLDEN_IMAGE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                            "tree_cube300_k2.sbht")


def test_golden_tree_image_matches_oracle():
    """A tree image written by the GPU build (tests/golden/make_tree_image.py)
    holds the oracle's tree: leaf order, node ranges and fp32 masses, with a
    valid FNV-1a checksum."""
    import oracle
    import tree_image as TI
    sys.path.insert(0, os.path.dirname(GOLDEN_IMAGE))
    import make_tree_image as M

    img = open(GOLDEN_IMAGE, "rb").read()
    im = TI.parse(img)
    assert im["checksum"] == TI.fnv1a(img[:-8])
    pts, nrm, m, lo, hi = M.inputs()
    assert (im["dim"], im["split_log"], im["n"]) == (3, M.K, M.N)
    o = oracle.Tree(pts, nrm, m, lo, hi, split_log=M.K)
    assert im["num_nodes"] == o.num_nodes
    np.testing.assert_array_equal(im["perm"].astype(np.int64), o.leaf_order())
    on = o.nodes()
    np.testing.assert_array_equal(im["n_start"].astype(np.int64), on["start"])
    np.testing.assert_array_equal(im["n_end"].astype(np.int64), on["end"])
    np.testing.assert_array_equal(im["nodeC"][:, 3].view(np.uint32), on["mass"].view(np.uint32))
    np.testing.assert_array_equal(im["n_diam"].view(np.uint32), on["diam"].view(np.uint32))


def test_golden_tree_image_validation_on_host():
    """The loader validates on the host: the intact image passes validation
    (and then needs a device), damaged but correctly sealed copies fail."""
    import ctypes

    import sbh_capi as C
    import sbh_gpu
    import tree_image as TI

    if sbh_gpu.device_count() > 0:
        pytest.skip("a GPU is visible (the GPU tests load images)")
    L = C.lib()
    img = open(GOLDEN_IMAGE, "rb").read()
    im = TI.parse(img)
    h = ctypes.c_void_p()
    assert L.sbh_tree_deserialize(img, len(img), ctypes.byref(h)) == C.ERR_NODEVICE
    cases = [("nodeL", 4, im["num_nodes"], "child"),       # root's first child out of range
             ("dfsL", 16 * 3 + 8, 1, "rope"),              # rope pointing backwards
             ("n_skip", 0, im["num_nodes"], "chain"),      # chain target out of range
             ("perm", 0, 10**6, "leaf order"),
             ("n_end", 0, 5, "")]                           # root no longer holds every point
    # a one-point leaf's depth-first record must hold its point
    one = int(np.flatnonzero(im["dfsL"][:, 3] == (0x80000000 | 1))[0])
    cases.append(("dfsC", 16 * one, im["dfsC"][one, 0].view(np.uint32) ^ 1, "one-point leaf"))
    for name, off, val, msg in cases:
        bad = bytearray(img)
        at = im["offsets"][name] + off
        bad[at:at + 4] = np.uint32(val).tobytes()
        bad = bytes(TI.reseal(bad))
        assert L.sbh_tree_deserialize(bad, len(bad), ctypes.byref(h)) == C.ERR_INVALID, name
        assert msg in L.sbh_last_error().decode(), (name, L.sbh_last_error())
    bad = bytearray(img)
    bad[100] ^= 1
    assert L.sbh_tree_deserialize(bytes(bad), len(bad), ctypes.byref(h)) == C.ERR_INVALID
    assert "checksum" in L.sbh_last_error().decode()


def _create_double(lib, dim=3, k=1, leaf=1, depth=100, pts=None, m=None, lo=None, hi=None,
                   alpha=1.0):
    n = 0 if pts is None else len(pts)
    pts = np.zeros((max(n, 1), dim)) if pts is None else np.asarray(pts, np.float64)
    m = np.ones(max(n, 1)) if m is None else np.asarray(m, np.float64)
    lo = np.zeros(dim) if lo is None else np.asarray(lo, np.float64)
    hi = np.ones(dim) if hi is None else np.asarray(hi, np.float64)
    h = ctypes.c_void_p()
    f = lib.sbh_tree_create_double
    f.argtypes = [ctypes.c_int32] * 4 + [ctypes.c_int64] + [ctypes.c_void_p] * 5 + [
        ctypes.c_double, ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p]
    return f(dim, k, leaf, depth, n, pts.ctypes.data, None, m.ctypes.data, lo.ctypes.data,
             hi.ctypes.data, alpha, ctypes.byref(h), None)


@pytest.mark.parametrize("kw", [
    dict(dim=1), dict(k=4), dict(leaf=0), dict(pts=np.zeros((0, 3))),
    dict(pts=[[0.5, np.inf, 0.5]]), dict(pts=[[0.5, 0.5, 0.5]], m=[0.0]),
    dict(pts=[[0.5, 0.5, 0.5]], lo=[0, 0, 1], hi=[1, 1, 1]),
    dict(pts=[[0.5, 0.5, 0.5]], alpha=float("nan")),
])
def test_invalid_double_tree_arguments(kw):
    lib = ctypes.CDLL(LIB)
    assert _create_double(lib, **kw) == 1  # SBH_ERR_INVALID, before touching a device


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_valid_double_tree_without_gpu_reports_no_device():
    lib = ctypes.CDLL(LIB)
    assert _create_double(lib, pts=[[0.5, 0.5, 0.5]]) == 4  # SBH_ERR_NODEVICE


def test_double_build_module_keeps_reference_names():
    import sbh_gpu64

    assert sbh_gpu64.fp == 64
    for c in ("QuadtreeGPU1", "QuadtreeGPU2", "QuadtreeGPU3", "OctreeGPU1", "OctreeGPU2",
              "OctreeGPU3"):
        assert hasattr(sbh_gpu64, c)
    for kern in ("gravity", "wn", "smoothdist", "helmholtz_x", "gravity_grad_z"):
        for fn in ("eval_brute_force_", "eval_barnes_hut_", "multi_level_prefix_control_variate_"):
            assert hasattr(sbh_gpu64, fn + kern)
    pts = np.random.default_rng(1).random((20, 3))
    with pytest.raises(ValueError):  # validated on the host, before any device work
        sbh_gpu64.OctreeGPU1(pts, np.zeros_like(pts), -np.ones(20), [0, 0, 0], [1, 1, 1])
    with pytest.raises(ValueError):
        sbh_gpu64.QuadtreeGPU1(pts, np.zeros_like(pts), np.ones(20), [0, 0], [1, 1])


def test_double_eval_validates_on_host():
    import sbh_capi as C

    L = C.lib()
    p = C.EvalParamsDouble(2.0, 1, 0, 1, 0)
    out = np.zeros(4)
    q = np.zeros(12)
    assert L.sbh_eval_double(None, C.BARNES_HUT, 0, p, q.ctypes.data, 4, out.ctypes.data, None,
                             None) == C.ERR_INVALID
    isd = ctypes.c_int32()
    assert L.sbh_tree_is_double(None, ctypes.byref(isd)) == C.ERR_INVALID
