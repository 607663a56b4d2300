"""CPU oracle for the Stochastic Barnes-Hut fast-summation path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in
__graft_entry__.py and by the CPU-baseline leg of bench.py, always as the
checker.  The product (the HIP library and the sbh / sbh_gpu modules) never
imports it.

ctypes binding of oracle/build/liboracle.so (built from oracle/oracle.cpp by
oracle/Makefile).  Every array crossing the boundary is converted the way the
reference's pybind11 layer converts it: float64 -> float32 with
round-to-nearest (include/utils.h:52-59, src/sbh.cpp:183-212).  A tree built
with dtype=numpy.float64 is the reference's double build instead (USE_DOUBLE,
include/utils.h:13-29): every input, moment, term and result in float64.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

GRAVITY, WINDING_NUMBER, SMOOTH_DIST = 0, 1, 2  # include/kernel.h:24-28
KERNELS = {"gravity": GRAVITY, "wn": WINDING_NUMBER, "smoothdist": SMOOTH_DIST,
           "helmholtz_x": 3, "helmholtz_y": 4, "gravity_grad_x": 5, "gravity_grad_y": 6,
           "gravity_grad_z": 7}


def build() -> str:
    """Compile the oracle library with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    vp, i32, i64, f32, u64 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                              ctypes.c_float, ctypes.c_uint64)
    f64 = ctypes.c_double
    lib.ot_build.restype = vp
    lib.ot_build.argtypes = [i32, i32, i32, i32, i64, vp, vp, vp, vp, vp, f32]
    lib.ot_build_double.restype = vp
    lib.ot_build_double.argtypes = [i32, i32, i32, i32, i64, vp, vp, vp, vp, vp, f64]
    lib.ot_barnes_hut_double.argtypes = [vp, i32, f64, i32, vp, i64, vp, vp, i32]
    lib.ot_stochastic_double.argtypes = [vp, i32, i32, u64, i32, vp, i64, vp, vp, i32]
    lib.ot_kernel_pairs_double.argtypes = [i32, i32, i64, vp, vp, vp, f64, vp, vp]
    lib.ot_free.argtypes = [vp]
    lib.ot_num_nodes.restype = i64
    lib.ot_num_nodes.argtypes = [vp]
    lib.ot_max_depth.restype = i32
    lib.ot_max_depth.argtypes = [vp]
    lib.ot_leaf_order.argtypes = [vp, vp]
    lib.ot_nodes.argtypes = [vp] + [vp] * 16
    lib.ot_brute_force.argtypes = [vp, i32, i32, vp, i64, vp, i32]
    lib.ot_barnes_hut.argtypes = [vp, i32, f32, i32, vp, i64, vp, vp, i32]
    lib.ot_stochastic.argtypes = [vp, i32, i32, u64, i32, vp, i64, vp, vp, i32]
    lib.ot_kernel_pairs.argtypes = [i32, i32, i64, vp, vp, vp, f32, vp, vp, vp]
    lib.ot_philox.argtypes = [vp, vp, vp]
    lib.ot_path_words.argtypes = [u64, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.c_uint32, ctypes.c_uint32, vp]
    _lib = lib
    return lib


def _f32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).astype(np.float32))
    if shape is not None:
        a = a.reshape(shape)
    return a


def _f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None:
        a = a.reshape(shape)
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data


def kernel_id(kernel) -> int:
    return KERNELS[kernel] if isinstance(kernel, str) else int(kernel)


class Tree:
    """Reference tree build (include/tree3d.h:338-371, tree2d.h) on the host.

    dtype float32: the reference's default build; float64: its double build
    (every input, moment, term and result in float64)."""

    def __init__(self, points, normals, masses, min_corner, max_corner, alpha=1.0,
                 split_log=1, leaf_size=1, max_depth=100, dtype=np.float32):
        lib = _load()
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.float32, np.float64):
            raise ValueError("dtype must be float32 or float64")
        conv = _f64 if self.dtype == np.float64 else _f32
        pts = conv(points)
        n, d = pts.shape
        nrm = conv(normals, (n, d)) if normals is not None else np.zeros((n, d), self.dtype)
        m = conv(masses, (n,))
        lo = conv(min_corner, (d,))
        hi = conv(max_corner, (d,))
        self.dim, self.n, self.split_log = d, n, split_log
        self.leaf_size, self.max_depth = leaf_size, max_depth
        self.points, self.normals, self.masses = pts, nrm, m
        self.min_corner, self.max_corner = lo, hi
        if self.dtype == np.float64:
            self.alpha = float(alpha)
            self._h = lib.ot_build_double(d, split_log, leaf_size, max_depth, n, _ptr(pts),
                                          _ptr(nrm), _ptr(m), _ptr(lo), _ptr(hi),
                                          ctypes.c_double(alpha))
        else:
            self.alpha = float(np.float32(alpha))
            self._h = lib.ot_build(d, split_log, leaf_size, max_depth, n, _ptr(pts), _ptr(nrm),
                                   _ptr(m), _ptr(lo), _ptr(hi), ctypes.c_float(alpha))

    @property
    def is_double(self) -> bool:
        return self.dtype == np.float64

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.ot_free(self._h)
            self._h = None

    @property
    def num_nodes(self) -> int:
        return int(_lib.ot_num_nodes(self._h))

    @property
    def depth(self) -> int:
        return int(_lib.ot_max_depth(self._h))

    def leaf_order(self) -> np.ndarray:
        out = np.empty(self.n, np.int64)
        _lib.ot_leaf_order(self._h, _ptr(out))
        return out

    def nodes(self) -> dict:
        """Node arrays; bounds, diameters and moments in the build's dtype."""
        nn, d, rt = self.num_nodes, self.dim, self.dtype
        r = dict(
            start=np.empty(nn, np.int64), end=np.empty(nn, np.int64),
            first_child=np.empty(nn, np.int64), nchild=np.empty(nn, np.int32),
            depth=np.empty(nn, np.int32), parent=np.empty(nn, np.int64),
            leaf=np.empty(nn, np.uint8), bmin=np.empty((nn, d), rt),
            bmax=np.empty((nn, d), rt), diam=np.empty(nn, rt),
            centroid=np.empty((nn, d), rt), mass=np.empty(nn, rt),
            normal=np.empty((nn, d), rt), centroid64=np.empty((nn, d), np.float64),
            mass64=np.empty(nn, np.float64), normal64=np.empty((nn, d), np.float64))
        keys = ["start", "end", "first_child", "nchild", "depth", "parent", "leaf", "bmin",
                "bmax", "diam", "centroid", "mass", "normal", "centroid64", "mass64", "normal64"]
        _lib.ot_nodes(self._h, *[_ptr(r[k]) for k in keys])
        return r

    def _queries(self, queries):
        if self.is_double:
            return _f64(queries, (-1, self.dim))
        return _f32(queries, (-1, self.dim))

    def brute_force(self, kernel, queries, fp64=False, nthreads=0) -> np.ndarray:
        """Kahan sum in leaf order in the build's dtype (include/brute_force_gpu.h:
        45-63); fp64=True (float builds): float64 sum of the float64 terms."""
        q = self._queries(queries)
        out = np.empty(len(q), np.float64)
        if self.is_double and fp64:
            raise ValueError("fp64=True applies to float builds")
        _lib.ot_brute_force(self._h, kernel_id(kernel), 1 if fp64 else 0, _ptr(q), len(q),
                            _ptr(out), nthreads)
        return out if (fp64 or self.is_double) else out.astype(np.float32)

    def abs_sum(self, kernel, queries, nthreads=0) -> np.ndarray:
        """float64 sum of |term| per query: the condition of the sum."""
        if self.is_double:
            raise ValueError("abs_sum applies to float builds")
        q = self._queries(queries)
        out = np.empty(len(q), np.float64)
        _lib.ot_brute_force(self._h, kernel_id(kernel), 2, _ptr(q), len(q), _ptr(out), nthreads)
        return out

    def barnes_hut(self, kernel, queries, beta, vote_group=1, nthreads=0, counts=False):
        q = self._queries(queries)
        out = np.empty(len(q), self.dtype)
        cnt = np.empty(len(q), np.int32)
        if self.is_double:
            _lib.ot_barnes_hut_double(self._h, kernel_id(kernel), ctypes.c_double(beta),
                                      vote_group, _ptr(q), len(q), _ptr(out), _ptr(cnt),
                                      nthreads)
        else:
            _lib.ot_barnes_hut(self._h, kernel_id(kernel), ctypes.c_float(beta), vote_group,
                               _ptr(q), len(q), _ptr(out), _ptr(cnt), nthreads)
        return (out, cnt) if counts else out

    def stochastic(self, kernel, queries, samples_per_subdomain, seed=0, group=64,
                   nthreads=0, counts=False):
        q = self._queries(queries)
        out = np.empty(len(q), self.dtype)
        cnt = np.empty(len(q), np.int32)
        fn = _lib.ot_stochastic_double if self.is_double else _lib.ot_stochastic
        fn(self._h, kernel_id(kernel), int(samples_per_subdomain), ctypes.c_uint64(seed),
           int(group), _ptr(q), len(q), _ptr(out), _ptr(cnt), nthreads)
        return (out, cnt) if counts else out


def kernel_pairs(dim, kernel, pos, nrm, mass, alpha, q):
    """fp32 and float64 single-pair kernel values (include/kernel.h:31-59)."""
    lib = _load()
    pos, nrm, q = _f32(pos, (-1, dim)), _f32(nrm, (-1, dim)), _f32(q, (-1, dim))
    mass = _f32(mass, (-1,))
    n = len(pos)
    o32, o64 = np.empty(n, np.float32), np.empty(n, np.float64)
    lib.ot_kernel_pairs(dim, kernel_id(kernel), n, _ptr(pos), _ptr(nrm), _ptr(mass),
                        ctypes.c_float(alpha), _ptr(q), _ptr(o32), _ptr(o64))
    return o32, o64


def kernel_pairs_double(dim, kernel, pos, nrm, mass, alpha, q):
    """Single-pair kernel values of the double build (include/kernel.h:31-83)."""
    lib = _load()
    pos, nrm, q = _f64(pos, (-1, dim)), _f64(nrm, (-1, dim)), _f64(q, (-1, dim))
    mass = _f64(mass, (-1,))
    n = len(pos)
    out = np.empty(n, np.float64)
    lib.ot_kernel_pairs_double(dim, kernel_id(kernel), n, _ptr(pos), _ptr(nrm), _ptr(mass),
                               ctypes.c_double(alpha), _ptr(q), _ptr(out))
    return out


def philox(counter, key) -> np.ndarray:
    lib = _load()
    c = np.ascontiguousarray(counter, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.empty(4, np.uint32)
    lib.ot_philox(_ptr(c), _ptr(k), _ptr(out))
    return out


def path_words(seed, group, subdomain, sample, nwords) -> np.ndarray:
    lib = _load()
    out = np.empty(nwords, np.uint32)
    lib.ot_path_words(ctypes.c_uint64(seed), group, subdomain, sample, nwords, _ptr(out))
    return out
