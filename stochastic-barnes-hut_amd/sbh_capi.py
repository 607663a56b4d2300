"""ctypes binding of libsbh_hip.so's C ABI (include/sbh_hip.h).

This is the binding a host program in any language would write; Python
callers use it when their query and result buffers already live on the GPU
(SBH_EVAL_DEVICE_POINTERS), for example torch tensors on cuda:0.  The
pybind11 module sbh_gpu is the numpy-facing plugin over the same ABI.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsbh_hip.so")

OK, ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_NODEVICE, ERR_BUDGET, ERR_INTERNAL = range(7)
GRAVITY, WINDING_NUMBER, SMOOTH_DIST = 0, 1, 2
HELMHOLTZ_X, HELMHOLTZ_Y, GRAVITY_GRAD_X, GRAVITY_GRAD_Y, GRAVITY_GRAD_Z = 3, 4, 5, 6, 7
KERNELS = {"gravity": GRAVITY, "wn": WINDING_NUMBER, "smoothdist": SMOOTH_DIST,
           "helmholtz_x": HELMHOLTZ_X, "helmholtz_y": HELMHOLTZ_Y,
           "gravity_grad_x": GRAVITY_GRAD_X, "gravity_grad_y": GRAVITY_GRAD_Y,
           "gravity_grad_z": GRAVITY_GRAD_Z}
BRUTE_FORCE, BARNES_HUT, STOCHASTIC = 0, 1, 2
TREE_MOMENTS_F64 = 1
EVAL_DEVICE_POINTERS, EVAL_SORT_QUERIES, EVAL_SPLIT_COUNTS = 1, 2, 4
EVAL_EXACT_TERMS, EVAL_MFMA, EVAL_PER_QUERY_PATHS, EVAL_MFMA_BF16 = 8, 16, 32, 64
EVAL_SMOOTH_DISTANCE, EVAL_NO_SPLIT, EVAL_WORK_COUNTS = 128, 256, 512


class EvalParams(ctypes.Structure):
    _fields_ = [("beta", ctypes.c_float), ("samples_per_subdomain", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("vote_group", ctypes.c_int32),
                ("flags", ctypes.c_int32)]


class EvalParamsDouble(ctypes.Structure):
    """sbh_eval_params_double (double-build trees)."""
    _fields_ = [("beta", ctypes.c_double), ("samples_per_subdomain", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("vote_group", ctypes.c_int32),
                ("flags", ctypes.c_int32)]


class SbhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def lib():
    """Load libsbh_hip.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run stochastic-barnes-hut_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
        L.sbh_last_error.restype = ctypes.c_char_p
        L.sbh_version.restype = i32
        L.sbh_device_count.argtypes = [ctypes.POINTER(i32)]
        L.sbh_tree_create.argtypes = [i32, i32, i32, i32, i64, vp, vp, vp, vp, vp, f32, i32,
                                      ctypes.POINTER(vp), ctypes.POINTER(f32)]
        L.sbh_tree_destroy.argtypes = [vp]
        L.sbh_tree_info.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64),
                                    ctypes.POINTER(i32), ctypes.POINTER(i64)]
        L.sbh_tree_leaf_order.argtypes = [vp, vp]
        L.sbh_tree_nodes.argtypes = [vp] + [vp] * 14
        L.sbh_eval.argtypes = [vp, i32, i32, ctypes.POINTER(EvalParams), vp, i64, vp, vp,
                               ctypes.POINTER(f32)]
        L.sbh_eval_noalloc.argtypes = L.sbh_eval.argtypes
        L.sbh_eval_f64.argtypes = [vp, i32, vp, i64, vp, i32, ctypes.POINTER(f32)]
        L.sbh_query_order.argtypes = [vp, vp, i64, vp, i32]
        L.sbh_tree_stream.argtypes = [vp, ctypes.POINTER(vp)]
        L.sbh_tree_serialize.argtypes = [vp, vp, i64, ctypes.POINTER(i64)]
        L.sbh_tree_deserialize.argtypes = [vp, i64, ctypes.POINTER(vp)]
        L.sbh_tree_save.argtypes = [vp, ctypes.c_char_p]
        L.sbh_tree_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
        L.sbh_tree_layout.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                      ctypes.POINTER(i32), ctypes.POINTER(f32)]
        f64 = ctypes.c_double
        L.sbh_tree_create_double.argtypes = [i32, i32, i32, i32, i64, vp, vp, vp, vp, vp, f64,
                                             ctypes.POINTER(vp), ctypes.POINTER(f32)]
        L.sbh_tree_is_double.argtypes = [vp, ctypes.POINTER(i32)]
        L.sbh_tree_nodes_double.argtypes = [vp] + [vp] * 6
        L.sbh_eval_double.argtypes = [vp, i32, i32, ctypes.POINTER(EvalParamsDouble), vp, i64,
                                      vp, vp, ctypes.POINTER(f32)]
        _lib = L
    return _lib


def check(status):
    if status != OK:
        raise SbhError(status, lib().sbh_last_error().decode())


def _f32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a).astype(np.float32))
    return a.reshape(shape) if shape is not None else a


class Tree:
    """Device tree handle (sbh_tree*)."""

    def __init__(self, points, normals, masses, min_corner, max_corner, alpha=1.0,
                 split_log=1, leaf_size=1, max_depth=100, moments64=False):
        L = lib()
        pts = _f32(points)
        n, d = pts.shape
        nrm = None if normals is None else _f32(normals, (n, d))
        m = _f32(masses, (n,))
        lo, hi = _f32(min_corner, (d,)), _f32(max_corner, (d,))
        self.dim, self.n, self.split_log = d, n, split_log
        h = ctypes.c_void_p()
        ms = ctypes.c_float()
        check(L.sbh_tree_create(d, split_log, leaf_size, max_depth, n, pts.ctypes.data,
                                None if nrm is None else nrm.ctypes.data, m.ctypes.data,
                                lo.ctypes.data, hi.ctypes.data, float(alpha),
                                TREE_MOMENTS_F64 if moments64 else 0, ctypes.byref(h),
                                ctypes.byref(ms)))
        self.h = h
        self.build_ms = ms.value

    @classmethod
    def _adopt(cls, h):
        self = cls.__new__(cls)
        self.h = h
        d, k, leaf = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().sbh_tree_layout(h, ctypes.byref(d), ctypes.byref(k), ctypes.byref(leaf), None))
        self.dim, self.split_log, self.build_ms = d.value, k.value, 0.0
        self.n = self.info()["num_points"]
        return self

    @classmethod
    def load(cls, path) -> "Tree":
        """A tree saved with save(path), resident on the current GPU."""
        h = ctypes.c_void_p()
        check(lib().sbh_tree_load(os.fsencode(path), ctypes.byref(h)))
        return cls._adopt(h)

    @classmethod
    def from_bytes(cls, image: bytes) -> "Tree":
        h = ctypes.c_void_p()
        check(lib().sbh_tree_deserialize(image, len(image), ctypes.byref(h)))
        return cls._adopt(h)

    def save(self, path):
        check(lib().sbh_tree_save(self.h, os.fsencode(path)))

    def to_bytes(self) -> bytes:
        size = ctypes.c_int64()
        check(lib().sbh_tree_serialize(self.h, None, 0, ctypes.byref(size)))
        buf = ctypes.create_string_buffer(size.value)
        check(lib().sbh_tree_serialize(self.h, buf, size.value, ctypes.byref(size)))
        return buf.raw[:size.value]

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().sbh_tree_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        np_, nn, bytes_ = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        depth = ctypes.c_int32()
        check(lib().sbh_tree_info(self.h, ctypes.byref(np_), ctypes.byref(nn),
                                  ctypes.byref(depth), ctypes.byref(bytes_)))
        return dict(num_points=np_.value, num_nodes=nn.value, depth=depth.value,
                    device_bytes=bytes_.value, build_ms=self.build_ms)

    def stream(self) -> int:
        s = ctypes.c_void_p()
        check(lib().sbh_tree_stream(self.h, ctypes.byref(s)))
        return s.value or 0

    def eval_device(self, method, kernel, q_ptr, nq, out_ptr, counts_ptr=0, beta=0.0,
                    samples_per_subdomain=1, seed=0, vote_group=1, flags=0) -> float:
        """Evaluate with device pointers (queries column-major d x nq); returns
        the hipEvent-timed kernel milliseconds."""
        p = EvalParams(beta, samples_per_subdomain, seed, vote_group,
                       flags | EVAL_DEVICE_POINTERS)
        ms = ctypes.c_float()
        k = KERNELS[kernel] if isinstance(kernel, str) else kernel
        check(lib().sbh_eval(self.h, method, k, ctypes.byref(p), ctypes.c_void_p(q_ptr), nq,
                             ctypes.c_void_p(out_ptr), ctypes.c_void_p(counts_ptr or None),
                             ctypes.byref(ms)))
        return ms.value

    def eval_device_noalloc(self, method, kernel, q_ptr, nq, out_ptr, counts_ptr=0, beta=0.0,
                            samples_per_subdomain=1, seed=0, vote_group=1, flags=0) -> float:
        """eval_device through sbh_eval_noalloc: no device allocation, scratch
        from the library's fixed per-GPU array."""
        p = EvalParams(beta, samples_per_subdomain, seed, vote_group,
                       flags | EVAL_DEVICE_POINTERS)
        ms = ctypes.c_float()
        k = KERNELS[kernel] if isinstance(kernel, str) else kernel
        check(lib().sbh_eval_noalloc(self.h, method, k, ctypes.byref(p), ctypes.c_void_p(q_ptr),
                                     nq, ctypes.c_void_p(out_ptr),
                                     ctypes.c_void_p(counts_ptr or None), ctypes.byref(ms)))
        return ms.value

    def eval_host(self, method, kernel, queries, beta=0.0, samples_per_subdomain=1, seed=0,
                  vote_group=1, flags=0, counts=False):
        q = _f32(queries, (-1, self.dim))
        soa = np.ascontiguousarray(q.T)
        nq = len(q)
        out = np.empty(nq, np.float32)
        cnt = (np.empty(nq * (2 if flags & (EVAL_SPLIT_COUNTS | EVAL_WORK_COUNTS) else 1), np.int32)
               if counts else None)
        p = EvalParams(beta, samples_per_subdomain, seed, vote_group, flags)
        ms = ctypes.c_float()
        k = KERNELS[kernel] if isinstance(kernel, str) else kernel
        check(lib().sbh_eval(self.h, method, k, ctypes.byref(p), soa.ctypes.data, nq,
                             out.ctypes.data, None if cnt is None else cnt.ctypes.data,
                             ctypes.byref(ms)))
        return (out, ms.value, cnt) if counts else (out, ms.value)

    def eval_f64_device(self, kernel, q_ptr, nq, out_ptr) -> float:
        ms = ctypes.c_float()
        k = KERNELS[kernel] if isinstance(kernel, str) else kernel
        check(lib().sbh_eval_f64(self.h, k, ctypes.c_void_p(q_ptr), nq, ctypes.c_void_p(out_ptr),
                                 EVAL_DEVICE_POINTERS, ctypes.byref(ms)))
        return ms.value

    def eval_f64_host(self, kernel, queries):
        q = _f32(queries, (-1, self.dim))
        soa = np.ascontiguousarray(q.T)
        out = np.empty(len(q), np.float64)
        ms = ctypes.c_float()
        k = KERNELS[kernel] if isinstance(kernel, str) else kernel
        check(lib().sbh_eval_f64(self.h, k, soa.ctypes.data, len(q), out.ctypes.data, 0,
                                 ctypes.byref(ms)))
        return out, ms.value


class TreeDouble(Tree):
    """Double-build tree handle (sbh_tree_create_double): float64 inputs,
    tree, terms and results, evaluated by sbh_eval_double."""

    def __init__(self, points, normals, masses, min_corner, max_corner, alpha=1.0,
                 split_log=1, leaf_size=1, max_depth=100):
        L = lib()
        pts = np.ascontiguousarray(np.asarray(points, np.float64))
        n, d = pts.shape
        nrm = None if normals is None else np.ascontiguousarray(
            np.asarray(normals, np.float64).reshape(n, d))
        m = np.ascontiguousarray(np.asarray(masses, np.float64).reshape(n))
        lo = np.ascontiguousarray(np.asarray(min_corner, np.float64).reshape(d))
        hi = np.ascontiguousarray(np.asarray(max_corner, np.float64).reshape(d))
        self.dim, self.n, self.split_log = d, n, split_log
        h = ctypes.c_void_p()
        ms = ctypes.c_float()
        check(L.sbh_tree_create_double(d, split_log, leaf_size, max_depth, n, pts.ctypes.data,
                                       None if nrm is None else nrm.ctypes.data, m.ctypes.data,
                                       lo.ctypes.data, hi.ctypes.data, float(alpha),
                                       ctypes.byref(h), ctypes.byref(ms)))
        self.h = h
        self.build_ms = ms.value

    def eval_device(self, method, kernel, q_ptr, nq, out_ptr, counts_ptr=0, beta=0.0,
                    samples_per_subdomain=1, seed=0, vote_group=1, flags=0) -> float:
        """Evaluate with device pointers (float64 queries column-major d x nq,
        float64 results); returns the hipEvent-timed kernel milliseconds."""
        p = EvalParamsDouble(beta, samples_per_subdomain, seed, vote_group,
                             flags | EVAL_DEVICE_POINTERS)
        ms = ctypes.c_float()
        k = KERNELS[kernel] if isinstance(kernel, str) else kernel
        check(lib().sbh_eval_double(self.h, method, k, ctypes.byref(p), ctypes.c_void_p(q_ptr),
                                    nq, ctypes.c_void_p(out_ptr),
                                    ctypes.c_void_p(counts_ptr or None), ctypes.byref(ms)))
        return ms.value

    def eval_host(self, method, kernel, queries, beta=0.0, samples_per_subdomain=1, seed=0,
                  vote_group=1, flags=0, counts=False):
        q = np.asarray(queries, np.float64).reshape(-1, self.dim)
        soa = np.ascontiguousarray(q.T)
        nq = len(q)
        out = np.empty(nq, np.float64)
        cnt = np.empty(nq * (2 if flags & EVAL_SPLIT_COUNTS else 1), np.int32) if counts else None
        p = EvalParamsDouble(beta, samples_per_subdomain, seed, vote_group, flags)
        ms = ctypes.c_float()
        k = KERNELS[kernel] if isinstance(kernel, str) else kernel
        check(lib().sbh_eval_double(self.h, method, k, ctypes.byref(p), soa.ctypes.data, nq,
                                    out.ctypes.data, None if cnt is None else cnt.ctypes.data,
                                    ctypes.byref(ms)))
        return (out, ms.value, cnt) if counts else (out, ms.value)

    def eval_f64_device(self, *a, **kw):
        raise SbhError(ERR_INVALID, "a double-build tree evaluates in float64 already: eval_*")

    eval_f64_host = eval_f64_device
