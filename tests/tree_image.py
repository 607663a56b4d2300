"""Reader and writer of the tree image format (version 2) written by
sbh_tree_save / sbh_tree_serialize, following its documented layout.  Used by
the tests to check images against the oracle's tree and to make damaged
images."""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"SBHTREE\0"
VERSION = 2
HEADER = struct.Struct("<8sI5i2Ii f6f I")

# (name, dtype, values per element, length basis) in file order
SECTIONS = [
    ("pts", np.float32, 4, "n"), ("nrm", np.float32, 4, "n"), ("perm", np.uint32, 1, "n"),
    ("n_start", np.uint32, 1, "nn"), ("n_end", np.uint32, 1, "nn"),
    ("n_parent", np.uint32, 1, "nn"), ("n_first", np.uint32, 1, "nn"),
    ("n_nchild", np.uint32, 1, "nn"), ("n_next", np.uint32, 1, "nn"),
    ("n_depth", np.int32, 1, "nn"), ("n_lo", np.float32, 4, "nn"), ("n_hi", np.float32, 4, "nn"),
    ("n_diam", np.float32, 1, "nn"), ("nodeC", np.float32, 4, "nn"),
    ("nodeN", np.float32, 4, "nn"), ("nodeL", np.uint32, 4, "nn"),
    ("n_m64", np.float64, 8, "nn"), ("n_skip", np.uint32, 1, "nn"),
    ("n_size", np.uint32, 1, "nn"), ("n_dfs", np.uint32, 1, "nn"),
    ("dfsC", np.float32, 4, "nn"), ("dfsN", np.float32, 4, "nn"), ("dfsL", np.uint32, 4, "nn"),
]


def fnv1a(data: bytes) -> int:
    """FNV-1a 64 (offset basis 14695981039346656037, prime 1099511628211)."""
    a = np.frombuffer(data, np.uint8)
    h = 14695981039346656037
    prime = 1099511628211
    mask = (1 << 64) - 1
    for b in a.tolist():
        h = ((h ^ b) * prime) & mask
    return h


def parse(img: bytes) -> dict:
    """Header fields, level table and arrays (with their byte offsets)."""
    (magic, version, dim, k, leaf, max_depth, m64, n, nn, depth, alpha,
     *corners, levels) = HEADER.unpack_from(img, 0)
    assert magic == MAGIC and version == VERSION
    off = HEADER.size
    level_begin = np.frombuffer(img, np.uint32, levels + 1, off)
    off += 4 * (levels + 1)
    out = dict(dim=dim, split_log=k, leaf_size=leaf, max_depth=max_depth, moments64=m64, n=n,
               num_nodes=nn, depth=depth, alpha=alpha, lo=corners[:3], hi=corners[3:],
               level_begin=level_begin, offsets={})
    for name, dt, per, basis in SECTIONS:
        count = (n if basis == "n" else nn) * per
        (nbytes,) = struct.unpack_from("<Q", img, off)
        off += 8
        assert nbytes == count * np.dtype(dt).itemsize, name
        out["offsets"][name] = off
        arr = np.frombuffer(img, dt, count, off)
        out[name] = arr.reshape(-1, per) if per > 1 else arr
        off += nbytes
    (checksum,) = struct.unpack_from("<Q", img, off)
    assert off + 8 == len(img)
    out["checksum"] = checksum
    return out


def reseal(img: bytearray) -> bytearray:
    """Recompute the trailing checksum after an edit."""
    img[-8:] = struct.pack("<Q", fnv1a(bytes(img[:-8])))
    return img
