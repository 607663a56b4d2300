"""Seeded synthetic workloads (numpy float64, cast to float32 at the plugin
boundary as the reference's pybind11 layer does).

  uniform_cube      3D/2D gravity: U[0,1)^d points, masses U[0.5,1.5],
                    queries 0.5 + 1.2 (U - 0.5)  (the 16384-point CPU case)
  plummer           Plummer sphere, r = (u^(-2/3) - 1)^(-1/2), equal masses 1/N
  gmm2d             8-component 2D Gaussian mixture
  ring2d            2D winding number: dipoles on a jittered circle
  noisy_icosphere   subdivided icosphere, vertices moved radially by 3-octave
                    noise; sources are area-weighted normal dipoles at triangle
                    centroids (winding number / smooth distance)

The reference's own mesh generators (python/source.py:118-155) need igl and
gpytoolbox, which are not available; these replace them with closed-form
shapes of the same role.

save_case / load_case store a workload and any expected results as one
.npz file (plain arrays, no pickled objects): the on-disk fixture format of
the golden vectors under tests/golden/.
"""
from __future__ import annotations

import numpy as np


def bounds(points: np.ndarray, pad: float = 0.01):
    """Root cell as the reference's driver chooses it (python/example.py:23-24)."""
    return points.min(axis=0) - pad, points.max(axis=0) + pad


def uniform_cube(n: int, q: int, dim: int = 3, seed: int = 0):
    rng = np.random.default_rng(seed)
    pts = rng.random((n, dim))
    masses = rng.uniform(0.5, 1.5, n)
    queries = 0.5 + 1.2 * (rng.random((q, dim)) - 0.5)
    return pts, np.zeros_like(pts), masses, queries


def _isotropic(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def plummer_points(n: int, seed: int, rmax: float | None = None) -> np.ndarray:
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    if rmax is not None:
        # truncate the mass profile at rmax: M(<r) = r^3 / (1 + r^2)^(3/2)
        umax = rmax ** 3 / (1.0 + rmax ** 2) ** 1.5
        u = u * umax
    u = np.clip(u, 1e-12, None)
    r = (u ** (-2.0 / 3.0) - 1.0) ** -0.5
    return r[:, None] * _isotropic(rng, n)


def plummer(n: int, q: int, seed_sources: int = 1, seed_queries: int = 2,
            rmax: float | None = None):
    pts = plummer_points(n, seed_sources, rmax)
    queries = plummer_points(q, seed_queries, rmax)
    masses = np.full(n, 1.0 / n)
    return pts, np.zeros_like(pts), masses, queries


def gmm2d(n: int, q: int, seed: int = 3, components: int = 8):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-1.0, 1.0, (components, 2))
    sig = rng.uniform(0.05, 0.2, components)
    ci = rng.integers(0, components, n)
    pts = means[ci] + sig[ci, None] * rng.normal(size=(n, 2))
    cq = rng.integers(0, components, q)
    queries = means[cq] + sig[cq, None] * rng.normal(size=(q, 2))
    masses = np.full(n, 1.0 / n)
    return pts, np.zeros_like(pts), masses, queries


def icosphere(level: int):
    t = (1.0 + 5 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    V = np.array(V, dtype=np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array(F, dtype=np.int64)
    for _ in range(level):
        nv = len(V)
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mids = V[uniq[:, 0]] + V[uniq[:, 1]]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        V = np.concatenate([V, mids])
        nf = len(F)
        a = nv + inv[:nf]
        b = nv + inv[nf:2 * nf]
        c = nv + inv[2 * nf:]
        F = np.concatenate([
            np.stack([F[:, 0], a, c], 1), np.stack([F[:, 1], b, a], 1),
            np.stack([F[:, 2], c, b], 1), np.stack([a, b, c], 1)])
    return V, F


def _value_noise(p: np.ndarray, freq: float, rng) -> np.ndarray:
    """Smooth lattice noise on the unit sphere (trilinear interpolation of
    random lattice values)."""
    g = 8
    lattice = rng.uniform(-1.0, 1.0, (g, g, g))
    x = (p * freq + 4.0) % g
    i0 = np.floor(x).astype(np.int64)
    f = x - i0
    f = f * f * (3 - 2 * f)
    i1 = (i0 + 1) % g
    out = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                ix = i1[:, 0] if dx else i0[:, 0]
                iy = i1[:, 1] if dy else i0[:, 1]
                iz = i1[:, 2] if dz else i0[:, 2]
                w = ((f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) *
                     (f[:, 2] if dz else 1 - f[:, 2]))
                out = out + w * lattice[ix, iy, iz]
    return out


def ring2d(n: int, q: int, seed: int = 0, radius: float = 0.4, jitter: float = 0.05):
    """2D winding-number sources: dipoles on a jittered circle (the closed
    curve python/sampler.py:5-20 samples), outward normals, masses = segment
    lengths; queries uniform in the square [-1.5 r, 1.5 r]^2."""
    rng = np.random.default_rng(seed)
    ang = np.sort(rng.uniform(0.0, 2.0 * np.pi, n))
    r = radius * (1.0 + jitter * rng.normal(size=n))
    pts = r[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1)
    e = np.roll(pts, -1, axis=0) - pts
    nrm = np.stack([e[:, 1], -e[:, 0]], 1)
    length = np.linalg.norm(nrm, axis=1)
    nrm /= length[:, None]
    queries = rng.uniform(-1.5 * radius, 1.5 * radius, (q, 2))
    return pts, nrm, length, queries


def noisy_icosphere(level: int, q: int, amplitude: float = 0.1, seed: int = 4,
                    expand: float = 0.1):
    """Dipole sources of a closed noisy sphere; returns (points, normals, masses,
    queries) with masses = triangle areas and unit outward normals."""
    rng = np.random.default_rng(seed)
    V, F = icosphere(level)
    noise = (_value_noise(V, 1.0, rng) + 0.5 * _value_noise(V, 2.0, rng) +
             0.25 * _value_noise(V, 4.0, rng)) / 1.75
    V = V * (1.0 + amplitude * noise)[:, None]
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    cr = np.cross(b - a, c - a)
    area2 = np.linalg.norm(cr, axis=1)
    normals = cr / area2[:, None]
    centroids = (a + b + c) / 3.0
    masses = 0.5 * area2
    lo, hi = V.min(axis=0), V.max(axis=0)
    ext = hi - lo
    qlo, qhi = lo - expand * 0.5 * ext, hi + expand * 0.5 * ext
    queries = qlo + (qhi - qlo) * rng.random((q, 3))
    return centroids, normals, masses, queries


CASE_FORMAT = 1
CASE_INPUTS = ("points", "normals", "masses", "min_corner", "max_corner", "queries")


def save_case(path: str, points, normals, masses, min_corner, max_corner, queries,
              alpha: float = 1.0, split_log: int = 1, leaf_size: int = 1, **results) -> None:
    """Writes a workload (inputs as float32, the precision the plugins
    compute in) and named result arrays to one compressed .npz file."""
    arrays = dict(points=points, normals=normals, masses=masses, min_corner=min_corner,
                  max_corner=max_corner, queries=queries)
    out = {k: np.ascontiguousarray(np.asarray(v, np.float32)) for k, v in arrays.items()}
    out["meta"] = np.array([CASE_FORMAT, split_log, leaf_size], np.int64)
    out["alpha"] = np.array([alpha], np.float32)
    for k, v in results.items():
        if k in out:
            raise ValueError(f"result name {k!r} collides with an input")
        out["result_" + k] = np.asarray(v)
    np.savez_compressed(path, **out)


def load_case(path: str) -> dict:
    """Reads a file written by save_case: the inputs, alpha, split_log,
    leaf_size and a dict of results.  Loads without unpickling."""
    with np.load(path, allow_pickle=False) as z:
        meta = z["meta"]
        if int(meta[0]) != CASE_FORMAT:
            raise ValueError(f"{path}: unsupported case format {int(meta[0])}")
        case = {k: z[k] for k in CASE_INPUTS}
        case["alpha"] = float(z["alpha"][0])
        case["split_log"], case["leaf_size"] = int(meta[1]), int(meta[2])
        case["results"] = {k[len("result_"):]: z[k] for k in z.files if k.startswith("result_")}
    return case
