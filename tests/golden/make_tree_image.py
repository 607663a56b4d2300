"""Writes tests/golden/tree_cube300_k2.sbht: the image (sbh_tree_save format,
version 2) of the octree k=2 the library builds on the GPU for 300 points of
the seeded unit-cube workload.  Run on a machine with a GPU:

    python tests/golden/make_tree_image.py [output path]

The CPU tests read it back: its leaf order, node ranges and fp32 node records
must equal the oracle's tree for the same inputs, and the loader must accept
it (up to the missing GPU) and refuse damaged copies before touching a device.
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "stochastic-barnes-hut_amd")]

import workloads as W  # noqa: E402
import sbh_gpu  # noqa: E402

N, SEED, K = 300, 5, 2


def inputs():
    pts, nrm, m, _ = W.uniform_cube(N, 1, dim=3, seed=SEED)
    lo, hi = W.bounds(pts)
    return pts, nrm, m, lo, hi


if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "tree_cube300_k2.sbht")
    pts, nrm, m, lo, hi = inputs()
    t = sbh_gpu.OctreeGPU2(pts, nrm, m, lo, hi)
    t.save(out)
    print(f"wrote {out}: {os.path.getsize(out)} bytes, {t.info()['num_nodes']} nodes")
