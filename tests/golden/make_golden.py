"""Writes the golden vectors under tests/golden/ (format: workloads.save_case,
one .npz per case) from the oracle, the CPU restatement of the reference's
algorithms:

    python tests/golden/make_golden.py

Each case holds a seeded workload and the oracle's leaf order, float64
all-pairs sums, per-query and 32-vote Barnes-Hut values (beta 2), and the
stochastic estimator's values and term counts with 64-query and 1-query path
groups.  The CPU tests re-derive them from the oracle (a regression pin of
the oracle itself); the GPU tests compare the HIP path's exact mode with them
bit for bit.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "stochastic-barnes-hut_amd")]

import oracle  # noqa: E402
import workloads as W  # noqa: E402

BETA, SPL, SEED = 2.0, 4, 3


def cases():
    """name -> (kernel, inputs, alpha, split_log, leaf_size)"""
    out = {}
    pts, nrm, m, q = W.uniform_cube(2000, 128, dim=3, seed=101)
    out["cube3d_gravity_k2"] = ("gravity", (pts, nrm, m, q), 1.0, 2, 1)
    pts, nrm, m, q = W.ring2d(1500, 128, seed=102)
    out["ring2d_wn_k1"] = ("wn", (pts, nrm, m, q), 1.0, 1, 1)
    pts, nrm, m, q = W.noisy_icosphere(3, 128, seed=103)
    out["icosphere3_smoothdist_k1_leaf4"] = ("smoothdist", (pts, nrm, m, q), 10.0, 1, 4)
    return out


def compute(kernel, inputs, alpha, k, leaf):
    pts, nrm, m, q = inputs
    lo, hi = W.bounds(pts)
    t = oracle.Tree(pts, nrm, m, lo, hi, alpha=alpha, split_log=k, leaf_size=leaf)
    res = dict(leaf_order=t.leaf_order(),
               brute_force_f64=t.brute_force(kernel, q, fp64=True))
    res["bh"], res["bh_counts"] = t.barnes_hut(kernel, q, BETA, vote_group=1, counts=True)
    res["bh_vote32"] = t.barnes_hut(kernel, q, BETA, vote_group=32)
    res["stochastic"], res["stochastic_counts"] = t.stochastic(kernel, q, SPL, seed=SEED,
                                                               group=64, counts=True)
    res["stochastic_per_query"], res["stochastic_per_query_counts"] = t.stochastic(
        kernel, q, SPL, seed=SEED, group=1, counts=True)
    return lo, hi, res


if __name__ == "__main__":
    for name, (kernel, inputs, alpha, k, leaf) in cases().items():
        lo, hi, res = compute(kernel, inputs, alpha, k, leaf)
        pts, nrm, m, q = inputs
        path = os.path.join(HERE, name + ".npz")
        W.save_case(path, pts, nrm, m, lo, hi, q, alpha=alpha, split_log=k, leaf_size=leaf,
                    kernel=np.array(kernel), **res)
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")
