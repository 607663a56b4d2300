"""Golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle): the oracle still reproduces them bit for bit (CPU), and the
HIP path's exact mode reproduces them on the GPU."""
import glob
import os
import sys

import numpy as np
import pytest

import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))
sys.path.insert(0, GOLDEN)
import make_golden as MG  # noqa: E402


def _ids():
    return [os.path.basename(c)[:-4] for c in CASES]


def test_golden_cases_exist():
    assert {os.path.basename(c)[:-4] for c in CASES} == set(MG.cases())


@pytest.mark.parametrize("path", CASES, ids=_ids())
def test_case_format_round_trip(tmp_path, path):
    case = W.load_case(path)
    out = os.path.join(tmp_path, "copy.npz")
    W.save_case(out, case["points"], case["normals"], case["masses"], case["min_corner"],
                case["max_corner"], case["queries"], alpha=case["alpha"],
                split_log=case["split_log"], leaf_size=case["leaf_size"], **case["results"])
    again = W.load_case(out)
    for k in W.CASE_INPUTS:
        np.testing.assert_array_equal(again[k], case[k])
    assert again["results"].keys() == case["results"].keys()


@pytest.mark.parametrize("path", CASES, ids=_ids())
def test_oracle_reproduces_golden(path):
    """The cases are regenerated from their seeds and the oracle must give
    the stored vectors exactly (the oracle and the generators are pinned)."""
    name = os.path.basename(path)[:-4]
    kernel, inputs, alpha, k, leaf = MG.cases()[name]
    case = W.load_case(path)
    np.testing.assert_array_equal(case["points"], np.asarray(inputs[0], np.float32))
    _, _, res = MG.compute(kernel, inputs, alpha, k, leaf)
    assert str(case["results"]["kernel"]) == kernel
    for key, val in res.items():
        np.testing.assert_array_equal(case["results"][key], val, key)


def _gpu_tree(sbh_gpu, case):
    cls = getattr(sbh_gpu, ("OctreeGPU%d" if case["points"].shape[1] == 3 else "QuadtreeGPU%d")
                  % case["split_log"])
    return cls(case["points"], case["normals"], case["masses"], case["min_corner"],
               case["max_corner"], alpha=case["alpha"], leaf_size=case["leaf_size"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", CASES, ids=_ids())
def test_hip_path_reproduces_golden(sbh_gpu, path):
    case = W.load_case(path)
    r = case["results"]
    kern = str(r["kernel"])
    q = case["queries"]
    print(f"[case] golden {os.path.basename(path)}", flush=True)
    g = _gpu_tree(sbh_gpu, case)
    np.testing.assert_array_equal(g.leaf_order(), r["leaf_order"])
    bf, _ = getattr(sbh_gpu, "eval_brute_force_" + kern)(g, q)
    den = np.maximum(np.abs(r["brute_force_f64"]), 1.0) if kern == "wn" else np.abs(
        r["brute_force_f64"])
    assert (np.abs(bf - r["brute_force_f64"]) / den).max() <= 1e-5
    exact = kern != "smoothdist"  # device expf vs glibc expf: ulps
    bh = getattr(sbh_gpu, "eval_barnes_hut_" + kern)
    st = getattr(sbh_gpu, "multi_level_prefix_control_variate_" + kern)
    for got, key in ((bh(g, q, MG.BETA, vote_group=1, exact_terms=True, return_counts=True),
                      "bh"),
                     (bh(g, q, MG.BETA, vote_group=32, exact_terms=True), "bh_vote32"),
                     (st(g, q, MG.SPL, seed=MG.SEED, exact_terms=True, return_counts=True),
                      "stochastic"),
                     (st(g, q, MG.SPL, seed=MG.SEED, exact_terms=True, return_counts=True,
                         per_query_paths=True), "stochastic_per_query")):
        vals = got[0]
        if exact:
            np.testing.assert_array_equal(vals, r[key], key)
        else:
            assert (np.abs(vals - r[key]) / np.abs(r[key])).max() <= 1e-5, key
        if len(got) == 3:
            np.testing.assert_array_equal(got[2], r[key + "_counts"], key)
