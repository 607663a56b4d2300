"""Benchmark: queries/s at matched 1e-2 mean relative error on one MI355X.

Headline workload (3D gravity, the 1M-source GPU case): N = 2^20 sources drawn
from a Plummer sphere (scale radius 1, equal masses summing to 1) and
Q = 2^20 queries from the same distribution; fp32; synthetic data generated
here with seeded numpy.

A step is one evaluation of all Q queries by the stochastic control-variate
estimator at the smallest sample budget whose mean relative error (against a
float64 all-pairs reference on the first --err-queries queries) is <= 1e-2.
Brute force and Barnes-Hut (theta 0.3-1.0) are swept beside it.  Inputs are
resident in HBM before timing (torch tensors on cuda:0, passed to the C ABI as
device pointers); kernel times are hipEvent-timed on the tree's stream.

Also measured (unless disabled): variants (truncated Plummer, 2D Gaussian
mixture, winding number and smooth distance on a noisy icosphere), the
serving loop (8 resident trees, Poisson request stream, p50/p99 latency) and
the CPU baseline.  One JSON line goes to stdout; progress to stderr.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "stochastic-barnes-hut_amd")
sys.path[:0] = [ROOT, PKG]

METRIC = "queries/s at 1e-2 rel-err (stoch-BH vs BH vs brute, 1 GPU); p50/p99 latency"
TARGET_ERR = 1e-2
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
# record bytes per term for 3D gravity: node = centroid 12 + mass 4 + diameter 4 +
# child range 8, padded to 32; point = position 12 + mass 4
NODE_RECORD_BYTES, POINT_RECORD_BYTES = 32, 16
# V_MFMA_F32_16X16X4_F32: 16*16*4 multiply-adds; one per 16 x 16 query-source block
MFMA_FLOPS_PER_PAIR = 8
# V_MFMA_F32_16X16X16_BF16: 16*16*16 multiply-adds per block; dense bf16 peak
MFMA_BF16_FLOPS_PER_PAIR = 32
BF16_PEAK_TFLOPS = 2500.0
FP64_PEAK_TFLOPS = 78.6
# algorithmic flops per pair (add/sub/mul = 1, sqrt/rcp/exp = 1, divide = 2)
FLOPS_PER_PAIR = {(3, "gravity"): 13, (3, "wn"): 22, (3, "smoothdist"): 12,
                  (2, "gravity"): 10, (2, "wn"): 14, (2, "smoothdist"): 9}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def percentile(x, p):
    return float(np.percentile(np.asarray(x), p)) if len(x) else None


# --------------------------------------------------------------------------
# workloads
# --------------------------------------------------------------------------
def make_workload(name, n, q, rmax=None):
    import workloads as W

    kernel, alpha = "gravity", 1.0
    if name == "plummer3d":
        pts, nrm, m, qs = W.plummer(n, q, seed_sources=1, seed_queries=2, rmax=rmax)
        desc = ("3D gravity, Plummer sphere N=%d sources (equal masses 1/N), Q=%d Plummer "
                "queries%s" % (n, q, "" if rmax is None else f", truncated at r<={rmax}"))
    elif name == "gmm2d":
        pts, nrm, m, qs = W.gmm2d(n, q, seed=3)
        desc = f"2D gravity, 8-component Gaussian mixture N={n}, Q={q}"
    elif name in ("icosphere_wn", "icosphere_smoothdist"):
        level = {1310720: 8, 327680: 7, 81920: 6}.get(n, 8)
        pts, nrm, m, qs = W.noisy_icosphere(level, q, seed=4)
        kernel = "wn" if name.endswith("wn") else "smoothdist"
        alpha = 10.0 if kernel == "smoothdist" else 1.0
        desc = (f"3D {'winding number' if kernel == 'wn' else 'smooth distance (alpha=10)'}, "
                f"noisy icosphere level {level} ({len(pts)} area-weighted dipoles), "
                f"Q={q} queries in the 10%-expanded bbox")
    elif name in ("cube3d_gravity", "cube3d_smoothdist"):
        pts, nrm, m, qs = W.uniform_cube(n, q, dim=3, seed=0)
        kernel = "gravity" if name.endswith("gravity") else "smoothdist"
        alpha = 10.0 if kernel == "smoothdist" else 1.0
        desc = (f"3D {'gravity' if kernel == 'gravity' else 'smooth distance (alpha=10)'}, "
                f"uniform unit cube N={n} (masses U[0.5,1.5]), Q={q} queries in [-0.1,1.1]^3")
    else:
        raise ValueError(name)
    lo, hi = W.bounds(pts)
    return dict(name=name, desc=desc, dim=pts.shape[1], kernel=kernel, alpha=alpha, pts=pts,
                nrm=nrm if kernel == "wn" else None, m=m, queries=qs, lo=lo, hi=hi)


def measure(wl, args, thetas, spls, headline_steps=0, warmup=0, want_counts=False):
    """Sweep brute force, Barnes-Hut and the stochastic estimator on one
    workload; optionally time the stochastic configuration at the target
    error for headline_steps steps."""
    import torch

    import sbh_capi as C

    dev = torch.device("cuda:0")
    dim, kern = wl["dim"], wl["kernel"]
    t1 = C.Tree(wl["pts"], wl["nrm"], wl["m"], wl["lo"], wl["hi"], alpha=wl["alpha"], split_log=1)
    t2 = C.Tree(wl["pts"], wl["nrm"], wl["m"], wl["lo"], wl["hi"], alpha=wl["alpha"], split_log=2)
    i1, i2 = t1.info(), t2.info()
    def root_children(t, info):
        nch = np.empty(info["num_nodes"], np.uint32)
        C.check(C.lib().sbh_tree_nodes(t.h, None, None, None, nch.ctypes.data, *([None] * 10)))
        return int(nch[0])

    nsub, nsub1 = root_children(t2, i2), root_children(t1, i1)

    def build_roofline(t, info):
        """Algorithmic bytes of a build: at every level a point's record is read
        to find its cell (16 B) and its index and cell digit written (4 + 2 B),
        for each internal node it lies in; each node's records are written
        once (~96 B)."""
        nn = info["num_nodes"]
        st, en, nch = (np.empty(nn, np.uint32) for _ in range(3))
        C.check(C.lib().sbh_tree_nodes(t.h, st.ctypes.data, en.ctypes.data, None,
                                       nch.ctypes.data, *([None] * 10)))
        inner = nch > 0
        point_levels = int((en[inner].astype(np.int64) - st[inner]).sum())
        nbytes = 22 * point_levels + 96 * nn
        return {"build_ms": t.build_ms, "point_levels": point_levels, "bytes": nbytes,
                "gb_per_s": nbytes / (t.build_ms / 1e3) / 1e9,
                "hbm_frac": nbytes / (t.build_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                "note": "wall time of sbh_tree_create incl. host input packing and copies"}
    log(f"[{wl['name']}] octree1 {i1['num_nodes']} nodes depth {i1['depth']} "
        f"{t1.build_ms:.1f} ms; octree2 {i2['num_nodes']} nodes depth {i2['depth']} "
        f"{t2.build_ms:.1f} ms; {nsub} subdomains (k=2), {nsub1} (k=1)")

    q32 = wl["queries"].astype(np.float32)
    nq = len(q32)
    qd = torch.from_numpy(np.ascontiguousarray(q32.T)).to(dev)
    out = torch.empty(nq, dtype=torch.float32, device=dev)
    cnt = torch.empty(2 * nq, dtype=torch.int32, device=dev)
    ne = min(args.err_queries, nq)
    qe = torch.from_numpy(np.ascontiguousarray(q32[:ne].T)).to(dev)
    ref = torch.empty(ne, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ref_ms = t1.eval_f64_device(kern, qe.data_ptr(), ne, ref.data_ptr())
    ref_np = ref.cpu().numpy()
    den = np.maximum(np.abs(ref_np), 1.0) if kern == "wn" else np.abs(ref_np)
    if kern == "smoothdist":
        # the distance form -log(w)/alpha, relative to max(|d|, 1/alpha)
        ref_d = -np.log(ref_np) / wl["alpha"]
        den_d = np.maximum(np.abs(ref_d), 1.0 / wl["alpha"])

    def err_of():
        r = out[:ne].double().cpu().numpy()
        return float(np.mean(np.abs(r - ref_np) / den))

    def dist_err_of():
        if kern != "smoothdist":
            return None
        r = out[:ne].double().cpu().numpy()
        d = -np.log(np.maximum(r, 1e-300)) / wl["alpha"]
        return float(np.mean(np.abs(d - ref_d) / den_d))

    def run(tree, method, steps, warm, **kw):
        for _ in range(warm):
            tree.eval_device(method, kern, qd.data_ptr(), nq, out.data_ptr(), **kw)
        torch.cuda.synchronize()
        wall, kms = [], []
        for _ in range(steps):
            t0 = time.perf_counter()
            kms.append(tree.eval_device(method, kern, qd.data_ptr(), nq, out.data_ptr(), **kw))
            torch.cuda.synchronize()
            wall.append((time.perf_counter() - t0) * 1e3)
        return wall, kms

    F = FLOPS_PER_PAIR[(dim, kern)]
    res = {"workload": wl["desc"], "n_sources": len(wl["pts"]), "n_queries": nq,
           "subdomains": nsub, "err_queries": ne, "fp64_reference_ms": ref_ms,
           "build_ms": {"octree1": t1.build_ms, "octree2": t2.build_ms},
           "build": {"octree1": build_roofline(t1, i1), "octree2": build_roofline(t2, i2)},
           "trees": {"octree1": i1, "octree2": i2}, "barnes_hut": [], "stochastic": []}

    _, kms = run(t1, C.BRUTE_FORCE, args.brute_steps, 1)
    ms = float(np.mean(kms))
    res["brute_force"] = {"kernel_ms": ms, "queries_per_s": nq / (ms / 1e3),
                          "mean_rel_err": err_of(), "distance_rel_err": dist_err_of(),
                          "tflops": len(wl["pts"]) * nq * F / (ms / 1e3) / 1e12}
    log(f"  brute force {ms:.1f} ms err {res['brute_force']['mean_rel_err']:.2e} "
        f"{res['brute_force']['tflops']:.1f} TFLOP/s")
    if dim == 3 and kern in ("gravity", "smoothdist"):
        # matrix-core distances; the timed region includes its Hilbert query sort
        _, kms = run(t1, C.BRUTE_FORCE, args.brute_steps, 1, flags=C.EVAL_MFMA)
        ms = float(np.mean(kms))
        res["brute_force_mfma"] = {"kernel_ms": ms, "queries_per_s": nq / (ms / 1e3),
                                   "mean_rel_err": err_of(),
                                   "tflops": len(wl["pts"]) * nq * F / (ms / 1e3) / 1e12}
        log(f"  brute force (MFMA) {ms:.1f} ms err "
            f"{res['brute_force_mfma']['mean_rel_err']:.2e} "
            f"{res['brute_force_mfma']['tflops']:.1f} TFLOP/s equivalent")
    for theta in thetas:
        _, kms = run(t1, C.BARNES_HUT, args.sweep_steps, 1, beta=1.0 / theta, vote_group=1,
                     flags=C.EVAL_SORT_QUERIES)
        ms = float(np.mean(kms))
        res["barnes_hut"].append({"theta": theta, "sorted": True, "kernel_ms": ms,
                                  "queries_per_s": nq / (ms / 1e3), "mean_rel_err": err_of(),
                                  "distance_rel_err": dist_err_of()})
        log(f"  BH theta={theta}: {ms:.2f} ms err {res['barnes_hut'][-1]['mean_rel_err']:.2e}"
            + ("" if kern != "smoothdist"
               else f" (distance {res['barnes_hut'][-1]['distance_rel_err']:.2e})"))
    trees = {2: (t2, nsub), 1: (t1, nsub1)}

    def stochastic_row(k, spl, sort, refined=False):
        tk, nsk = trees[k]
        flags = C.EVAL_SORT_QUERIES if sort else 0
        _, kms = run(tk, C.STOCHASTIC, args.sweep_steps, 1, samples_per_subdomain=spl,
                     seed=1, flags=flags)
        ms = float(np.mean(kms))
        row = {"tree_k": k, "samples_per_subdomain": spl, "samples_per_query": spl * nsk,
               "sorted": sort, "kernel_ms": ms, "queries_per_s": nq / (ms / 1e3),
               "mean_rel_err": err_of(), "distance_rel_err": dist_err_of()}
        if refined:
            row["refined"] = True
        res["stochastic"].append(row)
        log(f"  SBH k={k} spl={spl} sort={sort}: {ms:.2f} ms err {row['mean_rel_err']:.2e}"
            + ("" if kern != "smoothdist" else f" (distance {row['distance_rel_err']:.2e})"))
        return row

    for k, spl in [(2, s) for s in spls] + [(1, s) for s in spls]:
        for sort in (False, True):
            stochastic_row(k, spl, sort)
    # the grid brackets the target; for the headline workload, bisect the
    # sample budget of the fastest configuration that meets it (same tree and
    # query order) between the grid's last budget that misses and its first
    # that meets the target, to multiples of 32 samples per subdomain
    crit = "distance_rel_err" if kern == "smoothdist" else "mean_rel_err"
    ok = [r for r in res["stochastic"] if r[crit] <= TARGET_ERR]
    if ok and args.refine and headline_steps:
        best = min(ok, key=lambda r: r["kernel_ms"])
        k, sort, hi = best["tree_k"], best["sorted"], best["samples_per_subdomain"]
        below = [r["samples_per_subdomain"] for r in res["stochastic"]
                 if r["tree_k"] == k and r["sorted"] == sort and r[crit] > TARGET_ERR
                 and r["samples_per_subdomain"] < hi]
        lo = max(below) if below else 0
        while hi - lo > 32:
            mid = max(lo + 32, (lo + hi) // 64 * 32)
            if stochastic_row(k, mid, sort, refined=True)[crit] <= TARGET_ERR:
                hi = mid
            else:
                lo = mid

    def at_target(rows):
        ok = [r for r in rows if r["mean_rel_err"] <= TARGET_ERR]
        if ok:
            return min(ok, key=lambda r: r["kernel_ms"]), True
        return min(rows, key=lambda r: r["mean_rel_err"]), False

    head, head_ok = at_target(res["stochastic"])
    bh, bh_ok = at_target(res["barnes_hut"])
    if kern == "smoothdist":
        # the distance form -log(w)/alpha the reference's driver reports
        ok = [r for r in res["stochastic"] if r["distance_rel_err"] <= TARGET_ERR]
        res["stochastic_at_target_distance"] = (
            dict(min(ok, key=lambda r: r["kernel_ms"]), target_met=True) if ok else
            dict(min(res["stochastic"], key=lambda r: r["distance_rel_err"]), target_met=False))
    res["stochastic_at_target"] = dict(head, target_met=head_ok)
    res["barnes_hut_at_target"] = dict(bh, target_met=bh_ok)
    res["stochastic_over_brute_force"] = head["queries_per_s"] / res["brute_force"]["queries_per_s"]
    res["barnes_hut_over_brute_force"] = bh["queries_per_s"] / res["brute_force"]["queries_per_s"]

    hflags = C.EVAL_SORT_QUERIES if head["sorted"] else 0
    hkw = dict(samples_per_subdomain=head["samples_per_subdomain"], seed=1, flags=hflags)
    th = trees[head["tree_k"]][0]
    if headline_steps:
        wall, kms = run(th, C.STOCHASTIC, headline_steps, warmup, **hkw)
        res["headline"] = {"wall_ms": wall, "kernel_ms": kms, "mean_rel_err": err_of()}
    if want_counts and head["sorted"]:
        # the estimator kernel alone (for the roofline): the queries permuted
        # into the Hilbert order beforehand and evaluated in that order, which
        # is exactly the launch a sorted evaluation makes after its sort
        order = torch.empty(nq, dtype=torch.int32, device=dev)
        C.check(C.lib().sbh_query_order(th.h, ctypes.c_void_p(qd.data_ptr()), nq,
                                        ctypes.c_void_p(order.data_ptr()),
                                        C.EVAL_DEVICE_POINTERS))
        qs = qd[:, order.long()].contiguous()
        torch.cuda.synchronize()
        kk = [th.eval_device(C.STOCHASTIC, kern, qs.data_ptr(), nq, out.data_ptr(),
                             samples_per_subdomain=hkw["samples_per_subdomain"], seed=1)
              for _ in range(max(1, headline_steps))]
        res["kernel_only_ms"] = float(np.mean(kk))
        res["kernel_only_ms_all"] = [float(v) for v in kk]
        # the work that launch performs (same seed, same 64-query groups):
        # kernel terms evaluated after the node cache's reuse, path levels stepped
        th.eval_device(C.STOCHASTIC, kern, qs.data_ptr(), nq, out.data_ptr(),
                       counts_ptr=cnt.data_ptr(), samples_per_subdomain=hkw["samples_per_subdomain"],
                       seed=1, flags=C.EVAL_WORK_COUNTS)
        c = cnt.cpu().numpy().astype(np.int64)
        res["work"] = {"terms": int(c[:nq].sum()), "path_levels": int(c[nq:].sum())}
        del qs, order
    elif want_counts and headline_steps:
        res["kernel_only_ms"] = float(np.mean(res["headline"]["kernel_ms"]))
    if want_counts:
        # the CPU plugin's semantics at the same budget: one path per query
        _, kms = run(th, C.STOCHASTIC, args.sweep_steps, 1,
                     samples_per_subdomain=hkw["samples_per_subdomain"], seed=1,
                     flags=C.EVAL_SORT_QUERIES | C.EVAL_PER_QUERY_PATHS)
        ms = float(np.mean(kms))
        res["stochastic_per_query_paths"] = {
            "tree_k": head["tree_k"], "samples_per_subdomain": hkw["samples_per_subdomain"],
            "sorted": True, "kernel_ms": ms, "queries_per_s": nq / (ms / 1e3),
            "mean_rel_err": err_of()}
        log(f"  SBH per-query paths: {ms:.2f} ms err "
            f"{res['stochastic_per_query_paths']['mean_rel_err']:.2e}")
        th.eval_device(C.STOCHASTIC, kern, qd.data_ptr(), nq, out.data_ptr(),
                       counts_ptr=cnt.data_ptr(), samples_per_subdomain=hkw["samples_per_subdomain"],
                       seed=1, flags=hflags | C.EVAL_SPLIT_COUNTS)
        c = cnt.cpu().numpy().astype(np.int64)
        res["terms"] = {"node_terms": int(c[:nq].sum()), "point_terms": int(c[nq:].sum()),
                        "flops_per_term": F}
    t1.close()
    t2.close()
    del qd, out, cnt, qe, ref
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--q", type=int, default=1 << 20)
    ap.add_argument("--err-queries", type=int, default=16384)
    ap.add_argument("--sweep-steps", type=int, default=3)
    ap.add_argument("--brute-steps", type=int, default=2)
    ap.add_argument("--spl", type=int, nargs="+", default=[1, 4, 16, 64, 256, 512, 768, 1024],
                    help="samples per subdomain swept for the stochastic estimator")
    ap.add_argument("--rmax", type=float, default=None,
                    help="truncate the headline Plummer sphere at this radius")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="approximate CPU work of the baseline leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-refine", dest="refine", action="store_false",
                    help="report the best grid configuration without refining its sample "
                         "budget")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-variants", action="store_true", help="skip the variant workloads")
    ap.add_argument("--serve-requests", type=int, default=600)
    ap.add_argument("--serve-walk-k", type=int, choices=[1, 2], default=1,
                    help="split_log of the trees serving Barnes-Hut requests (the estimator's "
                         "trees are k=2)")
    ap.add_argument("--serve-load", type=float, default=0.5,
                    help="Poisson offered load as a fraction of the measured burst capacity")
    ap.add_argument("--tree-k", type=int, default=2,
                    help="profiling mode: split_log of the stochastic estimator's tree")
    ap.add_argument("--no-sort", action="store_true",
                    help="profiling mode: evaluate the stochastic estimator in input order")
    ap.add_argument("--profile-kernel",
                    choices=["stochastic", "bh", "brute", "mfma", "double", "serve", "build",
                             "precision", "c1"],
                    default=None,
                    help="profiling mode: build the headline trees, run only this method "
                         "--steps times (stochastic at --spl[0], Barnes-Hut at --theta) and exit")
    ap.add_argument("--theta", type=float, default=1.0)
    ap.add_argument("--kernel", choices=["gravity", "smoothdist"], default="gravity",
                    help="profiling mode precision: the kernel of the 2^18 uniform-cube case")
    ap.add_argument("--detail", default=None,
                    help="where the full record (sweeps, variants, serving, CPU legs) is written "
                         "(default sbh_bench_detail.json in the system's temporary directory; "
                         "profiling modes write none unless given)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="with --profile-kernel: also run a plain read+write kernel over this "
                         "many MiB (a counter calibration kernel of known traffic)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); this build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import torch

    assert torch.cuda.is_available(), "no GPU visible"
    log(f"generating Plummer N={args.n} Q={args.q}")
    wl = make_workload("plummer3d", args.n, args.q, rmax=args.rmax)

    if args.profile_kernel:
        profile(wl, args)
        return
    if args.detail is None:
        args.detail = os.path.join(tempfile.gettempdir(), "sbh_bench_detail.json")

    head = measure(wl, args, (0.3, 0.4, 0.5, 0.6, 0.7, 1.0), args.spl,
                   headline_steps=args.steps, warmup=args.warmup, want_counts=True)
    hs = head["stochastic_at_target"]
    wall, kms = head["headline"]["wall_ms"], head["headline"]["kernel_ms"]
    ms_per_step = float(np.mean(wall))
    value = head["n_queries"] / (ms_per_step / 1e3)
    roofline = estimator_roofline(head)

    variants = None
    if not args.no_variants:
        variants = []
        specs = [("plummer3d", 1 << 20, 1 << 20, 10.0, (0.5, 1.0), (1, 4, 8, 16)),
                 ("gmm2d", 1 << 20, 1 << 20, None, (0.5, 1.0), (1, 4, 16, 64)),
                 ("icosphere_wn", 1310720, 1 << 21, None, (0.5, 1.0), (1, 4, 16, 64)),
                 ("icosphere_smoothdist", 1310720, 1 << 21, None, (0.5, 1.0),
                  (1, 4, 16, 64, 256, 1024, 1536))]
        vargs = argparse.Namespace(**vars(args))
        vargs.sweep_steps, vargs.brute_steps = 2, 1
        for name, n, q, rmax, thetas, spls in specs:
            log(f"variant {name}")
            v = measure(make_workload(name, n, q, rmax=rmax), vargs, thetas, spls)
            variants.append({k: v.get(k) for k in (
                "workload", "n_sources", "n_queries", "subdomains", "build_ms", "brute_force",
                "brute_force_mfma", "stochastic_at_target", "stochastic_at_target_distance",
                "barnes_hut_at_target", "stochastic_over_brute_force",
                "barnes_hut_over_brute_force", "barnes_hut", "stochastic")})

    precision = None if args.no_variants else precision_variant(args)
    double_build = None if args.no_variants else double_build_variant(args)
    serving = None if args.no_serve else serving_bench(args)
    cpu = cpu_c1 = None
    if not args.no_cpu:
        cpu = cpu_baseline(wl, hs["samples_per_subdomain"], hs["tree_k"], args.cpu_seconds, value)
        cpu_c1 = cpu_baseline_c1(cpu["cores"])

    bf = head["brute_force"]
    bh = head["barnes_hut_at_target"]
    line = {
        "metric": METRIC, "value": value, "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded numpy)",
        "config": {
            "workload": "3D gravity, untruncated Plummer sphere (scale radius 1, equal masses "
                        "summing to 1), 2^20 sources and 2^20 queries, stochastic Barnes-Hut at "
                        "<=1e-2 mean rel err vs float64 brute force",
            "n_sources": head["n_sources"], "n_queries": head["n_queries"],
            "plummer_rmax": args.rmax, "method": "stochastic",
            "tree": f"octree k={hs['tree_k']}, leaf size 1",
            "samples_per_subdomain": hs["samples_per_subdomain"],
            "samples_per_query": hs["samples_per_query"], "sorted_queries": hs["sorted"],
            "mean_rel_err": head["headline"]["mean_rel_err"], "target_err_met": hs["target_met"],
            "err_queries": head["err_queries"],
        },
        "latency_ms": ({"p50": serving["poisson"]["p50_ms"], "p99": serving["poisson"]["p99_ms"],
                        "kind": "serving request latency (submit to results in host memory), "
                                "Poisson arrivals at {:.0%} of burst capacity".format(
                                    args.serve_load)} if serving else None),
        "step_latency_ms": {"p50": percentile(wall, 50), "p99": percentile(wall, 99),
                            "kind": "one 2^20-query evaluation incl. its query sort"},
        "roofline": roofline,
        "cpu_baseline": ({k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample",
                                              "host", "gpu_over_cpu")} if cpu else None),
        "speedup_vs_brute_force": value / bf["queries_per_s"],
        "brute_force": {"kernel_ms": bf["kernel_ms"], "queries_per_s": bf["queries_per_s"],
                        "mean_rel_err": bf["mean_rel_err"]},
        "barnes_hut": {"theta": bh["theta"], "kernel_ms": bh["kernel_ms"],
                       "queries_per_s": bh["queries_per_s"], "mean_rel_err": bh["mean_rel_err"],
                       "target_err_met": bh["target_met"],
                       "over_brute_force": head["barnes_hut_over_brute_force"]},
        "serving": ({"burst_queries_per_s": serving["burst"]["queries_per_s"],
                     "poisson_queries_per_s": serving["poisson"]["queries_per_s"],
                     "p50_ms": serving["poisson"]["p50_ms"], "p99_ms": serving["poisson"]["p99_ms"],
                     "requests": serving["requests"]} if serving else None),
        "detail": args.detail,
    }
    detail = dict(line, headline=head, variants=variants, precision_variant=precision,
                  double_build=double_build, serving_detail=serving, cpu_baseline_full=cpu,
                  cpu_baseline_c1=cpu_c1)
    os.makedirs(os.path.dirname(os.path.abspath(args.detail)), exist_ok=True)
    with open(args.detail, "w") as f:
        json.dump(detail, f, indent=1, default=float)
    log(f"detail written to {args.detail}")
    log("summary: " + json.dumps(summary_of(detail), default=float))
    text = json.dumps(line, default=float)
    assert len(text) < 4096, len(text)
    print(text, flush=True)


def summary_of(d):
    """A few figures of the detail record, for the log."""
    out = {}
    for v in d.get("variants") or []:
        sa = v["stochastic_at_target"]
        out[v["workload"][:40]] = {
            "brute_ms": v["brute_force"]["kernel_ms"],
            "bh_x": v["barnes_hut_over_brute_force"], "stoch_x": v["stochastic_over_brute_force"],
            "stoch_err": sa["mean_rel_err"], "stoch_ok": sa["target_met"]}
    return out


# algorithmic flops per path level of the estimator: the opening ratio (3 sub,
# 3 products summed = 5, sqrt, times the diameter's reciprocal = 10), the
# roulette probability and draw (reciprocal, product, min, compare = 4), the
# path probability (1 - p, product = 2), the level's weight and its sum
# (product, reciprocal, 2 products, add = 5) and the running beta (max = 1)
FLOPS_PER_PATH_LEVEL = 22


def estimator_roofline(head):
    """Roofline of the dominant kernel (the sorted-sample estimator launch,
    split into tasks per 64-query block and range of root subdomains, and the
    pass adding the ranges' parts, timed together with hipEvents on their
    stream): algorithmic flops = the kernel
    terms it evaluates x flops per term + the path levels it steps x flops per
    level, counted by its work-count mode; peak = MI355X FP32 (157.3 TFLOP/s).
    The counters put it on VALU issue; HBM traffic is a few percent of peak."""
    kms = head["kernel_only_ms"]
    F = head["terms"]["flops_per_term"]
    w = head["work"]
    flops = w["terms"] * F + w["path_levels"] * FLOPS_PER_PATH_LEVEL
    achieved = flops / (kms / 1e3) / 1e12
    ref_bytes = (head["terms"]["node_terms"] * NODE_RECORD_BYTES
                 + head["terms"]["point_terms"] * POINT_RECORD_BYTES)
    traffic = valu = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    hs = head["stochastic_at_target"]
    key = "k_stochastic_plummer_1m_k%d_spl%d_%s" % (
        hs["tree_k"], hs["samples_per_subdomain"], "sorted" if hs["sorted"] else "unsorted")
    if os.path.exists(tpath):
        rec = json.load(open(tpath))
        traffic = rec.get(key)
        valu = rec.get("valu_issue", {}).get(key)
    return {
        "bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
        "kernel": "k_stochastic_sorted_split<3, GRAVITY> + k_reduce_stoch_split",
        "kernel_ms": kms,
        "units": {"terms": w["terms"], "flops_per_term": F, "path_levels": w["path_levels"],
                  "flops_per_level": FLOPS_PER_PATH_LEVEL},
        "frac_non_packed_fp32": achieved / (FP32_PEAK_TFLOPS / 2),
        "valu_issue_frac": valu,
        "hbm_frac": (traffic / (kms / 1e3) / 1e9 / HBM_PEAK_GBS if traffic else None),
        "reference_terms": head["terms"]["node_terms"] + head["terms"]["point_terms"],
        "reference_term_bytes_over_hbm_peak": ref_bytes / (kms / 1e3) / 1e9 / HBM_PEAK_GBS,
    }


def precision_variant(args):
    """Brute force on the 2^18 x 2^18 uniform cube (gravity, smooth distance):
    the direct fp32 kernel and the matrix-core kernel, each with its float64
    error per result over all queries."""
    import torch

    import sbh_capi as C

    dev = torch.device("cuda:0")
    rows = []
    for name in ("cube3d_gravity", "cube3d_smoothdist"):
        wl = make_workload(name, 1 << 18, 1 << 18)
        kern = wl["kernel"]
        t = C.Tree(wl["pts"], None, wl["m"], wl["lo"], wl["hi"], alpha=wl["alpha"], split_log=1)
        q32 = wl["queries"].astype(np.float32)
        nq = len(q32)
        qd = torch.from_numpy(np.ascontiguousarray(q32.T)).to(dev)
        out = torch.empty(nq, dtype=torch.float32, device=dev)
        ref = torch.empty(nq, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        ref_ms = t.eval_f64_device(kern, qd.data_ptr(), nq, ref.data_ptr())
        ref_np = ref.cpu().numpy()
        F = FLOPS_PER_PAIR[(3, kern)]
        pairs = float(len(wl["pts"])) * nq
        row = {"workload": wl["desc"], "fp64_ms": ref_ms}
        for label, flags in (("direct", 0), ("mfma", C.EVAL_MFMA),
                             ("mfma_bf16", C.EVAL_MFMA_BF16)):
            t.eval_device(C.BRUTE_FORCE, kern, qd.data_ptr(), nq, out.data_ptr(), flags=flags)
            torch.cuda.synchronize()
            kms = [t.eval_device(C.BRUTE_FORCE, kern, qd.data_ptr(), nq, out.data_ptr(),
                                 flags=flags) for _ in range(max(1, args.brute_steps))]
            ms = float(np.mean(kms))
            rel = np.abs(out.double().cpu().numpy() - ref_np) / np.abs(ref_np)
            row[label] = {"kernel_ms": ms, "queries_per_s": nq / (ms / 1e3),
                          "tflops": pairs * F / (ms / 1e3) / 1e12,
                          "frac_fp32_peak": pairs * F / (ms / 1e3) / 1e12 / FP32_PEAK_TFLOPS,
                          "mean_rel_err": float(rel.mean()), "max_rel_err": float(rel.max())}
            if label == "mfma":
                mf = pairs * MFMA_FLOPS_PER_PAIR / (ms / 1e3) / 1e12
                row[label]["matrix_tflops"] = mf
                row[label]["matrix_frac"] = mf / FP32_PEAK_TFLOPS
            elif label == "mfma_bf16":
                mf = pairs * MFMA_BF16_FLOPS_PER_PAIR / (ms / 1e3) / 1e12
                row[label]["matrix_tflops"] = mf
                row[label]["matrix_frac"] = mf / BF16_PEAK_TFLOPS
            log(f"[{name}] brute force {label}: {ms:.2f} ms, "
                f"{row[label]['tflops']:.1f} TFLOP/s, mean rel err {rel.mean():.2e}, "
                f"max {rel.max():.2e}")
        row["mfma_speedup"] = row["direct"]["kernel_ms"] / row["mfma"]["kernel_ms"]
        row["mfma_bf16_speedup"] = row["direct"]["kernel_ms"] / row["mfma_bf16"]["kernel_ms"]
        rows.append(row)
        t.close()
        del qd, out, ref
        torch.cuda.empty_cache()
    return rows


def double_build_variant(args):
    """The double build (sbh_gpu64, float64 throughout) on the headline
    workload: Barnes-Hut (theta 1.0, per query) and the estimator (512 samples
    per subdomain) over all queries in Hilbert order, brute force on the first
    2^18 queries; errors against the double brute force on the first
    err_queries queries."""
    import sbh_gpu64

    wl = make_workload("plummer3d", args.n, args.q, rmax=args.rmax)
    pts, q = wl["pts"], wl["queries"]
    g = sbh_gpu64.OctreeGPU2(pts, np.zeros_like(pts), wl["m"], wl["lo"], wl["hi"])
    e = min(args.err_queries, len(q))
    ref, _ = sbh_gpu64.eval_brute_force_gravity(g, q[:e])
    nb = min(1 << 18, len(q))
    sbh_gpu64.eval_brute_force_gravity(g, q[:nb])
    _, bms = sbh_gpu64.eval_brute_force_gravity(g, q[:nb])
    out = {"workload": wl["desc"] + " (float64 inputs, tree, terms and sums)",
           "build_ms": g.build_ms,
           "brute_force": {"kernel_ms": bms, "queries": nb, "queries_per_s": nb / (bms / 1e3),
                           "tflops": float(len(pts)) * nb * FLOPS_PER_PAIR[(3, "gravity")]
                           / (bms / 1e3) / 1e12}}
    out["brute_force"]["frac_fp64_peak"] = out["brute_force"]["tflops"] / FP64_PEAK_TFLOPS

    def timed(fn, *a, **kw):
        fn(*a, **kw)
        v, ms = fn(*a, **kw)
        err = float(np.mean(np.abs(v[:e] - ref) / np.abs(ref)))
        return {"kernel_ms": ms, "queries_per_s": len(q) / (ms / 1e3), "mean_rel_err": err}

    out["barnes_hut"] = dict(theta=1.0, **timed(sbh_gpu64.eval_barnes_hut_gravity, g, q, 1.0,
                                                vote_group=1, sort_queries=True))
    out["stochastic"] = dict(samples_per_subdomain=512, **timed(
        sbh_gpu64.multi_level_prefix_control_variate_gravity, g, q, 512, seed=1,
        sort_queries=True))
    log(f"double build: brute {bms:.1f} ms / {nb} q, BH {out['barnes_hut']['kernel_ms']:.1f} ms "
        f"(err {out['barnes_hut']['mean_rel_err']:.2e}), stochastic "
        f"{out['stochastic']['kernel_ms']:.1f} ms (err {out['stochastic']['mean_rel_err']:.2e})")
    del g
    return out


def profile(wl, args):
    """Profiling mode: only the requested method on the headline trees."""
    import torch

    import sbh_capi as C

    dev = torch.device("cuda:0")
    if args.profile_kernel == "serve":
        # the serving measurement alone (--serve-requests requests)
        r = serving_bench(args)
        log(f"serve: burst {r['burst']['queries_per_s'] / 1e6:.1f} M q/s, poisson p50 "
            f"{r['poisson']['p50_ms']:.1f} ms, p99 {r['poisson']['p99_ms']:.1f} ms")
        if args.detail:
            os.makedirs(os.path.dirname(os.path.abspath(args.detail)), exist_ok=True)
            with open(args.detail, "w") as f:
                json.dump(r, f, indent=1, default=float)
        return
    if args.profile_kernel == "precision":
        # the precision case: 2^18 x 2^18 uniform cube, brute force direct,
        # with fp32 and with bf16 matrix-core operands, --steps times each
        pw = make_workload("cube3d_" + args.kernel, 1 << 18, 1 << 18)
        t = C.Tree(pw["pts"], None, pw["m"], pw["lo"], pw["hi"], alpha=pw["alpha"])
        q32 = pw["queries"].astype(np.float32)
        qd = torch.from_numpy(np.ascontiguousarray(q32.T)).to(dev)
        out = torch.empty(len(q32), dtype=torch.float32, device=dev)
        for label, flags in (("direct", 0), ("mfma", C.EVAL_MFMA), ("mfma_bf16", C.EVAL_MFMA_BF16)):
            for _ in range(args.steps):
                ms = t.eval_device(C.BRUTE_FORCE, args.kernel, qd.data_ptr(), len(q32),
                                   out.data_ptr(), flags=flags)
                log(f"precision {args.kernel} {label}: {ms:.3f} ms")
        torch.cuda.synchronize()
        return
    if args.profile_kernel == "c1":
        # the reference's CPU-runnable case on the GPU: brute force, Barnes-Hut
        # (beta 2, leaf sizes 8 and 1, per query, sorted) and the estimator
        # (octree k=2, 4 samples per subdomain, sorted), --steps times each
        import sbh_gpu
        import workloads as W

        pts, _, m, q = W.uniform_cube(16384, 16384, seed=0)
        lo, hi = W.bounds(pts)
        nrm = np.zeros_like(pts)
        q32 = q.astype(np.float32)
        g8 = sbh_gpu.OctreeGPU1(pts, nrm, m, lo, hi, leaf_size=8)
        g1 = sbh_gpu.OctreeGPU1(pts, nrm, m, lo, hi)
        g2 = sbh_gpu.OctreeGPU2(pts, nrm, m, lo, hi)
        for _ in range(args.steps):
            log(f"c1 brute: {sbh_gpu.eval_brute_force_gravity(g1, q32)[1]:.3f} ms")
            for name, g in (("leaf 8", g8), ("leaf 1", g1)):
                _, ms = sbh_gpu.eval_barnes_hut_gravity(g, q32, 2.0, vote_group=1,
                                                        sort_queries=True)
                log(f"c1 barnes-hut {name}: {ms:.3f} ms")
            _, ms = sbh_gpu.multi_level_prefix_control_variate_gravity(g2, q32, 4, seed=1,
                                                                      sort_queries=True)
            log(f"c1 stochastic: {ms:.3f} ms")
        return
    if args.profile_kernel == "build":
        # the headline tree (octree k = --tree-k) built --steps times
        for _ in range(args.steps):
            t = C.Tree(wl["pts"], None, wl["m"], wl["lo"], wl["hi"], split_log=args.tree_k)
            log(f"build k={args.tree_k}: {t.build_ms:.2f} ms, {t.info()['num_nodes']} nodes")
            t.close()
        return
    if args.profile_kernel == "double":
        # the double build: Barnes-Hut at --theta and the estimator at --spl[0],
        # Hilbert-sorted, --steps times each
        import sbh_gpu64

        pts, q = wl["pts"], wl["queries"]
        g = sbh_gpu64.OctreeGPU2(pts, np.zeros_like(pts), wl["m"], wl["lo"], wl["hi"])
        for _ in range(args.steps):
            _, ms = sbh_gpu64.eval_barnes_hut_gravity(g, q, 1.0 / args.theta, vote_group=1,
                                                      sort_queries=True)
            log(f"double bh: {ms:.3f} ms")
            _, ms = sbh_gpu64.multi_level_prefix_control_variate_gravity(
                g, q, args.spl[0], seed=1, sort_queries=True)
            log(f"double stochastic: {ms:.3f} ms")
        return
    if args.copy_mib:
        a = torch.empty(args.copy_mib << 18, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)  # one read and one write of copy_mib MiB
        torch.cuda.synchronize()
        log(f"read+write {args.copy_mib} MiB x3 (calibration)")
        del a, b
    t = C.Tree(wl["pts"], None, wl["m"], wl["lo"], wl["hi"],
               split_log=args.tree_k if args.profile_kernel == "stochastic" else 1)
    q32 = wl["queries"].astype(np.float32)
    nq = len(q32)
    qd = torch.from_numpy(np.ascontiguousarray(q32.T)).to(dev)
    out = torch.empty(nq, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    for _ in range(args.steps):
        if args.profile_kernel == "stochastic":
            ms = t.eval_device(C.STOCHASTIC, "gravity", qd.data_ptr(), nq, out.data_ptr(),
                               samples_per_subdomain=args.spl[0], seed=1,
                               flags=0 if args.no_sort else C.EVAL_SORT_QUERIES)
        elif args.profile_kernel == "bh":
            ms = t.eval_device(C.BARNES_HUT, "gravity", qd.data_ptr(), nq, out.data_ptr(),
                               beta=1.0 / args.theta, vote_group=1, flags=C.EVAL_SORT_QUERIES)
        elif args.profile_kernel == "mfma":
            ms = t.eval_device(C.BRUTE_FORCE, "gravity", qd.data_ptr(), nq, out.data_ptr(),
                               flags=C.EVAL_MFMA)
        else:
            ms = t.eval_device(C.BRUTE_FORCE, "gravity", qd.data_ptr(), nq, out.data_ptr())
        log(f"{args.profile_kernel}: {ms:.3f} ms")
    torch.cuda.synchronize()


SERVE_TREES = [  # name, dim, kernel, points
    ("plummer_r10_3d_gravity_64k", 3, "gravity", 1 << 16),
    ("gmm_2d_gravity_128k", 2, "gravity", 1 << 17),
    ("icosphere7_3d_wn_328k", 3, "wn", 327680),
    ("cube_3d_smoothdist_512k", 3, "smoothdist", 1 << 19),
    ("plummer_r10_3d_gravity_1m", 3, "gravity", 1 << 20),
    ("gmm_2d_smoothdist_2m", 2, "smoothdist", 1 << 21),
    ("cube_3d_gravity_4m", 3, "gravity", 1 << 22),
    ("plummer_r10_3d_smoothdist_4m", 3, "smoothdist", 1 << 22),
]


def _serve_sources(name, n):
    import workloads as W

    if name.startswith("plummer"):
        return W.plummer(n, 1, seed_sources=11, rmax=10.0)[:3]
    if name.startswith("gmm"):
        return W.gmm2d(n, 1, seed=12)[:3]
    if name.startswith("icosphere"):
        return W.noisy_icosphere(7, 1, seed=13)[:3]
    return W.uniform_cube(n, 1, seed=14)[:3]


def serving_bench(args):
    """Eight resident trees (64k-4M sources, 2D/3D, all three kernels); requests
    of log-uniform size 1e3-1e6 queries, each picking a tree, Barnes-Hut
    (theta 0.5) or the stochastic estimator (128 samples per query) and a
    seed.  A burst of all requests measures capacity; then a Poisson stream at
    --serve-load of that capacity measures latency."""
    import sbh_gpu

    rng = np.random.default_rng(21)
    srv = sbh_gpu.Server(4 << 30, 4)
    trees = []
    for name, dim, kern, n in SERVE_TREES:
        pts, nrm, m = _serve_sources(name, n)
        lo, hi = pts.min(0) - 0.01, pts.max(0) + 0.01
        alpha = 10.0 if kern == "smoothdist" else 1.0
        cls = sbh_gpu.OctreeGPU2 if dim == 3 else sbh_gpu.QuadtreeGPU2
        tr = cls(pts, nrm, m, lo, hi, alpha=alpha)
        nsub = int(tr.nodes()["nchild"][0])
        rec = dict(name=name, dim=dim, kernel=kern, n=len(pts), lo=lo, hi=hi,
                   id=srv.add_tree(tr), tree=tr, subdomains=nsub)
        # Barnes-Hut walks a 2^d-ary tree of the same sources (as the 1M-source
        # measurements do), the estimator samples the 4^d-ary one
        rec["walk_tree"] = tr
        rec["walk_id"] = rec["id"]
        if args.serve_walk_k == 1:
            cls1 = sbh_gpu.OctreeGPU1 if dim == 3 else sbh_gpu.QuadtreeGPU1
            rec["walk_tree"] = cls1(pts, nrm, m, lo, hi, alpha=alpha)
            rec["walk_id"] = srv.add_tree(rec["walk_tree"])
        trees.append(rec)
        log(f"serve tree {name}: {len(pts)} points, {nsub} subdomains, build {tr.build_ms:.1f} ms")
    nreq = args.serve_requests
    reqs = []
    for _ in range(nreq):
        t = trees[int(rng.integers(len(trees)))]
        q = int(np.exp(rng.uniform(np.log(1e3), np.log(1e6))))
        pts = t["lo"] + (t["hi"] - t["lo"]) * rng.random((q, t["dim"]))
        method = "stochastic" if rng.random() < 0.5 else "barnes_hut"
        spl = max(1, 128 // max(1, t["subdomains"]))
        reqs.append((t, method, pts.astype(np.float32), spl, int(rng.integers(1 << 40))))
    total_q = sum(len(r[2]) for r in reqs)

    def run(arrivals):
        hs = []
        t0 = time.perf_counter()
        for (t, method, pts, spl, seed), at in zip(reqs, arrivals):
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            hs.append(srv.submit(t["walk_id"] if method == "barnes_hut" else t["id"], method,
                                 t["kernel"], pts, beta=2.0, samples_per_subdomain=spl,
                                 seed=seed))
        lats = [h.result(600.0)[1] for h in hs]
        el = time.perf_counter() - t0
        return np.asarray(lats), el

    for h in [srv.submit(t["walk_id"] if m_ == "barnes_hut" else t["id"], m_, t["kernel"],
                         p[:2000], beta=2.0, samples_per_subdomain=s_, seed=1)
              for (t, m_, p, s_, _) in reqs[:16]]:
        h.result(600.0)
    lat_b, el_b = run(np.zeros(nreq))
    cap = total_q / el_b
    log(f"serve burst: {nreq} requests, {total_q} queries in {el_b:.2f} s -> {cap / 1e6:.1f} M q/s")
    rate = args.serve_load * cap / (total_q / nreq)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    lat_p, el_p = run(arrivals)
    log(f"serve poisson: rate {rate:.1f} req/s, p50 {np.percentile(lat_p, 50):.1f} ms, "
        f"p99 {np.percentile(lat_p, 99):.1f} ms")
    st = srv.stats()
    # the tail: each of the slowest requests once more, alone on the server
    tail = []
    for i in np.argsort(lat_p)[::-1][:6]:
        t, method, pts, spl, seed = reqs[i]
        h = srv.submit(t["walk_id"] if method == "barnes_hut" else t["id"], method, t["kernel"],
                       pts, beta=2.0, samples_per_subdomain=spl, seed=seed)
        alone = h.result(600.0)[1]
        tail.append({"latency_ms": float(lat_p[i]), "alone_ms": alone, "queries": len(pts),
                     "method": method, "tree": t["name"]})
        log(f"serve tail: {len(pts)} queries {method} on {t['name']}: {lat_p[i]:.1f} ms in the "
            f"stream, {alone:.1f} ms alone")
    per_request = [{"queries": len(r[2]), "method": r[1], "tree": r[0]["name"],
                    "arrival_s": float(a), "latency_ms": float(l)}
                   for r, a, l in zip(reqs, arrivals, lat_p)]
    # shut the engine down before the next leg: its scheduler thread and
    # streams end here, not at interpreter exit
    import gc

    tree_info = [{k: t[k] for k in ("name", "dim", "kernel", "n", "subdomains")} for t in trees]
    resident = len({id(t["tree"]) for t in trees} | {id(t["walk_tree"]) for t in trees})
    del srv, trees, reqs
    gc.collect()
    return {
        "trees": tree_info,
        "requests": nreq, "queries": int(total_q),
        "request_sizes": "log-uniform 1e3..1e6 queries",
        "methods": "barnes_hut theta=0.5 / stochastic 128 samples per query, 50/50",
        "walk_tree_k": args.serve_walk_k,
        "resident_trees": resident,
        "burst": {"queries_per_s": cap, "elapsed_s": el_b,
                  "p50_ms": float(np.percentile(lat_b, 50)),
                  "p99_ms": float(np.percentile(lat_b, 99))},
        "poisson": {"rate_req_per_s": rate, "offered_load": args.serve_load,
                    "queries_per_s": total_q / el_p, "elapsed_s": el_p,
                    "p50_ms": float(np.percentile(lat_p, 50)),
                    "p99_ms": float(np.percentile(lat_p, 99))},
        "tail": tail, "per_request": per_request,
        "launches": st["launches"], "peak_request_bytes": st["peak_bytes"],
        "budget_bytes": st["budget_bytes"], "resident_tree_bytes": st["tree_bytes"],
    }


def host_cpus():
    """The host's CPUs as this process sees them: os.cpu_count() (what the
    reference's std::thread::hardware_concurrency() returns, src/sbh.cpp:32-36),
    the CPUs this process may run on (affinity) and the cgroup CPU quota."""
    info = {"os_cpu_count": os.cpu_count() or 1}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["affinity"] = info["os_cpu_count"]
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    info["cgroup_cpu_quota"] = quota
    model = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    info["model"] = model
    allowed = info["affinity"]
    if quota:
        allowed = min(allowed, max(1, int(round(quota))))
    info["allowed"] = allowed
    return info


def _best_of(fn, runs):
    best = float("inf")
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        best = min(best, time.perf_counter() - t0)
    return best


def cpu_baseline(wl, spl, k, budget_s, gpu_qps):
    """CPU baseline leg: the oracle (a C++ restatement of the reference's CPU
    estimator with the same threading model and its per-query sample paths)
    on this machine's host cores, on a bounded sample of the headline
    workload.  Timed with the reference's thread count (all CPUs the OS
    reports) and with the CPUs this process is allowed (affinity and cgroup
    quota) where that differs; best of 3 after a warm-up each; the faster
    setting is the baseline."""
    import oracle

    host = host_cpus()
    settings = sorted({host["os_cpu_count"], host["allowed"]})
    t0 = time.perf_counter()
    ot = oracle.Tree(wl["pts"], None, wl["m"], wl["lo"], wl["hi"], split_log=k)
    build_s = time.perf_counter() - t0
    q32 = wl["queries"].astype(np.float32)
    probe = q32[:1024]
    t0 = time.perf_counter()
    ot.stochastic("gravity", probe, spl, seed=1, group=1, nthreads=max(settings))
    per_q = (time.perf_counter() - t0) / len(probe)
    # about budget_s of wall time over all timed runs (3 per setting)
    nsample = int(min(len(q32), max(1024, budget_s / (3 * len(settings)) / max(per_q, 1e-9))))
    sample = q32[:nsample]
    timings = {}
    for th in settings:
        ot.stochastic("gravity", sample[:256], spl, seed=1, group=1, nthreads=th)  # warm-up
        timings[th] = _best_of(
            lambda: ot.stochastic("gravity", sample, spl, seed=1, group=1, nthreads=th), 3)
        log(f"cpu baseline: {nsample / timings[th]:.0f} q/s on {th} threads "
            f"({nsample} queries, best of 3: {timings[th]:.2f} s)")
    best_th = min(timings, key=timings.get)
    qps = nsample / timings[best_th]
    log(f"cpu baseline host: {host}; oracle tree build {build_s:.1f} s")
    return {"value": qps, "unit": "queries/s", "cores": best_th, "kind": "port",
            "sample": f"stochastic estimator, {spl} samples per subdomain, octree k={k}, one "
                      f"sample path per query (the reference CPU sampler's draws); first "
                      f"{nsample} of the {len(q32)} Plummer queries against all "
                      f"{len(wl['pts'])} sources; best of 3; tree build excluded "
                      f"({build_s:.1f} s, one thread)",
            "host": host,
            "queries_per_s_by_threads": {str(t): nsample / v for t, v in timings.items()},
            "gpu_over_cpu": gpu_qps / qps}


def cpu_baseline_c1(threads):
    """CPU baseline leg on the reference's CPU-runnable case (3D gravity, 16384
    uniform sources with masses U[0.5,1.5], 16384 queries in [-0.1,1.1]^3):
    brute force, Barnes-Hut (beta 2, leaf sizes 8 and 1, per query) and the
    stochastic estimator (256 samples per query: octree k=2, 4 per subdomain),
    the oracle on the host cores (best of 5 after a warm-up) beside the GPU, each against the
    float64 all-pairs sum."""
    import oracle
    import sbh_gpu
    import workloads as W

    pts, _, m, q = W.uniform_cube(16384, 16384, seed=0)
    lo, hi = W.bounds(pts)
    nrm = np.zeros_like(pts)
    q32 = q.astype(np.float32)
    nq = len(q32)
    g1 = sbh_gpu.OctreeGPU1(pts, nrm, m, lo, hi)
    ref, _ = sbh_gpu.eval_brute_force_f64_gravity(g1, q32)

    def err(v):
        return float(np.mean(np.abs(np.asarray(v, np.float64) - ref) / np.abs(ref)))

    def cpu(fn):
        v = fn()  # warm-up
        return v, _best_of(fn, 5)

    def gpu(fn, *a, **kw):
        fn(*a, **kw)
        v, ms = fn(*a, **kw)
        return v, ms

    rows = {}

    def row(name, cv, cs, gv, gms, **extra):
        rows[name] = dict(cpu_s=cs, cpu_queries_per_s=nq / cs, cpu_mean_rel_err=err(cv),
                          gpu_kernel_ms=gms, gpu_queries_per_s=nq / (gms / 1e3),
                          gpu_mean_rel_err=err(gv), gpu_over_cpu=(nq / (gms / 1e3)) / (nq / cs),
                          **extra)
        log(f"C1 {name}: cpu {cs * 1e3:.1f} ms (err {err(cv):.2e}), gpu {gms:.3f} ms "
            f"(err {err(gv):.2e})")

    o1 = oracle.Tree(pts, None, m, lo, hi, split_log=1)
    cv, cs = cpu(lambda: o1.brute_force("gravity", q32, nthreads=threads))
    gv, gms = gpu(sbh_gpu.eval_brute_force_gravity, g1, q32)
    row("brute_force", cv, cs, gv, gms)
    for leaf in (8, 1):
        ot = o1 if leaf == 1 else oracle.Tree(pts, None, m, lo, hi, split_log=1, leaf_size=leaf)
        gt = g1 if leaf == 1 else sbh_gpu.OctreeGPU1(pts, nrm, m, lo, hi, leaf_size=leaf)
        cv, cs = cpu(lambda: ot.barnes_hut("gravity", q32, 2.0, vote_group=1, nthreads=threads))
        gv, gms = gpu(sbh_gpu.eval_barnes_hut_gravity, gt, q32, 2.0, vote_group=1,
                      sort_queries=True)
        row(f"barnes_hut_beta2_leaf{leaf}", cv, cs, gv, gms)
    o2 = oracle.Tree(pts, None, m, lo, hi, split_log=2)
    g2 = sbh_gpu.OctreeGPU2(pts, nrm, m, lo, hi)
    cv, cs = cpu(lambda: o2.stochastic("gravity", q32, 4, seed=1, group=1, nthreads=threads))
    gv, gms = gpu(sbh_gpu.multi_level_prefix_control_variate_gravity, g2, q32, 4, seed=1,
                  sort_queries=True)
    row("stochastic_256_per_query", cv, cs, gv, gms,
        paths="CPU: one sample path per query; GPU: paths shared by a wavefront's 64 queries")
    return {"workload": "3D gravity, uniform unit cube N=16384 (masses U[0.5,1.5]), Q=16384 "
                        "queries in [-0.1,1.1]^3, fp32", "cores": threads, "kind": "port",
            "methods": rows}


if __name__ == "__main__":
    main()
