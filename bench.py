"""Benchmark: queries/s at matched 1e-2 mean relative error on one MI355X.

Workload (default): 3D gravity, N = 2^20 sources drawn from a Plummer sphere
(scale radius 1, equal masses summing to 1) and Q = 2^20 queries from the same
distribution; fp32; synthetic data generated here.

A step is one evaluation of all Q queries by the headline method: the
stochastic control-variate estimator at the smallest sample budget whose mean
relative error (against a float64 all-pairs reference on the first
--err-queries queries) is <= 1e-2.  Brute force and Barnes-Hut (theta 0.3-0.7)
are swept and reported beside it at their own matched error.

Inputs are resident in HBM before timing (torch tensors on cuda:0, passed to
the C ABI as device pointers).  One JSON line is printed on stdout; progress
goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "stochastic-barnes-hut_amd")
sys.path[:0] = [ROOT, PKG]

METRIC = "queries/s at 1e-2 rel-err (stoch-BH vs BH vs brute, 1 GPU); p50/p99 latency"
TARGET_ERR = 1e-2
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
FLOPS_PER_PAIR_GRAVITY_3D = 13  # add/sub/mul = 1, sqrt/rcp = 1, divide = 2


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def percentile(x, p):
    return float(np.percentile(np.asarray(x), p)) if len(x) else None


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
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="approximate CPU work of the baseline leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spl", type=int, nargs="+", default=[1, 4, 16, 64, 256, 512],
                    help="samples per subdomain swept for the stochastic estimator")
    ap.add_argument("--rmax", type=float, default=None,
                    help="truncate the Plummer sphere at this radius (default: untruncated)")
    ap.add_argument("--profile-kernel", choices=["stochastic", "bh", "brute"], default=None,
                    help="profiling mode: build the trees, run only this method --steps times "
                         "(stochastic at --spl[0], Barnes-Hut at --theta) and exit")
    ap.add_argument("--theta", type=float, default=1.0)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-load", type=float, default=0.5,
                    help="Poisson offered load as a fraction of the measured burst capacity")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run a plain device-to-device copy of this many MiB (a counter "
                         "calibration kernel of known traffic)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); this build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import torch

    import sbh_capi as C
    import workloads as W

    assert torch.cuda.is_available(), "no GPU visible"
    dev = torch.device("cuda:0")

    log(f"generating Plummer N={args.n} Q={args.q}")
    pts, nrm, m, queries = W.plummer(args.n, args.q, seed_sources=1, seed_queries=2,
                                     rmax=args.rmax)
    lo, hi = W.bounds(pts)

    log("building trees on the GPU")
    t1 = C.Tree(pts, None, m, lo, hi, split_log=1)   # Barnes-Hut (Octree1)
    t2 = C.Tree(pts, None, m, lo, hi, split_log=2)   # stochastic (Octree2)
    i1, i2 = t1.info(), t2.info()
    # a second build of each, timed warm
    t1b = C.Tree(pts, None, m, lo, hi, split_log=1)
    t2b = C.Tree(pts, None, m, lo, hi, split_log=2)
    build_ms = {"octree1": t1b.build_ms, "octree2": t2b.build_ms}
    t1b.close()
    t2b.close()
    log(f"octree1 nodes={i1['num_nodes']} depth={i1['depth']} build={build_ms['octree1']:.1f} ms; "
        f"octree2 nodes={i2['num_nodes']} depth={i2['depth']} build={build_ms['octree2']:.1f} ms")

    nq = args.q
    q32 = queries.astype(np.float32)
    qd = torch.from_numpy(np.ascontiguousarray(q32.T)).to(dev)           # SoA [3, nq]
    out = torch.empty(nq, dtype=torch.float32, device=dev)
    cnt = torch.empty(2 * nq, dtype=torch.int32, device=dev)
    ne = min(args.err_queries, nq)
    qe = torch.from_numpy(np.ascontiguousarray(q32[:ne].T)).to(dev)
    ref = torch.empty(ne, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    if args.copy_mib:
        a = torch.empty(args.copy_mib << 18, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {args.copy_mib} MiB x3 (calibration)")
        del a, b

    if args.profile_kernel:
        which = args.profile_kernel
        for _ in range(args.steps):
            if which == "stochastic":
                ms = t2.eval_device(C.STOCHASTIC, "gravity", qd.data_ptr(), nq, out.data_ptr(),
                                    samples_per_subdomain=args.spl[0], seed=1)
            elif which == "bh":
                ms = t1.eval_device(C.BARNES_HUT, "gravity", qd.data_ptr(), nq, out.data_ptr(),
                                    beta=1.0 / args.theta, vote_group=1,
                                    flags=C.EVAL_SORT_QUERIES)
            else:
                ms = t1.eval_device(C.BRUTE_FORCE, "gravity", qd.data_ptr(), nq, out.data_ptr())
            log(f"{which}: {ms:.3f} ms")
        torch.cuda.synchronize()
        return

    log(f"float64 reference on {ne} queries")
    ref_ms = t1.eval_f64_device("gravity", qe.data_ptr(), ne, ref.data_ptr())
    ref_np = ref.cpu().numpy()

    def err_of(res):
        r = res[:ne].double().cpu().numpy()
        return float(np.mean(np.abs(r - ref_np) / np.abs(ref_np)))

    def run(tree, method, steps, warmup, **kw):
        for _ in range(warmup):
            tree.eval_device(method, "gravity", qd.data_ptr(), nq, out.data_ptr(), **kw)
        torch.cuda.synchronize()
        wall, kern = [], []
        for _ in range(steps):
            t0 = time.perf_counter()
            kern.append(tree.eval_device(method, "gravity", qd.data_ptr(), nq, out.data_ptr(), **kw))
            torch.cuda.synchronize()
            wall.append((time.perf_counter() - t0) * 1e3)
        return wall, kern

    results = {"brute_force": None, "barnes_hut": [], "stochastic": []}

    log("brute force")
    wall, kern = run(t1, C.BRUTE_FORCE, args.brute_steps, 1)
    bf_err = err_of(out)
    bf_ms = float(np.mean(kern))
    pairs = float(args.n) * nq
    results["brute_force"] = {
        "queries_per_s": nq / (bf_ms / 1e3), "kernel_ms": bf_ms, "mean_rel_err": bf_err,
        "tflops": pairs * FLOPS_PER_PAIR_GRAVITY_3D / (bf_ms / 1e3) / 1e12}
    log(f"  {bf_ms:.1f} ms  err {bf_err:.2e}  {results['brute_force']['tflops']:.1f} TFLOP/s")

    for theta in (0.3, 0.4, 0.5, 0.6, 0.7, 1.0):
        for sort in (True,):
            flags = C.EVAL_SORT_QUERIES if sort else 0
            wall, kern = run(t1, C.BARNES_HUT, args.sweep_steps, 1, beta=1.0 / theta,
                             vote_group=1, flags=flags)
            e = err_of(out)
            ms = float(np.mean(kern))
            results["barnes_hut"].append({"theta": theta, "sorted": sort, "kernel_ms": ms,
                                          "queries_per_s": nq / (ms / 1e3), "mean_rel_err": e})
            log(f"  BH theta={theta} sort={sort}: {ms:.2f} ms err {e:.2e}")

    nsub = None
    for spl in args.spl:
        for sort in (False, True):
            flags = C.EVAL_SORT_QUERIES if sort else 0
            wall, kern = run(t2, C.STOCHASTIC, args.sweep_steps, 1, samples_per_subdomain=spl,
                             seed=1, flags=flags)
            e = err_of(out)
            ms = float(np.mean(kern))
            results["stochastic"].append({"samples_per_subdomain": spl, "sorted": sort,
                                          "kernel_ms": ms, "queries_per_s": nq / (ms / 1e3),
                                          "mean_rel_err": e})
            log(f"  SBH spl={spl} sort={sort}: {ms:.2f} ms err {e:.2e}")

    # root children = subdomains of the stochastic tree
    import sbh_capi
    nchild = np.empty(i2["num_nodes"], np.uint32)
    sbh_capi.check(sbh_capi.lib().sbh_tree_nodes(t2.h, None, None, None, nchild.ctypes.data,
                                                 *([None] * 10)))
    nsub = int(nchild[0])

    def at_target(rows):
        ok = [r for r in rows if r["mean_rel_err"] <= TARGET_ERR]
        if ok:
            return min(ok, key=lambda r: r["kernel_ms"]), True
        return min(rows, key=lambda r: r["mean_rel_err"]), False

    head, head_ok = at_target(results["stochastic"])
    bh_best, _ = at_target(results["barnes_hut"])

    log(f"headline: stochastic spl={head['samples_per_subdomain']} sort={head['sorted']}")
    hflags = C.EVAL_SORT_QUERIES if head["sorted"] else 0
    hkw = dict(samples_per_subdomain=head["samples_per_subdomain"], seed=1, flags=hflags)
    wall, kern = run(t2, C.STOCHASTIC, args.steps, args.warmup, **hkw)
    head_err = err_of(out)
    ms_per_step = float(np.mean(wall))
    value = nq / (ms_per_step / 1e3)

    # roofline of the estimator kernel: FP32 work of the terms it evaluates
    t2.eval_device(C.STOCHASTIC, "gravity", qd.data_ptr(), nq, out.data_ptr(),
                   counts_ptr=cnt.data_ptr(), flags=hflags | C.EVAL_SPLIT_COUNTS,
                   samples_per_subdomain=head["samples_per_subdomain"], seed=1)
    c = cnt.cpu().numpy().astype(np.int64)
    node_terms, point_terms = int(c[:nq].sum()), int(c[nq:].sum())
    # each term is one pair evaluation; the control variate adds one more
    # node term per accepted level (prev_result), not counted in num_contribs
    terms = node_terms + point_terms
    flops = terms * FLOPS_PER_PAIR_GRAVITY_3D
    kernel_ms = float(np.mean(kern))
    achieved = flops / (kernel_ms / 1e3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("stochastic_gravity_plummer_1m")
        except Exception:
            traffic = None
    roofline = {"bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
                "kernel": "k_stochastic<3,GRAVITY>", "terms_per_launch": terms,
                "node_terms": node_terms, "point_terms": point_terms,
                "flops_per_term": FLOPS_PER_PAIR_GRAVITY_3D, "kernel_ms": kernel_ms}

    serving = None
    if not args.no_serve:
        del t1, t2
        torch.cuda.empty_cache()
        serving = serving_bench(args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pts, m, lo, hi, q32, head["samples_per_subdomain"], args.cpu_seconds,
                           value)

    line = {
        "metric": METRIC, "value": value, "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "3D gravity, Plummer sphere N=2^20 sources (equal masses 1/N), "
                        "Q=2^20 Plummer queries, stochastic Barnes-Hut at <=1e-2 mean rel err"
                        + ("" if args.rmax is None else f", truncated at r<={args.rmax}"),
            "plummer_rmax": args.rmax,
            "n_sources": args.n, "n_queries": nq, "method": "stochastic",
            "samples_per_subdomain": head["samples_per_subdomain"], "subdomains": nsub,
            "samples_per_query": head["samples_per_subdomain"] * nsub,
            "sorted_queries": head["sorted"], "mean_rel_err": head_err,
            "err_queries": ne, "tree": "octree k=2, leaf size 1",
            "target_err_met": head_ok,
        },
        "latency_ms": ({"p50": serving["poisson"]["p50_ms"], "p99": serving["poisson"]["p99_ms"],
                        "kind": "serving request latency, submit to results in host memory, "
                                "Poisson arrivals at {:.0%} of burst capacity".format(
                                    args.serve_load)}
                       if serving else None),
        "step_latency_ms": {"p50": percentile(wall, 50), "p99": percentile(wall, 99),
                            "kind": "per 2^20-query evaluation, wall clock"},
        "serving": serving,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "speedup_vs_brute_force": value / results["brute_force"]["queries_per_s"],
        "barnes_hut_at_target": bh_best,
        "methods": results,
        "build_ms": build_ms,
        "trees": {"octree1": i1, "octree2": i2},
        "fp64_reference_ms": ref_ms,
    }
    print(json.dumps(line), flush=True)


SERVE_TREES = [  # name, dim, kernel, points, generator (as in the 1M-point cases)
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
        cls = sbh_gpu.OctreeGPU2 if dim == 3 else sbh_gpu.QuadtreeGPU2
        tr = cls(pts, nrm, m, lo, hi, alpha=10.0 if kern == "smoothdist" else 1.0)
        nsub = int(tr.nodes()["nchild"][0])
        trees.append(dict(name=name, dim=dim, kernel=kern, n=len(pts), lo=lo, hi=hi,
                          id=srv.add_tree(tr), tree=tr, subdomains=nsub))
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
            hs.append(srv.submit(t["id"], method, t["kernel"], pts, beta=2.0,
                                 samples_per_subdomain=spl, seed=seed))
        lats = [h.result(600.0)[1] for h in hs]
        el = time.perf_counter() - t0
        return np.asarray(lats), el

    # warm-up (first launches compile nothing, but touch every code path)
    for h in [srv.submit(t["id"], m_, t["kernel"], p[:2000], beta=2.0, samples_per_subdomain=s_,
                         seed=1) for (t, m_, p, s_, _) in reqs[:16]]:
        h.result(600.0)
    lat_b, el_b = run(np.zeros(nreq))
    cap = total_q / el_b
    log(f"serve burst: {nreq} requests, {total_q} queries in {el_b:.2f} s -> {cap / 1e6:.1f} M q/s")
    mean_q = total_q / nreq
    rate = args.serve_load * cap / mean_q
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    lat_p, el_p = run(arrivals)
    log(f"serve poisson: rate {rate:.1f} req/s, p50 {np.percentile(lat_p, 50):.1f} ms, "
        f"p99 {np.percentile(lat_p, 99):.1f} ms")
    st = srv.stats()
    return {
        "trees": [{k: t[k] for k in ("name", "dim", "kernel", "n", "subdomains")} for t in trees],
        "requests": nreq, "queries": int(total_q),
        "request_sizes": "log-uniform 1e3..1e6 queries",
        "methods": "barnes_hut theta=0.5 / stochastic 128 samples per query, 50/50",
        "burst": {"queries_per_s": cap, "elapsed_s": el_b,
                  "p50_ms": float(np.percentile(lat_b, 50)),
                  "p99_ms": float(np.percentile(lat_b, 99))},
        "poisson": {"rate_req_per_s": rate, "offered_load": args.serve_load,
                    "queries_per_s": total_q / el_p, "elapsed_s": el_p,
                    "p50_ms": float(np.percentile(lat_p, 50)),
                    "p99_ms": float(np.percentile(lat_p, 99))},
        "launches": st["launches"], "peak_request_bytes": st["peak_bytes"],
        "budget_bytes": st["budget_bytes"], "resident_tree_bytes": st["tree_bytes"],
    }


def cpu_baseline(pts, m, lo, hi, q32, spl, budget_s, gpu_qps):
    """The oracle (a C++ restatement of the reference's CPU estimator with the
    same threading model) timed on this machine's host cores on a bounded
    sample of the same workload."""
    import oracle

    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    ot = oracle.Tree(pts, None, m, lo, hi, split_log=2)
    build_s = time.perf_counter() - t0
    probe = q32[:2048]
    t0 = time.perf_counter()
    ot.stochastic("gravity", probe, spl, seed=1, nthreads=threads)
    per_q = (time.perf_counter() - t0) / len(probe)
    nsample = int(min(len(q32), max(2048, budget_s / max(per_q, 1e-9))))
    t0 = time.perf_counter()
    ot.stochastic("gravity", q32[:nsample], spl, seed=1, nthreads=threads)
    el = time.perf_counter() - t0
    qps = nsample / el
    log(f"cpu baseline: {qps:.0f} q/s on {threads} threads ({nsample} queries, {el:.1f} s); "
        f"oracle tree build {build_s:.1f} s")
    return {"value": qps, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"stochastic estimator, samples_per_subdomain={spl}, first {nsample} of the "
                      f"2^20 Plummer queries against the 2^20-source octree (k=2); tree build "
                      f"excluded ({build_s:.1f} s single-threaded)",
            "gpu_over_cpu": gpu_qps / qps}


if __name__ == "__main__":
    main()
