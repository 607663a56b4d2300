// Launchers for the evaluation kernels (eval.hip) and device utilities
// (tree_build.hip).  All pointers are device pointers; queries are SoA columns
// [x..., y..., (z...)] of length nq.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sbh_runtime.hpp"
#include "sbh_walk.hpp"

namespace sbh {

inline TreeView view_of(const DeviceTree& t) {
  return TreeView{t.nodeC.get(), t.nodeL.get(), t.nodeN.get(), t.dfsC.get(), t.dfsL.get(),
                  t.dfsN.get(), t.n_start.get(), t.n_end.get(), t.n_skip.get(), t.pts.get(),
                  t.nrm.get(), t.dfsR.get(), t.bfsR.get(), t.stepR.get(), t.kidC.get(),
                  t.kidN.get(), t.alpha,
                  t.alpha * 1.4426950408889634f,
                  t.dens.get(),
                  {t.lo[0], t.lo[1], t.lo[2]},
                  {t.dens_inv[0], t.dens_inv[1], t.dens_inv[2]},
                  t.dens_diag,
                  t.dens_g};
}

uint32_t brute_split_count(uint32_t n, uint32_t nq);
void run_brute(const DeviceTree& t, int kernel, const float* q, uint32_t nq, float* out,
               double* partial, uint32_t nsplit, hipStream_t s);
// Matrix-core squared distances (3D gravity / smooth distance) with fp32 or
// split bf16 operands; `order` is a spatial order of the queries (query_order).
uint32_t brute_mfma_split_count(uint32_t n, uint32_t nq);
void run_brute_mfma(const DeviceTree& t, int kernel, bool bf16, const float* q, uint32_t nq,
                    const uint32_t* order, float* out, double* partial, uint32_t nsplit,
                    hipStream_t s);
void run_brute_f64(const DeviceTree& t, int kernel, const float* q, uint32_t nq, double* out,
                   double* partial, uint32_t nsplit, hipStream_t s);
// counts: total terms per query, or with split != 0 node terms in counts[0, nq)
// and point terms in counts[nq, 2 nq).  exact: every term with correctly
// rounded fp32 operations in the reference's order (otherwise hardware
// sqrt / reciprocal / exp2).
// partial (optional): bh_split_count(...) * nq floats; launches too small to
// fill the GPU then split the per-query fast walk over a frontier of subtrees
// (bh_split: the task count and the frontier's task table).
struct WalkSplit {
  uint32_t n = 1;
  const WalkTask* tasks = nullptr;
};
WalkSplit bh_split(const DeviceTree& t, uint32_t nq, int group, bool exact);
uint32_t bh_split_count(const DeviceTree& t, uint32_t nq, int group, bool exact, bool counts);
void run_bh(const DeviceTree& t, int kernel, float beta, int group, const float* q, uint32_t nq,
            const uint32_t* order, float* out, int32_t* counts, int split, bool exact,
            float* partial, hipStream_t s);
// per_query: one sample path per query keyed by its index (otherwise the 64
// queries of a wavefront share their paths, keyed by the wavefront's index).
// partial (optional): stochastic_split_count(...) * nq * 2 floats (and with
// counts as many int32 again); the sorted-sample estimator then runs one task
// per 64-query block and range of root subdomains (DeviceTree::sub_tasks).
uint32_t stochastic_split_count(const DeviceTree& t, uint32_t nq, bool per_query, bool exact);
// work: counts (2 nq values) receive the work counts of the sorted-sample
// estimator (kernel terms evaluated, then path levels stepped); never split.
void run_stochastic(const DeviceTree& t, int kernel, int spl, uint64_t seed, bool per_query,
                    const float* q, uint32_t nq, const uint32_t* order, float* out,
                    int32_t* counts, int split, bool exact, float* partial, hipStream_t s,
                    bool work = false);

// p[0, n) <- v, on the stream (a kernel: no host copy, no memset node).
void fill_i32(int32_t* p, uint32_t n, int32_t v, hipStream_t s);

// Smooth distance from its sum, in place: w <- -log(w) / alpha.
void run_smooth_distance(float* w, uint32_t n, float alpha, hipStream_t s);

// Double-build trees (eval_double.hip): brute force (Kahan, tree order),
// Barnes-Hut (vote group 1, 32 or 64) and the stochastic estimator in float64.
void run_brute_double(const DeviceTree& t, int kernel, const double* q, uint32_t nq, double* out,
                      hipStream_t s);
// order (optional): query evaluated in each slot (query_order_double).
void run_bh_double(const DeviceTree& t, int kernel, double beta, int group, const double* q,
                   uint32_t nq, const uint32_t* order, double* out, int32_t* counts, int split,
                   hipStream_t s);
// memo_diff (optional): 64 * nq doubles of level-1 memo; memo_cnt (optional,
// with counts): 128 * nq int32.
void run_stochastic_double(const DeviceTree& t, int kernel, int spl, uint64_t seed,
                           bool per_query, const double* q, uint32_t nq, const uint32_t* order,
                           double* out, int32_t* counts, int split, double* memo_diff,
                           int32_t* memo_cnt, hipStream_t s);
// Spatial order of float64 queries (their float rounding's Hilbert order);
// qf holds nq * dim floats, scratch sort_scratch_u32(nq) values.
void query_order_double(const DeviceTree& t, const double* q, uint32_t nq, float* qf,
                        uint32_t* order, uint32_t* scratch, hipStream_t s);
void run_smooth_distance_double(double* w, uint32_t n, double alpha, hipStream_t s);

// Spatial (Hilbert-curve) order of the queries; scratch holds sort_scratch_u32(nq)
// uint32 values.  compact: smaller workgroups, for a sort that shares the GPU
// with other running launches (the same order).
size_t sort_scratch_u32(uint32_t nq);
void query_order(const DeviceTree& t, const float* q, uint32_t nq, uint32_t* order,
                 uint32_t* scratch, hipStream_t s, bool compact = false);

// Exclusive scan of n uint32 values; scratch holds exclusive_scan_scratch(n)
// values; *total_dev (optional) receives the sum.
size_t exclusive_scan_scratch(uint32_t n);
void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total_dev,
                    uint32_t* scratch, hipStream_t s);

}  // namespace sbh
