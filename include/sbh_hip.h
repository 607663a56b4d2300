/*
 * sbh_hip.h - C ABI of the MI355X (gfx950) Stochastic Barnes-Hut library,
 * libsbh_hip.so.
 *
 * This is the boundary under the pybind11 plugin module `sbh_gpu`
 * (src/sbh_gpu.cu:566-661 in the reference) and under any other host
 * language: plain pointers and sizes, integer status codes, no exceptions
 * across the boundary, no framework types.
 *
 * Conventions
 *  - Every function returns SBH_OK (0) or an SBH_ERR_* code; the message of
 *    the last failure on the calling thread is sbh_last_error().
 *  - Host pointers are borrowed for the duration of the call only.  With
 *    SBH_EVAL_DEVICE_POINTERS the query / result / count pointers are device
 *    pointers instead and the call neither copies nor allocates them.
 *  - A tree owns its device memory and one HIP stream; calls on one tree are
 *    serialised by the caller (thread-compatible, not thread-safe).
 *  - Queries are a column-major d x q matrix: x[0..q), y[0..q), z[0..q)
 *    (the reference's Eigen column-major query copy, src/sbh_gpu.cu:217).
 */
#ifndef SBH_HIP_H
#define SBH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sbh_tree sbh_tree;

/* Status codes. */
#define SBH_OK 0
#define SBH_ERR_INVALID 1   /* bad argument (shape, NaN, empty, unsupported mode) */
#define SBH_ERR_HIP 2       /* HIP runtime failure */
#define SBH_ERR_NOMEM 3     /* device allocation failed */
#define SBH_ERR_NODEVICE 4  /* no usable GPU */
#define SBH_ERR_BUDGET 5    /* serving admission refused: HBM budget exceeded */
#define SBH_ERR_INTERNAL 6

/* Kernel functions: include/kernel.h:24-28 (KernelFuncID). */
#define SBH_GRAVITY 0
#define SBH_WINDING_NUMBER 1
#define SBH_SMOOTH_DIST 2
/* Kernels the reference defines but binds to no plugin (include/kernel.h:61-83):
 * helmholtzKernelX / Y, and gravityGrad as one scalar kernel per component
 * (-m / (d + 0.1)^3 * (q - p)_axis; Z in 3D only).  Evaluated by every method;
 * not by the serving engine or the matrix-core brute force. */
#define SBH_HELMHOLTZ_X 3
#define SBH_HELMHOLTZ_Y 4
#define SBH_GRAVITY_GRAD_X 5
#define SBH_GRAVITY_GRAD_Y 6
#define SBH_GRAVITY_GRAD_Z 7

/* Evaluation methods. */
#define SBH_BRUTE_FORCE 0 /* src/sbh_gpu.cu:212-322  evalBruteForce_*            */
#define SBH_BARNES_HUT 1  /* src/sbh_gpu.cu:324-434  evalBarnesHut_*             */
#define SBH_STOCHASTIC 2  /* src/sbh_gpu.cu:436-550  multiLevelPrefixControlVariateSampler_* */

/* sbh_tree_create flags. */
#define SBH_TREE_MOMENTS_F64 1 /* node moments aggregated in float64 (default: the
                                  reference's fp32 order, include/tree3d.h:255-327) */

/* sbh_eval flags. */
#define SBH_EVAL_DEVICE_POINTERS 1 /* queries / results / counts are device pointers */
#define SBH_EVAL_SORT_QUERIES 2    /* evaluate in spatial (Hilbert-curve) order; results are
                                      returned in input order.  Barnes-Hut with
                                      vote_group 1 gives identical values either way */
#define SBH_EVAL_SPLIT_COUNTS 4    /* counts holds 2q values: node terms in [0, q),
                                      point terms in [q, 2q) */
#define SBH_EVAL_EXACT_TERMS 8     /* Barnes-Hut / stochastic: every term, and the
                                      sampler's opening ratios, roulette probabilities
                                      and weights, with correctly rounded fp32
                                      operations in the reference CPU plugin's order
                                      (bit-reproducible against it); default is the
                                      hardware sqrt / reciprocal / exp2 the reference's
                                      fast-math GPU build uses (CMakeLists.txt:27).
                                      Barnes-Hut opening tests are exact in both modes */
#define SBH_EVAL_MFMA 16           /* brute force, 3D gravity / smooth distance: squared
                                      distances |q'|^2 + |p'|^2 - 2 q'.p' on the matrix
                                      cores (V_MFMA_F32_16X16X4_F32) in coordinates
                                      centred per 64-query wavefront, near-field pairs
                                      recomputed directly */
#define SBH_EVAL_MFMA_BF16 64      /* brute force, 3D gravity / smooth distance: the same
                                      expansion with bf16 operands, every coordinate
                                      split into bf16 hi + lo parts and the products
                                      h.h + h.l + l.h + l.l accumulated in fp32
                                      (V_MFMA_F32_16X16X16_BF16); pairs closer than 4x
                                      the wavefront radius recomputed directly */
#define SBH_EVAL_SMOOTH_DISTANCE 128 /* smooth distance: results are -log(w) / alpha,
                                      the distance the reference's driver derives
                                      from the sum w (python/example.py:79); a
                                      non-positive stochastic estimate gives NaN or
                                      +inf as the driver's numpy transform would */
#define SBH_EVAL_PER_QUERY_PATHS 32 /* stochastic: one sample path per query keyed by the
                                      query's index, as the reference CPU sampler draws
                                      them (include/tree_sampler.h:48-104); default: the
                                      64 queries of a wavefront share their paths, as the
                                      reference GPU sampler shares them per block
                                      (include/tree_sampler_gpu.h:41-71).  Per-query
                                      results do not depend on the query order */
#define SBH_EVAL_NO_SPLIT 256      /* Barnes-Hut (vote_group 1) / stochastic (shared
                                      paths), fast terms: run the whole-walk kernels
                                      at any launch size.  By default Barnes-Hut
                                      launches of fewer than 2^18 queries split each
                                      walk into tasks over up to 256 disjoint subtrees
                                      that cover the tree (a balanced frontier chosen
                                      by the launch size), and estimator launches on
                                      trees with two or more internal root subdomains
                                      split into ranges of subdomains at any size;
                                      partial sums are added in depth-first task
                                      order, which visits the same terms but rounds
                                      the sum differently.  With this flag the values
                                      are those of the whole walks */
#define SBH_EVAL_WORK_COUNTS 512   /* stochastic, fast shared-path estimator on trees of
                                      at most 64 children per node, with counts: counts
                                      holds 2q values describing the work the launch
                                      performed for each query after its reuse of
                                      node differences across sorted samples: kernel
                                      terms evaluated in [0, q), sample-path levels
                                      stepped in [q, 2q).  Values are unchanged; the
                                      launch is never split */

typedef struct sbh_eval_params {
  float beta;                    /* Barnes-Hut opening ratio (1/theta); +inf = exact */
  int32_t samples_per_subdomain; /* stochastic: samples per root child (>= 1) */
  uint64_t seed;                 /* stochastic: Philox key */
  int32_t vote_group;            /* Barnes-Hut: 1 = per query (src/sbh.cpp walk),
                                    32 = 32-query warp vote (include/barnes_hut_gpu.h:34),
                                    64 = wavefront vote */
  int32_t flags;                 /* SBH_EVAL_* */
} sbh_eval_params;

/* Message of the last failure on this thread ("" if none). */
const char* sbh_last_error(void);

/* Library ABI version (major * 100 + minor). */
int32_t sbh_version(void);

/* Number of visible GPUs. */
int sbh_device_count(int32_t* count);

/*
 * Build a tree on the current GPU.  Replaces the CPU build plus device
 * mirror of the reference: PyOctree / PyQuadtree (include/py_tree.h:30-74)
 * followed by PyOctreeGPU / PyQuadtreeGPU (include/py_tree_gpu.h:39-80).
 *   dim         2 or 3
 *   split_log   k in 1..3: 2^k splits per axis (the class suffix of
 *               OctreeK / QuadtreeK, src/sbh.cpp:171-177)
 *   leaf_size   LEAF_SIZE template argument (the reference binds 1)
 *   max_depth   100 in the reference (include/py_tree.h:44,72)
 *   points      n x dim row-major; normals n x dim row-major or NULL (zeros)
 *   masses      n; finite and > 0
 *   min_corner / max_corner  dim floats; the root cell
 *   alpha       smooth-distance exponent (include/py_tree.h:39,67)
 *   build_ms    optional: wall time of the build
 */
int sbh_tree_create(int32_t dim, int32_t split_log, int32_t leaf_size, int32_t max_depth,
                    int64_t n, const float* points, const float* normals, const float* masses,
                    const float* min_corner, const float* max_corner, float alpha, int32_t flags,
                    sbh_tree** out, float* build_ms);

int sbh_tree_destroy(sbh_tree* tree);

/* Sizes: points, nodes, deepest level, device bytes held.  Any output may be NULL. */
int sbh_tree_info(const sbh_tree* tree, int64_t* num_points, int64_t* num_nodes, int32_t* depth,
                  int64_t* device_bytes);

/* Leaf order: out[i] = input row of the i-th point in tree order (host, n values).
 * The reference exposes this implicitly by permuting leaf_data in place
 * (include/tree3d.h:281-289). */
int sbh_tree_leaf_order(const sbh_tree* tree, int64_t* out);

/*
 * Node arrays in breadth-first node order (include/tree3d.h:212-225 fields).
 * Host outputs, each NULL or sized num_nodes (x dim for vectors, x 8 for
 * moments64 = {cx, cy, cz, mass, nx, ny, nz, 0} in float64).  first_child is
 * 0xFFFFFFFF for leaves; rope is the depth-first successor used by the walk.
 */
int sbh_tree_nodes(const sbh_tree* tree, uint32_t* start, uint32_t* end, uint32_t* first_child,
                   uint32_t* num_children, int32_t* depth, uint32_t* parent, uint32_t* rope,
                   float* min_corner, float* max_corner, float* diameter, float* centroid,
                   float* mass, float* normal, double* moments64);

/*
 * Evaluate q queries.  results: q floats; counts (optional): kernel
 * evaluations per query (num_contribs of include/barnes_hut.h:29-50 and
 * include/tree_sampler.h:110-171); kernel_ms (optional): device time of the
 * evaluation kernels, hipEvent-timed on the tree's stream as the reference
 * times its launches (src/sbh_gpu.cu:227-241).
 */
int sbh_eval(sbh_tree* tree, int32_t method, int32_t kernel, const sbh_eval_params* params,
             const float* queries, int64_t q, float* results, int32_t* counts, float* kernel_ms);

/*
 * sbh_eval for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments, kernels and results as
 * sbh_eval (bit-identical), but params->flags must include
 * SBH_EVAL_DEVICE_POINTERS and the call makes no device allocation: the
 * scratch of a sorted launch (launch order and sort work space, 26 bytes a
 * query) and of split launches (partial sums) comes from one static array of
 * SBH_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  An
 * evaluation that needs more fails with SBH_ERR_INVALID and a message giving
 * the size (e.g. sorted launches above about 10M queries, or a sorted
 * estimator launch split over 12 subdomain ranges above about 2M;
 * SBH_EVAL_NO_SPLIT removes the partial sums).  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (two trees on two threads or streams) race
 * and must be serialised by the caller.
 */
#define SBH_NOALLOC_SCRATCH_BYTES (256ull << 20)
int sbh_eval_noalloc(sbh_tree* tree, int32_t method, int32_t kernel,
                     const sbh_eval_params* params, const float* queries, int64_t q,
                     float* results, int32_t* counts, float* kernel_ms);

/* float64 all-pairs sums of the fp32 kernel formulas (error reference). */
int sbh_eval_f64(sbh_tree* tree, int32_t kernel, const float* queries, int64_t q,
                 double* results, int32_t flags, float* kernel_ms);

/* The permutation SBH_EVAL_SORT_QUERIES applies: order[i] = query evaluated in
 * slot i (host or device pointers per flags). */
int sbh_query_order(sbh_tree* tree, const float* queries, int64_t q, uint32_t* order,
                    int32_t flags);

/* Stream the tree's work is issued on (hipStream_t). */
int sbh_tree_stream(sbh_tree* tree, void** stream);

/*
 * Double build.  The reference's USE_DOUBLE configuration (include/utils.h:
 * 13-29: FLOAT = double for every point, bound, moment, term and result,
 * literals FP(x) = x, PI = M_PI), selected per tree instead of at compile
 * time.  The tree is built from float64 inputs with the same partition,
 * breadth-first numbering and leaf order rules in float64; evaluations take
 * float64 queries and return float64 results, each term and sum with one
 * rounding per operation in the reference's order.  Stochastic paths use the
 * same Philox streams as the float build; the roulette uniform is the whole
 * 32-bit word, ((w + 1) * 2^-32).  A double-build tree is evaluated by
 * sbh_eval_double only (sbh_eval, sbh_eval_f64, sbh_query_order, tree images
 * and the serving engine take float-build trees and refuse it).
 */
typedef struct sbh_eval_params_double {
  double beta;                   /* Barnes-Hut opening ratio (1/theta); +inf = exact */
  int32_t samples_per_subdomain; /* stochastic: samples per root child (>= 1) */
  uint64_t seed;                 /* stochastic: Philox key */
  int32_t vote_group;            /* Barnes-Hut: 1, 32 or 64 (as sbh_eval_params) */
  int32_t flags;                 /* SBH_EVAL_DEVICE_POINTERS, _SORT_QUERIES (the
                                    Hilbert order of the queries' float rounding),
                                    _SPLIT_COUNTS, _PER_QUERY_PATHS, _SMOOTH_DISTANCE;
                                    _EXACT_TERMS is accepted and changes nothing */
} sbh_eval_params_double;
/* PyOctree / PyQuadtree + PyOctreeGPU / PyQuadtreeGPU of a USE_DOUBLE build
 * (include/py_tree.h:30-74, include/py_tree_gpu.h:39-80); arguments as
 * sbh_tree_create, in float64. */
int sbh_tree_create_double(int32_t dim, int32_t split_log, int32_t leaf_size,
                           int32_t max_depth, int64_t n, const double* points,
                           const double* normals, const double* masses,
                           const double* min_corner, const double* max_corner, double alpha,
                           sbh_tree** out, float* build_ms);
/* *is_double = 1 for a tree made by sbh_tree_create_double, else 0. */
int sbh_tree_is_double(const sbh_tree* tree, int32_t* is_double);
/* A double-build tree's float64 node arrays (breadth-first order; NULL or
 * sized num_nodes, x dim for vectors).  The index arrays come from
 * sbh_tree_nodes, whose float outputs must then be NULL (moments64 holds the
 * same float64 moments). */
int sbh_tree_nodes_double(const sbh_tree* tree, double* min_corner, double* max_corner,
                          double* diameter, double* centroid, double* mass, double* normal);
/* evalBruteForce_* / evalBarnesHut_* / multiLevelPrefixControlVariateSampler_*
 * (src/sbh_gpu.cu:212-550) of a USE_DOUBLE build: results q doubles; counts
 * and kernel_ms as sbh_eval.  Brute force is the reference's Kahan sum in
 * tree order. */
int sbh_eval_double(sbh_tree* tree, int32_t method, int32_t kernel,
                    const sbh_eval_params_double* params, const double* queries, int64_t q,
                    double* results, int32_t* counts, float* kernel_ms);
/*
 * Tree images.  No counterpart in the reference (its trees exist only inside
 * the Python objects that built them, include/py_tree_gpu.h:23-81): a built
 * tree is written to a byte image (format version 2: header, every node and
 * point array, FNV-1a checksum) and loaded back resident on the current GPU
 * without rebuilding, e.g. for restarts of a serving process.  A loaded tree
 * evaluates exactly as the tree that was saved.  Loading checks the checksum,
 * the sizes and every index the walks follow before any device memory is
 * written, and fails with SBH_ERR_INVALID on a damaged image.
 *   sbh_tree_serialize: with buffer == NULL (or capacity too small) only
 *   *size is set (and SBH_OK returned for NULL, SBH_ERR_INVALID otherwise).
 */
int sbh_tree_serialize(const sbh_tree* tree, void* buffer, int64_t capacity, int64_t* size);
/* Shape of a tree (e.g. one just loaded): dim, split_log, leaf_size, alpha. */
int sbh_tree_layout(const sbh_tree* tree, int32_t* dim, int32_t* split_log, int32_t* leaf_size,
                    float* alpha);
int sbh_tree_deserialize(const void* buffer, int64_t size, sbh_tree** out);
int sbh_tree_save(const sbh_tree* tree, const char* path);
int sbh_tree_load(const char* path, sbh_tree** out);

/*
 * Serving.  No counterpart in the reference (it re-copies queries per call
 * and runs synchronously, src/sbh_gpu.cu:212-248): resident trees, requests
 * queued and admitted in arrival order while their working sets (queries,
 * results, their Hilbert-curve order and its sort scratch: about
 * (dim + 7) * 4 bytes per query) fit in the server's device budget, each
 * admitted batch run as one work-queue launch over the requests' sorted
 * queries, up to max_inflight launches overlapping on separate streams.
 * Barnes-Hut runs per query (vote_group is ignored), the stochastic estimator
 * keys paths on (seed, 64-query block of the request's sorted order): a
 * request's values are identical alone and in any mix, and bit-identical to
 * those of sbh_eval with SBH_EVAL_SORT_QUERIES plus the request's flags: a
 * request has its walks split into tasks as a one-shot launch of its size
 * does, unless params.flags holds SBH_EVAL_NO_SPLIT.  Terms use the fast kernels; params.flags may hold only
 * SBH_EVAL_SORT_QUERIES and SBH_EVAL_NO_SPLIT, anything else is refused with
 * SBH_ERR_INVALID.
 *
 * Trees must outlive the server.  `queries` (column-major dim x q) and
 * `results` (q floats) are host memory that must stay valid until the request
 * is done, failed or cancelled.  sbh_server_destroy cancels queued requests,
 * waits for running ones, and wakes threads blocked in sbh_server_wait (which
 * then return the request's state); no call on the server may start once
 * sbh_server_destroy has been entered.
 */
typedef struct sbh_server sbh_server;

#define SBH_REQ_QUEUED 0
#define SBH_REQ_RUNNING 1
#define SBH_REQ_DONE 2
#define SBH_REQ_FAILED 3
#define SBH_REQ_CANCELLED 4

typedef struct sbh_request {
  int32_t tree_id;
  int32_t method; /* SBH_BARNES_HUT or SBH_STOCHASTIC */
  int32_t kernel;
  sbh_eval_params params;
  const float* queries;
  int64_t q;
  float* results;
} sbh_request;

typedef struct sbh_server_stats {
  int64_t completed, launches, queued;
  int64_t bytes_in_use, peak_bytes, budget_bytes, tree_bytes;
} sbh_server_stats;

int sbh_server_create(int64_t budget_bytes, int32_t max_inflight, sbh_server** out);
int sbh_server_destroy(sbh_server* server); /* waits for in-flight launches */
int sbh_server_add_tree(sbh_server* server, sbh_tree* tree, int32_t* tree_id);
/* SBH_ERR_BUDGET if the request alone exceeds the budget. */
int sbh_server_submit(sbh_server* server, const sbh_request* request, int64_t* request_id);
/* state: SBH_REQ_*; latency_ms (optional): submit -> results in place, once done. */
int sbh_server_poll(sbh_server* server, int64_t request_id, int32_t* state, double* latency_ms);
/* timeout_s < 0 waits without limit. */
int sbh_server_wait(sbh_server* server, int64_t request_id, double timeout_s, int32_t* state,
                    double* latency_ms);
/* Cancels a request that is still queued; *cancelled = 0 if it already started. */
int sbh_server_cancel(sbh_server* server, int64_t request_id, int32_t* cancelled);
/* Drops the bookkeeping of a finished request. */
int sbh_server_release(sbh_server* server, int64_t request_id);
int sbh_server_stats_get(sbh_server* server, sbh_server_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* SBH_HIP_H */
