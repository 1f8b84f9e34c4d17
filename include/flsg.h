/*
 * flsg.h -- C ABI of libflsg.so, the MI355X-native local graph-diffusion push
 * solvers.
 *
 * The reference (JiaheBai/Faster-Local-Solver-for-GDEs) has no C/FFI layer:
 * its boundary is a set of module-level Python functions (numba @njit /
 * CuPy).  Each entry point below names the reference function whose
 * computation it replaces; the Python drop-in modules (ppr_solver.py,
 * katz_solver.py, hk_solver.py, InstantGNN.py) keep the reference's Python
 * signatures and call these through ctypes.
 *
 * Conventions
 *   - every function returns int status: FLSG_OK (0) or an FLSG_E* code;
 *     flsg_last_error() gives a thread-local message for the last failure;
 *   - buffers named d_* are device pointers (e.g. torch ROCm tensors); h_*
 *     are host pointers;  no torch or HIP types appear in the signatures --
 *     streams are passed as void* (a hipStream_t, NULL = default stream);
 *   - device CSR: int32 indptr[n+1], int32 indices[nnz_cap], int32 degree[n];
 *     the live adjacency of u is indices[indptr[u] : indptr[u]+degree[u]]
 *     (for a plain CSR degree[u] == indptr[u+1]-indptr[u]; for the slack CSR
 *     of the dynamic workloads it is the live row length);
 *   - a graph handle borrows the caller's CSR buffers (they must outlive it)
 *     and owns the solver workspaces it allocates; calls on one handle are
 *     not thread-safe and are serialised on the stream passed in.
 */
#ifndef FLSG_H
#define FLSG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLSG_OK 0
#define FLSG_EARG 1      /* invalid argument (message says which)            */
#define FLSG_EASSERT 2   /* the reference's `assert eps*degree[s] <= 1.`     */
#define FLSG_EHIP 3      /* a HIP runtime call failed                         */
#define FLSG_ENOMEM 4    /* device allocation failed / admission refused      */
#define FLSG_ECAP 5      /* an output or trace capacity was too small         */
#define FLSG_ENOCONV 6   /* iteration cap reached without meeting eps         */

/* problem kinds */
#define FLSG_PPR 0       /* Q = I - (1-alpha) A D^-1, b = alpha e_s           */
#define FLSG_KATZ 1      /* (I - alpha A) x = e_s, returns x - e_s for GD     */
#define FLSG_HK 2        /* heat kernel, level-synchronous Taylor push        */

/* methods */
#define FLSG_GD 0        /* (local) gradient descent = Jacobi push            */
#define FLSG_CHEB 1      /* (local) Chebyshev                                 */

/* precision */
#define FLSG_FP64 0      /* plain FP64 path (parity path)                     */
#define FLSG_FP32 1      /* FP32 r/x, FP64 threshold test + FP64 certificate  */

typedef struct flsg_graph_s *flsg_graph;

/* Solver parameters (one query or one batch). */
typedef struct {
    int32_t kind;        /* FLSG_PPR / FLSG_KATZ / FLSG_HK                   */
    int32_t method;      /* FLSG_GD / FLSG_CHEB                               */
    int32_t local;       /* 1 = local (frontier) solver, 0 = global sweeps    */
    int32_t precision;   /* FLSG_FP64 / FLSG_FP32                             */
    double alpha;        /* PPR teleport / Katz beta                          */
    double eps;          /* threshold                                         */
    double l_max;        /* Katz Chebyshev eigen-bounds of I - alpha A        */
    double mu_min;
    double tau;          /* heat-kernel t                                     */
    int64_t n_terms;     /* heat-kernel Taylor order N (<=0: estimate_n)      */
    int64_t max_iters;   /* iteration cap (<=0: 1e6)                          */
} flsg_params;

/* Per-query result record. */
typedef struct {
    int64_t oper;        /* the reference's operation count                   */
    int64_t iters;       /* frontier iterations / global sweeps / HK levels   */
    int64_t oper_cert;   /* batched FP32: support edges of the FP64 residual
                            recompute + edges of the FP64 correction push     */
    int64_t n_corr;      /* nodes pushed by the FP64 correction               */
    int64_t supp;        /* |supp x|                                          */
    int64_t nodes;       /* sum over iterations of the frontier size          */
    double cert;         /* max_v |r_v| / threshold_v after the solve (< 1)   */
    double ms;           /* device time of the solve, milliseconds            */
    double ms_push;      /* batched: time of the FP32 push (before the FP64 certificate) */
    int32_t status;      /* FLSG_OK or error code for this query              */
    int32_t touched;     /* batched: records the query touched (and reset)    */
    double ms_recompute; /* batched FP32: time from the start to the end of the FP64
                            residual recompute                                */
    double ms_correct;   /* batched FP32: time from the start to the end of the FP64
                            correction push (then: output and record reset)   */
    double r_l1;         /* batched: sum over the touched records of |r_v| plus
                            its FP32 error bound -- the L1 norm of the certified
                            residual, so ||x - x*||_1 <= r_l1 / alpha for PPR
                            (Katz: r_l1 / (1 - beta d_max))                  */
    double ms_node;      /* batched push, summed over its iterations: node phases  */
    double ms_pairs;     /* binned iterations: count + write passes of the pairs   */
    double ms_bins;      /* binned iterations: the ranges' sums and adds            */
    double ms_direct;    /* directly pushed iterations: the edge phases            */
    int64_t bin_iters;   /* iterations pushed binned                               */
    int64_t corr_edges;  /* batched FP32: edges of the FP64 correction push (part of oper_cert) */
} flsg_result;

/* Optional per-iteration frontier trace (device buffers): iteration t
 * processed d_nodes[h_off[t] : h_off[t+1]] (the order within an iteration is
 * the GPU's, sets are what parity compares). */
typedef struct {
    int32_t *d_nodes;    /* device, capacity cap_nodes                        */
    int64_t cap_nodes;
    int64_t *h_off;      /* host, capacity cap_iters + 1                      */
    int64_t cap_iters;
    int64_t n_iters;     /* out */
    int32_t overflow;    /* out */
    int32_t pad;
} flsg_trace;

int flsg_version(void);
const char *flsg_last_error(void);

/* Device properties the host side needs (CU count, free memory). */
int flsg_device_info(int32_t *num_cus, int64_t *free_bytes, int64_t *total_bytes);

/* Graph handle over a caller-owned device CSR.  `slack` = 1 marks a slack
 * CSR (live row length in degree[]).  Replaces the whole-graph host->device
 * copy the reference does inside every GPU query (ppr_solver.py:103-105,
 * katz_solver.py:102-104): the graph is uploaded once and stays resident. */
int flsg_graph_create(int64_t n, int64_t nnz_cap, const int32_t *d_indptr, int32_t *d_indices,
                      int32_t *d_degree, int32_t slack, flsg_graph *out);
int flsg_graph_destroy(flsg_graph g);

/* Single-source solve on the whole GPU.  Replaces, by params:
 *   PPR  GD  local : ppr_gd(..., local=True)   ppr_solver.py:208-288,321-328
 *                    and ppr_gd_local_gpu        ppr_solver.py:82-146
 *   PPR  GD  global: ppr_gd(..., local=False)  ppr_solver.py:290-320
 *                    and ppr_gd_standard_gpu     ppr_solver.py:26-79
 *   PPR  CHEB      : ppr_cheby                  ppr_solver.py:582-661
 *   KATZ GD        : katz_gd / katz_gd_*_gpu    katz_solver.py:24-248
 *   KATZ CHEB      : katz_cheby                 katz_solver.py:340-431
 *   HK             : hk_sor_queue(omega=1)      hk_solver.py:206-295
 *                    (local = 0: its natural-order level sweep, :265-279)
 * d_x, d_r: dense FP64 outputs of n entries (HK: d_r holds 2n, the two level
 * residuals as the reference returns them).  trace may be NULL. */
int flsg_solve(flsg_graph g, const flsg_params *p, int64_t seed, double *d_x, double *d_r,
               flsg_result *h_res, flsg_trace *trace, void *stream);

/* Heat kernel with the reference's extra outputs: the Taylor order N
 * (hk_solver.py:225) in *est_n, and per level j the sums
 * h_lvl_sums[2j] = sum |r_j,u| pushed, h_lvl_sums[2j+1] = residual mass sent to
 * level j+1 (the reference's beta_t diagnostic, hk_solver.py:250,286-290).
 * h_lvl_sums (host) may be NULL; when it is not, *est_n on entry holds its
 * capacity in doubles, and a capacity below 2N fails with FLSG_ECAP.
 * Replaces hk_sor_queue with omega=1 (hk_solver.py:206-295): params local = 1
 * the queue push (:236-264), local = 0 the level sweep over all nodes
 * (:265-279; one sweep per level, oper = N (m + n), level sums zero: the
 * reference's beta_t stays [1.0] there). */
int flsg_hk_solve(flsg_graph g, const flsg_params *p, int64_t seed, double *d_x, double *d_r,
                  flsg_result *h_res, double *h_lvl_sums, int64_t *est_n, flsg_trace *trace, void *stream);

/* The reference's scalar helpers of the heat kernel, host code:
 * estimate_coefficient(t, k) = t^(k+1) / (k+1)!  (hk_solver.py:15-25) and the
 * Taylor order estimate_n(t, eps) (hk_solver.py:28-48; its loop factor uses
 * the initial k).  Arguments are not checked (the drop-in asserts t >= 1,
 * eps > 0, k >= 0 as the reference does). */
double flsg_hk_estimate_coefficient(double t, int64_t k);
int64_t flsg_hk_estimate_n(double t, double eps);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The local GD push of flsg_solve (PPR or Katz, FP64,
 * params local = 1, method = FLSG_GD), with the same results, run directly on
 * the caller's d_x / d_r (n doubles each, zeroed here by a fill kernel) and
 * with every other buffer in a fixed device array the library holds per GPU
 * until the process ends (flsg_solve_local_noalloc_scratch_bytes).  Calls no
 * hipMalloc / hipFree.  Graphs beyond 2^20 nodes or 2^24 column entries are
 * rejected with FLSG_EARG; the graph's device must be current.  Every call on
 * one GPU shares that array: calls that overlap on one GPU race.  h_res gets
 * oper, iters, nodes and ms (host wall time). */
int flsg_solve_local_noalloc(flsg_graph g, const flsg_params *p, int64_t seed, double *d_x, double *d_r,
                             flsg_result *h_res, void *stream);
int64_t flsg_solve_local_noalloc_scratch_bytes(void);

/* Batched solve: many seeds in one persistent launch (one workgroup owns a
 * query at a time, queries are claimed dynamically).  Replaces the
 * reference's per-seed fan-out solve_a_graph_all / solve_a_graph_gpu
 * (ppr_solver.py:149-164,511-579; katz_solver.py:149-164,605-666).
 * Outputs are sparse: query q's estimate is
 *   d_out_ids/d_out_vals[h_res[q].. offsets ..]  -- d_out_off[q] (device,
 *   int64) is the start and h_res[q].supp the length.
 * out_cap bounds the total number of (id, value) pairs.  PPR/KATZ GD only,
 * p->precision FLSG_FP32 (FP32 push + FP64 certificate) or FLSG_FP64.  When
 * some queries fail (bad seed or parameters: FLSG_EARG; output arena full:
 * FLSG_ECAP; iteration cap: FLSG_ENOCONV) h_res is still filled for every
 * query and the call returns the first failing query's status. */
int flsg_batch_solve(flsg_graph g, const flsg_params *p, const int64_t *d_seeds, int64_t n_queries,
                     int32_t *d_out_ids, double *d_out_vals, int64_t *d_out_off, int64_t out_cap,
                     flsg_result *h_res, void *stream);

/* Size of the per-query device state a batched launch keeps resident, in
 * bytes per workgroup slot (the unit of memory-footprint admission): the
 * per-node arrays and lists of an n-node graph; the binned push adds 16
 * bytes per column entry of the graph to every slot. */
int64_t flsg_batch_slot_bytes(int64_t n);

/* Number of resident workgroup slots the handle uses for batched solves
 * (0 = choose from CU count and free memory) and the workgroup size
 * (256, 512 or 1024).  No counterpart in the reference (tuning knobs). */
int flsg_batch_set_slots(flsg_graph g, int32_t slots);
int flsg_batch_set_block(flsg_graph g, int32_t block);

/* Dynamic LDS bytes per workgroup of the batched kernel.  In iterations
 * pushed directly they hold the residuals of nodes [0, bytes/4) (FP32 push)
 * or [0, bytes/8) (FP64 push): on degree-relabelled graphs these are the
 * hubs, whose pushes then use LDS atomics; in binned iterations they hold
 * the slice of 2^s residuals of one destination range (2^s the largest power
 * of two that fits; binning needs 2^s >= 2048); in the certificate, x32 of
 * the first ids.  -1 = default for the workgroup size (136 KiB at 1024,
 * 48 KiB at 512, 24 KiB at 256), 0 = off, else a multiple of 8 up to the
 * device's LDS per workgroup less the kernel's static LDS (~145,000 bytes on
 * gfx950).  No counterpart in the reference (tuning knob). */
int flsg_batch_set_hub_cache(flsg_graph g, int32_t bytes);

/* Binned push of the batched kernel: an iteration of at least min_edges
 * edges (0 = never) writes its pushes as (destination, value) pairs into one
 * bin per destination range and sums each bin in LDS before one add per
 * destination into its residual (in id order over the range when the bin is
 * dense); a bin with fewer than dense_pairs pairs is applied with memory-side
 * atomics instead.  Defaults 65536 and 64.  No counterpart in the reference
 * (tuning knob, set by the autotuner). */
int flsg_batch_set_binning(flsg_graph g, int64_t min_edges, int32_t dense_pairs);

/* Load rule of the binned push: in flsg_batch_solve and serving sessions an
 * iteration is binned only while at least min_busy workgroups of the launch
 * run a query (the binned push pays off when many slots push at once --
 * memory-side atomics are then the chip-wide limit -- and the direct push is
 * quicker when few do).  0 = the default, a quarter of the device's compute
 * units (64 on MI355X); 1 bins whenever an iteration is large enough.
 * flsg_batch_launch does not count its workgroups and bins by size alone.
 * No counterpart in the reference (tuning knob). */
int flsg_batch_set_bin_load(flsg_graph g, int32_t min_busy);

/* Degree bins of the batched edge phases: a list entry with more than
 * split_edges edges (0 = 256) is processed as pieces of piece_edges edges
 * (0 = 1024; raised to keep at most 255 pieces per entry), each piece by one
 * wavefront; smaller entries are flattened 64 at a time by a wavefront.  No
 * counterpart in the reference (tuning knob, set by the autotuner). */
int flsg_batch_set_degree_bins(flsg_graph g, int32_t split_edges, int32_t piece_edges);

/* Allocate the slot arena (slots <= 0: the set / default count). */
int flsg_batch_reserve(flsg_graph g, int32_t slots, void *stream);

/* Internal-to-caller node id map applied to batched output ids (device,
 * n entries; NULL = identity) -- used when the graph was relabelled (e.g. by
 * degree) before upload.  The map must outlive the handle's use of it. */
int flsg_graph_set_out_map(flsg_graph g, const int32_t *d_out_map);

/* One batched query for flsg_batch_launch. */
typedef struct {
    int64_t seed;
    int32_t kind;        /* FLSG_PPR / FLSG_KATZ (local GD) / FLSG_HK         */
    int32_t precision;   /* PPR / Katz: FLSG_FP32 (FP32 push, FP64 threshold
                            tests, FP64 recompute + correction push) or FLSG_FP64
                            (FP64 push); HK levels are FP32 either way       */
    double alpha;        /* PPR alpha / Katz beta (unused for HK)             */
    double eps;          /* unused for HK (the launch's flsg_hk_config)        */
} flsg_query;

/* Heat-kernel parameters shared by the HK queries of one batched launch
 * (FP32 levels, FP64 thresholds; no certificate -- the level push drops
 * sub-threshold residual by design, hk_solver.py:282-285). */
typedef struct {
    double tau;
    double eps;
    int64_t n_terms;     /* <= 0: estimate_n(tau, eps/2); at most 96          */
} flsg_hk_config;

/* Asynchronous batched launch over slots [slot_base, slot_base + n_slots) of
 * the reserved arena, over a query list in device memory: queries may mix
 * PPR, Katz and (with hk non-NULL) heat-kernel queries, precisions and
 * parameters.  d_ctl: 2 uint64 of device scratch owned by this launch; d_res
 * (device, n_queries records) receives the results; nothing is
 * synchronised.  Launches whose slot ranges are disjoint may run
 * concurrently on different streams.  The graph must not be modified while
 * a launch runs.  flsg_stream_poll tells when it has ended; the serving
 * session (flsg_session_*) admits queries into a running launch instead. */
int flsg_batch_launch(flsg_graph g, const flsg_query *d_queries, int64_t n_queries, int32_t slot_base,
                      int32_t n_slots, int32_t *d_out_ids, double *d_out_vals, int64_t *d_out_off, int64_t out_cap,
                      uint64_t *d_ctl, flsg_result *d_res, int64_t max_iters, const flsg_hk_config *hk,
                      void *stream);

/* *h_done = 1 when every operation queued on `stream` so far (a
 * flsg_batch_launch and the caller's copies of its results) has completed,
 * 0 while any is still running; never blocks.  A fault of that work is
 * returned as FLSG_EHIP.  No counterpart in the reference (it has no
 * asynchronous launches). */
int flsg_stream_poll(void *stream, int32_t *h_done);

/* Serving session: continuous admission into ONE persistent batched launch.
 * No counterpart in the reference (it has no serving loop; its per-query GPU
 * functions, ppr_solver.py:82-146, copy the graph and launch per query).
 *
 * flsg_session_create allocates a ring of query_cap query / result records
 * and done flags in host-coherent memory, and an output arena of out_cap
 * (id, value) entries on the device.  flsg_session_open launches k_batch over
 * slots [slot_base, slot_base + n_slots) of the reserved arena on the
 * session's own non-blocking stream: every workgroup claims the next ring
 * index and waits (bounded) until the host publishes it; a slot takes the
 * next published query as soon as its previous query is done, so queries
 * that arrive while others run start as slots free up.  The launch serves at
 * most query_cap queries; it ends when the host closes the session and every
 * published query is done, or -- a safety net -- when a workgroup has seen
 * nothing published for idle_seconds (then every workgroup ends and
 * flsg_session_poll reports the idle exit; queries published but not done
 * must be submitted again in a new launch).  Heat-kernel settings and the
 * iteration cap are per launch.
 *
 * flsg_session_submit copies n records into the ring and publishes them
 * (*h_first = the ring index of the first); FLSG_ECAP when the ring is full.
 * flsg_session_done copies the done flags of [first, first + count) (1 = the
 * query's result record and output are written); flsg_session_result copies
 * a done query's record and output offset (-1: no output).  Neither waits.
 * flsg_session_close publishes nothing more: the launch ends after the
 * published queries.  flsg_session_poll: *h_ended = 1 once the launch has
 * ended (never blocks), *h_idle_exits the idle exits of the ended launch.
 * flsg_session_output copies output entries to the host, only once the
 * launch has ended; the arena is kept until the next open.
 *
 * While a session is open the graph must not be modified, and nothing may
 * wait for the whole device (hipDeviceSynchronize, torch.cuda.synchronize):
 * the launch runs until the session is closed.  Destroy sessions before
 * their graph. */
typedef struct flsg_session_s *flsg_session;
int flsg_session_create(flsg_graph g, int64_t query_cap, int64_t out_cap, flsg_session *out);
int flsg_session_open(flsg_session s, int32_t slot_base, int32_t n_slots, const flsg_hk_config *hk,
                      int64_t max_iters, double idle_seconds);
int flsg_session_submit(flsg_session s, const flsg_query *h_queries, int64_t n, int64_t *h_first);
int flsg_session_done(flsg_session s, int64_t first, int64_t count, int32_t *h_done);
int flsg_session_result(flsg_session s, int64_t i, flsg_result *h_res, int64_t *h_out_off);
int flsg_session_close(flsg_session s);
int flsg_session_poll(flsg_session s, int32_t *h_ended, int64_t *h_idle_exits);
int flsg_session_output(flsg_session s, int64_t off, int64_t count, int32_t *h_ids, double *h_vals);
int flsg_session_destroy(flsg_session s);

/* Rounding-error bound the batched FP32 certificate applies to a boundary
 * node (reached by the push, never pushed) of degree d whose FP32 residual is
 * r32 after `iters` push iterations: |r_FP64 - r32| <= bound, r_FP64 being the
 * FP64 residual of the returned x (0 for r32 == 0; |r32| is used, the sign
 * marks a residual the correction push has moved into x).  The bound counts
 * the push values the node can have received, m = min(d iters, pushes):
 * `pushes` is the query's number of pushes (frontier entries over all
 * iterations), a valid cap when no row of the graph repeats an entry (a push
 * sends one value per row entry); pushes <= 0 means no cap (m = d iters).
 * Valid while 2 m 2^-24 <= 1/4 (+inf beyond); the kernels recompute every
 * node the bound does not certify in FP64.  Host function, no GPU needed.  No
 * counterpart in the reference (the certificate's error bound, exposed for
 * tests). */
double flsg_boundary_bound(float r32, int32_t d, int64_t iters, int64_t pushes);

/* The error bound a pushed node of the batched FP32 path is certified with
 * when it is not recomputed: |r_v - r32_v| <= flsg_support_bound(x32_v,
 * r32_v, d_v, iters, b_v, pushes), r_v the exact residual of the returned x32
 * and b_v the right-hand side at v (alpha or 1 at the seed, else 0); pushes
 * caps the adds as in flsg_boundary_bound.  Host-callable, for tests.  No
 * counterpart in the reference. */
double flsg_support_bound(float x32, float r32, int32_t d, int64_t iters, double b, int64_t pushes);

/* Number of slot-arena records not back in the between-queries state (every
 * field zero except the degree copy) -- 0 after any completed launch.
 * Synchronises the stream.  No counterpart in the reference (a check of the
 * arena's reset invariant, for tests). */
int flsg_batch_arena_dirty(flsg_graph g, int64_t *h_count, void *stream);

/* After flsg_graph_insert_edges on a handle with a slot arena: refresh the
 * degree copy the slots keep for the listed nodes. */
int flsg_batch_refresh_degrees(flsg_graph g, const int32_t *d_keys, int64_t n_keys, void *stream);

/* ---------------------------------------------------------------- coloured SOR */

/* Graph-coloured SOR / Gauss-Southwell push.  Replaces
 *   ppr_sor(..., omega, local=True/False)     ppr_solver.py:331-461
 *   katz_sor(..., omega, local=True/False)    katz_solver.py:252-337
 * (omega = 1 is the APPR / AKatz push of the solve_a_graph_all drivers,
 * ppr_solver.py:559-570, katz_solver.py:630-640).
 * Colours are visited in order and the nodes of one colour, which share no
 * edge, push at once; the local variant sweeps its frontier this way, the
 * global variant every live node.  Both stop on the reference's certificate
 * |r_u| < eps*alpha*d_u (Katz: eps*d_u) for every live u.  p->kind is
 * FLSG_PPR or FLSG_KATZ (p->alpha = beta; x excludes e_s as katz_sor's),
 * p->precision FLSG_FP64, p->local selects the variant; 0 < omega < 2.
 * d_x, d_r: dense FP64 outputs of n entries.  h_res->oper is the reference's
 * operation count (local: sum of degree over pushed nodes; global: n + m per
 * sweep), iters the frontier passes / sweeps, nodes the pushed nodes.
 * The colouring is computed on first use (and after edge insertions) and is
 * not part of h_res->ms.  Graphs with self-loops are rejected (FLSG_EARG). */
int flsg_sor_solve(flsg_graph g, const flsg_params *p, double omega, int64_t seed, double *d_x, double *d_r,
                   flsg_result *h_res, void *stream);

/* Graph-coloured SOR push of one InstantGNN feature column on a (slack) CSR
 * that may hold self-loops.  Replaces instantGNNSorPush (InstantGNN.py:74-105;
 * instantGNNSorParallel, :108-134, calls it per column).  d_p, d_r: FP64 n,
 * updated in place.  d_eps_vec = eps * d^(1-beta), d_dpow1 = d^(1-beta),
 * d_dpowb = d^beta (device, n doubles each; the caller computes them as the
 * reference does, np.power).  The frontier starts from d_nodes (n_nodes
 * distinct ids) or, when d_nodes is NULL, from every node with
 * eps_vec <= |r| and r != 0.  A pushed node u moves res = omega r_u:
 * p_u += alpha res, and (1-alpha) res / d_u^(1-beta) / d_v^beta to each live
 * neighbour v (its self-loops included); an isolated node or one whose only
 * edge is a self-loop takes p_u = r_u, r_u = 0.  Ends with |r_u| < eps_vec[u]
 * (r_u = 0 where eps_vec is 0).  max_iters <= 0: 1e5. */
int flsg_sor_ignn_column(flsg_graph g, double alpha, double omega, const double *d_eps_vec,
                         const double *d_dpow1, const double *d_dpowb, double *d_p, double *d_r,
                         const int32_t *d_nodes, int64_t n_nodes, int64_t max_iters, flsg_result *h_res,
                         void *stream);

/* The colouring the SOR solve uses (Jones-Plassmann rounds, priority
 * (degree, hash), first-fit colours): d_colour_out (device, n int32, may be
 * NULL), *h_num_colours, *h_rounds.  No counterpart in the reference. */
int flsg_graph_colouring(flsg_graph g, int32_t *d_colour_out, int32_t *h_num_colours, int64_t *h_rounds,
                         void *stream);

/* ---------------------------------------------------------------- dynamic graphs */

/* In-place insertion of k undirected edges (d_us[i], d_vs[i]) into a slack-CSR
 * handle.  Replaces updetaGraph(onlyinsert=True) (InstantGNN.py:137-180,
 * ppr_solver.py:723-766) with the same bytes: pairs are taken in batch order,
 * a pair whose v is already in u's live row (present before the batch, or
 * inserted by an earlier pair of it, in either orientation) is skipped, and
 * each row's new neighbours are appended in batch order.  Nothing is written
 * and an error is returned if an endpoint is outside [0, n) (FLSG_EARG), if a
 * self-loop would be inserted (FLSG_EARG: the reference raises the degree by
 * two but writes one slot, leaving an unwritten entry in the live row; a
 * self-loop already present is skipped), or if a row's slack would overflow
 * (FLSG_ECAP).  Outputs (any may be NULL): d_keys_out (device, up to 2k
 * entries) the endpoints that gained edges in the reference's insert_neighbor
 * key order (first insertion), *n_keys their number, d_old_degree (device, n)
 * the degrees before the batch, *n_inserted the number of edges added.  The
 * new neighbours of u -- its insert_neighbor list -- are
 * indices[indptr[u]+old_degree[u] : indptr[u]+degree[u]].  The slot arena's
 * degree copies of those rows are refreshed in the same call. */
int flsg_graph_insert_edges(flsg_graph g, const int32_t *d_us, const int32_t *d_vs, int64_t k, int32_t *d_keys_out,
                            int64_t *n_keys, int32_t *d_old_degree, int64_t *n_inserted, void *stream);

/* ---------------------------------------------------------------- InstantGNN */

typedef struct {
    int32_t precision;       /* FLSG_FP64 plain / FLSG_FP32 push + FP64 certificate */
    int32_t degrees_changed; /* recompute d^beta after an edge insertion          */
    int64_t cols;            /* feature columns C (row-major n x C state)         */
    double alpha;
    double eps;              /* threshold eps * d_u^(1-beta)                      */
    double beta;
    int64_t max_rounds;      /* <= 0: 1e6                                         */
} flsg_ignn_params;

/* Feature propagation P ~ sum_k alpha (1-alpha)^k (D^-b A D^(b-1))^k X for all
 * C columns at once.  d_P, d_R: FP64 n x C row-major state, updated in place
 * (initially P = 0, R = X).  The frontier starts from d_nodes (device,
 * n_nodes entries) or from all nodes when d_nodes is NULL.  FP32 precision runs
 * the rounds in FP32 and then restores |R| < eps d^(1-b) in FP64 from the
 * recomputed residual R = X - P/alpha + ((1-alpha)/alpha) M P (needs d_X,
 * FP64 n x C).  Replaces instantGNNPush / instantGNNParallel
 * (InstantGNN.py:12-71). */
int flsg_ignn_propagate(flsg_graph g, const flsg_ignn_params *p, const double *d_X, double *d_P, double *d_R,
                        const int32_t *d_nodes, int64_t n_nodes, flsg_result *h_res, void *stream);

/* After flsg_graph_insert_edges: rescale / correct P and R of the endpoints
 * (d_keys) so the propagation invariant holds on the new graph; follow with
 * flsg_ignn_propagate(..., d_keys, n_keys, ...).  Replaces the adjustment part
 * of instantGNNUpdateBatch (InstantGNN.py:298-318). */
int flsg_ignn_update(flsg_graph g, const flsg_ignn_params *p, const double *d_X, double *d_P, double *d_R,
                     const int32_t *d_keys, int64_t n_keys, const int32_t *d_old_degree, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FLSG_H */
