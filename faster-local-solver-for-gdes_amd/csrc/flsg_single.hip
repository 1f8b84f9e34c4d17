// flsg_single.hip -- single-source solves spread over the whole GPU.
//
// Local push (PPR / Katz gradient descent, heat-kernel levels).  One
// iteration = three kernels:
//   k_node  : for every frontier node u (frontier = list of ids):
//               tmp = r_u; r_u = 0; x_u += tmp; val_u = push value;
//             nodes with degree > kHub are also split into edge "pieces";
//             adds vol(frontier) to the reference's operation count.
//   k_push  : work items (claimed through an atomic counter) are either a
//             chunk of kNT regular frontier entries -- their edges are
//             flattened with an LDS prefix sum so every thread streams
//             consecutive col-indices regardless of the degree mix -- or one
//             kPiece-edge piece of a hub.  Each edge does
//               old = atomicAdd(r_v, val);   new = old + val;
//             and v joins the next frontier iff old < thr_v <= new.  All
//             pushes are >= 0, so r_v only grows and exactly one atomic sees
//             the crossing: the frontier is built without a mark array and
//             without an O(n) scan (the reference GPU path scans all n nodes
//             every iteration, ppr_solver.py:139).  A crossing node u is
//             appended to sub-list u mod 256, each with its own counter;
//   k_front_compact : joins the sub-lists into the next frontier array.
// The host launches iterations in blocks and reads the frontier size back
// once per block; launches past the end see an empty frontier and exit.
//
// Global sweeps (ppr_gd/katz_gd local=False, ppr_cheby/katz_cheby
// local=False) are pull-based: r_new[v] = sum over the sorted row of v of the
// neighbours' push values, accumulated sequentially in row order -- the same
// order the reference's `for u in range(n)` scatter produces -- so the FP64
// sweeps round exactly like the reference.
#include <math.h>

#include <string.h>

#include <chrono>
#include <mutex>
#include <vector>

#include "flsg_internal.h"

namespace flsg {

constexpr int kNT = 256;          // threads per block
constexpr int kHub = 4096;        // degree above which a node is split into pieces
constexpr int kPiece = 4096;      // edges per hub piece
constexpr int kIterBlock = 8;     // iterations launched between host checks (even: parity restarts at 0)
static_assert(kIterBlock % 2 == 0, "a block must end on parity 0");
constexpr int kRowLight = 64;     // global sweeps: rows up to this length are pulled by one thread,
constexpr int kRowWave = 1024;    // up to this by one wave, longer ones by one block
constexpr int kPB = 1024;         // threads of the long-row block
constexpr int kPRound = kPB - kWave;   // entries per block round

enum { K_PPR = 0, K_KATZ = 1, K_HK = 2 };

struct PushArgs {
    const int32_t *indptr;
    const int32_t *indices;
    const int32_t *degree;
    int32_t *front0;
    int32_t *front1;
    double *val;
    int64_t *pieces;
    int64_t pieces_cap;
    SolveCounters *ctr;
    int32_t *fsub;          // next-frontier sub-lists (kFrontSub x fsub_cap)
    int64_t fsub_cap;
    unsigned int *fsubcnt;  // kFrontCntWords
    double *x;
    double *r;          // residual read by the node phase (HK: level j)
    double *r_push;     // residual pushed into (HK: level j+1; else == r)
    int32_t *trace;     // may be null
    int64_t trace_cap;
    int64_t *trace_cnt; // per-iteration frontier sizes (device), may be null
    int64_t trace_iters_cap;
    double *lvl_sums;   // HK diagnostics: [2*j] sum |r_u| processed, [2*j+1] mass pushed
    // push value / threshold constants
    double alpha;       // PPR alpha / Katz beta
    double om_alpha;    // 1 - alpha
    double thr_scale;   // PPR: eps*alpha, Katz: eps, HK: exp(tau)*eps
    double thr_div1;    // HK: N
    double thr_div2;    // HK: psi[j+1]
    double node_div1;   // HK: N       (skip test at level j)
    double node_div2;   // HK: psi[j]
    double hk_coef;     // HK: tau / (j+1)
    int last_level;     // HK: pushes go to x (j+1 == N)
    int level;          // HK level j
};

// threshold of node v for the frontier test, bit-identical to the reference's
// eps_vec expression (left-to-right evaluation).
template <int KIND>
__device__ __forceinline__ double push_threshold(const PushArgs &a, int d) {
    if constexpr (KIND == K_HK) return ((a.thr_scale * (double)d) / a.thr_div1) / a.thr_div2;
    else return a.thr_scale * (double)d;
}

template <int KIND>
__device__ __forceinline__ double node_threshold(const PushArgs &a, int d) {
    if constexpr (KIND == K_HK) return ((a.thr_scale * (double)d) / a.node_div1) / a.node_div2;
    else return a.thr_scale * (double)d;
}

template <int KIND>
__device__ __forceinline__ double push_value(const PushArgs &a, double tmp, int d) {
    if constexpr (KIND == K_PPR) return a.om_alpha * tmp / (double)d;    // ppr_solver.py:282
    else if constexpr (KIND == K_KATZ) return a.alpha * tmp;               // katz_solver.py:212
    else {
        if (a.last_level) return tmp / (double)d;                          // hk_solver.py:255
        return a.hk_coef * tmp / (double)d;                                // hk_solver.py:249
    }
}

// ------------------------------------------------------------------ node phase
template <int KIND>
__global__ __launch_bounds__(kNT) void k_node(PushArgs a, int p) {
    __shared__ unsigned long long s_vol[kNT / kWave];
    __shared__ double s_sum[2][kNT / kWave];
    SolveCounters *c = a.ctr;
    const unsigned int cnt = c->cnt[p];
    const unsigned long long iters = c->iters;
    // the sub-list counters this iteration's push appends to (last read by
    // the compaction two launches back)
    if (blockIdx.x == 0)
        for (int t = threadIdx.x; t < kFrontSub; t += kNT) a.fsubcnt[((p ^ 1) * kFrontSub + t) * kFrontCs] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->cnt[p ^ 1] = 0;
        c->npieces[p ^ 1] = 0;
        c->work[p ^ 1] = 0;
        if (cnt > 0) {
            c->iters = iters + 1;
            if (a.trace_cnt && (int64_t)iters < a.trace_iters_cap) a.trace_cnt[iters] = cnt;
        }
    }
    if (cnt == 0) return;
    const int32_t *F = p ? a.front1 : a.front0;
    const unsigned long long toff = c->trace_off;
    unsigned long long vol = 0;
    double abs_sum = 0.0, mass_sum = 0.0;
    for (unsigned int i = blockIdx.x * kNT + threadIdx.x; i < cnt; i += gridDim.x * kNT) {
        const int u = F[i];
        const int d = a.degree[u];
        const double tmp = a.r[u];
        double v = 0.0;
        bool skip = false;
        if constexpr (KIND == K_HK) skip = fabs(tmp) < node_threshold<KIND>(a, d);   // hk_solver.py:244
        if (!skip) {
            a.r[u] = tmp - tmp;                 // ppr_solver.py:277 (r_u -= r_u)
            a.x[u] = a.x[u] + tmp;              // ppr_solver.py:276
            v = push_value<KIND>(a, tmp, d);
            vol += (unsigned long long)d;
            if constexpr (KIND == K_HK) {
                abs_sum += fabs(tmp);
                if (!a.last_level) mass_sum += v * (double)d;
            }
            if (d > kHub) {
                const unsigned int npc = (unsigned int)((d + kPiece - 1) / kPiece);
                const unsigned int base = atomicAdd(&c->npieces[p], npc);
                if ((int64_t)base + npc <= a.pieces_cap) {
                    for (unsigned int k = 0; k < npc; k++)
                        a.pieces[base + k] = ((int64_t)i << 32) | (int64_t)k;
                } else {
                    atomicOr(&c->overflow, 2u);
                }
            }
        }
        a.val[i] = v;
        if (a.trace) {
            if (toff + i < (unsigned long long)a.trace_cap) a.trace[toff + i] = u;
            else atomicOr(&c->overflow, 1u);
        }
    }
    // block reduction of vol (+ HK sums)
    vol = wave_sum(vol);
    if constexpr (KIND == K_HK) {
        abs_sum = wave_sum(abs_sum);
        mass_sum = wave_sum(mass_sum);
    }
    const int wid = threadIdx.x / kWave;
    if (__lane_id() == 0) {
        s_vol[wid] = vol;
        s_sum[0][wid] = abs_sum;
        s_sum[1][wid] = mass_sum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        double s0 = 0.0, s1 = 0.0;
        for (int w = 0; w < kNT / kWave; w++) {
            t += s_vol[w];
            s0 += s_sum[0][w];
            s1 += s_sum[1][w];
        }
        if (t) atomicAdd(&c->oper, t);
        if constexpr (KIND == K_HK) {
            if (a.lvl_sums) {
                atomic_add_dev(&a.lvl_sums[2 * a.level], s0);
                atomic_add_dev(&a.lvl_sums[2 * a.level + 1], s1);
            }
        }
    }
}

// one edge: returns true iff this atomic moved r_v across its threshold
template <int KIND>
__device__ __forceinline__ bool push_edge(const PushArgs &a, int v, double val) {
    if constexpr (KIND == K_HK) {
        if (a.last_level) {
            atomic_add_dev(&a.x[v], val);
            return false;
        }
    }
    const double old = atomic_add_dev(&a.r_push[v], val);
    const double nw = old + val;
    const double thr = push_threshold<KIND>(a, a.degree[v]);
    return old < thr && thr <= nw;
}

// a node that crossed its threshold joins the next frontier through sub-list
// (u mod kFrontSub): one counter per sub-list instead of one for all, since
// returning atomics on a single address serialise (~11 ns each on MI355X)
__device__ __forceinline__ void append_front(const PushArgs &a, unsigned int *subcnt, int u) {
    const int sub = u & (kFrontSub - 1);
    const unsigned int slot = atomicAdd(subcnt + sub * kFrontCs, 1u);
    // a node crosses at most once per iteration (pushes are >= 0), so the
    // sub-list cannot fill; the guard keeps a broken input from writing past it
    if ((int64_t)slot < a.fsub_cap) a.fsub[sub * a.fsub_cap + slot] = u;
    else atomicOr(&a.ctr->overflow, 4u);
}

// ------------------------------------------------------------------ push phase
template <int KIND>
__global__ __launch_bounds__(kNT) void k_push(PushArgs a, int p) {
    __shared__ int s_incl[kNT];
    __shared__ int s_start[kNT];
    __shared__ double s_val[kNT];
    __shared__ int s_scan[kNT / kWave];
    __shared__ unsigned int s_item;
    SolveCounters *c = a.ctr;
    const unsigned int cnt = c->cnt[p];
    if (cnt == 0) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) c->trace_off = c->trace_off + cnt;
    const int32_t *F = p ? a.front1 : a.front0;
    unsigned int *subcnt = a.fsubcnt + (p ^ 1) * kFrontSub * kFrontCs;
    const unsigned int nchunks = (cnt + kNT - 1) / kNT;
    unsigned int npc = c->npieces[p];
    if ((int64_t)npc > a.pieces_cap) npc = (unsigned int)a.pieces_cap;
    const unsigned int total = nchunks + npc;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(&c->work[p], 1u);
        __syncthreads();
        const unsigned int item = s_item;
        __syncthreads();
        if (item >= total) break;
        if (item < nchunks) {
            const unsigned int i = item * kNT + tid;
            int d = 0, st = 0;
            double v = 0.0;
            if (i < cnt) {
                const int u = F[i];
                d = a.degree[u];
                st = a.indptr[u];
                v = a.val[i];
                if (d > kHub || v == 0.0) d = 0;   // hubs go through pieces; skipped nodes push nothing
            }
            int tot;
            const int incl = block_incl_scan<kNT>(d, s_scan, &tot);
            s_incl[tid] = incl;
            s_start[tid] = st;
            s_val[tid] = v;
            __syncthreads();
            for (int e0 = 0; e0 < tot; e0 += kNT) {
                const int e = e0 + tid;
                bool cross = false;
                int vtx = 0;
                if (e < tot) {
                    const int k = upper_bound_lds(s_incl, 0, kNT, e);
                    const int off = e - (k ? s_incl[k - 1] : 0);
                    vtx = a.indices[s_start[k] + off];
                    cross = push_edge<KIND>(a, vtx, s_val[k]);
                }
                if (cross) append_front(a, subcnt, vtx);
            }
            __syncthreads();
        } else {
            const int64_t pc = a.pieces[item - nchunks];
            const unsigned int i = (unsigned int)(pc >> 32);
            const int k = (int)(pc & 0xffffffff);
            const int u = F[i];
            const int d = a.degree[u];
            const int st = a.indptr[u];
            const double v = a.val[i];
            const int beg = k * kPiece;
            const int end = min(d, beg + kPiece);
            for (int e0 = beg; e0 < end; e0 += kNT) {
                const int e = e0 + tid;
                bool cross = false;
                int vtx = 0;
                if (e < end) {
                    vtx = a.indices[st + e];
                    cross = push_edge<KIND>(a, vtx, v);
                }
                if (cross) append_front(a, subcnt, vtx);
            }
        }
    }
}

// next-frontier sub-lists -> the frontier array of parity p ^ 1 and its size
__global__ __launch_bounds__(kFrontSub) void k_front_compact(PushArgs a, int p) {
    __shared__ unsigned int s_part[kFrontSub / kWave];
    const unsigned int *subcnt = a.fsubcnt + (p ^ 1) * kFrontSub * kFrontCs;
    int32_t *Fn = p ? a.front0 : a.front1;
    const int t = threadIdx.x, sub = blockIdx.x;
    const unsigned int c = min((int64_t)subcnt[t * kFrontCs], a.fsub_cap);
    const unsigned int below = wave_sum(t < sub ? c : 0u);
    if (__lane_id() == 0) s_part[t / kWave] = below;
    __syncthreads();
    unsigned int off = 0;
    for (int w = 0; w < kFrontSub / kWave; w++) off += s_part[w];
    const unsigned int mine = min((int64_t)subcnt[sub * kFrontCs], a.fsub_cap);
    for (unsigned int j = t; j < mine; j += kFrontSub) Fn[off + j] = a.fsub[sub * a.fsub_cap + j];
    if (sub == kFrontSub - 1 && t == 0) a.ctr->cnt[p ^ 1] = off + mine;
}

__global__ void k_init_local(SolveCounters *c, int32_t *front0, double *r, int64_t s, double r0) {
    c->cnt[0] = 1;
    c->cnt[1] = 0;
    c->npieces[0] = 0;
    c->npieces[1] = 0;
    c->work[0] = 0;
    c->work[1] = 0;
    c->act[0] = 0;
    c->act[1] = 0;
    c->oper = 0;
    c->iters = 0;
    c->trace_off = 0;
    c->overflow = 0;
    front0[0] = (int32_t)s;
    r[s] = r0;
}

__global__ void k_add_scalar(double *x, int64_t s, double v) { x[s] = x[s] + v; }

__global__ void k_scale_div(double *x, int64_t n, double div) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = x[i] / div;
}

// HK: clear level-j residual after the level ran (hk_solver.py:282-285); a level
// whose queue was empty ends the loop without clearing, like the reference.
__global__ void k_clear_level(const SolveCounters *c, int p, double *r, int64_t n) {
    if (c->cnt[p] == 0) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        r[i] = 0.0;
}

// ------------------------------------------------------------------ global sweeps
// GD sweep, node part: x += r; val_u = push value (0 for isolated nodes).
template <int KIND>
__global__ __launch_bounds__(kNT) void k_gsweep_node(const int32_t *degree, double *x, const double *r,
                                                     double *val, int64_t n, double alpha, double om_alpha,
                                                     SolveCounters *c, int p) {
    const unsigned int act = c->act[p];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->act[p ^ 1] = 0;   // a finished solve stays finished in later launches
        if (act) c->iters = c->iters + 1;
    }
    if (act == 0) return;
    for (int64_t u = blockIdx.x * (int64_t)kNT + threadIdx.x; u < n; u += (int64_t)gridDim.x * kNT) {
        const double ru = r[u];
        x[u] = x[u] + ru;
        const int d = degree[u];
        double v = 0.0;
        if (d > 0) {
            if constexpr (KIND == K_PPR) v = om_alpha * ru / (double)d;
            else v = alpha * ru;
        }
        val[u] = v;
    }
}

// Pull: out[v] = sum_{u in row(v)} val[u], sequential in row order; counts the
// nodes with thr_v <= |out_v| into act[p^1].  mode 0: r_new = sum;
// mode 1 (PPR Chebyshev): r_new = r + (sum - dx); mode 2 (Katz Chebyshev):
// r_new = r + sum - dx; mode 3 (last heat-kernel level, hk_solver.py:270-272):
// r_new is x, the sum starts from x_v and takes v's own residual r_old[v] at
// its place in the sorted row (before the entries u >= v), the order in which
// the reference's natural-order scatter adds into x_v; no count.
template <int MODE>
__device__ __forceinline__ double sweep_result(double s, const double *r_old, const double *dx, int64_t v) {
    if constexpr (MODE == 0 || MODE == 3) return s;
    else if constexpr (MODE == 1) return r_old[v] + (s - dx[v]);
    else return r_old[v] + s - dx[v];
}

// first position in the sorted row indices[st, st + d) holding an id >= v
__device__ __forceinline__ int row_lower_bound(const int32_t *indices, int st, int d, int v) {
    int lo = 0, hi = d;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (indices[st + mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// term j of a row's sum; mode 3 rows have d + 1 terms, v's own residual at k
template <int MODE>
__device__ __forceinline__ double row_term(const int32_t *indices, const double *val, const double *r_old, int st,
                                           int k, int v, int j) {
    if constexpr (MODE == 3) {
        if (j == k) return r_old[v];
        return val[indices[st + (j < k ? j : j - 1)]];
    } else {
        return val[indices[st + j]];
    }
}

// rows of at most kRowLight entries: one thread each (longer rows are pulled
// by k_gsweep_pull_long, from the lists built by k_row_classes)
template <int MODE>
__global__ __launch_bounds__(kNT) void k_gsweep_pull(const int32_t *indptr, const int32_t *indices,
                                                     const int32_t *degree, const double *val,
                                                     const double *r_old, const double *dx, double *r_new,
                                                     int64_t n, double thr_scale, SolveCounters *c, int p) {
    __shared__ unsigned int s_cnt[kNT / kWave];
    if (c->act[p] == 0) return;
    unsigned int act = 0;
    for (int64_t v = blockIdx.x * (int64_t)kNT + threadIdx.x; v < n; v += (int64_t)gridDim.x * kNT) {
        const int d = degree[v];
        if (d > kRowLight) continue;
        const int st = indptr[v];
        double s = 0.0;
        if constexpr (MODE == 3) {
            s = r_new[v];
            bool own = false;
            for (int j = 0; j < d; j++) {
                const int u = indices[st + j];
                if (!own && u >= v) {
                    s += r_old[v];
                    own = true;
                }
                s += val[u];
            }
            if (!own) s += r_old[v];
        } else {
#pragma unroll 8
            for (int j = 0; j < d; j++) s += val[indices[st + j]];
        }
        const double rn = sweep_result<MODE>(s, r_old, dx, v);
        r_new[v] = rn;
        if (MODE != 3 && d > 0 && thr_scale * (double)d <= fabs(rn)) act++;
    }
    act = wave_sum(act);
    if (__lane_id() == 0) s_cnt[threadIdx.x / kWave] = act;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int t = 0;
        for (int w = 0; w < kNT / kWave; w++) t += s_cnt[w];
        if (t) atomicAdd(&c->act[p ^ 1], t);
    }
}

// adds s_add[0 .. cnt) to s in order; s_add holds -0.0 up to the next multiple
// of 8 (the identity of FP64 addition), and two register sets alternate so
// each group's LDS reads are in flight while the previous group is added
__device__ __forceinline__ void lds_chain(const double *s_add, int cnt, double &s) {
    const int ng = (cnt + 7) >> 3;
    double ga[8], gb[8];
#pragma unroll
    for (int k = 0; k < 8; k++) ga[k] = s_add[k];
    for (int g = 0; g < ng; g += 2) {
        const int ob = (g + 1 < ng ? g + 1 : g) * 8;
#pragma unroll
        for (int k = 0; k < 8; k++) gb[k] = s_add[ob + k];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < 8; k++) s += ga[k];
        if (g + 1 >= ng) break;
        const int oa = (g + 2 < ng ? g + 2 : g) * 8;
#pragma unroll
        for (int k = 0; k < 8; k++) ga[k] = s_add[oa + k];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < 8; k++) s += gb[k];
    }
}

// rows longer than kRowLight: each sum is still one FP64 add after another in
// row order, but its loads are spread over a wave (rows up to kRowWave: 64
// entries per chunk staged in LDS) or a block (longer rows: 15 waves load
// 960-entry rounds into LDS, double-buffered, while wave 0 adds the last).
// Rows are assigned by static strides of uniform values, and the count of
// active rows is reduced per block (one atomic per block).
template <int MODE>
__global__ __launch_bounds__(kPB) void k_gsweep_pull_long(const int32_t *indptr, const int32_t *indices,
                                                          const int32_t *degree, const double *val,
                                                          const double *r_old, const double *dx, double *r_new,
                                                          const int32_t *rows_wave, int64_t n_wave,
                                                          const int32_t *rows_block, int64_t n_block,
                                                          double thr_scale, SolveCounters *c, int p) {
    __shared__ double s_ad[2][kPRound];
    __shared__ double s_w[kPB / kWave][kWave];
    __shared__ unsigned int s_act[kPB / kWave];
    if (__builtin_amdgcn_readfirstlane((int)c->act[p]) == 0) return;
    const int t = threadIdx.x, lane = t & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(t / kWave);
    unsigned int act = 0;
    for (int64_t item = blockIdx.x; item < n_block; item += gridDim.x) {
        const int v = __builtin_amdgcn_readfirstlane(rows_block[item]);
        const int d0 = __builtin_amdgcn_readfirstlane(degree[v]);
        const int st = __builtin_amdgcn_readfirstlane(indptr[v]);
        const int k = MODE == 3 ? __builtin_amdgcn_readfirstlane(row_lower_bound(indices, st, d0, v)) : 0;
        const int d = d0 + (MODE == 3 ? 1 : 0);   // terms of the sum
        double s = MODE == 3 ? r_new[v] : 0.0;
        if (wid > 0) {
            const int j = t - kWave;
            s_ad[0][j] = j < d ? row_term<MODE>(indices, val, r_old, st, k, v, j) : -0.0;
        }
        __syncthreads();
        int b = 0;
        for (int j0 = 0; j0 < d; j0 += kPRound) {
            if (wid > 0) {
                if (j0 + kPRound < d) {
                    const int j = j0 + kPRound + (t - kWave);
                    s_ad[b ^ 1][t - kWave] = j < d ? row_term<MODE>(indices, val, r_old, st, k, v, j) : -0.0;
                }
            } else {
                lds_chain(s_ad[b], d - j0 < kPRound ? d - j0 : kPRound, s);
            }
            __syncthreads();
            b ^= 1;
        }
        if (t == 0) {
            const double rn = sweep_result<MODE>(s, r_old, dx, v);
            r_new[v] = rn;
            if (MODE != 3 && thr_scale * (double)d0 <= fabs(rn)) act++;
        }
    }
    const int64_t nw = (int64_t)gridDim.x * (kPB / kWave);
    for (int64_t item = (int64_t)blockIdx.x * (kPB / kWave) + wid; item < n_wave; item += nw) {
        const int v = __builtin_amdgcn_readfirstlane(rows_wave[item]);
        const int d0 = __builtin_amdgcn_readfirstlane(degree[v]);
        const int st = __builtin_amdgcn_readfirstlane(indptr[v]);
        const int k = MODE == 3 ? __builtin_amdgcn_readfirstlane(row_lower_bound(indices, st, d0, v)) : 0;
        const int d = d0 + (MODE == 3 ? 1 : 0);
        double s = MODE == 3 ? r_new[v] : 0.0;
        for (int j0 = 0; j0 < d; j0 += kWave) {
            const int j = j0 + lane;
            s_w[wid][lane] = j < d ? row_term<MODE>(indices, val, r_old, st, k, v, j) : -0.0;
            __builtin_amdgcn_wave_barrier();
            lds_chain(s_w[wid], d - j0 < kWave ? d - j0 : kWave, s);
            __builtin_amdgcn_wave_barrier();
        }
        if (lane == 0) {
            const double rn = sweep_result<MODE>(s, r_old, dx, v);
            r_new[v] = rn;
            if (MODE != 3 && thr_scale * (double)d0 <= fabs(rn)) act++;
        }
    }
    act = wave_sum(act);
    if (lane == 0) s_act[wid] = act;
    __syncthreads();
    if (t == 0) {
        unsigned int tot = 0;
        for (int w = 0; w < kPB / kWave; w++) tot += s_act[w];
        if (tot) atomicAdd(&c->act[p ^ 1], tot);
    }
}

// row classes of the global sweeps (they depend on the degrees only)
__global__ __launch_bounds__(kNT) void k_row_classes(const int32_t *degree, int64_t n, int32_t *rows_wave,
                                                     int32_t *rows_block, unsigned int *cnt) {
    for (int64_t base = blockIdx.x * (int64_t)kNT; base < n; base += (int64_t)gridDim.x * kNT) {
        const int64_t v = base + threadIdx.x;
        const int d = v < n ? degree[v] : 0;
        const long long sw = wave_append(d > kRowLight && d <= kRowWave, cnt);
        if (sw >= 0) rows_wave[sw] = (int32_t)v;
        const long long sb = wave_append(d > kRowWave, cnt + 1);
        if (sb >= 0) rows_block[sb] = (int32_t)v;
    }
}

// Chebyshev sweep, node part (coefficients per sweep precomputed in FP64 on
// the host, exactly as the reference's scalar recurrences).
// mode 1 (PPR):  t = b1*r + b2*dx  with b1 = beta, b2 = beta-1   (ppr_solver.py:610)
// mode 2 (Katz): t = b1*r - b2*dx                                (katz_solver.py:369)
template <int KIND, int MODE>
__global__ __launch_bounds__(kNT) void k_csweep_node(const int32_t *degree, double *x, const double *r,
                                                     double *dx, double *val, int64_t n, const double *coef,
                                                     double alpha, double om_alpha, SolveCounters *c, int p,
                                                     int64_t sweep) {
    const unsigned int act = c->act[p];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->act[p ^ 1] = 0;   // a finished solve stays finished in later launches
        if (act) c->iters = c->iters + 1;
    }
    if (act == 0) return;
    const double b1 = coef[2 * sweep], b2 = coef[2 * sweep + 1];
    for (int64_t u = blockIdx.x * (int64_t)kNT + threadIdx.x; u < n; u += (int64_t)gridDim.x * kNT) {
        double t;
        if constexpr (MODE == 1) t = b1 * r[u] + b2 * dx[u];
        else t = b1 * r[u] - b2 * dx[u];
        x[u] = x[u] + t;
        dx[u] = t;
        const int d = degree[u];
        double v = 0.0;
        if (d > 0) {
            if constexpr (KIND == K_PPR) v = om_alpha * t / (double)d;
            else v = alpha * t;
        }
        val[u] = v;
    }
}

static int grid_for(flsg_graph g, int per_cu) { return g->num_cus * per_cu; }

static PushArgs base_args(flsg_graph g, double *x, double *r) {
    PushArgs a{};
    a.indptr = g->indptr;
    a.indices = g->indices;
    a.degree = g->degree;
    a.front0 = g->front[0];
    a.front1 = g->front[1];
    a.val = g->val;
    a.pieces = g->pieces;
    a.pieces_cap = g->pieces_cap;
    a.ctr = g->ctr;
    a.fsub = g->fsub;
    a.fsub_cap = g->fsub_cap;
    a.fsubcnt = g->fsubcnt;
    a.x = x;
    a.r = r;
    a.r_push = r;
    return a;
}

struct TraceBufs {
    int64_t *d_cnt = nullptr;
    int64_t cap = 0;
};

static int setup_trace(flsg_graph g, flsg_trace *tr, PushArgs &a, hipStream_t st) {
    if (!tr) return FLSG_OK;
    if (tr->cap_iters > g->trace_dev_off_cap) {
        if (g->trace_dev_off) FLSG_HIP(hipFree(g->trace_dev_off));
        g->trace_dev_off = nullptr;
        FLSG_HIP(hipMalloc((void **)&g->trace_dev_off, sizeof(int64_t) * (tr->cap_iters + 1)));
        g->trace_dev_off_cap = tr->cap_iters;
    }
    a.trace = tr->d_nodes;
    a.trace_cap = tr->cap_nodes;
    a.trace_cnt = g->trace_dev_off;
    a.trace_iters_cap = tr->cap_iters;
    (void)st;
    return FLSG_OK;
}

static int finish_trace(flsg_graph g, flsg_trace *tr, int64_t iters, hipStream_t st) {
    if (!tr) return FLSG_OK;
    const int64_t k = iters < tr->cap_iters ? iters : tr->cap_iters;
    std::vector<int64_t> cnt(k > 0 ? k : 1);
    if (k > 0) FLSG_HIP(hipMemcpyAsync(cnt.data(), g->trace_dev_off, sizeof(int64_t) * k, hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    tr->h_off[0] = 0;
    for (int64_t t = 0; t < k; t++) tr->h_off[t + 1] = tr->h_off[t] + cnt[t];
    tr->n_iters = k;
    tr->overflow = (iters > tr->cap_iters) || (g->h_ctr->overflow & 1u) ||
                   (k > 0 && tr->h_off[k] > tr->cap_nodes);
    return FLSG_OK;
}

// ------------------------------------------------------------------ captured iteration blocks
// One block of kIterBlock iterations (3 kernels each, parity alternating, the
// same arguments throughout a solve) is captured once into a hipGraph and
// replayed until the frontier is empty.  The graph holds kernel nodes only
// (counters are cleared by k_node, not by memsets).  The executable stays
// cached on the handle while the arguments are unchanged (callers that reuse
// their x / r buffers skip the capture).
struct BlockGraph {
    hipStream_t cs = nullptr;       // capture stream (nothing runs on it)
    hipGraphExec_t exec = nullptr;
    int kind = -1;
    int grid = 0;
    PushArgs args{};
};

void single_graphs_free(flsg_graph g) {
    BlockGraph *c = (BlockGraph *)g->graphs;
    if (!c) return;
    if (c->exec) (void)hipGraphExecDestroy(c->exec);
    if (c->cs) (void)hipStreamDestroy(c->cs);
    delete c;
    g->graphs = nullptr;
}

template <int KIND>
static int push_block_exec(flsg_graph g, const PushArgs &a, int grid, hipGraphExec_t *out) {
    BlockGraph *c = (BlockGraph *)g->graphs;
    if (!c) {
        c = new BlockGraph();
        g->graphs = c;
        FLSG_HIP(hipStreamCreateWithFlags(&c->cs, hipStreamNonBlocking));
    }
    if (c->exec && c->kind == KIND && c->grid == grid && memcmp(&c->args, &a, sizeof(PushArgs)) == 0) {
        *out = c->exec;
        return FLSG_OK;
    }
    if (c->exec) {
        FLSG_HIP(hipGraphExecDestroy(c->exec));
        c->exec = nullptr;
    }
    hipGraph_t gr = nullptr;
    FLSG_HIP(hipStreamBeginCapture(c->cs, hipStreamCaptureModeThreadLocal));
    for (int b = 0; b < kIterBlock; b++) {
        const int p = b & 1;   // blocks start at even iterations
        hipLaunchKernelGGL(k_node<KIND>, dim3(grid), dim3(kNT), 0, c->cs, a, p);
        hipLaunchKernelGGL(k_push<KIND>, dim3(grid), dim3(kNT), 0, c->cs, a, p);
        hipLaunchKernelGGL(k_front_compact, dim3(kFrontSub), dim3(kFrontSub), 0, c->cs, a, p);
    }
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(c->cs, &gr);
    FLSG_HIP(le);
    FLSG_HIP(ce);
    const hipError_t ie = hipGraphInstantiate(&c->exec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    FLSG_HIP(ie);
    c->kind = KIND;
    c->grid = grid;
    memcpy(&c->args, &a, sizeof(PushArgs));
    *out = c->exec;
    return FLSG_OK;
}

// ------------------------------------------------------------------ local GD / Katz GD
template <int KIND>
static int run_local_push(flsg_graph g, PushArgs a, int64_t max_iters, hipStream_t st, int64_t *iters_out) {
    const int grid = grid_for(g, 8);
    hipGraphExec_t exec = nullptr;
    int rc = push_block_exec<KIND>(g, a, grid, &exec);
    if (rc) return rc;
    for (;;) {
        FLSG_HIP(hipGraphLaunch(exec, st));
        FLSG_HIP(hipMemcpyAsync(g->h_ctr, g->ctr, sizeof(SolveCounters), hipMemcpyDeviceToHost, st));
        FLSG_HIP(hipStreamSynchronize(st));
        if (g->h_ctr->cnt[0] == 0) break;   // the next block's first frontier (kIterBlock is even)
        if ((int64_t)g->h_ctr->iters >= max_iters) {
            *iters_out = (int64_t)g->h_ctr->iters;
            return fail(FLSG_ENOCONV, "iteration cap reached");
        }
    }
    if (g->h_ctr->overflow & 2u) return fail(FLSG_ECAP, "hub piece list overflow");
    if (g->h_ctr->overflow & 4u) return fail(FLSG_ECAP, "next-frontier sub-list overflow (a node crossed twice in one iteration)");
    *iters_out = (int64_t)g->h_ctr->iters;
    return FLSG_OK;
}

static int solve_local_gd(flsg_graph g, const flsg_params *p, int64_t s, double *d_x, double *d_r,
                          flsg_result *res, flsg_trace *tr, hipStream_t st) {
    const int64_t n = g->n;
    // the push runs on the handle's x / r (fixed addresses: the captured
    // iteration block is reused across solves) and copies them out at the end
    double *lx = g->local_x, *lr = g->local_r;
    FLSG_HIP(hipMemsetAsync(lx, 0, sizeof(double) * n, st));
    FLSG_HIP(hipMemsetAsync(lr, 0, sizeof(double) * n, st));
    PushArgs a = base_args(g, lx, lr);
    a.alpha = p->alpha;
    a.om_alpha = 1.0 - p->alpha;
    const bool ppr = p->kind == FLSG_PPR;
    a.thr_scale = ppr ? p->eps * p->alpha : p->eps;
    int rc = setup_trace(g, tr, a, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_init_local, dim3(1), dim3(1), 0, st, g->ctr, g->front[0], lr, s, ppr ? p->alpha : 1.0);
    int64_t iters = 0;
    rc = ppr ? run_local_push<K_PPR>(g, a, p->max_iters, st, &iters)
             : run_local_push<K_KATZ>(g, a, p->max_iters, st, &iters);
    if (rc) return rc;
    if (!ppr) hipLaunchKernelGGL(k_add_scalar, dim3(1), dim3(1), 0, st, lx, s, -1.0);   // katz_solver.py:243
    FLSG_HIP(hipMemcpyAsync(d_x, lx, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    FLSG_HIP(hipMemcpyAsync(d_r, lr, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    res->oper = (int64_t)g->h_ctr->oper;
    res->iters = iters;
    res->nodes = (int64_t)g->h_ctr->trace_off;
    return finish_trace(g, tr, iters, st);
}

// ------------------------------------------------------------------ heat kernel
double host_estimate_coefficient(double t, int64_t k) {
    double t0 = t;
    if (k == 0) return t0;
    for (int64_t i = 0; i < k; i++) t0 = t0 * (t / (double)(i + 2));
    return t0;
}

// hk_solver.py:28-48 (k = ceil(t-1) used as an integer count)
int64_t host_estimate_n(double t, double eps) {
    int64_t k = (int64_t)ceil(t - 1.);
    double kf = (double)k;
    double cur_est = host_estimate_coefficient(t, k) * ((kf + 2.) / (kf + 2. - t));
    if (cur_est < eps * exp(t)) return k;
    int64_t est_k = k;
    for (int64_t i = k; i < 1000 * (k + 1); i++) {
        double tmp = ((kf + 2. - t) * t / pow(kf + 2., 2.)) * ((kf + 3.) / (kf + 3. - t));
        cur_est = cur_est * tmp;
        est_k += 1;
        if (cur_est < eps * exp(t)) return est_k;
    }
    return est_k;
}

static int solve_hk_global(flsg_graph g, const flsg_params *p, int64_t s, int64_t N, double *d_x, double *d_r,
                           flsg_result *res, hipStream_t st);

// per-level diagnostics of the local heat-kernel push: 2N doubles on the handle
static int ensure_hk_levels(flsg_graph g, int64_t N) {
    if (g->hk_lvl_cap >= 2 * N) return FLSG_OK;
    if (g->hk_lvl) FLSG_HIP(hipFree(g->hk_lvl));
    g->hk_lvl = nullptr;
    g->hk_lvl_cap = 0;
    FLSG_HIP(hipMalloc((void **)&g->hk_lvl, sizeof(double) * 2 * N));
    g->hk_lvl_cap = 2 * N;
    return FLSG_OK;
}

static int solve_hk(flsg_graph g, const flsg_params *p, int64_t s, double *d_x, double *d_r,
                    flsg_result *res, flsg_trace *tr, hipStream_t st, double *h_lvl_sums, int64_t lvl_cap,
                    int64_t *est_n_out) {
    const int64_t n = g->n;
    if (p->tau < 1.0) return fail(FLSG_EARG, "heat kernel needs tau >= 1");
    const int64_t N = p->n_terms > 0 ? p->n_terms : host_estimate_n(p->tau, .5 * p->eps);
    if (est_n_out) *est_n_out = N;
    if (h_lvl_sums && lvl_cap < 2 * N)
        return fail(FLSG_ECAP, "level sums need " + std::to_string(2 * N) + " doubles, the buffer holds " +
                                   std::to_string(lvl_cap));
    if (!p->local) {
        if (tr) return fail(FLSG_EARG, "the global heat-kernel sweep has no frontier trace");
        if (h_lvl_sums) memset(h_lvl_sums, 0, sizeof(double) * 2 * N);   // the local push's diagnostic
        return solve_hk_global(g, p, s, N, d_x, d_r, res, st);
    }
    std::vector<double> psi(N + 1);
    psi[N] = 1.;
    for (int64_t i = N - 1; i >= 0; i--) psi[i] = psi[i + 1] * p->tau / ((double)i + 1.) + 1.;
    FLSG_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, st));
    FLSG_HIP(hipMemsetAsync(d_r, 0, sizeof(double) * 2 * n, st));
    // per-level diagnostics
    int rc0 = ensure_hk_levels(g, N);
    if (rc0) return rc0;
    double *d_lvl = g->hk_lvl;
    FLSG_HIP(hipMemsetAsync(d_lvl, 0, sizeof(double) * 2 * N, st));
    PushArgs a = base_args(g, d_x, d_r);
    a.thr_scale = exp(p->tau) * p->eps;
    a.thr_div1 = (double)N;
    a.node_div1 = (double)N;
    a.lvl_sums = d_lvl;
    int rc = setup_trace(g, tr, a, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_init_local, dim3(1), dim3(1), 0, st, g->ctr, g->front[0], d_r, s, 1.0);
    const int grid = grid_for(g, 8);
    int64_t j = 0;
    bool done = false;
    while (!done && j < N) {
        for (int b = 0; b < kIterBlock && j < N; b++, j++) {
            const int par = (int)(j & 1);
            a.r = d_r + (int64_t)par * n;
            a.r_push = d_r + (int64_t)(par ^ 1) * n;
            a.level = (int)j;
            a.node_div2 = psi[j];
            a.thr_div2 = psi[j + 1];
            a.hk_coef = p->tau / ((double)j + 1.);
            a.last_level = (j + 1 == N);
            hipLaunchKernelGGL(k_node<K_HK>, dim3(grid), dim3(kNT), 0, st, a, par);
            hipLaunchKernelGGL(k_push<K_HK>, dim3(grid), dim3(kNT), 0, st, a, par);
            hipLaunchKernelGGL(k_front_compact, dim3(kFrontSub), dim3(kFrontSub), 0, st, a, par);
            hipLaunchKernelGGL(k_clear_level, dim3(grid), dim3(kNT), 0, st, (const SolveCounters *)g->ctr, par,
                               a.r, n);
        }
        FLSG_HIP(hipGetLastError());
        FLSG_HIP(hipMemcpyAsync(g->h_ctr, g->ctr, sizeof(SolveCounters), hipMemcpyDeviceToHost, st));
        FLSG_HIP(hipStreamSynchronize(st));
        if (g->h_ctr->cnt[j & 1] == 0) done = true;
    }
    if (g->h_ctr->overflow & 2u) return fail(FLSG_ECAP, "hub piece list overflow");
    if (g->h_ctr->overflow & 4u) return fail(FLSG_ECAP, "next-frontier sub-list overflow (a node crossed twice in one iteration)");
    // x /= e^t  (hk_solver.py:291)
    hipLaunchKernelGGL(k_scale_div, dim3(grid), dim3(kNT), 0, st, d_x, n, exp(p->tau));
    res->oper = (int64_t)g->h_ctr->oper;
    res->iters = (int64_t)g->h_ctr->iters;
    if (h_lvl_sums) FLSG_HIP(hipMemcpyAsync(h_lvl_sums, d_lvl, sizeof(double) * 2 * N, hipMemcpyDeviceToHost, st));
    return finish_trace(g, tr, res->iters, st);
}

// ------------------------------------------------------------------ global sweeps
__global__ void k_init_global(SolveCounters *c) {
    c->act[0] = 0;
    c->act[1] = 0;
    c->iters = 0;
    c->oper = 0;
    c->overflow = 0;
}

// act[0] = #{v : degree_v > 0 and thr_v <= |r_v|}
__global__ __launch_bounds__(kNT) void k_count_active(const int32_t *degree, const double *r, int64_t n,
                                                       double thr_scale, SolveCounters *c) {
    __shared__ unsigned int s_cnt[kNT / kWave];
    unsigned int act = 0;
    for (int64_t v = blockIdx.x * (int64_t)kNT + threadIdx.x; v < n; v += (int64_t)gridDim.x * kNT) {
        const int d = degree[v];
        if (d > 0 && thr_scale * (double)d <= fabs(r[v])) act++;
    }
    act = wave_sum(act);
    if (__lane_id() == 0) s_cnt[threadIdx.x / kWave] = act;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int t = 0;
        for (int w = 0; w < kNT / kWave; w++) t += s_cnt[w];
        if (t) atomicAdd(&c->act[0], t);
    }
}

// Chebyshev initial state (ppr_solver.py:588-601, katz_solver.py:346-354):
// x = dx = x0 e_s, r = r_s0 e_s, then r[v] += val for v in row(s) in order.
__global__ void k_cheb_init(const int32_t *indptr, const int32_t *indices, const int32_t *degree, double *x,
                            double *dx, double *r, int64_t s, double x0, double r_s0, double val) {
    x[s] = x0;
    dx[s] = x0;
    r[s] = r_s0;
    const int st = indptr[s], d = degree[s];
    for (int j = 0; j < d; j++) r[indices[st + j]] += val;
}

static int ensure_coef(flsg_graph g, int64_t k) {
    if (k <= g->coef_cap) return FLSG_OK;
    if (g->coef) FLSG_HIP(hipFree(g->coef));
    g->coef = nullptr;
    FLSG_HIP(hipMalloc((void **)&g->coef, sizeof(double) * 2 * k));
    g->coef_cap = k;
    return FLSG_OK;
}

// Chebyshev coefficients of sweeps [0, k): PPR (ppr_solver.py:608-609) gives
// (beta, beta-1); Katz (katz_solver.py:366-368) gives (beta1, beta2).
static void cheb_coefs(const flsg_params *p, int64_t k, std::vector<double> &c) {
    c.resize(2 * k);
    if (p->kind == FLSG_PPR) {
        const double alpha = p->alpha;
        double delta_t = (1. - alpha);
        for (int64_t i = 0; i < k; i++) {
            delta_t = 1. / (2. / (1. - alpha) - delta_t);
            double beta = 2. * delta_t / (1. - alpha);
            c[2 * i] = beta;
            c[2 * i + 1] = beta - 1.;
        }
    } else {
        const double L = p->l_max, mu = p->mu_min;
        double delta_t = (L - mu) / (L + mu);
        for (int64_t i = 0; i < k; i++) {
            delta_t = 1. / (2. * ((L + mu) / (L - mu)) - delta_t);
            c[2 * i] = 4. * delta_t / (L - mu);
            c[2 * i + 1] = (1. - 2. * delta_t * (L + mu) / (L - mu));
        }
    }
}

// rows pulled by a wave / a block in the global sweeps; rebuilt after the
// degrees change
static int ensure_row_lists(flsg_graph g, hipStream_t st) {
    if (g->rows_valid) return FLSG_OK;
    const int64_t n = g->n;
    if (!g->rows_wave) {
        FLSG_HIP(hipMalloc((void **)&g->rows_wave, sizeof(int32_t) * (n + 1)));
        FLSG_HIP(hipMalloc((void **)&g->rows_block, sizeof(int32_t) * (n + 1)));
        FLSG_HIP(hipMalloc((void **)&g->rows_cnt, sizeof(unsigned int) * 2));
    }
    FLSG_HIP(hipMemsetAsync(g->rows_cnt, 0, sizeof(unsigned int) * 2, st));
    hipLaunchKernelGGL(k_row_classes, dim3(grid_for(g, 8)), dim3(kNT), 0, st, g->degree, n, g->rows_wave,
                       g->rows_block, g->rows_cnt);
    unsigned int h[2] = {0, 0};
    FLSG_HIP(hipMemcpyAsync(h, g->rows_cnt, sizeof(h), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    g->n_rows_wave = h[0];
    g->n_rows_block = h[1];
    g->rows_valid = 1;
    return FLSG_OK;
}

static int solve_global(flsg_graph g, const flsg_params *p, int64_t s, double *d_x, double *d_r,
                        flsg_result *res, hipStream_t st) {
    const int64_t n = g->n;
    const bool ppr = p->kind == FLSG_PPR;
    const bool cheb = p->method == FLSG_CHEB;
    const double alpha = p->alpha, om = 1.0 - p->alpha;
    const double thr_scale = ppr ? p->eps * p->alpha : p->eps;
    const int grid = grid_for(g, 8);
    const int64_t max_sweeps = p->max_iters > 0 ? p->max_iters : 100000;
    double *val = g->work_b, *dx = g->work_c;
    int rc0 = ensure_row_lists(g, st);
    if (rc0) return rc0;
    const int lgrid = grid_for(g, 2);
    FLSG_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, st));
    FLSG_HIP(hipMemsetAsync(d_r, 0, sizeof(double) * n, st));
    hipLaunchKernelGGL(k_init_global, dim3(1), dim3(1), 0, st, g->ctr);
    int32_t deg_s = 0;
    FLSG_HIP(hipMemcpyAsync(&deg_s, g->degree + s, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    if (deg_s <= 0) return fail(FLSG_EARG, "seed node has degree 0");
    std::vector<double> coef;
    int64_t coef_done = 0;
    if (cheb) {
        FLSG_HIP(hipMemsetAsync(dx, 0, sizeof(double) * n, st));
        if (ppr) {
            // r_s = alpha; x = dx = r; push of s replaces r (r_s -> 0)
            const double valv = (1. - alpha) * alpha / (double)deg_s;
            hipLaunchKernelGGL(k_cheb_init, dim3(1), dim3(1), 0, st, g->indptr, g->indices, g->degree, d_x, dx, d_r,
                               s, alpha, 0.0, valv);
        } else {
            const double L = p->l_max, mu = p->mu_min;
            hipLaunchKernelGGL(k_cheb_init, dim3(1), dim3(1), 0, st, g->indptr, g->indices, g->degree, d_x, dx, d_r,
                               s, 2. / (L + mu), 1. - 2. / (L + mu), 2. * alpha / (L + mu));
        }
    } else {
        const double r0 = ppr ? alpha : 1.0;
        FLSG_HIP(hipMemcpyAsync(d_r + s, &r0, sizeof(double), hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(k_count_active, dim3(grid), dim3(kNT), 0, st, g->degree, d_r, n, thr_scale, g->ctr);
    int64_t k = 0;
    for (;;) {
        if (cheb && k + kIterBlock > coef_done) {
            const int64_t want = coef_done + 4096;
            int rc = ensure_coef(g, want);
            if (rc) return rc;
            cheb_coefs(p, want, coef);
            FLSG_HIP(hipMemcpyAsync(g->coef, coef.data(), sizeof(double) * 2 * want, hipMemcpyHostToDevice, st));
            FLSG_HIP(hipStreamSynchronize(st));
            coef_done = want;
        }
        for (int b = 0; b < kIterBlock; b++, k++) {
            const int par = (int)(k & 1);
            if (!cheb) {
                if (ppr)
                    hipLaunchKernelGGL(k_gsweep_node<K_PPR>, dim3(grid), dim3(kNT), 0, st, g->degree, d_x, d_r, val, n,
                                       alpha, om, g->ctr, par);
                else
                    hipLaunchKernelGGL(k_gsweep_node<K_KATZ>, dim3(grid), dim3(kNT), 0, st, g->degree, d_x, d_r, val,
                                       n, alpha, om, g->ctr, par);
                hipLaunchKernelGGL(k_gsweep_pull<0>, dim3(grid), dim3(kNT), 0, st, g->indptr, g->indices, g->degree,
                                   val, d_r, dx, d_r, n, thr_scale, g->ctr, par);
                hipLaunchKernelGGL(k_gsweep_pull_long<0>, dim3(lgrid), dim3(kPB), 0, st, g->indptr, g->indices,
                                   g->degree, val, d_r, dx, d_r, g->rows_wave, g->n_rows_wave, g->rows_block,
                                   g->n_rows_block, thr_scale, g->ctr, par);
            } else if (ppr) {
                hipLaunchKernelGGL((k_csweep_node<K_PPR, 1>), dim3(grid), dim3(kNT), 0, st, g->degree, d_x, d_r, dx,
                                   val, n, g->coef, alpha, om, g->ctr, par, k);
                hipLaunchKernelGGL(k_gsweep_pull<1>, dim3(grid), dim3(kNT), 0, st, g->indptr, g->indices, g->degree,
                                   val, d_r, dx, d_r, n, thr_scale, g->ctr, par);
                hipLaunchKernelGGL(k_gsweep_pull_long<1>, dim3(lgrid), dim3(kPB), 0, st, g->indptr, g->indices,
                                   g->degree, val, d_r, dx, d_r, g->rows_wave, g->n_rows_wave, g->rows_block,
                                   g->n_rows_block, thr_scale, g->ctr, par);
            } else {
                hipLaunchKernelGGL((k_csweep_node<K_KATZ, 2>), dim3(grid), dim3(kNT), 0, st, g->degree, d_x, d_r, dx,
                                   val, n, g->coef, alpha, om, g->ctr, par, k);
                hipLaunchKernelGGL(k_gsweep_pull<2>, dim3(grid), dim3(kNT), 0, st, g->indptr, g->indices, g->degree,
                                   val, d_r, dx, d_r, n, thr_scale, g->ctr, par);
                hipLaunchKernelGGL(k_gsweep_pull_long<2>, dim3(lgrid), dim3(kPB), 0, st, g->indptr, g->indices,
                                   g->degree, val, d_r, dx, d_r, g->rows_wave, g->n_rows_wave, g->rows_block,
                                   g->n_rows_block, thr_scale, g->ctr, par);
            }
        }
        FLSG_HIP(hipGetLastError());
        FLSG_HIP(hipMemcpyAsync(g->h_ctr, g->ctr, sizeof(SolveCounters), hipMemcpyDeviceToHost, st));
        FLSG_HIP(hipStreamSynchronize(st));
        if (g->h_ctr->act[k & 1] == 0) break;
        if ((int64_t)g->h_ctr->iters >= max_sweeps) return fail(FLSG_ENOCONV, "sweep cap reached");
    }
    const int64_t sweeps = (int64_t)g->h_ctr->iters;
    res->iters = sweeps;
    res->oper = sweeps * (n + g->m_live);   // ppr_solver.py:320 (oper += n + m)
    if (!ppr && !cheb) hipLaunchKernelGGL(k_add_scalar, dim3(1), dim3(1), 0, st, d_x, s, -1.0);   // katz_solver.py:243
    return FLSG_OK;
}

// last heat-kernel level: val_u = r_u / d_u (hk_solver.py:271)
__global__ __launch_bounds__(kNT) void k_hk_last_val(const int32_t *degree, const double *r, double *val, int64_t n) {
    for (int64_t u = blockIdx.x * (int64_t)kNT + threadIdx.x; u < n; u += (int64_t)gridDim.x * kNT) {
        const int d = degree[u];
        val[u] = d > 0 ? r[u] / (double)d : 0.0;
    }
}

// hk_sor_queue(local=False), omega = 1 (hk_solver.py:265-279): one sweep per
// level.  Level j < N-1: x += r_j, r_{j+1}[v] = sum over the sorted row of v
// of ((tau/(j+1)) r_j[u]) / d_u (the natural-order scatter's order of adds,
// as in the GD sweeps); level N-1: x_v gets r_j[u] / d_u of its neighbours
// and its own r_j[v] in the scatter's order (pull mode 3).  With omega = 1 a
// sweep leaves the level's residual zero, so the reference's repeat test
// ends each level after one sweep once its isolated nodes are left out of
// it (their threshold is 0: the reference never ends on such a graph).
// oper = N (m + n); the returned two level residuals are zero, as the
// reference's are on return.
static int solve_hk_global(flsg_graph g, const flsg_params *p, int64_t s, int64_t N, double *d_x, double *d_r,
                           flsg_result *res, hipStream_t st) {
    const int64_t n = g->n;
    const int grid = grid_for(g, 8);
    const int lgrid = grid_for(g, 2);
    int rc = ensure_row_lists(g, st);
    if (rc) return rc;
    double *r = g->work_a, *val = g->work_b;
    FLSG_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, st));
    FLSG_HIP(hipMemsetAsync(d_r, 0, sizeof(double) * 2 * n, st));
    FLSG_HIP(hipMemsetAsync(r, 0, sizeof(double) * n, st));
    const double one = 1.0;
    FLSG_HIP(hipMemcpyAsync(r + s, &one, sizeof(double), hipMemcpyHostToDevice, st));   // hk_solver.py:223
    hipLaunchKernelGGL(k_init_global, dim3(1), dim3(1), 0, st, g->ctr);
    // the sweeps run while act[parity] > 0: with threshold 0 it counts every
    // node of degree > 0
    hipLaunchKernelGGL(k_count_active, dim3(grid), dim3(kNT), 0, st, g->degree, r, n, 0.0, g->ctr);
    for (int64_t j = 0; j < N; j++) {
        const int par = (int)(j & 1);
        if (j + 1 < N) {
            const double coef = p->tau / ((double)j + 1.);
            hipLaunchKernelGGL(k_gsweep_node<K_PPR>, dim3(grid), dim3(kNT), 0, st, g->degree, d_x, r, val, n, 0.0,
                               coef, g->ctr, par);
            hipLaunchKernelGGL(k_gsweep_pull<0>, dim3(grid), dim3(kNT), 0, st, g->indptr, g->indices, g->degree, val,
                               r, (const double *)nullptr, r, n, 0.0, g->ctr, par);
            hipLaunchKernelGGL(k_gsweep_pull_long<0>, dim3(lgrid), dim3(kPB), 0, st, g->indptr, g->indices, g->degree,
                               val, r, (const double *)nullptr, r, g->rows_wave, g->n_rows_wave, g->rows_block,
                               g->n_rows_block, 0.0, g->ctr, par);
        } else {
            hipLaunchKernelGGL(k_hk_last_val, dim3(grid), dim3(kNT), 0, st, g->degree, r, val, n);
            hipLaunchKernelGGL(k_gsweep_pull<3>, dim3(grid), dim3(kNT), 0, st, g->indptr, g->indices, g->degree, val,
                               r, (const double *)nullptr, d_x, n, 0.0, g->ctr, par);
            hipLaunchKernelGGL(k_gsweep_pull_long<3>, dim3(lgrid), dim3(kPB), 0, st, g->indptr, g->indices, g->degree,
                               val, r, (const double *)nullptr, d_x, g->rows_wave, g->n_rows_wave, g->rows_block,
                               g->n_rows_block, 0.0, g->ctr, par);
        }
        if ((j & 63) == 63) FLSG_HIP(hipGetLastError());
    }
    FLSG_HIP(hipGetLastError());
    // x /= e^t  (hk_solver.py:291)
    hipLaunchKernelGGL(k_scale_div, dim3(grid), dim3(kNT), 0, st, d_x, n, exp(p->tau));
    res->iters = N;
    res->oper = N * (g->m_live + n);   // hk_solver.py:275 (oper += m + n per sweep)
    return FLSG_OK;
}

int solve_local_cheb(flsg_graph g, const flsg_params *p, int64_t s, double *d_x, double *d_r, flsg_result *res,
                     flsg_trace *tr, hipStream_t st);   // flsg_cheb.hip

}  // namespace flsg

using namespace flsg;

extern "C" int flsg_solve(flsg_graph g, const flsg_params *p, int64_t seed, double *d_x, double *d_r,
                          flsg_result *h_res, flsg_trace *trace, void *stream) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (!p || !d_x || !d_r || !h_res) return fail(FLSG_EARG, "NULL argument");
    if (seed < 0 || seed >= g->n) return fail(FLSG_EARG, "seed out of range");
    if (p->precision != FLSG_FP64) return fail(FLSG_EARG, "single-source solves run the FP64 path");
    hipStream_t st = (hipStream_t)stream;
    rc = ensure_single_workspace(g);
    if (rc) return rc;
    int32_t deg_s = 0;
    FLSG_HIP(hipMemcpyAsync(&deg_s, g->degree + seed, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    if (deg_s <= 0) return fail(FLSG_EARG, "seed node has degree 0 (the reference divides by zero)");
    if (p->kind == FLSG_PPR && p->method == FLSG_GD && !(p->eps * (double)deg_s <= 1.))
        return fail(FLSG_EASSERT, "assert eps * degree[s_node] <= 1. failed");   // ppr_solver.py:213
    flsg_params q = *p;
    if (q.max_iters <= 0) q.max_iters = 100000;
    *h_res = flsg_result{};
    hipEvent_t e0, e1;
    FLSG_HIP(hipEventCreate(&e0));
    FLSG_HIP(hipEventCreate(&e1));
    FLSG_HIP(hipEventRecord(e0, st));
    if (q.kind == FLSG_HK) {
        int64_t est_n = 0;
        rc = solve_hk(g, &q, seed, d_x, d_r, h_res, trace, st, nullptr, 0, &est_n);
        h_res->n_corr = est_n;   // HK reports the Taylor order here
    } else if (q.kind == FLSG_PPR || q.kind == FLSG_KATZ) {
        if (!q.local) rc = solve_global(g, &q, seed, d_x, d_r, h_res, st);
        else if (q.method == FLSG_GD) rc = solve_local_gd(g, &q, seed, d_x, d_r, h_res, trace, st);
        else rc = solve_local_cheb(g, &q, seed, d_x, d_r, h_res, trace, st);
    } else {
        rc = fail(FLSG_EARG, "unknown problem kind");
    }
    hipError_t es = hipEventRecord(e1, st);
    if (es == hipSuccess) es = hipEventSynchronize(e1);   // also reports a fault of the solve's kernels
    float ms = 0.f;
    if (es == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == FLSG_OK && es != hipSuccess) rc = fail(FLSG_EHIP, std::string("solve: ") + hipGetErrorString(es));
    h_res->ms = ms;
    h_res->status = rc;
    return rc;
}

extern "C" int flsg_hk_solve(flsg_graph g, const flsg_params *p, int64_t seed, double *d_x, double *d_r,
                             flsg_result *h_res, double *h_lvl_sums, int64_t *est_n, flsg_trace *trace,
                             void *stream) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (!p || !d_x || !d_r || !h_res) return fail(FLSG_EARG, "NULL argument");
    if (seed < 0 || seed >= g->n) return fail(FLSG_EARG, "seed out of range");
    hipStream_t st = (hipStream_t)stream;
    rc = ensure_single_workspace(g);
    if (rc) return rc;
    int32_t deg_s = 0;
    FLSG_HIP(hipMemcpyAsync(&deg_s, g->degree + seed, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    if (deg_s <= 0) return fail(FLSG_EARG, "seed node has degree 0 (the reference divides by zero)");
    *h_res = flsg_result{};
    hipEvent_t e0, e1;
    FLSG_HIP(hipEventCreate(&e0));
    FLSG_HIP(hipEventCreate(&e1));
    FLSG_HIP(hipEventRecord(e0, st));
    if (h_lvl_sums && !est_n) return fail(FLSG_EARG, "level sums need their capacity in *est_n");
    const int64_t lvl_cap = h_lvl_sums ? *est_n : 0;
    rc = solve_hk(g, p, seed, d_x, d_r, h_res, trace, st, h_lvl_sums, lvl_cap, est_n);
    hipError_t es = hipEventRecord(e1, st);
    if (es == hipSuccess) es = hipEventSynchronize(e1);   // also reports a fault of the solve's kernels
    float ms = 0.f;
    if (es == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == FLSG_OK && es != hipSuccess) rc = fail(FLSG_EHIP, std::string("solve: ") + hipGetErrorString(es));
    h_res->ms = ms;
    h_res->status = rc;
    return rc;
}

// ------------------------------------------------------------------ allocation-free local push
// The local GD / Katz GD push of flsg_solve for callers that cannot allocate
// device memory: the same kernels, run directly on the caller's x / r (no
// captured graph, no handle workspace), with the frontier lists, push values,
// hub pieces, next-frontier sub-lists and counters in `scratch`, a file-scope
// device array of a fixed size (graphs of up to kNoAllocMaxN nodes and
// kNoAllocMaxNnz column entries).  Each GPU holds its own copy until the
// process ends; its address is looked up once per device.  Every call on a
// GPU shares it, so calls that overlap on one GPU race.
namespace flsg {

constexpr int64_t kNoAllocMaxN = (int64_t)1 << 20;
constexpr int64_t kNoAllocMaxNnz = (int64_t)1 << 24;
constexpr int64_t kNoAllocPieces = kNoAllocMaxNnz / 1024 + kNoAllocMaxN + 16;
constexpr int64_t kNoAllocFsubCap = (kNoAllocMaxN + kFrontSub - 1) / kFrontSub;

__host__ __device__ constexpr int64_t na_align(int64_t b) { return (b + 255) / 256 * 256; }
constexpr int64_t kNaFront0 = 0;
constexpr int64_t kNaFront1 = kNaFront0 + na_align(4 * (kNoAllocMaxN + 1));
constexpr int64_t kNaVal = kNaFront1 + na_align(4 * (kNoAllocMaxN + 1));
constexpr int64_t kNaPieces = kNaVal + na_align(8 * (kNoAllocMaxN + 1));
constexpr int64_t kNaFsub = kNaPieces + na_align(8 * kNoAllocPieces);
constexpr int64_t kNaFsubCnt = kNaFsub + na_align(4 * kFrontSub * kNoAllocFsubCap);
constexpr int64_t kNaCtr = kNaFsubCnt + na_align(4 * kFrontCntWords);
constexpr int64_t kNoAllocBytes = kNaCtr + na_align((int64_t)sizeof(SolveCounters));

}  // namespace flsg

static __device__ __align__(256) unsigned char scratch[flsg::kNoAllocBytes];

namespace flsg {

static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[64] = {};   // the array's device address, per device

static int scratch_ptr(int dev, unsigned char **out) {
    if (dev < 0 || dev >= 64) return fail(FLSG_EARG, "device id out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        FLSG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(FLSG_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[dev];
    return FLSG_OK;
}

__global__ void k_fill_zero2(double *a, double *b, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        a[i] = 0.0;
        b[i] = 0.0;
    }
}

template <int KIND>
static int noalloc_push(flsg_graph g, PushArgs a, int64_t max_iters, hipStream_t st, SolveCounters *h,
                        int64_t *iters_out) {
    const int grid = grid_for(g, 8);
    for (;;) {
        for (int b = 0; b < kIterBlock; b++) {
            const int p = b & 1;
            hipLaunchKernelGGL(k_node<KIND>, dim3(grid), dim3(kNT), 0, st, a, p);
            hipLaunchKernelGGL(k_push<KIND>, dim3(grid), dim3(kNT), 0, st, a, p);
            hipLaunchKernelGGL(k_front_compact, dim3(kFrontSub), dim3(kFrontSub), 0, st, a, p);
        }
        FLSG_HIP(hipGetLastError());
        FLSG_HIP(hipMemcpyAsync(h, a.ctr, sizeof(SolveCounters), hipMemcpyDeviceToHost, st));
        FLSG_HIP(hipStreamSynchronize(st));
        if (h->cnt[0] == 0) break;
        if ((int64_t)h->iters >= max_iters) {
            *iters_out = (int64_t)h->iters;
            return fail(FLSG_ENOCONV, "iteration cap reached");
        }
    }
    if (h->overflow & 2u) return fail(FLSG_ECAP, "hub piece list overflow");
    if (h->overflow & 4u) return fail(FLSG_ECAP, "next-frontier sub-list overflow");
    *iters_out = (int64_t)h->iters;
    return FLSG_OK;
}

}  // namespace flsg

extern "C" int flsg_solve_local_noalloc(flsg_graph g, const flsg_params *p, int64_t seed, double *d_x, double *d_r,
                                        flsg_result *h_res, void *stream) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (!p || !d_x || !d_r || !h_res) return fail(FLSG_EARG, "NULL argument");
    if ((p->kind != FLSG_PPR && p->kind != FLSG_KATZ) || p->method != FLSG_GD || !p->local ||
        p->precision != FLSG_FP64)
        return fail(FLSG_EARG, "allocation-free solve: local gradient descent, PPR or Katz, FP64");
    if (g->n > kNoAllocMaxN || g->nnz_cap > kNoAllocMaxNnz)
        return fail(FLSG_EARG, "allocation-free solve: the graph exceeds its fixed scratch (at most " +
                                   std::to_string(kNoAllocMaxN) + " nodes and " + std::to_string(kNoAllocMaxNnz) +
                                   " column entries)");
    if (seed < 0 || seed >= g->n) return fail(FLSG_EARG, "seed out of range");
    int dev = -1;
    FLSG_HIP(hipGetDevice(&dev));
    if (dev != g->device) return fail(FLSG_EARG, "allocation-free solve: make the graph's device current first");
    unsigned char *base = nullptr;
    rc = scratch_ptr(g->device, &base);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const auto t0 = std::chrono::steady_clock::now();
    int32_t deg_s = 0;
    FLSG_HIP(hipMemcpyAsync(&deg_s, g->degree + seed, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    if (deg_s <= 0) return fail(FLSG_EARG, "seed node has degree 0 (the reference divides by zero)");
    if (p->kind == FLSG_PPR && !(p->eps * (double)deg_s <= 1.))
        return fail(FLSG_EASSERT, "assert eps * degree[s_node] <= 1. failed");   // ppr_solver.py:213
    const int64_t n = g->n;
    PushArgs a{};
    a.indptr = g->indptr;
    a.indices = g->indices;
    a.degree = g->degree;
    a.front0 = reinterpret_cast<int32_t *>(base + kNaFront0);
    a.front1 = reinterpret_cast<int32_t *>(base + kNaFront1);
    a.val = reinterpret_cast<double *>(base + kNaVal);
    a.pieces = reinterpret_cast<int64_t *>(base + kNaPieces);
    a.pieces_cap = g->nnz_cap / 1024 + n + 16;
    a.ctr = reinterpret_cast<SolveCounters *>(base + kNaCtr);
    a.fsub = reinterpret_cast<int32_t *>(base + kNaFsub);
    a.fsub_cap = (n + kFrontSub - 1) / kFrontSub;
    a.fsubcnt = reinterpret_cast<unsigned int *>(base + kNaFsubCnt);
    a.x = d_x;
    a.r = d_r;
    a.r_push = d_r;
    a.alpha = p->alpha;
    a.om_alpha = 1.0 - p->alpha;
    const bool ppr = p->kind == FLSG_PPR;
    a.thr_scale = ppr ? p->eps * p->alpha : p->eps;
    // x and r start at zero; the counters and the seed's frontier entry are set
    // by k_init_local, and each iteration clears the sub-list counters it fills
    hipLaunchKernelGGL(k_fill_zero2, dim3(g->num_cus * 4), dim3(256), 0, st, d_x, d_r, n);
    hipLaunchKernelGGL(k_init_local, dim3(1), dim3(1), 0, st, a.ctr, a.front0, d_r, seed, ppr ? p->alpha : 1.0);
    FLSG_HIP(hipGetLastError());
    const int64_t max_iters = p->max_iters > 0 ? p->max_iters : 100000;
    SolveCounters h{};
    int64_t iters = 0;
    *h_res = flsg_result{};
    rc = ppr ? noalloc_push<K_PPR>(g, a, max_iters, st, &h, &iters) : noalloc_push<K_KATZ>(g, a, max_iters, st, &h, &iters);
    if (rc == FLSG_OK && !ppr) hipLaunchKernelGGL(k_add_scalar, dim3(1), dim3(1), 0, st, d_x, seed, -1.0);   // katz_solver.py:243
    const hipError_t es = hipStreamSynchronize(st);
    if (rc == FLSG_OK && es != hipSuccess) rc = fail(FLSG_EHIP, std::string("solve: ") + hipGetErrorString(es));
    h_res->oper = (int64_t)h.oper;
    h_res->iters = iters;
    h_res->nodes = (int64_t)h.trace_off;
    h_res->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    h_res->status = rc;
    return rc;
}

extern "C" int64_t flsg_solve_local_noalloc_scratch_bytes(void) { return flsg::kNoAllocBytes; }

extern "C" double flsg_hk_estimate_coefficient(double t, int64_t k) { return flsg::host_estimate_coefficient(t, k); }

extern "C" int64_t flsg_hk_estimate_n(double t, double eps) { return flsg::host_estimate_n(t, eps); }
