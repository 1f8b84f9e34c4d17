// flsg_ignn.hip -- InstantGNN feature propagation, all feature columns at once.
//
// Reference: instantGNNPush / instantGNNParallel (InstantGNN.py:12-71) run a
// FIFO push per feature column:
//   p_u += alpha r_u;  r_v += (1-alpha) r_u / (d_u^(1-beta) d_v^beta)  (v in N(u));  r_u = 0
// until |r_u| < eps_u = eps d_u^(1-beta) for every u, with the special case
// "degree 0 or a lone self-loop: p_u = r_u, r_u = 0".  The push keeps the
// invariant  p + alpha (I - (1-alpha) M)^-1 r = alpha (I - (1-alpha) M)^-1 X,
// M_vu = 1 / (d_v^beta d_u^(1-beta)), so any push order ends within the same
// certificate; this file runs it as synchronous rounds over a node frontier
// (a node is active if any of its columns is above threshold).
//
// Layout: P, R are n x C row-major (node-major), so one (u -> v) push moves
// C contiguous values: a wavefront handles one edge for 64 columns and its
// atomics are fully coalesced.  C = 1 is the per-column plain path (a lane per
// edge).  FP32 rounds are followed by an FP64 certificate: the residual is
// recomputed as R = X - P/alpha + ((1-alpha)/alpha) M P, and FP64 rounds push
// until |R| < eps_u holds in FP64.
//
// Dynamic updates (instantGNNUpdateBatch, InstantGNN.py:298-329): after edges
// are inserted (flsg_graph_insert_edges), p and r of every endpoint are
// rescaled / corrected so the invariant holds on the new graph, then pushed.
#include <math.h>

#include <vector>

#include "flsg_internal.h"

namespace flsg {

constexpr int kIT = 256;

struct IgnnWork {
    int64_t n = 0;
    int64_t cols = 0;
    double *dpb = nullptr;       // d^beta
    double *dpbi = nullptr;      // 1 / d^beta (0 for d = 0)
    double *dp1b = nullptr;      // d^(1-beta)
    uint32_t *pcs = nullptr;     // pieces of long frontier rows: (entry << 8) | piece
    unsigned int *pcnt = nullptr;  // [0] pieces, [1] work items claimed
    double beta = -1.0;
    int32_t *front[2] = {nullptr, nullptr};
    uint32_t *flag = nullptr;
    float *P32 = nullptr;        // n x C (FP32 rounds)
    float *R32 = nullptr;
    double *V = nullptr;         // per (frontier entry, column) push value, n x C doubles
    unsigned int *cnt = nullptr; // [2]
    unsigned long long *oper = nullptr;
    double *maxr = nullptr;
};

struct IgnnArgs {
    const int32_t *indptr;
    const int32_t *indices;
    const int32_t *degree;
    const double *dpb;
    const double *dpbi;
    const double *dp1b;
    uint32_t *pcs;
    unsigned int *pcnt;
    int64_t n;
    int cols;
    double alpha;
    double eps;
    uint32_t *flag;
    unsigned long long *oper;
};

// the reference's special case (InstantGNN.py:23-26)
__device__ __forceinline__ bool ignn_special(const IgnnArgs &a, int u, int d) {
    return d == 0 || (d == 1 && a.indices[a.indptr[u]] == u);
}

__global__ void k_ignn_pow(const int32_t *degree, int64_t n, double beta, double *dpb, double *dpbi, double *dp1b) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double d = (double)degree[i];
        const double b = pow(d, beta);
        dpb[i] = b;
        dpbi[i] = b > 0.0 ? 1.0 / b : 0.0;
        dp1b[i] = pow(d, 1. - beta);
    }
}

// rows longer than this are split into pieces of this many edges (one
// wavefront each) in the wide edge phase
constexpr int kPieceI = 256;

// frontier = nodes (all, or the listed ones) with some |R[u][c]| >= eps_u
template <typename T>
__global__ __launch_bounds__(kIT) void k_ignn_seed(IgnnArgs a, const T *R, const int32_t *list, int64_t nlist,
                                                   int32_t *front, unsigned int *cnt) {
    const int64_t total = list ? nlist : a.n;
    for (int64_t base = blockIdx.x * (int64_t)kIT; base < total; base += (int64_t)gridDim.x * kIT) {
        const int64_t i = base + threadIdx.x;
        bool act = false;
        int u = 0;
        if (i < total) {
            u = list ? list[i] : (int)i;
            const double thr = a.dp1b[u] * a.eps;
            for (int c = 0; c < a.cols && !act; c++) act = fabs((double)R[(int64_t)u * a.cols + c]) >= thr;
            if (act) act = atomicOr(&a.flag[u], 1u) == 0u;
        }
        const long long slot = wave_append(act, cnt);
        if (slot >= 0) front[slot] = u;
    }
}

// node phase: one (frontier entry, column) per thread
template <typename T>
__global__ __launch_bounds__(kIT) void k_ignn_node(IgnnArgs a, const int32_t *front, unsigned int k, T *P, T *R,
                                                   double *V) {
    __shared__ unsigned long long s_ops[kIT / kWave];
    const int C = a.cols;
    const int64_t total = (int64_t)k * C;
    unsigned long long ops = 0;
    for (int64_t t = blockIdx.x * (int64_t)kIT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kIT) {
        const int64_t i = t / C;
        const int c = (int)(t - i * C);
        const int u = front[i];
        const int d = a.degree[u];
        if (c == 0) {
            a.flag[u] = 0u;
            if (a.pcs && d > kPieceI && !ignn_special(a, u, d)) {
                const unsigned int np = (unsigned int)((d + kPieceI - 1) / kPieceI);
                const unsigned int b = atomicAdd(&a.pcnt[0], np);
                for (unsigned int q = 0; q < np; q++) a.pcs[b + q] = ((uint32_t)i << 8) | q;
            }
        }
        const int64_t o = (int64_t)u * C + c;
        const double r = (double)R[o];
        double v = 0.0;
        if (ignn_special(a, u, d)) {
            if (fabs(r) >= a.dp1b[u] * a.eps) {     // InstantGNN.py:20-26
                P[o] = (T)r;
                R[o] = (T)0;
            }
        } else if (fabs(r) >= a.dp1b[u] * a.eps) {
            P[o] = (T)((double)P[o] + a.alpha * r);             // InstantGNN.py:27-28
            v = (1. - a.alpha) * r / a.dp1b[u];                 // InstantGNN.py:29
            R[o] = (T)0;
            ops += (unsigned long long)d;
        }
        V[t] = v;
    }
    ops = wave_sum(ops);
    if (__lane_id() == 0) s_ops[threadIdx.x / kWave] = ops;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int w = 0; w < kIT / kWave; w++) s += s_ops[w];
        if (s) atomicAdd(a.oper, s);
    }
}

// edge phase: frontier edges flattened per chunk of kIT entries; a group of
// `lanes` = min(C, 64) lanes handles one edge (columns strided by lanes).
template <typename T>
__global__ __launch_bounds__(kIT) void k_ignn_push(IgnnArgs a, const int32_t *front, unsigned int k, T *R,
                                                   const double *V, int32_t *next, unsigned int *ncnt) {
    __shared__ int s_incl[kIT];
    __shared__ int s_start[kIT];
    __shared__ int s_scan[kIT / kWave];
    const int C = a.cols;
    int lanes = kWave;                          // power of two <= min(C, 64)
    while (lanes > C) lanes >>= 1;
    const int groups = kIT / lanes;             // edges in flight per block
    const int g = threadIdx.x / lanes;
    const int l = threadIdx.x - g * lanes;
    for (int64_t base = blockIdx.x * (int64_t)kIT; base < k; base += (int64_t)gridDim.x * kIT) {
        const int64_t i = base + threadIdx.x;
        int d = 0, st = 0;
        if (i < k) {
            const int u = front[i];
            d = a.degree[u];
            st = a.indptr[u];
            if (ignn_special(a, u, d)) d = 0;
        }
        int tot;
        const int incl = block_incl_scan<kIT>(d, s_scan, &tot);
        s_incl[threadIdx.x] = incl;
        s_start[threadIdx.x] = st;
        __syncthreads();
        for (int e0 = 0; e0 < tot; e0 += groups) {
            const int e = e0 + g;
            bool hot = false;
            int v = 0;
            if (e < tot) {
                const int kk = upper_bound_lds(s_incl, 0, kIT, e);
                const int off = e - (kk ? s_incl[kk - 1] : 0);
                v = a.indices[s_start[kk] + off];
                const int64_t ent = base + kk;
                const double inv = 1.0 / a.dpb[v];
                const double thr = a.dp1b[v] * a.eps;
                for (int c = l; c < C; c += lanes) {
                    const double pv = V[ent * C + c];
                    if (pv != 0.0) {
                        const T add = (T)(pv * inv);                     // InstantGNN.py:32
                        const T old = atomic_add_dev(&R[(int64_t)v * C + c], add);
                        const double nw = (double)old + (double)add;
                        hot = hot || fabs(nw) >= thr;
                    }
                }
            }
            // one flag per edge group: any lane saw a value at/above threshold
            bool enq = false;
            const unsigned long long m = __ballot(hot);
            if (e < tot && l == 0) {
                const int shift = (threadIdx.x % kWave);
                const unsigned long long gm = (lanes >= kWave) ? ~0ull : (((1ull << lanes) - 1) << shift);
                if (m & gm) enq = atomicOr(&a.flag[v], 1u) == 0u;
            }
            const long long slot = wave_append(enq, ncnt);
            if (slot >= 0) next[slot] = v;
        }
        __syncthreads();
    }
}

// Wide edge phase (C >= 32): one wavefront per frontier row (rows longer
// than kPieceI edges in pieces), claimed through a work counter so every
// wavefront of the grid shares the frontier whatever its size.  Lane l owns
// columns l, l+64, ... of the row's push values; a wavefront issues the column
// atomics of kUI edges before it consumes any result, and a target joins the
// next frontier when one of its columns reached its threshold (first setter
// of its flag appends it).
constexpr int kUI = 4;

template <typename T>
__global__ __launch_bounds__(kIT) void k_ignn_push_wide(IgnnArgs a, const int32_t *front, unsigned int k, T *R,
                                                        const double *V, int32_t *next, unsigned int *ncnt) {
    const int lane = __lane_id();
    const int C = a.cols;
    const unsigned int nitems = k + a.pcnt[0];
    for (;;) {
        unsigned int item = 0;
        if (lane == 0) item = atomicAdd(&a.pcnt[1], 1u);
        item = __shfl(item, 0);
        if (item >= nitems) break;
        unsigned int i;
        int beg = 0;
        if (item < k) {
            i = item;
        } else {
            const uint32_t pc = a.pcs[item - k];
            i = pc >> 8;
            beg = (int)(pc & 255u) * kPieceI;
        }
        const int u = front[i];
        const int d = a.degree[u];
        if (ignn_special(a, u, d)) continue;
        if (item < k && d > kPieceI) continue;      // done by its pieces
        const int end = item < k ? d : min(d, beg + kPieceI);
        const int st = a.indptr[u];
        const double *Vi = V + (int64_t)i * C;
        for (int cb = 0; cb < C; cb += 2 * kWave) {
            const int c0 = cb + lane, c1 = cb + kWave + lane;
            const double pv0 = c0 < C ? Vi[c0] : 0.0;
            const double pv1 = c1 < C ? Vi[c1] : 0.0;
            if (__ballot(pv0 != 0.0 || pv1 != 0.0) == 0ull) continue;
            for (int e0 = beg; e0 < end; e0 += kUI) {
                int vv[kUI];
                double inv[kUI], thr[kUI];
#pragma unroll
                for (int q = 0; q < kUI; q++) {
                    const int e = e0 + q;
                    vv[q] = e < end ? a.indices[st + e] : -1;
                }
#pragma unroll
                for (int q = 0; q < kUI; q++) {
                    inv[q] = vv[q] >= 0 ? a.dpbi[vv[q]] : 0.0;
                    thr[q] = vv[q] >= 0 ? a.dp1b[vv[q]] * a.eps : 0.0;
                }
                T add0[kUI], add1[kUI], o0[kUI], o1[kUI];
#pragma unroll
                for (int q = 0; q < kUI; q++) {
                    add0[q] = (T)(pv0 * inv[q]);                          // InstantGNN.py:32
                    add1[q] = (T)(pv1 * inv[q]);
                    o0[q] = (T)0;
                    o1[q] = (T)0;
                }
#pragma unroll
                for (int q = 0; q < kUI; q++) {
                    if (vv[q] >= 0 && pv0 != 0.0) o0[q] = atomic_add_dev(&R[(int64_t)vv[q] * C + c0], add0[q]);
                    if (vv[q] >= 0 && pv1 != 0.0) o1[q] = atomic_add_dev(&R[(int64_t)vv[q] * C + c1], add1[q]);
                }
#pragma unroll
                for (int q = 0; q < kUI; q++) {
                    const bool hot = vv[q] >= 0 &&
                                     ((pv0 != 0.0 && fabs((double)o0[q] + (double)add0[q]) >= thr[q]) ||
                                      (pv1 != 0.0 && fabs((double)o1[q] + (double)add1[q]) >= thr[q]));
                    if (__ballot(hot) != 0ull && lane == 0 && atomicOr(&a.flag[vv[q]], 1u) == 0u)
                        next[atomicAdd(ncnt, 1u)] = vv[q];
                }
            }
        }
    }
}

// FP64 residual recompute of the invariant: R = X - P/alpha + ((1-alpha)/alpha) M P
// (pull over the live row of v).  Special nodes: R = X - P.
__global__ __launch_bounds__(kIT) void k_ignn_recompute(IgnnArgs a, const double *X, const double *P, double *R) {
    const int C = a.cols;
    const int64_t total = a.n * C;
    const double ca = (1. - a.alpha) / a.alpha;
    for (int64_t t = blockIdx.x * (int64_t)kIT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kIT) {
        const int64_t v = t / C;
        const int c = (int)(t - v * C);
        const int d = a.degree[v];
        if (ignn_special(a, (int)v, d)) {
            R[t] = X[t] - P[t];
            continue;
        }
        const int st = a.indptr[v];
        double s = 0.0;
        for (int j = 0; j < d; j++) {
            const int u = a.indices[st + j];
            s += P[(int64_t)u * C + c] / a.dp1b[u];
        }
        R[t] = X[t] - P[t] / a.alpha + ca * s / a.dpb[v];
    }
}

__global__ void k_f64_to_f32(const double *in, float *out, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

__global__ void k_f32_to_f64(const float *in, double *out, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (double)in[i];
}

// max over nodes of max_c |R[u][c]| / eps_u  (certificate value)
__global__ __launch_bounds__(kIT) void k_ignn_maxratio(IgnnArgs a, const double *R, double *out) {
    __shared__ double s_m[kIT / kWave];
    const int C = a.cols;
    const int64_t total = a.n * C;
    double m = 0.0;
    for (int64_t t = blockIdx.x * (int64_t)kIT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kIT) {
        const int64_t v = t / C;
        const double thr = a.dp1b[v] * a.eps;
        if (thr > 0.0) m = fmax(m, fabs(R[t]) / thr);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if (__lane_id() == 0) s_m[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kIT / kWave; w++) t = fmax(t, s_m[w]);
        // non-negative doubles order like their bit patterns
        atomicMax((unsigned long long *)out, (unsigned long long)__double_as_longlong(t));
    }
}

// dynamic update adjustment (InstantGNN.py:303-318), one (key, column) per
// thread.  New neighbours of u are the row tail indices[indptr[u]+old_deg[u] :
// indptr[u]+deg[u]] written by the insert kernel.  Two passes: scale (all keys)
// then correct.
__global__ __launch_bounds__(kIT) void k_ignn_update_scale(IgnnArgs a, const int32_t *keys, int64_t nk,
                                                           const int32_t *old_deg, double beta, double *P, double *R) {
    const int C = a.cols;
    const int64_t total = nk * C;
    for (int64_t t = blockIdx.x * (int64_t)kIT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kIT) {
        const int64_t i = t / C;
        const int c = (int)(t - i * C);
        const int u = keys[i];
        const double du = (double)a.degree[u], od = (double)old_deg[u];
        const int64_t o = (int64_t)u * C + c;
        const double pu = P[o] * pow(du / od, 1. - beta);
        P[o] = pu;
        R[o] = R[o] + pu * (pow(od, 1. - beta) - pow(du, 1. - beta)) / pow(du, 1. - beta) * 1. / a.alpha;
    }
}

__global__ __launch_bounds__(kIT) void k_ignn_update_correct(IgnnArgs a, const int32_t *keys, int64_t nk,
                                                             const int32_t *old_deg, double beta, const double *P,
                                                             double *R, const double *X) {
    const int C = a.cols;
    const int64_t total = nk * C;
    for (int64_t t = blockIdx.x * (int64_t)kIT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kIT) {
        const int64_t i = t / C;
        const int c = (int)(t - i * C);
        const int u = keys[i];
        const double du = (double)a.degree[u], od = (double)old_deg[u];
        const int64_t o = (int64_t)u * C + c;
        double dr = (P[o] + a.alpha * R[o] - a.alpha * X[o]) * (pow(od, beta) - pow(du, beta)) / pow(du, beta);
        const int st = a.indptr[u];
        for (int j = old_deg[u]; j < a.degree[u]; j++) {
            const int v = a.indices[st + j];
            dr += (1. - a.alpha) * P[(int64_t)v * C + c] / (pow(du, beta) * pow((double)a.degree[v], 1. - beta));
        }
        R[o] = R[o] + dr / a.alpha;
    }
}

static void free_ignn(IgnnWork *w) {
    if (!w) return;
    (void)hipFree(w->dpb);
    (void)hipFree(w->dpbi);
    (void)hipFree(w->dp1b);
    (void)hipFree(w->pcs);
    (void)hipFree(w->pcnt);
    (void)hipFree(w->front[0]);
    (void)hipFree(w->front[1]);
    (void)hipFree(w->flag);
    (void)hipFree(w->P32);
    (void)hipFree(w->R32);
    (void)hipFree(w->V);
    (void)hipFree(w->cnt);
    (void)hipFree(w->oper);
    (void)hipFree(w->maxr);
    delete w;
}

void ignn_free(flsg_graph g) {
    free_ignn((IgnnWork *)g->ignn);
    g->ignn = nullptr;
}

static int ensure_ignn(flsg_graph g, int64_t cols, bool fp32, hipStream_t st) {
    IgnnWork *w = (IgnnWork *)g->ignn;
    const int64_t n = g->n;
    if (w && (w->cols != cols || (fp32 && !w->P32))) {
        free_ignn(w);
        w = nullptr;
        g->ignn = nullptr;
    }
    if (!w) {
        w = new IgnnWork();
        g->ignn = w;
        w->n = n;
        w->cols = cols;
        FLSG_HIP(hipMalloc((void **)&w->dpb, sizeof(double) * n));
        FLSG_HIP(hipMalloc((void **)&w->dpbi, sizeof(double) * n));
        FLSG_HIP(hipMalloc((void **)&w->dp1b, sizeof(double) * n));
        // frontier rows are distinct nodes: at most n + nnz/kPieceI pieces
        FLSG_HIP(hipMalloc((void **)&w->pcs, sizeof(uint32_t) * (size_t)(n + g->nnz_cap / kPieceI + 1)));
        FLSG_HIP(hipMalloc((void **)&w->pcnt, sizeof(unsigned int) * 2));
        FLSG_HIP(hipMalloc((void **)&w->front[0], sizeof(int32_t) * (n + 1)));
        FLSG_HIP(hipMalloc((void **)&w->front[1], sizeof(int32_t) * (n + 1)));
        FLSG_HIP(hipMalloc((void **)&w->flag, sizeof(uint32_t) * n));
        FLSG_HIP(hipMemsetAsync(w->flag, 0, sizeof(uint32_t) * n, st));
        FLSG_HIP(hipMalloc((void **)&w->V, sizeof(double) * n * cols));
        FLSG_HIP(hipMalloc((void **)&w->cnt, sizeof(unsigned int) * 2));
        FLSG_HIP(hipMalloc((void **)&w->oper, sizeof(unsigned long long)));
        FLSG_HIP(hipMalloc((void **)&w->maxr, sizeof(double)));
        if (fp32) {
            FLSG_HIP(hipMalloc((void **)&w->P32, sizeof(float) * n * cols));
            FLSG_HIP(hipMalloc((void **)&w->R32, sizeof(float) * n * cols));
        }
    }
    return FLSG_OK;
}

static IgnnArgs ignn_args(flsg_graph g, IgnnWork *w, const flsg_ignn_params *p) {
    IgnnArgs a{};
    a.indptr = g->indptr;
    a.indices = g->indices;
    a.degree = g->degree;
    a.dpb = w->dpb;
    a.dpbi = w->dpbi;
    a.dp1b = w->dp1b;
    a.pcs = p->cols >= 32 ? w->pcs : nullptr;   // wide edge phase only
    a.pcnt = w->pcnt;
    a.n = g->n;
    a.cols = (int)p->cols;
    a.alpha = p->alpha;
    a.eps = p->eps;
    a.flag = w->flag;
    a.oper = w->oper;
    return a;
}

// rounds until the frontier is empty; the frontier must already be in front[0]
// with its size in cnt[0] (and flags set for its members)
template <typename T>
static int ignn_rounds(flsg_graph g, IgnnWork *w, const IgnnArgs &a, T *P, T *R, int64_t max_rounds,
                       hipStream_t st, int64_t *rounds_out) {
    const int grid = g->num_cus * 8;
    unsigned int k = 0;
    FLSG_HIP(hipMemcpyAsync(&k, w->cnt, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    int cur = 0;
    int64_t rounds = 0;
    while (k > 0) {
        if (rounds >= max_rounds) return fail(FLSG_ENOCONV, "InstantGNN push: round cap reached");
        FLSG_HIP(hipMemsetAsync(w->cnt + (cur ^ 1), 0, sizeof(unsigned int), st));
        FLSG_HIP(hipMemsetAsync(w->pcnt, 0, sizeof(unsigned int) * 2, st));
        hipLaunchKernelGGL(k_ignn_node<T>, dim3(grid), dim3(kIT), 0, st, a, w->front[cur], k, P, R, w->V);
        if (a.pcs)
            hipLaunchKernelGGL(k_ignn_push_wide<T>, dim3(grid), dim3(kIT), 0, st, a, w->front[cur], k, R,
                               (const double *)w->V, w->front[cur ^ 1], w->cnt + (cur ^ 1));
        else
            hipLaunchKernelGGL(k_ignn_push<T>, dim3(grid), dim3(kIT), 0, st, a, w->front[cur], k, R,
                               (const double *)w->V, w->front[cur ^ 1], w->cnt + (cur ^ 1));
        FLSG_HIP(hipGetLastError());
        FLSG_HIP(hipMemcpyAsync(&k, w->cnt + (cur ^ 1), sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        FLSG_HIP(hipStreamSynchronize(st));
        cur ^= 1;
        rounds++;
    }
    // the final frontier lives in front[cur]; move it to front[0] position for callers
    *rounds_out = rounds;
    return FLSG_OK;
}

}  // namespace flsg

using namespace flsg;

extern "C" int flsg_ignn_propagate(flsg_graph g, const flsg_ignn_params *p, const double *d_X, double *d_P,
                                   double *d_R, const int32_t *d_nodes, int64_t n_nodes, flsg_result *h_res,
                                   void *stream) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (!p || !d_P || !d_R || !h_res) return fail(FLSG_EARG, "NULL argument");
    if (p->cols <= 0 || p->cols > 4096) return fail(FLSG_EARG, "cols out of range");
    const bool fp32 = p->precision == FLSG_FP32;
    if (fp32 && !d_X) return fail(FLSG_EARG, "the FP32 path needs X for its FP64 certificate");
    hipStream_t st = (hipStream_t)stream;
    rc = ensure_ignn(g, p->cols, fp32, st);
    if (rc) return rc;
    IgnnWork *w = (IgnnWork *)g->ignn;
    const int64_t n = g->n, NC = g->n * p->cols;
    const int grid = g->num_cus * 8;
    if (w->beta != p->beta || p->degrees_changed) {
        hipLaunchKernelGGL(k_ignn_pow, dim3(grid), dim3(kIT), 0, st, g->degree, n, p->beta, w->dpb, w->dpbi, w->dp1b);
        w->beta = p->beta;
    }
    const int64_t max_rounds = p->max_rounds > 0 ? p->max_rounds : 1000000;
    IgnnArgs a = ignn_args(g, w, p);
    *h_res = flsg_result{};
    hipEvent_t e0, e1;
    FLSG_HIP(hipEventCreate(&e0));
    FLSG_HIP(hipEventCreate(&e1));
    FLSG_HIP(hipEventRecord(e0, st));
    FLSG_HIP(hipMemsetAsync(w->oper, 0, sizeof(unsigned long long), st));
    int64_t rounds = 0, rounds64 = 0;
    unsigned long long oper32 = 0;
    if (fp32) {
        hipLaunchKernelGGL(k_f64_to_f32, dim3(grid), dim3(kIT), 0, st, (const double *)d_P, w->P32, NC);
        hipLaunchKernelGGL(k_f64_to_f32, dim3(grid), dim3(kIT), 0, st, (const double *)d_R, w->R32, NC);
        FLSG_HIP(hipMemsetAsync(w->cnt, 0, sizeof(unsigned int) * 2, st));
        hipLaunchKernelGGL(k_ignn_seed<float>, dim3(grid), dim3(kIT), 0, st, a, (const float *)w->R32, d_nodes,
                           n_nodes, w->front[0], w->cnt);
        rc = ignn_rounds<float>(g, w, a, w->P32, w->R32, max_rounds, st, &rounds);
        if (rc) return rc;
        FLSG_HIP(hipMemcpyAsync(&oper32, w->oper, sizeof(oper32), hipMemcpyDeviceToHost, st));
        hipLaunchKernelGGL(k_f32_to_f64, dim3(grid), dim3(kIT), 0, st, (const float *)w->P32, d_P, NC);
        hipLaunchKernelGGL(k_ignn_recompute, dim3(grid), dim3(kIT), 0, st, a, d_X, (const double *)d_P, d_R);
        FLSG_HIP(hipMemsetAsync(w->oper, 0, sizeof(unsigned long long), st));
        FLSG_HIP(hipMemsetAsync(w->cnt, 0, sizeof(unsigned int) * 2, st));
        hipLaunchKernelGGL(k_ignn_seed<double>, dim3(grid), dim3(kIT), 0, st, a, (const double *)d_R,
                           (const int32_t *)nullptr, (int64_t)0, w->front[0], w->cnt);
        rc = ignn_rounds<double>(g, w, a, d_P, d_R, max_rounds, st, &rounds64);
        if (rc) return rc;
    } else {
        FLSG_HIP(hipMemsetAsync(w->cnt, 0, sizeof(unsigned int) * 2, st));
        hipLaunchKernelGGL(k_ignn_seed<double>, dim3(grid), dim3(kIT), 0, st, a, (const double *)d_R, d_nodes,
                           n_nodes, w->front[0], w->cnt);
        rc = ignn_rounds<double>(g, w, a, d_P, d_R, max_rounds, st, &rounds);
        if (rc) return rc;
    }
    FLSG_HIP(hipMemsetAsync(w->maxr, 0, sizeof(double), st));
    hipLaunchKernelGGL(k_ignn_maxratio, dim3(grid), dim3(kIT), 0, st, a, (const double *)d_R, w->maxr);
    unsigned long long oper64 = 0;
    double cert = 0.0;
    FLSG_HIP(hipMemcpyAsync(&oper64, w->oper, sizeof(oper64), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipMemcpyAsync(&cert, w->maxr, sizeof(double), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipEventRecord(e1, st));
    FLSG_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h_res->oper = fp32 ? (int64_t)oper32 : (int64_t)oper64;
    h_res->oper_cert = fp32 ? (int64_t)oper64 : 0;
    h_res->iters = rounds;
    h_res->n_corr = rounds64;
    h_res->cert = cert;
    h_res->ms = ms;
    h_res->status = FLSG_OK;
    return FLSG_OK;
}

extern "C" int flsg_ignn_update(flsg_graph g, const flsg_ignn_params *p, const double *d_X, double *d_P,
                                double *d_R, const int32_t *d_keys, int64_t n_keys, const int32_t *d_old_degree,
                                void *stream) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (!p || !d_X || !d_P || !d_R || (n_keys > 0 && (!d_keys || !d_old_degree)))
        return fail(FLSG_EARG, "NULL argument");
    hipStream_t st = (hipStream_t)stream;
    rc = ensure_ignn(g, p->cols, p->precision == FLSG_FP32, st);
    if (rc) return rc;
    IgnnWork *w = (IgnnWork *)g->ignn;
    const int grid = g->num_cus * 8;
    hipLaunchKernelGGL(k_ignn_pow, dim3(grid), dim3(kIT), 0, st, g->degree, g->n, p->beta, w->dpb, w->dpbi, w->dp1b);
    w->beta = p->beta;
    if (n_keys <= 0) return FLSG_OK;
    IgnnArgs a = ignn_args(g, w, p);
    hipLaunchKernelGGL(k_ignn_update_scale, dim3(grid), dim3(kIT), 0, st, a, d_keys, n_keys, d_old_degree, p->beta,
                       d_P, d_R);
    hipLaunchKernelGGL(k_ignn_update_correct, dim3(grid), dim3(kIT), 0, st, a, d_keys, n_keys, d_old_degree,
                       p->beta, (const double *)d_P, d_R, d_X);
    FLSG_HIP(hipGetLastError());
    return FLSG_OK;
}
