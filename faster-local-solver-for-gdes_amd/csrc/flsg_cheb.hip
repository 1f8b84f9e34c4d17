// flsg_cheb.hip -- local Chebyshev (PPR and Katz) with the reference's exact
// sequential semantics, on the GPU.
//
// The reference's local Chebyshev (ppr_solver.py:619-657, katz_solver.py:384-427)
// walks its queue u_0..u_{k-1} in order; for each u_i it adds val_i to r_v for
// every v in row(u_i) (checking |r_v| >= thr_v after each add, enqueueing v on
// its first crossing), then subtracts dx_{u_i} from r_{u_i} (no check).
// Residuals change sign, so which nodes cross -- and the order of the next
// queue -- depends on that sequence.  Per iteration this file reproduces it:
//   1. queue update (parallel): t = b1 r_u (+|-) b2 dx_u, x_u += t, dx_u = t,
//      push value, and the operation count;
//   2. events: entry i owns positions P_i .. P_i + d_i of a flat event space
//      (P = exclusive scan of d_i + 1): position P_i + j is "add val_i to row
//      entry j", position P_i + d_i is "subtract dx_{u_i} from r_{u_i}";
//      key = (destination << pbits) | position;
//   3. radix sort of the keys (hipCUB) groups every destination's events in
//      the reference's order;
//   4. one thread per destination replays its events sequentially in FP64
//      (bit-identical sums) and records the position of the first crossing;
//   5. the crossing nodes, sorted by that position, are the next queue (or,
//      when there are more than n/4 of them, ascending ids, ppr_solver.py:639).
#include <math.h>

#include <hipcub/hipcub.hpp>
#include <vector>

#include "flsg_internal.h"

namespace flsg {

constexpr int kCT = 256;

struct ChebWork {
    int64_t cap = 0;               // events capacity (m + n)
    uint64_t *keys[2] = {nullptr, nullptr};
    int32_t *vals[2] = {nullptr, nullptr};
    int64_t *P = nullptr;          // n + 1
    int32_t *dplus = nullptr;      // n
    uint64_t *cross[2] = {nullptr, nullptr};   // n
    int32_t *queue = nullptr;      // n
    unsigned int *cnt = nullptr;   // crossing count
    unsigned long long *oper = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    double *dx = nullptr;          // n
    double *val = nullptr;         // n
};

static int bits_for(uint64_t v) {
    int b = 1;
    while (b < 63 && (1ull << b) <= v) b++;
    return b;
}

struct ChebCoefArgs {
    const int32_t *degree;
    const int32_t *queue;
    double *x;
    double *r;
    double *dx;
    double *val;
    int32_t *dplus;
    unsigned long long *oper;
    double b1, b2, alpha, om_alpha;
    int kind;   // FLSG_PPR / FLSG_KATZ
    int k;
};

// step 1 (ppr_solver.py:641-646 / katz_solver.py:406-411 and the push value of :650 / :415)
__global__ __launch_bounds__(kCT) void k_cheb_update(ChebCoefArgs a) {
    __shared__ unsigned long long s_vol[kCT / kWave];
    unsigned long long vol = 0;
    for (int i = blockIdx.x * kCT + threadIdx.x; i < a.k; i += gridDim.x * kCT) {
        const int u = a.queue[i];
        const int d = a.degree[u];
        double t;
        if (a.kind == FLSG_PPR) t = a.b1 * a.r[u] + a.b2 * a.dx[u];
        else t = a.b1 * a.r[u] - a.b2 * a.dx[u];
        a.x[u] = a.x[u] + t;
        a.dx[u] = t;
        a.val[i] = (a.kind == FLSG_PPR) ? a.om_alpha * t / (double)d : a.alpha * t;
        a.dplus[i] = d + 1;
        vol += (unsigned long long)d;
    }
    vol = wave_sum(vol);
    if (__lane_id() == 0) s_vol[threadIdx.x / kWave] = vol;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kCT / kWave; w++) t += s_vol[w];
        if (t) atomicAdd(a.oper, t);
    }
}

// step 2: one thread per event position
__global__ __launch_bounds__(kCT) void k_cheb_events(const int32_t *indptr, const int32_t *indices,
                                                     const int32_t *degree, const int32_t *queue, const int64_t *P,
                                                     int k, int64_t E, int pbits, uint64_t *keys, int32_t *vals) {
    for (int64_t pos = blockIdx.x * (int64_t)kCT + threadIdx.x; pos < E; pos += (int64_t)gridDim.x * kCT) {
        // entry i with P[i] <= pos < P[i+1]
        int lo = 0, hi = k;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (P[mid] <= pos) lo = mid;
            else hi = mid;
        }
        const int i = lo;
        const int u = queue[i];
        const int64_t j = pos - P[i];
        const int d = degree[u];
        const int v = (j < d) ? indices[indptr[u] + j] : u;
        keys[pos] = ((uint64_t)(uint32_t)v << pbits) | (uint64_t)pos;
        vals[pos] = i;
    }
}

// step 4: sequential replay per destination
__global__ __launch_bounds__(kCT) void k_cheb_replay(const uint64_t *keys, const int32_t *vals, int64_t E,
                                                     int pbits, const int64_t *P, const int32_t *degree,
                                                     const double *val, const double *dx, double *r,
                                                     double thr_scale, int vbits, uint64_t *cross,
                                                     unsigned int *cnt) {
    const uint64_t pmask = (pbits >= 64) ? ~0ull : ((1ull << pbits) - 1);
    for (int64_t e = blockIdx.x * (int64_t)kCT + threadIdx.x; e < E; e += (int64_t)gridDim.x * kCT) {
        const uint64_t key = keys[e];
        const int v = (int)(key >> pbits);
        if (e > 0 && (int)(keys[e - 1] >> pbits) == v) continue;   // not the head of v's events
        const double thr = thr_scale * (double)degree[v];
        double rv = r[v];
        bool marked = false;
        uint64_t first = 0;
        for (int64_t f = e; f < E; f++) {
            const uint64_t kf = keys[f];
            if ((int)(kf >> pbits) != v) break;
            const int64_t pos = (int64_t)(kf & pmask);
            const int i = vals[f];
            if (pos == P[i + 1] - 1) {
                rv += -dx[v];                                   // ppr_solver.py:657
            } else {
                rv += val[i];                                   // ppr_solver.py:652
                if (!marked && thr <= fabs(rv)) {               // ppr_solver.py:653
                    marked = true;
                    first = (uint64_t)pos;
                }
            }
        }
        r[v] = rv;
        if (marked) {
            const unsigned int slot = atomicAdd(cnt, 1u);
            cross[slot] = (first << vbits) | (uint64_t)(uint32_t)v;
        }
    }
}

__global__ void k_cheb_extract(const uint64_t *sorted, unsigned int c, int vbits, int by_id, int32_t *queue) {
    const uint64_t vmask = (1ull << vbits) - 1;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x)
        queue[i] = (int32_t)(by_id ? sorted[i] : (sorted[i] & vmask));
}

__global__ void k_queue_to_keys(const int32_t *queue, unsigned int c, uint64_t *ids) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x)
        ids[i] = (uint64_t)(uint32_t)queue[i];
}

__global__ void k_cheb_ids(const uint64_t *cross, unsigned int c, int vbits, uint64_t *ids) {
    const uint64_t vmask = (1ull << vbits) - 1;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x)
        ids[i] = cross[i] & vmask;
}

// initial queue: row(s) in order, thr <= |r| (ppr_solver.py:626-630)
__global__ void k_cheb_init_queue(const int32_t *indptr, const int32_t *indices, const int32_t *degree,
                                  const double *r, int64_t s, double thr_scale, int32_t *queue, unsigned int *cnt) {
    // rows are duplicate-free, so the reference's q_mark test never fires here
    const int st = indptr[s], d = degree[s];
    unsigned int c = 0;
    for (int j = 0; j < d; j++) {
        const int v = indices[st + j];
        if (thr_scale * (double)degree[v] <= fabs(r[v])) queue[c++] = v;
    }
    *cnt = c;
}

__global__ void k_cheb_init_state(const int32_t *indptr, const int32_t *indices, const int32_t *degree, double *x,
                                  double *dx, double *r, int64_t s, double x0, double r_s0, double val) {
    x[s] = x0;
    dx[s] = x0;
    r[s] = r_s0;
    const int st = indptr[s], d = degree[s];
    for (int j = 0; j < d; j++) r[indices[st + j]] += val;
}

static void free_cheb(ChebWork *w) {
    if (!w) return;
    (void)hipFree(w->keys[0]);
    (void)hipFree(w->keys[1]);
    (void)hipFree(w->vals[0]);
    (void)hipFree(w->vals[1]);
    (void)hipFree(w->P);
    (void)hipFree(w->dplus);
    (void)hipFree(w->cross[0]);
    (void)hipFree(w->cross[1]);
    (void)hipFree(w->queue);
    (void)hipFree(w->cnt);
    (void)hipFree(w->oper);
    (void)hipFree(w->tmp);
    (void)hipFree(w->dx);
    (void)hipFree(w->val);
    delete w;
}

void cheb_free(flsg_graph g) {
    free_cheb((ChebWork *)g->cheb);
    g->cheb = nullptr;
}

static int ensure_cheb(flsg_graph g, hipStream_t st) {
    const int64_t n = g->n;
    const int64_t cap = g->nnz_cap + n + 1;
    ChebWork *w = (ChebWork *)g->cheb;
    if (w && w->cap >= cap) return FLSG_OK;
    free_cheb(w);
    w = new ChebWork();
    g->cheb = w;
    w->cap = cap;
    FLSG_HIP(hipMalloc((void **)&w->keys[0], sizeof(uint64_t) * cap));
    FLSG_HIP(hipMalloc((void **)&w->keys[1], sizeof(uint64_t) * cap));
    FLSG_HIP(hipMalloc((void **)&w->vals[0], sizeof(int32_t) * cap));
    FLSG_HIP(hipMalloc((void **)&w->vals[1], sizeof(int32_t) * cap));
    FLSG_HIP(hipMalloc((void **)&w->P, sizeof(int64_t) * (n + 2)));
    FLSG_HIP(hipMalloc((void **)&w->dplus, sizeof(int32_t) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&w->cross[0], sizeof(uint64_t) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&w->cross[1], sizeof(uint64_t) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&w->queue, sizeof(int32_t) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&w->cnt, sizeof(unsigned int)));
    FLSG_HIP(hipMalloc((void **)&w->oper, sizeof(unsigned long long)));
    FLSG_HIP(hipMalloc((void **)&w->dx, sizeof(double) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&w->val, sizeof(double) * (n + 1)));
    // temp storage: the largest of the three primitives at full size
    size_t b1 = 0, b2 = 0, b3 = 0;
    FLSG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b1, (const int64_t *)nullptr, (int64_t *)nullptr, (int)(n + 1),
                                              st));
    FLSG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b2, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                                (const int32_t *)nullptr, (int32_t *)nullptr, (int)cap, 0, 64, st));
    FLSG_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, b3, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                               (int)(n + 1), 0, 64, st));
    size_t b = b1 > b2 ? b1 : b2;
    b = b > b3 ? b : b3;
    w->tmp_bytes = b;
    FLSG_HIP(hipMalloc(&w->tmp, b));
    return FLSG_OK;
}

// dplus (int32) -> P (int64) exclusive scan through a widening iterator
struct WidenInt {
    __host__ __device__ int64_t operator()(const int32_t &v) const { return (int64_t)v; }
};

int solve_local_cheb(flsg_graph g, const flsg_params *p, int64_t s, double *d_x, double *d_r, flsg_result *res,
                     flsg_trace *tr, hipStream_t st) {
    const int64_t n = g->n;
    int rc = ensure_cheb(g, st);
    if (rc) return rc;
    ChebWork *w = (ChebWork *)g->cheb;
    const bool ppr = p->kind == FLSG_PPR;
    const double alpha = p->alpha;
    const double thr_scale = ppr ? p->eps * p->alpha : p->eps;
    FLSG_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, st));
    FLSG_HIP(hipMemsetAsync(d_r, 0, sizeof(double) * n, st));
    FLSG_HIP(hipMemsetAsync(w->dx, 0, sizeof(double) * n, st));
    FLSG_HIP(hipMemsetAsync(w->oper, 0, sizeof(unsigned long long), st));
    int32_t deg_s = 0;
    FLSG_HIP(hipMemcpyAsync(&deg_s, g->degree + s, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    double L = 0, mu = 0;
    if (ppr) {
        // r_s = alpha, x = dx = alpha e_s, r replaced by the push of s (r_s = 0)
        hipLaunchKernelGGL(k_cheb_init_state, dim3(1), dim3(1), 0, st, g->indptr, g->indices, g->degree, d_x, w->dx,
                           d_r, s, alpha, 0.0, (1. - alpha) * alpha / (double)deg_s);
    } else {
        L = p->l_max;
        mu = p->mu_min;
        hipLaunchKernelGGL(k_cheb_init_state, dim3(1), dim3(1), 0, st, g->indptr, g->indices, g->degree, d_x, w->dx,
                           d_r, s, 2. / (L + mu), 1. - 2. / (L + mu), 2. * alpha / (L + mu));
    }
    hipLaunchKernelGGL(k_cheb_init_queue, dim3(1), dim3(1), 0, st, g->indptr, g->indices, g->degree, d_r, s,
                       thr_scale, w->queue, w->cnt);
    unsigned int k = 0;
    FLSG_HIP(hipMemcpyAsync(&k, w->cnt, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    const int vbits = bits_for((uint64_t)n);
    const int grid = g->num_cus * 4;
    if ((double)k > (double)n / 4.0) {   // ppr_solver.py:639 applies to the first queue too
        size_t tb = w->tmp_bytes;
        hipLaunchKernelGGL(k_queue_to_keys, dim3(grid), dim3(kCT), 0, st, w->queue, k, w->cross[1]);
        FLSG_HIP(hipcub::DeviceRadixSort::SortKeys(w->tmp, tb, w->cross[1], w->cross[0], (int)k, 0, vbits, st));
        hipLaunchKernelGGL(k_cheb_extract, dim3(grid), dim3(kCT), 0, st, w->cross[0], k, vbits, 1, w->queue);
    }
    double delta_t = ppr ? (1. - alpha) : (L - mu) / (L + mu);
    int64_t iters = 0, trace_nodes = 0, total_nodes = 0;
    std::vector<int64_t> tcounts;
    const int64_t max_iters = p->max_iters > 0 ? p->max_iters : 100000;
    while (k > 0) {
        if (iters >= max_iters) return fail(FLSG_ENOCONV, "iteration cap reached");
        // coefficients (ppr_solver.py:636-637, katz_solver.py:400-402)
        double b1, b2;
        if (ppr) {
            delta_t = 1. / (2. / (1. - alpha) - delta_t);
            double beta = 2. * delta_t / (1. - alpha);
            b1 = beta;
            b2 = beta - 1.;
        } else {
            delta_t = 1. / (2. * ((L + mu) / (L - mu)) - delta_t);
            b1 = 4. * delta_t / (L - mu);
            b2 = (1. - 2. * delta_t * (L + mu) / (L - mu));
        }
        // the queue is in first-crossing order; above n/4 entries the reference
        // switches to ascending ids (already applied when the queue was built)
        if (tr && tr->d_nodes) {
            if (trace_nodes + k <= tr->cap_nodes)
                FLSG_HIP(hipMemcpyAsync(tr->d_nodes + trace_nodes, w->queue, sizeof(int32_t) * k,
                                        hipMemcpyDeviceToDevice, st));
            trace_nodes += k;
            tcounts.push_back(k);
        }
        total_nodes += k;
        ChebCoefArgs ca{};
        ca.degree = g->degree;
        ca.queue = w->queue;
        ca.x = d_x;
        ca.r = d_r;
        ca.dx = w->dx;
        ca.val = w->val;
        ca.dplus = w->dplus;
        ca.oper = w->oper;
        ca.b1 = b1;
        ca.b2 = b2;
        ca.alpha = alpha;
        ca.om_alpha = 1. - alpha;
        ca.kind = p->kind;
        ca.k = (int)k;
        hipLaunchKernelGGL(k_cheb_update, dim3(grid), dim3(kCT), 0, st, ca);
        // P = exclusive scan of (d + 1), P[k] = number of events
        hipcub::TransformInputIterator<int64_t, WidenInt, const int32_t *> it(w->dplus, WidenInt());
        FLSG_HIP(hipMemsetAsync(w->dplus + k, 0, sizeof(int32_t), st));
        size_t tb = w->tmp_bytes;
        FLSG_HIP(hipcub::DeviceScan::ExclusiveSum(w->tmp, tb, it, w->P, (int)k + 1, st));
        int64_t E = 0;
        FLSG_HIP(hipMemcpyAsync(&E, w->P + k, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        FLSG_HIP(hipStreamSynchronize(st));
        if (E > w->cap) return fail(FLSG_ECAP, "event buffer too small");
        const int pbits = bits_for((uint64_t)E);
        if (pbits + vbits > 64) return fail(FLSG_ECAP, "event keys exceed 64 bits");
        hipLaunchKernelGGL(k_cheb_events, dim3(grid), dim3(kCT), 0, st, g->indptr, g->indices, g->degree, w->queue,
                           w->P, (int)k, E, pbits, w->keys[0], w->vals[0]);
        tb = w->tmp_bytes;
        FLSG_HIP(hipcub::DeviceRadixSort::SortPairs(w->tmp, tb, w->keys[0], w->keys[1], w->vals[0], w->vals[1],
                                                    (int)E, 0, pbits + vbits, st));
        FLSG_HIP(hipMemsetAsync(w->cnt, 0, sizeof(unsigned int), st));
        hipLaunchKernelGGL(k_cheb_replay, dim3(grid), dim3(kCT), 0, st, w->keys[1], w->vals[1], E, pbits, w->P,
                           g->degree, w->val, w->dx, d_r, thr_scale, vbits, w->cross[0], w->cnt);
        unsigned int c = 0;
        FLSG_HIP(hipMemcpyAsync(&c, w->cnt, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        FLSG_HIP(hipStreamSynchronize(st));
        iters++;
        if (c > 0) {
            const bool by_id = (double)c > (double)n / 4.0;   // ppr_solver.py:639
            tb = w->tmp_bytes;
            const uint64_t *sorted = w->cross[1];
            if (by_id) {
                hipLaunchKernelGGL(k_cheb_ids, dim3(grid), dim3(kCT), 0, st, w->cross[0], c, vbits, w->cross[1]);
                FLSG_HIP(hipcub::DeviceRadixSort::SortKeys(w->tmp, tb, w->cross[1], w->cross[0], (int)c, 0, vbits,
                                                           st));
                sorted = w->cross[0];
            } else {
                FLSG_HIP(hipcub::DeviceRadixSort::SortKeys(w->tmp, tb, w->cross[0], w->cross[1], (int)c, 0,
                                                           pbits + vbits, st));
            }
            hipLaunchKernelGGL(k_cheb_extract, dim3(grid), dim3(kCT), 0, st, sorted, c, vbits, by_id ? 1 : 0,
                               w->queue);
        }
        k = c;
    }
    unsigned long long oper = 0;
    FLSG_HIP(hipMemcpyAsync(&oper, w->oper, sizeof(oper), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    res->oper = (int64_t)oper;
    res->iters = iters;
    res->nodes = total_nodes;
    if (tr) {
        tr->h_off[0] = 0;
        const int64_t ni = (int64_t)tcounts.size() < tr->cap_iters ? (int64_t)tcounts.size() : tr->cap_iters;
        for (int64_t t = 0; t < ni; t++) tr->h_off[t + 1] = tr->h_off[t] + tcounts[t];
        tr->n_iters = ni;
        tr->overflow = (trace_nodes > tr->cap_nodes) || ((int64_t)tcounts.size() > tr->cap_iters);
    }
    return FLSG_OK;
}

}  // namespace flsg
