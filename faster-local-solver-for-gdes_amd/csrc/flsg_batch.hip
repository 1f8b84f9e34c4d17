// flsg_batch.hip -- many queries in one persistent launch, FP32 push with an
// FP64 threshold test and an FP64 residual certificate.
//
// Layout.  Workgroup b owns "slot" b for the whole launch: dense per-slot
// vectors of n entries (r32, x32, frontier double buffer, push values,
// support list, r64, x64, flags) that are all-zero between queries -- every
// query resets exactly the entries it touched.  A slot is only ever touched by
// its own workgroup (one CU, one XCD), so its residual atomics are
// workgroup-scoped and execute in that XCD's L2; kernel boundaries publish the
// slot state to the next launch.
//
// Per query (workgroup-local, barriers are __syncthreads only):
//   1. FP32 local push (gradient descent = Jacobi push, ppr_solver.py:246-288 /
//      katz_solver.py:183-224): node phase (r_u -> 0, x_u += r_u, push
//      value), then an edge phase whose edges are flattened with an LDS prefix
//      sum over chunks of the frontier; v joins the next frontier iff the
//      atomic that moved r_v from below to above thr_v = eps*alpha*d_v (FP64
//      compare) was this one.
//   2. FP64 certificate: r64 = b - Q x recomputed from the FP32 x over
//      supp(x) and its neighbours, every node with |r64_v| >= thr_v queued,
//      then an FP64 correction push (same push, mark-based queue because the
//      recomputed residual has both signs) until none is left: on return
//      |r_v| < thr_v for every v, in FP64, on every query.
//   3. Sparse output (support ids + FP64 values) and reset of the slot.
// Queries are claimed through a global counter, so uneven queries balance
// across workgroups.
#include <math.h>

#include <vector>

#include "flsg_internal.h"

namespace flsg {

constexpr int kBT = 512;             // threads per workgroup (8 waves)
constexpr unsigned kInQ = 1u;        // flags: in the correction queue
constexpr unsigned kSeen = 2u;       // flags: listed in the support

struct BatchArena {
    int64_t n = 0;
    int32_t slots = 0;
    float *r32 = nullptr;
    float *x32 = nullptr;
    int32_t *q0 = nullptr;
    int32_t *q1 = nullptr;
    double *vq = nullptr;
    int32_t *sup = nullptr;
    double *r64 = nullptr;
    double *x64 = nullptr;
    uint32_t *flags = nullptr;
    unsigned int *next_query = nullptr;
    unsigned long long *out_used = nullptr;
    flsg_result *d_res = nullptr;
    int64_t d_res_cap = 0;
};

struct BatchArgs {
    const int32_t *indptr;
    const int32_t *indices;
    const int32_t *degree;
    int64_t n;
    const int64_t *seeds;
    int64_t nq;
    float *r32;
    float *x32;
    int32_t *q0;
    int32_t *q1;
    double *vq;
    int32_t *sup;
    double *r64;
    double *x64;
    uint32_t *flags;
    unsigned int *next_query;
    int32_t *out_ids;
    double *out_vals;
    int64_t *out_off;
    int64_t out_cap;
    unsigned long long *out_used;
    flsg_result *res;
    int kind;            // FLSG_PPR / FLSG_KATZ
    double alpha;
    double om_alpha;
    double thr_scale;    // PPR eps*alpha, Katz eps
    double r0;           // PPR alpha, Katz 1
    int64_t max_iters;
};

__device__ __forceinline__ double bpush_value(const BatchArgs &a, double tmp, int d) {
    return a.kind == FLSG_PPR ? a.om_alpha * tmp / (double)d : a.alpha * tmp;
}

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v, unsigned long long *sh) {
    v = wave_sum(v);
    const int wid = threadIdx.x / kWave;
    if (__lane_id() == 0) sh[wid] = v;
    __syncthreads();
    unsigned long long t = 0;
    for (int w = 0; w < kBT / kWave; w++) t += sh[w];
    __syncthreads();
    return t;
}

__device__ __forceinline__ double block_max(double v, double *sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    const int wid = threadIdx.x / kWave;
    if (__lane_id() == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kBT / kWave; w++) t = fmax(t, sh[w]);
    __syncthreads();
    return t;
}

// Edge flattening over list[0:cnt): calls f(k_entry, v, e_flat) for every edge
// of every entry; `weight(i)` gives the entry's degree (0 = skip) and start.
// All threads of the block must call it.  LDS: s_incl/s_start of kBT ints.
struct Flat {
    int *s_incl;
    int *s_start;
    int *s_scan;
};

__device__ __forceinline__ uint64_t now_ticks() { return __builtin_amdgcn_s_memrealtime(); }

__global__ __launch_bounds__(kBT) void k_batch(BatchArgs a) {
    __shared__ int s_incl[kBT];
    __shared__ int s_start[kBT];
    __shared__ double s_val[kBT];
    __shared__ int s_scan[kBT / kWave];
    __shared__ unsigned long long s_red[kBT / kWave];
    __shared__ double s_redd[kBT / kWave];
    __shared__ unsigned int s_q, s_cnt, s_ncnt, s_sup;
    __shared__ unsigned long long s_base;

    const int tid = threadIdx.x;
    const int64_t off = (int64_t)blockIdx.x * a.n;
    float *r32 = a.r32 + off;
    float *x32 = a.x32 + off;
    int32_t *qa = a.q0 + off;
    int32_t *qb = a.q1 + off;
    double *vq = a.vq + off;
    int32_t *sup = a.sup + off;
    double *r64 = a.r64 + off;
    double *x64 = a.x64 + off;
    uint32_t *flags = a.flags + off;
    const int32_t *indptr = a.indptr;
    const int32_t *indices = a.indices;
    const int32_t *degree = a.degree;

    for (;;) {
        if (tid == 0) s_q = atomicAdd(a.next_query, 1u);
        __syncthreads();
        const unsigned int q = s_q;
        __syncthreads();
        if ((int64_t)q >= a.nq) break;
        const uint64_t t_start = now_ticks();
        const int s = (int)a.seeds[q];
        if (tid == 0) {
            r32[s] = (float)a.r0;
            qa[0] = s;
            s_cnt = 1;
            s_sup = 0;
        }
        __syncthreads();
        int32_t *qc = qa, *qn = qb;
        unsigned long long oper = 0;   // identical in all threads
        unsigned long long fnodes = 0;
        int64_t iters = 0;
        int status = FLSG_OK;

        // ---------------- 1. FP32 local push
        for (;;) {
            const unsigned int cnt = s_cnt;
            if (cnt == 0) break;
            if (iters >= a.max_iters) {
                status = FLSG_ENOCONV;
                break;
            }
            iters++;
            fnodes += cnt;
            __syncthreads();
            if (tid == 0) s_ncnt = 0;
            unsigned long long vol = 0;
            for (unsigned int i = tid; i < cnt; i += kBT) {
                const int u = qc[i];
                const int d = degree[u];
                const float tmp = r32[u];
                r32[u] = 0.0f;
                const float xo = x32[u];
                x32[u] = xo + tmp;
                if (xo == 0.0f) {
                    const unsigned int pos = atomicAdd(&s_sup, 1u);
                    sup[pos] = u;
                }
                vq[i] = (double)(float)bpush_value(a, (double)tmp, d);
                vol += (unsigned long long)d;
            }
            oper += block_sum_u64(vol, s_red);   // contains __syncthreads
            for (unsigned int base = 0; base < cnt; base += kBT) {
                const unsigned int i = base + tid;
                int d = 0, st = 0;
                double v = 0.0;
                if (i < cnt) {
                    const int u = qc[i];
                    d = degree[u];
                    st = indptr[u];
                    v = vq[i];
                }
                int tot;
                const int incl = block_incl_scan<kBT>(d, s_scan, &tot);
                s_incl[tid] = incl;
                s_start[tid] = st;
                s_val[tid] = v;
                __syncthreads();
                for (int e0 = 0; e0 < tot; e0 += kBT) {
                    const int e = e0 + tid;
                    bool cross = false;
                    int vtx = 0;
                    if (e < tot) {
                        const int k = upper_bound_lds(s_incl, 0, kBT, e);
                        const int o = e - (k ? s_incl[k - 1] : 0);
                        vtx = indices[s_start[k] + o];
                        const float val = (float)s_val[k];
                        const float old = atomic_add_wg(&r32[vtx], val);
                        const float nw = old + val;
                        const double thr = a.thr_scale * (double)degree[vtx];
                        cross = (double)old < thr && thr <= (double)nw;
                    }
                    const long long slot = wave_append(cross, &s_ncnt);
                    if (slot >= 0) qn[slot] = vtx;
                }
                __syncthreads();
            }
            if (tid == 0) s_cnt = s_ncnt;
            int32_t *t = qc;
            qc = qn;
            qn = t;
            __syncthreads();
        }
        __syncthreads();
        const unsigned int S32 = s_sup;

        // ---------------- 2a. r64 = pushes of x32 over supp (and reset r32 there)
        unsigned long long cert_ops = 0;
        for (unsigned int base = 0; base < S32; base += kBT) {
            const unsigned int i = base + tid;
            int d = 0, st = 0;
            double v = 0.0;
            if (i < S32) {
                const int u = sup[i];
                d = degree[u];
                st = indptr[u];
                v = bpush_value(a, (double)x32[u], d);
            }
            int tot;
            const int incl = block_incl_scan<kBT>(d, s_scan, &tot);
            s_incl[tid] = incl;
            s_start[tid] = st;
            s_val[tid] = v;
            __syncthreads();
            cert_ops += (unsigned long long)tot;
            for (int e = tid; e < tot; e += kBT) {
                const int k = upper_bound_lds(s_incl, 0, kBT, e);
                const int o = e - (k ? s_incl[k - 1] : 0);
                const int vtx = indices[s_start[k] + o];
                atomic_add_wg(&r64[vtx], s_val[k]);
                r32[vtx] = 0.0f;
            }
            __syncthreads();
        }
        // 2b. r64_u -= x_u on the support; x64 = x32; x32 reset
        for (unsigned int i = tid; i < S32; i += kBT) {
            const int u = sup[i];
            const double xd = (double)x32[u];
            r64[u] = r64[u] - xd;
            x64[u] = xd;
            x32[u] = 0.0f;
            flags[u] = kSeen;
        }
        __syncthreads();
        if (tid == 0) {
            r64[s] = r64[s] + a.r0;
            r32[s] = 0.0f;
            s_cnt = 0;
        }
        __syncthreads();
        // 2c. queue every candidate (supp and its neighbours) with |r64| >= thr
        for (unsigned int base = 0; base < S32; base += kBT) {
            const unsigned int i = base + tid;
            int d = 0, st = 0;
            if (i < S32) {
                const int u = sup[i];
                d = degree[u] + 1;      // the node itself + its row
                st = indptr[u];
                s_val[tid] = (double)u;
            }
            int tot;
            const int incl = block_incl_scan<kBT>(d, s_scan, &tot);
            s_incl[tid] = incl;
            s_start[tid] = st;
            __syncthreads();
            cert_ops += (unsigned long long)tot;
            for (int e0 = 0; e0 < tot; e0 += kBT) {
                const int e = e0 + tid;
                bool enq = false;
                int vtx = 0;
                if (e < tot) {
                    const int k = upper_bound_lds(s_incl, 0, kBT, e);
                    const int o = e - (k ? s_incl[k - 1] : 0);
                    vtx = (o == 0) ? (int)s_val[k] : indices[s_start[k] + o - 1];
                    const double thr = a.thr_scale * (double)degree[vtx];
                    if (fabs(r64[vtx]) >= thr) {
                        const uint32_t old = __hip_atomic_fetch_or(&flags[vtx], kInQ, __ATOMIC_RELAXED,
                                                                    __HIP_MEMORY_SCOPE_WORKGROUP);
                        enq = (old & kInQ) == 0;
                    }
                }
                const long long slot = wave_append(enq, &s_cnt);
                if (slot >= 0) qc[slot] = vtx;
            }
            __syncthreads();
        }
        // 2d. FP64 correction push
        unsigned long long n_corr = 0;
        int64_t corr_iters = 0;
        for (;;) {
            const unsigned int cnt = s_cnt;
            if (cnt == 0) break;
            if (corr_iters >= a.max_iters) {
                status = FLSG_ENOCONV;
                break;
            }
            corr_iters++;
            __syncthreads();
            if (tid == 0) s_ncnt = 0;
            unsigned long long pushed = 0;
            for (unsigned int i = tid; i < cnt; i += kBT) {
                const int u = qc[i];
                const int d = degree[u];
                const uint32_t fl = __hip_atomic_fetch_and(&flags[u], ~kInQ, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                const double tmp = r64[u];
                double v = 0.0;
                if (fabs(tmp) >= a.thr_scale * (double)d) {
                    r64[u] = 0.0;
                    x64[u] = x64[u] + tmp;
                    if ((fl & kSeen) == 0) {
                        flags[u] = kSeen;   // kInQ was just cleared and only this thread holds u
                        const unsigned int pos = atomicAdd(&s_sup, 1u);
                        sup[pos] = u;
                    }
                    v = bpush_value(a, tmp, d);
                    pushed += 1;
                }
                vq[i] = v;
            }
            n_corr += block_sum_u64(pushed, s_red);
            for (unsigned int base = 0; base < cnt; base += kBT) {
                const unsigned int i = base + tid;
                int d = 0, st = 0;
                double v = 0.0;
                if (i < cnt) {
                    v = vq[i];
                    if (v != 0.0) {
                        const int u = qc[i];
                        d = degree[u];
                        st = indptr[u];
                    }
                }
                int tot;
                const int incl = block_incl_scan<kBT>(d, s_scan, &tot);
                s_incl[tid] = incl;
                s_start[tid] = st;
                s_val[tid] = v;
                __syncthreads();
                cert_ops += (unsigned long long)tot;
                for (int e0 = 0; e0 < tot; e0 += kBT) {
                    const int e = e0 + tid;
                    bool enq = false;
                    int vtx = 0;
                    if (e < tot) {
                        const int k = upper_bound_lds(s_incl, 0, kBT, e);
                        const int o = e - (k ? s_incl[k - 1] : 0);
                        vtx = indices[s_start[k] + o];
                        const double val = s_val[k];
                        const double old = atomic_add_wg(&r64[vtx], val);
                        const double nw = old + val;
                        if (fabs(nw) >= a.thr_scale * (double)degree[vtx]) {
                            const uint32_t of = __hip_atomic_fetch_or(&flags[vtx], kInQ, __ATOMIC_RELAXED,
                                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                            enq = (of & kInQ) == 0;
                        }
                    }
                    const long long slot = wave_append(enq, &s_ncnt);
                    if (slot >= 0) qn[slot] = vtx;
                }
                __syncthreads();
            }
            if (tid == 0) s_cnt = s_ncnt;
            int32_t *t = qc;
            qc = qn;
            qn = t;
            __syncthreads();
        }
        __syncthreads();
        const unsigned int S = s_sup;
        // ---------------- 3. output, certificate value, reset
        if (tid == 0) {
            const unsigned long long base = atomicAdd(a.out_used, (unsigned long long)S);
            s_base = base;
        }
        __syncthreads();
        const unsigned long long obase = s_base;
        const bool fits = obase + S <= (unsigned long long)a.out_cap;
        double cmax = 0.0;
        for (unsigned int base = 0; base < S; base += kBT) {
            const unsigned int i = base + tid;
            int d = 0, st = 0;
            if (i < S) {
                const int u = sup[i];
                d = degree[u] + 1;
                st = indptr[u];
                s_val[tid] = (double)u;
                if (fits) {
                    a.out_ids[obase + i] = u;
                    a.out_vals[obase + i] = x64[u];
                }
            }
            int tot;
            const int incl = block_incl_scan<kBT>(d, s_scan, &tot);
            s_incl[tid] = incl;
            s_start[tid] = st;
            __syncthreads();
            cert_ops += (unsigned long long)tot;
            for (int e = tid; e < tot; e += kBT) {
                const int k = upper_bound_lds(s_incl, 0, kBT, e);
                const int o = e - (k ? s_incl[k - 1] : 0);
                const int vtx = (o == 0) ? (int)s_val[k] : indices[s_start[k] + o - 1];
                const double rv = r64[vtx];
                if (rv != 0.0) {
                    const double thr = a.thr_scale * (double)degree[vtx];
                    cmax = fmax(cmax, fabs(rv) / thr);
                }
            }
            __syncthreads();
        }
        // reset after the certificate value is taken (separate sweep: a
        // neighbour may be shared by several support nodes)
        for (unsigned int base = 0; base < S; base += kBT) {
            const unsigned int i = base + tid;
            int d = 0, st = 0;
            if (i < S) {
                const int u = sup[i];
                d = degree[u] + 1;
                st = indptr[u];
                s_val[tid] = (double)u;
            }
            int tot;
            const int incl = block_incl_scan<kBT>(d, s_scan, &tot);
            s_incl[tid] = incl;
            s_start[tid] = st;
            __syncthreads();
            for (int e = tid; e < tot; e += kBT) {
                const int k = upper_bound_lds(s_incl, 0, kBT, e);
                const int o = e - (k ? s_incl[k - 1] : 0);
                const int vtx = (o == 0) ? (int)s_val[k] : indices[s_start[k] + o - 1];
                r64[vtx] = 0.0;
                flags[vtx] = 0u;
                if (o == 0) x64[vtx] = 0.0;
            }
            __syncthreads();
        }
        cmax = block_max(cmax, s_redd);
        if (tid == 0) {
            flsg_result rr{};
            rr.oper = (int64_t)oper;
            rr.iters = iters;
            rr.oper_cert = (int64_t)cert_ops;
            rr.n_corr = (int64_t)n_corr;
            rr.supp = (int64_t)S;
            rr.nodes = (int64_t)fnodes;
            rr.cert = cmax;
            rr.ms = (double)(now_ticks() - t_start) * 1e-5;   // 100 MHz constant clock
            rr.status = fits ? status : FLSG_ECAP;
            a.res[q] = rr;
            a.out_off[q] = fits ? (int64_t)obase : -1;
        }
        __syncthreads();
    }
}

void batch_free(flsg_graph g) {
    BatchArena *A = g->arena;
    if (!A) return;
    hipFree(A->r32);
    hipFree(A->x32);
    hipFree(A->q0);
    hipFree(A->q1);
    hipFree(A->vq);
    hipFree(A->sup);
    hipFree(A->r64);
    hipFree(A->x64);
    hipFree(A->flags);
    hipFree(A->next_query);
    hipFree(A->out_used);
    hipFree(A->d_res);
    delete A;
    g->arena = nullptr;
}

int64_t slot_bytes(int64_t n) {
    // r32 x32 q0 q1 sup flags (4 B each) + vq r64 x64 (8 B each)
    return n * (6 * 4 + 3 * 8);
}

static int ensure_arena(flsg_graph g, int32_t slots, hipStream_t st) {
    BatchArena *A = g->arena;
    if (A && A->slots == slots && A->n == g->n) return FLSG_OK;
    batch_free(g);
    A = new BatchArena();
    g->arena = A;
    A->n = g->n;
    A->slots = slots;
    const size_t ns = (size_t)g->n * (size_t)slots;
    FLSG_HIP(hipMalloc((void **)&A->r32, sizeof(float) * ns));
    FLSG_HIP(hipMalloc((void **)&A->x32, sizeof(float) * ns));
    FLSG_HIP(hipMalloc((void **)&A->q0, sizeof(int32_t) * ns));
    FLSG_HIP(hipMalloc((void **)&A->q1, sizeof(int32_t) * ns));
    FLSG_HIP(hipMalloc((void **)&A->vq, sizeof(double) * ns));
    FLSG_HIP(hipMalloc((void **)&A->sup, sizeof(int32_t) * ns));
    FLSG_HIP(hipMalloc((void **)&A->r64, sizeof(double) * ns));
    FLSG_HIP(hipMalloc((void **)&A->x64, sizeof(double) * ns));
    FLSG_HIP(hipMalloc((void **)&A->flags, sizeof(uint32_t) * ns));
    FLSG_HIP(hipMalloc((void **)&A->next_query, sizeof(unsigned int)));
    FLSG_HIP(hipMalloc((void **)&A->out_used, sizeof(unsigned long long)));
    FLSG_HIP(hipMemsetAsync(A->r32, 0, sizeof(float) * ns, st));
    FLSG_HIP(hipMemsetAsync(A->x32, 0, sizeof(float) * ns, st));
    FLSG_HIP(hipMemsetAsync(A->r64, 0, sizeof(double) * ns, st));
    FLSG_HIP(hipMemsetAsync(A->x64, 0, sizeof(double) * ns, st));
    FLSG_HIP(hipMemsetAsync(A->flags, 0, sizeof(uint32_t) * ns, st));
    return FLSG_OK;
}

static int32_t default_slots(flsg_graph g) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return g->num_cus;
    const int64_t per = slot_bytes(g->n);
    int64_t fit = (int64_t)(fr / 2) / (per > 0 ? per : 1);
    int64_t want = (int64_t)g->num_cus * 2;
    if (fit < want) want = fit;
    if (want < 1) want = 1;
    return (int32_t)want;
}

}  // namespace flsg

using namespace flsg;

extern "C" int64_t flsg_batch_slot_bytes(int64_t n) { return slot_bytes(n); }

extern "C" int flsg_batch_set_slots(flsg_graph g, int32_t slots) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (slots < 0) return fail(FLSG_EARG, "slots < 0");
    g->batch_slots = slots;
    return FLSG_OK;
}

extern "C" int flsg_batch_solve(flsg_graph g, const flsg_params *p, const int64_t *d_seeds, int64_t n_queries,
                                int32_t *d_out_ids, double *d_out_vals, int64_t *d_out_off, int64_t out_cap,
                                flsg_result *h_res, void *stream) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (!p || !d_seeds || !h_res || !d_out_off || (out_cap > 0 && (!d_out_ids || !d_out_vals)))
        return fail(FLSG_EARG, "NULL argument");
    if (n_queries <= 0) return FLSG_OK;
    if (n_queries >= (int64_t)UINT32_MAX / 2) return fail(FLSG_EARG, "too many queries");
    if (p->kind != FLSG_PPR && p->kind != FLSG_KATZ) return fail(FLSG_EARG, "batched solves: PPR or Katz");
    if (p->method != FLSG_GD || !p->local) return fail(FLSG_EARG, "batched solves: local gradient descent");
    if (p->precision != FLSG_FP32) return fail(FLSG_EARG, "batched solves run the FP32 + FP64-certificate path");
    hipStream_t st = (hipStream_t)stream;
    const int32_t slots = g->batch_slots > 0 ? g->batch_slots : default_slots(g);
    rc = ensure_arena(g, slots, st);
    if (rc) return rc;
    BatchArena *A = g->arena;
    if (A->d_res_cap < n_queries) {
        if (A->d_res) FLSG_HIP(hipFree(A->d_res));
        A->d_res = nullptr;
        FLSG_HIP(hipMalloc((void **)&A->d_res, sizeof(flsg_result) * n_queries));
        A->d_res_cap = n_queries;
    }
    FLSG_HIP(hipMemsetAsync(A->next_query, 0, sizeof(unsigned int), st));
    FLSG_HIP(hipMemsetAsync(A->out_used, 0, sizeof(unsigned long long), st));
    BatchArgs a{};
    a.indptr = g->indptr;
    a.indices = g->indices;
    a.degree = g->degree;
    a.n = g->n;
    a.seeds = d_seeds;
    a.nq = n_queries;
    a.r32 = A->r32;
    a.x32 = A->x32;
    a.q0 = A->q0;
    a.q1 = A->q1;
    a.vq = A->vq;
    a.sup = A->sup;
    a.r64 = A->r64;
    a.x64 = A->x64;
    a.flags = A->flags;
    a.next_query = A->next_query;
    a.out_ids = d_out_ids;
    a.out_vals = d_out_vals;
    a.out_off = d_out_off;
    a.out_cap = out_cap;
    a.out_used = A->out_used;
    a.res = A->d_res;
    a.kind = p->kind;
    a.alpha = p->alpha;
    a.om_alpha = 1.0 - p->alpha;
    a.thr_scale = p->kind == FLSG_PPR ? p->eps * p->alpha : p->eps;
    a.r0 = p->kind == FLSG_PPR ? p->alpha : 1.0;
    a.max_iters = p->max_iters > 0 ? p->max_iters : 20000;
    hipLaunchKernelGGL(k_batch, dim3(slots), dim3(kBT), 0, st, a);
    FLSG_HIP(hipGetLastError());
    FLSG_HIP(hipMemcpyAsync(h_res, A->d_res, sizeof(flsg_result) * n_queries, hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n_queries; i++)
        if (h_res[i].status != FLSG_OK) return fail(h_res[i].status, "a batched query failed (see its status)");
    return FLSG_OK;
}
