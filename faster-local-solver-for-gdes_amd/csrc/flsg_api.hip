// flsg_api.hip -- graph handles, error reporting, device queries.
#include <stdio.h>

#include <string>

#include "flsg_internal.h"

namespace flsg {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int check_graph(flsg_graph g) {
    if (!g) return fail(FLSG_EARG, "graph handle is NULL");
    return FLSG_OK;
}

// out[0] = sum of degree, out[1] = largest row capacity indptr[u+1] - indptr[u]
__global__ void k_sum_degree(const int32_t *indptr, const int32_t *degree, int64_t n, unsigned long long *out) {
    unsigned long long t = 0, m = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        t += (unsigned long long)degree[i];
        const unsigned long long c = (unsigned long long)(indptr[i + 1] - indptr[i]);
        m = c > m ? c : m;
    }
    t = wave_sum(t);
    if (__lane_id() == 0 && t) atomicAdd(&out[0], t);
    if (m) atomicMax(&out[1], m);
}

int recompute_m(flsg_graph g, hipStream_t st) {
    unsigned long long *d = nullptr;
    unsigned long long h[2] = {0, 0};
    FLSG_HIP(hipMalloc((void **)&d, sizeof(unsigned long long) * 2));
    FLSG_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long) * 2, st));
    hipLaunchKernelGGL(k_sum_degree, dim3(1024), dim3(256), 0, st, g->indptr, g->degree, g->n, d);
    FLSG_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    FLSG_HIP(hipFree(d));
    g->m_live = (int64_t)h[0];
    g->d_max_hint = (int64_t)h[1];
    return FLSG_OK;
}

int ensure_single_workspace(flsg_graph g) {
    if (g->front[0]) return FLSG_OK;
    const int64_t n = g->n;
    FLSG_HIP(hipSetDevice(g->device));
    FLSG_HIP(hipMalloc((void **)&g->front[0], sizeof(int32_t) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&g->front[1], sizeof(int32_t) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&g->val, sizeof(double) * (n + 1)));
    // hub pieces: every node of degree > 0 contributes at most ceil(d / piece) pieces;
    // 1 + m / 1024 + n bounds it for any piece size >= 1024.
    g->pieces_cap = g->nnz_cap / 1024 + g->n + 16;
    FLSG_HIP(hipMalloc((void **)&g->pieces, sizeof(int64_t) * g->pieces_cap));
    FLSG_HIP(hipMalloc((void **)&g->work_a, sizeof(double) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&g->work_b, sizeof(double) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&g->work_c, sizeof(double) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&g->local_x, sizeof(double) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&g->local_r, sizeof(double) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&g->ctr, sizeof(SolveCounters)));
    FLSG_HIP(hipHostMalloc((void **)&g->h_ctr, sizeof(SolveCounters), hipHostMallocDefault));
    // next-frontier sub-lists: node u goes to sub-list u mod kFrontSub, so each
    // holds at most ceil(n / kFrontSub) entries per iteration
    g->fsub_cap = (n + kFrontSub - 1) / kFrontSub;
    FLSG_HIP(hipMalloc((void **)&g->fsub, sizeof(int32_t) * kFrontSub * g->fsub_cap));
    FLSG_HIP(hipMalloc((void **)&g->fsubcnt, sizeof(unsigned int) * kFrontCntWords));
    FLSG_HIP(hipMemset(g->fsubcnt, 0, sizeof(unsigned int) * kFrontCntWords));
    return FLSG_OK;
}

}  // namespace flsg

using namespace flsg;

extern "C" {

int flsg_version(void) { return 2; }

const char *flsg_last_error(void) { return g_last_error.c_str(); }

int flsg_device_info(int32_t *num_cus, int64_t *free_bytes, int64_t *total_bytes) {
    int dev = 0;
    FLSG_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    FLSG_HIP(hipGetDeviceProperties(&prop, dev));
    size_t fr = 0, tot = 0;
    FLSG_HIP(hipMemGetInfo(&fr, &tot));
    if (num_cus) *num_cus = prop.multiProcessorCount;
    if (free_bytes) *free_bytes = (int64_t)fr;
    if (total_bytes) *total_bytes = (int64_t)tot;
    return FLSG_OK;
}

int flsg_graph_create(int64_t n, int64_t nnz_cap, const int32_t *d_indptr, int32_t *d_indices,
                      int32_t *d_degree, int32_t slack, flsg_graph *out) {
    if (!out) return fail(FLSG_EARG, "out is NULL");
    *out = nullptr;
    if (n <= 0 || n >= (int64_t)INT32_MAX) return fail(FLSG_EARG, "n out of range");
    if (nnz_cap < 0 || nnz_cap >= (int64_t)INT32_MAX)
        return fail(FLSG_EARG, "nnz must be < 2^31 (int32 indptr)");
    if (!d_indptr || !d_degree || (nnz_cap > 0 && !d_indices))
        return fail(FLSG_EARG, "NULL CSR buffer");
    flsg_graph g = new flsg_graph_s();
    g->n = n;
    g->nnz_cap = nnz_cap;
    g->indptr = d_indptr;
    g->indices = d_indices;
    g->degree = d_degree;
    g->slack = slack;
    if (hipGetDevice(&g->device) != hipSuccess) {
        delete g;
        return fail(FLSG_EHIP, "hipGetDevice failed");
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, g->device) != hipSuccess) {
        delete g;
        return fail(FLSG_EHIP, "hipGetDeviceProperties failed");
    }
    g->num_cus = prop.multiProcessorCount;
    int rc = recompute_m(g, nullptr);
    if (rc) {
        delete g;
        return rc;
    }
    *out = g;
    return FLSG_OK;
}

int flsg_graph_destroy(flsg_graph g) {
    if (!g) return FLSG_OK;
    (void)hipSetDevice(g->device);
    (void)hipDeviceSynchronize();
    (void)hipFree(g->front[0]);
    (void)hipFree(g->front[1]);
    (void)hipFree(g->fsub);
    (void)hipFree(g->fsubcnt);
    (void)hipFree(g->rows_wave);
    (void)hipFree(g->rows_block);
    (void)hipFree(g->rows_cnt);
    (void)hipFree(g->val);
    (void)hipFree(g->pieces);
    (void)hipFree(g->work_a);
    (void)hipFree(g->work_b);
    (void)hipFree(g->work_c);
    (void)hipFree(g->hk_lvl);
    (void)hipFree(g->local_x);
    (void)hipFree(g->local_r);
    (void)hipFree(g->coef);
    (void)hipFree(g->ctr);
    (void)hipFree(g->trace_dev_off);
    if (g->h_ctr) (void)hipHostFree(g->h_ctr);
    batch_free(g);
    cheb_free(g);
    ignn_free(g);
    dyn_free(g);
    sor_free(g);
    single_graphs_free(g);
    delete g;
    return FLSG_OK;
}

}  // extern "C"

extern "C" int flsg_stream_poll(void *stream, int32_t *h_done) {
    if (!h_done) return fail(FLSG_EARG, "h_done is NULL");
    const hipError_t e = hipStreamQuery((hipStream_t)stream);
    if (e == hipSuccess) {
        *h_done = 1;
        return FLSG_OK;
    }
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();   // not-ready is a state, not an error to keep
        *h_done = 0;
        return FLSG_OK;
    }
    *h_done = 0;
    return fail(FLSG_EHIP, std::string("stream poll: ") + hipGetErrorString(e));
}
