// flsg_dyn.hip -- in-place edge insertion into a device-resident slack CSR.
//
// Reference: updetaGraph with onlyinsert=True (InstantGNN.py:137-180,
// ppr_solver.py:723-766): for each (u, v) in order, skip it if v is already in
// u's live row; otherwise append v to row u and u to row v (slot
// indptr[u] + degree[u]) and bump both degrees.  On the device:
//   1. pairs are canonicalised to (min, max) and radix-sorted (stable), so a
//      pair repeated in the batch -- in either orientation -- is inserted once,
//      as the sequential loop does;
//   2. each distinct pair is checked against the current live rows (the
//      shorter of the two rows is scanned);
//   3. per-node insertion counts are checked against the row capacities
//      (indptr[u+1] - indptr[u]); if any row would overflow nothing is written
//      and FLSG_ECAP is returned (the reference would overwrite the next row);
//   4. accepted pairs claim slots with atomics on degree and are written.
// The endpoints with new edges ("insert_neighbor" keys) and the old degrees are
// returned; the new neighbours of u are indices[indptr[u]+old_deg[u] :
// indptr[u]+deg[u]].  Self-loops are skipped (the reference would append u
// twice to its own row).
#include <hipcub/hipcub.hpp>

#include "flsg_internal.h"

namespace flsg {

constexpr int kDT = 256;

struct DynWork {
    int64_t cap = 0;
    uint64_t *keys[2] = {nullptr, nullptr};
    int32_t *cnt = nullptr;        // n
    uint32_t *flags = nullptr;     // [0] overflow, [1] key count, [2] inserted
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    int64_t n = 0;
};

__global__ void k_dyn_canon(const int32_t *us, const int32_t *vs, int64_t k, uint64_t *keys) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < k; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t u = (uint32_t)us[e], v = (uint32_t)vs[e];
        const uint32_t a = u < v ? u : v, b = u < v ? v : u;
        keys[e] = ((uint64_t)a << 32) | b;
    }
}

// distinct, non-loop pairs that are not yet edges: count per endpoint; the
// accepted flag is kept in bit 63 of the key
__global__ void k_dyn_check(const int32_t *indptr, const int32_t *indices, const int32_t *degree, uint64_t *keys,
                            int64_t k, int32_t *cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < k; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = keys[i];
        // repeated pair: the first one wins (mask the accepted bit another thread may set)
        if (i > 0 && (keys[i - 1] & ~(1ull << 63)) == key) continue;
        const int a = (int)(key >> 32), b = (int)(key & 0xffffffffu);
        if (a == b) continue;
        const int da = degree[a], db = degree[b];
        const int x = da <= db ? a : b, y = da <= db ? b : a;
        const int st = indptr[x], dx = da <= db ? da : db;
        bool found = false;
        for (int j = 0; j < dx && !found; j++) found = indices[st + j] == y;
        if (found) continue;
        atomicAdd(&cnt[a], 1);
        atomicAdd(&cnt[b], 1);
        keys[i] = key | (1ull << 63);
    }
}

__global__ void k_dyn_capacity(const int32_t *indptr, const int32_t *degree, const int32_t *cnt, int64_t n,
                               uint32_t *flags) {
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < n; u += (int64_t)gridDim.x * blockDim.x) {
        const int c = cnt[u];
        if (c > 0 && (int64_t)degree[u] + c > (int64_t)indptr[u + 1] - indptr[u]) atomicOr(&flags[0], 1u);
    }
}

__global__ void k_dyn_commit(const int32_t *indptr, int32_t *indices, int32_t *degree, const uint64_t *keys,
                             int64_t k, uint32_t *flags) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < k; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = keys[i];
        if (!(key >> 63)) continue;
        const int a = (int)((key >> 32) & 0x7fffffffu), b = (int)(key & 0xffffffffu);
        const int sa = atomicAdd(&degree[a], 1);
        indices[indptr[a] + sa] = b;
        const int sb = atomicAdd(&degree[b], 1);
        indices[indptr[b] + sb] = a;
        atomicAdd(&flags[2], 1u);
    }
}

__global__ void k_dyn_keys(int32_t *cnt, int64_t n, int32_t *keys_out, uint32_t *flags, int clear_only) {
    for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < n; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t u = base + threadIdx.x;
        bool has = false;
        if (u < n) {
            has = cnt[u] > 0;
            if (has) cnt[u] = 0;
        }
        if (clear_only) continue;
        const long long slot = wave_append(has, &flags[1]);
        if (slot >= 0) keys_out[slot] = (int32_t)u;
    }
}

void dyn_free(flsg_graph g) {
    DynWork *w = (DynWork *)g->dyn;
    if (!w) return;
    (void)hipFree(w->keys[0]);
    (void)hipFree(w->keys[1]);
    (void)hipFree(w->cnt);
    (void)hipFree(w->flags);
    (void)hipFree(w->tmp);
    delete w;
    g->dyn = nullptr;
}

static int ensure_dyn(flsg_graph g, int64_t k, hipStream_t st) {
    DynWork *w = (DynWork *)g->dyn;
    if (w && w->cap >= k) return FLSG_OK;
    const int64_t cap = k > 1024 ? k : 1024;
    dyn_free(g);
    w = new DynWork();
    g->dyn = w;
    w->cap = cap;
    w->n = g->n;
    FLSG_HIP(hipMalloc((void **)&w->keys[0], sizeof(uint64_t) * cap));
    FLSG_HIP(hipMalloc((void **)&w->keys[1], sizeof(uint64_t) * cap));
    FLSG_HIP(hipMalloc((void **)&w->cnt, sizeof(int32_t) * g->n));
    FLSG_HIP(hipMemsetAsync(w->cnt, 0, sizeof(int32_t) * g->n, st));
    FLSG_HIP(hipMalloc((void **)&w->flags, sizeof(uint32_t) * 4));
    size_t b = 0;
    FLSG_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, b, (const uint64_t *)nullptr, (uint64_t *)nullptr, (int)cap,
                                               0, 64, st));
    w->tmp_bytes = b;
    FLSG_HIP(hipMalloc(&w->tmp, b));
    return FLSG_OK;
}

}  // namespace flsg

using namespace flsg;

extern "C" int flsg_graph_insert_edges(flsg_graph g, const int32_t *d_us, const int32_t *d_vs, int64_t k,
                                       int32_t *d_keys_out, int64_t *n_keys, int32_t *d_old_degree,
                                       int64_t *n_inserted, void *stream) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (!g->slack) return fail(FLSG_EARG, "edge insertion needs a slack CSR handle");
    if (k < 0 || (k > 0 && (!d_us || !d_vs))) return fail(FLSG_EARG, "bad edge list");
    if (k >= (int64_t)INT32_MAX) return fail(FLSG_EARG, "too many edges in one batch");
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = g->n;
    if (d_old_degree) FLSG_HIP(hipMemcpyAsync(d_old_degree, g->degree, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
    if (n_keys) *n_keys = 0;
    if (n_inserted) *n_inserted = 0;
    if (k == 0) return FLSG_OK;
    rc = ensure_dyn(g, k, st);
    if (rc) return rc;
    DynWork *w = (DynWork *)g->dyn;
    const int grid = g->num_cus * 4;
    FLSG_HIP(hipMemsetAsync(w->flags, 0, sizeof(uint32_t) * 4, st));
    hipLaunchKernelGGL(k_dyn_canon, dim3(grid), dim3(kDT), 0, st, d_us, d_vs, k, w->keys[0]);
    size_t tb = w->tmp_bytes;
    FLSG_HIP(hipcub::DeviceRadixSort::SortKeys(w->tmp, tb, w->keys[0], w->keys[1], (int)k, 0, 64, st));
    hipLaunchKernelGGL(k_dyn_check, dim3(grid), dim3(kDT), 0, st, g->indptr, g->indices, g->degree, w->keys[1], k,
                       w->cnt);
    hipLaunchKernelGGL(k_dyn_capacity, dim3(grid), dim3(kDT), 0, st, g->indptr, g->degree, w->cnt, n, w->flags);
    uint32_t hf[4] = {0, 0, 0, 0};
    FLSG_HIP(hipMemcpyAsync(hf, w->flags, sizeof(hf), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    if (hf[0]) {
        hipLaunchKernelGGL(k_dyn_keys, dim3(grid), dim3(kDT), 0, st, w->cnt, n, (int32_t *)nullptr, w->flags, 1);
        FLSG_HIP(hipStreamSynchronize(st));
        return fail(FLSG_ECAP, "a row's slack capacity would overflow; nothing was inserted");
    }
    hipLaunchKernelGGL(k_dyn_commit, dim3(grid), dim3(kDT), 0, st, g->indptr, g->indices, g->degree, w->keys[1], k,
                       w->flags);
    hipLaunchKernelGGL(k_dyn_keys, dim3(grid), dim3(kDT), 0, st, w->cnt, n, d_keys_out, w->flags,
                       d_keys_out ? 0 : 1);
    g->rows_valid = 0;          // the global sweeps' row classes follow the degrees
    sor_invalidate(g);          // so does the SOR colouring
    rc = colx_refresh(g, st);   // degree-carrying column copy of the batched path, if built
    if (rc) return rc;
    FLSG_HIP(hipMemcpyAsync(hf, w->flags, sizeof(hf), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    if (n_keys) *n_keys = hf[1];
    if (n_inserted) *n_inserted = hf[2];
    g->m_live += 2 * (int64_t)hf[2];
    return FLSG_OK;
}
