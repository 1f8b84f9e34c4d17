// flsg_dyn.hip -- in-place edge insertion into a device-resident slack CSR,
// with the byte layout of the reference's sequential loop.
//
// Reference: updetaGraph with onlyinsert=True (InstantGNN.py:137-180,
// ppr_solver.py:723-766): for each (u, v) in batch order, skip it if v is
// already in u's live row; otherwise write v at indices[indptr[u] + degree[u]]
// and u at indices[indptr[v] + degree[v]], bump both degrees, and append v to
// insert_neighbor[u] and u to insert_neighbor[v] (a dict in first-insertion
// order).  On the device:
//   1. every pair gets its own check, "v not in u's live row" on the graph
//      before the batch, and its canonical key (min, max); endpoints outside
//      [0, n) fail the call;
//   2. a stable radix sort by canonical key groups the repeats of a pair in
//      batch order: the first repeat whose check passes is inserted and the
//      others are skipped -- once it is in, both rows hold the pair, so every
//      later repeat (either orientation) finds it, and before that no other
//      pair of the batch adds it;
//   3. each inserted pair makes two events, (u, batch position, side 0) with
//      value v and (v, position, side 1) with value u; a radix sort by
//      (row, position, side) lays each row's events out in the reference's
//      append order, so event i of row x goes to slot degree_old[x] + (i -
//      first event of x);
//   4. rows that would pass their slack capacity fail the call before anything
//      is written (the reference would overwrite the next row);
//   5. the endpoints that gained edges are returned in the order their first
//      event happened in the batch -- the key order of insert_neighbor -- and
//      the appended tail of each row is its insert_neighbor list.
// A self-loop that would be inserted (u == v, u not yet in its own row) fails
// the call: the reference writes u into one slot and raises the degree by two,
// leaving an unwritten entry inside the live row.  A self-loop already present
// is skipped as the reference skips it.
#include <hipcub/hipcub.hpp>

#include "flsg_internal.h"

namespace flsg {

constexpr int kDT = 256;
constexpr uint32_t kDynRange = 1u;     // flags[0]: an endpoint outside [0, n)
constexpr uint32_t kDynLoop = 2u;      // flags[0]: a self-loop would be inserted
constexpr uint32_t kDynCap = 4u;       // flags[0]: a row's slack capacity would overflow

struct DynWork {
    int64_t cap = 0;                // batch capacity (pairs)
    uint64_t *ka = nullptr;         // 2 * cap: sort keys (pairs, then events)
    uint64_t *kb = nullptr;         // 2 * cap
    int32_t *pa = nullptr;          // cap: batch positions (sort values)
    int32_t *pb = nullptr;          // cap
    int32_t *rows = nullptr;        // 2 * cap: rows with new edges, insert_neighbor order
    uint8_t *chk = nullptr;         // cap: v_e not in u_e's row before the batch
    uint8_t *acc = nullptr;         // cap: pair e is inserted
    int32_t *first = nullptr;       // n: index of a row's first event
    int32_t *old = nullptr;         // n: degrees before the batch
    uint32_t *flags = nullptr;      // [0] error bits, [1] rows with new edges, [2] pairs inserted
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    int64_t n = 0;
};

// per pair: range check, canonical key, and the reference's membership test
// (v in the live row of u, on the graph before the batch)
__global__ void k_dyn_check(const int32_t *us, const int32_t *vs, int64_t k, const int32_t *indptr,
                            const int32_t *indices, const int32_t *degree, int64_t n, uint64_t *keys, int32_t *pos,
                            uint8_t *chk, uint32_t *flags) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < k; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t u = us[e], v = vs[e];
        pos[e] = (int32_t)e;
        if (u < 0 || v < 0 || (int64_t)u >= n || (int64_t)v >= n) {
            atomicOr(&flags[0], kDynRange);
            keys[e] = 0;
            chk[e] = 0;
            continue;
        }
        const uint32_t a = (uint32_t)(u < v ? u : v), b = (uint32_t)(u < v ? v : u);
        keys[e] = ((uint64_t)a << 32) | b;
        const int st = indptr[u], du = degree[u];
        bool found = false;
        for (int j = 0; j < du && !found; j++) found = indices[st + j] == v;
        chk[e] = found ? 0 : 1;
        if (!found && u == v) atomicOr(&flags[0], kDynLoop);
    }
}

// sorted pairs: the first repeat of a canonical pair (batch order) whose check
// passes is inserted; its two events are keyed (row << 32) | (2 e + side)
__global__ void k_dyn_accept(const uint64_t *keys, const int32_t *pos, int64_t k, const int32_t *us,
                             const int32_t *vs, const uint8_t *chk, uint8_t *acc, uint64_t *ev, uint32_t *flags) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < k; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t e = pos[i];
        bool ok = chk[e] != 0;
        for (int64_t j = i - 1; ok && j >= 0 && keys[j] == keys[i]; j--) ok = chk[pos[j]] == 0;
        acc[e] = ok ? 1 : 0;
        uint64_t e0 = ~0ull, e1 = ~0ull;   // skipped pairs sort past every event
        if (ok) {
            atomicAdd(&flags[2], 1u);
            e0 = ((uint64_t)(uint32_t)us[e] << 32) | (2ull * (uint32_t)e);
            e1 = ((uint64_t)(uint32_t)vs[e] << 32) | (2ull * (uint32_t)e + 1ull);
        }
        ev[2 * e] = e0;
        ev[2 * e + 1] = e1;
    }
}

// sorted events [0, ne): first event index of each row, capacity check at a
// row's last event, and the row's key for the insert_neighbor order
// ((2 e + side) of its first event, then the row)
__global__ void k_dyn_rows(const uint64_t *ev, int64_t ne, const int32_t *indptr, const int32_t *old, int32_t *first,
                           uint64_t *rowkey, uint32_t *flags) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ne; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(ev[i] >> 32);
        const bool head = i == 0 || (uint32_t)(ev[i - 1] >> 32) != x;
        if (head) {
            first[x] = (int32_t)i;
            rowkey[atomicAdd(&flags[1], 1u)] = ((ev[i] & 0xffffffffull) << 32) | x;
        }
    }
}

__global__ void k_dyn_capacity(const uint64_t *ev, int64_t ne, const int32_t *indptr, const int32_t *old,
                               const int32_t *first, uint32_t *flags) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ne; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(ev[i] >> 32);
        const bool tail = i + 1 == ne || (uint32_t)(ev[i + 1] >> 32) != x;
        if (tail && (int64_t)old[x] + (i - first[x] + 1) > (int64_t)indptr[x + 1] - indptr[x])
            atomicOr(&flags[0], kDynCap);
    }
}

// event i of row x: the other endpoint of its pair into slot old[x] + (i - first[x]);
// a row's last event sets its new degree
__global__ void k_dyn_commit(const uint64_t *ev, int64_t ne, const int32_t *us, const int32_t *vs,
                             const int32_t *indptr, const int32_t *old, const int32_t *first, int32_t *indices,
                             int32_t *degree) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ne; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = ev[i];
        const uint32_t x = (uint32_t)(key >> 32);
        const uint32_t lo = (uint32_t)(key & 0xffffffffull);
        const uint32_t e = lo >> 1;
        const int32_t val = (lo & 1u) ? us[e] : vs[e];
        const int slot = old[x] + (int)(i - first[x]);
        indices[indptr[x] + slot] = val;
        const bool tail = i + 1 == ne || (uint32_t)(ev[i + 1] >> 32) != x;
        if (tail) degree[x] = slot + 1;
    }
}

__global__ void k_dyn_keys_out(const uint64_t *rowkey, int64_t nr, int32_t *keys_out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nr; i += (int64_t)gridDim.x * blockDim.x)
        keys_out[i] = (int32_t)(rowkey[i] & 0xffffffffull);
}

void dyn_free(flsg_graph g) {
    DynWork *w = (DynWork *)g->dyn;
    if (!w) return;
    (void)hipFree(w->ka);
    (void)hipFree(w->kb);
    (void)hipFree(w->pa);
    (void)hipFree(w->pb);
    (void)hipFree(w->rows);
    (void)hipFree(w->chk);
    (void)hipFree(w->acc);
    (void)hipFree(w->first);
    (void)hipFree(w->old);
    (void)hipFree(w->flags);
    (void)hipFree(w->tmp);
    delete w;
    g->dyn = nullptr;
}

static int ensure_dyn(flsg_graph g, int64_t k) {
    DynWork *w = (DynWork *)g->dyn;
    if (w && w->cap >= k && w->n == g->n) return FLSG_OK;
    const int64_t cap = k > 1024 ? k : 1024;
    dyn_free(g);
    w = new DynWork();
    g->dyn = w;
    w->cap = cap;
    w->n = g->n;
    FLSG_HIP(hipMalloc((void **)&w->ka, sizeof(uint64_t) * 2 * cap));
    FLSG_HIP(hipMalloc((void **)&w->kb, sizeof(uint64_t) * 2 * cap));
    FLSG_HIP(hipMalloc((void **)&w->pa, sizeof(int32_t) * cap));
    FLSG_HIP(hipMalloc((void **)&w->pb, sizeof(int32_t) * cap));
    FLSG_HIP(hipMalloc((void **)&w->rows, sizeof(int32_t) * 2 * cap));
    FLSG_HIP(hipMalloc((void **)&w->chk, cap));
    FLSG_HIP(hipMalloc((void **)&w->acc, cap));
    FLSG_HIP(hipMalloc((void **)&w->first, sizeof(int32_t) * (g->n > 0 ? g->n : 1)));
    FLSG_HIP(hipMalloc((void **)&w->old, sizeof(int32_t) * (g->n > 0 ? g->n : 1)));
    FLSG_HIP(hipMalloc((void **)&w->flags, sizeof(uint32_t) * 4));
    size_t b1 = 0, b2 = 0;
    FLSG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                                (const int32_t *)nullptr, (int32_t *)nullptr, (int)cap, 0, 64));
    FLSG_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, b2, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                               (int)(2 * cap), 0, 64));
    w->tmp_bytes = b1 > b2 ? b1 : b2;
    FLSG_HIP(hipMalloc(&w->tmp, w->tmp_bytes));
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
    if (k >= (int64_t)INT32_MAX / 2) return fail(FLSG_EARG, "too many edges in one batch");
    if (g->n >= ((int64_t)1 << 31)) return fail(FLSG_EARG, "edge insertion needs n < 2^31");
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = g->n;
    if (d_old_degree) FLSG_HIP(hipMemcpyAsync(d_old_degree, g->degree, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
    if (n_keys) *n_keys = 0;
    if (n_inserted) *n_inserted = 0;
    if (k == 0) return FLSG_OK;
    rc = ensure_dyn(g, k);
    if (rc) return rc;
    DynWork *w = (DynWork *)g->dyn;
    const int grid = g->num_cus * 4;
    uint32_t hf[4] = {0, 0, 0, 0};
    FLSG_HIP(hipMemsetAsync(w->flags, 0, sizeof(uint32_t) * 4, st));
    FLSG_HIP(hipMemcpyAsync(w->old, g->degree, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_dyn_check, dim3(grid), dim3(kDT), 0, st, d_us, d_vs, k, (const int32_t *)g->indptr,
                       (const int32_t *)g->indices, (const int32_t *)g->degree, n, w->ka, w->pa, w->chk, w->flags);
    FLSG_HIP(hipGetLastError());
    // repeats of a pair side by side, in batch order (radix sort is stable)
    size_t tb = w->tmp_bytes;
    FLSG_HIP(hipcub::DeviceRadixSort::SortPairs(w->tmp, tb, w->ka, w->kb, w->pa, w->pb, (int)k, 0, 64, st));
    hipLaunchKernelGGL(k_dyn_accept, dim3(grid), dim3(kDT), 0, st, (const uint64_t *)w->kb, (const int32_t *)w->pb,
                       k, d_us, d_vs, (const uint8_t *)w->chk, w->acc, w->ka, w->flags);
    FLSG_HIP(hipGetLastError());
    // events by (row, batch position, side): each row's appends in the reference's order
    tb = w->tmp_bytes;
    FLSG_HIP(hipcub::DeviceRadixSort::SortKeys(w->tmp, tb, w->ka, w->kb, (int)(2 * k), 0, 64, st));
    FLSG_HIP(hipMemcpyAsync(hf, w->flags, sizeof(hf), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    if (hf[0] & kDynRange) return fail(FLSG_EARG, "edge endpoint outside [0, n); nothing was inserted");
    if (hf[0] & kDynLoop)
        return fail(FLSG_EARG, "a self-loop would be inserted (the reference leaves an unwritten entry in the "
                               "live row for it); nothing was inserted");
    const int64_t ne = 2 * (int64_t)hf[2];   // two events per inserted pair
    if (ne == 0) return FLSG_OK;
    hipLaunchKernelGGL(k_dyn_rows, dim3(grid), dim3(kDT), 0, st, (const uint64_t *)w->kb, ne,
                       (const int32_t *)g->indptr, (const int32_t *)w->old, w->first, w->ka, w->flags);
    hipLaunchKernelGGL(k_dyn_capacity, dim3(grid), dim3(kDT), 0, st, (const uint64_t *)w->kb, ne,
                       (const int32_t *)g->indptr, (const int32_t *)w->old, (const int32_t *)w->first, w->flags);
    FLSG_HIP(hipGetLastError());
    FLSG_HIP(hipMemcpyAsync(hf, w->flags, sizeof(hf), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    if (hf[0] & kDynCap) return fail(FLSG_ECAP, "a row's slack capacity would overflow; nothing was inserted");
    hipLaunchKernelGGL(k_dyn_commit, dim3(grid), dim3(kDT), 0, st, (const uint64_t *)w->kb, ne, d_us, d_vs,
                       (const int32_t *)g->indptr, (const int32_t *)w->old, (const int32_t *)w->first, g->indices,
                       g->degree);
    FLSG_HIP(hipGetLastError());
    const int64_t nr = hf[1];
    // rows in the order of their first event: the insert_neighbor key order
    tb = w->tmp_bytes;
    FLSG_HIP(hipcub::DeviceRadixSort::SortKeys(w->tmp, tb, w->ka, w->kb, (int)nr, 0, 64, st));
    hipLaunchKernelGGL(k_dyn_keys_out, dim3(grid), dim3(kDT), 0, st, (const uint64_t *)w->kb, nr, w->rows);
    FLSG_HIP(hipGetLastError());
    if (d_keys_out) FLSG_HIP(hipMemcpyAsync(d_keys_out, w->rows, sizeof(int32_t) * nr, hipMemcpyDeviceToDevice, st));
    g->rows_valid = 0;          // the global sweeps' row classes follow the degrees
    sor_invalidate(g);          // so does the SOR colouring
    rc = colx_refresh(g, st);   // degree-carrying column copy of the batched path, if built
    if (rc) return rc;
    // the batched slots' degree copies of the rows that changed
    rc = batch_degrees_changed(g, w->rows, nr, st);
    if (rc) return rc;
    FLSG_HIP(hipStreamSynchronize(st));
    if (n_keys) *n_keys = nr;
    if (n_inserted) *n_inserted = ne / 2;
    g->m_live += ne;
    return FLSG_OK;
}
