// flsg_internal.h -- shared device helpers and the graph-handle layout.
//
// Target: gfx950 (MI355X), wave64.  All kernels are hand-written HIP; the FP64
// parity path is compiled with -ffp-contract=off so every multiply/add rounds
// exactly like the reference's numba code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/flsg.h"

namespace flsg {

constexpr int kWave = 64;

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define FLSG_HIP(call)                                                               \
    do {                                                                             \
        hipError_t _e = (call);                                                      \
        if (_e != hipSuccess)                                                        \
            return ::flsg::fail(FLSG_EHIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

// ---------------------------------------------------------------- device counters
// One block of per-solve counters living in device memory.
// next-frontier sub-lists of the single-source push: kFrontSub counters per
// parity, each on its own 64-byte line (returning atomics on one address
// serialise)
constexpr int kFrontSub = 256;
constexpr int kFrontCs = 16;
constexpr int kFrontCntWords = 2 * kFrontSub * kFrontCs;

struct SolveCounters {
    unsigned int cnt[2];        // frontier size by parity
    unsigned int npieces[2];    // hub pieces by parity
    unsigned int work[2];       // dynamic work-item counter by parity
    unsigned int act[2];        // global sweeps: active count by parity
    unsigned long long oper;    // reference operation count
    unsigned long long iters;   // non-empty iterations / sweeps
    unsigned long long trace_off;
    unsigned int overflow;      // trace / piece overflow flags
    unsigned int pad;
};

// ---------------------------------------------------------------- graph handle
struct BatchArena;   // flsg_batch.hip

}  // namespace flsg

struct flsg_graph_s {
    int64_t n = 0;
    int64_t nnz_cap = 0;
    const int32_t *indptr = nullptr;
    int32_t *indices = nullptr;
    int32_t *degree = nullptr;
    int32_t slack = 0;
    int64_t m_live = 0;             // sum of degree (the reference's m)
    int64_t d_max_hint = 0;         // largest row capacity (bounds every live degree)
    int device = 0;
    int num_cus = 0;
    // single-query workspace (lazily allocated)
    int32_t *front[2] = {nullptr, nullptr};
    int32_t *rows_wave = nullptr;           // global sweeps: rows pulled by a wave / a block
    int32_t *rows_block = nullptr;
    unsigned int *rows_cnt = nullptr;
    int64_t n_rows_wave = 0, n_rows_block = 0;
    int rows_valid = 0;                     // cleared when degrees change
    int32_t *fsub = nullptr;                // next-frontier sub-lists (flsg_single.hip)
    int64_t fsub_cap = 0;
    unsigned int *fsubcnt = nullptr;        // their counters, by parity
    double *val = nullptr;          // per frontier entry (float or double)
    int64_t *pieces = nullptr;      // hub pieces: (frontier index << 32) | piece no.
    int64_t pieces_cap = 0;
    double *work_a = nullptr;       // n doubles: scratch vectors for global sweeps
    double *work_b = nullptr;
    double *work_c = nullptr;
    double *coef = nullptr;         // per-sweep coefficients (Chebyshev)
    int64_t coef_cap = 0;
    flsg::SolveCounters *ctr = nullptr;
    flsg::SolveCounters *h_ctr = nullptr;   // pinned mirror
    double *hk_lvl = nullptr;               // heat-kernel per-level sums (2 per level)
    int64_t hk_lvl_cap = 0;
    int64_t *trace_dev_off = nullptr;       // device per-iteration offsets
    int64_t trace_dev_off_cap = 0;
    // ordered local Chebyshev workspace (flsg_cheb.hip)
    void *cheb = nullptr;
    // InstantGNN propagation workspace (flsg_ignn.hip)
    void *ignn = nullptr;
    // edge-insertion workspace (flsg_dyn.hip)
    void *dyn = nullptr;
    // coloured SOR: colouring and frontier workspace (flsg_sor.hip)
    void *sor = nullptr;
    // captured iteration blocks of the single-source push (flsg_single.hip)
    // and the handle-owned x / r they run on (fixed addresses, so the
    // captured graph is reused across solves)
    void *graphs = nullptr;
    double *local_x = nullptr;
    double *local_r = nullptr;
    // batched workspace
    int2 *colx = nullptr;                   // nnz_cap column entries {v, degree[v]} (batched edge phases)
    int64_t *vol2 = nullptr;                // 2n: three-hop volume per node, then two-hop (batched claim order)
    int32_t vol2_valid = 0;                 // vol2 follows the current degrees
    flsg::BatchArena *arena = nullptr;
    int32_t batch_slots = 0;
    int32_t batch_block = 1024;
    int32_t batch_hub_bytes = -1;           // LDS hub cache per workgroup (-1 = default for the block)
    int32_t batch_hub_w = 0;                // degree bin: entries with more edges are split (0 = 256)
    int32_t batch_piece_w = 0;              // edges per piece (0 = 1024, doubled to <= 255 pieces per entry)
    int64_t batch_bin_min = 65536;          // binned push for iterations of at least this many edges
    int32_t batch_dense_min = 64;           // binned push: bins with fewer pairs use memory-side atomics
    int32_t batch_bin_busy = 0;             // binned push only with this many busy workgroups (0 = CUs / 4)
    int32_t undirected = -1;                // every edge (u, v) has its (v, u): -1 unknown, 0 no, 1 yes
    int32_t simple = -1;                    // no row repeats an entry (set with undirected)
    const int32_t *out_map = nullptr;       // internal id -> caller id for batched outputs
};

namespace flsg {

int ensure_single_workspace(flsg_graph g);
int recompute_m(flsg_graph g, hipStream_t st);
int64_t host_estimate_n(double t, double eps);
int check_graph(flsg_graph g);
void batch_free(flsg_graph g);
int colx_refresh(flsg_graph g, hipStream_t st);
// after an edge insertion (the batched path keeps no per-slot degree copies)
int batch_degrees_changed(flsg_graph g, const int32_t *d_rows, int64_t n_rows, hipStream_t st);
void cheb_free(flsg_graph g);
void ignn_free(flsg_graph g);
void dyn_free(flsg_graph g);
void sor_free(flsg_graph g);
void sor_invalidate(flsg_graph g);
void single_graphs_free(flsg_graph g);

// ---------------------------------------------------------------- device helpers
__device__ __forceinline__ unsigned int lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned long long lanemask_lt() {
    unsigned int l = __lane_id();
    return (l == 0) ? 0ull : ((~0ull) >> (64 - l));
}

// Wave-aggregated append: every lane of the wave must call it (pred may be
// false).  Returns the slot for lanes with pred, -1 otherwise.
__device__ __forceinline__ long long wave_append(bool pred, unsigned int *counter) {
    unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return -1;
    int leader = __ffsll((long long)mask) - 1;
    unsigned int base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(counter, (unsigned int)__popcll(mask));
    base = (unsigned int)__builtin_amdgcn_readlane((int)base, leader);   // no LDS round trip
    if (!pred) return -1;
    return (long long)base + __popcll(mask & lanemask_lt());
}

// The same for U predicates per lane with one atomic: slot[u] for the lanes
// with pred[u] (slots of item u follow those of items < u), -1 otherwise.
template <int U>
__device__ __forceinline__ void wave_append_n(const bool *pred, unsigned int *counter, long long *slot) {
    unsigned long long m[U];
    unsigned int tot = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
        m[u] = __ballot(pred[u]);
        tot += (unsigned int)__popcll(m[u]);
    }
    if (tot == 0) {
#pragma unroll
        for (int u = 0; u < U; u++) slot[u] = -1;
        return;
    }
    const int leader = __ffsll((long long)__ballot(1)) - 1;
    unsigned int base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(counter, tot);
    base = (unsigned int)__builtin_amdgcn_readlane((int)base, leader);
    const unsigned long long lt = lanemask_lt();
#pragma unroll
    for (int u = 0; u < U; u++) {
        slot[u] = pred[u] ? (long long)base + __popcll(m[u] & lt) : -1;
        base += (unsigned int)__popcll(m[u]);
    }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ double atomic_add_wg(double *p, double v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ float atomic_add_wg(float *p, float v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ double atomic_add_dev(double *p, double v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float atomic_add_dev(float *p, float v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block-wide inclusive scan of one int per thread (blockDim.x == NT).
// `sh` needs NT/64 ints of LDS.  Returns inclusive prefix; *total = block sum.
template <int NT>
__device__ __forceinline__ int block_incl_scan(int v, int *sh, int *total) {
    const int lane = __lane_id();
    const int wid = threadIdx.x / kWave;
    int x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) sh[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int w = (lane < NT / kWave) ? sh[lane] : 0;
#pragma unroll
        for (int off = 1; off < NT / kWave; off <<= 1) {
            int y = __shfl_up(w, off);
            if (lane >= off) w += y;
        }
        if (lane < NT / kWave) sh[lane] = w;
    }
    __syncthreads();
    int res = x + (wid > 0 ? sh[wid - 1] : 0);
    *total = sh[NT / kWave - 1];
    __syncthreads();
    return res;
}

// First index k in [lo, hi) with incl[k] > e (incl is non-decreasing).
__device__ __forceinline__ int upper_bound_lds(const int *incl, int lo, int hi, int e) {
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (incl[mid] > e) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

}  // namespace flsg
