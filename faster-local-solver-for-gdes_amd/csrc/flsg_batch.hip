// flsg_batch.hip -- many queries in one persistent launch: PPR / Katz local
// push (FP32 with an FP64 threshold test and FP64 certificate, or FP64) and the
// heat-kernel level push, mixed freely within a launch.
//
// Layout.  Workgroup b of a launch owns slot (slot_base + b) for the whole
// launch.  A slot holds, per node (and for 64 sink entries after the n nodes,
// the targets of idle lanes):
//     rx32  {float r, x}    FP32 push state: residual and estimate, 8 bytes
//     rx64  {double r, x}   FP64 push state (FP64 push, FP32 certificate)
//     fl    uint32          certificate flags
//     hk2   double          heat-kernel level residuals (FP32: two levels as floats)
// so the FP32 push moves 8-byte node records -- half of what the FP64 push
// moves -- with a node's residual and estimate on one line, and the
// certificate's FP64 fields stay out of its lines; degrees and row starts
// come from the graph's own arrays (shared by every slot).
// Every field is zero between queries: a query lists each entry it touches
// and resets exactly those.  Launches with disjoint slot ranges may run
// concurrently on different streams.
//
// Push of one PPR / Katz iteration (gradient descent = Jacobi push,
// ppr_solver.py:246-288 / katz_solver.py:183-224): the node phase moves r_u to
// x_u and leaves each frontier entry's push value; then the edges, in one of
// two ways chosen per iteration:
//   * direct: each edge is one atomic add into r_v -- an LDS atomic for the
//     hub ids [0, H) the LDS hub cache holds (degree-ranked ids: the hubs),
//     a memory-side atomic for the rest (every global atomic on this part is
//     executed on the memory side, at ~20 G/s whatever its locality);
//   * binned (iterations of at least bin_min edges): the edges are written as
//     (v, value) pairs into one bin per destination range of R = 2^s ids (a
//     count pass, a scan, a write pass), and each bin is then summed into an
//     LDS slice of R residuals (LDS atomics) and added to r in one coalesced
//     sweep over the touched part of the range -- sparse bins (fewer than
//     dense_min pairs) take memory-side atomics instead.  An iteration that
//     touches ~10% of the graph costs ~30 streamed bytes per edge instead of
//     a memory-side atomic per edge.
// v joins the next frontier iff the add that moved r_v from below thr_v to at
// least thr_v = eps*alpha*d_v (FP64 compare) was this one: pushes are >= 0,
// so exactly one add sees the crossing (a bin's summed add counts as one).
// The FP32 path then certifies its result in FP64 (below), and one pass over
// the touched entries writes the certificate value, the output and the reset.
// Queries are claimed through a per-launch counter, so uneven queries balance
// across workgroups.
#include <math.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <type_traits>
#include <mutex>
#include <utility>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "flsg_internal.h"

namespace flsg {

constexpr int kMaxHkTerms = 96;      // heat-kernel Taylor order cap of a batched launch
constexpr unsigned kInQ = 1u;        // flags: in the correction queue
constexpr unsigned kSeen = 2u;       // flags: listed in the support by the correction push (x32 == 0)
constexpr unsigned kTouched = 4u;    // flags: in the query's touched list (reset at the end)
constexpr unsigned kExact = 8u;      // flags: r64 is the exact FP64 residual of x32 (no FP32 error bound)
constexpr unsigned kR64 = 16u;       // flags: the entry's FP64 fields may be nonzero (read and reset at the end)
constexpr int kSink = 64;            // per-slot sink entries (degree 0) for the idle lanes of an edge batch
constexpr int kBinShiftMin = 11;     // binned push: at least 2^11 ids per destination range
constexpr int kMaxBins = 512;        // binned push: at most this many destination ranges (static LDS)

// Boundary certificate of the FP32 push.  A node the push reached but never
// pushed (a boundary node) holds r32, the FP32 sum of the m push values it
// received.  m <= d * iters (undirected graph: a node's in-degree is its
// degree d; each neighbour pushes at most once per iteration), and on a graph
// with no repeated entry in a row also m <= F, the query's pushes (frontier
// entries over all iterations): a push of u adds one value to each entry of
// u's row, so at most one to v.  adds_max(d, iters, F) is that minimum (F =
// +inf where rows may repeat an entry).  Against the FP64 residual of the
// returned x -- the pushed nodes' x32 -- r32 is off by at most
//     g(m, iters) * r32 + (m + 1) * 2^-125,
// with u = 2^-24, u' = u + 2^-51, gm(k) = 2ku / (1 - 2ku) and
//     g = (gm(iters) + u' + gm(m) + u' gm(m)) / (1 - u' - gm(m) - u' gm(m)),
// because, writing a_j = c_u rho for the exact push terms (rho the FP32
// residuals the neighbours moved) and A = sum a_j:
//   - every FP32 add (memory-side or LDS atomic, or a bin's summed add) lands
//     within 2u of its exact result, and r32 is a summation tree of its m
//     terms whatever order the adds took, so each term is carried within gm(m);
//   - every push value fl32(fl64(fl64(1 - alpha) rho) / d) is within u' of a_j;
//   - so r32 = sum a_j (1 + eta_j), |eta_j| <= u' + gm(m) + u' gm(m), and the
//     true residual sum_u c_u x32_u = sum a_j (1 + theta_j), |theta_j| <=
//     gm(iters) (each x32_u an FP32 sum of at most iters terms), whence
//     |r - r32| <= A (gm(iters) + max|eta|) and A <= r32 / (1 - max|eta|);
//   - a flushed denormal result loses at most 2^-126 per add.
// It holds while 2 m u < 1; lazy_valid asks 2 m u <= 1/4.  Such a node keeps
// its FP32 residual (plus the bound) when that certifies it; the rest are
// recomputed in FP64 like the support.
constexpr int kLazyDeg = 64;                 // boundary nodes of higher degree are recomputed
constexpr double kNearFrac = 0.99609375;     // 1 - 2^-8: boundary residuals above this fraction of
                                             // the threshold are recomputed
constexpr double kLazyGmax = 0.0039;         // g(kLazyDeg, iters) limit: kNearFrac * (1 + g) < 1 - 1e-6

__host__ __device__ __forceinline__ double lazy_gm(double k) {
    const double t = 2.0 * k * 0x1p-24;
    return t / (1.0 - t);
}

// the most push values a node of degree d can have received (see above)
__host__ __device__ __forceinline__ double adds_max(int d, int64_t it, double F) {
    return fmin((double)d * (double)it, F);
}

__host__ __device__ __forceinline__ double lazy_g(double m, int64_t it) {
    const double up = 0x1p-24 + 0x1p-51;
    const double gm = lazy_gm(m);
    const double eta = up + gm + up * gm;
    return (lazy_gm((double)it) + eta) / (1.0 - eta);
}

// where lazy_g is used: 2 m u <= 1/4
__host__ __device__ __forceinline__ bool lazy_valid(double m) { return 2.0 * m * 0x1p-24 <= 0.25; }

// error bound of an entry whose FP32 residual field holds rl: > 0 a boundary
// residual still pending, < 0 one the correction push has moved into x (its
// error stays), +-0 an exact (FP64) entry
// (m: adds_max of the entry)
__host__ __device__ __forceinline__ double lazy_bound(float rl, double m, int64_t it) {
    if (rl == 0.0f) return 0.0;
    if (!lazy_valid(m)) return __builtin_inf();   // never certifies
    return lazy_g(m, it) * fabs((double)rl) + (m + 1.0) * 0x1p-125;
}

// Error bound of a pushed node v certified without the FP64 pull.  The true
// residual of the returned x32 is r_v = b_v - x32_v + S_v with
// S_v = sum over neighbours u of c_u x32_u (c_u = (1-alpha)/d_u, Katz beta),
// and the FP32 push tracked r32_v; with u = 2^-24, K pushes at most per node
// and every FP32 add within 2u of its exact result:
//   r_v - r32_v = -db_v - H_v - dx_v - E_v + sum_u c_u dx_u, where
//   db_v  = fl32(b_v) - b_v                       |db| <= 2u b
//   dx_v  = x32_v - (sum of the r moved out of v)   |dx_v| <= 2uK x32_v (1 + 4uK)
//   E_v   = received push values - sum_u c_u (moved out of u)
//   H_v   = rounding of every add into r32_v    |H_v| <= 2u m max(x32_v(1+4uK), r32_v) + m 2^-126
// (at most m = adds_max(d, K, F) adds: one per received push value, a bin's
// summed add included; every partial sum of r32_v is at most the residual the
// node next moves into x32_v, or its final r32_v), and |E_v| + sum_u c_u |dx_u| <= (2K + 1.01) u (1 + 4uK) S_v,
// while S_v = r_v - b_v + x32_v <= |r32_v| + B + x32_v -- the residual
// identity itself bounds the neighbour sum, so no gather is needed.  Solving
// for B:
//   B (1 - g) <= 2u b + 2uK x (1+4uK) + g (|r32| + x) + 2u m max(x(1+4uK), |r32|) + m 2^-126,
//   g = (2K + 1.01) u (1 + 4uK),
// returned with a 1.01 margin.  It is small next to eps*alpha*d_v unless
// x32_v is large (about x32 > 2e-3 at eps 1e-7 and m = d K): the seed and
// its near neighbourhood stay in the FP64 pull; the cap m <= F lets hubs, whose
// d K far exceeds the query's pushes, keep their FP32 values.
__host__ __device__ __forceinline__ double sup_bound(float x32, float rl, double m, int64_t K, double b) {
    const double u = 0x1p-24;
    const double x = fabs((double)x32), r = fabs((double)rl), k = (double)K;
    const double c = 1.0 + 4.0 * u * k;
    const double g = (2.0 * k + 1.01) * u * c;
    const double num = 2.0 * u * b + 2.0 * u * k * x * c + g * (r + x) + 2.0 * u * m * fmax(x * c, r) +
                       m * 0x1p-126;
    return 1.01 * num / (1.0 - g);
}

// Error bound of an entry of the FP32 certificate, from its flags fl, FP32
// residual field rl (> 0: pending, < 0: moved into x64 by the correction push,
// the bound staying that of |rl|), FP32 estimate x32 and b = b_v: 0 for an
// exact entry (r64 recomputed in FP64), sup_bound for a pushed node (x32 > 0),
// the boundary bound otherwise; F caps the adds (adds_max).  Pure function of
// the entry and the query's push: the certificate, the correction push and the
// output pass compute the same value.
__host__ __device__ __forceinline__ double rec_bound(uint32_t fl, float rl, float x32, int d, int64_t it,
                                                     double b, double F) {
    if (fl & kExact) return 0.0;
    const double m = adds_max(d, it, F);
    if (x32 != 0.0f) return sup_bound(x32, rl, m, it, b);
    return lazy_bound(rl, m, it);
}

// One slot's arrays (see the file header).
struct SlotView {
    float2 *rx32;
    uint32_t *fl;
    double2 *rx64;
    double *hk2;
};

// an array of T with a stride: the residual / estimate halves of a record
// array, or a plain array
template <typename T>
struct Strided {
    T *p;
    int s;
    __device__ __forceinline__ T &operator[](int64_t i) const { return p[i * s]; }
};

// {residual, estimate} of one node as one vector access (one load or store
// instruction instead of two to the same record)
template <typename T>
struct Rec2;
template <>
struct Rec2<float> {
    using type = float2;
};
template <>
struct Rec2<double> {
    using type = double2;
};
template <typename T>
__device__ __forceinline__ typename Rec2<T>::type rec2(T r, T x) {
    return typename Rec2<T>::type{r, x};
}

// A support-list entry of the PPR / Katz push: the node id (< 2^24) with its
// degree in the top 8 bits (255: 255 or more, read from the graph), so the
// certificate's scan of the list needs no degree load for most entries.
__device__ __forceinline__ int32_t sup_entry(int v, int d) {
    return (int32_t)((uint32_t)v | ((uint32_t)min(d, 255) << 24));
}

// LDS add that stops at 2^32 - 1
__device__ __forceinline__ void atomic_add_sat(unsigned int *p, unsigned int v) {
    const unsigned int old = atomicAdd(p, v);
    if (old + v < old) atomicMax(p, ~0u);
}

// Whether sup_bound certifies a pushed node {rl, x32}: |r32| + bound <
// (1 - 2^-8) thr.  Such a node keeps its FP32 residual and estimate as they
// are (nothing is written: the correction push and the output pass recompute
// the same bound from the entry); the others join the FP64 pull.
__device__ __forceinline__ bool support_certified(float rl, float x32, int d, int64_t iters, double b,
                                                            double thr_scale, double F) {
    const double bnd = sup_bound(x32, rl, adds_max(d, iters, F), iters, b);
    return fabs((double)rl) + bnd < kNearFrac * thr_scale * (double)d;
}

// rec_bound out of line (the correction push and the output pass)
__device__ __attribute__((noinline)) double rec_bound_ool(uint32_t fl, float rl, float x32, int d, int64_t it,
                                                          double b, double F) {
    return rec_bound(fl, rl, x32, d, it, b, F);
}

// A frontier entry as its node phase leaves it for the edge phase: row start,
// edge count (0 = nothing to push) and push value -- one 16-byte load.
struct __attribute__((aligned(16))) Ent {
    int32_t st;
    int32_t d;
    double v;
};

struct BatchArena {
    int64_t n = 0;
    int32_t slots = 0;
    int64_t rn = 0;              // per-node array stride: n + kSink
    float2 *rx32 = nullptr;      // slots * rn
    uint32_t *fl = nullptr;      // slots * rn
    double2 *rx64 = nullptr;     // slots * rn
    double *hk2 = nullptr;       // slots * rn
    int32_t *q0 = nullptr;       // slots * n
    int32_t *q1 = nullptr;
    Ent *ent = nullptr;          // slots * n: frontier entries (heat kernel: two touched lists)
    int32_t *sup = nullptr;      // slots * n
    uint32_t *pcs = nullptr;     // slots * pcs_stride: hub pieces of the current list
    int64_t pcs_stride = 0;
    int32_t *tl = nullptr;       // slots * n: entries touched by the current query
    int32_t *nl = nullptr;       // slots * n: FP32 push: boundary candidates for an exact recompute
    uint4 *pairs = nullptr;      // slots * pcap 16-byte entries: binned push pairs (FP32 pairs
                                 // are 8 bytes and use the first half of a slot's entries)
    int64_t pcap = 0;            // pairs per slot (0: no binned push)
    uint64_t *ctl = nullptr;     // launch control for the synchronous entry point
    flsg_result *d_res = nullptr;
    int64_t d_res_cap = 0;
    flsg_query *d_q = nullptr;
    int64_t d_q_cap = 0;
    int32_t *d_order = nullptr;  // claim order of the synchronous entry point
    int64_t *d_cost = nullptr;   // query cost estimates (for that order)
};

struct BatchArgs {
    const int32_t *indptr;
    const int32_t *indices;
    const int32_t *degree;
    const int2 *colx;            // {v, degree[v]} per column entry
    int64_t n;
    const flsg_query *queries;   // per-query (seed, kind, alpha, eps)
    int64_t nq;
    const int32_t *order;        // k-th claim runs query order[k] (NULL = k)
    float2 *rx32;                // slot arrays (see BatchArena)
    uint32_t *fl;
    double2 *rx64;
    double *hk2;
    int64_t rn;
    int32_t *q0;
    int32_t *q1;
    Ent *ent;
    int32_t *sup;
    uint32_t *pcs;
    int64_t pcs_stride;
    int32_t *tl;
    int32_t *nl;
    uint4 *pairs;
    int64_t pcap;
    int64_t bin_min;             // binned push for iterations of at least this many edges
    int32_t dense_min;           // bins with fewer pairs use memory-side atomics
    int32_t piece_w;             // edges per hub piece
    int32_t hub_w;               // entries with more edges go through pieces
    int32_t hub_bytes;           // dynamic LDS: hub cache / bin slice (0 = none)
    int32_t slot_base;
    int32_t undirected;          // 1 if every edge (u, v) has its (v, u): FP32 certificate precondition
    int32_t simple;              // 1 if no row repeats an entry: the FP32 bounds cap adds by the pushes
    uint64_t *ctl;               // [0] next query (low 32 bits used), [1] output entries used
    // workgroups of the launch running a query (NULL: not counted); an
    // iteration is binned only while at least bin_busy_min of them run: the
    // binned push pays off when many slots push at once (memory-side atomics
    // are the chip-wide limit), the direct push is quicker when few do
    unsigned int *busy;
    int32_t bin_busy_min;
    const int32_t *out_map;      // internal id -> caller id (NULL = identity)
    int32_t *out_ids;
    double *out_vals;
    int64_t *out_off;
    int64_t out_cap;
    flsg_result *res;
    int64_t max_iters;
    // serving session (all NULL for a launch over a fixed query list): queries
    // [0, *pub) are published by the host into host-coherent memory while the
    // launch runs; *closed != 0 ends it once every published query is claimed;
    // done[q] is set (system scope) when query q's record and output are
    // written; a workgroup that sees nothing published for idle_ticks (100 MHz)
    // exits on its own and counts itself in ctl[2] (ctl[3], ctl[4]: see
    // session_claim)
    const uint64_t *pub;
    const uint32_t *closed;
    uint32_t *done;
    uint64_t idle_ticks;
    // heat kernel (shared by the launch's HK queries)
    int32_t hk_n;                // Taylor order N
    double hk_tau;
    double hk_ete;               // exp(tau) * eps
    double hk_psi[kMaxHkTerms + 1];
};

struct QParams {
    int kind;
    bool fp64;                   // FP64 push (certificate by construction) instead of FP32 + recompute
    double alpha, om_alpha, thr_scale, r0;
};

__device__ __forceinline__ QParams qparams(const flsg_query &q) {
    QParams p;
    p.kind = q.kind;
    p.fp64 = q.precision == FLSG_FP64;
    p.alpha = q.alpha;
    p.om_alpha = 1.0 - q.alpha;
    p.thr_scale = q.kind == FLSG_PPR ? q.eps * q.alpha : q.eps;   // ppr_solver.py:221 / katz_solver.py:179
    p.r0 = q.kind == FLSG_PPR ? q.alpha : 1.0;                     // ppr_solver.py:219 / katz_solver.py:177
    return p;
}

__device__ __forceinline__ double bpush_value(const QParams &p, double tmp, int d) {
    return p.kind == FLSG_PPR ? p.om_alpha * tmp / (double)d : p.alpha * tmp;
}

template <int NT>
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v, unsigned long long *sh) {
    v = wave_sum(v);
    const int wid = threadIdx.x / kWave;
    if (__lane_id() == 0) sh[wid] = v;
    __syncthreads();
    unsigned long long t = 0;
    for (int w = 0; w < NT / kWave; w++) t += sh[w];
    __syncthreads();
    return t;
}

template <int NT>
__device__ __forceinline__ double block_sum_f64(double v, double *sh) {
    v = wave_sum(v);
    const int wid = threadIdx.x / kWave;
    if (__lane_id() == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < NT / kWave; w++) t += sh[w];
    __syncthreads();
    return t;
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double *sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    const int wid = threadIdx.x / kWave;
    if (__lane_id() == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < NT / kWave; w++) t = fmax(t, sh[w]);
    __syncthreads();
    return t;
}

__device__ __forceinline__ uint64_t now_ticks() { return __builtin_amdgcn_s_memrealtime(); }

__device__ __forceinline__ uint32_t fl_or(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t fl_and(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// a value other waves of the workgroup wrote before the last barrier (a
// vector load: the scalar cache is not kept coherent with vector stores)
__device__ __forceinline__ int32_t load_wg(const int32_t *p) {
    return __hip_atomic_load(const_cast<int32_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

constexpr int kUnroll = 4;           // edges in flight per lane
constexpr int kOutUnroll = 8;        // entries in flight per lane in the output pass
constexpr int kScanUnroll = 4;       // entries in flight per lane in the certificate's support scan
constexpr int kBinUnroll = 4;        // pairs / destinations in flight per lane in the binned push

struct WaveLds {
    int incl[kWave];
    int start[kWave];
    double val[kWave];
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// appends of U touched destinations per lane: bit 0 of bits[u] -> next[]
// (counter *s_next), bit 1 -> tl[] (*s_tl; the entry vt[u], the id with its
// degree, sup_entry), bit 2 -> nl[] (*s_nl, entries past nl_cap dropped but
// counted); one LDS atomic per list and wave; every lane of the wave calls it
template <int U>
__device__ __forceinline__ void append_bits(const int *bits, const int *v, const int *vt, int32_t *next,
                                            unsigned int *s_next,
                                            int32_t *tl, unsigned int *s_tl, int32_t *nl, unsigned int *s_nl,
                                            unsigned int nl_cap) {
    bool pr[U];
    long long sl[U];
#pragma unroll
    for (int u = 0; u < U; u++) pr[u] = (bits[u] & 1) != 0;
    wave_append_n<U>(pr, s_next, sl);
#pragma unroll
    for (int u = 0; u < U; u++)
        if (sl[u] >= 0) next[sl[u]] = v[u];
#pragma unroll
    for (int u = 0; u < U; u++) pr[u] = (bits[u] & 2) != 0;
    wave_append_n<U>(pr, s_tl, sl);
#pragma unroll
    for (int u = 0; u < U; u++)
        if (sl[u] >= 0) tl[sl[u]] = vt[u];
    if (s_nl) {
#pragma unroll
        for (int u = 0; u < U; u++) pr[u] = (bits[u] & 4) != 0;
        wave_append_n<U>(pr, s_nl, sl);
#pragma unroll
        for (int u = 0; u < U; u++)
            if (sl[u] >= 0 && sl[u] < (long long)nl_cap) nl[sl[u]] = v[u];
    }
}

// Edges of the entries [0, cnt) of a list, wavefront by wavefront with no
// block barrier: a wave claims either a group of 64 consecutive entries (its
// edges are flattened by a wave prefix sum; entries with more than hub_w edges
// contribute nothing there) or one piece of a hub entry (pcs[], built by the
// caller: (entry << 8) | piece).
// Column entries are read from colx ({v, degree[v]}, 8 bytes, coalesced), so
// the threshold test of an edge needs no load of its own.
//   load(i, st, d, v): row start, edge count (0 = nothing) and value of entry i
//   pre(vtx, v, S &, stage): issue one edge's LDS (stage 0) or memory-side
//                           (stage 1) operations, results into S
//   post(vtx, dv, v, S, act) -> int: append bits (append_bits); the result is
//                           masked by act, side effects must test it
// A lane issues pre() for all its kUnroll edges before post() consumes any
// result, so kUnroll atomics per lane are in flight at once.  Both run
// unpredicated (a predicated atomic whose result is used makes the compiler
// wait inside the predicated block, one round trip per edge): an idle lane
// aims at its own sink entry (id sink + lane, degree 0) with value -0.0,
// which leaves any value unchanged under IEEE addition.  Touched-list (tl)
// entries carry the node's degree (sup_entry); with kEnc the next[] and nl[]
// entries do too.
template <typename S, bool kEnc = false, typename Load, typename Pre, typename Post>
__device__ __forceinline__ void wave_edges(unsigned int cnt, const uint32_t *pcs, unsigned int npc, int piece_w, int hub_w,
                                           unsigned int *s_work, WaveLds &W, const int2 *colx, Load load,
                                           Pre pre, Post post, int32_t *next, unsigned int *s_next, int32_t *tl,
                                           unsigned int *s_tl, int sink, int32_t *nl = nullptr,
                                           unsigned int *s_nl = nullptr, unsigned int nl_cap = 0) {
    const int lane = __lane_id();
    const unsigned int ngroups = (cnt + kWave - 1) / kWave;
    const unsigned int nitems = ngroups + npc;
    for (;;) {
        unsigned int item = 0;
        if (lane == 0) item = atomicAdd(s_work, 1u);
        item = __shfl(item, 0);
        if (item >= nitems) break;
        int st = 0, d = 0, beg = 0, end = 0, tot = 0;
        double v = 0.0;
        const bool group = item < ngroups;
        if (group) {
            const unsigned int i = item * kWave + lane;
            if (i < cnt) {
                load(i, st, d, v);
                if (d > hub_w) d = 0;
            }
            int incl = d;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const int y = __shfl_up(incl, o);
                if (lane >= o) incl += y;
            }
            tot = __shfl(incl, kWave - 1);
            W.incl[lane] = incl;
            W.start[lane] = st;
            W.val[lane] = v;
            wave_sync();
            end = tot;
        } else {
            const uint32_t pc = pcs[item - ngroups];
            load(pc >> 8, st, d, v);
            beg = (int)(pc & 255u) * piece_w;
            end = min(d, beg + piece_w);
        }
        for (int e0 = beg; e0 < end; e0 += kWave * kUnroll) {
            int vtx[kUnroll], dv[kUnroll];
            double vv[kUnroll];
            bool act[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const int e = e0 + u * kWave + lane;
                act[u] = e < end;
                int2 c = make_int2(sink + lane, 0);
                vv[u] = -0.0;
                if (act[u]) {
                    vv[u] = v;
                    if (group) {
                        const int k = upper_bound_lds(W.incl, 0, kWave, e);
                        const int o = e - (k ? W.incl[k - 1] : 0);
                        c = colx[W.start[k] + o];
                        vv[u] = W.val[k];
                    } else {
                        c = colx[st + e];
                    }
                }
                vtx[u] = c.x;
                dv[u] = c.y;
            }
            // stage 0 issues the LDS-side operations of all kUnroll edges, stage 1
            // the memory-side ones: a result register the compiler shares between
            // the two kinds then only ever waits for an LDS op, never for memory
            S res[kUnroll];
#pragma unroll
            for (int st2 = 0; st2 < 2; st2++) {
#pragma unroll
                for (int u = 0; u < kUnroll; u++) pre(vtx[u], vv[u], res[u], st2);
            }
            int cr[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) cr[u] = post(vtx[u], dv[u], vv[u], res[u], act[u]) & (act[u] ? 7 : 0);
            int ev[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) ev[u] = sup_entry(vtx[u], dv[u]);
            append_bits<kUnroll>(cr, kEnc ? ev : vtx, ev, next, s_next, tl, s_tl, nl, s_nl, nl_cap);
        }
        if (group) wave_sync();
    }
}

// The same walk over the edges of a list's entries for the binned push's count
// and write passes: f(col, value, act) for every edge slot (act false for the
// idle ones), reading the 4-byte column indices.
template <typename Load, typename F>
__device__ __forceinline__ void wave_cols(unsigned int cnt, const uint32_t *pcs, unsigned int npc, int piece_w,
                                          int hub_w, unsigned int *s_work, WaveLds &W, const int32_t *cols, Load load,
                                          F f) {
    const int lane = __lane_id();
    const unsigned int ngroups = (cnt + kWave - 1) / kWave;
    const unsigned int nitems = ngroups + npc;
    for (;;) {
        unsigned int item = 0;
        if (lane == 0) item = atomicAdd(s_work, 1u);
        item = __shfl(item, 0);
        if (item >= nitems) break;
        int st = 0, d = 0, beg = 0, end = 0;
        double v = 0.0;
        const bool group = item < ngroups;
        if (group) {
            const unsigned int i = item * kWave + lane;
            if (i < cnt) {
                load(i, st, d, v);
                if (d > hub_w) d = 0;
            }
            int incl = d;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const int y = __shfl_up(incl, o);
                if (lane >= o) incl += y;
            }
            W.incl[lane] = incl;
            W.start[lane] = st;
            W.val[lane] = v;
            wave_sync();
            end = __shfl(incl, kWave - 1);
        } else {
            const uint32_t pc = pcs[item - ngroups];
            load(pc >> 8, st, d, v);
            beg = (int)(pc & 255u) * piece_w;
            end = min(d, beg + piece_w);
        }
        for (int e0 = beg; e0 < end; e0 += kWave * kUnroll) {
            int c[kUnroll];
            double vv[kUnroll];
            bool act[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const int e = e0 + u * kWave + lane;
                act[u] = e < end;
                c[u] = 0;
                vv[u] = 0.0;
                if (act[u]) {
                    if (group) {
                        const int k = upper_bound_lds(W.incl, 0, kWave, e);
                        const int o = e - (k ? W.incl[k - 1] : 0);
                        c[u] = cols[W.start[k] + o];
                        vv[u] = W.val[k];
                    } else {
                        c[u] = cols[st + e];
                        vv[u] = v;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++) f(c[u], vv[u], act[u]);
        }
        if (group) wave_sync();
    }
}

// hub pieces of list entry i (degree d): (i << 8) | piece
__device__ __forceinline__ void add_pieces(unsigned int i, int d, int pw, int hw, uint32_t *pcs,
                                           unsigned int *s_npc) {
    if (d > hw) {
        const unsigned int np = (unsigned int)((d + pw - 1) / pw);
        const unsigned int b = atomicAdd(s_npc, np);
        for (unsigned int k = 0; k < np; k++) pcs[b + k] = (i << 8) | k;
    }
}

__device__ __forceinline__ bool pristine(float v) { return __float_as_uint(v) == 0u; }
__device__ __forceinline__ bool pristine(double v) { return __double_as_longlong(v) == 0ll; }

// FP64 pull of the certificate: for the listed rows v, sum over the row's
// column entries u of push(x32_u) -- a gather per column entry where a
// scatter would take a memory-side atomic.  Rows of at most hub_w edges go 64
// at a time per wavefront: the row's lane gets node and row start from
// row_of(i, v, st), the edges are flattened over the wave, their terms summed
// in LDS, and the row's lane then finishes the entry (finish(ok, v, x32_v,
// d_v, sum), called by every lane).  Longer rows are listed in ent[j] =
// {start, degree, v} for j < *s_nlong, with their pieces in pcs[] for
// pull_pieces().  x32 of the nodes below xn comes from the LDS copy xc[].
// ops counts the rows' edges.
template <typename Row, typename PushOf, typename Finish>
__device__ __forceinline__ void pull_rows(unsigned int cnt, Row row_of, const SlotView &S, const int32_t *degree,
                                          const int2 *colx, const float *xc, int xn, int hub_w, int piece_w,
                                          Ent *ent, uint32_t *pcs, unsigned int *s_npc, unsigned int *s_nlong,
                                          unsigned int *s_work, WaveLds &W, PushOf push_of, Finish finish,
                                          unsigned long long &ops) {
    const int lane = __lane_id();
    const unsigned int ngroups = (cnt + kWave - 1) / kWave;
    for (;;) {
        unsigned int item = 0;
        if (lane == 0) item = atomicAdd(s_work, 1u);
        item = __shfl(item, 0);
        if (item >= ngroups) break;
        const unsigned int i = item * kWave + lane;
        int v = -1, st = 0, d = 0;
        float xv0 = 0.0f;
        if (i < cnt) {
            row_of(i, v, st);
            d = degree[v];
            xv0 = S.rx32[v].y;
        }
        ops += (unsigned long long)d;
        const bool lng = d > hub_w;
        if (lng) {
            const unsigned int j = atomicAdd(s_nlong, 1u);
            Ent e;
            e.st = st;
            e.d = d;
            e.v = (double)v;
            ent[j] = e;
            add_pieces(j, d, piece_w, hub_w, pcs, s_npc);
        }
        int incl = lng ? 0 : d;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        const int tot = __shfl(incl, kWave - 1);
        W.incl[lane] = incl;
        W.start[lane] = st;
        W.val[lane] = 0.0;
        wave_sync();
        for (int e0 = 0; e0 < tot; e0 += kWave * kUnroll) {
            int k[kUnroll];
            int2 c[kUnroll];
            float xv[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const int e = e0 + u * kWave + lane;
                k[u] = -1;
                c[u] = make_int2(0, 0);
                if (e < tot) {
                    k[u] = upper_bound_lds(W.incl, 0, kWave, e);
                    const int o = e - (k[u] ? W.incl[k[u] - 1] : 0);
                    c[u] = colx[W.start[k[u]] + o];
                }
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
                xv[u] = k[u] < 0 ? 0.0f : c[u].x < xn ? xc[c[u].x] : S.rx32[c[u].x].y;
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
                if (xv[u] != 0.0f) atomicAdd(&W.val[k[u]], push_of(xv[u], c[u].y));
        }
        wave_sync();
        finish(v >= 0 && !lng, v, xv0, d, W.val[lane]);
        wave_sync();
    }
}

// The pieces of the long rows pull_rows() listed: one wavefront per piece, one
// atomic per piece into the row's r64 (zero before).
template <typename PushOf>
__device__ __forceinline__ void pull_pieces(unsigned int npc, const uint32_t *pcs, const Ent *ent, int piece_w,
                                            const SlotView &S, const int2 *colx, const float *xc, int xn,
                                            unsigned int *s_work, PushOf push_of) {
    const int lane = __lane_id();
    for (;;) {
        unsigned int item = 0;
        if (lane == 0) item = atomicAdd(s_work, 1u);
        item = __shfl(item, 0);
        if (item >= npc) break;
        const uint32_t pc = pcs[item];
        const Ent e = ent[pc >> 8];
        const int beg = (int)(pc & 255u) * piece_w;
        const int end = min(e.d, beg + piece_w);
        double acc = 0.0;
        for (int e0 = beg; e0 < end; e0 += kWave * kUnroll) {
            int2 c[kUnroll];
            float xv[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const int ed = e0 + u * kWave + lane;
                c[u] = ed < end ? colx[e.st + ed] : make_int2(-1, 0);
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
                xv[u] = c[u].x < 0 ? 0.0f : c[u].x < xn ? xc[c[u].x] : S.rx32[c[u].x].y;
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
                if (xv[u] != 0.0f) acc += push_of(xv[u], c[u].y);
        }
        acc = wave_sum(acc);
        if (lane == 0 && acc != 0.0) atomic_add_wg(&S.rx64[(int)e.v].x, acc);
    }
}

// Touched-entry bookkeeping of the main push: an entry untouched by the
// current query has r == +0.0 (bit pattern 0); the node phase leaves -0.0 in
// the residual it moves to x and pushes are >= 0, so r never returns to +0.0
// once it has left it: "old == +0.0 and new != 0" in an add's result marks
// exactly the first touch, with no flag atomic.  The correction push
// (residuals of both signs) falls back to the flag for entries the main push
// never reached.
// an edge atomic's result (from the hub cache or from memory)
template <typename T>
struct ResT {
    T v;
    bool lds;
};

// write pass of the binned push: a hub's add (LDS cache) or a pair's position
template <typename T>
struct BinRes {
    T v;
    unsigned int p;
    bool lds;
};

// a binned push pair: key (below) and value, written and read as one access
template <typename T>
struct Pair;
template <>
struct alignas(8) Pair<float> {
    unsigned int key;
    float val;
};
template <>
struct alignas(16) Pair<double> {
    unsigned int key;
    unsigned int pad;
    double val;
};
__device__ __forceinline__ Pair<float> make_pair(unsigned int k, float v) { return Pair<float>{k, v}; }
__device__ __forceinline__ Pair<double> make_pair(unsigned int k, double v) { return Pair<double>{k, 0u, v}; }

// A pair's key: the destination's offset in its range (low `shift` bits) and
// its degree above them, saturated at deg_sat(shift) (a saturated degree is
// read from the graph in the drain), so the drain needs no degree load.  The
// saturation value keeps every key below 0xffffffff, the no-pair mark.
__device__ __forceinline__ unsigned int deg_sat(int shift) { return (1u << (32 - shift)) - 2u; }
__device__ __forceinline__ unsigned int pair_key(int v, int dv, int shift) {
    return ((unsigned int)v & ((1u << shift) - 1u)) | (min((unsigned int)dv, deg_sat(shift)) << shift);
}

// Heat-kernel fields by precision: level residuals (parity 0 / 1) and the
// estimate.  FP32: both levels in hk2 (as floats), the estimate in rx32.x;
// FP64: the even level in rx64.r, the odd one in hk2, the estimate in rx64.x.
template <bool F64>
struct HkF;
template <>
struct HkF<false> {
    using T = float;
    static __device__ __forceinline__ Strided<float> lvl(const SlotView &S, int64_t rn, bool odd) {
        return Strided<float>{reinterpret_cast<float *>(S.hk2) + (odd ? rn : 0), 1};
    }
    static __device__ __forceinline__ Strided<float> est(const SlotView &S) { return Strided<float>{&S.rx32[0].y, 2}; }
};
template <>
struct HkF<true> {
    using T = double;
    static __device__ __forceinline__ Strided<double> lvl(const SlotView &S, int64_t rn, bool odd) {
        return odd ? Strided<double>{S.hk2, 1} : Strided<double>{&S.rx64[0].x, 2};
    }
    static __device__ __forceinline__ Strided<double> est(const SlotView &S) { return Strided<double>{&S.rx64[0].y, 2}; }
};

// Binned edge phase of one push iteration (see the file header), with the
// next iteration's node phase folded into it for the nodes it sends across
// their threshold.
//   write pass: a push into a hub id (v < hub_n) is an LDS atomic on the hub
//     cache, whose result decides the crossing at once (-> next[], the nodes
//     whose node phase the next iteration runs); any other push becomes a
//     (v, value) pair in the region of its destination range k (ids
//     [k 2^shift, (k+1) 2^shift)): pairs[indptr[k 2^shift] ...].  An
//     iteration pushes each column entry at most once (the frontier is a set)
//     and, on an undirected graph, the entries pointing into a range number
//     its row volume, so every range has room with no counting pass.  The
//     ranges' write cursors (then their ends) are bins[] (static LDS).
//   then the hub cache goes back to the record array (the dynamic LDS becomes
//     the slice of one range: 2^shift residual sums, zero between ranges);
//   per range: the pairs are summed into the slice (LDS atomics), and then
//     drained by the pairs themselves: the first pair of each destination to
//     swap its slice entry with 0 gets the summed add and applies it to the
//     record in one read and one write.  A destination it sends across its
//     threshold has all of this iteration's pushes in (they all come through
//     its range), so its node phase runs right there: x += r, r -> -0, and its
//     frontier entry {row start, degree, push value} goes to ent[] of the next
//     iteration (counter *s_nf; its hub pieces to pcs[], counter *s_npc; its
//     volume to *s_vol and *s_fvol).  Ranges with fewer than dense_min pairs
//     take one memory-side atomic per pair instead (crossings -> next[]).
// r and x are the residual and estimate halves of the slot's {r, x} records
// (kRX: one record array, r.p[2 v] and r.p[2 v + 1], read and written as one
// vector; otherwise two strided arrays, accessed apart);
// post(old, new, d_v) -> append bits of one destination's add;
// push(r, d) -> the push value of a node moving r with degree d.
template <int NT, typename T, bool kRX, typename Load, typename Post, typename Push>
__device__ __forceinline__ void binned_edges(unsigned int cnt, const uint32_t *pcs, unsigned int npc, int pw, int hw,
                                             unsigned int *s_work, WaveLds &W, const int2 *colx,
                                             const int32_t *indptr, const int32_t *degree, Load load, int shift,
                                             int K, int64_t n, T *hub, int hub_n, void *lds, unsigned int *bins,
                                             Pair<T> *pairs, Strided<T> r, Strided<T> x, int dense_min,
                                             Post post, Push push,
                                             Ent *ent, uint32_t *pcs_next, unsigned int *s_npc, unsigned int *s_nf,
                                             unsigned long long *s_vol, unsigned long long *s_fvol, int32_t *sup,
                                             unsigned int *s_sup, int32_t *next, unsigned int *s_next,
                                             int32_t *tl, unsigned int *s_tl, int32_t *nl, unsigned int *s_nl,
                                             unsigned int nl_cap, uint64_t &t_pairs) {
    const int tid = threadIdx.x;
    const int lane = __lane_id();
    T *slice = reinterpret_cast<T *>(lds);
    // 1. write pass
    for (int k = tid; k < K; k += NT) bins[k] = (unsigned int)indptr[min((int64_t)k << shift, n)];
    if (tid == 0) *s_work = 0;
    __syncthreads();
    {
        const int lane0 = lane;
        wave_edges<BinRes<T>>(
            cnt, pcs, npc, pw, hw, s_work, W, colx, load,
            [&](int vtx, double val, BinRes<T> &res, int stage) {
                if (stage != 0) return;
                res.lds = vtx < hub_n;
                res.v = (T)0;
                res.p = 0;
                if (vtx < hub_n) res.v = atomicAdd(&hub[vtx], (T)val);
                else if (vtx < n) res.p = atomicAdd(&bins[vtx >> shift], 1u);
            },
            [&](int vtx, int dv, double val, const BinRes<T> &res, bool) -> int {
                if (res.lds) return post(res.v, res.v + (T)val, dv);
                if (vtx < n) {   // idle lanes aim past n
                    pairs[res.p] = make_pair(pair_key(vtx, dv, shift), (T)val);   // one store
                }
                return 0;
            },
            next, s_next, tl, s_tl, (int)n, nl, s_nl, nl_cap);
        (void)lane0;
    }
    __syncthreads();
    // 2. the hub cache back to the records; the dynamic LDS becomes the slice
    for (int v = tid; v < hub_n; v += NT) {
        const T h = hub[v];
        if (!pristine(h)) r[v] = h;
    }
    if (tid == 0) {
        *s_nf = 0;
        *s_npc = 0;
        *s_fvol = 0;
    }
    __syncthreads();
    {
        uint4 *z = reinterpret_cast<uint4 *>(lds);
        const int nz = (int)((((int64_t)1 << shift) * (int64_t)sizeof(T)) / 16);
        for (int i = tid; i < nz; i += NT) z[i] = make_uint4(0u, 0u, 0u, 0u);
    }
    __syncthreads();
    t_pairs = now_ticks();
    unsigned long long fvol = 0;
    // 3. range by range.  Every loop issues the loads of kBinUnroll items per
    // lane before it uses any (one memory round trip per batch, not per item).
    const unsigned int rmask = (1u << shift) - 1u;
    const unsigned int dsat = deg_sat(shift);
    // one destination's summed add dl (non-zero) into the record of v, with
    // the next iteration's node phase when it crosses; called by every lane
    auto apply = [&](const T *dl, const int *v, const int *dv, const T *ro, const T *xo) {
        int bits[kBinUnroll];
        bool fuse[kBinUnroll], first[kBinUnroll];
        double pv_u[kBinUnroll];
#pragma unroll
        for (int u = 0; u < kBinUnroll; u++) {
            bits[u] = 0;
            fuse[u] = false;
            pv_u[u] = 0.0;
            if (!pristine(dl[u])) {
                const T nw = ro[u] + dl[u];
                bits[u] = post(ro[u], nw, dv[u]);
                fuse[u] = (bits[u] & 1) != 0;
                if (fuse[u]) {
                    // the next iteration's node phase of v (ppr_solver.py:273-282)
                    if constexpr (kRX) {
                        reinterpret_cast<typename Rec2<T>::type *>(r.p)[v[u]] = rec2<T>(-(T)0, xo[u] + nw);
                    } else {
                        r[v[u]] = -(T)0;
                        x[v[u]] = xo[u] + nw;
                    }
                    pv_u[u] = push(nw, dv[u]);
                    fvol += (unsigned long long)dv[u];
                } else {
                    r[v[u]] = nw;
                }
            }
            bits[u] &= 6;
            first[u] = fuse[u] && xo[u] == (T)0;   // first push of a node: list it in the support
        }
        int vt[kBinUnroll];
#pragma unroll
        for (int u = 0; u < kBinUnroll; u++) vt[u] = sup_entry(v[u], dv[u]);
        append_bits<kBinUnroll>(bits, v, vt, next, s_next, tl, s_tl, nl, s_nl, nl_cap);
        long long f[kBinUnroll], sk[kBinUnroll];
        wave_append_n<kBinUnroll>(fuse, s_nf, f);
        wave_append_n<kBinUnroll>(first, s_sup, sk);
#pragma unroll
        for (int u = 0; u < kBinUnroll; u++) {
            if (f[u] >= 0) {
                Ent e;
                e.st = -v[u] - 1;   // row start read where the entry is used
                e.d = dv[u];
                e.v = pv_u[u];
                ent[f[u]] = e;
                add_pieces((unsigned int)f[u], dv[u], pw, hw, pcs_next, s_npc);
            }
            if (sk[u] >= 0) sup[sk[u]] = sup_entry(v[u], dv[u]);
        }
    };
    for (int k = 0; k < K; k++) {
        const int64_t lo = (int64_t)k << shift;
        const int c0 = indptr[lo];
        const int c1 = (int)bins[k];
        if (c1 == c0) continue;
        if (c1 - c0 < dense_min) {
            // few pairs: one memory-side atomic each (idle lanes add -0 to a sink)
            for (int p0 = c0; p0 < c1; p0 += NT * kBinUnroll) {
                int v[kBinUnroll], dv[kBinUnroll];
                T val[kBinUnroll], old[kBinUnroll];
#pragma unroll
                for (int u = 0; u < kBinUnroll; u++) {
                    const int p = p0 + u * NT + tid;
                    v[u] = (int)n + lane;
                    dv[u] = 0;
                    val[u] = -(T)0;
                    if (p < c1) {
                        const Pair<T> q = pairs[p];
                        v[u] = (int)lo + (int)(q.key & rmask);
                        dv[u] = (int)(q.key >> shift);
                        val[u] = q.val;
                    }
                }
#pragma unroll
                for (int u = 0; u < kBinUnroll; u++) old[u] = atomic_add_wg(&r[v[u]], val[u]);
#pragma unroll
                for (int u = 0; u < kBinUnroll; u++)
                    if ((unsigned int)dv[u] == dsat) dv[u] = degree[v[u]];
                int bits[kBinUnroll];
#pragma unroll
                for (int u = 0; u < kBinUnroll; u++) bits[u] = v[u] < n ? post(old[u], old[u] + val[u], dv[u]) : 0;
                int vt[kBinUnroll];
#pragma unroll
                for (int u = 0; u < kBinUnroll; u++) vt[u] = sup_entry(v[u], dv[u]);
                append_bits<kBinUnroll>(bits, v, vt, next, s_next, tl, s_tl, nl, s_nl, nl_cap);
            }
            continue;
        }
        // the range's pairs summed into the slice
        const bool one = c1 - c0 <= NT * kBinUnroll;
        unsigned int key[kBinUnroll];   // pair keys; 0xffffffff: no pair
        for (int p0 = c0; p0 < c1; p0 += NT * kBinUnroll) {
            T val[kBinUnroll];
#pragma unroll
            for (int u = 0; u < kBinUnroll; u++) {
                const int p = p0 + u * NT + tid;
                key[u] = ~0u;
                val[u] = (T)0;
                if (p < c1) {
                    const Pair<T> q = pairs[p];
                    key[u] = q.key;
                    val[u] = q.val;
                }
            }
#pragma unroll
            for (int u = 0; u < kBinUnroll; u++)
                if (key[u] != ~0u) atomicAdd(&slice[key[u] & rmask], val[u]);
        }
        __syncthreads();
        // drained by the pairs themselves: the first pair of a destination
        // to swap its slice entry with 0 applies the summed add
        for (int p0 = c0; p0 < c1; p0 += NT * kBinUnroll) {
            if (!one) {
#pragma unroll
                for (int u = 0; u < kBinUnroll; u++) {
                    const int p = p0 + u * NT + tid;
                    key[u] = p < c1 ? pairs[p].key : ~0u;
                }
            }
            T dl[kBinUnroll], ro[kBinUnroll], xo[kBinUnroll];
            int v[kBinUnroll], dv[kBinUnroll];
#pragma unroll
            for (int u = 0; u < kBinUnroll; u++) {
                dl[u] = (T)0;
                v[u] = 0;
                dv[u] = 0;
                if (key[u] != ~0u) {
                    v[u] = (int)lo + (int)(key[u] & rmask);
                    dv[u] = (int)(key[u] >> shift);
                    dl[u] = __hip_atomic_exchange(&slice[key[u] & rmask], (T)0, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
#pragma unroll
            for (int u = 0; u < kBinUnroll; u++) {
                ro[u] = (T)0;
                xo[u] = (T)0;
                if (!pristine(dl[u])) {
                    if constexpr (kRX) {
                        const auto w = reinterpret_cast<const typename Rec2<T>::type *>(r.p)[v[u]];
                        ro[u] = w.x;
                        xo[u] = w.y;
                    } else {
                        ro[u] = r[v[u]];
                        xo[u] = x[v[u]];
                    }
                    if ((unsigned int)dv[u] == dsat) dv[u] = degree[v[u]];
                }
            }
            apply(dl, v, dv, ro, xo);
        }
        __syncthreads();
    }
    if (fvol) {
        atomicAdd(s_vol, fvol);
        atomicAdd(s_fvol, fvol);
    }
}

// Heat kernel query (hk_sor_queue with omega = 1, hk_solver.py:206-295), with
// FP32 or FP64 level residuals (the query's precision) and FP64 threshold
// tests.  Level j's residual lives in the level array of parity j & 1 -- for
// the hub cache's nodes in one of two LDS level arrays.  Each level keeps
// the list of nodes it touched; the node phase walks that list, pushes every
// node at or above the level threshold (level residuals only grow, so this is
// the set the reference's crossing queue holds, hk_solver.py:258) and clears
// every entry (the reference zeroes the whole level, hk_solver.py:282-285),
// so no separate clearing pass is needed.
template <int NT, bool F64>
__device__ __forceinline__ void hk_query(const BatchArgs &a, const flsg_query &qq, unsigned int q, const SlotView &S,
                                         int32_t *qa, int32_t *qb, Ent *ent, int32_t *sup, uint32_t *pcs, WaveLds &W,
                                         double *hub_raw, unsigned int *s_cnt, unsigned int *s_ncnt,
                                         unsigned int *s_sup, unsigned int *s_npc, unsigned int *s_work,
                                         unsigned int *s_tn, unsigned long long *s_vol, unsigned long long *s_base,
                                         uint64_t t_start, int sink, uint4 *pairs_raw, unsigned int *bins) {
    using F = HkF<F64>;
    using T = typename F::T;
    const int tid = threadIdx.x;
    const int N = a.hk_n;
    const int s = (int)qq.seed;
    const int pw = a.piece_w;
    const int hw = a.hub_w;
    const int32_t *indptr = a.indptr;
    const Strided<T> est = F::est(S);
    T *hub = reinterpret_cast<T *>(hub_raw);
    // hub cache: one level array of hub_n entries, then kWave sink slots.  A
    // level's node phase reads and clears its hub entries before its edge
    // phase adds the next level's into the same array (a block barrier apart)
    const int cap = a.hub_bytes / (int)sizeof(T);
    const int hub_n = cap > 2 * kWave ? (int)min((int64_t)(cap - kWave), a.n) : 0;
    const int hsink = hub_n > 0 ? hub_n - sink : INT32_MIN;   // LDS index of an idle lane = vtx + hsink
    for (int i = tid; i < a.hub_bytes / 8; i += NT) hub_raw[i] = 0.0;
    // large levels under load take the binned push (as the PPR push does):
    // destination ranges of 2^shift ids summed in the LDS
    int shift = 0;
    while (((int64_t)sizeof(T) << (shift + 1)) <= (int64_t)a.hub_bytes) shift++;
    const int K = (int)((a.n + ((int64_t)1 << shift) - 1) >> shift);
    const bool can_bin = a.undirected && pairs_raw && a.pcap >= (int64_t)a.indptr[a.n] && shift >= kBinShiftMin &&
                         K <= kMaxBins && hub_n > 0;
    __shared__ unsigned long long s_lvol, s_dvol, s_dfvol;   // this level's push volume; unused sums
    __shared__ unsigned int s_dnf, s_hbusy;
    bool hub_lds = true;   // the hub cache holds the current level's hub residuals
    __syncthreads();
    if (tid == 0) {
        if (s < hub_n) hub[s] = (T)1;   // hk_solver.py:223
        else F::lvl(S, a.rn, false)[s] = (T)1;
        qa[0] = sup_entry(s, a.degree[s]);   // level lists carry the degree
        *s_cnt = 1;
        *s_sup = 0;
        *s_vol = 0;
        *s_tn = 0;
    }
    __syncthreads();
    int32_t *lc = qa, *ln = qb;
    int64_t levels = 0;
    auto load_ent = [&](unsigned int i, int &st, int &d, double &v) {
        const Ent e = ent[i];
        st = e.st;
        d = e.d;
        v = e.v;
    };
    for (int j = 0; j < N; j++) {
        const unsigned int cnt = *s_cnt;
        if (cnt == 0) break;   // hk_solver.py:237
        levels++;
        const bool odd = (j & 1) != 0;
        const bool last = (j + 1 == N);
        const double coef = a.hk_tau / ((double)j + 1.);
        const double psic = a.hk_psi[j];
        T *hc = hub;   // level j's hub entries, then level j + 1's
        const Strided<T> lv = F::lvl(S, a.rn, odd);
        const Strided<T> lnx = F::lvl(S, a.rn, !odd);
        const int nb = 0;   // LDS base of level j + 1
        __shared__ unsigned int s_np;   // this level's pushers, listed compactly in ent[]
        unsigned int busy_now = 0;
        if (tid == 0 && a.busy) busy_now = __hip_atomic_load(a.busy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        if (tid == 0) {
            *s_ncnt = 0;
            *s_npc = 0;
            *s_work = 0;
            s_np = 0;
            s_lvol = 0;
        }
        __syncthreads();
        unsigned long long vol = 0;
        unsigned int pushers = 0;
        // every node the level touched: its level residual (and degree, carried
        // by the list entry below 255) decides whether it pushes; only pushers
        // load their estimate and row start and get a frontier entry
        for (unsigned int i0 = tid; i0 < cnt; i0 += NT * kUnroll) {
            int uu[kUnroll], ss[kUnroll], dd[kUnroll];
            T ro[kUnroll], xo[kUnroll];
            bool pu[kUnroll];
#pragma unroll
            for (int k = 0; k < kUnroll; k++) {
                const unsigned int i = i0 + k * NT;
                uu[k] = -1;
                dd[k] = 0;
                if (i < cnt) {
                    const uint32_t e = (uint32_t)lc[i];
                    uu[k] = (int)(e & 0xffffffu);
                    dd[k] = (int)(e >> 24);
                }
            }
#pragma unroll
            for (int k = 0; k < kUnroll; k++) {
                ro[k] = (T)0;
                if (uu[k] >= 0) {
                    if (dd[k] == 255) dd[k] = a.degree[uu[k]];
                    ro[k] = hub_lds && uu[k] < hub_n ? hc[uu[k]] : lv[uu[k]];
                }
            }
#pragma unroll
            for (int k = 0; k < kUnroll; k++)   // hk_solver.py:244
                pu[k] = uu[k] >= 0 && !(fabs((double)ro[k]) < ((a.hk_ete * (double)dd[k]) / (double)N) / psic);
#pragma unroll
            for (int k = 0; k < kUnroll; k++) {
                xo[k] = (T)0;
                ss[k] = 0;
                if (pu[k]) {
                    xo[k] = est[uu[k]];
                    ss[k] = indptr[uu[k]];
                }
            }
            long long slot[kUnroll];
            wave_append_n<kUnroll>(pu, &s_np, slot);
#pragma unroll
            for (int k = 0; k < kUnroll; k++) {
                if (uu[k] < 0) continue;
                const int u = uu[k];
                const int d = dd[k];
                const T tmp = ro[k];
                if (hub_lds && u < hub_n) hc[u] = (T)0;            // hk_solver.py:282-285
                else lv[u] = (T)0;
                if (pu[k]) {
                    est[u] = xo[k] + tmp;
                    if (pristine(xo[k])) sup[atomicAdd(s_sup, 1u)] = u;
                    Ent e;
                    e.st = ss[k];
                    e.d = d;
                    e.v = last ? (double)tmp / (double)d : coef * (double)tmp / (double)d;   // hk_solver.py:249,255
                    ent[slot[k]] = e;
                    vol += (unsigned long long)d;
                    pushers++;
                    add_pieces((unsigned int)slot[k], d, pw, hw, pcs, s_npc);
                }
            }
        }
        if (vol) {
            atomicAdd(s_vol, vol);
            atomicAdd(&s_lvol, vol);
        }
        if (pushers) atomicAdd(s_tn, pushers);
        if (tid == 0) s_hbusy = busy_now;
        __syncthreads();
        if (!hub_lds) {
            // a binned level left its hub residuals in memory (read above) and
            // stale copies in the LDS: the next level's hub cache starts at zero
            for (int i = tid; i < a.hub_bytes / 8; i += NT) hub_raw[i] = 0.0;
            hub_lds = true;
            __syncthreads();
        }
        const bool binned = !last && can_bin && s_lvol >= (unsigned long long)a.bin_min &&
                            (!a.busy || s_hbusy >= (unsigned int)a.bin_busy_min);
        if (binned) {
            // the next level's residual through pairs per destination range; first
            // touch lists the node for level j + 1 (the hub residuals end in memory)
            uint64_t tp = 0;
            binned_edges<NT, T, false>(
                s_np, pcs, *s_npc, pw, hw, s_work, W, a.colx, a.indptr, a.degree, load_ent, shift, K, a.n, hub,
                hub_n, hub_raw, bins, reinterpret_cast<Pair<T> *>(pairs_raw), lnx, lnx, a.dense_min,
                [&](T old, T nw, int) -> int { return (pristine(old) && nw != (T)0) ? 2 : 0; },
                [&](T, int) -> double { return 0.0; }, ent, pcs, s_npc, &s_dnf, &s_dvol, &s_dfvol, sup, s_sup, ln,
                s_ncnt, ln, s_ncnt, (int32_t *)nullptr, (unsigned int *)nullptr, 0u, tp);
            hub_lds = false;
        } else if (!last) {
            // next level's residual; first touch lists the node for level j + 1
            wave_edges<ResT<T>, true>(
                s_np, pcs, *s_npc, pw, hw, s_work, W, a.colx, load_ent,
                [&](int vtx, double val, ResT<T> &r, int stage) {
                    const int li = vtx < hub_n ? nb + vtx : vtx >= sink ? vtx + hsink : -1;
                    if (stage == 0) {
                        r.lds = li >= 0;
                        if (li >= 0) r.v = atomicAdd(&hub[li], (T)val);
                    } else if (li < 0) {
                        r.v = atomic_add_wg(&lnx[vtx], (T)val);
                    }
                },
                [&](int, int, double val, const ResT<T> &r, bool) -> int {
                    return (pristine(r.v) && r.v + (T)val != (T)0) ? 2 : 0;
                },
                ln, s_ncnt, ln, s_ncnt, sink);
        } else {
            // last level pushes straight into x (hk_solver.py:254-256)
            wave_edges<T>(
                s_np, pcs, *s_npc, pw, hw, s_work, W, a.colx, load_ent,
                [&](int vtx, double val, T &old, int stage) {
                    if (stage == 1) old = atomic_add_wg(&est[vtx], (T)val);
                },
                [&](int, int, double val, T old, bool) -> int {
                    return (pristine(old) && old + (T)val != (T)0) ? 2 : 0;
                },
                ln, s_ncnt, sup, s_sup, sink);
        }
        __syncthreads();
        if (tid == 0) *s_cnt = *s_ncnt;
        int32_t *t = lc;
        lc = ln;
        ln = t;
        __syncthreads();
    }
    const unsigned int Sn = *s_sup;
    if (tid == 0) *s_base = atomicAdd((unsigned long long *)&a.ctl[1], (unsigned long long)Sn);
    __syncthreads();
    const unsigned long long obase = *s_base;
    const bool fits = obase + Sn <= (unsigned long long)a.out_cap;
    const double et = exp(a.hk_tau);
    for (unsigned int i0 = tid; i0 < Sn; i0 += NT * kUnroll) {
        int uu[kUnroll], id[kUnroll];
        T xv[kUnroll];
#pragma unroll
        for (int k = 0; k < kUnroll; k++) {
            const unsigned int i = i0 + k * NT;
            uu[k] = i < Sn ? (int)((uint32_t)sup[i] & 0xffffffu) : -1;   // the last level's entries carry a degree
        }
#pragma unroll
        for (int k = 0; k < kUnroll; k++) {
            id[k] = 0;
            xv[k] = (T)0;
            if (uu[k] >= 0) {
                xv[k] = est[uu[k]];
                id[k] = a.out_map ? a.out_map[uu[k]] : uu[k];
            }
        }
#pragma unroll
        for (int k = 0; k < kUnroll; k++) {
            if (uu[k] < 0) continue;
            const unsigned int i = i0 + k * NT;
            if (fits) {
                a.out_ids[obase + i] = id[k];
                a.out_vals[obase + i] = (double)xv[k] / et;   // hk_solver.py:291
            }
            est[uu[k]] = (T)0;
        }
    }
    __syncthreads();
    if (tid == 0) {
        const uint64_t t_end = now_ticks();
        flsg_result rr{};
        rr.oper = (int64_t)*s_vol;
        rr.iters = levels;
        rr.supp = (int64_t)Sn;
        rr.nodes = (int64_t)*s_tn;
        rr.ms = (double)(t_end - t_start) * 1e-5;
        rr.ms_push = rr.ms;
        rr.ms_recompute = rr.ms;
        rr.ms_correct = rr.ms;
        rr.status = fits ? FLSG_OK : FLSG_ECAP;
        a.res[q] = rr;
        a.out_off[q] = fits ? (int64_t)obase : -1;
    }
    __syncthreads();
}

// Local push of one PPR / Katz query in residual type T (float: the FP32
// push, double: the FP64 one), from the seed's residual to an empty frontier;
// rb is the slot's {r, x} record array seen as T (r_v = rb[2v], x_v = rb[2v+1]).  An iteration's frontier is ent[0, F) -- entries a binned
// iteration already prepared (its drain ran their node phase) -- followed by
// the nodes listed in qc[0, cu), whose node phase runs at the iteration's
// start.  Returns the status (FLSG_ENOCONV past the iteration cap); *iters,
// *fnodes and s_vol (operation count) are filled.
template <int NT, typename T>
__device__ __forceinline__ int local_push(const BatchArgs &a, const QParams &P, int s, T *rb, int32_t *qa,
                                          int32_t *qb, Ent *ent, int32_t *sup, uint32_t *pcs,
                                          int32_t *tl, int32_t *nl, Pair<T> *pairs, WaveLds &W,
                                          double *s_hub, unsigned int *bins, unsigned int *s_cnt,
                                          unsigned int *s_ncnt, unsigned int *s_sup, unsigned int *s_npc,
                                          unsigned int *s_work, unsigned int *s_tl, unsigned int *s_nl,
                                          unsigned int *s_nf, unsigned long long *s_vol, unsigned long long *s_ivol,
                                          unsigned long long *s_fvol, int sink, int64_t *iters_out,
                                          unsigned long long *fnodes_out, int32_t **final_q, uint64_t *tick) {
    const int tid = threadIdx.x;
    const int pw = a.piece_w;
    const int hw = a.hub_w;
    const bool f32 = sizeof(T) == 4;
    const Strided<T> r{rb, 2};        // residual and estimate halves of the records
    const Strided<T> x{rb + 1, 2};
    T *hub = reinterpret_cast<T *>(s_hub);
    // cache capacity in elements; its last kWave entries take the identity
    // adds of idle lanes (so they cost no memory-side atomic)
    const int hub_cap = a.hub_bytes / (int)sizeof(T);
    const int hub_n = hub_cap > kWave ? (int)min((int64_t)(hub_cap - kWave), a.n) : 0;
    const int hub_sink = hub_cap > kWave ? hub_cap - kWave - sink : INT32_MIN;   // + vtx of an idle lane
    // binned push: destination ranges of 2^shift ids in the dynamic LDS, at
    // most kMaxBins of them (the pair regions are sized by the ranges' row
    // volumes: undirected graphs only)
    int shift = 0;
    while (((int64_t)sizeof(T) << (shift + 1)) <= (int64_t)a.hub_bytes) shift++;
    const int K = (int)((a.n + ((int64_t)1 << shift) - 1) >> shift);
    const bool can_bin = a.undirected && a.pcap >= (int64_t)a.indptr[a.n] && shift >= kBinShiftMin &&
                         K <= kMaxBins;
    for (int i = tid; i < a.hub_bytes / 8; i += NT) s_hub[i] = 0.0;
    __syncthreads();
    if (tid == 0) {
        if (s < hub_n) hub[s] = (T)P.r0;
        else r[s] = (T)P.r0;
        tl[0] = sup_entry(s, a.degree[s]);
        qa[0] = s;
        *s_cnt = 1;
        *s_nf = 0;
        *s_npc = 0;
        *s_fvol = 0;
        *s_sup = 0;
        *s_tl = 1;
        *s_nl = 0;
        *s_vol = 0;
    }
    __syncthreads();
    bool hub_on = true;   // the hub cache holds the residuals of [0, hub_n)
    int32_t *qc = qa, *qn = qb;
    unsigned long long fnodes = 0;
    int64_t iters = 0;
    int status = FLSG_OK;
    // entries a binned iteration prepared hold -(v + 1) for the row start
    auto load_ent = [&](unsigned int i, int &st, int &d, double &v) {
        const Ent e = ent[i];
        st = e.st >= 0 ? e.st : a.indptr[-e.st - 1];
        d = e.d;
        v = e.v;
    };
    // one destination's add: crossing (-> next frontier), first touch (->
    // touched list) and, FP32, boundary candidates for an exact recompute:
    // first touch of a high-degree node, or a residual that entered
    // [kNearFrac * thr, thr) with this add
    auto post = [&](T old, T nw, int dv) -> int {
        const double thr = P.thr_scale * (double)dv;
        const bool cross = (double)old < thr && thr <= (double)nw;
        const bool first = pristine(old) && nw != (T)0;
        int bits = (cross ? 1 : 0) | (first ? 2 : 0);
        if (f32) {
            const double cn = kNearFrac * thr;
            const bool near = dv > kLazyDeg ? first : (!cross && (double)old < cn && cn <= (double)nw);
            bits |= near ? 4 : 0;
        }
        return bits;
    };
    // push value of a node moving r (ppr_solver.py:282, katz_solver.py:212);
    // the FP32 push rounds it to FP32
    auto push = [&](T r, int d) -> double {
        const double v = bpush_value(P, (double)r, d);
        return f32 ? (double)(float)v : v;
    };
    int32_t *nlp = f32 ? nl : nullptr;
    unsigned int *s_nlp = f32 ? s_nl : nullptr;
    for (;;) {
        const unsigned int cu = *s_cnt;   // nodes still to run their node phase
        const unsigned int F = *s_nf;     // entries already prepared
        const unsigned int cnt = F + cu;
        if (cnt == 0) break;
        if (iters >= a.max_iters) {
            status = FLSG_ENOCONV;
            break;
        }
        iters++;
        fnodes += cnt;
        const uint64_t t_it = now_ticks();
        unsigned int busy_now = 0;
        if (tid == 0 && a.busy) busy_now = __hip_atomic_load(a.busy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        if (tid == 0) {
            *s_ncnt = 0;
            *s_work = 0;
            *s_ivol = *s_fvol;
        }
        __syncthreads();
        // node phase: r_u -> 0, x_u += r_u, push value (ppr_solver.py:273-282);
        // every entry of the unrolled frontier slots is read before any is written
        unsigned long long vol = 0;
        for (unsigned int i00 = 0; i00 < cu; i00 += NT * kUnroll) {
            const unsigned int i0 = i00 + tid;
            int uu[kUnroll];
#pragma unroll
            for (int j = 0; j < kUnroll; j++) {
                const unsigned int i = i0 + j * NT;
                uu[j] = i < cu ? qc[i] : -1;
            }
            T rr[kUnroll], xx[kUnroll];
            int dd[kUnroll], ss[kUnroll];
#pragma unroll
            for (int j = 0; j < kUnroll; j++) {
                rr[j] = (T)0;
                xx[j] = (T)0;
                dd[j] = 0;
                ss[j] = 0;
                if (uu[j] >= 0) {
                    const int u = uu[j];
                    if (hub_on && u < hub_n) {
                        xx[j] = x[u];
                        rr[j] = hub[u];
                    } else {
                        const auto w = reinterpret_cast<const typename Rec2<T>::type *>(rb)[u];
                        rr[j] = w.x;
                        xx[j] = w.y;
                    }
                    dd[j] = a.degree[u];
                    ss[j] = a.indptr[u];
                }
            }
#pragma unroll
            for (int j = 0; j < kUnroll; j++) {
                if (uu[j] < 0) continue;
                const int u = uu[j];
                const unsigned int i = F + i0 + j * NT;
                Ent e;
                e.st = ss[j];
                e.d = dd[j];
                if (hub_on && u < hub_n) {
                    hub[u] = -(T)0;
                    x[u] = xx[j] + rr[j];
                } else {
                    reinterpret_cast<typename Rec2<T>::type *>(rb)[u] = rec2<T>(-(T)0, xx[j] + rr[j]);
                }
                e.v = push(rr[j], dd[j]);
                ent[i] = e;
                vol += (unsigned long long)dd[j];
                add_pieces(i, dd[j], pw, hw, pcs, s_npc);
            }
            // first push of a node: list it in the support (one LDS atomic per wave)
            bool fp[kUnroll];
            long long k[kUnroll];
#pragma unroll
            for (int j = 0; j < kUnroll; j++) fp[j] = uu[j] >= 0 && xx[j] == (T)0;
            wave_append_n<kUnroll>(fp, s_sup, k);
#pragma unroll
            for (int j = 0; j < kUnroll; j++) {
                if (k[j] >= 0) sup[k[j]] = sup_entry(uu[j], dd[j]);
            }
        }
        if (vol) {
            atomicAdd(s_vol, vol);
            atomicAdd(s_ivol, vol);
        }
        __shared__ unsigned int s_busy;
        if (tid == 0) s_busy = busy_now;
        __syncthreads();
        const uint64_t t0 = now_ticks();
        tick[0] += t0 - t_it;
        const bool binned = can_bin && *s_ivol >= (unsigned long long)a.bin_min &&
                            (!a.busy || s_busy >= (unsigned int)a.bin_busy_min);
        const unsigned int npc = *s_npc;
        if (!hub_on) {
            for (int v = tid; v < hub_n; v += NT) hub[v] = r[v];
            hub_on = true;
            __syncthreads();
        }
        if (binned) {
            uint64_t tp = t0;
            binned_edges<NT, T, true>(cnt, pcs, npc, pw, hw, s_work, W, a.colx, a.indptr, a.degree, load_ent, shift, K,
                                a.n, hub, hub_n, s_hub, bins, pairs, r, x, a.dense_min, post, push, ent, pcs,
                                s_npc, s_nf, s_vol, s_fvol, sup, s_sup, qn, s_ncnt, tl, s_tl, nlp, s_nlp,
                                (unsigned int)a.n, tp);
            hub_on = false;
            const uint64_t t1 = now_ticks();
            tick[1] += tp - t0;
            tick[2] += t1 - tp;
            tick[4] += 1;
        } else {
            // edge phase: v joins the next frontier iff this atomic moved r_v across thr_v
            wave_edges<ResT<T>>(
                cnt, pcs, npc, pw, hw, s_work, W, a.colx, load_ent,
                [&](int vtx, double val, ResT<T> &res, int stage) {
                    const int li = vtx < hub_n ? vtx : vtx >= sink ? vtx + hub_sink : -1;
                    if (stage == 0) {
                        res.lds = li >= 0;
                        if (li >= 0) res.v = atomicAdd(&hub[li], (T)val);
                    } else if (li < 0) {
                        res.v = atomic_add_wg(&r[vtx], (T)val);
                    }
                },
                [&](int, int dv, double val, const ResT<T> &res, bool) -> int {
                    return post(res.v, res.v + (T)val, dv);
                },
                qn, s_ncnt, tl, s_tl, sink, nlp, s_nlp, (unsigned int)a.n);
            __syncthreads();
            if (tid == 0) {
                *s_nf = 0;
                *s_npc = 0;
                *s_fvol = 0;
            }
            tick[3] += now_ticks() - t0;
        }
        __syncthreads();
        if (tid == 0) *s_cnt = *s_ncnt;
        int32_t *t = qc;
        qc = qn;
        qn = t;
        __syncthreads();
    }
    __syncthreads();
    if (hub_on) {
        // hub residuals back into the records (touched hubs have nonzero bits)
        for (int v = tid; v < hub_n; v += NT) {
            const T h = hub[v];
            if (!pristine(h)) r[v] = h;
        }
    }
    __syncthreads();
    *iters_out = iters;
    *fnodes_out = fnodes;
    *final_q = qc;
    return status;
}

// Serving session: the next query index for this workgroup (thread 0).  The
// index is claimed at once; the workgroup then waits until the host has
// published it -- or, once the host has closed the session with the index
// still unpublished, or after idle_ticks with nothing published (then every
// workgroup of the launch ends: ctl[2] != 0), returns a value >= nq, which
// ends the workgroup.  Every wait is bounded.
// Only the workgroup waiting for the first unpublished index reads the host's
// count and close flag (each read crosses PCIe); it mirrors them into device
// memory (ctl[3]: published count, ctl[4]: closed), where the other waiting
// workgroups read them.  The device words are polled with relaxed loads (an
// acquire at device scope invalidates caches the other slots of the XCD are
// using); one acquire fence orders a claimed index before its query is read.
__device__ __forceinline__ unsigned int session_claim(const BatchArgs &a) {
    unsigned long long *expired = (unsigned long long *)&a.ctl[2];
    unsigned long long *mpub = (unsigned long long *)&a.ctl[3];
    unsigned long long *mclosed = (unsigned long long *)&a.ctl[4];
    auto dload = [](unsigned long long *p) {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    auto claimed = [](unsigned int c) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        return c;
    };
    // once one workgroup has given up waiting, indices it claimed may be
    // published and never run: every workgroup ends, and the host re-submits
    // whatever was published but not done in a new launch
    if (dload(expired) != 0ull) return 0xffffffffu;
    const unsigned int c = atomicAdd((unsigned int *)&a.ctl[0], 1u);
    if ((int64_t)c >= a.nq) return c;   // past this session's query capacity
    const uint64_t t0 = now_ticks();
    for (unsigned int it = 0;; it++) {
        if (dload(expired) != 0ull) return 0xffffffffu;
        unsigned long long p = dload(mpub);
        if ((uint64_t)c < p) return claimed(c);
        if ((uint64_t)c == p) {
            // the poller: host count (and, every 8th round, the close flag)
            const unsigned long long hp = __hip_atomic_load(a.pub, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
            if (hp > p) {
                __hip_atomic_fetch_max(mpub, hp, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                continue;
            }
            if ((it & 7u) == 7u && __hip_atomic_load(a.closed, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) != 0u) {
                // the host publishes before it closes: read the count once more
                const unsigned long long hp2 =
                    __hip_atomic_load(a.pub, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_fetch_max(mpub, hp2, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(mclosed, 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                if ((uint64_t)c < hp2) return claimed(c);
                return 0xffffffffu;
            }
        } else if (dload(mclosed) != 0ull) {
            // closed: the mirror was raised to the final count before the flag
            if ((uint64_t)c < dload(mpub)) return claimed(c);
            return 0xffffffffu;
        }
        if (now_ticks() - t0 > a.idle_ticks) {
            atomicAdd(expired, 1ull);
            return 0xffffffffu;
        }
        // the poller reads every ~3.4 us (s_sleep 127 = 8128 clocks), the
        // others the device mirror every ~10 us
        __builtin_amdgcn_s_sleep(127);
        if ((uint64_t)c != p) {
            __builtin_amdgcn_s_sleep(127);
            __builtin_amdgcn_s_sleep(127);
        }
    }
}

// A published query record, read from host-coherent memory with system-scope
// loads (the record's line may have been read before the host wrote it).
__device__ __forceinline__ flsg_query session_query(const BatchArgs &a, unsigned int q) {
    static_assert(sizeof(flsg_query) == 32, "flsg_query is four 8-byte words");
    const unsigned long long *w = reinterpret_cast<const unsigned long long *>(a.queries + q);
    unsigned long long v[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        v[k] = __hip_atomic_load(const_cast<unsigned long long *>(w + k), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    flsg_query qq;
    __builtin_memcpy(&qq, v, sizeof(qq));
    return qq;
}

// End of a query: every thread's writes done; in a serving session thread 0
// then makes them visible system-wide and sets the query's done flag.
__device__ __forceinline__ void query_done(const BatchArgs &a, unsigned int q) {
    __syncthreads();
    if (a.busy && threadIdx.x == 0) atomicSub(a.busy, 1u);
    if (a.done && threadIdx.x == 0) {
        __threadfence_system();
        __hip_atomic_store(a.done + q, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_batch(BatchArgs a) {
    __shared__ WaveLds s_w[NT / kWave];
    __shared__ unsigned long long s_red[NT / kWave];
    __shared__ double s_redd[NT / kWave];
    __shared__ unsigned int s_q, s_cnt, s_ncnt, s_sup, s_tn, s_nlong, s_work, s_npc, s_tl, s_nl, s_ex, s_nf, s_dirty;
    __shared__ unsigned int s_pvol;   // FP32 certificate: volume of the pulled rows (saturating)
    __shared__ unsigned long long s_base, s_vol, s_ivol, s_fvol;
    __shared__ unsigned int s_bins[kMaxBins];   // binned push: write cursors, then ends of the pair regions
    // dynamic LDS: the hub cache (residuals of nodes [0, H) for the whole
    // query -- ids are degree-ranked on relabelled graphs, so these take the
    // largest share of the pushes: LDS atomics instead of memory-side ones),
    // or the binned push's counters and destination slice, or the
    // certificate's copy of the hubs' x32
    extern __shared__ double s_hub[];

    const int tid = threadIdx.x;
    WaveLds &W = s_w[tid / kWave];
    const int64_t slot = (int64_t)(a.slot_base + (int)blockIdx.x);
    const int64_t off = slot * a.n;
    SlotView S;
    S.rx32 = a.rx32 + slot * a.rn;
    S.fl = a.fl + slot * a.rn;
    S.rx64 = a.rx64 + slot * a.rn;
    S.hk2 = a.hk2 + slot * a.rn;
    const int sink = (int)a.n;
    int32_t *qa = a.q0 + off;
    int32_t *qb = a.q1 + off;
    Ent *ent = a.ent + off;
    int32_t *sup = a.sup + off;
    uint32_t *pcs = a.pcs + slot * a.pcs_stride;
    int32_t *tl = a.tl + off;
    int32_t *nl = a.nl + off;
    uint4 *pairs = a.pairs ? a.pairs + slot * a.pcap : nullptr;
    const int32_t *indptr = a.indptr;
    const int32_t *degree = a.degree;
    const int2 *colx = a.colx;
    const int pw = a.piece_w;
    const int hw = a.hub_w;

    for (;;) {
        if (tid == 0) s_q = a.pub ? session_claim(a) : atomicAdd((unsigned int *)&a.ctl[0], 1u);
        __syncthreads();
        const unsigned int c = s_q;
        __syncthreads();
        if ((int64_t)c >= a.nq) break;
        if (tid == 0 && a.busy) atomicAdd(a.busy, 1u);   // back down in query_done
        const unsigned int q = a.order ? (unsigned int)a.order[c] : c;
        const uint64_t t_start = now_ticks();
        const flsg_query qq = a.pub ? session_query(a, q) : a.queries[q];
        // a query the launch cannot run fails on its own record, the other
        // queries of the launch go on: seed outside the graph, unknown kind,
        // a heat-kernel query in a launch without heat-kernel settings, PPR /
        // Katz parameters outside the range where every push is >= 0
        // (0 < alpha < 1 for PPR, beta > 0 for Katz, eps > 0), or an FP32
        // push on a graph that is not undirected (its certificate's error
        // bounds count a node's in-edges by its degree)
        const bool bad_kind = qq.kind != FLSG_PPR && qq.kind != FLSG_KATZ && qq.kind != FLSG_HK;
        const bool bad_par = qq.kind == FLSG_HK ? a.hk_n <= 0
                                                : !(qq.eps > 0.0) || !(qq.alpha > 0.0) ||
                                                      (qq.kind == FLSG_PPR && !(qq.alpha < 1.0)) ||
                                                      (qq.precision != FLSG_FP64 && !a.undirected);
        if (qq.seed < 0 || qq.seed >= a.n || bad_kind || bad_par) {
            if (tid == 0) {
                flsg_result rr{};
                rr.status = FLSG_EARG;
                a.res[q] = rr;
                a.out_off[q] = -1;
            }
            query_done(a, q);
            continue;
        }
        if (qq.kind == FLSG_HK) {
            if (qq.precision == FLSG_FP64)
                hk_query<NT, true>(a, qq, q, S, qa, qb, ent, sup, pcs, W, s_hub, &s_cnt, &s_ncnt, &s_sup, &s_npc,
                                   &s_work, &s_tn, &s_vol, &s_base, t_start, sink, pairs, s_bins);
            else
                hk_query<NT, false>(a, qq, q, S, qa, qb, ent, sup, pcs, W, s_hub, &s_cnt, &s_ncnt, &s_sup, &s_npc,
                                    &s_work, &s_tn, &s_vol, &s_base, t_start, sink, pairs, s_bins);
            query_done(a, q);
            continue;
        }
        const QParams P = qparams(qq);
        const int s = (int)qq.seed;
        int64_t iters = 0;
        unsigned long long fnodes = 0;
        int32_t *qc = qa;
        int status;
        uint64_t tick[5] = {0, 0, 0, 0, 0};   // node, pairs, bins, direct (ticks), binned iterations
        // ---------------- 1. local push (ppr_solver.py:246-288, katz_solver.py:183-224)
        if (P.fp64)
            status = local_push<NT, double>(a, P, s, reinterpret_cast<double *>(S.rx64), qa, qb, ent, sup, pcs,
                                            tl, nl, reinterpret_cast<Pair<double> *>(pairs), W, s_hub, s_bins, &s_cnt, &s_ncnt, &s_sup, &s_npc,
                                            &s_work, &s_tl, &s_nl, &s_nf, &s_vol, &s_ivol, &s_fvol, sink, &iters,
                                            &fnodes, &qc, tick);
        else
            status = local_push<NT, float>(a, P, s, reinterpret_cast<float *>(S.rx32), qa, qb, ent, sup, pcs,
                                           tl, nl, reinterpret_cast<Pair<float> *>(pairs), W, s_hub, s_bins, &s_cnt,
                                           &s_ncnt, &s_sup, &s_npc, &s_work, &s_tl, &s_nl, &s_nf, &s_vol, &s_ivol,
                                           &s_fvol, sink, &iters, &fnodes, &qc, tick);
        int32_t *qn = qc == qa ? qb : qa;
        const unsigned long long oper = s_vol;
        const uint64_t t_push = now_ticks();
        unsigned long long cert_ops = 0, corr_ops = 0;
        unsigned long long n_corr = 0;
        // the adds cap of the FP32 error bounds: the query's pushes, on a
        // graph whose rows repeat no entry
        const double Fq = a.simple ? (double)fnodes : __builtin_inf();
        uint64_t t_rec = t_push, t_corr = t_push;
        auto load_ent = [&](unsigned int i, int &st, int &d, double &v) {
            const Ent e = ent[i];
            st = e.st;
            d = e.d;
            v = e.v;
        };
        if (!P.fp64) {
            // ---------------- 2. FP64 certificate of the FP32 result
            // Exact FP64 residual r64 = b - x + push(x) of the support (from the
            // FP32 x) and of the boundary nodes the error bound (lazy_bound)
            // cannot certify; every other boundary node keeps its FP32 residual
            // and the bound.  The exact set: the support, then the boundary
            // candidates nl[] the push listed (degree > kLazyDeg, or residual
            // within 2^-8 of the threshold) -- or every touched boundary node
            // when the bound does not hold at this iteration count or nl[]
            // overflowed.
            const unsigned int S32 = s_sup;
            const unsigned int NL = s_nl;
            const unsigned int T32 = s_tl;
            const bool lazy = (int64_t)NL <= a.n && lazy_g(adds_max(kLazyDeg, iters, Fq), iters) <= kLazyGmax;
            // the dynamic LDS may hold x32 of the first ids for the pull (filled
            // below when the pulled rows are long enough to pay for it)
            const int xn_max = (int)min((int64_t)(a.hub_bytes / 4), a.n);
            float *xc = reinterpret_cast<float *>(s_hub);
            if (tid == 0) {
                s_cnt = 0;
                s_ncnt = 0;
                s_npc = 0;
                s_work = 0;
                s_nlong = 0;
                s_pvol = lazy ? 0u : ~0u;   // every touched boundary row pulled: fill
            }
            __syncthreads();
            {
                const unsigned int lim = lazy ? NL : T32;
                const int32_t *src = lazy ? nl : tl;
                // kScanUnroll entries per lane, their record and degree loads
                // issued together (one memory round trip per batch)
                for (unsigned int i0 = 0; i0 < lim; i0 += NT * kScanUnroll) {
                    int vv[kScanUnroll], dd[kScanUnroll];
                    float2 rx[kScanUnroll];
#pragma unroll
                    for (int j = 0; j < kScanUnroll; j++) {
                        const unsigned int i = i0 + j * NT + tid;
                        vv[j] = i < lim ? (int)((uint32_t)src[i] & 0xffffffu) : -1;   // tl entries carry a degree
                    }
#pragma unroll
                    for (int j = 0; j < kScanUnroll; j++) {
                        rx[j] = vv[j] >= 0 ? S.rx32[vv[j]] : make_float2(0.0f, 0.0f);
                        dd[j] = (lazy && vv[j] >= 0) ? degree[vv[j]] : 0;
                    }
#pragma unroll
                    for (int j = 0; j < kScanUnroll; j++) {
                        int v = vv[j];
                        if (v >= 0) {
                            if (!pristine(rx[j].y)) {
                                v = -1;   // pushed: already in the support list
                            } else if (lazy) {
                                // a candidate listed for its degree (> kLazyDeg) whose
                                // final FP32 residual the bound still certifies keeps it
                                const int d = dd[j];
                                const float rv = rx[j].x;
                                const double m = adds_max(d, iters, Fq);
                                if (d > kLazyDeg && lazy_valid(m) &&
                                    fabs((double)rv) + lazy_bound(rv, m, iters) <
                                        kNearFrac * P.thr_scale * (double)d)
                                    v = -1;
                                else
                                    atomic_add_sat(&s_pvol, (unsigned int)d);
                            }
                        }
                        const long long k = wave_append(v >= 0, &s_ncnt);
                        if (k >= 0) qn[k] = v;
                    }
                }
                // the seed's residual was set, not added: never listed by the push
                if (lazy && tid == 0) {
                    const int ds = degree[s];
                    if (pristine(S.rx32[s].y) && (ds > kLazyDeg || P.r0 >= kNearFrac * P.thr_scale * (double)ds))
                        qn[atomicAdd(&s_ncnt, 1u)] = s;
                }
            }
            __syncthreads();
            // pushed nodes (x32 > 0) the local bound certifies (sup_bound) keep
            // their FP32 residual and estimate untouched; the others are listed
            // in nl[] (the boundary candidates are consumed above) for the pull.
            // Read-only: over the support list, or -- when the query pushed more
            // than a quarter of the graph -- streamed over the slot's records
            // in id order (coalesced).  No block barrier in the loop.
            if (tid == 0) s_ex = 0;
            __syncthreads();
            {
                const bool sstream = (int64_t)S32 * 4 > a.n;
                const unsigned int slim = sstream ? (unsigned int)a.n : S32;
                for (unsigned int i0 = 0; i0 < slim; i0 += NT * kScanUnroll) {
                    int vv[kScanUnroll];
                    float2 rx[kScanUnroll];
                    // a support-list entry carries its degree (below 255);
                    // streamed, every entry's degree loads with its record
                    // (coalesced, one memory round trip per batch)
                    int dd[kScanUnroll];
#pragma unroll
                    for (int j = 0; j < kScanUnroll; j++) {
                        const unsigned int i = i0 + j * NT + tid;
                        vv[j] = -1;
                        dd[j] = 0;
                        if (i < slim) {
                            if (sstream) {
                                vv[j] = (int)i;
                            } else {
                                const uint32_t e = (uint32_t)sup[i];
                                vv[j] = (int)(e & 0xffffffu);
                                dd[j] = (int)(e >> 24);
                            }
                        }
                    }
#pragma unroll
                    for (int j = 0; j < kScanUnroll; j++)
                        if (vv[j] >= 0 && (sstream || dd[j] == 255)) dd[j] = degree[vv[j]];
#pragma unroll
                    for (int j = 0; j < kScanUnroll; j++) rx[j] = vv[j] >= 0 ? S.rx32[vv[j]] : make_float2(0.0f, 0.0f);
#pragma unroll
                    for (int j = 0; j < kScanUnroll; j++) {
                        bool exact = false;
                        if (rx[j].y != 0.0f)
                            exact = !support_certified(rx[j].x, rx[j].y, dd[j], iters, vv[j] == s ? P.r0 : 0.0,
                                                       P.thr_scale, Fq);
                        if (exact) atomic_add_sat(&s_pvol, (unsigned int)dd[j]);
                        const long long k = wave_append(exact, &s_ex);
                        if (k >= 0) nl[k] = vv[j];
                    }
                }
            }
            __syncthreads();
            const unsigned int E32 = s_ex;
            const unsigned int NLIST = E32 + s_ncnt;
            // x32 of the first xn ids into LDS when the pulled rows' volume is
            // at least an eighth of xn (a pulled entry into those ids then reads
            // LDS instead of a memory line; the copy reads xn / 16 lines)
            const int xn = s_pvol >= (unsigned int)xn_max / 8u ? xn_max : 0;
            for (int v0 = tid; v0 < xn; v0 += NT * 4) {
                float t[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int v = v0 + k * NT;
                    t[k] = v < xn ? S.rx32[v].y : 0.0f;
                }
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (v0 + k * NT < xn) xc[v0 + k * NT] = t[k];
            }
            // every entry whose flags the certificate sets is listed in
            // nl[0, s_dirty): the exact support rows nl[0, E32), then the
            // other exact rows and the entries the correction push reaches
            // first (at most the touched entries, so at most n)
            if (tid == 0) s_dirty = E32;
            __syncthreads();
            // a listed entry, finished: r64 = b - x32 + (pulled pushes) in FP64,
            // x64 = 0 (x64 holds the correction push's FP64 additions to x32),
            // FP32 residual field -0.0, flags kExact | kR64, queued for the
            // correction push if |r64| >= thr
            auto finish = [&](bool ok, int v, float x32v, int d, double acc) {
                bool enq = false;
                if (ok) {
                    const double xd = (double)x32v;
                    double rt = acc - xd;
                    if (v == s) rt = rt + P.r0;
                    enq = fabs(rt) >= P.thr_scale * (double)d;
                    S.rx64[v] = make_double2(rt, 0.0);
                    S.rx32[v].x = -0.0f;
                    S.fl[v] = kExact | kR64 | (enq ? kInQ : 0u);
                    // a boundary row (never pushed) is not in nl[0, E32)
                    if (pristine(x32v)) nl[atomicAdd(&s_dirty, 1u)] = v;
                }
                const long long sl = wave_append(enq, &s_cnt);
                if (sl >= 0) qc[sl] = v;
            };
            auto push_of = [&](float xu, int du) { return bpush_value(P, (double)xu, du); };
            // 2a. rows of the support and of the listed boundary nodes
            pull_rows(
                NLIST,
                [&](unsigned int i, int &v, int &st) {
                    v = i < E32 ? nl[i] : qn[i - E32];
                    st = indptr[v];
                },
                S, degree, colx, xc, xn, hw, pw, ent, pcs, &s_npc, &s_nlong, &s_work, W, push_of, finish, cert_ops);
            __syncthreads();
            if (tid == 0) s_work = 0;
            __syncthreads();
            // 2b. pieces of the long rows, then their entries
            pull_pieces(s_npc, pcs, ent, pw, S, colx, xc, xn, &s_work, push_of);
            __syncthreads();
            {
                const unsigned int nlong = s_nlong;
                for (unsigned int j0 = 0; j0 < nlong; j0 += NT) {
                    const unsigned int j = j0 + tid;
                    int v = -1, d = 0;
                    float xv = 0.0f;
                    double acc = 0.0;
                    if (j < nlong) {
                        v = (int)ent[j].v;
                        xv = S.rx32[v].y;
                        d = degree[v];
                        acc = S.rx64[v].x;
                    }
                    finish(j < nlong, v, xv, d, acc);
                }
            }
            __syncthreads();
            t_rec = now_ticks();
            // 2d. FP64 correction push (mark-based queue: the recomputed residual has both signs)
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
                if (tid == 0) {
                    s_ncnt = 0;
                    s_npc = 0;
                    s_work = 0;
                    s_vol = 0;
                }
                __syncthreads();
                unsigned long long pushed = 0, cvol = 0;
                for (unsigned int i = tid; i < cnt; i += NT) {
                    const int u = qc[i];
                    const int d = degree[u];
                    const uint32_t fl = fl_and(&S.fl[u], ~kInQ);
                    // a pending FP32 residual (rl > 0) joins the FP64 one; once
                    // pushed its error bound stays (rl < 0)
                    const float2 rx = S.rx32[u];
                    const float rl = rx.x;
                    const double tmp = S.rx64[u].x + (rl > 0.0f ? (double)rl : 0.0);
                    Ent e;
                    e.st = 0;
                    e.d = 0;
                    e.v = 0.0;
                    if (fabs(tmp) + rec_bound_ool(fl, rl, rx.y, d, iters, u == s ? P.r0 : 0.0, Fq) >=
                        P.thr_scale * (double)d) {
                        S.rx64[u].x = 0.0;
                        if (rl > 0.0f) S.rx32[u].x = -rl;
                        S.rx64[u].y = S.rx64[u].y + tmp;
                        // a node the FP32 push never pushed joins the support
                        if ((fl & kSeen) == 0 && rx.y == 0.0f) {
                            fl_or(&S.fl[u], kSeen);
                            sup[atomicAdd(&s_sup, 1u)] = sup_entry(u, d);
                        }
                        e.st = indptr[u];
                        e.d = d;
                        e.v = bpush_value(P, tmp, d);
                        pushed += 1;
                        cvol += (unsigned long long)d;
                        add_pieces(i, d, pw, hw, pcs, &s_npc);
                    }
                    ent[i] = e;
                }
                if (pushed) atomicAdd(&s_vol, pushed);
                cert_ops += cvol;
                corr_ops += cvol;
                __syncthreads();
                n_corr += s_vol;
                wave_edges<double>(
                    cnt, pcs, s_npc, pw, hw, &s_work, W, colx, load_ent,
                    [&](int vtx, double val, double &old, int stage) {
                        if (stage == 1) old = atomic_add_wg(&S.rx64[vtx].x, val);
                    },
                    [&](int vtx, int dv, double val, double old, bool act) -> int {
                        if (!act) return 0;
                        const double nw = old + val;
                        const float2 rx = S.rx32[vtx];
                        const float rl = rx.x;
                        int res = 0;
                        // entries the main push never reached still hold r == +0.0;
                        // every entry the correction push adds to is flagged kR64
                        const bool fresh = old == 0.0 && pristine(rl);
                        const uint32_t of = fl_or(&S.fl[vtx], kR64 | (fresh ? kTouched : 0u));
                        if (fresh && (of & kTouched) == 0) res = 2;
                        if (of == 0u) nl[atomicAdd(&s_dirty, 1u)] = vtx;   // first flags of this entry
                        const double eff = nw + (rl > 0.0f ? (double)rl : 0.0);
                        const double bv = rec_bound(of, rl, rx.y, dv, iters, vtx == s ? P.r0 : 0.0, Fq);
                        if (fabs(eff) + bv >= P.thr_scale * (double)dv && (fl_or(&S.fl[vtx], kInQ) & kInQ) == 0)
                            res |= 1;
                        return res;
                    },
                    qn, &s_ncnt, tl, &s_tl, sink);
                __syncthreads();
                if (tid == 0) s_cnt = s_ncnt;
                int32_t *t = qc;
                qc = qn;
                qn = t;
                __syncthreads();
            }
            __syncthreads();
            t_corr = now_ticks();
        }
        const unsigned int Sn = s_sup;
        const unsigned int T = s_tl;
        // ---------------- 3. output, certificate value, reset of every touched entry
        if (tid == 0) s_base = atomicAdd((unsigned long long *)&a.ctl[1], (unsigned long long)Sn);
        __syncthreads();
        const unsigned long long obase = s_base;
        const bool fits = obase + Sn <= (unsigned long long)a.out_cap;
        // one pass over the touched entries: certificate value, output of the
        // support (listed by flag in the FP32 path, by x != 0 in the FP64 path,
        // whose x only grows; in any order), reset; the loads of kOutUnroll
        // entries are issued before any is used.  When the query touched more
        // than a quarter of the graph, the pass streams over all n entries of
        // the slot (coalesced, untouched ones -- all bits zero -- are skipped)
        // instead of gathering the touched list's entries one by one.
        if (tid == 0) s_ncnt = 0;
        __syncthreads();
        double cmax = 0.0, rl1 = 0.0;
        const bool stream = (int64_t)T * 4 > a.n;
        const unsigned int lim = stream ? (unsigned int)a.n : T;
        // mode 0 (FP32): the entries the certificate flagged, nl[0, s_dirty):
        //   flags read, FP64 fields read where kR64; every one reset here;
        // mode 1 (FP32): the rest, streamed or from the touched list, with no
        //   flags to read (every flagged entry was listed); entries whose
        //   {r32, x32} is all zero bits are untouched or already done;
        // mode 2 (FP64 push): only {r64, x64} are read and cleared.
        auto out_pass = [&](auto mode_c, const unsigned int cnt) {
            constexpr int mode = decltype(mode_c)::value;
            for (unsigned int i0 = 0; i0 < cnt; i0 += NT * kOutUnroll) {
                int vv[kOutUnroll], dg[kOutUnroll];
                float r32v[kOutUnroll], x32v[kOutUnroll];
                uint32_t flv[kOutUnroll];
                double r64v[kOutUnroll], x64v[kOutUnroll];
                // touched-list entries carry the degree (255: read it); the
                // flagged entries nl[] are plain ids and a streamed pass reads
                // every entry's degree (coalesced): all issued with the records,
                // before any output store (an int store may alias degree[])
#pragma unroll
                for (int j = 0; j < kOutUnroll; j++) {
                    const unsigned int i = i0 + j * NT + tid;
                    vv[j] = -1;
                    dg[j] = 255;
                    if (i < cnt) {
                        if (mode == 0) {
                            vv[j] = nl[i];
                        } else if (stream) {
                            vv[j] = (int)i;
                        } else {
                            const uint32_t e = (uint32_t)tl[i];
                            vv[j] = (int)(e & 0xffffffu);
                            dg[j] = (int)(e >> 24);
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < kOutUnroll; j++)
                    if (vv[j] >= 0 && dg[j] == 255) dg[j] = degree[vv[j]];
                if (mode == 2) {
#pragma unroll
                    for (int j = 0; j < kOutUnroll; j++) {
                        r32v[j] = 0.0f;
                        x32v[j] = 0.0f;
                        flv[j] = 0u;
                        r64v[j] = 0.0;
                        x64v[j] = 0.0;
                        if (vv[j] >= 0) {
                            const double2 w = S.rx64[vv[j]];
                            r64v[j] = w.x;
                            x64v[j] = w.y;
                        }
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < kOutUnroll; j++) {
                        r32v[j] = 0.0f;
                        x32v[j] = 0.0f;
                        flv[j] = 0u;
                        if (vv[j] >= 0) {
                            const float2 w = S.rx32[vv[j]];
                            r32v[j] = w.x;
                            x32v[j] = w.y;
                            if (mode == 0) flv[j] = S.fl[vv[j]];
                        }
                    }
#pragma unroll
                    for (int j = 0; j < kOutUnroll; j++) {
                        r64v[j] = 0.0;
                        x64v[j] = 0.0;
                        if (flv[j] & kR64) {
                            const double2 w = S.rx64[vv[j]];
                            r64v[j] = w.x;
                            x64v[j] = w.y;
                        }
                    }
                }
                // one entry at a time: certificate value, reset, then its output
                // slot (every lane runs the wave append), so only the entries still
                // to come stay live across the bound's call
#pragma unroll
                for (int j = 0; j < kOutUnroll; j++) {
                    bool is_sup = false;
                    int id = vv[j];
                    // FP32 push: x32 plus the correction push's FP64 additions
                    const double outv = mode == 2 ? x64v[j] : (double)x32v[j] + x64v[j];
                    bool live = vv[j] >= 0;
                    if (live && mode == 1)
                        live = (__float_as_uint(r32v[j]) | __float_as_uint(x32v[j])) != 0u;
                    else if (live && mode == 2 && stream)
                        live = !(pristine(r64v[j]) && pristine(x64v[j]));
                    if (live) {
                        const int v = vv[j];
                        const int d = dg[j];
                        // FP32 push: an entry may hold its residual in FP32 (pending,
                        // rl > 0) and carries the error bound of its FP32 values
                        const float rl = r32v[j];
                        const double res = r64v[j] + (rl > 0.0f ? (double)rl : 0.0);
                        const double bnd =
                            mode == 2 ? 0.0 : rec_bound_ool(flv[j], rl, x32v[j], d, iters, v == s ? P.r0 : 0.0, Fq);
                        if (res != 0.0 || bnd != 0.0) {
                            cmax = fmax(cmax, (fabs(res) + bnd) / (P.thr_scale * (double)d));
                            rl1 += fabs(res) + bnd;
                        }
                        // the support: FP32-pushed nodes (x32 > 0) and those the
                        // correction push pushed first (kSeen)
                        is_sup = mode == 2 ? x64v[j] != 0.0 : (x32v[j] != 0.0f || (flv[j] & kSeen) != 0u);
                        if (is_sup && a.out_map) id = a.out_map[v];
                        if (mode != 2) {
                            S.rx32[v] = make_float2(0.0f, 0.0f);
                            if (flv[j] != 0u) S.fl[v] = 0u;
                            if (flv[j] & kR64) S.rx64[v] = make_double2(0.0, 0.0);
                        } else {
                            S.rx64[v] = make_double2(0.0, 0.0);
                        }
                    }
                    const long long k = wave_append(is_sup, &s_ncnt);
                    if (k >= 0 && fits && (unsigned long long)k < Sn) {
                        a.out_ids[obase + k] = id;
                        a.out_vals[obase + k] = outv;
                    }
                }
            }
        };
        if (P.fp64) {
            out_pass(std::integral_constant<int, 2>{}, lim);
        } else {
            out_pass(std::integral_constant<int, 0>{}, s_dirty);
            __syncthreads();   // the listed entries are reset before the main pass skips them
            out_pass(std::integral_constant<int, 1>{}, lim);
        }
        cmax = block_max<NT>(cmax, s_redd);
        rl1 = block_sum_f64<NT>(rl1, s_redd);
        const unsigned long long cops = block_sum_u64<NT>(cert_ops, s_red);
        const unsigned long long kops = block_sum_u64<NT>(corr_ops, s_red);
        if (tid == 0) {
            const uint64_t t_end = now_ticks();
            flsg_result rr{};
            rr.oper = (int64_t)oper;
            rr.iters = iters;
            rr.oper_cert = (int64_t)cops;
            rr.n_corr = (int64_t)n_corr;
            rr.supp = (int64_t)min(s_ncnt, Sn);   // entries written (== Sn: every support entry is listed)
            rr.nodes = (int64_t)fnodes;
            rr.cert = cmax;
            rr.r_l1 = rl1;
            rr.ms = (double)(t_end - t_start) * 1e-5;       // 100 MHz constant clock
            rr.ms_push = (double)(t_push - t_start) * 1e-5;
            rr.ms_recompute = (double)(t_rec - t_start) * 1e-5;
            rr.ms_correct = (double)(t_corr - t_start) * 1e-5;
            rr.status = fits ? status : FLSG_ECAP;
            rr.touched = (int32_t)T;
            rr.ms_node = (double)tick[0] * 1e-5;
            rr.ms_pairs = (double)tick[1] * 1e-5;
            rr.ms_bins = (double)tick[2] * 1e-5;
            rr.ms_direct = (double)tick[3] * 1e-5;
            rr.bin_iters = (int64_t)tick[4];
            rr.corr_edges = (int64_t)kops;
            a.res[q] = rr;
            a.out_off[q] = fits ? (int64_t)obase : -1;
        }
        query_done(a, q);
    }
}

// every slot array not in the between-queries state (all zero), counted into *cnt
template <typename T>
__global__ void k_nonzero(const T *p, int64_t total, unsigned long long *cnt) {
    unsigned long long c = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        c += pristine(p[i]) ? 0ull : 1ull;
    if (c) atomicAdd(cnt, c);
}
__global__ void k_nonzero_u32(const uint32_t *p, int64_t total, unsigned long long *cnt) {
    unsigned long long c = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        c += p[i] ? 1ull : 0ull;
    if (c) atomicAdd(cnt, c);
}

// cost estimate of each query for the claim order: the sum of vol[] over the
// seed's neighbours (vol = the three-hop volumes: the seed's four-hop
// volume; -1 for a seed outside the graph); one wavefront per seed
__global__ void k_seed_cost(const int64_t *seeds, int64_t nq, const int32_t *indptr, const int32_t *degree,
                            const int2 *colx, const int64_t *vol2, int64_t n, int64_t *out) {
    const int lane = __lane_id();
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t i = blockIdx.x * (int64_t)(blockDim.x / kWave) + threadIdx.x / kWave; i < nq; i += waves) {
        const int64_t v = seeds[i];
        long long c = -1;
        if (v >= 0 && v < n) {
            const int st = indptr[v];
            const int d = degree[v];
            c = 0;
            for (int e = lane; e < d; e += kWave) c += vol2[colx[st + e].x];
            c = wave_sum(c);
        }
        if (lane == 0) out[i] = c;
    }
}

// vout[u] = sum over u's neighbours w of vin[w] (vin = NULL: of their degrees),
// one wavefront per row
__global__ void k_vol_hop(const int32_t *indptr, const int32_t *degree, const int2 *colx, int64_t n,
                          const int64_t *vin, int64_t *vout) {
    const int lane = __lane_id();
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t u = blockIdx.x * (int64_t)(blockDim.x / kWave) + threadIdx.x / kWave; u < n; u += waves) {
        const int st = indptr[u];
        const int d = degree[u];
        long long c = 0;
        if (vin) {
            for (int e = lane; e < d; e += kWave) c += vin[colx[st + e].x];
        } else {
            for (int e = lane; e < d; e += kWave) c += colx[st + e].y;
        }
        c = wave_sum(c);
        if (lane == 0) vout[u] = c;
    }
}

__global__ void k_fill_queries(flsg_query *q, const int64_t *seeds, int64_t nq, int32_t kind, double alpha, double eps,
                               int32_t precision) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nq; i += (int64_t)gridDim.x * blockDim.x) {
        flsg_query x;
        x.seed = seeds[i];
        x.kind = kind;
        x.precision = precision;
        x.alpha = alpha;
        x.eps = eps;
        q[i] = x;
    }
}

// column entries with the neighbour's degree; slack entries (id outside
// [0, n)) carry degree 0 and are never read by a live row
__global__ void k_colx(const int32_t *indices, const int32_t *degree, int64_t n, int64_t nnz, int2 *colx) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        const int v = indices[e];
        colx[e] = make_int2(v, (v >= 0 && v < n) ? degree[v] : 0);
    }
}

// live entries (u, v) as sort keys u:v and v:u (dead slack slots: ~0 in both)
__global__ void k_edge_keys(const int32_t *indptr, const int32_t *indices, const int32_t *degree, int64_t n,
                            int64_t nnz, uint64_t *fwd, uint64_t *rev) {
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < n; u += (int64_t)gridDim.x * blockDim.x) {
        const int st = indptr[u], en = indptr[u + 1], d = degree[u];
        for (int e = st; e < en; e++) {
            uint64_t a = ~0ull, b = ~0ull;
            if (e - st < d) {
                const uint32_t v = (uint32_t)indices[e];
                a = ((uint64_t)u << 32) | v;
                b = ((uint64_t)v << 32) | (uint64_t)u;
            }
            fwd[e] = a;
            rev[e] = b;
        }
    }
}

// cnt[0]: positions where the sorted forward and reverse keys differ (0: every
// (u, v) has its (v, u)); cnt[1]: live keys equal to the one before (0: no
// row repeats an entry).  Unused slack entries sort last as ~0.
__global__ void k_count_diff(const uint64_t *a, const uint64_t *b, int64_t m, unsigned long long *cnt) {
    unsigned long long c = 0, r = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        c += a[i] != b[i] ? 1ull : 0ull;
        r += (i > 0 && a[i] != ~0ull && a[i] == a[i - 1]) ? 1ull : 0ull;
    }
    if (c) atomicAdd(&cnt[0], c);
    if (r) atomicAdd(&cnt[1], r);
}

int colx_refresh(flsg_graph g, hipStream_t st) {
    g->vol2_valid = 0;   // the degrees changed (or the copy is new)
    if (!g->colx || g->nnz_cap == 0) return FLSG_OK;
    hipLaunchKernelGGL(k_colx, dim3(g->num_cus * 8), dim3(256), 0, st, (const int32_t *)g->indices,
                       (const int32_t *)g->degree, g->n, g->nnz_cap, g->colx);
    FLSG_HIP(hipGetLastError());
    return FLSG_OK;
}

// built once per handle (synchronously, so launches on other streams see it);
// edge insertions refresh it in place
static int colx_ensure(flsg_graph g, hipStream_t st) {
    if (g->colx) return FLSG_OK;
    FLSG_HIP(hipMalloc((void **)&g->colx, sizeof(int2) * (size_t)(g->nnz_cap > 0 ? g->nnz_cap : 1)));
    int rc = colx_refresh(g, st);
    if (rc) return rc;
    FLSG_HIP(hipStreamSynchronize(st));
    return FLSG_OK;
}

// Three-hop volume of every node, vol2[u] = sum over u's neighbours w of
// their two-hop volumes (the sums of w's neighbours' degrees), for the claim
// order of flsg_batch_solve; rebuilt after edge insertions.  The buffer holds
// 2n entries: the two-hop volumes go to its second half.
static int vol2_ensure(flsg_graph g, hipStream_t st) {
    if (g->vol2 && g->vol2_valid) return FLSG_OK;
    const size_t nn = (size_t)std::max<int64_t>(g->n, 1);
    if (!g->vol2) FLSG_HIP(hipMalloc((void **)&g->vol2, sizeof(int64_t) * 2 * nn));
    hipLaunchKernelGGL(k_vol_hop, dim3(g->num_cus * 8), dim3(256), 0, st, (const int32_t *)g->indptr,
                       (const int32_t *)g->degree, (const int2 *)g->colx, g->n, (const int64_t *)nullptr,
                       g->vol2 + nn);
    hipLaunchKernelGGL(k_vol_hop, dim3(g->num_cus * 8), dim3(256), 0, st, (const int32_t *)g->indptr,
                       (const int32_t *)g->degree, (const int2 *)g->colx, g->n, (const int64_t *)(g->vol2 + nn),
                       g->vol2);
    FLSG_HIP(hipGetLastError());
    g->vol2_valid = 1;
    return FLSG_OK;
}

// Whether every live entry (u, v) has its (v, u) -- the FP32 certificate's
// error bounds count a node's incoming pushes by its degree -- and whether no
// row repeats an entry (then they also cap them by the query's pushes).
// Decided once per handle by sorting the entries both ways (edge insertions
// add both directions of pairs not yet present, so both stay valid).
static int undirected_ensure(flsg_graph g, hipStream_t st) {
    if (g->undirected >= 0) return FLSG_OK;
    const int64_t m = g->nnz_cap;
    if (m == 0) {
        g->undirected = 1;
        g->simple = 1;
        return FLSG_OK;
    }
    if (m >= INT32_MAX) return fail(FLSG_EARG, "graph too large for the symmetry check");
    uint64_t *k[4] = {nullptr, nullptr, nullptr, nullptr};
    unsigned long long *d_cnt = nullptr;
    void *tmp = nullptr;
    size_t tb = 0;
    hipError_t e = hipSuccess;
    for (int i = 0; i < 4 && e == hipSuccess; i++) e = hipMalloc((void **)&k[i], sizeof(uint64_t) * (size_t)m);
    if (e == hipSuccess) e = hipMalloc((void **)&d_cnt, 2 * sizeof(unsigned long long));
    if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortKeys(nullptr, tb, (const uint64_t *)nullptr, (uint64_t *)nullptr, (int)m, 0,
                                              64, st);
    if (e == hipSuccess) e = hipMalloc(&tmp, tb);
    unsigned long long h[2] = {0, 0};
    if (e == hipSuccess) e = hipMemsetAsync(d_cnt, 0, 2 * sizeof(unsigned long long), st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_edge_keys, dim3(g->num_cus * 4), dim3(256), 0, st, g->indptr,
                           (const int32_t *)g->indices, (const int32_t *)g->degree, g->n, m, k[0], k[1]);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortKeys(tmp, tb, k[0], k[2], (int)m, 0, 64, st);
    if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortKeys(tmp, tb, k[1], k[3], (int)m, 0, 64, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_count_diff, dim3(g->num_cus * 4), dim3(256), 0, st, (const uint64_t *)k[2],
                           (const uint64_t *)k[3], m, d_cnt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h, d_cnt, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    for (int i = 0; i < 4; i++) (void)hipFree(k[i]);
    (void)hipFree(d_cnt);
    (void)hipFree(tmp);
    FLSG_HIP(e);
    g->undirected = h[0] == 0 ? 1 : 0;
    g->simple = h[1] == 0 ? 1 : 0;
    return FLSG_OK;
}

void batch_free(flsg_graph g) {
    (void)hipFree(g->colx);
    g->colx = nullptr;
    (void)hipFree(g->vol2);
    g->vol2 = nullptr;
    g->vol2_valid = 0;
    BatchArena *A = g->arena;
    if (!A) return;
    (void)hipFree(A->rx32);
    (void)hipFree(A->fl);
    (void)hipFree(A->rx64);
    (void)hipFree(A->hk2);
    (void)hipFree(A->q0);
    (void)hipFree(A->q1);
    (void)hipFree(A->ent);
    (void)hipFree(A->sup);
    (void)hipFree(A->pcs);
    (void)hipFree(A->tl);
    (void)hipFree(A->nl);
    (void)hipFree(A->pairs);
    (void)hipFree(A->ctl);
    (void)hipFree(A->d_res);
    (void)hipFree(A->d_q);
    (void)hipFree(A->d_order);
    (void)hipFree(A->d_cost);
    delete A;
    g->arena = nullptr;
}

// bytes per slot for n nodes (per-node arrays and lists), without the binned
// push's pair buffer (12 bytes per column entry of the graph)
int64_t slot_bytes(int64_t n) {
    // rx32 {r, x} and fl (8 + 4 B), rx64 {r, x} and hk2 (16 + 8 B) over n + kSink entries; q0 q1 sup tl nl (4 B) +
    // frontier entry (16 B)
    return (n + kSink) * (3 * 4 + 3 * 8) + n * (5 * 4 + 16);
}

static int64_t pair_bytes(flsg_graph g) { return g->nnz_cap * (int64_t)sizeof(uint4); }

static int ensure_arena(flsg_graph g, int32_t slots, hipStream_t st) {
    BatchArena *A = g->arena;
    if (A && A->slots == slots && A->n == g->n) return FLSG_OK;
    batch_free(g);
    A = new BatchArena();
    g->arena = A;
    A->n = g->n;
    A->slots = slots;
    A->rn = g->n + kSink;
    const size_t ns = (size_t)g->n * (size_t)slots;
    const size_t nr = (size_t)A->rn * (size_t)slots;
    FLSG_HIP(hipMalloc((void **)&A->rx32, sizeof(float2) * nr));
    FLSG_HIP(hipMalloc((void **)&A->fl, sizeof(uint32_t) * nr));
    FLSG_HIP(hipMalloc((void **)&A->rx64, sizeof(double2) * nr));
    FLSG_HIP(hipMalloc((void **)&A->hk2, sizeof(double) * nr));
    FLSG_HIP(hipMalloc((void **)&A->q0, sizeof(int32_t) * ns));
    FLSG_HIP(hipMalloc((void **)&A->q1, sizeof(int32_t) * ns));
    FLSG_HIP(hipMalloc((void **)&A->ent, sizeof(Ent) * ns));
    FLSG_HIP(hipMalloc((void **)&A->sup, sizeof(int32_t) * ns));
    // a list's entries are distinct nodes and every split entry has more than
    // 64 edges per piece, so a list never needs more than n + nnz/64 pieces
    A->pcs_stride = g->n + g->nnz_cap / kWave + 1;
    FLSG_HIP(hipMalloc((void **)&A->pcs, sizeof(uint32_t) * (size_t)A->pcs_stride * (size_t)slots));
    FLSG_HIP(hipMalloc((void **)&A->tl, sizeof(int32_t) * ns));
    FLSG_HIP(hipMalloc((void **)&A->nl, sizeof(int32_t) * ns));
    // binned push: one iteration's pairs (at most one per column entry)
    A->pcap = g->nnz_cap > 0 ? g->nnz_cap : 1;
    FLSG_HIP(hipMalloc((void **)&A->pairs, sizeof(uint4) * (size_t)A->pcap * (size_t)slots));
    FLSG_HIP(hipMalloc((void **)&A->ctl, sizeof(uint64_t) * 3));   // [2]: busy workgroups
    FLSG_HIP(hipMemsetAsync(A->rx32, 0, sizeof(float2) * nr, st));
    FLSG_HIP(hipMemsetAsync(A->fl, 0, sizeof(uint32_t) * nr, st));
    FLSG_HIP(hipMemsetAsync(A->rx64, 0, sizeof(double2) * nr, st));
    FLSG_HIP(hipMemsetAsync(A->hk2, 0, sizeof(double) * nr, st));
    return FLSG_OK;
}

static int32_t default_slots(flsg_graph g) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return g->num_cus;
    const int64_t per = slot_bytes(g->n) + pair_bytes(g);
    int64_t fit = (int64_t)(fr / 2) / (per > 0 ? per : 1);
    int64_t want = (int64_t)g->num_cus;
    if (fit < want) want = fit;
    if (want < 1) want = 1;
    return (int32_t)want;
}

// the host-coherent words of a serving session a launch polls (see BatchArgs)
struct SessionWords {
    const uint64_t *pub;
    const uint32_t *closed;
    uint32_t *done;
    uint64_t idle_ticks;
};

static int launch(flsg_graph g, const flsg_query *d_q, int64_t nq, int32_t slot_base, int32_t n_slots,
                  int32_t block, const int32_t *d_out_map, int32_t *d_out_ids, double *d_out_vals,
                  int64_t *d_out_off, int64_t out_cap, uint64_t *d_ctl, flsg_result *d_res, int64_t max_iters,
                  const flsg_hk_config *hk, hipStream_t st, const int32_t *d_order = nullptr,
                  const SessionWords *sw = nullptr, unsigned int *d_busy = nullptr) {
    BatchArena *A = g->arena;
    if (!A) return fail(FLSG_EARG, "no batch arena: call flsg_batch_reserve first");
    if (slot_base < 0 || n_slots <= 0 || slot_base + n_slots > A->slots)
        return fail(FLSG_EARG, "slot range outside the reserved arena");
    int rc = colx_ensure(g, st);
    if (rc) return rc;
    rc = undirected_ensure(g, st);
    if (rc) return rc;
    FLSG_HIP(hipMemsetAsync(d_ctl, 0, sizeof(uint64_t) * (sw ? 5 : 2), st));
    if (d_busy) FLSG_HIP(hipMemsetAsync(d_busy, 0, sizeof(unsigned int), st));
    BatchArgs a{};
    if (sw) {
        a.pub = sw->pub;
        a.closed = sw->closed;
        a.done = sw->done;
        a.idle_ticks = sw->idle_ticks;
    }
    a.indptr = g->indptr;
    a.indices = g->indices;
    a.degree = g->degree;
    a.colx = g->colx;
    a.n = g->n;
    a.queries = d_q;
    a.nq = nq;
    a.order = d_order;
    a.rx32 = A->rx32;
    a.fl = A->fl;
    a.rx64 = A->rx64;
    a.hk2 = A->hk2;
    a.rn = A->rn;
    a.q0 = A->q0;
    a.q1 = A->q1;
    a.ent = A->ent;
    a.sup = A->sup;
    a.pcs = A->pcs;
    a.pcs_stride = A->pcs_stride;
    a.tl = A->tl;
    a.nl = A->nl;
    a.pairs = A->pairs;
    a.pcap = A->pcap;
    a.bin_min = g->batch_bin_min;
    a.dense_min = g->batch_dense_min;
    a.undirected = g->undirected;
    a.simple = g->simple;
    // hub pieces are encoded (entry << 8) | piece: entries < 2^24, < 256 pieces per node
    if (g->n >= (1 << 24)) return fail(FLSG_EARG, "batched solves need n < 2^24");
    // degree bins: entries with more than hub_w edges are split into pieces of
    // piece_w edges (at most 255 per entry)
    int pw = g->batch_piece_w > 0 ? g->batch_piece_w : 1024;
    while ((int64_t)pw * 255 < (int64_t)g->d_max_hint) pw <<= 1;
    a.piece_w = pw;
    a.hub_w = g->batch_hub_w > 0 ? g->batch_hub_w : 256;
    a.slot_base = slot_base;
    a.ctl = d_ctl;
    a.busy = d_busy;
    // binned iterations while at least a quarter of the device's compute
    // units run a query (PPR at eps 1e-7 on the 2.39M-node graph: the binned
    // push is 1.5x faster than the direct one with 256 slots busy, even with
    // 64, 14% slower with 16, 25% slower for a lone query)
    a.bin_busy_min = g->batch_bin_busy > 0 ? g->batch_bin_busy : std::max<int32_t>(1, g->num_cus / 4);
    a.out_map = d_out_map;
    a.out_ids = d_out_ids;
    a.out_vals = d_out_vals;
    a.out_off = d_out_off;
    a.out_cap = out_cap;
    a.res = d_res;
    a.max_iters = max_iters > 0 ? max_iters : 20000;
    if (hk) {
        if (hk->tau < 1.0 || !(hk->eps > 0.0)) return fail(FLSG_EARG, "heat kernel: tau >= 1 and eps > 0");
        const int64_t N = hk->n_terms > 0 ? hk->n_terms : host_estimate_n(hk->tau, .5 * hk->eps);   // hk_solver.py:225
        if (N > kMaxHkTerms) return fail(FLSG_EARG, "heat kernel: Taylor order above the batched cap (96)");
        a.hk_n = (int32_t)N;
        a.hk_tau = hk->tau;
        a.hk_ete = exp(hk->tau) * hk->eps;
        a.hk_psi[N] = 1.;
        for (int64_t i = N - 1; i >= 0; i--) a.hk_psi[i] = a.hk_psi[i + 1] * hk->tau / ((double)i + 1.) + 1.;
    }
    const int grid = n_slots;
    int hub = g->batch_hub_bytes;
    if (hub < 0) hub = block == 1024 ? 136 * 1024 : block == 512 ? 48 * 1024 : 24 * 1024;
    a.hub_bytes = hub;
    const int bi = block == 1024 ? 0 : block == 256 ? 1 : 2;
    const void *fn = bi == 0   ? reinterpret_cast<const void *>(&k_batch<1024>)
                     : bi == 1 ? reinterpret_cast<const void *>(&k_batch<256>)
                               : reinterpret_cast<const void *>(&k_batch<512>);
    // the dynamic-LDS limit is raised once per device and kernel (a call per
    // launch would sit on the serving loop's launch path)
    static std::mutex attr_mu;
    static std::map<std::pair<int, int>, int> attr_set;
    {
        std::lock_guard<std::mutex> lk(attr_mu);
        int &have = attr_set[std::make_pair(g->device, bi)];
        if (hub > have) {
            FLSG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, hub));
            have = hub;
        }
    }
    if (block == 1024) hipLaunchKernelGGL(k_batch<1024>, dim3(grid), dim3(1024), hub, st, a);
    else if (block == 256) hipLaunchKernelGGL(k_batch<256>, dim3(grid), dim3(256), hub, st, a);
    else hipLaunchKernelGGL(k_batch<512>, dim3(grid), dim3(512), hub, st, a);
    FLSG_HIP(hipGetLastError());
    return FLSG_OK;
}

// after an edge insertion: nothing per slot follows the degrees any more
// (they are read from the graph), so only the handle-level state is touched
int batch_degrees_changed(flsg_graph g, const int32_t *d_rows, int64_t n_rows, hipStream_t st) {
    (void)g;
    (void)d_rows;
    (void)n_rows;
    (void)st;
    return FLSG_OK;
}

}  // namespace flsg

using namespace flsg;

extern "C" int64_t flsg_batch_slot_bytes(int64_t n) { return slot_bytes(n); }

// pushes <= 0: no cap (m = d iters)
static double host_adds(int32_t d, int64_t iters, int64_t pushes) {
    return adds_max(d, iters, pushes > 0 ? (double)pushes : __builtin_inf());
}

extern "C" double flsg_boundary_bound(float r32, int32_t d, int64_t iters, int64_t pushes) {
    return lazy_bound(r32, host_adds(d, iters, pushes), iters);
}

extern "C" double flsg_support_bound(float x32, float r32, int32_t d, int64_t iters, double b, int64_t pushes) {
    return sup_bound(x32, r32, host_adds(d, iters, pushes), iters, b);
}

extern "C" int flsg_batch_set_slots(flsg_graph g, int32_t slots) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (slots < 0) return fail(FLSG_EARG, "slots < 0");
    g->batch_slots = slots;
    return FLSG_OK;
}

extern "C" int flsg_batch_set_degree_bins(flsg_graph g, int32_t split_edges, int32_t piece_edges) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (split_edges != 0 && (split_edges < kWave || split_edges > (1 << 20)))
        return fail(FLSG_EARG, "split_edges: 0 or 64..2^20");
    if (piece_edges != 0 && (piece_edges < kWave || piece_edges > (1 << 24)))
        return fail(FLSG_EARG, "piece_edges: 0 or 64..2^24");
    g->batch_hub_w = split_edges;
    g->batch_piece_w = piece_edges;
    return FLSG_OK;
}

extern "C" int flsg_batch_set_bin_load(flsg_graph g, int32_t min_busy) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (min_busy < 0) return fail(FLSG_EARG, "bin load: min_busy >= 0");
    g->batch_bin_busy = min_busy;
    return FLSG_OK;
}

extern "C" int flsg_batch_set_binning(flsg_graph g, int64_t min_edges, int32_t dense_pairs) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (min_edges < 0 || dense_pairs < 0) return fail(FLSG_EARG, "binning: min_edges >= 0 and dense_pairs >= 0");
    g->batch_bin_min = min_edges == 0 ? INT64_MAX : min_edges;
    g->batch_dense_min = dense_pairs;
    return FLSG_OK;
}

extern "C" int flsg_batch_set_hub_cache(flsg_graph g, int32_t bytes) {
    int rc = check_graph(g);
    if (rc) return rc;
    // the largest: the device's LDS per workgroup (160 KiB on gfx950) less
    // the static LDS of the 1024-thread kernel
    int lds_max = 0;
    FLSG_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, g->device));
    hipFuncAttributes fa;
    FLSG_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&k_batch<1024>)));
    const int64_t cap = ((int64_t)lds_max - (int64_t)fa.sharedSizeBytes) / 8 * 8;
    if (bytes < -1 || bytes > cap || (bytes > 0 && bytes % 8 != 0))
        return fail(FLSG_EARG, "hub cache bytes: -1, 0 or a multiple of 8 up to " + std::to_string(cap) +
                                   " (LDS per workgroup less the kernel's static LDS)");
    g->batch_hub_bytes = bytes;
    return FLSG_OK;
}

extern "C" int flsg_batch_set_block(flsg_graph g, int32_t block) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (block != 256 && block != 512 && block != 1024) return fail(FLSG_EARG, "block must be 256, 512 or 1024");
    g->batch_block = block;
    return FLSG_OK;
}

extern "C" int flsg_graph_set_out_map(flsg_graph g, const int32_t *d_out_map) {
    int rc = check_graph(g);
    if (rc) return rc;
    g->out_map = d_out_map;
    return FLSG_OK;
}

extern "C" int flsg_batch_reserve(flsg_graph g, int32_t slots, void *stream) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (slots <= 0) slots = g->batch_slots > 0 ? g->batch_slots : default_slots(g);
    g->batch_slots = slots;
    rc = ensure_arena(g, slots, (hipStream_t)stream);
    if (rc) return rc;
    rc = colx_ensure(g, (hipStream_t)stream);
    if (rc) return rc;
    return undirected_ensure(g, (hipStream_t)stream);
}

extern "C" int flsg_batch_refresh_degrees(flsg_graph g, const int32_t *d_keys, int64_t n_keys, void *stream) {
    int rc = check_graph(g);
    if (rc) return rc;
    return batch_degrees_changed(g, d_keys, n_keys, (hipStream_t)stream);
}

extern "C" int flsg_batch_arena_dirty(flsg_graph g, int64_t *h_count, void *stream) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (!h_count) return fail(FLSG_EARG, "NULL argument");
    *h_count = 0;
    BatchArena *A = g->arena;
    if (!A) return FLSG_OK;
    hipStream_t st = (hipStream_t)stream;
    unsigned long long *d_cnt = nullptr;
    FLSG_HIP(hipMalloc((void **)&d_cnt, sizeof(unsigned long long)));
    unsigned long long h = 0;
    const int64_t tot = A->rn * A->slots;
    const dim3 gr(g->num_cus * 8), bl(256);
    hipError_t e = hipMemcpyAsync(d_cnt, &h, sizeof(h), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_nonzero<float>, gr, bl, 0, st, (const float *)A->rx32, 2 * tot, d_cnt);
        hipLaunchKernelGGL(k_nonzero_u32, gr, bl, 0, st, (const uint32_t *)A->fl, tot, d_cnt);
        hipLaunchKernelGGL(k_nonzero<double>, gr, bl, 0, st, (const double *)A->rx64, 2 * tot, d_cnt);
        hipLaunchKernelGGL(k_nonzero<double>, gr, bl, 0, st, (const double *)A->hk2, tot, d_cnt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d_cnt, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_cnt);
    FLSG_HIP(e);
    *h_count = (int64_t)h;
    return FLSG_OK;
}

extern "C" int flsg_batch_launch(flsg_graph g, const flsg_query *d_queries, int64_t n_queries, int32_t slot_base,
                                 int32_t n_slots, int32_t *d_out_ids, double *d_out_vals, int64_t *d_out_off,
                                 int64_t out_cap, uint64_t *d_ctl, flsg_result *d_res, int64_t max_iters,
                                 const flsg_hk_config *hk, void *stream) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (!d_queries || !d_ctl || !d_res || !d_out_off || (out_cap > 0 && (!d_out_ids || !d_out_vals)))
        return fail(FLSG_EARG, "NULL argument");
    if (n_queries <= 0) return FLSG_OK;
    if (n_queries >= (int64_t)UINT32_MAX / 2) return fail(FLSG_EARG, "too many queries");
    return launch(g, d_queries, n_queries, slot_base, n_slots, g->batch_block, g->out_map, d_out_ids, d_out_vals,
                  d_out_off, out_cap, d_ctl, d_res, max_iters, hk, (hipStream_t)stream);
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
    if (p->precision != FLSG_FP32 && p->precision != FLSG_FP64) return fail(FLSG_EARG, "unknown precision");
    hipStream_t st = (hipStream_t)stream;
    rc = undirected_ensure(g, st);
    if (rc) return rc;
    if (p->precision == FLSG_FP32 && !g->undirected)
        return fail(FLSG_EARG, "batched FP32 push: the graph is not undirected (its certificate bounds count a "
                               "node's incoming pushes by its degree); use FLSG_FP64");
    const int32_t slots = g->batch_slots > 0 ? g->batch_slots : default_slots(g);
    rc = ensure_arena(g, slots, st);
    if (rc) return rc;
    BatchArena *A = g->arena;
    if (A->d_res_cap < n_queries) {
        if (A->d_res) FLSG_HIP(hipFree(A->d_res));
        if (A->d_q) FLSG_HIP(hipFree(A->d_q));
        if (A->d_order) FLSG_HIP(hipFree(A->d_order));
        if (A->d_cost) FLSG_HIP(hipFree(A->d_cost));
        A->d_res = nullptr;
        A->d_q = nullptr;
        A->d_order = nullptr;
        A->d_cost = nullptr;
        FLSG_HIP(hipMalloc((void **)&A->d_res, sizeof(flsg_result) * n_queries));
        FLSG_HIP(hipMalloc((void **)&A->d_q, sizeof(flsg_query) * n_queries));
        FLSG_HIP(hipMalloc((void **)&A->d_order, sizeof(int32_t) * n_queries));
        FLSG_HIP(hipMalloc((void **)&A->d_cost, sizeof(int64_t) * n_queries));
        A->d_res_cap = n_queries;
        A->d_q_cap = n_queries;
    }
    hipLaunchKernelGGL(k_fill_queries, dim3(64), dim3(256), 0, st, A->d_q, d_seeds, n_queries, p->kind, p->alpha,
                       p->eps, p->precision);
    // claim order: the costliest queries first (started first, the long ones
    // leave the short ones to fill the launch's tail); the cost estimate is the
    // seed's four-hop volume (the sum over its neighbours of their three-hop
    // volumes), which ranks the push work of eps = 1e-7 PPR queries on the
    // 2.39M-node power-law graph with Spearman 0.87 (three-hop 0.87, two-hop
    // 0.84, seed degree 0.40); scheduled greedily by it onto 256 slots with
    // time proportional to work, the makespan is 1.064x the perfectly balanced
    // one (three-hop 1.078x, two-hop 1.19x, the true work 1.062x)
    std::vector<int64_t> cost((size_t)n_queries);
    std::vector<int32_t> order((size_t)n_queries);
    rc = colx_ensure(g, st);
    if (rc) return rc;
    rc = vol2_ensure(g, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_seed_cost, dim3(256), dim3(256), 0, st, d_seeds, n_queries, (const int32_t *)g->indptr,
                       (const int32_t *)g->degree, (const int2 *)g->colx, (const int64_t *)g->vol2, g->n,
                       A->d_cost);
    FLSG_HIP(hipGetLastError());
    FLSG_HIP(hipMemcpyAsync(cost.data(), A->d_cost, sizeof(int64_t) * n_queries, hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n_queries; i++) order[(size_t)i] = (int32_t)i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return cost[(size_t)x] > cost[(size_t)y]; });
    FLSG_HIP(hipMemcpyAsync(A->d_order, order.data(), sizeof(int32_t) * n_queries, hipMemcpyHostToDevice, st));
    const int32_t used = (int32_t)(n_queries < slots ? n_queries : slots);
    rc = launch(g, A->d_q, n_queries, 0, used, g->batch_block, g->out_map, d_out_ids, d_out_vals, d_out_off, out_cap,
                A->ctl, A->d_res, p->max_iters, nullptr, st, A->d_order, nullptr,
                reinterpret_cast<unsigned int *>(A->ctl + 2));
    if (rc) {
        (void)hipStreamSynchronize(st);   // the order copy reads a host vector of this frame
        return rc;
    }
    FLSG_HIP(hipMemcpyAsync(h_res, A->d_res, sizeof(flsg_result) * n_queries, hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n_queries; i++)
        if (h_res[i].status != FLSG_OK) return fail(h_res[i].status, "a batched query failed (see its status)");
    return FLSG_OK;
}

// ---------------------------------------------------------------------------
// Serving sessions: continuous admission into one persistent batched launch.
// The host writes query records into a host-coherent ring and publishes them
// by raising a count the launch's idle workgroups poll; each workgroup takes
// the next published query as soon as its slot is free, writes the result
// record and the done flag into host-coherent memory, and the host reads
// them without synchronising the launch's stream.

struct flsg_session_s {
    flsg_graph g = nullptr;
    int device = 0;
    hipStream_t st = nullptr;        // non-blocking: the open launch never holds up other streams
    int64_t qcap = 0;
    int64_t out_cap = 0;
    flsg_query *h_q = nullptr;       // host-coherent, mapped (ring of qcap records)
    flsg_result *h_res = nullptr;
    int64_t *h_off = nullptr;
    uint32_t *h_done = nullptr;
    uint64_t *h_pub = nullptr;
    uint32_t *h_closed = nullptr;
    flsg_query *d_q = nullptr;       // device addresses of the same memory
    flsg_result *d_res = nullptr;
    int64_t *d_off = nullptr;
    uint32_t *d_done = nullptr;
    uint64_t *d_pub = nullptr;
    uint32_t *d_closed = nullptr;
    uint64_t *d_ctl = nullptr;       // device: [0] claims, [1] output entries used, [2] idle exits,
                                     // [3] / [4] device mirrors of the published count / close flag,
                                     // [5] workgroups running a query (low 32 bits)
    int32_t *d_ids = nullptr;        // output arena (device)
    double *d_vals = nullptr;
    int running = 0;                 // a launch was enqueued and has not been seen to end
    int closed = 0;
    int64_t published = 0;
};

namespace {

template <typename T>
int host_coherent(T **h, T **d, int64_t count) {
    FLSG_HIP(hipHostMalloc((void **)h, sizeof(T) * (size_t)std::max<int64_t>(count, 1),
                           hipHostMallocCoherent | hipHostMallocMapped));
    FLSG_HIP(hipHostGetDevicePointer((void **)d, *h, 0));
    return FLSG_OK;
}

void session_free(flsg_session s) {
    if (s->st) (void)hipStreamSynchronize(s->st);
    (void)hipHostFree(s->h_q);
    (void)hipHostFree(s->h_res);
    (void)hipHostFree(s->h_off);
    (void)hipHostFree(s->h_done);
    (void)hipHostFree(s->h_pub);
    (void)hipHostFree(s->h_closed);
    (void)hipFree(s->d_ctl);
    (void)hipFree(s->d_ids);
    (void)hipFree(s->d_vals);
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
}

// whether the session's launch has ended (the stream is idle); a fault of
// the launch is returned as FLSG_EHIP
int session_ended(flsg_session s, bool *ended) {
    const hipError_t e = hipStreamQuery(s->st);
    if (e == hipSuccess) {
        *ended = true;
        s->running = 0;
        return FLSG_OK;
    }
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();
        *ended = false;
        return FLSG_OK;
    }
    *ended = false;
    return fail(FLSG_EHIP, std::string("serving session: ") + hipGetErrorString(e));
}

}  // namespace

extern "C" int flsg_session_create(flsg_graph g, int64_t query_cap, int64_t out_cap, flsg_session *out) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (!out) return fail(FLSG_EARG, "NULL argument");
    if (query_cap <= 0 || query_cap >= (int64_t)UINT32_MAX / 2) return fail(FLSG_EARG, "query_cap out of range");
    if (out_cap < 0) return fail(FLSG_EARG, "out_cap < 0");
    *out = nullptr;
    int device = 0;
    FLSG_HIP(hipGetDevice(&device));
    flsg_session s = new flsg_session_s();
    s->g = g;
    s->qcap = query_cap;
    s->out_cap = out_cap;
    s->device = device;
    rc = FLSG_OK;
    auto chk = [&](int r) {
        if (r && !rc) rc = r;
    };
    if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) chk(fail(FLSG_EHIP, "stream create"));
    if (!rc) chk(host_coherent(&s->h_q, &s->d_q, query_cap));
    if (!rc) chk(host_coherent(&s->h_res, &s->d_res, query_cap));
    if (!rc) chk(host_coherent(&s->h_off, &s->d_off, query_cap));
    if (!rc) chk(host_coherent(&s->h_done, &s->d_done, query_cap));
    if (!rc) chk(host_coherent(&s->h_pub, &s->d_pub, 1));
    if (!rc) chk(host_coherent(&s->h_closed, &s->d_closed, 1));
    if (!rc && hipMalloc((void **)&s->d_ctl, sizeof(uint64_t) * 6) != hipSuccess) chk(fail(FLSG_ENOMEM, "ctl"));
    if (!rc && out_cap > 0) {
        if (hipMalloc((void **)&s->d_ids, sizeof(int32_t) * out_cap) != hipSuccess ||
            hipMalloc((void **)&s->d_vals, sizeof(double) * out_cap) != hipSuccess)
            chk(fail(FLSG_ENOMEM, "serving session: output arena"));
    }
    if (rc) {
        session_free(s);
        return rc;
    }
    *s->h_pub = 0;
    *s->h_closed = 1;   // no launch yet
    s->closed = 1;
    *out = s;
    return FLSG_OK;
}

extern "C" int flsg_session_open(flsg_session s, int32_t slot_base, int32_t n_slots, const flsg_hk_config *hk,
                                 int64_t max_iters, double idle_seconds) {
    if (!s) return fail(FLSG_EARG, "NULL session");
    if (s->running) return fail(FLSG_EARG, "serving session: the previous launch has not ended");
    if (!(idle_seconds > 0.0) || idle_seconds > 3600.0) return fail(FLSG_EARG, "idle_seconds outside (0, 3600]");
    flsg_graph g = s->g;
    int rc = check_graph(g);
    if (rc) return rc;
    std::memset(s->h_done, 0, sizeof(uint32_t) * (size_t)s->qcap);
    std::memset(s->h_res, 0, sizeof(flsg_result) * (size_t)s->qcap);
    __atomic_store_n(s->h_pub, (uint64_t)0, __ATOMIC_RELEASE);
    __atomic_store_n(s->h_closed, 0u, __ATOMIC_RELEASE);
    s->published = 0;
    SessionWords sw{s->d_pub, s->d_closed, s->d_done, (uint64_t)(idle_seconds * 1e8)};
    rc = launch(g, s->d_q, s->qcap, slot_base, n_slots, g->batch_block, g->out_map, s->d_ids, s->d_vals, s->d_off,
                s->out_cap, s->d_ctl, s->d_res, max_iters, hk, s->st, nullptr, &sw,
                reinterpret_cast<unsigned int *>(s->d_ctl + 5));
    if (rc) {
        __atomic_store_n(s->h_closed, 1u, __ATOMIC_RELEASE);
        s->closed = 1;
        (void)hipStreamSynchronize(s->st);
        return rc;
    }
    s->running = 1;
    s->closed = 0;
    return FLSG_OK;
}

extern "C" int flsg_session_submit(flsg_session s, const flsg_query *h_queries, int64_t n, int64_t *h_first) {
    if (!s || (n > 0 && !h_queries)) return fail(FLSG_EARG, "NULL argument");
    if (s->closed) return fail(FLSG_EARG, "serving session: not open");
    if (n < 0) return fail(FLSG_EARG, "n < 0");
    if (s->published + n > s->qcap) return fail(FLSG_ECAP, "serving session: query ring full (close and reopen)");
    if (h_first) *h_first = s->published;
    std::memcpy(s->h_q + s->published, h_queries, sizeof(flsg_query) * (size_t)n);
    s->published += n;
    // the records before the count (the launch reads the count with acquire)
    __atomic_store_n(s->h_pub, (uint64_t)s->published, __ATOMIC_RELEASE);
    return FLSG_OK;
}

extern "C" int flsg_session_done(flsg_session s, int64_t first, int64_t count, int32_t *h_done) {
    if (!s || (count > 0 && !h_done)) return fail(FLSG_EARG, "NULL argument");
    if (first < 0 || count < 0 || first + count > s->qcap) return fail(FLSG_EARG, "query range outside the ring");
    for (int64_t i = 0; i < count; i++)
        h_done[i] = (int32_t)__atomic_load_n(s->h_done + first + i, __ATOMIC_ACQUIRE);
    return FLSG_OK;
}

extern "C" int flsg_session_result(flsg_session s, int64_t i, flsg_result *h_res, int64_t *h_out_off) {
    if (!s || !h_res) return fail(FLSG_EARG, "NULL argument");
    if (i < 0 || i >= s->published) return fail(FLSG_EARG, "query index not published");
    if (__atomic_load_n(s->h_done + i, __ATOMIC_ACQUIRE) == 0u) return fail(FLSG_EARG, "query not done");
    std::memcpy(h_res, s->h_res + i, sizeof(flsg_result));
    if (h_out_off) *h_out_off = s->h_off[i];
    return FLSG_OK;
}

extern "C" int flsg_session_close(flsg_session s) {
    if (!s) return fail(FLSG_EARG, "NULL session");
    __atomic_store_n(s->h_closed, 1u, __ATOMIC_RELEASE);
    s->closed = 1;
    return FLSG_OK;
}

extern "C" int flsg_session_poll(flsg_session s, int32_t *h_ended, int64_t *h_idle_exits) {
    if (!s || !h_ended) return fail(FLSG_EARG, "NULL argument");
    bool ended = true;
    if (s->running) {
        int rc = session_ended(s, &ended);
        if (rc) return rc;
    }
    *h_ended = ended ? 1 : 0;
    if (h_idle_exits) {
        *h_idle_exits = 0;
        if (ended) {
            uint64_t c[3] = {0, 0, 0};
            FLSG_HIP(hipMemcpyAsync(c, s->d_ctl, sizeof(c), hipMemcpyDeviceToHost, s->st));
            FLSG_HIP(hipStreamSynchronize(s->st));
            *h_idle_exits = (int64_t)c[2];
        }
    }
    return FLSG_OK;
}

extern "C" int flsg_session_output(flsg_session s, int64_t off, int64_t count, int32_t *h_ids, double *h_vals) {
    if (!s || (count > 0 && (!h_ids || !h_vals))) return fail(FLSG_EARG, "NULL argument");
    if (off < 0 || count < 0 || off + count > s->out_cap) return fail(FLSG_EARG, "range outside the output arena");
    bool ended = true;
    if (s->running) {
        int rc = session_ended(s, &ended);
        if (rc) return rc;
    }
    if (!ended) return fail(FLSG_EARG, "serving session: outputs are read after the launch has ended");
    if (count == 0) return FLSG_OK;
    FLSG_HIP(hipMemcpyAsync(h_ids, s->d_ids + off, sizeof(int32_t) * count, hipMemcpyDeviceToHost, s->st));
    FLSG_HIP(hipMemcpyAsync(h_vals, s->d_vals + off, sizeof(double) * count, hipMemcpyDeviceToHost, s->st));
    FLSG_HIP(hipStreamSynchronize(s->st));
    return FLSG_OK;
}

extern "C" int flsg_session_destroy(flsg_session s) {
    if (!s) return FLSG_OK;
    if (!s->closed) flsg_session_close(s);
    session_free(s);
    return FLSG_OK;
}
