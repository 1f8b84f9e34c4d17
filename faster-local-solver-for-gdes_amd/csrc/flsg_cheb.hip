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
//      key = (destination << pbits) | (2 position + [subtraction]) and the
//      sort payload is the FP64 addend itself, so the replay reads every
//      event's value in sorted order instead of gathering it;
//   3. radix sort of the keys (hipCUB) groups every destination's events in
//      the reference's order;
//   4. every destination's events are replayed sequentially in FP64
//      (bit-identical sums) by a thread, a wave or a block depending on
//      how many there are, recording the position of the first crossing;
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
    double *vals[2] = {nullptr, nullptr};   // sorted payload: each event's addend
    int64_t *P = nullptr;          // n + 1
    int32_t *dplus = nullptr;      // n
    uint64_t *cross[2] = {nullptr, nullptr};   // n: compacted, sorted crossings
    uint64_t *xsub = nullptr;      // crossing sub-lists (kSub x xcap)
    int32_t *queue = nullptr;      // n
    unsigned int *cnt = nullptr;   // crossings, medium and heavy destinations (3)
    uint32_t *seg[2] = {nullptr, nullptr};   // medium / heavy sub-lists (kSub x lcap each)
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
                                                     const double *val, const double *dx, int k, int64_t E, int pbits,
                                                     uint64_t *keys, double *vals) {
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
        const bool sub = j >= d;
        const int v = sub ? u : indices[indptr[u] + j];
        keys[pos] = ((uint64_t)(uint32_t)v << pbits) | ((uint64_t)pos << 1) | (sub ? 1u : 0u);
        vals[pos] = sub ? -dx[u] : val[i];      // ppr_solver.py:657 / :652
    }
}

// step 4: sequential replay per destination.  A destination's events must be
// summed one after another in FP64 (bit-identical to the reference), but the
// loads behind them need not be: destinations are split by event count into
//   light  (<= kLight events): the head's lane, from its wave's lanes;
//   medium (<= kBig):          one wave, 64 events loaded per chunk and
//                              staged in LDS, summed from there in order;
//   heavy  (>  kBig):          one 1024-thread block, 15 waves load rounds of
//                              960 events into LDS while wave 0 sums the last.
// Positions are < 2^31, so bit 31 of a position word marks a checked add
// (every add except the subtraction of dx).
constexpr int kLight = 16;
constexpr int kBig = 256;
constexpr int kHB = 1024;
constexpr int kRound = kHB - kWave;
static_assert(kRound % kWave == 0, "a round is whole chunks");
constexpr uint32_t kChk = 0x80000000u;
// Appends go to kSub sub-lists per list, each with its own counter on its own
// 64-byte line: returning atomics on one address serialise (about 11 ns each
// on MI355X, one per appending wave), on kSub addresses they overlap.
// Counters: crossings, medium, heavy (kSub each), then the crossing total.
constexpr int kSub = 256;
constexpr int kCs = 16;
constexpr int kCntWords = (3 * kSub + 1) * kCs;
__host__ __device__ __forceinline__ int sub_word(int list, int s) { return (list * kSub + s) * kCs; }

struct ReplayArgs {
    const uint64_t *keys;
    const double *vals;   // addend of each sorted event
    int64_t E;
    int pbits;   // bits of 2 position + [subtraction] in a key
    int vbits;
    const int32_t *degree;
    double *r;
    double thr_scale;
    uint64_t *cross;     // kSub sub-lists of xcap crossing keys
    int64_t xcap;
    unsigned int *cnt;   // kCntWords
    uint32_t *med;       // kSub sub-lists of lcap first-event indices of medium destinations
    uint32_t *big;       // same, heavy destinations
    int64_t lcap;
};

__device__ __forceinline__ void emit_cross(const ReplayArgs &a, int sub, int v, uint32_t first) {
    const unsigned int slot = atomicAdd(a.cnt + sub_word(0, sub), 1u);
    a.cross[sub * a.xcap + slot] = ((uint64_t)first << a.vbits) | (uint64_t)(uint32_t)v;
}

// event f exists and belongs to destination v
__device__ __forceinline__ bool e_same(const ReplayArgs &a, int64_t f, int v) {
    return f < a.E && (int)(a.keys[f] >> a.pbits) == v;
}

// position word of a key: the position, with kChk unless the event is the
// subtraction of dx (not checked, ppr_solver.py:657)
__device__ __forceinline__ uint32_t pos_word(const ReplayArgs &a, uint64_t kf) {
    const uint32_t low = (uint32_t)(kf & ((1ull << a.pbits) - 1));
    return (low >> 1) | ((low & 1u) ? 0u : kChk);
}

// event f of destination v: its addend and position word; false past the segment
__device__ __forceinline__ bool load_event(const ReplayArgs &a, int v, int64_t f, double &ad, uint32_t &pc) {
    ad = -0.0;   // exact identity of FP64 addition, so padding never changes a sum
    pc = 0;
    if (f >= a.E) return false;
    const uint64_t kf = a.keys[f];
    if ((int)(kf >> a.pbits) != v) return false;
    ad = a.vals[f];
    pc = pos_word(a, kf);
    return true;
}

// adds the wave's first cnt addends to rv in lane order; lane j keeps the
// partial sum after its add, the value the reference tests
__device__ __forceinline__ void wave_serial(double ad, uint32_t pc, int lane, int cnt, double thr, double &rv,
                                            bool &marked, uint32_t &first) {
    const int lo = __double2loint(ad), hi = __double2hiint(ad);
    double mine = 0.0;
#pragma unroll 4
    for (int j = 0; j < cnt; j++) {
        rv += __hiloint2double(__builtin_amdgcn_readlane(hi, j), __builtin_amdgcn_readlane(lo, j));
        mine = (lane == j) ? rv : mine;
    }
    const unsigned long long hit = __ballot(lane < cnt && (pc & kChk) && thr <= fabs(mine));
    if (!marked && hit) {
        marked = true;
        first = (uint32_t)__builtin_amdgcn_readlane((int)pc, __ffsll((long long)hit) - 1) & ~kChk;
    }
}

// the same sum with the addends read from LDS at uniform addresses (one
// broadcast read per add instead of two readlanes) and only the running
// maximum of |partial sum| kept; any crossing in the chunk shows as
// max >= thr, and then the chunk is summed again lane by lane to find the
// first checked crossing, which happens at most once per destination
__device__ __forceinline__ void chunk_sum(const double *s_add, double ad, uint32_t pc, int lane, int cnt, double thr,
                                          double &rv, bool &marked, uint32_t &first) {
    // s_add[cnt .. round up to 8) holds -0.0, so whole groups of 8 are summed;
    // two register sets alternate so each group's reads are in flight while
    // the previous group is added
    const double r0 = rv;
    double m = 0.0;
    const int ng = (cnt + 7) >> 3;
    double ga[8], gb[8];
#pragma unroll
    for (int k = 0; k < 8; k++) ga[k] = s_add[k];
    for (int g = 0; g < ng; g += 2) {
        const int ob = (g + 1 < ng ? g + 1 : g) * 8;
#pragma unroll
        for (int k = 0; k < 8; k++) gb[k] = s_add[ob + k];
        __builtin_amdgcn_sched_barrier(0);   // keep the reads ahead of the adds
#pragma unroll
        for (int k = 0; k < 8; k++) {
            rv += ga[k];
            m = fmax(m, fabs(rv));
        }
        if (g + 1 >= ng) break;
        const int oa = (g + 2 < ng ? g + 2 : g) * 8;
#pragma unroll
        for (int k = 0; k < 8; k++) ga[k] = s_add[oa + k];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            rv += gb[k];
            m = fmax(m, fabs(rv));
        }
    }
    if (!marked && thr <= m) {
        rv = r0;
        wave_serial(ad, pc, lane, cnt, thr, rv, marked, first);
    }
}

// one thread per event; a wave covers 64 events and looks kLight events past
// its end (lanes 0..kLight-1 hold those as a second event), so every segment
// whose head is in the wave and that has at most kLight events lies in the
// window.  Such a segment is summed by its head lane from the window's
// addends through shuffles; longer ones go to the medium or heavy list, and
// a segment whose head is in an earlier wave belongs to that wave.
__global__ __launch_bounds__(kCT) void k_cheb_replay(ReplayArgs a) {
    const int64_t f = blockIdx.x * (int64_t)kCT + threadIdx.x;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t f0 = f - lane;
    const bool in = f < a.E;
    const uint64_t key = in ? a.keys[f] : 0;
    const int v = in ? (int)(key >> a.pbits) : -1;
    // the look-ahead events f0 + 64 .. f0 + 64 + kLight - 1
    const int64_t f2 = f0 + kWave + lane;
    const bool in2 = lane < kLight && f2 < a.E;
    const uint64_t key2 = in2 ? a.keys[f2] : 0;
    const int v2 = in2 ? (int)(key2 >> a.pbits) : -3;
    int vp = __shfl_up(v, 1);
    int vn = __shfl_down(v, 1);
    int v2n = __shfl_down(v2, 1);
    const int v2_first = __shfl(v2, 0);   // event f0 + 64 (-3 past the end)
    if (lane == 0) vp = (in && f > 0) ? (int)(a.keys[f - 1] >> a.pbits) : -2;
    if (lane == kWave - 1) vn = v2_first;
    if (lane == kLight - 1) v2n = (f2 + 1 < a.E) ? (int)(a.keys[f2 + 1] >> a.pbits) : -4;
    const bool head = in && vp != v;
    const unsigned long long hm = __ballot(head);
    const unsigned long long tm = __ballot(in && vn != v);
    const unsigned long long tm2 = __ballot(in2 && v2n != v2);   // bits 0 .. kLight-1
    // my segment: head lane h (or -1: it began in an earlier wave) and tail
    // index t in the window (64 + j for look-ahead event j; 128 if beyond)
    const unsigned long long below = (lane == kWave - 1) ? ~0ull : ((2ull << lane) - 1);
    const unsigned long long hb = hm & below, tb = tm & ~(below >> 1);
    const int h = hb ? 63 - __clzll((long long)hb) : -1;
    const int t = tb ? __ffsll((long long)tb) - 1 : (tm2 ? kWave + __ffsll((long long)tm2) - 1 : 128);
    const bool light = in && h >= 0 && t - h < kLight;
    // the segment that runs into the look-ahead, if light, takes those events
    const int t2 = tm2 ? __ffsll((long long)tm2) - 1 : -1;
    const bool cross_light = __ballot(light && t >= kWave) != 0;
    double ad = -0.0, ad2 = -0.0;
    uint32_t pc = 0, pc2 = 0;
    if (light) {
        ad = a.vals[f];
        pc = pos_word(a, key);
    }
    if (cross_light && lane <= t2) {
        ad2 = a.vals[f2];
        pc2 = pos_word(a, key2);
    }
    const bool lh = head && light;
    const int len = lh ? t - lane + 1 : 0;
    double rv = 0.0, thr = 0.0;
    if (lh) {
        rv = a.r[v];
        thr = a.thr_scale * (double)a.degree[v];
    }
    bool marked = false;
    uint32_t first = 0;
    const int lo = __double2loint(ad), hi = __double2hiint(ad);
    const int lo2 = __double2loint(ad2), hi2 = __double2hiint(ad2);
    for (int j = 0; __ballot(j < len); j++) {
        const int src = lane + j;                   // < 64 + kLight for the lanes that use it
        const int s1 = src & (kWave - 1);
        const bool second = src >= kWave;
        const int l1 = __shfl(lo, s1), h1 = __shfl(hi, s1), p1 = __shfl((int)pc, s1);
        const int l2 = __shfl(lo2, s1), h2 = __shfl(hi2, s1), p2 = __shfl((int)pc2, s1);
        const double aj = second ? __hiloint2double(h2, l2) : __hiloint2double(h1, l1);
        const uint32_t pj = (uint32_t)(second ? p2 : p1);
        if (j < len) {
            rv += aj;
            if ((pj & kChk) && !marked && thr <= fabs(rv)) {
                marked = true;
                first = pj & ~kChk;
            }
        }
    }
    if (lh) a.r[v] = rv;
    // one aggregated append per wave and list, into the wave's sub-lists (a
    // wave adds at most 64 entries to each)
    const int sub = (int)((f0 / kWave) & (kSub - 1));
    const long long cs = wave_append(lh && marked, a.cnt + sub_word(0, sub));
    if (cs >= 0) a.cross[sub * a.xcap + cs] = ((uint64_t)first << a.vbits) | (uint64_t)(uint32_t)v;
    // more than kLight events: at most 64 + kLight when the tail is in the
    // window, else probe the event kBig places on
    const bool longh = head && !light;
    const bool heavy = longh && t == 128 && e_same(a, f + kBig, v);
    const long long ms = wave_append(longh && !heavy, a.cnt + sub_word(1, sub));
    if (ms >= 0) a.med[sub * a.lcap + ms] = (uint32_t)f;
    const long long bs = wave_append(heavy, a.cnt + sub_word(2, sub));
    if (bs >= 0) a.big[sub * a.lcap + bs] = (uint32_t)f;
}

// one medium destination on one wave; s_w is the wave's own 64 LDS slots
__device__ void replay_wave(const ReplayArgs &a, uint32_t e, int lane, double *s_w, int sub) {
    const int v = (int)(a.keys[e] >> a.pbits);
    const double thr = a.thr_scale * (double)a.degree[v];
    double rv = a.r[v];
    bool marked = false;
    uint32_t first = 0;
    for (int64_t f0 = e;; f0 += kWave) {
        double ad;
        uint32_t pc;
        const unsigned long long ok = __ballot(load_event(a, v, f0 + lane, ad, pc));
        s_w[lane] = ad;
        __builtin_amdgcn_wave_barrier();
        chunk_sum(s_w, ad, pc, lane, __popcll(ok), thr, rv, marked, first);
        __builtin_amdgcn_wave_barrier();
        if (ok != ~0ull) break;
    }
    if (lane == 0) {
        a.r[v] = rv;
        if (marked) emit_cross(a, sub, v, first);
    }
}

// entry `item` of kSub sub-lists of capacity cap, given their inclusive
// prefix counts
__device__ __forceinline__ uint32_t sub_item(const uint32_t *lists, int64_t cap, const unsigned int *pre,
                                             unsigned int item) {
    int lo = 0, hi = kSub - 1;   // first sub-list whose prefix exceeds item
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] > item) hi = mid;
        else lo = mid + 1;
    }
    const unsigned int before = lo ? pre[lo - 1] : 0u;
    return lists[lo * cap + (item - before)];
}

// heavy destinations (one per block), then medium ones (one per wave).  Both
// schedules are static strides of uniform values: a work counter claimed by
// thread 0 at the top of the loop would put "if (t == 0)" at both ends of the
// body, and the compiler then threads the other threads' back edge straight
// to the barrier, leaving thread 0 outside a loop that never sees a new item.
__global__ __launch_bounds__(kHB) void k_cheb_replay_heavy(ReplayArgs a) {
    __shared__ double s_ad[2][kRound];
    __shared__ uint32_t s_pc[2][kRound];
    __shared__ double s_w[kHB / kWave][kWave];
    __shared__ unsigned int s_pre[2][kSub];   // inclusive prefix sums: medium, heavy sub-list counts
    const int t = threadIdx.x, lane = t & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(t / kWave);
    if (t < 2 * kSub) s_pre[t / kSub][t % kSub] = a.cnt[sub_word(1 + t / kSub, t % kSub)];
    __syncthreads();
    for (int d = 1; d < kSub; d <<= 1) {
        unsigned int y = 0;
        if (t < 2 * kSub && t % kSub >= d) y = s_pre[t / kSub][t % kSub - d];
        __syncthreads();
        if (t < 2 * kSub) s_pre[t / kSub][t % kSub] += y;
        __syncthreads();
    }
    // uniform by construction: the item loops below hold barriers
    const unsigned int nmed = (unsigned int)__builtin_amdgcn_readfirstlane((int)s_pre[0][kSub - 1]);
    const unsigned int nbig = (unsigned int)__builtin_amdgcn_readfirstlane((int)s_pre[1][kSub - 1]);
    for (unsigned int item = blockIdx.x; item < nbig; item += gridDim.x) {
        const uint32_t e = sub_item(a.big, a.lcap, s_pre[1], item);
        const int v = (int)(a.keys[e] >> a.pbits);
        const double thr = a.thr_scale * (double)a.degree[v];
        double rv = (wid == 0) ? a.r[v] : 0.0;
        bool marked = false;
        uint32_t first = 0;
        bool ok = false;
        if (wid > 0) {
            double ad;
            uint32_t pc;
            ok = load_event(a, v, (int64_t)e + (t - kWave), ad, pc);
            s_ad[0][t - kWave] = ad;   // -0.0 past the segment
            s_pc[0][t - kWave] = pc;
        }
        int c = __syncthreads_count(ok);   // valid events of a round are a prefix
        int64_t f0 = e;
        int b = 0;
        for (;;) {
            ok = false;
            if (wid > 0) {
                if (c == kRound) {
                    double ad;
                    uint32_t pc;
                    ok = load_event(a, v, f0 + kRound + (t - kWave), ad, pc);
                    s_ad[b ^ 1][t - kWave] = ad;
                    s_pc[b ^ 1][t - kWave] = pc;
                }
            } else {
                for (int q = 0; q < c; q += kWave) {
                    const int idx = q + lane;
                    double ad = -0.0;
                    uint32_t pc = 0;
                    if (idx < c) {
                        ad = s_ad[b][idx];
                        pc = s_pc[b][idx];
                    }
                    chunk_sum(&s_ad[b][q], ad, pc, lane, c - q < kWave ? c - q : kWave, thr, rv, marked, first);
                }
            }
            const int c2 = __syncthreads_count(ok);
            if (c < kRound) break;
            c = c2;
            f0 += kRound;
            b ^= 1;
        }
        if (t == 0) {
            a.r[v] = rv;
            if (marked) emit_cross(a, (int)(item & (kSub - 1)), v, first);
        }
    }
    const unsigned int waves = gridDim.x * (kHB / kWave);
    for (unsigned int item = blockIdx.x * (kHB / kWave) + wid; item < nmed; item += waves)
        replay_wave(a, sub_item(a.med, a.lcap, s_pre[0], item), lane, s_w[wid], (int)(item & (kSub - 1)));
}

// crossing sub-lists -> one array (sub-list order), and the total
__global__ __launch_bounds__(kSub) void k_cheb_xcompact(const uint64_t *xsub, int64_t xcap, const unsigned int *cnt,
                                                         uint64_t *out, unsigned int *total) {
    __shared__ unsigned int s_part[kSub / kWave];
    const int t = threadIdx.x, sub = blockIdx.x;
    const unsigned int c = cnt[sub_word(0, t)];
    unsigned int below = wave_sum(t < sub ? c : 0u);
    if (__lane_id() == 0) s_part[t / kWave] = below;
    __syncthreads();
    unsigned int off = 0;
    for (int w = 0; w < kSub / kWave; w++) off += s_part[w];
    const unsigned int mine = cnt[sub_word(0, sub)];
    for (unsigned int j = t; j < mine; j += kSub) out[off + j] = xsub[sub * xcap + j];
    if (sub == kSub - 1 && t == 0) *total = off + mine;
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

// sub-list capacities for E events: the head kernel's wave w appends at most
// 64 entries to each list's sub-list w mod kSub; the heavy kernel adds one
// crossing per medium or heavy destination (at most n of each) to sub-list
// item mod kSub
static void sub_caps(int64_t E, int64_t n, int64_t *lcap, int64_t *xcap) {
    const int64_t waves = (E + kWave - 1) / kWave;
    *lcap = (int64_t)kWave * ((waves + kSub - 1) / kSub);
    *xcap = *lcap + 2 * ((n + kSub - 1) / kSub);
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
    (void)hipFree(w->seg[0]);
    (void)hipFree(w->seg[1]);
    (void)hipFree(w->xsub);
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
    FLSG_HIP(hipMalloc((void **)&w->vals[0], sizeof(double) * cap));
    FLSG_HIP(hipMalloc((void **)&w->vals[1], sizeof(double) * cap));
    FLSG_HIP(hipMalloc((void **)&w->P, sizeof(int64_t) * (n + 2)));
    FLSG_HIP(hipMalloc((void **)&w->dplus, sizeof(int32_t) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&w->cross[0], sizeof(uint64_t) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&w->cross[1], sizeof(uint64_t) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&w->queue, sizeof(int32_t) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&w->cnt, sizeof(unsigned int) * kCntWords));
    int64_t lcap = 0, xcap = 0;
    sub_caps(cap, n, &lcap, &xcap);
    FLSG_HIP(hipMalloc((void **)&w->seg[0], sizeof(uint32_t) * kSub * lcap));
    FLSG_HIP(hipMalloc((void **)&w->seg[1], sizeof(uint32_t) * kSub * lcap));
    FLSG_HIP(hipMalloc((void **)&w->xsub, sizeof(uint64_t) * kSub * xcap));
    FLSG_HIP(hipMalloc((void **)&w->oper, sizeof(unsigned long long)));
    FLSG_HIP(hipMalloc((void **)&w->dx, sizeof(double) * (n + 1)));
    FLSG_HIP(hipMalloc((void **)&w->val, sizeof(double) * (n + 1)));
    // temp storage: the largest of the three primitives at full size
    size_t b1 = 0, b2 = 0, b3 = 0;
    FLSG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b1, (const int64_t *)nullptr, (int64_t *)nullptr, (int)(n + 1),
                                              st));
    FLSG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b2, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                                (const double *)nullptr, (double *)nullptr, (int)cap, 0, 64, st));
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
        if (E >= (int64_t)kChk) return fail(FLSG_ECAP, "more than 2^31 events in one iteration");
        const int pbits = bits_for((uint64_t)(2 * E));   // 2 position + [subtraction]
        if (pbits + vbits > 64) return fail(FLSG_ECAP, "event keys exceed 64 bits");
        hipLaunchKernelGGL(k_cheb_events, dim3(grid), dim3(kCT), 0, st, g->indptr, g->indices, g->degree, w->queue,
                           w->P, w->val, w->dx, (int)k, E, pbits, w->keys[0], w->vals[0]);
        tb = w->tmp_bytes;
        FLSG_HIP(hipcub::DeviceRadixSort::SortPairs(w->tmp, tb, w->keys[0], w->keys[1], w->vals[0], w->vals[1],
                                                    (int)E, 0, pbits + vbits, st));
        FLSG_HIP(hipMemsetAsync(w->cnt, 0, sizeof(unsigned int) * kCntWords, st));
        int64_t lcap = 0, xcap = 0;
        sub_caps(E, n, &lcap, &xcap);
        ReplayArgs ra{};
        ra.keys = w->keys[1];
        ra.vals = w->vals[1];
        ra.E = E;
        ra.pbits = pbits;
        ra.vbits = vbits;
        ra.degree = g->degree;
        ra.r = d_r;
        ra.thr_scale = thr_scale;
        ra.cross = w->xsub;
        ra.xcap = xcap;
        ra.cnt = w->cnt;
        ra.med = w->seg[0];
        ra.big = w->seg[1];
        ra.lcap = lcap;
        hipLaunchKernelGGL(k_cheb_replay, dim3((unsigned int)((E + kCT - 1) / kCT)), dim3(kCT), 0, st, ra);
        hipLaunchKernelGGL(k_cheb_replay_heavy, dim3(g->num_cus * 2), dim3(kHB), 0, st, ra);
        unsigned int *d_total = w->cnt + 3 * kSub * kCs;
        hipLaunchKernelGGL(k_cheb_xcompact, dim3(kSub), dim3(kSub), 0, st, w->xsub, xcap, w->cnt, w->cross[0], d_total);
        unsigned int c = 0;
        FLSG_HIP(hipMemcpyAsync(&c, d_total, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
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
