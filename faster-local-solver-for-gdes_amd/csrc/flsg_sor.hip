// flsg_sor.hip -- graph-coloured SOR push: the GPU counterpart of the
// reference's ppr_sor (ppr_solver.py:331-461) and katz_sor
// (katz_solver.py:252-337), local (frontier) and global (sweeps), any omega
// in (0, 2) (omega = 1 is the Gauss-Southwell / APPR push of the
// solve_a_graph_all drivers, ppr_solver.py:559-570, katz_solver.py:630-640),
// and of instantGNNSorPush (InstantGNN.py:74-105) for one feature column.
//
// The reference's SOR is sequential: each push reads the residual the
// previous pushes left.  Two neighbours pushing at once would each move their
// own residual without seeing the other's contribution, which is not an SOR
// step (and can diverge for omega > 1).  Nodes that share no edge can push at
// once, so the graph is coloured once per handle and every iteration visits
// the colours in order:
//   * colouring (Jones-Plassmann rounds, first-fit colours): priority of u is
//     (degree[u], mix32(u)) as one 64-bit key; a round selects each uncoloured
//     node with no uncoloured neighbour of higher key (an independent set) and
//     gives it the smallest colour no coloured neighbour holds.  Rows longer
//     than kColLong are scanned by a wavefront.  On the 2.39M-node power-law
//     graph: 100 colours (the ~100 mutually adjacent hubs need one each), 269
//     rounds, three colours hold 98% of the nodes;
//   * an iteration buckets its frontier by colour (LDS histograms, one scan
//     block, scatter), then launches one kernel per non-empty colour c.  That
//     kernel pushes the colour-c nodes that are still active -- res = omega
//     r_u, (1-alpha) res / d_u atomically added to every neighbour, none of
//     which has colour c, so r_u of colour c is stable during the launch and
//     hub pieces recompute the same push value -- and applies the previous
//     colour's x_u += res, r_u -= res (atomically: the colour-c pushes add to
//     those residuals in the same launch).  The pushes are non-returning
//     atomics with kSU edges in flight per thread; nothing waits on them.
//   * after the last colour a check pass walks the pushed nodes and their
//     rows again (gathers, no atomics) and lists each one that is active
//     (eps*alpha*d_v <= |r_v|, a mark dedups) as the next frontier.  That is
//     the active set after the sweep whatever order the adds of one colour
//     took (residuals change sign for omega > 1, so a test after every add
//     would depend on it); only pushed nodes and their neighbours changed,
//     so an empty frontier certifies |r_v| < eps*alpha*d_v everywhere.
// One iteration of the local variant is a Gauss-Seidel sweep over the frontier
// in colour order (the reference's FIFO order differs), the global variant a
// full multicolour SOR sweep; both end on the reference's certificate.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "flsg_internal.h"

namespace flsg {

constexpr int kSNT = 256;        // threads per block
constexpr int kSHub = 128;       // entries with more edges are pushed in pieces, so one
                                 // 256-entry chunk holds at most 32K edges
constexpr int kSPiece = 1024;    // edges per piece (one round of kSU edges per thread)
constexpr int kColLong = 64;     // colouring: longer rows are scanned by a wavefront
constexpr int kLdsBins = 1024;   // bucket kernels: colours below this use LDS histograms
constexpr int kColRoundsPerCheck = 8;

struct SorCtr {
    unsigned int nf[2];          // frontier size by parity
    unsigned int left;           // colouring: live nodes still uncoloured
    unsigned int ncol;           // colouring: colours used (max colour + 1)
    unsigned long long rounds;   // colouring: rounds that coloured something
    unsigned long long oper;     // local: sum of degree over pushed (active) nodes
    unsigned long long nodes;    // local: pushed nodes
    unsigned int act;            // global: live nodes at or above the threshold
    unsigned int selfloop;       // a row holds its own node
    unsigned int nlong;          // colouring: rows longer than kColLong
    unsigned int pad;
};

struct SorWs {
    int32_t *colour = nullptr;   // n
    int32_t *sel = nullptr;      // n: selected in the current colouring round
    int32_t *longs = nullptr;    // n: live rows longer than kColLong
    int32_t *loops = nullptr;    // n: self-loops in each row
    int64_t nlong = 0;
    int ncol = 0;
    int64_t rounds = 0;
    int valid = 0;
    int selfloop = 0;
    // bucketing by colour
    unsigned int *cnts = nullptr;    // [2 * ncol_cap]: entries, pieces per colour
    unsigned int *offs = nullptr;    // [2 * (ncol_cap + 1)]: their exclusive offsets (+ totals)
    unsigned int *curs = nullptr;    // [2 * ncol_cap]: scatter cursors
    unsigned int *h_offs = nullptr;  // pinned mirror of offs
    int64_t ncol_cap = 0;
    // local frontier
    int32_t *front[2] = {nullptr, nullptr};
    int32_t *bucket = nullptr;       // n entries by colour
    double *res = nullptr;           // n: omega r_u of each entry (0 = not pushed)
    int32_t *mark = nullptr;         // n: listed in the next frontier
    int64_t *pieces = nullptr;       // (entry << 32) | piece
    int64_t pieces_cap = 0;
    // every live node by colour (global sweeps)
    int32_t *ball = nullptr;
    int64_t *pall = nullptr;
    std::vector<unsigned int> offs_all;   // host copy: entries then pieces offsets
    int all_valid = 0;
    SorCtr *ctr = nullptr;
    SorCtr *h_ctr = nullptr;         // pinned
};

void sor_free(flsg_graph g) {
    SorWs *w = (SorWs *)g->sor;
    if (!w) return;
    (void)hipFree(w->colour);
    (void)hipFree(w->sel);
    (void)hipFree(w->longs);
    (void)hipFree(w->loops);
    (void)hipFree(w->cnts);
    (void)hipFree(w->offs);
    (void)hipFree(w->curs);
    (void)hipHostFree(w->h_offs);
    (void)hipFree(w->front[0]);
    (void)hipFree(w->front[1]);
    (void)hipFree(w->bucket);
    (void)hipFree(w->res);
    (void)hipFree(w->mark);
    (void)hipFree(w->pieces);
    (void)hipFree(w->ball);
    (void)hipFree(w->pall);
    (void)hipFree(w->ctr);
    (void)hipHostFree(w->h_ctr);
    delete w;
    g->sor = nullptr;
}

void sor_invalidate(flsg_graph g) {
    SorWs *w = (SorWs *)g->sor;
    if (w) {
        w->valid = 0;
        w->all_valid = 0;
    }
}

// ------------------------------------------------------------------ colouring
__device__ __forceinline__ uint32_t mix32(uint32_t h) {   // MurmurHash3 fmix32, a bijection
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

__device__ __forceinline__ unsigned long long colour_key(int d, int u) {
    return ((unsigned long long)(uint32_t)d << 32) | (unsigned long long)mix32((uint32_t)u);
}

// colour 0 for isolated nodes, -1 (uncoloured) for the rest; counts live
// nodes and lists the long rows; counts each row's self-loops
__global__ __launch_bounds__(kSNT) void k_col_init(const int32_t *indptr, const int32_t *indices,
                                                   const int32_t *degree, int64_t n, int32_t *colour,
                                                   int32_t *longs, int32_t *loops, SorCtr *c) {
    unsigned int live = 0;
    bool loop = false;
    const int64_t stride = (int64_t)gridDim.x * kSNT;
    for (int64_t base = (int64_t)blockIdx.x * kSNT; base < n; base += stride) {
        const int64_t u = base + threadIdx.x;
        bool is_long = false;
        if (u < n) {
            const int d = degree[u];
            colour[u] = d == 0 ? 0 : -1;
            live += d != 0;
            is_long = d > kColLong;
            const int st = indptr[u];
            int nl = 0;
#pragma unroll 8
            for (int j = 0; j < d; j++) nl += indices[st + j] == (int)u;
            loops[u] = nl;
            loop |= nl > 0;
        }
        const long long slot = wave_append(is_long, &c->nlong);
        if (slot >= 0) longs[slot] = (int32_t)u;
    }
    live = wave_sum(live);
    const bool any_loop = __ballot(loop) != 0ull;
    if (__lane_id() == 0) {
        if (live) atomicAdd(&c->left, live);
        if (any_loop) c->selfloop = 1u;
    }
}

__global__ void k_col_round(SorCtr *c) {
    if (c->left > 0) c->rounds++;
}

// selection of the short rows: no uncoloured neighbour with a higher key
__global__ __launch_bounds__(kSNT) void k_col_select(const int32_t *indptr, const int32_t *indices,
                                                     const int32_t *degree, int64_t n, const int32_t *colour,
                                                     int32_t *sel) {
    for (int64_t u = (int64_t)blockIdx.x * kSNT + threadIdx.x; u < n; u += (int64_t)gridDim.x * kSNT) {
        const int d = degree[u];
        if (d > kColLong) continue;   // k_col_select_long
        int s = 0;
        if (colour[u] < 0) {
            const unsigned long long ku = colour_key(d, (int)u);
            const int st = indptr[u];
            s = 1;
            for (int j = 0; j < d; j++) {
                const int v = indices[st + j];
                if (v != (int)u && colour[v] < 0 && colour_key(degree[v], v) > ku) {
                    s = 0;
                    break;
                }
            }
        }
        sel[u] = s;
    }
}

// selection of the long rows, one wavefront per row, 256 entries per step
__global__ __launch_bounds__(kSNT) void k_col_select_long(const int32_t *indptr, const int32_t *indices,
                                                          const int32_t *degree, const int32_t *longs,
                                                          const SorCtr *c, const int32_t *colour, int32_t *sel) {
    const int lane = __lane_id();
    const int64_t nlong = c->nlong;
    const int64_t nw = (int64_t)gridDim.x * (kSNT / kWave);
    for (int64_t i = (int64_t)blockIdx.x * (kSNT / kWave) + threadIdx.x / kWave; i < nlong; i += nw) {
        const int u = longs[i];
        int s = 0;
        if (colour[u] < 0) {
            const int d = degree[u];
            const int st = indptr[u];
            const unsigned long long ku = colour_key(d, u);
            bool beaten = false;
            for (int j0 = 0; j0 < d && !beaten; j0 += 4 * kWave) {
                bool b = false;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int j = j0 + k * kWave + lane;
                    if (j < d) {
                        const int v = indices[st + j];
                        if (v != u && colour[v] < 0 && colour_key(degree[v], v) > ku) b = true;
                    }
                }
                beaten = __ballot(b) != 0ull;
            }
            s = beaten ? 0 : 1;
        }
        if (lane == 0) sel[u] = s;
    }
}

// first-fit colour of the selected short rows.  No neighbour of a selected
// node is selected, so the colours read here were all set in earlier rounds.
__global__ __launch_bounds__(kSNT) void k_col_assign(const int32_t *indptr, const int32_t *indices,
                                                     const int32_t *degree, int64_t n, int32_t *colour,
                                                     const int32_t *sel, SorCtr *c) {
    unsigned int done = 0, mx = 0;
    for (int64_t u = (int64_t)blockIdx.x * kSNT + threadIdx.x; u < n; u += (int64_t)gridDim.x * kSNT) {
        const int d = degree[u];
        if (d > kColLong || !sel[u]) continue;
        const int st = indptr[u];
        int col = -1;
        for (int base = 0; col < 0; base += 64) {
            unsigned long long used = 0;
            for (int j = 0; j < d; j++) {
                const int v = indices[st + j];
                if (v == (int)u) continue;
                const int cv = colour[v];
                if (cv >= base && cv < base + 64) used |= 1ull << (cv - base);
            }
            if (~used) col = base + __builtin_ctzll(~used);
        }
        colour[u] = col;
        done++;
        mx = max(mx, (unsigned int)(col + 1));
    }
    done = wave_sum(done);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, (unsigned int)__shfl_xor((int)mx, off));
    if (__lane_id() == 0 && done) {
        atomicSub(&c->left, done);
        atomicMax(&c->ncol, mx);
    }
}

__global__ __launch_bounds__(kSNT) void k_col_assign_long(const int32_t *indptr, const int32_t *indices,
                                                          const int32_t *degree, const int32_t *longs,
                                                          int32_t *colour, const int32_t *sel, SorCtr *c) {
    const int lane = __lane_id();
    const int64_t nlong = c->nlong;
    const int64_t nw = (int64_t)gridDim.x * (kSNT / kWave);
    for (int64_t i = (int64_t)blockIdx.x * (kSNT / kWave) + threadIdx.x / kWave; i < nlong; i += nw) {
        const int u = longs[i];
        if (!sel[u]) continue;
        const int d = degree[u];
        const int st = indptr[u];
        int col = -1;
        for (int base = 0; col < 0; base += 64) {
            unsigned long long used = 0;
            for (int j = lane; j < d; j += kWave) {
                const int v = indices[st + j];
                if (v == u) continue;
                const int cv = colour[v];
                if (cv >= base && cv < base + 64) used |= 1ull << (cv - base);
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) used |= (unsigned long long)__shfl_xor((long long)used, off);
            if (~used) col = base + __builtin_ctzll(~used);
        }
        if (lane == 0) {
            colour[u] = col;
            atomicSub(&c->left, 1u);
            atomicMax(&c->ncol, (unsigned int)(col + 1));
        }
    }
}

// ------------------------------------------------------------------ bucketing by colour
// Entries: F[0 .. *nfp) when F is given, else every live node (u = index).
__device__ __forceinline__ bool entry_at(const int32_t *F, unsigned int nF, const int32_t *degree, int64_t i,
                                         int &u) {
    if (i >= (int64_t)nF) return false;
    u = F ? F[i] : (int)i;
    return F ? true : degree[u] > 0;
}

__device__ __forceinline__ unsigned int n_pieces(int d) {
    return d > kSHub ? (unsigned int)((d + kSPiece - 1) / kSPiece) : 0u;
}

// counts per colour (entries, pieces); clears the entries' marks and zeroes
// *zero_me (the next frontier's counter) when given
__global__ __launch_bounds__(kSNT) void k_bkt_count(const int32_t *F, const unsigned int *nfp, unsigned int nF_all,
                                                    const int32_t *colour, const int32_t *degree,
                                                    unsigned int *cnts, int ncol, int32_t *mark,
                                                    unsigned int *zero_me) {
    __shared__ unsigned int h[kLdsBins];
    const unsigned int nF = nfp ? *nfp : nF_all;
    if (blockIdx.x == 0 && threadIdx.x == 0 && zero_me) *zero_me = 0;
    for (int b = threadIdx.x; b < kLdsBins; b += kSNT) h[b] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * kSNT + threadIdx.x; i < (int64_t)nF; i += (int64_t)gridDim.x * kSNT) {
        int u;
        if (!entry_at(F, nF, degree, i, u)) continue;
        if (mark) mark[u] = 0;
        const int c = colour[u];
        if (c < kLdsBins) atomicAdd(&h[c], 1u);
        else atomicAdd(&cnts[c], 1u);
        const unsigned int np = n_pieces(degree[u]);
        if (np) atomicAdd(&cnts[ncol + c], np);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kLdsBins && b < ncol; b += kSNT)
        if (h[b]) atomicAdd(&cnts[b], h[b]);
}

// exclusive offsets of both count arrays (one block); sets the scatter
// cursors and zeroes the counts for the next use
__global__ __launch_bounds__(1024) void k_bkt_scan(unsigned int *cnts, unsigned int *offs, unsigned int *curs,
                                                   int ncol) {
    __shared__ int sh[1024 / kWave];
    for (int part = 0; part < 2; part++) {
        unsigned int *cn = cnts + part * ncol;
        unsigned int *of = offs + part * (ncol + 1);
        unsigned int *cu = curs + part * ncol;
        unsigned int carry = 0;
        for (int b0 = 0; b0 < ncol; b0 += 1024) {
            const int c = b0 + threadIdx.x;
            const unsigned int v = c < ncol ? cn[c] : 0u;
            int tot;
            const int incl = block_incl_scan<1024>((int)v, sh, &tot);
            if (c < ncol) {
                const unsigned int ex = carry + (unsigned int)incl - v;
                of[c] = ex;
                cu[c] = ex;
                cn[c] = 0;
            }
            carry += (unsigned int)tot;
        }
        if (threadIdx.x == 0) of[ncol] = carry;
        __syncthreads();
    }
}

// scatter entries into their colour's range (LDS positions for the low
// colours, one global atomic per block and colour) and list hub pieces
__global__ __launch_bounds__(kSNT) void k_bkt_scatter(const int32_t *F, const unsigned int *nfp,
                                                      unsigned int nF_all, const int32_t *colour,
                                                      const int32_t *degree, unsigned int *curs, int ncol,
                                                      int32_t *B, int64_t *pieces) {
    __shared__ unsigned int h[kLdsBins];
    __shared__ unsigned int hb[kLdsBins];
    const unsigned int nF = nfp ? *nfp : nF_all;
    for (int b = threadIdx.x; b < kLdsBins; b += kSNT) h[b] = 0;
    __syncthreads();
    for (int64_t base = (int64_t)blockIdx.x * kSNT; base < (int64_t)nF; base += (int64_t)gridDim.x * kSNT) {
        const int64_t i = base + threadIdx.x;
        int u = 0, c = -1;
        unsigned int lp = 0;
        if (entry_at(F, nF, degree, i, u)) {
            c = colour[u];
            if (c < kLdsBins) lp = atomicAdd(&h[c], 1u);
        }
        __syncthreads();
        for (int b = threadIdx.x; b < kLdsBins && b < ncol; b += kSNT)
            if (h[b]) hb[b] = atomicAdd(&curs[b], h[b]);
        __syncthreads();
        if (c >= 0) {
            const unsigned int pos = c < kLdsBins ? hb[c] + lp : atomicAdd(&curs[c], 1u);
            B[pos] = u;
            const unsigned int np = n_pieces(degree[u]);
            if (np) {
                const unsigned int pb = atomicAdd(&curs[ncol + c], np);
                for (unsigned int k = 0; k < np; k++) pieces[pb + k] = ((int64_t)pos << 32) | (int64_t)k;
            }
        }
        __syncthreads();
        for (int b = threadIdx.x; b < kLdsBins; b += kSNT) h[b] = 0;
        __syncthreads();
    }
}

// ------------------------------------------------------------------ SOR push
struct SorArgs {
    const int32_t *indptr;
    const int32_t *indices;
    const int32_t *degree;
    double *x;
    double *r;
    const int32_t *B;        // entries by colour
    const int64_t *pieces;   // hub pieces by colour
    double *res;             // per entry: omega r_u pushed (0 = not pushed)
    int32_t *mark;           // local: listed in the next frontier
    int32_t *next;           // local: next frontier
    unsigned int *nnext;     // local: its counter
    SorCtr *ctr;
    const double *thr;       // InstantGNN: eps_vec (eps d^(1-beta)); else null (threshold ea * d)
    const double *dp1;       // InstantGNN: d^(1-beta)
    const double *dpb;       // InstantGNN: d^beta
    const int32_t *loops;    // self-loops per row
    double coef;             // PPR / InstantGNN: 1 - alpha (push = coef*res/d_u, /d_u^(1-beta)); Katz: beta
    double xcoef;            // x_u += xcoef * res: InstantGNN alpha, else 1
    double omega;
    double ea;               // threshold per unit degree: PPR eps*alpha, Katz eps
    int local;
    int katz;
    int ignn;
};

constexpr int kSU = 4;   // edges in flight per thread

enum { kPush = 0, kCheck = 1 };

// push value of u (InstantGNN: before the 1/d_v^beta of each edge)
__device__ __forceinline__ double sor_push_value(const SorArgs &a, double rs, int u, int d) {
    if (a.ignn) return a.coef * rs / a.dp1[u];  // InstantGNN.py:90
    return a.katz ? a.coef * rs                  // katz_solver.py:302
                  : a.coef * rs / (double)d;     // ppr_solver.py:407
}

// u is pushed (local: at or above its threshold; global: nonzero residual).
// InstantGNN: |r_u| >= eps_vec[u] (InstantGNN.py:82) and r_u != 0, so an
// isolated node (threshold 0) is taken once; other kinds need a live row.
__device__ __forceinline__ bool sor_active(const SorArgs &a, int u, int d, double ru) {
    if (a.ignn) return a.thr[u] <= fabs(ru) && ru != 0.0;
    return d > 0 && (a.local ? (a.ea * (double)d <= fabs(ru)) : (ru != 0.0));
}

// InstantGNN: an isolated node or one whose only edge is a self-loop takes
// p_u = r_u, r_u = 0 instead of pushing (InstantGNN.py:84-87)
__device__ __forceinline__ bool sor_absorbs(const SorArgs &a, int u, int d) {
    return a.ignn && (d == 0 || (d == 1 && a.loops[u] == 1));
}

// v active at the end of the sweep, listed once (mark)
__device__ __forceinline__ bool active_claim(const SorArgs &a, int v) {
    const int d = a.degree[v];
    const double rv = a.r[v];
    const bool act = a.ignn ? (a.thr[v] <= fabs(rv) && rv != 0.0) : (d > 0 && a.ea * (double)d <= fabs(rv));
    return act && a.mark[v] == 0 && atomicExch(&a.mark[v], 1) == 0;
}

// Work over bucketed entries [beg, end) (chunks of kSNT, their edges
// flattened by an LDS prefix sum) and hub pieces [pbeg, pend):
//   kPush  (one colour): finalize entries [fbeg, fend) of the previous colour
//          (x_u += res, r_u -= res), then push the entries still active --
//          res = omega r_u, its push value atomically added to every neighbour
//          (non-returning atomics, kSU edges in flight per thread);
//   kCheck (local, after the last colour, over every entry of the iteration):
//          each pushed node and each neighbour of one -- every node whose
//          residual changed -- joins the next frontier if active.
template <int MODE>
__global__ __launch_bounds__(kSNT) void k_sor_items(SorArgs a, unsigned int beg, unsigned int end,
                                                    unsigned int pbeg, unsigned int pend, unsigned int fbeg,
                                                    unsigned int fend) {
    __shared__ int s_incl[kSNT];
    __shared__ int s_start[kSNT];
    __shared__ int s_node[kSNT];
    __shared__ double s_val[kSNT];
    __shared__ int s_sh[kSNT / kWave];
    const int tid = threadIdx.x;
    if (MODE == kPush) {
        for (unsigned int i = fbeg + blockIdx.x * kSNT + tid; i < fend; i += gridDim.x * kSNT) {
            const double rs = a.res[i];
            if (rs != 0.0) {
                const int u = a.B[i];
                if (a.ignn) {
                    const int d = a.degree[u];
                    if (sor_absorbs(a, u, d)) {              // InstantGNN.py:85-86: nothing else writes r_u
                        a.x[u] = rs;
                        a.r[u] = 0.0;
                    } else {                                 // InstantGNN.py:89, 91 and the self-loops of :92-93
                        a.x[u] += a.xcoef * rs;
                        const double push_u = sor_push_value(a, rs, u, d);
                        (void)atomic_add_dev(&a.r[u], (double)a.loops[u] * (push_u / a.dpb[u]) - rs);
                    }
                } else {
                    a.x[u] += rs;                            // ppr_solver.py:405
                    (void)atomic_add_dev(&a.r[u], -rs);      // ppr_solver.py:406
                }
            }
        }
    }
    const unsigned int nchunks = (end - beg + kSNT - 1) / kSNT;
    const unsigned int nitems = nchunks + (pend - pbeg);
    unsigned long long oper = 0, nodes = 0;
    for (unsigned int item = blockIdx.x; item < nitems; item += gridDim.x) {
        int st = 0, e_beg = 0, e_end = 0, pu = 0;
        double val = 0.0;
        const bool chunk = item < nchunks;
        if (chunk) {
            const unsigned int i = beg + item * kSNT + tid;
            int cnt = 0;
            bool self = false;
            int u = 0;
            if (i < end) {
                u = a.B[i];
                const int d = a.degree[u];
                bool act;
                if (MODE == kPush) {
                    const double ru = a.r[u];
                    act = sor_active(a, u, d, ru);
                    double rs = 0.0;
                    if (act && sor_absorbs(a, u, d)) {
                        rs = ru;                             // applied by the finalize; no edges
                        act = false;
                    } else if (act) {
                        rs = a.omega * ru;                   // ppr_solver.py:404, katz_solver.py:299
                        val = sor_push_value(a, rs, u, d);
                        if (a.local) {
                            oper += (unsigned long long)d;   // ppr_solver.py:401
                            nodes++;
                        }
                    }
                    a.res[i] = rs;
                } else {
                    act = a.res[i] != 0.0;
                    self = act && active_claim(a, u);
                }
                if (act) {
                    st = a.indptr[u];
                    cnt = d > kSHub ? 0 : d;
                }
            }
            if (MODE == kCheck) {
                const long long slot = wave_append(self, a.nnext);
                if (slot >= 0) a.next[slot] = u;
            }
            int tot;
            const int incl = block_incl_scan<kSNT>(cnt, s_sh, &tot);
            s_incl[tid] = incl;
            s_start[tid] = st;
            s_node[tid] = u;
            s_val[tid] = val;
            __syncthreads();
            e_end = tot;
        } else {
            const int64_t pc = a.pieces[pbeg + (item - nchunks)];
            const unsigned int i = (unsigned int)(pc >> 32);
            const int k = (int)(pc & 0xffffffffLL);
            pu = a.B[i];
            const int d = a.degree[pu];
            if (MODE == kPush) {
                const double ru = a.r[pu];
                if (!sor_active(a, pu, d, ru)) continue;   // uniform: every thread read the same entry
                val = sor_push_value(a, a.omega * ru, pu, d);   // bit-equal to the chunk's value
            } else if (a.res[i] == 0.0) {
                continue;
            }
            st = a.indptr[pu];
            e_beg = k * kSPiece;
            e_end = min(d, e_beg + kSPiece);
        }
        for (int e0 = e_beg; e0 < e_end; e0 += kSNT * kSU) {
            int v[kSU];
            double vv[kSU];
            bool ok[kSU];
#pragma unroll
            for (int q = 0; q < kSU; q++) {
                const int e = e0 + q * kSNT + tid;
                ok[q] = e < e_end;
                v[q] = 0;
                vv[q] = val;
                if (ok[q]) {
                    int src = pu;
                    if (chunk) {
                        const int k = upper_bound_lds(s_incl, 0, kSNT, e);
                        const int o = e - (k ? s_incl[k - 1] : 0);
                        v[q] = a.indices[s_start[k] + o];
                        vv[q] = s_val[k];
                        src = s_node[k];
                    } else {
                        v[q] = a.indices[st + e];
                    }
                    if (a.ignn) {
                        // self-loops go through the finalize (this colour's
                        // residuals are read while it pushes); InstantGNN.py:93
                        ok[q] = v[q] != src;
                        if (MODE == kPush && ok[q]) vv[q] = vv[q] / a.dpb[v[q]];
                    }
                }
            }
            if (MODE == kPush) {
#pragma unroll
                for (int q = 0; q < kSU; q++)
                    if (ok[q]) (void)atomic_add_dev(&a.r[v[q]], vv[q]);   // ppr_solver.py:409
            } else {
                bool enq[kSU];
#pragma unroll
                for (int q = 0; q < kSU; q++) enq[q] = ok[q] && active_claim(a, v[q]);
#pragma unroll
                for (int q = 0; q < kSU; q++) {
                    const long long slot = wave_append(enq[q], a.nnext);
                    if (slot >= 0) a.next[slot] = v[q];
                }
            }
        }
        if (chunk) __syncthreads();   // s_incl / s_start / s_val are rewritten by the next chunk
    }
    if (MODE == kPush && a.local) {
        oper = wave_sum(oper);
        nodes = wave_sum(nodes);
        if (__lane_id() == 0 && nodes) {
            atomicAdd(&a.ctr->oper, oper);
            atomicAdd(&a.ctr->nodes, nodes);
        }
    }
}

__global__ void k_sor_init(double *r, int32_t *F, SorCtr *c, int64_t s, double r0, int local) {
    r[s] = r0;   // ppr_solver.py:339 (alpha), katz_solver.py:258 (1)
    if (local) {
        F[0] = (int32_t)s;
        c->nf[0] = 1;
    }
}

__global__ void k_sor_add(double *x, int64_t s, double v) { x[s] += v; }

// starting frontier of an InstantGNN column: the listed nodes (distinct), or
// every node with eps_vec <= |r| and r != 0 (InstantGNN.py:118-123)
__global__ __launch_bounds__(kSNT) void k_sor_seed(const int32_t *nodes, int64_t n_nodes, int64_t n,
                                                   const double *thr, const double *r, int32_t *F,
                                                   unsigned int *nf) {
    const int64_t cnt = nodes ? n_nodes : n;
    for (int64_t base = (int64_t)blockIdx.x * kSNT; base < cnt; base += (int64_t)gridDim.x * kSNT) {
        const int64_t i = base + threadIdx.x;
        bool act = false;
        int v = 0;
        if (i < cnt) {
            v = nodes ? nodes[i] : (int)i;
            act = nodes ? true : (thr[v] <= fabs(r[v]) && r[v] != 0.0);
        }
        const long long slot = wave_append(act, nf);
        if (slot >= 0) F[slot] = v;
    }
}

// live nodes at or above the threshold (the global loop's stopping test,
// ppr_solver.py:449-452; isolated nodes are never counted)
__global__ __launch_bounds__(kSNT) void k_sor_active(const double *r, const int32_t *degree, int64_t n, double ea,
                                                     unsigned int *act) {
    unsigned int k = 0;
    for (int64_t u = (int64_t)blockIdx.x * kSNT + threadIdx.x; u < n; u += (int64_t)gridDim.x * kSNT) {
        const int d = degree[u];
        k += (d > 0 && ea * (double)d <= fabs(r[u])) ? 1u : 0u;
    }
    k = wave_sum(k);
    if (__lane_id() == 0 && k) atomicAdd(act, k);
}

// ------------------------------------------------------------------ host side
static int sor_ws(flsg_graph g, SorWs **out) {
    SorWs *w = (SorWs *)g->sor;
    if (!w) {
        w = new SorWs();
        g->sor = w;
        const size_t n = (size_t)g->n;
        FLSG_HIP(hipMalloc((void **)&w->colour, sizeof(int32_t) * n));
        FLSG_HIP(hipMalloc((void **)&w->sel, sizeof(int32_t) * n));
        FLSG_HIP(hipMalloc((void **)&w->longs, sizeof(int32_t) * n));
        FLSG_HIP(hipMalloc((void **)&w->loops, sizeof(int32_t) * n));
        FLSG_HIP(hipMalloc((void **)&w->front[0], sizeof(int32_t) * n));
        FLSG_HIP(hipMalloc((void **)&w->front[1], sizeof(int32_t) * n));
        FLSG_HIP(hipMalloc((void **)&w->bucket, sizeof(int32_t) * n));
        FLSG_HIP(hipMalloc((void **)&w->res, sizeof(double) * n));
        FLSG_HIP(hipMalloc((void **)&w->mark, sizeof(int32_t) * n));
        FLSG_HIP(hipMalloc((void **)&w->ball, sizeof(int32_t) * n));
        FLSG_HIP(hipMalloc((void **)&w->ctr, sizeof(SorCtr)));
        FLSG_HIP(hipHostMalloc((void **)&w->h_ctr, sizeof(SorCtr), hipHostMallocDefault));
    }
    *out = w;
    return FLSG_OK;
}

static int sor_sync_ctr(SorWs *w, hipStream_t st) {
    FLSG_HIP(hipMemcpyAsync(w->h_ctr, w->ctr, sizeof(SorCtr), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    return FLSG_OK;
}

static int grid_of(flsg_graph g, int64_t items, int per_cu) {
    const int64_t cap = (int64_t)g->num_cus * per_cu;
    return (int)std::max<int64_t>(1, std::min<int64_t>(items, cap));
}

// Colours the graph (once per handle and degree state) and sizes the bucket
// buffers for the colour count.
static int ensure_colouring(flsg_graph g, SorWs *w, hipStream_t st) {
    if (w->valid) return FLSG_OK;
    const int64_t n = g->n;
    const int gridn = grid_of(g, (n + kSNT - 1) / kSNT, 8);
    FLSG_HIP(hipMemsetAsync(w->ctr, 0, sizeof(SorCtr), st));
    hipLaunchKernelGGL(k_col_init, dim3(gridn), dim3(kSNT), 0, st, g->indptr, g->indices, g->degree, n,
                       w->colour, w->longs, w->loops, w->ctr);
    int rc = sor_sync_ctr(w, st);
    if (rc) return rc;
    w->selfloop = w->h_ctr->selfloop != 0;
    w->nlong = w->h_ctr->nlong;
    const int gridl = grid_of(g, (w->nlong + kSNT / kWave - 1) / (kSNT / kWave), 8);
    int64_t launched = 0;
    while (w->h_ctr->left > 0) {
        if (launched > n + 1) return fail(FLSG_EHIP, "colouring made no progress");
        for (int k = 0; k < kColRoundsPerCheck; k++) {
            hipLaunchKernelGGL(k_col_round, dim3(1), dim3(1), 0, st, w->ctr);
            hipLaunchKernelGGL(k_col_select, dim3(gridn), dim3(kSNT), 0, st, g->indptr, g->indices, g->degree, n,
                               w->colour, w->sel);
            hipLaunchKernelGGL(k_col_select_long, dim3(gridl), dim3(kSNT), 0, st, g->indptr, g->indices, g->degree,
                               w->longs, w->ctr, w->colour, w->sel);
            hipLaunchKernelGGL(k_col_assign, dim3(gridn), dim3(kSNT), 0, st, g->indptr, g->indices, g->degree, n,
                               w->colour, w->sel, w->ctr);
            hipLaunchKernelGGL(k_col_assign_long, dim3(gridl), dim3(kSNT), 0, st, g->indptr, g->indices,
                               g->degree, w->longs, w->colour, w->sel, w->ctr);
        }
        launched += kColRoundsPerCheck;
        FLSG_HIP(hipGetLastError());
        rc = sor_sync_ctr(w, st);
        if (rc) return rc;
    }
    w->ncol = (int)std::max(1u, w->h_ctr->ncol);
    w->rounds = (int64_t)w->h_ctr->rounds;
    if (w->ncol > w->ncol_cap) {
        (void)hipFree(w->cnts);
        (void)hipFree(w->offs);
        (void)hipFree(w->curs);
        (void)hipHostFree(w->h_offs);
        w->cnts = w->offs = w->curs = w->h_offs = nullptr;
        w->ncol_cap = 0;
        FLSG_HIP(hipMalloc((void **)&w->cnts, sizeof(unsigned int) * 2 * w->ncol));
        FLSG_HIP(hipMalloc((void **)&w->offs, sizeof(unsigned int) * 2 * (w->ncol + 1)));
        FLSG_HIP(hipMalloc((void **)&w->curs, sizeof(unsigned int) * 2 * w->ncol));
        FLSG_HIP(hipHostMalloc((void **)&w->h_offs, sizeof(unsigned int) * 2 * (w->ncol + 1), hipHostMallocDefault));
        w->ncol_cap = w->ncol;
    }
    FLSG_HIP(hipMemsetAsync(w->cnts, 0, sizeof(unsigned int) * 2 * w->ncol, st));
    w->valid = 1;
    w->all_valid = 0;
    return FLSG_OK;
}

// bucket entries (F with its device count, or every live node) by colour;
// returns the offsets on the host (entries in [0, ncol], pieces after)
static int bucket(flsg_graph g, SorWs *w, const int32_t *F, const unsigned int *nfp, int32_t *mark,
                  unsigned int *zero_me, int32_t *B, int64_t *pieces, hipStream_t st) {
    const int ncol = w->ncol;
    const unsigned int nall = (unsigned int)g->n;
    const int grid = grid_of(g, (g->n + kSNT - 1) / kSNT, 4);
    hipLaunchKernelGGL(k_bkt_count, dim3(grid), dim3(kSNT), 0, st, F, nfp, nall, w->colour, g->degree, w->cnts,
                       ncol, mark, zero_me);
    hipLaunchKernelGGL(k_bkt_scan, dim3(1), dim3(1024), 0, st, w->cnts, w->offs, w->curs, ncol);
    hipLaunchKernelGGL(k_bkt_scatter, dim3(grid), dim3(kSNT), 0, st, F, nfp, nall, w->colour, g->degree, w->curs,
                       ncol, B, pieces);
    FLSG_HIP(hipGetLastError());
    FLSG_HIP(hipMemcpyAsync(w->h_offs, w->offs, sizeof(unsigned int) * 2 * (ncol + 1), hipMemcpyDeviceToHost, st));
    return sor_sync_ctr(w, st);
}

static int ensure_all_lists(flsg_graph g, SorWs *w, hipStream_t st) {
    if (w->all_valid) return FLSG_OK;
    const int ncol = w->ncol;
    // piece capacity: every live hub's pieces (a frontier never holds more)
    std::vector<unsigned int> h(2 * (ncol + 1));
    const int grid = grid_of(g, (g->n + kSNT - 1) / kSNT, 4);
    hipLaunchKernelGGL(k_bkt_count, dim3(grid), dim3(kSNT), 0, st, (const int32_t *)nullptr,
                       (const unsigned int *)nullptr, (unsigned int)g->n, w->colour, g->degree, w->cnts, ncol,
                       (int32_t *)nullptr, (unsigned int *)nullptr);
    hipLaunchKernelGGL(k_bkt_scan, dim3(1), dim3(1024), 0, st, w->cnts, w->offs, w->curs, ncol);
    FLSG_HIP(hipMemcpyAsync(h.data(), w->offs, sizeof(unsigned int) * h.size(), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    const int64_t np = (int64_t)h[2 * ncol + 1] + 1;
    if (np > w->pieces_cap) {
        (void)hipFree(w->pieces);
        (void)hipFree(w->pall);
        w->pieces = nullptr;
        w->pall = nullptr;
        w->pieces_cap = 0;
        FLSG_HIP(hipMalloc((void **)&w->pieces, sizeof(int64_t) * np));
        FLSG_HIP(hipMalloc((void **)&w->pall, sizeof(int64_t) * np));
        w->pieces_cap = np;
    }
    // the scan above left the cursors set: scatter straight away
    hipLaunchKernelGGL(k_bkt_scatter, dim3(grid), dim3(kSNT), 0, st, (const int32_t *)nullptr,
                       (const unsigned int *)nullptr, (unsigned int)g->n, w->colour, g->degree, w->curs, ncol,
                       w->ball, w->pall);
    FLSG_HIP(hipGetLastError());
    FLSG_HIP(hipStreamSynchronize(st));
    w->offs_all = h;
    w->all_valid = 1;
    return FLSG_OK;
}

// one pass over the colours of a bucketed list; offs = entry offsets
// [0..ncol] then piece offsets [0..ncol]
static int colour_pass(flsg_graph g, SorWs *w, SorArgs a, const unsigned int *offs, hipStream_t st) {
    const int ncol = w->ncol;
    const unsigned int *po = offs + ncol + 1;
    unsigned int fbeg = 0, fend = 0;
    for (int c = 0; c <= ncol; c++) {
        unsigned int beg = 0, end = 0, pbeg = 0, pend = 0;
        if (c < ncol) {
            beg = offs[c];
            end = offs[c + 1];
            pbeg = po[c];
            pend = po[c + 1];
            if (end == beg) continue;
        } else if (fend == fbeg) {
            break;
        }
        const int64_t items = (int64_t)(end - beg + kSNT - 1) / kSNT + (pend - pbeg);
        const int64_t fin = (int64_t)(fend - fbeg + kSNT - 1) / kSNT;
        const int grid = grid_of(g, std::max(items, fin), 8);
        hipLaunchKernelGGL(k_sor_items<kPush>, dim3(grid), dim3(kSNT), 0, st, a, beg, end, pbeg, pend, fbeg, fend);
        fbeg = beg;
        fend = end;
    }
    FLSG_HIP(hipGetLastError());
    return FLSG_OK;
}

// local iterations from the frontier in front[0] (count ctr->nf[0]) until
// it is empty
static int run_local(flsg_graph g, SorWs *w, SorArgs a, int64_t max_iters, hipStream_t st, int64_t *iters_out) {
    a.B = w->bucket;
    a.pieces = w->pieces;
    a.mark = w->mark;
    int64_t iters = 0;
    int par = 0;
    for (;;) {
        int rc = bucket(g, w, w->front[par], &w->ctr->nf[par], w->mark, &w->ctr->nf[par ^ 1], w->bucket,
                        w->pieces, st);
        if (rc) return rc;
        if (w->h_ctr->nf[par] == 0) break;
        if (iters >= max_iters) {
            *iters_out = iters;
            return fail(FLSG_ENOCONV, "iteration cap reached");
        }
        iters++;
        rc = colour_pass(g, w, a, w->h_offs, st);
        if (rc) return rc;
        // every entry of the iteration and all its pieces
        const unsigned int ne = w->h_offs[w->ncol], np = w->h_offs[2 * w->ncol + 1];
        a.next = w->front[par ^ 1];
        a.nnext = &w->ctr->nf[par ^ 1];
        const int gridc = grid_of(g, (int64_t)(ne + kSNT - 1) / kSNT + np, 8);
        hipLaunchKernelGGL(k_sor_items<kCheck>, dim3(gridc), dim3(kSNT), 0, st, a, 0u, ne, 0u, np, 0u, 0u);
        FLSG_HIP(hipGetLastError());
        par ^= 1;
    }
    *iters_out = iters;
    return FLSG_OK;
}

static int sor_solve(flsg_graph g, const flsg_params *p, double omega, int64_t s, double *d_x, double *d_r,
                     flsg_result *res, hipStream_t st) {
    SorWs *w = nullptr;
    int rc = sor_ws(g, &w);
    if (rc) return rc;
    rc = ensure_colouring(g, w, st);
    if (rc) return rc;
    if (w->selfloop) return fail(FLSG_EARG, "coloured SOR needs a graph without self-loops");
    rc = ensure_all_lists(g, w, st);   // also sizes the piece buffers
    if (rc) return rc;
    const int64_t n = g->n;
    const int64_t max_iters = p->max_iters > 0 ? p->max_iters : 100000;
    FLSG_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, st));
    FLSG_HIP(hipMemsetAsync(d_r, 0, sizeof(double) * n, st));
    FLSG_HIP(hipMemsetAsync(w->ctr, 0, sizeof(SorCtr), st));
    SorArgs a{};
    a.indptr = g->indptr;
    a.indices = g->indices;
    a.degree = g->degree;
    a.x = d_x;
    a.r = d_r;
    a.res = w->res;
    a.ctr = w->ctr;
    const bool katz = p->kind == FLSG_KATZ;
    a.loops = w->loops;
    a.xcoef = 1.0;
    a.coef = katz ? p->alpha : 1. - p->alpha;
    a.omega = omega;
    a.ea = katz ? p->eps : p->eps * p->alpha;   // katz_solver.py:259 eps * degree; ppr_solver.py:340
    a.local = p->local ? 1 : 0;
    a.katz = katz ? 1 : 0;
    const double r0 = katz ? 1.0 : p->alpha;
    int64_t iters = 0;
    if (p->local) {
        FLSG_HIP(hipMemsetAsync(w->mark, 0, sizeof(int32_t) * n, st));
        hipLaunchKernelGGL(k_sor_init, dim3(1), dim3(1), 0, st, d_r, w->front[0], w->ctr, s, r0, 1);
        rc = run_local(g, w, a, max_iters, st, &iters);
        if (rc) {
            res->iters = iters;
            return rc;
        }
        res->oper = (int64_t)w->h_ctr->oper;
        res->nodes = (int64_t)w->h_ctr->nodes;
    } else {
        hipLaunchKernelGGL(k_sor_init, dim3(1), dim3(1), 0, st, d_r, (int32_t *)nullptr, w->ctr, s, r0, 0);
        a.B = w->ball;
        a.pieces = w->pall;
        const int grida = grid_of(g, (n + kSNT - 1) / kSNT, 4);
        for (;;) {
            if (iters >= max_iters) {
                res->iters = iters;
                return fail(FLSG_ENOCONV, "iteration cap reached");
            }
            iters++;
            rc = colour_pass(g, w, a, w->offs_all.data(), st);
            if (rc) return rc;
            FLSG_HIP(hipMemsetAsync(&w->ctr->act, 0, sizeof(unsigned int), st));
            hipLaunchKernelGGL(k_sor_active, dim3(grida), dim3(kSNT), 0, st, d_r, g->degree, n, a.ea, &w->ctr->act);
            rc = sor_sync_ctr(w, st);
            if (rc) return rc;
            if (w->h_ctr->act == 0) break;
        }
        res->oper = iters * (n + g->m_live);   // ppr_solver.py:447-448 (oper += n + m per sweep)
    }
    if (katz) hipLaunchKernelGGL(k_sor_add, dim3(1), dim3(1), 0, st, d_x, s, -1.0);   // katz_solver.py:332
    res->iters = iters;
    return FLSG_OK;
}

}  // namespace flsg

using namespace flsg;

extern "C" int flsg_graph_colouring(flsg_graph g, int32_t *d_colour_out, int32_t *h_num_colours, int64_t *h_rounds,
                                    void *stream) {
    int rc = check_graph(g);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    SorWs *w = nullptr;
    rc = sor_ws(g, &w);
    if (rc) return rc;
    rc = ensure_colouring(g, w, st);
    if (rc) return rc;
    if (d_colour_out) {
        FLSG_HIP(hipMemcpyAsync(d_colour_out, w->colour, sizeof(int32_t) * g->n, hipMemcpyDeviceToDevice, st));
        FLSG_HIP(hipStreamSynchronize(st));
    }
    if (h_num_colours) *h_num_colours = w->ncol;
    if (h_rounds) *h_rounds = w->rounds;
    return FLSG_OK;
}

extern "C" int flsg_sor_solve(flsg_graph g, const flsg_params *p, double omega, int64_t seed, double *d_x,
                              double *d_r, flsg_result *h_res, void *stream) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (!p || !d_x || !d_r || !h_res) return fail(FLSG_EARG, "NULL argument");
    if (seed < 0 || seed >= g->n) return fail(FLSG_EARG, "seed out of range");
    if (p->kind != FLSG_PPR && p->kind != FLSG_KATZ) return fail(FLSG_EARG, "coloured SOR solves PPR or Katz");
    if (p->precision != FLSG_FP64) return fail(FLSG_EARG, "coloured SOR runs the FP64 path");
    if (!(omega > 0.0 && omega < 2.0)) return fail(FLSG_EARG, "omega must lie in (0, 2)");
    if (!(p->alpha > 0.0 && p->alpha < 1.0)) return fail(FLSG_EARG, "alpha (PPR) / beta (Katz) must lie in (0, 1)");
    hipStream_t st = (hipStream_t)stream;
    int32_t deg_s = 0;
    FLSG_HIP(hipMemcpyAsync(&deg_s, g->degree + seed, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FLSG_HIP(hipStreamSynchronize(st));
    if (deg_s <= 0) return fail(FLSG_EARG, "seed node has degree 0 (the reference divides by zero)");
    if (p->kind == FLSG_PPR && !(p->eps * (double)deg_s <= 1.))
        return fail(FLSG_EASSERT, "assert eps * degree[s_node] <= 1. failed");   // ppr_solver.py:333 (katz_sor has none)
    *h_res = flsg_result{};
    hipEvent_t e0, e1;
    FLSG_HIP(hipEventCreate(&e0));
    FLSG_HIP(hipEventCreate(&e1));
    // the colouring (once per handle) is set-up, outside the solve's time
    SorWs *w = nullptr;
    rc = sor_ws(g, &w);
    if (rc == FLSG_OK) rc = ensure_colouring(g, w, st);
    if (rc == FLSG_OK && !w->selfloop) rc = ensure_all_lists(g, w, st);
    if (rc == FLSG_OK) {
        FLSG_HIP(hipEventRecord(e0, st));
        rc = sor_solve(g, p, omega, seed, d_x, d_r, h_res, st);
    } else {
        FLSG_HIP(hipEventRecord(e0, st));
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

extern "C" int flsg_sor_ignn_column(flsg_graph g, double alpha, double omega, const double *d_eps_vec,
                                    const double *d_dpow1, const double *d_dpowb, double *d_p, double *d_r,
                                    const int32_t *d_nodes, int64_t n_nodes, int64_t max_iters,
                                    flsg_result *h_res, void *stream) {
    int rc = check_graph(g);
    if (rc) return rc;
    if (!d_eps_vec || !d_dpow1 || !d_dpowb || !d_p || !d_r || !h_res) return fail(FLSG_EARG, "NULL argument");
    if (n_nodes < 0 || n_nodes > g->n || (n_nodes > 0 && !d_nodes)) return fail(FLSG_EARG, "bad node list");
    if (!(omega > 0.0 && omega < 2.0)) return fail(FLSG_EARG, "omega must lie in (0, 2)");
    if (!(alpha > 0.0 && alpha < 1.0)) return fail(FLSG_EARG, "alpha must lie in (0, 1)");
    hipStream_t st = (hipStream_t)stream;
    *h_res = flsg_result{};
    SorWs *w = nullptr;
    rc = sor_ws(g, &w);
    if (rc == FLSG_OK) rc = ensure_colouring(g, w, st);
    if (rc == FLSG_OK) rc = ensure_all_lists(g, w, st);   // sizes the piece buffers
    if (rc) return rc;
    hipEvent_t e0, e1;
    FLSG_HIP(hipEventCreate(&e0));
    FLSG_HIP(hipEventCreate(&e1));
    FLSG_HIP(hipEventRecord(e0, st));
    const int64_t n = g->n;
    FLSG_HIP(hipMemsetAsync(w->ctr, 0, sizeof(SorCtr), st));
    FLSG_HIP(hipMemsetAsync(w->mark, 0, sizeof(int32_t) * n, st));
    const int64_t cnt = d_nodes ? n_nodes : n;
    if (cnt > 0)
        hipLaunchKernelGGL(k_sor_seed, dim3(grid_of(g, (cnt + kSNT - 1) / kSNT, 4)), dim3(kSNT), 0, st,
                           d_nodes, n_nodes, n, d_eps_vec, (const double *)d_r, w->front[0], &w->ctr->nf[0]);
    SorArgs a{};
    a.indptr = g->indptr;
    a.indices = g->indices;
    a.degree = g->degree;
    a.x = d_p;
    a.r = d_r;
    a.res = w->res;
    a.ctr = w->ctr;
    a.thr = d_eps_vec;
    a.dp1 = d_dpow1;
    a.dpb = d_dpowb;
    a.loops = w->loops;
    a.coef = 1. - alpha;
    a.xcoef = alpha;   // InstantGNN.py:89 p_u += alpha * res_u
    a.omega = omega;
    a.local = 1;
    a.ignn = 1;
    int64_t iters = 0;
    rc = run_local(g, w, a, max_iters > 0 ? max_iters : 100000, st, &iters);
    h_res->iters = iters;
    h_res->oper = (int64_t)w->h_ctr->oper;
    h_res->nodes = (int64_t)w->h_ctr->nodes;
    hipError_t es = hipEventRecord(e1, st);
    if (es == hipSuccess) es = hipEventSynchronize(e1);
    float ms = 0.f;
    if (es == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == FLSG_OK && es != hipSuccess) rc = fail(FLSG_EHIP, std::string("solve: ") + hipGetErrorString(es));
    h_res->ms = ms;
    h_res->status = rc;
    return rc;
}
