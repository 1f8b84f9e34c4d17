/*
 * graphgen.c -- deterministic synthetic graph generators (host only).
 *
 * The reference reads real graphs from scipy .npz files (demo.py:39,
 * instantGNN-arxiv-exp.py:174); those blobs are not available, so the
 * workloads are synthetic.  Everything here uses its own splitmix64 stream so
 * that a (size, seed) pair gives the same graph on every machine.
 *
 *   flsg_gen_chung_lu      Chung-Lu G(w): edge {u,v} present with probability
 *                          min(1, w_u w_v / sum(w)), no self-loops, sampled in
 *                          O(n + m) by geometric skipping (Miller & Hagberg,
 *                          "Efficient generation of networks with given
 *                          expected degrees", WAW 2011).
 *   flsg_gen_pref_attach   citation-style preferential attachment by node
 *                          index: node i cites earlier nodes, chosen with
 *                          probability proportional to (citations received + 1).
 *   flsg_spanning_forest   Kruskal over a given edge order (union-find), the
 *                          procedure of data_process_GNN.py:31-40.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t sm64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* uniform in (0, 1] */
static inline double unif(uint64_t *s) { return ((double)(sm64(s) >> 11) + 1.0) * (1.0 / 9007199254740992.0); }

/* w must be sorted in non-increasing order; edges are (u, v) with u < v in
 * index space of w.  Returns 0, or 1 if cap was too small (count holds the
 * number written). */
int flsg_gen_chung_lu(int64_t n, const double *w, uint64_t seed, int32_t *out_u, int32_t *out_v,
                      int64_t cap, int64_t *count) {
    double S = 0.0;
    for (int64_t i = 0; i < n; i++) S += w[i];
    uint64_t st = seed * 0x2545F4914F6CDD1Dull + 1;
    int64_t c = 0;
    for (int64_t u = 0; u + 1 < n; u++) {
        int64_t v = u + 1;
        double p = w[u] * w[v] / S;
        if (p > 1.0) p = 1.0;
        while (v < n && p > 0.0) {
            if (p != 1.0) {
                double r = unif(&st);
                double skip = floor(log(r) / log1p(-p));
                if (skip > (double)n) break;
                v += (int64_t)skip;
            }
            if (v < n) {
                double q = w[u] * w[v] / S;
                if (q > 1.0) q = 1.0;
                double r = unif(&st);
                if (r <= q / p) {
                    if (c >= cap) { *count = c; return 1; }
                    out_u[c] = (int32_t)u;
                    out_v[c] = (int32_t)v;
                    c++;
                }
                p = q;
                v += 1;
            }
        }
    }
    *count = c;
    return 0;
}

/* Citation graph: node i >= 1 makes k_i citations to distinct earlier nodes;
 * k_i is drawn so the expected total is m_target.  Targets come from an
 * urn holding every node once plus one copy per citation received. */
int flsg_gen_pref_attach(int64_t n, int64_t m_target, uint64_t seed, int32_t *out_src, int32_t *out_dst,
                         int64_t cap, int64_t *count) {
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 7;
    int32_t *urn = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + cap + 1));
    if (!urn) return 2;
    int64_t urn_len = 0, c = 0;
    double mean_k = (double)m_target / (double)(n - 1);
    int32_t picked[256];
    urn[urn_len++] = 0;
    for (int64_t i = 1; i < n; i++) {
        /* k_i: floor(mean) or ceil(mean) with the right probability, capped by i */
        int64_t k = (int64_t)mean_k;
        if (unif(&st) <= mean_k - (double)k) k += 1;
        if (k > i) k = i;
        if (k > 256) k = 256;
        int64_t np_ = 0;
        int64_t guard = 0;
        while (np_ < k && guard < 64 * k + 64) {
            guard++;
            int32_t t = urn[(int64_t)(unif(&st) * (double)urn_len) % urn_len];
            int dup = 0;
            for (int64_t a = 0; a < np_; a++)
                if (picked[a] == t) { dup = 1; break; }
            if (dup) continue;
            picked[np_++] = t;
        }
        for (int64_t a = 0; a < np_; a++) {
            if (c >= cap) { free(urn); *count = c; return 1; }
            out_src[c] = (int32_t)i;
            out_dst[c] = picked[a];
            c++;
            urn[urn_len++] = picked[a];
        }
        urn[urn_len++] = (int32_t)i;
    }
    free(urn);
    *count = c;
    return 0;
}

static int64_t uf_find(int64_t *par, int64_t x) {
    while (par[x] != x) {
        par[x] = par[par[x]];
        x = par[x];
    }
    return x;
}

/* Kruskal over edges taken in the given order (data_process_GNN.py:31-40):
 * in_tree[e] = 1 for the edges kept in the spanning forest. */
int flsg_spanning_forest(int64_t n, int64_t m, const int32_t *u, const int32_t *v, uint8_t *in_tree) {
    int64_t *par = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    if (!par) return 2;
    for (int64_t i = 0; i < n; i++) par[i] = i;
    int64_t kept = 0;
    for (int64_t e = 0; e < m; e++) {
        in_tree[e] = 0;
        if (kept == n - 1) continue;
        int64_t a = uf_find(par, u[e]), b = uf_find(par, v[e]);
        if (a != b) {
            par[a] = b;
            in_tree[e] = 1;
            kept++;
        }
    }
    free(par);
    return 0;
}
