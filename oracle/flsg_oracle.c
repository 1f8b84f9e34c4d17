/*
 * flsg_oracle.c -- CPU restatement of the reference's local graph-diffusion
 * solvers, used ONLY as the correctness checker (tests, smoke check) and as the
 * timed CPU baseline.  Nothing in the product links or calls this file.
 *
 * Every function follows the reference's numba source line by line (cited per
 * function) so that, on the same graph / seed / parameters, it reproduces the
 * reference's frontier order, operation counts and FP64 rounding:
 *   - compiled with -O2 -ffp-contract=off (numba without fastmath never fuses a
 *     multiply and an add), SSE2 doubles, no -ffast-math;
 *   - expressions are evaluated in the reference's left-to-right order.
 *
 * Deliberate, documented deviations (the reference hangs or raises there):
 *   - isolated nodes (degree 0) are never active and never pushed from in the
 *     global sweeps (the reference divides by zero / never terminates);
 *   - Katz's 4-hour wall-clock bail-out is not restated.
 *
 * Graph: CSR with int64 indptr[n+1], int32 indices, int32 degree[n]; the
 * solvers iterate indices[indptr[u] : indptr[u+1]] like the reference.
 *
 * Pinning status of each function is stated in the header of oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_OK 0
#define ORA_EASSERT 1   /* the reference's `assert eps * degree[s] <= 1.` failed */
#define ORA_ENOMEM 2
#define ORA_EARG 3
#define ORA_ETRACE 4    /* trace capacity exceeded (results still valid) */

/* Per-iteration frontier trace: iteration t processed
 * nodes[off[t] : off[t+1]] in this order. */
typedef struct {
    int64_t *off;      /* capacity cap_iters + 1 */
    int32_t *nodes;    /* capacity cap_nodes */
    int64_t cap_iters;
    int64_t cap_nodes;
    int64_t n_iters;   /* out */
    int64_t n_nodes;   /* out */
    int32_t overflow;  /* out */
} ora_trace;

static void trace_begin(ora_trace *tr) {
    if (!tr) return;
    tr->n_iters = 0;
    tr->n_nodes = 0;
    tr->overflow = 0;
    if (tr->cap_iters >= 0 && tr->off) tr->off[0] = 0;
}

static void trace_iter(ora_trace *tr, const int64_t *ids, int64_t cnt) {
    if (!tr) return;
    if (tr->n_iters >= tr->cap_iters || tr->n_nodes + cnt > tr->cap_nodes) {
        tr->overflow = 1;
        return;
    }
    for (int64_t i = 0; i < cnt; i++) tr->nodes[tr->n_nodes + i] = (int32_t)ids[i];
    tr->n_nodes += cnt;
    tr->n_iters += 1;
    tr->off[tr->n_iters] = tr->n_nodes;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ascending ids of marked nodes (np.nonzero(q_mark)[0]) */
static int64_t nonzero_marks(int64_t n, const uint8_t *mark, int64_t *out) {
    int64_t c = 0;
    for (int64_t i = 0; i < n; i++)
        if (mark[i]) out[c++] = i;
    return c;
}

/* Enqueue tests of the last local GD push on this thread whose |r_v| lay
 * within one ulp of the threshold eps_vec[v] (either side): the nodes whose
 * frontier membership a different but equally rounded summation order could
 * flip.  Read with ora_near_threshold() after the call. */
static _Thread_local int64_t g_near;

static inline int64_t near_ulp(double thr, double r) {
    const double a = fabs(r);
    return fabs(a - thr) <= nextafter(thr, INFINITY) - thr ? 1 : 0;
}

int64_t ora_near_threshold(void) { return g_near; }

/* ------------------------------------------------------------------------
 * PPR local / global gradient descent.
 * Restates ppr_solver.py:208-328 (ppr_gd, debug=False).
 * -------------------------------------------------------------------- */
int ora_ppr_gd(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
               int64_t s, double alpha, double eps, int local, double *xt, double *rt,
               int64_t *oper_out, double *time_out, ora_trace *tr) {
    g_near = 0;
    /* ppr_solver.py:213 */
    if (!(eps * (double)degree[s] <= 1.)) return ORA_EASSERT;
    if (degree[s] == 0) return ORA_EARG;
    double t0 = now_s();
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) m += degree[i];           /* :216 */
    memset(xt, 0, sizeof(double) * n);
    memset(rt, 0, sizeof(double) * n);
    rt[s] = alpha;                                              /* :219 */
    double *tmp_rt = (double *)calloc(n, sizeof(double));
    double *eps_vec = (double *)malloc(sizeof(double) * n);
    if (!tmp_rt || !eps_vec) { free(tmp_rt); free(eps_vec); return ORA_ENOMEM; }
    const double ea = eps * alpha;                              /* :221 (eps*alpha)*degree */
    for (int64_t i = 0; i < n; i++) eps_vec[i] = ea * (double)degree[i];
    int64_t oper = 0;
    trace_begin(tr);
    int rc = ORA_OK;
    if (local) {
        int64_t *queue = (int64_t *)malloc(sizeof(int64_t) * n);
        int64_t *st = (int64_t *)malloc(sizeof(int64_t) * n);
        uint8_t *q_mark = (uint8_t *)calloc(n, 1);
        if (!queue || !st || !q_mark) { free(queue); free(st); free(q_mark); free(tmp_rt); free(eps_vec); return ORA_ENOMEM; }
        queue[0] = s;                                           /* :238-244 */
        q_mark[s] = 1;
        int64_t rear = 1;
        st[0] = s;
        for (;;) {
            for (int64_t i = 0; i < rear; i++) oper += degree[queue[i]];   /* :264 */
            if (rear == 0) break;                                          /* :266 */
            if ((double)rear < (double)n / 4.0) {                          /* :269-272 */
                memcpy(st, queue, sizeof(int64_t) * rear);
            } else {
                nonzero_marks(n, q_mark, st);
            }
            trace_iter(tr, st, rear);
            for (int64_t i = 0; i < rear; i++) tmp_rt[i] = rt[st[i]];       /* :273 */
            for (int64_t i = 0; i < rear; i++) q_mark[st[i]] = 0;           /* :274 */
            for (int64_t i = 0; i < rear; i++) xt[st[i]] += tmp_rt[i];      /* :276 */
            for (int64_t i = 0; i < rear; i++) rt[st[i]] -= tmp_rt[i];      /* :277 */
            int64_t tmp_rear = rear;
            rear = 0;
            for (int64_t ind = 0; ind < tmp_rear; ind++) {                  /* :280-288 */
                int64_t u = st[ind];
                double val = (1. - alpha) * tmp_rt[ind] / (double)degree[u];
                for (int64_t j = indptr[u]; j < indptr[u + 1]; j++) {
                    int64_t v = indices[j];
                    rt[v] += val;
                    if (!q_mark[v]) g_near += near_ulp(eps_vec[v], rt[v]);
                    if (!q_mark[v] && eps_vec[v] <= fabs(rt[v])) {
                        queue[rear] = v;
                        q_mark[v] = 1;
                        rear += 1;
                    }
                }
            }
        }
        free(queue); free(st); free(q_mark);
    } else {
        for (;;) {                                                          /* :290-320 */
            int64_t act = 0;
            for (int64_t i = 0; i < n; i++)
                if (degree[i] > 0 && eps_vec[i] <= fabs(rt[i])) act++;
            if (act <= 0) break;
            for (int64_t i = 0; i < n; i++) xt[i] += rt[i];
            memset(tmp_rt, 0, sizeof(double) * n);
            for (int64_t u = 0; u < n; u++) {
                if (degree[u] == 0) continue;
                double val = (1. - alpha) * rt[u] / (double)degree[u];
                for (int64_t j = indptr[u]; j < indptr[u + 1]; j++) tmp_rt[indices[j]] += val;
            }
            memcpy(rt, tmp_rt, sizeof(double) * n);
            oper += n + m;
        }
    }
    free(tmp_rt); free(eps_vec);
    if (oper_out) *oper_out = oper;
    if (time_out) *time_out = now_s() - t0;
    if (tr && tr->overflow) rc = ORA_ETRACE;
    return rc;
}

/* ------------------------------------------------------------------------
 * Chebyshev, shared by PPR and Katz.  PPR's ppr_cheby (ppr_solver.py:582-661)
 * is Katz's katz_cheby (katz_solver.py:340-431) with L+mu = 2,
 * L-mu = 2(1-alpha) -- but the two files write the coefficients and the
 * residual update with different expressions, so each keeps its own code to
 * stay bit-exact.
 * -------------------------------------------------------------------- */

/* ppr_solver.py:582-661 */
int ora_ppr_cheby(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
                  int64_t s, double alpha, double eps, int local, double *xt, double *rt,
                  int64_t *oper_out, double *time_out, ora_trace *tr) {
    g_near = 0;
    if (degree[s] == 0) return ORA_EARG;
    double t0 = now_s();
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) m += degree[i];
    double *delta_xt = (double *)calloc(n, sizeof(double));
    double *tmp_vec = (double *)calloc(n, sizeof(double));
    double *eps_vec = (double *)malloc(sizeof(double) * n);
    if (!delta_xt || !tmp_vec || !eps_vec) { free(delta_xt); free(tmp_vec); free(eps_vec); return ORA_ENOMEM; }
    memset(xt, 0, sizeof(double) * n);
    memset(rt, 0, sizeof(double) * n);
    rt[s] = alpha;                                     /* :588-589 */
    memcpy(xt, rt, sizeof(double) * n);                /* :590 */
    memcpy(delta_xt, xt, sizeof(double) * n);          /* :591 */
    const double ea = eps * alpha;
    for (int64_t i = 0; i < n; i++) eps_vec[i] = ea * (double)degree[i];  /* :592 */
    double delta_t = (1. - alpha);                     /* :593 */
    {                                                   /* :596-601 */
        int64_t u = s;
        double val = (1. - alpha) * rt[u] / (double)degree[u];
        for (int64_t j = indptr[u]; j < indptr[u + 1]; j++) tmp_vec[indices[j]] += val;
    }
    memcpy(rt, tmp_vec, sizeof(double) * n);           /* :601 */
    int64_t oper = 0;
    trace_begin(tr);
    if (!local) {
        for (;;) {                                      /* :603-618 */
            int64_t act = 0;
            for (int64_t i = 0; i < n; i++)
                if (degree[i] > 0 && eps_vec[i] <= fabs(rt[i])) act++;
            if (act <= 0) break;
            oper += n + m;
            delta_t = 1. / (2. / (1. - alpha) - delta_t);
            double beta = 2. * delta_t / (1. - alpha);
            for (int64_t i = 0; i < n; i++) tmp_vec[i] = beta * rt[i] + (beta - 1.) * delta_xt[i];
            for (int64_t i = 0; i < n; i++) xt[i] += tmp_vec[i];
            memcpy(delta_xt, tmp_vec, sizeof(double) * n);
            memset(tmp_vec, 0, sizeof(double) * n);
            for (int64_t u = 0; u < n; u++) {
                if (degree[u] == 0) continue;
                double val = (1. - alpha) * delta_xt[u] / (double)degree[u];
                for (int64_t j = indptr[u]; j < indptr[u + 1]; j++) tmp_vec[indices[j]] += val;
            }
            for (int64_t i = 0; i < n; i++) rt[i] += (tmp_vec[i] - delta_xt[i]);
        }
    } else {
        int64_t *queue = (int64_t *)malloc(sizeof(int64_t) * n);
        int64_t *queue_pre = (int64_t *)malloc(sizeof(int64_t) * n);
        uint8_t *q_mark = (uint8_t *)calloc(n, 1);
        if (!queue || !queue_pre || !q_mark) {
            free(queue); free(queue_pre); free(q_mark); free(delta_xt); free(tmp_vec); free(eps_vec);
            return ORA_ENOMEM;
        }
        int64_t rear = 0;
        for (int64_t j = indptr[s]; j < indptr[s + 1]; j++) {   /* :626-630, u == s_node */
            int64_t v = indices[j];
            if (!q_mark[v] && eps_vec[v] <= fabs(rt[v])) {
                queue[rear] = v;
                q_mark[v] = 1;
                rear += 1;
            }
        }
        for (;;) {
            if (rear == 0) break;                                   /* :633 */
            for (int64_t i = 0; i < rear; i++) oper += degree[queue[i]];   /* :635 */
            delta_t = 1. / (2. / (1. - alpha) - delta_t);           /* :636 */
            double beta = 2. * delta_t / (1. - alpha);              /* :637 */
            if ((double)rear > (double)n / 4.0) nonzero_marks(n, q_mark, queue);   /* :639-640 */
            trace_iter(tr, queue, rear);
            for (int64_t ind = 0; ind < rear; ind++) {              /* :641-646 */
                int64_t u = queue[ind];
                double tmp = beta * rt[u] + (beta - 1.) * delta_xt[u];
                xt[u] += tmp;
                delta_xt[u] = tmp;
                q_mark[u] = 0;
                queue_pre[ind] = u;
            }
            int64_t rear_pre = rear;
            rear = 0;
            for (int64_t ind = 0; ind < rear_pre; ind++) {          /* :649-657 */
                int64_t u = queue_pre[ind];
                double val = (1. - alpha) * delta_xt[u] / (double)degree[u];
                for (int64_t j = indptr[u]; j < indptr[u + 1]; j++) {
                    int64_t v = indices[j];
                    rt[v] += val;
                    if (!q_mark[v]) g_near += near_ulp(eps_vec[v], rt[v]);
                    if (!q_mark[v] && eps_vec[v] <= fabs(rt[v])) {
                        queue[rear] = v;
                        q_mark[v] = 1;
                        rear += 1;
                    }
                }
                rt[u] += -delta_xt[u];
            }
        }
        free(queue); free(queue_pre); free(q_mark);
    }
    free(delta_xt); free(tmp_vec); free(eps_vec);
    if (oper_out) *oper_out = oper;
    if (time_out) *time_out = now_s() - t0;
    return (tr && tr->overflow) ? ORA_ETRACE : ORA_OK;
}

/* ------------------------------------------------------------------------
 * Katz gradient descent: katz_solver.py:167-248.  `alpha` is Katz's beta.
 * -------------------------------------------------------------------- */
int ora_katz_gd(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
                int64_t s, double alpha, double eps, int local, double *xt, double *rt,
                int64_t *oper_out, double *time_out, ora_trace *tr) {
    g_near = 0;
    double t0 = now_s();
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) m += degree[i];
    memset(xt, 0, sizeof(double) * n);
    memset(rt, 0, sizeof(double) * n);
    rt[s] = 1.;                                                 /* :177 */
    double *tmp_rt = (double *)calloc(n, sizeof(double));
    double *eps_vec = (double *)malloc(sizeof(double) * n);
    if (!tmp_rt || !eps_vec) { free(tmp_rt); free(eps_vec); return ORA_ENOMEM; }
    for (int64_t i = 0; i < n; i++) eps_vec[i] = eps * (double)degree[i];   /* :179 */
    int64_t oper = 0;
    trace_begin(tr);
    if (local) {
        int64_t *queue = (int64_t *)malloc(sizeof(int64_t) * n);
        int64_t *st = (int64_t *)malloc(sizeof(int64_t) * n);
        uint8_t *q_mark = (uint8_t *)calloc(n, 1);
        if (!queue || !st || !q_mark) { free(queue); free(st); free(q_mark); free(tmp_rt); free(eps_vec); return ORA_ENOMEM; }
        queue[0] = s;
        q_mark[s] = 1;
        int64_t rear = 1;
        st[0] = s;
        for (;;) {                                                  /* :193-224 */
            for (int64_t i = 0; i < rear; i++) oper += degree[queue[i]];
            if (rear == 0) break;
            if ((double)rear < (double)n / 4.0) memcpy(st, queue, sizeof(int64_t) * rear);
            else nonzero_marks(n, q_mark, st);
            trace_iter(tr, st, rear);
            for (int64_t i = 0; i < rear; i++) tmp_rt[i] = rt[st[i]];
            for (int64_t i = 0; i < rear; i++) q_mark[st[i]] = 0;
            for (int64_t i = 0; i < rear; i++) xt[st[i]] += tmp_rt[i];
            for (int64_t i = 0; i < rear; i++) rt[st[i]] -= tmp_rt[i];
            int64_t tmp_rear = rear;
            rear = 0;
            for (int64_t ind = 0; ind < tmp_rear; ind++) {
                int64_t u = st[ind];
                double val = alpha * tmp_rt[ind];                   /* :212 */
                for (int64_t j = indptr[u]; j < indptr[u + 1]; j++) {
                    int64_t v = indices[j];
                    rt[v] += val;
                    if (!q_mark[v]) g_near += near_ulp(eps_vec[v], rt[v]);
                    if (!q_mark[v] && eps_vec[v] <= fabs(rt[v])) {
                        queue[rear] = v;
                        q_mark[v] = 1;
                        rear += 1;
                    }
                }
            }
        }
        free(queue); free(st); free(q_mark);
    } else {
        for (;;) {                                                  /* :226-242 */
            int64_t act = 0;
            for (int64_t i = 0; i < n; i++)
                if (degree[i] > 0 && eps_vec[i] <= fabs(rt[i])) act++;
            if (act <= 0) break;
            for (int64_t i = 0; i < n; i++) xt[i] += rt[i];
            memset(tmp_rt, 0, sizeof(double) * n);
            for (int64_t u = 0; u < n; u++) {
                double val = alpha * rt[u];
                for (int64_t j = indptr[u]; j < indptr[u + 1]; j++) tmp_rt[indices[j]] += val;
            }
            memcpy(rt, tmp_rt, sizeof(double) * n);
            oper += n + m;
        }
    }
    xt[s] -= 1.;                                                    /* :243 */
    free(tmp_rt); free(eps_vec);
    if (oper_out) *oper_out = oper;
    if (time_out) *time_out = now_s() - t0;
    return (tr && tr->overflow) ? ORA_ETRACE : ORA_OK;
}

/* katz_solver.py:340-431.  Returns x including the e_s component (the
 * reference's katz_cheby does not subtract it, unlike katz_gd). */
int ora_katz_cheby(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
                   int64_t s, double alpha, double l_max, double mu_min, double eps, int local,
                   double *xt, double *rt, int64_t *oper_out, double *time_out, ora_trace *tr) {
    g_near = 0;
    double t0 = now_s();
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) m += degree[i];
    double *delta_xt = (double *)calloc(n, sizeof(double));
    double *tmp_vec = (double *)calloc(n, sizeof(double));
    double *eps_vec = (double *)malloc(sizeof(double) * n);
    if (!delta_xt || !tmp_vec || !eps_vec) { free(delta_xt); free(tmp_vec); free(eps_vec); return ORA_ENOMEM; }
    memset(xt, 0, sizeof(double) * n);
    memset(rt, 0, sizeof(double) * n);
    xt[s] = 2. / (l_max + mu_min);                                  /* :346 */
    memcpy(delta_xt, xt, sizeof(double) * n);                       /* :348 */
    rt[s] = 1. - 2. / (l_max + mu_min);                             /* :350 */
    {
        double val = 2. * alpha / (l_max + mu_min);                 /* :352 */
        for (int64_t j = indptr[s]; j < indptr[s + 1]; j++) rt[indices[j]] += val;
    }
    for (int64_t i = 0; i < n; i++) eps_vec[i] = eps * (double)degree[i];   /* :355 */
    double delta_t = (l_max - mu_min) / (l_max + mu_min);           /* :356 */
    int64_t oper = 0;
    trace_begin(tr);
    if (!local) {
        for (;;) {                                                  /* :362-382 */
            int64_t act = 0;
            for (int64_t i = 0; i < n; i++)
                if (degree[i] > 0 && eps_vec[i] <= fabs(rt[i])) act++;
            if (act <= 0) break;
            oper += n + m;
            delta_t = 1. / (2. * ((l_max + mu_min) / (l_max - mu_min)) - delta_t);
            double beta1 = 4. * delta_t / (l_max - mu_min);
            double beta2 = (1. - 2. * delta_t * (l_max + mu_min) / (l_max - mu_min));
            for (int64_t i = 0; i < n; i++) tmp_vec[i] = beta1 * rt[i] - beta2 * delta_xt[i];
            for (int64_t i = 0; i < n; i++) xt[i] += tmp_vec[i];
            memcpy(delta_xt, tmp_vec, sizeof(double) * n);
            memset(tmp_vec, 0, sizeof(double) * n);
            for (int64_t u = 0; u < n; u++) {
                double val = alpha * delta_xt[u];
                for (int64_t j = indptr[u]; j < indptr[u + 1]; j++) tmp_vec[indices[j]] += val;
            }
            for (int64_t i = 0; i < n; i++) rt[i] = rt[i] + tmp_vec[i] - delta_xt[i];
        }
    } else {
        int64_t *queue = (int64_t *)malloc(sizeof(int64_t) * n);
        int64_t *queue_pre = (int64_t *)malloc(sizeof(int64_t) * n);
        uint8_t *q_mark = (uint8_t *)calloc(n, 1);
        if (!queue || !queue_pre || !q_mark) {
            free(queue); free(queue_pre); free(q_mark); free(delta_xt); free(tmp_vec); free(eps_vec);
            return ORA_ENOMEM;
        }
        int64_t rear = 0;
        for (int64_t j = indptr[s]; j < indptr[s + 1]; j++) {       /* :390-394 */
            int64_t v = indices[j];
            if (!q_mark[v] && eps_vec[v] <= fabs(rt[v])) {
                queue[rear] = v;
                q_mark[v] = 1;
                rear += 1;
            }
        }
        for (;;) {                                                  /* :396-427 */
            for (int64_t i = 0; i < rear; i++) oper += degree[queue[i]];
            if (rear == 0) break;
            delta_t = 1. / (2. * ((l_max + mu_min) / (l_max - mu_min)) - delta_t);
            double beta1 = 4. * delta_t / (l_max - mu_min);
            double beta2 = (1. - 2. * delta_t * (l_max + mu_min) / (l_max - mu_min));
            if ((double)rear > (double)n / 4.0) nonzero_marks(n, q_mark, queue);
            trace_iter(tr, queue, rear);
            for (int64_t ind = 0; ind < rear; ind++) {
                int64_t u = queue[ind];
                double tmp = beta1 * rt[u] - beta2 * delta_xt[u];
                xt[u] += tmp;
                delta_xt[u] = tmp;
                q_mark[u] = 0;
                queue_pre[ind] = u;
            }
            int64_t rear_pre = rear;
            rear = 0;
            for (int64_t ind = 0; ind < rear_pre; ind++) {
                int64_t u = queue_pre[ind];
                double val = alpha * delta_xt[u];
                for (int64_t j = indptr[u]; j < indptr[u + 1]; j++) {
                    int64_t v = indices[j];
                    rt[v] += val;
                    if (!q_mark[v]) g_near += near_ulp(eps_vec[v], rt[v]);
                    if (!q_mark[v] && eps_vec[v] <= fabs(rt[v])) {
                        queue[rear] = v;
                        q_mark[v] = 1;
                        rear += 1;
                    }
                }
                rt[u] += -delta_xt[u];
            }
        }
        free(queue); free(queue_pre); free(q_mark);
    }
    free(delta_xt); free(tmp_vec); free(eps_vec);
    if (oper_out) *oper_out = oper;
    if (time_out) *time_out = now_s() - t0;
    return (tr && tr->overflow) ? ORA_ETRACE : ORA_OK;
}

/* ------------------------------------------------------------------------
 * Heat kernel.  hk_solver.py:15-48 (estimate_coefficient / estimate_n),
 * :51-80 (hk_taylor_approx), :206-295 (hk_sor_queue).
 * -------------------------------------------------------------------- */

/* hk_solver.py:15-25: t^{k+1}/(k+1)! */
double ora_estimate_coefficient(double t, int64_t k) {
    double t0 = t;
    if (k == 0) return t0;
    for (int64_t i = 0; i < k; i++) t0 = t0 * (t / (double)(i + 2));
    return t0;
}

/* hk_solver.py:28-48.  k = ceil(t-1) is used as an integer count; the loop's
 * per-step factor is built from the initial k, as written. */
int64_t ora_estimate_n(double t, double eps) {
    int64_t k = (int64_t)ceil(t - 1.);
    double kf = (double)k;
    double cur_est = ora_estimate_coefficient(t, k) * ((kf + 2.) / (kf + 2. - t));
    if (cur_est < eps * exp(t)) return k;
    int64_t est_k = k;
    for (int64_t i = k; i < 1000 * (k + 1); i++) {
        double tmp = ((kf + 2. - t) * t / pow(kf + 2., 2.)) * ((kf + 3.) / (kf + 3. - t));
        cur_est = cur_est * tmp;
        est_k += 1;
        if (cur_est < eps * exp(t)) return est_k;
    }
    return est_k;
}

/* hk_solver.py:51-80: x = e^{-t} sum_{k=0}^{N} (t^k/k!) (A D^-1)^k e_s.
 * n_terms <= 0 means N = estimate_n(t, eps). */
int ora_hk_taylor(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
                  int64_t s, double tau, double eps, int64_t n_terms, double *xt, int64_t *est_n_out) {
    int64_t est_n = n_terms > 0 ? n_terms : ora_estimate_n(tau, eps);
    double *cur = (double *)calloc(n, sizeof(double));
    double *tmp = (double *)calloc(n, sizeof(double));
    if (!cur || !tmp) { free(cur); free(tmp); return ORA_ENOMEM; }
    cur[s] = 1.;
    memset(xt, 0, sizeof(double) * n);
    for (int64_t i = 0; i < n; i++) xt[i] = xt[i] + cur[i];
    for (int64_t k = 0; k < est_n; k++) {
        memset(tmp, 0, sizeof(double) * n);
        for (int64_t u = 0; u < n; u++) {
            if (degree[u] == 0) continue;
            double val = cur[u] / (double)degree[u];
            for (int64_t j = indptr[u]; j < indptr[u + 1]; j++) tmp[indices[j]] += val;
        }
        double f = tau / ((double)k + 1.);
        for (int64_t i = 0; i < n; i++) cur[i] = f * tmp[i];
        for (int64_t i = 0; i < n; i++) xt[i] = xt[i] + cur[i];
    }
    double et = exp(tau);
    for (int64_t i = 0; i < n; i++) xt[i] = xt[i] / et;
    free(cur); free(tmp);
    if (est_n_out) *est_n_out = est_n;
    return ORA_OK;
}

/* hk_solver.py:206-295 (hk_sor_queue).  `rt` has 2*n entries (two levels).
 * n_terms <= 0 means N = estimate_n(tau, eps/2) as in the reference. */
int ora_hk_push(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
                int64_t s, double tau, double eps, double omega, int local, int64_t n_terms,
                double *xt, double *rt, double *oper_out, int64_t *est_n_out, double *time_out,
                ora_trace *tr) {
    if (degree[s] == 0) return ORA_EARG;
    double t0 = now_s();
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) m += degree[i];
    memset(xt, 0, sizeof(double) * n);
    memset(rt, 0, sizeof(double) * 2 * n);
    int64_t *queue = (int64_t *)malloc(sizeof(int64_t) * 2 * n);
    uint8_t *q_mark = (uint8_t *)calloc(2 * n, 1);
    double *eps_vec = (double *)malloc(sizeof(double) * n);
    int64_t est_n = n_terms > 0 ? n_terms : ora_estimate_n(tau, .5 * eps);   /* :225 */
    double *psi = (double *)calloc(est_n + 1, sizeof(double));
    int64_t *lvl = (int64_t *)malloc(sizeof(int64_t) * n);
    if (!queue || !q_mark || !eps_vec || !psi || !lvl) {
        free(queue); free(q_mark); free(eps_vec); free(psi); free(lvl);
        return ORA_ENOMEM;
    }
    int64_t rear[2] = {0, 0}, front[2] = {0, 0};
    queue[0] = s;                                                   /* :218-223 */
    q_mark[s] = 1;
    rear[0] = 1;
    rt[s] = 1.;
    psi[est_n] = 1.;                                                /* :226-229 */
    for (int64_t i = est_n - 1; i >= 0; i--) psi[i] = psi[i + 1] * tau / ((double)i + 1.) + 1.;
    const double ete = exp(tau) * eps;
    for (int64_t i = 0; i < n; i++) eps_vec[i] = ete * (double)degree[i] / (double)est_n;   /* :230 */
    double oper = 0.;
    trace_begin(tr);
    for (int64_t j = 0; j < est_n; j++) {
        int qc = (int)(j % 2), qn = (int)((j + 1) % 2);
        double *rc = rt + (int64_t)qc * n, *rn = rt + (int64_t)qn * n;
        int64_t *qcq = queue + (int64_t)qc * n, *qnq = queue + (int64_t)qn * n;
        uint8_t *mn = q_mark + (int64_t)qn * n, *mc = q_mark + (int64_t)qc * n;
        if (local) {
            if (rear[qc] == front[qc]) break;                       /* :237 */
            int64_t ntr = 0;
            while (front[qc] != rear[qc]) {                         /* :238-264 */
                int64_t uu = qcq[front[qc] % n];
                front[qc] += 1;
                if (fabs(rc[uu]) < eps_vec[uu] / psi[j]) continue;
                lvl[ntr++] = uu;
                oper += (double)degree[uu];
                double ru_j = omega * rc[uu];
                xt[uu] += ru_j;
                rc[uu] -= ru_j;
                double mass = (tau / ((double)j + 1.)) * ru_j / (double)degree[uu];
                for (int64_t e = indptr[uu]; e < indptr[uu + 1]; e++) {
                    int64_t vv = indices[e];
                    if (j + 1 == est_n) {
                        xt[vv] += ru_j / (double)degree[uu];
                        continue;
                    }
                    rn[vv] = rn[vv] + mass;
                    if (fabs(rn[vv]) >= eps_vec[vv] / psi[j + 1] && !mn[vv]) {
                        mn[vv] = 1;
                        qnq[rear[qn]] = vv;
                        rear[qn] += 1;
                    }
                }
                if (fabs(rc[uu]) >= eps_vec[uu] / psi[j] && !mn[uu]) {
                    qcq[rear[qc] % n] = uu;
                    rear[qc] += 1;
                }
            }
            trace_iter(tr, lvl, ntr);
        } else {
            for (;;) {                                              /* :265-279 */
                for (int64_t uu = 0; uu < n; uu++) {
                    if (degree[uu] == 0) continue;
                    double ru_j = omega * rc[uu];
                    xt[uu] += ru_j;
                    rc[uu] -= ru_j;
                    double mass = (tau / ((double)j + 1.)) * ru_j / (double)degree[uu];
                    for (int64_t e = indptr[uu]; e < indptr[uu + 1]; e++) {
                        int64_t vv = indices[e];
                        if (j + 1 == est_n) xt[vv] += ru_j / (double)degree[uu];
                        else rn[vv] = rn[vv] + mass;
                    }
                }
                oper += (double)(m + n);
                int64_t act = 0;
                for (int64_t i = 0; i < n; i++)
                    if (degree[i] > 0 && (eps_vec[i] / psi[j]) <= fabs(rc[i])) act++;
                if (act <= 0) break;
            }
        }
        front[qc] = 0;                                              /* :281-285 */
        rear[qc] = 0;
        memset(rc, 0, sizeof(double) * n);
        memset(mc, 0, n);
    }
    double et = exp(tau);
    for (int64_t i = 0; i < n; i++) xt[i] /= et;                   /* :291 */
    free(queue); free(q_mark); free(eps_vec); free(psi); free(lvl);
    if (oper_out) *oper_out = oper;
    if (est_n_out) *est_n_out = est_n;
    if (time_out) *time_out = now_s() - t0;
    return (tr && tr->overflow) ? ORA_ETRACE : ORA_OK;
}

/* ------------------------------------------------------------------------
 * InstantGNN propagation on a slack CSR (live adjacency of u is
 * indices[indptr[u] : indptr[u] + degree[u]]).
 * InstantGNN.py:12-43 (instantGNNPush), :46-71 (instantGNNParallel, per
 * column), :137-180 (updetaGraph, insert-only), :298-329
 * (instantGNNUpdateBatch, algo 'fwd...insert').
 * p is float32 in the initial pass (InstantGNN.py:55) and float64 in updates
 * (the driver's P array), so the push exists in both forms.
 * -------------------------------------------------------------------- */

#define IGNN_PUSH_BODY(PT)                                                                     \
    int64_t oper = 0;                                                                          \
    while (*rear != *front) {                                                                  \
        *front = (*front + 1) % qn;                                                            \
        int64_t u = queue[*front];                                                             \
        q_mark[u] = 0;                                                                         \
        if (fabs(r[u]) < eps_vec[u]) continue;                                                 \
        if ((degree[u] == 1 && indices[indptr[u]] == u) || (degree[u] == 0)) {                 \
            p[u] = (PT)r[u];                                                                   \
            r[u] = 0;                                                                          \
            continue;                                                                          \
        }                                                                                      \
        double res_u = r[u];                                                                   \
        p[u] = (PT)((double)p[u] + alpha * res_u);                                             \
        double push_u = (1. - alpha) * res_u / pow((double)degree[u], 1. - beta);              \
        r[u] = 0;                                                                              \
        oper += degree[u];                                                                     \
        for (int64_t e = indptr[u]; e < indptr[u] + degree[u]; e++) {                          \
            int64_t v = indices[e];                                                            \
            r[v] += push_u / pow((double)degree[v], beta);                                     \
            if (!q_mark[v] && fabs(r[v]) >= eps_vec[v]) {                                      \
                *rear = (*rear + 1) % qn;                                                      \
                queue[*rear] = v;                                                              \
                q_mark[v] = 1;                                                                 \
            }                                                                                  \
        }                                                                                      \
        if (!q_mark[u] && fabs(r[u]) >= eps_vec[u]) {                                          \
            *rear = (*rear + 1) % qn;                                                          \
            queue[*rear] = u;                                                                  \
            q_mark[u] = 1;                                                                     \
        }                                                                                      \
    }                                                                                          \
    return oper;

/* qn is the ring size (the reference passes n+2). Returns the number of edge
 * visits (sum of live degrees of the nodes pushed). */
static int64_t ignn_push_f32(int64_t qn, const int64_t *indptr, const int32_t *indices,
                             const int32_t *degree, double alpha, const double *eps_vec, float *p,
                             double *r, int64_t *queue, int64_t *front, int64_t *rear,
                             uint8_t *q_mark, double beta) {
    IGNN_PUSH_BODY(float)
}

static int64_t ignn_push_f64(int64_t qn, const int64_t *indptr, const int32_t *indices,
                             const int32_t *degree, double alpha, const double *eps_vec, double *p,
                             double *r, int64_t *queue, int64_t *front, int64_t *rear,
                             uint8_t *q_mark, double beta) {
    IGNN_PUSH_BODY(double)
}

/* One column of instantGNNParallel (InstantGNN.py:53-67): p (float32, out)
 * starts at zero, r (in/out) starts at the feature column. */
int ora_ignn_column(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
                    double alpha, double eps, double beta, float *p, double *r, int64_t *oper_out) {
    int64_t qn = n + 2;
    int64_t *queue = (int64_t *)calloc(qn, sizeof(int64_t));
    uint8_t *q_mark = (uint8_t *)calloc(n, 1);
    double *eps_vec = (double *)malloc(sizeof(double) * n);
    if (!queue || !q_mark || !eps_vec) { free(queue); free(q_mark); free(eps_vec); return ORA_ENOMEM; }
    for (int64_t i = 0; i < n; i++) eps_vec[i] = pow((double)degree[i], 1. - beta) * eps;
    for (int64_t i = 0; i < n; i++) p[i] = 0.f;
    int64_t front = 0, rear = 0;
    for (int64_t i = 0; i < n; i++) {
        if (fabs(r[i]) >= eps_vec[i]) {
            rear = (rear + 1) % qn;
            queue[rear] = i;
            q_mark[i] = 1;
        }
    }
    int64_t oper = ignn_push_f32(qn, indptr, indices, degree, alpha, eps_vec, p, r, queue, &front,
                                 &rear, q_mark, beta);
    free(queue); free(q_mark); free(eps_vec);
    if (oper_out) *oper_out = oper;
    return ORA_OK;
}

/* updetaGraph with onlyinsert=True (InstantGNN.py:137-180).  Mutates indices
 * and degree in place.  The returned insert_neighbor dict keeps first-insertion
 * key order: keys[k] has neighbours nbr[nbr_off[k] : nbr_off[k+1]].
 * Capacities: keys and nbr_off need 2*n_edges(+1) entries, nbr 2*n_edges. */
int ora_update_graph(int64_t n, const int64_t *indptr, int32_t *indices, int32_t *degree,
                     const int32_t *us, const int32_t *vs, int64_t n_edges, int64_t *keys,
                     int64_t *nbr_off, int64_t *nbr, int64_t *n_keys_out) {
    /* first pass: record accepted (u, v) pairs in order */
    int64_t *pu = (int64_t *)malloc(sizeof(int64_t) * (n_edges > 0 ? n_edges : 1));
    int64_t *pv = (int64_t *)malloc(sizeof(int64_t) * (n_edges > 0 ? n_edges : 1));
    int64_t *slot = (int64_t *)malloc(sizeof(int64_t) * n);
    int64_t *cnt = (int64_t *)calloc(2 * n_edges + 1, sizeof(int64_t));
    if (!pu || !pv || !slot || !cnt) { free(pu); free(pv); free(slot); free(cnt); return ORA_ENOMEM; }
    int64_t na = 0;
    for (int64_t e = 0; e < n_edges; e++) {
        int64_t u = us[e], v = vs[e];
        int insertion = 1;
        for (int64_t i = indptr[u]; i < indptr[u] + degree[u]; i++) {
            if (v == indices[i]) { insertion = 0; break; }
        }
        if (insertion) {
            if (indptr[u] + degree[u] >= indptr[u + 1] || indptr[v] + degree[v] >= indptr[v + 1]) {
                free(pu); free(pv); free(slot); free(cnt);
                return ORA_EARG;   /* slack exhausted: the reference would overwrite a neighbour row */
            }
            indices[indptr[u] + degree[u]] = (int32_t)v;
            indices[indptr[v] + degree[v]] = (int32_t)u;
            degree[u] += 1;
            degree[v] += 1;
            pu[na] = u;
            pv[na] = v;
            na++;
        }
    }
    /* dict in first-insertion order: for (u,v) the key u is touched before v */
    for (int64_t i = 0; i < n; i++) slot[i] = -1;
    int64_t nk = 0;
    for (int64_t a = 0; a < na; a++) {
        int64_t ends[2] = {pu[a], pv[a]};
        for (int k = 0; k < 2; k++) {
            int64_t x = ends[k];
            if (slot[x] < 0) { slot[x] = nk; keys[nk] = x; nk++; }
            cnt[slot[x]] += 1;
        }
    }
    nbr_off[0] = 0;
    for (int64_t k = 0; k < nk; k++) nbr_off[k + 1] = nbr_off[k] + cnt[k];
    for (int64_t k = 0; k < nk; k++) cnt[k] = nbr_off[k];
    for (int64_t a = 0; a < na; a++) {
        int64_t u = pu[a], v = pv[a];
        nbr[cnt[slot[u]]++] = v;
        nbr[cnt[slot[v]]++] = u;
    }
    free(pu); free(pv); free(slot); free(cnt);
    if (n_keys_out) *n_keys_out = nk;
    return ORA_OK;
}

/* instantGNNUpdateBatch, algo 'fwdbatchinsert' (InstantGNN.py:298-329) for one
 * column: p, r float64 (the driver's P, R), x the feature column. */
int ora_ignn_update_column(int64_t n, const int64_t *indptr, const int32_t *indices,
                           const int32_t *degree, const int32_t *old_degree, double *p, double *r,
                           const double *x, double alpha, double eps, double beta,
                           const int64_t *keys, const int64_t *nbr_off, const int64_t *nbr,
                           int64_t n_keys, int64_t *oper_out) {
    int64_t qn = n + 2;
    int64_t *queue = (int64_t *)calloc(qn, sizeof(int64_t));
    uint8_t *q_mark = (uint8_t *)calloc(n, 1);
    double *eps_vec = (double *)malloc(sizeof(double) * n);
    if (!queue || !q_mark || !eps_vec) { free(queue); free(q_mark); free(eps_vec); return ORA_ENOMEM; }
    for (int64_t i = 0; i < n; i++) eps_vec[i] = pow((double)degree[i], 1. - beta) * eps;  /* :338-339 */
    int64_t front = 0, rear = 0;
    for (int64_t k = 0; k < n_keys; k++) {                                       /* :303-305 */
        int64_t u = keys[k];
        double du = (double)degree[u], od = (double)old_degree[u];
        p[u] *= pow(du / od, 1. - beta);
        r[u] += p[u] * (pow(od, 1. - beta) - pow(du, 1. - beta)) / pow(du, 1. - beta) * 1. / alpha;
    }
    for (int64_t k = 0; k < n_keys; k++) {                                       /* :306-318 */
        int64_t u = keys[k];
        double du = (double)degree[u], od = (double)old_degree[u];
        double dr = (p[u] + alpha * r[u] - alpha * x[u]) * (pow(od, beta) - pow(du, beta)) / pow(du, beta);
        for (int64_t q = nbr_off[k]; q < nbr_off[k + 1]; q++) {
            int64_t v = nbr[q];
            dr += (1. - alpha) * p[v] / (pow(du, beta) * pow((double)degree[v], 1. - beta));
        }
        r[u] += dr / alpha;
        if (fabs(r[u]) >= eps_vec[u]) {
            rear = (rear + 1) % qn;
            queue[rear] = u;
            q_mark[u] = 1;
        }
    }
    int64_t oper = 0;
    if (rear != front)
        oper = ignn_push_f64(qn, indptr, indices, degree, alpha, eps_vec, p, r, queue, &front, &rear,
                             q_mark, beta);
    free(queue); free(q_mark); free(eps_vec);
    if (oper_out) *oper_out = oper;
    return ORA_OK;
}

/* ------------------------------------------------------------------------
 * CPU baseline driver: local GD / local Chebyshev over many seeds with OpenMP
 * (the reference fans seeds out over numba prange, ppr_solver.py:537,568).
 * kind 0 = PPR local GD, 1 = PPR local Chebyshev.  Per-seed oper and
 * seconds out; returns the first non-OK status.
 * -------------------------------------------------------------------- */
int ora_ppr_batch(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
                  const int64_t *seeds, int64_t n_seeds, double alpha, double eps, int kind,
                  int nthreads, int64_t *opers, double *times) {
    int status = ORA_OK;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double *x = (double *)malloc(sizeof(double) * n);
        double *r = (double *)malloc(sizeof(double) * n);
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < n_seeds; i++) {
            int rc;
            if (!x || !r) rc = ORA_ENOMEM;
            else if (kind == 0)
                rc = ora_ppr_gd(n, indptr, indices, degree, seeds[i], alpha, eps, 1, x, r, &opers[i], &times[i], NULL);
            else
                rc = ora_ppr_cheby(n, indptr, indices, degree, seeds[i], alpha, eps, 1, x, r, &opers[i], &times[i], NULL);
            if (rc != ORA_OK) {
#pragma omp critical
                { if (status == ORA_OK) status = rc; }
            }
        }
        free(x);
        free(r);
    }
    return status;
}

/* ------------------------------------------------------------------------
 * Checkers of sparse results (test and bench infrastructure, no reference
 * counterpart).  A result set is n_q sparse vectors: query q holds
 * ids[off[q] .. off[q] + cnt[q]) and vals[...] (caller's node ids).
 *
 * ora_ppr_batch_l1: per query, the FP64 local GD of the seed (ora_ppr_gd)
 * and the L1 distance of result q to it, sum_v |x_v - y_v|, and the oracle's
 * sum_v |x_v|; OpenMP over queries.
 * -------------------------------------------------------------------- */
int ora_ppr_batch_l1(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
                     const int64_t *seeds, int64_t n_q, double alpha, double eps, int nthreads,
                     const int64_t *off, const int64_t *cnt, const int32_t *ids, const double *vals,
                     double *l1_out, double *norm_out, int64_t *opers) {
    int status = ORA_OK;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double *x = (double *)malloc(sizeof(double) * n);
        double *r = (double *)malloc(sizeof(double) * n);
#pragma omp for schedule(dynamic, 1)
        for (int64_t q = 0; q < n_q; q++) {
            double t = 0.0;
            int rc = (!x || !r) ? ORA_ENOMEM
                                : ora_ppr_gd(n, indptr, indices, degree, seeds[q], alpha, eps, 1, x, r, &opers[q], &t, NULL);
            if (rc != ORA_OK) {
#pragma omp critical
                { if (status == ORA_OK) status = rc; }
                continue;
            }
            double norm = 0.0;
            for (int64_t v = 0; v < n; v++) norm += fabs(x[v]);
            /* sum over all v of |x_v - y_v|: start from sum |x_v| and correct
             * the entries the result holds */
            double l1 = norm;
            for (int64_t k = off[q]; k < off[q] + cnt[q]; k++) {
                const int32_t v = ids[k];
                l1 += fabs(x[v] - vals[k]) - fabs(x[v]);
            }
            l1_out[q] = l1;
            norm_out[q] = norm;
        }
        free(x);
        free(r);
    }
    return status;
}

/* ora_residual_cert: per query, the residual of result x (a sparse vector)
 * recomputed in FP64 from the graph alone, for PPR (kind 0, a = alpha):
 *     r = a e_s - x + (1 - a) A D^-1 x,   thr_v = eps a d_v,
 * or Katz (kind 1, a = beta; x the full solution, e_s included):
 *     r = e_s - x + a A x,                   thr_v = eps d_v,
 * scattered over the live adjacency of each support node; cert_out[q] =
 * max over touched v of |r_v| / thr_v (nodes of degree 0 count only when
 * r_v != 0, as +inf) and rl1_out[q] = sum |r_v|.  OpenMP over queries.
 * -------------------------------------------------------------------- */
int ora_residual_cert(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree, int kind,
                      double a, double eps, const int64_t *seeds, int64_t n_q, int nthreads, const int64_t *off,
                      const int64_t *cnt, const int32_t *ids, const double *vals, double *cert_out,
                      double *rl1_out) {
    int status = ORA_OK;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double *rv = (double *)calloc(n, sizeof(double));
        uint8_t *mark = (uint8_t *)calloc(n, 1);
        int64_t *list = (int64_t *)malloc(sizeof(int64_t) * n);
#pragma omp for schedule(dynamic, 1)
        for (int64_t q = 0; q < n_q; q++) {
            if (!rv || !mark || !list) {
#pragma omp critical
                { if (status == ORA_OK) status = ORA_ENOMEM; }
                continue;
            }
            int64_t nt = 0;
#define ORA_TOUCH(v_) do { if (!mark[v_]) { mark[v_] = 1; list[nt++] = (v_); } } while (0)
            const int64_t s = seeds[q];
            if (s < 0 || s >= n) {
#pragma omp critical
                { if (status == ORA_OK) status = ORA_EARG; }
                continue;
            }
            ORA_TOUCH(s);
            for (int64_t k = off[q]; k < off[q] + cnt[q]; k++) {
                const int64_t u = ids[k];
                const double xu = vals[k];
                ORA_TOUCH(u);
                if (kind == 0) {
                    rv[u] -= xu;
                    if (degree[u] > 0) {
                        const double c = (1.0 - a) * xu / (double)degree[u];
                        for (int64_t e = indptr[u]; e < indptr[u] + degree[u]; e++) {
                            const int64_t v = indices[e];
                            ORA_TOUCH(v);
                            rv[v] += c;
                        }
                    }
                } else {
                    rv[u] -= xu;
                    const double c = a * xu;
                    for (int64_t e = indptr[u]; e < indptr[u] + degree[u]; e++) {
                        const int64_t v = indices[e];
                        ORA_TOUCH(v);
                        rv[v] += c;
                    }
                }
            }
            rv[s] += kind == 0 ? a : 1.0;
#undef ORA_TOUCH
            const double ts = kind == 0 ? eps * a : eps;
            double cmax = 0.0, l1 = 0.0;
            for (int64_t i = 0; i < nt; i++) {
                const int64_t v = list[i];
                const double rr = fabs(rv[v]);
                l1 += rr;
                if (rr > 0.0) {
                    const double c = degree[v] > 0 ? rr / (ts * (double)degree[v]) : INFINITY;
                    if (c > cmax) cmax = c;
                }
                rv[v] = 0.0;
                mark[v] = 0;
            }
            cert_out[q] = cmax;
            rl1_out[q] = l1;
        }
        free(rv);
        free(mark);
        free(list);
    }
    return status;
}

int ora_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------
 * SOR / Gauss-Southwell push (ppr_solver.py:331-461, ppr_sor).
 *
 * local = 1: FIFO queue of capacity n+2 used circularly, with the iteration
 * flag node n (:356-366); a popped node is skipped while eps_vec[u] > |r_u|
 * (:398-399); an active node adds degree[u] to oper (a float, :401), moves
 * res = omega*r_u into x_u, pushes (1-alpha)*res/degree[u] to its live
 * neighbours (:402-413) and is re-queued while still active (:415-419).
 * local = 0: natural-order sweeps u = 0..n-1 updating in place (:421-429),
 * oper += n + m per sweep, stop once no node is active (:449-452).
 * Deviation (as for the global GD sweeps above): isolated nodes are not
 * counted as active -- their eps_vec is 0, so the reference's count never
 * reaches 0 on a graph with one; they are skipped in the sweep, which is exact
 * since their residual stays 0.
 * iters_out: local = non-empty flag passes (iterations), global = sweeps.
 * -------------------------------------------------------------------- */
int ora_ppr_sor(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
                int64_t s, double alpha, double eps, double omega, int local, double *xt, double *rt,
                double *oper_out, int64_t *iters_out, double *time_out) {
    if (!(eps * (double)degree[s] <= 1.)) return ORA_EASSERT;   /* :333 */
    if (degree[s] == 0) return ORA_EARG;
    double t0 = now_s();
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) m += degree[i];            /* :336 */
    memset(xt, 0, sizeof(double) * n);
    memset(rt, 0, sizeof(double) * n);
    rt[s] = alpha;                                               /* :339 */
    double *eps_vec = (double *)malloc(sizeof(double) * n);
    if (!eps_vec) return ORA_ENOMEM;
    const double ea = eps * alpha;                               /* :340 */
    for (int64_t i = 0; i < n; i++) eps_vec[i] = ea * (double)degree[i];
    double oper = 0.;
    int64_t iters = 0;
    if (local) {
        const int64_t q_max = n + 2;                             /* :354-356 */
        int64_t *queue = (int64_t *)malloc(sizeof(int64_t) * q_max);
        uint8_t *q_mark = (uint8_t *)calloc(q_max, 1);
        if (!queue || !q_mark) { free(queue); free(q_mark); free(eps_vec); return ORA_ENOMEM; }
        int64_t front = 0, rear = 0;
        queue[rear] = n; rear = (rear + 1) % q_max; q_mark[n] = 1;            /* :358-361 */
        queue[rear] = s; rear = (rear + 1) % q_max; q_mark[s] = 1;            /* :363-365 */
        for (;;) {
            const int64_t u = queue[front];                                    /* :371-373 */
            front = (front + 1) % q_max;
            q_mark[u] = 0;
            if (u == n) {                                                      /* :375-396 */
                if (rear - front == 0) break;
                queue[rear] = n; rear = (rear + 1) % q_max; q_mark[u] = 1;
                iters++;
                continue;
            }
            if (eps_vec[u] > fabs(rt[u])) continue;                            /* :398-399 */
            oper += (double)degree[u];                                         /* :401 */
            const double res = omega * rt[u];                                  /* :404-406 */
            xt[u] += res;
            rt[u] -= res;
            const double push_res = (1. - alpha) * res / (double)degree[u];   /* :407 */
            for (int64_t j = indptr[u]; j < indptr[u] + degree[u]; j++) {     /* :408-413 */
                const int64_t v = indices[j];
                rt[v] += push_res;
                if (!q_mark[v] && eps_vec[v] <= fabs(rt[v])) {
                    queue[rear] = v; rear = (rear + 1) % q_max; q_mark[v] = 1;
                }
            }
            if (!q_mark[u] && eps_vec[u] <= fabs(rt[u])) {                    /* :415-419 */
                queue[rear] = u; rear = (rear + 1) % q_max; q_mark[u] = 1;
            }
        }
        free(queue); free(q_mark);
    } else {
        for (;;) {
            for (int64_t u = 0; u < n; u++) {                                  /* :421-429 */
                if (degree[u] == 0) continue;
                const double res = omega * rt[u];
                xt[u] += res;
                rt[u] -= res;
                const double push_res = (1. - alpha) * res / (double)degree[u];
                for (int64_t j = indptr[u]; j < indptr[u] + degree[u]; j++) rt[indices[j]] += push_res;
            }
            oper += (double)(n + m);                                           /* :447-448 */
            iters++;
            int64_t act = 0;                                                   /* :449-452 */
            for (int64_t i = 0; i < n; i++)
                if (degree[i] > 0 && eps_vec[i] <= fabs(rt[i])) act++;
            if (act <= 0) break;
        }
    }
    free(eps_vec);
    if (oper_out) *oper_out = oper;
    if (iters_out) *iters_out = iters;
    if (time_out) *time_out = now_s() - t0;
    return ORA_OK;
}

/* ------------------------------------------------------------------------
 * Katz SOR / Gauss-Southwell push (katz_solver.py:252-337, katz_sor): the
 * same control flow as ppr_sor with r_s = 1 (:258), eps_vec = eps*degree
 * (:259), push value alpha*res (alpha is Katz's beta, :302, :321), full CSR
 * rows (:303, :322), no assert, and x_s -= 1 at the end (:332).  The 4-hour
 * wall-clock bail-outs (:288, :330) are not restated; isolated nodes are not
 * counted as active, as in ora_ppr_sor.
 * -------------------------------------------------------------------- */
int ora_katz_sor(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
                 int64_t s, double alpha, double eps, double omega, int local, double *xt, double *rt,
                 double *oper_out, int64_t *iters_out, double *time_out) {
    if (degree[s] == 0) return ORA_EARG;
    double t0 = now_s();
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) m += degree[i];            /* :255 */
    memset(xt, 0, sizeof(double) * n);
    memset(rt, 0, sizeof(double) * n);
    rt[s] = 1.0;                                                 /* :258 */
    double *eps_vec = (double *)malloc(sizeof(double) * n);
    if (!eps_vec) return ORA_ENOMEM;
    for (int64_t i = 0; i < n; i++) eps_vec[i] = eps * (double)degree[i];   /* :259 */
    double oper = 0.;
    int64_t iters = 0;
    if (local) {
        const int64_t q_max = n + 2;                             /* :264-274 */
        int64_t *queue = (int64_t *)malloc(sizeof(int64_t) * q_max);
        uint8_t *q_mark = (uint8_t *)calloc(q_max, 1);
        if (!queue || !q_mark) { free(queue); free(q_mark); free(eps_vec); return ORA_ENOMEM; }
        int64_t front = 0, rear = 0;
        queue[rear] = n; rear = (rear + 1) % q_max; q_mark[n] = 1;
        queue[rear] = s; rear = (rear + 1) % q_max; q_mark[s] = 1;
        for (;;) {
            const int64_t u = queue[front];                                    /* :277-279 */
            front = (front + 1) % q_max;
            q_mark[u] = 0;
            if (u == n) {                                                      /* :281-293 */
                if (rear - front == 0) break;
                queue[rear] = n; rear = (rear + 1) % q_max; q_mark[u] = 1;
                iters++;
                continue;
            }
            if (eps_vec[u] > fabs(rt[u])) continue;                            /* :295-296 */
            oper += (double)degree[u];                                         /* :298 */
            const double res = omega * rt[u];                                  /* :299-302 */
            xt[u] += res;
            rt[u] -= res;
            const double push_res = alpha * res;
            for (int64_t j = indptr[u]; j < indptr[u + 1]; j++) {             /* :303-308 */
                const int64_t v = indices[j];
                rt[v] += push_res;
                if (!q_mark[v] && eps_vec[v] <= fabs(rt[v])) {
                    queue[rear] = v; rear = (rear + 1) % q_max; q_mark[v] = 1;
                }
            }
            if (!q_mark[u] && eps_vec[u] <= fabs(rt[u])) {                    /* :310-313 */
                queue[rear] = u; rear = (rear + 1) % q_max; q_mark[u] = 1;
            }
        }
        free(queue); free(q_mark);
    } else {
        for (;;) {
            for (int64_t u = 0; u < n; u++) {                                  /* :317-323 */
                if (degree[u] == 0) continue;
                const double res = omega * rt[u];
                xt[u] += res;
                rt[u] -= res;
                const double push_res = alpha * res;
                for (int64_t j = indptr[u]; j < indptr[u + 1]; j++) rt[indices[j]] += push_res;
            }
            oper += (double)(n + m);                                           /* :324 */
            iters++;
            int64_t act = 0;                                                   /* :325-326 */
            for (int64_t i = 0; i < n; i++)
                if (degree[i] > 0 && eps_vec[i] <= fabs(rt[i])) act++;
            if (act <= 0) break;
        }
    }
    xt[s] -= 1.;                                                               /* :332 */
    free(eps_vec);
    if (oper_out) *oper_out = oper;
    if (iters_out) *iters_out = iters;
    if (time_out) *time_out = now_s() - t0;
    return ORA_OK;
}

/* ------------------------------------------------------------------------
 * Graph colouring and the coloured SOR push: NOT the reference's algorithm.
 * They restate, one step after another, what libflsg's GPU SOR variant does,
 * so its results can be checked exactly (the colouring) or to rounding (the
 * push).  The reference's ppr_sor above is the semantic anchor: both stop on
 * the same certificate (every live |r_u| < eps*alpha*d_u).
 *
 * Colouring: Jones-Plassmann rounds.  Priority of u = (degree[u], mix32(u))
 * compared as one 64-bit key (mix32 is a bijection, so keys are distinct).
 * Isolated nodes get colour 0 up front.  A round selects every uncoloured u
 * whose live neighbours (self-loops ignored) include no uncoloured node of
 * higher key -- an independent set -- and gives each the smallest colour no
 * coloured neighbour holds.  Returns the number of colours; *rounds_out the
 * number of rounds.
 * -------------------------------------------------------------------- */
static inline uint32_t mix32(uint32_t h) {   /* MurmurHash3 fmix32 (bijective) */
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    return h;
}
static inline uint64_t colour_key(const int32_t *degree, int64_t u) {
    return ((uint64_t)(uint32_t)degree[u] << 32) | (uint64_t)mix32((uint32_t)u);
}

int64_t ora_colour(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
                   int32_t *colour, int64_t *rounds_out) {
    uint8_t *sel = (uint8_t *)calloc(n, 1);
    if (!sel) return -1;
    int64_t left = 0, ncol = 1, rounds = 0;
    for (int64_t u = 0; u < n; u++) {
        colour[u] = degree[u] == 0 ? 0 : -1;
        left += degree[u] != 0;
    }
    while (left > 0) {
        rounds++;
        for (int64_t u = 0; u < n; u++) {
            sel[u] = 0;
            if (colour[u] >= 0) continue;
            const uint64_t ku = colour_key(degree, u);
            int ok = 1;
            for (int64_t j = indptr[u]; j < indptr[u] + degree[u] && ok; j++) {
                const int64_t v = indices[j];
                if (v != u && colour[v] < 0 && colour_key(degree, v) > ku) ok = 0;
            }
            sel[u] = (uint8_t)ok;
        }
        for (int64_t u = 0; u < n; u++) {
            if (!sel[u]) continue;
            /* smallest colour not held by a coloured neighbour, 64 at a time */
            for (int64_t base = 0;; base += 64) {
                uint64_t used = 0;
                for (int64_t j = indptr[u]; j < indptr[u] + degree[u]; j++) {
                    const int64_t v = indices[j];
                    const int64_t c = colour[v];
                    if (v != u && !sel[v] && c >= base && c < base + 64) used |= 1ull << (c - base);
                }
                if (~used) {
                    colour[u] = (int32_t)(base + __builtin_ctzll(~used));
                    break;
                }
            }
            if (colour[u] + 1 > ncol) ncol = colour[u] + 1;
            left--;
        }
    }
    free(sel);
    if (rounds_out) *rounds_out = rounds;
    return ncol;
}

/* Coloured SOR push (local or global), sequential restatement of the GPU
 * variant.  Per iteration (local: the current frontier F; global: every live
 * node), colours are visited in increasing order; within colour c every
 * listed node u with eps_vec[u] <= |r_u| (global: every live node) takes
 * res = omega*r_u and pushes (1-alpha)*res/degree[u] to its neighbours --
 * none of which has colour c, so the nodes of one colour are independent.
 * After the colour, each such u applies x_u += res, r_u -= res (the GPU
 * applies it in the next colour's launch).  Local: the next frontier is every
 * node whose residual changed in the iteration (listed at its first change)
 * and satisfies eps_vec[v] <= |r_v| when the iteration ends -- the active set
 * after the sweep, whatever the order of the adds within a colour.  Nodes of
 * F are processed in F's order within a colour. */
int ora_sor_coloured(int kind, int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
                     const int32_t *colour, int64_t ncol, int64_t s, double alpha, double eps,
                     double omega, int local, int64_t max_iters, double *xt, double *rt,
                     int64_t *oper_out, int64_t *iters_out, int64_t *nodes_out) {
    const int katz = kind == 1;   /* 0: PPR (ppr_sor), 1: Katz (katz_sor, alpha = beta) */
    if (!katz && !(eps * (double)degree[s] <= 1.)) return ORA_EASSERT;
    if (degree[s] == 0) return ORA_EARG;
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) m += degree[i];
    memset(xt, 0, sizeof(double) * n);
    memset(rt, 0, sizeof(double) * n);
    rt[s] = katz ? 1.0 : alpha;
    const double ea = katz ? eps : eps * alpha;
    double *eps_vec = (double *)malloc(sizeof(double) * n);
    int64_t *F = (int64_t *)malloc(sizeof(int64_t) * n);
    int64_t *G = (int64_t *)malloc(sizeof(int64_t) * n);
    int64_t *B = (int64_t *)malloc(sizeof(int64_t) * n);
    double *res = (double *)malloc(sizeof(double) * n);
    int64_t *coff = (int64_t *)calloc(ncol + 1, sizeof(int64_t));
    uint8_t *mark = (uint8_t *)calloc(n, 1);
    if (!eps_vec || !F || !G || !B || !res || !coff || !mark) {
        free(eps_vec); free(F); free(G); free(B); free(res); free(coff); free(mark);
        return ORA_ENOMEM;
    }
    for (int64_t i = 0; i < n; i++) eps_vec[i] = ea * (double)degree[i];
    int64_t nf = 0, oper = 0, iters = 0, nodes = 0;
    if (local) {
        F[nf++] = s;
    } else {
        for (int64_t u = 0; u < n; u++)
            if (degree[u] > 0) F[nf++] = u;
    }
    int rc = ORA_OK;
    while (nf > 0) {
        if (iters >= max_iters) { rc = ORA_EARG; break; }
        iters++;
        /* stable bucket of F by colour */
        memset(coff, 0, sizeof(int64_t) * (ncol + 1));
        for (int64_t i = 0; i < nf; i++) coff[colour[F[i]] + 1]++;
        for (int64_t c = 0; c < ncol; c++) coff[c + 1] += coff[c];
        for (int64_t i = 0; i < nf; i++) B[coff[colour[F[i]]]++] = F[i];
        for (int64_t c = ncol; c > 0; c--) coff[c] = coff[c - 1];
        coff[0] = 0;
        int64_t ng = 0;
        for (int64_t c = 0; c < ncol; c++) {
            for (int64_t i = coff[c]; i < coff[c + 1]; i++) {
                const int64_t u = B[i];
                res[i] = 0.0;
                if (local && eps_vec[u] > fabs(rt[u])) continue;
                if (local) { oper += degree[u]; nodes++; }
                const double rs = omega * rt[u];
                res[i] = rs;
                const double push_res = katz ? alpha * rs : (1. - alpha) * rs / (double)degree[u];
                for (int64_t j = indptr[u]; j < indptr[u] + degree[u]; j++) {
                    const int64_t v = indices[j];
                    rt[v] += push_res;
                    if (local && !mark[v]) { mark[v] = 1; G[ng++] = v; }   /* touched */
                }
            }
            for (int64_t i = coff[c]; i < coff[c + 1]; i++) {
                const int64_t u = B[i];
                if (res[i] == 0.0) continue;   /* inactive (local) or zero residual */
                xt[u] += res[i];
                rt[u] -= res[i];
                if (local && !mark[u]) { mark[u] = 1; G[ng++] = u; }
            }
        }
        if (local) {
            nf = 0;
            for (int64_t i = 0; i < ng; i++) {
                const int64_t v = G[i];
                mark[v] = 0;
                if (degree[v] > 0 && eps_vec[v] <= fabs(rt[v])) F[nf++] = v;
            }
        } else {
            oper += n + m;
            int64_t act = 0;
            for (int64_t i = 0; i < n; i++)
                if (degree[i] > 0 && eps_vec[i] <= fabs(rt[i])) act++;
            if (act == 0) nf = 0;
        }
    }
    if (katz) xt[s] -= 1.;   /* katz_solver.py:332 */
    free(eps_vec); free(F); free(G); free(B); free(res); free(coff); free(mark);
    if (oper_out) *oper_out = oper;
    if (iters_out) *iters_out = iters;
    if (nodes_out) *nodes_out = nodes;
    return rc;
}

/* ------------------------------------------------------------------------
 * InstantGNN SOR push for one feature column (InstantGNN.py:74-105,
 * instantGNNSorPush).  Ring queue of length qn (the drivers pass n + 2),
 * entries front+1 .. rear; a popped u is skipped while |r_u| < eps_vec[u]
 * (:82-83); an isolated node, or one whose only edge is a self-loop, takes
 * p_u = r_u, r_u = 0 (:84-87); otherwise res = omega r_u, p_u += alpha res,
 * push = (1-alpha) res / d_u^(1-beta), r_u = (1-omega) r_u, and every live
 * neighbour (self-loops included) gets push / d_v^beta with the enqueue test
 * (:88-97); u is re-queued while active (:98-101).  pow() is the host libm's,
 * as numba's ** on floats.  Returns the sum of degrees pushed in *oper_out.
 * -------------------------------------------------------------------- */
int ora_ignn_sor_push(int64_t qn, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
                      double alpha, const double *eps_vec, double omega, double *p, double *r,
                      int64_t *queue, int64_t front, int64_t rear, uint8_t *q_mark, double beta,
                      int64_t *oper_out) {
    int64_t oper = 0;
    while (rear != front) {
        front = (front + 1) % qn;                                               /* :79-81 */
        const int64_t u = queue[front];
        q_mark[u] = 0;
        if (fabs(r[u]) < eps_vec[u]) continue;                                  /* :82-83 */
        if ((degree[u] == 1 && indices[indptr[u]] == u) || degree[u] == 0) {   /* :84-87 */
            p[u] = r[u];
            r[u] = 0;
            continue;
        }
        oper += degree[u];
        const double res_u = omega * r[u];                                      /* :88-91 */
        p[u] += alpha * res_u;
        const double push_u = (1. - alpha) * res_u / pow((double)degree[u], 1 - beta);
        r[u] = (1 - omega) * r[u];
        for (int64_t j = indptr[u]; j < indptr[u] + degree[u]; j++) {          /* :92-97 */
            const int64_t v = indices[j];
            r[v] += push_u / pow((double)degree[v], beta);
            if (!q_mark[v] && fabs(r[v]) >= eps_vec[v]) {
                rear = (rear + 1) % qn;
                queue[rear] = v;
                q_mark[v] = 1;
            }
        }
        if (!q_mark[u] && fabs(r[u]) >= eps_vec[u]) {                         /* :98-101 */
            rear = (rear + 1) % qn;
            queue[rear] = u;
            q_mark[u] = 1;
        }
    }
    if (oper_out) *oper_out = oper;
    return ORA_OK;
}

/* Coloured InstantGNN SOR for one column: the GPU variant, restated.  The
 * frontier starts as F0 (distinct nodes) and each iteration visits colours in
 * order: a listed node u of colour c with eps_vec[u] <= |r_u| and r_u != 0
 * either absorbs (isolated or self-loop only: p_u = r_u, r_u = 0) or takes
 * res = omega r_u and sends (1-alpha) res / dp1[u] / dpb[v] to each live
 * neighbour v != u; after the colour each pushed u applies p_u += alpha res
 * and r_u += loops[u] * ((1-alpha) res / dp1[u] / dpb[u]) - res (its
 * self-loops).  The next frontier is every pushed node or neighbour of one
 * with eps_vec <= |r| and r != 0 at the end of the iteration.  dp1 / dpb are
 * d^(1-beta) / d^beta, loops[u] the self-loops in u's row. */
int ora_ignn_sor_coloured(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *degree,
                          const int32_t *colour, int64_t ncol, double alpha, double omega,
                          const double *eps_vec, const double *dp1, const double *dpb, const int32_t *loops,
                          const int32_t *F0, int64_t nF0, double *p, double *r, int64_t max_iters,
                          int64_t *oper_out, int64_t *iters_out) {
    int64_t *F = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    int64_t *G = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    int64_t *B = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    double *res = (double *)malloc(sizeof(double) * (n + 1));
    int64_t *coff = (int64_t *)calloc(ncol + 1, sizeof(int64_t));
    uint8_t *mark = (uint8_t *)calloc(n, 1);
    uint8_t *seen = (uint8_t *)calloc(n, 1);
    if (!F || !G || !B || !res || !coff || !mark || !seen) {
        free(F); free(G); free(B); free(res); free(coff); free(mark); free(seen);
        return ORA_ENOMEM;
    }
    int64_t nf = 0, oper = 0, iters = 0;
    for (int64_t i = 0; i < nF0; i++) F[nf++] = F0[i];
    int rc = ORA_OK;
    while (nf > 0) {
        if (iters >= max_iters) { rc = ORA_EARG; break; }
        iters++;
        for (int64_t i = 0; i < nf; i++) mark[F[i]] = 0;
        memset(coff, 0, sizeof(int64_t) * (ncol + 1));
        for (int64_t i = 0; i < nf; i++) coff[colour[F[i]] + 1]++;
        for (int64_t c = 0; c < ncol; c++) coff[c + 1] += coff[c];
        for (int64_t i = 0; i < nf; i++) B[coff[colour[F[i]]]++] = F[i];
        for (int64_t c = ncol; c > 0; c--) coff[c] = coff[c - 1];
        coff[0] = 0;
        int64_t ng = 0;   /* pushed nodes and their neighbours, listed once */
        for (int64_t c = 0; c < ncol; c++) {
            for (int64_t i = coff[c]; i < coff[c + 1]; i++) {
                const int64_t u = B[i];
                res[i] = 0.0;
                if (!(eps_vec[u] <= fabs(r[u]) && r[u] != 0.0)) continue;
                const int d = degree[u];
                if (d == 0 || (d == 1 && loops[u] == 1)) {
                    res[i] = r[u];
                    continue;
                }
                oper += d;
                const double rs = omega * r[u];
                res[i] = rs;
                const double push_u = (1. - alpha) * rs / dp1[u];
                for (int64_t j = indptr[u]; j < indptr[u] + d; j++) {
                    const int64_t v = indices[j];
                    if (v == u) continue;
                    r[v] += push_u / dpb[v];
                    if (!seen[v]) { seen[v] = 1; G[ng++] = v; }
                }
            }
            for (int64_t i = coff[c]; i < coff[c + 1]; i++) {
                const int64_t u = B[i];
                const double rs = res[i];
                if (rs == 0.0) continue;
                const int d = degree[u];
                if (d == 0 || (d == 1 && loops[u] == 1)) {
                    p[u] = rs;
                    r[u] = 0.0;
                } else {
                    p[u] += alpha * rs;
                    const double push_u = (1. - alpha) * rs / dp1[u];
                    r[u] += (double)loops[u] * (push_u / dpb[u]) - rs;
                }
                if (!seen[u]) { seen[u] = 1; G[ng++] = u; }
            }
        }
        nf = 0;
        for (int64_t i = 0; i < ng; i++) {
            const int64_t v = G[i];
            seen[v] = 0;
            if (eps_vec[v] <= fabs(r[v]) && r[v] != 0.0 && !mark[v]) { mark[v] = 1; F[nf++] = v; }
        }
    }
    free(F); free(G); free(B); free(res); free(coff); free(mark); free(seen);
    if (oper_out) *oper_out = oper;
    if (iters_out) *iters_out = iters;
    return rc;
}
