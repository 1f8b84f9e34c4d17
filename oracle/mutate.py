"""Mutation check of the oracle's pins (test infrastructure, CPU only).

Builds altered copies of flsg_oracle.c -- each with one plausible
restatement slip -- into a temporary directory, runs the oracle's pin tests
(tests/test_oracle.py, tests/test_oracle_pins.py) against each copy through
FLSG_ORACLE_LIB, and prints which tests fail.  A slip no test catches is
reported as SURVIVED and the script exits non-zero.

    python oracle/mutate.py            # every mutation
    python oracle/mutate.py cheb_prev_beta katz_gd_beta_thr
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SRC = os.path.join(HERE, "flsg_oracle.c")
CFLAGS = ["-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-fopenmp", "-std=c11", "-shared"]
TESTS = ["tests/test_oracle.py", "tests/test_oracle_pins.py"]

# marks of the previous queue, for the "stale dx forgotten" slips
_PRELUDE = "static unsigned char MUT_INPREV[1 << 20];\n"


def _r(old, new, nth=0):
    return (old, new, nth)


# name -> (what the slip is, [(old text, new text, which occurrence)])
MUTATIONS = {
    # ---- PPR local Chebyshev (ppr_solver.py:632-657)
    "cheb_prev_beta": ("local PPR Chebyshev uses the previous iteration's beta", [
        _r("            delta_t = 1. / (2. / (1. - alpha) - delta_t);           /* :636 */\n"
           "            double beta = 2. * delta_t / (1. - alpha);              /* :637 */\n",
           "            double beta = 2. * delta_t / (1. - alpha);\n"
           "            delta_t = 1. / (2. / (1. - alpha) - delta_t);\n")]),
    "cheb_enqueue_after_sub": ("local PPR Chebyshev tests u after r_u -= dx_u", [
        _r("                rt[u] += -delta_xt[u];\n",
           "                rt[u] += -delta_xt[u];\n"
           "                if (!q_mark[u] && eps_vec[u] <= fabs(rt[u])) { queue[rear] = u; q_mark[u] = 1; rear += 1; }\n",
           0)]),
    "cheb_always_sort": ("local PPR Chebyshev sorts every queue", [
        _r("if ((double)rear > (double)n / 4.0) nonzero_marks(n, q_mark, queue);   /* :639-640 */",
           "if (1) nonzero_marks(n, q_mark, queue);")]),
    "cheb_never_sort": ("local PPR Chebyshev never sorts the queue", [
        _r("if ((double)rear > (double)n / 4.0) nonzero_marks(n, q_mark, queue);   /* :639-640 */",
           "if (0) nonzero_marks(n, q_mark, queue);")]),
    "cheb_forget_stale_dx": ("local PPR Chebyshev drops dx of a node that was not in the previous queue", [
        _r("                double tmp = beta * rt[u] + (beta - 1.) * delta_xt[u];\n",
           "                double tmp = beta * rt[u] + (beta - 1.) * (MUT_INPREV[u] ? delta_xt[u] : 0.);\n"),
        _r("            int64_t rear_pre = rear;\n",
           "            int64_t rear_pre = rear;\n"
           "            memset(MUT_INPREV, 0, sizeof MUT_INPREV);\n"
           "            for (int64_t k_ = 0; k_ < rear_pre; k_++) MUT_INPREV[queue_pre[k_]] = 1;\n", 0),
        _r("        int64_t rear = 0;\n        for (int64_t j = indptr[s]; j < indptr[s + 1]; j++) {   /* :626-630, u == s_node */\n",
           "        int64_t rear = 0;\n        memset(MUT_INPREV, 0, sizeof MUT_INPREV); MUT_INPREV[s] = 1;\n"
           "        for (int64_t j = indptr[s]; j < indptr[s + 1]; j++) {\n")]),
    "cheb_global_prev_beta": ("global PPR Chebyshev uses the previous sweep's beta", [
        _r("            delta_t = 1. / (2. / (1. - alpha) - delta_t);\n"
           "            double beta = 2. * delta_t / (1. - alpha);\n"
           "            for (int64_t i = 0; i < n; i++) tmp_vec[i] = beta * rt[i] + (beta - 1.) * delta_xt[i];\n",
           "            double beta = 2. * delta_t / (1. - alpha);\n"
           "            delta_t = 1. / (2. / (1. - alpha) - delta_t);\n"
           "            for (int64_t i = 0; i < n; i++) tmp_vec[i] = beta * rt[i] + (beta - 1.) * delta_xt[i];\n")]),
    "cheb_global_sign": ("global PPR Chebyshev with (1 - beta) for (beta - 1)", [
        _r("tmp_vec[i] = beta * rt[i] + (beta - 1.) * delta_xt[i];", "tmp_vec[i] = beta * rt[i] + (1. - beta) * delta_xt[i];")]),
    # ---- Katz local / global Chebyshev (katz_solver.py:396-427, 362-382)
    "katz_cheb_prev_beta": ("local Katz Chebyshev uses the previous iteration's coefficients", [
        _r("            delta_t = 1. / (2. * ((l_max + mu_min) / (l_max - mu_min)) - delta_t);\n"
           "            double beta1 = 4. * delta_t / (l_max - mu_min);\n"
           "            double beta2 = (1. - 2. * delta_t * (l_max + mu_min) / (l_max - mu_min));\n"
           "            if ((double)rear > (double)n / 4.0)",
           "            double beta1 = 4. * delta_t / (l_max - mu_min);\n"
           "            double beta2 = (1. - 2. * delta_t * (l_max + mu_min) / (l_max - mu_min));\n"
           "            delta_t = 1. / (2. * ((l_max + mu_min) / (l_max - mu_min)) - delta_t);\n"
           "            if ((double)rear > (double)n / 4.0)")]),
    "katz_cheb_enqueue_after_sub": ("local Katz Chebyshev tests u after r_u -= dx_u", [
        _r("                rt[u] += -delta_xt[u];\n",
           "                rt[u] += -delta_xt[u];\n"
           "                if (!q_mark[u] && eps_vec[u] <= fabs(rt[u])) { queue[rear] = u; q_mark[u] = 1; rear += 1; }\n",
           1)]),
    "katz_cheb_always_sort": ("local Katz Chebyshev sorts every queue", [
        _r("            if ((double)rear > (double)n / 4.0) nonzero_marks(n, q_mark, queue);\n            trace_iter(tr, queue, rear);\n"
           "            for (int64_t ind = 0; ind < rear; ind++) {\n                int64_t u = queue[ind];\n"
           "                double tmp = beta1",
           "            if (1) nonzero_marks(n, q_mark, queue);\n            trace_iter(tr, queue, rear);\n"
           "            for (int64_t ind = 0; ind < rear; ind++) {\n                int64_t u = queue[ind];\n"
           "                double tmp = beta1")]),
    "katz_cheb_never_sort": ("local Katz Chebyshev never sorts the queue", [
        _r("            if ((double)rear > (double)n / 4.0) nonzero_marks(n, q_mark, queue);\n            trace_iter(tr, queue, rear);\n"
           "            for (int64_t ind = 0; ind < rear; ind++) {\n                int64_t u = queue[ind];\n"
           "                double tmp = beta1",
           "            if (0) nonzero_marks(n, q_mark, queue);\n            trace_iter(tr, queue, rear);\n"
           "            for (int64_t ind = 0; ind < rear; ind++) {\n                int64_t u = queue[ind];\n"
           "                double tmp = beta1")]),
    "katz_cheb_forget_stale_dx": ("local Katz Chebyshev drops dx of a node that was not in the previous queue", [
        _r("                double tmp = beta1 * rt[u] - beta2 * delta_xt[u];\n",
           "                double tmp = beta1 * rt[u] - beta2 * (MUT_INPREV[u] ? delta_xt[u] : 0.);\n"),
        _r("            int64_t rear_pre = rear;\n",
           "            int64_t rear_pre = rear;\n"
           "            memset(MUT_INPREV, 0, sizeof MUT_INPREV);\n"
           "            for (int64_t k_ = 0; k_ < rear_pre; k_++) MUT_INPREV[queue_pre[k_]] = 1;\n", 1),
        _r("        for (int64_t j = indptr[s]; j < indptr[s + 1]; j++) {       /* :390-394 */\n",
           "        memset(MUT_INPREV, 0, sizeof MUT_INPREV); MUT_INPREV[s] = 1;\n"
           "        for (int64_t j = indptr[s]; j < indptr[s + 1]; j++) {\n")]),
    "katz_cheb_global_prev_beta": ("global Katz Chebyshev uses the previous sweep's coefficients", [
        _r("            delta_t = 1. / (2. * ((l_max + mu_min) / (l_max - mu_min)) - delta_t);\n"
           "            double beta1 = 4. * delta_t / (l_max - mu_min);\n"
           "            double beta2 = (1. - 2. * delta_t * (l_max + mu_min) / (l_max - mu_min));\n"
           "            for (int64_t i = 0; i < n; i++) tmp_vec[i]",
           "            double beta1 = 4. * delta_t / (l_max - mu_min);\n"
           "            double beta2 = (1. - 2. * delta_t * (l_max + mu_min) / (l_max - mu_min));\n"
           "            delta_t = 1. / (2. * ((l_max + mu_min) / (l_max - mu_min)) - delta_t);\n"
           "            for (int64_t i = 0; i < n; i++) tmp_vec[i]")]),
    # ---- Katz GD (katz_solver.py:179)
    "katz_gd_beta_thr": ("Katz GD threshold eps*beta*d instead of eps*d", [
        _r("eps_vec[i] = eps * (double)degree[i];   /* :179 */", "eps_vec[i] = eps * alpha * (double)degree[i];")]),
    "katz_gd_push_scaled": ("Katz GD push value beta*r/d instead of beta*r", [
        _r("double val = alpha * tmp_rt[ind];                   /* :212 */",
           "double val = alpha * tmp_rt[ind] / (double)degree[u];")]),
    # ---- heat kernel (hk_solver.py:225-291)
    "hk_pop_psi_next": ("heat-kernel pop test with psi_{j+1}", [
        _r("if (fabs(rc[uu]) < eps_vec[uu] / psi[j]) continue;", "if (fabs(rc[uu]) < eps_vec[uu] / psi[j + 1]) continue;")]),
    "hk_enqueue_psi_cur": ("heat-kernel enqueue test with psi_j", [
        _r("if (fabs(rn[vv]) >= eps_vec[vv] / psi[j + 1] && !mn[vv])", "if (fabs(rn[vv]) >= eps_vec[vv] / psi[j] && !mn[vv])")]),
    "hk_psi_shift": ("heat-kernel psi recurrence with t/(i+2)", [
        _r("psi[i] = psi[i + 1] * tau / ((double)i + 1.) + 1.;", "psi[i] = psi[i + 1] * tau / ((double)i + 2.) + 1.;")]),
    "hk_no_div_n": ("heat-kernel threshold without the 1/N factor", [
        _r("eps_vec[i] = ete * (double)degree[i] / (double)est_n;", "eps_vec[i] = ete * (double)degree[i];")]),
    "hk_last_level_factor": ("heat-kernel last level pushes with the t/(j+1) factor", [
        _r("                        xt[vv] += ru_j / (double)degree[uu];\n                        continue;",
           "                        xt[vv] += mass;\n                        continue;")]),
    "hk_eps_half": ("heat kernel estimates N from eps instead of eps/2", [
        _r("ora_estimate_n(tau, .5 * eps);   /* :225 */", "ora_estimate_n(tau, eps);")]),
    # ---- heat kernel, global level sweep (hk_solver.py:265-279)
    "hk_global_last_level_factor": ("global heat-kernel sweep: last level with the t/(j+1) factor", [
        _r("if (j + 1 == est_n) xt[vv] += ru_j / (double)degree[uu];", "if (j + 1 == est_n) xt[vv] += mass;")]),
    "hk_global_coef_shift": ("global heat-kernel sweep: level coefficient t/(j+2)", [
        _r("double mass = (tau / ((double)j + 1.)) * ru_j / (double)degree[uu];",
           "double mass = (tau / ((double)j + 2.)) * ru_j / (double)degree[uu];", 1)]),
    "hk_global_oper_m": ("global heat-kernel sweep counts m instead of m + n per sweep", [
        _r("oper += (double)(m + n);", "oper += (double)m;")]),
}


def apply(src: str, edits) -> str:
    for old, new, nth in edits:
        idx = [m.start() for m in re.finditer(re.escape(old), src)]
        if len(idx) <= nth:
            raise ValueError(f"mutation site not found ({len(idx)} matches): {old[:80]!r}")
        i = idx[nth]
        src = src[:i] + new + src[i + len(old):]
    return _PRELUDE + src


def run(name: str, tmp: str) -> list[str]:
    desc, edits = MUTATIONS[name]
    with open(SRC) as f:
        src = apply(f.read(), edits)
    c = os.path.join(tmp, name + ".c")
    so = os.path.join(tmp, name + ".so")
    with open(c, "w") as f:
        f.write(src)
    subprocess.run(["gcc", *CFLAGS, "-I", HERE, "-o", so, c, "-lm"], check=True, capture_output=True)
    env = dict(os.environ, FLSG_ORACLE_LIB=so)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "--timeout", "300",
                        "-rf", *TESTS], cwd=REPO, env=env, capture_output=True, text=True)
    return sorted({ln.split(" ")[1].split(" - ")[0] for ln in r.stdout.splitlines() if ln.startswith("FAILED ")})


def main(argv) -> int:
    names = argv or list(MUTATIONS)
    survived = 0
    with tempfile.TemporaryDirectory() as tmp:
        for name in names:
            failed = run(name, tmp)
            desc = MUTATIONS[name][0]
            if failed:
                print(f"{name}: {desc}\n    caught by {len(failed)}: " + ", ".join(failed), flush=True)
            else:
                survived += 1
                print(f"{name}: {desc}\n    SURVIVED", flush=True)
    return 1 if survived else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
