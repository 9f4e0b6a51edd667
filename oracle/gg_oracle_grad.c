/*
 * gg_oracle_grad.c -- CPU restatement of the differentiated and batched forms
 * that stand beside the reference's apps.  TEST INFRASTRUCTURE ONLY (see
 * gg_oracle.h): only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.
 *
 * These forms have no counterpart in the reference; they are defined on the
 * reference's weighted graph (out-edge CSR with per-edge weights, the
 * WEIGHTED layout of graph-numa.h:74-89 and 170-185 with the weight as a
 * separate FP32 array) and its two traversal directions:
 *
 *   SpMV            y[d] = sum_{e=(s->d)} w_e x[s]               (app/SPMV.C:17-21)
 *   SpMV adjoint    gx[s] = sum_{e=(s->d)} w_e gy[d],  gw[e] = x[s] gy[d]
 *   weighted PR     r <- (1-a) v + a P r,  P[d][s] = w_{s->d} / S_s,
 *                   S_s = sum of s's out-weights, all vertices active
 *                   (the PageRankBit.C form, app/PageRankBit.C:145, with
 *                   row-normalised weights; dangling sources contribute 0)
 *   PR adjoint      lambda <- u + a P^T lambda  (implicit function theorem
 *                   at the fixed point r) until the L1 step is at most
 *                   eps times the L1 norm of the new iterate,
 *                   dL/dw_e = (a r_s / S_s) (lambda_d - mu_s),
 *                   mu_s = (P^T lambda)_s = (lambda_s - u_s) / a,
 *                   dL/dv = (1-a) lambda
 *
 * Sums run sequentially in CSR order in FP64.  Pins (tests/test_oracle_grad.py):
 * scipy sparse products and direct solves of (I - aP) r = (1-a) v and
 * (I - aP^T) lambda = u, dense brute-force derivatives of P on tiny graphs,
 * and central finite differences of the loss.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "gg_oracle.h"

void ggo_out_wsum(uint32_t n, const uint32_t *off, const float *w, double *S)
{
    for (uint32_t s = 0; s < n; s++) {
        double a = 0;
        for (uint32_t e = off[s]; e < off[s + 1]; e++) a += (double)w[e];
        S[s] = a;
    }
}

void ggo_spmv_csr(uint32_t n, const uint32_t *off, const uint32_t *ids, const float *w,
                  const double *x, double *y)
{
    memset(y, 0, sizeof(double) * n);
    for (uint32_t s = 0; s < n; s++)
        for (uint32_t e = off[s]; e < off[s + 1]; e++) y[ids[e]] += (w ? (double)w[e] : 1.0) * x[s];
}

void ggo_spmv_bwd(uint32_t n, const uint32_t *off, const uint32_t *ids, const float *w,
                  const double *x, const double *gy, double *gx, double *gw)
{
    for (uint32_t s = 0; s < n; s++) {
        double a = 0;
        for (uint32_t e = off[s]; e < off[s + 1]; e++) {
            const uint32_t d = ids[e];
            a += (w ? (double)w[e] : 1.0) * gy[d];
            if (gw) gw[e] = x[s] * gy[d];
        }
        if (gx) gx[s] = a;
    }
}

int ggo_wpagerank(uint32_t n, const uint32_t *off, const uint32_t *ids, const float *w,
                  double alpha, double eps, int max_iter, const double *v, double *r,
                  double *l1_hist)
{
    double *S = malloc(sizeof(double) * n), *c = malloc(sizeof(double) * n);
    double *acc = malloc(sizeof(double) * n);
    if (!S || !c || !acc) {
        free(S), free(c), free(acc);
        return -1;
    }
    ggo_out_wsum(n, off, w, S);
    const double q = 1 / (double)n;
    for (uint32_t i = 0; i < n; i++) r[i] = v ? v[i] : q;
    int it = 0;
    while (it < max_iter) {
        it++;
        for (uint32_t s = 0; s < n; s++) c[s] = S[s] > 0 ? r[s] / S[s] : 0.0;
        memset(acc, 0, sizeof(double) * n);
        for (uint32_t s = 0; s < n; s++)
            for (uint32_t e = off[s]; e < off[s + 1]; e++) acc[ids[e]] += (double)w[e] * c[s];
        double l1 = 0;
        for (uint32_t d = 0; d < n; d++) {
            const double rn = (1 - alpha) * (v ? v[d] : q) + alpha * acc[d];
            l1 += fabs(rn - r[d]);
            r[d] = rn;
        }
        if (l1_hist) l1_hist[it - 1] = l1;
        if (l1 < eps) break;
    }
    free(S), free(c), free(acc);
    return it;
}

int ggo_wpagerank_adjoint(uint32_t n, const uint32_t *off, const uint32_t *ids, const float *w,
                          double alpha, double eps, int max_iter, const double *r, const double *u,
                          double *lambda, double *gw, double *gv)
{
    double *S = malloc(sizeof(double) * n), *nl = malloc(sizeof(double) * n);
    if (!S || !nl) {
        free(S), free(nl);
        return -1;
    }
    ggo_out_wsum(n, off, w, S);
    for (uint32_t i = 0; i < n; i++) lambda[i] = u[i];
    int it = 0;
    while (it < max_iter) {
        it++;
        double l1 = 0, norm = 0;
        for (uint32_t s = 0; s < n; s++) {
            double a = 0;
            for (uint32_t e = off[s]; e < off[s + 1]; e++) a += (double)w[e] * lambda[ids[e]];
            nl[s] = S[s] > 0 ? u[s] + (alpha * a) / S[s] : u[s];
            l1 += fabs(nl[s] - lambda[s]);
            norm += fabs(nl[s]);
        }
        memcpy(lambda, nl, sizeof(double) * n);
        if (l1 <= eps * norm) break;
    }
    if (gw)
        for (uint32_t s = 0; s < n; s++) {
            if (!(S[s] > 0)) {
                for (uint32_t e = off[s]; e < off[s + 1]; e++) gw[e] = 0;
                continue;
            }
            const double coef = (alpha * r[s]) / S[s];
            const double mu = (lambda[s] - u[s]) / alpha;
            for (uint32_t e = off[s]; e < off[s + 1]; e++) gw[e] = coef * (lambda[ids[e]] - mu);
        }
    if (gv)
        for (uint32_t i = 0; i < n; i++) gv[i] = (1 - alpha) * lambda[i];
    free(S), free(nl);
    return it;
}
