/*
 * gg_oracle_apps.c -- CPU restatement of the reference apps beyond the four
 * on the hot path (the "next" apps on the same engine).  TEST
 * INFRASTRUCTURE ONLY (see gg_oracle.h).
 *
 *   ggo_bc   app/BC.C:154-226: single-source betweenness contributions.
 *            Forward: level by level from 'root', NumPaths[d] += NumPaths[s]
 *            over edges s->d with s in the frontier and d not yet visited
 *            (BC_F, 28-77; a vertex joins the next level on its first
 *            update); backward on the transposed graph, deepest level
 *            first: Dependencies[d] += Dependencies[s] for s in level r+1
 *            and d not yet visited (BC_Back_F, 80-124), then level r adds
 *            inverseNumPaths (BC_Back_Vertex_F, 141-152); finally
 *            (Dependencies - inv) / inv (222).  fType is float in the
 *            reference; this restatement computes in FP64.
 *   ggo_sssp app/BellmanFord.C: distances by relaxation rounds over the
 *            frontier of changed vertices (integer weights).
 *   ggo_prdelta app/PageRankDelta.C:124-176: each round the active
 *            vertices push Delta[s]/outdeg(s) along their out-edges
 *            (PR_F, 27-72) into nghSum; then every vertex updates p and
 *            Delta (PR_Vertex_F_FirstRound in round 1, PR_Vertex_F after,
 *            74-110) and stays active while |Delta| > eps2 * p; stop when
 *            sum |Delta| < eps.
 *   ggo_prconv app/PageRankConverage.C:227-310: add_factor = 0.85 p/outdeg
 *            pushed along out-edges (PR_F, 34-88), then p_next += (1 -
 *            sum p_next)/n (PR_Vertex_Csum), L1 = sum |p_curr - p_next|,
 *            p_next /= sum p_next (PR_Vertex_norm), swap.
 *
 * Pins (tests/test_oracle_apps.py): textbook Brandes dependencies and path
 * counts on small graphs; textbook Bellman-Ford / Dijkstra distances;
 * PageRankDelta with eps2 = 0 against PageRank power iteration (the two are
 * the same recurrence then) and with a huge eps2 against one power step;
 * PageRankConverage's first iteration in closed form and its converged
 * ranks against a direct solve of the dangling-corrected PageRank system.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "gg_oracle.h"

int ggo_bc(uint32_t n, const uint32_t *off, const uint32_t *ids, uint32_t root, double *score,
           int32_t *level, double *sigma_out)
{
    if (root >= n) return -1;
    const uint64_t m = off[n];
    uint32_t *toff = malloc(sizeof(uint32_t) * ((size_t)n + 1));
    uint32_t *tids = malloc(sizeof(uint32_t) * (m ? m : 1));
    double *sigma = calloc(n, sizeof(double)), *dep = calloc(n, sizeof(double)), *inv = malloc(sizeof(double) * n);
    uint8_t *vis = calloc(n, 1);
    uint32_t *cur = malloc(sizeof(uint32_t) * n), *nxt = malloc(sizeof(uint32_t) * n);
    if (!toff || !tids || !sigma || !dep || !inv || !vis || !cur || !nxt) {
        free(toff), free(tids), free(sigma), free(dep), free(inv), free(vis), free(cur), free(nxt);
        return -1;
    }
    ggo_transpose(n, m, off, ids, toff, tids, NULL);
    for (uint32_t v = 0; v < n; v++) level[v] = -1;
    sigma[root] = 1.0;
    vis[root] = 1;
    level[root] = 0;
    cur[0] = root;
    uint32_t nc = 1;
    int round = 0;
    while (nc) {
        round++;
        uint32_t nn = 0;
        for (uint32_t i = 0; i < nc; i++) {
            const uint32_t s = cur[i];
            for (uint32_t e = off[s]; e < off[s + 1]; e++) {
                const uint32_t d = ids[e];
                if (vis[d]) continue;
                const double old = sigma[d];
                sigma[d] += sigma[s];
                if (old == 0.0) nxt[nn++] = d;
            }
        }
        for (uint32_t i = 0; i < nn; i++) {
            vis[nxt[i]] = 1;
            level[nxt[i]] = round;
        }
        uint32_t *t = cur;
        cur = nxt;
        nxt = t;
        nc = nn;
    }
    for (uint32_t v = 0; v < n; v++) inv[v] = 1.0 / sigma[v];
    memset(vis, 0, n);
    for (uint32_t v = 0; v < n; v++)
        if (level[v] == round - 1) {
            vis[v] = 1;
            dep[v] += inv[v];
        }
    for (int r = round - 2; r >= 0; r--) {
        for (uint32_t s = 0; s < n; s++) {
            if (level[s] != r + 1) continue;
            for (uint32_t e = toff[s]; e < toff[s + 1]; e++) {
                const uint32_t d = tids[e];
                if (!vis[d]) dep[d] += dep[s];
            }
        }
        for (uint32_t v = 0; v < n; v++)
            if (level[v] == r) {
                vis[v] = 1;
                dep[v] += inv[v];
            }
    }
    for (uint32_t v = 0; v < n; v++) score[v] = (dep[v] - inv[v]) / inv[v];
    if (sigma_out) memcpy(sigma_out, sigma, sizeof(double) * n);
    free(toff), free(tids), free(sigma), free(dep), free(inv), free(vis), free(cur), free(nxt);
    return round;
}

int ggo_sssp(uint32_t n, const uint32_t *off, const uint32_t *ids, const float *w, uint32_t root, int64_t *dist)
{
    /* app/BellmanFord.C:100-140 with integer edge lengths (intE): rounds over
     * the frontier of vertices whose distance fell in the previous round;
     * after n rounds with a non-empty frontier (a negative cycle) every
     * distance becomes -(INT_MAX/2). */
    if (root >= n) return -1;
    const int64_t inf = 2147483647 / 2;
    uint8_t *vis = calloc(n, 1);
    uint32_t *cur = malloc(sizeof(uint32_t) * n), *nxt = malloc(sizeof(uint32_t) * n);
    if (!vis || !cur || !nxt) {
        free(vis), free(cur), free(nxt);
        return -1;
    }
    for (uint32_t v = 0; v < n; v++) dist[v] = inf;
    dist[root] = 0;
    cur[0] = root;
    uint32_t nc = 1;
    int64_t round = 0;
    while (nc) {
        if (round == (int64_t)n) {
            for (uint32_t v = 0; v < n; v++) dist[v] = -inf;
            break;
        }
        uint32_t nn = 0;
        for (uint32_t i = 0; i < nc; i++) {
            const uint32_t s = cur[i];
            for (uint32_t e = off[s]; e < off[s + 1]; e++) {
                const uint32_t d = ids[e];
                const int64_t nd = dist[s] + (int64_t)w[e];
                if (dist[d] > nd) {
                    dist[d] = nd;
                    if (!vis[d]) {
                        vis[d] = 1;
                        nxt[nn++] = d;
                    }
                }
            }
        }
        for (uint32_t i = 0; i < nn; i++) vis[nxt[i]] = 0;
        uint32_t *t = cur;
        cur = nxt;
        nxt = t;
        nc = nn;
        round++;
    }
    free(vis), free(cur), free(nxt);
    return (int)(round > 2147483647 ? 2147483647 : round);
}

int ggo_prconv(uint32_t n, const uint32_t *off, const uint32_t *ids, double damping, double eps, int max_iter,
               double *p)
{
    if (!n) return 0;
    double *cur = malloc(sizeof(double) * n), *nxt = calloc(n, sizeof(double));
    double *add = malloc(sizeof(double) * n);
    if (!cur || !nxt || !add) {
        free(cur), free(nxt), free(add);
        return -1;
    }
    for (uint32_t v = 0; v < n; v++) cur[v] = 1.0 / (double)n;
    int count = 0;
    double l1 = 2.0;
    while (count < max_iter && l1 > eps) {
        for (uint32_t v = 0; v < n; v++) add[v] = damping / (double)(off[v + 1] - off[v]) * cur[v];
        for (uint32_t s = 0; s < n; s++)
            for (uint32_t e = off[s]; e < off[s + 1]; e++) nxt[ids[e]] += add[s];
        double sum = 0.0;
        for (uint32_t v = 0; v < n; v++) sum += nxt[v];
        const double csum = (1.0 - sum) / (double)n;
        for (uint32_t v = 0; v < n; v++) nxt[v] += csum;
        l1 = 0.0;
        for (uint32_t v = 0; v < n; v++) l1 += fabs(cur[v] - nxt[v]);
        sum = 0.0;
        for (uint32_t v = 0; v < n; v++) sum += nxt[v];
        for (uint32_t v = 0; v < n; v++) nxt[v] /= sum;
        for (uint32_t v = 0; v < n; v++) cur[v] = 0.0;
        count++;
        double *t = cur;
        cur = nxt;
        nxt = t;
    }
    memcpy(p, cur, sizeof(double) * n);
    free(cur), free(nxt), free(add);
    return count;
}

int ggo_prdelta(uint32_t n, const uint32_t *off, const uint32_t *ids, double damping, double eps, double eps2,
                int max_rounds, double *p)
{
    if (!n) return 0;
    double *delta = malloc(sizeof(double) * n), *ngh = calloc(n, sizeof(double));
    uint8_t *act = malloc(n);
    if (!delta || !ngh || !act) {
        free(delta), free(ngh), free(act);
        return -1;
    }
    const double one_over_n = 1.0 / (double)n;
    const double added = (1.0 - damping) * one_over_n;
    for (uint32_t v = 0; v < n; v++) {
        p[v] = 0.0;
        delta[v] = one_over_n;
        act[v] = 1;
    }
    int round = 0;
    while (round < max_rounds) {
        round++;
        for (uint32_t s = 0; s < n; s++) {
            if (!act[s]) continue;
            const double c = delta[s] / (double)(off[s + 1] - off[s]);
            for (uint32_t e = off[s]; e < off[s + 1]; e++) ngh[ids[e]] += c;
        }
        double l1 = 0.0;
        for (uint32_t v = 0; v < n; v++) {
            if (round == 1) {
                delta[v] = damping * (p[v] + ngh[v]) + added - p[v];
                p[v] += delta[v];
                delta[v] -= one_over_n;
            } else {
                delta[v] = ngh[v] * damping;
                p[v] += delta[v];
            }
            act[v] = fabs(delta[v]) > eps2 * p[v];
            l1 += fabs(delta[v]);
        }
        if (l1 < eps) break;
        memset(ngh, 0, sizeof(double) * n);
    }
    free(delta), free(ngh), free(act);
    return round;
}
