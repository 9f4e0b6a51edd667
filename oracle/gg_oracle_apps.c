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
 *
 * Pins (tests/test_oracle_apps.py): textbook Brandes dependencies and path
 * counts on small graphs; textbook Bellman-Ford / Dijkstra distances.
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
