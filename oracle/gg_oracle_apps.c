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
 *            (Dependencies - inv) / inv (222).  A backward step from a
 *            level whose forward out-degree sum is 0 is skipped, as the
 *            reference's edgeMap returns empty() on the Levels' stored
 *            counts (ligra-numa.h:946-947): on a directed graph whose
 *            deepest level has no out-edges, that level's dependencies do
 *            not reach the level above (textbook Brandes would add them).
 *            fType is float in the reference; this restatement computes in
 *            FP64.
 *   ggo_sssp app/BellmanFord.C: distances by relaxation rounds over the
 *            frontier of changed vertices (integer weights).
 *   ggo_prdelta app/PageRankDelta.C:124-176: each round the active
 *            vertices push Delta[s]/outdeg(s) along their out-edges
 *            (PR_F, 27-72) into nghSum; then every vertex updates p and
 *            Delta (PR_Vertex_F_FirstRound in round 1, PR_Vertex_F after,
 *            74-110) and stays active while |Delta| > eps2 * p; stop when
 *            sum |Delta| < eps.
 *   ggo_vebo graphtools/VEBO.C:84-321: vertices sorted by in-degree
 *            (descending, ties by id); degree by degree, vertices go to the
 *            partition with the fewest edges (the lowest index on ties), a
 *            run of (left)/P at once while the edge spread exceeds the
 *            degree; zero-degree vertices then even out the vertex counts;
 *            each partition's ids are given degree group by degree group.
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
 * ranks against a direct solve of the dangling-corrected PageRank system;
 * VEBO against a hand-worked order and its invariants (a permutation;
 * in-degrees non-increasing inside each partition; partition edge counts
 * within the largest degree of each other; vertex counts n/P or n/P + 1
 * when there are enough zero-degree vertices).
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
        /* edgeMap returns empty() when the frontier's stored out-degree sum
         * is 0 (ligra-numa.h:946-947); the Levels keep the sums of the
         * forward graph, so after GA.transpose() the deepest level is not
         * propagated when none of its vertices has a forward out-edge */
        uint64_t fwd_out = 0;
        for (uint32_t s = 0; s < n; s++)
            if (level[s] == r + 1) fwd_out += off[s + 1] - off[s];
        for (uint32_t s = 0; s < n && fwd_out; s++) {
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

typedef struct {
    uint32_t v, d;
} vebo_pair;

static int vebo_cmp(const void *a, const void *b)
{
    const vebo_pair *x = (const vebo_pair *)a, *y = (const vebo_pair *)b;
    if (x->d != y->d) return x->d > y->d ? -1 : 1; /* in-degree, high to low */
    return x->v < y->v ? -1 : x->v > y->v;         /* then by id */
}

int ggo_vebo(uint32_t n, const uint32_t *indeg, int P, uint32_t *new_of_old, uint64_t *starts)
{
    if (!n || P < 1 || (uint32_t)P > n) return -1;
    vebo_pair *chk = malloc(sizeof(vebo_pair) * n);
    for (uint32_t v = 0; v < n; v++) chk[v] = (vebo_pair){v, indeg[v]};
    qsort(chk, n, sizeof(vebo_pair), vebo_cmp);
    /* degree groups, high to low (VEBO.C:95-144); the zero group counts only
     * its first vertex, and only when it is the first group (no edges) */
    uint32_t c = 1, czero = 0;
    for (uint32_t i = 1; i < n; i++) {
        if (chk[i].d != chk[i - 1].d) c++;
        if (chk[i].d == 0) czero++;
    }
    uint32_t *gdeg = calloc(c, sizeof(uint32_t)), *gcnt = calloc(c, sizeof(uint32_t));
    int64_t zgroup = -1;
    {
        uint32_t a = 0;
        for (uint32_t i = 0; i < n; i++) {
            if (i && chk[i].d != chk[i - 1].d) a++;
            if (chk[i].d == 0 && zgroup < 0) zgroup = a;
            if (i == 0 || chk[i].d != 0) {
                gcnt[a]++;
                gdeg[a] = chk[i].d;
            }
        }
    }
    const uint32_t zg = zgroup < 0 ? 0u : (uint32_t)zgroup; /* degreeN[0] */
    int64_t *edges = calloc((size_t)P, sizeof(int64_t));
    int64_t *vsize = calloc((size_t)P, sizeof(int64_t));
    int64_t *per = calloc((size_t)P * c, sizeof(int64_t)); /* perDeg[p][group] */
    /* phase 1 (VEBO.C:179-216) */
    for (uint32_t i = 0; i < c; i++) {
        const int64_t deg = gdeg[i], same = gcnt[i];
        for (int64_t k = 0; k < same;) {
            int64_t mx = edges[0], mn = edges[0];
            int mp = 0;
            for (int j = 1; j < P; j++) {
                if (edges[j] < mn) {
                    mn = edges[j];
                    mp = j;
                }
                if (edges[j] > mx) mx = edges[j];
            }
            int64_t take = 1;
            if (mx - mn > deg) {
                const int64_t remain = (same - k) / P;
                if (remain > 1) take = remain;
            }
            edges[mp] += deg * take;
            per[(size_t)mp * c + i] += take;
            vsize[mp] += take;
            k += take;
        }
    }
    /* phase 2 (VEBO.C:235-284): zero-degree vertices */
    const int64_t avg = n / (uint32_t)P, extra = n % (uint32_t)P, dzero = czero / (uint32_t)P;
    int64_t *add = calloc((size_t)P, sizeof(int64_t));
    int64_t placed = 0;
    if (extra == 0 || dzero != 0) {
        for (int j = 0; j < P; j++)
            if (vsize[j] < avg) add[j] = avg - vsize[j];
        if (extra != 0)
            for (int64_t j = 0; j < extra; j++) add[j]++;
        for (int j = 0; j < P; j++) placed += add[j];
    }
    if ((extra == 0 || dzero != 0) && placed == (int64_t)czero) {
        for (int j = 0; j < P; j++) {
            vsize[j] += add[j];
            per[(size_t)j * c + zg] += add[j];
        }
    } else {
        for (uint32_t z = 0; z < czero; z++) {
            int64_t mnv = vsize[0];
            int mp = 0;
            for (int j = 1; j < P; j++)
                if (vsize[j] < mnv) {
                    mnv = vsize[j];
                    mp = j;
                }
            vsize[mp]++;
            per[(size_t)mp * c + zg]++;
        }
    }
    /* ids: partition starts, then group by group, partition by partition,
     * the next vertices of the sorted order (VEBO.C:289-321) */
    int64_t *start = calloc((size_t)P, sizeof(int64_t));
    for (int j = 1; j < P; j++) start[j] = start[j - 1] + vsize[j - 1];
    int64_t *fill = calloc((size_t)P, sizeof(int64_t));
    uint64_t real = 0;
    int rc = 0;
    for (uint32_t i = 0; i < c && !rc; i++)
        for (int j = 0; j < P && !rc; j++)
            for (int64_t k = 0; k < per[(size_t)j * c + i]; k++) {
                if (real >= n) {
                    rc = -2;
                    break;
                }
                new_of_old[chk[real].v] = (uint32_t)(start[j] + fill[j]);
                fill[j]++;
                real++;
            }
    if (!rc && real != n) rc = -2;
    if (!rc && starts) {
        for (int j = 0; j < P; j++) starts[j] = (uint64_t)start[j];
        starts[P] = n;
    }
    free(chk), free(gdeg), free(gcnt), free(edges), free(vsize), free(per), free(add), free(start), free(fill);
    return rc;
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
