/*
 * gg_oracle.c -- sequential C restatement of GraphGrind's traversal engine.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * Every function cites the reference file:line it follows; the header
 * lists the check that pins each one.
 */
#include "gg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================
 * Synthetic inputs (build decision; the reference has no generator).
 * ==================================================================== */

/* splitmix64 finaliser (S. Vigna, public domain description). */
uint64_t ggo_splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* Quadrant thresholds on the top 32 bits of a draw: a=.57, b=.19, c=.19,
 * d=.05, each scaled by 2^32 and rounded to the nearest integer. */
#define RMAT_TA 2448131359u
#define RMAT_TAB (RMAT_TA + 816043786u)
#define RMAT_TABC (RMAT_TAB + 816043786u)

int ggo_rmat_raw(int scale, uint64_t nedges, uint64_t seed, uint32_t *src,
                 uint32_t *dst)
{
    if (scale < 1 || scale > 31) return -1;
#pragma omp parallel for schedule(static)
    for (uint64_t e = 0; e < nedges; e++) {
        uint32_t s = 0, d = 0;
        for (int l = scale - 1; l >= 0; l--) {
            uint64_t z = ggo_splitmix64((seed << 40) ^ (e << 6) ^ (uint64_t)l);
            uint32_t r = (uint32_t)(z >> 32);
            uint32_t sb, db;
            if (r < RMAT_TA) { sb = 0; db = 0; }
            else if (r < RMAT_TAB) { sb = 0; db = 1; }
            else if (r < RMAT_TABC) { sb = 1; db = 0; }
            else { sb = 1; db = 1; }
            s |= sb << l;
            d |= db << l;
        }
        src[e] = s;
        dst[e] = d;
    }
    return 0;
}

/* LSD radix sort of 64-bit keys, 16-bit digits. */
static void radix_sort_u64(uint64_t *a, uint64_t n)
{
    uint64_t *tmp = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *cnt = (uint64_t *)malloc(65536 * sizeof(uint64_t));
    uint64_t *from = a, *to = tmp;
    for (int pass = 0; pass < 4; pass++) {
        int sh = pass * 16;
        memset(cnt, 0, 65536 * sizeof(uint64_t));
        for (uint64_t i = 0; i < n; i++) cnt[(from[i] >> sh) & 0xffff]++;
        uint64_t run = 0;
        for (int b = 0; b < 65536; b++) { uint64_t c = cnt[b]; cnt[b] = run; run += c; }
        for (uint64_t i = 0; i < n; i++) to[cnt[(from[i] >> sh) & 0xffff]++] = from[i];
        uint64_t *t = from; from = to; to = t;
    }
    /* four passes: result is back in a */
    free(tmp);
    free(cnt);
}

int64_t ggo_build_csr(uint32_t n, uint64_t nraw, uint32_t *src, uint32_t *dst,
                      int symmetric, uint32_t *off, uint32_t *ids)
{
    uint64_t cap = symmetric ? 2 * nraw : nraw;
    uint64_t *key = (uint64_t *)malloc((cap ? cap : 1) * sizeof(uint64_t));
    if (!key) return -1;
    uint64_t k = 0;
    for (uint64_t e = 0; e < nraw; e++) {
        uint32_t s = src[e], d = dst[e];
        if (s >= n || d >= n) { free(key); return -2; }
        if (s == d) continue;
        key[k++] = ((uint64_t)s << 32) | d;
        if (symmetric) key[k++] = ((uint64_t)d << 32) | s;
    }
    radix_sort_u64(key, k);
    uint64_t m = 0;
    memset(off, 0, ((size_t)n + 1) * sizeof(uint32_t));
    for (uint64_t i = 0; i < k; i++) {
        if (i > 0 && key[i] == key[i - 1]) continue;
        uint32_t s = (uint32_t)(key[i] >> 32);
        ids[m++] = (uint32_t)key[i];
        off[s + 1]++;
    }
    for (uint32_t v = 0; v < n; v++) off[v + 1] += off[v];
    free(key);
    if (m > 0xffffffffULL) return -3;
    return (int64_t)m;
}

int ggo_transpose(uint32_t n, uint64_t m, const uint32_t *off,
                  const uint32_t *ids, uint32_t *toff, uint32_t *tids,
                  uint32_t *perm)
{
    memset(toff, 0, ((size_t)n + 1) * sizeof(uint32_t));
    for (uint64_t e = 0; e < m; e++) toff[ids[e] + 1]++;
    for (uint32_t v = 0; v < n; v++) toff[v + 1] += toff[v];
    uint32_t *pos = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
    if (!pos) return -1;
    memcpy(pos, toff, ((size_t)n + 1) * sizeof(uint32_t));
    /* stable: sources visited ascending, so each in-list is sorted by src */
    for (uint32_t s = 0; s < n; s++)
        for (uint32_t e = off[s]; e < off[s + 1]; e++) {
            uint32_t k = pos[ids[e]]++;
            tids[k] = s;
            if (perm) perm[k] = e;
        }
    free(pos);
    return 0;
}

float ggo_weight(uint64_t seed, uint64_t pos)
{
    uint64_t z = ggo_splitmix64((seed << 40) ^ (pos << 6) ^ 63u);
    uint32_t k = (uint32_t)(z >> 40);
    return (float)(k + 1u) * (1.0f / 16777216.0f);
}

double ggo_uniform_pm1(uint64_t seed, uint64_t pos)
{
    uint64_t z = ggo_splitmix64((seed << 40) ^ (pos << 6) ^ 62u);
    uint32_t k = (uint32_t)(z >> 40);
    return (double)k / 8388608.0 - 1.0;
}

/* ======================================================================
 * Partitioning (graph-numa.h:1177-1357, partitioner.h:153-161)
 * ==================================================================== */

int ggo_partition_sizes(uint64_t count, const uint32_t *deg, uint64_t m,
                        int P, int mode, int relabel, uint64_t *sizes)
{
    if (P < 1) return -1;
    for (int p = 0; p < P; p++) sizes[p] = 0;
    if (relabel) {
        /* graph-numa.h:1180-1194 / 1211-1224 / 1261-1273 / 1290-1304:
         * cut wherever the next degree is larger, at most P-1 cuts */
        int64_t counter = 0;
        for (uint64_t j = 0; j < count; j++) {
            sizes[counter]++;
            if (j + 1 < count && deg[j + 1] > deg[j] && counter < P - 1) counter++;
        }
        /* 1224/1304 assert counter+1 == P; with NDEBUG the trailing
         * partitions are simply empty.  1 = "the asserting build aborts" */
        return (mode == 0 && counter + 1 != P) ? 1 : 0;
    }
    if (mode == 1) {
        /* graph-numa.h:1198-1204 / 1277-1283 */
        uint64_t chunk = 0;
        for (int p = 0; p < P - 1; p++) { sizes[p] = count / (uint64_t)P; chunk += sizes[p]; }
        sizes[P - 1] = count - chunk;
        return 0;
    }
    /* graph-numa.h:1328-1339 (and 1242-1252): greedy edge balance */
    int64_t avg = (int64_t)(m / (uint64_t)P);
    int64_t *edges = (int64_t *)calloc((size_t)P, sizeof(int64_t));
    if (!edges) return -1;
    int64_t counter = 0;
    int rc = 0;
    for (uint64_t i = 0; i < count; i++) {
        edges[counter] += deg[i];
        sizes[counter]++;
        /* degrees[i+1] is read past the end at i = count-1 in the
         * reference; the restatement reads 0 there */
        int64_t nxt = (i + 1 < count) ? (int64_t)deg[i + 1] : 0;
        if (edges[counter] < avg && (double)(nxt + edges[counter]) > 1.1 * (double)avg)
            counter++;
        if (counter >= P) { rc = -2; break; } /* reference overruns its arrays */
        if (edges[counter] >= avg && counter < P - 1) counter++;
    }
    /* 1253/1353 assert counter+1 == P.  When the greedy runs out of edges
     * early the sizes are final and the trailing partitions stay empty (the
     * -DNDEBUG build); report 1 so callers know the asserting build aborts. */
    if (rc == 0 && counter + 1 != P) rc = 1;
    free(edges);
    return rc;
}

uint64_t ggo_round_up_pow2(uint64_t v)
{
    /* graph-numa.h:301-313 */
    --v;
    for (unsigned sh = 1; sh < 64; sh <<= 1) v |= v >> sh;
    return ++v;
}

static void hil_rot(int64_t s, int64_t *x, int64_t *y, int64_t rx, int64_t ry)
{
    /* graph-numa.h:403-414 */
    if (ry == 0) {
        if (rx == 1) { *x = s - 1 - *x; *y = s - 1 - *y; }
        int64_t t = *x; *x = *y; *y = t;
    }
}

int64_t ggo_hilbert_xy2d(int64_t side, int64_t x, int64_t y)
{
    /* graph-numa.h:375-386, signed 64-bit arithmetic as intT=long */
    int64_t d = 0;
    for (int64_t s = side / 2; s > 0; s /= 2) {
        int64_t rx = (x & s) > 0;
        int64_t ry = (y & s) > 0;
        d += s * s * ((3 * rx) ^ ry);
        hil_rot(s, &x, &y, rx, ry);
    }
    return d;
}

/* ======================================================================
 * Graph with partitions (graph-numa.h:626-685)
 * ==================================================================== */

struct ggo_graph {
    uint32_t n;
    uint64_t m;
    int symmetric;
    int P;
    int part_vertex;
    uint32_t *off, *ids;   /* CSR                          */
    uint32_t *toff, *tids; /* CSC, in-lists ascending src  */
    double *w, *tw;        /* weights (CSR / CSC order) or NULL */
    uint64_t *coo_starts;  /* P+1 vertex ids               */
    uint64_t *coo_eoff;    /* P+1 offsets into coo arrays  */
    uint32_t *coo_src, *coo_dst;
    double *coo_w;
    uint64_t ncomp;        /* compressed in-vertices       */
    uint32_t *comp;        /* ids with in-degree > 0       */
    uint64_t *csc_starts;  /* P+1 positions in comp[]      */
};

static int g_last_error = 0;
int ggo_graph_error(void) { return g_last_error; }

typedef struct { int64_t key; uint32_t src, dst; double w; uint64_t ord; } hedge;

static int hedge_cmp(const void *a, const void *b)
{
    const hedge *x = (const hedge *)a, *y = (const hedge *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    /* ties: keep in-list order (std::sort leaves them unspecified) */
    return x->ord < y->ord ? -1 : (x->ord > y->ord);
}

void ggo_graph_free(ggo_graph *g)
{
    if (!g) return;
    free(g->off); free(g->ids); free(g->toff); free(g->tids);
    free(g->w); free(g->tw);
    free(g->coo_starts); free(g->coo_eoff); free(g->coo_src); free(g->coo_dst);
    free(g->coo_w); free(g->comp); free(g->csc_starts);
    free(g);
}

ggo_graph *ggo_graph_new(uint32_t n, uint64_t m, const uint32_t *off,
                         const uint32_t *ids, const double *w, int symmetric,
                         int P, int part_vertex, int hilbert, int relabel)
{
    g_last_error = 0;
    if (n == 0 || P < 1 || off[n] != m) { g_last_error = -1; return NULL; }
    ggo_graph *g = (ggo_graph *)calloc(1, sizeof(ggo_graph));
    g->n = n; g->m = m; g->symmetric = symmetric; g->P = P; g->part_vertex = part_vertex;
    size_t mb = (size_t)(m ? m : 1);
    g->off = (uint32_t *)malloc(((size_t)n + 1) * 4);
    g->ids = (uint32_t *)malloc(mb * 4);
    g->toff = (uint32_t *)malloc(((size_t)n + 1) * 4);
    g->tids = (uint32_t *)malloc(mb * 4);
    memcpy(g->off, off, ((size_t)n + 1) * 4);
    memcpy(g->ids, ids, (size_t)m * 4);
    uint32_t *perm = (uint32_t *)malloc(mb * 4);
    ggo_transpose(n, m, g->off, g->ids, g->toff, g->tids, perm);
    if (w) {
        g->w = (double *)malloc(mb * 8);
        g->tw = (double *)malloc(mb * 8);
        memcpy(g->w, w, (size_t)m * 8);
        for (uint64_t k = 0; k < m; k++) g->tw[k] = w[perm[k]];
    }
    free(perm);

    /* in-degrees: '-P dest' partitions on in-degree (ligra-numa.h:1110-1111) */
    uint32_t *indeg = (uint32_t *)malloc((size_t)n * 4);
    for (uint32_t v = 0; v < n; v++) indeg[v] = g->toff[v + 1] - g->toff[v];

    /* COO partitioner sizes + starts (graph-numa.h:651-655) */
    uint64_t *sizes = (uint64_t *)malloc((size_t)P * 8);
    int rc = ggo_partition_sizes(n, indeg, m, P, part_vertex ? 1 : 0, relabel, sizes);
    if (rc < 0) { g_last_error = rc; free(sizes); free(indeg); ggo_graph_free(g); return NULL; }
    g_last_error = rc;
    g->coo_starts = (uint64_t *)malloc(((size_t)P + 1) * 8);
    g->coo_starts[0] = 0;
    for (int p = 0; p < P; p++) g->coo_starts[p + 1] = g->coo_starts[p] + sizes[p];

    /* COO lists (graph-numa.h:661-675, 1116-1143), edge mode only */
    g->coo_eoff = (uint64_t *)calloc((size_t)P + 1, 8);
    if (!part_vertex) {
        g->coo_src = (uint32_t *)malloc(mb * 4);
        g->coo_dst = (uint32_t *)malloc(mb * 4);
        if (w) g->coo_w = (double *)malloc(mb * 8);
        uint64_t k = 0;
        for (int p = 0; p < P; p++) {
            g->coo_eoff[p] = k;
            uint64_t lo = g->coo_starts[p], hi = g->coo_starts[p + 1];
            uint64_t cnt = (uint64_t)g->toff[hi] - g->toff[lo];
            hedge *hb = (hedge *)malloc((size_t)(cnt ? cnt : 1) * sizeof(hedge));
            uint64_t j = 0;
            for (uint64_t v = lo; v < hi; v++)
                for (uint32_t e = g->toff[v]; e < g->toff[v + 1]; e++) {
                    hb[j].src = g->tids[e];
                    hb[j].dst = (uint32_t)v;
                    hb[j].w = w ? g->tw[e] : 1.0;
                    hb[j].ord = j;
                    hb[j].key = 0;
                    j++;
                }
            if (hilbert) {
                /* HilbertEdgeSort(num_edges): side = roundUpPow2(count) */
                int64_t side = (int64_t)ggo_round_up_pow2(cnt);
                for (uint64_t t = 0; t < cnt; t++)
                    hb[t].key = ggo_hilbert_xy2d(side, hb[t].src, hb[t].dst);
            } else {
                /* CSRDestSort (graph-numa.h:338-350): by (src, dst) */
                for (uint64_t t = 0; t < cnt; t++)
                    hb[t].key = ((int64_t)hb[t].src << 32) | hb[t].dst;
            }
            qsort(hb, (size_t)cnt, sizeof(hedge), hedge_cmp);
            for (uint64_t t = 0; t < cnt; t++) {
                g->coo_src[k] = hb[t].src;
                g->coo_dst[k] = hb[t].dst;
                if (w) g->coo_w[k] = hb[t].w;
                k++;
            }
            free(hb);
        }
        g->coo_eoff[P] = k;
    }

    /* compressed CSC + CSC partitions (graph-numa.h:678-683, 1079-1114) */
    uint64_t nc = 0;
    for (uint32_t v = 0; v < n; v++) nc += indeg[v] != 0;
    g->ncomp = nc;
    g->comp = (uint32_t *)malloc((size_t)(nc ? nc : 1) * 4);
    uint32_t *cdeg = (uint32_t *)malloc((size_t)(nc ? nc : 1) * 4);
    uint64_t c = 0;
    for (uint32_t v = 0; v < n; v++)
        if (indeg[v]) { g->comp[c] = v; cdeg[c] = indeg[v]; c++; }
    rc = ggo_partition_sizes(nc, cdeg, m, P, part_vertex ? 1 : 0, relabel, sizes);
    free(cdeg);
    free(indeg);
    if (rc < 0) { g_last_error = rc - 10; free(sizes); ggo_graph_free(g); return NULL; }
    if (rc > 0) g_last_error |= 2;
    g->csc_starts = (uint64_t *)malloc(((size_t)P + 1) * 8);
    g->csc_starts[0] = 0;
    for (int p = 0; p < P; p++) g->csc_starts[p + 1] = g->csc_starts[p] + sizes[p];
    free(sizes);
    return g;
}

int ggo_graph_starts(const ggo_graph *g, int which, uint64_t *starts)
{
    const uint64_t *src = which == 0 ? g->coo_starts : g->csc_starts;
    memcpy(starts, src, ((size_t)g->P + 1) * 8);
    return 0;
}

uint64_t ggo_graph_ncompressed(const ggo_graph *g) { return g->ncomp; }

int64_t ggo_graph_coo_count(const ggo_graph *g, int p)
{
    if (p < 0 || p >= g->P || !g->coo_src) return -1;
    return (int64_t)(g->coo_eoff[p + 1] - g->coo_eoff[p]);
}

int64_t ggo_graph_coo(const ggo_graph *g, int p, uint32_t *src, uint32_t *dst)
{
    int64_t cnt = ggo_graph_coo_count(g, p);
    if (cnt < 0) return cnt;
    memcpy(src, g->coo_src + g->coo_eoff[p], (size_t)cnt * 4);
    memcpy(dst, g->coo_dst + g->coo_eoff[p], (size_t)cnt * 4);
    return cnt;
}

/* ======================================================================
 * Frontier (partitioned_vertices, ligra-numa.h:91-272)
 * ==================================================================== */

typedef struct {
    int64_t n;       /* numVertices (0 for empty())       */
    int64_t d_m;
    uint8_t *d;      /* dense flags or NULL               */
    int has_dense;
    int64_t *s;      /* sparse list or NULL               */
    int64_t num_out_edges;
    int bit;
} fr_t;

static void fr_free(fr_t *f)
{
    free(f->d); free(f->s);
    f->d = NULL; f->s = NULL;
}

static uint32_t outdeg(const ggo_graph *g, int64_t v) { return g->off[v + 1] - g->off[v]; }

/* ---- edge functors of the four apps under app/, run sequentially ---- */
enum { OP_PR = 0, OP_BFS = 1, OP_CC = 2, OP_SPMV = 3 };
typedef struct {
    int kind;
    double *p_curr, *p_next;   /* PR / SPMV */
    int64_t *a;                /* BFS Parents, CC IDs */
    int64_t *b;                /* CC prevIDs */
} op_t;

static int op_cond(const op_t *f, int64_t d)
{
    if (f->kind == OP_BFS) return f->a[d] == -1; /* app/BFS.C:70-74 */
    return 1;                                    /* cond_true       */
}

/* update(s,d[,w]): app/PageRank.C:39-43, BFS.C:35-43, Components.C:38-50,
 * SPMV.C:17-21 */
static int op_update(const ggo_graph *g, const op_t *f, int64_t s, int64_t d, double w)
{
    switch (f->kind) {
    case OP_PR: f->p_next[d] += f->p_curr[s] / (double)outdeg(g, s); return 1;
    case OP_BFS:
        if (f->a[d] == -1) { f->a[d] = s; return 1; }
        return 0;
    case OP_CC: {
        int64_t orig = f->a[d];
        if (f->a[s] < orig) {
            f->a[d] = orig < f->a[s] ? orig : f->a[s];
            if (orig == f->b[d]) return 1;
        }
        return 0;
    }
    default: f->p_next[d] += f->p_curr[s] * w; return 1;
    }
}

/* updateAtomic(s,d[,w]) run sequentially -- PageRank.C:44-48 (writeAdd),
 * BFS.C:44-48 (CAS), Components.C:51-56 (writeMin), SPMV.C:22-27 */
static int op_update_atomic(const ggo_graph *g, const op_t *f, int64_t s, int64_t d, double w)
{
    switch (f->kind) {
    case OP_PR: f->p_next[d] += f->p_curr[s] / (double)outdeg(g, s); return 1;
    case OP_BFS:
        if (f->a[d] == -1) { f->a[d] = s; return 1; }
        return 0;
    case OP_CC: {
        int64_t orig = f->a[d];
        int r = 0;
        if (f->a[d] > f->a[s]) { f->a[d] = f->a[s]; r = 1; } /* writeMin utils.h:445 */
        return r && orig == f->b[d];
    }
    default: f->p_next[d] += f->p_curr[s] * w; return 1;
    }
}

static int op_use_cache(const op_t *f) { return f->kind != OP_BFS; }

/* cache protocol: PageRank.C:50-64, Components.C:57-79, SPMV.C:29-42. */
typedef struct { double v; int64_t id, prev; } cache_t;
static void op_create_cache(const op_t *f, cache_t *c, int64_t d)
{
    if (f->kind == OP_CC) { c->id = f->a[d]; c->prev = f->b[d]; }
    else c->v = f->p_next[d];
}
static int op_update_cache(const ggo_graph *g, const op_t *f, cache_t *c, int64_t s, double w)
{
    switch (f->kind) {
    case OP_PR: c->v += f->p_curr[s] / (double)outdeg(g, s); return 1;
    case OP_CC: {
        int64_t orig = c->id;
        if (f->b[s] < orig) {
            c->id = orig < f->b[s] ? orig : f->b[s];
            if (orig == c->prev) return 1;
        }
        return 0;
    }
    default: c->v += f->p_curr[s] * w; return 1;
    }
}
static void op_commit_cache(const op_t *f, cache_t *c, int64_t d)
{
    if (f->kind == OP_CC) f->a[d] = c->id;
    else f->p_next[d] = c->v;
}

/* ---- frontier conversions ---- */
static void to_dense(const ggo_graph *g, fr_t *F)
{
    /* ligra-numa.h:228-244 */
    if (!F->d) {
        F->d = (uint8_t *)calloc((size_t)g->n, 1);
        for (int64_t i = 0; i < F->d_m; i++) F->d[F->s[i]] = 1;
    }
    F->has_dense = 1;
}

static int to_sparse(const ggo_graph *g, fr_t *F)
{
    /* ligra-numa.h:246-260 (packIndex) */
    if (!F->s) {
        int64_t cnt = 0;
        for (int64_t i = 0; i < (int64_t)g->n; i++) cnt += F->d[i] != 0;
        if (cnt != F->d_m) return -1; /* "bad stored value of m" + abort */
        F->s = (int64_t *)malloc((size_t)(cnt ? cnt : 1) * 8);
        int64_t k = 0;
        for (int64_t i = 0; i < (int64_t)g->n; i++) if (F->d[i]) F->s[k++] = i;
    }
    F->has_dense = 0;
    return 0;
}

/* ---- the three traversals ---- */
static void dense_csc(const ggo_graph *g, const op_t *f, const uint8_t *vert, int bit,
                      uint8_t *next, uint64_t lo, uint64_t hi)
{
    /* ligra-numa.h:589-639 (bit) and 676-726 (frontier), dest partitioning,
     * sequential loop (PART96), parallel=0 so the inner loop never splits */
    for (uint64_t i = lo; i < hi; i++) {
        int64_t id = g->comp[i];
        if (!op_cond(f, id)) continue;
        uint32_t b = g->toff[id], e = g->toff[id + 1];
        if (op_use_cache(f)) {
            cache_t c = {0, 0, 0};
            op_create_cache(f, &c, id);
            for (uint32_t j = b; j < e; j++) {
                int64_t ngh = g->tids[j];
                double w = g->tw ? g->tw[j] : 1.0;
                if ((bit || vert[ngh]) && op_update_cache(g, f, &c, ngh, w)) next[id] = 1;
                if (!op_cond(f, id)) break;
            }
            op_commit_cache(f, &c, id);
        } else {
            for (uint32_t j = b; j < e; j++) {
                int64_t ngh = g->tids[j];
                double w = g->tw ? g->tw[j] : 1.0;
                if ((bit || vert[ngh]) && op_update(g, f, ngh, id, w)) next[id] = 1;
                if (!op_cond(f, id)) break;
            }
        }
    }
}

static void dense_coo(const ggo_graph *g, const op_t *f, const uint8_t *vert, int bit,
                      uint8_t *next, int p)
{
    /* ligra-numa.h:804-866 with PART96: sequential, non-atomic edgeOpIn */
    for (uint64_t k = g->coo_eoff[p]; k < g->coo_eoff[p + 1]; k++) {
        int64_t s = g->coo_src[k], d = g->coo_dst[k];
        double w = g->coo_w ? g->coo_w[k] : 1.0;
        if (!op_cond(f, d)) continue;
        if ((bit || vert[s]) && op_update(g, f, s, d, w)) next[d] = 1;
    }
}

static int64_t sparse_push(const ggo_graph *g, const op_t *f, const fr_t *F, int64_t **out)
{
    /* ligra-numa.h:868-916: degrees, plusScan, per-edge slot, filter */
    int64_t mcount = F->d_m;
    uint64_t *offs = (uint64_t *)malloc((size_t)(mcount + 1) * 8);
    uint64_t tot = 0;
    for (int64_t i = 0; i < mcount; i++) { offs[i] = tot; tot += outdeg(g, F->s[i]); }
    int64_t *slots = (int64_t *)malloc((size_t)(tot ? tot : 1) * 8);
    for (int64_t k = 0; k < mcount; k++) {
        int64_t v = F->s[k];
        uint64_t o = offs[k];
        uint32_t b = g->off[v], e = g->off[v + 1];
        for (uint32_t j = b; j < e; j++) {
            int64_t ngh = g->ids[j];
            double w = g->w ? g->w[j] : 1.0;
            /* edgeOpFwd (379-390): cond && updateAtomic */
            int ok = op_cond(f, ngh) && op_update_atomic(g, f, v, ngh, w);
            slots[o + (j - b)] = ok ? ngh : -1;
        }
    }
    int64_t cnt = 0;
    for (uint64_t i = 0; i < tot; i++) if (slots[i] >= 0) slots[cnt++] = slots[i];
    free(offs);
    *out = slots;
    return cnt;
}

/* edgeMap (ligra-numa.h:919-1019).  F is not modified. */
static fr_t edge_map(const ggo_graph *g, const fr_t *Fin, const op_t *f,
                     int64_t threshold, ggo_trace *tr, int *err)
{
    fr_t out;
    memset(&out, 0, sizeof(out));
    *err = 0;
    if (threshold == -1) threshold = (int64_t)(g->m / 20);
    int64_t dense_thr = (int64_t)(g->m / 2);
    if (Fin->n != (int64_t)g->n) { *err = -4; return out; } /* "Sizes Don't match" */
    int64_t mcount = Fin->d_m, tod = Fin->num_out_edges;
    if (tr) { tr->in_active = mcount; tr->in_out_edges = tod; tr->path = -1; }
    if (tod == 0) return out; /* partitioned_vertices::empty() */
    fr_t F = *Fin;            /* edgeMap takes its frontier by value */
    int own_d = 0, own_s = 0;
    if (mcount + tod > threshold) {
        if (!F.d) own_d = 1;
        to_dense(g, &F);
        out.n = g->n;
        out.d = (uint8_t *)calloc((size_t)g->n, 1);
        out.has_dense = 1;
        /* Partitions run in parallel, as the reference's map_partitionL
         * does.  Each partition owns a destination range and sums its
         * edges in order, so values do not depend on the thread count.
         * The exception is CC's COO path, whose update reads IDs[s] of
         * other partitions while they are written (Gauss-Seidel): it stays
         * sequential so its round count is reproducible. */
        const int par = !(f->kind == OP_CC);
        if (mcount + tod > dense_thr && !g->part_vertex) {
#pragma omp parallel for schedule(dynamic, 1) if (par)
            for (int p = 0; p < g->P; p++) dense_coo(g, f, F.d, F.bit, out.d, p);
            if (tr) tr->path = GGO_PATH_COO;
        } else {
#pragma omp parallel for schedule(dynamic, 1)
            for (int p = 0; p < g->P; p++)
                dense_csc(g, f, F.d, F.bit, out.d, g->csc_starts[p], g->csc_starts[p + 1]);
            if (tr) tr->path = GGO_PATH_CSC;
        }
        int64_t dm = 0, oe = 0;
#pragma omp parallel for reduction(+ : dm, oe) schedule(static)
        for (int64_t i = 0; i < (int64_t)g->n; i++)
            if (out.d[i]) { dm++; oe += outdeg(g, i); }
        out.d_m = dm;
        out.num_out_edges = oe;
    } else {
        if (!F.s) own_s = 1;
        if (to_sparse(g, &F) < 0) { *err = -5; if (own_s) free(F.s); return out; }
        out.n = g->n;
        out.d_m = sparse_push(g, f, &F, &out.s);
        int64_t oe = 0;
        for (int64_t i = 0; i < out.d_m; i++) oe += outdeg(g, out.s[i]);
        out.num_out_edges = out.d_m ? oe : 0;
        if (tr) tr->path = GGO_PATH_SPARSE;
    }
    if (own_d) free(F.d);
    if (own_s) free(F.s);
    if (tr) { tr->out_active = out.d_m; tr->out_out_edges = out.num_out_edges; }
    return out;
}

/* vertexMap (ligra-numa.h:1023-1049) over F; kind selects the functor. */
enum { VM_PR = 0, VM_RESET = 1, VM_CC_COPY = 2 };
static void vertex_apply(int kind, int64_t i, double *a, double *b, int64_t *ia,
                         int64_t *ib, double damping, double added)
{
    (void)a;
    switch (kind) {
    case VM_PR: b[i] = damping * b[i] + added; break; /* PageRank.C:82-86 */
    case VM_RESET: a[i] = 0.0; break;                  /* PageRank.C:95-99 */
    default: ib[i] = ia[i]; break;                     /* Components.C:94-98 */
    }
}
static void vertex_map(const ggo_graph *g, const fr_t *F, int kind, double *a, double *b,
                       int64_t *ia, int64_t *ib, double damping, double added)
{
    if (F->has_dense) {
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < (int64_t)g->n; j++)
            if (F->bit || F->d[j]) vertex_apply(kind, j, a, b, ia, ib, damping, added);
    } else {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < F->d_m; i++)
            vertex_apply(kind, F->s[i], a, b, ia, ib, damping, added);
    }
}

static fr_t fr_bits(const ggo_graph *g)
{
    /* partitioned_vertices::bits (ligra-numa.h:169-180) */
    fr_t F;
    memset(&F, 0, sizeof(F));
    F.n = g->n;
    F.d = (uint8_t *)malloc((size_t)g->n);
    memset(F.d, 1, (size_t)g->n); /* never read while bit is set */
    F.bit = 1;
    F.has_dense = 1;
    F.d_m = g->n;
    F.num_out_edges = (int64_t)g->m;
    return F;
}

/* sequence::plusReduce (utils.h:159-195) */
static double reduce_serial(const double *a, uint64_t s, uint64_t e)
{
    double r = a[s];
    for (uint64_t j = s + 1; j < e; j++) r = r + a[j];
    return r;
}
double ggo_plus_reduce(const double *a, uint64_t n)
{
    uint64_t l = n ? 1 + (n - 1) / 1024 : 1;
    if (l <= 1) return reduce_serial(a, 0, n);
    double *sums = (double *)malloc((size_t)l * 8);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < l; i++) {
        uint64_t s = i * 1024, e = s + 1024 < n ? s + 1024 : n;
        sums[i] = reduce_serial(a, s, e);
    }
    double r = ggo_plus_reduce(sums, l);
    free(sums);
    return r;
}

/* ======================================================================
 * Apps
 * ==================================================================== */

int ggo_pagerank(const ggo_graph *g, int max_iter, double *out, int *iters,
                 ggo_trace *trace)
{
    return ggo_pagerank_ex(g, max_iter, 0, out, iters, trace);
}

int ggo_pagerank_ex(const ggo_graph *g, int max_iter, int bit_frontiers, double *out, int *iters,
                    ggo_trace *trace)
{
    /* app/PageRank.C:102-151; bit_frontiers: app/PageRankBit.C:102-150,
     * whose loop marks each edgeMap output all-active (output.bit = true,
     * 145) while its stored d_m / num_out_edges still steer the branch */
    int64_t n = g->n;
    const double damping = 0.85, epsilon = 0.0000001;
    double one_over_n = 1 / (double)n;
    double *pc = (double *)malloc((size_t)n * 8), *pn = (double *)malloc((size_t)n * 8);
    for (int64_t j = 0; j < n; j++) { pc[j] = one_over_n; pn[j] = 0; }
    fr_t F = fr_bits(g);
    int count = 0, err = 0, broke = 0;
    while (count < max_iter) {
        count++;
        op_t f = { OP_PR, pc, pn, NULL, NULL };
        fr_t output = edge_map(g, &F, &f, (int64_t)(g->m / 20), trace ? &trace[count - 1] : NULL, &err);
        if (err) { fr_free(&F); free(pc); free(pn); return err; }
        vertex_map(g, &F, VM_PR, pc, pn, NULL, NULL, damping, (1 - damping) * (1 / (double)n));
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < n; j++) pc[j] = fabs(pc[j] - pn[j]); /* 136 */
        double L1 = ggo_plus_reduce(pc, (uint64_t)n);
        if (L1 < epsilon) { broke = 1; fr_free(&output); break; }
        vertex_map(g, &F, VM_RESET, pc, NULL, NULL, NULL, 0, 0);
        double *t = pc; pc = pn; pn = t;
        fr_free(&F);
        F = output;
        if (bit_frontiers && F.n) F.bit = 1;
    }
    /* the x of the last executed iteration */
    memcpy(out, broke ? pn : pc, (size_t)n * 8);
    if (iters) *iters = count;
    fr_free(&F);
    free(pc); free(pn);
    return 0;
}

int ggo_bfs(const ggo_graph *g, uint32_t root, int32_t *parents, ggo_trace *trace, int max_trace)
{
    /* app/BFS.C:77-101 */
    int64_t n = g->n;
    if (root >= g->n) return -6;
    int64_t *P = (int64_t *)malloc((size_t)n * 8);
    for (int64_t j = 0; j < n; j++) P[j] = -1;
    P[root] = root;
    fr_t F;
    memset(&F, 0, sizeof(F));
    F.n = n; F.d_m = 1; F.s = (int64_t *)malloc(8); F.s[0] = root;
    F.num_out_edges = outdeg(g, root);
    int levels = 0, err = 0;
    while (F.d_m != 0) {
        op_t f = { OP_BFS, NULL, NULL, P, NULL };
        fr_t output = edge_map(g, &F, &f, (int64_t)(g->m / 20),
                               (trace && levels < max_trace) ? &trace[levels] : NULL, &err);
        if (err) { fr_free(&F); free(P); return err; }
        levels++;
        fr_free(&F);
        F = output;
    }
    fr_free(&F);
    for (int64_t j = 0; j < n; j++) parents[j] = (int32_t)P[j];
    free(P);
    return levels;
}

int ggo_cc(const ggo_graph *g, uint32_t *ids, ggo_trace *trace, int max_trace)
{
    /* app/Components.C:101-126 */
    int64_t n = g->n;
    int64_t *I = (int64_t *)malloc((size_t)n * 8), *Pv = (int64_t *)malloc((size_t)n * 8);
    for (int64_t j = 0; j < n; j++) { I[j] = j; Pv[j] = j; }
    fr_t F = fr_bits(g);
    int rounds = 0, err = 0;
    while (F.d_m != 0) {
        vertex_map(g, &F, VM_CC_COPY, NULL, NULL, I, Pv, 0, 0);
        op_t f = { OP_CC, NULL, NULL, I, Pv };
        fr_t output = edge_map(g, &F, &f, (int64_t)(g->m / 20),
                               (trace && rounds < max_trace) ? &trace[rounds] : NULL, &err);
        if (err) { fr_free(&F); free(I); free(Pv); return err; }
        rounds++;
        fr_free(&F);
        F = output;
    }
    fr_free(&F);
    for (int64_t j = 0; j < n; j++) ids[j] = (uint32_t)I[j];
    free(I); free(Pv);
    return rounds;
}

int ggo_spmv(const ggo_graph *g, const double *x, double *y, ggo_trace *trace)
{
    /* app/SPMV.C:64-92 */
    int64_t n = g->n;
    double one_over_n = 1 / (double)n;
    double *pc = (double *)malloc((size_t)n * 8), *pn = (double *)malloc((size_t)n * 8);
    for (int64_t j = 0; j < n; j++) { pc[j] = x ? x[j] : one_over_n; pn[j] = 0; }
    fr_t F = fr_bits(g);
    op_t f = { OP_SPMV, pc, pn, NULL, NULL };
    int err = 0;
    fr_t output = edge_map(g, &F, &f, (int64_t)(g->m / 20), trace, &err);
    if (!err) {
        vertex_map(g, &F, VM_RESET, pc, NULL, NULL, NULL, 0, 0);
        memcpy(y, pn, (size_t)n * 8); /* after swap(p_curr,p_next): y in p_curr */
    }
    fr_free(&output);
    fr_free(&F);
    free(pc); free(pn);
    return err;
}

int ggo_edgemap_once(const ggo_graph *g, int op, const uint8_t *in_dense,
                     const uint32_t *in_sparse, int64_t in_count, int bit,
                     int64_t threshold, double *a_f64, double *b_f64,
                     int64_t *a_i64, int64_t *b_i64, uint8_t *out_dense,
                     int64_t *out_dm, int64_t *out_edges, int *path)
{
    fr_t F;
    memset(&F, 0, sizeof(F));
    F.n = g->n;
    F.bit = bit;
    int64_t oe = 0;
    if (in_dense) {
        F.d = (uint8_t *)malloc((size_t)g->n);
        memcpy(F.d, in_dense, (size_t)g->n);
        F.has_dense = 1;
        int64_t c = 0;
        for (int64_t i = 0; i < (int64_t)g->n; i++) if (F.d[i] || bit) { c++; oe += outdeg(g, i); }
        F.d_m = c;
    } else {
        F.s = (int64_t *)malloc((size_t)(in_count ? in_count : 1) * 8);
        for (int64_t i = 0; i < in_count; i++) { F.s[i] = in_sparse[i]; oe += outdeg(g, in_sparse[i]); }
        F.d_m = in_count;
    }
    F.num_out_edges = oe;
    op_t f = { op, a_f64, b_f64, a_i64, b_i64 };
    ggo_trace tr;
    int err = 0;
    fr_t out = edge_map(g, &F, &f, threshold, &tr, &err);
    if (!err) {
        memset(out_dense, 0, (size_t)g->n);
        if (out.d) memcpy(out_dense, out.d, (size_t)g->n);
        else for (int64_t i = 0; i < out.d_m; i++) out_dense[out.s[i]] = 1;
        *out_dm = out.d_m;
        *out_edges = out.num_out_edges;
        *path = tr.path;
    }
    fr_free(&out);
    fr_free(&F);
    return err;
}
