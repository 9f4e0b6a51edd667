/*
 * gg_oracle.h -- CPU restatement of GraphGrind's traversal engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product in graphgrind_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in the product links it.
 *
 * It restates, sequentially and in plain C, what the reference computes on
 * the hot path (citations are to the reference repository):
 *   - edgeMap's three-way switch              graphgrind/ligra-numa.h:919-1019
 *   - sparse push  (edgeMapSparseWithG)       graphgrind/ligra-numa.h:868-916
 *   - dense CSC pull (edgeMapDenseCSC, dest)  graphgrind/ligra-numa.h:549-729
 *   - dense COO (edgeMapDense, PART96 order)  graphgrind/ligra-numa.h:803-866
 *   - vertexMap                               graphgrind/ligra-numa.h:1023-1049
 *   - frontier stats (GoutDegree + reduce)    graphgrind/ligra-numa.h:275-302,985-988,1016
 *   - partition sizing (edge / vertex)        graphgrind/graph-numa.h:1177-1357
 *   - COO partition build + Hilbert order     graphgrind/graph-numa.h:353-415,482-490,1116-1143
 *   - the four apps' Compute loops            app/PageRank.C, app/BFS.C,
 *                                             app/Components.C, app/SPMV.C
 * plus the synthetic-input generator shared by every test and benchmark
 * (R-MAT via splitmix64, a build decision: the reference ships no generator).
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   rmat/csr/transpose   numpy re-derivation of the same spec, scipy transpose
 *   partition sizes      oracle/_ref (the reference's own partitioned_graph)
 *   hilbert key          Wikipedia xy2d on in-range points; oracle/_ref COO lists
 *   pagerank             oracle/_ref dumps; closed-form iteration 1; dense-matrix
 *                        simulation of the frontier-gated recurrence on tiny graphs
 *   bfs                  textbook BFS levels; parent validity; oracle/_ref levels
 *   cc                   union-find minimum labels; oracle/_ref labels
 *   spmv                 scipy.sparse matvec; oracle/_ref dumps
 * Error convention: functions return >= 0 on success, negative on error.
 */
#ifndef GG_ORACLE_H
#define GG_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- synthetic inputs ------------------------------------------------- */
uint64_t ggo_splitmix64(uint64_t x);
/* Raw R-MAT edge e (0 <= e < ef*2^scale): 'scale' quadrant draws, one
 * splitmix64 per level keyed by (seed<<40)^(e<<6)^level. */
int ggo_rmat_raw(int scale, uint64_t nedges, uint64_t seed,
                 uint32_t *src, uint32_t *dst);
/* Sort (src,dst), drop self loops, optionally add reverse edges, dedupe.
 * src/dst must have room for 2*nraw entries when symmetric.  Writes CSR
 * offsets (n+1) and ids (m) and returns m (or <0 on error). */
int64_t ggo_build_csr(uint32_t n, uint64_t nraw, uint32_t *src, uint32_t *dst,
                      int symmetric, uint32_t *off, uint32_t *ids);
/* In-edge (CSC) twin, in-neighbours ascending by source (IO.h:43-51,303).
 * perm[k] = CSR position of CSC edge k (may be NULL). */
int ggo_transpose(uint32_t n, uint64_t m, const uint32_t *off,
                  const uint32_t *ids, uint32_t *toff, uint32_t *tids,
                  uint32_t *perm);
/* Per-position uniform draws used for weights and dense vectors. */
float ggo_weight(uint64_t seed, uint64_t pos);      /* (k+1)/2^24, k 24-bit */
double ggo_uniform_pm1(uint64_t seed, uint64_t pos); /* k/2^23 - 1           */

/* ---- partitioning ------------------------------------------------------ */
/* Sizes of P partitions over 'count' items with degree deg[] (count items)
 * and edge total m.  mode 0: partitionByDegree / cscpartitionByDegree
 * (graph-numa.h:1208-1256,1287-1357); mode 1: by vertex count
 * (graph-numa.h:1177-1206,1258-1285).  relabel != 0 selects the '-o' rule.
 * Returns 0; 1 when the greedy ends before using all P partitions (the
 * trailing ones are empty; the reference's assert at graph-numa.h:1353
 * aborts there unless built with NDEBUG); -2 where the greedy overruns P
 * (the reference writes past its arrays). */
int ggo_partition_sizes(uint64_t count, const uint32_t *deg, uint64_t m,
                        int P, int mode, int relabel, uint64_t *sizes);
/* Hilbert index of (x,y) on a side x side grid (graph-numa.h:375-414). */
int64_t ggo_hilbert_xy2d(int64_t side, int64_t x, int64_t y);
/* roundUpPow2 of graph-numa.h:301-313 for a 64-bit unsigned value. */
uint64_t ggo_round_up_pow2(uint64_t v);

/* ---- engine ------------------------------------------------------------ */
typedef struct ggo_graph ggo_graph;
/* Copies off/ids (and weights when w != NULL, CSR order).  Builds the CSC
 * twin, P partitions (mode 0 edge / 1 vertex; relabel != 0: the driver's -o,
 * partition_relabel in graph-numa.h:1180-1194, 1290-1305), COO lists (edge
 * mode only, Hilbert order when hilbert != 0 else CSR order). */
ggo_graph *ggo_graph_new(uint32_t n, uint64_t m, const uint32_t *off,
                         const uint32_t *ids, const double *w, int symmetric,
                         int P, int part_vertex, int hilbert, int relabel);
void ggo_graph_free(ggo_graph *g);
int ggo_graph_error(void); /* status of the last ggo_graph_new (<0 error, bit0/bit1: COO/CSC sizing underran) */
/* COO partition starts (P+1 vertex ids) and CSC partition starts (P+1
 * positions in the compressed in-vertex list). */
int ggo_graph_starts(const ggo_graph *g, int which, uint64_t *starts);
uint64_t ggo_graph_ncompressed(const ggo_graph *g);
/* Edges of COO partition p in traversal order; returns the count. */
int64_t ggo_graph_coo(const ggo_graph *g, int p, uint32_t *src, uint32_t *dst);
int64_t ggo_graph_coo_count(const ggo_graph *g, int p);

/* Per-edgeMap trace record written by the app runners. */
typedef struct {
    int32_t path;           /* 0 sparse push, 1 dense CSC, 2 dense COO, -1 empty */
    int32_t pad;
    int64_t in_active;      /* frontier d_m going in   */
    int64_t in_out_edges;   /* frontier num_out_edges  */
    int64_t out_active;     /* output d_m              */
    int64_t out_out_edges;  /* output num_out_edges    */
} ggo_trace;
#define GGO_PATH_SPARSE 0
#define GGO_PATH_CSC 1
#define GGO_PATH_COO 2

/* app/PageRank.C:102-151.  out = x of the last executed iteration. */
int ggo_pagerank(const ggo_graph *g, int max_iter, double *out, int *iters,
                 ggo_trace *trace);
/* ggo_pagerank with bit_frontiers != 0: app/PageRankBit.C (every edgeMap
 * output marked all-active). */
int ggo_pagerank_ex(const ggo_graph *g, int max_iter, int bit_frontiers, double *out, int *iters,
                    ggo_trace *trace);
/* app/BFS.C:77-101.  parents: -1 unreached.  Returns number of levels. */
int ggo_bfs(const ggo_graph *g, uint32_t root, int32_t *parents,
            ggo_trace *trace, int max_trace);
/* app/Components.C:101-126.  Returns number of rounds. */
int ggo_cc(const ggo_graph *g, uint32_t *ids, ggo_trace *trace, int max_trace);
/* app/SPMV.C:64-92: y[d] = sum_{s->d} w * x[s] with x given; one edgeMap
 * over an all-active frontier.  x = NULL means x = 1/n as in SPMV.C:78. */
int ggo_spmv(const ggo_graph *g, const double *x, double *y, ggo_trace *trace);

/* sequence::plusReduce (utils.h:159-195): 1024-element serial blocks,
 * recursive over block sums. */
double ggo_plus_reduce(const double *a, uint64_t n);

/* Generic single edgeMap for the frontier/edge-op tests.  op: 0 PR, 1 BFS,
 * 2 CC, 3 SPMV.  The frontier is given dense (bytes) or sparse (list); the
 * output frontier is returned dense in out_dense and as (d_m, out_edges). */
int ggo_edgemap_once(const ggo_graph *g, int op, const uint8_t *in_dense,
                     const uint32_t *in_sparse, int64_t in_count, int bit,
                     int64_t threshold, double *a_f64, double *b_f64,
                     int64_t *a_i64, int64_t *b_i64, uint8_t *out_dense,
                     int64_t *out_dm, int64_t *out_edges, int *path);

/* ---- differentiated forms (gg_oracle_grad.c; no reference counterpart) --- */
/* S[s] = sum of s's out-edge weights, CSR order, FP64. */
void ggo_out_wsum(uint32_t n, const uint32_t *off, const float *w, double *S);
/* y[d] = sum_{e=(s->d)} w_e x[s] (w NULL: 1). */
void ggo_spmv_csr(uint32_t n, const uint32_t *off, const uint32_t *ids, const float *w,
                  const double *x, double *y);
/* Backward of ggo_spmv_csr: gx[s] = sum_e w_e gy[d], gw[e] = x[s] gy[d]
 * (CSR order); gx or gw may be NULL. */
void ggo_spmv_bwd(uint32_t n, const uint32_t *off, const uint32_t *ids, const float *w,
                  const double *x, const double *gy, double *gx, double *gw);
/* Row-normalised weighted PageRank, all vertices active: r = v (1/n when v
 * is NULL), then r <- (1-alpha) v + alpha P r until sum|r_new - r| < eps or
 * max_iter.  Returns the iteration count. */
int ggo_wpagerank(uint32_t n, const uint32_t *off, const uint32_t *ids, const float *w,
                  double alpha, double eps, int max_iter, const double *v, double *r,
                  double *l1_hist);
/* Implicit adjoint at r for dL/dr = u: lambda = u, then
 * lambda <- u + alpha P^T lambda until sum|dlambda| <= eps * sum|lambda| or
 * max_iter; gw[e] (CSR order) and gv as in gg_oracle_grad.c.  Returns the
 * iteration count. */
int ggo_wpagerank_adjoint(uint32_t n, const uint32_t *off, const uint32_t *ids, const float *w,
                          double alpha, double eps, int max_iter, const double *r, const double *u,
                          double *lambda, double *gw, double *gv);

/* ---- apps beyond the four (gg_oracle_apps.c) ------------------------------ */
/* app/BC.C:154-226 from 'root' on the out-edge CSR (the backward phase runs
 * on the true transpose).  score = (Dependencies - inv) / inv, NaN where
 * unreached; level = BFS round of first visit (-1 unreached); sigma_out (may
 * be NULL) = NumPaths.  FP64.  Returns the number of forward rounds. */
int ggo_bc(uint32_t n, const uint32_t *off, const uint32_t *ids, uint32_t root, double *score,
           int32_t *level, double *sigma_out);
/* app/BellmanFord.C:100-140 with integer edge lengths (w truncated to
 * integers, as intE): dist = shortest path length, INT_MAX/2 unreached,
 * all -(INT_MAX/2) after a negative cycle.  Returns the round count. */
int ggo_sssp(uint32_t n, const uint32_t *off, const uint32_t *ids, const float *w, uint32_t root, int64_t *dist);
/* app/PageRankDelta.C:124-176: delta-propagating PageRank; only vertices
 * whose last change exceeded eps2 * p push their change.  Stops when the
 * L1 norm of the changes falls below eps, or after max_rounds.  p = ranks.
 * Returns the number of rounds. */
int ggo_prdelta(uint32_t n, const uint32_t *off, const uint32_t *ids, double damping, double eps, double eps2,
                int max_rounds, double *p);
/* app/PageRankConverage.C:227-310: PageRank with the lost (dangling) mass
 * spread uniformly, p normalised each iteration; at most max_iter (the
 * app's MaxIter = 100) iterations while the L1 change exceeds eps.
 * p = ranks.  Returns the iteration count. */
int ggo_prconv(uint32_t n, const uint32_t *off, const uint32_t *ids, double damping, double eps, int max_iter,
               double *p);
/* graphtools/VEBO.C:84-321: the VEBO vertex order for P partitions from the
 * in-degrees.  new_of_old[v] = new id of vertex v; starts (P+1 entries, may
 * be NULL) = first new id of each partition.  Returns 0, or -1 on bad
 * arguments.  When the reference's zero-degree placement would not place
 * exactly the zero-degree vertices (its counts then overrun its arrays),
 * the zero-degree vertices go one by one to the partition with the fewest
 * vertices (the reference's own rule for fewer such vertices than
 * partitions). */
int ggo_vebo(uint32_t n, const uint32_t *indeg, int P, uint32_t *new_of_old, uint64_t *starts);

#ifdef __cplusplus
}
#endif
#endif
