/*
 * graphgrind_amd.h -- C ABI of the MI355X (gfx950) traversal engine.
 *
 * Drop-in boundary for GraphGrind's hot path.  Every entry point below
 * replaces a piece of the reference's header-only C++ template API; the
 * comment on each names the reference interface (file:line) it stands for.
 * The C++ front end (include/gg/ligra-numa.h) maps the reference's app-facing
 * API onto these calls; tests and benchmarks bind them through ctypes.
 *
 * Conventions
 *  - Every function returns an int status: GG_OK (0) or a negative GG_ERR_*.
 *    The library never aborts; gg_last_error() returns the message of the
 *    last failing call on the calling thread.  Where the reference abort()s
 *    or asserts, the matching call returns an error instead.
 *  - Handles own device memory.  Host arrays passed in are copied during the
 *    call and not retained.  "dev" pointers are device (HBM) addresses
 *    obtained from gg_malloc (or any other hipMalloc'd memory on the
 *    context's device).
 *  - A context is one device and one HIP stream; it is not thread-safe.
 *    Calls that return host-visible results synchronize the stream.
 *  - Vertex ids and edge offsets are uint32 (n < 2^32, m < 2^32).
 */
#ifndef GRAPHGRIND_AMD_H
#define GRAPHGRIND_AMD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define GG_OK 0
#define GG_ERR_INVALID -1     /* bad argument / null handle                 */
#define GG_ERR_HIP -2         /* HIP runtime failure                        */
#define GG_ERR_OOM -3         /* device allocation failed                   */
#define GG_ERR_SIZE -4        /* frontier/graph size mismatch               */
#define GG_ERR_PARTITION -5   /* partition greedy overruns P                */
#define GG_ERR_UNSUPPORTED -6 /* option not provided by this build          */
#define GG_ERR_STATE -7       /* object in the wrong state for the call     */

/* ---- element types ------------------------------------------------------ */
#define GG_F64 0
#define GG_F32 1
#define GG_U32 2
#define GG_I32 3
#define GG_U8 4

typedef struct gg_ctx_s *gg_ctx_t;
typedef struct gg_graph_s *gg_graph_t;
typedef struct gg_frontier_s *gg_frontier_t;

/* ---- context ------------------------------------------------------------ */
/* Replaces the Cilk runtime + NUMA setup of parallel.h:27-46 and
 * partitioner.h:15-30: one context per GPU, one stream. */
int gg_ctx_create(int device, gg_ctx_t *out);
int gg_ctx_destroy(gg_ctx_t ctx);
int gg_ctx_sync(gg_ctx_t ctx);
/* Thread-local text of the last error (never NULL). */
const char *gg_last_error(void);
/* Library version string. */
const char *gg_version(void);

/* ---- device memory: mmap_ptr<T> (mm.h:246-325) -------------------------- */
int gg_malloc(gg_ctx_t ctx, uint64_t bytes, void **dev);
int gg_free(gg_ctx_t ctx, void *dev);
int gg_memcpy_h2d(gg_ctx_t ctx, void *dev, const void *host, uint64_t bytes);
int gg_memcpy_d2h(gg_ctx_t ctx, void *host, const void *dev, uint64_t bytes);
int gg_memcpy_d2d(gg_ctx_t ctx, void *dst, const void *src, uint64_t bytes);
/* Page-locked host memory for the host side of an mmap_ptr (the shadow the
 * apps' host loops read and write, part_allocate / Interleave_allocate,
 * mm.h:60-130): copies to and from HBM then run at full link speed. */
int gg_host_alloc(gg_ctx_t ctx, uint64_t bytes, void **host);
int gg_host_free(gg_ctx_t ctx, void *host);
/* map_vertexL init loops (app/PageRank.C:124-125, BFS.C:88,
 * Components.C:113): dev[i] = value for i < count. */
int gg_fill(gg_ctx_t ctx, void *dev, int dtype, uint64_t count, double value);
/* dev[i] = i (Components.C:113: IDs[j] = j). */
int gg_iota(gg_ctx_t ctx, void *dev, int dtype, uint64_t count);

/* ---- graph: wholeGraph + partitioned_graph (graph-numa.h:500-756) ------ */
/* From host CSR (offsets n+1, ids m, optional FP32 weights in CSR order).
 * Builds the in-edge twin on the device (in-lists ascending by source, as
 * IO.h:43-51,303 produce) and the out/in degree arrays.  No partitions yet. */
int gg_graph_create_csr(gg_ctx_t ctx, uint32_t n, uint64_t m, const uint32_t *off,
                        const uint32_t *ids, const float *w, int symmetric,
                        gg_graph_t *out);
/* Seeded R-MAT graph generated, deduplicated and laid out on the device
 * (a=.57 b=.19 c=.19 d=.05, splitmix64 draws keyed by (seed, edge, level);
 * self loops dropped; symmetric adds reverse edges).  weighted: FP32 weights
 * (k+1)/2^24 keyed by (seed, CSR position).  The reference has no
 * generator; this feeds its apps the inputs the benchmarks are quoted on. */
int gg_graph_create_rmat(gg_ctx_t ctx, int scale, int edge_factor, uint64_t seed,
                         int symmetric, int weighted, gg_graph_t *out);
int gg_graph_destroy(gg_graph_t g);
/* GA.n / GA.m / isSymmetric (graph-numa.h:633, 511). */
int gg_graph_info(gg_graph_t g, uint32_t *n, uint64_t *m, int *symmetric, int *weighted);
/* Copies of the device layout (tests, interop).  Any pointer may be NULL. */
int gg_graph_copy_csr(gg_graph_t g, uint32_t *off, uint32_t *ids, float *w);
int gg_graph_copy_csc(gg_graph_t g, uint32_t *toff, uint32_t *tids, float *tw);
/* V[v].getOutDegree() / getInDegree() for all v (graph-numa.h:187-194). */
int gg_graph_degrees(gg_graph_t g, int in_edges, uint32_t *out);
/* Replaces the int weights of a graph (CSR order, m FP32 values). */
int gg_graph_set_weights(gg_graph_t g, const float *w_csr);
/* The same from a device array (m FP32 values in CSR order). */
int gg_graph_set_weights_dev(gg_graph_t g, const float *w_dev);
/* Batching (no reference counterpart): packs 'count' graphs of one context
 * into one arena graph, a block-diagonal union whose CSR is the parts' CSRs
 * back to back.  vbase[i] / ebase[i] (count+1 entries each, may be NULL)
 * receive graph i's first vertex id and first edge position in the arena.
 * The arena is weighted when every part is.  Per-graph results are slices
 * of the arena's arrays. */
int gg_graph_create_arena(gg_ctx_t ctx, int count, const gg_graph_t *parts, gg_graph_t *out,
                          uint32_t *vbase, uint64_t *ebase);

/* GA.transpose() (graph-numa.h:547-555, 721-725): swaps the out-edge CSR
 * and the in-edge twin (with weights, degrees and schedules), so later
 * edgeMaps run on the transposed graph; calling it again restores the
 * graph.  A symmetric graph is left as it is, as in the reference.  The
 * reference flips only the vertex views of its CSC graph and leaves the
 * COO partitions as they were; here every structure follows the transpose:
 * a partitioned graph is re-partitioned by the new destinations (same P and
 * options) and a relabelled plan (gg_graph_reorder) is dropped. */
int gg_graph_transpose(gg_graph_t g);

/* VEBO order (graphtools/VEBO.C:56-321, the reference's offline reorder
 * tool): the new id of every vertex (new_of_old, n entries, host) for P
 * partitions from the graph's in-degrees, and the first new id of each
 * partition (starts, P+1 entries, host, may be NULL).  Each partition gets
 * about m/P in-edges and n/P vertices, its vertices in descending
 * in-degree.  Where the reference's zero-degree placement would not place
 * exactly the zero-degree vertices (its counts then overrun its arrays),
 * they go one by one to the partition with the fewest vertices. */
int gg_vebo_order(gg_graph_t g, int P, uint32_t *new_of_old, uint64_t *starts);

/* The graph renumbered by new_of_old (a permutation of 0..n-1, host): new
 * vertex new_of_old[v] gets v's out-edges, each list in its old order with
 * its heads renumbered, weights carried (graphtools/VEBO.C:333-356).  *out
 * is a new, unpartitioned graph on the same context. */
int gg_graph_relabel(gg_graph_t g, const uint32_t *new_of_old, gg_graph_t *out);

/* Internal vertex renumbering for the pull kernels (no reference
 * counterpart in the engine; the reference reorders offline with VEBO,
 * graphtools/VEBO.C, and '-o').  mode 1: vertices by out-degree descending,
 * so the sources most edges read sit together in L2; mode 0 drops it.
 * Used, when present, by gg_pagerank, gg_bfs, gg_cc, gg_spmv, gg_spmv_bwd,
 * gg_wpagerank, gg_wpagerank_adjoint, gg_bc, gg_sssp and gg_msbfs, which
 * map their inputs in and results back to the caller's numbering (integer
 * results are the same, BFS parents stay a valid tree, only the FP
 * summation order changes); gg_edge_map and gg_vertex_map ignore it.
 * Costs about 3 extra copies of the edge ids and weights in device memory
 * (the relabelled in-edge twin, plus a relabelled out-edge CSR built on the
 * first call that needs it).  Rebuilt when the weights change; drops the
 * graph's captured loops. */
int gg_graph_reorder(gg_graph_t g, int mode);

/* Raw device addresses of a graph's layout, for interop (no reference
 * counterpart: the reference's arrays are host memory its apps read
 * directly).  which 0: the graph as given; which 1: the renumbered plan of
 * gg_graph_reorder (perm = new id of each old vertex).  Pointers stay valid
 * until the graph is destroyed or re-planned. */
typedef struct {
    const uint32_t *toff;       /* n+1 in-edge offsets          */
    const uint32_t *tids;       /* m in-edge sources            */
    const uint32_t *outdeg;     /* n out-degrees                */
    const uint32_t *perm;       /* n (which 1) or NULL          */
    const uint32_t *tile_row;   /* tiles+1 merge-path row starts  */
    const uint32_t *tile_edge;  /* tiles+1 merge-path edge starts */
    uint32_t tiles;
    uint32_t n;
    uint64_t m;
    const uint32_t *row_work;   /* row blocks: 8 uint32 each (r0, r1, e0, e1,
                                   kind, slot, 0, 0); NULL when not built  */
    uint32_t row_blocks;
    uint32_t row_long;          /* rows split into chunks                */
} gg_graph_view;
int gg_graph_device_view(gg_graph_t g, int which, gg_graph_view *out);

/* Device partitioner: partitioned_graph ctor (graph-numa.h:642-685) with
 * partitionByDegree/partitionByVertex (1177-1206, 1287-1357),
 * cscpartitionByDegree/ByVertex (1208-1285), COOPartitionByDest
 * (1116-1143) and hilbert_sort (482-490).  by_vertex: '-v vertex';
 * relabel: '-o'; hilbert: EDGES_HILBERT (else CSR order).  Partitions are
 * by destination ('-P dest'); '-P source' returns GG_ERR_UNSUPPORTED.
 * Returns GG_ERR_PARTITION where the reference greedy overruns P. */
int gg_graph_partition(gg_graph_t g, int P, int by_vertex, int relabel, int hilbert);
/* which 0: COO partition starts (P+1 vertex ids, partitioner.h:153-165);
 * which 1: CSC partition starts (P+1 positions in the compressed in-vertex
 * list, graph-numa.h:1092-1107).  *underran (may be NULL) is set when the
 * greedy used fewer than P partitions (the reference asserts there). */
int gg_graph_partition_starts(gg_graph_t g, int which, uint64_t *starts, int *underran);
/* COO edge lists (get_edge_list_partition, graph-numa.h:696-699): all
 * partitions back to back; part_off has P+1 entries.  Any may be NULL. */
int gg_graph_coo(gg_graph_t g, uint32_t *src, uint32_t *dst, uint64_t *part_off);

/* ---- frontier: partitioned_vertices (ligra-numa.h:91-272) -------------- */
#define GG_FRONTIER_EMPTY 0  /* empty()  119-130                         */
#define GG_FRONTIER_SINGLE 1 /* create(n, v, outdeg(v)) 103-116         */
#define GG_FRONTIER_ALL 2    /* bits(part, n, m) 169-180: all active     */
int gg_frontier_create(gg_graph_t g, int kind, uint32_t v, gg_frontier_t *out);
/* From host flags (dense bool[n], 183-194) or a host id list (197-208). */
int gg_frontier_from_host(gg_graph_t g, const uint8_t *dense, const uint32_t *sparse,
                          uint64_t count, gg_frontier_t *out);
int gg_frontier_destroy(gg_frontier_t f);
/* numNonzeros (268-271), num_out_edges, has_dense, bit, numRows. */
int gg_frontier_info(gg_frontier_t f, uint64_t *d_m, uint64_t *out_edges, int *is_dense,
                     int *bit, uint32_t *n);
/* The public field 'bit' (ligra-numa.h:91-101), written by apps such as
 * app/PageRankBit.C:145 ('output.bit = true'): with bit set every vertex
 * counts as active and the flags are not read; d_m and num_out_edges keep
 * their stored values, which edgeMap's branch choice still reads
 * (ligra-numa.h:929-957).  Not allowed on an empty() frontier. */
int gg_frontier_set_bit(gg_frontier_t f, int bit);
/* toDense (228-244) / toSparse (246-260; list ascending like packIndex). */
int gg_frontier_to_dense(gg_frontier_t f);
int gg_frontier_to_sparse(gg_frontier_t f);
/* Host copies: dense flags (n bytes) / sparse list (cap entries, *count). */
int gg_frontier_copy_dense(gg_frontier_t f, uint8_t *out);
int gg_frontier_copy_sparse(gg_frontier_t f, uint32_t *out, uint64_t cap, uint64_t *count);

/* ---- edgeMap (ligra-numa.h:919-1019) ------------------------------------ */
/* The app functors become an enumerated operator set; a and b are the
 * functor's device arrays.  See include/gg/ligra.h for the trait mapping. */
#define GG_OP_PAGERANK 0 /* PR_F, app/PageRank.C:28-70 and
                            app/PageRankDelta.C:27-72.  a=p_curr (Delta)
                            b=p_next (nghSum), F64 or F32;
                            p_next[d] += p_curr[s]/outdeg(s); any sign   */
#define GG_OP_BFS 1      /* BFS_F, app/BFS.C:26-75.  a=Parents (I32, -1 =
                            unvisited); cond = Parents[d] == -1           */
#define GG_OP_CC 2       /* CC_F, app/Components.C:26-86.  a=IDs b=prevIDs
                            (U32); min-label, frontier = changed          */
#define GG_OP_SPMV 3     /* SPMV_F, app/SPMV.C:7-48.  a=p_curr b=p_next
                            (F64 or F32); p_next[d] += p_curr[s]*w        */
#define GG_OP_SSSP 4     /* BF_F, app/BellmanFord.C:28-92.  a=ShortestPathLen
                            b=Visited (I32 both); d joins the output when
                            SP[s] + w < SP[d] lowers SP[d] and Visited[d]
                            was 0 (then set to 1).  Integer lengths: the
                            graph's weights truncated (1 when unweighted).
                            The dense branch relaxes from a copy of the
                            frontier's distances (a Jacobi step).       */
#define GG_OP_PATHSUM 5  /* BC_F / BC_Back_F, app/BC.C:28-124.  a=NumPaths
                            or Dependencies (F64 or F32), b=Visited (U8);
                            cond = !Visited[d]; a[d] += a[s]; d joins the
                            output when a[d] was 0 before this edgeMap   */
#define GG_OP_SUM 6      /* PR_F of app/PageRankConverage.C:34-88.
                            a=add_factor b=p_next (F64 or F32);
                            p_next[d] += a[s], edge weights not used     */
typedef struct {
    int32_t kind;
    int32_t dtype; /* GG_F64 / GG_F32 for PAGERANK, SPMV, PATHSUM, SUM */
    void *a;
    void *b;
} gg_op_t;
#define GG_PATH_EMPTY -1
#define GG_PATH_SPARSE 0 /* edgeMapSparseWithG 868-916 */
#define GG_PATH_CSC 1    /* edgeMapDenseCSC    549-729 */
#define GG_PATH_COO 2    /* edgeMapDense (COO) 803-866 */
/* threshold -1 means m/20 (932).  remdups: remDuplicates (317-348) on the
 * sparse output.  *out is a new frontier owned by the caller. */
int gg_edgemap(gg_graph_t g, gg_frontier_t in, const gg_op_t *op, int64_t threshold,
               int remdups, gg_frontier_t *out, int *path);

/* ---- vertexMap (ligra-numa.h:1023-1049) / map_vertexL ------------------- */
#define GG_VOP_PAGERANK 0 /* PR_Vertex_F (PageRank.C:73-87): b[i] = c0*b[i] + c1 */
#define GG_VOP_RESET 1    /* PR_Vertex_Reset (90-100): a[i] = 0                */
#define GG_VOP_CC_COPY 2  /* CC_Vertex_F (Components.C:88-99): b[i] = a[i]      */
#define GG_VOP_ABSDIFF 3  /* map_vertexL of PageRank.C:136: a[i] = |a[i]-b[i]| */
typedef struct {
    int32_t kind;
    int32_t dtype;
    void *a;
    void *b;
    double c0;
    double c1;
} gg_vop_t;
/* f == NULL applies to every vertex (map_vertexL over the partitioner). */
int gg_vertexmap(gg_graph_t g, gg_frontier_t f, const gg_vop_t *op);
/* vertexFilter (1053-1082): keep active i with a[i] (U8 flags, n) != 0. */
int gg_vertexfilter(gg_graph_t g, gg_frontier_t f, const uint8_t *flags_dev,
                    gg_frontier_t *out);
/* sequence::plusReduce (utils.h:192): fixed-order device sum. */
int gg_reduce_sum(gg_ctx_t ctx, const void *dev, int dtype, uint64_t count, double *out);

/* ---- fused app loops (the Compute() of each app, on device) ------------- */
typedef struct {
    int32_t iters;          /* iterations / levels / rounds executed       */
    int32_t paths[64];      /* edgeMap path of the first 64 steps          */
    uint64_t edges_scanned; /* edges read by the traversal kernels         */
    double ms;              /* hipEvent time of the loop (incl. init)      */
} gg_run_info;
/* app/PageRank.C:102-151: exactly the reference's frontier-gated loop,
 * max_iter iterations or until L1 < eps.  ranks_dev receives the x of the
 * last iteration (F64 or F32 per dtype). */
int gg_pagerank(gg_graph_t g, int dtype, int max_iter, double damping, double eps,
                void *ranks_dev, gg_run_info *info);
/* app/BFS.C:77-101 with the direction switch of edgeMap; parents_dev (I32
 * n, -1 unreached); levels_dev (I32 n, -1 unreached) may be NULL. */
int gg_bfs(gg_graph_t g, uint32_t root, int64_t threshold, int32_t *parents_dev,
           int32_t *levels_dev, gg_run_info *info);
/* app/Components.C:101-126: min-label propagation to the fixed point. */
int gg_cc(gg_graph_t g, uint32_t *labels_dev, gg_run_info *info);
/* app/SPMV.C:64-92 generalised: y[d] = sum_{s->d} w(s->d) x[s] over the
 * in-edge twin (unit weights when the graph has none). */
int gg_spmv(gg_graph_t g, int dtype, const void *x_dev, void *y_dev, gg_run_info *info);
/* gg_spmv without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops).  Same kernels and result as gg_spmv.  The x vector
 * permuted into a renumbered plan's order (gg_graph_reorder) goes to a
 * 64 MiB device array the library holds per GPU until the process ends, so
 * n * sizeof(value) must be at most 64 MiB with a plan (GG_ERR_SIZE
 * otherwise: R-MAT scale 24 fits in FP32, scale 23 in FP64).  Without a plan
 * the graph's row-block schedule must exist, and where gg_spmv takes the
 * source-segmented pull (GG_OPT_XCD_SEGMENTS) its layout must (one gg_spmv
 * call builds either; GG_ERR_STATE otherwise).  Asynchronous on the context's stream: no
 * synchronisation and no timing.  Calls that overlap on one GPU share the
 * array and race. */
int gg_spmv_noalloc(gg_graph_t g, int dtype, const void *x_dev, void *y_dev);

/* ---- further apps on the engine --------------------------------------------- */
/* app/BC.C:154-226 from 'root': forward pulls level by level summing
 * NumPaths, backward pulls over the out-edge CSR (the true transpose)
 * summing Dependencies; scores_dev (F64 or F32 per dtype, n) receives
 * (Dependencies - inv) / inv, NaN where unreached; levels_dev (I32 n, may be
 * NULL) the forward level, -1 unreached.  info->iters = forward rounds.
 * The reference's transpose() leaves the COO lists unflipped
 * (graph-numa.h:721-725), so its backward phase on a directed graph can
 * run along the wrong direction; this follows the algorithm. */
int gg_bc(gg_graph_t g, uint32_t root, int dtype, void *scores_dev, int32_t *levels_dev, gg_run_info *info);
/* app/BellmanFord.C:100-140 from 'root' with integer edge lengths (the
 * graph's weights truncated to integers, as intE; unit lengths when the
 * graph has none): dist_dev (I32 n) = shortest distance, INT_MAX/2 when
 * unreached, all -(INT_MAX/2) after a negative cycle (n rounds).  Each
 * round is a Jacobi relaxation from the last round's frontier. */
int gg_sssp(gg_graph_t g, uint32_t root, int32_t *dist_dev, gg_run_info *info);

/* ---- differentiated forms ------------------------------------------------
 * No counterpart in the reference: gradients of SPMV.C's product and of the
 * all-active PageRank (app/PageRankBit.C:145) with row-normalised weights,
 * on the same weighted graph.  Arrays are device pointers in 'dtype' (F64
 * or F32); per-edge arrays are in the CSR edge order the graph was created
 * with; weights stay FP32 on the graph. */
/* Backward of gg_spmv for a loss with dL/dy = gy: gx[s] = sum_{e=(s->d)}
 * w_e gy[d] (a pull over the out-edge CSR, the transpose traversal) and
 * gw[e] = x[s] gy[d].  gx or gw may be NULL (x is needed only for gw). */
int gg_spmv_bwd(gg_graph_t g, int dtype, const void *x_dev, const void *gy_dev, void *gx_dev,
                void *gw_dev, gg_run_info *info);
/* Weighted PageRank: r = v, then r <- (1-alpha) v + alpha P r with
 * P[d][s] = w_{s->d} / S_s (S_s = FP64 sum of s's out-weights; dangling
 * sources contribute nothing) until sum|r_new - r| < eps or max_iter.
 * v_dev NULL means the uniform teleport 1/n.  Needs edge weights. */
int gg_wpagerank(gg_graph_t g, int dtype, double alpha, double eps, int max_iter,
                 const void *v_dev, void *r_dev, gg_run_info *info);
/* eps < 0 (here and in gg_wpagerank_adjoint): exactly max_iter iterations
 * with no stop rule.  Such a loop is captured into a hipGraph on first use
 * and replayed as one launch while the buffers and parameters stay the same
 * (one captured forward and one adjoint per graph).  Over an arena graph
 * this runs every packed graph's personalised PageRank (teleport v = each
 * graph's seed distribution) in one replay. */
/* Implicit adjoint of gg_wpagerank at its output r for dL/dr = u:
 *   lambda = u, lambda <- u + alpha P^T lambda (pull over the out-edge CSR)
 *   until sum|lambda_new - lambda| <= eps * sum|lambda_new| or max_iter;
 *   gw[e] = (alpha r_s / S_s) (lambda_d - (lambda_s - u_s) / alpha)
 *   gv    = (1 - alpha) lambda            (dL/d teleport)
 * gw, gv and lambda (the converged adjoint) may each be NULL. */
int gg_wpagerank_adjoint(gg_graph_t g, int dtype, double alpha, double eps, int max_iter,
                         const void *r_dev, const void *u_dev, void *gw_dev, void *gv_dev,
                         void *lambda_dev, gg_run_info *info);

/* ---- multi-source BFS -----------------------------------------------------
 * No counterpart in the reference: the forward (BFS) phase of app/BC.C for
 * up to 32 sources at once, one 32-bit mask per vertex.  levels_dev is I32
 * [nsrc][n]: levels_dev[k*n + v] = hop distance of v from sources[k]
 * (host array), -1 when unreached.  info->iters = levels run; info->paths
 * (first 64 levels): GG_PATH_SPARSE where the level pushed from a frontier of
 * at most 4096 vertices (symmetric graphs only), GG_PATH_CSC where it pulled. */
int gg_msbfs(gg_graph_t g, int nsrc, const uint32_t *sources, int32_t *levels_dev, gg_run_info *info);
/* No counterpart in the reference: the forward phase of app/BC.C (BC_F and
 * BC_Vertex_F, 28-58, with NumPaths[start] = 1 at 196) for up to 32 sources
 * at once, the multi-source levels of gg_msbfs plus each source's path
 * counts.  paths_dev (F64 or F32 per dtype) is [nsrc][n]: paths_dev[k*n + v]
 * = NumPaths of v from sources[k] (host array), 0 when unreached; levels_dev
 * (I32 [nsrc][n], may be NULL) as gg_msbfs.  Sums run in in-list order, a
 * fixed order.  info->iters = levels run. */
int gg_msbc_forward(gg_graph_t g, int nsrc, const uint32_t *sources, int dtype, void *paths_dev,
                    int32_t *levels_dev, gg_run_info *info);

/* ---- tuning ---------------------------------------------------------------
 * No counterpart in the reference (its only knobs are -c/-v/-o): schedule
 * choices of the device pull, set per graph; the autotuner sweeps them. */
#define GG_OPT_PULL_SCHEDULE 0      /* 0 merge-path tiles, 1 row blocks (default) */
#define GG_OPT_ROW_EDGES_PER_LANE 1 /* row blocks: target edges per lane of a row group (8) */
/* The in-edge order of edgeMap's dense COO branch (ligra-numa.h:953-961,
 * taken when d_m + out-degree > m/2 on an edge-partitioned graph).  Both
 * are destination-owned pulls (one owner per destination, no atomics, a
 * fixed summation order): 0 (default) each destination's in-edges by source
 * id (the in-edge twin); 1 each destination's in-edges in its COO
 * partition's Hilbert order (edgeMapDense, ligra-numa.h:803-866), the order
 * in which the reference's PART96 loop adds them.  Graphs with a renumbered
 * plan run the plan's pull either way. */
#define GG_OPT_DENSE_COO 2
/* The fused PageRank loop and gg_spmv on a graph with a renumbered plan:
 * the source-segmented pull, in which each XCD gathers from one eighth of
 * the sources (its own L2's slice, the slice's hottest sources cached in
 * each CU's LDS) and a combine pass adds the per-group partial sums of each
 * row in group order, or one pull over the whole twin.  0 never, 1
 * (default) when m >= 2^24 (below that the source vector fits the L2s
 * anyway), 2 always.  Both have a fixed summation order. */
#define GG_OPT_XCD_SEGMENTS 3
int gg_graph_set_option(gg_graph_t g, int option, int64_t value);
int gg_graph_get_option(gg_graph_t g, int option, int64_t *value);
/* Autotuner: times 'trials' pulls of a probe operator (one gathered value
 * per edge, one write per row, in 'dtype') under each candidate schedule --
 * merge-path tiles, and row blocks at 1, 2, 4, 8, 16, 32 edges per lane --
 * then sets the fastest on the graph and reports all of them. */
typedef struct {
    int32_t schedule;            /* chosen GG_OPT_PULL_SCHEDULE             */
    int32_t edges_per_lane;      /* chosen GG_OPT_ROW_EDGES_PER_LANE        */
    double ms;                   /* its mean pull time                      */
    int32_t candidates;
    int32_t cand_schedule[8];
    int32_t cand_edges_per_lane[8];
    double cand_ms[8];
} gg_tune_info;
int gg_autotune_pull(gg_graph_t g, int dtype, int trials, gg_tune_info *out);

/* ---- timing------------------------------------------------------------- */
/* Average device time (ms) of the last fused loop's dominant kernel and
 * its launch count, measured with hipEvents on the context stream. */
int gg_last_kernel_time(gg_ctx_t ctx, double *avg_ms, int *launches);

#ifdef __cplusplus
}
#endif
#endif
