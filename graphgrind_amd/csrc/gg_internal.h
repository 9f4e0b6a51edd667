// Internal declarations shared by the HIP translation units of libgraphgrind_amd.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "graphgrind_amd.h"

namespace gg {

// ---- error plumbing --------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define GG_HIP(expr)                                                               \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess)                                                      \
            return ::gg::fail(_e == hipErrorOutOfMemory ? GG_ERR_OOM : GG_ERR_HIP, \
                              std::string(#expr) + ": " + hipGetErrorString(_e));  \
    } while (0)

#define GG_TRY(expr)              \
    do {                          \
        int _rc = (expr);         \
        if (_rc != GG_OK) return _rc; \
    } while (0)

// Work-group size of every kernel in the library (4 wave64s).
constexpr int kBlock = 256;
// Items (rows + edges) per work-group of the merge-path pull.
constexpr int kPullIPT = 8;
constexpr int kPullIPB = kBlock * kPullIPT;

inline uint32_t blocks_for(uint64_t items, uint64_t per_block)
{
    uint64_t b = (items + per_block - 1) / per_block;
    return (uint32_t)(b ? b : 1);
}

// ---- device buffer owned by a context --------------------------------------
struct Buf {
    void *p = nullptr;
    uint64_t bytes = 0;
};

struct Ctx;
int buf_alloc(Ctx *c, Buf &b, uint64_t bytes);
int buf_ensure(Ctx *c, Buf &b, uint64_t bytes); // grow-only
void buf_free(Buf &b);

// ---- merge-path schedule of one CSR-like array ------------------------------
// For offsets off[0..rows] and rows+edges items cut into kPullIPB-item tiles:
// tile t starts at diagonal t*kPullIPB at (row start_row[t], edge start_edge[t]).
struct PullSched {
    uint32_t rows = 0;
    uint64_t edges = 0;
    uint32_t tiles = 0;
    Buf start_row;  // uint32[tiles+1]
    Buf start_edge; // uint32[tiles+1]
    Buf carry_row;  // uint32[tiles]  row in progress at tile end
    Buf carry_val;  // 8 bytes [tiles] partial of that row
    Buf head_val;   // 8 bytes [tiles] partial of the first row if it began earlier
};

// ---- row-block pull schedule (CSR-Adaptive) -----------------------------------
// Consecutive complete rows are packed into blocks of at most kRowEdges
// edges and kRowCap rows (Greathouse & Daga, "Efficient Sparse Matrix-Vector
// Multiplication on GPUs using the CSR Storage Format", SC 2014): a block
// streams its edges' gathered values into LDS with coalesced loads and then
// reduces each row without a scan, with 2^kind lanes per row (1 lane = one
// thread per row; up to a whole wave), chosen from the block's average
// degree.  A row with more than kRowEdges edges is split into chunks, one
// block each; its partials are combined in chunk order by the fix-up kernel.
constexpr int kRowEdges = kBlock * kPullIPT; // edges staged per block
constexpr int kRowCap = 2 * kBlock;          // rows per block
constexpr uint32_t kRowLong = 0xffu;         // kind of a long-row chunk
struct RowWork {
    uint32_t r0, r1;   // rows [r0, r1) (long: the row in r0)
    uint32_t e0, e1;   // edges [e0, e1)
    uint32_t kind;     // log2(lanes per row), or kRowLong
    uint32_t slot;     // long: partial slot of this chunk
    uint32_t pad[2];
};
struct RowSched {
    uint32_t blocks = 0; // work items, one per block
    uint32_t nlong = 0;  // rows split into chunks
    uint32_t slots = 0;  // chunks of those rows
    Buf work;            // RowWork[blocks]
    Buf long_rows;       // uint32[3*nlong]: row, first slot, end slot
    Buf slot_val;        // 8 bytes [slots]
};

// ---- source-segmented pull (one segment per XCD) -------------------------------
// MI355X has 8 XCDs, each with its own 4 MB L2; a plain pull has every XCD
// gather from the whole source vector, so each L2 caches the same hot lines
// and misses the rest.  Splitting the sources into kXSegs groups (source s in
// group (s >> kXSegShift) % kXSegs, whole 128-byte lines of values) and
// letting each XCD pull only its group's in-edges makes the 8 L2s hold 8
// different slices.  Each (row, group) pair with at least one edge is a
// "virtual row" of its own CSC; its partial sum is written once, and a
// combine pass adds a row's partials in group order (a fixed order) and
// finishes the row.
constexpr int kXSegs = 8;
constexpr int kXSegShift = 5;     // 32 sources: one 128-byte line of FP32, two of FP64
constexpr int kXCombRows = 1024;  // rows per combine work-group (4 per thread)
constexpr int kXLaneEdges = 16;                 // edges a lane takes in a short item of the partial pass
constexpr int kXWaveEdges = 64 * kXLaneEdges;   // edges one wave covers in a step
constexpr int kXShortEdges = kXWaveEdges - 4;   // short items: whole rows, at most this many edges (the
                                                // wave's window starts up to 3 edges early, on a 16-byte line)
constexpr int kXPartThreads = 1024;             // partial-pass work-group (one per CU)
constexpr uint32_t kXHotBytes = 156u * 1024u;   // its LDS cache of the group's hottest sources
static_assert(kXSegs == 8, "the local numbering below assumes 8 groups");
// a source's index within its group (its chunk number there, then its place
// in the chunk), and back
__host__ __device__ inline uint32_t xseg_local(uint32_t s)
{
    return ((s >> (kXSegShift + 3)) << kXSegShift) | (s & ((1u << kXSegShift) - 1u));
}
__host__ __device__ inline uint32_t xseg_global(uint32_t loc, uint32_t grp)
{
    return ((loc >> kXSegShift) << (kXSegShift + 3)) | (grp << kXSegShift) | (loc & ((1u << kXSegShift) - 1u));
}

// ---- relabelled pull plan ----------------------------------------------------
// An internal renumbering of the vertices (hot sources first) with its own
// in-edge twin and schedule; kernels run in the new numbering and results
// are mapped back.  Built by gg_graph_reorder, like the reference's offline
// VEBO reordering but kept inside the library.
// In-lists longer than kHubChunk edges cut into chunks of kHubChunk, so a
// per-vertex pass can spread one vertex's in-edges over many waves: items
// are (vertex, first edge, end edge) triplets, hubs (vertex, first item,
// end item), both in vertex order.
constexpr uint32_t kHubChunk = 1024;
struct HubChunks {
    bool ok = false;
    uint32_t nitems = 0, nhubs = 0;
    Buf items, hubs;
};
void hubs_free(HubChunks &h);
int ensure_hub_chunks(Ctx *c, const uint32_t *toff_dev, uint32_t n, HubChunks &h);

struct PullPlan {
    int mode = 0;        // 0 none, 1 out-degree descending
    Buf perm;            // uint32[n]: new id of old vertex
    Buf inv;             // uint32[n]: old id of new vertex
    Buf toff, tids, tw;  // relabelled in-edge twin (tw when weighted)
    Buf outdeg;          // uint32[n] in the new numbering
    PullSched sched;
    RowSched rows;
    // Out-edge CSR in the new numbering, built on first use by the
    // transpose-side kernels: row perm[v] holds v's out-edges in their
    // original order (heads renamed), so the caller's edge e of source v sits
    // at coff[perm[v]] + (e - off[v]); cdelta[perm[v]] = off[v] - coff[perm[v]]
    // (mod 2^32) maps it back.
    Buf coff, cids, cw, cdelta;
    Buf indeg; // uint32[n] in-degrees in the new numbering
    Buf first_in; // uint32[n]: first source of each in-list (BFS pulls), built on first use
    HubChunks hubs; // long in-lists in chunks (multi-source path counts), built on first use
    PullSched csched;
    RowSched crows;
};

// The source-segmented form of a plan's in-edge twin (built on first use).
struct XSeg {
    bool ok = false;
    uint64_t vrows = 0;                  // virtual rows: (row, group) pairs with an edge
    uint64_t row_start[kXSegs + 1] = {}; // first virtual row of each group
    uint32_t item_start[kXSegs + 1] = {};// first row-block work item of each group
    uint32_t comb_blocks = 0;            // combine work-groups (kXCombRows rows each)
    uint32_t grid = 0;                   // partial-pass work-groups: kXSegs x (CUs / kXSegs)
    Buf vtoff, vtids;                    // virtual CSC: offsets [vrows+1], sources [m + kXWaveEdges] (zero-padded;
                                         // bit 31 set on each virtual row's first edge)
    Buf vw;                              // f32 [m + kXWaveEdges]: the edges' weights in vtids order (weighted
                                         // graphs; zero-padded)
    Buf mask;                            // uint32[n]: bit k when the row has group-k in-edges
    Buf cursor;                          // uint32[comb_blocks*kXSegs]: group-k virtual rows before block b
    Buf partial;                         // 8 bytes [vrows]
    Buf meta;                            // uint32[2*(kXSegs+1)]: item_start, then row_start (device)
    RowSched rows;                       // wave items over the virtual CSC (no item spans two groups)
};
void xseg_free(XSeg &x);

// In-edge arrays + schedule a pull runs over: the row-block schedule when
// 'rows' is set, else the merge-path tiles.
struct CscView {
    const uint32_t *toff = nullptr;
    const uint32_t *tids = nullptr;
    PullSched *sched = nullptr;
    RowSched *rows = nullptr;
    // work-groups of the main pull kernel (per-block residual partials)
    uint32_t blocks() const { return rows ? rows->blocks : sched->tiles; }
    // work-groups of the fix-up kernel (per-block folded residuals)
    uint32_t fix_blocks() const
    {
        return rows ? blocks_for(rows->nlong > rows->blocks ? rows->nlong : rows->blocks, kBlock)
                    : blocks_for(sched->tiles, kBlock);
    }
};
// Row-block work items over CSR offsets: runs of whole rows with at most
// item_edges edges and row_cap rows, or item_edges-edge chunks of a longer
// row; no item spans a row listed in 'breaks' (ascending).
int rows_build(Ctx *c, RowSched &s, const uint32_t *d_off, uint32_t rows, uint64_t edges, int edges_per_lane,
               uint32_t item_edges = kRowEdges, uint32_t row_cap = kRowCap,
               const std::vector<uint64_t> *breaks = nullptr);
void rows_free(RowSched &s);

// ---- captured fixed-iteration loops ---------------------------------------------
// A loop with a fixed iteration count has no data-dependent exit, so it is
// captured once into a hipGraph and replayed as one launch.  The capture
// bakes in buffer addresses and parameters; 'key' records them and the
// buffers the loop owns live here, so a call with the same key replays.
struct GraphExec {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t key[10] = {};
    Buf own[6];
};
void graph_exec_free(GraphExec &x);

// ---- per-context counters ---------------------------------------------------
// u64 [0..3]: totals (active, out_edges, scanned, cursor); from u64
// kCntStripeBase on: kStripes stripes of kStripeStride u64 (128 bytes)
// holding partial active / out_edges / scanned counts.
constexpr int kStripes = 64;
constexpr int kStripeStride = 16;
constexpr int kCntStripeBase = 64;
constexpr uint64_t kCounterBytes = 8ull * (kCntStripeBase + kStripes * kStripeStride);
struct Ctx;
// Pooled buffers of a context: take reuses a pooled buffer of at least
// 'bytes' (and at most about twice that) or allocates one; give returns the
// buffer to the pool (or frees it when the pool is full or the context is
// gone).  ensure grows a buffer through the pool.
int pool_take(Ctx *c, Buf &b, uint64_t bytes);
int pool_ensure(Ctx *c, Buf &b, uint64_t bytes);
void pool_give(Ctx *c, Buf &b);
void pool_drain(Ctx *c);
int counters_zero(Ctx *c);                           // totals and stripes
// Copies 'bytes' (<= kPinBytes) from device memory to host memory through
// the context's pinned buffer and waits for the stream: a read-back into
// pageable memory goes through a driver staging copy and takes longer.
constexpr size_t kPinBytes = 1024;
int read_back(Ctx *c, void *host, const void *dev, size_t bytes);
int counters_read(Ctx *c, unsigned long long out[4]); // folds the stripes first

// ---- context ----------------------------------------------------------------
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    Buf counters;   // uint64 counters: totals + stripes (kCounterBytes)
    void *hpin = nullptr; // kPinBytes of pinned host memory for small read-backs
    Buf tmp;        // grow-only temporary (sort space etc.)
    // Pool of device buffers released by per-call temporaries and frontiers
    // (size -> pointers).  They are reused by later requests on this
    // context instead of hipFree + hipMalloc: hipFree waits for the whole
    // device and unmaps, hundreds of microseconds per edgeMap level.  A
    // pooled buffer is only used on this context's stream, so stream order
    // makes reuse safe.
    std::multimap<uint64_t, void *> pool;
    uint64_t pool_bytes = 0;
    std::mutex pool_mu; // frontiers may be released from another host thread
    double last_kernel_ms = 0;
    int last_kernel_launches = 0;
};

// ---- graph ------------------------------------------------------------------
struct Graph {
    Ctx *ctx = nullptr;
    uint32_t n = 0;
    uint64_t m = 0;
    bool symmetric = false;
    bool weighted = false;
    // CSR (out-edges) and in-edge twin
    Buf off, ids, w;     // uint32[n+1], uint32[m], float[m]
    Buf toff, tids, tw;  // uint32[n+1], uint32[m], float[m]
    Buf perm;            // uint32[m]: CSC position -> CSR position
    Buf outdeg, indeg;   // uint32[n]
    // partitions
    int P = 0;
    bool part_vertex = false, hilbert = true, relabel = false;
    bool coo_underran = false, csc_underran = false;
    std::vector<uint64_t> coo_starts, csc_starts; // host copies (P+1)
    std::vector<uint64_t> coo_eoff;               // host (P+1)
    Buf comp;           // uint32[ncomp] compressed in-vertices
    uint64_t ncomp = 0;
    Buf coo_src, coo_dst; // uint32[m] (edge mode only)
    Buf coo_w;            // float[m] when weighted
    // the in-edge twin's rows with each destination's edges in COO (Hilbert)
    // order: sources and weights, built on first use by the COO traversal
    Buf coo_tids, coo_tw;
    bool coo_pull_ok = false;
    // first source of each in-list (BFS pulls), built on first use
    Buf first_in;
    HubChunks hubs; // long in-lists in chunks (multi-source path counts), built on first use
    // merge-path schedule over the in-edge twin
    PullSched pull;
    PullPlan plan;
    // transpose traversals (adjoints): schedule over the out-edge CSR and the
    // FP64 out-weight sums, built on first use
    PullSched csr_pull;
    bool csr_pull_ok = false;
    Buf wsum;            // double[n]
    bool wsum_ok = false;
    // row blocks over the in-edge twin and over the out-edge CSR, built on
    // first use
    RowSched rows, csr_rows;
    bool rows_ok = false, csr_rows_ok = false;
    // captured fixed-iteration loops (weighted PageRank forward / adjoint)
    GraphExec fwd_exec, adj_exec;
    // source-segmented form of the plan's twin, and whether the fused
    // PageRank loop and SpMV use it (GG_OPT_XCD_SEGMENTS)
    XSeg xseg;
    int xcd_segments = 1;
    // workspace of the fused PageRank loop, kept between calls (grow-only)
    Buf pr_ws[12];
    bool transposed = false; // gg_graph_transpose flips the two directions
    int bfs_levels_hint = 0;  // depth of the last BFS on the graph (sizes the next one's first batch of levels)
    int cc_rounds_hint = 0;   // rounds of the last Components call on the graph (likewise)
    int msbfs_levels_hint = 0; // levels of the last multi-source BFS or BC forward on the graph (likewise)
    // tuning (gg_graph_set_option)
    int pull_schedule = 1;  // 0 merge-path tiles, 1 row blocks where built
    int edges_per_lane = 8; // row blocks: target edges per lane of a row group
    int dense_coo = 0;      // 1: the dense COO branch streams the COO lists
};
// Drops captured loops (their addresses or schedules changed).
void graph_exec_drop(Graph *g);
// The COO-ordered in-lists (coo_tids / coo_tw), built on first use from the
// COO partitions; requires edge partitions.
int ensure_coo_pull(Graph *g);
// The source-segmented form of the plan's twin (g->xseg), built on first use;
// requires a plan.
int ensure_xseg(Graph *g);
// Row blocks of the in-edge twin / of the out-edge CSR (built on first use).
int ensure_rows(Graph *g);
int ensure_csr_rows(Graph *g);

// ---- frontier ---------------------------------------------------------------
struct Frontier {
    Graph *g = nullptr;
    Ctx *ctx = nullptr;  // outlives the graph: destroy needs no graph
    uint32_t n = 0;      // numRows (0 for the empty() frontier)
    bool bit = false;    // all active, flags not read
    bool has_dense = false;
    bool dense_valid = false;  // bytes + bitmap hold the set
    bool sparse_valid = false; // list holds the set
    uint64_t d_m = 0;
    uint64_t out_edges = 0;
    // The out-degree array out_edges was summed from over exactly the
    // members (list entries, duplicates included) when the engine made it;
    // NULL when unknown.  A sparse push over the same degrees sizes itself
    // from out_edges instead of reading the scan total back.
    const void *deg_src = nullptr;
    Buf bytes;  // uint8[n]
    Buf bitmap; // uint32[(n+31)/32]
    Buf list;   // uint32[cap]
};

// The relabelled out-edge CSR of the graph's plan (built on first use), as
// a pull view with the graph's schedule kind; requires a plan.
int plan_csr_view(Graph *g, CscView *v);
// The graph in its plan's numbering as a Graph sharing the plan's arrays
// (both edge directions, degrees, weights, schedules); it owns nothing:
// release it with plan_alias_release, never graph_free.
int plan_alias(Graph *g, Graph &t);
void plan_alias_release(Graph &t);

// ---- shared device helpers (defined in gg_core.hip) ------------------------
int sched_build(Ctx *c, PullSched &s, const uint32_t *d_off, uint32_t rows, uint64_t edges);
// A graph from a CSR already on the device (takes the buffers; w may be
// empty): builds degrees, the in-edge twin and the pull schedule.
int graph_from_device(Ctx *c, uint32_t n, uint64_t m, Buf &off, Buf &ids, Buf &w, bool symmetric, Graph **out);
int fill_bytes(Ctx *c, void *p, uint64_t bytes, uint8_t v);
int fill_u32(Ctx *c, void *p, uint64_t count, uint32_t v);
int bitmap_from_bytes(Ctx *c, const uint8_t *bytes, uint32_t *bitmap, uint32_t n);
int sync(Ctx *c);

// Frontier helpers (gg_traverse.hip)
int frontier_alloc(Graph *g, Frontier *f);
int frontier_make_dense(Frontier *f);
int frontier_make_sparse(Frontier *f);
int frontier_stats_dense(Frontier *f);

} // namespace gg

struct gg_ctx_s : gg::Ctx {};
struct gg_graph_s : gg::Graph {};
struct gg_frontier_s : gg::Frontier {};

namespace gg {
// Contexts are always allocated as gg_ctx_s, so the downcast is exact.
inline gg_ctx_t abi(Ctx *c) { return static_cast<gg_ctx_t>(c); }
} // namespace gg
