// gg_more.hip -- apps beyond the four on the same engine: betweenness
// contributions (app/BC.C) and Bellman-Ford distances (app/BellmanFord.C),
// each a loop of destination-owned pulls.
//
// BC from 'root' (BC.C:154-226) on the generic edgeMap with the path-sum
// operator (GG_OP_PATHSUM).  Forward level L: every vertex not yet visited
// adds NumPaths of its in-neighbours in level L-1 and, if it had none
// before, joins level L (BC_F 28-77).  Backward on the transposed view,
// deepest level first: Dependencies of level r+1 flow to every vertex not
// yet visited (BC_Back_F 80-124), then level r adds inverseNumPaths
// (BC_Back_Vertex_F 141-152); finally score = (Dependencies - inv) / inv
// (222), NaN where unreached.  Each level takes the reference's branch by
// the threshold m/20: the pull (fixed summation order) when the frontier is
// large, the push (atomic adds, as BC_F::updateAtomic) when it is small.
//
// Bellman-Ford from 'root' (BellmanFord.C:100-140), integer edge lengths,
// on the generic edgeMap with the relaxation operator (GG_OP_SSSP): each
// round the vertices whose distance fell last round relax their out-edges,
// by push (atomicMin, then the Visited CAS) or by the pull over a gated copy
// of the frontier's distances, whichever the threshold m/20 picks.  Integer
// minima do not depend on the order, so distances are exact either way.
// n rounds with a non-empty frontier mean a negative cycle: every distance
// becomes -(INT_MAX/2), as the reference does.
#include <cstring>
#include <limits>
#include <vector>

#include "gg_kernels.h"

namespace gg {

template <class V>
__global__ void k_bc_init(V *__restrict__ sigma, uint8_t *__restrict__ vis, int32_t *__restrict__ level,
                          uint32_t root, uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    sigma[v] = v == root ? V(1) : V(0); // NumPaths (BC.C:167-172)
    vis[v] = v == root;
    level[v] = v == root ? 0 : -1;
}

// BC_Vertex_F (BC.C:127-137) over a level's output: mark visited and
// record the level
__global__ void k_bc_mark(const uint8_t *__restrict__ F, uint8_t *__restrict__ vis, int32_t *__restrict__ level,
                          int32_t L, uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n && F[v]) {
        vis[v] = 1;
        level[v] = L;
    }
}

// inverseNumPaths (BC.C:196-198), Dependencies and Visited reset (193, 202)
template <class V>
__global__ void k_bc_inv(const V *__restrict__ sigma, V *__restrict__ inv, V *__restrict__ dep,
                         uint8_t *__restrict__ vis, uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    inv[v] = V(1) / sigma[v];
    dep[v] = V(0);
    vis[v] = 0;
}

// BC_Back_Vertex_F (BC.C:141-152) over a level's frontier
template <class V>
__global__ void k_bc_back_mark(const uint8_t *__restrict__ F, uint8_t *__restrict__ vis, V *__restrict__ dep,
                               const V *__restrict__ inv, uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n && F[v]) {
        vis[v] = 1;
        dep[v] = dep[v] + inv[v];
    }
}

template <class V>
__global__ void k_bc_final(const V *__restrict__ dep, const V *__restrict__ inv, V *__restrict__ score,
                           const int32_t *__restrict__ level, int32_t *__restrict__ levels_out,
                           const uint32_t *__restrict__ to_caller, uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint32_t o = to_caller ? to_caller[v] : v; // the caller's id of v
    score[o] = (dep[v] - inv[v]) / inv[v];
    if (levels_out) levels_out[o] = level[v];
}

// One-entry frontier list {root} with the root's out-degree after it (for
// the read-back of the frontier's statistics).
__global__ void k_root_list(uint32_t root, const uint32_t *__restrict__ outdeg, uint32_t *__restrict__ list)
{
    if (threadIdx.x == 0) {
        list[0] = root;
        list[1] = outdeg[root];
    }
}

static int root_frontier(Ctx *c, Graph *g, uint32_t root, Frontier *F)
{
    k_root_list<<<1, 64, 0, c->stream>>>(root, (const uint32_t *)g->outdeg.p, (uint32_t *)F->list.p);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

// The graph with its two edge directions swapped (GA.transpose()), sharing
// g's arrays and schedules: edgeMap on it pushes along in-edges and pulls
// over the out-edge CSR.  Unpartitioned, so its dense branch is the pull.
// It owns nothing: release it with alias_release, never graph_free.
static int transpose_alias(Graph *g, Graph &t)
{
    GG_TRY(ensure_csr_sched(g));
    if (g->pull_schedule == 1) {
        GG_TRY(ensure_rows(g));
        GG_TRY(ensure_csr_rows(g));
    }
    t.ctx = g->ctx;
    t.n = g->n;
    t.m = g->m;
    t.symmetric = g->symmetric;
    t.weighted = g->weighted;
    t.off = g->toff;
    t.ids = g->tids;
    t.w = g->tw;
    t.toff = g->off;
    t.tids = g->ids;
    t.tw = g->w;
    t.outdeg = g->indeg;
    t.indeg = g->outdeg;
    t.pull = g->csr_pull;
    t.csr_pull = g->pull;
    t.csr_pull_ok = true;
    t.rows = g->csr_rows;
    t.rows_ok = g->csr_rows_ok;
    t.csr_rows = g->rows;
    t.csr_rows_ok = g->rows_ok;
    t.pull_schedule = g->pull_schedule;
    t.edges_per_lane = g->edges_per_lane;
    t.transposed = !g->transposed;
    return GG_OK;
}

static void alias_release(Graph &t)
{
    Buf *bs[] = {&t.off, &t.ids, &t.w, &t.toff, &t.tids, &t.tw, &t.outdeg, &t.indeg};
    for (Buf *b : bs) *b = Buf{};
    t.pull = PullSched{};
    t.csr_pull = PullSched{};
    t.rows = RowSched{};
    t.csr_rows = RowSched{};
}

// BC.C Compute (154-226) on the generic edgeMap: each level takes the
// reference's branch (sparse push with atomic adds, or the destination-
// owned pull) by its threshold m/20; the backward phase runs on the
// transposed view.
// On a plan alias, 'root' is in the plan's numbering and to_caller (the
// plan's inverse permutation) maps scores and levels back.
template <class V>
static int bc_impl(Graph *g, uint32_t root, V *score, int32_t *levels_out, gg_run_info *info,
                   const uint32_t *to_caller)
{
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    const int dt = sizeof(V) == 8 ? GG_F64 : GG_F32;
    Buf sigma, vis, level, dep, inv;
    int rc = pool_take(c, sigma, sizeof(V) * (uint64_t)n);
    if (!rc) rc = pool_take(c, vis, (uint64_t)n + 64);
    if (!rc) rc = pool_take(c, level, 4ull * n);
    if (!rc) rc = pool_take(c, dep, sizeof(V) * (uint64_t)n);
    if (!rc) rc = pool_take(c, inv, sizeof(V) * (uint64_t)n);
    std::vector<Frontier *> levels;
    Graph tg;
    bool alias = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    EdgeMapExtra ex;
    int round = 0;
    const uint32_t nb = blocks_for(n, kBlock);
    if (!rc) {
        (void)hipEventRecord(e0, c->stream);
        k_bc_init<V><<<nb, kBlock, 0, c->stream>>>((V *)sigma.p, (uint8_t *)vis.p, (int32_t *)level.p, root, n);
        // partitioned_vertices::create(n, start, ...) (BC.C:175)
        Frontier *F = new Frontier();
        frontier_alloc(g, F);
        uint32_t od = 0;
        rc = pool_take(c, F->list, 8);
        if (!rc) rc = root_frontier(c, g, root, F);
        if (!rc) rc = read_back(c, &od, (uint32_t *)F->list.p + 1, 4);
        F->d_m = 1;
        F->out_edges = od;
        F->deg_src = g->outdeg.p;
        F->sparse_valid = true;
        levels.push_back(F);
        // forward (BC.C:183-191)
        while (!rc && F->d_m != 0) {
            round++;
            gg_op_t op = {GG_OP_PATHSUM, dt, sigma.p, vis.p};
            Frontier *out = nullptr;
            int path = 0;
            if ((rc = edgemap_core(g, F, &op, -1, 0, &out, &path, &ex))) break;
            if (info && round <= 64) info->paths[round - 1] = path;
            levels.push_back(out);
            if (out->d_m) {
                if ((rc = frontier_make_dense(out))) break;
                k_bc_mark<<<nb, kBlock, 0, c->stream>>>((const uint8_t *)out->bytes.p, (uint8_t *)vis.p,
                                                        (int32_t *)level.p, round, n);
            }
            F = out;
            if (round > (int)n) rc = fail(GG_ERR_STATE, "gg_bc: more levels than vertices");
        }
        // backward on the transpose, deepest level first (BC.C:193-218)
        if (!rc) {
            k_bc_inv<V><<<nb, kBlock, 0, c->stream>>>((const V *)sigma.p, (V *)inv.p, (V *)dep.p, (uint8_t *)vis.p, n);
            F = levels[round - 1];
            rc = frontier_make_dense(F);
            if (!rc)
                k_bc_back_mark<V><<<nb, kBlock, 0, c->stream>>>((const uint8_t *)F->bytes.p, (uint8_t *)vis.p,
                                                                (V *)dep.p, (const V *)inv.p, n);
        }
        if (!rc && round >= 2) {
            rc = transpose_alias(g, tg);
            alias = !rc;
        }
        for (int r = round - 2; r >= 0 && !rc; r--) {
            gg_op_t op = {GG_OP_PATHSUM, dt, dep.p, vis.p};
            Frontier *out = nullptr;
            F->g = &tg;
            rc = edgemap_core(&tg, F, &op, -1, 0, &out, nullptr, &ex);
            F->g = g;
            if (rc) break;
            frontier_delete(out);
            F = levels[r];
            if ((rc = frontier_make_dense(F))) break;
            k_bc_back_mark<V><<<nb, kBlock, 0, c->stream>>>((const uint8_t *)F->bytes.p, (uint8_t *)vis.p, (V *)dep.p,
                                                            (const V *)inv.p, n);
        }
        if (!rc) {
            k_bc_final<V><<<nb, kBlock, 0, c->stream>>>((const V *)dep.p, (const V *)inv.p, score,
                                                         (const int32_t *)level.p, levels_out, to_caller, n);
            if (hipGetLastError() != hipSuccess) rc = fail(GG_ERR_HIP, "gg_bc: launch");
        }
        (void)hipEventRecord(e1, c->stream);
        if (!rc) rc = sync(c);
    }
    float ms = 0;
    if (!rc) (void)hipEventElapsedTime(&ms, e0, e1);
    if (!rc && info) {
        info->iters = round;
        info->ms = ms;
        info->edges_scanned = ex.scanned;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (alias) alias_release(tg);
    for (Frontier *f : levels) frontier_delete(f);
    Buf *all[] = {&sigma, &vis, &level, &dep, &inv};
    for (Buf *b : all) pool_give(c, *b);
    return rc;
}

// ---- Bellman-Ford ---------------------------------------------------------

constexpr int32_t kDistInf = 2147483647 / 2; // INT_MAX/2 (BellmanFord.C:105)

__global__ void k_sssp_init(int32_t *__restrict__ dist, int32_t *__restrict__ vis, uint32_t root, uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    dist[v] = v == root ? 0 : kDistInf;
    vis[v] = 0;
}

// BF_Vertex_F (BellmanFord.C:94-103) over a round's output
__global__ void k_sssp_unvisit(const uint8_t *__restrict__ F, int32_t *__restrict__ vis, uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n && F[v]) vis[v] = 0;
}

__global__ void k_sssp_negcycle(int32_t *__restrict__ dist, uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) dist[v] = -kDistInf;
}

} // namespace gg

using namespace gg;

extern "C" int gg_bc(gg_graph_t g, uint32_t root, int dtype, void *scores_dev, int32_t *levels_dev, gg_run_info *info)
{
    if (!g || !scores_dev) return fail(GG_ERR_INVALID, "gg_bc: NULL argument");
    if (root >= g->n) return fail(GG_ERR_INVALID, "gg_bc: root out of range");
    if (dtype != GG_F64 && dtype != GG_F32) return fail(GG_ERR_INVALID, "gg_bc: dtype must be F64 or F32");
    GG_HIP(hipSetDevice(g->ctx->device));
    if (info) std::memset(info, 0, sizeof(*info));
    // With a renumbered plan both phases run on the plan's relabelled
    // arrays; scores and levels are written back in the caller's numbering.
    Graph t;
    uint32_t r = root;
    const uint32_t *to_caller = nullptr;
    Graph *run = g;
    if (g->plan.mode != 0 && g->m) {
        GG_TRY(plan_alias(g, t));
        int rc = read_back(g->ctx, &r, (const uint32_t *)g->plan.perm.p + root, 4);
        if (rc) {
            plan_alias_release(t);
            return rc;
        }
        to_caller = (const uint32_t *)g->plan.inv.p;
        run = &t;
    }
    const int rc = dtype == GG_F64 ? bc_impl<double>(run, r, (double *)scores_dev, levels_dev, info, to_caller)
                                   : bc_impl<float>(run, r, (float *)scores_dev, levels_dev, info, to_caller);
    if (run == &t) plan_alias_release(t);
    return rc;
}

// BellmanFord.C Compute (108-142) on g from 'root'.
static int sssp_impl(Graph *g, uint32_t root, int32_t *dist_dev, gg_run_info *info)
{
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    Buf vis;
    int rc = pool_take(c, vis, 4ull * n);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const uint32_t nb = blocks_for(n, kBlock);
    int64_t round = 0;
    EdgeMapExtra ex;
    Frontier *F = nullptr;
    if (!rc) {
        (void)hipEventRecord(e0, c->stream);
        k_sssp_init<<<nb, kBlock, 0, c->stream>>>(dist_dev, (int32_t *)vis.p, root, n);
        // partitioned_vertices::create(n, start, ...) (BellmanFord.C:121)
        F = new Frontier();
        frontier_alloc(g, F);
        uint32_t od = 0;
        rc = pool_take(c, F->list, 8);
        if (!rc) rc = root_frontier(c, g, root, F);
        if (!rc) rc = read_back(c, &od, (uint32_t *)F->list.p + 1, 4);
        F->d_m = 1;
        F->out_edges = od;
        F->deg_src = g->outdeg.p;
        F->sparse_valid = true;
        // rounds (BellmanFord.C:124-139): each takes the reference's branch
        // by the threshold m/20 (atomicMin push or the Jacobi pull)
        while (!rc && F->d_m != 0) {
            if (round == (int64_t)n) { // negative cycle (BellmanFord.C:126-132)
                k_sssp_negcycle<<<nb, kBlock, 0, c->stream>>>(dist_dev, n);
                break;
            }
            gg_op_t op = {GG_OP_SSSP, GG_I32, dist_dev, vis.p};
            Frontier *out = nullptr;
            int path = 0;
            if ((rc = edgemap_core(g, F, &op, -1, 0, &out, &path, &ex))) break;
            if (info && round < 64) info->paths[round] = path;
            if (out->d_m) {
                if ((rc = frontier_make_dense(out))) {
                    frontier_delete(out);
                    break;
                }
                k_sssp_unvisit<<<nb, kBlock, 0, c->stream>>>((const uint8_t *)out->bytes.p, (int32_t *)vis.p, n);
            }
            frontier_delete(F);
            F = out;
            round++;
        }
        (void)hipEventRecord(e1, c->stream);
        if (!rc && hipGetLastError() != hipSuccess) rc = fail(GG_ERR_HIP, "gg_sssp: launch");
        if (!rc) rc = sync(c);
    }
    float ms = 0;
    if (!rc) (void)hipEventElapsedTime(&ms, e0, e1);
    if (!rc && info) {
        info->iters = (int32_t)round;
        info->ms = ms;
        info->edges_scanned = ex.scanned;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    frontier_delete(F);
    pool_give(c, vis);
    return rc;
}

extern "C" int gg_sssp(gg_graph_t g, uint32_t root, int32_t *dist_dev, gg_run_info *info)
{
    if (!g || !dist_dev) return fail(GG_ERR_INVALID, "gg_sssp: NULL argument");
    if (root >= g->n) return fail(GG_ERR_INVALID, "gg_sssp: root out of range");
    Ctx *c = g->ctx;
    GG_HIP(hipSetDevice(c->device));
    if (info) std::memset(info, 0, sizeof(*info));
    if (g->plan.mode == 0 || g->m == 0) return sssp_impl(g, root, dist_dev, info);
    // With a renumbered plan the rounds run on the plan's relabelled arrays
    // (weights carried) and the distances are mapped back.
    Graph t;
    Buf dp;
    uint32_t r = 0;
    int rc = plan_alias(g, t);
    if (!rc) rc = pool_take(c, dp, 4ull * g->n);
    if (!rc) rc = read_back(c, &r, (const uint32_t *)g->plan.perm.p + root, 4);
    if (!rc) rc = sssp_impl(&t, r, (int32_t *)dp.p, info);
    if (!rc) rc = gather_idx(c, (const int32_t *)dp.p, (const uint32_t *)g->plan.perm.p, dist_dev, g->n);
    if (!rc) rc = sync(c);
    plan_alias_release(t);
    pool_give(c, dp);
    return rc;
}
