// gg_apps.hip -- the four apps' Compute() loops on the device.
//
// gg_pagerank runs PageRank.C's frontier-gated iteration (app/PageRank.C:
// 102-151) with edgeMap, both vertexMaps, the L1 reduction and the reset
// fused into one destination-owned pull per iteration:
//   x[d]  += sum over active in-neighbours s of c[s]/outdeg(s)   (PR_F)
//   x[d]   = 0.85*x[d] + 0.15/n            if d in F               (PR_Vertex_F)
//   L1    += |c[d] - x[d]|;  c'[d] = d in F ? 0 : |c[d]-x[d]|       (136-142)
//   next   = {d : some active in-neighbour}; contributions for the next
//            iteration written in the same epilogue
// The three rank buffers rotate (c, x, c') -> (x, c', free) exactly as
// swap(p_curr, p_next) does, with the reset copy kept apart so no kernel
// reads an array it writes.  A device flag carries the L1 < eps break.
#include <cstring>

#include "gg_kernels.h"

namespace gg {

template <class V>
struct OpPRFused {
    using acc_t = V;
    const V *contrib_in;
    V *contrib_out;
    const V *c;
    V *x;
    V *y;
    const uint8_t *Fin;
    int bit;
    uint8_t *Fout;
    const uint32_t *outdeg;
    V damping, added;
    const int *done;
    __device__ bool skip() const { return *done != 0; }
    __device__ V identity() const { return -V(0); }
    __device__ V combine(V a, V b) const { return a + b; }
    __device__ V gather(uint32_t, uint32_t s) const { return contrib_in[s]; }
    // per-destination inputs, prefetched with the gathers; every per-vertex
    // array is touched once per iteration, so it is streamed past L2 to
    // keep the gathered contributions resident
    struct row_t {
        V x, c;
        uint32_t od;
        uint8_t f;
    };
    __device__ row_t load_row(uint32_t d) const
    {
        row_t r;
        r.x = __builtin_nontemporal_load(&x[d]);
        r.c = __builtin_nontemporal_load(&c[d]);
        r.od = __builtin_nontemporal_load(&outdeg[d]);
        // unconditional load (a branch here would drain the load queue);
        // Fin always points at a valid array, 'bit' forces membership
        r.f = __builtin_nontemporal_load(&Fin[d]) | (uint8_t)bit;
        return r;
    }
    __device__ void store_row(uint32_t d, const row_t &r, V acc, LocalStats &st) const
    {
        V xv = r.x;
        const bool any = !signbit(acc);
        if (any) xv = xv + acc;
        const bool inF = r.f != 0;
        if (inF) xv = damping * xv + added;
        const V diff = fabs(r.c - xv);
        st.l1 += (double)diff;
        __builtin_nontemporal_store(inF ? V(0) : diff, &y[d]);
        __builtin_nontemporal_store(xv, &x[d]);
        __builtin_nontemporal_store((uint8_t)any, &Fout[d]);
        __builtin_nontemporal_store((any && r.od) ? xv / (V)r.od : -V(0), &contrib_out[d]);
        if (any) {
            st.active++;
            st.out_edges += r.od;
        }
    }
};

// The rank vector of the last executed iteration (known on the device from
// the stop flags) into the caller's array and numbering: ranks[v] =
// R[it % 3][perm ? perm[v] : v], it = the stopping iteration or max_iter.
template <class V>
__global__ void k_pr_result(V *__restrict__ out, const V *r0, const V *r1, const V *r2,
                            const uint32_t *__restrict__ perm, const int *__restrict__ flags, int max_iter, uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int it = flags[0] ? flags[1] : max_iter;
    const V *in = it % 3 == 0 ? r0 : it % 3 == 1 ? r1 : r2;
    out[v] = in[perm ? perm[v] : v];
}

template <class V>
__global__ void k_pr_init(V *c, V *x, V *contrib, const uint32_t *__restrict__ outdeg, uint32_t n, V one_over_n)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    c[v] = one_over_n;
    x[v] = V(0);
    const uint32_t od = outdeg[v];
    contrib[v] = od ? one_over_n / (V)od : -V(0);
}

// Per-iteration bookkeeping in one block: L1 over the tile partials in index
// order, the break flag, and the frontier statistics of this iteration.
__global__ __launch_bounds__(kBlock) void k_pr_check(const double *__restrict__ l1_part, uint32_t np,
                                                     const double *__restrict__ l1_fix, uint32_t nf,
                                                     double eps, int iter, int *done, int *done_iter,
                                                     unsigned long long *counters,
                                                     unsigned long long *stats, double *l1_hist)
{
    if (*done) return;
    double s = 0;
    (void)l1_part;
    (void)np; // folded into l1_fix by k_pull_fixup
    for (uint32_t i = threadIdx.x; i < nf; i += kBlock) s += l1_fix[i];
    s = wave_sum_f64(s);
    // this iteration's frontier statistics: the counter stripes (the first
    // wave reads one stripe per lane and clears it)
    unsigned long long a = 0, o = 0;
    if (threadIdx.x < kStripes) {
        unsigned long long *st = counters + kCntStripeBase + threadIdx.x * kStripeStride;
        a = st[CNT_ACTIVE];
        o = st[CNT_OUT];
        st[CNT_ACTIVE] = 0;
        st[CNT_OUT] = 0;
    }
    a = wave_sum_u64(a);
    o = wave_sum_u64(o);
    __shared__ double w[kBlock / 64];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < kBlock / 64; k++) t += w[k];
        l1_hist[iter - 1] = t;
        stats[2 * (iter - 1)] = a;
        stats[2 * (iter - 1) + 1] = o;
        if (t < eps) {
            *done = 1;
            *done_iter = iter;
        }
    }
}

// The fused loop's buffers, views of the graph's workspace (allocated on the
// first call and kept: per-call allocation of seven n-sized arrays would
// cost more than a PageRank iteration).
struct PRBuffers {
    Buf *R, *K, *F, &l1p, &l1f, &flags, &stats, &l1h;
    explicit PRBuffers(Graph *g)
        : R(g->pr_ws), K(g->pr_ws + 3), F(g->pr_ws + 5), l1p(g->pr_ws[7]), l1f(g->pr_ws[8]),
          flags(g->pr_ws[9]), stats(g->pr_ws[10]), l1h(g->pr_ws[11])
    {
    }
};

template <class V>
static int pagerank_fused(Graph *g, int max_iter, double damping, double eps, void *ranks, gg_run_info *info,
                          bool *needs_generic)
{
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const bool use_plan = g->plan.mode != 0;
    CscView view;
    if (use_plan) view = plan_view(g);
    else GG_TRY(pull_view(g, &view));
    if (!use_plan && g->dense_coo && g->coo_src.p && g->P && !g->part_vertex && m) {
        // GG_OPT_DENSE_COO: every iteration takes the COO branch, whose
        // in-lists keep each destination's edges in COO (Hilbert) order
        GG_TRY(ensure_coo_pull(g));
        view.tids = (const uint32_t *)g->coo_tids.p;
    }
    const uint32_t *outdeg = use_plan ? (const uint32_t *)g->plan.outdeg.p : (const uint32_t *)g->outdeg.p;
    // the source-segmented pull (GG_OPT_XCD_SEGMENTS): each XCD gathers from
    // its own slice of the contributions, a combine pass finishes the rows
    // (sources below 2^31, and the gathered vector below 2^31 bytes for the
    // partial pass's buffer loads)
    const bool xseg = use_plan && m && (uint64_t)n * sizeof(V) < (1ull << 31) &&
                      (g->xcd_segments == 2 || (g->xcd_segments == 1 && m >= (1ull << 24)));
    if (xseg) GG_TRY(ensure_xseg(g));
    XSeg &X = g->xseg;
    const uint32_t xhot = kXHotBytes / sizeof(V); // sources the partial pass caches in LDS
    if (xseg)
        GG_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_xseg_part<SegSum<V>>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)(xhot * sizeof(V))));
    const uint32_t nblk = view.blocks();
    const uint32_t nfix = xseg ? std::max(view.fix_blocks(), X.comb_blocks) : view.fix_blocks();
    PRBuffers B(g);
    int rc = GG_OK;
    for (int i = 0; i < 3 && !rc; i++) rc = buf_ensure(c, B.R[i], sizeof(V) * (uint64_t)n);
    for (int i = 0; i < 2 && !rc; i++) rc = buf_ensure(c, B.K[i], sizeof(V) * (uint64_t)n);
    for (int i = 0; i < 2 && !rc; i++) rc = buf_ensure(c, B.F[i], (uint64_t)n + 64);
    if (!rc) rc = buf_ensure(c, B.l1p, 8ull * nblk);
    if (!rc) rc = buf_ensure(c, B.l1f, 8ull * nfix);
    if (!rc) rc = buf_ensure(c, B.flags, 64);
    if (!rc) rc = buf_ensure(c, B.stats, 16ull * (max_iter + 1));
    if (!rc) rc = buf_ensure(c, B.l1h, 8ull * (max_iter + 1));
    if (rc) return rc;
    int *done = (int *)B.flags.p;
    int *done_iter = done + 1;
    unsigned long long *cnt = (unsigned long long *)c->counters.p;
    const V one_over_n = (V)(1 / (double)n);
    const V added = (V)((1 - damping) * (1 / (double)n)); // PageRank.C:80
    const int nev = 2 * max_iter + 2;
    std::vector<hipEvent_t> ev(nev);
    for (auto &e : ev) (void)hipEventCreate(&e);
    // timed region: init + iterations (the reference times Compute() whole)
    GG_HIP(hipEventRecord(ev[0], c->stream));
    fill_bytes(c, B.flags.p, 64, 0);
    counters_zero(c);
    k_pr_init<V><<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>((V *)B.R[0].p, (V *)B.R[1].p, (V *)B.K[0].p,
                                                                   outdeg, n, one_over_n);
    fill_bytes(c, B.F[1].p, n, 1);
    for (int it = 1; it <= max_iter; it++) {
        OpPRFused<V> op;
        op.c = (const V *)B.R[(it - 1) % 3].p;
        op.x = (V *)B.R[it % 3].p;
        op.y = (V *)B.R[(it + 1) % 3].p;
        op.contrib_in = (const V *)B.K[(it - 1) % 2].p;
        op.contrib_out = (V *)B.K[it % 2].p;
        op.bit = it == 1;
        // iteration 1 runs on bits(): its flags are not read (bit forces them)
        op.Fin = (const uint8_t *)B.F[it == 1 ? 1 : (it - 2) % 2].p;
        op.Fout = (uint8_t *)B.F[(it - 1) % 2].p;
        op.outdeg = outdeg;
        op.damping = (V)damping;
        op.added = added;
        op.done = done;
        if (xseg) {
            // group partials (each group's items on one XCD, its hottest
            // sources in LDS), long-row chunks, then the combine with the
            // loop's epilogue
            const uint32_t *meta = (const uint32_t *)X.meta.p;
            OpSegPartial<SegSum<V>> sop{(V *)X.partial.p, done};
            GG_HIP(hipEventRecord(ev[2 * it], c->stream));
            k_xseg_part<SegSum<V>><<<X.grid, kXPartThreads, xhot * sizeof(V), c->stream>>>(
                op.contrib_in, (V *)X.partial.p, (V *)X.rows.slot_val.p, (const uint32_t *)X.vtids.p,
                (const RowWork *)X.rows.work.p, meta, n, xhot, done, nullptr);
            if (X.rows.nlong)
                k_rows_fix<OpSegPartial<SegSum<V>>><<<blocks_for(X.rows.nlong, kBlock), kBlock, 0, c->stream>>>(
                    sop, (const uint32_t *)X.rows.long_rows.p, X.rows.nlong, (const V *)X.rows.slot_val.p, 0,
                    nullptr, nullptr, nullptr);
            k_xseg_combine<OpPRFused<V>><<<X.comb_blocks, kBlock, 0, c->stream>>>(
                op, (const V *)X.partial.p, (const uint32_t *)X.mask.p, (const uint32_t *)X.cursor.p,
                meta + kXSegs + 1, X.vrows, n, cnt, (double *)B.l1f.p);
            GG_HIP(hipEventRecord(ev[2 * it + 1], c->stream));
            GG_HIP(hipGetLastError());
            k_pr_check<<<1, kBlock, 0, c->stream>>>((const double *)B.l1p.p, 0, (const double *)B.l1f.p,
                                                     X.comb_blocks, eps, it, done, done_iter, cnt,
                                                     (unsigned long long *)B.stats.p, (double *)B.l1h.p);
            continue;
        }
        rc = launch_pull(c, view, op, cnt, (double *)B.l1p.p, (double *)B.l1f.p, ev[2 * it], ev[2 * it + 1]);
        if (rc) break;
        k_pr_check<<<1, kBlock, 0, c->stream>>>((const double *)B.l1p.p, nblk, (const double *)B.l1f.p, nfix,
                                                 eps, it, done, done_iter, cnt,
                                                 (unsigned long long *)B.stats.p, (double *)B.l1h.p);
    }
    // the rank vector in the caller's numbering (un-permuted with a plan);
    // the kernel picks the last executed iteration's buffer from the flags
    if (!rc)
        k_pr_result<V><<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>(
            (V *)ranks, (const V *)B.R[0].p, (const V *)B.R[1].p, (const V *)B.R[2].p,
            use_plan ? (const uint32_t *)g->plan.perm.p : nullptr, done, max_iter, n);
    GG_HIP(hipEventRecord(ev[1], c->stream));
    GG_HIP(hipGetLastError());
    int hflags[2] = {0, 0};
    std::vector<unsigned long long> st(2ull * (max_iter + 1));
    if (!rc) rc = read_back(c, hflags, B.flags.p, 8);
    if (!rc) {
        const size_t sb = 16ull * max_iter;
        rc = sb <= kPinBytes ? read_back(c, st.data(), B.stats.p, sb) : gg_memcpy_d2h(abi(c), st.data(), B.stats.p, sb);
    }
    int iters = hflags[0] ? hflags[1] : max_iter;
    if (!rc) {
        float ms = 0, kms = 0, tot = 0;
        (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
        for (int it = 1; it <= iters; it++) {
            (void)hipEventElapsedTime(&kms, ev[2 * it], ev[2 * it + 1]);
            tot += kms;
        }
        c->last_kernel_ms = tot;
        c->last_kernel_launches = iters;
        if (info) {
            info->iters = iters;
            info->ms = ms;
            info->edges_scanned = (uint64_t)iters * m;
            for (int it = 1; it <= iters && it <= 64; it++) info->paths[it - 1] = GG_PATH_COO;
        }
        // The fused loop takes the dense branch every iteration.  Check the
        // decisions edgeMap would have made (ligra-numa.h:947,953): iteration
        // it > 1 reads the statistics of iteration it-1's output.
        const int64_t thr = (int64_t)(m / 20), dthr = (int64_t)(m / 2);
        for (int it = 2; it <= iters; it++) {
            const uint64_t dm = st[2 * (it - 2)], oe = st[2 * (it - 2) + 1];
            if (oe == 0 || (int64_t)(dm + oe) <= thr) *needs_generic = true;
            if (info && it <= 64)
                info->paths[it - 1] = (int64_t)(dm + oe) > dthr && g->P && !g->part_vertex ? GG_PATH_COO
                                                                                           : GG_PATH_CSC;
        }
        if (info && iters >= 1)
            info->paths[0] = g->P && !g->part_vertex ? GG_PATH_COO : GG_PATH_CSC;
        // with needs_generic set, the caller reruns the generic loop, which
        // overwrites 'ranks'
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}

// PageRank.C's loop through the generic edgeMap/vertexMap entry points; used
// when a frontier turns small enough for the sparse branch, whose duplicate
// outputs the fused loop does not model.
template <class V>
static int pagerank_generic(Graph *g, int max_iter, double damping, double eps, void *ranks, gg_run_info *info)
{
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    const int dt = sizeof(V) == 8 ? GG_F64 : GG_F32;
    Buf pc, pn, red;
    GG_TRY(pool_take(c, pc, sizeof(V) * (uint64_t)n));
    GG_TRY(pool_take(c, pn, sizeof(V) * (uint64_t)n));
    GG_TRY(pool_take(c, red, 8));
    GG_TRY(gg_fill(abi(c), pc.p, dt, n, 1 / (double)n));
    GG_TRY(gg_fill(abi(c), pn.p, dt, n, 0));
    Frontier *F = new Frontier();
    frontier_alloc(g, F);
    F->bit = true;
    F->has_dense = true;
    F->d_m = n;
    F->out_edges = g->m;
    int count = 0, rc = GG_OK;
    bool broke = false;
    while (count < max_iter) {
        count++;
        gg_op_t op = {GG_OP_PAGERANK, dt, pc.p, pn.p};
        Frontier *out = nullptr;
        int path = 0;
        if ((rc = edgemap_core(g, F, &op, (int64_t)(g->m / 20), 0, &out, &path, nullptr))) break;
        if (info && count <= 64) info->paths[count - 1] = path;
        gg_vop_t v1 = {GG_VOP_PAGERANK, dt, pc.p, pn.p, damping, (1 - damping) * (1 / (double)n)};
        if ((rc = vertexmap_core(g, F, &v1))) break;
        gg_vop_t v2 = {GG_VOP_ABSDIFF, dt, pc.p, pn.p, 0, 0};
        if ((rc = vertexmap_core(g, nullptr, &v2))) break;
        double L1 = 0;
        if ((rc = reduce_sum_core(c, pc.p, dt, n, (double *)red.p))) break;
        if ((rc = gg_memcpy_d2h(abi(c), &L1, red.p, 8))) break;
        if (L1 < eps) {
            broke = true;
            frontier_delete(out);
            break;
        }
        gg_vop_t v3 = {GG_VOP_RESET, dt, pc.p, nullptr, 0, 0};
        if ((rc = vertexmap_core(g, F, &v3))) break;
        std::swap(pc, pn);
        frontier_delete(F);
        F = out;
        if (F->n == 0 && count < max_iter) {
            rc = fail(GG_ERR_SIZE, "PageRank: edgeMap returned empty() and the next call has mismatched sizes "
                                   "(ligra-numa.h:937-942)");
            break;
        }
    }
    if (!rc) rc = gg_memcpy_d2d(abi(c), ranks, broke ? pn.p : pc.p, sizeof(V) * (uint64_t)n);
    if (info) {
        info->iters = count;
        info->edges_scanned = (uint64_t)count * g->m;
    }
    frontier_delete(F);
    pool_give(c, pc);
    pool_give(c, pn);
    pool_give(c, red);
    return rc;
}

} // namespace gg

using namespace gg;

extern "C" int gg_pagerank(gg_graph_t g, int dtype, int max_iter, double damping, double eps, void *ranks_dev,
                           gg_run_info *info)
{
    if (!g || !ranks_dev) return fail(GG_ERR_INVALID, "gg_pagerank: NULL argument");
    if (max_iter < 1 || max_iter > 100000) return fail(GG_ERR_INVALID, "gg_pagerank: max_iter out of range");
    if (dtype != GG_F64 && dtype != GG_F32) return fail(GG_ERR_INVALID, "gg_pagerank: dtype must be F64 or F32");
    GG_HIP(hipSetDevice(g->ctx->device));
    if (info) std::memset(info, 0, sizeof(*info));
    bool generic = false;
    int rc = dtype == GG_F64 ? pagerank_fused<double>(g, max_iter, damping, eps, ranks_dev, info, &generic)
                             : pagerank_fused<float>(g, max_iter, damping, eps, ranks_dev, info, &generic);
    if (rc || !generic) return rc;
    return dtype == GG_F64 ? pagerank_generic<double>(g, max_iter, damping, eps, ranks_dev, info)
                           : pagerank_generic<float>(g, max_iter, damping, eps, ranks_dev, info);
}

extern "C" int gg_bfs(gg_graph_t g, uint32_t root, int64_t threshold, int32_t *parents, int32_t *levels,
                      gg_run_info *info)
{
    if (!g || !parents) return fail(GG_ERR_INVALID, "gg_bfs: NULL argument");
    if (root >= g->n) return fail(GG_ERR_INVALID, "gg_bfs: root out of range");
    Ctx *c = g->ctx;
    GG_HIP(hipSetDevice(c->device));
    if (info) std::memset(info, 0, sizeof(*info));
    // app/BFS.C:86-101: Parents[root] = root, the root's frontier, edgeMap
    // until the frontier is empty, each level's direction chosen on the
    // device.  With a renumbered plan the levels run on the plan's
    // relabelled arrays (every in-list starts with its largest out-degree,
    // the likeliest frontier member, and vertices of like degree share
    // waves); the loop's first level past the end maps parents and levels
    // back to the caller's numbering.
    const bool planned = g->plan.mode != 0 && g->m;
    Buf pp, lp;
    Graph t;
    uint32_t r2 = root;
    int rc = ensure_first_in(g);
    if (!rc && planned) {
        rc = pool_take(c, pp, 4ull * g->n);
        if (!rc && levels) rc = pool_take(c, lp, 4ull * g->n);
        if (!rc) rc = plan_alias(g, t);
        if (!rc) rc = read_back(c, &r2, (const uint32_t *)g->plan.perm.p + root, 4);
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, c->stream);
    if (!rc && planned) {
        const BfsMapBack mb{parents, levels, (const uint32_t *)g->plan.perm.p, (const uint32_t *)g->plan.inv.p};
        rc = bfs_device_loop(&t, r2, (const uint32_t *)g->plan.first_in.p, threshold, (int32_t *)pp.p,
                             levels ? (int32_t *)lp.p : nullptr, info, &mb, &g->bfs_levels_hint);
    } else if (!rc) {
        rc = bfs_device_loop(g, root, (const uint32_t *)g->first_in.p, threshold, parents, levels, info, nullptr,
                             &g->bfs_levels_hint);
    }
    (void)hipEventRecord(e1, c->stream);
    if (!rc) rc = sync(c);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (planned) plan_alias_release(t);
    pool_give(c, pp);
    pool_give(c, lp);
    if (info) info->ms = ms;
    return rc;
}

extern "C" int gg_cc(gg_graph_t g, uint32_t *labels, gg_run_info *info)
{
    if (!g || !labels) return fail(GG_ERR_INVALID, "gg_cc: NULL argument");
    Ctx *c = g->ctx;
    GG_HIP(hipSetDevice(c->device));
    if (info) std::memset(info, 0, sizeof(*info));
    // Components.C's loop (101-126) on the device, each round's branch
    // chosen there.  With a renumbered plan the loop runs on the plan's
    // relabelled arrays (the dense rounds gather labels of hot vertices
    // from a narrow range); label values stay the caller's vertex ids, so
    // the minimum of each component, the fixed point, is the same, and
    // positions are mapped back in the round after the last.
    const bool planned = g->plan.mode != 0 && g->m;
    Buf prev, lp;
    GG_TRY(pool_take(c, prev, 4ull * g->n));
    if (planned) {
        int r = pool_take(c, lp, 4ull * g->n);
        if (r) {
            pool_give(c, prev);
            return r;
        }
    }
    Graph t;
    int rc = ensure_first_in(g);
    if (!rc && planned) rc = plan_alias(g, t);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, c->stream);
    if (!rc) {
        if (planned) { // Components.C:113 IDs[i] = i: the caller's ids in the new numbering
            // the round after the last maps the labels back
            rc = cc_device_loop(&t, (const uint32_t *)g->plan.first_in.p, (const uint32_t *)g->plan.inv.p,
                                (uint32_t *)lp.p, (uint32_t *)prev.p, info, &g->cc_rounds_hint,
                                (const uint32_t *)g->plan.perm.p, labels);
        } else {
            rc = cc_device_loop(g, (const uint32_t *)g->first_in.p, nullptr, labels, (uint32_t *)prev.p, info,
                                &g->cc_rounds_hint);
        }
    }
    (void)hipEventRecord(e1, c->stream);
    if (!rc) rc = sync(c);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (planned) plan_alias_release(t);
    pool_give(c, prev);
    pool_give(c, lp);
    if (info) info->ms = ms;
    return rc;
}

namespace gg {
template <class V>
struct OpSpmvY {
    using acc_t = V;
    const V *x;
    const float *tw;
    V *y;
    const uint32_t *inv = nullptr; // relabelled twin: row d is vertex inv[d]
    __device__ bool skip() const { return false; }
    __device__ V identity() const { return V(0); }
    __device__ V combine(V a, V b) const { return a + b; }
    __device__ V gather(uint32_t e, uint32_t s) const { return tw ? x[s] * (V)tw[e] : x[s]; }
    __device__ void finish(uint32_t d, V acc, LocalStats &) const { y[inv ? inv[d] : d] = acc; }
};
} // namespace gg

namespace gg {
// SpMV over the source-segmented form of the plan's in-edges: the group
// partials (each group's items on one XCD, its hottest sources of x in LDS,
// each term x[s] times the edge's weight on a weighted graph), the long
// rows' chunks, then the combine, which adds each row's partials in group
// order and writes y in the caller's numbering.  x is in the plan's
// numbering.
template <class V>
int spmv_xseg(Graph *g, const V *x, V *y)
{
    Ctx *c = g->ctx;
    XSeg &X = g->xseg;
    const uint32_t n = g->n;
    const uint32_t xhot = kXHotBytes / sizeof(V);
    const bool weighted = X.vw.p != nullptr;
    auto part = weighted ? &k_xseg_part<SegSum<V>, true> : &k_xseg_part<SegSum<V>, false>;
    GG_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(part), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)(xhot * sizeof(V))));
    const uint32_t *meta = (const uint32_t *)X.meta.p;
    part<<<X.grid, kXPartThreads, xhot * sizeof(V), c->stream>>>(
        x, (V *)X.partial.p, (V *)X.rows.slot_val.p, (const uint32_t *)X.vtids.p, (const RowWork *)X.rows.work.p,
        meta, n, xhot, nullptr, (const float *)X.vw.p);
    if (X.rows.nlong)
        k_rows_fix<OpSegPartial<SegSum<V>>><<<blocks_for(X.rows.nlong, kBlock), kBlock, 0, c->stream>>>(
            OpSegPartial<SegSum<V>>{(V *)X.partial.p, nullptr}, (const uint32_t *)X.rows.long_rows.p, X.rows.nlong,
            (const V *)X.rows.slot_val.p, 0, nullptr, nullptr, nullptr);
    k_xseg_combine<OpSpmvY<V>><<<X.comb_blocks, kBlock, 0, c->stream>>>(
        OpSpmvY<V>{x, nullptr, y, (const uint32_t *)g->plan.inv.p}, (const V *)X.partial.p,
        (const uint32_t *)X.mask.p, (const uint32_t *)X.cursor.p, meta + kXSegs + 1, X.vrows, n, nullptr, nullptr);
    GG_HIP(hipGetLastError());
    return GG_OK;
}
} // namespace gg

extern "C" int gg_spmv(gg_graph_t g, int dtype, const void *x, void *y, gg_run_info *info)
{
    if (!g || !x || !y) return fail(GG_ERR_INVALID, "gg_spmv: NULL argument");
    if (dtype != GG_F64 && dtype != GG_F32) return fail(GG_ERR_INVALID, "gg_spmv: dtype must be F64 or F32");
    Ctx *c = g->ctx;
    GG_HIP(hipSetDevice(c->device));
    if (info) std::memset(info, 0, sizeof(*info));
    const float *tw = g->weighted ? (const float *)g->tw.p : nullptr;
    CscView pv;
    // With a renumbered plan the pull runs over the plan's relabelled twin
    // (rows of similar in-degree together, hot sources together): x is
    // permuted into the new numbering first and y scattered back per row.
    const bool planned = g->plan.mode != 0 && g->m;
    // the source-segmented pull, as PageRank.C's loop takes it (sources
    // below 2^31, and x below 2^31 bytes for the partial pass's buffer loads)
    const bool xseg = planned && (uint64_t)g->n * (dtype == GG_F64 ? 8ull : 4ull) < (1ull << 31) &&
                      (g->xcd_segments == 2 || (g->xcd_segments == 1 && g->m >= (1ull << 24)));
    if (xseg) GG_TRY(ensure_xseg(g));
    const uint32_t *inv = nullptr;
    Buf xp;
    if (planned) {
        pv = plan_view(g);
        if (g->weighted) tw = (const float *)g->plan.tw.p;
        inv = (const uint32_t *)g->plan.inv.p;
        GG_TRY(pool_take(c, xp, (dtype == GG_F64 ? 8ull : 4ull) * g->n));
    } else {
        GG_TRY(pull_view(g, &pv));
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, c->stream);
    int rc = GG_OK;
    if (planned) {
        rc = dtype == GG_F64 ? gather_idx(c, (const double *)x, inv, (double *)xp.p, g->n)
                             : gather_idx(c, (const float *)x, inv, (float *)xp.p, g->n);
        x = xp.p;
    }
    if (!rc && xseg)
        rc = dtype == GG_F64 ? spmv_xseg<double>(g, (const double *)x, (double *)y)
                             : spmv_xseg<float>(g, (const float *)x, (float *)y);
    else if (!rc)
        rc = dtype == GG_F64
                 ? launch_pull(c, pv, OpSpmvY<double>{(const double *)x, tw, (double *)y, inv}, nullptr, nullptr,
                               nullptr)
                 : launch_pull(c, pv, OpSpmvY<float>{(const float *)x, tw, (float *)y, inv}, nullptr, nullptr,
                               nullptr);
    (void)hipEventRecord(e1, c->stream);
    if (!rc) rc = sync(c);
    pool_give(c, xp);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->last_kernel_ms = ms;
    c->last_kernel_launches = 1;
    if (info) {
        info->iters = 1;
        info->ms = ms;
        info->edges_scanned = g->m;
        info->paths[0] = GG_PATH_CSC;
    }
    return rc;
}

// ---------------------------------------------------------------------------
// gg_spmv_noalloc: gg_spmv without any device allocation, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops).  The
// one scratch buffer gg_spmv takes from the pool, x permuted into a
// renumbered plan's order, lives in a file-scope device array instead.
// ---------------------------------------------------------------------------
namespace gg {
namespace {
constexpr uint64_t kScratchBytes = 64ull << 20; // n * sizeof(value) at most
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
// per-device address of 'scratch' (each GPU holds its own copy); the host
// name of the array is only a stand-in and is never passed to a kernel
void *g_scratch_ptr[kScratchDevices] = {};
std::mutex g_scratch_mu;

int scratch_for(int device, void **out)
{
    if (device < 0 || device >= kScratchDevices)
        return fail(GG_ERR_UNSUPPORTED, "gg_spmv_noalloc: device ordinal above " + std::to_string(kScratchDevices - 1));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        // set-up on this device's first call, with the device current
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GG_ERR_HIP, std::string("gg_spmv_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return GG_OK;
}
} // namespace
} // namespace gg

extern "C" int gg_spmv_noalloc(gg_graph_t g, int dtype, const void *x, void *y)
{
    if (!g || !x || !y) return fail(GG_ERR_INVALID, "gg_spmv_noalloc: NULL argument");
    if (dtype != GG_F64 && dtype != GG_F32) return fail(GG_ERR_INVALID, "gg_spmv_noalloc: dtype must be F64 or F32");
    Ctx *c = g->ctx;
    GG_HIP(hipSetDevice(c->device));
    const uint64_t vbytes = (dtype == GG_F64 ? 8ull : 4ull) * g->n;
    const bool planned = g->plan.mode != 0 && g->m;
    // the path gg_spmv takes (the same condition), so the result is its own
    const bool xseg = planned && vbytes < (1ull << 31) &&
                      (g->xcd_segments == 2 || (g->xcd_segments == 1 && g->m >= (1ull << 24)));
    const float *tw = g->weighted ? (const float *)g->tw.p : nullptr;
    CscView pv;
    const uint32_t *inv = nullptr;
    void *xp = nullptr;
    if (planned) {
        if (vbytes > kScratchBytes)
            return fail(GG_ERR_SIZE, "gg_spmv_noalloc: n * value size = " + std::to_string(vbytes) +
                                         " bytes exceeds the " + std::to_string(kScratchBytes) +
                                         "-byte scratch array; use gg_spmv");
        // the segmented layout is built by the first segmented pull on the
        // plan (an allocation): it must exist already
        if (xseg && !g->xseg.ok)
            return fail(GG_ERR_STATE, "gg_spmv_noalloc: the plan's source-segmented layout is not built yet "
                                      "(one gg_spmv call builds it)");
        GG_TRY(scratch_for(c->device, &xp));
        pv = plan_view(g);
        if (g->weighted) tw = (const float *)g->plan.tw.p;
        inv = (const uint32_t *)g->plan.inv.p;
    } else {
        // the row-block schedule is built by the first pull on the graph
        // (an allocation): it must exist already
        if (g->pull_schedule == 1 && !g->rows_ok && g->m)
            return fail(GG_ERR_STATE, "gg_spmv_noalloc: the graph's row-block schedule is not built yet "
                                      "(one gg_spmv call builds it)");
        GG_TRY(pull_view(g, &pv));
    }
    int rc = GG_OK;
    if (planned) {
        rc = dtype == GG_F64 ? gather_idx(c, (const double *)x, inv, (double *)xp, g->n)
                             : gather_idx(c, (const float *)x, inv, (float *)xp, g->n);
        x = xp;
    }
    if (!rc && xseg)
        rc = dtype == GG_F64 ? spmv_xseg<double>(g, (const double *)x, (double *)y)
                             : spmv_xseg<float>(g, (const float *)x, (float *)y);
    else if (!rc)
        rc = dtype == GG_F64
                 ? launch_pull(c, pv, OpSpmvY<double>{(const double *)x, tw, (double *)y, inv}, nullptr, nullptr,
                               nullptr)
                 : launch_pull(c, pv, OpSpmvY<float>{(const float *)x, tw, (float *)y, inv}, nullptr, nullptr,
                               nullptr);
    return rc;
}
