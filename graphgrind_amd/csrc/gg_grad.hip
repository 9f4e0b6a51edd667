// gg_grad.hip -- differentiated forms of the traversal engine.
//
// The reference has no gradients; these entry points stand beside its SpMV
// (app/SPMV.C:7-92) and all-active PageRank (app/PageRankBit.C:145) on the
// same weighted graph, and reuse the engine's two traversal directions:
//
//   forward   pull over the in-edge twin (CSC): one owner per destination
//   adjoint   pull over the out-edge CSR: one owner per source, the
//             transpose traversal, with the same merge-path tiles and the
//             same fixed combination order (no floating-point atomics)
//   dL/dw     one edge-parallel pass over the CSR tiles, written in the
//             caller's (CSR) edge order
//
// SpMV:      y[d] = sum_{e=(s->d)} w_e x[s];  gx[s] = sum_e w_e gy[d];
//            gw[e] = x[s] gy[d]
// PageRank:  r <- (1-a) v + a P r with P[d][s] = w_{s->d} / S_s (row-
//            normalised; dangling sources contribute nothing), from r = v
//            until sum|r_new - r| < eps.  Implicit adjoint at r for
//            dL/dr = u: lambda <- u + a P^T lambda until
//            sum|lambda_new - lambda| <= eps sum|lambda_new|; then
//            dL/dw_e = (a r_s / S_s) (lambda_d - (lambda_s - u_s)/a),
//            dL/dv   = (1-a) lambda.
#include <algorithm>
#include <cstring>

#include "gg_kernels.h"

namespace gg {

// ---------------------------------------------------------------------------
// Edge-parallel pass over merge-path tiles of a CSR-like array: every edge
// of tile t is handled once, with coalesced index loads and stores; the row
// an edge belongs to comes from a binary search over the tile's row ends in
// LDS, whose per-row inputs (erow) are staged there too.
//   Op: erow_t erow(row); gval_t gval(edge, nbr); edge(edge, erow, gval)
// ---------------------------------------------------------------------------
template <class Op>
__global__ __launch_bounds__(kBlock) void k_edge(Op op, const uint32_t *__restrict__ off,
                                                 const uint32_t *__restrict__ ids,
                                                 const uint32_t *__restrict__ srow,
                                                 const uint32_t *__restrict__ sedge, uint32_t rows)
{
    using ER = typename Op::erow_t;
    using GV = typename Op::gval_t;
    __shared__ uint32_t s_end[kPullIPB + 1];
    __shared__ ER s_row[kPullIPB + 1];
    const uint32_t tid = threadIdx.x, t = blockIdx.x;
    const uint32_t r0 = srow[t], r1 = srow[t + 1];
    const uint32_t e0 = sedge[t], e1 = sedge[t + 1];
    const uint32_t ne = e1 - e0;
    if (ne == 0) return;
    // rows r0 .. r1 hold this tile's edges (r1 only when it exists); a tile
    // with edges has r0 < rows, and off[r0 + 1] >= e0 on the merge path
    const uint32_t nst = min(r1 - r0 + 1, rows - r0);
    for (uint32_t j = tid; j < nst; j += kBlock) {
        s_end[j] = off[r0 + j + 1] - e0;
        s_row[j] = op.erow(r0 + j);
    }
    __syncthreads();
    const uint32_t nel = ne - 1;
    uint32_t nb[kPullIPT];
    GV gv[kPullIPT];
#pragma unroll
    for (int k = 0; k < kPullIPT; k++) nb[k] = __builtin_nontemporal_load(&ids[e0 + min(tid + k * kBlock, nel)]);
#pragma unroll
    for (int k = 0; k < kPullIPT; k++) gv[k] = op.gval(e0 + min(tid + k * kBlock, nel), nb[k]);
#pragma unroll
    for (int k = 0; k < kPullIPT; k++) {
        const uint32_t j = tid + k * kBlock;
        if (j < ne) {
            // first staged row whose end lies past edge j (the last staged
            // row always does)
            uint32_t lo = 0, hi = nst - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_end[mid] > j) hi = mid;
                else lo = mid + 1;
            }
            op.edge(e0 + j, s_row[lo], gv[k]);
        }
    }
}

// S[s] = sum of s's out-edge weights (FP64, one owner per source)
struct OpWSum {
    using acc_t = double;
    const float *w;
    double *S;
    __device__ bool skip() const { return false; }
    __device__ double identity() const { return 0.0; }
    __device__ double combine(double a, double b) const { return a + b; }
    __device__ double gather(uint32_t e, uint32_t) const { return (double)w[e]; }
    __device__ void finish(uint32_t s, double acc, LocalStats &) const { S[s] = acc; }
};

// gx[s] = sum_{e=(s->d)} w_e gy[d]
template <class V>
struct OpSpmvBwdX {
    using acc_t = V;
    const float *w; // NULL: unit weights
    const V *gy;
    V *gx;
    __device__ bool skip() const { return false; }
    __device__ V identity() const { return V(0); }
    __device__ V combine(V a, V b) const { return a + b; }
    const uint32_t *inv = nullptr; // relabelled CSR: row s is vertex inv[s]
    __device__ V gather(uint32_t e, uint32_t d) const
    {
        return w ? (V)__builtin_nontemporal_load(&w[e]) * gy[d] : gy[d];
    }
    __device__ void finish(uint32_t s, V acc, LocalStats &) const { gx[inv ? inv[s] : s] = acc; }
};

// Both outputs in one pass over the out-edge CSR (row-block schedule):
// while staging edge e = (s -> d), gw[e] = x[s] gy[d] is written (coalesced,
// in edge order) and w_e gy[d] is staged for gx[s] (the same products in the
// same order as OpSpmvBwdX), so the random gy gathers are made once for both.
// Over the plan's relabelled CSR (x and gy in the new numbering), edge e of
// row s is the caller's edge e + delta[s] and row s is vertex inv[s].
template <class V>
struct OpSpmvBwdFused {
    using acc_t = V;
    const float *w; // NULL: unit weights
    const V *gy, *x;
    V *gx, *gw;
    const uint32_t *delta = nullptr, *inv = nullptr;
    __device__ bool skip() const { return false; }
    __device__ V identity() const { return V(0); }
    __device__ V combine(V a, V b) const { return a + b; }
    __device__ V gather(uint32_t e, uint32_t d) const { return w ? (V)w[e] * gy[d] : gy[d]; }
    __device__ V stage_edge(uint32_t e, uint32_t s, uint32_t d) const
    {
        const V g = gy[d];
        __builtin_nontemporal_store(x[s] * g, &gw[delta ? e + delta[s] : e]);
        return w ? (V)__builtin_nontemporal_load(&w[e]) * g : g;
    }
    __device__ void finish(uint32_t s, V acc, LocalStats &) const { gx[inv ? inv[s] : s] = acc; }
};

// gw[e] = x[s] gy[d]
template <class V>
struct EdgeSpmvGW {
    using erow_t = V;
    using gval_t = V;
    const V *x, *gy;
    V *gw;
    __device__ V erow(uint32_t s) const { return x[s]; }
    __device__ V gval(uint32_t, uint32_t d) const { return gy[d]; }
    __device__ void edge(uint32_t e, V xs, V gd) const { __builtin_nontemporal_store(xs * gd, &gw[e]); }
};

// One forward iteration: r_new[d] = beta v[d] + alpha sum_{e=(s->d)} w_e c[s]
// with c = r/S; r is updated in place (each row has one owner, and the
// gathered values are the separate contribution array).
template <class V, bool kV>
struct OpWPR {
    using acc_t = V;
    const float *tw;
    const V *cin;
    V *cout;
    V *r;
    const V *v;
    V vconst;
    const double *S;
    V alpha, beta;
    const int *done;
    __device__ bool skip() const { return *done != 0; }
    __device__ V identity() const { return V(0); }
    __device__ V combine(V a, V b) const { return a + b; }
    __device__ V gather(uint32_t e, uint32_t s) const { return (V)__builtin_nontemporal_load(&tw[e]) * cin[s]; }
    struct row_t {
        V rold, vd;
        double S;
    };
    __device__ row_t load_row(uint32_t d) const
    {
        row_t q;
        q.rold = __builtin_nontemporal_load(&r[d]);
        if constexpr (kV) q.vd = __builtin_nontemporal_load(&v[d]);
        else q.vd = vconst;
        q.S = __builtin_nontemporal_load(&S[d]);
        return q;
    }
    __device__ void store_row(uint32_t d, const row_t &q, V acc, LocalStats &st) const
    {
        const V rn = beta * q.vd + alpha * acc;
        st.l1 += (double)fabs(rn - q.rold);
        __builtin_nontemporal_store(rn, &r[d]);
        cout[d] = q.S > 0 ? rn / (V)q.S : V(0);
    }
};

// One adjoint iteration over the CSR: lambda_new[s] = u[s] + alpha (sum_e
// w_e lambda[d]) / S[s]; aux accumulates |lambda_new| for the relative rule.
template <class V>
struct OpWPRAdj {
    static constexpr bool kAux = true;
    using acc_t = V;
    const float *w;
    const V *lin;
    V *lout;
    const V *u;
    const double *S;
    V alpha;
    const int *done;
    __device__ bool skip() const { return *done != 0; }
    __device__ V identity() const { return V(0); }
    __device__ V combine(V a, V b) const { return a + b; }
    __device__ V gather(uint32_t e, uint32_t d) const { return (V)__builtin_nontemporal_load(&w[e]) * lin[d]; }
    struct row_t {
        V lold, ud;
        double S;
    };
    __device__ row_t load_row(uint32_t s) const
    {
        row_t q;
        q.lold = lin[s];
        q.ud = __builtin_nontemporal_load(&u[s]);
        q.S = __builtin_nontemporal_load(&S[s]);
        return q;
    }
    __device__ void store_row(uint32_t s, const row_t &q, V acc, LocalStats &st) const
    {
        const V ln = q.S > 0 ? q.ud + (alpha * acc) / (V)q.S : q.ud;
        st.l1 += (double)fabs(ln - q.lold);
        st.aux += (double)fabs(ln);
        lout[s] = ln;
    }
};

// dL/dw_e = coef_s (lambda_d - mu_s), coef_s = alpha r_s / S_s,
// mu_s = (lambda_s - u_s) / alpha
template <class V>
struct EdgeWPRGrad {
    struct erow_t {
        V coef, mu;
    };
    using gval_t = V;
    const V *r, *lam, *u;
    const double *S;
    V alpha;
    V *gw;
    __device__ erow_t erow(uint32_t s) const
    {
        const double Ss = S[s];
        const V rs = r[s], ls = lam[s], us = u[s];
        erow_t q;
        q.coef = Ss > 0 ? (alpha * rs) / (V)Ss : V(0);
        q.mu = Ss > 0 ? (ls - us) / alpha : V(0);
        return q;
    }
    __device__ V gval(uint32_t, uint32_t d) const { return lam[d]; }
    __device__ void edge(uint32_t e, const erow_t &q, V ld) const
    {
        __builtin_nontemporal_store(q.coef * (ld - q.mu), &gw[e]);
    }
};

template <class V>
__global__ void k_wpr_init(V *__restrict__ r, V *__restrict__ cin, const V *__restrict__ v, V vconst,
                           const double *__restrict__ S, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V ri = v ? v[i] : vconst;
    const double Si = S[i];
    r[i] = ri;
    cin[i] = Si > 0 ? ri / (V)Si : V(0);
}

template <class V>
__global__ void k_scale(V *__restrict__ out, const V *__restrict__ in, V a, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a * in[i];
}

// Convergence check of one iteration: the step's L1 (and, for the relative
// rule, the new iterate's L1) from the per-fixup-block partials, summed in
// a fixed order; sets flags[0] and the stopping iteration in flags[1].
__global__ __launch_bounds__(kBlock) void k_fp_check(const double *__restrict__ l1_fix, uint32_t nf, int relative,
                                                     double eps, int iter, int *flags, double *hist)
{
    if (flags[0]) return;
    double s = 0, a = 0;
    for (uint32_t i = threadIdx.x; i < nf; i += kBlock) {
        s += l1_fix[i];
        if (relative) a += l1_fix[nf + i];
    }
    s = wave_sum_f64(s);
    a = wave_sum_f64(a);
    __shared__ double ws[kBlock / 64], wa[kBlock / 64];
    if ((threadIdx.x & 63) == 0) {
        ws[threadIdx.x >> 6] = s;
        wa[threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ts = 0, ta = 0;
        for (int k = 0; k < kBlock / 64; k++) {
            ts += ws[k];
            ta += wa[k];
        }
        hist[2 * (iter - 1)] = ts;
        hist[2 * (iter - 1) + 1] = ta;
        const bool stop = relative ? ts <= eps * ta : ts < eps;
        if (stop) {
            flags[0] = 1;
            flags[1] = iter;
        }
    }
}

// ---- lazily built transpose-side state ----------------------------------------


static int ensure_wsum(Graph *g)
{
    if (g->wsum_ok) return GG_OK;
    if (!g->weighted) return fail(GG_ERR_STATE, "graph has no edge weights");
    CscView cv;
    GG_TRY(csr_pull_view(g, &cv));
    GG_TRY(buf_ensure(g->ctx, g->wsum, 8ull * g->n));
    GG_TRY(launch_pull(g->ctx, cv, OpWSum{(const float *)g->w.p, (double *)g->wsum.p}, nullptr, nullptr,
                       nullptr));
    g->wsum_ok = true;
    return GG_OK;
}

template <class Op>
static int launch_edge(Ctx *c, Graph *g, const Op &op)
{
    const PullSched &s = g->csr_pull;
    k_edge<Op><<<s.tiles, kBlock, 0, c->stream>>>(op, (const uint32_t *)g->off.p, (const uint32_t *)g->ids.p,
                                                   (const uint32_t *)s.start_row.p, (const uint32_t *)s.start_edge.p,
                                                   g->n);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

// Events of one fixed-point solve: [0] start, [1] end, [2it], [2it+1]
// around iteration it's pull.
struct SolveEvents {
    std::vector<hipEvent_t> ev;
    explicit SolveEvents(int max_iter) : ev(2 * (size_t)max_iter + 2)
    {
        for (auto &e : ev) (void)hipEventCreate(&e);
    }
    ~SolveEvents()
    {
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
};

// Runs iterations 1..max_iter of a pull whose operator 'make(it)' builds,
// with the check after each; iterations are enqueued in growing chunks and
// the stop flag is read between chunks (launches after the stop skip).
template <class Make>
static int run_fixed_point(Graph *g, const CscView &view, int max_iter, int relative, double eps, Make make,
                           SolveEvents &E, int *iters, std::vector<double> *hist_out)
{
    Ctx *c = g->ctx;
    const uint32_t nfix = view.fix_blocks();
    Buf l1p, l1f, flags, hist;
    int rc = pool_take(c, l1p, 16ull * view.blocks());
    if (!rc) rc = pool_take(c, l1f, 16ull * nfix);
    if (!rc) rc = pool_take(c, flags, 64);
    if (!rc) rc = pool_take(c, hist, 16ull * max_iter);
    int done_iter = max_iter;
    if (!rc) {
        fill_bytes(c, flags.p, 64, 0);
        int it = 0, chunk = 8;
        int hflags[2] = {0, 0};
        while (it < max_iter && !rc) {
            const int k1 = std::min(max_iter, it + chunk);
            while (it < k1 && !rc) {
                it++;
                rc = launch_pull(c, view, make(it, (const int *)flags.p), nullptr, (double *)l1p.p,
                                 (double *)l1f.p, E.ev[2 * it], E.ev[2 * it + 1]);
                if (rc) break;
                k_fp_check<<<1, kBlock, 0, c->stream>>>((const double *)l1f.p, nfix, relative, eps, it,
                                                         (int *)flags.p, (double *)hist.p);
            }
            if (rc) break;
            GG_HIP(hipGetLastError());
            rc = read_back(c, hflags, flags.p, 8);
            if (!rc && hflags[0]) {
                done_iter = hflags[1];
                break;
            }
            chunk = std::min(chunk * 2, 32);
        }
        if (!rc) {
            *iters = done_iter;
            if (hist_out) {
                hist_out->assign(2 * (size_t)done_iter, 0.0);
                rc = gg_memcpy_d2h(abi(c), hist_out->data(), hist.p, 16ull * done_iter);
            }
        }
    }
    pool_give(c, l1p);
    pool_give(c, l1f);
    pool_give(c, flags);
    pool_give(c, hist);
    return rc;
}

static void record_times(Ctx *c, SolveEvents &E, int iters, gg_run_info *info, uint64_t m)
{
    float ms = 0, kms = 0, tot = 0;
    (void)hipEventElapsedTime(&ms, E.ev[0], E.ev[1]);
    for (int it = 1; it <= iters; it++) {
        (void)hipEventElapsedTime(&kms, E.ev[2 * it], E.ev[2 * it + 1]);
        tot += kms;
    }
    c->last_kernel_ms = tot;
    c->last_kernel_launches = iters;
    if (info) {
        info->iters = iters;
        info->ms = ms;
        info->edges_scanned = (uint64_t)iters * m;
    }
}

template <class V>
__global__ void k_copy(V *__restrict__ out, const V *__restrict__ in, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// The numbering a weighted PageRank solve runs in: the graph's own, or the
// renumbered plan (its relabelled in-edge twin for the forward, its
// relabelled CSR for the adjoint: rows of similar degree together, hot
// sources together), with vertex vectors permuted in and results back.
struct WprSpace {
    CscView view;
    const float *w = nullptr;                       // weights in the view's edge order
    const uint32_t *perm = nullptr, *inv = nullptr; // set when planned
    bool planned() const { return perm != nullptr; }
};

static int wpr_space(Graph *g, bool adjoint, WprSpace *sp)
{
    *sp = WprSpace{};
    if (g->plan.mode && g->m) {
        if (adjoint) {
            GG_TRY(plan_csr_view(g, &sp->view));
            sp->w = (const float *)g->plan.cw.p;
        } else {
            sp->view = plan_view(g);
            sp->w = (const float *)g->plan.tw.p;
        }
        sp->perm = (const uint32_t *)g->plan.perm.p;
        sp->inv = (const uint32_t *)g->plan.inv.p;
        return GG_OK;
    }
    if (adjoint) {
        GG_TRY(csr_pull_view(g, &sp->view));
        sp->w = (const float *)g->w.p;
    } else {
        GG_TRY(pull_view(g, &sp->view));
        sp->w = (const float *)g->tw.p;
    }
    return GG_OK;
}

// The weight sums and one vertex vector x (may be NULL) in the space's
// numbering: the graph's own arrays, or permuted copies in bs / bx (n
// doubles / n values, allocated by the caller when planned).
template <class V>
static int wpr_inputs(Graph *g, const WprSpace &sp, const V *x, const Buf &bs, const Buf &bx, const double **S,
                      const V **xs)
{
    *S = (const double *)g->wsum.p;
    *xs = x;
    if (!sp.planned()) return GG_OK;
    GG_TRY(gather_idx(g->ctx, (const double *)g->wsum.p, sp.inv, (double *)bs.p, g->n));
    *S = (const double *)bs.p;
    if (x) {
        GG_TRY(gather_idx(g->ctx, x, sp.inv, (V *)bx.p, g->n));
        *xs = (const V *)bx.p;
    }
    return GG_OK;
}

template <class V, bool kV>
static OpWPR<V, kV> wpr_op(const WprSpace &sp, const double *S, const Buf *K, int it, V *r, const V *v, V vconst,
                           double alpha, const int *done)
{
    OpWPR<V, kV> op;
    op.tw = sp.w;
    op.cin = (const V *)K[(it - 1) % 2].p;
    op.cout = (V *)K[it % 2].p;
    op.r = r;
    op.v = v;
    op.vconst = vconst;
    op.S = S;
    op.alpha = (V)alpha;
    op.beta = (V)(1 - alpha);
    op.done = done;
    return op;
}

template <class V>
static OpWPRAdj<V> adj_op(const WprSpace &sp, const double *S, const Buf *L, int it, const V *u, double alpha,
                          const int *done)
{
    OpWPRAdj<V> op;
    op.w = sp.w;
    op.lin = (const V *)L[(it - 1) % 2].p;
    op.lout = (V *)L[it % 2].p;
    op.u = u;
    op.S = S;
    op.alpha = (V)alpha;
    op.done = done;
    return op;
}

static uint64_t dbits(double x)
{
    uint64_t b;
    std::memcpy(&b, &x, 8);
    return b;
}

// Captures what 'body' enqueues on the context stream into X (a hipGraph and
// its executable); the stream leaves capture mode on every path.
template <class Body>
static int capture(Ctx *c, Body body, GraphExec &X)
{
    GG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = body();
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &gr);
    if (rc != GG_OK || e != hipSuccess) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc != GG_OK ? rc : fail(GG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    const hipError_t e2 = hipGraphInstantiate(&X.exec, gr, nullptr, nullptr, 0);
    if (e2 != hipSuccess) {
        (void)hipGraphDestroy(gr);
        X.exec = nullptr;
        return fail(GG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
    }
    X.graph = gr;
    return GG_OK;
}

// Replays a captured loop between two events; ms = the replay's device time.
static int replay(Ctx *c, GraphExec &X, float *ms)
{
    hipEvent_t e0, e1;
    GG_HIP(hipEventCreate(&e0));
    GG_HIP(hipEventCreate(&e1));
    int rc = GG_OK;
    if (hipEventRecord(e0, c->stream) != hipSuccess || hipGraphLaunch(X.exec, c->stream) != hipSuccess ||
        hipEventRecord(e1, c->stream) != hipSuccess)
        rc = fail(GG_ERR_HIP, "hipGraphLaunch");
    if (!rc) rc = sync(c);
    if (!rc) (void)hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

static void fixed_info(Ctx *c, float ms, int iters, uint64_t m, gg_run_info *info)
{
    c->last_kernel_ms = ms;
    c->last_kernel_launches = iters;
    if (info) {
        info->iters = iters;
        info->ms = ms;
        info->edges_scanned = (uint64_t)iters * m;
    }
}

// Exactly 'iters' forward iterations (no stop rule), captured once per
// (buffers, parameters) and replayed as one hipGraph launch.
template <class V>
static int wpagerank_fixed(Graph *g, const WprSpace &sp, double alpha, int iters, const V *v, V *r,
                           gg_run_info *info)
{
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    const V vconst = (V)(1 / (double)n);
    GraphExec &X = g->fwd_exec;
    const uint64_t key[10] = {1, sizeof(V), (uint64_t)(uintptr_t)r, (uint64_t)(uintptr_t)v, dbits(alpha),
                              (uint64_t)iters, sp.planned() ? 1u : 0u, 0, 0, 0};
    if (!X.exec || std::memcmp(X.key, key, sizeof key) != 0) {
        graph_exec_free(X);
        // own: [0], [1] contributions, [2] flags; planned: [3] ranks,
        // [4] weight sums, [5] teleport vector in the plan's numbering
        GG_TRY(buf_alloc(c, X.own[0], sizeof(V) * (uint64_t)n));
        GG_TRY(buf_alloc(c, X.own[1], sizeof(V) * (uint64_t)n));
        GG_TRY(buf_alloc(c, X.own[2], 64));
        if (sp.planned()) {
            GG_TRY(buf_alloc(c, X.own[3], sizeof(V) * (uint64_t)n));
            GG_TRY(buf_alloc(c, X.own[4], 8ull * n));
            if (v) GG_TRY(buf_alloc(c, X.own[5], sizeof(V) * (uint64_t)n));
        }
        GG_TRY(fill_bytes(c, X.own[2].p, 64, 0));
        GG_TRY(sync(c));
        const int *done = (const int *)X.own[2].p;
        auto body = [&]() -> int {
            const double *S;
            const V *vv;
            GG_TRY(wpr_inputs(g, sp, v, X.own[4], X.own[5], &S, &vv));
            V *rr = sp.planned() ? (V *)X.own[3].p : r;
            k_wpr_init<V><<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>(rr, (V *)X.own[0].p, vv, vconst, S, n);
            for (int it = 1; it <= iters; it++) {
                const int rc = v ? launch_pull(c, sp.view, wpr_op<V, true>(sp, S, X.own, it, rr, vv, vconst, alpha, done),
                                               nullptr, nullptr, nullptr)
                                 : launch_pull(c, sp.view, wpr_op<V, false>(sp, S, X.own, it, rr, vv, vconst, alpha, done),
                                               nullptr, nullptr, nullptr);
                if (rc) return rc;
            }
            if (sp.planned()) GG_TRY(gather_idx(c, (const V *)rr, sp.perm, r, n));
            return GG_OK;
        };
        const int rc = capture(c, body, X);
        if (rc) {
            graph_exec_free(X);
            return rc;
        }
        std::memcpy(X.key, key, sizeof key);
    }
    float ms = 0;
    GG_TRY(replay(c, X, &ms));
    fixed_info(c, ms, iters, g->m, info);
    return GG_OK;
}

template <class V>
static int wpagerank_impl(Graph *g, double alpha, double eps, int max_iter, const V *v, V *r, gg_run_info *info)
{
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    GG_TRY(ensure_wsum(g));
    WprSpace sp;
    GG_TRY(wpr_space(g, false, &sp));
    if (eps < 0) return wpagerank_fixed<V>(g, sp, alpha, max_iter, v, r, info);
    SolveEvents E(max_iter);
    Buf K[2], T[3]; // T (planned): ranks, weight sums, teleport vector in the plan's numbering
    int rc = pool_take(c, K[0], sizeof(V) * (uint64_t)n);
    if (!rc) rc = pool_take(c, K[1], sizeof(V) * (uint64_t)n);
    if (!rc && sp.planned()) {
        rc = pool_take(c, T[0], sizeof(V) * (uint64_t)n);
        if (!rc) rc = pool_take(c, T[1], 8ull * n);
        if (!rc && v) rc = pool_take(c, T[2], sizeof(V) * (uint64_t)n);
    }
    const V vconst = (V)(1 / (double)n);
    if (!rc) {
        (void)hipEventRecord(E.ev[0], c->stream);
        const double *S = nullptr;
        const V *vv = nullptr;
        rc = wpr_inputs(g, sp, v, T[1], T[2], &S, &vv);
        V *rr = sp.planned() ? (V *)T[0].p : r;
        int iters = 0;
        if (!rc) {
            k_wpr_init<V><<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>(rr, (V *)K[0].p, vv, vconst, S, n);
            auto make_v = [&](int it, const int *done) {
                return wpr_op<V, true>(sp, S, K, it, rr, vv, vconst, alpha, done);
            };
            auto make_u = [&](int it, const int *done) {
                return wpr_op<V, false>(sp, S, K, it, rr, vv, vconst, alpha, done);
            };
            rc = v ? run_fixed_point(g, sp.view, max_iter, 0, eps, make_v, E, &iters, nullptr)
                   : run_fixed_point(g, sp.view, max_iter, 0, eps, make_u, E, &iters, nullptr);
        }
        if (!rc && sp.planned()) rc = gather_idx(c, (const V *)rr, sp.perm, r, n);
        if (!rc) {
            (void)hipEventRecord(E.ev[1], c->stream);
            rc = sync(c);
        }
        if (!rc) record_times(c, E, iters, info, g->m);
    }
    pool_give(c, K[0]);
    pool_give(c, K[1]);
    for (Buf &b : T) pool_give(c, b);
    return rc;
}

// Enqueues the gradient outputs from the converged adjoint 'lam'.
template <class V>
static int adjoint_outputs(Graph *g, double alpha, const V *r, const V *u, const V *lam, V *gw, V *gv, V *lam_out)
{
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    if (gw && g->m) {
        EdgeWPRGrad<V> op;
        op.r = r;
        op.lam = lam;
        op.u = u;
        op.S = (const double *)g->wsum.p;
        op.alpha = (V)alpha;
        op.gw = gw;
        GG_TRY(launch_edge(c, g, op));
    }
    if (gv) k_scale<V><<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>(gv, lam, (V)(1 - alpha), n);
    if (lam_out) k_copy<V><<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>(lam_out, lam, n);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

template <class V>
static int wpagerank_adjoint_fixed(Graph *g, const WprSpace &sp, double alpha, int iters, const V *r, const V *u,
                                   V *gw, V *gv, V *lam_out, gg_run_info *info)
{
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    GraphExec &X = g->adj_exec;
    const uint64_t key[10] = {2, sizeof(V), (uint64_t)(uintptr_t)r, (uint64_t)(uintptr_t)u,
                              (uint64_t)(uintptr_t)gw, (uint64_t)(uintptr_t)gv, (uint64_t)(uintptr_t)lam_out,
                              dbits(alpha), (uint64_t)iters, sp.planned() ? 1u : 0u};
    if (!X.exec || std::memcmp(X.key, key, sizeof key) != 0) {
        graph_exec_free(X);
        // own: [0], [1] adjoint iterates, [2] flags; planned: [3] weight
        // sums, [4] seed vector u, [5] the solution in the caller's numbering
        GG_TRY(buf_alloc(c, X.own[0], sizeof(V) * (uint64_t)n));
        GG_TRY(buf_alloc(c, X.own[1], sizeof(V) * (uint64_t)n));
        GG_TRY(buf_alloc(c, X.own[2], 64));
        if (sp.planned()) {
            GG_TRY(buf_alloc(c, X.own[3], 8ull * n));
            GG_TRY(buf_alloc(c, X.own[4], sizeof(V) * (uint64_t)n));
            GG_TRY(buf_alloc(c, X.own[5], sizeof(V) * (uint64_t)n));
        }
        GG_TRY(fill_bytes(c, X.own[2].p, 64, 0));
        GG_TRY(sync(c));
        const int *done = (const int *)X.own[2].p;
        auto body = [&]() -> int {
            const double *S;
            const V *uu;
            GG_TRY(wpr_inputs(g, sp, u, X.own[3], X.own[4], &S, &uu));
            k_copy<V><<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>((V *)X.own[0].p, uu, n);
            for (int it = 1; it <= iters; it++)
                GG_TRY(launch_pull(c, sp.view, adj_op<V>(sp, S, X.own, it, uu, alpha, done), nullptr, nullptr,
                                   nullptr));
            const V *lam = (const V *)X.own[iters % 2].p;
            if (sp.planned()) {
                GG_TRY(gather_idx(c, lam, sp.perm, (V *)X.own[5].p, n));
                lam = (const V *)X.own[5].p;
            }
            return adjoint_outputs<V>(g, alpha, r, u, lam, gw, gv, lam_out);
        };
        const int rc = capture(c, body, X);
        if (rc) {
            graph_exec_free(X);
            return rc;
        }
        std::memcpy(X.key, key, sizeof key);
    }
    float ms = 0;
    GG_TRY(replay(c, X, &ms));
    fixed_info(c, ms, iters, g->m, info);
    return GG_OK;
}

template <class V>
static int wpagerank_adjoint_impl(Graph *g, double alpha, double eps, int max_iter, const V *r, const V *u, V *gw,
                                  V *gv, V *lam_out, gg_run_info *info)
{
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    GG_TRY(ensure_wsum(g));
    GG_TRY(ensure_csr_sched(g)); // the dL/dw pass runs over the caller's CSR
    WprSpace sp;
    GG_TRY(wpr_space(g, true, &sp));
    if (eps < 0) return wpagerank_adjoint_fixed<V>(g, sp, alpha, max_iter, r, u, gw, gv, lam_out, info);
    SolveEvents E(max_iter);
    Buf L[2], T[3]; // T (planned): weight sums, u, the solution in the caller's numbering
    int rc = pool_take(c, L[0], sizeof(V) * (uint64_t)n);
    if (!rc) rc = pool_take(c, L[1], sizeof(V) * (uint64_t)n);
    if (!rc && sp.planned()) {
        rc = pool_take(c, T[0], 8ull * n);
        if (!rc) rc = pool_take(c, T[1], sizeof(V) * (uint64_t)n);
        if (!rc) rc = pool_take(c, T[2], sizeof(V) * (uint64_t)n);
    }
    if (!rc) {
        (void)hipEventRecord(E.ev[0], c->stream);
        const double *S = nullptr;
        const V *uu = nullptr;
        rc = wpr_inputs(g, sp, u, T[0], T[1], &S, &uu);
        int iters = 0;
        if (!rc) {
            k_copy<V><<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>((V *)L[0].p, uu, n);
            auto make = [&](int it, const int *done) { return adj_op<V>(sp, S, L, it, uu, alpha, done); };
            rc = run_fixed_point(g, sp.view, max_iter, 1, eps, make, E, &iters, nullptr);
        }
        const V *lam = (const V *)L[iters % 2].p;
        if (!rc && sp.planned()) {
            rc = gather_idx(c, lam, sp.perm, (V *)T[2].p, n);
            lam = (const V *)T[2].p;
        }
        if (!rc) rc = adjoint_outputs<V>(g, alpha, r, u, lam, gw, gv, lam_out);
        if (!rc) {
            (void)hipEventRecord(E.ev[1], c->stream);
            rc = sync(c);
        }
        if (!rc) record_times(c, E, iters, info, g->m);
    }
    pool_give(c, L[0]);
    pool_give(c, L[1]);
    for (Buf &b : T) pool_give(c, b);
    return rc;
}

} // namespace gg

using namespace gg;

static int check_args(gg_graph_t g, int dtype, const char *who)
{
    if (!g) return fail(GG_ERR_INVALID, std::string(who) + ": NULL graph");
    if (dtype != GG_F64 && dtype != GG_F32) return fail(GG_ERR_INVALID, std::string(who) + ": dtype must be F64 or F32");
    return GG_OK;
}

extern "C" int gg_spmv_bwd(gg_graph_t g, int dtype, const void *x, const void *gy, void *gx, void *gw,
                           gg_run_info *info)
{
    GG_TRY(check_args(g, dtype, "gg_spmv_bwd"));
    if (!gy || (gw && !x)) return fail(GG_ERR_INVALID, "gg_spmv_bwd: NULL argument");
    Ctx *c = g->ctx;
    GG_HIP(hipSetDevice(c->device));
    if (info) std::memset(info, 0, sizeof(*info));
    GG_TRY(ensure_csr_sched(g));
    CscView cv;
    GG_TRY(csr_pull_view(g, &cv));
    const float *w = g->weighted ? (const float *)g->w.p : nullptr;
    // both outputs from one pass when the row-block schedule runs the pull
    const bool fused = gx && gw && g->m && cv.rows;
    const bool f64 = dtype == GG_F64;
    const size_t vs = f64 ? 8 : 4;
    // With a renumbered plan the dL/dx pull runs over the plan's relabelled
    // CSR, with x and gy permuted into the new numbering first and gx
    // scattered back; the separate dL/dw pass stays on the caller's CSR.
    const bool planned = g->plan.mode != 0 && g->m && gx;
    const void *gy_in = gy, *x_in = x;
    Buf gyp, xp;
    const uint32_t *delta = nullptr, *inv = nullptr;
    if (planned) {
        GG_TRY(plan_csr_view(g, &cv));
        if (g->weighted) w = (const float *)g->plan.cw.p;
        inv = (const uint32_t *)g->plan.inv.p;
        delta = (const uint32_t *)g->plan.cdelta.p;
        GG_TRY(pool_take(c, gyp, vs * g->n));
        if (fused) {
            int rt = pool_take(c, xp, vs * g->n);
            if (rt) {
                pool_give(c, gyp);
                return rt;
            }
        }
    }
    const bool fused_now = fused && (!planned || cv.rows);
    hipEvent_t e0, e1, k0, k1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventCreate(&k0);
    (void)hipEventCreate(&k1);
    int rc = GG_OK;
    (void)hipEventRecord(e0, c->stream);
    if (planned) {
        rc = f64 ? gather_idx(c, (const double *)gy_in, inv, (double *)gyp.p, g->n)
                 : gather_idx(c, (const float *)gy_in, inv, (float *)gyp.p, g->n);
        if (!rc && fused_now)
            rc = f64 ? gather_idx(c, (const double *)x_in, inv, (double *)xp.p, g->n)
                     : gather_idx(c, (const float *)x_in, inv, (float *)xp.p, g->n);
        gy = gyp.p;
        if (fused_now) x = xp.p;
    }
    if (!rc && fused_now) {
        rc = f64 ? launch_pull(c, cv,
                               OpSpmvBwdFused<double>{w, (const double *)gy, (const double *)x, (double *)gx,
                                                      (double *)gw, delta, inv},
                               nullptr, nullptr, nullptr, k0, k1)
                 : launch_pull(c, cv,
                               OpSpmvBwdFused<float>{w, (const float *)gy, (const float *)x, (float *)gx,
                                                     (float *)gw, delta, inv},
                               nullptr, nullptr, nullptr, k0, k1);
    } else if (!rc && gx) {
        rc = f64 ? launch_pull(c, cv, OpSpmvBwdX<double>{w, (const double *)gy, (double *)gx, inv}, nullptr,
                               nullptr, nullptr, k0, k1)
                 : launch_pull(c, cv, OpSpmvBwdX<float>{w, (const float *)gy, (float *)gx, inv}, nullptr,
                               nullptr, nullptr, k0, k1);
    }
    if (!rc && gw && g->m && !fused_now) {
        rc = f64 ? launch_edge(c, g, EdgeSpmvGW<double>{(const double *)x_in, (const double *)gy_in, (double *)gw})
                 : launch_edge(c, g, EdgeSpmvGW<float>{(const float *)x_in, (const float *)gy_in, (float *)gw});
    }
    (void)hipEventRecord(e1, c->stream);
    if (!rc) rc = sync(c);
    pool_give(c, gyp);
    pool_give(c, xp);
    float ms = 0, kms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (gx) (void)hipEventElapsedTime(&kms, k0, k1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(k0);
    (void)hipEventDestroy(k1);
    c->last_kernel_ms = kms;
    c->last_kernel_launches = gx ? 1 : 0;
    if (info) {
        info->iters = 1;
        info->ms = ms;
        info->edges_scanned = g->m;
        info->paths[0] = GG_PATH_CSC;
    }
    return rc;
}

extern "C" int gg_wpagerank(gg_graph_t g, int dtype, double alpha, double eps, int max_iter, const void *v,
                            void *r, gg_run_info *info)
{
    GG_TRY(check_args(g, dtype, "gg_wpagerank"));
    if (!r) return fail(GG_ERR_INVALID, "gg_wpagerank: NULL rank array");
    if (max_iter < 1 || max_iter > 100000) return fail(GG_ERR_INVALID, "gg_wpagerank: max_iter out of range");
    if (!(alpha >= 0 && alpha < 1)) return fail(GG_ERR_INVALID, "gg_wpagerank: alpha must lie in [0, 1)");
    if (!g->weighted) return fail(GG_ERR_STATE, "gg_wpagerank: graph has no edge weights");
    GG_HIP(hipSetDevice(g->ctx->device));
    if (info) std::memset(info, 0, sizeof(*info));
    return dtype == GG_F64 ? wpagerank_impl<double>(g, alpha, eps, max_iter, (const double *)v, (double *)r, info)
                           : wpagerank_impl<float>(g, alpha, eps, max_iter, (const float *)v, (float *)r, info);
}

extern "C" int gg_wpagerank_adjoint(gg_graph_t g, int dtype, double alpha, double eps, int max_iter, const void *r,
                                    const void *u, void *gw, void *gv, void *lambda, gg_run_info *info)
{
    GG_TRY(check_args(g, dtype, "gg_wpagerank_adjoint"));
    if (!r || !u) return fail(GG_ERR_INVALID, "gg_wpagerank_adjoint: NULL argument");
    if (max_iter < 1 || max_iter > 100000) return fail(GG_ERR_INVALID, "gg_wpagerank_adjoint: max_iter out of range");
    if (!(alpha > 0 && alpha < 1)) return fail(GG_ERR_INVALID, "gg_wpagerank_adjoint: alpha must lie in (0, 1)");
    if (!g->weighted) return fail(GG_ERR_STATE, "gg_wpagerank_adjoint: graph has no edge weights");
    GG_HIP(hipSetDevice(g->ctx->device));
    if (info) std::memset(info, 0, sizeof(*info));
    return dtype == GG_F64
               ? wpagerank_adjoint_impl<double>(g, alpha, eps, max_iter, (const double *)r, (const double *)u,
                                                (double *)gw, (double *)gv, (double *)lambda, info)
               : wpagerank_adjoint_impl<float>(g, alpha, eps, max_iter, (const float *)r, (const float *)u,
                                               (float *)gw, (float *)gv, (float *)lambda, info);
}
