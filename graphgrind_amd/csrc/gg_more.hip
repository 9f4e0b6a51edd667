// gg_more.hip -- apps beyond the four on the same engine: betweenness
// contributions (app/BC.C) and Bellman-Ford distances (app/BellmanFord.C),
// each a loop of destination-owned pulls.
//
// BC from 'root' (BC.C:154-226).  Forward level L: every vertex not yet
// visited sums NumPaths over its in-neighbours in level L-1 and, if any,
// joins level L (BC_F 28-77; the sum starts at -0.0 so its sign says
// whether a frontier in-neighbour exists).  Backward on the transpose,
// deepest level first: a pull over the out-edge CSR sums Dependencies of
// out-neighbours in level r+1 into every vertex not yet visited (BC_Back_F
// 80-124), then level r adds inverseNumPaths (BC_Back_Vertex_F 141-152);
// finally score = (Dependencies - inv) / inv (222), NaN where unreached.
//
// Bellman-Ford from 'root' (BellmanFord.C:100-140), integer edge lengths:
// each round the vertices whose distance fell last round relax their
// out-edges; a pull takes, per vertex, the minimum of dist[s] + w over
// in-neighbours s in the frontier (read from a gated copy, so the round
// is a Jacobi step with a deterministic result).  n rounds with a
// non-empty frontier mean a negative cycle: every distance becomes
// -(INT_MAX/2), as the reference does.
#include <cstring>
#include <limits>

#include "gg_kernels.h"

namespace gg {

template <class V>
__global__ void k_bc_gate(const int32_t *__restrict__ level, const V *__restrict__ val, int32_t want, V none,
                          V *__restrict__ gate, uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) gate[v] = level[v] == want ? val[v] : none;
}

template <class V>
struct OpBCFwd {
    using acc_t = V;
    const V *gate; // NumPaths of level L-1, -0.0 elsewhere
    V *sigma;
    int32_t *level;
    int32_t L;
    __device__ bool skip() const { return false; }
    __device__ V identity() const { return -V(0); }
    __device__ V combine(V a, V b) const { return a + b; }
    __device__ V gather(uint32_t, uint32_t s) const { return gate[s]; }
    struct row_t {
        int32_t lvl;
    };
    __device__ row_t load_row(uint32_t d) const { return row_t{level[d]}; }
    __device__ void store_row(uint32_t d, const row_t &q, V acc, LocalStats &st) const
    {
        if (q.lvl < 0 && !signbit(acc)) {
            sigma[d] = acc; // NumPaths[d] was 0: NumPaths[d] += sum
            level[d] = L;
            st.active++;
        }
    }
};

template <class V>
struct OpBCBack {
    using acc_t = V;
    const V *gate; // Dependencies of level r+1, 0 elsewhere
    V *dep;
    const V *inv;
    const int32_t *level;
    int32_t r;
    __device__ bool skip() const { return false; }
    __device__ V identity() const { return V(0); }
    __device__ V combine(V a, V b) const { return a + b; }
    __device__ V gather(uint32_t, uint32_t s) const { return gate[s]; }
    struct row_t {
        int32_t lvl;
        V dep, inv;
    };
    __device__ row_t load_row(uint32_t d) const { return row_t{level[d], dep[d], inv[d]}; }
    __device__ void store_row(uint32_t d, const row_t &q, V acc, LocalStats &) const
    {
        // visited so far in the backward sweep: levels deeper than r
        if (q.lvl > r) return;
        V x = q.dep + acc;
        if (q.lvl == r) x = x + q.inv;
        dep[d] = x;
    }
};

template <class V>
__global__ void k_bc_init(V *__restrict__ sigma, int32_t *__restrict__ level, V *__restrict__ dep, uint32_t root,
                          uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    sigma[v] = v == root ? V(1) : V(0);
    level[v] = v == root ? 0 : -1;
    dep[v] = V(0);
}

template <class V>
__global__ void k_bc_inv(const V *__restrict__ sigma, const int32_t *__restrict__ level, int32_t deepest,
                         V *__restrict__ inv, V *__restrict__ dep, uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const V iv = V(1) / sigma[v];
    inv[v] = iv;
    if (level[v] == deepest) dep[v] = dep[v] + iv;
}

template <class V>
__global__ void k_bc_final(const V *__restrict__ dep, const V *__restrict__ inv, V *__restrict__ score, uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) score[v] = (dep[v] - inv[v]) / inv[v];
}

template <class V>
static int bc_impl(Graph *g, uint32_t root, V *score, int32_t *levels_out, gg_run_info *info)
{
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    CscView fwd, bwd;
    GG_TRY(pull_view(g, &fwd));
    GG_TRY(csr_pull_view(g, &bwd));
    Buf sigma, level, dep, inv, gate;
    int rc = pool_take(c, sigma, sizeof(V) * (uint64_t)n);
    if (!rc) rc = pool_take(c, level, 4ull * n);
    if (!rc) rc = pool_take(c, dep, sizeof(V) * (uint64_t)n);
    if (!rc) rc = pool_take(c, inv, sizeof(V) * (uint64_t)n);
    if (!rc) rc = pool_take(c, gate, sizeof(V) * (uint64_t)n);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    int round = 0;
    const uint32_t nb = blocks_for(n, kBlock);
    if (!rc) {
        (void)hipEventRecord(e0, c->stream);
        k_bc_init<V><<<nb, kBlock, 0, c->stream>>>((V *)sigma.p, (int32_t *)level.p, (V *)dep.p, root, n);
        unsigned long long *cnt = (unsigned long long *)c->counters.p;
        // forward: one pull per level until a level finds no vertex
        while (!rc) {
            round++;
            k_bc_gate<V><<<nb, kBlock, 0, c->stream>>>((const int32_t *)level.p, (const V *)sigma.p, round - 1,
                                                       -V(0), (V *)gate.p, n);
            rc = counters_zero(c);
            if (!rc)
                rc = launch_pull(c, fwd, OpBCFwd<V>{(const V *)gate.p, (V *)sigma.p, (int32_t *)level.p, round}, cnt,
                                 nullptr, nullptr);
            unsigned long long cn[4] = {0, 0, 0, 0};
            if (!rc) rc = counters_read(c, cn);
            if (!rc && info && round <= 64) info->paths[round - 1] = GG_PATH_CSC;
            if (!rc && cn[CNT_ACTIVE] == 0) break;
            if (!rc && round > (int)n) rc = fail(GG_ERR_STATE, "gg_bc: more levels than vertices");
        }
        // backward over the transpose, from the deepest level (round - 1)
        if (!rc) {
            k_bc_inv<V><<<nb, kBlock, 0, c->stream>>>((const V *)sigma.p, (const int32_t *)level.p, round - 1,
                                                      (V *)inv.p, (V *)dep.p, n);
            for (int r = round - 2; r >= 0 && !rc; r--) {
                k_bc_gate<V><<<nb, kBlock, 0, c->stream>>>((const int32_t *)level.p, (const V *)dep.p, r + 1, V(0),
                                                           (V *)gate.p, n);
                rc = launch_pull(c, bwd,
                                 OpBCBack<V>{(const V *)gate.p, (V *)dep.p, (const V *)inv.p,
                                             (const int32_t *)level.p, r},
                                 nullptr, nullptr, nullptr);
            }
        }
        if (!rc) {
            k_bc_final<V><<<nb, kBlock, 0, c->stream>>>((const V *)dep.p, (const V *)inv.p, score, n);
            if (levels_out)
                (void)hipMemcpyAsync(levels_out, level.p, 4ull * n, hipMemcpyDeviceToDevice, c->stream);
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
        info->edges_scanned = (uint64_t)(2 * round - 1) * g->m;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    Buf *all[] = {&sigma, &level, &dep, &inv, &gate};
    for (Buf *b : all) pool_give(c, *b);
    return rc;
}

// ---- Bellman-Ford ---------------------------------------------------------

constexpr int32_t kDistInf = 2147483647 / 2;    // INT_MAX/2 (BellmanFord.C:105)
constexpr int32_t kGateOff = std::numeric_limits<int32_t>::min();

__global__ void k_sssp_gate(const int32_t *__restrict__ dist, const uint8_t *__restrict__ F, int32_t *__restrict__ gate,
                            uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) gate[v] = F[v] ? dist[v] : kGateOff;
}

struct OpSSSP {
    using acc_t = int32_t;
    const int32_t *gate;
    const float *tw; // lengths in in-edge order (integral values), NULL: 1
    int32_t *dist;
    uint8_t *next;
    __device__ bool skip() const { return false; }
    __device__ int32_t identity() const { return std::numeric_limits<int32_t>::max(); }
    __device__ int32_t combine(int32_t a, int32_t b) const { return a < b ? a : b; }
    __device__ int32_t gather(uint32_t e, uint32_t s) const
    {
        const int32_t gs = gate[s];
        const int32_t w = tw ? (int32_t)__builtin_nontemporal_load(&tw[e]) : 1;
        return gs == kGateOff ? std::numeric_limits<int32_t>::max() : gs + w;
    }
    struct row_t {
        int32_t d;
    };
    __device__ row_t load_row(uint32_t v) const { return row_t{dist[v]}; }
    __device__ void store_row(uint32_t v, const row_t &q, int32_t acc, LocalStats &st) const
    {
        const bool fell = acc < q.d;
        if (fell) {
            dist[v] = acc;
            st.active++;
        }
        next[v] = fell;
    }
};

__global__ void k_sssp_init(int32_t *__restrict__ dist, uint8_t *__restrict__ F, uint32_t root, uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    dist[v] = v == root ? 0 : kDistInf;
    F[v] = v == root;
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
    return dtype == GG_F64 ? bc_impl<double>(g, root, (double *)scores_dev, levels_dev, info)
                           : bc_impl<float>(g, root, (float *)scores_dev, levels_dev, info);
}

extern "C" int gg_sssp(gg_graph_t g, uint32_t root, int32_t *dist_dev, gg_run_info *info)
{
    if (!g || !dist_dev) return fail(GG_ERR_INVALID, "gg_sssp: NULL argument");
    if (root >= g->n) return fail(GG_ERR_INVALID, "gg_sssp: root out of range");
    Ctx *c = g->ctx;
    GG_HIP(hipSetDevice(c->device));
    if (info) std::memset(info, 0, sizeof(*info));
    const uint32_t n = g->n;
    CscView view;
    GG_TRY(pull_view(g, &view));
    Buf F[2], gate;
    int rc = pool_take(c, F[0], (uint64_t)n + 64);
    if (!rc) rc = pool_take(c, F[1], (uint64_t)n + 64);
    if (!rc) rc = pool_take(c, gate, 4ull * n);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const uint32_t nb = blocks_for(n, kBlock);
    int64_t round = 0;
    if (!rc) {
        (void)hipEventRecord(e0, c->stream);
        k_sssp_init<<<nb, kBlock, 0, c->stream>>>(dist_dev, (uint8_t *)F[0].p, root, n);
        unsigned long long *cnt = (unsigned long long *)c->counters.p;
        const float *tw = g->weighted ? (const float *)g->tw.p : nullptr;
        unsigned long long active = 1;
        while (!rc && active) {
            if (round == (int64_t)n) { // negative cycle (BellmanFord.C:116-122)
                k_sssp_negcycle<<<nb, kBlock, 0, c->stream>>>(dist_dev, n);
                break;
            }
            const uint8_t *Fin = (const uint8_t *)F[round % 2].p;
            uint8_t *Fout = (uint8_t *)F[(round + 1) % 2].p;
            k_sssp_gate<<<nb, kBlock, 0, c->stream>>>(dist_dev, Fin, (int32_t *)gate.p, n);
            rc = counters_zero(c);
            if (!rc) rc = launch_pull(c, view, OpSSSP{(const int32_t *)gate.p, tw, dist_dev, Fout}, cnt, nullptr, nullptr);
            unsigned long long cn[4] = {0, 0, 0, 0};
            if (!rc) rc = counters_read(c, cn);
            active = cn[CNT_ACTIVE];
            if (!rc && info && round < 64) info->paths[round] = GG_PATH_CSC;
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
        info->edges_scanned = (uint64_t)round * g->m;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    pool_give(c, F[0]);
    pool_give(c, F[1]);
    pool_give(c, gate);
    return rc;
}
