// gg_batch.hip -- many sources / many graphs in one pass, and the pull
// autotuner.
//
// No counterpart in the reference.  Multi-source BFS is the forward phase
// of app/BC.C (BC_F, 28-77) run for up to 32 sources at once: every vertex
// carries a 32-bit mask, one bit per source, and one pull per level ORs the
// frontier masks of its in-neighbours:
//   next[v]  = (OR_{u in N_in(v)} frontier[u]) & ~seen[v]
//   seen[v] |= next[v];  level_k[v] = level for every bit k of next[v]
// OR is order-free, so the result is deterministic.  Many independent graphs
// are batched by packing them into one arena graph (gg_graph_create_arena);
// the fixed-iteration weighted PageRank and its adjoint then run over the
// arena as one captured loop.
#include <cstring>

#include "gg_kernels.h"

namespace gg {

struct OpMSBFS {
    using acc_t = uint32_t;
    const uint32_t *fin;
    uint32_t *fout;
    uint32_t *seen;
    int32_t *levels; // vmajor: [n][nsrc] (a vertex's levels in one run), else [nsrc][n]
    uint32_t n;
    int32_t level;
    uint32_t nsrc;
    int vmajor;
    const int *done; // the run's flags: [0] set once a level gains no bit, [3] a push level
    uint32_t full;   // every source's bit
    // a sparse frontier (flags[2] clear): the gathers read fin only where
    // bit s of fbm (fin[s] != 0) is set, a 2 MiB bitmap at n = 2^24 that
    // stays in L2, instead of a gather from the n-word array per edge
    const uint32_t *fbm;
    const int *dense;
    int gate;        // flags[gate] set: the other pull runs this level
    __device__ bool skip() const { return done[0] != 0 || done[3] != 0 || done[gate] != 0; }
    // a vertex every source has reached gains nothing: the row-block pull
    // skips the gathers of a work item whose rows all have
    __device__ bool row_done(uint32_t v) const { return seen[v] == full; }
    __device__ uint32_t identity() const { return 0u; }
    __device__ uint32_t combine(uint32_t a, uint32_t b) const { return a | b; }
    __device__ uint32_t gather(uint32_t, uint32_t s) const
    {
        return *dense || bit_test(fbm, s) ? fin[s] : 0u;
    }
    struct row_t {
        uint32_t seen;
    };
    __device__ row_t load_row(uint32_t v) const { return row_t{seen[v]}; }
    __device__ void store_row(uint32_t v, const row_t &q, uint32_t acc, LocalStats &st) const
    {
        const uint32_t nw = acc & ~q.seen;
        fout[v] = nw;
        if (nw) {
            seen[v] = q.seen | nw;
            st.active++;
            st.out_edges += (q.seen | nw) == full; // reached by every source now
            uint32_t b = nw;
            while (b) {
                const int k = __builtin_ctz(b);
                b &= b - 1;
                levels[vmajor ? (uint64_t)v * nsrc + k : (uint64_t)k * n + v] = level;
            }
        }
    }
};

// flag words of a multi-source run (k_msbfs_check): which pull a dense
// level takes
constexpr int kSegRun = 6, kRowsRun = 7, kSegSkip = 8;

// perm (may be NULL): the renumbered plan's new id of each caller's vertex
__global__ void k_msbfs_init(const uint32_t *__restrict__ src, int nsrc, uint32_t n, uint32_t *__restrict__ fin,
                             uint32_t *__restrict__ seen, int32_t *__restrict__ levels,
                             const uint32_t *__restrict__ perm, int vmajor, uint32_t *__restrict__ fbm,
                             uint32_t *__restrict__ list, int *__restrict__ flags, int push_ok)
{
    // one thread per source; sources may repeat, so bits are ORed atomically
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsrc) return;
    const uint32_t v = perm ? perm[src[k]] : src[k];
    atomicOr(&fin[v], 1u << k);
    atomicOr(&seen[v], 1u << k);
    atomicOr(&fbm[v >> 5], 1u << (v & 31));
    list[k] = v; // a repeated source is pushed twice: OR is idempotent
    levels[vmajor ? (uint64_t)v * nsrc + k : (uint64_t)k * n + v] = 0;
    if (k == 0) {
        flags[3] = push_ok; // level 1 pushes from the sources on a symmetric graph
        flags[4] = nsrc;
        flags[kRowsRun] = 1; // and pulls, if it does, by row blocks
        flags[kSegSkip] = 1;
    }
}

// After each level, one wave: the vertices that gained a bit (the counter
// stripes, one per lane, cleared for the next level).  None: flags[0] = 1
// and flags[1] = this level, which ends the loop (every later level's
// kernels return at once).  flags[2] = 1 when they are n/32 or more: the
// next level's gathers then skip the bitmap test (at that density most
// tests pass and only add a dependent load per edge).  flags[3] = 1 when
// they are at most kPushCap and the graph is symmetric: the next level
// pushes from their list (flags[4] entries, rebuilt by k_msbfs_bits)
// instead of pulling over every in-edge.  The first dense level after
// sparse ones pulls through the source-segmented layout (flags[kSegRun];
// where it is built) when fewer than a quarter of this level's new vertices
// were reached by every source: most rows then still gather, and the
// row-block pull (flags[kRowsRun]), which skips work items whose rows all
// have been, gains little.  Later dense levels skip enough items that the
// row blocks are as fast.
// flags[kSegSkip]: the segmented partial pass does not run (the run ended,
// a push, or row blocks).  flags[kBranchLog + l - 1] records the branch of
// level l (up to 64): 0 row blocks, 1 push, 2 segmented.
constexpr uint32_t kPushCap = 4096;
constexpr int kBranchLog = 16, kFlagWords = kBranchLog + 64;
static_assert(4 * kFlagWords <= kPinBytes, "the flags are read back through the pinned buffer");

__global__ __launch_bounds__(64) void k_msbfs_check(unsigned long long *__restrict__ counters, int level,
                                                    uint32_t n, int push_ok, int seg_ok, int *__restrict__ flags)
{
    if (flags[0]) return;
    unsigned long long a = 0, full = 0;
    if (threadIdx.x < kStripes) {
        unsigned long long *st = counters + kCntStripeBase + threadIdx.x * kStripeStride;
        a = st[CNT_ACTIVE];
        full = st[CNT_OUT]; // the new vertices now reached by every source
        st[CNT_ACTIVE] = 0;
        st[CNT_OUT] = 0;
        st[CNT_SCANNED] = 0;
    }
    a = wave_sum_u64(a);
    full = wave_sum_u64(full);
    if (threadIdx.x == 0) {
        if (level <= 64) flags[kBranchLog + level - 1] = flags[3] ? 1 : flags[kSegRun] ? 2 : 0;
        if (a == 0) {
            flags[0] = 1;
            flags[1] = level;
        }
        const int push = push_ok && a <= kPushCap;
        const int seg = seg_ok && !push && a * 32 >= n && !flags[2] && full * 4 < a;
        flags[2] = a * 32 >= n;
        flags[3] = push;
        flags[4] = 0;
        flags[kSegRun] = seg;
        flags[kRowsRun] = !seg;
        flags[kSegSkip] = a == 0 || !seg;
    }
}

// bm bit v = (f[v] != 0), a wave ballot per 64 vertices: after a level
// whose new vertices are a sparse frontier (flags[2] clear), for the next
// level's gathers.  Before a push level (flags[3]) the same vertices are
// appended to list, flags[4] counting them (a list in no fixed order: the
// push ORs).
__global__ __launch_bounds__(kBlock) void k_msbfs_bits(const uint32_t *__restrict__ f, uint32_t n,
                                                       uint32_t *__restrict__ bm, uint32_t *__restrict__ list,
                                                       int *__restrict__ flags)
{
    if (flags[0] || (flags[2] && !flags[3])) return;
    const uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool in = v < n && f[v] != 0;
    const unsigned long long b = __ballot(in);
    const uint32_t lane = threadIdx.x & 63, w = (uint32_t)((v - lane) >> 5) + lane;
    if (!flags[2] && lane < 2 && (uint64_t)w * 32 < n) bm[w] = (uint32_t)(b >> (32 * lane));
    if (flags[3] && b) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd((unsigned int *)&flags[4], (unsigned int)__popcll(b));
        base = __shfl(base, 0, 64);
        if (in) list[base + __popcll(b & ((1ull << lane) - 1))] = (uint32_t)v;
    }
}

// A push level (flags[3]): every frontier vertex u of the list ORs its mask
// into nxt[v] of its neighbours v (the in-list of a symmetric graph is its
// out-list), bits v has already seen left out.  Each list entry's edges are
// cut into kPushParts equal runs, so a hub is spread over work-groups.
constexpr uint32_t kPushParts = 16, kPushStep = 8;
constexpr uint32_t kMsGrid = 2048; // work-groups of the push and finish passes (grid-stride)

__global__ __launch_bounds__(kBlock) void k_msbfs_push(const uint32_t *__restrict__ toff,
                                                       const uint32_t *__restrict__ tids,
                                                       const uint32_t *__restrict__ list,
                                                       const uint32_t *__restrict__ fin,
                                                       const uint32_t *__restrict__ seen, uint32_t *__restrict__ nxt,
                                                       const int *__restrict__ flags)
{
    if (flags[0] || !flags[3]) return;
    const uint32_t parts = (uint32_t)flags[4] * kPushParts;
    for (uint32_t p = blockIdx.x; p < parts; p += gridDim.x) {
        const uint32_t u = list[p / kPushParts], part = p % kPushParts;
        const uint32_t mask = fin[u];
        const uint32_t b = toff[u];
        const uint64_t len = toff[u + 1] - b;
        const uint32_t lo = b + (uint32_t)(len * part / kPushParts);
        const uint32_t hi = b + (uint32_t)(len * (part + 1) / kPushParts);
        // 8 edges a thread per step: their ids, then their seen words, then
        // the ORs, so a hub's run is not one chain of dependent loads per edge
        for (uint32_t i0 = lo + threadIdx.x; i0 < hi; i0 += kBlock * kPushStep) {
            uint32_t v[kPushStep], add[kPushStep];
#pragma unroll
            for (uint32_t q = 0; q < kPushStep; q++) {
                const uint32_t i = i0 + q * kBlock;
                v[q] = i < hi ? tids[i] : 0u;
            }
#pragma unroll
            for (uint32_t q = 0; q < kPushStep; q++)
                add[q] = i0 + q * kBlock < hi ? mask & ~seen[v[q]] : 0u;
#pragma unroll
            for (uint32_t q = 0; q < kPushStep; q++)
                if (add[q]) atomicOr(&nxt[v[q]], add[q]);
        }
    }
}

// The rest of a push level, over every vertex: the new bits (nxt, cleared
// for the next push), seen, levels, fout and the count of vertices with a
// new bit, as the pull's rows would have written them.
__global__ __launch_bounds__(kBlock) void k_msbfs_finish(uint32_t *__restrict__ nxt, uint32_t *__restrict__ seen,
                                                         uint32_t *__restrict__ fout, int32_t *__restrict__ levels,
                                                         uint32_t n, int32_t level, uint32_t nsrc, int vmajor,
                                                         uint32_t full, unsigned long long *__restrict__ counters,
                                                         const int *__restrict__ flags)
{
    if (flags[0] || !flags[3]) return;
    LocalStats st;
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (uint64_t)gridDim.x * kBlock) {
        const uint32_t x = nxt[v];
        uint32_t nw = 0;
        if (x) {
            nxt[v] = 0;
            const uint32_t sn = seen[v];
            nw = x & ~sn;
            if (nw) {
                seen[v] = sn | nw;
                st.active++;
                st.out_edges += (sn | nw) == full;
                uint32_t bb = nw;
                while (bb) {
                    const uint32_t k = __builtin_ctz(bb);
                    bb &= bb - 1;
                    levels[vmajor ? v * nsrc + k : (uint64_t)k * n + v] = level;
                }
            }
        }
        fout[v] = nw;
    }
    (void)block_stats(st, counters);
}

// BC_F's path counts (app/BC.C:28-47) for the sources of a multi-source
// level: P is vertex-major, P[v*32 + k] = source k's NumPaths of v.  Half a
// wave per vertex v with a new bit: lane k adds P[u*32 + k] over v's
// in-neighbours u whose previous-level mask holds bit k (BC_F::update for
// the frontier of source k, in in-list order: a fixed order), and writes it
// when bit k is new at v.  The in-list is read kPathStep entries a step:
// lane j of the half-wave loads entry j's id and then its mask, so a step
// costs two dependent loads whatever its length; every lane then takes the
// step's ids and masks by shuffles and issues its kPathStep count loads
// together (one line of counts per entry for the half-wave).  In-lists
// longer than kHubChunk are left to the chunked pass below.
constexpr int kPathStep = 32;

// lane k's sum of P[u*32 + k] over in-list entries [b, e) whose mask
// fin[u] & nw holds bit k, in list order (k: the lane in its half-wave)
template <class V>
__device__ __forceinline__ V msbc_sum(const uint32_t *__restrict__ tids, const uint32_t *__restrict__ fin,
                                      const V *__restrict__ P, uint32_t b, uint32_t e, uint32_t nw, uint32_t k)
{
    static_assert(kPathStep == 32, "one entry per lane of a half-wave");
    V acc = V(0);
    for (uint32_t j = b; j < e; j += kPathStep) {
        const bool in = j + k < e;
        const uint32_t u = in ? tids[j + k] : 0u;
        const uint32_t m = in ? fin[u] & nw : 0u;
        V c[kPathStep];
#pragma unroll
        for (int q = 0; q < kPathStep; q++) {
            const uint32_t uq = __shfl(u, q, 32), mq = __shfl(m, q, 32);
            c[q] = ((mq >> k) & 1u) ? P[(uint64_t)uq * 32 + k] : V(0);
        }
#pragma unroll
        for (int q = 0; q < kPathStep; q++) acc += c[q];
    }
    return acc;
}

template <class V>
__global__ __launch_bounds__(kBlock) void k_msbc_paths(const uint32_t *__restrict__ toff,
                                                       const uint32_t *__restrict__ tids,
                                                       const uint32_t *__restrict__ fin,
                                                       const uint32_t *__restrict__ fout, V *__restrict__ P,
                                                       uint32_t n, const int *__restrict__ done)
{
    if (*done) return;
    // a wave reads 64 vertices' new masks (and their in-list bounds) at
    // once, one per lane; its two halves then take the vertices with a new
    // source two at a time.  Lane l of wave w takes vertex first + l*W (W
    // waves): consecutive ids have alike degrees under a hot-first
    // numbering, so each wave gets a spread of them.
    const uint32_t lane = threadIdx.x & 63, k = lane & 31, half = lane >> 5;
    const uint64_t w0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t W = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t first = w0; first < n; first += 64 * W) {
        const uint64_t vl = first + lane * W;
        uint32_t f = 0, lb = 0, le = 0;
        if (vl < n) {
            f = fout[vl];
            if (f) {
                lb = toff[vl];
                le = toff[vl + 1];
            }
        }
        unsigned long long bal = __ballot(f != 0 && le - lb <= kHubChunk); // longer: the chunked pass
        while (bal) {
            const int i0 = __builtin_ctzll(bal);
            bal &= bal - 1;
            int i1 = -1;
            if (bal) {
                i1 = __builtin_ctzll(bal);
                bal &= bal - 1;
            }
            const int mine = half ? i1 : i0;
            const int from = mine < 0 ? 0 : mine;
            const uint32_t nw = __shfl(f, from, 64), b = __shfl(lb, from, 64), e = __shfl(le, from, 64);
            if (mine >= 0) {
                const V acc = msbc_sum(tids, fin, P, b, e, nw, k);
                if ((nw >> k) & 1u) P[(first + mine * W) * 32 + k] = acc;
            }
        }
    }
}

// The hubs' chunks: half a wave per chunk item (v, e0, e1), lane k's sum
// into slot[item*32 + k]; then each hub's slots added in chunk order.
template <class V>
__global__ __launch_bounds__(kBlock) void k_msbc_hub_items(const uint32_t *__restrict__ items, uint32_t nitems,
                                                           const uint32_t *__restrict__ tids,
                                                           const uint32_t *__restrict__ fin,
                                                           const uint32_t *__restrict__ fout,
                                                           const V *__restrict__ P, V *__restrict__ slot,
                                                           const int *__restrict__ done)
{
    if (*done) return;
    const uint32_t k = threadIdx.x & 31;
    const uint64_t h0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 5;
    const uint64_t nh = ((uint64_t)gridDim.x * blockDim.x) >> 5;
    for (uint64_t i = h0; i < nitems; i += nh) {
        const uint32_t v = items[3 * i], b = items[3 * i + 1], e = items[3 * i + 2];
        const uint32_t nw = fout[v];
        if (!nw) continue;
        slot[i * 32 + k] = msbc_sum(tids, fin, P, b, e, nw, k);
    }
}

template <class V>
__global__ void k_msbc_hub_fix(const uint32_t *__restrict__ hubs, uint32_t nhubs, const uint32_t *__restrict__ fout,
                               const V *__restrict__ slot, V *__restrict__ P, const int *__restrict__ done)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (*done || t >= (uint64_t)nhubs * 32) return;
    const uint32_t h = (uint32_t)(t >> 5), k = (uint32_t)(t & 31);
    const uint32_t v = hubs[3 * h], i0 = hubs[3 * h + 1], i1 = hubs[3 * h + 2];
    if (!((fout[v] >> k) & 1u)) return;
    V acc = V(0);
    for (uint32_t i = i0; i < i1; i++) acc += slot[(uint64_t)i * 32 + k];
    P[(uint64_t)v * 32 + k] = acc;
}

// sources: P[src_k*32 + k] = 1 (BC.C:196 NumPaths[start] = 1)
template <class V>
__global__ void k_msbc_init(const uint32_t *__restrict__ src, int nsrc, const uint32_t *__restrict__ perm,
                            V *__restrict__ P)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsrc) return;
    const uint32_t v = perm ? perm[src[k]] : src[k];
    P[(uint64_t)v * 32 + k] = V(1);
}

// The output of a run whose levels are vertex-major in the traversal's
// numbering (lp[w*nsrc + k], set where seen[w] holds bit k): levels[k][v]
// (-1 unreached) and, with P, paths[k][v] (0 unreached) in the caller's
// numbering, w = perm[v] (v without perm).  A work-group takes kOutTile
// vertices: it reads each one's nsrc levels and 32 counts as runs (16-byte
// loads for 32 sources, all of a thread's loads issued together),
// transposes them in LDS and writes each source's kOutTile entries
// coalesced.  kLv / kPa: levels / paths wanted (LDS only for those).
constexpr uint32_t kOutTile = 128;

template <class T>
__device__ __forceinline__ void out_stage(T (*s)[kOutTile + 1], const T *__restrict__ src, const uint32_t *s_w,
                                          const uint32_t *s_s, uint32_t stride, T none)
{
    // 32 values of a vertex as 32 / E loads of 16 bytes
    constexpr uint32_t E = 16 / sizeof(T), Q = 32 / E, R = kOutTile * Q / kBlock;
    struct alignas(16) Vec {
        T x[E];
    };
    Vec v[R];
#pragma unroll
    for (uint32_t r = 0; r < R; r++) {
        const uint32_t j = threadIdx.x + r * kBlock, vi = j / Q;
        if (s_s[vi]) v[r] = *(const Vec *)(src + (uint64_t)s_w[vi] * stride + (j % Q) * E);
    }
#pragma unroll
    for (uint32_t r = 0; r < R; r++) {
        const uint32_t j = threadIdx.x + r * kBlock, vi = j / Q, k0 = (j % Q) * E;
#pragma unroll
        for (uint32_t i = 0; i < E; i++) s[k0 + i][vi] = (s_s[vi] >> (k0 + i)) & 1u ? v[r].x[i] : none;
    }
}

template <class V, bool kLv, bool kPa>
__global__ __launch_bounds__(kBlock) void k_msbfs_out(const int32_t *__restrict__ lp,
                                                      const uint32_t *__restrict__ seen, const V *__restrict__ P,
                                                      const uint32_t *__restrict__ perm, uint32_t nsrc, uint32_t n,
                                                      int32_t *__restrict__ levels, V *__restrict__ paths)
{
    static_assert(kBlock % kOutTile == 0, "whole tiles of threads");
    __shared__ int32_t s_l[kLv ? 32 : 1][kOutTile + 1];
    __shared__ V s_p[kPa ? 32 : 1][kOutTile + 1];
    __shared__ uint32_t s_w[kOutTile], s_s[kOutTile];
    const uint32_t tid = threadIdx.x;
    const uint64_t v0 = (uint64_t)blockIdx.x * kOutTile;
    if (tid < kOutTile) {
        const uint64_t v = v0 + tid;
        uint32_t w = 0, sn = 0; // past n: no bit, nothing read
        if (v < n) {
            w = perm ? perm[v] : (uint32_t)v;
            sn = seen[w];
        }
        s_w[tid] = w;
        s_s[tid] = sn;
    }
    __syncthreads();
    if (nsrc == 32) {
        if constexpr (kLv) out_stage<int32_t>(s_l, lp, s_w, s_s, 32, -1);
        if constexpr (kPa) out_stage<V>(s_p, P, s_w, s_s, 32, V(0));
    } else {
        for (uint32_t j = tid; j < kOutTile * nsrc; j += kBlock) {
            const uint32_t vi = j / nsrc, k = j - vi * nsrc;
            const uint64_t w = s_w[vi];
            const bool hit = (s_s[vi] >> k) & 1u;
            if constexpr (kLv) s_l[k][vi] = hit ? lp[w * nsrc + k] : -1;
            if constexpr (kPa) s_p[k][vi] = hit ? P[w * 32 + k] : V(0);
        }
    }
    __syncthreads();
    const uint32_t vi = tid % kOutTile;
    const uint64_t v = v0 + vi;
    if (v >= n) return;
    for (uint32_t k = tid / kOutTile; k < nsrc; k += kBlock / kOutTile) {
        if constexpr (kLv) levels[(uint64_t)k * n + v] = s_l[k][vi];
        if constexpr (kPa) paths[(uint64_t)k * n + v] = s_p[k][vi];
    }
}

} // namespace gg

using namespace gg;

namespace {
// The multi-source level loop shared by gg_msbfs and gg_msbc_forward: levels
// lv[k][v] in the traversal's numbering; with P, the path counts too.
template <class V>
int msbfs_run(Graph *g, int nsrc, const uint32_t *sources, int32_t *levels, V *paths, gg_run_info *info)
{
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    const bool planned = g->plan.mode != 0 && g->m;
    CscView view;
    if (planned) view = plan_view(g);
    else GG_TRY(pull_view(g, &view));
    const uint32_t *perm = planned ? (const uint32_t *)g->plan.perm.p : nullptr;
    // the source-segmented layout of the plan's in-edges, as PageRank.C's
    // loop takes it (masks below 2^31 bytes for the partial pass's loads)
    const int seg_ok = planned && (uint64_t)n * 4 < (1ull << 31) &&
                       (g->xcd_segments == 2 || (g->xcd_segments == 1 && g->m >= (1ull << 24)));
    const uint32_t xhot = kXHotBytes / 4;
    if (seg_ok) {
        GG_TRY(ensure_xseg(g));
        GG_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_xseg_part<SegOr<uint32_t>>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)(xhot * 4)));
    }
    Buf F[2], seen, src, lp, pb, hslot, fb, B[2], fl, nx;
    const uint64_t bmb = 4ull * ((n + 31) / 32);
    int rc = pool_take(c, F[0], 4ull * n);
    if (!rc) rc = pool_take(c, fl, 4ull * std::max<uint32_t>(n, 32)); // the frontier list (the sources first)
    if (!rc) rc = pool_take(c, nx, 4ull * n);
    if (!rc) rc = pool_take(c, B[0], bmb);
    if (!rc) rc = pool_take(c, B[1], bmb);
    if (!rc) rc = pool_take(c, fb, 4 * kFlagWords);
    if (!rc) rc = pool_take(c, F[1], 4ull * n);
    if (!rc) rc = pool_take(c, seen, 4ull * n);
    if (!rc) rc = pool_take(c, src, 4ull * nsrc);
    if (!rc && (planned || paths)) rc = pool_take(c, lp, 4ull * nsrc * n);
    if (!rc && paths) rc = pool_take(c, pb, sizeof(V) * 32ull * n);
    // the long in-lists' chunks (cached per graph or plan)
    HubChunks *hub = planned ? &g->plan.hubs : &g->hubs;
    if (!rc && paths) rc = ensure_hub_chunks(c, view.toff, n, *hub);
    if (!rc && paths && hub->nitems) rc = pool_take(c, hslot, sizeof(V) * 32ull * hub->nitems);
    // levels in the traversal's numbering: with a plan or path counts,
    // vertex-major in lp (a vertex's levels in one run, mapped out by
    // k_msbfs_out with the counts); else straight into the caller's array
    const int vmajor = planned || paths;
    const int push_ok = g->symmetric; // in-lists are the out-lists
    int32_t *lv = vmajor ? (int32_t *)lp.p : levels;
    V *P = paths ? (V *)pb.p : nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    int level = 0;
    if (!rc) {
        (void)hipEventRecord(e0, c->stream);
        rc = fill_u32(c, F[0].p, n, 0);
        if (!rc) rc = fill_u32(c, seen.p, n, 0);
        if (!rc) rc = fill_bytes(c, B[0].p, bmb, 0);
        if (!rc) rc = fill_u32(c, nx.p, n, 0);
        if (!rc) rc = fill_bytes(c, fb.p, 4 * kFlagWords, 0);
        if (!rc && !vmajor) rc = gg_fill(abi(c), lv, GG_I32, (uint64_t)nsrc * n, -1.0);
        if (!rc && hipMemcpyAsync(src.p, sources, 4ull * nsrc, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            rc = fail(GG_ERR_HIP, "multi-source BFS: upload sources");
        if (!rc) {
            k_msbfs_init<<<1, 64, 0, c->stream>>>((const uint32_t *)src.p, nsrc, n, (uint32_t *)F[0].p,
                                                  (uint32_t *)seen.p, lv, perm, vmajor, (uint32_t *)B[0].p,
                                                  (uint32_t *)fl.p, (int *)fb.p, push_ok);
            if (P) k_msbc_init<V><<<1, 64, 0, c->stream>>>((const uint32_t *)src.p, nsrc, perm, P);
            if (hipGetLastError() != hipSuccess) rc = fail(GG_ERR_HIP, "multi-source BFS: init");
        }
        unsigned long long *cnt = (unsigned long long *)c->counters.p;
        const uint32_t pgrid = (uint32_t)std::min<uint64_t>(blocks_for(n, 4 * kBlock), 1u << 16);
        int *flags = (int *)fb.p;
        const uint32_t full = nsrc == 32 ? ~0u : (1u << nsrc) - 1;
        const uint32_t fgrid = std::min<uint32_t>(blocks_for(n, kBlock), kMsGrid);
        if (!rc) rc = counters_zero(c);
        // Levels are enqueued in batches.  A level is the push and the two
        // pulls, of which the flags let one run (k_msbfs_check chose it at
        // the level before), with BC the path counts, then k_msbfs_check and
        // k_msbfs_bits; the flags are read back once per batch.  The first
        // batch is the last call's depth, else 8, then 16, 32 and 64; at
        // most n levels.
        int batch = g->msbfs_levels_hint > 0 ? std::min(g->msbfs_levels_hint, 64) : 8;
        int enq = 0, hf[kFlagWords] = {};
        while (!rc && enq < (int)n) {
            const int upto = (int)std::min<int64_t>((int64_t)enq + batch, (int64_t)n);
            for (int lvl = enq + 1; !rc && lvl <= upto; lvl++) {
                const uint32_t *fin = (const uint32_t *)F[(lvl - 1) % 2].p;
                uint32_t *fout = (uint32_t *)F[lvl % 2].p;
                OpMSBFS op{fin, fout, (uint32_t *)seen.p, lv, n, lvl, (uint32_t)nsrc, vmajor, flags, full,
                           (const uint32_t *)B[(lvl - 1) % 2].p, flags + 2, kSegRun};
                // a push level (small frontier, symmetric graph) or the pull:
                // the kernels of the branch not taken return at once
                k_msbfs_push<<<kMsGrid, kBlock, 0, c->stream>>>(view.toff, view.tids, (const uint32_t *)fl.p, fin,
                                                                 (const uint32_t *)seen.p, (uint32_t *)nx.p, flags);
                k_msbfs_finish<<<fgrid, kBlock, 0, c->stream>>>((uint32_t *)nx.p, (uint32_t *)seen.p, fout, lv, n,
                                                                lvl, (uint32_t)nsrc, vmajor, full, cnt, flags);
                if (hipGetLastError() != hipSuccess) rc = fail(GG_ERR_HIP, "multi-source BFS: push");
                if (!rc) rc = launch_pull(c, view, op, cnt, nullptr, nullptr);
                if (!rc && seg_ok) {
                    // the segmented pull: partials of fin's ORs per (row,
                    // source group), long rows' chunks, then each row's
                    // partials through OpMSBFS's store
                    XSeg &X = g->xseg;
                    const uint32_t *meta = (const uint32_t *)X.meta.p;
                    k_xseg_part<SegOr<uint32_t>><<<X.grid, kXPartThreads, xhot * 4, c->stream>>>(
                        fin, (uint32_t *)X.partial.p, (uint32_t *)X.rows.slot_val.p, (const uint32_t *)X.vtids.p,
                        (const RowWork *)X.rows.work.p, meta, n, xhot, flags + kSegSkip, nullptr);
                    if (X.rows.nlong)
                        k_rows_fix<OpSegPartial<SegOr<uint32_t>>>
                            <<<blocks_for(X.rows.nlong, kBlock), kBlock, 0, c->stream>>>(
                                OpSegPartial<SegOr<uint32_t>>{(uint32_t *)X.partial.p, flags + kSegSkip},
                                (const uint32_t *)X.rows.long_rows.p, X.rows.nlong, (const uint32_t *)X.rows.slot_val.p,
                                0, nullptr, nullptr, nullptr);
                    OpMSBFS sop = op;
                    sop.gate = kRowsRun;
                    k_xseg_combine<OpMSBFS><<<X.comb_blocks, kBlock, 0, c->stream>>>(
                        sop, (const uint32_t *)X.partial.p, (const uint32_t *)X.mask.p, (const uint32_t *)X.cursor.p,
                        meta + kXSegs + 1, X.vrows, n, cnt, nullptr);
                    if (hipGetLastError() != hipSuccess) rc = fail(GG_ERR_HIP, "multi-source BFS: segmented pull");
                }
                if (!rc && P) {
                    k_msbc_paths<V><<<pgrid, kBlock, 0, c->stream>>>(view.toff, view.tids, fin, fout, P, n, flags);
                    if (hub->nitems) {
                        const uint32_t igrid =
                            (uint32_t)std::min<uint64_t>(((uint64_t)hub->nitems * 32 + kBlock - 1) / kBlock, 1u << 16);
                        k_msbc_hub_items<V><<<igrid, kBlock, 0, c->stream>>>(
                            (const uint32_t *)hub->items.p, hub->nitems, view.tids, fin, fout, P, (V *)hslot.p, flags);
                        k_msbc_hub_fix<V><<<blocks_for((uint64_t)hub->nhubs * 32, kBlock), kBlock, 0, c->stream>>>(
                            (const uint32_t *)hub->hubs.p, hub->nhubs, fout, (const V *)hslot.p, P, flags);
                    }
                    if (hipGetLastError() != hipSuccess) rc = fail(GG_ERR_HIP, "multi-source BC: path counts");
                }
                if (!rc) {
                    k_msbfs_check<<<1, 64, 0, c->stream>>>(cnt, lvl, n, push_ok, seg_ok, flags);
                    k_msbfs_bits<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>(fout, n, (uint32_t *)B[lvl % 2].p,
                                                                                (uint32_t *)fl.p, flags);
                    if (hipGetLastError() != hipSuccess) rc = fail(GG_ERR_HIP, "multi-source BFS: level check");
                }
            }
            enq = upto;
            if (!rc) rc = read_back(c, hf, flags, sizeof(hf));
            if (rc || hf[0]) break;
            batch = std::min(batch * 2, 64);
        }
        level = hf[0] ? hf[1] : enq;
        if (!rc) g->msbfs_levels_hint = level;
        if (!rc && info)
            for (int l = 1; l <= level && l <= 64; l++)
                info->paths[l - 1] = hf[kBranchLog + l - 1] == 1 ? GG_PATH_SPARSE : GG_PATH_CSC;
        if (!rc && vmajor && (levels || P)) {
            const uint32_t ob = (uint32_t)((n + kOutTile - 1) / kOutTile);
            const uint32_t *sn = (const uint32_t *)seen.p;
            if (levels && P)
                k_msbfs_out<V, true, true><<<ob, kBlock, 0, c->stream>>>(lv, sn, P, perm, nsrc, n, levels, paths);
            else if (levels)
                k_msbfs_out<V, true, false><<<ob, kBlock, 0, c->stream>>>(lv, sn, P, perm, nsrc, n, levels, paths);
            else
                k_msbfs_out<V, false, true><<<ob, kBlock, 0, c->stream>>>(lv, sn, P, perm, nsrc, n, levels, paths);
            if (hipGetLastError() != hipSuccess) rc = fail(GG_ERR_HIP, "multi-source BFS: output");
        }
        (void)hipEventRecord(e1, c->stream);
        if (!rc) rc = sync(c);
    }
    float ms = 0;
    if (!rc) (void)hipEventElapsedTime(&ms, e0, e1);
    if (info && !rc) {
        info->iters = level;
        info->ms = ms;
        info->edges_scanned = (uint64_t)level * g->m * (P ? 2 : 1);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    pool_give(c, F[0]);
    pool_give(c, F[1]);
    pool_give(c, seen);
    pool_give(c, src);
    pool_give(c, lp);
    pool_give(c, pb);
    pool_give(c, hslot);
    pool_give(c, fb);
    pool_give(c, B[0]);
    pool_give(c, B[1]);
    pool_give(c, fl);
    pool_give(c, nx);
    return rc;
}
} // namespace

extern "C" int gg_msbc_forward(gg_graph_t g, int nsrc, const uint32_t *sources, int dtype, void *paths_dev,
                               int32_t *levels_dev, gg_run_info *info)
{
    if (!g || !sources || !paths_dev) return fail(GG_ERR_INVALID, "gg_msbc_forward: NULL argument");
    if (nsrc < 1 || nsrc > 32) return fail(GG_ERR_INVALID, "gg_msbc_forward: 1 to 32 sources");
    if (dtype != GG_F64 && dtype != GG_F32) return fail(GG_ERR_INVALID, "gg_msbc_forward: dtype must be F64 or F32");
    for (int k = 0; k < nsrc; k++)
        if (sources[k] >= g->n) return fail(GG_ERR_INVALID, "gg_msbc_forward: source out of range");
    GG_HIP(hipSetDevice(g->ctx->device));
    if (info) std::memset(info, 0, sizeof(*info));
    return dtype == GG_F64 ? msbfs_run<double>(g, nsrc, sources, levels_dev, (double *)paths_dev, info)
                           : msbfs_run<float>(g, nsrc, sources, levels_dev, (float *)paths_dev, info);
}

extern "C" int gg_msbfs(gg_graph_t g, int nsrc, const uint32_t *sources, int32_t *levels, gg_run_info *info)
{
    if (!g || !sources || !levels) return fail(GG_ERR_INVALID, "gg_msbfs: NULL argument");
    if (nsrc < 1 || nsrc > 32) return fail(GG_ERR_INVALID, "gg_msbfs: 1 to 32 sources");
    for (int k = 0; k < nsrc; k++)
        if (sources[k] >= g->n) return fail(GG_ERR_INVALID, "gg_msbfs: source out of range");
    GG_HIP(hipSetDevice(g->ctx->device));
    if (info) std::memset(info, 0, sizeof(*info));
    // With a renumbered plan the levels run over the plan's relabelled twin
    // and the per-source levels are mapped back at the end.
    return msbfs_run<float>(g, nsrc, sources, levels, nullptr, info);
}

namespace gg {
// The autotuner's probe: one pull of a gathered value per edge and one
// write per row, the access pattern of every PageRank-like iteration.
template <class V>
struct OpProbe {
    using acc_t = V;
    const V *x;
    V *y;
    __device__ bool skip() const { return false; }
    __device__ V identity() const { return V(0); }
    __device__ V combine(V a, V b) const { return a + b; }
    __device__ V gather(uint32_t, uint32_t s) const { return x[s]; }
    __device__ void finish(uint32_t d, V acc, LocalStats &) const { y[d] = acc; }
};

template <class V>
static int tune_time(Graph *g, const V *x, V *y, int trials, double *ms)
{
    Ctx *c = g->ctx;
    CscView view;
    GG_TRY(pull_view(g, &view));
    hipEvent_t e0, e1;
    GG_HIP(hipEventCreate(&e0));
    GG_HIP(hipEventCreate(&e1));
    int rc = launch_pull(c, view, OpProbe<V>{x, y}, nullptr, nullptr, nullptr); // warm-up
    if (!rc) rc = (hipEventRecord(e0, c->stream) == hipSuccess) ? GG_OK : fail(GG_ERR_HIP, "event");
    for (int t = 0; t < trials && !rc; t++) rc = launch_pull(c, view, OpProbe<V>{x, y}, nullptr, nullptr, nullptr);
    if (!rc) rc = (hipEventRecord(e1, c->stream) == hipSuccess) ? GG_OK : fail(GG_ERR_HIP, "event");
    if (!rc) rc = sync(c);
    float f = 0;
    if (!rc) (void)hipEventElapsedTime(&f, e0, e1);
    *ms = f / trials;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}
} // namespace gg

extern "C" int gg_autotune_pull(gg_graph_t g, int dtype, int trials, gg_tune_info *out)
{
    if (!g || !out) return fail(GG_ERR_INVALID, "gg_autotune_pull: NULL argument");
    if (dtype != GG_F64 && dtype != GG_F32) return fail(GG_ERR_INVALID, "gg_autotune_pull: dtype must be F64 or F32");
    if (trials < 1 || trials > 1000) return fail(GG_ERR_INVALID, "gg_autotune_pull: trials in [1, 1000]");
    Ctx *c = g->ctx;
    GG_HIP(hipSetDevice(c->device));
    std::memset(out, 0, sizeof(*out));
    const size_t vs = dtype == GG_F64 ? 8 : 4;
    Buf x, y;
    GG_TRY(pool_take(c, x, vs * g->n));
    int rc = pool_take(c, y, vs * g->n);
    if (!rc) rc = gg_fill(abi(c), x.p, dtype, g->n, 1.0);
    // candidates: merge-path tiles, then row blocks at 1..32 edges per lane
    static const int kEpl[] = {1, 2, 4, 8, 16, 32};
    const int ncand = 1 + (int)(sizeof(kEpl) / sizeof(kEpl[0]));
    int best = -1;
    double best_ms = 0;
    const int keep_sched = g->pull_schedule, keep_epl = g->edges_per_lane;
    for (int k = 0; k < ncand && !rc; k++) {
        const int sched = k == 0 ? 0 : 1, epl = k == 0 ? keep_epl : kEpl[k - 1];
        rc = gg_graph_set_option(g, GG_OPT_PULL_SCHEDULE, sched);
        if (!rc) rc = gg_graph_set_option(g, GG_OPT_ROW_EDGES_PER_LANE, epl);
        double ms = 0;
        if (!rc)
            rc = dtype == GG_F64 ? tune_time<double>(g, (const double *)x.p, (double *)y.p, trials, &ms)
                                 : tune_time<float>(g, (const float *)x.p, (float *)y.p, trials, &ms);
        if (rc) break;
        out->cand_schedule[k] = sched;
        out->cand_edges_per_lane[k] = epl;
        out->cand_ms[k] = ms;
        if (best < 0 || ms < best_ms) {
            best = k;
            best_ms = ms;
        }
    }
    pool_give(c, x);
    pool_give(c, y);
    if (rc) {
        (void)gg_graph_set_option(g, GG_OPT_PULL_SCHEDULE, keep_sched);
        (void)gg_graph_set_option(g, GG_OPT_ROW_EDGES_PER_LANE, keep_epl);
        return rc;
    }
    out->candidates = ncand;
    out->schedule = out->cand_schedule[best];
    out->edges_per_lane = out->cand_edges_per_lane[best];
    out->ms = best_ms;
    GG_TRY(gg_graph_set_option(g, GG_OPT_PULL_SCHEDULE, out->schedule));
    return gg_graph_set_option(g, GG_OPT_ROW_EDGES_PER_LANE, out->edges_per_lane);
}
