// gg_traverse.hip -- frontier objects, edgeMap and vertexMap of
// libgraphgrind_amd.
//
// edgeMap (ligra-numa.h:919-1019) keeps the reference's three-way decision
// on (d_m + out_edges) against m/20 and m/2 and runs one of:
//   sparse push   load-balanced over the frontier's out-edges, CAS /
//                 atomicMin / atomicAdd as the app's updateAtomic, output
//                 compacted with wave ballots (edgeMapSparseWithG 868-916)
//   dense CSC     merge-path pull over the in-edge twin, destination-owned
//                 (edgeMapDenseCSC 549-729); BFS uses a bottom-up kernel
//                 with the early exit of 711-712
//   dense COO     Hilbert-ordered COO partitions streamed edge-parallel for
//                 the order-free min operators (BFS, CC); the sum operators
//                 (PageRank, SpMV) use the destination-owned pull so their
//                 floating-point sums keep a fixed order (edgeMapDense
//                 803-866 under PART96)
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <cstring>

#include "gg_kernels.h"

namespace gg {

// ============================================================================
// Operators of the generic edgeMap
// ============================================================================

// "No contribution": a NaN of a fixed payload that the sum operators'
// combine() treats as their identity.  A source outside the frontier
// gathers it, so "no frontier in-neighbour" stays apart from every sum of
// real values, negative and zero sums included (PageRankDelta.C pushes
// negative changes).  Arithmetic never produces this payload.
template <class V>
struct PathNone;
template <>
struct PathNone<float> {
    using bits_t = uint32_t;
    static constexpr bits_t kBits = 0x7fa5a5a5u;
};
template <>
struct PathNone<double> {
    using bits_t = uint64_t;
    static constexpr bits_t kBits = 0x7ff5a5a5a5a5a5a5ull;
};
template <class V>
__device__ __forceinline__ bool path_none(V x)
{
    typename PathNone<V>::bits_t b;
    __builtin_memcpy(&b, &x, sizeof(V));
    return b == PathNone<V>::kBits;
}
template <class V>
__device__ __forceinline__ V path_none_value()
{
    const typename PathNone<V>::bits_t b = PathNone<V>::kBits;
    V x;
    __builtin_memcpy(&x, &b, sizeof(V));
    return x;
}

// PR_F (app/PageRank.C:28-70, app/PageRankDelta.C:27-72): contributions
// a[s]/outdeg(s) come from a prepass, "no contribution" for sources outside
// the frontier; a row whose sum is not that value had a frontier source.
template <class V>
struct OpPRSum {
    using acc_t = V;
    const V *contrib;
    V *p_next;
    uint8_t *next;
    const uint32_t *outdeg;
    __device__ bool skip() const { return false; }
    __device__ V identity() const { return path_none_value<V>(); }
    __device__ V combine(V a, V b) const { return path_none(a) ? b : path_none(b) ? a : a + b; }
    __device__ V gather(uint32_t, uint32_t s) const { return contrib[s]; }
    __device__ void finish(uint32_t d, V acc, LocalStats &st) const
    {
        const bool any = !path_none(acc);
        if (any) {
            p_next[d] = p_next[d] + acc;
            st.active++;
            st.out_edges += outdeg[d];
        }
        next[d] = any;
    }
};

// SPMV_F (app/SPMV.C:7-48) with an all-active frontier (SPMV.C:81).
template <class V>
struct OpSpmvBit {
    using acc_t = V;
    const V *x;
    const float *tw; // in-edge order, NULL = unit weights
    V *p_next;
    uint8_t *next;
    const uint32_t *outdeg, *indeg;
    __device__ bool skip() const { return false; }
    __device__ V identity() const { return V(0); }
    __device__ V combine(V a, V b) const { return a + b; }
    __device__ V gather(uint32_t e, uint32_t s) const { return tw ? x[s] * (V)tw[e] : x[s]; }
    __device__ void finish(uint32_t d, V acc, LocalStats &st) const
    {
        const bool any = indeg[d] != 0;
        if (any) {
            p_next[d] = p_next[d] + acc;
            st.active++;
            st.out_edges += outdeg[d];
        }
        next[d] = any;
    }
};

// SPMV_F (and PageRankConverage's PR_F, tw NULL) with a frontier: sources
// outside it gather "no contribution", so a row's sum says whether any
// frontier source reached it.
template <class V>
struct OpSpmvFrontier {
    using acc_t = V;
    const V *x;
    const float *tw;
    const uint32_t *bm;
    V *p_next;
    uint8_t *next;
    const uint32_t *outdeg;
    __device__ bool skip() const { return false; }
    __device__ V identity() const { return path_none_value<V>(); }
    __device__ V combine(V a, V b) const { return path_none(a) ? b : path_none(b) ? a : a + b; }
    __device__ V gather(uint32_t e, uint32_t s) const
    {
        if (!bit_test(bm, s)) return path_none_value<V>();
        return tw ? x[s] * (V)tw[e] : x[s];
    }
    __device__ void finish(uint32_t d, V acc, LocalStats &st) const
    {
        const bool any = !path_none(acc);
        if (any) {
            p_next[d] = p_next[d] + acc;
            st.active++;
            st.out_edges += outdeg[d];
        }
        next[d] = any;
    }
};

// CC_F cache path (app/Components.C:57-79): Jacobi over prevIDs of the
// frontier's sources; g[s] = prevIDs[s] for active s, else UINT_MAX.
struct OpCCMin {
    using acc_t = uint32_t;
    const uint32_t *g;
    uint32_t *ids;
    const uint32_t *prev;
    uint8_t *next;
    const uint32_t *outdeg;
    const uint32_t *bm; // the frontier's bitmap (unread when 'bit')
    int bit;
    __device__ bool skip() const { return false; }
    __device__ uint32_t identity() const { return 0xffffffffu; }
    __device__ uint32_t combine(uint32_t a, uint32_t b) const { return a < b ? a : b; }
    // the frontier bit first (a bitmap that stays in L2): only the
    // frontier's sources read their label
    __device__ uint32_t gather(uint32_t, uint32_t s) const
    {
        return (bit || bit_test(bm, s)) ? g[s] : 0xffffffffu;
    }
    // label 0 is the least there is: such a row cannot change
    __device__ bool row_done(uint32_t d) const { return ids[d] == 0u; }
    __device__ void finish(uint32_t d, uint32_t acc, LocalStats &st) const
    {
        const uint32_t cur = ids[d];
        const bool dec = acc < cur;
        const bool ch = dec && cur == prev[d];
        if (dec) ids[d] = acc;
        if (ch) {
            st.active++;
            st.out_edges += outdeg[d];
        }
        next[d] = ch;
    }
};

// ============================================================================
// Prepass / vertex kernels
// ============================================================================

template <class V>
__global__ void k_pr_contrib(const V *__restrict__ p_curr, const uint32_t *__restrict__ outdeg,
                             const uint8_t *__restrict__ F, int bit, V *__restrict__ contrib, uint32_t n)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint32_t od = outdeg[v];
    contrib[v] = ((bit || F[v]) && od) ? p_curr[v] / (V)od : path_none_value<V>();
}

__global__ void k_cc_gate(const uint32_t *__restrict__ prev, const uint8_t *__restrict__ F, int bit,
                          uint32_t *__restrict__ g, uint32_t n)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) g[v] = (bit || F[v]) ? prev[v] : 0xffffffffu;
}

// BFS bottom-up (edgeMapDenseCSC, bit=false, no cache: ligra-numa.h:706-713):
// an unvisited vertex scans its in-list (ascending source) and takes the
// first frontier source as parent, then stops because cond() turns false.
__global__ __launch_bounds__(kBlock) void k_bfs_pull(const uint32_t *__restrict__ toff,
                                                     const uint32_t *__restrict__ tids,
                                                     const uint32_t *__restrict__ bm, int bit,
                                                     int32_t *__restrict__ parents,
                                                     int32_t *__restrict__ levels, int32_t level,
                                                     uint8_t *__restrict__ next,
                                                     const uint32_t *__restrict__ outdeg, uint32_t n,
                                                     unsigned long long *__restrict__ counters)
{
    LocalStats st;
    unsigned long long scanned = 0;
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < n) {
        uint8_t nx = 0;
        if (parents[d] == -1) {
            const uint32_t b = toff[d], e = toff[d + 1];
            for (uint32_t j = b; j < e; j++) {
                const uint32_t s = tids[j];
                scanned++;
                if (bit || bit_test(bm, s)) {
                    parents[d] = (int32_t)s;
                    if (levels) levels[d] = level;
                    nx = 1;
                    st.active++;
                    st.out_edges += outdeg[d];
                    break;
                }
            }
        }
        next[d] = nx;
    }
    // one counter add per work-group (not per wave): the striped counters
    // take about 6.5e4 adds per level instead of 2.6e5
    __shared__ unsigned long long s_scan[kBlock / 64];
    unsigned long long sc = wave_sum_u64(scanned);
    if ((threadIdx.x & 63) == 0) s_scan[threadIdx.x >> 6] = sc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < kBlock / 64; k++) t += s_scan[k];
        cnt_add(counters, CNT_SCANNED, t);
    }
    block_stats(st, counters);
}

// BC_F / BC_Back_F (app/BC.C:28-124) on the dense branch: a[d] += sum of
// the frontier in-neighbours' values for d not yet visited; d joins the
// output when a[d] was 0 (its first update in the reference's loop).
// Sources outside the frontier gather the "no contribution" NaN.
template <class V>
__global__ void k_pathsum_gate(const V *__restrict__ a, const uint8_t *__restrict__ F, int bit, V *__restrict__ gate,
                               uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) gate[v] = (bit || F[v]) ? a[v] : path_none_value<V>();
}

template <class V>
struct OpPathSum {
    using acc_t = V;
    const V *gate;
    V *a;
    const uint8_t *vis;
    uint8_t *next;
    const uint32_t *outdeg;
    const uint32_t *bm; // the frontier's bitmap (unread when 'bit')
    int bit;
    __device__ bool skip() const { return false; }
    __device__ V identity() const { return path_none_value<V>(); }
    __device__ V combine(V x, V y) const { return path_none(x) ? y : path_none(y) ? x : x + y; }
    // the frontier bit first (a 2 MiB bitmap that stays in L2): only the
    // frontier's sources read their gated value
    __device__ V gather(uint32_t, uint32_t s) const
    {
        return (bit || bit_test(bm, s)) ? gate[s] : path_none_value<V>();
    }
    // a visited row cannot change: a work item of visited rows only skips
    __device__ bool row_done(uint32_t d) const { return vis[d] != 0; }
    __device__ void finish(uint32_t d, V acc, LocalStats &st) const
    {
        bool emit = false;
        if (!vis[d] && !path_none(acc)) {
            const V old = a[d];
            a[d] = old + acc;
            emit = old == V(0);
        }
        next[d] = emit;
        if (emit) {
            st.active++;
            st.out_edges += outdeg[d];
        }
    }
};

// BF_F (app/BellmanFord.C:28-92) on the dense branch: the frontier's
// distances are copied first (gate, INT_MIN elsewhere), so every vertex
// relaxes against last round's values and the round is deterministic.
__global__ void k_sp_gate(const int32_t *__restrict__ sp, const uint8_t *__restrict__ F, int bit,
                          int32_t *__restrict__ gate, uint32_t n)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) gate[v] = (bit || F[v]) ? sp[v] : INT32_MIN;
}

struct OpSPRelax {
    using acc_t = int32_t;
    const int32_t *gate;
    const float *tw; // in-edge order, NULL: unit lengths
    int32_t *sp;
    int32_t *vis;
    uint8_t *next;
    const uint32_t *outdeg;
    const uint32_t *bm; // the frontier's bitmap (unread when 'bit')
    int bit;
    __device__ bool skip() const { return false; }
    __device__ int32_t identity() const { return INT32_MAX; }
    __device__ int32_t combine(int32_t a, int32_t b) const { return a < b ? a : b; }
    // the frontier bit first (a bitmap that stays in L2): only the
    // frontier's sources read their distance and the edge's length
    __device__ int32_t gather(uint32_t e, uint32_t s) const
    {
        if (!bit && !bit_test(bm, s)) return INT32_MAX;
        const int32_t g = gate[s];
        const int32_t w = tw ? (int32_t)__builtin_nontemporal_load(&tw[e]) : 1;
        return g == INT32_MIN ? INT32_MAX : g + w;
    }
    __device__ void finish(uint32_t d, int32_t acc, LocalStats &st) const
    {
        bool emit = false;
        if (acc < sp[d]) {
            sp[d] = acc;
            if (vis[d] == 0) {
                vis[d] = 1;
                emit = true;
            }
        }
        next[d] = emit;
        if (emit) {
            st.active++;
            st.out_edges += outdeg[d];
        }
    }
};

// Count and out-degree sum of a dense flag array (GoutDegree + reduce,
// ligra-numa.h:985-988).
__global__ __launch_bounds__(kBlock) void k_dense_stats(const uint8_t *__restrict__ F,
                                                        const uint32_t *__restrict__ outdeg,
                                                        uint32_t n, unsigned long long *counters)
{
    LocalStats st;
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (; v < n; v += stride)
        if (F[v]) {
            st.active++;
            st.out_edges += outdeg[v];
        }
    block_stats(st, counters);
}

// ---- sparse push (edgeMapSparseWithG, ligra-numa.h:868-916) -----------------

enum { PUSH_PR = 0, PUSH_BFS = 1, PUSH_CC = 2, PUSH_SPMV = 3, PUSH_SSSP = 4, PUSH_PATHSUM = 5 };

struct PushArgs {
    const uint32_t *list;   // frontier ids
    const uint32_t *offs;   // exclusive scan of their out-degrees (k+1)
    uint32_t k;
    uint64_t total;
    const uint32_t *off, *ids, *outdeg;
    const float *w;         // CSR weights or NULL
    int kind, f32;
    void *a, *b;            // op arrays
    int32_t *levels;
    int32_t level;
    uint32_t *out;          // output list (capacity total)
    unsigned long long *counters;
};

constexpr int kPushTile = kBlock * 4;
constexpr int kPushLds = 2048;

__device__ __forceinline__ uint32_t upper_idx(const uint32_t *a, uint32_t lo, uint32_t hi, uint64_t e)
{
    // last i in [lo, hi) with a[i] <= e
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)a[mid] <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The same search by a whole wave (every lane calls it with the same
// arguments; a[lo] <= e): 64 probes a step, so a list of 2^22 entries takes
// four dependent loads instead of 22.
__device__ __forceinline__ uint32_t wave_upper_idx(const uint32_t *a, uint32_t lo, uint32_t hi, uint64_t e)
{
    const uint32_t lane = threadIdx.x & 63;
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + 63) / 64;
        const uint32_t idx = lo + lane * step;
        const bool ok = idx < hi && (uint64_t)a[idx] <= e;
        const uint64_t m = __ballot(ok); // lane 0 (idx = lo) is always set
        const uint32_t last = 63u - (uint32_t)__builtin_clzll(m);
        lo += last * step;
        hi = min(lo + step, hi);
    }
    return lo;
}

// One update of edge (v -> ngh) by the functor of A.kind (the apps'
// updateAtomic: BFS_F, CC_F, PR_F, SPMV_F, BF_F, BC_F), w[pos] its weight;
// returns whether ngh joins the output frontier.  Shared by the sparse push
// and the COO traversal.
__device__ __forceinline__ bool edge_update(const PushArgs &A, uint32_t v, uint32_t ngh, uint64_t pos)
{
    bool emit = false;
    switch (A.kind) {
    case PUSH_BFS: {
        int32_t *par = (int32_t *)A.a;
        if (par[ngh] == -1 && atomicCAS(&par[ngh], -1, (int32_t)v) == -1) {
            emit = true;
            if (A.levels) A.levels[ngh] = A.level;
        }
        break;
    }
    case PUSH_CC: {
        uint32_t *ids = (uint32_t *)A.a;
        const uint32_t *prev = (const uint32_t *)A.b;
        const uint32_t val = prev[v];
        if (val < ids[ngh]) {
            const uint32_t old = atomicMin(&ids[ngh], val);
            emit = val < old && old == prev[ngh];
        }
        break;
    }
    case PUSH_PATHSUM: { // BC_F::updateAtomic: an add; first update emits
        const uint8_t *vis = (const uint8_t *)A.b;
        if (!vis[ngh]) {
            if (A.f32) {
                float *x = (float *)A.a;
                emit = unsafeAtomicAdd(&x[ngh], x[v]) == 0.0f;
            } else {
                double *x = (double *)A.a;
                emit = unsafeAtomicAdd(&x[ngh], x[v]) == 0.0;
            }
        }
        break;
    }
    case PUSH_SSSP: { // BF_F::updateAtomic: writeMin && CAS(Visited, 0, 1)
        int32_t *sp = (int32_t *)A.a;
        int32_t *vis = (int32_t *)A.b;
        const int32_t nd = sp[v] + (A.w ? (int32_t)A.w[pos] : 1);
        if (nd < sp[ngh]) {
            const int32_t old = atomicMin(&sp[ngh], nd);
            if (nd < old) emit = atomicCAS(&vis[ngh], 0, 1) == 0;
        }
        break;
    }
    case PUSH_PR:
        if (A.f32) {
            const float *pc = (const float *)A.a;
            unsafeAtomicAdd(&((float *)A.b)[ngh], pc[v] / (float)A.outdeg[v]);
        } else {
            const double *pc = (const double *)A.a;
            unsafeAtomicAdd(&((double *)A.b)[ngh], pc[v] / (double)A.outdeg[v]);
        }
        emit = true;
        break;
    default: { // SPMV
        const float wv = A.w ? A.w[pos] : 1.0f;
        if (A.f32) unsafeAtomicAdd(&((float *)A.b)[ngh], ((const float *)A.a)[v] * wv);
        else unsafeAtomicAdd(&((double *)A.b)[ngh], ((const double *)A.a)[v] * (double)wv);
        emit = true;
        break;
    }
    }
    return emit;
}

__global__ __launch_bounds__(kBlock) void k_push(PushArgs A)
{
    __shared__ uint32_t s_off[kPushLds];
    __shared__ uint32_t s_lo, s_hi;
    // the edge range is the host's total clamped to the scan's own
    const uint64_t total = min(A.total, (uint64_t)A.offs[A.k]);
    const uint64_t t0 = (uint64_t)blockIdx.x * kPushTile;
    if (t0 >= total) return; // whole work-group: no barrier is skipped by part of it
    const uint64_t t1 = min(t0 + (uint64_t)kPushTile, total);
    // the tile's first and last entries: two waves search at once
    const uint32_t wv = threadIdx.x >> 6;
    if (wv == 0) {
        const uint32_t r = wave_upper_idx(A.offs, 0, A.k, t0);
        if (threadIdx.x == 0) s_lo = r;
    } else if (wv == 1) {
        const uint32_t r = wave_upper_idx(A.offs, 0, A.k, t1 - 1) + 1;
        if (threadIdx.x == 64) s_hi = r;
    }
    __syncthreads();
    const uint32_t lo = s_lo, hi = s_hi;
    const bool in_lds = hi - lo <= (uint32_t)kPushLds;
    if (in_lds)
        for (uint32_t i = threadIdx.x; i < hi - lo; i += kBlock) s_off[i] = A.offs[lo + i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (int it = 0; it < kPushTile / kBlock; it++) {
        const uint64_t e = t0 + (uint64_t)it * kBlock + threadIdx.x;
        bool emit = false;
        uint32_t ngh = 0;
        if (e < t1) {
            uint32_t kk;
            if (in_lds) kk = lo + upper_idx(s_off, 0, hi - lo, e);
            else kk = upper_idx(A.offs, lo, hi, e);
            const uint32_t v = A.list[kk];
            const uint32_t pos = A.off[v] + (uint32_t)(e - A.offs[kk]);
            ngh = A.ids[pos];
            emit = edge_update(A, v, ngh, pos);
        }
        // wave-ballot compaction of the emitted destinations
        const uint64_t mask = __ballot(emit);
        const uint32_t cnt = (uint32_t)__popcll(mask);
        unsigned long long od = emit ? A.outdeg[ngh] : 0;
        od = wave_sum_u64(od);
        unsigned long long base = 0;
        if (lane == 0 && cnt) {
            base = atomicAdd(&A.counters[CNT_CURSOR], (unsigned long long)cnt);
            cnt_add(A.counters, CNT_ACTIVE, (unsigned long long)cnt);
            cnt_add(A.counters, CNT_OUT, od);
        }
        base = __shfl(base, 0, 64);
        if (emit) {
            const uint64_t below = mask & ((1ULL << lane) - 1ULL);
            A.out[base + __popcll(below)] = ngh;
        }
    }
}

// remDuplicates (ligra-numa.h:317-348): the lowest slot of each id wins.
__global__ void k_dedup_mark(const uint32_t *__restrict__ list, uint64_t k, uint32_t *__restrict__ flags)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < k) atomicMin(&flags[list[i]], (uint32_t)i);
}

__global__ void k_dedup_keep(const uint32_t *__restrict__ list, uint64_t k, uint32_t *__restrict__ flags,
                             const uint32_t *__restrict__ outdeg, uint32_t *__restrict__ out,
                             unsigned long long *__restrict__ counters)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool keep = false;
    uint32_t v = 0;
    if (i < k) {
        v = list[i];
        keep = flags[v] == (uint32_t)i;
    }
    __syncthreads();
    if (keep) flags[v] = 0xffffffffu;
    const int lane = threadIdx.x & 63;
    const uint64_t mask = __ballot(keep);
    const uint32_t cnt = (uint32_t)__popcll(mask);
    unsigned long long od = wave_sum_u64(keep ? outdeg[v] : 0ull);
    unsigned long long base = 0;
    if (lane == 0 && cnt) {
        base = atomicAdd(&counters[CNT_CURSOR], (unsigned long long)cnt);
        cnt_add(counters, CNT_ACTIVE, (unsigned long long)cnt);
        cnt_add(counters, CNT_OUT, od);
    }
    base = __shfl(base, 0, 64);
    if (keep) out[base + __popcll(mask & ((1ULL << lane) - 1ULL))] = v;
}

// ---- frontier conversions --------------------------------------------------

__global__ void k_scatter_flags(const uint32_t *__restrict__ list, uint64_t k, uint8_t *__restrict__ F)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < k) F[list[i]] = 1;
}

// Ordered compaction (packIndex, utils.h:404): block b handles 4096 flags.
constexpr int kPackPer = kBlock * 16;
__global__ __launch_bounds__(kBlock) void k_pack_count(const uint8_t *__restrict__ F, uint32_t n,
                                                       uint32_t *__restrict__ bcount)
{
    const uint32_t b0 = blockIdx.x * kPackPer;
    uint32_t c = 0;
    for (int i = 0; i < 16; i++) {
        const uint32_t v = b0 + i * kBlock + threadIdx.x;
        if (v < n && F[v]) c++;
    }
    __shared__ uint32_t s[kBlock];
    s[threadIdx.x] = c;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) bcount[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(kBlock) void k_pack_write(const uint8_t *__restrict__ F, uint32_t n,
                                                       const uint32_t *__restrict__ bbase,
                                                       uint32_t *__restrict__ out)
{
    // thread t owns flags [b0 + 16t, b0 + 16t + 16) so the output keeps id order
    const uint32_t b0 = blockIdx.x * kPackPer + threadIdx.x * 16;
    uint32_t c = 0;
    for (int i = 0; i < 16; i++) {
        const uint32_t v = b0 + i;
        if (v < n && F[v]) c++;
    }
    __shared__ uint32_t s[kBlock];
    s[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        uint32_t add = threadIdx.x >= (uint32_t)off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t pos = bbase[blockIdx.x] + s[threadIdx.x] - c;
    for (int i = 0; i < 16; i++) {
        const uint32_t v = b0 + i;
        if (v < n && F[v]) out[pos++] = v;
    }
}

// ---- vertexMap kernels ---------------------------------------------------------

template <class V>
__device__ __forceinline__ void vop_apply(int kind, uint32_t i, void *a, void *b, double c0, double c1)
{
    V *pa = (V *)a;
    V *pb = (V *)b;
    switch (kind) {
    case GG_VOP_PAGERANK: pb[i] = (V)c0 * pb[i] + (V)c1; break;
    case GG_VOP_RESET: pa[i] = V(0); break;
    case GG_VOP_ABSDIFF: pa[i] = fabs(pa[i] - pb[i]); break;
    default: break;
    }
}

template <class V>
__global__ void k_vmap(int kind, int mode, const uint8_t *__restrict__ F,
                       const uint32_t *__restrict__ list, uint64_t count, void *a, void *b, double c0,
                       double c1)
{
    // mode 0: all vertices, 1: dense flags, 2: sparse list
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) {
        uint32_t v;
        if (mode == 2) v = list[i];
        else {
            v = (uint32_t)i;
            if (mode == 1 && !F[v]) continue;
        }
        if (kind == GG_VOP_CC_COPY) ((uint32_t *)b)[v] = ((const uint32_t *)a)[v];
        else vop_apply<V>(kind, v, a, b, c0, c1);
    }
}

// vertexFilter (ligra-numa.h:1053-1082): out[i] = active(i) && keep[i]
__global__ void k_vfilter(int mode, const uint8_t *__restrict__ F, const uint8_t *__restrict__ keep,
                          uint8_t *__restrict__ out, uint32_t n)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) out[v] = (mode == 0 || F[v]) && keep[v];
}

// Fixed-order sum: block b sums a fixed strided slice, partials reduced by
// one block in index order.
template <class V>
__global__ __launch_bounds__(kBlock) void k_sum_part(const V *__restrict__ a, uint64_t n, double *part)
{
    double s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kBlock)
        s += (double)a[i];
    s = wave_sum_f64(s);
    __shared__ double w[kBlock / 64];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < kBlock / 64; k++) t += w[k];
        part[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kBlock) void k_sum_final(const double *__restrict__ part, uint32_t np,
                                                      double *out)
{
    double s = 0;
    for (uint32_t i = threadIdx.x; i < np; i += kBlock) s += part[i];
    s = wave_sum_f64(s);
    __shared__ double w[kBlock / 64];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < kBlock / 64; k++) t += w[k];
        *out = t;
    }
}

// ============================================================================
// Frontier helpers
// ============================================================================

int frontier_alloc(Graph *g, Frontier *f)
{
    f->g = g;
    f->ctx = g->ctx;
    f->n = g->n;
    return GG_OK;
}

static int ensure_bytes(Frontier *f)
{
    return pool_ensure(f->ctx, f->bytes, ((uint64_t)f->n + 64) & ~63ull);
}

static int ensure_bitmap(Frontier *f)
{
    return pool_ensure(f->ctx, f->bitmap, 4ull * ((f->n + 31) / 32 + 1));
}

// Dense flags (+ bitmap) from whatever the frontier holds.
int frontier_make_dense(Frontier *f)
{
    Ctx *c = f->g->ctx;
    if (f->bit) {
        // flags already held (a frontier whose bit was set later) stay: the
        // dense kernels take 'bit' and do not read them
        if (f->dense_valid) return GG_OK;
        GG_TRY(ensure_bytes(f));
        GG_TRY(fill_bytes(c, f->bytes.p, f->n, 1));
        GG_TRY(ensure_bitmap(f));
        GG_TRY(bitmap_from_bytes(c, (uint8_t *)f->bytes.p, (uint32_t *)f->bitmap.p, f->n));
        f->dense_valid = true;
        return GG_OK;
    }
    if (f->dense_valid) return GG_OK;
    if (!f->sparse_valid) return fail(GG_ERR_STATE, "frontier has no representation");
    GG_TRY(ensure_bytes(f));
    GG_TRY(fill_bytes(c, f->bytes.p, f->n, 0));
    if (f->d_m)
        k_scatter_flags<<<blocks_for(f->d_m, kBlock), kBlock, 0, c->stream>>>(
            (const uint32_t *)f->list.p, f->d_m, (uint8_t *)f->bytes.p);
    GG_TRY(ensure_bitmap(f));
    GG_TRY(bitmap_from_bytes(c, (uint8_t *)f->bytes.p, (uint32_t *)f->bitmap.p, f->n));
    GG_HIP(hipGetLastError());
    f->dense_valid = true;
    return GG_OK;
}

int frontier_make_sparse(Frontier *f)
{
    Ctx *c = f->g->ctx;
    if (f->sparse_valid) return GG_OK;
    if (f->bit && !f->dense_valid) {
        // bits() frontier: its flags are not set in the reference (toSparse
        // on it would read uninitialised memory); list every vertex
        GG_TRY(pool_ensure(c, f->list, 4ull * f->n + 4));
        GG_TRY(gg_iota(abi(c), f->list.p, GG_U32, f->n));
        f->sparse_valid = true;
        return GG_OK;
    }
    if (!f->dense_valid) return fail(GG_ERR_STATE, "frontier has no representation");
    uint32_t nb = blocks_for(f->n, kPackPer);
    Buf cnt, base;
    GG_TRY(pool_take(c, cnt, 4ull * (nb + 1)));
    GG_TRY(pool_take(c, base, 4ull * (nb + 1)));
    k_pack_count<<<nb, kBlock, 0, c->stream>>>((const uint8_t *)f->bytes.p, f->n, (uint32_t *)cnt.p);
    size_t tb = 0;
    GG_HIP(rocprim::exclusive_scan(nullptr, tb, (uint32_t *)cnt.p, (uint32_t *)base.p, 0u, (size_t)nb,
                                   rocprim::plus<uint32_t>(), c->stream));
    GG_TRY(buf_ensure(c, c->tmp, tb + 16));
    GG_HIP(rocprim::exclusive_scan(c->tmp.p, tb, (uint32_t *)cnt.p, (uint32_t *)base.p, 0u, (size_t)nb,
                                   rocprim::plus<uint32_t>(), c->stream));
    // capacity n whatever d_m says, so a stale count cannot overrun the list
    GG_TRY(pool_ensure(c, f->list, 4ull * f->n + 4));
    k_pack_write<<<nb, kBlock, 0, c->stream>>>((const uint8_t *)f->bytes.p, f->n,
                                                (const uint32_t *)base.p, (uint32_t *)f->list.p);
    GG_HIP(hipGetLastError());
    uint32_t last[2] = {0, 0};
    GG_HIP(hipMemcpyAsync(c->hpin, (uint32_t *)base.p + (nb - 1), 4, hipMemcpyDeviceToHost, c->stream));
    GG_HIP(hipMemcpyAsync((uint32_t *)c->hpin + 1, (uint32_t *)cnt.p + (nb - 1), 4, hipMemcpyDeviceToHost,
                          c->stream));
    GG_TRY(sync(c));
    std::memcpy(last, c->hpin, 8);
    pool_give(c, cnt);
    pool_give(c, base);
    if ((uint64_t)last[0] + last[1] != f->d_m) // toSparse's "bad stored value of m" (ligra-numa.h:251-255)
        return fail(GG_ERR_STATE, "toSparse: stored count " + std::to_string(f->d_m) + " but " +
                                      std::to_string((uint64_t)last[0] + last[1]) + " flags set");
    f->sparse_valid = true;
    return GG_OK;
}

__global__ void k_counters_fold(unsigned long long *counters)
{
    const int t = threadIdx.x;
    unsigned long long *st = counters + kCntStripeBase + t * kStripeStride;
    unsigned long long a = st[CNT_ACTIVE], o = st[CNT_OUT], sc = st[CNT_SCANNED];
    st[CNT_ACTIVE] = 0;
    st[CNT_OUT] = 0;
    st[CNT_SCANNED] = 0;
    a = wave_sum_u64(a);
    o = wave_sum_u64(o);
    sc = wave_sum_u64(sc);
    if (t == 0) {
        counters[CNT_ACTIVE] += a;
        counters[CNT_OUT] += o;
        counters[CNT_SCANNED] += sc;
    }
}

int counters_zero(Ctx *c)
{
    return fill_bytes(c, c->counters.p, kCounterBytes, 0);
}

int counters_read(Ctx *c, unsigned long long out[4])
{
    k_counters_fold<<<1, kStripes, 0, c->stream>>>((unsigned long long *)c->counters.p);
    GG_HIP(hipGetLastError());
    return read_back(c, out, c->counters.p, 32);
}


int frontier_stats_dense(Frontier *f)
{
    Ctx *c = f->g->ctx;
    GG_TRY(counters_zero(c));
    uint32_t grid = std::min<uint32_t>(blocks_for(f->n, kBlock), 4096);
    k_dense_stats<<<grid, kBlock, 0, c->stream>>>((const uint8_t *)f->bytes.p,
                                                   (const uint32_t *)f->g->outdeg.p, f->n,
                                                   (unsigned long long *)c->counters.p);
    GG_HIP(hipGetLastError());
    unsigned long long cn[4];
    GG_TRY(counters_read(c, cn));
    f->d_m = cn[CNT_ACTIVE];
    f->out_edges = cn[CNT_OUT];
    return GG_OK;
}

static void frontier_free(Frontier *f)
{
    pool_give(f->ctx, f->bytes);
    pool_give(f->ctx, f->bitmap);
    pool_give(f->ctx, f->list);
}

// ============================================================================
// edgeMap core (shared with the fused app loops)
// ============================================================================

int edgemap_core(Graph *g, Frontier *in, const gg_op_t *op, int64_t threshold, int remdups,
                 Frontier **pout, int *path, EdgeMapExtra *ex)
{
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    if (threshold == -1) threshold = (int64_t)(m / 20);
    const int64_t dense_thr = (int64_t)(m / 2);
    if (in->n != n)
        return fail(GG_ERR_SIZE, "edgeMap: Sizes Don't match (frontier rows " + std::to_string(in->n) +
                                     ", graph n " + std::to_string(n) + ")");
    const bool f32 = op->dtype == GG_F32;
    Frontier *out = new Frontier();
    frontier_alloc(g, out);
    auto bail = [&](int rc) {
        frontier_free(out);
        delete out;
        return rc;
    };
    if (in->out_edges == 0) { // partitioned_vertices::empty() (ligra-numa.h:947)
        out->n = 0;
        out->d_m = 0;
        *pout = out;
        if (path) *path = GG_PATH_EMPTY;
        return GG_OK;
    }
    int rc;
    if ((int64_t)(in->d_m + in->out_edges) > threshold) {
        // ---------------- dense ----------------
        if (!in->bit && (rc = frontier_make_dense(in))) return bail(rc);
        if ((rc = ensure_bytes(out))) return bail(rc);
        const bool coo = (int64_t)(in->d_m + in->out_edges) > dense_thr && g->P && !g->part_vertex;
        if (path) *path = coo ? GG_PATH_COO : GG_PATH_CSC;
        // In-edge lists the pull reads: the twin (in-lists by source id) or,
        // on the COO branch with GG_OPT_DENSE_COO, the same rows with each
        // destination's edges in its COO partition's Hilbert order, the
        // order in which the reference's PART96 loop adds them
        // (ligra-numa.h:803-866); either way one owner per destination.
        const uint32_t *tids_v = (const uint32_t *)g->tids.p;
        const float *tw_v = (const float *)g->tw.p;
        CscView pv;
        if ((rc = pull_view(g, &pv))) return bail(rc);
        if (coo && g->dense_coo && g->coo_src.p && m) {
            if ((rc = ensure_coo_pull(g))) return bail(rc);
            tids_v = (const uint32_t *)g->coo_tids.p;
            tw_v = (const float *)g->coo_tw.p;
            pv.tids = tids_v;
        }
        if ((rc = counters_zero(c))) return bail(rc);
        unsigned long long *cnt = (unsigned long long *)c->counters.p;
        const int bit = in->bit ? 1 : 0;
        const uint8_t *F = (const uint8_t *)in->bytes.p;
        const uint32_t *bm = (const uint32_t *)in->bitmap.p;
        uint8_t *next = (uint8_t *)out->bytes.p;
        const uint32_t *outdeg = (const uint32_t *)g->outdeg.p;
        Buf tmpv;
        switch (op->kind) {
        case GG_OP_PAGERANK: {
            if ((rc = pool_take(c, tmpv, (f32 ? 4ull : 8ull) * n))) return bail(rc);
            if (f32) {
                k_pr_contrib<float><<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>(
                    (const float *)op->a, outdeg, F, bit, (float *)tmpv.p, n);
                rc = launch_pull(c, pv, OpPRSum<float>{(const float *)tmpv.p, (float *)op->b, next, outdeg},
                                 cnt, nullptr, nullptr);
            } else {
                k_pr_contrib<double><<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>(
                    (const double *)op->a, outdeg, F, bit, (double *)tmpv.p, n);
                rc = launch_pull(c, pv, OpPRSum<double>{(const double *)tmpv.p, (double *)op->b, next, outdeg},
                                 cnt, nullptr, nullptr);
            }
            break;
        }
        case GG_OP_SPMV:
        case GG_OP_SUM: {
            const float *tw = (g->weighted && op->kind == GG_OP_SPMV) ? tw_v : nullptr;
            if (bit) {
                if (f32)
                    rc = launch_pull(c, pv, OpSpmvBit<float>{(const float *)op->a, tw, (float *)op->b, next, outdeg,
                                                            (const uint32_t *)g->indeg.p},
                                     cnt, nullptr, nullptr);
                else
                    rc = launch_pull(c, pv, OpSpmvBit<double>{(const double *)op->a, tw, (double *)op->b, next,
                                                             outdeg, (const uint32_t *)g->indeg.p},
                                     cnt, nullptr, nullptr);
            } else if (f32) {
                rc = launch_pull(c, pv, OpSpmvFrontier<float>{(const float *)op->a, tw, bm, (float *)op->b, next,
                                                             outdeg},
                                 cnt, nullptr, nullptr);
            } else {
                rc = launch_pull(c, pv, OpSpmvFrontier<double>{(const double *)op->a, tw, bm, (double *)op->b, next,
                                                              outdeg},
                                 cnt, nullptr, nullptr);
            }
            break;
        }
        case GG_OP_CC: {
            // The COO and CSC branches compute the same Jacobi round (the
            // minimum prevID over active in-neighbours); the destination-
            // owned pull reads 4 bytes per edge where the COO lists read 8
            // and needs no atomics, so both branches run it.
            if ((rc = pool_take(c, tmpv, 4ull * n))) return bail(rc);
            k_cc_gate<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>((const uint32_t *)op->b, F, bit,
                                                                        (uint32_t *)tmpv.p, n);
            rc = launch_pull(c, pv, OpCCMin{(const uint32_t *)tmpv.p, (uint32_t *)op->a,
                                           (const uint32_t *)op->b, next, outdeg, bm, bit},
                             cnt, nullptr, nullptr);
            break;
        }
        case GG_OP_BFS: {
            // Both dense branches visit the same vertices in this level; the
            // bottom-up pull stops at an unvisited vertex's first frontier
            // in-neighbour (ligra-numa.h:711-712) where the COO lists scan
            // every edge, so both branches run the pull.  Levels are
            // identical; parents are a valid choice either way (the
            // reference's own choice depends on its thread schedule).
            {
                k_bfs_pull<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>(
                    (const uint32_t *)g->toff.p, tids_v, bm, bit, (int32_t *)op->a,
                    ex ? ex->levels : nullptr, ex ? ex->level : 0, next, outdeg, n, cnt);
            }
            rc = GG_OK;
            break;
        }
        case GG_OP_PATHSUM: {
            // the frontier's values are copied first (the identity NaN
            // elsewhere): the sources are visited and the updated vertices
            // are not, so the copy changes no value but keeps the pull free
            // of aliasing between gathered and written entries
            if ((rc = pool_take(c, tmpv, (f32 ? 4ull : 8ull) * n))) return bail(rc);
            if (f32) {
                k_pathsum_gate<float><<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>((const float *)op->a, F, bit,
                                                                                       (float *)tmpv.p, n);
                rc = launch_pull(c, pv, OpPathSum<float>{(const float *)tmpv.p, (float *)op->a,
                                                         (const uint8_t *)op->b, next, outdeg, bm, bit},
                                 cnt, nullptr, nullptr);
            } else {
                k_pathsum_gate<double><<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>(
                    (const double *)op->a, F, bit, (double *)tmpv.p, n);
                rc = launch_pull(c, pv, OpPathSum<double>{(const double *)tmpv.p, (double *)op->a,
                                                          (const uint8_t *)op->b, next, outdeg, bm, bit},
                                 cnt, nullptr, nullptr);
            }
            break;
        }
        case GG_OP_SSSP: {
            if ((rc = pool_take(c, tmpv, 4ull * n))) return bail(rc);
            k_sp_gate<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>((const int32_t *)op->a, F, bit,
                                                                        (int32_t *)tmpv.p, n);
            const float *tw = g->weighted ? tw_v : nullptr;
            rc = launch_pull(c, pv, OpSPRelax{(const int32_t *)tmpv.p, tw, (int32_t *)op->a, (int32_t *)op->b, next,
                                              outdeg, bm, bit},
                             cnt, nullptr, nullptr);
            break;
        }
        default: return bail(fail(GG_ERR_INVALID, "edgeMap: unknown operator kind"));
        }
        if (rc) { pool_give(c, tmpv); return bail(rc); }
        GG_HIP(hipGetLastError());
        out->has_dense = true;
        out->dense_valid = true;
        {
            unsigned long long cn[4];
            if ((rc = counters_read(c, cn))) return bail(rc);
            out->d_m = cn[CNT_ACTIVE];
            out->out_edges = cn[CNT_OUT];
            out->deg_src = g->outdeg.p;
            if (ex) ex->scanned += op->kind == GG_OP_BFS ? cn[CNT_SCANNED] : m;
        }
        pool_give(c, tmpv);
        if ((rc = ensure_bitmap(out))) return bail(rc);
        if ((rc = bitmap_from_bytes(c, (uint8_t *)out->bytes.p, (uint32_t *)out->bitmap.p, n))) return bail(rc);
        *pout = out;
        return GG_OK;
    }
    // ---------------- sparse push ----------------
    if (path) *path = GG_PATH_SPARSE;
    if ((rc = frontier_make_sparse(in))) return bail(rc);
    const uint32_t k = (uint32_t)in->d_m;
    Buf offs;
    if ((rc = pool_take(c, offs, 4ull * (k + 1)))) return bail(rc);
    {
        struct DegOf {
            const uint32_t *list, *outdeg;
            uint32_t k;
            __host__ __device__ uint32_t operator()(uint32_t i) const { return i < k ? outdeg[list[i]] : 0u; }
        };
        auto it = rocprim::make_transform_iterator(rocprim::make_counting_iterator<uint32_t>(0),
                                                   DegOf{(const uint32_t *)in->list.p,
                                                         (const uint32_t *)g->outdeg.p, k});
        size_t tb = 0;
        GG_HIP(rocprim::exclusive_scan(nullptr, tb, it, (uint32_t *)offs.p, 0u, (size_t)k + 1,
                                       rocprim::plus<uint32_t>(), c->stream));
        if ((rc = buf_ensure(c, c->tmp, tb + 16))) { pool_give(c, offs); return bail(rc); }
        GG_HIP(rocprim::exclusive_scan(c->tmp.p, tb, it, (uint32_t *)offs.p, 0u, (size_t)k + 1,
                                       rocprim::plus<uint32_t>(), c->stream));
    }
    // The scan total is the frontier's out-degree sum.  When the engine
    // summed it over these very entries and degrees, no read-back is
    // needed; k_push clamps its range to the device total either way.
    uint64_t total = in->out_edges;
    if (in->deg_src != g->outdeg.p || total > 0xffffffffull) {
        uint32_t total32 = 0;
        if ((rc = read_back(c, &total32, (uint32_t *)offs.p + k, 4))) { pool_give(c, offs); return bail(rc); }
        total = total32;
    }
    if ((rc = pool_ensure(c, out->list, 4ull * (total ? total : 1)))) { pool_give(c, offs); return bail(rc); }
    if ((rc = counters_zero(c))) { pool_give(c, offs); return bail(rc); }
    if (total) {
        PushArgs A;
        A.list = (const uint32_t *)in->list.p;
        A.offs = (const uint32_t *)offs.p;
        A.k = k;
        A.total = total;
        A.off = (const uint32_t *)g->off.p;
        A.ids = (const uint32_t *)g->ids.p;
        A.outdeg = (const uint32_t *)g->outdeg.p;
        A.w = (g->weighted && op->kind != GG_OP_SUM) ? (const float *)g->w.p : nullptr;
        A.kind = op->kind == GG_OP_BFS ? PUSH_BFS
                 : op->kind == GG_OP_CC ? PUSH_CC
                 : op->kind == GG_OP_PAGERANK ? PUSH_PR
                 : op->kind == GG_OP_SSSP ? PUSH_SSSP
                 : op->kind == GG_OP_PATHSUM ? PUSH_PATHSUM
                                              : PUSH_SPMV;
        A.f32 = f32;
        A.a = op->a;
        A.b = op->b;
        A.levels = ex ? ex->levels : nullptr;
        A.level = ex ? ex->level : 0;
        A.out = (uint32_t *)out->list.p;
        A.counters = (unsigned long long *)c->counters.p;
        k_push<<<blocks_for(total, kPushTile), kBlock, 0, c->stream>>>(A);
        GG_HIP(hipGetLastError());
    }
    if (ex) ex->scanned += total;
    unsigned long long cn[4];
    if ((rc = counters_read(c, cn))) { pool_give(c, offs); return bail(rc); }
    pool_give(c, offs);
    out->d_m = cn[CNT_CURSOR];
    out->out_edges = cn[CNT_OUT];
    out->deg_src = g->outdeg.p;
    if (remdups && out->d_m) {
        Buf flags, list2;
        if ((rc = pool_take(c, flags, 4ull * n))) return bail(rc);
        if ((rc = pool_take(c, list2, 4ull * out->d_m))) { pool_give(c, flags); return bail(rc); }
        fill_u32(c, flags.p, n, 0xffffffffu);
        counters_zero(c);
        k_dedup_mark<<<blocks_for(out->d_m, kBlock), kBlock, 0, c->stream>>>((const uint32_t *)out->list.p,
                                                                              out->d_m, (uint32_t *)flags.p);
        k_dedup_keep<<<blocks_for(out->d_m, kBlock), kBlock, 0, c->stream>>>(
            (const uint32_t *)out->list.p, out->d_m, (uint32_t *)flags.p, (const uint32_t *)g->outdeg.p,
            (uint32_t *)list2.p, (unsigned long long *)c->counters.p);
        GG_HIP(hipGetLastError());
        if ((rc = counters_read(c, cn))) { pool_give(c, flags); pool_give(c, list2); return bail(rc); }
        std::swap(out->list, list2);
        pool_give(c, flags);
        pool_give(c, list2);
        out->d_m = cn[CNT_CURSOR];
        out->out_edges = cn[CNT_OUT];
    }
    if (out->d_m == 0) out->out_edges = 0; // ligra-numa.h:1013-1014
    out->has_dense = false;
    out->sparse_valid = true;
    *pout = out;
    return GG_OK;
}

int vertexmap_core(Graph *g, Frontier *f, const gg_vop_t *op)
{
    Ctx *c = g->ctx;
    int mode;
    uint64_t count;
    const uint8_t *F = nullptr;
    const uint32_t *list = nullptr;
    if (!f || f->bit) {
        mode = 0;
        count = g->n;
    } else if (f->n == 0) {
        return GG_OK; // empty() frontier
    } else if (f->has_dense) {
        mode = 1;
        count = g->n;
        GG_TRY(frontier_make_dense(f));
        F = (const uint8_t *)f->bytes.p;
    } else {
        mode = 2;
        count = f->d_m;
        list = (const uint32_t *)f->list.p;
    }
    if (!count) return GG_OK;
    uint32_t grid = (uint32_t)std::min<uint64_t>((count + kBlock - 1) / kBlock, 65536);
    if (op->dtype == GG_F32 && op->kind != GG_VOP_CC_COPY)
        k_vmap<float><<<grid, kBlock, 0, c->stream>>>(op->kind, mode, F, list, count, op->a, op->b, op->c0,
                                                       op->c1);
    else
        k_vmap<double><<<grid, kBlock, 0, c->stream>>>(op->kind, mode, F, list, count, op->a, op->b, op->c0,
                                                        op->c1);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

int reduce_sum_core(Ctx *c, const void *dev, int dtype, uint64_t count, double *dev_out)
{
    const uint32_t np = 1024;
    Buf part;
    GG_TRY(pool_take(c, part, 8ull * np));
    if (dtype == GG_F32)
        k_sum_part<float><<<np, kBlock, 0, c->stream>>>((const float *)dev, count, (double *)part.p);
    else
        k_sum_part<double><<<np, kBlock, 0, c->stream>>>((const double *)dev, count, (double *)part.p);
    k_sum_final<<<1, kBlock, 0, c->stream>>>((const double *)part.p, np, dev_out);
    GG_HIP(hipGetLastError());
    GG_TRY(sync(c));
    pool_give(c, part);
    return GG_OK;
}

// ============================================================================
// BFS.C's loop (app/BFS.C:77-101) with every edgeMap direction choice
// (ligra-numa.h:929-957) made on the device.  The host enqueues levels in
// batches without reading anything back; each level's kernels run only if
// the level's action word selects them, and one read-back per batch tells
// whether the frontier emptied.  Same branches and levels as the edgeMap
// path.  The visited set is a bitmap (2 MiB at 2^24 vertices, held in L2):
// a sparse level pushes and claims each new vertex with an atomic OR on its
// bit (the first claimant is its parent, as BFS_F::updateAtomic's CAS on
// Parents, app/BFS.C:36-40); a dense level pulls bottom-up, each unvisited
// vertex taking its first in-neighbour in the frontier (BFS_F::update and
// cond, app/BFS.C:31-47).  Frontiers are lists after a push and bitmaps
// after a pull.
// ============================================================================

namespace {
// state words (uint64)
enum {
    BD_ACT = 0,     // this level's actions (BA_* bits)
    BD_MODE = 1,    // this level's branch: 0 sparse, 1 dense, 2 empty, 3 done
    BD_DM = 2,      // frontier size entering this level
    BD_OUT = 3,     // its out-degree sum
    BD_LEVELS = 4,  // levels run
    BD_SCANNED = 5, // edges read
    BD_PATH = 8     // GG_PATH_* of level L at BD_PATH + L - 1 (first 64 levels)
};
constexpr int kBdWords = BD_PATH + 64;
enum : unsigned long long {
    BA_BM_LIST = 1, // dense input from a list: set its bits in the (cleared) input bitmap
    BA_PACK = 4,    // sparse input from a bitmap: pack it into a list
    BA_PUSH = 8,    // sparse: scan the degrees and push
    BA_PULL = 16,   // dense: bottom-up pull
    BA_UNMAP = 32   // the search has ended: parents and levels into the caller's numbering
};
enum { BM_SPARSE = 0, BM_DENSE = 1, BM_EMPTY = 2, BM_DONE = 3 };
constexpr uint32_t kNoVertex = 0xffffffffu;

// Arrays of one search.
struct BfsdArgs {
    const uint32_t *off, *ids;   // out-edge CSR (push)
    const uint32_t *toff, *tids; // in-edge twin (pull)
    const uint32_t *first;       // first entry of each in-list, kNoVertex if empty
    const uint32_t *outdeg;
    uint32_t *vis;               // visited bitmap
    int32_t *parents, *levels;   // levels nullable
    unsigned long long *st, *counters;
    uint32_t n, nwords;
    // with a renumbered plan: the caller's outputs and the permutation
    // (caller id -> search id) and its inverse; NULL otherwise
    int32_t *out_parents, *out_levels;
    const uint32_t *perm, *inv;
};

// Start state (app/BFS.C:86-92): Parents = -1 but the root's, the visited
// bitmap and the first input bitmap hold the root alone, the root's list.
// Each word has one writer.
// Level 1 runs no conversion: the start also writes the root list's degree
// offsets and clears level 1's output bitmap.
__global__ void k_bfsd_start(BfsdArgs A, uint32_t root, uint32_t *__restrict__ list, uint32_t *__restrict__ fbm0,
                             uint32_t *__restrict__ fbm1, uint32_t *__restrict__ offs)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kBdWords)
        A.st[i] = i == BD_DM ? 1ull : i == BD_OUT ? (unsigned long long)A.outdeg[root] : 0ull;
    if (i == 0) {
        list[0] = root;
        offs[0] = 0;
        offs[1] = A.outdeg[root];
    }
    if (i < A.nwords) {
        const uint32_t b = (root >> 5) == i ? 1u << (root & 31) : 0u;
        A.vis[i] = b;
        fbm0[i] = b;
        fbm1[i] = 0;
    }
    if (i < A.n) {
        const bool r = i == root;
        A.parents[i] = r ? (int32_t)root : -1;
        if (A.levels) A.levels[i] = r ? 0 : -1;
    }
}

// Level L's decision from the previous level's output (one wave): the
// output statistics from the counter stripes (cleared here), the branch
// (ligra-numa.h:929-957: empty when no out-edges, dense when d_m + out-
// degree > threshold, COO label when also > m/2 on edge partitions), and
// the conversion the input needs (the previous output is a list after a
// push, a bitmap after a pull).
__global__ void k_bfsd_begin(unsigned long long *__restrict__ st, unsigned long long *__restrict__ counters,
                             int level, int64_t thr, int64_t dense_thr, int coo_ok, int unmap,
                             unsigned long long dense_scan)
{
    const int lane = threadIdx.x;
    unsigned long long a = 0, o = 0, sc = 0;
    if (lane < kStripes) {
        unsigned long long *p = counters + kCntStripeBase + lane * kStripeStride;
        a = p[CNT_ACTIVE];
        o = p[CNT_OUT];
        sc = p[CNT_SCANNED];
        p[CNT_ACTIVE] = 0;
        p[CNT_OUT] = 0;
        p[CNT_SCANNED] = 0;
    }
    a = wave_sum_u64(a);
    o = wave_sum_u64(o);
    sc = wave_sum_u64(sc);
    if (lane != 0) return;
    const unsigned long long prev = st[BD_MODE];
    if (prev == BM_DONE) {
        st[BD_ACT] = 0;
        return;
    }
    if (level > 1) {
        st[BD_DM] = a;
        st[BD_OUT] = prev == BM_EMPTY ? 0 : o;
        st[BD_SCANNED] += sc;
    }
    counters[CNT_CURSOR] = 0;
    const unsigned long long dm = st[BD_DM], out = st[BD_OUT];
    if (dm == 0 || (level > 1 && prev == BM_EMPTY)) {
        // the search has ended: this level's pull launch maps the results
        // back to the caller's numbering (once)
        st[BD_MODE] = BM_DONE;
        st[BD_ACT] = unmap ? BA_UNMAP : 0ull;
        return;
    }
    st[BD_LEVELS] = (unsigned long long)level;
    int path;
    unsigned long long act = 0, mode;
    if (out == 0) {
        mode = BM_EMPTY; // partitioned_vertices::empty() (ligra-numa.h:947)
        path = GG_PATH_EMPTY;
    } else if ((int64_t)(dm + out) > thr) {
        mode = BM_DENSE;
        path = ((int64_t)(dm + out) > dense_thr && coo_ok) ? GG_PATH_COO : GG_PATH_CSC;
        // the first level's input bitmap is set by the start kernel
        act = BA_PULL | ((level > 1 && prev == BM_SPARSE) ? BA_BM_LIST : 0ull);
        st[BD_SCANNED] += dense_scan; // a dense pass that does not count its own edges
    } else {
        mode = BM_SPARSE;
        path = GG_PATH_SPARSE;
        act = BA_PUSH | (prev == BM_DENSE ? BA_PACK : 0ull);
        st[BD_SCANNED] += out;
    }
    st[BD_MODE] = mode;
    st[BD_ACT] = act;
    if (level <= 64) st[BD_PATH + level - 1] = (unsigned long long)(long long)path;
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v, int lane)
{
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// exclusive prefix of a kBlock-thread block's values (and the block total)
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t *total)
{
    __shared__ uint32_t s_w[kBlock / 64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t inc = wave_incl_scan_u32(v, lane);
    if (lane == 63) s_w[wv] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int k = 0; k < kBlock / 64; k++) {
        if (k < wv) before += s_w[k];
        all += s_w[k];
    }
    __syncthreads(); // s_w is written again by the next call
    *total = all;
    return before + inc - v;
}

// ---------------------------------------------------------------------------
// A sparse level's input list and the push's degree offsets, in three
// launches gated on BA_PUSH (k_fl_parts, k_fl_scan, k_fl_write).  The input
// is the previous level's list (after a push) or its flags (after a dense
// level, BA_PACK: a bitmap for BFS, a byte per vertex for Components), from
// which the list is written in vertex order.  Parts: kFlParts slices of a
// list, or kBlock flag groups (one a thread) of the flags.  offs[i] is the
// first out-edge slot of list entry i and offs[d_m] the total.
// ---------------------------------------------------------------------------
constexpr uint32_t kFlParts = 1024;

__device__ __forceinline__ uint64_t fl_chunk(uint64_t dm)
{
    const uint64_t per = (dm + kFlParts - 1) / kFlParts;
    return (per + kBlock - 1) / kBlock * kBlock;
}

// 32 vertices a group from a bitmap (bits at and past n are clear)
struct BitFlags {
    static constexpr uint32_t kPer = 32;
    const uint32_t *bm;
    uint32_t groups; // ceil(n / 32)
    __device__ __forceinline__ uint32_t mask(uint32_t g) const { return bm[g]; }
};

// 16 vertices a group from one byte each (nonzero: set); the array is
// padded to whole groups and the bytes at and past n are ignored
struct ByteFlags {
    static constexpr uint32_t kPer = 16;
    const uint8_t *f;
    uint32_t n, groups; // ceil(n / 16)
    __device__ __forceinline__ uint32_t mask(uint32_t g) const
    {
        const uint4 q = *(const uint4 *)(f + 16ull * g);
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            // bit 7 of each byte: that byte is nonzero
            const uint32_t y = (((w[j] & 0x7f7f7f7fu) + 0x7f7f7f7fu) | w[j]) & 0x80808080u;
            m |= (((y >> 7) & 1u) | ((y >> 14) & 2u) | ((y >> 21) & 4u) | ((y >> 28) & 8u)) << (4 * j);
        }
        const uint32_t have = n - 16u * g; // > 0 for g < groups
        return have < 16 ? m & ((1u << have) - 1u) : m;
    }
};

// Out-degrees of a flag group's vertices base .. base + kPer - 1 (one
// 64- or 128-byte line, read whole; the last group element by element up to
// n), all zero for an empty group.
template <uint32_t kPer>
__device__ __forceinline__ void group_degs(uint32_t m, uint32_t base, uint32_t n,
                                           const uint32_t *__restrict__ outdeg, uint32_t (&deg)[kPer])
{
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) deg[j] = 0;
    if (!m) return;
    if (base + kPer <= n) {
        const uint4 *p = (const uint4 *)(outdeg + base); // base is a multiple of kPer
#pragma unroll
        for (uint32_t q = 0; q < kPer / 4; q++) {
            const uint4 x = p[q];
            deg[4 * q] = x.x;
            deg[4 * q + 1] = x.y;
            deg[4 * q + 2] = x.z;
            deg[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kPer; j++)
            if (base + j < n) deg[j] = outdeg[base + j];
    }
}

// sum of deg[j] over the set bits j of m
template <uint32_t kPer>
__device__ __forceinline__ uint32_t masked_sum(uint32_t m, const uint32_t (&deg)[kPer])
{
    uint32_t s = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) s += ((m >> j) & 1u) ? deg[j] : 0u;
    return s;
}

// The block's sums of c and d into pc[blockIdx.x], pd[blockIdx.x].
__device__ __forceinline__ void block_sums2(uint32_t c, uint32_t d, uint32_t *__restrict__ pc,
                                            uint32_t *__restrict__ pd)
{
    __shared__ unsigned long long s_c[kBlock / 64], s_d[kBlock / 64];
    const unsigned long long wc = wave_sum_u64(c), wd = wave_sum_u64(d);
    if ((threadIdx.x & 63) == 0) {
        s_c[threadIdx.x >> 6] = wc;
        s_d[threadIdx.x >> 6] = wd;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tc = 0, td = 0;
        for (int k = 0; k < kBlock / 64; k++) {
            tc += s_c[k];
            td += s_d[k];
        }
        pc[blockIdx.x] = (uint32_t)tc;
        pd[blockIdx.x] = (uint32_t)td;
    }
}

// Stage 1.  Besides, outside BA_PUSH: with BA_UNMAP, out[v] = cp_src[perm[v]]
// (Components' labels back into the caller's order); with BA_BM_LIST (a
// dense level after a push), the list's bits into 'fbm' (BFS's input
// bitmap, cleared by that push's level) and prevIDs copied over it (cp_*,
// Components).  With BA_PUSH: 'fclr' cleared (the bitmap a dense next level
// takes its input in), and per part the frontier count (flags) and the
// out-degree sum; over a list, prevIDs copied as well (cp_*).
template <class Fl>
__global__ __launch_bounds__(kBlock) void k_fl_parts(Fl fl, const uint32_t *__restrict__ list,
                                                     const uint32_t *__restrict__ outdeg,
                                                     const unsigned long long *__restrict__ st,
                                                     uint32_t *__restrict__ pcnt, uint32_t *__restrict__ pdeg,
                                                     uint32_t *__restrict__ fbm, uint32_t *__restrict__ fclr,
                                                     uint32_t nwords, const uint32_t *__restrict__ cp_src,
                                                     uint32_t *__restrict__ cp_dst,
                                                     const uint32_t *__restrict__ perm, uint32_t *__restrict__ out,
                                                     uint32_t n)
{
    const unsigned long long act = st[BD_ACT];
    const uint64_t gid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    if (act & BA_UNMAP) {
        if (!out) return;
        // 8 independent gathers a thread in flight
        constexpr int kU = 8;
        for (uint64_t v0 = gid; v0 < n; v0 += stride * kU) {
            uint32_t p[kU], x[kU];
#pragma unroll
            for (int k = 0; k < kU; k++) {
                const uint64_t v = v0 + k * stride;
                p[k] = v < n ? perm[v] : 0u;
            }
#pragma unroll
            for (int k = 0; k < kU; k++) x[k] = cp_src[p[k]];
#pragma unroll
            for (int k = 0; k < kU; k++) {
                const uint64_t v = v0 + k * stride;
                if (v < n) out[v] = x[k];
            }
        }
        return;
    }
    if (act & BA_BM_LIST) {
        const uint64_t k = st[BD_DM];
        for (uint64_t i = gid; i < k; i += stride) {
            const uint32_t v = list[i];
            if (fbm) atomicOr(&fbm[v >> 5], 1u << (v & 31));
            if (cp_dst) cp_dst[v] = cp_src[v];
        }
        return;
    }
    if (!(act & BA_PUSH)) return;
    if (fclr)
        for (uint64_t w = gid; w < nwords; w += stride) fclr[w] = 0;
    uint32_t c = 0, d = 0;
    if (act & BA_PACK) {
        if ((uint64_t)blockIdx.x * kBlock >= fl.groups) return; // the whole work-group
        if (gid < fl.groups) {
            const uint32_t m = fl.mask((uint32_t)gid);
            uint32_t deg[Fl::kPer];
            group_degs<Fl::kPer>(m, (uint32_t)gid * Fl::kPer, n, outdeg, deg);
            c = __popc(m);
            d = masked_sum<Fl::kPer>(m, deg);
        }
    } else {
        if (blockIdx.x >= kFlParts) return; // the whole work-group
        const uint64_t dm = st[BD_DM], ch = fl_chunk(dm);
        const uint64_t a = (uint64_t)blockIdx.x * ch, b = min(a + ch, dm);
        for (uint64_t i = a + threadIdx.x; i < b; i += kBlock) {
            const uint32_t v = list[i];
            d += outdeg[v];
            if (cp_dst) cp_dst[v] = cp_src[v];
        }
    }
    block_sums2(c, d, pcnt, pdeg);
}

// Stage 2 (one work-group): exclusive scans of the parts' counts and
// degree sums in place; offs[d_m] gets the total out-degree.
__global__ __launch_bounds__(kBlock) void k_fl_scan(uint32_t *__restrict__ pcnt, uint32_t *__restrict__ pdeg,
                                                    const unsigned long long *__restrict__ st, uint32_t flag_parts,
                                                    uint32_t *__restrict__ offs)
{
    const unsigned long long act = st[BD_ACT];
    if (!(act & BA_PUSH)) return;
    const uint32_t parts = (act & BA_PACK) ? flag_parts : kFlParts;
    uint32_t cc = 0, cd = 0;
    for (uint32_t b = 0; b < parts; b += kBlock) { // block-uniform bounds
        const uint32_t i = b + threadIdx.x;
        const uint32_t c = i < parts ? pcnt[i] : 0u, d = i < parts ? pdeg[i] : 0u;
        uint32_t all_c, all_d;
        const uint32_t ec = block_excl_scan_u32(c, &all_c);
        const uint32_t ed = block_excl_scan_u32(d, &all_d);
        if (i < parts) {
            pcnt[i] = cc + ec;
            pdeg[i] = cd + ed;
        }
        cc += all_c;
        cd += all_d;
    }
    if (threadIdx.x == 0) offs[st[BD_DM]] = cd;
}

// Stage 3: the list (from flags) and the offsets.
template <class Fl>
__global__ __launch_bounds__(kBlock) void k_fl_write(Fl fl, uint32_t *__restrict__ list,
                                                     const uint32_t *__restrict__ outdeg,
                                                     const unsigned long long *__restrict__ st,
                                                     const uint32_t *__restrict__ pcnt,
                                                     const uint32_t *__restrict__ pdeg, uint32_t *__restrict__ offs,
                                                     uint32_t n)
{
    constexpr uint32_t kStage = 2048; // entries staged per pass
    __shared__ uint32_t s_v[kStage], s_o[kStage];
    const unsigned long long act = st[BD_ACT];
    if (!(act & BA_PUSH)) return;
    if (act & BA_PACK) {
        if ((uint64_t)blockIdx.x * kBlock >= fl.groups) return; // the whole work-group
        const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
        const uint32_t base = g * Fl::kPer;
        const uint32_t m = g < fl.groups ? fl.mask(g) : 0u;
        uint32_t deg[Fl::kPer];
        group_degs<Fl::kPer>(m, base, n, outdeg, deg);
        uint32_t total, dall;
        const uint32_t li0 = block_excl_scan_u32(__popc(m), &total);
        const uint32_t dp0 = pdeg[blockIdx.x] + block_excl_scan_u32(masked_sum<Fl::kPer>(m, deg), &dall);
        const uint32_t out0 = pcnt[blockIdx.x];
        // the work-group's entries are one run of the list: they are placed
        // in LDS kStage at a time and written out coalesced
        for (uint32_t p0 = 0; p0 < total; p0 += kStage) { // block-uniform bounds
            uint32_t li = li0, dp = dp0;
#pragma unroll
            for (uint32_t j = 0; j < Fl::kPer; j++)
                if ((m >> j) & 1u) {
                    if (li >= p0 && li < p0 + kStage) {
                        s_v[li - p0] = base + j;
                        s_o[li - p0] = dp;
                    }
                    li++;
                    dp += deg[j];
                }
            __syncthreads();
            const uint32_t cnt = min(kStage, total - p0);
            for (uint32_t i = threadIdx.x; i < cnt; i += kBlock) {
                list[out0 + p0 + i] = s_v[i];
                offs[out0 + p0 + i] = s_o[i];
            }
            __syncthreads(); // the stage is filled again by the next pass
        }
        return;
    }
    if (blockIdx.x >= kFlParts) return; // the whole work-group
    const uint64_t dm = st[BD_DM], ch = fl_chunk(dm);
    const uint64_t a = (uint64_t)blockIdx.x * ch, b = min(a + ch, dm);
    uint32_t carry = pdeg[blockIdx.x];
    for (uint64_t t0 = a; t0 < b; t0 += kBlock) { // block-uniform bounds
        const uint64_t i = t0 + threadIdx.x;
        const uint32_t d = i < b ? outdeg[list[i]] : 0u;
        uint32_t all;
        const uint32_t ex = block_excl_scan_u32(d, &all);
        if (i < b) offs[i] = carry + ex;
        carry += all;
    }
}

// Launch size of the three stages: whichever is larger of the flag parts
// and the list parts.
inline uint32_t fl_grid(uint32_t groups)
{
    return std::max<uint32_t>((groups + kBlock - 1) / kBlock, kFlParts);
}

// Sparse level's update of edge (v -> ngh) for BFS.C (BFS_F::updateAtomic,
// app/BFS.C:49-56): the head is claimed through its visited bit (read
// first: a stale read only sends the edge on to the atomic, which decides)
// and gets its parent and level; returns whether ngh joins the output.
struct BfsVisit {
    uint32_t *vis;
    int32_t *parents, *levels;
    int32_t level;
    __device__ __forceinline__ bool operator()(uint32_t v, uint32_t ngh) const
    {
        const uint32_t w = ngh >> 5, bit = 1u << (ngh & 31);
        if (!(vis[w] & bit) && !(atomicOr(&vis[w], bit) & bit)) {
            parents[ngh] = (int32_t)v;
            if (levels) levels[ngh] = level;
            return true;
        }
        return false;
    }
};

// Sparse level: the frontier's out-edges in kPushTile tiles (grid-stride),
// each edge updated by U; the heads it returns true for are appended to the
// output list, one cursor add per tile (a block scan places them).
template <class U>
__global__ __launch_bounds__(kBlock) void k_dpush(U upd, const uint32_t *__restrict__ off,
                                                  const uint32_t *__restrict__ ids,
                                                  const uint32_t *__restrict__ outdeg,
                                                  const unsigned long long *__restrict__ st,
                                                  unsigned long long *__restrict__ counters,
                                                  const uint32_t *__restrict__ list,
                                                  const uint32_t *__restrict__ offs, uint32_t *__restrict__ out,
                                                  uint64_t cap)
{
    constexpr int kIt = kPushTile / kBlock;
    __shared__ uint32_t s_off[kPushLds];
    __shared__ uint32_t s_lo, s_hi;
    __shared__ unsigned long long s_base;
    __shared__ unsigned long long s_od[kBlock / 64];
    if (!(st[BD_ACT] & BA_PUSH)) return;
    const uint32_t k = (uint32_t)st[BD_DM];
    const uint64_t total = min(cap, (uint64_t)offs[k]);
    for (uint64_t t0 = (uint64_t)blockIdx.x * kPushTile; t0 < total; t0 += (uint64_t)gridDim.x * kPushTile) {
        const uint64_t t1 = min(t0 + (uint64_t)kPushTile, total);
        // the tile's first and last entries: two waves search at once
        const uint32_t wv = threadIdx.x >> 6;
        if (wv == 0) {
            const uint32_t r = wave_upper_idx(offs, 0, k, t0);
            if (threadIdx.x == 0) s_lo = r;
        } else if (wv == 1) {
            const uint32_t r = wave_upper_idx(offs, 0, k, t1 - 1) + 1;
            if (threadIdx.x == 64) s_hi = r;
        }
        __syncthreads();
        const uint32_t lo = s_lo, hi = s_hi;
        const bool in_lds = hi - lo <= (uint32_t)kPushLds;
        if (in_lds)
            for (uint32_t i = threadIdx.x; i < hi - lo; i += kBlock) s_off[i] = offs[lo + i];
        __syncthreads();
        uint32_t got[kIt];
        uint32_t c = 0;
        unsigned long long od = 0;
#pragma unroll
        for (int it = 0; it < kIt; it++) {
            const uint64_t e = t0 + (uint64_t)it * kBlock + threadIdx.x;
            got[it] = kNoVertex;
            if (e < t1) {
                const uint32_t kk = in_lds ? lo + upper_idx(s_off, 0, hi - lo, e) : upper_idx(offs, lo, hi, e);
                const uint32_t v = list[kk];
                const uint32_t ngh = ids[off[v] + (uint32_t)(e - offs[kk])];
                if (upd(v, ngh)) {
                    got[it] = ngh;
                    c++;
                    od += outdeg[ngh];
                }
            }
        }
        uint32_t all;
        uint32_t pos = block_excl_scan_u32(c, &all);
        od = wave_sum_u64(od);
        if ((threadIdx.x & 63) == 0) s_od[threadIdx.x >> 6] = od;
        if (threadIdx.x == 0 && all) s_base = atomicAdd(&counters[CNT_CURSOR], (unsigned long long)all);
        __syncthreads();
        if (threadIdx.x == 0 && all) {
            unsigned long long t = 0;
            for (int q = 0; q < kBlock / 64; q++) t += s_od[q];
            cnt_add(counters, CNT_ACTIVE, (unsigned long long)all);
            cnt_add(counters, CNT_OUT, t);
        }
        const unsigned long long base = s_base;
#pragma unroll
        for (int it = 0; it < kIt; it++)
            if (got[it] != kNoVertex) out[base + pos++] = got[it];
        __syncthreads(); // s_lo, s_hi, s_off, s_base and s_od are written again by the next tile
    }
}

// Dense level: every unvisited vertex looks for its first in-neighbour in
// the input bitmap: the in-list's first entry from a coalesced array, then
// the rest of a short list kPullStep entries at a time (independent loads,
// so a step costs two round trips whatever its length); a rest longer than
// kPullShort is walked by the whole wave, 64 entries a step, one such lane
// after another (the first hit in list order is the lowest lane of the
// first step with a hit).  Each wave owns 64 consecutive vertices, so it
// writes their two words of the visited and output bitmaps whole.
constexpr uint32_t kPullShort = 16;
constexpr int kPullStep = 8;

__global__ __launch_bounds__(kBlock) void k_bfsd_pull(BfsdArgs A, const uint32_t *__restrict__ fin,
                                                      uint32_t *__restrict__ fout, int32_t level)
{
    const unsigned long long act = A.st[BD_ACT];
    if (act & BA_UNMAP) {
        // parents[v] = inv[pp[perm[v]]] (or -1), levels[v] = lp[perm[v]]: a
        // gather over the caller's order (the reads of the plan-numbered
        // arrays are random but sit in the MALL; every write is coalesced),
        // kU vertices a thread with their loads in flight together
        constexpr int kU = 8;
        const uint32_t stride = gridDim.x * kBlock;
        for (uint32_t v0 = blockIdx.x * kBlock + threadIdx.x; v0 < A.n; v0 += stride * kU) {
            uint32_t k[kU];
            int32_t p[kU];
#pragma unroll
            for (int j = 0; j < kU; j++) {
                const uint64_t v = v0 + (uint64_t)j * stride;
                k[j] = v < A.n ? A.perm[v] : 0u;
            }
#pragma unroll
            for (int j = 0; j < kU; j++) p[j] = A.parents[k[j]];
            uint32_t q[kU];
#pragma unroll
            for (int j = 0; j < kU; j++) q[j] = A.inv[p[j] < 0 ? 0 : p[j]];
#pragma unroll
            for (int j = 0; j < kU; j++) {
                const uint64_t v = v0 + (uint64_t)j * stride;
                if (v < A.n) {
                    A.out_parents[v] = p[j] < 0 ? -1 : (int32_t)q[j];
                    if (A.out_levels) A.out_levels[v] = A.levels[k[j]];
                }
            }
        }
        return;
    }
    if (!(act & BA_PULL)) return;
    LocalStats ls;
    unsigned long long scanned = 0;
    const int lane = threadIdx.x & 63;
    for (uint32_t base = blockIdx.x * kBlock; base < A.n; base += gridDim.x * kBlock) {
        const uint32_t d = base + threadIdx.x;
        bool was = false, longl = false;
        uint32_t par = kNoVertex, b = 0, e = 0;
        if (d < A.n) {
            was = bit_test(A.vis, d);
            const uint32_t s0 = was ? kNoVertex : A.first[d];
            if (s0 != kNoVertex) {
                scanned++;
                if (bit_test(fin, s0)) {
                    par = s0;
                } else {
                    b = A.toff[d] + 1;
                    e = A.toff[d + 1];
                    if (e - b > kPullShort) {
                        longl = true;
                    } else {
                        for (uint32_t j = b; j < e && par == kNoVertex; j += kPullStep) {
                            uint32_t src[kPullStep];
#pragma unroll
                            for (int k = 0; k < kPullStep; k++) src[k] = A.tids[min(j + k, e - 1)];
                            bool hit[kPullStep];
#pragma unroll
                            for (int k = 0; k < kPullStep; k++) hit[k] = bit_test(fin, src[k]);
#pragma unroll
                            for (int k = kPullStep - 1; k >= 0; k--)
                                if (j + k < e && hit[k]) par = src[k];
                            // entries read: up to the hit, or the step
                            uint32_t used = min((uint32_t)kPullStep, e - j);
#pragma unroll
                            for (int k = kPullStep - 1; k >= 0; k--)
                                if (j + k < e && hit[k]) used = k + 1;
                            scanned += used;
                        }
                    }
                }
            }
        }
        uint64_t lm = __ballot(longl);
        while (lm) { // wave-uniform
            const int L = __builtin_ctzll(lm);
            lm &= lm - 1;
            const uint32_t lb = __shfl(b, L, 64), le = __shfl(e, L, 64);
            uint32_t found = kNoVertex, cnt = 0;
            for (uint32_t j0 = lb; j0 < le; j0 += 64) {
                const uint32_t j = j0 + lane;
                const uint32_t s = j < le ? A.tids[j] : kNoVertex;
                const bool hit = s != kNoVertex && bit_test(fin, s);
                const uint64_t hm = __ballot(hit);
                if (hm) {
                    const int p = __builtin_ctzll(hm);
                    found = __shfl(s, p, 64);
                    cnt += (uint32_t)p + 1;
                    break;
                }
                cnt += min(64u, le - j0);
            }
            if (lane == L) {
                par = found;
                scanned += cnt;
            }
        }
        const bool now = par != kNoVertex;
        if (now) {
            A.parents[d] = (int32_t)par;
            if (A.levels) A.levels[d] = level;
            ls.active++;
            ls.out_edges += A.outdeg[d];
        }
        const uint64_t nb = __ballot(now), vb = __ballot(was || now);
        const uint32_t w0 = (base + (threadIdx.x & ~63u)) >> 5;
        if (lane == 0 && w0 < A.nwords) {
            fout[w0] = (uint32_t)nb;
            A.vis[w0] = (uint32_t)vb;
        }
        if (lane == 32 && w0 + 1 < A.nwords) {
            fout[w0 + 1] = (uint32_t)(nb >> 32);
            A.vis[w0 + 1] = (uint32_t)(vb >> 32);
        }
    }
    __shared__ unsigned long long s_scan[kBlock / 64];
    const unsigned long long sc = wave_sum_u64(scanned);
    if (lane == 0) s_scan[threadIdx.x >> 6] = sc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < kBlock / 64; k++) t += s_scan[k];
        cnt_add(A.counters, CNT_SCANNED, t);
    }
    block_stats(ls, A.counters);
}

// first[d] = the first source of d's in-list, kNoVertex when it is empty
__global__ void k_first_in(const uint32_t *__restrict__ toff, const uint32_t *__restrict__ tids, uint32_t n,
                           uint32_t *__restrict__ first)
{
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const uint32_t b = toff[d];
    first[d] = b < toff[d + 1] ? tids[b] : kNoVertex;
}
} // namespace

int ensure_first_in(Graph *g)
{
    const bool planned = g->plan.mode != 0;
    Buf &first = planned ? g->plan.first_in : g->first_in;
    if (first.p) return GG_OK;
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    GG_TRY(buf_alloc(c, first, 4ull * std::max<uint32_t>(n, 1)));
    if (n)
        k_first_in<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>(
            (const uint32_t *)(planned ? g->plan.toff.p : g->toff.p),
            (const uint32_t *)(planned ? g->plan.tids.p : g->tids.p), n, (uint32_t *)first.p);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

int bfs_device_loop(Graph *g, uint32_t root, const uint32_t *first, int64_t threshold, int32_t *parents,
                    int32_t *levels, gg_run_info *info, const BfsMapBack *mb, int *level_hint)
{
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    if (threshold == -1) threshold = (int64_t)(m / 20);
    const int64_t dense_thr = (int64_t)(m / 2);
    const int coo_ok = g->P && !g->part_vertex;
    // a push emits each vertex at most once, and the scan's edges are at
    // most the threshold when the level pushes
    const uint64_t lcap = (uint64_t)n + 1;
    const uint64_t ecap = (uint64_t)std::max<int64_t>(threshold, 0);
    const uint32_t nwords = (n + 31) / 32;
    const uint32_t flag_parts = (nwords + kBlock - 1) / kBlock, flg = fl_grid(nwords);
    Buf st, list[2], fbm[2], vis, offs, pcnt, dpart;
    auto give = [&]() {
        Buf *bs[] = {&st, &list[0], &list[1], &fbm[0], &fbm[1], &vis, &offs, &pcnt, &dpart};
        for (Buf *b : bs) pool_give(c, *b);
    };
    int rc = pool_take(c, st, 8ull * kBdWords);
    if (!rc) rc = pool_take(c, list[0], 4 * lcap);
    if (!rc) rc = pool_take(c, list[1], 4 * lcap);
    if (!rc) rc = pool_take(c, fbm[0], 4ull * (nwords + 1));
    if (!rc) rc = pool_take(c, fbm[1], 4ull * (nwords + 1));
    if (!rc) rc = pool_take(c, vis, 4ull * (nwords + 1));
    if (!rc) rc = pool_take(c, offs, 4ull * ((uint64_t)n + 2));
    if (!rc) rc = pool_take(c, pcnt, 4ull * flg);
    if (!rc) rc = pool_take(c, dpart, 4ull * flg);
    if (rc) { give(); return rc; }
    unsigned long long *S = (unsigned long long *)st.p;
    BfsdArgs A{};
    A.off = (const uint32_t *)g->off.p;
    A.ids = (const uint32_t *)g->ids.p;
    A.toff = (const uint32_t *)g->toff.p;
    A.tids = (const uint32_t *)g->tids.p;
    A.first = first;
    A.outdeg = (const uint32_t *)g->outdeg.p;
    A.vis = (uint32_t *)vis.p;
    A.parents = parents;
    A.levels = levels;
    A.st = S;
    A.counters = (unsigned long long *)c->counters.p;
    A.n = n;
    A.nwords = nwords;
    if (mb) {
        A.out_parents = mb->parents;
        A.out_levels = mb->levels;
        A.perm = mb->perm;
        A.inv = mb->inv;
    }
    if ((rc = counters_zero(c))) { give(); return rc; }
    k_bfsd_start<<<blocks_for(std::max<uint32_t>(n, kBdWords), kBlock), kBlock, 0, c->stream>>>(
        A, root, (uint32_t *)list[0].p, (uint32_t *)fbm[0].p, (uint32_t *)fbm[1].p, (uint32_t *)offs.p);
    const uint32_t push_grid = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((ecap + kPushTile - 1) / kPushTile, 1),
                                                            4096);
    // the pull walks the vertices grid-stride: a few waves of work-groups
    // per CU, so a level that skips it (or only maps back) costs little
    int cus = 0;
    GG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    const uint32_t pull_grid = std::max<uint32_t>(std::min<uint32_t>(blocks_for(n, kBlock), (uint32_t)cus * 16u), 1);
    int level = 0;
    bool done = false;
    // the first batch: the last search's depth on this graph plus the level
    // that finds the frontier empty, else 8; then 16, 32, 64
    int batch = level_hint && *level_hint > 0 ? std::min(*level_hint + 1, 64) : 8;
    unsigned long long h[kBdWords];
    while (!done && !rc) {
        for (int b = 0; b < batch && !rc; b++) {
            level++;
            // level L reads list[(L+1)&1] / fbm[(L+1)&1] (what level L-1
            // wrote; the root's for L = 1) and writes list[L&1] / fbm[L&1]
            const int in = (level + 1) & 1, ou = level & 1;
            uint32_t *lin = (uint32_t *)list[in].p, *lout = (uint32_t *)list[ou].p;
            uint32_t *fin = (uint32_t *)fbm[in].p, *fout = (uint32_t *)fbm[ou].p;
            k_bfsd_begin<<<1, 64, 0, c->stream>>>(S, A.counters, level, threshold, dense_thr, coo_ok, mb != nullptr,
                                                  0ull);
            if (level > 1) {
                // dense input from the previous push's list; sparse input
                // (from the previous pull's bitmap) and its degree offsets,
                // fout cleared for a dense next level
                const BitFlags fl{fin, nwords};
                k_fl_parts<<<flg, kBlock, 0, c->stream>>>(fl, lin, A.outdeg, S, (uint32_t *)pcnt.p,
                                                          (uint32_t *)dpart.p, fin, fout, nwords, nullptr, nullptr,
                                                          nullptr, nullptr, n);
                k_fl_scan<<<1, kBlock, 0, c->stream>>>((uint32_t *)pcnt.p, (uint32_t *)dpart.p, S, flag_parts,
                                                       (uint32_t *)offs.p);
                k_fl_write<<<flg, kBlock, 0, c->stream>>>(fl, lin, A.outdeg, S, (const uint32_t *)pcnt.p,
                                                          (const uint32_t *)dpart.p, (uint32_t *)offs.p, n);
            }
            k_dpush<<<push_grid, kBlock, 0, c->stream>>>(BfsVisit{A.vis, A.parents, A.levels, level}, A.off, A.ids,
                                                         A.outdeg, S, A.counters, lin, (const uint32_t *)offs.p,
                                                         lout, ecap);
            // dense: bottom-up pull
            k_bfsd_pull<<<pull_grid, kBlock, 0, c->stream>>>(A, fin, fout, level);
            if (hipGetLastError() != hipSuccess) rc = fail(GG_ERR_HIP, "gg_bfs: level kernels");
        }
        if (rc) break;
        static_assert(sizeof h <= kPinBytes, "the state words fit the pinned read-back buffer");
        if ((rc = read_back(c, h, S, sizeof h))) break;
        done = h[BD_MODE] == BM_DONE;
        batch = std::min(std::max(batch, 8) * 2, 64);
        // every level that runs reaches a new vertex, so a search still
        // going after n + 2 enqueued levels has gone wrong
        if (!done && level > (int)n + 2) rc = fail(GG_ERR_STATE, "gg_bfs: more levels than vertices");
    }
    if (!rc) {
        if (level_hint) *level_hint = (int)h[BD_LEVELS];
        if (info) {
            info->iters = (int)h[BD_LEVELS];
            info->edges_scanned = h[BD_SCANNED];
            for (int L = 1; L <= info->iters && L <= 64; L++) info->paths[L - 1] = (int)(long long)h[BD_PATH + L - 1];
        }
    }
    give();
    return rc;
}

// ============================================================================
// Components.C's Compute() on the device (app/Components.C:101-126).  Each
// round's edgeMap branch is chosen by k_bfsd_begin from the previous round's
// striped counters, as for BFS, with no read-back per round.  A dense round
// (k_ccd_pull) gives each row the minimum over its in-neighbours' current
// labels: a vertex outside the frontier holds a label its neighbours
// already took, so reading it changes nothing, and labels only decrease, so
// a label read while its owner rewrites it is the old or the new one, both
// upper bounds of the component's least.  A sparse round pushes the
// frontier's round-start labels with atomicMin (CC_F::updateAtomic,
// app/Components.C:37-46); the first lowering of a vertex in the round
// appends it to the next frontier.  prevIDs holds the round-start labels.
// CC_Vertex_F's copy over the frontier (app/Components.C:88-99) runs in
// three places: in the degree scan of a sparse round, in a list pass before
// a dense round that follows a sparse one, and in the dense round's own row
// writes.  The dense round copies every row it lowers, so after it prevIDs
// equals IDs everywhere.
// ============================================================================
namespace {
// CC_F::updateAtomic on edge (v -> ngh): the source's round-start label
// into the head with atomicMin; the lowering that finds the head still at
// its round-start label is its first in the round.
struct CcVisit {
    uint32_t *ids;
    const uint32_t *prev;
    __device__ __forceinline__ bool operator()(uint32_t v, uint32_t ngh) const
    {
        const uint32_t val = prev[v];
        if (val < ids[ngh]) {
            const uint32_t old = atomicMin(&ids[ngh], val);
            return val < old && old == prev[ngh];
        }
        return false;
    }
};

// Dense round, bottom-up: each row not already at label 0 takes the
// minimum over its in-neighbours' current labels, walking its in-list in
// order and stopping at label 0, the least there is (the minimum is then
// exact without the rest).  The in-list's first entry comes from a
// coalesced array; a short rest is read kPullStep entries a step, a rest over
// kPullShort entries by the whole wave, 64 a step.  With a renumbered plan
// the lists start with their hottest sources, and those rows run first in
// the grid-stride order, so most rows stop at their first entry.  A
// row's owner writes its label (and prevIDs) and flag; the round's
// statistics and scanned entries go to the counter stripes.
__global__ __launch_bounds__(kBlock) void k_ccd_pull(const uint32_t *__restrict__ first,
                                                     const uint32_t *__restrict__ toff,
                                                     const uint32_t *__restrict__ tids, uint32_t *__restrict__ ids,
                                                     uint32_t *__restrict__ prev, uint8_t *__restrict__ next,
                                                     const uint32_t *__restrict__ outdeg,
                                                     unsigned long long *__restrict__ counters,
                                                     const unsigned long long *__restrict__ st, uint32_t n)
{
    if (!(st[BD_ACT] & BA_PULL)) return;
    LocalStats ls;
    unsigned long long scanned = 0;
    const int lane = threadIdx.x & 63;
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const uint32_t d = base + threadIdx.x;
        uint32_t cur = 0, acc = 0xffffffffu, b = 0, e = 0;
        bool longl = false;
        if (d < n) {
            cur = ids[d];
            const uint32_t s0 = cur ? first[d] : kNoVertex;
            if (s0 != kNoVertex) {
                scanned++;
                acc = ids[s0];
                if (acc != 0) {
                    b = toff[d] + 1;
                    e = toff[d + 1];
                    if (e - b > kPullShort) {
                        longl = true;
                    } else {
                        for (uint32_t j = b; j < e && acc != 0; j += kPullStep) {
                            uint32_t src[kPullStep];
#pragma unroll
                            for (int k = 0; k < kPullStep; k++) src[k] = tids[min(j + k, e - 1)];
                            uint32_t lab[kPullStep];
#pragma unroll
                            for (int k = 0; k < kPullStep; k++) lab[k] = ids[src[k]];
#pragma unroll
                            for (int k = 0; k < kPullStep; k++) acc = min(acc, lab[k]); // clamped repeats are harmless
                            scanned += min((uint32_t)kPullStep, e - j);
                        }
                    }
                }
            }
        }
        uint64_t lm = __ballot(longl);
        while (lm) { // wave-uniform
            const int L = __builtin_ctzll(lm);
            lm &= lm - 1;
            const uint32_t lb = __shfl(b, L, 64), le = __shfl(e, L, 64);
            uint32_t wmin = 0xffffffffu, cnt = 0;
            for (uint32_t j0 = lb; j0 < le; j0 += 64) {
                const uint32_t j = j0 + lane;
                const uint32_t l = j < le ? ids[tids[j]] : 0xffffffffu;
                wmin = min(wmin, l);
                cnt += min(64u, le - j0);
                if (__ballot(l == 0)) break;
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) wmin = min(wmin, (uint32_t)__shfl_xor((int)wmin, off, 64));
            if (lane == L) {
                acc = min(acc, wmin);
                scanned += cnt;
            }
        }
        if (d < n) {
            const bool dec = acc < cur;
            if (dec) {
                ids[d] = acc;
                prev[d] = acc;
                ls.active++;
                ls.out_edges += outdeg[d];
            }
            next[d] = dec;
        }
    }
    __shared__ unsigned long long s_scan[kBlock / 64];
    const unsigned long long sc = wave_sum_u64(scanned);
    if (lane == 0) s_scan[threadIdx.x >> 6] = sc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < kBlock / 64; k++) t += s_scan[k];
        cnt_add(counters, CNT_SCANNED, t);
    }
    block_stats(ls, counters);
}

// Start state (app/Components.C:110-116): IDs = prevIDs = the vertex ids
// ('init', the caller's ids of a renumbered plan, else the position), the
// frontier all vertices with d_m = n and m out-edges.
__global__ void k_ccd_start(unsigned long long *__restrict__ st, uint32_t n, uint64_t m,
                            const uint32_t *__restrict__ init, uint32_t *__restrict__ ids,
                            uint32_t *__restrict__ prev)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kBdWords) st[i] = i == BD_DM ? (unsigned long long)n : i == BD_OUT ? (unsigned long long)m : 0ull;
    if (i < n) {
        const uint32_t l = init ? init[i] : i;
        ids[i] = l;
        prev[i] = l;
    }
}

} // namespace

int cc_device_loop(Graph *g, const uint32_t *first, const uint32_t *init, uint32_t *labels, uint32_t *prev,
                   gg_run_info *info, int *round_hint, const uint32_t *perm, uint32_t *out)
{
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    if (!n) return GG_OK;
    // Components.C:123 passes m/20; the first round (all n vertices, m
    // out-edges) is then dense, or empty without edges, never sparse
    const int64_t threshold = (int64_t)(m / 20);
    const int64_t dense_thr = (int64_t)(m / 2);
    const int coo_ok = g->P && !g->part_vertex;
    const uint64_t lcap = (uint64_t)n + 1;
    const uint64_t ecap = (uint64_t)std::max<int64_t>(threshold, 0);
    const uint32_t groups = (n + 15) / 16; // 16-flag groups of the dense rounds' output
    const uint32_t flag_parts = (groups + kBlock - 1) / kBlock, flg = fl_grid(groups);
    Buf st, list[2], flags, offs, pcnt, dpart;
    auto give = [&]() {
        Buf *bs[] = {&st, &list[0], &list[1], &flags, &offs, &pcnt, &dpart};
        for (Buf *b : bs) pool_give(c, *b);
    };
    int rc = pool_take(c, st, 8ull * kBdWords);
    if (!rc) rc = pool_take(c, list[0], 4 * lcap);
    if (!rc) rc = pool_take(c, list[1], 4 * lcap);
    if (!rc) rc = pool_take(c, flags, 16ull * groups); // whole 16-byte groups
    if (!rc) rc = pool_take(c, offs, 4ull * ((uint64_t)n + 2));
    if (!rc) rc = pool_take(c, pcnt, 4ull * flg);
    if (!rc) rc = pool_take(c, dpart, 4ull * flg);
    if (rc) { give(); return rc; }
    unsigned long long *S = (unsigned long long *)st.p;
    unsigned long long *cnt = (unsigned long long *)c->counters.p;
    const uint32_t *outdeg = (const uint32_t *)g->outdeg.p;
    uint8_t *next = (uint8_t *)flags.p;
    if ((rc = counters_zero(c))) { give(); return rc; }
    k_ccd_start<<<blocks_for(std::max<uint32_t>(n, kBdWords), kBlock), kBlock, 0, c->stream>>>(S, n, m, init,
                                                                                                  labels, prev);
    const uint32_t push_grid = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((ecap + kPushTile - 1) / kPushTile, 1),
                                                            4096);
    int cus = 0;
    GG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    const uint32_t pull_grid = std::max<uint32_t>(std::min<uint32_t>(blocks_for(n, kBlock), (uint32_t)cus * 16u), 1);
    const uint32_t *toff = (const uint32_t *)g->toff.p, *tids = (const uint32_t *)g->tids.p;
    int level = 0;
    bool done = false;
    // the first batch: the last call's rounds on this graph plus the round
    // that finds the frontier empty, else 8; then 16, 32, 64
    int batch = round_hint && *round_hint > 0 ? std::min(*round_hint + 1, 64) : 8;
    const int unmap = out != nullptr;
    unsigned long long h[kBdWords];
    while (!done && !rc) {
        for (int b = 0; b < batch && !rc; b++) {
            level++;
            k_bfsd_begin<<<1, 64, 0, c->stream>>>(S, cnt, level, threshold, dense_thr, coo_ok, unmap, 0ull);
            if (level == 1) {
                // the whole vertex set with its m out-edges: dense (or empty
                // without edges)
                k_ccd_pull<<<pull_grid, kBlock, 0, c->stream>>>(first, toff, tids, labels, prev, next, outdeg, cnt, S,
                                                                n);
                if (hipGetLastError() != hipSuccess) rc = fail(GG_ERR_HIP, "gg_cc: round kernels");
                continue;
            }
            // round L reads list[(L+1)&1] (what round L-1 pushed, or the
            // pack of its flags) and pushes into list[L&1]
            uint32_t *lin = (uint32_t *)list[(level + 1) & 1].p, *lout = (uint32_t *)list[level & 1].p;
            // dense after sparse: prevIDs of the pushed list; after the last
            // round: the labels into the caller's order; sparse: the input
            // list (from the dense round's flags after a dense round, with
            // prevIDs copied over a pushed one) and its degree offsets
            const ByteFlags fl{next, n, groups};
            k_fl_parts<<<flg, kBlock, 0, c->stream>>>(fl, lin, outdeg, S, (uint32_t *)pcnt.p, (uint32_t *)dpart.p,
                                                      nullptr, nullptr, 0u, labels, prev, perm, out, n);
            k_fl_scan<<<1, kBlock, 0, c->stream>>>((uint32_t *)pcnt.p, (uint32_t *)dpart.p, S, flag_parts,
                                                   (uint32_t *)offs.p);
            k_fl_write<<<flg, kBlock, 0, c->stream>>>(fl, lin, outdeg, S, (const uint32_t *)pcnt.p,
                                                      (const uint32_t *)dpart.p, (uint32_t *)offs.p, n);
            k_dpush<<<push_grid, kBlock, 0, c->stream>>>(CcVisit{labels, prev}, (const uint32_t *)g->off.p,
                                                         (const uint32_t *)g->ids.p, outdeg, S, cnt, lin,
                                                         (const uint32_t *)offs.p, lout, ecap);
            // dense: the live-label pull
            k_ccd_pull<<<pull_grid, kBlock, 0, c->stream>>>(first, toff, tids, labels, prev, next, outdeg, cnt, S, n);
            if (hipGetLastError() != hipSuccess) rc = fail(GG_ERR_HIP, "gg_cc: round kernels");
        }
        if (rc) break;
        static_assert(sizeof h <= kPinBytes, "the state words fit the pinned read-back buffer");
        if ((rc = read_back(c, h, S, sizeof h))) break;
        done = h[BD_MODE] == BM_DONE;
        batch = std::min(std::max(batch, 8) * 2, 64);
        // a component's least label moves on by an edge at least every
        // round, so a run still going after n + 2 enqueued rounds has gone
        // wrong
        if (!done && level > (int)n + 2) rc = fail(GG_ERR_STATE, "gg_cc: more rounds than vertices");
    }
    if (!rc) {
        if (round_hint) *round_hint = (int)h[BD_LEVELS];
        if (info) {
            info->iters = (int)h[BD_LEVELS];
            info->edges_scanned = h[BD_SCANNED];
            for (int L = 1; L <= info->iters && L <= 64; L++) info->paths[L - 1] = (int)(long long)h[BD_PATH + L - 1];
        }
    }
    give();
    return rc;
}

} // namespace gg

using namespace gg;

// ============================================================================
// C ABI: frontier
// ============================================================================

extern "C" int gg_frontier_create(gg_graph_t g, int kind, uint32_t v, gg_frontier_t *out)
{
    if (!g || !out) return fail(GG_ERR_INVALID, "gg_frontier_create: NULL argument");
    gg_frontier_s *f = new gg_frontier_s();
    frontier_alloc(g, f);
    Ctx *c = g->ctx;
    int rc = GG_OK;
    switch (kind) {
    case GG_FRONTIER_EMPTY:
        f->n = 0; // numVertices = 0 as in empty() (ligra-numa.h:124)
        break;
    case GG_FRONTIER_SINGLE: {
        if (v >= g->n) { delete f; return fail(GG_ERR_INVALID, "gg_frontier_create: vertex out of range"); }
        rc = pool_take(c, f->list, 8);
        if (!rc) rc = gg_memcpy_h2d(abi(c), f->list.p, &v, 4);
        uint32_t od = 0;
        if (!rc) rc = gg_memcpy_d2h(abi(c), &od, (uint32_t *)g->outdeg.p + v, 4);
        f->d_m = 1;
        f->out_edges = od;
        f->sparse_valid = true;
        break;
    }
    case GG_FRONTIER_ALL:
        f->bit = true;
        f->has_dense = true;
        f->d_m = g->n;
        f->out_edges = g->m;
        break;
    default: delete f; return fail(GG_ERR_INVALID, "gg_frontier_create: bad kind");
    }
    if (rc) { frontier_free(f); delete f; return rc; }
    *out = f;
    return GG_OK;
}

extern "C" int gg_frontier_from_host(gg_graph_t g, const uint8_t *dense, const uint32_t *sparse,
                                     uint64_t count, gg_frontier_t *out)
{
    if (!g || !out || (!dense && !sparse && count)) return fail(GG_ERR_INVALID, "gg_frontier_from_host: NULL argument");
    gg_frontier_s *f = new gg_frontier_s();
    frontier_alloc(g, f);
    Ctx *c = g->ctx;
    int rc = GG_OK;
    if (dense) {
        rc = ensure_bytes(f);
        if (!rc) rc = gg_memcpy_h2d(abi(c), f->bytes.p, dense, g->n);
        if (!rc) rc = ensure_bitmap(f);
        if (!rc) rc = bitmap_from_bytes(c, (uint8_t *)f->bytes.p, (uint32_t *)f->bitmap.p, g->n);
        f->has_dense = true;
        f->dense_valid = true;
        if (!rc) rc = frontier_stats_dense(f);
    } else {
        for (uint64_t i = 0; i < count; i++)
            if (sparse[i] >= g->n) { delete f; return fail(GG_ERR_INVALID, "gg_frontier_from_host: id out of range"); }
        rc = pool_take(c, f->list, 4 * (count ? count : 1));
        if (!rc && count) rc = gg_memcpy_h2d(abi(c), f->list.p, sparse, 4 * count);
        f->d_m = count;
        f->sparse_valid = true;
        if (!rc && count) {
            std::vector<uint32_t> od(g->n);
            rc = gg_memcpy_d2h(abi(c), od.data(), g->outdeg.p, 4ull * g->n);
            uint64_t s = 0;
            for (uint64_t i = 0; i < count; i++) s += od[sparse[i]];
            f->out_edges = s;
        }
    }
    if (rc) { frontier_free(f); delete f; return rc; }
    *out = f;
    return GG_OK;
}

extern "C" int gg_frontier_destroy(gg_frontier_t f)
{
    if (!f) return fail(GG_ERR_INVALID, "gg_frontier_destroy: NULL");
    // the buffers go back to the context's pool (or are freed if the
    // context is gone); work queued on them finishes first by stream order
    frontier_free(f);
    delete f;
    return GG_OK;
}

extern "C" int gg_frontier_info(gg_frontier_t f, uint64_t *d_m, uint64_t *out_edges, int *is_dense,
                                int *bit, uint32_t *n)
{
    if (!f) return fail(GG_ERR_INVALID, "gg_frontier_info: NULL");
    if (d_m) *d_m = f->d_m;
    if (out_edges) *out_edges = f->out_edges;
    if (is_dense) *is_dense = f->has_dense;
    if (bit) *bit = f->bit;
    if (n) *n = f->n;
    return GG_OK;
}

extern "C" int gg_frontier_set_bit(gg_frontier_t f, int bit)
{
    if (!f) return fail(GG_ERR_INVALID, "gg_frontier_set_bit: NULL");
    if (f->n == 0) return fail(GG_ERR_STATE, "gg_frontier_set_bit: empty() frontier has no rows");
    if (f->bit == (bit != 0)) return GG_OK;
    if (bit) {
        // dense branches now treat every vertex as active; the stored
        // members (flags or list) stay, as the reference's d does, for a
        // sparse branch chosen from the stored d_m / num_out_edges
        f->bit = true;
        return GG_OK;
    }
    // clearing the flag: the members become every vertex explicitly
    GG_TRY(frontier_make_dense(f));
    f->bit = false;
    return sync(f->g->ctx);
}

extern "C" int gg_frontier_to_dense(gg_frontier_t f)
{
    if (!f) return fail(GG_ERR_INVALID, "gg_frontier_to_dense: NULL");
    if (f->n == 0) return GG_OK;
    GG_TRY(frontier_make_dense(f));
    f->has_dense = true;
    return sync(f->g->ctx);
}

extern "C" int gg_frontier_to_sparse(gg_frontier_t f)
{
    if (!f) return fail(GG_ERR_INVALID, "gg_frontier_to_sparse: NULL");
    if (f->n == 0) return GG_OK;
    if (f->bit) return fail(GG_ERR_STATE, "toSparse on a bits() frontier reads unset flags (ligra-numa.h:250)");
    GG_TRY(frontier_make_sparse(f));
    f->has_dense = false;
    return sync(f->g->ctx);
}

extern "C" int gg_frontier_copy_dense(gg_frontier_t f, uint8_t *out)
{
    if (!f || !out) return fail(GG_ERR_INVALID, "gg_frontier_copy_dense: NULL");
    if (f->n == 0) return GG_OK;
    GG_TRY(frontier_make_dense(f));
    return gg_memcpy_d2h(abi(f->g->ctx), out, f->bytes.p, f->n);
}

extern "C" int gg_frontier_copy_sparse(gg_frontier_t f, uint32_t *out, uint64_t cap, uint64_t *count)
{
    if (!f || (!out && cap)) return fail(GG_ERR_INVALID, "gg_frontier_copy_sparse: NULL");
    if (f->n == 0) {
        if (count) *count = 0;
        return GG_OK;
    }
    GG_TRY(frontier_make_sparse(f));
    uint64_t k = f->bit ? f->n : f->d_m;
    if (count) *count = k;
    if (cap < k) return fail(GG_ERR_SIZE, "gg_frontier_copy_sparse: capacity too small");
    return gg_memcpy_d2h(abi(f->g->ctx), out, f->list.p, 4 * k);
}

// ============================================================================
// C ABI: edgeMap / vertexMap / reduce
// ============================================================================

static int check_op(gg_graph_t g, const gg_op_t *op)
{
    if (!op || !op->a) return fail(GG_ERR_INVALID, "edgeMap: operator or its first array is NULL");
    switch (op->kind) {
    case GG_OP_PAGERANK:
    case GG_OP_SPMV:
    case GG_OP_SUM:
        if (!op->b) return fail(GG_ERR_INVALID, "edgeMap: p_next is NULL");
        if (op->dtype != GG_F64 && op->dtype != GG_F32) return fail(GG_ERR_INVALID, "edgeMap: dtype must be F64 or F32");
        break;
    case GG_OP_CC:
        if (!op->b) return fail(GG_ERR_INVALID, "edgeMap: prevIDs is NULL");
        break;
    case GG_OP_BFS: break;
    case GG_OP_SSSP:
        if (!op->b) return fail(GG_ERR_INVALID, "edgeMap: Visited is NULL");
        break;
    case GG_OP_PATHSUM:
        if (!op->b) return fail(GG_ERR_INVALID, "edgeMap: Visited is NULL");
        if (op->dtype != GG_F64 && op->dtype != GG_F32) return fail(GG_ERR_INVALID, "edgeMap: dtype must be F64 or F32");
        break;
    default: return fail(GG_ERR_INVALID, "edgeMap: unknown operator kind");
    }
    (void)g;
    return GG_OK;
}

extern "C" int gg_edgemap(gg_graph_t g, gg_frontier_t in, const gg_op_t *op, int64_t threshold,
                          int remdups, gg_frontier_t *out, int *path)
{
    if (!g || !in || !out) return fail(GG_ERR_INVALID, "gg_edgemap: NULL argument");
    GG_TRY(check_op(g, op));
    GG_HIP(hipSetDevice(g->ctx->device));
    Frontier *o = nullptr;
    int rc = edgemap_core(g, in, op, threshold, remdups, &o, path, nullptr);
    if (rc) return rc;
    gg_frontier_s *fo = new gg_frontier_s();
    static_cast<Frontier &>(*fo) = std::move(*o);
    delete o;
    *out = fo;
    return sync(g->ctx);
}

extern "C" int gg_vertexmap(gg_graph_t g, gg_frontier_t f, const gg_vop_t *op)
{
    if (!g || !op || !op->a) return fail(GG_ERR_INVALID, "gg_vertexmap: NULL argument");
    if ((op->kind == GG_VOP_PAGERANK || op->kind == GG_VOP_ABSDIFF || op->kind == GG_VOP_CC_COPY) && !op->b)
        return fail(GG_ERR_INVALID, "gg_vertexmap: second array is NULL");
    if (op->kind < GG_VOP_PAGERANK || op->kind > GG_VOP_ABSDIFF) return fail(GG_ERR_INVALID, "gg_vertexmap: bad kind");
    if (f && f->n != 0 && f->n != g->n) return fail(GG_ERR_SIZE, "gg_vertexmap: frontier size mismatch");
    GG_TRY(vertexmap_core(g, f, op));
    return sync(g->ctx);
}

extern "C" int gg_vertexfilter(gg_graph_t g, gg_frontier_t f, const uint8_t *flags_dev, gg_frontier_t *out)
{
    if (!g || !f || !flags_dev || !out) return fail(GG_ERR_INVALID, "gg_vertexfilter: NULL argument");
    if (f->n != g->n) return fail(GG_ERR_SIZE, "gg_vertexfilter: frontier size mismatch");
    Ctx *c = g->ctx;
    int mode = f->bit ? 0 : 1;
    if (!f->bit) GG_TRY(frontier_make_dense(f));
    gg_frontier_s *o = new gg_frontier_s();
    frontier_alloc(g, o);
    int rc = ensure_bytes(o);
    if (!rc) {
        k_vfilter<<<blocks_for(g->n, kBlock), kBlock, 0, c->stream>>>(mode, (const uint8_t *)f->bytes.p,
                                                                      flags_dev, (uint8_t *)o->bytes.p, g->n);
        o->has_dense = true;
        o->dense_valid = true;
        rc = ensure_bitmap(o);
    }
    if (!rc) rc = bitmap_from_bytes(c, (uint8_t *)o->bytes.p, (uint32_t *)o->bitmap.p, g->n);
    if (!rc) rc = frontier_stats_dense(o);
    if (rc) { frontier_free(o); delete o; return rc; }
    *out = o;
    return GG_OK;
}

extern "C" int gg_reduce_sum(gg_ctx_t c, const void *dev, int dtype, uint64_t count, double *out)
{
    if (!c || !out || (count && !dev)) return fail(GG_ERR_INVALID, "gg_reduce_sum: NULL argument");
    if (dtype != GG_F64 && dtype != GG_F32) return fail(GG_ERR_INVALID, "gg_reduce_sum: F64 or F32 only");
    Buf r;
    GG_TRY(pool_take(c, r, 8));
    int rc = reduce_sum_core(c, dev, dtype, count, (double *)r.p);
    if (!rc) rc = gg_memcpy_d2h(abi(c), out, r.p, 8);
    pool_give(c, r);
    return rc;
}
