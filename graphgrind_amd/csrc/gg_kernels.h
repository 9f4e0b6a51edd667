// gg_kernels.h -- device kernel templates shared by the traversal and app
// translation units.  Included by .hip files only.
#pragma once
#include <algorithm>
#include <type_traits>

#include "gg_internal.h"

namespace gg {

// Per-thread statistics of an edgeMap output frontier (GoutDegree,
// ligra-numa.h:275-288): active count, their out-degree sum, plus an FP
// residual used by the fused PageRank loop.
struct LocalStats {
    unsigned long long active = 0;
    unsigned long long out_edges = 0;
    double l1 = 0.0;
    double aux = 0.0; // second residual term, reduced only for ops with kAux
};

// Operators that set 'static constexpr bool kAux = true' have LocalStats::aux
// reduced beside l1 (written after the l1 partials in the same arrays).
template <class Op, class = void>
struct has_aux : std::false_type {};
template <class Op>
struct has_aux<Op, std::void_t<decltype(Op::kAux)>> : std::bool_constant<Op::kAux> {};

// counters layout (unsigned long long): [0] active, [1] out_edges,
// [2] edges scanned, [3] list cursor.  Kernels with many work-groups add
// their active / out_edges / scanned counts into one of kStripes stripes
// (each on its own 128-byte line, picked by the work-group index) instead of
// the totals: same-address atomics from ~10^5 work-groups serialise at one
// L2 channel.  counters_fold() (host) and the PageRank check kernel sum the
// stripes into the totals.
enum { CNT_ACTIVE = 0, CNT_OUT = 1, CNT_SCANNED = 2, CNT_CURSOR = 3 };

__device__ __forceinline__ void cnt_add(unsigned long long *counters, int which, unsigned long long v)
{
    if (v) atomicAdd(&counters[kCntStripeBase + (blockIdx.x & (kStripes - 1)) * kStripeStride + which], v);
}

// Sums the stripes into counters[0..2] and clears them (one work-group of
// kStripes threads).
__global__ void k_counters_fold(unsigned long long *counters);

// dst[i] = src[idx[i]] for i < count (a permutation of a vertex vector, or
// ids renamed through a permutation).
template <class T>
__global__ void k_gather_idx(const T *__restrict__ src, const uint32_t *__restrict__ idx, T *__restrict__ dst,
                             uint64_t count)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) dst[i] = src[__builtin_nontemporal_load(&idx[i])];
}

template <class T>
inline int gather_idx(Ctx *c, const T *src, const uint32_t *idx, T *dst, uint64_t count)
{
    if (!count) return GG_OK;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((count + kBlock - 1) / kBlock, 1u << 20);
    k_gather_idx<T><<<grid, kBlock, 0, c->stream>>>(src, idx, dst, count);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_sum_f64(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block reduction of LocalStats in a fixed tree; thread 0 adds the counts to
// the global counters and returns the block's l1 (meaningful in thread 0);
// with kAux it also writes the block's aux sum to *aux_tot.
template <bool kAux = false>
__device__ __forceinline__ double block_stats(const LocalStats &st, unsigned long long *counters,
                                              double *aux_tot = nullptr)
{
    __shared__ unsigned long long s_a[kBlock / 64], s_o[kBlock / 64];
    __shared__ double s_l[kBlock / 64], s_x[kBlock / 64];
    unsigned long long a = wave_sum_u64(st.active);
    unsigned long long o = wave_sum_u64(st.out_edges);
    double l = wave_sum_f64(st.l1);
    double x = 0;
    if constexpr (kAux) x = wave_sum_f64(st.aux);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        s_a[w] = a;
        s_o[w] = o;
        s_l[w] = l;
        if constexpr (kAux) s_x[w] = x;
    }
    __syncthreads();
    double tot = 0;
    if (threadIdx.x == 0) {
        unsigned long long ta = 0, to = 0;
        double tx = 0;
        for (int k = 0; k < kBlock / 64; k++) {
            ta += s_a[k];
            to += s_o[k];
            tot += s_l[k];
            if constexpr (kAux) tx += s_x[k];
        }
        if constexpr (kAux) *aux_tot = tx;
        if (counters) {
            cnt_add(counters, CNT_ACTIVE, ta);
            cnt_add(counters, CNT_OUT, to);
        }
    }
    return tot;
}

__device__ __forceinline__ bool bit_test(const uint32_t *bm, uint32_t v)
{
    return (bm[v >> 5] >> (v & 31)) & 1u;
}

// Optional per-row prefetch.  An operator that defines row_t, load_row(d)
// and store_row(d, row, acc, stats) has its per-destination inputs loaded in
// the staging phase together with the gathers (so the epilogue waits on no
// memory); other operators are finished through finish(d, acc, stats).
struct NoRow {};
// Optional stage_edge(e, row, s): the value staged for edge e of row 'row'
// with head s, in place of gather(e, s), for operators that need the row
// while staging (the row-block pull then loads the item's row starts
// first).  An operator may also write a per-edge output there, coalesced,
// as the SpMV backward writes dL/dw.
template <class Op, class = void>
struct has_stage_edge : std::false_type {};
template <class Op>
struct has_stage_edge<Op, std::void_t<decltype(std::declval<const Op &>().stage_edge(0u, 0u, 0u))>>
    : std::true_type {};

// Optional row_done(d): true when row d cannot change whatever its
// in-neighbours hold (a CC label already 0).  The row-block pull then skips
// the gathers of a work item whose rows are all done.
template <class Op, class = void>
struct has_row_done : std::false_type {};
template <class Op>
struct has_row_done<Op, std::void_t<decltype(std::declval<const Op &>().row_done(0u))>> : std::true_type {};

template <class Op, class = void>
struct has_row : std::false_type {};
template <class Op>
struct has_row<Op, std::void_t<typename Op::row_t>> : std::true_type {};
template <class Op, bool = has_row<Op>::value>
struct row_of {
    using type = NoRow;
};
template <class Op>
struct row_of<Op, true> {
    using type = typename Op::row_t;
};

template <class Op>
__device__ __forceinline__ typename row_of<Op>::type row_load(const Op &op, uint32_t d)
{
    if constexpr (has_row<Op>::value) return op.load_row(d);
    else return NoRow{};
}

template <class Op, class R, class A>
__device__ __forceinline__ void row_store(const Op &op, uint32_t d, const R &r, A acc, LocalStats &st)
{
    if constexpr (has_row<Op>::value) op.store_row(d, r, acc, st);
    else op.finish(d, acc, st);
}

// ---------------------------------------------------------------------------
// Merge-path pull over an in-edge (CSC-like) array: every destination row is
// owned by exactly one work-group, so per-row results have a fixed
// combination order (no floating-point atomics).  Tile t covers kPullIPB
// consecutive items of the merge of row ends and edges.
//   1. staging: each thread issues its index loads, its per-row prefetches
//      and then its gathers (kPullIPT independent requests per lane), and
//      parks gathered values and row ends in LDS;
//   2. each thread walks kPullIPT consecutive merge items; a block-wide
//      segmented scan joins rows that cross thread boundaries; row sums go
//      to registers, then to LDS;
//   3. epilogue: rows are finished in a coalesced pass, row i by thread
//      i mod kBlock, with its prefetched inputs.
// Rows crossing tile boundaries are finished by k_pull_fixup from per-tile
// carries.  l1_part receives each tile's residual (and, for kAux ops, the
// aux sums in l1_part[tiles ..]).
//
// Op: acc_t, identity(), combine(earlier, later), gather(edge, src),
//     finish(row, acc, LocalStats&) [or row_t/load_row/store_row].
// ---------------------------------------------------------------------------
template <class Op>
__global__ __launch_bounds__(kBlock) void k_pull(Op op, const uint32_t *__restrict__ toff,
                                                 const uint32_t *__restrict__ tids,
                                                 const uint32_t *__restrict__ srow,
                                                 const uint32_t *__restrict__ sedge,
                                                 uint32_t *__restrict__ carry_row,
                                                 typename Op::acc_t *__restrict__ carry_val,
                                                 typename Op::acc_t *__restrict__ head_val,
                                                 unsigned long long *__restrict__ counters,
                                                 double *__restrict__ l1_part)
{
    using A = typename Op::acc_t;
    using R = typename row_of<Op>::type;
    __shared__ A s_val[kPullIPB];
    __shared__ uint32_t s_end[kPullIPB];
    __shared__ uint32_t s_key[kBlock];
    __shared__ A s_sv[kBlock];
    if (op.skip()) return;
    const uint32_t tid = threadIdx.x, t = blockIdx.x;
    const uint32_t r0 = srow[t], r1 = srow[t + 1];
    const uint32_t e0 = sedge[t], e1 = sedge[t + 1];
    const uint32_t nr = r1 - r0, ne = e1 - e0;
    const bool head_open = toff[r0] < e0; // first row has edges in earlier tiles
    // 1. staging (the in-edge arrays are streamed once: non-temporal loads
    //    keep them from evicting the gathered source values from L2)
    // Every load below is unconditional, with its index clamped into the
    // tile: a load under a divergent branch makes the compiler wait for it at
    // the branch join, which would serialise the round trips.  Only the
    // block-uniform emptiness tests branch.
    R rin[kPullIPT];
    {
        const uint32_t nel = ne ? ne - 1 : 0u, nrl = nr ? nr - 1 : 0u;
        uint32_t src[kPullIPT];
        uint32_t rend[kPullIPT];
        A val[kPullIPT];
        if (ne) {
#pragma unroll
            for (int k = 0; k < kPullIPT; k++)
                src[k] = __builtin_nontemporal_load(&tids[e0 + min(tid + k * kBlock, nel)]);
        }
        if (nr) {
#pragma unroll
            for (int k = 0; k < kPullIPT; k++) {
                const uint32_t jr = min(tid + k * kBlock, nrl);
                rend[k] = __builtin_nontemporal_load(&toff[r0 + jr + 1]);
                rin[k] = row_load(op, r0 + jr);
            }
        }
        if (ne) {
#pragma unroll
            for (int k = 0; k < kPullIPT; k++) val[k] = op.gather(e0 + min(tid + k * kBlock, nel), src[k]);
        }
#pragma unroll
        for (int k = 0; k < kPullIPT; k++) {
            const uint32_t j = tid + k * kBlock;
            if (j < ne) s_val[j] = val[k];
            if (j < nr) s_end[j] = rend[k] - e0;
        }
    }
    __syncthreads();
    // 2. merge-path walk
    const uint32_t tot = nr + ne;
    const uint32_t diag = min(tid * (uint32_t)kPullIPT, tot);
    const uint32_t nitems = min(diag + (uint32_t)kPullIPT, tot) - diag;
    uint32_t lo = diag > ne ? diag - ne : 0u, hi = min(diag, nr);
    while (lo < hi) {
        const uint32_t p = (lo + hi) >> 1;
        if (s_end[p] <= diag - p - 1) lo = p + 1;
        else hi = p;
    }
    const uint32_t ri0 = lo, ej0 = diag - lo;
    uint32_t ri = ri0, ej = ej0;
    A acc = op.identity();
#pragma unroll
    for (int k = 0; k < kPullIPT; k++) {
        if ((uint32_t)k < nitems) {
            if (ri < nr && ej >= s_end[ri]) {
                acc = op.identity();
                ri++;
            } else {
                acc = op.combine(acc, s_val[ej]);
                ej++;
            }
        }
    }
    s_key[tid] = ri;
    s_sv[tid] = acc;
    __syncthreads();
    // inclusive segmented scan over threads (keys are non-decreasing)
    for (uint32_t off = 1; off < kBlock; off <<= 1) {
        const bool take = tid >= off;
        uint32_t k2 = 0;
        A v2 = op.identity();
        if (take) {
            k2 = s_key[tid - off];
            v2 = s_sv[tid - off];
        }
        __syncthreads();
        if (take && k2 == s_key[tid]) s_sv[tid] = op.combine(v2, s_sv[tid]);
        __syncthreads();
    }
    A run = op.identity();
    if (tid > 0 && s_key[tid - 1] == ri0) run = s_sv[tid - 1];
    // pass B: row sums of the rows ending in this thread's items
    uint32_t rloc[kPullIPT];
    A rsum[kPullIPT];
    ri = ri0;
    ej = ej0;
#pragma unroll
    for (int k = 0; k < kPullIPT; k++) {
        rloc[k] = 0xffffffffu;
        rsum[k] = op.identity();
        if ((uint32_t)k < nitems) {
            if (ri < nr && ej >= s_end[ri]) {
                rloc[k] = ri;
                rsum[k] = run;
                run = op.identity();
                ri++;
            } else {
                run = op.combine(run, s_val[ej]);
                ej++;
            }
        }
    }
    if (nitems && diag + nitems == tot) {
        carry_row[t] = r0 + ri;
        carry_val[t] = run;
    }
    __syncthreads(); // s_val is free: reuse it for the row sums
#pragma unroll
    for (int k = 0; k < kPullIPT; k++)
        if (rloc[k] != 0xffffffffu) s_val[rloc[k]] = rsum[k];
    __syncthreads();
    // 3. coalesced epilogue
    LocalStats st;
#pragma unroll
    for (int k = 0; k < kPullIPT; k++) {
        const uint32_t i = tid + k * kBlock;
        if (i < nr) {
            if (i == 0 && head_open) head_val[t] = s_val[0];
            else row_store(op, r0 + i, rin[k], s_val[i], st);
        }
    }
    constexpr bool kAux = has_aux<Op>::value;
    double aux = 0;
    double l1 = block_stats<kAux>(st, counters, &aux);
    if (l1_part && tid == 0) {
        l1_part[t] = l1;
        if constexpr (kAux) l1_part[gridDim.x + t] = aux;
    }
}

// Finishes rows that began in an earlier tile: row r0 of tile t when its
// end lies in t.  Partials are combined in tile order.  Block b also folds
// the per-tile residuals l1_part[256b .. 256b+255] into l1_fix[b] in a fixed
// tree, so the per-iteration check reads one value per fixup block.
template <class Op>
__global__ __launch_bounds__(kBlock) void k_pull_fixup(Op op, const uint32_t *__restrict__ toff,
                                                       const uint32_t *__restrict__ srow,
                                                       const uint32_t *__restrict__ sedge,
                                                       const uint32_t *__restrict__ carry_row,
                                                       const typename Op::acc_t *__restrict__ carry_val,
                                                       const typename Op::acc_t *__restrict__ head_val,
                                                       uint32_t tiles,
                                                       unsigned long long *__restrict__ counters,
                                                       const double *__restrict__ l1_part,
                                                       double *__restrict__ l1_fix)
{
    using A = typename Op::acc_t;
    if (op.skip()) return;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    LocalStats st;
    if (t < tiles) {
        const uint32_t r0 = srow[t], r1 = srow[t + 1], e0 = sedge[t];
        if (r1 > r0 && toff[r0] < e0) {
            uint32_t tb = t;
            while (tb > 0 && carry_row[tb - 1] == r0) tb--;
            A acc = op.identity();
            for (uint32_t k = tb; k < t; k++) acc = op.combine(acc, carry_val[k]);
            acc = op.combine(acc, head_val[t]);
            row_store(op, r0, row_load(op, r0), acc, st);
        }
        if (l1_part) {
            st.l1 += l1_part[t];
            if constexpr (has_aux<Op>::value) st.aux += l1_part[tiles + t];
        }
    }
    constexpr bool kAux = has_aux<Op>::value;
    double aux = 0;
    double l1 = block_stats<kAux>(st, counters, &aux);
    if (l1_fix && threadIdx.x == 0) {
        l1_fix[blockIdx.x] = l1;
        if constexpr (kAux) l1_fix[gridDim.x + blockIdx.x] = aux;
    }
}

// Lane shuffle of any 4- or 8-byte value within segments of 'width' lanes.
template <class T>
__device__ __forceinline__ T shfl_down_any(T v, int off, int width = 64)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte values");
    if constexpr (sizeof(T) == 4) {
        int x;
        __builtin_memcpy(&x, &v, 4);
        x = __shfl_down(x, off, width);
        __builtin_memcpy(&v, &x, 4);
    } else {
        long long x;
        __builtin_memcpy(&x, &v, 8);
        x = __shfl_down(x, off, width);
        __builtin_memcpy(&v, &x, 8);
    }
    return v;
}

// Wave reduction in a fixed tree; the result is in lane 0.
template <class Op, class A>
__device__ __forceinline__ A wave_combine(const Op &op, A v)
{
    for (int off = 32; off > 0; off >>= 1) v = op.combine(v, shfl_down_any(v, off));
    return v;
}

// ---------------------------------------------------------------------------
// Row-block pull (CSR-Adaptive).  Work item b (one per work-group):
//   rows [r0,r1) with all their edges [e0,e1) (<= kRowEdges): the gathered
//     values are staged in LDS with coalesced loads; then groups of
//     G = 2^kind lanes take one row at a time, each lane combining every
//     G-th edge in edge order and the group combining its lanes in a fixed
//     tree (G = 1: one thread per row, edges in order);
//   long (kind kRowLong): edges [e0,e1) of one row with more than kRowEdges
//     edges; the block's partial (lane-strided, then a fixed tree) goes to
//     slot_val[slot]; k_rows_fix combines a row's slots in order.
// Each row is finished by exactly one thread, so results are deterministic.
// ---------------------------------------------------------------------------
// The XCD (accelerator complex) this work-group runs on: HW_REG_XCC_ID
// (hardware register 20), bits [3:0].
__device__ __forceinline__ uint32_t xcc_id()
{
    return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u;
}

template <class Op>
__global__ __launch_bounds__(kBlock) void k_rows(Op op, const uint32_t *__restrict__ toff,
                                                 const uint32_t *__restrict__ tids,
                                                 const RowWork *__restrict__ work,
                                                 typename Op::acc_t *__restrict__ slot_val,
                                                 unsigned long long *__restrict__ counters,
                                                 double *__restrict__ l1_part)
{
    using A = typename Op::acc_t;
    using R = typename row_of<Op>::type;
    constexpr int kRPT = kRowCap / kBlock;
    __shared__ A s_val[kRowEdges];
    __shared__ A s_sum[kRowCap];
    __shared__ uint32_t s_rs[kRowCap + 1];
    __shared__ A s_w[kBlock / 64];
    if (op.skip()) return;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const RowWork w = work[blockIdx.x];
    uint32_t ne = w.e1 - w.e0;
    if constexpr (has_row_done<Op>::value) {
        // every row of the item done: no gathers; rows finish with the
        // identity (a long row's chunk leaves the identity in its slot)
        int done;
        if (w.kind == kRowLong) {
            done = op.row_done(w.r0);
        } else {
            bool d = true;
            for (uint32_t i = tid; i < w.r1 - w.r0; i += kBlock) d = d && op.row_done(w.r0 + i);
            done = __syncthreads_and(d);
        }
        if (done) ne = 0;
    }
    const uint32_t nel = ne ? ne - 1 : 0u;
    // gathered values of this block's edges (clamped, unconditional loads:
    // a load under a divergent branch is waited for at the branch join; the
    // emptiness test is block-uniform)
    A val[kPullIPT];
    if (ne) {
        uint32_t src[kPullIPT];
#pragma unroll
        for (int k = 0; k < kPullIPT; k++)
            src[k] = __builtin_nontemporal_load(&tids[w.e0 + min(tid + k * kBlock, nel)]);
        if constexpr (has_stage_edge<Op>::value) {
            if (w.kind == kRowLong) {
#pragma unroll
                for (int k = 0; k < kPullIPT; k++)
                    val[k] = op.stage_edge(w.e0 + min(tid + k * kBlock, nel), w.r0, src[k]);
            } else {
                // row of each staged edge: the last row starting at or
                // before it (empty rows share their start with the next)
                const uint32_t nr0 = w.r1 - w.r0;
                for (uint32_t i = tid; i < nr0; i += kBlock)
                    s_rs[i] = __builtin_nontemporal_load(&toff[w.r0 + i]) - w.e0;
                if (tid == 0) s_rs[nr0] = ne;
                __syncthreads();
#pragma unroll
                for (int k = 0; k < kPullIPT; k++) {
                    const uint32_t j = min(tid + k * kBlock, nel);
                    uint32_t lo = 0, hi = nr0;
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (s_rs[mid] <= j) lo = mid;
                        else hi = mid;
                    }
                    val[k] = op.stage_edge(w.e0 + j, w.r0 + lo, src[k]);
                }
                __syncthreads(); // the row starts are written again below
            }
        } else {
#pragma unroll
            for (int k = 0; k < kPullIPT; k++) val[k] = op.gather(w.e0 + min(tid + k * kBlock, nel), src[k]);
        }
    }
    if (w.kind == kRowLong) {
        A acc = op.identity();
#pragma unroll
        for (int k = 0; k < kPullIPT; k++)
            if (tid + k * kBlock < ne) acc = op.combine(acc, val[k]);
        acc = wave_combine(op, acc);
        if (lane == 0) s_w[wv] = acc;
        __syncthreads();
        if (tid == 0) {
            A t = s_w[0];
            for (int k = 1; k < kBlock / 64; k++) t = op.combine(t, s_w[k]);
            slot_val[w.slot] = t;
            if (l1_part) {
                l1_part[blockIdx.x] = 0.0;
                if constexpr (has_aux<Op>::value) l1_part[gridDim.x + blockIdx.x] = 0.0;
            }
        }
        return;
    }
    const uint32_t nr = w.r1 - w.r0;
    // row starts (relative to e0) in LDS; no per-row state is held in
    // registers across the staging, which keeps the kernel's register
    // footprint, and so its occupancy, that of a plain gather
#pragma unroll
    for (int k = 0; k < kRPT; k++) {
        const uint32_t i = tid + k * kBlock;
        if (i < nr) s_rs[i] = ne ? __builtin_nontemporal_load(&toff[w.r0 + i]) - w.e0 : 0u;
    }
    if (tid == 0) s_rs[nr] = ne;
#pragma unroll
    for (int k = 0; k < kPullIPT; k++) {
        const uint32_t j = tid + k * kBlock;
        if (j < ne) s_val[j] = val[k];
    }
    __syncthreads();
    const uint32_t lg = w.kind;
    A acc[kRPT];
    if (lg == 0) {
#pragma unroll
        for (int k = 0; k < kRPT; k++) {
            acc[k] = op.identity();
            const uint32_t i = tid + k * kBlock;
            if (i < nr) {
                const uint32_t a = s_rs[i], b = s_rs[i + 1];
                for (uint32_t j = a; j < b; j++) acc[k] = op.combine(acc[k], s_val[j]);
            }
        }
    } else {
        const uint32_t G = 1u << lg, q = tid & (G - 1);
        for (uint32_t i = tid >> lg; i < nr; i += kBlock >> lg) {
            const uint32_t a = s_rs[i], b = s_rs[i + 1];
            A t = op.identity();
            for (uint32_t j = a + q; j < b; j += G) t = op.combine(t, s_val[j]);
            for (uint32_t off = G >> 1; off > 0; off >>= 1) t = op.combine(t, shfl_down_any(t, off, G));
            if (q == 0) s_sum[i] = t;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kRPT; k++) {
            const uint32_t i = tid + k * kBlock;
            acc[k] = i < nr ? s_sum[i] : op.identity();
        }
    }
    // epilogue: the rows' own inputs are loaded now, all before any store
    LocalStats st;
    R rin[kRPT];
    const uint32_t nrl = nr - 1;
#pragma unroll
    for (int k = 0; k < kRPT; k++) rin[k] = row_load(op, w.r0 + min(tid + k * kBlock, nrl));
#pragma unroll
    for (int k = 0; k < kRPT; k++)
        if (tid + k * kBlock < nr) row_store(op, w.r0 + tid + k * kBlock, rin[k], acc[k], st);
    constexpr bool kAux = has_aux<Op>::value;
    double aux = 0;
    double l1 = block_stats<kAux>(st, counters, &aux);
    if (l1_part && tid == 0) {
        l1_part[blockIdx.x] = l1;
        if constexpr (kAux) l1_part[gridDim.x + blockIdx.x] = aux;
    }
}

// Reductions of a source-segmented pull: the value type, the identity and
// the combine (applied in a fixed order).  SegSum: the fused PageRank
// loop's sums, whose identity -0.0 marks "no active source"; SegMin: a
// minimum (measured for Components' dense rounds, which keep the live-label
// row-block pull: reading round-start labels costs them an extra round).
template <class T>
struct SegSum {
    using V = T;
    static __device__ __forceinline__ T identity() { return -T(0); }
    static __device__ __forceinline__ T combine(T a, T b) { return a + b; }
};
template <class T>
struct SegMin {
    using V = T;
    static __device__ __forceinline__ T identity() { return ~T(0); }
    static __device__ __forceinline__ T combine(T a, T b) { return a < b ? a : b; }
};

template <class T>
struct SegOr {
    using V = T;
    static __device__ __forceinline__ T identity() { return T(0); }
    static __device__ __forceinline__ T combine(T a, T b) { return a | b; }
};

// Finishes a long virtual row for k_rows_fix: its chunks' reductions
// combined in chunk order go to partial[v].
template <class P>
struct OpSegPartial {
    using acc_t = typename P::V;
    acc_t *partial;
    const int *done;
    __device__ bool skip() const { return done && *done != 0; }
    __device__ acc_t identity() const { return P::identity(); }
    __device__ acc_t combine(acc_t a, acc_t b) const { return P::combine(a, b); }
    __device__ acc_t gather(uint32_t, uint32_t) const { return P::identity(); }
    __device__ void finish(uint32_t v, acc_t acc, LocalStats &) const { partial[v] = acc; }
};

// Cross-lane moves on the DPP path (no LDS traffic, so no wait on the
// LDS queue): update_dpp returns 'old' in lanes whose source lane is out of
// range or whose row is masked off.  Controls: row_shr:k = 0x110 + k,
// row_bcast:15 = 0x142, row_bcast:31 = 0x143, wave_shr:1 = 0x138.
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, kCtrl, kRowMask, 0xf, false);
}
template <int kCtrl, int kRowMask = 0xf, class T>
__device__ __forceinline__ T dpp_val(T old, T v)
{
    if constexpr (sizeof(T) == 8) {
        const uint64_t o = __builtin_bit_cast(uint64_t, old), x = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = dpp_u32<kCtrl, kRowMask>((uint32_t)o, (uint32_t)x);
        const uint32_t hi = dpp_u32<kCtrl, kRowMask>((uint32_t)(o >> 32), (uint32_t)(x >> 32));
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    } else {
        return __builtin_bit_cast(T, dpp_u32<kCtrl, kRowMask>(__builtin_bit_cast(uint32_t, old),
                                                              __builtin_bit_cast(uint32_t, v)));
    }
}
// inclusive wave scan (sum): in-row shifts 1, 2, 4, 8, then row 0's total
// into row 1 and row 2's into row 3, then the first half's into rows 2, 3
__device__ __forceinline__ uint32_t wave_incl_sum_dpp(uint32_t x)
{
    x += dpp_u32<0x111>(0u, x);
    x += dpp_u32<0x112>(0u, x);
    x += dpp_u32<0x114>(0u, x);
    x += dpp_u32<0x118>(0u, x);
    x += dpp_u32<0x142, 0xa>(0u, x);
    x += dpp_u32<0x143, 0xc>(0u, x);
    return x;
}
// one step of the inclusive segmented sum (S, f): a lane whose segment
// started in it keeps S, else adds the earlier lanes' S in front
template <int kCtrl, int kRowMask, class P>
__device__ __forceinline__ void seg_step(typename P::V &S, uint32_t &f)
{
    const typename P::V pv = dpp_val<kCtrl, kRowMask>(P::identity(), S);
    const uint32_t pf = dpp_u32<kCtrl, kRowMask>(0u, f);
    if (!f) S = P::combine(pv, S);
    f |= pf;
}

// Partial pass of a source-segmented pull (GG_OPT_XCD_SEGMENTS): one
// persistent work-group of kXPartThreads per CU.  Work-groups are dispatched
// to the XCDs round-robin by index, so work-group b serves group b % kXSegs
// and every work-group of a group shares one XCD's L2.  Each work-group
// first copies the group's 'hot' lowest-numbered sources (with a
// hot-sources-first plan, its highest out-degrees) into LDS; its waves then
// take the group's work items in turn, gathering from LDS where the source
// is cached and from global memory (a buffer load, which for a cached
// source is given an offset out of range and so reads nothing) otherwise,
// and reduce with P (SegSum, SegMin).
// A wave covers kXWaveEdges edges from its item's first edge rounded down
// to 4, 16 consecutive ones a lane read as four 16-byte vectors, and masks
// off the positions outside the item.  vtids carries each virtual row's
// start in bit 31 of its first source (sources are below 2^31).  Long-row
// chunks sum in the lane and then in a fixed wave tree into slot_val
// (finished by k_rows_fix); a short item (whole rows, at most kXShortEdges
// edges) adds within the lane by row and joins rows that span lanes with a
// segmented scan across the wave: fixed orders throughout, so the partials
// are deterministic.  The partials are plain (cached) stores: stores share
// the load counter, and waiting for the next item's loads would otherwise
// wait for the far acknowledgements of streaming stores.  Which work-group
// serves a group and which XCD it runs on change no value.  kWeighted: each
// edge's term is the gathered value times the edge's weight from vw (the
// weights in vtids order), as SPMV_F's update (app/SPMV.C:17-21).
template <class P, bool kWeighted = false>
__global__ __launch_bounds__(kXPartThreads) void k_xseg_part(const typename P::V *__restrict__ vals,
                                                             typename P::V *__restrict__ partial,
                                                             typename P::V *__restrict__ slot_val,
                                                             const uint32_t *__restrict__ vtids,
                                                             const RowWork *__restrict__ work,
                                                             const uint32_t *__restrict__ item_start, uint32_t n,
                                                             uint32_t hot, const int *__restrict__ done,
                                                             const float *__restrict__ vw)
{
    static_assert(kXLaneEdges == 16, "a lane reads its edges as four 16-byte vectors");
    using V = typename P::V;
    extern __shared__ unsigned char s_raw[];
    V *s_hot = reinterpret_cast<V *>(s_raw);
    if (done && *done) return;
    constexpr uint32_t kWaves = kXPartThreads / 64;
    const uint32_t grp = blockIdx.x % kXSegs, wig = blockIdx.x / kXSegs, G = gridDim.x / kXSegs;
    for (uint32_t i = threadIdx.x; i < hot; i += kXPartThreads) {
        const uint32_t s = xseg_global(i, grp);
        s_hot[i] = s < n ? vals[s] : V(0);
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t i0 = item_start[grp], cnt = item_start[grp + 1] - i0;
    // buffer view of vals: n * sizeof(V) bytes (below 2^31, checked by the
    // host); an offset of 2^31 is out of range and returns 0 without a read
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<V *>(vals), 0, (int)(n * sizeof(V)), 0x00020000);
    auto gather = [&](uint32_t s, bool in) -> V {
        const uint32_t loc = xseg_local(s);
        const bool h = in && loc < hot;
        const V a = s_hot[h ? loc : 0u];
        V g;
        const uint32_t off = (in && !h) ? s * (uint32_t)sizeof(V) : 0x80000000u;
        if constexpr (sizeof(V) == 8)
            g = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(rsrc, off, 0, 0));
        else
            g = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(rsrc, off, 0, 0));
        return h ? a : g;
    };
    const uint32_t stride = G * kWaves;
    for (uint32_t t = wig * kWaves + wv; t < cnt; t += stride) {
        const RowWork w = work[i0 + t];
        // the wave's window [A0, A0 + kXWaveEdges) and this lane's part of it
        const uint32_t A0 = w.e0 & ~3u;
        const uint32_t base = A0 + lane * kXLaneEdges;
        const uint32_t lo = w.e0 > base ? w.e0 - base : 0u;
        const uint32_t hi = w.e1 > base ? min(w.e1 - base, (uint32_t)kXLaneEdges) : 0u;
        const uint32_t vmask = hi > lo ? ((1u << hi) - 1u) & ~((1u << lo) - 1u) : 0u;
        uint32_t src[kXLaneEdges];
        {
            typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
            const u32x4 *p4 = reinterpret_cast<const u32x4 *>(vtids + base); // vtids is padded past its end
#pragma unroll
            for (int q = 0; q < kXLaneEdges / 4; q++) {
                const u32x4 u = p4[q]; // (a streaming load here measured slower)
                src[4 * q] = u.x;
                src[4 * q + 1] = u.y;
                src[4 * q + 2] = u.z;
                src[4 * q + 3] = u.w;
            }
        }
        uint32_t bits = 0;
#pragma unroll
        for (int j = 0; j < kXLaneEdges; j++) {
            bits |= (src[j] >> 31) << j;
            src[j] &= 0x7fffffffu;
        }
        bits &= vmask;
        V val[kXLaneEdges];
#pragma unroll
        for (int j = 0; j < kXLaneEdges; j++) val[j] = gather(src[j], (vmask >> j) & 1u);
        if constexpr (kWeighted) {
            // each edge's term is the gathered value times its weight, read
            // from the weights in the same order as the ids
            const float4 *w4 = reinterpret_cast<const float4 *>(vw + base); // padded like vtids
#pragma unroll
            for (int q = 0; q < kXLaneEdges / 4; q++) {
                const float4 wq = w4[q];
                val[4 * q] *= (V)wq.x;
                val[4 * q + 1] *= (V)wq.y;
                val[4 * q + 2] *= (V)wq.z;
                val[4 * q + 3] *= (V)wq.w;
            }
        }
        if (w.kind == kRowLong) {
            V acc = P::identity();
#pragma unroll
            for (int j = 0; j < kXLaneEdges; j++)
                if ((vmask >> j) & 1u) acc = P::combine(acc, val[j]);
            for (int off = 32; off > 0; off >>= 1) acc = P::combine(acc, __shfl_down(acc, off, 64));
            if (lane == 0) slot_val[w.slot] = acc;
            continue;
        }
        // rows started in earlier lanes of the item
        const uint32_t nst = __popc(bits);
        const uint32_t inc = wave_incl_sum_dpp(nst);
        // r: the row open before the lane's next edge (w.r0 - 1 before the first)
        uint32_t r = w.r0 + (inc - nst) - 1;
        // (selects rather than nested branches: only the store branches)
        V acc = P::identity(), head = P::identity();
        bool seen = false;
#pragma unroll
        for (int j = 0; j < kXLaneEdges; j++) {
            const bool valid = (vmask >> j) & 1u;
            const bool start = (bits >> j) & 1u; // set only where valid
            if (start && seen) partial[r] = acc; // a row inside the lane ends
            head = (start && !seen) ? acc : head;
            seen = seen || start;
            r += start ? 1u : 0u;
            acc = start ? P::identity() : acc;
            acc = valid ? P::combine(acc, val[j]) : acc;
        }
        // S: the open row's sum up to the lane's end (segmented inclusive scan)
        V S = acc;
        uint32_t f = seen ? 1u : 0u;
        seg_step<0x111, 0xf, P>(S, f);
        seg_step<0x112, 0xf, P>(S, f);
        seg_step<0x114, 0xf, P>(S, f);
        seg_step<0x118, 0xf, P>(S, f);
        seg_step<0x142, 0xa, P>(S, f);
        seg_step<0x143, 0xc, P>(S, f);
        const V C = dpp_val<0x138, 0xf>(P::identity(), S); // lane 0: the identity
        // the row open at the lane's start ends at its first row start
        if (seen && lane > 0) partial[w.r0 + (inc - nst) - 1] = P::combine(C, head);
        // the item's last row ends with its last edge
        if (lane == (w.e1 - 1 - A0) / kXLaneEdges) partial[r] = S;
    }
}

// Combine pass of a source-segmented pull: row d adds the partials of its
// virtual rows in group order 0..kXSegs-1 and is finished by Op (its
// row_load / store_row, or finish).  Group k's virtual rows are in row
// order, so the position of row d's is the number of group-k virtual rows
// before it: the block's cursor plus the rows of the block before d with
// bit k in their masks (wave ballots, then the earlier waves' counts).
// l1_part receives each block's residual (and, for kAux ops, its aux sum in
// l1_part[gridDim.x + b]).
template <class Op>
__global__ __launch_bounds__(kBlock) void k_xseg_combine(Op op, const typename Op::acc_t *__restrict__ partial,
                                                         const uint32_t *__restrict__ mask,
                                                         const uint32_t *__restrict__ cursor,
                                                         const uint32_t *__restrict__ row_start, uint64_t vrows,
                                                         uint32_t n, unsigned long long *__restrict__ counters,
                                                         double *__restrict__ l1_part)
{
    using A = typename Op::acc_t;
    using R = typename row_of<Op>::type;
    constexpr int kPer = kXCombRows / kBlock;
    constexpr int kW = kBlock / 64;
    __shared__ uint32_t s_wcnt[kW][kXSegs];
    __shared__ uint32_t s_base[kXSegs];
    if (op.skip()) return;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    if (tid < kXSegs) s_base[tid] = row_start[tid] + cursor[(uint64_t)blockIdx.x * kXSegs + tid];
    LocalStats st;
    const uint64_t vlast = vrows ? vrows - 1 : 0;
    for (int q = 0; q < kPer; q++) {
        const uint32_t d = blockIdx.x * kXCombRows + q * kBlock + tid;
        const bool in = d < n;
        const uint32_t dd = in ? d : n - 1;
        const uint32_t mk = in ? __builtin_nontemporal_load(&mask[dd]) : 0u;
        uint32_t lrank[kXSegs];
#pragma unroll
        for (int k = 0; k < kXSegs; k++) {
            const uint64_t bal = __ballot((mk >> k) & 1u);
            lrank[k] = (uint32_t)__popcll(bal & lt);
            if (lane == 0) s_wcnt[wv][k] = (uint32_t)__popcll(bal);
        }
        __syncthreads(); // s_wcnt and (first pass) s_base written
        // every group's partial is loaded (clamped, unconditional: the
        // neighbouring lanes read the same lines) and added when the bit is set
        A part[kXSegs];
#pragma unroll
        for (int k = 0; k < kXSegs; k++) {
            uint32_t pre = 0;
            for (int w = 0; w < (int)wv; w++) pre += s_wcnt[w][k];
            const uint64_t pos = min((uint64_t)s_base[k] + pre + lrank[k], vlast);
            part[k] = partial[pos];
        }
        const R rin = row_load(op, dd);
        A acc = op.identity();
#pragma unroll
        for (int k = 0; k < kXSegs; k++)
            if ((mk >> k) & 1u) acc = op.combine(acc, part[k]);
        __syncthreads(); // every lane has read s_base and s_wcnt
        if (tid < kXSegs) {
            uint32_t tot = 0;
            for (int w = 0; w < kW; w++) tot += s_wcnt[w][tid];
            s_base[tid] += tot;
        }
        if (in) row_store(op, d, rin, acc, st);
        __syncthreads(); // s_base updated before the next rows read it
    }
    constexpr bool kAux = has_aux<Op>::value;
    double aux = 0;
    const double l1 = block_stats<kAux>(st, counters, &aux);
    if (l1_part && tid == 0) {
        l1_part[blockIdx.x] = l1;
        if constexpr (kAux) l1_part[gridDim.x + blockIdx.x] = aux;
    }
}

// Finishes the long rows (their chunk partials combined in chunk order) and
// folds the per-block residuals l1_part[256b .. 256b+255] into l1_fix[b].
template <class Op>
__global__ __launch_bounds__(kBlock) void k_rows_fix(Op op, const uint32_t *__restrict__ long_rows, uint32_t nlong,
                                                     const typename Op::acc_t *__restrict__ slot_val,
                                                     uint32_t nblocks, unsigned long long *__restrict__ counters,
                                                     const double *__restrict__ l1_part,
                                                     double *__restrict__ l1_fix)
{
    using A = typename Op::acc_t;
    if (op.skip()) return;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    LocalStats st;
    if (t < nlong) {
        const uint32_t row = long_rows[3 * t], s0 = long_rows[3 * t + 1], s1 = long_rows[3 * t + 2];
        A acc = op.identity();
        for (uint32_t k = s0; k < s1; k++) acc = op.combine(acc, slot_val[k]);
        row_store(op, row, row_load(op, row), acc, st);
    }
    if (l1_part && t < nblocks) {
        st.l1 += l1_part[t];
        if constexpr (has_aux<Op>::value) st.aux += l1_part[nblocks + t];
    }
    constexpr bool kAux = has_aux<Op>::value;
    double aux = 0;
    double l1 = block_stats<kAux>(st, counters, &aux);
    if (l1_fix && threadIdx.x == 0) {
        l1_fix[blockIdx.x] = l1;
        if constexpr (kAux) l1_fix[gridDim.x + blockIdx.x] = aux;
    }
}

template <class Op>
int launch_pull(Ctx *c, const CscView &v, const Op &op, unsigned long long *counters, double *l1_part,
                double *l1_fix, hipEvent_t ev_a = nullptr, hipEvent_t ev_b = nullptr)
{
    using A = typename Op::acc_t;
    static_assert(sizeof(A) <= 8, "carry buffers hold 8-byte partials");
    if (v.rows) {
        RowSched &r = *v.rows;
        if (ev_a) GG_HIP(hipEventRecord(ev_a, c->stream));
        k_rows<Op><<<r.blocks, kBlock, 0, c->stream>>>(op, v.toff, v.tids, (const RowWork *)r.work.p,
                                                        (A *)r.slot_val.p, counters, l1_part);
        if (ev_b) GG_HIP(hipEventRecord(ev_b, c->stream));
        k_rows_fix<Op><<<v.fix_blocks(), kBlock, 0, c->stream>>>(op, (const uint32_t *)r.long_rows.p, r.nlong,
                                                                 (const A *)r.slot_val.p, r.blocks, counters,
                                                                 l1_part, l1_fix);
        GG_HIP(hipGetLastError());
        return GG_OK;
    }
    PullSched &s = *v.sched;
    if (ev_a) GG_HIP(hipEventRecord(ev_a, c->stream));
    k_pull<Op><<<s.tiles, kBlock, 0, c->stream>>>(
        op, v.toff, v.tids, (const uint32_t *)s.start_row.p, (const uint32_t *)s.start_edge.p,
        (uint32_t *)s.carry_row.p, (A *)s.carry_val.p, (A *)s.head_val.p, counters, l1_part);
    if (ev_b) GG_HIP(hipEventRecord(ev_b, c->stream));
    k_pull_fixup<Op><<<blocks_for(s.tiles, kBlock), kBlock, 0, c->stream>>>(
        op, v.toff, (const uint32_t *)s.start_row.p, (const uint32_t *)s.start_edge.p,
        (const uint32_t *)s.carry_row.p, (const A *)s.carry_val.p, (const A *)s.head_val.p, s.tiles, counters,
        l1_part, l1_fix);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

inline CscView graph_view(Graph *g)
{
    CscView v;
    v.toff = (const uint32_t *)g->toff.p;
    v.tids = (const uint32_t *)g->tids.p;
    v.sched = &g->pull;
    return v;
}

// Transpose traversal: a pull over the out-edge CSR (rows = sources), with
// merge-path tiles built on first use.
inline int ensure_csr_sched(Graph *g)
{
    if (g->csr_pull_ok) return GG_OK;
    GG_TRY(sched_build(g->ctx, g->csr_pull, (const uint32_t *)g->off.p, g->n, g->m));
    g->csr_pull_ok = true;
    return GG_OK;
}

inline CscView csr_view(Graph *g)
{
    CscView v;
    v.toff = (const uint32_t *)g->off.p;
    v.tids = (const uint32_t *)g->ids.p;
    v.sched = &g->csr_pull;
    return v;
}

// The out-edge CSR with the graph's chosen schedule (the transpose
// traversal): row blocks or merge-path tiles, built on first use.
inline int csr_pull_view(Graph *g, CscView *v)
{
    GG_TRY(ensure_csr_sched(g));
    *v = csr_view(g);
    if (g->pull_schedule == 1) {
        GG_TRY(ensure_csr_rows(g));
        v->rows = &g->csr_rows;
    }
    return GG_OK;
}

// The in-edge twin with the graph's chosen schedule: row blocks (built on
// first use) or merge-path tiles.
inline int pull_view(Graph *g, CscView *v)
{
    *v = graph_view(g);
    if (g->pull_schedule == 1) {
        GG_TRY(ensure_rows(g));
        v->rows = &g->rows;
    }
    return GG_OK;
}

inline CscView plan_view(Graph *g)
{
    CscView v;
    v.toff = (const uint32_t *)g->plan.toff.p;
    v.tids = (const uint32_t *)g->plan.tids.p;
    v.sched = &g->plan.sched;
    if (g->pull_schedule == 1 && g->plan.rows.blocks) v.rows = &g->plan.rows;
    return v;
}

// BFS level bookkeeping carried through edgemap_core by the fused loops.
struct EdgeMapExtra {
    int32_t *levels = nullptr;
    int32_t level = 0;
    unsigned long long scanned = 0;
};

int edgemap_core(Graph *g, Frontier *in, const gg_op_t *op, int64_t threshold, int remdups,
                 Frontier **pout, int *path, EdgeMapExtra *ex);
// BFS.C's loop with the direction choice made on the device (no read-back
// per level) over g's arrays from 'root', parents and levels (nullable) in
// g's numbering; 'first' holds the first entry of each of g's in-lists
// (ensure_first_in).  'level_hint' (nullable) holds the depth of the last
// search on the graph, which sizes the first batch of levels, and receives
// this search's.
// With a renumbered plan: the caller's parents and levels (nullable) and
// the plan's permutation and inverse, filled when the search ends.
struct BfsMapBack {
    int32_t *parents, *levels;
    const uint32_t *perm, *inv;
};
int bfs_device_loop(Graph *g, uint32_t root, const uint32_t *first, int64_t threshold, int32_t *parents,
                    int32_t *levels, gg_run_info *info, const BfsMapBack *mb = nullptr, int *level_hint = nullptr);
// Components.C's loop with each round's branch chosen on the device, over
// g's arrays ('first': the first entry of each of g's in-lists, as
// ensure_first_in builds): labels (and prev, n entries of scratch) start at
// init[v] (the vertex's id in the caller's numbering; NULL: v) and end at
// each component's least label.  'round_hint' (nullable) holds the rounds of the
// last call on the graph, which size the first batch, and receives this
// call's.  With 'out' set (a renumbered plan) the round after the last
// writes out[v] = labels[perm[v]] (the caller's order).
int cc_device_loop(Graph *g, const uint32_t *first, const uint32_t *init, uint32_t *labels, uint32_t *prev,
                   gg_run_info *info, int *round_hint = nullptr, const uint32_t *perm = nullptr,
                   uint32_t *out = nullptr);
// The first entry of each in-list of the graph's traversal arrays (the
// plan's when it has one; 0xffffffff for an empty list), built on first use.
int ensure_first_in(Graph *g);
int vertexmap_core(Graph *g, Frontier *f, const gg_vop_t *op);
int reduce_sum_core(Ctx *c, const void *dev, int dtype, uint64_t count, double *dev_out);

inline void frontier_delete(Frontier *f)
{
    if (!f) return;
    pool_give(f->ctx, f->bytes);
    pool_give(f->ctx, f->bitmap);
    pool_give(f->ctx, f->list);
    delete f;
}

} // namespace gg
