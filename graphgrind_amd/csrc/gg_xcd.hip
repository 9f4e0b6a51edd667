// gg_xcd.hip -- the source-segmented form of a plan's in-edge twin, for the
// pull that gives each XCD one slice of the source vector (see XSeg in
// gg_internal.h).  Built once per plan on first use:
//   1. key each in-edge by (source group, row), stably radix-sorted, so each
//      group's edges are grouped by row with the row's source order kept
//      (and, on a weighted graph, the same sort moves the weights);
//   2. each run of equal keys is a virtual row: offsets, the row's group mask,
//      the group boundaries;
//   3. wave work items over the virtual rows (no item spans two groups),
//      the first work item of each group, and for each combine block of
//      kXCombRows rows the number of each group's virtual rows before it.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <algorithm>
#include <cstring>

#include "gg_kernels.h"

namespace gg {

void xseg_free(XSeg &x)
{
    Buf *bs[] = {&x.vtoff, &x.vtids, &x.vw, &x.mask, &x.cursor, &x.partial, &x.meta};
    for (Buf *b : bs) buf_free(*b);
    rows_free(x.rows);
    x.ok = false;
    x.vrows = 0;
    x.comb_blocks = 0;
    x.grid = 0;
    std::memset(x.row_start, 0, sizeof(x.row_start));
    std::memset(x.item_start, 0, sizeof(x.item_start));
}

namespace {

// key = group << 32 | row for every in-edge (row by binary search in toff)
__global__ void k_xseg_keys(const uint32_t *__restrict__ toff, const uint32_t *__restrict__ tids, uint32_t n,
                            uint64_t m, uint64_t *__restrict__ keys)
{
    uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; e < m; e += stride) {
        uint32_t lo = 0, hi = n; // last row with toff[row] <= e
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if ((uint64_t)toff[mid] <= e) lo = mid;
            else hi = mid;
        }
        const uint64_t grp = (tids[e] >> kXSegShift) % kXSegs;
        keys[e] = grp << 32 | lo;
    }
}

struct RunHead {
    const uint64_t *keys;
    __host__ __device__ uint32_t operator()(uint64_t e) const { return (e == 0 || keys[e] != keys[e - 1]) ? 1u : 0u; }
};

// the first edge of each run opens virtual row vidx[e] (and is marked in
// bit 31 of its source)
__global__ void k_xseg_runs(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ vidx, uint64_t m,
                            uint32_t *__restrict__ vtoff, uint64_t *__restrict__ vkey, uint32_t *__restrict__ mask,
                            uint32_t *__restrict__ vtids)
{
    uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; e < m; e += stride) {
        const uint64_t k = keys[e];
        if (e == 0 || k != keys[e - 1]) {
            const uint32_t v = vidx[e];
            vtoff[v] = (uint32_t)e;
            vkey[v] = k;
            atomicOr(&mask[(uint32_t)k], 1u << (uint32_t)(k >> 32));
            vtids[e] |= 0x80000000u;
        }
    }
}

// first virtual row with key >= target, within [lo, hi)
__device__ __forceinline__ uint64_t key_lower_bound(const uint64_t *__restrict__ vkey, uint64_t lo, uint64_t hi,
                                                    uint64_t target)
{
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (vkey[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// group boundaries: row_start[k] = first virtual row of group k
__global__ void k_xseg_groups(const uint64_t *__restrict__ vkey, uint64_t vrows, uint64_t *__restrict__ row_start)
{
    const uint32_t k = threadIdx.x;
    if (k <= (uint32_t)kXSegs) row_start[k] = key_lower_bound(vkey, 0, vrows, (uint64_t)k << 32);
}

// cursor[b*kXSegs + k] = group-k virtual rows whose row is below b*kXCombRows
__global__ void k_xseg_cursors(const uint64_t *__restrict__ vkey, const uint64_t *__restrict__ row_start,
                               uint32_t blocks, uint32_t *__restrict__ cursor)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)blocks * kXSegs) return;
    const uint32_t b = (uint32_t)(t / kXSegs), k = (uint32_t)(t % kXSegs);
    const uint64_t lo = row_start[k], hi = row_start[k + 1];
    const uint64_t target = (uint64_t)k << 32 | ((uint64_t)b * kXCombRows);
    cursor[t] = (uint32_t)(key_lower_bound(vkey, lo, hi, target) - lo);
}

} // namespace

int ensure_xseg(Graph *g)
{
    XSeg &x = g->xseg;
    if (x.ok) return GG_OK;
    if (!g->plan.mode) return fail(GG_ERR_STATE, "source-segmented pull: the graph has no plan");
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    if (!m || !n) return fail(GG_ERR_STATE, "source-segmented pull: empty graph");
    if (n > 0x7fffffffu) return fail(GG_ERR_SIZE, "source-segmented pull: more than 2^31 vertices");
    xseg_free(x);
    Buf k0, k1, vidx, vkey, rs;
    auto cleanup = [&]() {
        Buf *bs[] = {&k0, &k1, &vidx, &vkey, &rs};
        for (Buf *b : bs) buf_free(*b);
    };
    auto bail = [&](int rc) {
        cleanup();
        xseg_free(x);
        return rc;
    };
    int rc = buf_alloc(c, k0, 8 * m);
    if (!rc) rc = buf_alloc(c, k1, 8 * m);
    // the partial pass reads whole lanes' edges past an item's end
    const uint64_t pad_edges = m + kXWaveEdges;
    if (!rc) rc = buf_alloc(c, x.vtids, 4 * pad_edges);
    if (!rc) rc = buf_alloc(c, vidx, 4 * m);
    if (!rc) rc = buf_alloc(c, rs, 8ull * (kXSegs + 1));
    if (!rc) rc = buf_alloc(c, x.mask, 4ull * n);
    if (rc) return bail(rc);
    const uint32_t grid = (uint32_t)std::min<uint64_t>((m + kBlock - 1) / kBlock, 1u << 16);
    // 1. keys and the stable sort (sources ride along as the values)
    k_xseg_keys<<<grid, kBlock, 0, c->stream>>>((const uint32_t *)g->plan.toff.p, (const uint32_t *)g->plan.tids.p,
                                                 n, m, (uint64_t *)k0.p);
    size_t tb = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, tb, (uint64_t *)k0.p, (uint64_t *)k1.p,
                                             (const uint32_t *)g->plan.tids.p, (uint32_t *)x.vtids.p, (size_t)m, 0,
                                             32 + 3, c->stream);
    if (e == hipSuccess && !(rc = buf_ensure(c, c->tmp, tb + 16)))
        e = rocprim::radix_sort_pairs(c->tmp.p, tb, (uint64_t *)k0.p, (uint64_t *)k1.p,
                                      (const uint32_t *)g->plan.tids.p, (uint32_t *)x.vtids.p, (size_t)m, 0, 32 + 3,
                                      c->stream);
    if (rc) return bail(rc);
    if (e != hipSuccess) return bail(fail(GG_ERR_HIP, std::string("source-segmented pull: sort: ") + hipGetErrorString(e)));
    if (g->weighted && g->plan.tw.p) {
        // the weights in the same order: the same stable sort of the same
        // keys (still in k0), the weights riding along
        if ((rc = buf_alloc(c, x.vw, 4 * pad_edges))) return bail(rc);
        if ((rc = fill_bytes(c, (float *)x.vw.p + m, 4ull * kXWaveEdges, 0))) return bail(rc);
        Buf kout;
        if ((rc = buf_alloc(c, kout, 8 * m))) return bail(rc);
        tb = 0;
        e = rocprim::radix_sort_pairs(nullptr, tb, (const uint64_t *)k0.p, (uint64_t *)kout.p,
                                      (const float *)g->plan.tw.p, (float *)x.vw.p, (size_t)m, 0, 32 + 3, c->stream);
        if (e == hipSuccess && !(rc = buf_ensure(c, c->tmp, tb + 16)))
            e = rocprim::radix_sort_pairs(c->tmp.p, tb, (const uint64_t *)k0.p, (uint64_t *)kout.p,
                                          (const float *)g->plan.tw.p, (float *)x.vw.p, (size_t)m, 0, 32 + 3,
                                          c->stream);
        if (!rc) rc = sync(c); // kout is freed next
        buf_free(kout);
        if (rc) return bail(rc);
        if (e != hipSuccess)
            return bail(fail(GG_ERR_HIP, std::string("source-segmented pull: weight sort: ") + hipGetErrorString(e)));
    }
    // 2. virtual rows: runs of equal keys
    {
        auto it = rocprim::make_transform_iterator(rocprim::make_counting_iterator<uint64_t>(0),
                                                   RunHead{(const uint64_t *)k1.p});
        tb = 0;
        e = rocprim::exclusive_scan(nullptr, tb, it, (uint32_t *)vidx.p, 0u, (size_t)m, rocprim::plus<uint32_t>(),
                                    c->stream);
        if (e == hipSuccess && !(rc = buf_ensure(c, c->tmp, tb + 16)))
            e = rocprim::exclusive_scan(c->tmp.p, tb, it, (uint32_t *)vidx.p, 0u, (size_t)m,
                                        rocprim::plus<uint32_t>(), c->stream);
        if (rc) return bail(rc);
        if (e != hipSuccess) return bail(fail(GG_ERR_HIP, std::string("source-segmented pull: scan: ") + hipGetErrorString(e)));
    }
    uint32_t last_idx = 0;
    uint64_t last_keys[2] = {0, 0};
    if ((rc = read_back(c, &last_idx, (const uint32_t *)vidx.p + (m - 1), 4))) return bail(rc);
    if (m >= 2) {
        if ((rc = read_back(c, last_keys, (const uint64_t *)k1.p + (m - 2), 16))) return bail(rc);
    }
    const uint64_t vrows = (uint64_t)last_idx + ((m == 1 || last_keys[1] != last_keys[0]) ? 1u : 0u);
    x.vrows = vrows;
    if (!rc) rc = buf_alloc(c, x.vtoff, 4ull * (vrows + 1));
    if (!rc) rc = buf_alloc(c, vkey, 8ull * vrows);
    if (!rc) rc = buf_alloc(c, x.partial, 8ull * std::max<uint64_t>(vrows, 1));
    if (rc) return bail(rc);
    if ((rc = fill_bytes(c, x.mask.p, 4ull * n, 0))) return bail(rc);
    if ((rc = fill_bytes(c, (uint32_t *)x.vtids.p + m, 4ull * kXWaveEdges, 0))) return bail(rc);
    k_xseg_runs<<<grid, kBlock, 0, c->stream>>>((const uint64_t *)k1.p, (const uint32_t *)vidx.p, m,
                                                 (uint32_t *)x.vtoff.p, (uint64_t *)vkey.p, (uint32_t *)x.mask.p,
                                                 (uint32_t *)x.vtids.p);
    GG_HIP(hipGetLastError());
    if ((rc = fill_u32(c, (uint32_t *)x.vtoff.p + vrows, 1, (uint32_t)m))) return bail(rc);
    k_xseg_groups<<<1, 64, 0, c->stream>>>((const uint64_t *)vkey.p, vrows, (uint64_t *)rs.p);
    GG_HIP(hipGetLastError());
    if ((rc = read_back(c, x.row_start, rs.p, 8ull * (kXSegs + 1)))) return bail(rc);
    // 3. wave items over the virtual rows, each group's first work item,
    //    the combine cursors
    {
        const std::vector<uint64_t> breaks(x.row_start + 1, x.row_start + kXSegs);
        if ((rc = rows_build(c, x.rows, (const uint32_t *)x.vtoff.p, (uint32_t)vrows, m, g->edges_per_lane,
                             kXShortEdges, kXShortEdges, &breaks)))
            return bail(rc);
    }
    {
        std::vector<RowWork> work(x.rows.blocks);
        GG_HIP(hipMemcpyAsync(work.data(), x.rows.work.p, sizeof(RowWork) * work.size(), hipMemcpyDeviceToHost,
                              c->stream));
        if ((rc = sync(c))) return bail(rc);
        for (int k = 0; k <= kXSegs; k++) {
            // the first item whose first row is in group k or later
            const uint64_t r = x.row_start[k];
            x.item_start[k] = (uint32_t)(std::lower_bound(work.begin(), work.end(), r,
                                                          [](const RowWork &w, uint64_t v) { return w.r0 < v; }) -
                                         work.begin());
        }
        x.item_start[kXSegs] = x.rows.blocks;
        // one partial-pass work-group per CU, kXSegs groups of them
        int cus = 0;
        GG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        x.grid = (uint32_t)std::max(1, cus / kXSegs) * kXSegs;
    }
    x.comb_blocks = blocks_for(n, kXCombRows);
    if (!rc) rc = buf_alloc(c, x.cursor, 4ull * x.comb_blocks * kXSegs);
    if (!rc) rc = buf_alloc(c, x.meta, 4ull * 2 * (kXSegs + 1));
    if (rc) return bail(rc);
    k_xseg_cursors<<<blocks_for((uint64_t)x.comb_blocks * kXSegs, kBlock), kBlock, 0, c->stream>>>(
        (const uint64_t *)vkey.p, (const uint64_t *)rs.p, x.comb_blocks, (uint32_t *)x.cursor.p);
    GG_HIP(hipGetLastError());
    {
        uint32_t meta[2 * (kXSegs + 1)];
        for (int k = 0; k <= kXSegs; k++) {
            meta[k] = x.item_start[k];
            meta[kXSegs + 1 + k] = (uint32_t)x.row_start[k];
        }
        GG_HIP(hipMemcpyAsync(x.meta.p, meta, sizeof(meta), hipMemcpyHostToDevice, c->stream));
        if ((rc = sync(c))) return bail(rc); // 'meta' leaves scope here
    }
    if ((rc = sync(c))) return bail(rc);
    cleanup();
    x.ok = true;
    return GG_OK;
}

} // namespace gg
