// gg_vebo.hip -- VEBO vertex ordering and graph relabelling.
//
// graphtools/VEBO.C (56-395) is the reference's offline reorder tool: it
// numbers vertices so that P consecutive id ranges hold about the same
// number of in-edges and of vertices, with each range's vertices in
// descending in-degree.  Here the order is computed in the library:
//   - the (in-degree, id) sort runs on the device (one 64-bit radix sort);
//   - the greedy placement is sequential by nature (each step depends on the
//     partition loads so far) and runs on the host, with a heap for the
//     least-loaded partition instead of the reference's scan over all P
//     partitions per vertex (same choices: fewest edges, lowest index);
//   - gg_graph_relabel builds the renumbered graph on the device.
// The steps follow VEBO.C:84-321; the zero-degree placement falls back to
// the reference's one-by-one rule when its fill-to-average rule would not
// place exactly the zero-degree vertices (VEBO.C:235-284 then writes past
// its arrays).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <algorithm>
#include <cstring>
#include <functional>
#include <queue>
#include <vector>

#include "gg_kernels.h"

namespace gg {

// key = (~indeg << 32) | v: ascending keys are in-degree high to low, then id
__global__ void k_vebo_keys(const uint32_t *__restrict__ indeg, uint32_t n, uint64_t *__restrict__ keys)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) keys[v] = ((uint64_t)(0xffffffffu - indeg[v]) << 32) | v;
}

// new out-degree of new vertex i = out-degree of its old vertex
__global__ void k_relabel_deg(const uint32_t *__restrict__ off, const uint32_t *__restrict__ old_of_new, uint32_t n,
                              uint32_t *__restrict__ deg)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint32_t o = old_of_new[i];
        deg[i] = off[o + 1] - off[o];
    }
}

// deg[i] for i < n, 0 for the extra element of an n+1 exclusive scan
struct VeboPadDeg {
    const uint32_t *deg;
    uint32_t n;
    __host__ __device__ uint32_t operator()(uint32_t i) const { return i < n ? deg[i] : 0u; }
};

__global__ void k_invert_u32(const uint32_t *__restrict__ fwd, uint32_t n, uint32_t *__restrict__ inv)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) inv[fwd[v]] = v;
}

// Edge p of the new CSR: its row i by binary search in noff, then the
// matching edge of the old row, its head renumbered (VEBO.C:343-356 keeps
// each list in its old order).
__global__ void k_relabel_edges(const uint32_t *__restrict__ off, const uint32_t *__restrict__ ids,
                                const float *__restrict__ w, const uint32_t *__restrict__ noff,
                                const uint32_t *__restrict__ old_of_new, const uint32_t *__restrict__ new_of_old,
                                uint32_t n, uint64_t m, uint32_t *__restrict__ nids, float *__restrict__ nw)
{
    uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; p < m; p += stride) {
        uint32_t lo = 0, hi = n; // last row with noff[row] <= p
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (noff[mid] <= p) lo = mid;
            else hi = mid;
        }
        const uint32_t o = old_of_new[lo];
        const uint64_t e = off[o] + (p - noff[lo]);
        nids[p] = new_of_old[ids[e]];
        if (w) nw[p] = w[e];
    }
}

} // namespace gg

using namespace gg;

extern "C" int gg_vebo_order(gg_graph_t g, int P, uint32_t *new_of_old, uint64_t *starts)
{
    if (!g || !new_of_old) return fail(GG_ERR_INVALID, "gg_vebo_order: NULL argument");
    const uint32_t n = g->n;
    if (P < 1 || (uint32_t)P > n) return fail(GG_ERR_INVALID, "gg_vebo_order: need 1 <= P <= n");
    Ctx *c = g->ctx;
    GG_HIP(hipSetDevice(c->device));
    // (in-degree desc, id asc) order (VEBO.C:88-100), sorted on the device
    Buf k0, k1;
    int rc = pool_take(c, k0, 8ull * n);
    if (!rc) rc = pool_take(c, k1, 8ull * n);
    std::vector<uint64_t> keys(n);
    if (!rc) {
        k_vebo_keys<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>((const uint32_t *)g->indeg.p, n,
                                                                       (uint64_t *)k0.p);
        size_t tb = 0;
        if (hipGetLastError() != hipSuccess ||
            rocprim::radix_sort_keys(nullptr, tb, (uint64_t *)k0.p, (uint64_t *)k1.p, (size_t)n, 0, 64,
                                     c->stream) != hipSuccess)
            rc = fail(GG_ERR_HIP, "gg_vebo_order: sort");
        if (!rc) rc = buf_ensure(c, c->tmp, tb + 16);
        if (!rc && rocprim::radix_sort_keys(c->tmp.p, tb, (uint64_t *)k0.p, (uint64_t *)k1.p, (size_t)n, 0, 64,
                                            c->stream) != hipSuccess)
            rc = fail(GG_ERR_HIP, "gg_vebo_order: sort");
        if (!rc && hipMemcpyAsync(keys.data(), k1.p, 8ull * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = fail(GG_ERR_HIP, "gg_vebo_order: readback");
        if (!rc) rc = sync(c);
    }
    pool_give(c, k0);
    pool_give(c, k1);
    if (rc) return rc;
    auto deg_of = [&](uint32_t i) { return 0xffffffffu - (uint32_t)(keys[i] >> 32); };
    auto vtx_of = [&](uint32_t i) { return (uint32_t)keys[i]; };
    // degree groups, high to low (VEBO.C:95-144); the zero group counts only
    // a first vertex of degree 0 (a graph with no edges)
    std::vector<uint32_t> gdeg, gcnt;
    uint64_t czero = 0;
    int64_t zgroup = -1;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t d = deg_of(i);
        if (i == 0 || d != deg_of(i - 1)) {
            gdeg.push_back(d);
            gcnt.push_back(0);
        }
        if (d == 0 && zgroup < 0) zgroup = (int64_t)gdeg.size() - 1;
        if (i > 0 && d == 0) czero++;
        if (i == 0 || d != 0) gcnt.back()++;
    }
    const size_t cg = gdeg.size();
    const size_t zg = zgroup < 0 ? 0 : (size_t)zgroup;
    std::vector<int64_t> edges(P, 0), vsize(P, 0);
    std::vector<int64_t> per((size_t)P * cg, 0); // per[p * cg + group]
    // phase 1 (VEBO.C:179-216): least-loaded partition, lowest index on ties
    using Item = std::pair<int64_t, int>;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
    for (int j = 0; j < P; j++) heap.push({0, j});
    int64_t emax = 0;
    for (size_t i = 0; i < cg; i++) {
        const int64_t deg = gdeg[i], same = gcnt[i];
        for (int64_t k = 0; k < same;) {
            const Item top = heap.top();
            heap.pop();
            int64_t take = 1;
            if (emax - top.first > deg) {
                const int64_t remain = (same - k) / P;
                if (remain > 1) take = remain;
            }
            const int j = top.second;
            edges[j] += deg * take;
            per[(size_t)j * cg + i] += take;
            vsize[j] += take;
            emax = std::max(emax, edges[j]);
            heap.push({edges[j], j});
            k += take;
        }
    }
    // phase 2 (VEBO.C:235-284): zero-degree vertices even out the counts
    const int64_t avg = n / (uint32_t)P, extra = n % (uint32_t)P, dzero = (int64_t)(czero / (uint32_t)P);
    std::vector<int64_t> add(P, 0);
    int64_t placed = -1;
    if (extra == 0 || dzero != 0) {
        placed = 0;
        for (int j = 0; j < P; j++) {
            if (vsize[j] < avg) add[j] = avg - vsize[j];
            if (j < extra) add[j]++;
            placed += add[j];
        }
    }
    if (placed == (int64_t)czero) {
        for (int j = 0; j < P; j++) {
            vsize[j] += add[j];
            per[(size_t)j * cg + zg] += add[j];
        }
    } else {
        std::priority_queue<Item, std::vector<Item>, std::greater<Item>> vh;
        for (int j = 0; j < P; j++) vh.push({vsize[j], j});
        for (uint64_t z = 0; z < czero; z++) {
            const Item top = vh.top();
            vh.pop();
            const int j = top.second;
            vsize[j]++;
            per[(size_t)j * cg + zg]++;
            vh.push({vsize[j], j});
        }
    }
    // ids: partition starts, then group by group and partition by
    // partition, the next vertices of the sorted order (VEBO.C:289-321)
    std::vector<int64_t> start(P, 0), fill(P, 0);
    for (int j = 1; j < P; j++) start[j] = start[j - 1] + vsize[j - 1];
    uint64_t real = 0;
    for (size_t i = 0; i < cg; i++)
        for (int j = 0; j < P; j++)
            for (int64_t k = 0; k < per[(size_t)j * cg + i]; k++) {
                if (real >= n) return fail(GG_ERR_STATE, "gg_vebo_order: partition counts exceed n");
                new_of_old[vtx_of((uint32_t)real)] = (uint32_t)(start[j] + fill[j]);
                fill[j]++;
                real++;
            }
    if (real != n) return fail(GG_ERR_STATE, "gg_vebo_order: partition counts do not cover n");
    if (starts) {
        for (int j = 0; j < P; j++) starts[j] = (uint64_t)start[j];
        starts[P] = n;
    }
    return GG_OK;
}

extern "C" int gg_graph_relabel(gg_graph_t g, const uint32_t *new_of_old, gg_graph_t *out)
{
    if (!g || !new_of_old || !out) return fail(GG_ERR_INVALID, "gg_graph_relabel: NULL argument");
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    {
        std::vector<uint8_t> seen(n, 0);
        for (uint32_t v = 0; v < n; v++) {
            if (new_of_old[v] >= n || seen[new_of_old[v]])
                return fail(GG_ERR_INVALID, "gg_graph_relabel: not a permutation (vertex " + std::to_string(v) + ")");
            seen[new_of_old[v]] = 1;
        }
    }
    Ctx *c = g->ctx;
    GG_HIP(hipSetDevice(c->device));
    Buf fwd, inv, deg, noff, nids, nw;
    auto cleanup = [&](int rc) {
        Buf *tmp[] = {&fwd, &inv, &deg};
        for (Buf *b : tmp) pool_give(c, *b);
        buf_free(noff);
        buf_free(nids);
        buf_free(nw);
        return rc;
    };
    int rc = pool_take(c, fwd, 4ull * n);
    if (!rc) rc = pool_take(c, inv, 4ull * n);
    if (!rc) rc = pool_take(c, deg, 4ull * n);
    if (!rc) rc = buf_alloc(c, noff, 4ull * (n + 1));
    if (!rc) rc = buf_alloc(c, nids, 4ull * (m ? m : 1));
    if (!rc && g->weighted) rc = buf_alloc(c, nw, 4ull * (m ? m : 1));
    if (rc) return cleanup(rc);
    if (hipMemcpyAsync(fwd.p, new_of_old, 4ull * n, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return cleanup(fail(GG_ERR_HIP, "gg_graph_relabel: upload"));
    k_invert_u32<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>((const uint32_t *)fwd.p, n, (uint32_t *)inv.p);
    k_relabel_deg<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>((const uint32_t *)g->off.p, (const uint32_t *)inv.p,
                                                                    n, (uint32_t *)deg.p);
    if (hipGetLastError() != hipSuccess) return cleanup(fail(GG_ERR_HIP, "gg_graph_relabel: launch"));
    {
        auto in = rocprim::make_transform_iterator(rocprim::make_counting_iterator<uint32_t>(0),
                                                   VeboPadDeg{(const uint32_t *)deg.p, n});
        size_t tb = 0;
        if (rocprim::exclusive_scan(nullptr, tb, in, (uint32_t *)noff.p, 0u, (size_t)n + 1, rocprim::plus<uint32_t>(),
                                    c->stream) != hipSuccess)
            return cleanup(fail(GG_ERR_HIP, "gg_graph_relabel: scan"));
        if ((rc = buf_ensure(c, c->tmp, tb + 16))) return cleanup(rc);
        if (rocprim::exclusive_scan(c->tmp.p, tb, in, (uint32_t *)noff.p, 0u, (size_t)n + 1, rocprim::plus<uint32_t>(),
                                    c->stream) != hipSuccess)
            return cleanup(fail(GG_ERR_HIP, "gg_graph_relabel: scan"));
    }
    if (m) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>((m + kBlock - 1) / kBlock, 1u << 20);
        k_relabel_edges<<<grid, kBlock, 0, c->stream>>>(
            (const uint32_t *)g->off.p, (const uint32_t *)g->ids.p, g->weighted ? (const float *)g->w.p : nullptr,
            (const uint32_t *)noff.p, (const uint32_t *)inv.p, (const uint32_t *)fwd.p, n, m, (uint32_t *)nids.p,
            g->weighted ? (float *)nw.p : nullptr);
        if (hipGetLastError() != hipSuccess) return cleanup(fail(GG_ERR_HIP, "gg_graph_relabel: launch"));
    }
    if ((rc = sync(c))) return cleanup(rc);
    Graph *ng = nullptr;
    rc = graph_from_device(c, n, m, noff, nids, nw, g->symmetric, &ng);
    if (rc) return cleanup(rc);
    cleanup(GG_OK);
    *out = static_cast<gg_graph_s *>(ng);
    return GG_OK;
}
