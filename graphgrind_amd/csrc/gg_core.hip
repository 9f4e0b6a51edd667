// gg_core.hip -- context, device memory, graph construction and the device
// partitioner of libgraphgrind_amd (MI355X / gfx950).
//
// Graph construction replaces the reference's load path (IO.h:136-361,
// wholeGraph graph-numa.h:500-557) and partitioned_graph constructor
// (graph-numa.h:642-685) with device kernels: the in-edge twin comes from a
// stable device radix sort, the partition greedy runs on the device as P
// binary searches over the degree prefix sum, and COO partitions are
// Hilbert-ordered by a stable segmented radix sort of Hilbert keys.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <cmath>
#include <cstring>
#include <mutex>
#include <set>

#include "gg_internal.h"

namespace gg {

static thread_local std::string t_err = "no error";

void set_error(const std::string &msg) { t_err = msg; }

int fail(int code, const std::string &msg)
{
    t_err = msg;
    return code;
}

int buf_alloc(Ctx *c, Buf &b, uint64_t bytes)
{
    (void)c;
    buf_free(b);
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(GG_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    b.bytes = bytes;
    return GG_OK;
}

int buf_ensure(Ctx *c, Buf &b, uint64_t bytes)
{
    if (b.p && b.bytes >= bytes) return GG_OK;
    return buf_alloc(c, b, bytes);
}

void buf_free(Buf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

// Contexts alive now: a frontier may be destroyed after its context, and
// then its buffers are freed rather than pooled.
static std::mutex g_ctx_mu;
static std::set<Ctx *> g_live_ctx;
constexpr uint64_t kPoolCap = 8ull << 30; // bytes kept per context

void pool_drain(Ctx *c)
{
    std::lock_guard<std::mutex> lk(c->pool_mu);
    for (auto &kv : c->pool) (void)hipFree(kv.second);
    c->pool.clear();
    c->pool_bytes = 0;
}

int pool_take(Ctx *c, Buf &b, uint64_t bytes)
{
    if (b.p) pool_give(c, b);
    if (bytes == 0) bytes = 8;
    {
        std::lock_guard<std::mutex> lk(c->pool_mu);
        auto it = c->pool.lower_bound(bytes);
        if (it != c->pool.end() && it->first <= std::max<uint64_t>(2 * bytes, bytes + (1ull << 20))) {
            b.p = it->second;
            b.bytes = it->first;
            c->pool_bytes -= it->first;
            c->pool.erase(it);
            return GG_OK;
        }
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess && !c->pool.empty()) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(c->stream);
        pool_drain(c);
        e = hipMalloc(&b.p, bytes);
    }
    if (e != hipSuccess) {
        b.p = nullptr;
        b.bytes = 0;
        return fail(GG_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    b.bytes = bytes;
    return GG_OK;
}

int pool_ensure(Ctx *c, Buf &b, uint64_t bytes)
{
    if (b.p && b.bytes >= bytes) return GG_OK;
    return pool_take(c, b, bytes);
}

void pool_give(Ctx *c, Buf &b)
{
    if (!b.p) return;
    {
        // the registry lock is held while pooling, so the context cannot be
        // destroyed in between
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        if (c && g_live_ctx.count(c)) {
            std::lock_guard<std::mutex> pk(c->pool_mu);
            if (c->pool_bytes + b.bytes <= kPoolCap) {
                c->pool.emplace(b.bytes, b.p);
                c->pool_bytes += b.bytes;
                b.p = nullptr;
                b.bytes = 0;
                return;
            }
        }
    }
    buf_free(b);
}

int sync(Ctx *c)
{
    GG_HIP(hipStreamSynchronize(c->stream));
    return GG_OK;
}

int read_back(Ctx *c, void *host, const void *dev, size_t bytes)
{
    if (bytes > kPinBytes) return fail(GG_ERR_INVALID, "read_back: more than the pinned buffer holds");
    GG_HIP(hipMemcpyAsync(c->hpin, dev, bytes, hipMemcpyDeviceToHost, c->stream));
    GG_TRY(sync(c));
    std::memcpy(host, c->hpin, bytes);
    return GG_OK;
}

// ---- small kernels -----------------------------------------------------------

__global__ void k_fill_u8(uint8_t *p, uint64_t n, uint8_t v)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <class T>
__global__ void k_fill(T *p, uint64_t n, T v)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <class T>
__global__ void k_iota(T *p, uint64_t n)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = (T)i;
}

static uint32_t grid_for(uint64_t n)
{
    uint64_t b = (n + kBlock - 1) / kBlock;
    if (b > 65536) b = 65536; // grid-stride beyond this
    return (uint32_t)(b ? b : 1);
}

int fill_bytes(Ctx *c, void *p, uint64_t bytes, uint8_t v)
{
    if (!bytes) return GG_OK;
    k_fill_u8<<<grid_for(bytes), kBlock, 0, c->stream>>>((uint8_t *)p, bytes, v);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

int fill_u32(Ctx *c, void *p, uint64_t count, uint32_t v)
{
    if (!count) return GG_OK;
    k_fill<uint32_t><<<grid_for(count), kBlock, 0, c->stream>>>((uint32_t *)p, count, v);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

// 32 flag bytes -> one bitmap word; one thread per word, coalesced reads.
__global__ void k_bitmap_from_bytes(const uint8_t *__restrict__ bytes, uint32_t *__restrict__ bm,
                                    uint32_t n)
{
    uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nw = (n + 31) / 32;
    if (w >= nw) return;
    uint32_t base = w * 32, word = 0;
    if (base + 32 <= n) {
        const uint4 *p = reinterpret_cast<const uint4 *>(bytes + base);
        uint4 a = p[0], b = p[1];
        uint32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        for (int k = 0; k < 8; k++)
            for (int j = 0; j < 4; j++)
                if ((v[k] >> (8 * j)) & 0xff) word |= 1u << (k * 4 + j);
    } else {
        for (uint32_t j = 0; base + j < n; j++)
            if (bytes[base + j]) word |= 1u << j;
    }
    bm[w] = word;
}

int bitmap_from_bytes(Ctx *c, const uint8_t *bytes, uint32_t *bitmap, uint32_t n)
{
    uint32_t nw = (n + 31) / 32;
    if (!nw) return GG_OK;
    // 16-byte loads need 16-byte alignment of every 32-byte group
    k_bitmap_from_bytes<<<blocks_for(nw, kBlock), kBlock, 0, c->stream>>>(bytes, bitmap, n);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

// ---- merge-path schedule -------------------------------------------------------

// Largest x in [max(diag-edges,0), min(diag,rows)] with off[x] <= diag-x,
// i.e. the row coordinate where diagonal 'diag' crosses the merge path of
// row ends off[1..rows] against edge indices 0..edges-1.
__device__ __forceinline__ uint32_t merge_path_row(const uint32_t *off, uint32_t rows,
                                                   uint64_t edges, uint64_t diag)
{
    uint64_t lo = diag > edges ? diag - edges : 0;
    uint64_t hi = diag < rows ? diag : rows;
    while (lo < hi) {
        uint64_t pivot = (lo + hi) >> 1;
        if ((uint64_t)off[pivot + 1] <= diag - pivot - 1) lo = pivot + 1;
        else hi = pivot;
    }
    return (uint32_t)lo;
}

__global__ void k_sched(const uint32_t *__restrict__ off, uint32_t rows, uint64_t edges,
                        uint32_t tiles, uint32_t *srow, uint32_t *sedge)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > tiles) return;
    uint64_t total = (uint64_t)rows + edges;
    uint64_t diag = (uint64_t)t * kPullIPB;
    if (diag > total) diag = total;
    uint32_t r = merge_path_row(off, rows, edges, diag);
    srow[t] = r;
    sedge[t] = (uint32_t)(diag - r);
}

int sched_build(Ctx *c, PullSched &s, const uint32_t *d_off, uint32_t rows, uint64_t edges)
{
    s.rows = rows;
    s.edges = edges;
    s.tiles = blocks_for((uint64_t)rows + edges, kPullIPB);
    GG_TRY(buf_ensure(c, s.start_row, 4ull * (s.tiles + 1)));
    GG_TRY(buf_ensure(c, s.start_edge, 4ull * (s.tiles + 1)));
    GG_TRY(buf_ensure(c, s.carry_row, 4ull * s.tiles));
    GG_TRY(buf_ensure(c, s.carry_val, 8ull * s.tiles));
    GG_TRY(buf_ensure(c, s.head_val, 8ull * s.tiles));
    k_sched<<<blocks_for(s.tiles + 1, kBlock), kBlock, 0, c->stream>>>(
        d_off, rows, edges, s.tiles, (uint32_t *)s.start_row.p, (uint32_t *)s.start_edge.p);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

void graph_exec_free(GraphExec &x)
{
    if (x.exec) (void)hipGraphExecDestroy(x.exec);
    if (x.graph) (void)hipGraphDestroy(x.graph);
    x.exec = nullptr;
    x.graph = nullptr;
    for (auto &k : x.key) k = 0;
    for (auto &b : x.own) buf_free(b);
}

void graph_exec_drop(Graph *g)
{
    graph_exec_free(g->fwd_exec);
    graph_exec_free(g->adj_exec);
}

int ensure_csr_rows(Graph *g)
{
    if (g->csr_rows_ok) return GG_OK;
    GG_TRY(rows_build(g->ctx, g->csr_rows, (const uint32_t *)g->off.p, g->n, g->m, g->edges_per_lane));
    g->csr_rows_ok = true;
    return GG_OK;
}

int ensure_rows(Graph *g)
{
    if (g->rows_ok) return GG_OK;
    GG_TRY(rows_build(g->ctx, g->rows, (const uint32_t *)g->toff.p, g->n, g->m, g->edges_per_lane));
    g->rows_ok = true;
    return GG_OK;
}

void rows_free(RowSched &s)
{
    buf_free(s.work);
    buf_free(s.long_rows);
    buf_free(s.slot_val);
    s.blocks = s.nlong = s.slots = 0;
}

// Greedy packing of rows in index order (host pass over the offsets; part of
// the untimed graph preparation).  A block closes when the next row would
// exceed kRowEdges edges or kRowCap rows, or when the next row is on the
// other side of kRowBig edges than the block's rows (so one long row does
// not hold up a block of short ones); rows above kRowEdges edges become
// chunked long rows.  Lanes per row: the power of two G <= 64 that gives
// each lane about 'edges_per_lane' of the block's average row.
void hubs_free(HubChunks &h)
{
    buf_free(h.items);
    buf_free(h.hubs);
    h.nitems = h.nhubs = 0;
    h.ok = false;
}

int ensure_hub_chunks(Ctx *c, const uint32_t *toff_dev, uint32_t n, HubChunks &h)
{
    if (h.ok) return GG_OK;
    hubs_free(h);
    std::vector<uint32_t> off((size_t)n + 1);
    GG_HIP(hipMemcpyAsync(off.data(), toff_dev, 4ull * (n + 1), hipMemcpyDeviceToHost, c->stream));
    GG_TRY(sync(c));
    std::vector<uint32_t> items, hubs;
    for (uint32_t v = 0; v < n; v++) {
        if (off[v + 1] - off[v] <= kHubChunk) continue;
        hubs.push_back(v);
        hubs.push_back((uint32_t)(items.size() / 3));
        for (uint32_t e = off[v]; e < off[v + 1]; e += kHubChunk) {
            items.push_back(v);
            items.push_back(e);
            items.push_back(std::min<uint32_t>(e + kHubChunk, off[v + 1]));
        }
        hubs.push_back((uint32_t)(items.size() / 3));
    }
    h.nitems = (uint32_t)(items.size() / 3);
    h.nhubs = (uint32_t)(hubs.size() / 3);
    GG_TRY(buf_alloc(c, h.items, 4ull * std::max<size_t>(items.size(), 1)));
    GG_TRY(buf_alloc(c, h.hubs, 4ull * std::max<size_t>(hubs.size(), 1)));
    if (!items.empty())
        GG_HIP(hipMemcpyAsync(h.items.p, items.data(), 4 * items.size(), hipMemcpyHostToDevice, c->stream));
    if (!hubs.empty())
        GG_HIP(hipMemcpyAsync(h.hubs.p, hubs.data(), 4 * hubs.size(), hipMemcpyHostToDevice, c->stream));
    GG_TRY(sync(c)); // the host vectors leave scope
    h.ok = true;
    return GG_OK;
}

int rows_build(Ctx *c, RowSched &s, const uint32_t *d_off, uint32_t rows, uint64_t edges, int edges_per_lane,
               uint32_t item_edges, uint32_t row_cap, const std::vector<uint64_t> *breaks)
{
    rows_free(s);
    std::vector<uint32_t> off((size_t)rows + 1);
    GG_HIP(hipMemcpyAsync(off.data(), d_off, 4ull * (rows + 1), hipMemcpyDeviceToHost, c->stream));
    GG_TRY(sync(c));
    if ((uint64_t)off[rows] != edges) return fail(GG_ERR_SIZE, "rows_build: offsets do not end at the edge count");
    const uint32_t kRowBig = std::min<uint32_t>(256, item_edges / 2);
    std::vector<RowWork> work;
    std::vector<uint32_t> longs;
    work.reserve(edges / item_edges + rows / row_cap + 16);
    size_t bi = 0; // next forced break (items never span one)
    uint32_t slots = 0;
    uint32_t r0 = 0, nr = 0;
    uint64_t ne = 0;
    bool big = false;
    auto close = [&]() {
        if (!nr) return;
        RowWork w{};
        w.r0 = r0;
        w.r1 = r0 + nr;
        w.e0 = off[r0];
        w.e1 = off[r0 + nr];
        const uint64_t per = (ne + (uint64_t)nr * edges_per_lane - 1) / ((uint64_t)nr * edges_per_lane);
        uint32_t lg = 0;
        while (lg < 6 && (1ull << lg) < per) lg++;
        w.kind = lg;
        work.push_back(w);
        nr = 0;
        ne = 0;
    };
    for (uint32_t r = 0; r < rows; r++) {
        if (breaks) {
            while (bi < breaks->size() && (*breaks)[bi] < r) bi++;
            if (bi < breaks->size() && (*breaks)[bi] == r) close();
        }
        const uint32_t d = off[r + 1] - off[r];
        if (d > item_edges) {
            close();
            longs.push_back(r);
            longs.push_back(slots);
            for (uint32_t e = off[r]; e < off[r + 1]; e += item_edges) {
                RowWork w{};
                w.r0 = r;
                w.r1 = r + 1;
                w.e0 = e;
                w.e1 = std::min<uint32_t>(e + item_edges, off[r + 1]);
                w.kind = kRowLong;
                w.slot = slots++;
                work.push_back(w);
            }
            longs.push_back(slots);
            continue;
        }
        const bool b = d > kRowBig;
        if (nr && (ne + d > (uint64_t)item_edges || nr + 1 > row_cap || b != big)) close();
        if (!nr) {
            r0 = r;
            big = b;
        }
        nr++;
        ne += d;
    }
    close();
    if (work.size() >= (1ull << 31)) return fail(GG_ERR_SIZE, "rows_build: too many blocks");
    s.blocks = (uint32_t)work.size();
    s.nlong = (uint32_t)(longs.size() / 3);
    s.slots = slots;
    GG_TRY(buf_alloc(c, s.work, sizeof(RowWork) * (uint64_t)std::max<size_t>(work.size(), 1)));
    GG_TRY(buf_alloc(c, s.long_rows, 4ull * std::max<size_t>(longs.size(), 1)));
    GG_TRY(buf_alloc(c, s.slot_val, 8ull * std::max<uint32_t>(slots, 1)));
    if (!work.empty())
        GG_HIP(hipMemcpyAsync(s.work.p, work.data(), sizeof(RowWork) * work.size(), hipMemcpyHostToDevice, c->stream));
    if (!longs.empty())
        GG_HIP(hipMemcpyAsync(s.long_rows.p, longs.data(), 4 * longs.size(), hipMemcpyHostToDevice, c->stream));
    return sync(c);
}

} // namespace gg

using namespace gg;

// =============================================================================
// C ABI: context and memory
// =============================================================================

extern "C" const char *gg_last_error(void) { return t_err.c_str(); }

extern "C" const char *gg_version(void) { return "graphgrind_amd 0.1 (gfx950)"; }

extern "C" int gg_ctx_create(int device, gg_ctx_t *out)
{
    if (!out) return fail(GG_ERR_INVALID, "gg_ctx_create: out is NULL");
    int count = 0;
    GG_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(GG_ERR_INVALID, "gg_ctx_create: no device " + std::to_string(device));
    GG_HIP(hipSetDevice(device));
    gg_ctx_s *c = new gg_ctx_s();
    c->device = device;
    // default-priority stream
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(GG_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    (void)hipEventCreate(&c->ev[0]);
    (void)hipEventCreate(&c->ev[1]);
    int rc = buf_alloc(c, c->counters, kCounterBytes);
    if (!rc) rc = fill_bytes(c, c->counters.p, kCounterBytes, 0);
    if (!rc && hipHostMalloc(&c->hpin, kPinBytes, hipHostMallocDefault) != hipSuccess) {
        c->hpin = nullptr;
        rc = fail(GG_ERR_HIP, "gg_ctx_create: hipHostMalloc of the read-back buffer failed");
    }
    if (rc) {
        buf_free(c->counters);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    {
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        g_live_ctx.insert(c);
    }
    *out = c;
    return GG_OK;
}

extern "C" int gg_ctx_destroy(gg_ctx_t c)
{
    if (!c) return fail(GG_ERR_INVALID, "gg_ctx_destroy: NULL");
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    {
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        g_live_ctx.erase(c);
    }
    pool_drain(c);
    buf_free(c->counters);
    buf_free(c->tmp);
    if (c->hpin) (void)hipHostFree(c->hpin);
    (void)hipEventDestroy(c->ev[0]);
    (void)hipEventDestroy(c->ev[1]);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return GG_OK;
}

extern "C" int gg_ctx_sync(gg_ctx_t c)
{
    if (!c) return fail(GG_ERR_INVALID, "gg_ctx_sync: NULL");
    return sync(c);
}

extern "C" int gg_malloc(gg_ctx_t c, uint64_t bytes, void **dev)
{
    if (!c || !dev) return fail(GG_ERR_INVALID, "gg_malloc: NULL argument");
    GG_HIP(hipSetDevice(c->device));
    hipError_t e = hipMalloc(dev, bytes ? bytes : 8);
    if (e != hipSuccess) {
        *dev = nullptr;
        return fail(GG_ERR_OOM, "gg_malloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    return GG_OK;
}

extern "C" int gg_host_alloc(gg_ctx_t c, uint64_t bytes, void **host)
{
    if (!c || !host) return fail(GG_ERR_INVALID, "gg_host_alloc: NULL argument");
    GG_HIP(hipSetDevice(c->device));
    hipError_t e = hipHostMalloc(host, bytes ? bytes : 8, hipHostMallocDefault);
    if (e != hipSuccess) {
        *host = nullptr;
        return fail(GG_ERR_OOM, "gg_host_alloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    return GG_OK;
}

extern "C" int gg_host_free(gg_ctx_t c, void *host)
{
    if (!c) return fail(GG_ERR_INVALID, "gg_host_free: NULL ctx");
    if (!host) return GG_OK;
    GG_HIP(hipStreamSynchronize(c->stream));
    GG_HIP(hipHostFree(host));
    return GG_OK;
}

extern "C" int gg_free(gg_ctx_t c, void *dev)
{
    if (!c) return fail(GG_ERR_INVALID, "gg_free: NULL ctx");
    if (!dev) return GG_OK;
    GG_HIP(hipStreamSynchronize(c->stream));
    GG_HIP(hipFree(dev));
    return GG_OK;
}

extern "C" int gg_memcpy_h2d(gg_ctx_t c, void *dev, const void *host, uint64_t bytes)
{
    if (!c || (bytes && (!dev || !host))) return fail(GG_ERR_INVALID, "gg_memcpy_h2d: NULL argument");
    if (!bytes) return GG_OK;
    GG_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, c->stream));
    return sync(c);
}

extern "C" int gg_memcpy_d2h(gg_ctx_t c, void *host, const void *dev, uint64_t bytes)
{
    if (!c || (bytes && (!dev || !host))) return fail(GG_ERR_INVALID, "gg_memcpy_d2h: NULL argument");
    if (!bytes) return GG_OK;
    GG_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

extern "C" int gg_memcpy_d2d(gg_ctx_t c, void *dst, const void *src, uint64_t bytes)
{
    if (!c || (bytes && (!dst || !src))) return fail(GG_ERR_INVALID, "gg_memcpy_d2d: NULL argument");
    if (!bytes) return GG_OK;
    GG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return sync(c);
}

extern "C" int gg_fill(gg_ctx_t c, void *dev, int dtype, uint64_t count, double value)
{
    if (!c || (count && !dev)) return fail(GG_ERR_INVALID, "gg_fill: NULL argument");
    if (!count) return GG_OK;
    uint32_t g = grid_for(count);
    switch (dtype) {
    case GG_F64: k_fill<double><<<g, kBlock, 0, c->stream>>>((double *)dev, count, value); break;
    case GG_F32: k_fill<float><<<g, kBlock, 0, c->stream>>>((float *)dev, count, (float)value); break;
    case GG_U32: k_fill<uint32_t><<<g, kBlock, 0, c->stream>>>((uint32_t *)dev, count, (uint32_t)(int64_t)value); break;
    case GG_I32: k_fill<int32_t><<<g, kBlock, 0, c->stream>>>((int32_t *)dev, count, (int32_t)value); break;
    case GG_U8: k_fill_u8<<<g, kBlock, 0, c->stream>>>((uint8_t *)dev, count, (uint8_t)value); break;
    default: return fail(GG_ERR_INVALID, "gg_fill: bad dtype");
    }
    GG_HIP(hipGetLastError());
    return sync(c);
}

extern "C" int gg_iota(gg_ctx_t c, void *dev, int dtype, uint64_t count)
{
    if (!c || (count && !dev)) return fail(GG_ERR_INVALID, "gg_iota: NULL argument");
    if (!count) return GG_OK;
    uint32_t g = grid_for(count);
    switch (dtype) {
    case GG_U32: k_iota<uint32_t><<<g, kBlock, 0, c->stream>>>((uint32_t *)dev, count); break;
    case GG_I32: k_iota<int32_t><<<g, kBlock, 0, c->stream>>>((int32_t *)dev, count); break;
    case GG_F64: k_iota<double><<<g, kBlock, 0, c->stream>>>((double *)dev, count); break;
    case GG_F32: k_iota<float><<<g, kBlock, 0, c->stream>>>((float *)dev, count); break;
    default: return fail(GG_ERR_INVALID, "gg_iota: bad dtype");
    }
    GG_HIP(hipGetLastError());
    return sync(c);
}

// =============================================================================
// Graph construction
// =============================================================================

namespace gg {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// One raw R-MAT edge per thread; self loops become the sentinel key.
// key = (src << scale) | dst, sentinel = 1 << (2*scale).
__global__ void k_rmat_keys(int scale, uint64_t nraw, uint64_t seed, int symmetric,
                            uint64_t *__restrict__ keys)
{
    const uint32_t TA = 2448131359u, TAB = TA + 816043786u, TABC = TAB + 816043786u;
    uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t sentinel = 1ULL << (2 * scale);
    for (; e < nraw; e += stride) {
        uint32_t s = 0, d = 0;
        for (int l = scale - 1; l >= 0; l--) {
            uint32_t r = (uint32_t)(splitmix64((seed << 40) ^ (e << 6) ^ (uint64_t)l) >> 32);
            uint32_t sb = r >= TAB;
            uint32_t db = (r >= TA && r < TAB) || r >= TABC;
            s |= sb << l;
            d |= db << l;
        }
        keys[e] = (s == d) ? sentinel : (((uint64_t)s << scale) | d);
        if (symmetric) keys[nraw + e] = (s == d) ? sentinel : (((uint64_t)d << scale) | s);
    }
}

// Counts per high part of sorted keys without same-address atomics: the
// last element of a run adds its end position, the first subtracts its
// start, so cnt[h] ends as the run length (two atomics per distinct h; one
// atomicAdd per element serialised on the hubs' counters).
__device__ __forceinline__ void run_count(const uint64_t *__restrict__ keys, uint64_t m, int shift, uint64_t i,
                                          uint64_t h, uint32_t *__restrict__ cnt)
{
    const bool first = i == 0 || (keys[i - 1] >> shift) != h;
    const bool last = i + 1 == m || (keys[i + 1] >> shift) != h;
    if (first && last) cnt[h] = 1u;
    else if (last) atomicAdd(&cnt[h], (uint32_t)(i + 1));
    else if (first) atomicSub(&cnt[h], (uint32_t)i);
}

__global__ void k_decode_keys(const uint64_t *__restrict__ keys, uint64_t m, int shift,
                              uint32_t *__restrict__ ids, uint32_t *__restrict__ deg)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t mask = (1ULL << shift) - 1;
    for (; i < m; i += stride) {
        const uint64_t k = keys[i];
        ids[i] = (uint32_t)(k & mask);
        run_count(keys, m, shift, i, k >> shift, deg);
    }
}

__global__ void k_diff(const uint32_t *__restrict__ off, uint32_t *__restrict__ deg, uint32_t n)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) deg[v] = off[v + 1] - off[v];
}

// In-edge sort keys: (dst << bits) | src for CSR edge e, value e.
__global__ void k_csc_keys(const uint32_t *__restrict__ off, const uint32_t *__restrict__ ids,
                           uint32_t n, uint64_t m, int bits, uint64_t *__restrict__ keys,
                           uint32_t *__restrict__ vals)
{
    uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; e < m; e += stride) {
        // source = last v with off[v] <= e
        uint32_t lo = 0, hi = n;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (off[mid] <= e) lo = mid;
            else hi = mid;
        }
        keys[e] = ((uint64_t)ids[e] << bits) | lo;
        vals[e] = (uint32_t)e;
    }
}

__global__ void k_csc_decode(const uint64_t *__restrict__ keys, uint64_t m, int bits,
                             uint32_t *__restrict__ tids, uint32_t *__restrict__ indeg)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t mask = (1ULL << bits) - 1;
    for (; i < m; i += stride) {
        const uint64_t k = keys[i];
        tids[i] = (uint32_t)(k & mask);
        run_count(keys, m, bits, i, k >> bits, indeg);
    }
}

__global__ void k_gather_f32(const float *__restrict__ src, const uint32_t *__restrict__ idx,
                             float *__restrict__ dst, uint64_t m)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < m; i += stride) dst[i] = src[idx[i]];
}

__global__ void k_rmat_weights(uint64_t seed, float *__restrict__ w, uint64_t m)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < m; i += stride) {
        uint32_t k = (uint32_t)(splitmix64((seed << 40) ^ (i << 6) ^ 63u) >> 40);
        w[i] = (float)(k + 1u) * (1.0f / 16777216.0f);
    }
}

static int bits_for(uint64_t n)
{
    int b = 1;
    while ((1ULL << b) < n) b++;
    return b;
}

// deg[i] for i < n, 0 at i = n: input of the offsets scan
struct PadDeg {
    const uint32_t *deg;
    uint32_t n;
    __host__ __device__ uint32_t operator()(uint32_t i) const { return i < n ? deg[i] : 0u; }
};

// exclusive scan of deg[0..n) into off[0..n], off[n] = total
static int scan_offsets(Ctx *c, const uint32_t *deg, uint32_t *off, uint32_t n)
{
    size_t tb = 0;
    auto in = rocprim::make_transform_iterator(rocprim::make_counting_iterator<uint32_t>(0),
                                               PadDeg{deg, n});
    GG_HIP(rocprim::exclusive_scan(nullptr, tb, in, off, 0u, (size_t)n + 1,
                                   rocprim::plus<uint32_t>(), c->stream));
    GG_TRY(buf_ensure(c, c->tmp, tb + 16));
    GG_HIP(rocprim::exclusive_scan(c->tmp.p, tb, in, off, 0u, (size_t)n + 1,
                                   rocprim::plus<uint32_t>(), c->stream));
    return GG_OK;
}

// Builds outdeg, the in-edge twin (toff/tids/perm/tw/indeg) and the pull
// schedule from a graph whose CSR (off/ids[/w]) is already on the device.
static int build_twin(Graph *g)
{
    Ctx *c = g->ctx;
    uint32_t n = g->n;
    uint64_t m = g->m;
    GG_TRY(buf_alloc(c, g->outdeg, 4ull * n));
    GG_TRY(buf_alloc(c, g->indeg, 4ull * n));
    GG_TRY(buf_alloc(c, g->toff, 4ull * (n + 1)));
    GG_TRY(buf_alloc(c, g->tids, 4ull * (m ? m : 1)));
    GG_TRY(buf_alloc(c, g->perm, 4ull * (m ? m : 1)));
    if (n) k_diff<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>((uint32_t *)g->off.p,
                                                                    (uint32_t *)g->outdeg.p, n);
    GG_TRY(fill_u32(c, g->indeg.p, n, 0));
    if (m) {
        int bits = bits_for(n);
        Buf k0, k1, v1;
        GG_TRY(buf_alloc(c, k0, 8 * m));
        GG_TRY(buf_alloc(c, k1, 8 * m));
        GG_TRY(buf_alloc(c, v1, 4 * m));
        uint32_t grid = (uint32_t)std::min<uint64_t>((m + kBlock - 1) / kBlock, 1u << 20);
        k_csc_keys<<<grid, kBlock, 0, c->stream>>>((uint32_t *)g->off.p, (uint32_t *)g->ids.p, n,
                                                    m, bits, (uint64_t *)k0.p, (uint32_t *)v1.p);
        GG_HIP(hipGetLastError());
        size_t tb = 0;
        GG_HIP(rocprim::radix_sort_pairs(nullptr, tb, (uint64_t *)k0.p, (uint64_t *)k1.p,
                                         (uint32_t *)v1.p, (uint32_t *)g->perm.p, (size_t)m, 0,
                                         (unsigned)(2 * bits), c->stream));
        GG_TRY(buf_ensure(c, c->tmp, tb + 16));
        GG_HIP(rocprim::radix_sort_pairs(c->tmp.p, tb, (uint64_t *)k0.p, (uint64_t *)k1.p,
                                         (uint32_t *)v1.p, (uint32_t *)g->perm.p, (size_t)m, 0,
                                         (unsigned)(2 * bits), c->stream));
        k_csc_decode<<<grid, kBlock, 0, c->stream>>>((uint64_t *)k1.p, m, bits,
                                                      (uint32_t *)g->tids.p, (uint32_t *)g->indeg.p);
        GG_HIP(hipGetLastError());
        GG_TRY(sync(c));
        buf_free(k0);
        buf_free(k1);
        buf_free(v1);
        if (g->weighted) {
            GG_TRY(buf_alloc(c, g->tw, 4 * m));
            k_gather_f32<<<grid, kBlock, 0, c->stream>>>((float *)g->w.p, (uint32_t *)g->perm.p,
                                                          (float *)g->tw.p, m);
            GG_HIP(hipGetLastError());
        }
    }
    GG_TRY(scan_offsets(c, (uint32_t *)g->indeg.p, (uint32_t *)g->toff.p, n));
    GG_TRY(sched_build(c, g->pull, (uint32_t *)g->toff.p, n, m));
    return sync(c);
}

static void graph_free(Graph *g)
{
    Buf *bufs[] = {&g->off, &g->ids, &g->w, &g->toff, &g->tids, &g->tw, &g->perm, &g->outdeg,
                   &g->indeg, &g->comp, &g->coo_src, &g->coo_dst, &g->coo_w, &g->coo_tids, &g->coo_tw,
                   &g->first_in, &g->plan.first_in, &g->pull.start_row,
                   &g->pull.start_edge, &g->pull.carry_row, &g->pull.carry_val, &g->pull.head_val,
                   &g->plan.perm, &g->plan.inv, &g->plan.toff, &g->plan.tids, &g->plan.tw, &g->plan.outdeg,
                   &g->plan.coff, &g->plan.cids, &g->plan.cw, &g->plan.cdelta, &g->plan.indeg,
                   &g->plan.csched.start_row,
                   &g->plan.csched.start_edge, &g->plan.csched.carry_row, &g->plan.csched.carry_val,
                   &g->plan.csched.head_val,
                   &g->plan.sched.start_row, &g->plan.sched.start_edge, &g->plan.sched.carry_row,
                   &g->plan.sched.carry_val, &g->plan.sched.head_val, &g->csr_pull.start_row,
                   &g->csr_pull.start_edge, &g->csr_pull.carry_row, &g->csr_pull.carry_val,
                   &g->csr_pull.head_val, &g->wsum};
    for (Buf *b : bufs) buf_free(*b);
    hubs_free(g->hubs);
    hubs_free(g->plan.hubs);
    rows_free(g->plan.rows);
    rows_free(g->plan.crows);
    xseg_free(g->xseg);
    rows_free(g->rows);
    rows_free(g->csr_rows);
    g->rows_ok = false;
    g->csr_rows_ok = false;
    graph_exec_drop(g);
    for (Buf &b : g->pr_ws) buf_free(b);
    g->plan.mode = 0;
}

} // namespace gg

namespace gg {
int graph_from_device(Ctx *c, uint32_t n, uint64_t m, Buf &off, Buf &ids, Buf &w, bool symmetric, Graph **out)
{
    gg_graph_s *g = new gg_graph_s();
    g->ctx = c;
    g->n = n;
    g->m = m;
    g->symmetric = symmetric;
    g->weighted = w.p != nullptr;
    std::swap(g->off, off);
    std::swap(g->ids, ids);
    std::swap(g->w, w);
    int rc = build_twin(g);
    if (rc) {
        graph_free(g);
        delete g;
        return rc;
    }
    *out = g;
    return GG_OK;
}
} // namespace gg

extern "C" int gg_graph_create_csr(gg_ctx_t c, uint32_t n, uint64_t m, const uint32_t *off,
                                   const uint32_t *ids, const float *w, int symmetric,
                                   gg_graph_t *out)
{
    if (!c || !out || !off || (m && !ids)) return fail(GG_ERR_INVALID, "gg_graph_create_csr: NULL argument");
    if (n == 0) return fail(GG_ERR_INVALID, "gg_graph_create_csr: n must be > 0");
    if (m > 0xffffffffULL) return fail(GG_ERR_INVALID, "gg_graph_create_csr: m must be < 2^32");
    // the reference's loader checks header counts (IO.h:150-161); check the CSR itself
    if (off[0] != 0 || off[n] != m) return fail(GG_ERR_SIZE, "gg_graph_create_csr: offsets do not span [0, m]");
    for (uint32_t v = 0; v < n; v++)
        if (off[v + 1] < off[v]) return fail(GG_ERR_INVALID, "gg_graph_create_csr: offsets decrease at " + std::to_string(v));
    for (uint64_t e = 0; e < m; e++)
        if (ids[e] >= n) return fail(GG_ERR_INVALID, "gg_graph_create_csr: id out of range at edge " + std::to_string(e));
    GG_HIP(hipSetDevice(c->device));
    gg_graph_s *g = new gg_graph_s();
    g->ctx = c;
    g->n = n;
    g->m = m;
    g->symmetric = symmetric != 0;
    g->weighted = w != nullptr;
    int rc = buf_alloc(c, g->off, 4ull * (n + 1));
    if (!rc) rc = buf_alloc(c, g->ids, 4ull * (m ? m : 1));
    if (!rc && w) rc = buf_alloc(c, g->w, 4ull * (m ? m : 1));
    if (!rc && hipMemcpyAsync(g->off.p, off, 4ull * (n + 1), hipMemcpyHostToDevice, c->stream)) rc = fail(GG_ERR_HIP, "upload offsets");
    if (!rc && m && hipMemcpyAsync(g->ids.p, ids, 4ull * m, hipMemcpyHostToDevice, c->stream)) rc = fail(GG_ERR_HIP, "upload ids");
    if (!rc && w && m && hipMemcpyAsync(g->w.p, w, 4ull * m, hipMemcpyHostToDevice, c->stream)) rc = fail(GG_ERR_HIP, "upload weights");
    if (!rc) rc = build_twin(g);
    if (rc) {
        graph_free(g);
        delete g;
        return rc;
    }
    *out = g;
    return GG_OK;
}

namespace gg {
// out[i] = in[i] + add
__global__ void k_shift_u32(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, uint64_t count,
                            uint32_t add)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) out[i] = in[i] + add;
}
} // namespace gg

extern "C" int gg_graph_create_arena(gg_ctx_t c, int count, const gg_graph_t *parts, gg_graph_t *out,
                                     uint32_t *vbase, uint64_t *ebase)
{
    if (!c || !parts || !out || count < 1) return fail(GG_ERR_INVALID, "gg_graph_create_arena: bad argument");
    uint64_t n = 0, m = 0;
    bool weighted = true, symmetric = true;
    for (int i = 0; i < count; i++) {
        if (!parts[i] || parts[i]->ctx != c)
            return fail(GG_ERR_INVALID, "gg_graph_create_arena: graph " + std::to_string(i) + " missing or on another context");
        n += parts[i]->n;
        m += parts[i]->m;
        weighted = weighted && parts[i]->weighted;
        symmetric = symmetric && parts[i]->symmetric;
    }
    if (n >= 0xffffffffULL || m >= 0xffffffffULL)
        return fail(GG_ERR_SIZE, "gg_graph_create_arena: the arena needs n, m < 2^32");
    GG_HIP(hipSetDevice(c->device));
    gg_graph_s *g = new gg_graph_s();
    g->ctx = c;
    g->n = (uint32_t)n;
    g->m = m;
    g->symmetric = symmetric;
    g->weighted = weighted;
    int rc = buf_alloc(c, g->off, 4ull * (n + 1));
    if (!rc) rc = buf_alloc(c, g->ids, 4ull * (m ? m : 1));
    if (!rc && weighted) rc = buf_alloc(c, g->w, 4ull * (m ? m : 1));
    uint64_t vb = 0, eb = 0;
    for (int i = 0; i < count && !rc; i++) {
        const Graph *p = parts[i];
        if (vbase) vbase[i] = (uint32_t)vb;
        if (ebase) ebase[i] = eb;
        const uint32_t gv = blocks_for(std::max<uint64_t>(p->n, 1), kBlock);
        k_shift_u32<<<std::min<uint32_t>(gv, 1u << 16), kBlock, 0, c->stream>>>(
            (const uint32_t *)p->off.p, (uint32_t *)g->off.p + vb, p->n, (uint32_t)eb);
        if (p->m) {
            const uint32_t ge = blocks_for(p->m, kBlock);
            k_shift_u32<<<std::min<uint32_t>(ge, 1u << 16), kBlock, 0, c->stream>>>(
                (const uint32_t *)p->ids.p, (uint32_t *)g->ids.p + eb, p->m, (uint32_t)vb);
            if (weighted && hipMemcpyAsync((float *)g->w.p + eb, p->w.p, 4ull * p->m, hipMemcpyDeviceToDevice,
                                           c->stream) != hipSuccess)
                rc = fail(GG_ERR_HIP, "gg_graph_create_arena: weight copy");
        }
        vb += p->n;
        eb += p->m;
    }
    if (vbase) vbase[count] = (uint32_t)vb;
    if (ebase) ebase[count] = eb;
    if (!rc) rc = fill_u32(c, (uint32_t *)g->off.p + n, 1, (uint32_t)m);
    if (!rc && hipGetLastError() != hipSuccess) rc = fail(GG_ERR_HIP, "gg_graph_create_arena: launch");
    if (!rc) rc = build_twin(g);
    if (rc) {
        graph_free(g);
        delete g;
        return rc;
    }
    *out = g;
    return GG_OK;
}

extern "C" int gg_graph_create_rmat(gg_ctx_t c, int scale, int edge_factor, uint64_t seed,
                                    int symmetric, int weighted, gg_graph_t *out)
{
    if (!c || !out) return fail(GG_ERR_INVALID, "gg_graph_create_rmat: NULL argument");
    if (scale < 1 || scale > 30 || edge_factor < 1 || edge_factor > 1024)
        return fail(GG_ERR_INVALID, "gg_graph_create_rmat: scale in [1,30], edge_factor in [1,1024]");
    GG_HIP(hipSetDevice(c->device));
    uint64_t n = 1ULL << scale;
    uint64_t nraw = (uint64_t)edge_factor << scale;
    uint64_t nkeys = symmetric ? 2 * nraw : nraw;
    if (nkeys > 0xffffffffULL) return fail(GG_ERR_INVALID, "gg_graph_create_rmat: too many edges for uint32 offsets");
    gg_graph_s *g = new gg_graph_s();
    g->ctx = c;
    g->n = (uint32_t)n;
    g->symmetric = symmetric != 0;
    g->weighted = weighted != 0;
    Buf k0, k1, cnt;
    auto cleanup = [&](int rc) {
        buf_free(k0);
        buf_free(k1);
        buf_free(cnt);
        if (rc) {
            graph_free(g);
            delete g;
        }
        return rc;
    };
    int rc = buf_alloc(c, k0, 8 * nkeys);
    if (!rc) rc = buf_alloc(c, k1, 8 * nkeys);
    if (!rc) rc = buf_alloc(c, cnt, 8);
    if (rc) return cleanup(rc);
    uint32_t grid = (uint32_t)std::min<uint64_t>((nraw + kBlock - 1) / kBlock, 1u << 20);
    k_rmat_keys<<<grid, kBlock, 0, c->stream>>>(scale, nraw, seed, symmetric, (uint64_t *)k0.p);
    if (hipGetLastError() != hipSuccess) return cleanup(fail(GG_ERR_HIP, "k_rmat_keys launch"));
    size_t tb = 0;
    unsigned end_bit = (unsigned)(2 * scale + 1);
    if (rocprim::radix_sort_keys(nullptr, tb, (uint64_t *)k0.p, (uint64_t *)k1.p, (size_t)nkeys, 0,
                                 end_bit, c->stream) != hipSuccess)
        return cleanup(fail(GG_ERR_HIP, "radix_sort_keys size query"));
    if ((rc = buf_ensure(c, c->tmp, tb + 16))) return cleanup(rc);
    if (rocprim::radix_sort_keys(c->tmp.p, tb, (uint64_t *)k0.p, (uint64_t *)k1.p, (size_t)nkeys, 0,
                                 end_bit, c->stream) != hipSuccess)
        return cleanup(fail(GG_ERR_HIP, "radix_sort_keys"));
    // dedupe into k0
    tb = 0;
    if (rocprim::unique(nullptr, tb, (uint64_t *)k1.p, (uint64_t *)k0.p, (uint64_t *)cnt.p,
                        (size_t)nkeys, rocprim::equal_to<uint64_t>(), c->stream) != hipSuccess)
        return cleanup(fail(GG_ERR_HIP, "unique size query"));
    if ((rc = buf_ensure(c, c->tmp, tb + 16))) return cleanup(rc);
    if (rocprim::unique(c->tmp.p, tb, (uint64_t *)k1.p, (uint64_t *)k0.p, (uint64_t *)cnt.p,
                        (size_t)nkeys, rocprim::equal_to<uint64_t>(), c->stream) != hipSuccess)
        return cleanup(fail(GG_ERR_HIP, "unique"));
    uint64_t uniq = 0, last = 0;
    if (hipMemcpyAsync(&uniq, cnt.p, 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        return cleanup(fail(GG_ERR_HIP, "count readback"));
    if ((rc = sync(c))) return cleanup(rc);
    if (uniq &&
        hipMemcpyAsync(&last, (uint64_t *)k0.p + (uniq - 1), 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        return cleanup(fail(GG_ERR_HIP, "last key readback"));
    if ((rc = sync(c))) return cleanup(rc);
    uint64_t m = (uniq && last == (1ULL << (2 * scale))) ? uniq - 1 : uniq;
    g->m = m;
    Buf deg;
    rc = buf_alloc(c, g->off, 4 * (n + 1));
    if (!rc) rc = buf_alloc(c, g->ids, 4 * (m ? m : 1));
    if (!rc) rc = buf_alloc(c, deg, 4 * n);
    if (!rc) rc = fill_u32(c, deg.p, n, 0);
    if (rc) { buf_free(deg); return cleanup(rc); }
    if (m) {
        uint32_t g2 = (uint32_t)std::min<uint64_t>((m + kBlock - 1) / kBlock, 1u << 20);
        k_decode_keys<<<g2, kBlock, 0, c->stream>>>((uint64_t *)k0.p, m, scale, (uint32_t *)g->ids.p,
                                                     (uint32_t *)deg.p);
    }
    rc = scan_offsets(c, (uint32_t *)deg.p, (uint32_t *)g->off.p, (uint32_t)n);
    if (!rc) rc = sync(c);
    buf_free(deg);
    buf_free(k0);
    buf_free(k1);
    if (rc) return cleanup(rc);
    if (weighted) {
        rc = buf_alloc(c, g->w, 4 * (m ? m : 1));
        if (rc) return cleanup(rc);
        if (m) {
            uint32_t g2 = (uint32_t)std::min<uint64_t>((m + kBlock - 1) / kBlock, 1u << 20);
            k_rmat_weights<<<g2, kBlock, 0, c->stream>>>(seed, (float *)g->w.p, m);
        }
    }
    rc = build_twin(g);
    if (rc) return cleanup(rc);
    cleanup(0);
    *out = g;
    return GG_OK;
}

extern "C" int gg_graph_destroy(gg_graph_t g)
{
    if (!g) return fail(GG_ERR_INVALID, "gg_graph_destroy: NULL");
    (void)hipStreamSynchronize(g->ctx->stream);
    graph_free(g);
    delete g;
    return GG_OK;
}

extern "C" int gg_graph_info(gg_graph_t g, uint32_t *n, uint64_t *m, int *symmetric, int *weighted)
{
    if (!g) return fail(GG_ERR_INVALID, "gg_graph_info: NULL");
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (symmetric) *symmetric = g->symmetric;
    if (weighted) *weighted = g->weighted;
    return GG_OK;
}

static int copy_out(Ctx *c, void *host, const Buf &b, uint64_t bytes)
{
    if (!host || !bytes) return GG_OK;
    GG_HIP(hipMemcpyAsync(host, b.p, bytes, hipMemcpyDeviceToHost, c->stream));
    return GG_OK;
}

extern "C" int gg_graph_copy_csr(gg_graph_t g, uint32_t *off, uint32_t *ids, float *w)
{
    if (!g) return fail(GG_ERR_INVALID, "gg_graph_copy_csr: NULL");
    if (w && !g->weighted) return fail(GG_ERR_STATE, "gg_graph_copy_csr: graph has no weights");
    GG_TRY(copy_out(g->ctx, off, g->off, 4ull * (g->n + 1)));
    GG_TRY(copy_out(g->ctx, ids, g->ids, 4ull * g->m));
    if (w) GG_TRY(copy_out(g->ctx, w, g->w, 4ull * g->m));
    return sync(g->ctx);
}

extern "C" int gg_graph_copy_csc(gg_graph_t g, uint32_t *toff, uint32_t *tids, float *tw)
{
    if (!g) return fail(GG_ERR_INVALID, "gg_graph_copy_csc: NULL");
    if (tw && !g->weighted) return fail(GG_ERR_STATE, "gg_graph_copy_csc: graph has no weights");
    GG_TRY(copy_out(g->ctx, toff, g->toff, 4ull * (g->n + 1)));
    GG_TRY(copy_out(g->ctx, tids, g->tids, 4ull * g->m));
    if (tw) GG_TRY(copy_out(g->ctx, tw, g->tw, 4ull * g->m));
    return sync(g->ctx);
}

extern "C" int gg_graph_degrees(gg_graph_t g, int in_edges, uint32_t *out)
{
    if (!g || !out) return fail(GG_ERR_INVALID, "gg_graph_degrees: NULL");
    GG_TRY(copy_out(g->ctx, out, in_edges ? g->indeg : g->outdeg, 4ull * g->n));
    return sync(g->ctx);
}

namespace gg {
static int weights_changed(gg_graph_t g)
{
    Ctx *c = g->ctx;
    const uint64_t m = g->m;
    g->wsum_ok = false;
    graph_exec_drop(g);
    if (!m) return GG_OK;
    uint32_t grid = (uint32_t)std::min<uint64_t>((m + kBlock - 1) / kBlock, 1u << 20);
    k_gather_f32<<<grid, kBlock, 0, c->stream>>>((float *)g->w.p, (uint32_t *)g->perm.p, (float *)g->tw.p, m);
    GG_HIP(hipGetLastError());
    if (g->coo_w.p) {
        // COO weights follow the COO order: rebuild them by re-partitioning
        int rc = gg_graph_partition(g, g->P, g->part_vertex, g->relabel, g->hilbert);
        if (rc) return rc;
    }
    // the renumbered plan holds its own copies of the weights: rebuild it
    if (g->plan.mode) GG_TRY(gg_graph_reorder(g, g->plan.mode));
    return sync(c);
}
} // namespace gg

extern "C" int gg_graph_set_weights_dev(gg_graph_t g, const float *w_dev)
{
    if (!g || !w_dev) return fail(GG_ERR_INVALID, "gg_graph_set_weights_dev: NULL");
    Ctx *c = g->ctx;
    GG_HIP(hipSetDevice(c->device));
    const uint64_t m = g->m;
    if (!g->weighted) {
        GG_TRY(buf_alloc(c, g->w, 4 * (m ? m : 1)));
        GG_TRY(buf_alloc(c, g->tw, 4 * (m ? m : 1)));
        g->weighted = true;
    }
    if (m) GG_HIP(hipMemcpyAsync(g->w.p, w_dev, 4 * m, hipMemcpyDeviceToDevice, c->stream));
    return weights_changed(g);
}

extern "C" int gg_graph_set_weights(gg_graph_t g, const float *w_csr)
{
    if (!g || !w_csr) return fail(GG_ERR_INVALID, "gg_graph_set_weights: NULL");
    Ctx *c = g->ctx;
    uint64_t m = g->m;
    if (!g->weighted) {
        GG_TRY(buf_alloc(c, g->w, 4 * (m ? m : 1)));
        GG_TRY(buf_alloc(c, g->tw, 4 * (m ? m : 1)));
        g->weighted = true;
    }
    if (m) GG_HIP(hipMemcpyAsync(g->w.p, w_csr, 4 * m, hipMemcpyHostToDevice, c->stream));
    return weights_changed(g);
}

// =============================================================================
// Device partitioner (graph-numa.h:1177-1357) and COO build (1116-1143)
// =============================================================================

namespace gg {

// Edge-balanced greedy of partitionByDegree / cscpartitionByDegree on a
// prefix array pre[0..count] (pre[i+1]-pre[i] = degree of item i).  With
// E(s,i) = pre[i+1]-pre[s], a partition starting at s closes after
//   i1 = first i with E(s,i) >= avg                     (graph-numa.h:1337)
//   i2 = first i <= count-2 with E(s,i+1) > 1.1*avg,    (look-ahead, 1335)
//        when E(s,i2) < avg (then i2 < i1)
// whichever comes first; the look-ahead has no P guard, so firing it in the
// last partition is the reference's overrun.  The degree past the end
// (degrees[i+1] at i = count-1) is taken as 0.  One thread runs the P
// sequential cuts, each a pair of binary searches.
__global__ void k_partition_greedy(const uint32_t *__restrict__ pre, uint64_t count, uint64_t m,
                                   int P, uint64_t *__restrict__ starts, int *__restrict__ status)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int64_t avg = (int64_t)(m / (uint64_t)P);
    const double lim = 1.1 * (double)avg;
    uint64_t s = 0;
    int cuts = 0;
    int st = 0;
    starts[0] = 0;
    for (int c = 0; c < P; c++) {
        if (s >= count) { // nothing left: the remaining partitions are empty
            for (int k = c; k < P; k++) starts[k + 1] = count;
            break;
        }
        const uint64_t base = pre[s];
        // i2: first j in [s+2, count] with pre[j]-base > lim, i2 = j-2
        uint64_t lo = s + 2, hi = count + 1;
        while (lo < hi) {
            uint64_t mid = (lo + hi) >> 1;
            if ((double)(int64_t)((uint64_t)pre[mid] - base) > lim) hi = mid;
            else lo = mid + 1;
        }
        uint64_t i2 = lo <= count ? lo - 2 : count; // count = none
        bool fires = i2 < count && (int64_t)((uint64_t)pre[i2 + 1] - base) < avg;
        if (c == P - 1) {
            if (fires) st = -1; // greedy overruns its P partitions
            starts[P] = count;
            break;
        }
        // i1: first j in [s+1, count] with pre[j]-base >= avg, i1 = j-1
        lo = s + 1;
        hi = count + 1;
        while (lo < hi) {
            uint64_t mid = (lo + hi) >> 1;
            if ((int64_t)((uint64_t)pre[mid] - base) >= avg) hi = mid;
            else lo = mid + 1;
        }
        uint64_t i1 = lo <= count ? lo - 1 : count;
        uint64_t cut = fires ? i2 : i1;
        if (cut >= count) { // the rest stays in this partition
            starts[c + 1] = count;
            for (int k = c + 1; k < P; k++) starts[k + 1] = count;
            break;
        }
        cuts++;
        s = cut + 1;
        starts[c + 1] = s;
        if (s >= count) {
            for (int k = c + 1; k < P; k++) starts[k + 1] = count;
            break;
        }
    }
    // assert(counter+1 == P) at 1353: counter = number of increments
    if (st == 0 && cuts + 1 != P) st = 1;
    *status = st;
}

// Relabel ('-o') rule: cut after the first P-1 positions j with
// deg[j+1] > deg[j] (graph-numa.h:1188-1193, 1298-1303).
__global__ void k_partition_relabel(const uint32_t *__restrict__ pre, uint64_t count, int P,
                                    uint64_t *__restrict__ starts, int *__restrict__ status)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int c = 0;
    starts[0] = 0;
    for (uint64_t j = 0; j + 1 < count && c < P - 1; j++) {
        uint32_t dj = pre[j + 1] - pre[j], dn = pre[j + 2] - pre[j + 1];
        if (dn > dj) starts[++c] = j + 1;
    }
    int used = c;
    for (int k = c + 1; k <= P; k++) starts[k] = count;
    *status = (used + 1 != P) ? 1 : 0;
}

// Prefix of in-degrees over the compressed in-vertex list: cpre[k] =
// toff[comp[k]], cpre[nc] = m.
__global__ void k_compress_flags(const uint32_t *__restrict__ indeg, uint32_t n,
                                 uint32_t *__restrict__ flag)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) flag[v] = indeg[v] != 0;
}

__global__ void k_compress_scatter(const uint32_t *__restrict__ indeg, const uint32_t *__restrict__ pos,
                                   const uint32_t *__restrict__ toff, uint32_t n,
                                   uint32_t *__restrict__ comp, uint32_t *__restrict__ cpre)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n && indeg[v]) {
        comp[pos[v]] = v;
        cpre[pos[v]] = toff[v];
    }
}

__device__ __forceinline__ uint64_t hilbert_xy2d(int64_t side, int64_t x, int64_t y)
{
    // graph-numa.h:375-386 and rot 403-414, signed 64-bit as intT = long
    int64_t d = 0;
    for (int64_t s = side / 2; s > 0; s /= 2) {
        int64_t rx = (x & s) > 0;
        int64_t ry = (y & s) > 0;
        d += s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - x;
                y = s - 1 - y;
            }
            int64_t t = x;
            x = y;
            y = t;
        }
    }
    return (uint64_t)d;
}

// COO keys for the in-edges of destination partitions: edge k of the in-edge
// twin belongs to partition p (eoff[p] <= k < eoff[p+1]); its key is the
// Hilbert index of (src, dst) on the side x side grid with side =
// roundUpPow2(#edges of p) (HilbertEdgeSort(num_edges), graph-numa.h:358,487),
// or (src,dst) packed for CSR order (CSRDestSort, 338-350).
__global__ void k_coo_keys(const uint32_t *__restrict__ toff, const uint32_t *__restrict__ tids,
                           uint32_t n, uint64_t m, const uint64_t *__restrict__ eoff,
                           const int64_t *__restrict__ side, int P, int hilbert,
                           uint64_t *__restrict__ keys, uint32_t *__restrict__ vals,
                           uint32_t *__restrict__ edst)
{
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; k < m; k += stride) {
        uint32_t lo = 0, hi = n;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (toff[mid] <= k) lo = mid;
            else hi = mid;
        }
        uint32_t d = lo;
        int plo = 0, phi = P;
        while (phi - plo > 1) {
            int mid = (plo + phi) >> 1;
            if (eoff[mid] <= k) plo = mid;
            else phi = mid;
        }
        uint32_t s = tids[k];
        keys[k] = hilbert ? hilbert_xy2d(side[plo], s, d) : (((uint64_t)s << 32) | d);
        vals[k] = (uint32_t)k;
        edst[k] = d;
    }
}

__global__ void k_coo_emit(const uint32_t *__restrict__ order, const uint32_t *__restrict__ tids,
                           const uint32_t *__restrict__ edst, const float *__restrict__ tw,
                           uint64_t m, uint32_t *__restrict__ csrc, uint32_t *__restrict__ cdst,
                           float *__restrict__ cw)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < m; i += stride) {
        uint32_t k = order[i];
        csrc[i] = tids[k];
        cdst[i] = edst[k];
        if (cw) cw[i] = tw[k];
    }
}

static uint64_t round_up_pow2(uint64_t v)
{
    --v;
    for (unsigned sh = 1; sh < 64; sh <<= 1) v |= v >> sh;
    return ++v;
}

// Runs the partition sizing on a device prefix array.  Returns GG_OK,
// sets *underran; GG_ERR_PARTITION on overrun.
static int run_sizing(Ctx *c, const uint32_t *pre, uint64_t count, uint64_t m, int P, bool by_vertex,
                      bool relabel, std::vector<uint64_t> &starts, bool *underran)
{
    starts.assign((size_t)P + 1, 0);
    *underran = false;
    if (relabel) {
        Buf ds, st;
        GG_TRY(buf_alloc(c, ds, 8ull * (P + 1)));
        GG_TRY(buf_alloc(c, st, 8));
        k_partition_relabel<<<1, 64, 0, c->stream>>>(pre, count, P, (uint64_t *)ds.p, (int *)st.p);
        GG_HIP(hipGetLastError());
        int status = 0;
        GG_HIP(hipMemcpyAsync(starts.data(), ds.p, 8ull * (P + 1), hipMemcpyDeviceToHost, c->stream));
        GG_HIP(hipMemcpyAsync(&status, st.p, 4, hipMemcpyDeviceToHost, c->stream));
        GG_TRY(sync(c));
        buf_free(ds);
        buf_free(st);
        *underran = (status == 1) && !by_vertex;
        return GG_OK;
    }
    if (by_vertex) {
        // partitionByVertex / cscpartitionByVertex: count/P each, rest last
        uint64_t chunk = count / (uint64_t)P;
        for (int p = 0; p < P; p++) starts[p] = chunk * p;
        starts[P] = count;
        return GG_OK;
    }
    Buf ds, st;
    GG_TRY(buf_alloc(c, ds, 8ull * (P + 1)));
    GG_TRY(buf_alloc(c, st, 8));
    k_partition_greedy<<<1, 64, 0, c->stream>>>(pre, count, m, P, (uint64_t *)ds.p, (int *)st.p);
    GG_HIP(hipGetLastError());
    int status = 0;
    GG_HIP(hipMemcpyAsync(starts.data(), ds.p, 8ull * (P + 1), hipMemcpyDeviceToHost, c->stream));
    GG_HIP(hipMemcpyAsync(&status, st.p, 4, hipMemcpyDeviceToHost, c->stream));
    GG_TRY(sync(c));
    buf_free(ds);
    buf_free(st);
    if (status < 0)
        return fail(GG_ERR_PARTITION, "partition greedy overruns " + std::to_string(P) +
                                          " partitions (graph-numa.h:1335 has no bound check)");
    *underran = status == 1;
    return GG_OK;
}

} // namespace gg

extern "C" int gg_graph_partition(gg_graph_t g, int P, int by_vertex, int relabel, int hilbert)
{
    if (!g) return fail(GG_ERR_INVALID, "gg_graph_partition: NULL");
    if (P < 1 || (uint64_t)P > g->n) return fail(GG_ERR_INVALID, "gg_graph_partition: need 1 <= P <= n");
    Ctx *c = g->ctx;
    GG_HIP(hipSetDevice(c->device));
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    // COO partitioner over all vertices, in-degree (graph-numa.h:651-655)
    std::vector<uint64_t> coo, csc;
    bool u1 = false, u2 = false;
    GG_TRY(run_sizing(c, (uint32_t *)g->toff.p, n, m, P, by_vertex, relabel, coo, &u1));
    // compressed in-vertex list and its prefix (PartitionByDest 1085-1107)
    Buf flag, pos, cpre;
    GG_TRY(buf_alloc(c, flag, 4ull * n));
    GG_TRY(buf_alloc(c, pos, 4ull * (n + 1)));
    k_compress_flags<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>((uint32_t *)g->indeg.p, n,
                                                                       (uint32_t *)flag.p);
    GG_TRY(scan_offsets(c, (uint32_t *)flag.p, (uint32_t *)pos.p, n));
    uint32_t nc = 0;
    GG_HIP(hipMemcpyAsync(&nc, (uint32_t *)pos.p + n, 4, hipMemcpyDeviceToHost, c->stream));
    GG_TRY(sync(c));
    g->ncomp = nc;
    GG_TRY(buf_alloc(c, g->comp, 4ull * (nc ? nc : 1)));
    GG_TRY(buf_alloc(c, cpre, 4ull * (nc + 1)));
    k_compress_scatter<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>(
        (uint32_t *)g->indeg.p, (uint32_t *)pos.p, (uint32_t *)g->toff.p, n, (uint32_t *)g->comp.p,
        (uint32_t *)cpre.p);
    uint32_t mm = (uint32_t)m;
    GG_HIP(hipMemcpyAsync((uint32_t *)cpre.p + nc, &mm, 4, hipMemcpyHostToDevice, c->stream));
    GG_TRY(sync(c));
    int rc = run_sizing(c, (uint32_t *)cpre.p, nc, m, P, by_vertex, relabel, csc, &u2);
    buf_free(flag);
    buf_free(pos);
    buf_free(cpre);
    if (rc) return rc;
    g->P = P;
    g->part_vertex = by_vertex != 0;
    g->relabel = relabel != 0;
    g->hilbert = hilbert != 0;
    g->coo_starts = coo;
    g->csc_starts = csc;
    g->coo_underran = u1;
    g->csc_underran = u2;
    // COO lists (edge mode only, graph-numa.h:661-675)
    buf_free(g->coo_src);
    buf_free(g->coo_dst);
    buf_free(g->coo_w);
    buf_free(g->coo_tids);
    buf_free(g->coo_tw);
    g->coo_pull_ok = false;
    g->coo_eoff.assign((size_t)P + 1, 0);
    if (by_vertex || m == 0) return GG_OK;
    std::vector<uint32_t> toff_h((size_t)n + 1);
    GG_HIP(hipMemcpyAsync(toff_h.data(), g->toff.p, 4ull * (n + 1), hipMemcpyDeviceToHost, c->stream));
    GG_TRY(sync(c));
    std::vector<int64_t> side((size_t)P);
    std::vector<uint32_t> seg((size_t)P + 1);
    int maxbits = 1;
    for (int p = 0; p <= P; p++) {
        g->coo_eoff[p] = toff_h[coo[p]];
        seg[p] = (uint32_t)g->coo_eoff[p];
    }
    for (int p = 0; p < P; p++) {
        uint64_t cnt = g->coo_eoff[p + 1] - g->coo_eoff[p];
        side[p] = (int64_t)round_up_pow2(cnt);
        int b = 0;
        while (b < 62 && (1ULL << b) < (uint64_t)side[p]) b++;
        maxbits = std::max(maxbits, 2 * b);
    }
    Buf deoff, dside, dseg, k0, k1, v0, v1, edst;
    auto cleanup = [&]() {
        Buf *bs[] = {&deoff, &dside, &dseg, &k0, &k1, &v0, &v1, &edst};
        for (Buf *b : bs) buf_free(*b);
    };
    rc = buf_alloc(c, deoff, 8ull * (P + 1));
    if (!rc) rc = buf_alloc(c, dside, 8ull * P);
    if (!rc) rc = buf_alloc(c, dseg, 4ull * (P + 1));
    if (!rc) rc = buf_alloc(c, k0, 8 * m);
    if (!rc) rc = buf_alloc(c, k1, 8 * m);
    if (!rc) rc = buf_alloc(c, v0, 4 * m);
    if (!rc) rc = buf_alloc(c, v1, 4 * m);
    if (!rc) rc = buf_alloc(c, edst, 4 * m);
    if (!rc) rc = buf_alloc(c, g->coo_src, 4 * m);
    if (!rc) rc = buf_alloc(c, g->coo_dst, 4 * m);
    if (!rc && g->weighted) rc = buf_alloc(c, g->coo_w, 4 * m);
    if (rc) { cleanup(); return rc; }
    if (hipMemcpyAsync(deoff.p, g->coo_eoff.data(), 8ull * (P + 1), hipMemcpyHostToDevice, c->stream) ||
        hipMemcpyAsync(dside.p, side.data(), 8ull * P, hipMemcpyHostToDevice, c->stream) ||
        hipMemcpyAsync(dseg.p, seg.data(), 4ull * (P + 1), hipMemcpyHostToDevice, c->stream)) {
        cleanup();
        return fail(GG_ERR_HIP, "partition upload");
    }
    uint32_t grid = (uint32_t)std::min<uint64_t>((m + kBlock - 1) / kBlock, 1u << 20);
    k_coo_keys<<<grid, kBlock, 0, c->stream>>>((uint32_t *)g->toff.p, (uint32_t *)g->tids.p, n, m,
                                                (uint64_t *)deoff.p, (int64_t *)dside.p, P, hilbert,
                                                (uint64_t *)k0.p, (uint32_t *)v0.p, (uint32_t *)edst.p);
    size_t tb = 0;
    unsigned end_bit = hilbert ? (unsigned)std::min(64, maxbits) : 64u;
    hipError_t e = rocprim::segmented_radix_sort_pairs(
        nullptr, tb, (uint64_t *)k0.p, (uint64_t *)k1.p, (uint32_t *)v0.p, (uint32_t *)v1.p,
        (unsigned)m, (unsigned)P, (uint32_t *)dseg.p, (uint32_t *)dseg.p + 1, 0, end_bit, c->stream);
    if (e == hipSuccess) rc = buf_ensure(c, c->tmp, tb + 16);
    if (e == hipSuccess && !rc)
        e = rocprim::segmented_radix_sort_pairs(c->tmp.p, tb, (uint64_t *)k0.p, (uint64_t *)k1.p,
                                                (uint32_t *)v0.p, (uint32_t *)v1.p, (unsigned)m,
                                                (unsigned)P, (uint32_t *)dseg.p,
                                                (uint32_t *)dseg.p + 1, 0, end_bit, c->stream);
    if (e != hipSuccess || rc) {
        cleanup();
        return rc ? rc : fail(GG_ERR_HIP, std::string("segmented sort: ") + hipGetErrorString(e));
    }
    k_coo_emit<<<grid, kBlock, 0, c->stream>>>((uint32_t *)v1.p, (uint32_t *)g->tids.p,
                                                (uint32_t *)edst.p, (float *)g->tw.p, m,
                                                (uint32_t *)g->coo_src.p, (uint32_t *)g->coo_dst.p,
                                                (float *)g->coo_w.p);
    e = hipGetLastError();
    rc = sync(c);
    cleanup();
    if (e != hipSuccess) return fail(GG_ERR_HIP, "k_coo_emit");
    return rc;
}

namespace gg {
__global__ void k_coo_rows(const uint32_t *__restrict__ order, const uint32_t *__restrict__ csrc,
                           const float *__restrict__ cw, uint64_t m, uint32_t *__restrict__ tids,
                           float *__restrict__ tw)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < m; i += stride) {
        const uint32_t k = order[i];
        tids[i] = csrc[k];
        if (cw) tw[i] = cw[k];
    }
}

// The COO lists regrouped by destination: a stable radix sort of the COO
// positions by destination keeps each destination's edges in COO order,
// and the rows line up with the twin's offsets (a row's length is its
// in-degree either way).
int ensure_coo_pull(Graph *g)
{
    if (g->coo_pull_ok) return GG_OK;
    if (!g->coo_src.p || !g->m) return fail(GG_ERR_STATE, "COO traversal: graph has no COO partitions");
    Ctx *c = g->ctx;
    const uint64_t m = g->m;
    unsigned bits = 1;
    while (bits < 32 && (1ull << bits) < (uint64_t)g->n) bits++;
    Buf k0, k1, v0, v1;
    auto cleanup = [&]() {
        Buf *bs[] = {&k0, &k1, &v0, &v1};
        for (Buf *b : bs) buf_free(*b);
    };
    int rc = buf_alloc(c, k0, 4 * m);
    if (!rc) rc = buf_alloc(c, k1, 4 * m);
    if (!rc) rc = buf_alloc(c, v0, 4 * m);
    if (!rc) rc = buf_alloc(c, v1, 4 * m);
    if (!rc) rc = buf_alloc(c, g->coo_tids, 4 * m);
    if (!rc && g->coo_w.p) rc = buf_alloc(c, g->coo_tw, 4 * m);
    if (rc) { cleanup(); return rc; }
    if (hipMemcpyAsync(k0.p, g->coo_dst.p, 4 * m, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
        cleanup();
        return fail(GG_ERR_HIP, "COO traversal: key copy");
    }
    if ((rc = gg_iota(abi(c), v0.p, GG_U32, m))) { cleanup(); return rc; }
    size_t tb = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, tb, (uint32_t *)k0.p, (uint32_t *)k1.p, (uint32_t *)v0.p,
                                             (uint32_t *)v1.p, (size_t)m, 0, bits, c->stream);
    if (e == hipSuccess) rc = buf_ensure(c, c->tmp, tb + 16);
    if (e == hipSuccess && !rc)
        e = rocprim::radix_sort_pairs(c->tmp.p, tb, (uint32_t *)k0.p, (uint32_t *)k1.p, (uint32_t *)v0.p,
                                      (uint32_t *)v1.p, (size_t)m, 0, bits, c->stream);
    if (e != hipSuccess || rc) {
        cleanup();
        return rc ? rc : fail(GG_ERR_HIP, std::string("COO traversal sort: ") + hipGetErrorString(e));
    }
    const uint32_t grid = (uint32_t)std::min<uint64_t>((m + kBlock - 1) / kBlock, 1u << 20);
    k_coo_rows<<<grid, kBlock, 0, c->stream>>>((const uint32_t *)v1.p, (const uint32_t *)g->coo_src.p,
                                                (const float *)g->coo_w.p, m, (uint32_t *)g->coo_tids.p,
                                                (float *)g->coo_tw.p);
    e = hipGetLastError();
    rc = sync(c);
    cleanup();
    if (e != hipSuccess) return fail(GG_ERR_HIP, "k_coo_rows");
    if (!rc) g->coo_pull_ok = true;
    return rc;
}
} // namespace gg

extern "C" int gg_graph_partition_starts(gg_graph_t g, int which, uint64_t *starts, int *underran)
{
    if (!g || !starts) return fail(GG_ERR_INVALID, "gg_graph_partition_starts: NULL");
    if (!g->P) return fail(GG_ERR_STATE, "gg_graph_partition_starts: graph not partitioned");
    const std::vector<uint64_t> &s = which == 0 ? g->coo_starts : g->csc_starts;
    std::memcpy(starts, s.data(), 8 * s.size());
    if (underran) *underran = which == 0 ? g->coo_underran : g->csc_underran;
    return GG_OK;
}

extern "C" int gg_graph_coo(gg_graph_t g, uint32_t *src, uint32_t *dst, uint64_t *part_off)
{
    if (!g) return fail(GG_ERR_INVALID, "gg_graph_coo: NULL");
    if (!g->P) return fail(GG_ERR_STATE, "gg_graph_coo: graph not partitioned");
    if (g->part_vertex) return fail(GG_ERR_STATE, "gg_graph_coo: '-v vertex' builds no COO lists");
    if (part_off) std::memcpy(part_off, g->coo_eoff.data(), 8 * g->coo_eoff.size());
    GG_TRY(copy_out(g->ctx, src, g->coo_src, 4 * g->m));
    GG_TRY(copy_out(g->ctx, dst, g->coo_dst, 4 * g->m));
    return sync(g->ctx);
}

// =============================================================================
// Relabelled pull plan (hot sources first)
// =============================================================================

namespace gg {

// Sort key of vertex v: out-degree descending, id ascending.
__global__ void k_order_keys(const uint32_t *__restrict__ outdeg, uint32_t n, uint64_t *__restrict__ keys)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) keys[v] = ((uint64_t)(0xffffffffu - outdeg[v]) << 32) | v;
}

__global__ void k_order_emit(const uint64_t *__restrict__ keys, uint32_t n, const uint32_t *__restrict__ outdeg,
                             uint32_t *__restrict__ inv, uint32_t *__restrict__ perm,
                             uint32_t *__restrict__ outdeg2)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t v = (uint32_t)keys[k];
    inv[k] = v;
    perm[v] = k;
    outdeg2[k] = outdeg[v];
}

// Relabelled in-edge key of CSR edge e: (perm[dst] << bits) | perm[src].
__global__ void k_plan_keys(const uint32_t *__restrict__ off, const uint32_t *__restrict__ ids,
                            const uint32_t *__restrict__ perm, uint32_t n, uint64_t m, int bits,
                            uint64_t *__restrict__ keys, uint32_t *__restrict__ vals)
{
    uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; e < m; e += stride) {
        uint32_t lo = 0, hi = n;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (off[mid] <= e) lo = mid;
            else hi = mid;
        }
        keys[e] = ((uint64_t)perm[ids[e]] << bits) | perm[lo];
        vals[e] = (uint32_t)e;
    }
}

} // namespace gg

extern "C" int gg_graph_reorder(gg_graph_t g, int mode)
{
    if (!g) return fail(GG_ERR_INVALID, "gg_graph_reorder: NULL");
    if (mode != 0 && mode != 1) return fail(GG_ERR_INVALID, "gg_graph_reorder: mode must be 0 or 1");
    Ctx *c = g->ctx;
    GG_HIP(hipSetDevice(c->device));
    graph_exec_drop(g); // captured loops may hold the plan's arrays
    PullPlan &P = g->plan;
    Buf *pb[] = {&P.perm, &P.inv, &P.toff, &P.tids, &P.tw, &P.outdeg, &P.sched.start_row, &P.sched.start_edge,
                 &P.sched.carry_row, &P.sched.carry_val, &P.sched.head_val, &P.coff, &P.cids, &P.cw, &P.cdelta, &P.indeg,
                 &P.first_in, &P.csched.start_row, &P.csched.start_edge, &P.csched.carry_row, &P.csched.carry_val,
                 &P.csched.head_val};
    for (Buf *b : pb) buf_free(*b);
    hubs_free(P.hubs);
    rows_free(P.rows);
    rows_free(P.crows);
    xseg_free(g->xseg);
    P.mode = 0;
    if (mode == 0) return GG_OK;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    GG_TRY(buf_alloc(c, P.perm, 4ull * n));
    GG_TRY(buf_alloc(c, P.inv, 4ull * n));
    GG_TRY(buf_alloc(c, P.outdeg, 4ull * n));
    GG_TRY(buf_alloc(c, P.toff, 4ull * (n + 1)));
    GG_TRY(buf_alloc(c, P.tids, 4ull * (m ? m : 1)));
    {
        Buf k0, k1;
        GG_TRY(buf_alloc(c, k0, 8ull * n));
        GG_TRY(buf_alloc(c, k1, 8ull * n));
        k_order_keys<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>((const uint32_t *)g->outdeg.p, n,
                                                                       (uint64_t *)k0.p);
        size_t tb = 0;
        GG_HIP(rocprim::radix_sort_keys(nullptr, tb, (uint64_t *)k0.p, (uint64_t *)k1.p, (size_t)n, 0, 64,
                                        c->stream));
        GG_TRY(buf_ensure(c, c->tmp, tb + 16));
        GG_HIP(rocprim::radix_sort_keys(c->tmp.p, tb, (uint64_t *)k0.p, (uint64_t *)k1.p, (size_t)n, 0, 64,
                                        c->stream));
        k_order_emit<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>((const uint64_t *)k1.p, n,
                                                                       (const uint32_t *)g->outdeg.p,
                                                                       (uint32_t *)P.inv.p, (uint32_t *)P.perm.p,
                                                                       (uint32_t *)P.outdeg.p);
        GG_HIP(hipGetLastError());
        GG_TRY(sync(c));
        buf_free(k0);
        buf_free(k1);
    }
    Buf deg;
    GG_TRY(buf_alloc(c, deg, 4ull * n));
    GG_TRY(fill_u32(c, deg.p, n, 0));
    if (m) {
        int bits = bits_for(n);
        Buf k0, k1, v0, v1;
        GG_TRY(buf_alloc(c, k0, 8 * m));
        GG_TRY(buf_alloc(c, k1, 8 * m));
        GG_TRY(buf_alloc(c, v0, 4 * m));
        GG_TRY(buf_alloc(c, v1, 4 * m));
        uint32_t grid = (uint32_t)std::min<uint64_t>((m + kBlock - 1) / kBlock, 1u << 20);
        k_plan_keys<<<grid, kBlock, 0, c->stream>>>((const uint32_t *)g->off.p, (const uint32_t *)g->ids.p,
                                                     (const uint32_t *)P.perm.p, n, m, bits, (uint64_t *)k0.p,
                                                     (uint32_t *)v0.p);
        size_t tb = 0;
        GG_HIP(rocprim::radix_sort_pairs(nullptr, tb, (uint64_t *)k0.p, (uint64_t *)k1.p, (uint32_t *)v0.p,
                                         (uint32_t *)v1.p, (size_t)m, 0, (unsigned)(2 * bits), c->stream));
        GG_TRY(buf_ensure(c, c->tmp, tb + 16));
        GG_HIP(rocprim::radix_sort_pairs(c->tmp.p, tb, (uint64_t *)k0.p, (uint64_t *)k1.p, (uint32_t *)v0.p,
                                         (uint32_t *)v1.p, (size_t)m, 0, (unsigned)(2 * bits), c->stream));
        k_csc_decode<<<grid, kBlock, 0, c->stream>>>((const uint64_t *)k1.p, m, bits, (uint32_t *)P.tids.p,
                                                      (uint32_t *)deg.p);
        if (g->weighted) {
            GG_TRY(buf_alloc(c, P.tw, 4 * m));
            k_gather_f32<<<grid, kBlock, 0, c->stream>>>((const float *)g->w.p, (const uint32_t *)v1.p,
                                                          (float *)P.tw.p, m);
        }
        GG_HIP(hipGetLastError());
        GG_TRY(sync(c));
        buf_free(k0);
        buf_free(k1);
        buf_free(v0);
        buf_free(v1);
    }
    GG_TRY(scan_offsets(c, (const uint32_t *)deg.p, (uint32_t *)P.toff.p, n));
    GG_TRY(sched_build(c, P.sched, (const uint32_t *)P.toff.p, n, m));
    GG_TRY(sync(c));
    buf_free(deg);
    GG_TRY(rows_build(c, P.rows, (const uint32_t *)P.toff.p, n, m, g->edges_per_lane));
    P.mode = mode;
    return GG_OK;
}

namespace gg {
// Edge e of source v (found by binary search in off) moves to
// coff[perm[v]] + (e - off[v]) with its head renamed and its weight.
__global__ void k_plan_csr_edges(const uint32_t *__restrict__ off, const uint32_t *__restrict__ ids,
                                 const float *__restrict__ w, const uint32_t *__restrict__ perm,
                                 const uint32_t *__restrict__ coff, uint32_t n, uint64_t m,
                                 uint32_t *__restrict__ cids, float *__restrict__ cw)
{
    uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; e < m; e += stride) {
        uint32_t lo = 0, hi = n; // largest v with off[v] <= e
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (off[mid] <= e) lo = mid;
            else hi = mid;
        }
        const uint32_t to = coff[perm[lo]] + (uint32_t)(e - off[lo]);
        cids[to] = perm[ids[e]];
        if (w) cw[to] = w[e];
    }
}

// cdelta[k] = off[inv[k]] - coff[k] (mod 2^32)
__global__ void k_plan_csr_delta(const uint32_t *__restrict__ off, const uint32_t *__restrict__ inv,
                                 const uint32_t *__restrict__ coff, uint32_t n, uint32_t *__restrict__ cdelta)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) cdelta[k] = off[inv[k]] - coff[k];
}

int plan_csr_view(Graph *g, CscView *v)
{
    PullPlan &P = g->plan;
    if (!P.mode) return fail(GG_ERR_STATE, "no renumbered plan");
    Ctx *c = g->ctx;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    if (!P.coff.p) {
        Buf coff, cids, cw, cdelta, indeg;
        int rc = buf_alloc(c, coff, 4ull * (n + 1));
        if (!rc) rc = buf_alloc(c, cids, 4 * (m ? m : 1));
        if (!rc && g->weighted) rc = buf_alloc(c, cw, 4 * (m ? m : 1));
        if (!rc) rc = buf_alloc(c, cdelta, 4ull * n);
        if (!rc) rc = buf_alloc(c, indeg, 4ull * n);
        if (!rc) rc = scan_offsets(c, (const uint32_t *)P.outdeg.p, (uint32_t *)coff.p, n);
        if (!rc && n)
            k_diff<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>((const uint32_t *)P.toff.p,
                                                                     (uint32_t *)indeg.p, n);
        if (!rc && m) {
            const uint32_t grid = (uint32_t)std::min<uint64_t>((m + kBlock - 1) / kBlock, 1u << 20);
            k_plan_csr_edges<<<grid, kBlock, 0, c->stream>>>(
                (const uint32_t *)g->off.p, (const uint32_t *)g->ids.p, (const float *)(g->weighted ? g->w.p : nullptr),
                (const uint32_t *)P.perm.p, (const uint32_t *)coff.p, n, m, (uint32_t *)cids.p, (float *)cw.p);
        }
        if (!rc && n)
            k_plan_csr_delta<<<blocks_for(n, kBlock), kBlock, 0, c->stream>>>(
                (const uint32_t *)g->off.p, (const uint32_t *)P.inv.p, (const uint32_t *)coff.p, n,
                (uint32_t *)cdelta.p);
        if (!rc && hipGetLastError() != hipSuccess) rc = fail(GG_ERR_HIP, "plan CSR kernels failed to launch");
        if (!rc) rc = sched_build(c, P.csched, (const uint32_t *)coff.p, n, m);
        if (!rc) rc = sync(c);
        if (rc) {
            Buf *bs[] = {&coff, &cids, &cw, &cdelta, &indeg};
            for (Buf *b : bs) buf_free(*b);
            return rc;
        }
        P.coff = coff;
        P.cids = cids;
        P.cw = cw;
        P.cdelta = cdelta;
        P.indeg = indeg;
    }
    if (g->pull_schedule == 1 && !P.crows.blocks && n)
        GG_TRY(rows_build(c, P.crows, (const uint32_t *)P.coff.p, n, m, g->edges_per_lane));
    *v = CscView{};
    v->toff = (const uint32_t *)P.coff.p;
    v->tids = (const uint32_t *)P.cids.p;
    v->sched = &P.csched;
    if (g->pull_schedule == 1) v->rows = &P.crows;
    return GG_OK;
}
// The graph in its plan's numbering, sharing the plan's arrays: the
// relabelled in-edge twin with its schedules and the relabelled out-edge
// CSR (built on first use), degrees in the new numbering, the same
// partition settings, weights when the graph has them.  It owns nothing:
// release it with plan_alias_release.
int plan_alias(Graph *g, Graph &t)
{
    CscView cv;
    GG_TRY(plan_csr_view(g, &cv));
    PullPlan &P = g->plan;
    if (g->pull_schedule == 1 && !P.rows.blocks) return fail(GG_ERR_STATE, "plan has no row-block schedule");
    t.ctx = g->ctx;
    t.n = g->n;
    t.m = g->m;
    t.symmetric = g->symmetric;
    t.off = P.coff;
    t.ids = P.cids;
    t.toff = P.toff;
    t.tids = P.tids;
    t.outdeg = P.outdeg;
    t.indeg = P.indeg;
    t.weighted = g->weighted;
    t.w = P.cw;
    t.tw = P.tw;
    t.pull = P.sched;
    t.rows = P.rows;
    t.rows_ok = P.rows.blocks != 0;
    t.csr_pull = P.csched;
    t.csr_pull_ok = true;
    t.csr_rows = P.crows;
    t.csr_rows_ok = P.crows.blocks != 0;
    t.P = g->P;
    t.part_vertex = g->part_vertex;
    t.pull_schedule = g->pull_schedule;
    t.edges_per_lane = g->edges_per_lane;
    return GG_OK;
}

void plan_alias_release(Graph &t)
{
    Buf *bs[] = {&t.off, &t.ids, &t.w, &t.toff, &t.tids, &t.tw, &t.outdeg, &t.indeg};
    for (Buf *b : bs) *b = Buf{};
    t.pull = PullSched{};
    t.csr_pull = PullSched{};
    t.rows = RowSched{};
    t.csr_rows = RowSched{};
}

} // namespace gg

namespace gg {
// inv[perm[k]] = k
__global__ void k_invert_perm(const uint32_t *__restrict__ perm, uint32_t *__restrict__ inv, uint64_t m)
{
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; k < m; k += stride) inv[perm[k]] = (uint32_t)k;
}
} // namespace gg

extern "C" int gg_graph_transpose(gg_graph_t g)
{
    if (!g) return fail(GG_ERR_INVALID, "gg_graph_transpose: NULL");
    // a symmetric graph is its own transpose (graph-numa.h:547-555 skips it)
    if (g->symmetric) return GG_OK;
    Ctx *c = g->ctx;
    GG_HIP(hipSetDevice(c->device));
    // both directions' schedules exist before the swap
    if (!g->csr_pull_ok) {
        GG_TRY(sched_build(c, g->csr_pull, (const uint32_t *)g->off.p, g->n, g->m));
        g->csr_pull_ok = true;
    }
    Buf inv;
    if (g->m) {
        GG_TRY(buf_alloc(c, inv, 4ull * g->m));
        const uint32_t grid = (uint32_t)std::min<uint64_t>((g->m + kBlock - 1) / kBlock, 1u << 20);
        k_invert_perm<<<grid, kBlock, 0, c->stream>>>((const uint32_t *)g->perm.p, (uint32_t *)inv.p, g->m);
        GG_HIP(hipGetLastError());
        GG_TRY(sync(c));
        std::swap(g->perm, inv);
        buf_free(inv);
    }
    std::swap(g->off, g->toff);
    std::swap(g->ids, g->tids);
    std::swap(g->w, g->tw);
    std::swap(g->outdeg, g->indeg);
    std::swap(g->pull, g->csr_pull);
    std::swap(g->rows, g->csr_rows);
    std::swap(g->rows_ok, g->csr_rows_ok);
    g->csr_pull_ok = true;
    // state tied to one direction
    g->wsum_ok = false;
    buf_free(g->first_in);
    hubs_free(g->hubs);
    graph_exec_drop(g);
    // destination partitions and COO lists follow the new in-edges
    if (g->P) GG_TRY(gg_graph_partition(g, g->P, g->part_vertex, g->relabel, g->hilbert));
    if (g->plan.mode) GG_TRY(gg_graph_reorder(g, 0));
    g->transposed = !g->transposed;
    return GG_OK;
}

extern "C" int gg_graph_set_option(gg_graph_t g, int option, int64_t value)
{
    if (!g) return fail(GG_ERR_INVALID, "gg_graph_set_option: NULL");
    GG_HIP(hipSetDevice(g->ctx->device));
    switch (option) {
    case GG_OPT_PULL_SCHEDULE:
        if (value != 0 && value != 1) return fail(GG_ERR_INVALID, "gg_graph_set_option: schedule must be 0 or 1");
        if (g->pull_schedule != (int)value) graph_exec_drop(g);
        g->pull_schedule = (int)value;
        return GG_OK;
    case GG_OPT_ROW_EDGES_PER_LANE:
        if (value < 1 || value > kRowEdges)
            return fail(GG_ERR_INVALID, "gg_graph_set_option: edges per lane out of range");
        if (g->edges_per_lane != (int)value) {
            graph_exec_drop(g);
            rows_free(g->rows);
            rows_free(g->csr_rows);
            rows_free(g->plan.crows); // rebuilt on first use
            xseg_free(g->xseg);       // likewise
            g->rows_ok = false;
            g->csr_rows_ok = false;
        }
        g->edges_per_lane = (int)value;
        if (g->plan.mode)
            return rows_build(g->ctx, g->plan.rows, (const uint32_t *)g->plan.toff.p, g->n, g->m,
                              g->edges_per_lane);
        return GG_OK;
    case GG_OPT_DENSE_COO:
        if (value != 0 && value != 1) return fail(GG_ERR_INVALID, "gg_graph_set_option: dense COO must be 0 or 1");
        g->dense_coo = (int)value;
        return GG_OK;
    case GG_OPT_XCD_SEGMENTS:
        if (value < 0 || value > 2)
            return fail(GG_ERR_INVALID, "gg_graph_set_option: XCD segments must be 0, 1 or 2");
        g->xcd_segments = (int)value;
        return GG_OK;
    default:
        return fail(GG_ERR_INVALID, "gg_graph_set_option: unknown option " + std::to_string(option));
    }
}

extern "C" int gg_graph_get_option(gg_graph_t g, int option, int64_t *value)
{
    if (!g || !value) return fail(GG_ERR_INVALID, "gg_graph_get_option: NULL");
    switch (option) {
    case GG_OPT_PULL_SCHEDULE: *value = g->pull_schedule; return GG_OK;
    case GG_OPT_ROW_EDGES_PER_LANE: *value = g->edges_per_lane; return GG_OK;
    case GG_OPT_DENSE_COO: *value = g->dense_coo; return GG_OK;
    case GG_OPT_XCD_SEGMENTS: *value = g->xcd_segments; return GG_OK;
    default: return fail(GG_ERR_INVALID, "gg_graph_get_option: unknown option " + std::to_string(option));
    }
}

extern "C" int gg_graph_device_view(gg_graph_t g, int which, gg_graph_view *out)
{
    if (!g || !out) return fail(GG_ERR_INVALID, "gg_graph_device_view: NULL");
    if (which != 0 && which != 1) return fail(GG_ERR_INVALID, "gg_graph_device_view: which must be 0 or 1");
    if (which == 1 && !g->plan.mode) return fail(GG_ERR_STATE, "gg_graph_device_view: no renumbered plan");
    const bool p = which == 1;
    out->toff = (const uint32_t *)(p ? g->plan.toff.p : g->toff.p);
    out->tids = (const uint32_t *)(p ? g->plan.tids.p : g->tids.p);
    out->outdeg = (const uint32_t *)(p ? g->plan.outdeg.p : g->outdeg.p);
    out->perm = p ? (const uint32_t *)g->plan.perm.p : nullptr;
    const PullSched &s = p ? g->plan.sched : g->pull;
    out->tile_row = (const uint32_t *)s.start_row.p;
    out->tile_edge = (const uint32_t *)s.start_edge.p;
    out->tiles = s.tiles;
    out->n = g->n;
    out->m = g->m;
    const RowSched *rs = p ? &g->plan.rows : nullptr;
    out->row_work = rs && rs->blocks ? (const uint32_t *)rs->work.p : nullptr;
    out->row_blocks = rs ? rs->blocks : 0;
    out->row_long = rs ? rs->nlong : 0;
    return GG_OK;
}

extern "C" int gg_last_kernel_time(gg_ctx_t c, double *avg_ms, int *launches)
{
    if (!c) return fail(GG_ERR_INVALID, "gg_last_kernel_time: NULL");
    if (avg_ms) *avg_ms = c->last_kernel_launches ? c->last_kernel_ms / c->last_kernel_launches : 0;
    if (launches) *launches = c->last_kernel_launches;
    return GG_OK;
}
