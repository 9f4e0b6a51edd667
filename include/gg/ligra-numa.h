// gg/ligra-numa.h -- drop-in C++ front end of the MI355X traversal engine.
//
// Application sources written against GraphGrind's app-facing API
// (graphgrind/ligra-numa.h, graph-numa.h, partitioner.h, mm.h, utils.h)
// compile against this header unchanged: build with -I include/gg and
// link -lgraphgrind_amd.  The traversal runs on the GPU through the C ABI
// of include/graphgrind_amd.h; nothing here launches kernels itself.
//
// How the reference API maps onto the device:
//  - edgeMap / vertexMap take the app's functor structs.  A functor is
//    recognised by its data members (the reference functors are plain
//    structs of arrays) and becomes an enumerated device operator:
//      Parents                       -> GG_OP_BFS       (app/BFS.C:26-75)
//      IDs + prevIDs                 -> GG_OP_CC        (app/Components.C:26-86)
//      p_curr + p_next + V           -> GG_OP_PAGERANK  (app/PageRank.C:28-70)
//      p_curr + p_next (no V)        -> GG_OP_SPMV      (app/SPMV.C:7-48)
//      ... + add_factor              -> GG_OP_SUM       (app/PageRankConverage.C:34-88)
//      V + Delta + nghSum            -> GG_OP_PAGERANK  (app/PageRankDelta.C:27-72)
//      ShortestPathLen + Visited     -> GG_OP_SSSP      (app/BellmanFord.C:28-92)
//      NumPaths|Dependencies+Visited -> GG_OP_PATHSUM   (app/BC.C:28-124)
//    and for vertexMap
//      damping + addedConstant       -> GG_VOP_PAGERANK (app/PageRank.C:73-87)
//      IDs + prevIDs                 -> GG_VOP_CC_COPY  (app/Components.C:88-99)
//      p_curr only                   -> GG_VOP_RESET    (app/PageRank.C:90-100)
//    Other vertexMap functors run on the host over the frontier's members.
//  - mmap_ptr<T> (mm.h) is an HBM buffer with a host shadow.  Host code
//    (map_vertexL lambdas, operator[], get()) works on the shadow; device
//    operators work on HBM; each side is refreshed when the other wrote
//    last.  Those copies are the cost of running unmodified app source;
//    the fused device loops (gg_pagerank, gg_bfs, gg_cc, gg_spmv) avoid them.
//  - partitioned_graph / partitioned_vertices hold C-ABI handles.
//  - intT is 32-bit here: the device engine uses uint32 ids and offsets.
#ifndef GG_LIGRA_NUMA_H
#define GG_LIGRA_NUMA_H

#include <sys/mman.h>
#include <sys/stat.h>
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <string>
#include <type_traits>
#include <typeinfo>
#include <utility>
#include <vector>

#include "../graphgrind_amd.h"

using namespace std;

// ---- widths and parallel loops (parallel.h:27-106) ---------------------------
typedef int intT;
typedef unsigned int uintT;
typedef int intE;
typedef unsigned int uintE;
#define INT_T_MAX INT32_MAX
#define UINT_T_MAX UINT32_MAX
#ifdef _OPENMP
#include <omp.h>
#define parallel_for _Pragma("omp parallel for") for
#else
#define parallel_for for
#endif
#define parallel_main main
#define cilk_spawn
#define cilk_sync

namespace gg_front {

[[noreturn]] inline void die(const char *what)
{
    std::cerr << what << ": " << gg_last_error() << std::endl;
    std::abort();
}

inline void check(int rc, const char *what)
{
    if (rc != GG_OK) die(what);
}

inline gg_ctx_t &context()
{
    static gg_ctx_t ctx = nullptr;
    if (!ctx) {
        const char *dev = getenv("GG_DEVICE");
        check(gg_ctx_create(dev ? atoi(dev) : 0, &ctx), "gg_ctx_create");
    }
    return ctx;
}

// ---- host/device coherence of mmap_ptr buffers ---------------------------------
struct Buffer {
    void *host = nullptr;
    void *dev = nullptr;
    size_t bytes = 0;
    enum State { HOST, DEVICE, BOTH } state = HOST;
};

inline std::map<const void *, Buffer *> &registry()
{
    static std::map<const void *, Buffer *> r;
    return r;
}

// What keeping the host shadows coherent costs: copies each way and the time
// spent in host loops, printed by the driver after each round when the
// environment sets GG_FRONT_STATS (see report_traffic).
struct Traffic {
    uint64_t down_bytes = 0, up_bytes = 0, downs = 0, ups = 0;
    double down_s = 0, up_s = 0, host_s = 0;
};
inline Traffic &traffic()
{
    static Traffic t;
    return t;
}
inline double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline void to_host(Buffer *b)
{
    if (b->state == Buffer::DEVICE) {
        const double t0 = now_s();
        check(gg_memcpy_d2h(context(), b->host, b->dev, b->bytes), "download");
        Traffic &t = traffic();
        t.down_s += now_s() - t0;
        t.down_bytes += b->bytes;
        t.downs++;
        b->state = Buffer::BOTH;
    }
}

inline void to_device(Buffer *b)
{
    if (b->state == Buffer::HOST) {
        const double t0 = now_s();
        check(gg_memcpy_h2d(context(), b->dev, b->host, b->bytes), "upload");
        Traffic &t = traffic();
        t.up_s += now_s() - t0;
        t.up_bytes += b->bytes;
        t.ups++;
        b->state = Buffer::BOTH;
    }
}

// Every host phase may read and write any buffer.
inline double &host_phase_t0()
{
    static double t0 = 0;
    return t0;
}
inline void host_phase_begin()
{
    for (auto &kv : registry()) to_host(kv.second);
    host_phase_t0() = now_s();
}
inline void host_phase_end()
{
    traffic().host_s += now_s() - host_phase_t0();
    for (auto &kv : registry()) kv.second->state = Buffer::HOST;
}

// "gg_front: round R: ..." after a round when GG_FRONT_STATS is set: the
// round's HBM<->host copies (count, MB, ms) and its time in host loops.
inline void report_traffic(long round)
{
    Traffic &t = traffic();
    if (getenv("GG_FRONT_STATS"))
        std::printf("gg_front: round %ld: downloads %llu (%.1f MB, %.3f ms), uploads %llu (%.1f MB, %.3f ms), "
                    "host loops %.3f ms\n",
                    round, (unsigned long long)t.downs, t.down_bytes / 1e6, t.down_s * 1e3,
                    (unsigned long long)t.ups, t.up_bytes / 1e6, t.up_s * 1e3, t.host_s * 1e3);
    t = Traffic{};
}

// Device address of an app array (a host shadow pointer handed to a functor).
inline void *device_of(const void *host, bool written)
{
    auto it = registry().find(host);
    if (it == registry().end()) die("device operator on an array not allocated with mmap_ptr");
    Buffer *b = it->second;
    to_device(b);
    if (written) b->state = Buffer::DEVICE;
    return b->dev;
}

} // namespace gg_front

// ---- partitioner (partitioner.h:38-200) -----------------------------------------
class partitioner
{
    std::vector<intT> starts_;
    intT elements_ = 0;

public:
    partitioner() {}
    partitioner(intT P, intT elements) : starts_(P + 1, 0), elements_(elements) {}
    void set_starts(const std::vector<uint64_t> &s)
    {
        starts_.assign(s.begin(), s.end());
    }
    int get_num_per_node_partitions() const { return get_num_partitions(); }
    int get_num_partitions() const { return (int)starts_.size() - 1; }
    intT get_num_elements() const { return elements_; }
    intT start_of(intT p) const { return starts_[p]; }
    intT get_size(intT p) const { return starts_[p + 1] - starts_[p]; }
    intT partition_of(intT v) const
    {
        return (intT)(std::upper_bound(starts_.begin(), starts_.end(), v) - starts_.begin()) - 1;
    }
    intT offset_of(intT v) const { return v - starts_[partition_of(v)]; }
};

// ---- mmap_ptr<T> (mm.h:246-325) ---------------------------------------------------
template <typename T>
class mmap_ptr
{
    gg_front::Buffer *buf_ = nullptr;

public:
    mmap_ptr() {}
    explicit mmap_ptr(const partitioner &part) { part_allocate(part); }
    explicit mmap_ptr(size_t elements) { Interleave_allocate(elements); }
    mmap_ptr(size_t elements, size_t) { Interleave_allocate(elements); }

    void part_allocate(const partitioner &part) { Interleave_allocate(part.get_num_elements()); }
    void local_allocate(size_t elements, int) { Interleave_allocate(elements); }
    void Interleave_allocate(size_t elements)
    {
        buf_ = new gg_front::Buffer();
        buf_->bytes = std::max<size_t>(elements, 1) * sizeof(T);
        gg_front::check(gg_host_alloc(gg_front::context(), buf_->bytes, &buf_->host), "gg_host_alloc");
        gg_front::check(gg_malloc(gg_front::context(), buf_->bytes, &buf_->dev), "gg_malloc");
        buf_->state = gg_front::Buffer::HOST;
        gg_front::registry()[buf_->host] = buf_;
    }
    void del()
    {
        if (!buf_) return;
        gg_front::registry().erase(buf_->host);
        gg_free(gg_front::context(), buf_->dev);
        gg_host_free(gg_front::context(), buf_->host);
        delete buf_;
        buf_ = nullptr;
    }
    // Host view: refreshed from HBM when a device operator wrote last, and
    // considered written afterwards.
    // (the apps index an mmap_ptr once per element inside parallel host
    // loops: the state is written only when it changes, so those accesses
    // only read the shared Buffer record)
    operator T *()
    {
        if (!buf_) return nullptr;
        if (buf_->state != gg_front::Buffer::HOST) {
            gg_front::to_host(buf_);
            buf_->state = gg_front::Buffer::HOST;
        }
        return reinterpret_cast<T *>(buf_->host);
    }
    operator const T *() const
    {
        if (!buf_) return nullptr;
        gg_front::to_host(buf_);
        return reinterpret_cast<const T *>(buf_->host);
    }
    T *get() const
    {
        if (!buf_) return nullptr;
        if (buf_->state != gg_front::Buffer::HOST) {
            gg_front::to_host(buf_);
            buf_->state = gg_front::Buffer::HOST;
        }
        return reinterpret_cast<T *>(buf_->host);
    }
    explicit operator bool() const { return buf_ != nullptr; }
    size_t get_bytes() const { return buf_ ? buf_->bytes : 0; }
    T *device() const { return reinterpret_cast<T *>(gg_front::device_of(buf_->host, false)); }
};

// ---- vertex types (graph-numa.h:39-217): degrees only on the host ------------------
class symmetricVertex
{
    intT degree = 0;

public:
    symmetricVertex() {}
    intT getInDegree() const { return degree; }
    intT getOutDegree() const { return degree; }
    void setInDegree(intT d) { degree = d; }
    void setOutDegree(intT d) { degree = d; }
    void flipEdges() {}
};

class asymmetricVertex
{
    intT inDegree = 0, outDegree = 0;

public:
    asymmetricVertex() {}
    intT getInDegree() const { return inDegree; }
    intT getOutDegree() const { return outDegree; }
    void setInDegree(intT d) { inDegree = d; }
    void setOutDegree(intT d) { outDegree = d; }
    void flipEdges() { std::swap(inDegree, outDegree); }
};

// graph<vertex> (graph-numa.h:560-613): the vertex table apps index (V[v]).
template <class vertex>
class graph
{
public:
    intT n = 0, m = 0;
    vertex *V = nullptr;
    partitioner csc;
    bool isSymmetric = false;
    bool transposed = false;
    const partitioner &get_csc_partitioner() const { return csc; }
};

// partitioned_graph (graph-numa.h:626-756) holding the device graph.
template <class vertex>
class partitioned_graph
{
public:
    typedef vertex vertex_type;
    intT n = 0, m = 0;
    bool source = false, part_ver = false, part_relabel = false;
    gg_graph_t h = nullptr;
    partitioner coo_partition;
    graph<vertex> *WG = nullptr;

    partitioned_graph() {}
    // Takes ownership of a device graph and partitions it ('-c', '-v', '-o').
    partitioned_graph(gg_graph_t g, int P, bool by_source, bool by_vertex, bool relabel)
        : source(by_source), part_ver(by_vertex), part_relabel(relabel), h(g)
    {
        if (by_source) {
            std::cerr << "-P source: only destination partitioning is provided" << std::endl;
            std::abort();
        }
        uint32_t nn;
        uint64_t mm;
        int sym = 0;
        gg_front::check(gg_graph_info(g, &nn, &mm, &sym, nullptr), "gg_graph_info");
        n = (intT)nn;
        m = (intT)mm;
        gg_front::check(gg_graph_partition(g, P, by_vertex, relabel, 1), "gg_graph_partition");
        std::vector<uint64_t> s(P + 1);
        coo_partition = partitioner(P, n);
        gg_front::check(gg_graph_partition_starts(g, 0, s.data(), nullptr), "partition starts");
        coo_partition.set_starts(s);
        WG = new graph<vertex>();
        WG->n = n;
        WG->m = m;
        WG->isSymmetric = sym != 0;
        WG->V = new vertex[n];
        std::vector<uint32_t> od(n), id(n);
        gg_front::check(gg_graph_degrees(g, 0, od.data()), "out-degrees");
        gg_front::check(gg_graph_degrees(g, 1, id.data()), "in-degrees");
        for (intT v = 0; v < n; v++) {
            WG->V[v].setInDegree((intT)id[v]);
            WG->V[v].setOutDegree((intT)od[v]);
        }
        uint64_t nc = 0;
        gg_front::check(gg_graph_partition_starts(g, 1, s.data(), nullptr), "csc starts");
        nc = s[P];
        WG->csc = partitioner(P, (intT)nc);
        WG->csc.set_starts(s);
    }
    void del()
    {
        if (h) gg_graph_destroy(h);
        h = nullptr;
        if (WG) {
            delete[] WG->V;
            delete WG;
            WG = nullptr;
        }
    }
    graph<vertex> &get_partition() { return *WG; }
    const partitioner &get_partitioner() const { return coo_partition; }
    const partitioner &get_coo_partitioner() const { return coo_partition; }
    int get_num_coo_partitions() const { return coo_partition.get_num_partitions(); }
    // transpose() (graph-numa.h:547-555, 721-725): the device graph swaps
    // its two edge directions; the host vertex table flips its degrees.
    // Copies of this object share the device graph and the vertex table.
    bool transposed() const { return WG && WG->transposed; }
    void transpose()
    {
        gg_front::check(gg_graph_transpose(h), "transpose");
        if (WG && !WG->isSymmetric) {
            for (intT v = 0; v < n; v++) WG->V[v].flipEdges();
            WG->transposed = !WG->transposed;
        }
    }
};

// ---- partitioned_vertices (ligra-numa.h:91-272) -------------------------------------
class partitioned_vertices
{
public:
    gg_frontier_t h = nullptr;
    intT numVertices = 0;
    intT d_m = 0;
    intT num_out_edges = 0;
    bool bit = false;
    bool has_dense = false;

    void refresh()
    {
        uint64_t dm = 0, oe = 0;
        int dense = 0, b = 0;
        uint32_t nn = 0;
        gg_front::check(gg_frontier_info(h, &dm, &oe, &dense, &b, &nn), "gg_frontier_info");
        d_m = (intT)dm;
        num_out_edges = (intT)oe;
        has_dense = dense;
        bit = b;
        numVertices = (intT)nn;
    }
    static partitioned_vertices wrap(gg_frontier_t f)
    {
        partitioned_vertices pv;
        pv.h = f;
        pv.refresh();
        return pv;
    }
    template <class vertex>
    static partitioned_vertices create(partitioned_graph<vertex> &GA, intT v)
    {
        gg_frontier_t f;
        gg_front::check(gg_frontier_create(GA.h, GG_FRONTIER_SINGLE, (uint32_t)v, &f), "create");
        return wrap(f);
    }
    // create(n, v, outdeg) / bits(part, n, m) / empty(): the graph is the
    // one the driver loaded (one graph per process, as in the reference)
    static partitioned_vertices create(intT n, intT v, intT);
    static partitioned_vertices bits(const partitioner &, intT n, intT m);
    static partitioned_vertices empty();
    // dense(n, part) / sparse(n): n rows, nothing active (132-159);
    // boolean(n, flags, ...) / indice(n, ids, count, ...): from a bool
    // array or an id list (183-208); the active count and out-degree sum are
    // recomputed on the device
    static partitioned_vertices dense(intT n, const partitioner &);
    static partitioned_vertices sparse(intT n);
    static partitioned_vertices boolean(intT n, mmap_ptr<bool> flags, intT, intT);
    static partitioned_vertices indice(intT n, mmap_ptr<intT> ids, intT activeM, intT);
    void del()
    {
        if (h) gg_frontier_destroy(h);
        h = nullptr;
    }
    bool isEmpty() const { return d_m == 0; }
    intT numRows() const { return numVertices; }
    intT numNonzeros() const { return d_m; }
    void toDense(const partitioner &)
    {
        gg_front::check(gg_frontier_to_dense(h), "toDense");
        refresh();
    }
    void toSparse()
    {
        gg_front::check(gg_frontier_to_sparse(h), "toSparse");
        refresh();
    }
    // Apps may write the public field 'bit' (app/PageRankBit.C:145 sets it
    // on an edgeMap output); the engine's frontier follows it before use.
    void push_bit() const
    {
        if (!h || numVertices == 0) return;
        int b = 0;
        gg_front::check(gg_frontier_info(h, nullptr, nullptr, nullptr, &b, nullptr), "gg_frontier_info");
        if ((b != 0) != bit) gg_front::check(gg_frontier_set_bit(h, bit ? 1 : 0), "bit");
    }
    std::vector<uint8_t> members() const
    {
        std::vector<uint8_t> d((size_t)std::max<intT>(numVertices, 0));
        if (numVertices) gg_front::check(gg_frontier_copy_dense(h, d.data()), "copy frontier");
        return d;
    }
};

namespace gg_front {
inline gg_graph_t &current_graph()
{
    static gg_graph_t g = nullptr;
    return g;
}
} // namespace gg_front

inline partitioned_vertices partitioned_vertices::create(intT, intT v, intT)
{
    gg_frontier_t f;
    gg_front::check(gg_frontier_create(gg_front::current_graph(), GG_FRONTIER_SINGLE, (uint32_t)v, &f), "create");
    return wrap(f);
}
inline partitioned_vertices partitioned_vertices::bits(const partitioner &, intT, intT)
{
    gg_frontier_t f;
    gg_front::check(gg_frontier_create(gg_front::current_graph(), GG_FRONTIER_ALL, 0, &f), "bits");
    return wrap(f);
}
inline partitioned_vertices partitioned_vertices::dense(intT n, const partitioner &)
{
    std::vector<uint8_t> none((size_t)std::max<intT>(n, 1), 0);
    gg_frontier_t f;
    gg_front::check(gg_frontier_from_host(gg_front::current_graph(), none.data(), nullptr, 0, &f), "dense");
    return wrap(f);
}
inline partitioned_vertices partitioned_vertices::sparse(intT)
{
    gg_frontier_t f;
    gg_front::check(gg_frontier_from_host(gg_front::current_graph(), nullptr, nullptr, 0, &f), "sparse");
    return wrap(f);
}
inline partitioned_vertices partitioned_vertices::boolean(intT n, mmap_ptr<bool> flags, intT, intT)
{
    const bool *b = flags.get();
    std::vector<uint8_t> d((size_t)std::max<intT>(n, 1), 0);
    for (intT i = 0; i < n; i++) d[i] = b[i] ? 1 : 0;
    gg_frontier_t f;
    gg_front::check(gg_frontier_from_host(gg_front::current_graph(), d.data(), nullptr, 0, &f), "boolean");
    return wrap(f);
}
inline partitioned_vertices partitioned_vertices::indice(intT, mmap_ptr<intT> ids, intT activeM, intT)
{
    const intT *s = ids.get();
    std::vector<uint32_t> l((size_t)std::max<intT>(activeM, 0));
    for (intT i = 0; i < activeM; i++) l[i] = (uint32_t)s[i];
    gg_frontier_t f;
    gg_front::check(gg_frontier_from_host(gg_front::current_graph(), nullptr, l.data(), l.size(), &f), "indice");
    return wrap(f);
}
inline partitioned_vertices partitioned_vertices::empty()
{
    gg_frontier_t f;
    gg_front::check(gg_frontier_create(gg_front::current_graph(), GG_FRONTIER_EMPTY, 0, &f), "empty");
    return wrap(f);
}

// ---- functor traits ------------------------------------------------------------------
namespace gg_front {
#define GG_HAS_MEMBER(name)                                                                  \
    template <class F, class = void>                                                         \
    struct has_##name : std::false_type {};                                                  \
    template <class F>                                                                       \
    struct has_##name<F, std::void_t<decltype(std::declval<F &>().name)>> : std::true_type {};
GG_HAS_MEMBER(Parents)
GG_HAS_MEMBER(IDs)
GG_HAS_MEMBER(prevIDs)
GG_HAS_MEMBER(p_curr)
GG_HAS_MEMBER(p_next)
GG_HAS_MEMBER(V)
GG_HAS_MEMBER(damping)
GG_HAS_MEMBER(addedConstant)
GG_HAS_MEMBER(Delta)
GG_HAS_MEMBER(nghSum)
GG_HAS_MEMBER(ShortestPathLen)
GG_HAS_MEMBER(Visited)
GG_HAS_MEMBER(NumPaths)
GG_HAS_MEMBER(Dependencies)
GG_HAS_MEMBER(add_factor)
#undef GG_HAS_MEMBER

template <class P>
constexpr int dtype_of()
{
    using T = typename std::remove_cv<typename std::remove_pointer<P>::type>::type;
    return std::is_same<T, float>::value ? GG_F32 : GG_F64;
}

// ---- functor self-checks ------------------------------------------------------------
// Functors are recognised by their data members.  Before the first device
// launch for a functor type, its own update()/updateAtomic()/cond() (edgeMap)
// or operator() (vertexMap) run on a copy whose arrays point at small host
// scratch arrays with known values, and the results are compared with what
// the chosen device operator computes.  A recognised edgeMap functor whose
// body computes something else stops the program with a message (there is
// no host path for edgeMap); a vertexMap functor that fails runs on the host.
template <class F, class = void>
struct has_update3 : std::false_type {};
template <class F>
struct has_update3<F, std::void_t<decltype(std::declval<F &>().update(intT(0), intT(0), intE(0)))>>
    : std::true_type {};
template <class F, class = void>
struct has_update2 : std::false_type {};
template <class F>
struct has_update2<F, std::void_t<decltype(std::declval<F &>().update(intT(0), intT(0)))>> : std::true_type {};
template <class F, class = void>
struct has_atomic3 : std::false_type {};
template <class F>
struct has_atomic3<F, std::void_t<decltype(std::declval<F &>().updateAtomic(intT(0), intT(0), intE(0)))>>
    : std::true_type {};
template <class F, class = void>
struct has_atomic2 : std::false_type {};
template <class F>
struct has_atomic2<F, std::void_t<decltype(std::declval<F &>().updateAtomic(intT(0), intT(0)))>>
    : std::true_type {};
template <class F, class = void>
struct has_cond : std::false_type {};
template <class F>
struct has_cond<F, std::void_t<decltype(std::declval<F &>().cond(intT(0)))>> : std::true_type {};

// update / updateAtomic with the edge weight when the functor takes one
// (the device operators use weight 1 on unweighted graphs)
template <class F>
bool call_edge(F &f, bool atomic, intT s, intT d, intE w)
{
    // a functor with only one of the two forms is checked through it twice
    if (atomic || !(has_update3<F>::value || has_update2<F>::value)) {
        if constexpr (has_atomic3<F>::value) return f.updateAtomic(s, d, w);
        else if constexpr (has_atomic2<F>::value) return f.updateAtomic(s, d);
    }
    if constexpr (has_update3<F>::value) return f.update(s, d, w);
    else if constexpr (has_update2<F>::value) return f.update(s, d);
    else die("edgeMap: functor has neither update(s, d[, w]) nor updateAtomic(s, d[, w])");
}
template <class F>
constexpr intE weight_used()
{
    return (has_update3<F>::value || has_atomic3<F>::value) ? intE(4) : intE(1);
}
template <class F>
bool call_cond(F &f, intT d)
{
    if constexpr (has_cond<F>::value) return f.cond(d);
    else return true;
}
template <class P>
using elem_of = typename std::remove_cv<typename std::remove_pointer<P>::type>::type;

[[noreturn]] inline void functor_mismatch(const char *fname, const char *what)
{
    std::cerr << "edgeMap: functor " << fname << " has the data members of the device operator " << what
              << " but its update computes something else; the device engine cannot run it\n";
    std::exit(1);
}

// One scenario per operator: source 0, destinations 1..3, both the plain
// and the atomic update.  Values are exact in binary floating point.
template <int K, class F>
bool edge_op_agrees(const F &f)
{
    for (int atomic = 0; atomic < 2; atomic++) {
        F g = f;
        if constexpr (K == GG_OP_BFS) {
            using T = elem_of<decltype(g.Parents)>;
            T P[4] = {T(0), T(-1), T(5), T(-1)};
            g.Parents = P;
            if (!call_cond(g, 1) || call_cond(g, 2)) return false;
            if (!call_edge(g, atomic, 0, 1, 1) || P[1] != T(0)) return false;
            if (call_edge(g, atomic, 0, 2, 1) || P[2] != T(5)) return false;
        } else if constexpr (K == GG_OP_CC) {
            using T = elem_of<decltype(g.IDs)>;
            T I[4] = {T(2), T(9), T(1), T(9)}, Pv[4] = {T(2), T(9), T(1), T(7)};
            g.IDs = I;
            g.prevIDs = Pv;
            if (!call_edge(g, atomic, 0, 1, 1) || I[1] != T(2)) return false;
            if (call_edge(g, atomic, 0, 2, 1) || I[2] != T(1)) return false;
            if (call_edge(g, atomic, 0, 3, 1) || I[3] != T(2)) return false;
        } else if constexpr (K == GG_OP_SUM) {
            using T = elem_of<decltype(g.p_next)>;
            T A[4] = {T(2.5), T(0), T(0), T(0)}, N[4] = {T(0), T(1), T(0), T(0)};
            g.add_factor = A;
            g.p_next = N;
            call_edge(g, atomic, 0, 1, 1);
            if (N[1] != T(3.5)) return false;
        } else if constexpr (K == GG_OP_PAGERANK && has_Delta<F>::value) {
            // PageRankDelta's PR_F: nghSum[d] += Delta[s]/outdeg(s)
            using T = elem_of<decltype(g.Delta)>;
            using VT = elem_of<decltype(g.V)>;
            T D[4] = {T(-6), T(0), T(0), T(0)}, N[4] = {T(0), T(1), T(0), T(0)};
            VT V[4];
            V[0].setOutDegree(3);
            g.V = V;
            g.Delta = D;
            g.nghSum = N;
            call_edge(g, atomic, 0, 1, 1);
            if (N[1] != T(-1)) return false;
        } else if constexpr (K == GG_OP_PAGERANK) {
            // PR_F: p_next[d] += p_curr[s]/outdeg(s)
            using T = elem_of<decltype(g.p_next)>;
            using VT = elem_of<decltype(g.V)>;
            T C[4] = {T(-6), T(0), T(0), T(0)}, N[4] = {T(0), T(1), T(0), T(0)};
            VT V[4];
            V[0].setOutDegree(3);
            g.V = V;
            g.p_curr = C;
            g.p_next = N;
            call_edge(g, atomic, 0, 1, 1);
            if (N[1] != T(-1)) return false;
        } else if constexpr (K == GG_OP_SPMV) {
            // SPMV_F: p_next[d] += p_curr[s] * w
            using T = elem_of<decltype(g.p_next)>;
            T C[4] = {T(1.5), T(0), T(0), T(0)}, N[4] = {T(0), T(1), T(0), T(0)};
            g.p_curr = C;
            g.p_next = N;
            call_edge(g, atomic, 0, 1, weight_used<F>());
            if (N[1] != T(1) + T(1.5) * T(weight_used<F>())) return false;
        } else if constexpr (K == GG_OP_SSSP) {
            using T = elem_of<decltype(g.ShortestPathLen)>;
            using U = elem_of<decltype(g.Visited)>;
            const intE w = weight_used<F>();
            T L[4] = {T(2), T(10), T(2), T(10)};
            U Vis[4] = {U(1), U(0), U(0), U(1)};
            g.ShortestPathLen = L;
            g.Visited = Vis;
            if (!call_edge(g, atomic, 0, 1, w) || L[1] != T(2 + w) || Vis[1] != U(1)) return false;
            if (call_edge(g, atomic, 0, 2, w) || L[2] != T(2)) return false;
            if (call_edge(g, atomic, 0, 3, w) || L[3] != T(2 + w)) return false;
        } else if constexpr (K == GG_OP_PATHSUM) {
            using U = elem_of<decltype(g.Visited)>;
            U Vis[4] = {U(1), U(0), U(0), U(0)};
            g.Visited = Vis;
            auto run = [&](auto *A) {
                using T = std::remove_pointer_t<decltype(A)>;
                A[0] = T(2);
                A[1] = T(3);
                A[2] = T(0);
                A[3] = T(0);
                if (call_cond(g, 0) || !call_cond(g, 1)) return false;
                if (call_edge(g, atomic, 0, 1, 1) || A[1] != T(5)) return false;
                if (!call_edge(g, atomic, 0, 2, 1) || A[2] != T(2)) return false;
                return true;
            };
            if constexpr (has_NumPaths<F>::value) {
                elem_of<decltype(g.NumPaths)> A[4];
                g.NumPaths = A;
                if (!run(A)) return false;
            } else {
                elem_of<decltype(g.Dependencies)> A[4];
                g.Dependencies = A;
                if (!run(A)) return false;
            }
        } else {
            return false;
        }
    }
    return true;
}

template <int K, class F>
void check_edge_functor(const F &f)
{
    static bool checked = false;
    if (checked) return;
    if (!edge_op_agrees<K>(f)) {
        static const char *names[] = {"PAGERANK", "BFS", "CC", "SPMV", "SSSP", "PATHSUM", "SUM"};
        functor_mismatch(typeid(F).name(), names[K]);
    }
    checked = true;
}

template <class F>
gg_op_t edge_op(F &f)
{
    gg_op_t op{};
    if constexpr (has_Parents<F>::value) {
        check_edge_functor<GG_OP_BFS>(f);
        op.kind = GG_OP_BFS;
        op.dtype = GG_I32;
        op.a = device_of(f.Parents, true);
    } else if constexpr (has_IDs<F>::value && has_prevIDs<F>::value) {
        check_edge_functor<GG_OP_CC>(f);
        op.kind = GG_OP_CC;
        op.dtype = GG_U32;
        op.a = device_of(f.IDs, true);
        op.b = device_of(f.prevIDs, false);
    } else if constexpr (has_p_next<F>::value && has_add_factor<F>::value) {
        // PR_F of app/PageRankConverage.C:34-88: p_next[d] += add_factor[s]
        check_edge_functor<GG_OP_SUM>(f);
        op.kind = GG_OP_SUM;
        op.dtype = dtype_of<decltype(f.add_factor)>();
        op.a = device_of(f.add_factor, false);
        op.b = device_of(f.p_next, true);
    } else if constexpr (has_p_curr<F>::value && has_p_next<F>::value) {
        op.kind = has_V<F>::value ? GG_OP_PAGERANK : GG_OP_SPMV;
        check_edge_functor<has_V<F>::value ? GG_OP_PAGERANK : GG_OP_SPMV>(f);
        op.dtype = dtype_of<decltype(f.p_curr)>();
        op.a = device_of(f.p_curr, false);
        op.b = device_of(f.p_next, true);
    } else if constexpr (has_V<F>::value && has_Delta<F>::value && has_nghSum<F>::value) {
        // PR_F of app/PageRankDelta.C:27-72: nghSum[d] += Delta[s]/outdeg(s)
        check_edge_functor<GG_OP_PAGERANK>(f);
        op.kind = GG_OP_PAGERANK;
        op.dtype = dtype_of<decltype(f.Delta)>();
        op.a = device_of(f.Delta, false);
        op.b = device_of(f.nghSum, true);
    } else if constexpr (has_ShortestPathLen<F>::value && has_Visited<F>::value) {
        // BF_F of app/BellmanFord.C:28-92
        static_assert(sizeof(*f.ShortestPathLen) == 4 && sizeof(*f.Visited) == 4, "BF_F: 32-bit intT arrays");
        check_edge_functor<GG_OP_SSSP>(f);
        op.kind = GG_OP_SSSP;
        op.dtype = GG_I32;
        op.a = device_of(f.ShortestPathLen, true);
        op.b = device_of(f.Visited, true);
    } else if constexpr ((has_NumPaths<F>::value || has_Dependencies<F>::value) && has_Visited<F>::value) {
        // BC_F / BC_Back_F of app/BC.C:28-124: a[d] += a[s] for unvisited d
        static_assert(sizeof(*f.Visited) == 1, "BC_F: bool Visited");
        check_edge_functor<GG_OP_PATHSUM>(f);
        op.kind = GG_OP_PATHSUM;
        if constexpr (has_NumPaths<F>::value) {
            op.dtype = dtype_of<decltype(f.NumPaths)>();
            op.a = device_of(f.NumPaths, true);
        } else {
            op.dtype = dtype_of<decltype(f.Dependencies)>();
            op.a = device_of(f.Dependencies, true);
        }
        op.b = device_of(f.Visited, false);
    } else {
        static_assert(sizeof(F) == 0, "edgeMap: functor not recognised by the device engine");
    }
    return op;
}
} // namespace gg_front

namespace gg_front {
// vertexMap's self-check (see the edgeMap one above): the functor's own
// operator() on a copy over scratch arrays must do what the device
// operator does, else the functor runs on the host.  Checked once per type.
template <class F>
bool vertex_op_agrees(const F &add, int kind)
{
    static int verdict = -1;
    if (verdict >= 0) return verdict != 0;
    bool ok = false;
    F g = add;
    if constexpr (has_damping<F>::value && has_addedConstant<F>::value && has_p_next<F>::value) {
        if (kind == GG_VOP_PAGERANK) {
            using T = elem_of<decltype(g.p_next)>;
            T N[2] = {T(2), T(3)};
            g.p_next = N;
            g(0);
            const double want = g.damping * 2.0 + g.addedConstant;
            ok = std::fabs((double)N[0] - want) <= 1e-6 * std::fabs(want) && N[1] == T(3);
        }
    } else if constexpr (has_IDs<F>::value && has_prevIDs<F>::value) {
        if (kind == GG_VOP_CC_COPY) {
            using T = elem_of<decltype(g.IDs)>;
            T I[2] = {T(7), T(8)}, Pv[2] = {T(1), T(2)};
            g.IDs = I;
            g.prevIDs = Pv;
            g(0);
            ok = Pv[0] == T(7) && I[0] == T(7) && Pv[1] == T(2);
        }
    } else if constexpr (has_p_curr<F>::value) {
        if (kind == GG_VOP_RESET) {
            using T = elem_of<decltype(g.p_curr)>;
            T C[2] = {T(5), T(6)};
            g.p_curr = C;
            g(0);
            ok = C[0] == T(0) && C[1] == T(6);
        }
    }
    verdict = ok ? 1 : 0;
    return ok;
}
} // namespace gg_front

enum options { DENSE, DENSE_FORWARD };

inline bool cond_true(intT) { return 1; }

// edgeMap (ligra-numa.h:919-1019)
template <class F, class vertex>
partitioned_vertices edgeMap(partitioned_graph<vertex> GA, partitioned_vertices Localfrontier, F f,
                             intT threshold = -1, char option = DENSE, bool remDups = false)
{
    (void)option;
    Localfrontier.push_bit();
    gg_op_t op = gg_front::edge_op(f);
    gg_frontier_t out = nullptr;
    int path = 0;
    gg_front::check(gg_edgemap(GA.h, Localfrontier.h, &op, threshold, remDups, &out, &path), "edgeMap");
    return partitioned_vertices::wrap(out);
}

// vertexMap (ligra-numa.h:1023-1049)
template <class F>
void vertexMap(const partitioner &part, partitioned_vertices V, F add)
{
    using namespace gg_front;
    V.push_bit();
    gg_vop_t op{};
    bool device = true;
    if constexpr (has_damping<F>::value && has_addedConstant<F>::value && has_p_next<F>::value) {
        op.kind = GG_VOP_PAGERANK;
        op.dtype = dtype_of<decltype(add.p_next)>();
        op.a = device_of(add.p_next, true);
        op.b = op.a;
        op.c0 = add.damping;
        op.c1 = add.addedConstant;
    } else if constexpr (has_IDs<F>::value && has_prevIDs<F>::value) {
        op.kind = GG_VOP_CC_COPY;
        op.dtype = GG_U32;
        op.a = device_of(add.IDs, false);
        op.b = device_of(add.prevIDs, true);
    } else if constexpr (has_p_curr<F>::value && !has_p_next<F>::value) {
        op.kind = GG_VOP_RESET;
        op.dtype = dtype_of<decltype(add.p_curr)>();
        op.a = device_of(add.p_curr, true);
    } else {
        device = false;
    }
    if (device && vertex_op_agrees(add, op.kind)) {
        check(gg_vertexmap(current_graph(), V.h, &op), "vertexMap");
        return;
    }
    // any other functor: run it on the host over the members
    std::vector<uint8_t> d = V.bit ? std::vector<uint8_t>() : V.members();
    host_phase_begin();
    for (intT j = 0; j < part.get_num_elements(); j++)
        if (V.bit || d[j]) add(j);
    host_phase_end();
}

// vertexFilter (ligra-numa.h:1053-1082): the vertices of V for which F
// returns true, as a new dense frontier.  F runs on the host over V's
// members (the filters of the apps are arbitrary vertex code).
template <class vertex, class F>
partitioned_vertices vertexFilter(partitioned_graph<vertex> GA, partitioned_vertices V, F filter)
{
    using namespace gg_front;
    const intT n = GA.n;
    std::vector<uint8_t> d = V.bit ? std::vector<uint8_t>() : V.members();
    std::vector<uint8_t> keep((size_t)n, 0);
    host_phase_begin();
    for (intT j = 0; j < n; j++)
        if (V.bit || d[j]) keep[j] = filter(j) ? 1 : 0;
    host_phase_end();
    gg_frontier_t f;
    check(gg_frontier_from_host(GA.h, keep.data(), nullptr, 0, &f), "vertexFilter");
    return partitioned_vertices::wrap(f);
}

// map_partitionL / map_partition / map_vertex (partitioner.h:338-399):
// host loops over the partitions of 'part' (and over their vertices).
template <typename Fn>
void map_partitionL(const partitioner &part, Fn fn)
{
    gg_front::host_phase_begin();
    const intT np = part.get_num_partitions();
    parallel_for(intT p = 0; p < np; ++p) fn(p);
    gg_front::host_phase_end();
}
#define map_partition(vname, part, ...) map_partitionL((part), [&](intT vname) { __VA_ARGS__; })
#define map_vertex(vname, part, ...)                                                    \
    map_partitionL((part), [&](intT gg_part_) {                                        \
        for (intT vname = (part).start_of(gg_part_); vname < (part).start_of(gg_part_ + 1); ++vname) { \
            __VA_ARGS__;                                                                \
        }                                                                               \
    })

// map_vertexL (partitioner.h:376-386): a host loop over all vertices.
template <typename Fn>
void map_vertexL(const partitioner &part, Fn fn)
{
    gg_front::host_phase_begin();
    const intT n = part.get_num_elements();
    parallel_for(intT v = 0; v < n; ++v) fn(v);
    gg_front::host_phase_end();
}

// ---- utils.h pieces the apps reference ---------------------------------------------
// CAS (utils.h:423-442): 4- and 8-byte values, floats compared by bits
template <class ET>
inline bool CAS(ET *ptr, ET oldv, ET newv)
{
    if constexpr (std::is_floating_point<ET>::value) {
        using U = typename std::conditional<sizeof(ET) == 4, uint32_t, uint64_t>::type;
        U o, n;
        std::memcpy(&o, &oldv, sizeof(ET));
        std::memcpy(&n, &newv, sizeof(ET));
        return __sync_bool_compare_and_swap(reinterpret_cast<U *>(ptr), o, n);
    } else {
        return __sync_bool_compare_and_swap(ptr, oldv, newv);
    }
}
// The current value is read once, atomically: a plain read may be split by
// the compiler into one load for the arithmetic and another for the CAS's
// expected bits, and a write between the two then slips through the CAS.
template <class ET>
inline ET atomic_peek(ET *a)
{
    ET v;
    __atomic_load(a, &v, __ATOMIC_RELAXED);
    return v;
}
template <class ET>
inline bool writeMin(ET *a, ET b)
{
    ET c;
    bool r = 0;
    do c = atomic_peek(a);
    while (c > b && !(r = CAS(a, c, b)));
    return r;
}
// writeAdd (utils.h:468-477): a CAS loop of the element's own width, so
// 4-byte types touch only their own 4 bytes
template <class ET>
inline void writeAdd(ET *a, ET b)
{
    static_assert(sizeof(ET) == 4 || sizeof(ET) == 8, "writeAdd: 4- or 8-byte element");
    ET o, n;
    do {
        o = atomic_peek(a);
        n = o + b;
    } while (!CAS(a, o, n));
}

namespace sequence {
// plusReduce (utils.h:192): 1024-element serial blocks, then over the sums.
// The blocks are summed in parallel, each in order, so the result does not
// depend on the thread count.
template <class OT, class IT>
OT plusReduce(OT *A, IT n)
{
    if (n <= 0) return OT(0);
    const IT nb = (n + 1023) / 1024;
    std::vector<OT> sums((size_t)nb);
    parallel_for(IT b = 0; b < nb; b++)
    {
        const IT s = b * 1024, e = std::min<IT>(s + 1024, n);
        OT r = A[s];
        for (IT j = s + 1; j < e; j++) r = r + A[j];
        sums[(size_t)b] = r;
    }
    if (nb == 1) return sums[0];
    return plusReduce(sums.data(), nb);
}
} // namespace sequence

// ---- timer (gettime.h:31-143) ----------------------------------------------------------
struct timer {
    double total = 0;
    std::chrono::steady_clock::time_point t0;
    void start() { t0 = std::chrono::steady_clock::now(); }
    double next()
    {
        auto t1 = std::chrono::steady_clock::now();
        double d = std::chrono::duration<double>(t1 - t0).count();
        total += d;
        t0 = t1;
        return d;
    }
    // seconds since start(), added to the total (gettime.h's stop())
    double stop() { return next(); }
};
static timer _tm;
#define startTime() _tm.start();
#define nextTime(_string) std::cout << _string << " : " << _tm.next() << std::endl;
#define reportAvg(r) std::cout << "Average" << " : " << _tm.total / (r) << std::endl;

// ---- command line (parseCommandLine.h:32-113) -------------------------------------------
class commandLine
{
    int argc;
    char **argv;
    std::string usage;

public:
    commandLine(int ac, char **av, std::string u) : argc(ac), argv(av), usage(u) {}
    [[noreturn]] void badArgument()
    {
        std::cout << "usage: " << argv[0] << " " << usage << std::endl;
        std::exit(0);
    }
    char *getArgument(int i)
    {
        if (argc < 2 + i) badArgument();
        return argv[argc - 1 - i];
    }
    bool getOption(const std::string &o)
    {
        for (int i = 1; i < argc; i++)
            if ((std::string)argv[i] == o) return true;
        return false;
    }
    char *getOptionValue(const std::string &o)
    {
        for (int i = 1; i < argc - 1; i++)
            if ((std::string)argv[i] == o) return argv[i + 1];
        return nullptr;
    }
    long getOptionLongValue(const std::string &o, long d)
    {
        char *v = getOptionValue(o);
        return v ? std::atol(v) : d;
    }
};

// ---- input (IO.h:136-361 text PBBS format; 365-656 Galois .gr) ----------------------------
namespace gg_front {
struct HostCSR {
    uint32_t n = 0;
    uint64_t m = 0;
    std::vector<uint32_t> off, ids;
    std::vector<float> w;
};

inline bool read_text(const char *path, bool weighted, HostCSR &G)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    std::vector<char> buf;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    buf.resize((size_t)sz + 1);
    size_t got = std::fread(buf.data(), 1, (size_t)sz, f);
    std::fclose(f);
    buf[got] = 0;
    char *p = buf.data();
    auto word = [&](std::string &out) {
        while (*p && std::isspace((unsigned char)*p)) p++;
        char *b = p;
        while (*p && !std::isspace((unsigned char)*p)) p++;
        out.assign(b, p);
        return !out.empty();
    };
    std::string tok;
    if (!word(tok) || tok != (weighted ? "WeightedAdjacencyGraph" : "AdjacencyGraph")) {
        std::cout << "Bad input file" << std::endl;
        return false;
    }
    auto num = [&]() { return std::strtoull(p, &p, 10); };
    G.n = (uint32_t)num();
    G.m = num();
    G.off.resize(G.n + 1);
    for (uint32_t i = 0; i < G.n; i++) G.off[i] = (uint32_t)num();
    G.off[G.n] = (uint32_t)G.m;
    G.ids.resize(G.m);
    for (uint64_t e = 0; e < G.m; e++) G.ids[e] = (uint32_t)num();
    if (weighted) {
        G.w.resize(G.m);
        for (uint64_t e = 0; e < G.m; e++) G.w[e] = (float)std::strtoll(p, &p, 10);
    }
    return true;
}

inline bool read_galois(const char *path, bool weighted, HostCSR &G)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    uint64_t hdr[4];
    if (std::fread(hdr, 8, 4, f) != 4) return false;
    G.n = (uint32_t)hdr[2];
    G.m = hdr[3];
    std::vector<uint64_t> ends(G.n);
    bool ok = std::fread(ends.data(), 8, G.n, f) == G.n;
    G.ids.resize(G.m);
    ok = ok && std::fread(G.ids.data(), 4, G.m, f) == G.m;
    if (ok && (G.m & 1)) {
        uint32_t pad;
        ok = std::fread(&pad, 4, 1, f) == 1;
    }
    if (ok && weighted && hdr[1] == 4) {
        G.w.resize(G.m);
        std::vector<uint32_t> wi(G.m);
        ok = std::fread(wi.data(), 4, G.m, f) == G.m;
        for (uint64_t e = 0; e < G.m; e++) G.w[e] = (float)wi[e];
    }
    std::fclose(f);
    G.off.assign(G.n + 1, 0);
    for (uint32_t i = 0; i < G.n; i++) G.off[i + 1] = (uint32_t)ends[i];
    return ok;
}
} // namespace gg_front

template <class GraphType>
void Compute(GraphType &, long);

#ifndef WEIGHTED
#define GG_FRONT_WEIGHTED false
#else
#define GG_FRONT_WEIGHTED true
#endif

// driver (ligra-numa.h:1095-1208) with one extension: "-rmat S EF SEED"
// generates the seeded R-MAT graph on the device instead of reading a file.
// Tools that include this header for its readers or types and bring their
// own main() define GG_FRONT_NO_DRIVER.
#ifndef GG_FRONT_NO_DRIVER
int parallel_main(int argc, char *argv[])
{
    commandLine P(argc, argv, " [-s] [-b] [-r root] [-rounds R] [-c P] [-v edge|vertex] [-o] <inFile | -rmat S EF SEED>");
    bool symmetric = P.getOption("-s");
    bool binary = P.getOption("-b");
    long start = P.getOptionLongValue("-r", 100);
    long rounds = P.getOptionLongValue("-rounds", 3);
    int numOfCoo = (int)P.getOptionLongValue("-c", 384);
    char *part_how = P.getOptionValue("-P");
    char *vertex_edge = P.getOptionValue("-v");
    bool relabel = P.getOption("-o");
    bool part_src = false, part_vertex = false;
    if (part_how && strcmp(part_how, "dest") && strcmp(part_how, "source")) {
        std::cerr << "Illegal value for -P: \"" << part_how << "\". Allowed values: dest source. Default: dest\n";
        return 1;
    }
    part_src = part_how && !strcmp(part_how, "source");
    if (vertex_edge && strcmp(vertex_edge, "edge") && strcmp(vertex_edge, "vertex")) {
        std::cerr << "Illegal value for -v: \"" << vertex_edge << "\". Allowed values: edge vertex. Default: edge\n";
        return 1;
    }
    part_vertex = vertex_edge && !strcmp(vertex_edge, "vertex");
    gg_ctx_t ctx = gg_front::context();
    gg_graph_t g = nullptr;
    char *rm = P.getOptionValue("-rmat");
    if (rm) {
        int i = 1;
        while (i < argc && strcmp(argv[i], "-rmat")) i++;
        if (i + 3 >= argc) P.badArgument();
        gg_front::check(gg_graph_create_rmat(ctx, atoi(argv[i + 1]), atoi(argv[i + 2]),
                                             strtoull(argv[i + 3], nullptr, 10), symmetric, GG_FRONT_WEIGHTED, &g),
                        "gg_graph_create_rmat");
    } else {
        gg_front::HostCSR G;
        char *iFile = P.getArgument(0);
        bool ok = binary ? gg_front::read_galois(iFile, GG_FRONT_WEIGHTED, G)
                         : gg_front::read_text(iFile, GG_FRONT_WEIGHTED, G);
        if (!ok) {
            std::cerr << "cannot read graph " << iFile << std::endl;
            return 1;
        }
        gg_front::check(gg_graph_create_csr(ctx, G.n, G.m, G.off.data(), G.ids.data(),
                                            G.w.empty() ? nullptr : G.w.data(), symmetric, &g),
                        "gg_graph_create_csr");
    }
    gg_front::current_graph() = g;
    int rc = 0;
    auto run = [&](auto &PG) {
        // like the reference, 'start' (-r, default 100) is only checked by
        // the apps that use it: a frontier on a vertex >= n fails there
        for (long r = 0; r < rounds; r++) {
            gg_front::traffic() = gg_front::Traffic{};
            startTime();
            Compute(PG, start);
            nextTime("Running");
            gg_front::report_traffic(r);
            // an app that transposed the graph (BC.C) gets it back (1195)
            if (PG.transposed()) PG.transpose();
        }
        reportAvg(rounds);
    };
    if (symmetric) {
        partitioned_graph<symmetricVertex> PG(g, numOfCoo, part_src, part_vertex, relabel);
        run(PG);
        PG.del();
    } else {
        partitioned_graph<asymmetricVertex> PG(g, numOfCoo, part_src, part_vertex, relabel);
        run(PG);
        PG.del();
    }
    return rc;
}
#endif // GG_FRONT_NO_DRIVER

#endif
