// ref_harness.cpp -- runs the REFERENCE engine (compiled from its own sources
// under /root/reference by oracle/Makefile) on a given graph and dumps what
// the app computed.  TEST INFRASTRUCTURE ONLY: it produces golden vectors for
// tests/golden/ and is the "reference" CPU baseline of bench.py.
//
// Built once per app with -DGG_APP=<n>; the app source is included from the
// reference tree, and the reference's main() is renamed away with
// -Dmain=gg_ref_app_main so this file supplies its own driver.  The per-app
// loops below restate each app's Compute() (app/PageRank.C:102-151,
// app/BFS.C:77-101, app/Components.C:101-126, app/SPMV.C:64-92,
// app/BC.C:154-226, app/BellmanFord.C:105-146, app/PageRankDelta.C:124-176,
// app/PageRankConverage.C:238-310, app/PageRankBit.C:102-150) with the app's own functors and the
// reference's own edgeMap/vertexMap/vertexFilter, but keep the result
// instead of freeing it.  BellmanFord is built with -DWEIGHTED (its
// functor's update takes the edge length).
//
// Options beyond the reference driver's: -R builds the partitioned graph
// with partition_relabel (the driver's -o), -t FILE records the frontier's
// public d_m and num_out_edges (ligra-numa.h:91-272) before every edgeMap of
// PageRank, PageRankBit, BFS and Components (int64 pairs), and -p FILE
// writes the partition starts and COO lists.  Built with -DEDGES_HILBERT=0
// (ref_pagerank_csr) the COO lists are in the reference's CSR order.
//
// Input: an in-memory graph, either read from a .ggb file written by the
// tests (magic, n, m, flags, offsets, ids[, int weights]) or generated with
// the oracle's R-MAT generator.  The wholeGraph is filled directly (out-lists
// in CSR order, in-lists sorted by source as IO.h:43-51,303 produce them).
#if GG_APP == 1
#include "app/PageRank.C"
#elif GG_APP == 2
#include "app/BFS.C"
#elif GG_APP == 3
#include "app/Components.C"
#elif GG_APP == 4
#include "app/SPMV.C"
#elif GG_APP == 5
#include "app/BC.C"
#elif GG_APP == 6
#include "app/BellmanFord.C"
#elif GG_APP == 7
#include "app/PageRankDelta.C"
#elif GG_APP == 8
#include "app/PageRankConverage.C"
#elif GG_APP == 9
#include "app/PageRankBit.C"
#else
#error "GG_APP must be 1..9"
#endif
#undef main

#include <chrono>
#include <cstdio>
#include <vector>

#include "gg_oracle.h"

namespace {

// -t: (d_m, num_out_edges) of the frontier each edgeMap is called with
std::vector<int64_t> g_trace;
bool g_tracing = false;
inline void trace_frontier(const partitioned_vertices &F)
{
    if (g_tracing) {
        g_trace.push_back((int64_t)F.d_m);
        g_trace.push_back((int64_t)F.num_out_edges);
    }
}

struct HostGraph {
    uint64_t n = 0, m = 0;
    bool symmetric = false, weighted = false;
    std::vector<uint32_t> off, ids;
    std::vector<int32_t> w;
};

bool read_ggb(const char *path, HostGraph &g)
{
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    uint64_t hdr[4];
    bool ok = fread(hdr, 8, 4, f) == 4 && hdr[0] == 0x31424747ULL; /* "GGB1" */
    if (ok) {
        g.n = hdr[1];
        g.m = hdr[2];
        g.symmetric = hdr[3] & 1;
        g.weighted = (hdr[3] >> 1) & 1;
        g.off.resize(g.n + 1);
        g.ids.resize(g.m);
        ok = fread(g.off.data(), 4, g.n + 1, f) == g.n + 1 &&
             fread(g.ids.data(), 4, g.m, f) == g.m;
        if (ok && g.weighted) {
            g.w.resize(g.m);
            ok = fread(g.w.data(), 4, g.m, f) == g.m;
        }
    }
    fclose(f);
    return ok;
}

bool gen_rmat(int scale, int ef, uint64_t seed, bool sym, HostGraph &g)
{
    uint64_t n = 1ULL << scale, raw = (uint64_t)ef << scale;
    std::vector<uint32_t> src(sym ? 2 * raw : raw), dst(sym ? 2 * raw : raw);
    ggo_rmat_raw(scale, raw, seed, src.data(), dst.data());
    g.off.resize(n + 1);
    g.ids.resize(sym ? 2 * raw : raw);
    int64_t m = ggo_build_csr((uint32_t)n, raw, src.data(), dst.data(), sym,
                              g.off.data(), g.ids.data());
    if (m < 0) return false;
    g.n = n;
    g.m = (uint64_t)m;
    g.ids.resize(g.m);
    g.symmetric = sym;
    if (g.weighted) {
        g.w.resize(g.m);
        for (uint64_t e = 0; e < g.m; e++) g.w[e] = 1 + (int32_t)(ggo_splitmix64(seed ^ e) % 19);
    }
    return true;
}

// Fill the reference's wholeGraph from CSR (+ CSC built here).
template <class vertex>
wholeGraph<vertex> to_whole(const HostGraph &h)
{
    wholeGraph<vertex> G((intT)h.n, (intT)h.m, h.symmetric);
    intE *out = G.allocatedInplace;
#ifndef WEIGHTED
    for (uint64_t e = 0; e < h.m; e++) out[e] = (intE)h.ids[e];
#else
    for (uint64_t e = 0; e < h.m; e++) {
        out[2 * e] = (intE)h.ids[e];
        out[2 * e + 1] = h.weighted ? h.w[e] : 1;
    }
#endif
    vertex *V = G.V;
    for (uint64_t i = 0; i < h.n; i++) {
        V[i].setOutDegree(h.off[i + 1] - h.off[i]);
#ifndef WEIGHTED
        V[i].setOutNeighbors(out + h.off[i]);
#else
        V[i].setOutNeighbors(out + 2 * (uint64_t)h.off[i]);
#endif
    }
    if (!h.symmetric) {
        std::vector<uint32_t> toff(h.n + 1), tids(h.m), perm(h.m);
        ggo_transpose((uint32_t)h.n, h.m, h.off.data(), h.ids.data(), toff.data(),
                      tids.data(), perm.data());
        intE *in = G.inEdges;
#ifndef WEIGHTED
        for (uint64_t e = 0; e < h.m; e++) in[e] = (intE)tids[e];
#else
        for (uint64_t e = 0; e < h.m; e++) {
            in[2 * e] = (intE)tids[e];
            in[2 * e + 1] = h.weighted ? h.w[perm[e]] : 1;
        }
#endif
        for (uint64_t i = 0; i < h.n; i++) {
            V[i].setInDegree(toff[i + 1] - toff[i]);
#ifndef WEIGHTED
            V[i].setInNeighbors(in + toff[i]);
#else
            V[i].setInNeighbors(in + 2 * (uint64_t)toff[i]);
#endif
        }
    }
    return G;
}

#if GG_APP == 1 || GG_APP == 9
// PageRank.C:102-151; PageRankBit.C:102-150 differs only in marking each
// edgeMap output as all-active (output.bit = true)
template <class GraphType>
int run_app(GraphType &GA, long, std::vector<double> &res, long &iters)
{
    typedef typename GraphType::vertex_type vertex;
    const partitioner &part = GA.get_partitioner();
    graph<vertex> &WG = GA.get_partition();
    intT n = GA.n, m = GA.m;
    double one_over_n = 1 / (double)n;
    mmap_ptr<double> p_curr, p_next;
    p_curr.part_allocate(part);
    p_next.part_allocate(part);
    map_vertexL(part, [&](intT j) { p_curr[j] = one_over_n; });
    map_vertexL(part, [&](intT j) { p_next[j] = 0; });
    partitioned_vertices Frontier = partitioned_vertices::bits(part, n, m);
    int count = 0;
    bool broke = false;
    while (count < MaxIter) {
        count++;
        trace_frontier(Frontier);
        partitioned_vertices output = edgeMap(GA, Frontier, PR_F<vertex>(p_curr, p_next, WG.V), m / 20);
        vertexMap(part, Frontier, PR_Vertex_F(p_curr, p_next, 0.85, n));
        map_vertexL(part, [&](intT j) { p_curr[j] = fabs(p_curr[j] - p_next[j]); });
        double L1 = sequence::plusReduce(p_curr.get(), n);
        if (L1 < 0.0000001) { broke = true; output.del(); break; }
        vertexMap(part, Frontier, PR_Vertex_Reset(p_curr));
        swap(p_curr, p_next);
        Frontier.del();
#if GG_APP == 9
        output.bit = true; // app/PageRankBit.C:145
#endif
        Frontier = output;
    }
    double *r = broke ? p_next.get() : p_curr.get();
    res.assign(r, r + n);
    iters = count;
    Frontier.del();
    p_curr.del();
    p_next.del();
    return 0;
}
#elif GG_APP == 2
template <class GraphType>
int run_app(GraphType &GA, long start, std::vector<double> &res, long &iters)
{
    const intT n = GA.n, m = GA.m;
    const partitioner &part = GA.get_partitioner();
    mmap_ptr<intT> Parents;
    Parents.part_allocate(part);
    map_vertexL(part, [&](intT j) { Parents[j] = -1; });
    Parents[start] = start;
    partitioned_vertices Frontier = partitioned_vertices::create(n, start, GA.get_partition().V[start].getOutDegree());
    long levels = 0;
    while (!Frontier.isEmpty()) {
        trace_frontier(Frontier);
        partitioned_vertices output = edgeMap(GA, Frontier, BFS_F(Parents), m / 20);
        Frontier.del();
        Frontier = output;
        levels++;
    }
    res.resize(n);
    for (intT i = 0; i < n; i++) res[i] = (double)Parents[i];
    iters = levels;
    Frontier.del();
    Parents.del();
    return 0;
}
#elif GG_APP == 3
template <class GraphType>
int run_app(GraphType &GA, long, std::vector<double> &res, long &iters)
{
    const partitioner &part = GA.get_partitioner();
    intT n = GA.n, m = GA.m;
    mmap_ptr<intT> IDs, prevIDs;
    IDs.part_allocate(part);
    prevIDs.part_allocate(part);
    map_vertexL(part, [&](intT j) { IDs[j] = j; });
    partitioned_vertices Frontier = partitioned_vertices::bits(part, n, m);
    long rounds = 0;
    while (!Frontier.isEmpty()) {
        vertexMap(part, Frontier, CC_Vertex_F(IDs, prevIDs));
        trace_frontier(Frontier);
        partitioned_vertices output = edgeMap(GA, Frontier, CC_F(IDs, prevIDs), m / 20);
        Frontier.del();
        Frontier = output;
        rounds++;
    }
    res.resize(n);
    for (intT i = 0; i < n; i++) res[i] = (double)IDs[i];
    iters = rounds;
    Frontier.del();
    IDs.del();
    prevIDs.del();
    return 0;
}
#elif GG_APP == 5
template <class GraphType>
int run_app(GraphType &GA, long start, std::vector<double> &res, long &iters)
{
    const partitioner &part = GA.get_partitioner();
    intT n = GA.n, m = GA.m;
    mmap_ptr<fType> NumPaths;
    NumPaths.part_allocate(part);
    mmap_ptr<bool> Visited;
    Visited.part_allocate(part);
    map_vertexL(part, [&](intT j) { NumPaths[j] = 0.0; });
    map_vertexL(part, [&](intT j) { Visited[j] = false; });
    NumPaths[start] = 1.0;
    Visited[start] = true;
    partitioned_vertices Frontier = partitioned_vertices::create(n, start, GA.get_partition().V[start].getOutDegree());
    vector<partitioned_vertices> Levels;
    Levels.push_back(Frontier);
    intT round = 0;
    while (!Frontier.isEmpty()) {
        round++;
        partitioned_vertices output = edgeMap(GA, Frontier, BC_F(NumPaths, Visited), m / 20);
        vertexMap(part, output, BC_Vertex_F(Visited));
        Levels.push_back(output);
        Frontier = output;
    }
    mmap_ptr<fType> Dependencies;
    Dependencies.part_allocate(part);
    map_vertexL(part, [&](intT j) { Dependencies[j] = 0.0; });
    mmap_ptr<fType> inverseNumPaths;
    inverseNumPaths = NumPaths;
    map_vertexL(part, [&](intT j) { inverseNumPaths[j] = 1 / inverseNumPaths[j]; });
    Levels[round].del();
    map_vertexL(part, [&](intT j) { Visited[j] = false; });
    Frontier = Levels[round - 1];
    vertexMap(part, Frontier, BC_Back_Vertex_F(Visited, Dependencies, inverseNumPaths));
    GA.transpose();
    for (intT r = round - 2; r >= 0; r--) {
        partitioned_vertices output = edgeMap(GA, Frontier, BC_Back_F(Dependencies, Visited), m / 20);
        output.del();
        Frontier.del();
        Frontier = Levels[r];
        vertexMap(part, Frontier, BC_Back_Vertex_F(Visited, Dependencies, inverseNumPaths));
    }
    Frontier.del();
    map_vertexL(part, [&](intT j) { Dependencies[j] = (Dependencies[j] - inverseNumPaths[j]) / inverseNumPaths[j]; });
    res.resize(n);
    for (intT i = 0; i < n; i++) res[i] = (double)Dependencies[i];
    iters = round;
    inverseNumPaths.del();
    Visited.del();
    Dependencies.del();
    return 0;
}
#elif GG_APP == 6
template <class GraphType>
int run_app(GraphType &GA, long start, std::vector<double> &res, long &iters)
{
    intT n = GA.n, m = GA.m;
    const partitioner &part = GA.get_partitioner();
    mmap_ptr<intT> ShortestPathLen;
    ShortestPathLen.part_allocate(part);
    map_vertexL(part, [&](intT j) { ShortestPathLen[j] = INT_MAX / 2; });
    ShortestPathLen[start] = 0;
    mmap_ptr<intT> Visited;
    Visited.part_allocate(part);
    map_vertexL(part, [&](intT j) { Visited[j] = 0; });
    partitioned_vertices Frontier = partitioned_vertices::create(n, start, GA.get_partition().V[start].getOutDegree());
    intT round = 0;
    while (!Frontier.isEmpty()) {
        if (round == n) {
            map_vertexL(part, [&](intT j) { ShortestPathLen[j] = -(INT_MAX / 2); });
            break;
        }
        partitioned_vertices output = edgeMap(GA, Frontier, BF_F(ShortestPathLen, Visited), m / 20);
        vertexMap(part, output, BF_Vertex_F(Visited));
        Frontier.del();
        Frontier = output;
        round++;
    }
    res.resize(n);
    for (intT i = 0; i < n; i++) res[i] = (double)ShortestPathLen[i];
    iters = round;
    Frontier.del();
    Visited.del();
    ShortestPathLen.del();
    return 0;
}
#elif GG_APP == 7
template <class GraphType>
int run_app(GraphType &GA, long, std::vector<double> &res, long &iters)
{
    typedef typename GraphType::vertex_type vertex;
    const partitioner &part = GA.get_partitioner();
    const double damping = 0.85, epsilon = 0.0000001, epsilon2 = 0.01;
    intT n = GA.n, m = GA.m;
    double one_over_n = 1 / (double)n;
    mmap_ptr<double> p, nghSum, Delta;
    p.part_allocate(part);
    nghSum.part_allocate(part);
    Delta.part_allocate(part);
    map_vertexL(part, [&](intT j) { p[j] = 0.0; });
    map_vertexL(part, [&](intT j) { Delta[j] = one_over_n; });
    map_vertexL(part, [&](intT j) { nghSum[j] = 0.0; });
    partitioned_vertices Frontier = partitioned_vertices::bits(part, n, m);
    partitioned_vertices All = partitioned_vertices::bits(part, n, m);
    intT round = 0;
    while (1) {
        round++;
        partitioned_vertices output = edgeMap(GA, Frontier, PR_F<vertex>(GA.get_partition().V, Delta, nghSum), m / 20);
        output.del();
        partitioned_vertices active =
            (round == 1) ? vertexFilter(GA, All, PR_Vertex_F_FirstRound(p, Delta, nghSum, damping, one_over_n, epsilon2))
                         : vertexFilter(GA, All, PR_Vertex_F(p, Delta, nghSum, damping, epsilon2));
        map_vertexL(part, [&](intT j) { nghSum[j] = fabs(Delta[j]); });
        double L1_norm = sequence::plusReduce(nghSum.get(), n);
        if (L1_norm < epsilon) { active.del(); break; }
        vertexMap(part, All, PR_Vertex_Reset(nghSum));
        Frontier.del();
        Frontier = active;
    }
    res.assign(p.get(), p.get() + n);
    iters = round;
    Frontier.del();
    All.del();
    p.del();
    Delta.del();
    nghSum.del();
    return 0;
}
#elif GG_APP == 8
template <class GraphType>
int run_app(GraphType &GA, long, std::vector<double> &res, long &iters)
{
    typedef typename GraphType::vertex_type vertex;
    const partitioner &part = GA.get_partitioner();
    graph<vertex> &WG = GA.get_partition();
    intT n = GA.n, m = GA.m;
    const double epsilon = 0.0000001;
    double one_over_n = 1 / (double)n;
    mmap_ptr<double> p_curr, p_next, edge_factor, add_factor;
    p_curr.part_allocate(part);
    p_next.part_allocate(part);
    edge_factor.part_allocate(part);
    add_factor.part_allocate(part);
    map_vertexL(part, [&](intT j) { p_curr[j] = one_over_n; });
    map_vertexL(part, [&](intT j) { p_next[j] = 0; });
    map_vertexL(part, [&](intT j) { edge_factor[j] = (double)0.85 / (WG.V[j].getOutDegree()); });
    int count = 0;
    double L1_norm = 2;
    partitioned_vertices Frontier = partitioned_vertices::bits(part, n, m);
    while (count < MaxIter && L1_norm > epsilon) {
        map_vertexL(part, [&](intT j) { add_factor[j] = (double)edge_factor[j] * p_curr[j]; });
        partitioned_vertices output = edgeMap(GA, Frontier, PR_F<vertex>(p_curr, p_next, WG.V, add_factor), m / 20);
        vertexMap(part, Frontier, PR_Vertex_Csum(p_next, sum(part, p_next, n, 1), n));
        L1_norm = normdiff(part, p_curr, p_next, n);
        vertexMap(part, Frontier, PR_Vertex_norm(p_next, sum(part, p_next, n, 0), n));
        vertexMap(part, Frontier, PR_Vertex_Reset(p_curr));
        count++;
        swap(p_curr, p_next);
        output.del();
    }
    res.assign(p_curr.get(), p_curr.get() + n);
    iters = count;
    Frontier.del();
    p_curr.del();
    p_next.del();
    edge_factor.del();
    add_factor.del();
    return 0;
}
#else
static std::vector<double> g_x;
template <class GraphType>
int run_app(GraphType &GA, long, std::vector<double> &res, long &iters)
{
    typedef typename GraphType::vertex_type vertex;
    const partitioner &part = GA.get_partitioner();
    intT n = GA.n, m = GA.m;
    double one_over_n = 1 / (double)n;
    mmap_ptr<double> p_curr, p_next;
    p_curr.part_allocate(part);
    p_next.part_allocate(part);
    map_vertexL(part, [&](intT j) { p_curr[j] = g_x.empty() ? one_over_n : g_x[j]; });
    map_vertexL(part, [&](intT j) { p_next[j] = 0; });
    partitioned_vertices Frontier = partitioned_vertices::bits(part, n, m);
    partitioned_vertices output = edgeMap(GA, Frontier, SPMV_F<vertex>(p_curr, p_next), m / 20);
    vertexMap(part, Frontier, SPMV_Vertex_Reset(p_curr));
    swap(p_curr, p_next);
    res.assign(p_curr.get(), p_curr.get() + n);
    iters = 1;
    output.del();
    Frontier.del();
    p_curr.del();
    p_next.del();
    return 0;
}
#endif

template <class vertex>
int drive(HostGraph &h, int P, bool part_vertex, bool relabel, long start, int rounds, const char *out_path,
          const char *part_path, const char *trace_path)
{
    wholeGraph<vertex> G = to_whole<vertex>(h);
    partitioned_graph<vertex> PG(G, P, false, part_vertex, relabel);
    if (part_path) {
        // partition boundaries and COO lists, for pinning the restatement
        FILE *f = fopen(part_path, "wb");
        const partitioner &cp = PG.get_coo_partitioner();
        const partitioner &sp = PG.get_partition().get_csc_partitioner();
        for (int p = 0; p <= P; p++) { int64_t v = cp.start_of(p); fwrite(&v, 8, 1, f); }
        for (int p = 0; p <= P; p++) { int64_t v = sp.start_of(p); fwrite(&v, 8, 1, f); }
        if (!part_vertex)
            for (int p = 0; p < P; p++) {
                const EdgeList<Edge> &el = PG.get_edge_list_partition(p);
                int64_t cnt = (int64_t)el.get_num_edges();
                fwrite(&cnt, 8, 1, f);
                for (int64_t k = 0; k < cnt; k++) {
                    uint32_t sd[2] = {(uint32_t)el[k].getSource(), (uint32_t)el[k].getDestination()};
                    fwrite(sd, 4, 2, f);
                }
            }
        fclose(f);
    }
    std::vector<double> res;
    long iters = 0;
    std::vector<double> times;
    for (int r = 0; r < rounds; r++) {
        g_tracing = trace_path && r == 0;
        auto t0 = std::chrono::steady_clock::now();
        run_app(PG, start, res, iters);
        auto t1 = std::chrono::steady_clock::now();
        times.push_back(std::chrono::duration<double>(t1 - t0).count());
        if (PG.transposed()) PG.transpose();
    }
    printf("{\"n\": %llu, \"m\": %llu, \"iters\": %ld, \"times\": [", (unsigned long long)h.n,
           (unsigned long long)h.m, iters);
    for (size_t i = 0; i < times.size(); i++) printf("%s%.6f", i ? ", " : "", times[i]);
    printf("], \"threads\": %d}\n", omp_get_max_threads());
    if (trace_path) {
        FILE *f = fopen(trace_path, "wb");
        fwrite(g_trace.data(), 8, g_trace.size(), f);
        fclose(f);
    }
    if (out_path) {
        FILE *f = fopen(out_path, "wb");
        int64_t it = iters;
        fwrite(&it, 8, 1, f);
        fwrite(res.data(), 8, res.size(), f);
        fclose(f);
    }
    PG.del();
    G.del();
    return 0;
}

} // namespace

int main(int argc, char **argv)
{
    HostGraph h;
    int P = 384, rounds = 1;
    bool part_vertex = false;
    long start = 100;
    const char *in = nullptr, *out = nullptr, *part = nullptr, *xfile = nullptr, *trace = nullptr;
    bool relabel = false;
    int scale = 0, ef = 16;
    uint64_t seed = 1;
    bool sym = false;
#if GG_APP == 4 || GG_APP == 6
    h.weighted = true;
#endif
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() { return i + 1 < argc ? argv[++i] : (char *)""; };
        if (a == "-c") P = atoi(next());
        else if (a == "-v") part_vertex = std::string(next()) == "vertex";
        else if (a == "-r") start = atol(next());
        else if (a == "-rounds") rounds = atoi(next());
        else if (a == "-i") in = next();
        else if (a == "-o") out = next();
        else if (a == "-p") part = next();
        else if (a == "-x") xfile = next();
        else if (a == "-s") sym = true;
        else if (a == "-R") relabel = true;
        else if (a == "-t") trace = next();
        else if (a == "-rmat") { scale = atoi(next()); ef = atoi(next()); seed = strtoull(next(), 0, 10); }
        else { fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
    }
    if (in) {
        if (!read_ggb(in, h)) { fprintf(stderr, "cannot read %s\n", in); return 1; }
    } else if (scale > 0) {
        if (!gen_rmat(scale, ef, seed, sym, h)) { fprintf(stderr, "generator failed\n"); return 1; }
    } else {
        fprintf(stderr, "usage: %s [-c P] [-v edge|vertex] [-R] [-r root] [-rounds R] (-i g.ggb | -rmat S EF SEED [-s]) [-o out] [-p part.bin] [-t trace.bin]\n", argv[0]);
        return 2;
    }
#if GG_APP == 4
    if (xfile) {
        g_x.resize(h.n);
        FILE *f = fopen(xfile, "rb");
        if (!f || fread(g_x.data(), 8, h.n, f) != h.n) { fprintf(stderr, "bad -x file\n"); return 1; }
        fclose(f);
    }
#else
    (void)xfile;
#endif
    if (start < 0 || (uint64_t)start >= h.n) start = 0;
    if (h.symmetric) return drive<symmetricVertex>(h, P, part_vertex, relabel, start, rounds, out, part, trace);
    return drive<asymmetricVertex>(h, P, part_vertex, relabel, start, rounds, out, part, trace);
}
