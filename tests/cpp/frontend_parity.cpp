// Front-end parity program: app loops written against include/gg/ligra-numa.h
// (the drop-in C++ API) that write their results to a file, so the tests can
// compare the front end's device path with the oracle.
//
//   GG_APP=pagerank|bfs|cc|spmv GG_DUMP=out.bin frontend_parity[_w] <driver args>
//
// The functor structs follow the reference apps' interface (member names
// p_curr/p_next/V, Parents, IDs/prevIDs, damping/addedConstant), which is
// what the front end recognises; their host bodies restate the updates.
#include "ligra-numa.h"

namespace {

template <class vertex>
struct RankStep {  // PageRank edge functor (app/PageRank.C:28-70 interface)
    double *p_curr, *p_next;
    vertex *V;
    RankStep(double *c, double *x, vertex *v) : p_curr(c), p_next(x), V(v) {}
    bool update(intT s, intT d) { p_next[d] += p_curr[s] / V[s].getOutDegree(); return true; }
    bool updateAtomic(intT s, intT d) { writeAdd(&p_next[d], p_curr[s] / V[s].getOutDegree()); return true; }
    bool cond(intT) { return true; }
};

struct RankDamp {  // PR_Vertex_F interface (app/PageRank.C:73-87)
    double damping, addedConstant;
    double *p_curr, *p_next;
    RankDamp(double *c, double *x, double dmp, intT n)
        : damping(dmp), addedConstant((1 - dmp) * (1 / (double)n)), p_curr(c), p_next(x) {}
    bool operator()(intT i) { p_next[i] = damping * p_next[i] + addedConstant; return true; }
};

struct RankReset {  // PR_Vertex_Reset interface (app/PageRank.C:90-100)
    double *p_curr;
    explicit RankReset(double *c) : p_curr(c) {}
    bool operator()(intT i) { p_curr[i] = 0.0; return true; }
};

struct ParentStep {  // BFS_F interface (app/BFS.C:26-75)
    intT *Parents;
    explicit ParentStep(intT *p) : Parents(p) {}
    bool updateAtomic(intT s, intT d) { return CAS(&Parents[d], (intT)-1, s); }
    bool cond(intT d) { return Parents[d] == -1; }
};

struct LabelStep {  // CC_F interface (app/Components.C:26-86)
    intT *IDs, *prevIDs;
    LabelStep(intT *a, intT *b) : IDs(a), prevIDs(b) {}
    bool updateAtomic(intT s, intT d) { intT o = IDs[d]; return writeMin(&IDs[d], IDs[s]) && o == prevIDs[d]; }
    bool cond(intT) { return true; }
};

struct LabelCopy {  // CC_Vertex_F interface (app/Components.C:88-99)
    intT *IDs, *prevIDs;
    LabelCopy(intT *a, intT *b) : IDs(a), prevIDs(b) {}
    bool operator()(intT i) { prevIDs[i] = IDs[i]; return true; }
};

template <class vertex>
struct WeightedStep {  // SPMV_F interface (app/SPMV.C:7-48)
    double *p_curr, *p_next;
    WeightedStep(double *c, double *x) : p_curr(c), p_next(x) {}
    bool cond(intT) { return true; }
};

void dump(const std::vector<double> &v, long iters)
{
    const char *path = getenv("GG_DUMP");
    if (!path) return;
    FILE *f = fopen(path, "wb");
    int64_t it = iters;
    fwrite(&it, 8, 1, f);
    fwrite(v.data(), 8, v.size(), f);
    fclose(f);
}

} // namespace

template <class GraphType>
void Compute(GraphType &GA, long start)
{
    typedef typename GraphType::vertex_type vertex;
    const partitioner &part = GA.get_partitioner();
    const intT n = GA.n, m = GA.m;
    const std::string app = getenv("GG_APP") ? getenv("GG_APP") : "pagerank";
    std::vector<double> out(n);
    long iters = 0;
    if (app == "pagerank") {
        graph<vertex> &WG = GA.get_partition();
        mmap_ptr<double> c, x;
        c.part_allocate(part);
        x.part_allocate(part);
        const double q = 1 / (double)n;
        map_vertexL(part, [&](intT j) { c[j] = q; });
        map_vertexL(part, [&](intT j) { x[j] = 0; });
        partitioned_vertices F = partitioned_vertices::bits(part, n, m);
        bool broke = false;
        while (iters < 10) {
            iters++;
            partitioned_vertices next = edgeMap(GA, F, RankStep<vertex>(c, x, WG.V), m / 20);
            vertexMap(part, F, RankDamp(c, x, 0.85, n));
            map_vertexL(part, [&](intT j) { c[j] = fabs(c[j] - x[j]); });
            if (sequence::plusReduce(c.get(), n) < 1e-7) {
                broke = true;
                next.del();
                break;
            }
            vertexMap(part, F, RankReset(c));
            swap(c, x);
            F.del();
            F = next;
        }
        const double *r = broke ? x.get() : c.get();
        out.assign(r, r + n);
        F.del();
        c.del();
        x.del();
    } else if (app == "bfs") {
        mmap_ptr<intT> P;
        P.part_allocate(part);
        map_vertexL(part, [&](intT j) { P[j] = -1; });
        P[start] = (intT)start;
        partitioned_vertices F = partitioned_vertices::create(n, (intT)start, 0);
        while (!F.isEmpty()) {
            partitioned_vertices next = edgeMap(GA, F, ParentStep(P), m / 20);
            F.del();
            F = next;
            iters++;
        }
        const intT *p = P.get();
        for (intT i = 0; i < n; i++) out[i] = p[i];
        F.del();
        P.del();
    } else if (app == "cc") {
        mmap_ptr<intT> ids, prev;
        ids.part_allocate(part);
        prev.part_allocate(part);
        map_vertexL(part, [&](intT j) { ids[j] = j; });
        partitioned_vertices F = partitioned_vertices::bits(part, n, m);
        while (!F.isEmpty()) {
            vertexMap(part, F, LabelCopy(ids, prev));
            partitioned_vertices next = edgeMap(GA, F, LabelStep(ids, prev), m / 20);
            F.del();
            F = next;
            iters++;
        }
        const intT *l = ids.get();
        for (intT i = 0; i < n; i++) out[i] = l[i];
        F.del();
        ids.del();
        prev.del();
    } else {
        mmap_ptr<double> c, x;
        c.part_allocate(part);
        x.part_allocate(part);
        const double q = 1 / (double)n;
        map_vertexL(part, [&](intT j) { c[j] = q; });
        map_vertexL(part, [&](intT j) { x[j] = 0; });
        partitioned_vertices F = partitioned_vertices::bits(part, n, m);
        partitioned_vertices next = edgeMap(GA, F, WeightedStep<vertex>(c, x), m / 20);
        vertexMap(part, F, RankReset(c));
        swap(c, x);
        const double *y = c.get();
        out.assign(y, y + n);
        iters = 1;
        next.del();
        F.del();
        c.del();
        x.del();
    }
    dump(out, iters);
}
