// Front-end parity program: app loops written against include/gg/ligra-numa.h
// (the drop-in C++ API) that write their results to a file, so the tests can
// compare the front end's device path with the oracle.
//
//   GG_APP=pagerank|prbit|bfs|cc|spmv|bc|bc32|sssp|prdelta|prconv|factories GG_DUMP=out.bin
//       frontend_parity[_w] <driver args>
//
// The functor structs follow the reference apps' interface (member names
// p_curr/p_next/V, Parents, IDs/prevIDs, damping/addedConstant, NumPaths/
// Dependencies/Visited, ShortestPathLen/Visited, Delta/nghSum), which is
// what the front end recognises; their host bodies restate the updates.
// bc, sssp, prdelta and prconv follow the loops of app/BC.C,
// app/BellmanFord.C, app/PageRankDelta.C and app/PageRankConverage.C,
// including their host-side vertexMap / vertexFilter functors, BC's
// GA.transpose() and PageRankConverage's map_partition sums.
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
    bool update(intT s, intT d, intE w) { p_next[d] += p_curr[s] * w; return true; }
    bool updateAtomic(intT s, intT d, intE w) { writeAdd(&p_next[d], p_curr[s] * w); return true; }
    bool cond(intT) { return true; }
};

template <class T>
struct PathStep {  // BC_F interface (app/BC.C:28-77)
    T *NumPaths;
    bool *Visited;
    PathStep(T *a, bool *v) : NumPaths(a), Visited(v) {}
    bool update(intT s, intT d) { T o = NumPaths[d]; NumPaths[d] += NumPaths[s]; return o == T(0); }
    bool cond(intT d) { return !Visited[d]; }
};

template <class T>
struct DepStep {  // BC_Back_F interface (app/BC.C:80-124)
    T *Dependencies;
    bool *Visited;
    DepStep(T *a, bool *v) : Dependencies(a), Visited(v) {}
    bool update(intT s, intT d) { T o = Dependencies[d]; Dependencies[d] += Dependencies[s]; return o == T(0); }
    bool cond(intT d) { return !Visited[d]; }
};

struct MarkVisited {  // BC_Vertex_F (app/BC.C:127-137): runs on the host
    bool *Visited;
    explicit MarkVisited(bool *v) : Visited(v) {}
    bool operator()(intT i) { Visited[i] = true; return true; }
};

template <class T>
struct MarkBack {  // BC_Back_Vertex_F (app/BC.C:141-152): runs on the host
    bool *Visited;
    T *Dependencies, *inverseNumPaths;
    MarkBack(bool *v, T *d, T *i) : Visited(v), Dependencies(d), inverseNumPaths(i) {}
    bool operator()(intT i) { Visited[i] = true; Dependencies[i] += inverseNumPaths[i]; return true; }
};

struct RelaxStep {  // BF_F interface (app/BellmanFord.C:28-92)
    intT *ShortestPathLen, *Visited;
    RelaxStep(intT *a, intT *v) : ShortestPathLen(a), Visited(v) {}
    bool updateAtomic(intT s, intT d, intE w)
    {
        return writeMin(&ShortestPathLen[d], ShortestPathLen[s] + w) && CAS(&Visited[d], intT(0), intT(1));
    }
    bool cond(intT) { return true; }
};

struct ClearVisited {  // BF_Vertex_F (app/BellmanFord.C:94-103): host
    intT *Visited;
    explicit ClearVisited(intT *v) : Visited(v) {}
    bool operator()(intT i) { Visited[i] = 0; return true; }
};

template <class vertex>
struct DeltaStep {  // PR_F interface (app/PageRankDelta.C:27-72)
    vertex *V;
    double *Delta, *nghSum;
    DeltaStep(vertex *v, double *d, double *s) : V(v), Delta(d), nghSum(s) {}
    bool update(intT s, intT d) { nghSum[d] += Delta[s] / V[s].getOutDegree(); return true; }
    bool updateAtomic(intT s, intT d) { writeAdd(&nghSum[d], Delta[s] / V[s].getOutDegree()); return true; }
    bool cond(intT) { return true; }
};

struct DeltaFirst {  // PR_Vertex_F_FirstRound (app/PageRankDelta.C:74-93)
    double damping, addedConstant, one_over_n, epsilon2;
    double *p, *Delta, *nghSum;
    DeltaFirst(double *p_, double *d, double *s, double dmp, double q, double e2)
        : damping(dmp), addedConstant((1 - dmp) * q), one_over_n(q), epsilon2(e2), p(p_), Delta(d), nghSum(s) {}
    bool operator()(intT i)
    {
        Delta[i] = damping * (p[i] + nghSum[i]) + addedConstant - p[i];
        p[i] += Delta[i];
        Delta[i] -= one_over_n;
        return fabs(Delta[i]) > epsilon2 * p[i];
    }
};

struct DeltaNext {  // PR_Vertex_F (app/PageRankDelta.C:95-110)
    double damping, epsilon2;
    double *p, *Delta, *nghSum;
    DeltaNext(double *p_, double *d, double *s, double dmp, double e2)
        : damping(dmp), epsilon2(e2), p(p_), Delta(d), nghSum(s) {}
    bool operator()(intT i)
    {
        Delta[i] = nghSum[i] * damping;
        p[i] += Delta[i];
        return fabs(Delta[i]) > epsilon2 * p[i];
    }
};

struct SumReset {  // PR_Vertex_Reset (app/PageRankDelta.C:112-121): host
    double *nghSum;
    explicit SumReset(double *s) : nghSum(s) {}
    bool operator()(intT i) { nghSum[i] = 0.0; return true; }
};

template <class vertex>
struct FactorStep {  // PR_F interface of app/PageRankConverage.C:34-88
    double *p_curr, *p_next;
    vertex *V;
    double *add_factor;
    FactorStep(double *c, double *x, vertex *v, double *a) : p_curr(c), p_next(x), V(v), add_factor(a) {}
    bool update(intT s, intT d) { p_next[d] += add_factor[s]; return true; }
    bool updateAtomic(intT s, intT d) { writeAdd(&p_next[d], add_factor[s]); return true; }
    bool cond(intT) { return true; }
};

struct AddConst {  // PR_Vertex_Csum (app/PageRankConverage.C:107-119): host
    double addedConstant;
    double *p_next;
    AddConst(double *x, double c) : addedConstant(c), p_next(x) {}
    bool operator()(intT i) { p_next[i] = p_next[i] + addedConstant; return true; }
};

struct Normalise {  // PR_Vertex_norm (app/PageRankConverage.C:121-134): host
    double csum;
    double *p_next;
    Normalise(double *x, double c) : csum(c), p_next(x) {}
    bool operator()(intT i) { p_next[i] = p_next[i] / csum; return true; }
};

// Partitioned sums of app/PageRankConverage.C:166-235 (per-partition serial
// sums, then over the partitions in order).
double part_sum(const partitioner &part, double *a, intT n, bool scale)
{
    const int np = part.get_num_partitions();
    std::vector<double> ps(np);
    map_partition(k, part, {
        double d = 0.;
        for (intT i = part.start_of(k); i < part.start_of(k + 1); i++) d += a[i];
        ps[k] = d;
    });
    double d = 0.;
    for (int i = 0; i < np; i++) d += ps[i];
    return scale ? (1 - d) / n : d;
}

double part_normdiff(const partitioner &part, double *a, double *b)
{
    const int np = part.get_num_partitions();
    std::vector<double> ps(np);
    map_partition(k, part, {
        double d = 0.;
        for (intT i = part.start_of(k); i < part.start_of(k + 1); i++) d += fabs(a[i] - b[i]);
        ps[k] = d;
    });
    double d = 0.;
    for (int i = 0; i < np; i++) d += ps[i];
    return d;
}

// BC from 'start' (app/BC.C:154-226); returns the forward round count.
template <class T, class GraphType>
long run_bc(GraphType &GA, long start, std::vector<double> &out)
{
    const partitioner &part = GA.get_partitioner();
    const intT n = GA.n, m = GA.m;
    mmap_ptr<T> NumPaths;
    NumPaths.part_allocate(part);
    mmap_ptr<bool> Visited;
    Visited.part_allocate(part);
    map_vertexL(part, [&](intT j) { NumPaths[j] = 0.0; });
    map_vertexL(part, [&](intT j) { Visited[j] = false; });
    NumPaths[start] = 1.0;
    Visited[start] = true;
    partitioned_vertices Frontier = partitioned_vertices::create(n, (intT)start, 0);
    std::vector<partitioned_vertices> Levels;
    Levels.push_back(Frontier);
    long round = 0;
    while (!Frontier.isEmpty()) {
        round++;
        partitioned_vertices output = edgeMap(GA, Frontier, PathStep<T>(NumPaths, Visited), m / 20);
        vertexMap(part, output, MarkVisited(Visited));
        Levels.push_back(output);
        Frontier = output;
    }
    mmap_ptr<T> Dependencies;
    Dependencies.part_allocate(part);
    map_vertexL(part, [&](intT j) { Dependencies[j] = 0.0; });
    mmap_ptr<T> inverseNumPaths;
    inverseNumPaths = NumPaths;
    map_vertexL(part, [&](intT j) { inverseNumPaths[j] = 1 / inverseNumPaths[j]; });
    Levels[round].del();
    map_vertexL(part, [&](intT j) { Visited[j] = false; });
    Frontier = Levels[round - 1];
    vertexMap(part, Frontier, MarkBack<T>(Visited, Dependencies, inverseNumPaths));
    GA.transpose();
    for (long r = round - 2; r >= 0; r--) {
        partitioned_vertices output = edgeMap(GA, Frontier, DepStep<T>(Dependencies, Visited), m / 20);
        output.del();
        Frontier.del();
        Frontier = Levels[r];
        vertexMap(part, Frontier, MarkBack<T>(Visited, Dependencies, inverseNumPaths));
    }
    Frontier.del();
    map_vertexL(part, [&](intT j) { Dependencies[j] = (Dependencies[j] - inverseNumPaths[j]) / inverseNumPaths[j]; });
    const T *d = Dependencies.get();
    for (intT i = 0; i < n; i++) out[i] = (double)d[i];
    inverseNumPaths.del();
    Visited.del();
    Dependencies.del();
    return round;
}

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
    if (app == "pagerank" || app == "prbit") {
        // prbit: app/PageRankBit.C, which marks each output all-active
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
            if (app == "prbit") next.bit = true; // app/PageRankBit.C:145
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
    } else if (app == "bc") {
        iters = run_bc<double>(GA, start, out);
    } else if (app == "bc32") {
        iters = run_bc<float>(GA, start, out);
    } else if (app == "sssp") {
        // app/BellmanFord.C:105-146
        mmap_ptr<intT> SP, Visited;
        SP.part_allocate(part);
        Visited.part_allocate(part);
        map_vertexL(part, [&](intT j) { SP[j] = INT_MAX / 2; });
        SP[start] = 0;
        map_vertexL(part, [&](intT j) { Visited[j] = 0; });
        partitioned_vertices F = partitioned_vertices::create(n, (intT)start, 0);
        while (!F.isEmpty()) {
            if (iters == n) {
                map_vertexL(part, [&](intT j) { SP[j] = -(INT_MAX / 2); });
                break;
            }
            partitioned_vertices next = edgeMap(GA, F, RelaxStep(SP, Visited), m / 20);
            vertexMap(part, next, ClearVisited(Visited));
            F.del();
            F = next;
            iters++;
        }
        F.del();
        const intT *d = SP.get();
        for (intT i = 0; i < n; i++) out[i] = d[i];
        SP.del();
        Visited.del();
    } else if (app == "factories") {
        // partitioned_vertices factories (ligra-numa.h:103-208): out = the
        // active count of dense(n), sparse(n), boolean(every third vertex),
        // indice(vertices 0..9), then each frontier's out-edge sum
        mmap_ptr<bool> flags;
        flags.part_allocate(part);
        map_vertexL(part, [&](intT j) { flags[j] = j % 3 == 0; });
        mmap_ptr<intT> ids;
        ids.part_allocate(part);
        map_vertexL(part, [&](intT j) { ids[j] = j; });
        partitioned_vertices a = partitioned_vertices::dense(n, part);
        partitioned_vertices b = partitioned_vertices::sparse(n);
        partitioned_vertices c = partitioned_vertices::boolean(n, flags, 0, 0);
        partitioned_vertices d = partitioned_vertices::indice(n, ids, std::min<intT>(10, n), 0);
        out.assign(n, 0.0);
        partitioned_vertices *fs[4] = {&a, &b, &c, &d};
        for (int k = 0; k < 4 && 2 * k + 1 < n; k++) {
            out[2 * k] = fs[k]->numNonzeros();
            out[2 * k + 1] = fs[k]->num_out_edges;
            if (fs[k]->numRows() != n) out[2 * k] = -1;
        }
        iters = a.isEmpty() && b.isEmpty() && !c.isEmpty() && !d.isEmpty();
        a.del();
        b.del();
        c.del();
        d.del();
        flags.del();
        ids.del();
    } else if (app == "prconv") {
        // app/PageRankConverage.C:238-310
        graph<vertex> &WG = GA.get_partition();
        const double epsilon = 0.0000001, q = 1 / (double)n;
        mmap_ptr<double> c, x, ef, af;
        c.part_allocate(part);
        x.part_allocate(part);
        ef.part_allocate(part);
        af.part_allocate(part);
        map_vertexL(part, [&](intT j) { c[j] = q; });
        map_vertexL(part, [&](intT j) { x[j] = 0; });
        map_vertexL(part, [&](intT j) { ef[j] = (double)0.85 / (WG.V[j].getOutDegree()); });
        double l1 = 2;
        partitioned_vertices F = partitioned_vertices::bits(part, n, m);
        while (iters < 100 && l1 > epsilon) {
            map_vertexL(part, [&](intT j) { af[j] = (double)ef[j] * c[j]; });
            partitioned_vertices output = edgeMap(GA, F, FactorStep<vertex>(c, x, WG.V, af), m / 20);
            vertexMap(part, F, AddConst(x, part_sum(part, x, n, true)));
            l1 = part_normdiff(part, c, x);
            vertexMap(part, F, Normalise(x, part_sum(part, x, n, false)));
            vertexMap(part, F, RankReset(c));
            iters++;
            swap(c, x);
            output.del();
        }
        const double *r = c.get();
        out.assign(r, r + n);
        F.del();
        c.del();
        x.del();
        ef.del();
        af.del();
    } else if (app == "prdelta") {
        // app/PageRankDelta.C:124-176
        graph<vertex> &WG = GA.get_partition();
        const double damping = 0.85, epsilon = 0.0000001, epsilon2 = 0.01, q = 1 / (double)n;
        mmap_ptr<double> p, nghSum, Delta;
        p.part_allocate(part);
        nghSum.part_allocate(part);
        Delta.part_allocate(part);
        map_vertexL(part, [&](intT j) { p[j] = 0.0; });
        map_vertexL(part, [&](intT j) { Delta[j] = q; });
        map_vertexL(part, [&](intT j) { nghSum[j] = 0.0; });
        partitioned_vertices F = partitioned_vertices::bits(part, n, m);
        partitioned_vertices All = partitioned_vertices::bits(part, n, m);
        while (1) {
            iters++;
            partitioned_vertices output = edgeMap(GA, F, DeltaStep<vertex>(WG.V, Delta, nghSum), m / 20);
            output.del();
            partitioned_vertices active = iters == 1
                                              ? vertexFilter(GA, All, DeltaFirst(p, Delta, nghSum, damping, q, epsilon2))
                                              : vertexFilter(GA, All, DeltaNext(p, Delta, nghSum, damping, epsilon2));
            map_vertexL(part, [&](intT j) { nghSum[j] = fabs(Delta[j]); });
            const double l1 = sequence::plusReduce(nghSum.get(), n);
            if (l1 < epsilon) {
                active.del();
                break;
            }
            vertexMap(part, All, SumReset(nghSum));
            F.del();
            F = active;
        }
        const double *r = p.get();
        out.assign(r, r + n);
        F.del();
        All.del();
        p.del();
        Delta.del();
        nghSum.del();
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
