// Host check of the front end's functor self-checks.  Built once per
// reference app source (the app's source fed to the compiler unchanged
// ahead of this file, with -DGG_FRONT_NO_DRIVER and -DCHECK_<APP>), it runs
// the self-check on that app's own functors and expects agreement; built
// alone, it expects the functors below, which carry recognised data members
// but compute something else, to be refused.  Exit status 0 when every
// expectation holds; one line per checked functor on stdout.
#ifndef GG_FRONT_NO_DRIVER
#define GG_FRONT_NO_DRIVER
#endif
#include "ligra-numa.h"
#undef main

#include <cstdio>

namespace {
int failures = 0;
template <int K, class F>
void expect(const char *name, const F &f, bool want)
{
    const bool got = gg_front::edge_op_agrees<K>(f);
    std::printf("%s %s\n", name, got ? "agrees" : "refused");
    if (got != want) failures++;
}
template <class F>
void expect_v(const char *name, const F &f, int kind, bool want)
{
    const bool got = gg_front::vertex_op_agrees(f, kind);
    std::printf("%s %s\n", name, got ? "agrees" : "refused");
    if (got != want) failures++;
}

// recognised as PR_F by its members, but scales the contribution
template <class vertex>
struct DoubledRank {
    double *p_curr, *p_next;
    vertex *V;
    bool update(intT s, intT d)
    {
        p_next[d] += 2 * p_curr[s] / V[s].getOutDegree();
        return 1;
    }
    bool updateAtomic(intT s, intT d)
    {
        writeAdd(&p_next[d], 2 * p_curr[s] / V[s].getOutDegree());
        return 1;
    }
};
// recognised as BFS_F, but overwrites parents already set
struct AnyParent {
    intT *Parents;
    bool update(intT s, intT d)
    {
        Parents[d] = s;
        return 1;
    }
    bool updateAtomic(intT s, intT d) { return update(s, d); }
    bool cond(intT) { return true; }
};
// recognised as CC_F, but takes the maximum label
struct MaxLabel {
    intT *IDs, *prevIDs;
    bool update(intT s, intT d)
    {
        if (IDs[s] > IDs[d]) {
            IDs[d] = IDs[s];
            return 1;
        }
        return 0;
    }
    bool updateAtomic(intT s, intT d) { return update(s, d); }
};
// a vertex functor recognised as PR_Vertex_Reset whose body adds one
struct NotAReset {
    double *p_curr;
    bool operator()(intT i)
    {
        p_curr[i] += 1;
        return 1;
    }
};
} // namespace

int main()
{
#if defined(CHECK_PAGERANK)
    expect<GG_OP_PAGERANK>("PageRank.C PR_F", PR_F<asymmetricVertex>(nullptr, nullptr, nullptr), true);
    expect_v("PageRank.C PR_Vertex_F", PR_Vertex_F(nullptr, nullptr, 0.85, 1000), GG_VOP_PAGERANK, true);
    expect_v("PageRank.C PR_Vertex_Reset", PR_Vertex_Reset(nullptr), GG_VOP_RESET, true);
#elif defined(CHECK_BFS)
    expect<GG_OP_BFS>("BFS.C BFS_F", BFS_F(nullptr), true);
#elif defined(CHECK_COMPONENTS)
    expect<GG_OP_CC>("Components.C CC_F", CC_F(nullptr, nullptr), true);
    expect_v("Components.C CC_Vertex_F", CC_Vertex_F(nullptr, nullptr), GG_VOP_CC_COPY, true);
#elif defined(CHECK_SPMV)
    expect<GG_OP_SPMV>("SPMV.C SPMV_F", SPMV_F<asymmetricVertex>(nullptr, nullptr), true);
#elif defined(CHECK_BC)
    expect<GG_OP_PATHSUM>("BC.C BC_F", BC_F(nullptr, nullptr), true);
    expect<GG_OP_PATHSUM>("BC.C BC_Back_F", BC_Back_F(nullptr, nullptr), true);
#elif defined(CHECK_BELLMANFORD)
    expect<GG_OP_SSSP>("BellmanFord.C BF_F", BF_F(nullptr, nullptr), true);
#elif defined(CHECK_PAGERANKDELTA)
    expect<GG_OP_PAGERANK>("PageRankDelta.C PR_F", PR_F<asymmetricVertex>(nullptr, nullptr, nullptr), true);
#elif defined(CHECK_PAGERANKCONVERAGE)
    expect<GG_OP_SUM>("PageRankConverage.C PR_F",
                      PR_F<asymmetricVertex>(nullptr, nullptr, nullptr, nullptr), true);
#elif defined(CHECK_FRONTEND_PARITY)
    // this repository's own parity program (tests/cpp/frontend_parity.cpp)
#ifndef WEIGHTED
    expect<GG_OP_PAGERANK>("RankStep", RankStep<asymmetricVertex>(nullptr, nullptr, nullptr), true);
    expect<GG_OP_BFS>("ParentStep", ParentStep(nullptr), true);
    expect<GG_OP_CC>("LabelStep", LabelStep(nullptr, nullptr), true);
    expect<GG_OP_PATHSUM>("PathStep", PathStep<double>(nullptr, nullptr), true);
    expect<GG_OP_PATHSUM>("DepStep", DepStep<float>(nullptr, nullptr), true);
    expect<GG_OP_PAGERANK>("DeltaStep", DeltaStep<asymmetricVertex>(nullptr, nullptr, nullptr), true);
    expect<GG_OP_SUM>("FactorStep", FactorStep<asymmetricVertex>(nullptr, nullptr, nullptr, nullptr), true);
    expect_v("RankDamp", RankDamp(nullptr, nullptr, 0.85, 1000), GG_VOP_PAGERANK, true);
    expect_v("RankReset", RankReset(nullptr), GG_VOP_RESET, true);
    expect_v("LabelCopy", LabelCopy(nullptr, nullptr), GG_VOP_CC_COPY, true);
#else
    expect<GG_OP_SPMV>("WeightedStep", WeightedStep<asymmetricVertex>(nullptr, nullptr), true);
    expect<GG_OP_SSSP>("RelaxStep", RelaxStep(nullptr, nullptr), true);
#endif
#else
    expect<GG_OP_PAGERANK>("DoubledRank", DoubledRank<asymmetricVertex>{nullptr, nullptr, nullptr}, false);
    expect<GG_OP_BFS>("AnyParent", AnyParent{nullptr}, false);
    expect<GG_OP_CC>("MaxLabel", MaxLabel{nullptr, nullptr}, false);
    expect_v("NotAReset", NotAReset{nullptr}, GG_VOP_RESET, false);
#endif
    return failures ? 1 : 0;
}
