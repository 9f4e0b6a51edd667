// vebo -- the VEBO reorder tool on the GPU library (graphtools/VEBO.C:56-395
// is the reference's):
//   vebo [-s] [-r start] [-p P] <in.gr> <out.adj>
// Reads a Galois .gr graph, computes the VEBO order for P partitions
// (gg_vebo_order), renumbers the graph on the device (gg_graph_relabel) and
// writes it as a PBBS AdjacencyGraph text file (graphtools/graphIO.h:82-107).
// Like the reference it reports where vertex 'start' went.
#define GG_FRONT_NO_DRIVER 1
#include "ligra-numa.h"

int main(int argc, char **argv)
{
    commandLine P(argc, argv, "[-s] [-r <start>] [-p <part>] <inFile> <outFile>");
    const bool sym = P.getOption("-s");
    const long start = P.getOptionLongValue("-r", 100);
    const long part = P.getOptionLongValue("-p", 384);
    if (argc < 3) P.badArgument();
    const char *in = argv[argc - 2], *out = argv[argc - 1];
    gg_front::HostCSR G;
    if (!gg_front::read_galois(in, false, G)) {
        std::cerr << "cannot read graph " << in << std::endl;
        return 1;
    }
    gg_ctx_t ctx = gg_front::context();
    gg_graph_t g = nullptr, h = nullptr;
    gg_front::check(gg_graph_create_csr(ctx, G.n, G.m, G.off.data(), G.ids.data(), nullptr, sym, &g), "load");
    std::vector<uint32_t> newid(G.n);
    timer t;
    t.start();
    gg_front::check(gg_vebo_order(g, (int)part, newid.data(), nullptr), "gg_vebo_order");
    gg_front::check(gg_graph_relabel(g, newid.data(), &h), "gg_graph_relabel");
    std::cerr << "Reorder: " << t.stop() << std::endl;
    if (start >= 0 && start < (long)G.n)
        std::cerr << "Trace start vertex, new: " << newid[start] << " old: " << start << std::endl;
    std::vector<uint32_t> off(G.n + 1), ids(G.m ? G.m : 1);
    gg_front::check(gg_graph_copy_csr(h, off.data(), ids.data(), nullptr), "copy");
    FILE *f = fopen(out, "w");
    if (!f) {
        std::cerr << "cannot write " << out << std::endl;
        return 1;
    }
    fprintf(f, "AdjacencyGraph\n%u\n%llu\n", G.n, (unsigned long long)G.m);
    for (uint32_t v = 0; v < G.n; v++) fprintf(f, "%u\n", off[v]);
    for (uint64_t e = 0; e < G.m; e++) fprintf(f, "%u\n", ids[e]);
    fclose(f);
    gg_graph_destroy(h);
    gg_graph_destroy(g);
    return 0;
}
