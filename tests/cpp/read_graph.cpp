// Reads a graph file with the front end's readers (PBBS AdjacencyGraph /
// WeightedAdjacencyGraph text, IO.h:136-361, or Galois .gr binary with -b,
// IO.h:365-656) and writes the CSR it produced as raw arrays:
//   read_graph [-b] [-w] <in> <out.bin>
//   out.bin: u64 n, u64 m, u32 off[n+1], u32 ids[m], f32 w[m] (when -w)
// Host code only: no device call is made.
#define GG_FRONT_NO_DRIVER 1
#include "ligra-numa.h"

int main(int argc, char **argv)
{
    bool bin = false, w = false;
    int i = 1;
    for (; i < argc && argv[i][0] == '-'; i++) {
        if (!strcmp(argv[i], "-b")) bin = true;
        if (!strcmp(argv[i], "-w")) w = true;
    }
    if (i + 2 != argc) return 2;
    gg_front::HostCSR G;
    const bool ok = bin ? gg_front::read_galois(argv[i], w, G) : gg_front::read_text(argv[i], w, G);
    if (!ok) return 1;
    FILE *f = fopen(argv[i + 1], "wb");
    if (!f) return 3;
    uint64_t hdr[2] = {G.n, G.m};
    fwrite(hdr, 8, 2, f);
    fwrite(G.off.data(), 4, G.off.size(), f);
    fwrite(G.ids.data(), 4, G.ids.size(), f);
    if (w) fwrite(G.w.data(), 4, G.w.size(), f);
    fclose(f);
    return 0;
}
