// pcd_convert: drop-in for the reference `convert` (convert.cpp:281-337):
// text edge list -> PCD file with the same bytes.  The per-row adjacency
// dump the reference prints to stdout (convert.cpp:259-267) is written only
// with -v.
#include <cstdio>
#include <cstring>
#include <iostream>

#include "parcd.h"

using namespace std;

int main(int argc, char *argv[]) {
    bool verbose = false;
    int argi = 1;
    if (argc > 1 && strcmp(argv[1], "-v") == 0) { verbose = true; argi++; }
    if (argc - argi < 2) {
        cerr << "usage: " << argv[0] << " [-v] <input file> <output file>" << endl;
        return -1;
    }
    const char *in = argv[argi], *outp = argv[argi + 1];
    cout << "reading graph from file " << in << endl;
    int64_t ne = 0;
    int32_t *a = nullptr, *b = nullptr;
    int nn = 0;
    if (pcd_read_edge_text(in, &ne, &a, &b, &nn) != PCD_OK) {
        cerr << pcd_last_error() << endl;
        return -1;
    }
    cout << "found " << nn << " nodes and " << ne << " edges" << endl;
    int n = 0, m = 0;
    int32_t *ind = nullptr, *adj = nullptr, *map = nullptr;
    if (pcd_convert_edges(nn, ne, a, b, &n, &m, &ind, &adj, &map) != PCD_OK) {
        cerr << pcd_last_error() << endl;
        return -1;
    }
    cout << "final graph has " << n << " nodes and " << m << " edges" << endl;
    if (pcd_write_pcd(outp, n, m, ind, adj, map) != PCD_OK) {
        cerr << pcd_last_error() << endl;
        return -1;
    }
    cout << n << " " << m << endl;
    if (verbose) {
        for (int i = 0; i < n; i++) {
            cout << i << "] " << ind[i] << " " << ind[i + 1] << " | ";
            for (int j = ind[i]; j < ind[i + 1]; j++) cout << adj[j] << " ";
            cout << "\n";
        }
    }
    pcd_free(a); pcd_free(b); pcd_free(ind); pcd_free(adj); pcd_free(map);
    return 0;
}
