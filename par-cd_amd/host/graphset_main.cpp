// pcd_graphset: builds and reads graph sets (many PCD graphs in one file).
//
//   pcd_graphset pack OUT IN.pcd...     packs PCD files into a set, in order
//   pcd_graphset list SET               one line per member: index, n, m
//   pcd_graphset extract SET I OUT.pcd  writes member I as a PCD file
//
// Exit status 0 on success, 1 on any error (the library's message on stderr).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "parcd.h"

static int usage() {
    fprintf(stderr,
            "usage: pcd_graphset pack OUT IN.pcd...\n"
            "       pcd_graphset list SET\n"
            "       pcd_graphset extract SET INDEX OUT.pcd\n");
    return 1;
}

static int failed(const char *what) {
    fprintf(stderr, "%s: %s\n", what, pcd_last_error());
    return 1;
}

int main(int argc, char **argv) {
    if (argc < 3) return usage();
    if (!strcmp(argv[1], "pack")) {
        std::vector<const char *> in(argv + 3, argv + argc);
        if (pcd_graphset_pack(argv[2], (int)in.size(), in.data()) != PCD_OK) return failed("pack");
        return 0;
    }
    if (!strcmp(argv[1], "list") && argc == 3) {
        pcd_graphset *s = nullptr;
        if (pcd_graphset_open(argv[2], &s) != PCD_OK) return failed("open");
        int count = 0, rc = 0;
        pcd_graphset_count(s, &count);
        for (int i = 0; i < count && rc == 0; i++) {
            int n, m;
            if (pcd_graphset_graph(s, i, &n, &m, nullptr, nullptr, nullptr) != PCD_OK) {
                rc = failed("list");
            } else {
                printf("%d %d %d\n", i, n, m);
            }
        }
        pcd_graphset_close(s);
        return rc;
    }
    if (!strcmp(argv[1], "extract") && argc == 5) {
        pcd_graphset *s = nullptr;
        if (pcd_graphset_open(argv[2], &s) != PCD_OK) return failed("open");
        char *end = nullptr;
        long i = strtol(argv[3], &end, 10);
        int rc = 0;
        if (!*argv[3] || *end || i < 0 || i > 0x7fffffff) {
            fprintf(stderr, "extract: bad index %s\n", argv[3]);
            rc = 1;
        } else if (pcd_graphset_extract(s, (int)i, argv[4]) != PCD_OK) {
            rc = failed("extract");
        }
        pcd_graphset_close(s);
        return rc;
    }
    return usage();
}
