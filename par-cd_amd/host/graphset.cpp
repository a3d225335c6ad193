// Graph sets: many PCD graphs in one file, for batching and serving.
//
// Layout (native byte order, as the PCD format itself):
//   bytes 0-7     "PCDSET1\0"
//   bytes 8-15    int64 count
//   bytes 16-     count x {int64 offset, int64 bytes}: where member i starts
//                 and how long it is
//   members       each the exact bytes of one PCD file (graph.cpp:40-105
//                 reads them, convert.cpp:244-278 writes them): "PCD", n, m,
//                 indices[n + 1], adj[2m], map[n]
// A member starts 11 bytes before a 64-byte boundary, so that its arrays
// are 4-byte aligned and its indices 64-byte aligned in a mapping of the
// file; the gaps are zero bytes, members follow one another in table order
// and the file ends with the last member.  The layout is canonical: writing
// the graphs that a set holds gives the same bytes again.
//
// The reader maps the file read-only and hands out pointers into the
// mapping, so a batch or a server admits a set's graphs without copying
// them into separate host arrays first.
#include <sys/mman.h>
#include <sys/stat.h>
#include <fcntl.h>
#include <unistd.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "parcd.h"
#include "hostutil.hpp"

namespace {

constexpr char SET_MAGIC[8] = {'P', 'C', 'D', 'S', 'E', 'T', '1', '\0'};
constexpr int64_t SET_HEADER = 16, SET_ENTRY = 16, MEMBER_HEADER = 11, MEMBER_ALIGN = 64;

// first offset >= at where a member may start
int64_t member_start(int64_t at) {
    int64_t r = (at + MEMBER_HEADER) % MEMBER_ALIGN;
    return r == 0 ? at : at + (MEMBER_ALIGN - r);
}

int64_t member_bytes(int64_t n, int64_t m) { return MEMBER_HEADER + 4 * (2 * n + 1 + 2 * m); }

// the structural checks pcd_read_pcd makes on a graph's arrays
bool graph_ok(int n, int m, const int32_t *ind, const int32_t *adj) {
    if (n < 0 || m < 0 || !ind || (m > 0 && !adj)) return false;
    if (ind[0] != 0 || (int64_t)ind[n] != 2 * (int64_t)m) return false;
    for (int i = 0; i < n; i++)
        if (ind[i + 1] < ind[i]) return false;
    for (int64_t j = 0; j < 2 * (int64_t)m; j++)
        if (adj[j] < 0 || adj[j] >= n) return false;
    return true;
}

}  // namespace

struct pcd_graphset {
    const uint8_t *base = nullptr;
    int64_t size = 0;
    int count = 0;
    std::vector<int64_t> offset;
    // per member: 0 not checked yet, 1 valid, 2 malformed
    std::unique_ptr<std::atomic<uint8_t>[]> state;
};

struct pcd_graphset_writer {
    FILE *fp = nullptr;
    std::string path;
    int count = 0;
    int64_t at = 0;  // bytes written so far
    std::vector<int64_t> table;
};

using pcd::fail;

extern "C" int pcd_graphset_open(const char *path, pcd_graphset **out) {
    if (!path || !out) return fail(PCD_E_INVALID, "bad arguments");
    *out = nullptr;
    int fd = open(path, O_RDONLY);
    if (fd < 0) return fail(PCD_E_IO, std::string("failed to open ") + path);
    struct stat sb;
    if (fstat(fd, &sb) != 0) {
        close(fd);
        return fail(PCD_E_IO, std::string("failed to stat ") + path);
    }
    const int64_t size = sb.st_size;
    if (size < SET_HEADER) {
        close(fd);
        return fail(PCD_E_FORMAT, std::string("not a graph set (too short): ") + path);
    }
    void *p = mmap(nullptr, (size_t)size, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (p == MAP_FAILED) return fail(PCD_E_IO, std::string("failed to map ") + path);
    std::unique_ptr<pcd_graphset> s(new pcd_graphset);
    s->base = (const uint8_t *)p;
    s->size = size;
    auto bad = [&](const std::string &why) {
        munmap(p, (size_t)size);
        return fail(PCD_E_FORMAT, std::string("malformed graph set ") + path + ": " + why);
    };
    if (memcmp(s->base, SET_MAGIC, 8) != 0) return bad("invalid file type");
    int64_t count;
    memcpy(&count, s->base + 8, 8);
    if (count < 0 || count > 0x7fffffff || count > (size - SET_HEADER) / SET_ENTRY) return bad("bad member count");
    const int64_t table_end = SET_HEADER + SET_ENTRY * count;
    int64_t prev_end = table_end;
    s->count = (int)count;
    s->offset.resize((size_t)count);
    for (int64_t i = 0; i < count; i++) {
        int64_t off, len;
        memcpy(&off, s->base + SET_HEADER + SET_ENTRY * i, 8);
        memcpy(&len, s->base + SET_HEADER + SET_ENTRY * i + 8, 8);
        const std::string who = "member " + std::to_string(i);
        if (off != member_start(prev_end)) return bad(who + " is not where the layout places it");
        if (len < member_bytes(0, 0) || off > size || len > size - off) return bad(who + " runs past the end");
        for (int64_t b = prev_end; b < off; b++)
            if (s->base[b] != 0) return bad("non-zero bytes before " + who);
        const uint8_t *q = s->base + off;
        int32_t n, m;
        memcpy(&n, q + 3, 4);
        memcpy(&m, q + 7, 4);
        if (memcmp(q, "PCD", 3) != 0) return bad(who + " is not a PCD graph");
        if (n < 0 || m < 0 || member_bytes(n, m) != len) return bad(who + " sizes disagree with its length");
        s->offset[(size_t)i] = off;
        prev_end = off + len;
    }
    if (prev_end != size) return bad("trailing bytes");
    s->state.reset(new std::atomic<uint8_t>[(size_t)(count ? count : 1)]);
    for (int64_t i = 0; i < count; i++) s->state[(size_t)i].store(0);
    *out = s.release();
    return PCD_OK;
}

extern "C" int pcd_graphset_close(pcd_graphset *s) {
    if (!s) return fail(PCD_E_INVALID, "bad arguments");
    munmap((void *)s->base, (size_t)s->size);
    delete s;
    return PCD_OK;
}

extern "C" int pcd_graphset_count(const pcd_graphset *s, int *count) {
    if (!s || !count) return fail(PCD_E_INVALID, "bad arguments");
    *count = s->count;
    return PCD_OK;
}

extern "C" int pcd_graphset_graph(pcd_graphset *s, int i, int *n, int *m, const int32_t **indices, const int32_t **adj,
                             const int32_t **map) {
    if (!s || i < 0 || i >= s->count) return fail(PCD_E_INVALID, "bad arguments");
    const uint8_t *q = s->base + s->offset[(size_t)i];
    int32_t nn, mm;
    memcpy(&nn, q + 3, 4);
    memcpy(&mm, q + 7, 4);
    const int32_t *ind = (const int32_t *)(q + MEMBER_HEADER);
    const int32_t *a = ind + (int64_t)nn + 1;
    const int32_t *mp = a + 2 * (int64_t)mm;
    uint8_t st = s->state[(size_t)i].load(std::memory_order_acquire);
    if (st == 0) {
        st = graph_ok(nn, mm, ind, a) ? 1 : 2;
        s->state[(size_t)i].store(st, std::memory_order_release);
    }
    if (st != 1) return fail(PCD_E_FORMAT, "graph set member " + std::to_string(i) + " is not a valid CSR graph");
    if (n) *n = nn;
    if (m) *m = mm;
    if (indices) *indices = ind;
    if (adj) *adj = a;
    if (map) *map = mp;
    return PCD_OK;
}

extern "C" int pcd_graphset_extract(pcd_graphset *s, int i, const char *path) {
    if (!path) return fail(PCD_E_INVALID, "bad arguments");
    int rc = pcd_graphset_graph(s, i, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (rc != PCD_OK) return rc;
    int32_t nn, mm;
    const uint8_t *q = s->base + s->offset[(size_t)i];
    memcpy(&nn, q + 3, 4);
    memcpy(&mm, q + 7, 4);
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail(PCD_E_IO, std::string("failed to open ") + path);
    const size_t len = (size_t)member_bytes(nn, mm);
    bool ok = fwrite(q, 1, len, fp) == len;
    ok = (fclose(fp) == 0) && ok;
    return ok ? PCD_OK : fail(PCD_E_IO, std::string("failed to write ") + path);
}

extern "C" int pcd_graphset_writer_open(const char *path, int count, pcd_graphset_writer **out) {
    if (!path || count < 0 || !out) return fail(PCD_E_INVALID, "bad arguments");
    *out = nullptr;
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail(PCD_E_IO, std::string("failed to open ") + path);
    std::unique_ptr<pcd_graphset_writer> w(new pcd_graphset_writer);
    w->fp = fp;
    w->path = path;
    w->count = count;
    const int64_t c = count;
    bool ok = fwrite(SET_MAGIC, 1, 8, fp) == 8 && fwrite(&c, 8, 1, fp) == 1;
    std::vector<uint8_t> zeros((size_t)(SET_ENTRY * count), 0);
    ok = ok && fwrite(zeros.data(), 1, zeros.size(), fp) == zeros.size();
    if (!ok) {
        fclose(fp);
        remove(path);
        return fail(PCD_E_IO, std::string("failed to write ") + path);
    }
    w->at = SET_HEADER + SET_ENTRY * c;
    *out = w.release();
    return PCD_OK;
}

extern "C" int pcd_graphset_writer_add(pcd_graphset_writer *w, int n, int m, const int32_t *indices, const int32_t *adj,
                                  const int32_t *map) {
    if (!w) return fail(PCD_E_INVALID, "bad arguments");
    if ((int)(w->table.size() / 2) >= w->count) return fail(PCD_E_STATE, "the graph set is full");
    if (!graph_ok(n, m, indices, adj)) return fail(PCD_E_INVALID, "not a valid CSR graph");
    const int64_t off = member_start(w->at);
    static const uint8_t zeros[MEMBER_ALIGN] = {0};
    bool ok = fwrite(zeros, 1, (size_t)(off - w->at), w->fp) == (size_t)(off - w->at);
    ok = ok && fwrite("PCD", 1, 3, w->fp) == 3 && fwrite(&n, 4, 1, w->fp) == 1 && fwrite(&m, 4, 1, w->fp) == 1;
    ok = ok && fwrite(indices, 4, (size_t)n + 1, w->fp) == (size_t)n + 1;
    if (m > 0) ok = ok && fwrite(adj, 4, (size_t)2 * m, w->fp) == (size_t)2 * m;
    if (map) {
        ok = ok && fwrite(map, 4, (size_t)n, w->fp) == (size_t)n;
    } else {
        // no map given: vertex v keeps id v
        std::vector<int32_t> id((size_t)n);
        for (int v = 0; v < n; v++) id[(size_t)v] = v;
        ok = ok && fwrite(id.data(), 4, (size_t)n, w->fp) == (size_t)n;
    }
    if (!ok) return fail(PCD_E_IO, "failed to write " + w->path);
    const int64_t len = member_bytes(n, m);
    w->table.push_back(off);
    w->table.push_back(len);
    w->at = off + len;
    return PCD_OK;
}

extern "C" int pcd_graphset_writer_close(pcd_graphset_writer *w) {
    if (!w) return fail(PCD_E_INVALID, "bad arguments");
    std::unique_ptr<pcd_graphset_writer> own(w);
    if ((int)(w->table.size() / 2) != w->count) {
        fclose(w->fp);
        remove(w->path.c_str());
        return fail(PCD_E_STATE, "graph set closed with " + std::to_string(w->table.size() / 2) + " of " +
                                     std::to_string(w->count) + " graphs; file removed");
    }
    bool ok = fseek(w->fp, SET_HEADER, SEEK_SET) == 0;
    if (w->count > 0) ok = ok && fwrite(w->table.data(), 8, w->table.size(), w->fp) == w->table.size();
    ok = (fclose(w->fp) == 0) && ok;
    if (!ok) {
        remove(w->path.c_str());
        return fail(PCD_E_IO, "failed to write " + w->path);
    }
    return PCD_OK;
}

extern "C" int pcd_graphset_pack(const char *path, int count, const char *const *pcd_paths) {
    if (!path || count < 0 || (count > 0 && !pcd_paths)) return fail(PCD_E_INVALID, "bad arguments");
    pcd_graphset_writer *w = nullptr;
    int rc = pcd_graphset_writer_open(path, count, &w);
    if (rc != PCD_OK) return rc;
    for (int i = 0; i < count; i++) {
        int n, m;
        int32_t *ind, *adj, *mp;
        rc = pcd_read_pcd(pcd_paths[i], &n, &m, &ind, &adj, &mp);
        if (rc == PCD_OK) {
            rc = pcd_graphset_writer_add(w, n, m, ind, adj, mp);
            free(ind);
            free(adj);
            free(mp);
        }
        if (rc != PCD_OK) {
            std::string why = pcd_last_error();
            fclose(w->fp);
            remove(w->path.c_str());
            delete w;
            return fail(rc, why);
        }
    }
    return pcd_graphset_writer_close(w);
}
