// compress_labels (framework.cpp:146-170) on caller-owned device arrays,
// without any device allocation: the first-occurrence table, the new-id
// table, the scan partials and the error flag live in one file-scope
// device array.  For callers that cannot allocate: code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc.
//
// Same algorithm as the context path (partition.hip): first[l] = the least
// vertex with label l (atomicMin), a vertex is a "first" if it holds its
// label's first[] or NEW_LABEL, and the exclusive scan of those flags over
// v is the reference's running counter.  The output is bit-identical.
//
// Every call on one device shares that device's array and overwrites what
// the last call left there; calls that overlap in time on one device race.
#include <mutex>

#include "../host/hostutil.hpp"
#include "prims.h"

namespace pcd {
namespace {

constexpr int64_t NA_MAX = PCD_NOALLOC_MAX_VERTICES;
constexpr int64_t NA_TILES = (NA_MAX + SCAN_TILE - 1) / SCAN_TILE;
// word offsets inside the scratch array
constexpr int64_t NA_FIRST = 0;                      // first[l], l < n
constexpr int64_t NA_NEWID = NA_MAX;                 // newid[l]
constexpr int64_t NA_PARTS = 2 * NA_MAX;             // tile sums, then their scan and the total
constexpr int64_t NA_FLAGS = NA_PARTS + NA_TILES + 64;  // [0] = a label outside [-1, n)
constexpr int64_t NA_WORDS = NA_FLAGS + 64;

static __device__ __align__(256) int32_t scratch[NA_WORDS];

__global__ void k_na_clear(int n, int32_t *first, int32_t *flags) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) first[i] = INT32_MAX;
    if (blockIdx.x == 0 && threadIdx.x == 0) flags[0] = 0;
}

// atomicMin of the first vertex per label; labels outside [-1, n) set the
// flag and are left out of every table
__global__ void k_na_first(int n, const int32_t *labels, int32_t *first, int32_t *flags) {
    const int stride = gridDim.x * blockDim.x;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const int l = labels[v];
        if (l < -1 || l >= n) {
            flags[0] = 1;
        } else if (l >= 0 && first[l] > v) {
            atomicMin(&first[l], v);
        }
    }
}

struct NaFlag {
    int n;
    const int32_t *labels, *first;
    __device__ int operator()(int64_t v) const {
        const int l = labels[v];
        if (l < -1 || l >= n) return 0;
        return (l < 0 || first[l] == (int)v) ? 1 : 0;
    }
};

struct NaStore {
    const int32_t *labels;
    int32_t *out, *newid;
    __device__ void operator()(int64_t v, int ex, int flag) const {
        if (!flag) return;
        const int l = labels[v];
        if (l < 0) out[v] = ex;
        else newid[l] = ex;
    }
};

__global__ void k_na_relabel(int n, const int32_t *labels, const int32_t *newid, int32_t *out) {
    const int stride = gridDim.x * blockDim.x;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const int l = labels[v];
        if (l >= 0 && l < n) out[v] = newid[l];
        else if (l != -1) out[v] = -1;  // out of contract: reported through k = -1
    }
}

__global__ void k_na_finish(const int32_t *total, const int32_t *flags, int32_t *num_labels) {
    *num_labels = flags[0] ? -1 : *total;
}

// the device address of `scratch` on each device, taken once per device
std::mutex g_mu;
int32_t *g_ptr[64] = {nullptr};

int scratch_ptr(int32_t **out) {
    int dev = -1;
    PCD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(PCD_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(PCD_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        g_ptr[dev] = (int32_t *)p;
    }
    *out = g_ptr[dev];
    return PCD_OK;
}

}  // namespace
}  // namespace pcd

using namespace pcd;

extern "C" int pcd_compress_labels_noalloc(int n, const int32_t *labels, int32_t *out_labels, int32_t *num_labels,
                                           void *stream) {
    if (n < 0) return fail(PCD_E_INVALID, "n must be >= 0");
    if (n > NA_MAX)
        return fail(PCD_E_INVALID, "n = " + std::to_string(n) + " exceeds the allocation-free maximum of " +
                                       std::to_string(NA_MAX) + " vertices");
    if (!num_labels || (n > 0 && (!labels || !out_labels)))
        return fail(PCD_E_INVALID, "null device pointer");
    if (n > 0 && labels == out_labels) return fail(PCD_E_INVALID, "labels and out_labels must not be the same array");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return fail(PCD_E_NODEVICE, "no HIP device available");
    }
    int32_t *base = nullptr;
    PCD_TRY(scratch_ptr(&base));
    hipStream_t s = (hipStream_t)stream;
    int32_t *first = base + NA_FIRST, *newid = base + NA_NEWID, *parts = base + NA_PARTS, *flags = base + NA_FLAGS;
    const int grid = (int)std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 2048);
    hipLaunchKernelGGL(k_na_clear, dim3(grid), dim3(256), 0, s, n, first, flags);
    if (n > 0) {
        hipLaunchKernelGGL(k_na_first, dim3(grid), dim3(256), 0, s, n, labels, first, flags);
        const int tiles = (int)((n + SCAN_TILE - 1) / SCAN_TILE);
        const NaFlag load{n, labels, first};
        const NaStore store{labels, out_labels, newid};
        hipLaunchKernelGGL(k_scan_tiles<NaFlag>, dim3(tiles), dim3(SCAN_THREADS), 0, s, (int64_t)n, load, parts,
                           SCAN_ROUNDS);
        hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(SCAN_THREADS), 0, s, tiles, parts);
        hipLaunchKernelGGL((k_scan_apply<NaFlag, NaStore>), dim3(tiles), dim3(SCAN_THREADS), 0, s, (int64_t)n, load,
                           store, parts, SCAN_ROUNDS);
        hipLaunchKernelGGL(k_na_relabel, dim3(grid), dim3(256), 0, s, n, labels, newid, out_labels);
        hipLaunchKernelGGL(k_na_finish, dim3(1), dim3(1), 0, s, parts + tiles, flags, num_labels);
    } else {
        // parts[0] = 0 is the scan total of no tiles
        hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(SCAN_THREADS), 0, s, 0, parts);
        hipLaunchKernelGGL(k_na_finish, dim3(1), dim3(1), 0, s, parts, flags, num_labels);
    }
    PCD_HIP(hipGetLastError());
    return PCD_OK;
}
