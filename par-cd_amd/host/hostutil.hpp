// Internal helpers shared by the host-side C++ of the library.
#ifndef PARCD_HOSTUTIL_HPP
#define PARCD_HOSTUTIL_HPP

#include <cstdint>
#include <functional>
#include <string>

namespace pcd {

// Records msg as the calling thread's last error and returns code.
int fail(int code, const std::string &msg);
extern thread_local std::string g_last_error;

// Number of host worker threads (hardware threads, at most 16, or
// PCD_HOST_THREADS).
int host_threads();

// Runs fn(lo, hi) over [0, n) split into contiguous chunks on host threads.
void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &fn, int64_t grain);

}  // namespace pcd

#endif
