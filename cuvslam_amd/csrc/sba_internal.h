// Handle internals shared by the host runtimes of the visual and the inertial solves (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>

struct sba_handle;

namespace sba_internal {

hipStream_t stream_of(sba_handle* h);
int device_of(const sba_handle* h);
// Records msg as the handle's last error and returns code.
int set_error(sba_handle* h, int code, const std::string& msg);
// The handle's auxiliary device workspace (the inertial and the mono solves), at least `bytes` long (regrown to
// twice the need).
int aux_workspace(sba_handle* h, size_t bytes, char** out);
// Calls fn(i) for i in [0, n) on up to `threads` of the handle's host threads (i = 0, threads, 2 threads, ... on the
// first); every call has its own i, so results written by index do not depend on the thread count.
void parallel_for(sba_handle* h, int n, int threads, const std::function<void(int)>& fn);

}  // namespace sba_internal
