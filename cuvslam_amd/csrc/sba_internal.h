// Handle internals shared by the host runtimes of the visual and the inertial solves (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

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

}  // namespace sba_internal
