// C ABI plumbing: the last-error string, the version, the library's stream.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "common.hpp"
#include "kernels.hpp"

namespace sl3amd {

namespace {
thread_local std::string lastError;
}

void SetLastError(const std::string& message) { lastError = message; }

// One stream for the library's work on the current device, at the default
// priority (created on first use).
hipStream_t Stream() {
  static std::once_flag once;
  static hipStream_t stream = nullptr;
  std::call_once(once, [] {
    SL3_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  });
  return stream;
}

}  // namespace sl3amd

extern "C" {

const char* sl3_last_error(void) { return sl3amd::lastError.c_str(); }

const char* sl3_version(void) { return "sl3dpp_amd 0.1 (gfx950)"; }

int sl3_synchronize(void) {
  return sl3amd::Guard([] { SL3_HIP(hipStreamSynchronize(sl3amd::Stream())); });
}

}  // extern "C"
