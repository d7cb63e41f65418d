// C ABI plumbing: the last-error string, the version, the library's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace sl3amd {

namespace {
thread_local std::string lastError;
}

void SetLastError(const std::string& message) { lastError = message; }

// The library's stream: one per host thread, at the default priority
// (created on the thread's first call, kept until the process ends). Every
// entry point finishes its work on it before returning, so a handle made
// on one thread can be used on another, and operators stepped from
// different threads (independent realisations) run concurrently.
hipStream_t Stream() {
  thread_local hipStream_t stream = nullptr;
  if (stream == nullptr) SL3_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  return stream;
}

namespace {
struct Profiler {
  bool enabled = false;
  std::set<std::string> only;  // the kernels timed (empty: every kernel)
  std::mutex mutex;
  std::vector<hipEvent_t> pool;
  struct Pending {
    std::string name;
    int start, stop;
  };
  std::vector<Pending> pending;
  std::map<std::string, std::pair<double, long long>> totals;
  int nextEvent = 0;
  void Drain() {
    for (const auto& p : pending) {
      SL3_HIP(hipEventSynchronize(pool[p.stop]));
      float ms = 0.0f;
      SL3_HIP(hipEventElapsedTime(&ms, pool[p.start], pool[p.stop]));
      auto& t = totals[p.name];
      t.first += ms;
      t.second += 1;
    }
    pending.clear();
    nextEvent = 0;
  }
  int Take() {
    if (nextEvent + 2 > static_cast<int>(pool.size())) {
      if (pool.size() >= 4096) Drain();
      while (nextEvent + 2 > static_cast<int>(pool.size())) {
        hipEvent_t e;
        SL3_HIP(hipEventCreate(&e));
        pool.push_back(e);
      }
    }
    const int slot = nextEvent;
    nextEvent += 2;
    return slot;
  }
};
Profiler& GetProfiler() {
  static Profiler p;
  return p;
}
}  // namespace

ProfileScope::ProfileScope(const char* name, hipStream_t stream)
    : name_(name), stream_(stream) {
  auto& p = GetProfiler();
  if (!p.enabled) return;
  std::lock_guard<std::mutex> lock(p.mutex);
  if (!p.only.empty() && p.only.count(name_) == 0) return;
  // launches recorded into a graph are not timed one by one
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream_, &capture) != hipSuccess ||
      capture != hipStreamCaptureStatusNone) {
    return;
  }
  slot_ = p.Take();
  SL3_HIP(hipEventRecord(p.pool[slot_], stream_));
}

ProfileScope::~ProfileScope() {
  if (slot_ < 0) return;
  auto& p = GetProfiler();
  std::lock_guard<std::mutex> lock(p.mutex);
  if (hipEventRecord(p.pool[slot_ + 1], stream_) == hipSuccess) {
    p.pending.push_back({name_, slot_, slot_ + 1});
  }
}

}  // namespace sl3amd

extern "C" {

int sl3_profile_enable(int on) {
  return sl3amd::Guard([&] {
    auto& p = sl3amd::GetProfiler();
    std::lock_guard<std::mutex> lock(p.mutex);
    p.Drain();
    p.enabled = on != 0;
  });
}

int sl3_profile_select(const char* kernels) {
  return sl3amd::Guard([&] {
    auto& p = sl3amd::GetProfiler();
    std::lock_guard<std::mutex> lock(p.mutex);
    p.only.clear();
    if (kernels == nullptr) return;
    std::string cur;
    for (const char* c = kernels;; ++c) {
      if (*c == ',' || *c == ' ' || *c == 0) {
        if (!cur.empty()) p.only.insert(cur);
        cur.clear();
        if (*c == 0) break;
      } else {
        cur += *c;
      }
    }
  });
}

int sl3_profile_reset(void) {
  return sl3amd::Guard([&] {
    auto& p = sl3amd::GetProfiler();
    std::lock_guard<std::mutex> lock(p.mutex);
    p.Drain();
    p.totals.clear();
  });
}

int sl3_profile_read(const char* kernel, double* total_ms, int64_t* launches) {
  return sl3amd::Guard([&] {
    auto& p = sl3amd::GetProfiler();
    std::lock_guard<std::mutex> lock(p.mutex);
    p.Drain();
    const auto it = p.totals.find(kernel ? kernel : "");
    if (total_ms) *total_ms = it == p.totals.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == p.totals.end() ? 0 : it->second.second;
  });
}

int sl3_profile_names(char* buffer, int64_t capacity) {
  return sl3amd::Guard([&] {
    auto& p = sl3amd::GetProfiler();
    std::lock_guard<std::mutex> lock(p.mutex);
    p.Drain();
    std::string all;
    for (const auto& [name, t] : p.totals) all += name + "\n";
    if (buffer && capacity > 0) {
      const size_t n = std::min<size_t>(all.size(), static_cast<size_t>(capacity - 1));
      std::memcpy(buffer, all.data(), n);
      buffer[n] = 0;
    }
  });
}


const char* sl3_last_error(void) { return sl3amd::lastError.c_str(); }

const char* sl3_version(void) { return "sl3dpp_amd 0.1 (gfx950)"; }

int sl3_synchronize(void) {
  return sl3amd::Guard([] { SL3_HIP(hipStreamSynchronize(sl3amd::Stream())); });
}

int sl3_copy(void* dst_dev, const void* src_dev, int64_t bytes) {
  return sl3amd::Guard([&] {
    if (bytes < 0 || ((dst_dev == nullptr || src_dev == nullptr) && bytes > 0)) {
      throw sl3amd::ArgumentError("sl3_copy: null address or negative size");
    }
    if (bytes > 0) {
      SL3_HIP(hipMemcpyAsync(dst_dev, src_dev, static_cast<size_t>(bytes),
                             hipMemcpyDeviceToDevice, sl3amd::Stream()));
    }
  });
}

}  // extern "C"
