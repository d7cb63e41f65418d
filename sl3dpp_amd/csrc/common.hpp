// Shared plumbing of the HIP library: status codes, the last-error string,
// HIP error checks and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <future>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "sl3dpp_amd.h"

namespace sl3amd {

// The message of the last failed call on this thread (sl3_last_error).
void SetLastError(const std::string& message);

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ArgumentError : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};
// A factorisation that failed outright (the reference's std::runtime_error
// from DegreeOperator::Factorise, src/DegreeOperator.cpp:848-851).
struct NumericalError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define SL3_HIP(call)                                                        \
  do {                                                                       \
    hipError_t err_ = (call);                                                \
    if (err_ != hipSuccess) {                                                \
      throw ::sl3amd::HipError(std::string(#call) + ": " +                   \
                               hipGetErrorString(err_) + " (" __FILE__ ":" + \
                               std::to_string(__LINE__) + ")");              \
    }                                                                        \
  } while (0)

// Runs a C-ABI body, mapping exceptions to status codes.
template <class F>
int Guard(F&& body) {
  try {
    body();
    return SL3_OK;
  } catch (const ArgumentError& e) {
    SetLastError(e.what());
    return SL3_EINVAL;
  } catch (const NumericalError& e) {
    SetLastError(e.what());
    return SL3_ENUMERIC;
  } catch (const HipError& e) {
    SetLastError(e.what());
    return SL3_EHIP;
  } catch (const std::bad_alloc&) {
    SetLastError("out of host memory");
    return SL3_ENOMEM;
  } catch (const std::exception& e) {
    SetLastError(e.what());
    return SL3_EINTERNAL;
  }
}

// Kernel timing on the library's stream, enabled by sl3_profile_enable:
// device events around each launch of a named kernel, read back by
// sl3_profile_read. No cost when disabled.
class ProfileScope {
 public:
  ProfileScope(const char* name, hipStream_t stream);
  ~ProfileScope();
  ProfileScope(const ProfileScope&) = delete;
  ProfileScope& operator=(const ProfileScope&) = delete;

 private:
  const char* name_;
  hipStream_t stream_;
  int slot_ = -1;
};

// body(i) for i in [0, count) on up to 8 host threads.
template <class F>
void ParallelFor(int count, F&& body) {
  const int threads =
      std::max(1, std::min<int>(8, static_cast<int>(std::thread::hardware_concurrency())));
  std::vector<std::future<void>> jobs;
  std::atomic<int> next{0};
  for (int t = 0; t < threads; ++t) {
    jobs.push_back(std::async(std::launch::async, [&] {
      for (int i = next++; i < count; i = next++) body(i);
    }));
  }
  for (auto& j : jobs) j.get();
}

inline int CeilDiv(long long a, long long b) {
  return static_cast<int>((a + b - 1) / b);
}

// A device buffer owned by the library (freed on destruction).
template <class T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t count) { Resize(count); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : ptr_(o.ptr_), count_(o.count_) {
    o.ptr_ = nullptr;
    o.count_ = 0;
  }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      Free();
      ptr_ = o.ptr_;
      count_ = o.count_;
      o.ptr_ = nullptr;
      o.count_ = 0;
    }
    return *this;
  }
  ~DeviceArray() { Free(); }
  void Resize(size_t count) {
    if (count == count_) return;
    Free();
    if (count > 0) {
      SL3_HIP(hipMalloc(reinterpret_cast<void**>(&ptr_), count * sizeof(T)));
    }
    count_ = count;
  }
  void Upload(const T* host, size_t count, hipStream_t stream) {
    Resize(count);
    if (count > 0) {
      SL3_HIP(hipMemcpyAsync(ptr_, host, count * sizeof(T),
                             hipMemcpyHostToDevice, stream));
    }
  }
  T* data() const { return ptr_; }
  size_t size() const { return count_; }

 private:
  void Free() {
    if (ptr_ != nullptr) (void)hipFree(ptr_);
    ptr_ = nullptr;
    count_ = 0;
  }
  T* ptr_ = nullptr;
  size_t count_ = 0;
};

}  // namespace sl3amd
