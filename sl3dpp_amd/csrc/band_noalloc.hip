// The batched LU solve that allocates nothing (sl3_band_lu_solve_batched_
// noalloc): the same kernels as sl3_band_lu_solve_batched, with the
// per-call system descriptors and the solve's work list staged in a
// file-scope device array instead of device buffers allocated per call.
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "common.hpp"
#include "kernels.hpp"

namespace sl3amd {
namespace {

// The scratch: up to kMaxSystems descriptors at offset 0, up to kMaxTasks
// (system, first column) work items at kTasksOffset. One copy per device,
// held until the process ends; every call on a device shares it.
constexpr int kMaxSystems = 512;
constexpr int kMaxTasks = 4096;
constexpr int kWaveColumns = 64;  // right-hand sides per work item (the solve kernels' wave)
constexpr size_t kTasksOffset = 32768;
constexpr size_t kScratchBytes = 65536;
static_assert(sizeof(BandSystem) * kMaxSystems <= kTasksOffset, "descriptors fit");
static_assert(kTasksOffset + sizeof(int2) * kMaxTasks <= kScratchBytes, "work list fits");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

// The device address of this device's scratch, looked up once per device
// with that device current (the array's name is not a device address in
// host code).
unsigned char* ScratchOfCurrentDevice() {
  int device = 0;
  SL3_HIP(hipGetDevice(&device));
  if (device < 0 || device >= kMaxDevices) throw ArgumentError("device index out of range");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[device] == nullptr) {
    void* ptr = nullptr;
    const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || ptr == nullptr) {
      throw HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err));
    }
    g_scratch[device] = ptr;
  }
  return static_cast<unsigned char*>(g_scratch[device]);
}

void Require(bool ok, const char* what) {
  if (!ok) throw ArgumentError(what);
}

}  // namespace
}  // namespace sl3amd

using namespace sl3amd;

extern "C" int sl3_band_lu_solve_batched_noalloc(
    int count, const int* n, const int* kl, const int* ku, const int64_t* ab_off,
    const double* ab_dev, const int64_t* piv_off, const int* piv_dev, const int* nrhs,
    const int64_t* b_off, const int64_t* row_stride, const int64_t* col_stride,
    double* b_dev) {
  return Guard([&] {
    Require(count >= 0, "the system count must be non-negative");
    Require(count <= kMaxSystems, "at most 512 systems per call (the scratch's limit)");
    Require(count == 0 || (n && kl && ku && ab_off && ab_dev && piv_off && piv_dev && nrhs &&
                           b_off && b_dev),
            "null argument");
    // descriptors and work list on the host stack: no allocation
    BandSystem systems[kMaxSystems];
    int2 tasks[kMaxTasks];
    int taskCount = 0, klMax = 0, kuMax = 0;
    for (int s = 0; s < count; ++s) {
      BandSystem& sys = systems[s];
      sys = BandSystem{};
      sys.n = n[s];
      sys.kl = kl[s];
      sys.ku = ku[s];
      Require(sys.n >= 1, "a band system has at least one row");
      Require(sys.kl >= 0 && sys.ku >= 0, "band widths must be non-negative");
      Require(sys.kl + sys.ku <= 95, "band wider than the kernels' LDS staging (kl + ku <= 95)");
      sys.ab_off = ab_off[s];
      sys.piv_off = piv_off[s];
      sys.nrhs = nrhs[s];
      Require(sys.nrhs >= 0, "the right-hand side count must be non-negative");
      sys.b_off = b_off[s];
      sys.row_stride = row_stride ? row_stride[s] : 1;
      sys.col_stride = col_stride ? col_stride[s] : sys.n;
      klMax = sys.kl > klMax ? sys.kl : klMax;
      kuMax = sys.ku > kuMax ? sys.ku : kuMax;
      for (int c = 0; c < sys.nrhs; c += kWaveColumns) {
        Require(taskCount < kMaxTasks,
                "at most 4096 blocks of 64 right-hand sides per call (the scratch's limit)");
        tasks[taskCount++] = make_int2(s, c);
      }
    }
    if (taskCount == 0) return;
    hipStream_t stream = Stream();
    unsigned char* base = ScratchOfCurrentDevice();
    auto* systemsDev = reinterpret_cast<BandSystem*>(base);
    auto* tasksDev = reinterpret_cast<int2*>(base + kTasksOffset);
    SL3_HIP(hipMemcpyAsync(systemsDev, systems, sizeof(BandSystem) * count,
                           hipMemcpyHostToDevice, stream));
    SL3_HIP(hipMemcpyAsync(tasksDev, tasks, sizeof(int2) * taskCount, hipMemcpyHostToDevice,
                           stream));
    // the launch of sl3_band_lu_solve_batched: LDS for klMax + kuMax
    LaunchLuSolve(systemsDev, tasksDev, taskCount, klMax + kuMax, ab_dev, piv_dev, b_dev,
                  stream);
    // the host arrays above are staged by the copies: wait before they go
    SL3_HIP(hipStreamSynchronize(stream));
  });
}
