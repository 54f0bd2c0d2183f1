// runtime.hip -- MemoryPool<Mem::HIP> and Runtime entry points of the C ABI
// (util/memory_pool.hpp:53-377, util/memory_pool.cu:56-320, util/runtime.cpp:33-144).

#include "runtime.hpp"
#include "feat3_hip.h"

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>

namespace feat3_amd
{
namespace rt
{
  namespace
  {
    std::mutex g_mutex;
    std::unordered_map<int, Workspace> g_ws;   // per device
    int g_lanes = -1;                          // -1: not read from the environment yet
    int g_debug_sync = -1;

    // refcounted allocation map (memory_pool.cu:97-149)
    std::unordered_map<void*, std::uint64_t> g_refcount;
  }

  Workspace* workspace()
  {
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_mutex);
    auto it = g_ws.find(dev);
    return it != g_ws.end() ? &it->second : nullptr;
  }

  int create_workspace()
  {
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess) return kErrRuntime;
    std::lock_guard<std::mutex> lock(g_mutex);
    if(g_ws.count(dev)) return kOk;
    Workspace w{};
    if(hipMalloc(&w.partials, sizeof(double) * Workspace::kMaxPartials) != hipSuccess) return kErrRuntime;
    if(hipMalloc(&w.scalar, 64) != hipSuccess) { (void)hipFree(w.partials); return kErrRuntime; }
    if(hipEventCreateWithFlags(&w.done, hipEventDisableTiming) != hipSuccess)
    {
      (void)hipFree(w.partials);
      (void)hipFree(w.scalar);
      return kErrRuntime;
    }
    if(hipDeviceSynchronize() != hipSuccess) return kErrRuntime;
    g_ws[dev] = w;
    return kOk;
  }

  int check_launch()
  {
    if(hipGetLastError() != hipSuccess) return kErrLaunch;
    if(g_debug_sync < 0)
    {
      const char* e = std::getenv("FEAT3_HIP_DEBUG_SYNC");
      g_debug_sync = (e && std::strcmp(e, "1") == 0) ? 1 : 0;
    }
    if(g_debug_sync == 1)
    {
      if(hipDeviceSynchronize() != hipSuccess) return kErrLaunch;
      if(hipGetLastError() != hipSuccess) return kErrLaunch;
    }
    return kOk;
  }

  int csr_lanes_override()
  {
    if(g_lanes < 0)
    {
      const char* e = std::getenv("FEAT3_HIP_CSR_LANES");
      g_lanes = e ? std::atoi(e) : 0;
      if(g_lanes != 0 && g_lanes != 4 && g_lanes != 8 && g_lanes != 16 && g_lanes != 32 && g_lanes != 64)
        g_lanes = 0;
    }
    return g_lanes;
  }

  void set_csr_lanes_override(int g) { g_lanes = g; }

  // used by the C ABI below
  int register_alloc(void* p)
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    g_refcount[p] = 1;
    return kOk;
  }
} // namespace rt

namespace
{
  typedef double d2 __attribute__((ext_vector_type(2)));

  // read-rate calibration: every 16-byte word of the buffer read once with nontemporal loads,
  // four in flight per thread; the sum is stored only in a case the data never produces, which
  // keeps the loads without a reduction
  __global__ __launch_bounds__(256) void k_read_probe(const d2* __restrict__ a, std::uint64_t n2, double* out)
  {
    const std::uint64_t stride = std::uint64_t(gridDim.x) * 256;
    std::uint64_t i = std::uint64_t(blockIdx.x) * 256 + threadIdx.x;
    double s = 0.0;
    for(; i + 3 * stride < n2; i += 4 * stride)
    {
      d2 v[4];
#pragma unroll
      for(int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(a + i + u * stride);
#pragma unroll
      for(int u = 0; u < 4; ++u) s += v[u].x + v[u].y;
    }
    for(; i < n2; i += stride)
    {
      const d2 v = a[i];
      s += v.x + v.y;
    }
    if(s == -1.0e308) *out = s;
  }
} // namespace
} // namespace feat3_amd

using namespace feat3_amd;

extern "C" {

const char* feat3_hip_error_string(int code)
{
  switch(code)
  {
    case rt::kOk: return "success";
    case rt::kErrArg: return "invalid argument";
    case rt::kErrLaunch: return "kernel launch failed";
    case rt::kErrRuntime: return "HIP runtime error";
    case rt::kErrState: return "invalid object state";
    case rt::kErrNumeric: return "numerical breakdown during setup";
    default: return "unknown error";
  }
}

// MemoryPool<Mem::HIP>::initialise (memory_pool.cu:56-72): select the device, create the
// reduction workspace (once per device).
int feat3_hip_initialise(int device)
{
  if(hipSetDevice(device) != hipSuccess) return rt::kErrRuntime;
  const int rc = rt::create_workspace();
  if(rc) return rc;
  return rt::setup_noalloc_scratch();
}

// MemoryPool<Mem::HIP>::finalise (memory_pool.cu:74-95): reports leaked allocations
int feat3_hip_finalise(std::uint64_t* leaked)
{
  std::lock_guard<std::mutex> lock(rt::g_mutex);
  if(leaked) *leaked = rt::g_refcount.size();
  return rt::g_refcount.empty() ? rt::kOk : rt::kErrState;
}

int feat3_hip_device_count(int* n)
{
  return hipGetDeviceCount(n) == hipSuccess ? rt::kOk : rt::kErrRuntime;
}

// allocate_memory / increase_memory / release_memory (memory_pool.cu:97-149)
int feat3_hip_malloc(void** p, std::uint64_t bytes)
{
  if(!p) return rt::kErrArg;
  *p = nullptr;
  if(bytes == 0) return rt::kOk;
  if(hipMalloc(p, bytes) != hipSuccess) { *p = nullptr; return rt::kErrRuntime; }
  return rt::register_alloc(*p);
}

int feat3_hip_increase(void* p)
{
  std::lock_guard<std::mutex> lock(rt::g_mutex);
  auto it = rt::g_refcount.find(p);
  if(it == rt::g_refcount.end()) return rt::kErrArg;
  ++it->second;
  return rt::kOk;
}

int feat3_hip_free(void* p)
{
  if(!p) return rt::kOk;
  {
    std::lock_guard<std::mutex> lock(rt::g_mutex);
    auto it = rt::g_refcount.find(p);
    if(it == rt::g_refcount.end()) return rt::kErrArg;
    if(--it->second > 0) return rt::kOk;
    rt::g_refcount.erase(it);
  }
  return hipFree(p) == hipSuccess ? rt::kOk : rt::kErrRuntime;
}

// upload / download / copy (memory_pool.cu:158-175): asynchronous on the given stream
int feat3_hip_upload(void* stream, void* dst, const void* src, std::uint64_t bytes)
{
  if(bytes == 0) return rt::kOk;
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)) == hipSuccess
           ? rt::kOk : rt::kErrRuntime;
}
int feat3_hip_download(void* stream, void* dst, const void* src, std::uint64_t bytes)
{
  if(bytes == 0) return rt::kOk;
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)) == hipSuccess
           ? rt::kOk : rt::kErrRuntime;
}
int feat3_hip_copy(void* stream, void* dst, const void* src, std::uint64_t bytes)
{
  if(bytes == 0) return rt::kOk;
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)) == hipSuccess
           ? rt::kOk : rt::kErrRuntime;
}
// HBM read-rate calibration (no counterpart in the reference): reads bytes (a multiple of 16,
// 16-byte aligned) once; out is a device double that the kernel does not write for finite data
int feat3_hip_read_probe(void* stream, const void* p, std::uint64_t bytes, double* out)
{
  if(!p || !out || bytes % 16 != 0 || reinterpret_cast<std::uintptr_t>(p) % 16 != 0) return rt::kErrArg;
  if(bytes == 0) return rt::kOk;
  feat3_amd::k_read_probe<<<16384, 256, 0, static_cast<hipStream_t>(stream)>>>(static_cast<const feat3_amd::d2*>(p),
                                                                               bytes / 16, out);
  return hipGetLastError() == hipSuccess ? rt::kOk : rt::kErrRuntime;
}
int feat3_hip_synchronise(void* stream)
{
  return hipStreamSynchronize(static_cast<hipStream_t>(stream)) == hipSuccess ? rt::kOk : rt::kErrRuntime;
}

// lanes per CSR row for the drop-in apply kernels (0 = automatic); set by the autotuner
int feat3_hip_set_csr_lanes(int lanes)
{
  if(lanes != 0 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64) return rt::kErrArg;
  rt::set_csr_lanes_override(lanes);
  return rt::kOk;
}

} // extern "C"
