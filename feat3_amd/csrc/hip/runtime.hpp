// runtime.hpp -- MemoryPool<Mem::HIP> runtime support: error codes, launch checks and the
// per-device reduction workspace (util/memory_pool.cu:56-72 initialise; cuda_util.cu).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace feat3_amd
{
namespace rt
{
  // C ABI status codes (0 = success)
  enum : int
  {
    kOk = 0,
    kErrArg = 1,      // invalid argument: null pointer, unsupported size or format
    kErrLaunch = 2,   // kernel launch failed
    kErrRuntime = 3,  // other HIP runtime error (allocation, copy, synchronisation)
    kErrState = 4,    // object used in the wrong lifecycle state (e.g. apply before init)
    kErrNumeric = 5   // numerical breakdown detected during setup (e.g. singular coarse matrix)
  };

  // per-device scratch for deterministic grid reductions of the drop-in BLAS-1 calls; `done` is
  // recorded after every reduction that used it, and the next reduction's stream waits for it, so
  // that reductions issued on different streams take turns on the one set of partials
  struct Workspace
  {
    static constexpr unsigned kMaxPartials = 8192;
    void* partials;        // kMaxPartials doubles
    void* scalar;          // one double, target of the host-returning reductions
    hipEvent_t done;
  };

  // the workspace of the current device, created by feat3_hip_initialise (nullptr before that)
  Workspace* workspace();
  // creates the current device's workspace (feat3_hip_initialise); kOk if it exists already
  int create_workspace();
  // looks up the current device's address of the no-allocation dot product's scratch array
  // (feat3_hip_initialise); kOk if it is known already
  int setup_noalloc_scratch();

  // returns kOk or kErrLaunch after a kernel launch; with FEAT3_HIP_DEBUG_SYNC=1 in the
  // environment every launch is also synchronised and checked (FEAT_DEBUG_MODE analogue)
  int check_launch();

  // lanes per row forced by the autotuner or by FEAT3_HIP_CSR_LANES (0 = automatic)
  int csr_lanes_override();
  void set_csr_lanes_override(int g);
} // namespace rt
} // namespace feat3_amd
