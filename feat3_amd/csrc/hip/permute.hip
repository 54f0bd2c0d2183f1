// permute.hip -- renumbering of device-resident CSR matrices and vectors (setup; no counterpart in
// the reference, whose containers keep the numbering they were assembled in).
//
// A level assembled in one numbering (e.g. the lexicographic lattice order of the device
// assembly) is handed over in another (FEAT's hierarchical numbering), and the solver takes a
// level in the caller's numbering and works in the lattice order internally.  Both are the same
// operation: old row i -> row row_order[i], old column j -> column col_order[j], columns sorted
// ascending inside every new row, values moved unchanged.
//
// Kernels: one thread per old row writes its entry count at the new row's slot (with bit 31 set,
// so that a row order that is not a permutation leaves a slot without or with a doubled bit and
// is detected without scratch memory); the row pointers are summed on the host; one thread per
// old row then copies its entries to the new row, maps the columns and sorts them by insertion
// (rows of finite-element matrices hold a few dozen entries).

#include "device_common.hpp"
#include "runtime.hpp"
#include "feat3_hip.h"

#include <hip/hip_runtime.h>

#include <vector>

using namespace feat3_amd;
typedef std::uint32_t u32;
typedef std::uint64_t u64;

namespace
{
  constexpr u32 kHit = 0x80000000u;

  inline unsigned grid_for(u64 n)
  {
    const u64 b = (n + 255) / 256;
    return unsigned(b < 1 ? 1 : (b > 16384 ? 16384 : b));
  }

  __global__ void k_clear_u32(u64 n, u32* __restrict__ p)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) p[i] = 0u;
  }

  // slot[row_order[i]] += len(i) | bit 31 (a slot hit twice wraps bit 31 away); row pointers
  // that decrease or leave [0, nnz] are flagged before any entry is moved
  __global__ void k_row_lengths(u64 rows, u64 nnz, const u32* __restrict__ rp, const u32* __restrict__ row_order,
                                u32* slot, int* bad)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < rows; i += u64(gridDim.x) * blockDim.x)
    {
      const u32 r = row_order ? row_order[i] : u32(i);
      const u32 a = rp[i], b = rp[i + 1];
      if(u64(r) >= rows || a > b || u64(b) > nnz || (i == 0 && a != 0)) { atomicOr(bad, 1); continue; }
      atomicAdd(&slot[r], (b - a) | kHit);
    }
  }

  __global__ void k_move_rows(u64 rows, u64 cols, const u32* __restrict__ rp, const u32* __restrict__ col,
                              const double* __restrict__ val, const u32* __restrict__ row_order,
                              const u32* __restrict__ col_order, const u32* __restrict__ nrp, u32* __restrict__ ncol,
                              double* __restrict__ nval, int* bad)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < rows; i += u64(gridDim.x) * blockDim.x)
    {
      const u32 r = row_order ? row_order[i] : u32(i);
      const u64 a = rp[i], len = u64(rp[i + 1]) - a, o = nrp[r];
      for(u64 k = 0; k < len; ++k)
      {
        const u32 c0 = col[a + k];
        if(u64(c0) >= cols) { atomicOr(bad, 1); return; }
        const u32 c = col_order ? col_order[c0] : c0;
        if(u64(c) >= cols) { atomicOr(bad, 1); return; }
        const double v = val[a + k];
        // insertion into the sorted prefix o .. o + k (stable for equal columns)
        u64 p = k;
        while(p > 0 && ncol[o + p - 1] > c)
        {
          ncol[o + p] = ncol[o + p - 1];
          nval[o + p] = nval[o + p - 1];
          --p;
        }
        ncol[o + p] = c;
        nval[o + p] = v;
      }
    }
  }

  template<typename T>
  __global__ void k_scatter(u64 n, const u32* __restrict__ order, const T* __restrict__ src, T* __restrict__ dst)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      dst[order[i]] = src[i];
  }
  template<typename T>
  __global__ void k_gather(u64 n, const u32* __restrict__ order, const T* __restrict__ src, T* __restrict__ dst)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      dst[i] = src[order[i]];
  }
} // namespace

extern "C" {

int feat3_device_permute_csr(void* stream, uint64_t rows, uint64_t cols, const uint32_t* row_ptr,
                             const uint32_t* col_ind, const double* val, const uint32_t* row_order,
                             const uint32_t* col_order, uint32_t* out_row_ptr, uint32_t* out_col_ind, double* out_val)
{
  if(!row_ptr || !out_row_ptr || rows == 0 || rows >= (u64(1) << 31) || cols >= (u64(1) << 32)) return rt::kErrArg;
  hipStream_t s = static_cast<hipStream_t>(stream);
  // the error flag lives in the first slot of the output row pointer until the lengths are read
  u32 nnz = 0;
  if(hipMemcpyAsync(&nnz, row_ptr + rows, sizeof(u32), hipMemcpyDeviceToHost, s) != hipSuccess) return rt::kErrRuntime;
  if(hipStreamSynchronize(s) != hipSuccess) return rt::kErrRuntime;
  if(nnz && (!col_ind || !val || !out_col_ind || !out_val)) return rt::kErrArg;
  k_clear_u32<<<grid_for(rows + 1), 256, 0, s>>>(rows + 1, out_row_ptr);
  int* bad = reinterpret_cast<int*>(out_row_ptr);   // out_row_ptr[0]: 0 until a fault is flagged
  k_row_lengths<<<grid_for(rows), 256, 0, s>>>(rows, nnz, row_ptr, row_order, out_row_ptr + 1, bad);
  if(rt::check_launch()) return rt::kErrLaunch;
  std::vector<u32> h(rows + 1);
  if(hipMemcpyAsync(h.data(), out_row_ptr, sizeof(u32) * (rows + 1), hipMemcpyDeviceToHost, s) != hipSuccess)
    return rt::kErrRuntime;
  if(hipStreamSynchronize(s) != hipSuccess) return rt::kErrRuntime;
  if(h[0] != 0) return rt::kErrArg;
  u64 acc = 0;
  for(u64 r = 0; r < rows; ++r)
  {
    if(!(h[r + 1] & kHit)) return rt::kErrArg;   // no old row, or two, moved to this row
    acc += h[r + 1] & ~kHit;
    h[r + 1] = u32(acc);
  }
  h[0] = 0;
  if(acc != nnz) return rt::kErrArg;
  if(hipMemcpyAsync(out_row_ptr, h.data(), sizeof(u32) * (rows + 1), hipMemcpyHostToDevice, s) != hipSuccess)
    return rt::kErrRuntime;
  // a flag word for the column checks: the last output row pointer, restored afterwards
  int* bad2 = reinterpret_cast<int*>(out_row_ptr + rows);
  if(hipMemsetAsync(bad2, 0, sizeof(int), s) != hipSuccess) return rt::kErrRuntime;
  k_move_rows<<<grid_for(rows), 256, 0, s>>>(rows, cols, row_ptr, col_ind, val, row_order, col_order, out_row_ptr,
                                             out_col_ind, out_val, bad2);
  if(rt::check_launch()) return rt::kErrLaunch;
  int hb = 0;
  if(hipMemcpyAsync(&hb, bad2, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return rt::kErrRuntime;
  if(hipMemcpyAsync(out_row_ptr + rows, &h[rows], sizeof(u32), hipMemcpyHostToDevice, s) != hipSuccess)
    return rt::kErrRuntime;
  if(hipStreamSynchronize(s) != hipSuccess) return rt::kErrRuntime;   // h leaves scope
  return hb ? rt::kErrArg : rt::kOk;
}

int feat3_device_permute_vector(void* stream, uint64_t n, const uint32_t* order, int inverse, const double* src,
                                double* dst)
{
  if(n == 0) return rt::kOk;
  if(!order || !src || !dst || src == dst) return rt::kErrArg;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if(inverse) k_gather<double><<<grid_for(n), 256, 0, s>>>(n, order, src, dst);
  else k_scatter<double><<<grid_for(n), 256, 0, s>>>(n, order, src, dst);
  return rt::check_launch();
}

} // extern "C"
