// device_common.hpp -- device helpers shared by the CDNA4 kernels: value loads with
// precision conversion, wave64 reductions, deterministic grid reductions and the
// XCD-aware block remap.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace feat3_amd
{
namespace dev
{
  typedef _Float16 half;

  constexpr int kWave = 64;

  // ---- loads with conversion into the arithmetic type ----------------------------------
  template<typename DT_, typename VT_>
  __device__ __forceinline__ DT_ ld(const VT_* p, std::uint64_t i) { return DT_(p[i]); }
  template<>
  __device__ __forceinline__ float ld<float, half>(const half* p, std::uint64_t i) { return float(p[i]); }
  template<>
  __device__ __forceinline__ double ld<double, half>(const half* p, std::uint64_t i) { return double(float(p[i])); }

  // read-only views in the constant address space: a load through one whose address is
  // wave-uniform is issued as a scalar (SMEM) load through the scalar cache, so one load serves
  // the whole wave instead of 64 lanes requesting the same address
  template<typename T>
  using cptr = const __attribute__((address_space(4))) T*;
  template<typename T>
  __device__ __forceinline__ cptr<T> to_const(const T* p) { return (cptr<T>)(p); }
  template<typename DT_, typename VT_>
  __device__ __forceinline__ DT_ ldc(cptr<VT_> p, std::uint64_t i) { return DT_(p[i]); }
  template<>
  __device__ __forceinline__ float ldc<float, half>(cptr<half> p, std::uint64_t i) { return float(p[i]); }

  // nontemporal load with conversion into the arithmetic type (binary16: an ordinary load)
  template<typename DT_, typename VT_>
  __device__ __forceinline__ DT_ ld_nt(const VT_* p, std::uint64_t i) { return DT_(__builtin_nontemporal_load(p + i)); }
  template<>
  __device__ __forceinline__ float ld_nt<float, half>(const half* p, std::uint64_t i) { return float(p[i]); }
  template<>
  __device__ __forceinline__ double ld_nt<double, half>(const half* p, std::uint64_t i) { return double(float(p[i])); }

  // nontemporal streaming loads for matrix arrays that are read exactly once per sweep
  template<typename T>
  __device__ __forceinline__ T ld_stream(const T* p) { return __builtin_nontemporal_load(p); }
  template<>
  __device__ __forceinline__ half ld_stream<half>(const half* p) { return *p; }

  // Stages prod[k] = val[e0 + k] * x[col[e0 + k]] for k in [0, ne) by the NT threads of a
  // block: each thread issues the loads of U entries (values and columns, predicated, no
  // branches between them, so that all U are in flight), then their x gathers, then stores the
  // products.  Measured on the 3D Q1 L7 matrix: U = 8 entries of one element each beat
  // 16-byte loads of 2 or 4 consecutive entries per lane (whose partial first and last groups
  // need branches).
  template<int U, int NT, typename DT, typename VT, typename IT>
  __device__ __forceinline__ void stage_products(const VT* __restrict__ val, const IT* __restrict__ col, std::uint64_t e0,
                                                 std::uint64_t ne, const DT* __restrict__ x, DT* prod)
  {
    typedef std::uint64_t u64;
    for(u64 k0 = threadIdx.x; k0 < ne; k0 += u64(NT) * U)
    {
      DT v[U];
      u64 c[U];
#pragma unroll
      for(int j = 0; j < U; ++j)
      {
        const u64 k = k0 + u64(j) * NT;
        v[j] = k < ne ? DT(ld_stream(val + e0 + k)) : DT(0);
        c[j] = k < ne ? u64(ld_stream(col + e0 + k)) : 0;
      }
      DT xv[U];
#pragma unroll
      for(int j = 0; j < U; ++j) xv[j] = k0 + u64(j) * NT < ne ? x[c[j]] : DT(0);
#pragma unroll
      for(int j = 0; j < U; ++j)
        if(k0 + u64(j) * NT < ne) prod[k0 + u64(j) * NT] = v[j] * xv[j];
    }
  }

  // ---- sub-wave reductions (G lanes, G a power of two <= 64) ----------------------------
  template<int G, typename T>
  __device__ __forceinline__ T group_sum(T v)
  {
#pragma unroll
    for(int off = G / 2; off > 0; off >>= 1)
      v += __shfl_xor(v, off, G);
    return v;
  }

  template<typename T>
  __device__ __forceinline__ T wave_sum(T v) { return group_sum<kWave>(v); }

  // block sum (blockDim.x a multiple of 64, at most 1024); result valid in thread 0
  template<typename T>
  __device__ __forceinline__ T block_sum(T v)
  {
    __shared__ T red[16];
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    if(lane == 0) red[wid] = v;
    __syncthreads();
    const int nw = (blockDim.x + kWave - 1) / kWave;
    T s = T(0);
    if(wid == 0)
    {
      s = lane < nw ? red[lane] : T(0);
      s = wave_sum(s);
    }
    __syncthreads();
    return s;
  }

  // Deterministic two-kernel grid reduction.  Stage 1 (in the producing kernel): every block
  // stores its block sum to partials[blockIdx.x].  Stage 2 (a one-block kernel launched
  // after it on the same stream): sum_partials adds them in block order.  The kernel boundary
  // makes the partials visible across the XCDs' L2s; a per-block fence + ticket atomic instead
  // forces an L2 write-back in every block, which costs more than the reduced data.
  template<typename T>
  __device__ __forceinline__ void block_partial(T v, T* partials)
  {
    const T bs = block_sum(v);
    if(threadIdx.x == 0) partials[blockIdx.x] = bs;
  }

  // sum of partials[0, n) in a fixed order; valid in thread 0 (one block).  Each thread keeps
  // eight accumulators so that eight of its loads are in flight at once (the partials were
  // written by the previous kernel and come from memory, not from this CU's cache).
  template<typename T>
  __device__ __forceinline__ T sum_partials(const T* partials, unsigned int n)
  {
    constexpr unsigned int U = 8;
    T acc[U];
#pragma unroll
    for(unsigned int j = 0; j < U; ++j) acc[j] = T(0);
    const unsigned int step = blockDim.x;
    unsigned int i = threadIdx.x;
    for(; i + (U - 1) * step < n; i += U * step)
#pragma unroll
      for(unsigned int j = 0; j < U; ++j) acc[j] += partials[i + j * step];
    for(; i < n; i += step) acc[0] += partials[i];
    const T s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    return block_sum(s);
  }

  // Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8), each XCD with
  // its own L2.  Remapping gives every XCD a contiguous range of logical blocks, so the x
  // entries gathered by neighbouring rows are fetched into one L2 instead of eight.
  __device__ __forceinline__ unsigned int xcd_remap(unsigned int b, unsigned int nb)
  {
    constexpr unsigned int kXcd = 8;
    const unsigned int q = nb / kXcd, r = nb % kXcd;
    const unsigned int x = b % kXcd, i = b / kXcd;
    return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
  }

  // Dirichlet rows as a bitmask (bit i%32 of word i/32)
  __device__ __forceinline__ bool masked(const std::uint32_t* mask, std::uint64_t i)
  {
    return mask != nullptr && ((mask[i >> 5] >> (i & 31u)) & 1u);
  }
  // the same for a mask that is never null and a 32-bit row (row pointers are 32-bit, so every
  // row index is): without the pointer test the compiler keeps one copy of a row loop instead of a
  // copy per case, whose scalar state overflowed the SGPRs of the banded row kernels, and 32-bit
  // indices keep the epilogue's address arithmetic off 64-bit registers (FP32 value-coded defect
  // 117 -> 76 VGPRs, FP64 outer product 79 -> 61)
  __device__ __forceinline__ bool masked_nn(const std::uint32_t* __restrict__ mask, std::uint32_t i)
  {
    return (mask[i >> 5] >> (i & 31u)) & 1u;
  }
} // namespace dev
} // namespace feat3_amd
