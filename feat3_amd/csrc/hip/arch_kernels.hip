// arch_kernels.hip -- CDNA4 kernels behind the lafem_hip_* C ABI: the device counterparts of
// FEAT3's LAFEM::Arch::<Op><Mem_> statics (lafem/arch/apply.hpp:24-303 and siblings).
//
// Every entry point takes an explicit hipStream_t (as void*), raw device pointers and sizes,
// launches asynchronously and returns 0 or a feat3 error code.  Host-returning reductions
// (the *_value variants) synchronise the stream, as the reference's cuBLAS scalar returns
// do; the plain variants leave their result in device memory (graph-capturable).

#include "device_common.hpp"
#include "runtime.hpp"
#include "feat3_hip.h"

#include <cmath>
#include <limits>
#include <mutex>

using namespace feat3_amd;
using feat3_amd::dev::half;
typedef std::uint32_t u32;
typedef std::uint64_t u64;

namespace
{
  constexpr int kBlock = 256;

  inline unsigned int blocks_for(u64 work, int per_block = kBlock)
  {
    const u64 b = (work + u64(per_block) - 1) / u64(per_block);
    return unsigned(b == 0 ? 1 : b);
  }

  // lanes per row for the CSR kernels: the smallest power of two >= the mean row length,
  // clamped to [4, 64]; an explicit choice (autotuner) overrides it
  int lanes_for(u64 rows, u64 nnz)
  {
    const int forced = rt::csr_lanes_override();
    if(forced > 0) return forced;
    const double mean = rows ? double(nnz) / double(rows) : 1.0;
    int g = 4;
    while(g < 64 && double(g) < mean) g *= 2;
    return g;
  }

  // -------------------------------------------------------------------------------------
  // CSR: r = a*A*x + b*y (apply_generic.hpp:24-69); G lanes per row, group reduction
  // -------------------------------------------------------------------------------------
  template<int G, typename DT, typename VT, typename IT>
  __global__ __launch_bounds__(kBlock) void k_csr(DT* __restrict__ r, DT a, const DT* __restrict__ x, DT b,
                                                  const DT* y, const VT* __restrict__ val, const IT* __restrict__ col,
                                                  const IT* __restrict__ rp, u64 rows, int use_y)
  {
    const unsigned lb = dev::xcd_remap(blockIdx.x, gridDim.x);
    const u64 row = (u64(lb) * kBlock + threadIdx.x) / G;
    const int lane = threadIdx.x % G;
    if(row >= rows) return;
    const IT beg = rp[row], end = rp[row + 1];
    DT sum = DT(0);
    for(IT i = beg + IT(lane); i < end; i += IT(G))
      sum += dev::ld<DT>(val, i) * x[col[i]];
    sum = dev::group_sum<G>(sum);
    if(lane == 0)
      r[row] = use_y ? (sum * a) + (b * y[row]) : (sum * a);
  }

  template<typename DT, typename VT, typename IT>
  int launch_csr(hipStream_t s, DT* r, DT a, const DT* x, DT b, const DT* y, const VT* val, const IT* col,
                 const IT* rp, u64 rows, u64 nnz)
  {
    if(rows == 0) return 0;
    const int use_y = std::abs(double(b)) < double(std::numeric_limits<DT>::epsilon()) ? 0 : 1;
    const int g = lanes_for(rows, nnz);
    const unsigned nb = blocks_for(rows * u64(g));
#define F3_CSR_CASE(GG) case GG: k_csr<GG, DT, VT, IT><<<nb, kBlock, 0, s>>>(r, a, x, b, y, val, col, rp, rows, use_y); break;
    switch(g)
    {
      F3_CSR_CASE(4) F3_CSR_CASE(8) F3_CSR_CASE(16) F3_CSR_CASE(32) F3_CSR_CASE(64)
      default: return rt::kErrArg;
    }
#undef F3_CSR_CASE
    return rt::check_launch();
  }

  // -------------------------------------------------------------------------------------
  // CSR by streaming the entries of a run of rows (r = a*A*x + b*y, and the csrsb product
  // with BS-vectors): a block takes R consecutive rows; their entries [rp[r0], rp[r1]) are
  // contiguous, so the 256 threads read val and col with unit stride and several independent
  // loads each (every load instruction of a wave covers 64 consecutive entries), and store the
  // products val * x[col] in LDS; then thread t adds up row r0 + t from LDS in entry order.
  // A block whose entries exceed the LDS capacity sums its rows with 8 lanes per row instead.
  // R is chosen on the host from the mean row length so that a block's entries fill about
  // 85% of the capacity (16 KB of products: up to 8 blocks per CU).  Each thread issues the
  // loads of U entries (values and columns, then the x gathers) before using any of them.
  // -------------------------------------------------------------------------------------
  constexpr int kStreamBytes = 16 * 1024;   // products per block in LDS (8 blocks per CU)
  constexpr int kStreamLoads = 8;           // entries per thread whose loads are in flight together

  template<int BS, typename DT>
  constexpr int stream_cap() { return kStreamBytes / (BS * int(sizeof(DT))); }

  template<int BS, typename DT, typename VT, typename IT>
  __global__ __launch_bounds__(kBlock) void k_csr_stream(DT* __restrict__ r, DT a, const DT* __restrict__ x, DT b,
                                                         const DT* y, const VT* __restrict__ val, const IT* __restrict__ col,
                                                         const IT* __restrict__ rp, u64 rows, u64 R, int use_y)
  {
    constexpr int CAP = stream_cap<BS, DT>();
    __shared__ DT prod[CAP * BS];
    const u64 lb = dev::xcd_remap(blockIdx.x, gridDim.x);
    const u64 r0 = lb * R;
    if(r0 >= rows) return;
    const u64 r1 = r0 + R < rows ? r0 + R : rows;
    const u64 e0 = u64(rp[r0]), ne = u64(rp[r1]) - e0;
    if(ne <= u64(CAP))
    {
      // the bounds of this thread's first row, loaded with the entries
      const u64 my = r0 + threadIdx.x;
      const u64 mk0 = my < r1 ? u64(rp[my]) : 0, mk1 = my < r1 ? u64(rp[my + 1]) : 0;
      if constexpr(BS == 1)
        dev::stage_products<kStreamLoads, kBlock>(val, col, e0, ne, x, prod);
      else
      {
      constexpr int U = kStreamLoads / 2;
      for(u64 k0 = threadIdx.x; k0 < ne; k0 += u64(kBlock) * U)
      {
        DT v[U];
        u64 c[U];
#pragma unroll
        for(int j = 0; j < U; ++j)
        {
          const u64 k = k0 + u64(j) * kBlock;
          v[j] = k < ne ? DT(dev::ld_stream(val + e0 + k)) : DT(0);
          c[j] = k < ne ? u64(dev::ld_stream(col + e0 + k)) : 0;
        }
        DT xv[U][BS];
#pragma unroll
        for(int j = 0; j < U; ++j)
#pragma unroll
          for(int h = 0; h < BS; ++h) xv[j][h] = k0 + u64(j) * kBlock < ne ? x[c[j] * BS + h] : DT(0);
#pragma unroll
        for(int j = 0; j < U; ++j)
        {
          const u64 k = k0 + u64(j) * kBlock;
          if(k < ne)
#pragma unroll
            for(int h = 0; h < BS; ++h) prod[k * BS + h] = v[j] * xv[j][h];
        }
      }
      }
      __syncthreads();
      for(u64 row = r0 + threadIdx.x; row < r1; row += kBlock)
      {
        const bool first = row == my;
        const u64 k0 = (first ? mk0 : u64(rp[row])) - e0, k1 = (first ? mk1 : u64(rp[row + 1])) - e0;
        DT s[BS];
#pragma unroll
        for(int h = 0; h < BS; ++h) s[h] = DT(0);
        for(u64 k = k0; k < k1; ++k)
#pragma unroll
          for(int h = 0; h < BS; ++h) s[h] += prod[k * BS + h];
#pragma unroll
        for(int h = 0; h < BS; ++h)
          r[row * BS + h] = use_y ? (s[h] * a) + (b * y[row * BS + h]) : (s[h] * a);
      }
      return;
    }
    constexpr int G = 8;
    const int lane = int(threadIdx.x % G);
    for(u64 row = r0 + threadIdx.x / G; row < r1; row += kBlock / G)
    {
      DT s[BS];
#pragma unroll
      for(int h = 0; h < BS; ++h) s[h] = DT(0);
      for(u64 i = u64(rp[row]) + u64(lane); i < u64(rp[row + 1]); i += G)
      {
        const DT v = dev::ld<DT>(val, i);
        const u64 c = u64(col[i]);
#pragma unroll
        for(int h = 0; h < BS; ++h) s[h] += v * x[c * BS + h];
      }
#pragma unroll
      for(int h = 0; h < BS; ++h) s[h] = dev::group_sum<G>(s[h]);
      if(lane == 0)
      {
#pragma unroll
        for(int h = 0; h < BS; ++h)
          r[row * BS + h] = use_y ? (s[h] * a) + (b * y[row * BS + h]) : (s[h] * a);
      }
    }
  }

  template<int BS, typename DT, typename VT, typename IT>
  int launch_csr_stream(hipStream_t s, DT* r, DT a, const DT* x, DT b, const DT* y, const VT* val, const IT* col,
                        const IT* rp, u64 rows, u64 nnz)
  {
    if(rows == 0) return 0;
    const int use_y = std::abs(double(b)) < double(std::numeric_limits<DT>::epsilon()) ? 0 : 1;
    const double mean = double(nnz) / double(rows);
    const double fit = 0.85 * double(stream_cap<BS, DT>()) / (mean > 1.0 ? mean : 1.0);
    const u64 R = fit < 1.0 ? 1 : (fit > 1024.0 ? 1024 : u64(fit));
    const u64 nb = (rows + R - 1) / R;
    if(nb > 0x7fffffffull) return rt::kErrArg;
    k_csr_stream<BS, DT, VT, IT><<<unsigned(nb), kBlock, 0, s>>>(r, a, x, b, y, val, col, rp, rows, R, use_y);
    return rt::check_launch();
  }

  // transposed CSR: r = a*(b/a*y + A^T x) with atomic scatter (apply_generic.hpp:39-55)
  template<typename DT>
  __global__ void k_scale_copy(DT* r, const DT* y, DT s, int use_y, u64 n)
  {
    const u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if(i < n) r[i] = use_y ? s * y[i] : DT(0);
  }
  template<typename DT, typename VT, typename IT>
  __global__ void k_csr_scatter(DT* r, const DT* x, const VT* val, const IT* col, const IT* rp, u64 rows)
  {
    const u64 row = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if(row >= rows) return;
    const DT xr = x[row];
    for(IT i = rp[row]; i < rp[row + 1]; ++i)
      atomicAdd(&r[col[i]], dev::ld<DT>(val, i) * xr);
  }
  template<typename DT>
  __global__ void k_scale_inplace(DT* r, DT s, u64 n)
  {
    const u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if(i < n) r[i] = s * r[i];
  }

  template<typename DT, typename VT, typename IT>
  int launch_csr_t(hipStream_t s, DT* r, DT a, const DT* x, DT b, const DT* y, const VT* val, const IT* col,
                   const IT* rp, u64 rows, u64 cols)
  {
    const int use_y = std::abs(double(b)) < double(std::numeric_limits<DT>::epsilon()) ? 0 : 1;
    k_scale_copy<DT><<<blocks_for(cols), kBlock, 0, s>>>(r, y, b / a, use_y, cols);
    k_csr_scatter<DT, VT, IT><<<blocks_for(rows), kBlock, 0, s>>>(r, x, val, col, rp, rows);
    k_scale_inplace<DT><<<blocks_for(cols), kBlock, 0, s>>>(r, a, cols);
    return rt::check_launch();
  }

  // -------------------------------------------------------------------------------------
  // SELL-C (FEAT's SparseMatrixELL layout, sparse_matrix_ell.hpp:41-70,865-960) and its
  // sigma-sorted variant: one lane per row, chunk-column-major val/col (coalesced).
  // perm == nullptr: position p is row p (sigma = 1).  rl != nullptr: skip row padding.
  // -------------------------------------------------------------------------------------
  template<typename DT, typename VT, typename IT>
  __global__ __launch_bounds__(kBlock) void k_sell(DT* __restrict__ r, DT a, const DT* __restrict__ x, DT b,
                                                   const DT* y, const VT* __restrict__ val, const IT* __restrict__ col,
                                                   const IT* __restrict__ cs, const IT* __restrict__ cl,
                                                   const IT* __restrict__ rl, const IT* __restrict__ perm,
                                                   u64 C, u64 rows, int use_y)
  {
    const unsigned lb = dev::xcd_remap(blockIdx.x, gridDim.x);
    const u64 p = u64(lb) * kBlock + threadIdx.x;
    if(p >= rows) return;
    const u64 chunk = p / C, loc = p % C;
    const u64 base = u64(cs[chunk]) + loc;
    const u64 len = rl ? u64(rl[p]) : u64(cl[chunk]);
    DT sum = DT(0);
#pragma unroll 4
    for(u64 k = 0; k < len; ++k)
      sum += dev::ld<DT>(val, base + k * C) * x[col[base + k * C]];
    const u64 row = perm ? u64(perm[p]) : p;
    r[row] = use_y ? (sum * a) + (b * y[row]) : (sum * a);
  }

  // -------------------------------------------------------------------------------------
  // Banded (FEAT's SparseMatrixBanded layout, apply_generic.hpp:519-598): num_of_offsets
  // diagonals, val[d*rows + row] on diagonal d with column row + offsets[d] + 1 - rows; the
  // diagonals whose column lies outside [0, columns) do not contribute.  A thread takes kRpt
  // rows t, t + 256, ... of its block's range: a wave reads 64 consecutive values of each
  // diagonal per row set and 64 consecutive x, no indices, with kRpt * num_of_offsets
  // independent loads per thread (measured on 3D Q1 L7: FP32 1 -> 4 rows 52.7 -> 40.8 us,
  // FP64 1 -> 2 rows 108.8 -> 107.5 us, 4 rows slower).
  // -------------------------------------------------------------------------------------
  template<typename DT>
  constexpr int banded_rpt() { return sizeof(DT) == 4 ? 4 : 2; }

  template<typename DT, typename VT, typename IT, int kRpt = banded_rpt<DT>()>
  __global__ __launch_bounds__(kBlock) void k_banded(DT* __restrict__ r, DT a, const DT* __restrict__ x, DT b,
                                                     const DT* y, const VT* __restrict__ val, const IT* __restrict__ offsets,
                                                     u64 ndiag, u64 rows, u64 cols, int use_y)
  {
    const unsigned lb = dev::xcd_remap(blockIdx.x, gridDim.x);
    u64 row[kRpt];
    DT sum[kRpt];
#pragma unroll
    for(int j = 0; j < kRpt; ++j)
    {
      row[j] = (u64(lb) * kRpt + u64(j)) * kBlock + threadIdx.x;
      sum[j] = DT(0);
    }
    if(row[0] >= rows) return;
    for(u64 d = 0; d < ndiag; ++d)
    {
      // column (wraps past 2^64 below 0); an unconditional load from a stand-in column and a
      // select keep the loads of consecutive diagonals in flight together
      const u64 off = u64(offsets[d]) + 1 - rows;
#pragma unroll
      for(int j = 0; j < kRpt; ++j)
      {
        const u64 rr = row[j] < rows ? row[j] : row[0];
        const u64 c = rr + off;
        const bool in = c < cols;
        const DT t = dev::ld<DT>(val, d * rows + rr) * x[in ? c : 0];
        sum[j] += in ? t : DT(0);
      }
    }
#pragma unroll
    for(int j = 0; j < kRpt; ++j)
      if(row[j] < rows) r[row[j]] = use_y ? (b * y[row[j]]) + (a * sum[j]) : (a * sum[j]);
  }

  // -------------------------------------------------------------------------------------
  // BCSR with BH x BW row-major blocks (apply_generic.hpp:121-154): G lanes per block row
  // -------------------------------------------------------------------------------------
  template<int G, int BH, int BW, typename DT, typename IT>
  __global__ __launch_bounds__(kBlock) void k_csrb(DT* __restrict__ r, DT a, const DT* __restrict__ x, DT b,
                                                   const DT* y, const DT* __restrict__ val, const IT* __restrict__ col,
                                                   const IT* __restrict__ rp, u64 rows, int use_y)
  {
    const unsigned lb = dev::xcd_remap(blockIdx.x, gridDim.x);
    const u64 row = (u64(lb) * kBlock + threadIdx.x) / G;
    const int lane = threadIdx.x % G;
    if(row >= rows) return;
    DT s[BH];
#pragma unroll
    for(int h = 0; h < BH; ++h) s[h] = DT(0);
    for(IT i = rp[row] + IT(lane); i < rp[row + 1]; i += IT(G))
    {
      const DT* v = val + u64(i) * (BH * BW);
      const DT* xb = x + u64(col[i]) * BW;
      DT xv[BW];
#pragma unroll
      for(int w = 0; w < BW; ++w) xv[w] = xb[w];
#pragma unroll
      for(int h = 0; h < BH; ++h)
#pragma unroll
        for(int w = 0; w < BW; ++w)
          s[h] += v[h * BW + w] * xv[w];
    }
#pragma unroll
    for(int h = 0; h < BH; ++h)
      s[h] = dev::group_sum<G>(s[h]);
    if(lane == 0)
    {
#pragma unroll
      for(int h = 0; h < BH; ++h)
        r[row * BH + h] = use_y ? (s[h] * a) + (b * y[row * BH + h]) : (s[h] * a);
    }
  }

  template<int BH, int BW, typename DT, typename IT>
  int launch_csrb(hipStream_t s, DT* r, DT a, const DT* x, DT b, const DT* y, const DT* val, const IT* col,
                  const IT* rp, u64 rows, u64 nnzb)
  {
    if(rows == 0) return 0;
    const int use_y = std::abs(double(b)) < double(std::numeric_limits<DT>::epsilon()) ? 0 : 1;
    const int g = lanes_for(rows, nnzb);
    const unsigned nb = blocks_for(rows * u64(g));
#define F3_CSRB_CASE(GG) case GG: k_csrb<GG, BH, BW, DT, IT><<<nb, kBlock, 0, s>>>(r, a, x, b, y, val, col, rp, rows, use_y); break;
    switch(g)
    {
      F3_CSRB_CASE(4) F3_CSRB_CASE(8) F3_CSRB_CASE(16) F3_CSRB_CASE(32) F3_CSRB_CASE(64)
      default: return rt::kErrArg;
    }
#undef F3_CSRB_CASE
    return rt::check_launch();
  }

  // BCSR with any block shape bh x bw up to 9 x 9 (the reference's generic CUDA kernel takes
  // BlockHeight < 10 and any width, apply.hpp:277-281, apply.cu:203-231): the block loops are
  // unrolled to 9 and predicated on the runtime shape, so the partial sums stay in registers
  constexpr int kMaxBlock = 9;
  template<int G, typename DT, typename IT>
  __global__ __launch_bounds__(kBlock) void k_csrb_any(DT* __restrict__ r, DT a, const DT* __restrict__ x, DT b,
                                                       const DT* y, const DT* __restrict__ val, const IT* __restrict__ col,
                                                       const IT* __restrict__ rp, u64 rows, int bh, int bw, int use_y)
  {
    const unsigned lb = dev::xcd_remap(blockIdx.x, gridDim.x);
    const u64 row = (u64(lb) * kBlock + threadIdx.x) / G;
    const int lane = threadIdx.x % G;
    if(row >= rows) return;
    const u64 bb = u64(bh) * u64(bw);
    DT s[kMaxBlock];
#pragma unroll
    for(int h = 0; h < kMaxBlock; ++h) s[h] = DT(0);
    for(IT i = rp[row] + IT(lane); i < rp[row + 1]; i += IT(G))
    {
      const DT* v = val + u64(i) * bb;
      const DT* xb = x + u64(col[i]) * u64(bw);
      DT xv[kMaxBlock];
#pragma unroll
      for(int w = 0; w < kMaxBlock; ++w) xv[w] = w < bw ? xb[w] : DT(0);
#pragma unroll
      for(int h = 0; h < kMaxBlock; ++h)
#pragma unroll
        for(int w = 0; w < kMaxBlock; ++w)
          if(h < bh && w < bw) s[h] += v[h * bw + w] * xv[w];
    }
#pragma unroll
    for(int h = 0; h < kMaxBlock; ++h)
      s[h] = dev::group_sum<G>(s[h]);
    if(lane == 0)
    {
#pragma unroll
      for(int h = 0; h < kMaxBlock; ++h)
        if(h < bh) r[row * u64(bh) + h] = use_y ? (s[h] * a) + (b * y[row * u64(bh) + h]) : (s[h] * a);
    }
  }

  template<typename DT, typename IT>
  int launch_csrb_any(hipStream_t s, DT* r, DT a, const DT* x, DT b, const DT* y, const DT* val, const IT* col,
                      const IT* rp, u64 rows, u64 nnzb, int bh, int bw)
  {
    if(rows == 0) return 0;
    const int use_y = std::abs(double(b)) < double(std::numeric_limits<DT>::epsilon()) ? 0 : 1;
    const int g = lanes_for(rows, nnzb);
    const unsigned nb = blocks_for(rows * u64(g));
#define F3_CSRB_CASE(GG) \
  case GG: k_csrb_any<GG, DT, IT><<<nb, kBlock, 0, s>>>(r, a, x, b, y, val, col, rp, rows, bh, bw, use_y); break;
    switch(g)
    {
      F3_CSRB_CASE(4) F3_CSRB_CASE(8) F3_CSRB_CASE(16) F3_CSRB_CASE(32) F3_CSRB_CASE(64)
      default: return rt::kErrArg;
    }
#undef F3_CSRB_CASE
    return rt::check_launch();
  }

  // scalar CSR times a vector of BS-blocks (csrsb, apply_generic.hpp:156-181)
  template<int G, int BS, typename DT, typename VT, typename IT>
  __global__ __launch_bounds__(kBlock) void k_csrsb(DT* __restrict__ r, DT a, const DT* __restrict__ x, DT b,
                                                    const DT* y, const VT* __restrict__ val, const IT* __restrict__ col,
                                                    const IT* __restrict__ rp, u64 rows, int use_y)
  {
    const unsigned lb = dev::xcd_remap(blockIdx.x, gridDim.x);
    const u64 row = (u64(lb) * kBlock + threadIdx.x) / G;
    const int lane = threadIdx.x % G;
    if(row >= rows) return;
    DT s[BS];
#pragma unroll
    for(int h = 0; h < BS; ++h) s[h] = DT(0);
    for(IT i = rp[row] + IT(lane); i < rp[row + 1]; i += IT(G))
    {
      const DT v = dev::ld<DT>(val, i);
      const DT* xb = x + u64(col[i]) * BS;
#pragma unroll
      for(int h = 0; h < BS; ++h) s[h] += v * xb[h];
    }
#pragma unroll
    for(int h = 0; h < BS; ++h) s[h] = dev::group_sum<G>(s[h]);
    if(lane == 0)
    {
#pragma unroll
      for(int h = 0; h < BS; ++h)
        r[row * BS + h] = use_y ? (s[h] * a) + (b * y[row * BS + h]) : (s[h] * a);
    }
  }

  template<int BS, typename DT, typename VT, typename IT>
  int launch_csrsb(hipStream_t s, DT* r, DT a, const DT* x, DT b, const DT* y, const VT* val, const IT* col,
                   const IT* rp, u64 rows, u64 nnz)
  {
    if(rows == 0) return 0;
    if(rt::csr_lanes_override() <= 0) return launch_csr_stream<BS, DT, VT, IT>(s, r, a, x, b, y, val, col, rp, rows, nnz);
    const int use_y = std::abs(double(b)) < double(std::numeric_limits<DT>::epsilon()) ? 0 : 1;
    const int g = lanes_for(rows, nnz);
    const unsigned nb = blocks_for(rows * u64(g));
#define F3_CSRSB_CASE(GG) case GG: k_csrsb<GG, BS, DT, VT, IT><<<nb, kBlock, 0, s>>>(r, a, x, b, y, val, col, rp, rows, use_y); break;
    switch(g)
    {
      F3_CSRSB_CASE(4) F3_CSRSB_CASE(8) F3_CSRSB_CASE(16) F3_CSRSB_CASE(32) F3_CSRSB_CASE(64)
      default: return rt::kErrArg;
    }
#undef F3_CSRSB_CASE
    return rt::check_launch();
  }

  // -------------------------------------------------------------------------------------
  // BLAS-1 and friends (axpy_generic.hpp, component_*_generic.hpp, scale_generic.hpp)
  // -------------------------------------------------------------------------------------
  template<typename DT>
  __global__ void k_axpy(DT* r, DT a, const DT* x, const DT* y, u64 n)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      r[i] = (a * x[i]) + y[i];
  }
  // 16-byte accesses (V = 16 / sizeof(DT) elements) when r, x, y are aligned for them; the
  // n % V trailing elements by the first threads of the grid
  template<typename DT, int V>
  __global__ __launch_bounds__(kBlock) void k_axpy_vec(DT* r, DT a, const DT* x, const DT* y, u64 n)
  {
    typedef DT vt __attribute__((ext_vector_type(V)));
    const u64 nv = n / V;
    const u64 step = u64(gridDim.x) * blockDim.x;
    const u64 gid = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    for(u64 i = gid; i < nv; i += step)
    {
      const vt xv = reinterpret_cast<const vt*>(x)[i], yv = reinterpret_cast<const vt*>(y)[i];
      vt o;
#pragma unroll
      for(int q = 0; q < V; ++q) o[q] = (a * xv[q]) + yv[q];
      reinterpret_cast<vt*>(r)[i] = o;
    }
    const u64 t = nv * V + gid;
    if(t < n) r[t] = (a * x[t]) + y[t];
  }

  template<typename DT>
  __global__ void k_cprod(DT* r, const DT* x, const DT* y, u64 n)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      r[i] = x[i] * y[i];
  }
  template<typename DT>
  __global__ void k_cinv(DT* r, const DT* x, DT s, u64 n)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      r[i] = s / x[i];
  }
  template<typename DT>
  __global__ void k_scale(DT* r, const DT* x, DT s, u64 n)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      r[i] = x[i] * s;
  }
  template<typename T>
  __global__ void k_set(T* r, T v, u64 n)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      r[i] = v;
  }
  template<typename DO, typename DI>
  __global__ void k_convert(DO* r, const DI* x, u64 n)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      r[i] = DO(x[i]);
  }

  inline unsigned stream_blocks(u64 n)
  {
    const u64 b = (n + kBlock - 1) / kBlock;
    return unsigned(std::max<u64>(1, std::min<u64>(b, 8192)));
  }

  // dot / norm: deterministic two-kernel reduction into a device scalar (block partials,
  // then one block sums them in block order)
  template<typename DT>
  __global__ __launch_bounds__(kBlock) void k_dot(const DT* x, const DT* y, u64 n, DT* partials)
  {
    // four accumulators: four independent load pairs in flight per thread
    DT s0 = DT(0), s1 = DT(0), s2 = DT(0), s3 = DT(0);
    const u64 step = u64(gridDim.x) * blockDim.x;
    u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    for(; i + 3 * step < n; i += 4 * step)
    {
      s0 += x[i] * y[i];
      s1 += x[i + step] * y[i + step];
      s2 += x[i + 2 * step] * y[i + 2 * step];
      s3 += x[i + 3 * step] * y[i + 3 * step];
    }
    for(; i < n; i += step) s0 += x[i] * y[i];
    dev::block_partial((s0 + s1) + (s2 + s3), partials);
  }
  // the same with 16-byte loads of x and y (aligned for V = 16 / sizeof(DT) elements)
  template<typename DT, int V>
  __global__ __launch_bounds__(kBlock) void k_dot_vec(const DT* x, const DT* y, u64 n, DT* partials)
  {
    typedef DT vt __attribute__((ext_vector_type(V)));
    const u64 nv = n / V;
    const u64 step = u64(gridDim.x) * blockDim.x;
    const u64 gid = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    DT s[2][V];
#pragma unroll
    for(int q = 0; q < V; ++q) { s[0][q] = DT(0); s[1][q] = DT(0); }
    u64 i = gid;
    for(; i + step < nv; i += 2 * step)
    {
      const vt x0 = reinterpret_cast<const vt*>(x)[i], y0 = reinterpret_cast<const vt*>(y)[i];
      const vt x1 = reinterpret_cast<const vt*>(x)[i + step], y1 = reinterpret_cast<const vt*>(y)[i + step];
#pragma unroll
      for(int q = 0; q < V; ++q) { s[0][q] += x0[q] * y0[q]; s[1][q] += x1[q] * y1[q]; }
    }
    if(i < nv)
    {
      const vt x0 = reinterpret_cast<const vt*>(x)[i], y0 = reinterpret_cast<const vt*>(y)[i];
#pragma unroll
      for(int q = 0; q < V; ++q) s[0][q] += x0[q] * y0[q];
    }
    const u64 t = nv * V + gid;
    if(t < n) s[1][0] += x[t] * y[t];
    DT tot = DT(0);
#pragma unroll
    for(int q = 0; q < V; ++q) tot += s[0][q] + s[1][q];
    dev::block_partial(tot, partials);
  }
  template<typename DT, bool SQRT>
  __global__ __launch_bounds__(kBlock) void k_dot_fin(const DT* partials, unsigned nb, DT* out)
  {
    const DT t = dev::sum_partials(partials, nb);
    if(threadIdx.x == 0) *out = SQRT ? DT(sqrt(t)) : t;
  }

  // partials of x.y (or x.x) from at most 2048 blocks (two per SIMD: each thread reads several
  // elements, and the finalisation sums few partials), then the one-block finalisation into out
  template<typename DT, bool SQRT>
  int enqueue_dot(hipStream_t s, const DT* x, const DT* y, u64 n, DT* part, unsigned max_blocks, DT* out)
  {
    const unsigned nb = unsigned(std::min<u64>(std::min<u64>(stream_blocks(n), 2048), max_blocks));
    constexpr int V = 16 / int(sizeof(DT));
    if(reinterpret_cast<std::uintptr_t>(x) % 16 == 0 && reinterpret_cast<std::uintptr_t>(y) % 16 == 0)
      k_dot_vec<DT, V><<<nb, kBlock, 0, s>>>(x, y, n, part);
    else
      k_dot<DT><<<nb, kBlock, 0, s>>>(x, y, n, part);
    k_dot_fin<DT, SQRT><<<1, kBlock, 0, s>>>(part, nb, out);
    return rt::check_launch();
  }

  // the drop-in reductions share the device's workspace (feat3_hip_initialise): the stream waits
  // for the previous reduction that used it (on any stream) and records the workspace's event
  // after its own finalisation.  Inside a stream capture the captured graph orders its nodes
  // itself, and the event is left alone (capture does not take foreign events).
  template<typename DT, bool SQRT>
  int launch_dot(hipStream_t s, const DT* x, const DT* y, u64 n, DT* out, DT* host_out = nullptr)
  {
    rt::Workspace* ws = rt::workspace();
    if(!ws) return rt::kErrState;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if(hipStreamIsCapturing(s, &cap) != hipSuccess) return rt::kErrRuntime;
    const bool order = cap == hipStreamCaptureStatusNone;
    if(order && hipStreamWaitEvent(s, ws->done, 0) != hipSuccess) return rt::kErrRuntime;
    const int rc = enqueue_dot<DT, SQRT>(s, x, y, n, reinterpret_cast<DT*>(ws->partials), rt::Workspace::kMaxPartials, out);
    if(rc) return rc;
    // a host-returning reduction reads the workspace's scalar before the next one may write it
    if(host_out && hipMemcpyAsync(host_out, out, sizeof(DT), hipMemcpyDeviceToHost, s) != hipSuccess)
      return rt::kErrRuntime;
    if(order && hipEventRecord(ws->done, s) != hipSuccess) return rt::kErrRuntime;
    return rt::kOk;
  }

  template<typename DT, bool SQRT>
  int dot_value(hipStream_t s, const DT* x, const DT* y, u64 n, DT* host_out)
  {
    rt::Workspace* ws = rt::workspace();
    if(!ws) return rt::kErrState;
    DT* dres = reinterpret_cast<DT*>(ws->scalar);
    int rc = launch_dot<DT, SQRT>(s, x, y, n, dres, host_out);
    if(rc) return rc;
    if(hipStreamSynchronize(s) != hipSuccess) return rt::kErrRuntime;
    return 0;
  }

  // diagonal extraction (diagonal_generic.hpp:20-81)
  template<typename DT, typename IT>
  __global__ void k_diag_csr(DT* d, const DT* val, const IT* col, const IT* rp, u64 rows)
  {
    const u64 row = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if(row >= rows) return;
    DT v = DT(0);
    for(IT i = rp[row]; i < rp[row + 1]; ++i)
      if(u64(col[i]) == row) { v = val[i]; break; }
    d[row] = v;
  }
  template<typename DT, typename IT>
  __global__ void k_diag_ell(DT* d, const DT* val, const IT* col, const IT* cs, u64 C, u64 rows)
  {
    const u64 row = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if(row >= rows) return;
    const u64 chunk = row / C;
    DT v = DT(0);
    for(u64 p = u64(cs[chunk]) + row % C; p < u64(cs[chunk + 1]); p += C)
      if(u64(col[p]) == row) { v = val[p]; break; }
    d[row] = v;
  }
  template<typename DT, typename IT>
  __global__ void k_diag_csrb(DT* d, const DT* val, const IT* col, const IT* rp, u64 rows, int bh, int bw)
  {
    const u64 row = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if(row >= rows) return;
    for(int i = 0; i < bh; ++i) d[row * bh + i] = DT(0);
    for(IT c = rp[row]; c < rp[row + 1]; ++c)
      if(u64(col[c]) == row)
        for(int i = 0; i < bh; ++i) d[row * bh + i] = val[u64(c) * bh * bw + i + i * bw];
  }

  // unit filter (unit_filter_generic.hpp:21-37, unit_filter_blocked_generic.hpp)
  template<typename DT, typename IT>
  __global__ void k_filter_rhs(DT* v, const DT* sv, const IT* idx, u64 ue, int bs)
  {
    const u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if(i >= ue * u64(bs)) return;
    const u64 k = i / bs, c = i % bs;
    v[u64(idx[k]) * bs + c] = sv[k * bs + c];
  }
  template<typename DT, typename IT>
  __global__ void k_filter_def(DT* v, const IT* idx, u64 ue, int bs)
  {
    const u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if(i >= ue * u64(bs)) return;
    v[u64(idx[i / bs]) * bs + i % bs] = DT(0);
  }
} // namespace

// ---- scratch of the no-allocation dot product (lafem_hip_dot_noalloc_f64) -------------------
// The block partials of that entry point live in this file-scope device array: every GPU the
// library runs on holds its own copy until the process ends, and calls on one GPU share it.  In
// host code the array's name is a host-side stand-in, never a device address: host code uses
// only the address hipGetSymbolAddress returns, taken once per device with that device current
// (feat3_hip_initialise, or the entry point's first call on the device) and kept per device.
static constexpr unsigned kNoAllocPartials = 2048;   // the reduction grid's largest block count
static __device__ __align__(256) double scratch[kNoAllocPartials];

namespace
{
  constexpr int kMaxScratchDevices = 64;
  std::mutex g_scratch_mutex;
  double* g_scratch[kMaxScratchDevices] = {};

  // the current device's address of `scratch`, looked up at the device's first use
  int scratch_address(double** out)
  {
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess) return rt::kErrRuntime;
    if(dev < 0 || dev >= kMaxScratchDevices) return rt::kErrArg;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if(!g_scratch[dev])
    {
      void* p = nullptr;
      if(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return rt::kErrRuntime;
      g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return rt::kOk;
  }
} // namespace

namespace feat3_amd
{
namespace rt
{
  int setup_noalloc_scratch()
  {
    double* p = nullptr;
    return scratch_address(&p);
  }
} // namespace rt
} // namespace feat3_amd

// =========================================================================================
// C ABI
// =========================================================================================
#define S_ static_cast<hipStream_t>(stream)
#define F3_EPS(DT) std::numeric_limits<DT>::epsilon()

extern "C" {

// ---- Apply<Mem::HIP>::csr  (apply.hpp:271-274) ----------------------------------------
#define F3_DEF_CSR(SUF, DT, VT, IT)                                                                          \
  int lafem_hip_apply_csr_##SUF(void* stream, DT* r, DT a, const DT* x, DT b, const DT* y, const VT* val,      \
                                const IT* col_ind, const IT* row_ptr, u64 rows, u64 columns, u64 used_elements, \
                                int transposed)                                                               \
  {                                                                                                           \
    if(!r || !row_ptr || (used_elements > 0 && (!x || !val || !col_ind))) return rt::kErrArg;                 \
    if(transposed) return launch_csr_t<DT, VT, IT>(S_, r, a, x, b, y, val, col_ind, row_ptr, rows, columns);  \
    if(rt::csr_lanes_override() > 0)                                                                          \
      return launch_csr<DT, VT, IT>(S_, r, a, x, b, y, val, col_ind, row_ptr, rows, used_elements);           \
    return launch_csr_stream<1, DT, VT, IT>(S_, r, a, x, b, y, val, col_ind, row_ptr, rows, used_elements);   \
  }
F3_DEF_CSR(f64_u32, double, double, u32)
F3_DEF_CSR(f64_u64, double, double, u64)
F3_DEF_CSR(f32_u32, float, float, u32)
F3_DEF_CSR(f32_u64, float, float, u64)
#undef F3_DEF_CSR

int lafem_hip_apply_csr_h16f32_u32(void* stream, float* r, float a, const float* x, float b, const float* y,
                                   const void* val, const u32* col_ind, const u32* row_ptr, u64 rows, u64 columns,
                                   u64 used_elements)
{
  if(!r || !row_ptr || (used_elements > 0 && (!x || !val || !col_ind))) return rt::kErrArg;
  (void)columns;
  if(rt::csr_lanes_override() > 0)
    return launch_csr<float, half, u32>(S_, r, a, x, b, y, static_cast<const half*>(val), col_ind, row_ptr, rows, used_elements);
  return launch_csr_stream<1, float, half, u32>(S_, r, a, x, b, y, static_cast<const half*>(val), col_ind, row_ptr, rows,
                                                used_elements);
}

// ---- Apply<Mem::HIP>::ell  (apply.hpp:296) and the sigma-sorted SELL variant -------------
#define F3_DEF_ELL(SUF, DT, IT)                                                                              \
  int lafem_hip_apply_ell_##SUF(void* stream, DT* r, DT a, const DT* x, DT b, const DT* y, const DT* val,      \
                                const IT* col_ind, const IT* cs, const IT* cl, u64 C, u64 rows)               \
  {                                                                                                           \
    if(!r || !x || !val || !col_ind || !cs || !cl || C == 0) return rt::kErrArg;                              \
    if(rows == 0) return 0;                                                                                   \
    const int use_y = std::abs(double(b)) < double(F3_EPS(DT)) ? 0 : 1;                                       \
    k_sell<DT, DT, IT><<<blocks_for(rows), kBlock, 0, S_>>>(r, a, x, b, y, val, col_ind, cs, cl, nullptr,    \
                                                            nullptr, C, rows, use_y);                         \
    return rt::check_launch();                                                                                \
  }                                                                                                           \
  int lafem_hip_apply_sell_##SUF(void* stream, DT* r, DT a, const DT* x, DT b, const DT* y, const DT* val,     \
                                 const IT* col_ind, const IT* cs, const IT* cl, const IT* rl, const IT* perm,  \
                                 u64 C, u64 rows)                                                             \
  {                                                                                                           \
    if(!r || !x || !val || !col_ind || !cs || !cl || C == 0) return rt::kErrArg;                              \
    if(rows == 0) return 0;                                                                                   \
    const int use_y = std::abs(double(b)) < double(F3_EPS(DT)) ? 0 : 1;                                       \
    k_sell<DT, DT, IT><<<blocks_for(rows), kBlock, 0, S_>>>(r, a, x, b, y, val, col_ind, cs, cl, rl, perm, C, \
                                                            rows, use_y);                                     \
    return rt::check_launch();                                                                                \
  }
F3_DEF_ELL(f64_u32, double, u32)
F3_DEF_ELL(f64_u64, double, u64)
F3_DEF_ELL(f32_u32, float, u32)
F3_DEF_ELL(f32_u64, float, u64)
#undef F3_DEF_ELL

int lafem_hip_apply_sell_h16f32_u32(void* stream, float* r, float a, const float* x, float b, const float* y,
                                    const void* val, const u32* col_ind, const u32* cs, const u32* cl,
                                    const u32* rl, const u32* perm, u64 C, u64 rows)
{
  if(!r || !x || !val || !col_ind || !cs || !cl || C == 0) return rt::kErrArg;
  if(rows == 0) return 0;
  const int use_y = std::abs(b) < F3_EPS(float) ? 0 : 1;
  k_sell<float, half, u32><<<blocks_for(rows), kBlock, 0, S_>>>(r, a, x, b, y, static_cast<const half*>(val), col_ind,
                                                                cs, cl, rl, perm, C, rows, use_y);
  return rt::check_launch();
}

// ---- Apply<Mem::HIP>::banded  (apply.hpp:162,299) ----------------------------------------
#define F3_DEF_BANDED(SUF, DT, IT)                                                                           \
  int lafem_hip_apply_banded_##SUF(void* stream, DT* r, DT a, const DT* x, DT b, const DT* y, const DT* val,   \
                                   const IT* offsets, u64 num_of_offsets, u64 rows, u64 columns)              \
  {                                                                                                           \
    if(!r || (num_of_offsets > 0 && (!x || !val || !offsets))) return rt::kErrArg;                            \
    if(rows == 0) return 0;                                                                                   \
    const int use_y = std::abs(double(b)) < double(F3_EPS(DT)) ? 0 : 1;                                       \
    if(use_y && !y) return rt::kErrArg;                                                                       \
    /* no column: no diagonal contributes (and x holds nothing to stand in) */                                \
    k_banded<DT, DT, IT><<<blocks_for(rows, kBlock * banded_rpt<DT>()), kBlock, 0, S_>>>(r, a, x, b, y, val, offsets,                    \
                                                              columns ? num_of_offsets : 0, rows, columns,    \
                                                              use_y);                                         \
    return rt::check_launch();                                                                                \
  }
F3_DEF_BANDED(f64_u32, double, u32)
F3_DEF_BANDED(f64_u64, double, u64)
F3_DEF_BANDED(f32_u32, float, u32)
F3_DEF_BANDED(f32_u64, float, u64)
#undef F3_DEF_BANDED

// ---- Apply<Mem::HIP>::csrb  (apply.hpp:277) --------------------------------------------
#define F3_DEF_CSRB(SUF, DT, IT)                                                                             \
  int lafem_hip_apply_csrb_##SUF(void* stream, DT* r, DT a, const DT* x, DT b, const DT* y, const DT* val,     \
                                 const IT* col_ind, const IT* row_ptr, u64 rows, u64 columns,                 \
                                 u64 used_elements, int block_height, int block_width)                        \
  {                                                                                                           \
    (void)columns;                                                                                            \
    if(!r || !row_ptr || (used_elements > 0 && (!x || !val || !col_ind))) return rt::kErrArg;                 \
    if(block_height < 1 || block_height > kMaxBlock || block_width < 1 || block_width > kMaxBlock)           \
      return rt::kErrArg;                                                                                     \
    if(block_height == block_width)                                                                           \
      switch(block_height)                                                                                    \
      {                                                                                                       \
        case 1: return launch_csrb<1, 1, DT, IT>(S_, r, a, x, b, y, val, col_ind, row_ptr, rows, used_elements); \
        case 2: return launch_csrb<2, 2, DT, IT>(S_, r, a, x, b, y, val, col_ind, row_ptr, rows, used_elements); \
        case 3: return launch_csrb<3, 3, DT, IT>(S_, r, a, x, b, y, val, col_ind, row_ptr, rows, used_elements); \
        case 4: return launch_csrb<4, 4, DT, IT>(S_, r, a, x, b, y, val, col_ind, row_ptr, rows, used_elements); \
        default: break;                                                                                       \
      }                                                                                                       \
    return launch_csrb_any<DT, IT>(S_, r, a, x, b, y, val, col_ind, row_ptr, rows, used_elements, block_height, \
                                   block_width);                                                              \
  }
F3_DEF_CSRB(f64_u32, double, u32)
F3_DEF_CSRB(f64_u64, double, u64)
F3_DEF_CSRB(f32_u32, float, u32)
F3_DEF_CSRB(f32_u64, float, u64)
#undef F3_DEF_CSRB

// ---- Apply<Mem::HIP>::csrsb  (apply.hpp:283) -------------------------------------------
#define F3_DEF_CSRSB(SUF, DT, IT)                                                                            \
  int lafem_hip_apply_csrsb_##SUF(void* stream, DT* r, DT a, const DT* x, DT b, const DT* y, const DT* val,    \
                                  const IT* col_ind, const IT* row_ptr, u64 rows, u64 columns,                \
                                  u64 used_elements, int block_size)                                          \
  {                                                                                                           \
    (void)columns;                                                                                            \
    if(!r || !row_ptr || (used_elements > 0 && (!x || !val || !col_ind))) return rt::kErrArg;                 \
    switch(block_size)                                                                                        \
    {                                                                                                         \
      case 1: return launch_csrsb<1, DT, DT, IT>(S_, r, a, x, b, y, val, col_ind, row_ptr, rows, used_elements); \
      case 2: return launch_csrsb<2, DT, DT, IT>(S_, r, a, x, b, y, val, col_ind, row_ptr, rows, used_elements); \
      case 3: return launch_csrsb<3, DT, DT, IT>(S_, r, a, x, b, y, val, col_ind, row_ptr, rows, used_elements); \
      default: return rt::kErrArg;                                                                            \
    }                                                                                                         \
  }
F3_DEF_CSRSB(f64_u32, double, u32)
F3_DEF_CSRSB(f64_u64, double, u64)
F3_DEF_CSRSB(f32_u32, float, u32)
F3_DEF_CSRSB(f32_u64, float, u64)
#undef F3_DEF_CSRSB

// ---- BLAS-1 ---------------------------------------------------------------------------
#define F3_DEF_BLAS1(SUF, DT)                                                                                \
  int lafem_hip_axpy_##SUF(void* stream, DT* r, DT a, const DT* x, const DT* y, u64 n)                        \
  {                                                                                                           \
    if(n == 0) return 0;                                                                                      \
    if(!r || !x || !y) return rt::kErrArg;                                                                    \
    if(reinterpret_cast<std::uintptr_t>(r) % 16 == 0 && reinterpret_cast<std::uintptr_t>(x) % 16 == 0 &&         \
       reinterpret_cast<std::uintptr_t>(y) % 16 == 0)                                                         \
      k_axpy_vec<DT, 16 / int(sizeof(DT))><<<stream_blocks((n + 1) / 2), kBlock, 0, S_>>>(r, a, x, y, n);     \
    else                                                                                                      \
      k_axpy<DT><<<stream_blocks(n), kBlock, 0, S_>>>(r, a, x, y, n);                                         \
    return rt::check_launch();                                                                                \
  }                                                                                                           \
  int lafem_hip_dot_##SUF(void* stream, const DT* x, const DT* y, u64 n, DT* result_dev)                      \
  {                                                                                                           \
    if(!result_dev || (n > 0 && (!x || !y))) return rt::kErrArg;                                              \
    return launch_dot<DT, false>(S_, x, y, n, result_dev);                                                    \
  }                                                                                                           \
  int lafem_hip_dot_value_##SUF(void* stream, const DT* x, const DT* y, u64 n, DT* result_host)               \
  {                                                                                                           \
    if(!result_host || (n > 0 && (!x || !y))) return rt::kErrArg;                                             \
    return dot_value<DT, false>(S_, x, y, n, result_host);                                                    \
  }                                                                                                           \
  int lafem_hip_norm2_##SUF(void* stream, const DT* x, u64 n, DT* result_dev)                                 \
  {                                                                                                           \
    if(!result_dev || (n > 0 && !x)) return rt::kErrArg;                                                      \
    return launch_dot<DT, true>(S_, x, x, n, result_dev);                                                     \
  }                                                                                                           \
  int lafem_hip_norm2_value_##SUF(void* stream, const DT* x, u64 n, DT* result_host)                          \
  {                                                                                                           \
    if(!result_host || (n > 0 && !x)) return rt::kErrArg;                                                     \
    return dot_value<DT, true>(S_, x, x, n, result_host);                                                     \
  }                                                                                                           \
  int lafem_hip_component_product_##SUF(void* stream, DT* r, const DT* x, const DT* y, u64 n)                 \
  {                                                                                                           \
    if(n == 0) return 0;                                                                                      \
    if(!r || !x || !y) return rt::kErrArg;                                                                    \
    k_cprod<DT><<<stream_blocks(n), kBlock, 0, S_>>>(r, x, y, n);                                             \
    return rt::check_launch();                                                                                \
  }                                                                                                           \
  int lafem_hip_component_invert_##SUF(void* stream, DT* r, const DT* x, DT s, u64 n)                         \
  {                                                                                                           \
    if(n == 0) return 0;                                                                                      \
    if(!r || !x) return rt::kErrArg;                                                                          \
    k_cinv<DT><<<stream_blocks(n), kBlock, 0, S_>>>(r, x, s, n);                                              \
    return rt::check_launch();                                                                                \
  }                                                                                                           \
  int lafem_hip_scale_##SUF(void* stream, DT* r, const DT* x, DT s, u64 n)                                    \
  {                                                                                                           \
    if(n == 0) return 0;                                                                                      \
    if(!r || !x) return rt::kErrArg;                                                                          \
    k_scale<DT><<<stream_blocks(n), kBlock, 0, S_>>>(r, x, s, n);                                             \
    return rt::check_launch();                                                                                \
  }
F3_DEF_BLAS1(f64, double)
F3_DEF_BLAS1(f32, float)
#undef F3_DEF_BLAS1

// ---- Diagonal<Mem::HIP> -----------------------------------------------------------------
#define F3_DEF_DIAG(SUF, DT, IT)                                                                             \
  int lafem_hip_diagonal_csr_##SUF(void* stream, DT* diag, const DT* val, const IT* col_ind, const IT* row_ptr, \
                                   u64 rows)                                                                  \
  {                                                                                                           \
    if(rows == 0) return 0;                                                                                   \
    if(!diag || !val || !col_ind || !row_ptr) return rt::kErrArg;                                             \
    k_diag_csr<DT, IT><<<blocks_for(rows), kBlock, 0, S_>>>(diag, val, col_ind, row_ptr, rows);               \
    return rt::check_launch();                                                                                \
  }                                                                                                           \
  int lafem_hip_diagonal_ell_##SUF(void* stream, DT* diag, const DT* val, const IT* col_ind, const IT* cs,    \
                                   const IT* cl, u64 C, u64 rows)                                             \
  {                                                                                                           \
    (void)cl;                                                                                                 \
    if(rows == 0) return 0;                                                                                   \
    if(!diag || !val || !col_ind || !cs || C == 0) return rt::kErrArg;                                        \
    k_diag_ell<DT, IT><<<blocks_for(rows), kBlock, 0, S_>>>(diag, val, col_ind, cs, C, rows);                 \
    return rt::check_launch();                                                                                \
  }                                                                                                           \
  int lafem_hip_diagonal_csrb_##SUF(void* stream, DT* diag, const DT* val, const IT* col_ind,                 \
                                    const IT* row_ptr, u64 rows, int block_height, int block_width)           \
  {                                                                                                           \
    if(rows == 0) return 0;                                                                                   \
    if(!diag || !val || !col_ind || !row_ptr || block_height < 1 || block_height > 16 ||                     \
       block_width != block_height)                                                                           \
      return rt::kErrArg;   /* the diagonal of square blocks only */                                          \
    k_diag_csrb<DT, IT><<<blocks_for(rows), kBlock, 0, S_>>>(diag, val, col_ind, row_ptr, rows, block_height, \
                                                             block_width);                                    \
    return rt::check_launch();                                                                                \
  }
F3_DEF_DIAG(f64_u32, double, u32)
F3_DEF_DIAG(f64_u64, double, u64)
F3_DEF_DIAG(f32_u32, float, u32)
F3_DEF_DIAG(f32_u64, float, u64)
#undef F3_DEF_DIAG

// ---- UnitFilter<Mem::HIP> / UnitFilterBlocked<Mem::HIP> -------------------------------
#define F3_DEF_FILTER(SUF, DT, IT)                                                                           \
  int lafem_hip_unit_filter_rhs_##SUF(void* stream, DT* v, const DT* sv_elements, const IT* sv_indices, u64 ue, \
                                      int block_size)                                                         \
  {                                                                                                           \
    if(ue == 0) return 0;                                                                                     \
    if(!v || !sv_elements || !sv_indices || block_size < 1) return rt::kErrArg;                               \
    k_filter_rhs<DT, IT><<<blocks_for(ue * u64(block_size)), kBlock, 0, S_>>>(v, sv_elements, sv_indices, ue, \
                                                                              block_size);                    \
    return rt::check_launch();                                                                                \
  }                                                                                                           \
  int lafem_hip_unit_filter_def_##SUF(void* stream, DT* v, const IT* sv_indices, u64 ue, int block_size)      \
  {                                                                                                           \
    if(ue == 0) return 0;                                                                                     \
    if(!v || !sv_indices || block_size < 1) return rt::kErrArg;                                               \
    k_filter_def<DT, IT><<<blocks_for(ue * u64(block_size)), kBlock, 0, S_>>>(v, sv_indices, ue, block_size); \
    return rt::check_launch();                                                                                \
  }
F3_DEF_FILTER(f64_u32, double, u32)
F3_DEF_FILTER(f64_u64, double, u64)
F3_DEF_FILTER(f32_u32, float, u32)
F3_DEF_FILTER(f32_u64, float, u64)
#undef F3_DEF_FILTER

// ---- MemoryPool<Mem::HIP>::set_memory / convert  (memory_pool.cu:26-41,176-227) --------
#define F3_DEF_SET(SUF, T)                                                                                   \
  int lafem_hip_set_memory_##SUF(void* stream, T* dst, T value, u64 n)                                        \
  {                                                                                                           \
    if(n == 0) return 0;                                                                                      \
    if(!dst) return rt::kErrArg;                                                                              \
    k_set<T><<<stream_blocks(n), kBlock, 0, S_>>>(dst, value, n);                                             \
    return rt::check_launch();                                                                                \
  }
F3_DEF_SET(f64, double)
F3_DEF_SET(f32, float)
F3_DEF_SET(u32, u32)
F3_DEF_SET(u64, u64)
#undef F3_DEF_SET

#define F3_DEF_CONV(SUF, DO, DI)                                                                             \
  int lafem_hip_convert_##SUF(void* stream, DO* dst, const DI* src, u64 n)                                    \
  {                                                                                                           \
    if(n == 0) return 0;                                                                                      \
    if(!dst || !src) return rt::kErrArg;                                                                      \
    k_convert<DO, DI><<<stream_blocks(n), kBlock, 0, S_>>>(dst, src, n);                                      \
    return rt::check_launch();                                                                                \
  }
F3_DEF_CONV(f32_f64, float, double)
F3_DEF_CONV(f64_f32, double, float)
F3_DEF_CONV(u32_u64, u32, u64)
F3_DEF_CONV(u64_u32, u64, u32)
#undef F3_DEF_CONV

// <x, y> into a device scalar with the kernels and grid of lafem_hip_dot_f64 (identical
// result), the block partials in `scratch` instead of the device's workspace: no allocation
int lafem_hip_dot_noalloc_f64(void* stream, const double* x, const double* y, u64 n, double* result_dev)
{
  if(!result_dev || (n > 0 && (!x || !y))) return rt::kErrArg;
  double* part = nullptr;
  const int rc = scratch_address(&part);
  if(rc) return rc;
  return enqueue_dot<double, false>(S_, x, y, n, part, kNoAllocPartials, result_dev);
}

int lafem_hip_convert_h16_f32(void* stream, void* dst, const float* src, u64 n)
{
  if(n == 0) return 0;
  if(!dst || !src) return rt::kErrArg;
  k_convert<half, float><<<stream_blocks(n), kBlock, 0, S_>>>(static_cast<half*>(dst), src, n);
  return rt::check_launch();
}
int lafem_hip_convert_f32_h16(void* stream, float* dst, const void* src, u64 n)
{
  if(n == 0) return 0;
  if(!dst || !src) return rt::kErrArg;
  k_convert<float, half><<<stream_blocks(n), kBlock, 0, S_>>>(dst, static_cast<const half*>(src), n);
  return rt::check_launch();
}

} // extern "C"
