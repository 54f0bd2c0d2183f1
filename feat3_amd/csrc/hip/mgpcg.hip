// mgpcg.hip -- the device-resident PCG preconditioned by a multigrid V-cycle, with the
// Solver lifecycle of the reference (solver/base.hpp:184-316: init_symbolic, init_numeric,
// apply, correct, done) behind the feat3_mgpcg_* C ABI.
//
// Reference path restated (one GPU, no Global layer):
//   PCG::_apply_intern          solver/pcg.hpp:188-288
//   IterativeSolver criteria    solver/iterative.hpp:213-237,476-505,744-879
//   MultiGrid V-cycle           solver/multigrid.hpp:1268-1351,1524-1587,1711-2000
//   Richardson(Jacobi) smoother solver/richardson.hpp:187-228, jacobi_precond.hpp:137-165
//   ConvertPrecond              solver/convert_precond.hpp:100-119 (FP64 outer / FP32 inner)
//
// Layout in HBM: per level CSR (u32 row_ptr/col_ind, values in the inner storage type),
// optional SELL-C-sigma copy, Dirichlet row bitmask, dinv, and the four level vectors
// rhs/x/t/def in the inner precision; the outer PCG holds the FP64 finest matrix and
// x, r, p, q in FP64 plus a PcgState record of device scalars.  The whole PCG iteration
// (SpMV+dot, update+norm, V-cycle, dot, p update) is captured once into a hipGraph and
// replayed; every kernel returns immediately once the device status leaves "progress", so
// the host only polls the status one iteration behind.

#include "device_common.hpp"
#include "mg_kernels.hpp"
#include "runtime.hpp"
#include "feat3_hip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

using namespace feat3_amd;
using feat3_amd::dev::half;
typedef std::uint32_t u32;
typedef std::uint64_t u64;

namespace
{
  constexpr int kBlock = 256;

#define F3_CHECK(...) do { if((__VA_ARGS__) != hipSuccess) return rt::kErrRuntime; } while(0)
#define F3_RC(...) do { const int rc_ = (__VA_ARGS__); if(rc_) return rc_; } while(0)

  struct DevMem
  {
    void* p = nullptr;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() { if(p) (void)hipFree(p); }
    template<typename T> T* as() const { return static_cast<T*>(p); }
    int alloc(u64 bytes)
    {
      if(p) { (void)hipFree(p); p = nullptr; }
      if(bytes == 0) bytes = 16;
      return hipMalloc(&p, bytes) == hipSuccess ? rt::kOk : rt::kErrRuntime;
    }
  };

  template<typename VT> VT to_storage(double v) { return VT(v); }
  template<> half to_storage<half>(double v) { return half(float(v)); }

  // host copy of one CSR matrix (FP64 values), as pushed through the C ABI
  struct HostCsr
  {
    u64 rows = 0, cols = 0;
    std::vector<u32> rp, col;
    std::vector<double> val;
    // device-resident source (feat3_mgpcg_push_level_device): the arrays above stay empty
    const u32* drp = nullptr;
    const u32* dcol = nullptr;
    const double* dval = nullptr;
    u64 dnnz = 0;
    bool on_device() const { return drp != nullptr; }
    u64 nnz() const { return on_device() ? dnnz : (rp.empty() ? 0 : rp[rows]); }
  };

  // ---- setup kernels for device-resident level matrices -----------------------------------
  __global__ void k_validate_csr(u64 rows, u64 cols, u64 nnz, const u32* __restrict__ rp, const u32* __restrict__ col,
                                 int* bad)
  {
    for(u64 r = u64(blockIdx.x) * blockDim.x + threadIdx.x; r < rows; r += u64(gridDim.x) * blockDim.x)
    {
      const u64 a = rp[r], b = rp[r + 1];
      bool ok = a <= b && b <= nnz && (r > 0 || a == 0) && (r + 1 < rows || b == nnz);
      // strictly ascending columns inside the row, all in range
      for(u64 k = a; ok && k < b; ++k) ok = u64(col[k]) < cols && (k == a || col[k] > col[k - 1]);
      if(!ok) atomicOr(bad, 1);
    }
  }
  // Entries the copy of an inner-precision level matrix leaves out: |a_ij| <= eps2 |a_ii| with
  // eps2 the square of the inner precision's epsilon (FP32: 1.4e-14; FP64: 0, so only exact zeros
  // go).  Such a coupling changes a row sum by less than eps^2 of its diagonal term, far below the
  // rounding of the inner arithmetic itself (binary16 storage flushes such values to zero anyway).
  // The 3D Q1 Laplacian's 6 face couplings per row are roundoff residues of about 2e-17 of the
  // diagonal, so its FP32 band has 21 diagonals instead of 27 (22% fewer bytes per sweep).  The
  // test is written so that a NaN value, or any value in a row whose diagonal is NaN, is kept: a
  // non-finite matrix stays non-finite in every copy, and the solve reports it (aborted) instead
  // of solving another matrix.  Host and device builders evaluate the same expression on the same
  // FP64 values.
  __host__ __device__ inline bool band_keeps(double v, double diag, double eps2)
  {
    return !((v < 0.0 ? -v : v) <= eps2 * (diag < 0.0 ? -diag : diag));
  }
  __device__ inline double row_diagonal(const u32* rp, const u32* col, const double* val, u64 r)
  {
    for(u64 k = rp[r]; k < rp[r + 1]; ++k)
      if(u64(col[k]) == r) return val[k];
    return 0.0;
  }

  // entries a thinned copy keeps: band_keeps(v, diagonal of the row, eps2) -- eps2 = 0 keeps the
  // nonzeros (the transfers' compaction), eps2 = eps_inner^2 the couplings the inner arithmetic
  // can resolve (inner-precision level matrices)
  __global__ void k_count_kept(u64 rows, const u32* __restrict__ rp, const u32* __restrict__ col,
                               const double* __restrict__ val, double eps2, u32* __restrict__ cnt)
  {
    for(u64 r = u64(blockIdx.x) * blockDim.x + threadIdx.x; r < rows; r += u64(gridDim.x) * blockDim.x)
    {
      const double d = row_diagonal(rp, col, val, r);
      u32 c = 0;
      for(u64 k = rp[r]; k < rp[r + 1]; ++k) c += band_keeps(val[k], d, eps2) ? 1u : 0u;
      cnt[r] = c;
    }
  }
  __global__ void k_compact(u64 rows, const u32* __restrict__ rp, const u32* __restrict__ col, const double* __restrict__ val,
                            double eps2, const u32* __restrict__ nrp, u32* __restrict__ ncol, double* __restrict__ nval)
  {
    for(u64 r = u64(blockIdx.x) * blockDim.x + threadIdx.x; r < rows; r += u64(gridDim.x) * blockDim.x)
    {
      const double d = row_diagonal(rp, col, val, r);
      u64 j = nrp[r];
      for(u64 k = rp[r]; k < rp[r + 1]; ++k)
        if(band_keeps(val[k], d, eps2)) { ncol[j] = col[k]; nval[j] = val[k]; ++j; }
    }
  }
  template<typename VT>
  __global__ void k_to_storage(u64 n, const double* __restrict__ in, VT* __restrict__ out)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      out[i] = VT(float(in[i]));
  }
  template<>
  __global__ void k_to_storage<double>(u64 n, const double* __restrict__ in, double* __restrict__ out)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      out[i] = in[i];
  }
  // bit (col - row + rows - 1) for every kept entry: the set of diagonals.  Banded matrices set
  // the same few bits from every row, so a bit is tested with a plain load first and the atomic
  // issued only while it reads clear (a repeated OR is harmless; unconditional atomics on ~27
  // words serialised 57 M entries: 202 ms for 3D Q1 L7)
  __global__ void k_diag_bitmap(u64 rows, const u32* __restrict__ rp, const u32* __restrict__ col,
                                const double* __restrict__ val, double eps2, u32* bits)
  {
    for(u64 r = u64(blockIdx.x) * blockDim.x + threadIdx.x; r < rows; r += u64(gridDim.x) * blockDim.x)
    {
      const double d = row_diagonal(rp, col, val, r);
      for(u64 k = rp[r]; k < rp[r + 1]; ++k)
      {
        if(!band_keeps(val[k], d, eps2)) continue;
        const u64 o = u64(col[k]) + rows - 1 - r;
        const u32 m = 1u << (o & 31u);
        if((__atomic_load_n(&bits[o >> 5], __ATOMIC_RELAXED) & m) == 0u) atomicOr(&bits[o >> 5], m);
      }
    }
  }
  // chunk-major banded values from CSR (offsets stored as off + rows - 1, ascending); the left
  // out entries are skipped (their diagonal need not be stored; the band is zero-filled)
  template<typename VT> __device__ inline VT band_value(double v) { return VT(float(v)); }
  template<> __device__ inline double band_value<double>(double v) { return v; }
  template<typename VT>
  __global__ void k_banded_scatter(u64 rows, const u32* __restrict__ rp, const u32* __restrict__ col,
                                   const double* __restrict__ val, double eps2, const u32* __restrict__ boff, u32 nd,
                                   VT* __restrict__ bval)
  {
    for(u64 r = u64(blockIdx.x) * blockDim.x + threadIdx.x; r < rows; r += u64(gridDim.x) * blockDim.x)
    {
      const double d = row_diagonal(rp, col, val, r);
      for(u64 k = rp[r]; k < rp[r + 1]; ++k)
      {
        if(!band_keeps(val[k], d, eps2)) continue;
        const u32 key = u32(u64(col[k]) + rows - 1 - r);
        u32 lo = 0, hi = nd;
        while(lo < hi) { const u32 mid = (lo + hi) / 2; if(boff[mid] < key) lo = mid + 1; else hi = mid; }
        bval[(r >> 6) * u64(nd) * 64 + u64(lo) * 64 + (r & 63)] = band_value<VT>(val[k]);
      }
    }
  }
  // bit patterns for the exact comparison of band values
  template<typename VT> struct Bits;
  template<> struct Bits<double> { typedef std::uint64_t T; };
  template<> struct Bits<float> { typedef std::uint32_t T; };
  template<> struct Bits<half> { typedef std::uint16_t T; };

  // any lower-diagonal value of the chunk-major band that differs (bit for bit) from its mirror
  // image, the upper value stored at row c = r + off (off < 0), raises *bad
  template<typename VT>
  __global__ void k_band_sym_check(u64 padded, u32 nd, const VT* __restrict__ bval, const std::int64_t* __restrict__ offs,
                                   const u32* __restrict__ mirror, int* bad)
  {
    typedef typename Bits<VT>::T B;
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < padded * u64(nd); i += u64(gridDim.x) * blockDim.x)
    {
      const u64 r = i / nd;
      const u32 d = u32(i % nd);
      const std::int64_t o = offs[d];
      if(o >= 0 || std::int64_t(r) + o < 0) continue;
      const u64 c = u64(std::int64_t(r) + o);
      const B a = reinterpret_cast<const B*>(bval)[(r >> 6) * nd * 64 + u64(d) * 64 + (r & 63)];
      const B b = reinterpret_cast<const B*>(bval)[(c >> 6) * nd * 64 + u64(mirror[d]) * 64 + (c & 63)];
      if(a != b) atomicOr(bad, 1);
    }
  }
  // the center and upper diagonals (pos[k], k < ns) of the band into the symmetric storage
  template<typename VT>
  __global__ void k_band_sym_pack(u64 padded, u32 nd, const VT* __restrict__ bval, u32 ns, const u32* __restrict__ pos,
                                  VT* __restrict__ sval)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < padded * u64(ns); i += u64(gridDim.x) * blockDim.x)
    {
      const u64 r = i / ns;
      const u32 k = u32(i % ns);
      sval[(r >> 6) * ns * 64 + u64(k) * 64 + (r & 63)] = bval[(r >> 6) * nd * 64 + u64(pos[k]) * 64 + (r & 63)];
    }
  }

  // ---- value-coded band (mg::DictBandRows) ------------------------------------------------
  // The distinct values of each diagonal of a chunk-major band as bit patterns (zero-extended to
  // 64 bits): kDictSlots slots per diagonal, kDictEmpty where unused.  Each block collects into
  // its own LDS tables, then merges them into the global ones; a diagonal with more distinct
  // values than slots raises *over.  Slot order is that of first insertion (it does not change
  // any value the row sums read).
  constexpr unsigned long long kDictEmpty = ~0ull;
  __device__ inline bool dict_insert(unsigned long long* slots, unsigned long long key)
  {
    for(u32 j = 0; j < mg::kDictSlots; ++j)
    {
      const unsigned long long cur = __atomic_load_n(&slots[j], __ATOMIC_RELAXED);
      if(cur == key) return true;
      if(cur == kDictEmpty)
      {
        const unsigned long long prev = atomicCAS(&slots[j], kDictEmpty, key);
        if(prev == kDictEmpty || prev == key) return true;
      }
    }
    return false;
  }
  template<typename VT>
  __global__ void k_dict_collect(u64 rows, u32 nd, const VT* __restrict__ bval, unsigned long long* keys, int* over)
  {
    typedef typename Bits<VT>::T B;
    __shared__ unsigned long long lk[mg::kDictDiag * mg::kDictSlots];
    for(u32 i = threadIdx.x; i < nd * mg::kDictSlots; i += blockDim.x) lk[i] = kDictEmpty;
    __syncthreads();
    const u64 padded = (rows + 63) / 64 * 64, total = padded * nd;
    bool ok = true;
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += u64(gridDim.x) * blockDim.x)
    {
      const u64 r = (i / (u64(nd) * 64)) * 64 + (i & 63);
      if(r >= rows) continue;
      const u32 d = u32((i >> 6) % nd);
      ok = dict_insert(&lk[d * mg::kDictSlots], (unsigned long long)(reinterpret_cast<const B*>(bval)[i])) && ok;
    }
    __syncthreads();
    for(u32 t = threadIdx.x; t < nd * mg::kDictSlots; t += blockDim.x)
      if(lk[t] != kDictEmpty) ok = dict_insert(&keys[(t / mg::kDictSlots) * mg::kDictSlots], lk[t]) && ok;
    if(!ok) atomicOr(over, 1);
  }
  // the code words of every row: code k of diagonal d is the slot holding the row's value; a
  // value without a slot raises *over; rows past the matrix (chunk padding) get code 0
  template<typename VT>
  __global__ void k_dict_encode(u64 rows, u32 nd, u32 nw, const VT* __restrict__ bval, const unsigned long long* __restrict__ keys,
                                u32* __restrict__ idx, int* over)
  {
    typedef typename Bits<VT>::T B;
    __shared__ unsigned long long lk[mg::kDictDiag * mg::kDictSlots];
    for(u32 i = threadIdx.x; i < nd * mg::kDictSlots; i += blockDim.x) lk[i] = keys[i];
    __syncthreads();
    const u64 padded = (rows + 63) / 64 * 64;
    bool ok = true;
    for(u64 r = u64(blockIdx.x) * blockDim.x + threadIdx.x; r < padded; r += u64(gridDim.x) * blockDim.x)
      for(u32 w = 0; w < nw; ++w)
      {
        u32 code = 0;
        for(u32 j = 0; j < 8u && w * 8u + j < nd && r < rows; ++j)
        {
          const u32 d = w * 8u + j;
          const unsigned long long key = (unsigned long long)(reinterpret_cast<const B*>(bval)[(r >> 6) * u64(nd) * 64 + u64(d) * 64 + (r & 63)]);
          u32 k = mg::kDictSlots;
          for(u32 q = 0; q < mg::kDictSlots; ++q)
            if(lk[d * mg::kDictSlots + q] == key) { k = q; break; }
          if(k == mg::kDictSlots || key == kDictEmpty) { ok = false; k = 0; }
          code |= k << (4u * j);
        }
        idx[(r >> 6) * u64(nw) * 64 + u64(w) * 64 + (r & 63)] = code;
      }
    if(!ok) atomicOr(over, 1);
  }

  inline unsigned setup_blocks(u64 n)
  {
    const u64 b = (n + 255) / 256;
    return unsigned(b < 1 ? 1 : (b > 16384 ? 16384 : b));
  }

  // ---- tensor-product transfers (mg::TensorRows) -------------------------------------------
  // the entries of the listed rows in the arithmetic type, kTensorW per row: cnt[i] the row's
  // length (kTensorW + 1 for a longer row, whose entries are not copied)
  template<typename DT, typename VT>
  __global__ void k_tensor_gather(u32 nlist, const u32* __restrict__ list, const u32* __restrict__ rp,
                                  const u32* __restrict__ col, const VT* __restrict__ val, u32* __restrict__ cnt,
                                  u32* __restrict__ ocol, DT* __restrict__ oval)
  {
    constexpr u32 W = mg::kTensorW;
    for(u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < nlist; i += gridDim.x * blockDim.x)
    {
      const u32 a = rp[list[i]], n = rp[list[i] + 1] - a;
      cnt[i] = n > W ? W + 1 : n;
      for(u32 k = 0; k < n && k < W; ++k)
      {
        ocol[i * W + k] = col[a + k];
        oval[i * W + k] = dev::ld<DT, VT>(val, a + k);
      }
    }
  }
  // every CSR row equals its tensor-product row: the same number of entries, the same columns in
  // the same order and each value, converted to the arithmetic type, bit for bit the product
  // (wz * wy) * wx that TensorRows::row_sum forms; any difference raises *bad
  template<typename DT, typename VT>
  __global__ void k_tensor_check(mg::TensorRows<DT> T, const u32* __restrict__ rp, const u32* __restrict__ col,
                                 const VT* __restrict__ val, int* bad)
  {
#pragma clang fp contract(off)
    constexpr u32 W = mg::kTensorW;
    typedef typename Bits<DT>::T B;
    for(u64 p = u64(blockIdx.x) * blockDim.x + threadIdx.x; p < T.rows; p += u64(gridDim.x) * blockDim.x)
    {
      const u32 q = u32(p);
      const u32 ox = q % T.mo[0], t = q / T.mo[0];
      const u32 oy = t % T.mo[1], oz = t / T.mo[1];
      const u32 nz = T.ax[2].cnt[oz], ny = T.ax[1].cnt[oy], nx = T.ax[0].cnt[ox];
      u64 k = rp[p];
      const u64 end = rp[p + 1];
      bool ok = u64(nz) * ny * nx == end - k;
      for(u32 a = 0; ok && a < nz; ++a)
      {
        const u32 iz = T.ax[2].idx[oz * W + a];
        const DT wz = T.ax[2].w[oz * W + a];
        for(u32 b = 0; ok && b < ny; ++b)
        {
          const DT wzy = wz * T.ax[1].w[oy * W + b];
          const u64 base = u64(iz) + u64(T.ax[1].idx[oy * W + b]);
          for(u32 c = 0; ok && c < nx; ++c, ++k)
          {
            const DT v = wzy * T.ax[0].w[ox * W + c];
            ok = u64(col[k]) == base + T.ax[0].idx[ox * W + c] &&
                 __builtin_bit_cast(B, dev::ld<DT, VT>(val, k)) == __builtin_bit_cast(B, v);
          }
        }
      }
      if(!ok) atomicOr(bad, 1);
    }
  }

  struct HostLevel
  {
    HostCsr A, P, R;
    std::vector<u32> fidx;
  };

  // device CSR with an optional SELL-C-sigma copy (derived; the CSR arrays stay as given)
  template<typename VT>
  struct DevCsr
  {
    u64 rows = 0, cols = 0, nnz = 0;
    int lanes = 8;
    u64 rows_per_block = 0;   // row positions per block of the row kernels (0: by row_grid)
    DevMem rp, col, val;
    bool sell = false, sell_perm = false, sell_d16 = false;
    u64 C = 64, sigma = 1, sell_size = 0;
    DevMem cs, cl, perm, scol, sval;
    bool banded = false;
    u32 ndiag = 0;
    DevMem boff, bval;
    // symmetric band storage (mg::BandedRows<.., true>): nds stored diagonals, bsx[d] the stored
    // diagonal holding diagonal d; bval then holds nds diagonals
    bool bsym = false;
    u32 nds = 0;
    DevMem bsx;
    bool allow_sym = true;    // symmetric storage allowed (option "symmetric_band")
    // value-coded band (mg::DictBandRows, option "band_dict"): dnw code words per row in didx,
    // the tables in dtab, the offsets padded to dnw * 8 diagonals in doff, the smallest and
    // largest offset col - row in domin / domax; the band's values are then not kept
    bool allow_dict = true;
    bool dict = false;
    bool dtrip = false;              // 27 offsets in nine runs of three (DictBandRows::trip)
    bool allow_lane_shift = true;    // take those runs' outer x from the neighbouring lanes
    u32 dnw = 0;
    int domin = 0, domax = 0;
    DevMem didx, dtab, doff;
    // processing order of the row ranges (banded matrices): tile[logical block] = the row range
    // it evaluates, for the grid of tile_blocks blocks (0: ascending order)
    DevMem tile;
    unsigned tile_blocks = 0;
    // x staged in LDS by the banded row kernels (mg::BandStage, option band_stage_x): set by
    // build_stage when the windows of a round fit
    bool stage = false;
    mg::BandStage stg{};
    DevMem stg_xoff;
    // tensor-product evaluation (mg::TensorRows, transfers of structured hierarchies; set by
    // build_tensor_pair after every stored entry was checked): output and input lattice extents
    // per axis, per axis the factor tables cnt / idx / w (w in the arithmetic type, tw_bytes wide)
    bool tensor = false;
    u32 tmo[3] = {1, 1, 1}, tmi[3] = {1, 1, 1};
    u32 tw_bytes = 0;
    u32 tk = 0;   // the largest number of entries of a factor row
    DevMem tcnt[3], tidx[3], tw[3];

    // thin >= 0: upload a compact copy with the entries band_keeps(v, row diagonal, thin) only:
    // thin = 0 drops the stored exact zeros (the transfer matrices, whose two-level pattern holds
    // structural zeros; a zero adds nothing to a finite product), thin = eps_inner^2 also the
    // couplings below eps^2 of the diagonal that the inner arithmetic cannot resolve (inner-precision
    // level matrices); thin < 0 keeps every entry
    int upload(const HostCsr& h0, hipStream_t s, bool want_sell, u64 sell_C, u64 sell_sigma, int force_lanes,
               double thin = -1.0, bool allow_d16 = true, bool want_banded = false)
    {
      HostCsr compact;
      const HostCsr* hp = &h0;
      if(thin >= 0.0)
      {
        compact.rows = h0.rows; compact.cols = h0.cols;
        compact.rp.assign(h0.rows + 1, 0);
        for(u64 r = 0; r < h0.rows; ++r)
        {
          double d = 0.0;
          for(u32 k = h0.rp[r]; k < h0.rp[r + 1]; ++k)
            if(u64(h0.col[k]) == r) { d = h0.val[k]; break; }
          for(u32 k = h0.rp[r]; k < h0.rp[r + 1]; ++k)
            if(band_keeps(h0.val[k], d, thin)) { compact.col.push_back(h0.col[k]); compact.val.push_back(h0.val[k]); }
          compact.rp[r + 1] = u32(compact.col.size());
        }
        hp = &compact;
      }
      const HostCsr& h = *hp;
      rows = h.rows; cols = h.cols; nnz = h.nnz();
      F3_RC(rp.alloc(sizeof(u32) * (rows + 1)));
      F3_RC(col.alloc(sizeof(u32) * std::max<u64>(nnz, 1)));
      F3_RC(val.alloc(sizeof(VT) * std::max<u64>(nnz, 1)));
      std::vector<VT> v(nnz);
      for(u64 i = 0; i < nnz; ++i) v[i] = to_storage<VT>(h.val[i]);
      F3_CHECK(hipMemcpyAsync(rp.p, h.rp.data(), sizeof(u32) * (rows + 1), hipMemcpyHostToDevice, s));
      if(nnz)
      {
        F3_CHECK(hipMemcpyAsync(col.p, h.col.data(), sizeof(u32) * nnz, hipMemcpyHostToDevice, s));
        F3_CHECK(hipMemcpyAsync(val.p, v.data(), sizeof(VT) * nnz, hipMemcpyHostToDevice, s));
      }
      F3_CHECK(hipStreamSynchronize(s)); // host staging vectors go out of scope
      // lanes = 0: the streaming evaluator (CsrStreamRows); a forced lane count (tuner) selects
      // the G-lanes-per-row evaluator instead
      lanes = force_lanes > 0 ? force_lanes : 0;
      if(want_banded) F3_RC(build_banded(h, s));
      sell = want_sell && !banded;
      if(sell) F3_RC(build_sell(h, s, sell_C, sell_sigma, allow_d16));
      return rt::kOk;
    }

    // The same from a device-resident source (validated on the device: rows ascending, columns
    // strictly ascending and in range).  The banded copy and the compacted transfers are built
    // on the device; only the diagonal bitmap (2 rows bits) and the per-row entry counts of a
    // compaction come back to the host.  A matrix that needs the SELL-C-sigma copy is brought
    // to the host and takes upload().
    int upload_dev(const HostCsr& h0, hipStream_t s, bool want_sell, u64 sell_C, u64 sell_sigma, int force_lanes,
                   double thin, bool allow_d16, bool want_banded)
    {
      rows = h0.rows; cols = h0.cols;
      u64 n_in = h0.dnnz;
      DevMem bad;
      F3_RC(bad.alloc(sizeof(int)));
      F3_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
      if(rows) k_validate_csr<<<setup_blocks(rows), 256, 0, s>>>(rows, cols, n_in, h0.drp, h0.dcol, bad.template as<int>());
      int hbad = 0;
      F3_CHECK(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
      F3_CHECK(hipStreamSynchronize(s));
      if(hbad) return rt::kErrArg;
      // compaction (the thin rule of upload) into FP64 device arrays
      DevMem crp, ccol, cval;
      const u32* srp = h0.drp;
      const u32* scol_ = h0.dcol;
      const double* sval_ = h0.dval;
      if(thin >= 0.0 && rows)
      {
        DevMem cnt;
        F3_RC(cnt.alloc(sizeof(u32) * rows));
        k_count_kept<<<setup_blocks(rows), 256, 0, s>>>(rows, h0.drp, h0.dcol, h0.dval, thin, cnt.template as<u32>());
        std::vector<u32> hc(rows), hrp(rows + 1, 0);
        F3_CHECK(hipMemcpyAsync(hc.data(), cnt.p, sizeof(u32) * rows, hipMemcpyDeviceToHost, s));
        F3_CHECK(hipStreamSynchronize(s));
        for(u64 r = 0; r < rows; ++r) hrp[r + 1] = hrp[r] + hc[r];
        n_in = hrp[rows];
        F3_RC(crp.alloc(sizeof(u32) * (rows + 1)));
        F3_RC(ccol.alloc(sizeof(u32) * std::max<u64>(n_in, 1)));
        F3_RC(cval.alloc(sizeof(double) * std::max<u64>(n_in, 1)));
        F3_CHECK(hipMemcpyAsync(crp.p, hrp.data(), sizeof(u32) * (rows + 1), hipMemcpyHostToDevice, s));
        k_compact<<<setup_blocks(rows), 256, 0, s>>>(rows, h0.drp, h0.dcol, h0.dval, thin, crp.template as<u32>(),
                                                     ccol.template as<u32>(), cval.template as<double>());
        F3_CHECK(hipStreamSynchronize(s));   // hrp leaves scope
        srp = crp.template as<u32>();
        scol_ = ccol.template as<u32>();
        sval_ = cval.template as<double>();
      }
      nnz = n_in;
      F3_RC(rp.alloc(sizeof(u32) * (rows + 1)));
      F3_RC(col.alloc(sizeof(u32) * std::max<u64>(nnz, 1)));
      F3_RC(val.alloc(sizeof(VT) * std::max<u64>(nnz, 1)));
      F3_CHECK(hipMemcpyAsync(rp.p, srp, sizeof(u32) * (rows + 1), hipMemcpyDeviceToDevice, s));
      if(nnz)
      {
        F3_CHECK(hipMemcpyAsync(col.p, scol_, sizeof(u32) * nnz, hipMemcpyDeviceToDevice, s));
        k_to_storage<VT><<<setup_blocks(nnz), 256, 0, s>>>(nnz, sval_, val.template as<VT>());
        F3_RC(rt::check_launch());
      }
      lanes = force_lanes > 0 ? force_lanes : 0;
      banded = false;
      if(want_banded) F3_RC(build_banded_dev(srp, scol_, sval_, s));
      sell = want_sell && !banded;
      if(sell)
      {
        HostCsr h;
        h.rows = rows; h.cols = cols;
        h.rp.resize(rows + 1); h.col.resize(nnz); h.val.resize(nnz);
        F3_CHECK(hipMemcpyAsync(h.rp.data(), srp, sizeof(u32) * (rows + 1), hipMemcpyDeviceToHost, s));
        if(nnz)
        {
          F3_CHECK(hipMemcpyAsync(h.col.data(), scol_, sizeof(u32) * nnz, hipMemcpyDeviceToHost, s));
          F3_CHECK(hipMemcpyAsync(h.val.data(), sval_, sizeof(double) * nnz, hipMemcpyDeviceToHost, s));
        }
        F3_CHECK(hipStreamSynchronize(s));
        F3_RC(build_sell(h, s, sell_C, sell_sigma, allow_d16));
      }
      F3_CHECK(hipStreamSynchronize(s));   // the compaction buffers leave scope
      return rt::kOk;
    }

    // banded copy from device CSR arrays: the diagonal set from a device bitmap, the same
    // acceptance rule as build_banded, values scattered on the device
    int build_banded_dev(const u32* drp, const u32* dcol, const double* dval, hipStream_t s)
    {
      banded = false;
      // BandedRows indexes in 32 bits: at most kMaxDiag (64) diagonals of fewer than 2^26 rows
      if(rows != cols || rows == 0 || rows >= (u64(1) << 26)) return rt::kOk;
      const u64 nbits = 2 * rows - 1, nw = (nbits + 31) / 32;
      DevMem bits;
      F3_RC(bits.alloc(sizeof(u32) * nw));
      F3_CHECK(hipMemsetAsync(bits.p, 0, sizeof(u32) * nw, s));
      k_diag_bitmap<<<setup_blocks(rows), 256, 0, s>>>(rows, drp, dcol, dval, band_eps2(), bits.template as<u32>());
      std::vector<u32> hb(nw);
      F3_CHECK(hipMemcpyAsync(hb.data(), bits.p, sizeof(u32) * nw, hipMemcpyDeviceToHost, s));
      F3_CHECK(hipStreamSynchronize(s));
      std::vector<u32> ho;
      for(u64 w = 0; w < nw; ++w)
        for(u32 b = hb[w]; b; b &= b - 1)
        {
          if(ho.size() >= kMaxDiag) return rt::kOk;
          ho.push_back(u32(w * 32 + u64(__builtin_ctz(b))));
        }
      const bool dense = double(ho.size()) * double(rows) <= 1.25 * double(std::max<u64>(nnz, 1));
      if(!dense && !codable_fill(ho.size())) return rt::kOk;
      const u64 nd = ho.size(), padded = (rows + 63) / 64 * 64;
      F3_RC(boff.alloc(sizeof(u32) * nd));
      F3_RC(bval.alloc(sizeof(VT) * nd * padded));
      F3_CHECK(hipMemcpyAsync(boff.p, ho.data(), sizeof(u32) * nd, hipMemcpyHostToDevice, s));
      F3_CHECK(hipMemsetAsync(bval.p, 0, sizeof(VT) * nd * padded, s));
      k_banded_scatter<VT><<<setup_blocks(rows), 256, 0, s>>>(rows, drp, dcol, dval, band_eps2(), boff.template as<u32>(),
                                                             u32(nd), bval.template as<VT>());
      F3_RC(rt::check_launch());
      F3_CHECK(hipStreamSynchronize(s));   // ho leaves scope
      ndiag = u32(nd);
      banded = true;
      bsym = false;
      F3_RC(build_dict(s));
      if(dict) return rt::kOk;
      if(!dense) { drop_band(); return rt::kOk; }
      // symmetric storage when the offsets are symmetric and every value equals its mirror image
      // (FP64 values, as build_banded)
      if(sizeof(VT) != 8 || !allow_sym) return rt::kOk;
      std::vector<std::int64_t> offs(nd);
      for(u64 d = 0; d < nd; ++d) offs[d] = std::int64_t(ho[d]) - std::int64_t(rows) + 1;
      std::vector<u32> mirror(nd), pos, sx(nd);
      for(u64 d = 0; d < nd; ++d)
      {
        auto it = std::lower_bound(offs.begin(), offs.end(), -offs[d]);
        if(it == offs.end() || *it != -offs[d]) return rt::kOk;
        mirror[d] = u32(it - offs.begin());
        if(offs[d] >= 0) pos.push_back(u32(d));
      }
      for(u64 k = 0; k < pos.size(); ++k) sx[pos[k]] = u32(k);
      for(u64 d = 0; d < nd; ++d)
        if(offs[d] < 0) sx[d] = sx[mirror[d]];
      DevMem doffs, dmirror, dbad;
      F3_RC(doffs.alloc(sizeof(std::int64_t) * nd));
      F3_RC(dmirror.alloc(sizeof(u32) * nd));
      F3_RC(dbad.alloc(sizeof(int)));
      F3_CHECK(hipMemcpyAsync(doffs.p, offs.data(), sizeof(std::int64_t) * nd, hipMemcpyHostToDevice, s));
      F3_CHECK(hipMemcpyAsync(dmirror.p, mirror.data(), sizeof(u32) * nd, hipMemcpyHostToDevice, s));
      F3_CHECK(hipMemsetAsync(dbad.p, 0, sizeof(int), s));
      k_band_sym_check<VT><<<setup_blocks(padded * nd), 256, 0, s>>>(padded, u32(nd), bval.template as<VT>(),
                                                                    doffs.template as<std::int64_t>(),
                                                                    dmirror.template as<u32>(), dbad.template as<int>());
      F3_RC(rt::check_launch());
      int hbad = 0;
      F3_CHECK(hipMemcpyAsync(&hbad, dbad.p, sizeof(int), hipMemcpyDeviceToHost, s));
      F3_CHECK(hipStreamSynchronize(s));
      if(hbad) return rt::kOk;
      const u32 ns = u32(pos.size());
      DevMem dpos, sv;
      F3_RC(dpos.alloc(sizeof(u32) * ns));
      F3_RC(sv.alloc(sizeof(VT) * u64(ns) * padded));
      F3_RC(bsx.alloc(sizeof(u32) * nd));
      F3_CHECK(hipMemcpyAsync(dpos.p, pos.data(), sizeof(u32) * ns, hipMemcpyHostToDevice, s));
      F3_CHECK(hipMemcpyAsync(bsx.p, sx.data(), sizeof(u32) * nd, hipMemcpyHostToDevice, s));
      k_band_sym_pack<VT><<<setup_blocks(padded * ns), 256, 0, s>>>(padded, u32(nd), bval.template as<VT>(), ns,
                                                                   dpos.template as<u32>(), sv.template as<VT>());
      F3_RC(rt::check_launch());
      F3_CHECK(hipStreamSynchronize(s));   // host vectors leave scope
      std::swap(bval.p, sv.p);             // sv frees the full band
      nds = ns;
      bsym = true;
      return rt::kOk;
    }

    // Banded storage (the diagonals of SparseMatrixBanded, chunk-major) when the matrix has few distinct diagonals
    // (col - row) and the zero fill stays small: at most kMaxDiag diagonals and
    // ndiag * rows <= 1.25 * nnz.  Otherwise the matrix keeps its other format.
    static constexpr u32 kMaxDiag = 64;
    // eps^2 of the inner arithmetic (binary16 storage computes in FP32); FP64 bands keep every
    // nonzero (the outer operator and the FP64 path are the matrix as given)
    static double band_eps2()
    {
      const double e = 1.1920928955078125e-07;
      return sizeof(VT) == 8 ? 0.0 : e * e;
    }
    // A value-coded band stores 4 bits per slot, so it may carry more zero fill than a band of
    // values: up to 4 slots per stored entry (2D Q2: 25 diagonals at 57% fill, 13 bytes of codes
    // per row against 16 entries of 6 bytes in SELL-C-sigma); the band is only kept if the codes
    // are built
    bool codable_fill(u64 nd) const
    {
      return allow_dict && nd <= mg::kDictDiag && double(nd) * double(rows) <= 4.0 * double(std::max<u64>(nnz, 1));
    }
    void drop_band()
    {
      banded = false;
      ndiag = 0;
      DevMem a, b;
      std::swap(a.p, bval.p);
      std::swap(b.p, boff.p);
    }
    int build_banded(const HostCsr& h, hipStream_t s)
    {
      banded = false;
      // BandedRows indexes in 32 bits: at most kMaxDiag (64) diagonals of fewer than 2^26 rows
      if(rows != cols || rows == 0 || rows >= (u64(1) << 26)) return rt::kOk;
      const double eps2 = band_eps2();
      std::vector<double> dg(rows, 0.0);
      for(u64 r = 0; r < rows; ++r)
        for(u32 k = h.rp[r]; k < h.rp[r + 1]; ++k)
          if(u64(h.col[k]) == r) { dg[r] = h.val[k]; break; }
      std::vector<std::int64_t> offs;
      auto find = [&](std::int64_t o) { return std::lower_bound(offs.begin(), offs.end(), o); };
      for(u64 r = 0; r < rows; ++r)
        for(u32 k = h.rp[r]; k < h.rp[r + 1]; ++k)
        {
          if(!band_keeps(h.val[k], dg[r], eps2)) continue;
          const std::int64_t o = std::int64_t(h.col[k]) - std::int64_t(r);
          auto it = find(o);
          if(it == offs.end() || *it != o)
          {
            if(offs.size() >= kMaxDiag) return rt::kOk;
            offs.insert(it, o);
          }
        }
      const bool dense = double(offs.size()) * double(rows) <= 1.25 * double(std::max<u64>(nnz, 1));
      if(!dense && !codable_fill(offs.size())) return rt::kOk;
      const u64 nd = offs.size();
      // chunk-major: the nd diagonals of 64 consecutive rows are contiguous, so that a wave
      // reads one contiguous region (val[(p / 64) * nd * 64 + d * 64 + p % 64])
      const u64 padded = (rows + 63) / 64 * 64;
      std::vector<VT> hv(nd * padded, to_storage<VT>(0.0));
      std::vector<char> used(nd * padded, 0);
      for(u64 r = 0; r < rows; ++r)
        for(u32 k = h.rp[r]; k < h.rp[r + 1]; ++k)
        {
          if(!band_keeps(h.val[k], dg[r], eps2)) continue;
          const u64 d = u64(find(std::int64_t(h.col[k]) - std::int64_t(r)) - offs.begin());
          const u64 at = (r / 64) * nd * 64 + d * 64 + (r % 64);
          if(used[at]) return rt::kOk;   // duplicate entry: keep the general format
          used[at] = 1;
          hv[at] = to_storage<VT>(h.val[k]);
        }
      std::vector<u32> ho(nd);
      for(u64 d = 0; d < nd; ++d) ho[d] = u32(offs[d] + std::int64_t(rows) - 1);
      // the value-coded copy of the full band where it applies
      if(allow_dict)
      {
        F3_RC(boff.alloc(sizeof(u32) * nd));
        F3_RC(bval.alloc(sizeof(VT) * hv.size()));
        F3_CHECK(hipMemcpyAsync(boff.p, ho.data(), sizeof(u32) * nd, hipMemcpyHostToDevice, s));
        F3_CHECK(hipMemcpyAsync(bval.p, hv.data(), sizeof(VT) * hv.size(), hipMemcpyHostToDevice, s));
        F3_CHECK(hipStreamSynchronize(s));
        ndiag = u32(nd);
        bsym = false;
        nds = 0;
        banded = true;
        F3_RC(build_dict(s));
        if(dict) return rt::kOk;
      }
      if(!dense) { drop_band(); return rt::kOk; }
      // symmetric storage when the band equals its mirror image bit for bit, for FP64 values
      // (3D Q1 L7: FP64 sweep 107 -> 85 us, outer PCG product likewise).  FP32 and binary16 bands
      // stay full: their values are served partly from the Infinity Cache, and the mirrored
      // reads cost about what the halved stream saves (FP32 sweep 32.5 -> 32.3 us; binary16
      // 23.5 -> 32.0 us)
      std::vector<u32> sx;
      const std::vector<VT> sv = sizeof(VT) == 8 && allow_sym ? symmetric_band(hv, offs, padded, sx) : std::vector<VT>();
      const bool sym = !sx.empty();
      const std::vector<VT>& up = sym ? sv : hv;
      F3_RC(boff.alloc(sizeof(u32) * nd));
      F3_RC(bval.alloc(sizeof(VT) * up.size()));
      F3_CHECK(hipMemcpyAsync(boff.p, ho.data(), sizeof(u32) * nd, hipMemcpyHostToDevice, s));
      F3_CHECK(hipMemcpyAsync(bval.p, up.data(), sizeof(VT) * up.size(), hipMemcpyHostToDevice, s));
      if(sym)
      {
        F3_RC(bsx.alloc(sizeof(u32) * nd));
        F3_CHECK(hipMemcpyAsync(bsx.p, sx.data(), sizeof(u32) * nd, hipMemcpyHostToDevice, s));
      }
      F3_CHECK(hipStreamSynchronize(s));
      ndiag = u32(nd);
      bsym = sym;
      nds = sym ? u32(up.size() / padded) : 0;
      banded = true;
      return rt::kOk;
    }

    // the band's symmetric storage (center and upper diagonals, chunk-major) and sx, or an empty
    // sx when the offsets are not symmetric or some value differs from its mirror image
    // (compared bit for bit, so that the row sums are unchanged)
    static std::vector<VT> symmetric_band(const std::vector<VT>& hv, const std::vector<std::int64_t>& offs, u64 padded,
                                          std::vector<u32>& sx)
    {
      sx.clear();
      const u64 nd = offs.size(), rows_p = padded;
      std::vector<u64> mirror(nd), pos;
      for(u64 d = 0; d < nd; ++d)
      {
        auto it = std::lower_bound(offs.begin(), offs.end(), -offs[d]);
        if(it == offs.end() || *it != -offs[d]) return {};
        mirror[d] = u64(it - offs.begin());
        if(offs[d] >= 0) pos.push_back(d);
      }
      auto at = [&](u64 r, u64 d) { return (r / 64) * nd * 64 + d * 64 + (r % 64); };
      const std::int64_t nrows = std::int64_t(rows_p);
      for(u64 d = 0; d < nd; ++d)
      {
        if(offs[d] >= 0) continue;
        for(std::int64_t r = -offs[d]; r < nrows; ++r)
        {
          const u64 c = u64(r + offs[d]);
          if(std::memcmp(&hv[at(u64(r), d)], &hv[at(c, mirror[d])], sizeof(VT)) != 0) return {};
        }
      }
      const u64 ns = pos.size();
      std::vector<u64> sidx(nd);
      for(u64 k = 0; k < ns; ++k) sidx[pos[k]] = k;
      sx.resize(nd);
      for(u64 d = 0; d < nd; ++d) sx[d] = u32(offs[d] >= 0 ? sidx[d] : sidx[mirror[d]]);
      std::vector<VT> sv(ns * padded);
      for(u64 r = 0; r < padded; ++r)
        for(u64 k = 0; k < ns; ++k) sv[(r / 64) * ns * 64 + k * 64 + (r % 64)] = hv[at(r, pos[k])];
      return sv;
    }

    // SELL-C-sigma from the host converter feat3_sell_build (no permutation is stored for
    // sigma = 1); values converted to the storage type, columns as 16-bit row offsets when
    // they all fit.
    int build_sell(const HostCsr& h, hipStream_t s, u64 Cc, u64 sig, bool allow_d16)
    {
      C = Cc; sigma = std::max<u64>(1, sig);
      sell_perm = sigma > 1;
      u64 nch = 0;
      if(feat3_sell_sizes(rows, h.rp.data(), C, sigma, &nch, &sell_size)) return rt::kErrArg;
      std::vector<u32> hcs(nch + 1), hcl(std::max<u64>(nch, 1)), hpm(std::max<u64>(rows, 1)), hcol(std::max<u64>(sell_size, 1));
      std::vector<double> hv(std::max<u64>(sell_size, 1));
      std::vector<std::int16_t> hd16(allow_d16 ? std::max<u64>(sell_size, 1) : 0);
      int d16 = 0;
      if(feat3_sell_build(rows, cols, h.rp.data(), h.col.data(), h.val.data(), C, sigma, hcs.data(), hcl.data(),
                          hpm.data(), hcol.data(), hv.data(), allow_d16 ? hd16.data() : nullptr, &d16))
        return rt::kErrArg;
      sell_d16 = d16 != 0;
      std::vector<VT> hval(sell_size);
      for(u64 i = 0; i < sell_size; ++i) hval[i] = to_storage<VT>(hv[i]);
      F3_RC(cs.alloc(sizeof(u32) * (nch + 1)));
      F3_RC(cl.alloc(sizeof(u32) * std::max<u64>(nch, 1)));
      F3_RC(sval.alloc(sizeof(VT) * std::max<u64>(sell_size, 1)));
      F3_CHECK(hipMemcpyAsync(cs.p, hcs.data(), sizeof(u32) * (nch + 1), hipMemcpyHostToDevice, s));
      if(nch) F3_CHECK(hipMemcpyAsync(cl.p, hcl.data(), sizeof(u32) * nch, hipMemcpyHostToDevice, s));
      if(sell_perm)
      {
        F3_RC(perm.alloc(sizeof(u32) * std::max<u64>(rows, 1)));
        if(rows) F3_CHECK(hipMemcpyAsync(perm.p, hpm.data(), sizeof(u32) * rows, hipMemcpyHostToDevice, s));
      }
      if(sell_d16)
      {
        F3_RC(scol.alloc(sizeof(std::int16_t) * std::max<u64>(sell_size, 1)));
        if(sell_size) F3_CHECK(hipMemcpyAsync(scol.p, hd16.data(), sizeof(std::int16_t) * sell_size, hipMemcpyHostToDevice, s));
      }
      else
      {
        F3_RC(scol.alloc(sizeof(u32) * std::max<u64>(sell_size, 1)));
        if(sell_size) F3_CHECK(hipMemcpyAsync(scol.p, hcol.data(), sizeof(u32) * sell_size, hipMemcpyHostToDevice, s));
      }
      if(sell_size) F3_CHECK(hipMemcpyAsync(sval.p, hval.data(), sizeof(VT) * sell_size, hipMemcpyHostToDevice, s));
      F3_CHECK(hipStreamSynchronize(s));
      return rt::kOk;
    }

    // The value-coded copy of the full band in bval (ndiag diagonals, not the symmetric storage)
    // when every diagonal holds at most kDictSlots distinct values and the diagonals, padded to a
    // multiple of 8, fit kDictDiag; on success the band's values are released.  The tables take
    // the values' bit patterns, so the row sums read exactly the band's values.
    int build_dict(hipStream_t s)
    {
      dict = false;
      if(!allow_dict || !banded || bsym || ndiag == 0) return rt::kOk;
      const u32 nw = (ndiag + 7) / 8;
      if(nw * 8 > mg::kDictDiag) return rt::kOk;
      const u64 padded = (rows + 63) / 64 * 64, nk = u64(ndiag) * mg::kDictSlots;
      DevMem keys, over, ix;
      F3_RC(keys.alloc(sizeof(unsigned long long) * nk));
      F3_RC(over.alloc(sizeof(int)));
      F3_CHECK(hipMemsetAsync(keys.p, 0xFF, sizeof(unsigned long long) * nk, s));
      F3_CHECK(hipMemsetAsync(over.p, 0, sizeof(int), s));
      k_dict_collect<VT><<<setup_blocks(padded * ndiag), 256, 0, s>>>(rows, ndiag, bval.template as<VT>(),
                                                                      keys.template as<unsigned long long>(), over.template as<int>());
      F3_RC(rt::check_launch());
      int hover = 0;
      F3_CHECK(hipMemcpyAsync(&hover, over.p, sizeof(int), hipMemcpyDeviceToHost, s));
      F3_CHECK(hipStreamSynchronize(s));
      if(hover) return rt::kOk;
      F3_RC(ix.alloc(sizeof(u32) * padded * nw));
      k_dict_encode<VT><<<setup_blocks(padded), 256, 0, s>>>(rows, ndiag, nw, bval.template as<VT>(),
                                                             keys.template as<unsigned long long>(), ix.template as<u32>(),
                                                             over.template as<int>());
      F3_RC(rt::check_launch());
      std::vector<unsigned long long> hk(nk);
      std::vector<u32> ho(ndiag);
      F3_CHECK(hipMemcpyAsync(&hover, over.p, sizeof(int), hipMemcpyDeviceToHost, s));
      F3_CHECK(hipMemcpyAsync(hk.data(), keys.p, sizeof(unsigned long long) * nk, hipMemcpyDeviceToHost, s));
      F3_CHECK(hipMemcpyAsync(ho.data(), boff.p, sizeof(u32) * ndiag, hipMemcpyDeviceToHost, s));
      F3_CHECK(hipStreamSynchronize(s));
      if(hover) return rt::kOk;
      // tables (unused slots and the slots of the diagonals past ndiag hold 0)
      std::vector<VT> tab(u64(nw) * 8 * mg::kDictSlots, to_storage<VT>(0.0));
      for(u64 i = 0; i < nk; ++i)
        if(hk[i] != kDictEmpty)
        {
          const typename Bits<VT>::T b = typename Bits<VT>::T(hk[i]);
          std::memcpy(&tab[i], &b, sizeof(VT));
        }
      // the offsets padded with stand-ins whose column p + cols is outside the matrix for every row
      std::vector<u32> po(u64(nw) * 8, u32(rows - 1 + cols));
      std::copy(ho.begin(), ho.end(), po.begin());
      F3_RC(dtab.alloc(sizeof(VT) * tab.size()));
      F3_RC(doff.alloc(sizeof(u32) * po.size()));
      F3_CHECK(hipMemcpyAsync(dtab.p, tab.data(), sizeof(VT) * tab.size(), hipMemcpyHostToDevice, s));
      F3_CHECK(hipMemcpyAsync(doff.p, po.data(), sizeof(u32) * po.size(), hipMemcpyHostToDevice, s));
      F3_CHECK(hipStreamSynchronize(s));   // tab and po leave scope
      dtrip = ndiag == 27;
      for(u32 g = 0; g < 9 && dtrip; ++g) dtrip = ho[3 * g] + 1 == ho[3 * g + 1] && ho[3 * g + 1] + 1 == ho[3 * g + 2];
      domin = int(std::int64_t(ho.front()) + 1 - std::int64_t(rows));
      domax = int(std::int64_t(ho.back()) + 1 - std::int64_t(rows));
      std::swap(didx.p, ix.p);
      {
        DevMem released;
        std::swap(released.p, bval.p);
      }
      dnw = nw;
      dict = true;
      return rt::kOk;
    }

    // the launch grid of the banded row kernels (with_rows); staged: 512 rows per block (two
    // rounds of mg::kStageRows)
    mg::RowGrid band_grid() const
    {
      if(rows_per_block) return mg::row_grid_fixed(rows, rows_per_block);
      return stage ? mg::row_grid_fixed(rows, 2 * mg::kStageRows) : mg::row_grid(rows, 1);
    }

    // The LDS windows of a banded matrix for rounds of kStageRows rows: the column ranges
    // [o_d, o_d + R) of the diagonals (offsets o_d ascending) merged where they overlap or lie
    // less than 64 columns apart; staged only if at most kStageWin windows of in total at most
    // kStageBytes / 8 values result (FP64-sized, so the same windows serve FP32 arithmetic)
    int build_stage(hipStream_t s)
    {
      stage = false;
      if(!banded || dict || ndiag == 0) return rt::kOk;
      std::vector<u32> ho(ndiag);
      F3_CHECK(hipMemcpy(ho.data(), boff.p, sizeof(u32) * ndiag, hipMemcpyDeviceToHost));
      const std::int64_t R = mg::kStageRows;
      std::vector<std::int64_t> ws, we;   // windows [ws, we)
      std::vector<u32> win(ndiag);
      for(u32 d = 0; d < ndiag; ++d)
      {
        const std::int64_t o = std::int64_t(ho[d]) + 1 - std::int64_t(rows);
        if(ws.empty() || o > we.back() + 64) { ws.push_back(o); we.push_back(o + R); }
        else we.back() = std::max(we.back(), o + R);
        win[d] = u32(ws.size() - 1);
      }
      if(ws.size() > mg::kStageWin) return rt::kOk;
      u64 total = 0;
      mg::BandStage b{};
      b.nwin = u32(ws.size());
      for(std::size_t w = 0; w < ws.size(); ++w)
      {
        b.wstart[w] = int(ws[w]);
        b.wlen[w] = u32(we[w] - ws[w]);
        b.wbase[w] = u32(total);
        total += b.wlen[w];
      }
      if(total * 8 > mg::kStageBytes) return rt::kOk;
      std::vector<u32> xo(ndiag);
      for(u32 d = 0; d < ndiag; ++d)
        xo[d] = b.wbase[win[d]] + u32(std::int64_t(ho[d]) + 1 - std::int64_t(rows) - ws[win[d]]);
      F3_RC(stg_xoff.alloc(sizeof(u32) * ndiag));
      F3_CHECK(hipMemcpyAsync(stg_xoff.p, xo.data(), sizeof(u32) * ndiag, hipMemcpyHostToDevice, s));
      F3_CHECK(hipStreamSynchronize(s));
      b.xoff = stg_xoff.template as<u32>();
      stg = b;
      stage = true;
      return rt::kOk;
    }

    // Plane-strip order of the row ranges of a banded matrix.  Each XCD evaluates a contiguous
    // set of row ranges (dev::xcd_remap), in ascending order by default: with ~256 blocks in
    // flight per XCD its working set spans several node planes, and a row's reads at the
    // mirrored (symmetric storage) and gathered positions one plane below miss the 4 MB L2
    // about one time in five (3D Q1 L7 FP64 outer product: 1.22x its stored bytes).  Here the
    // ranges of an XCD are taken strip by strip (range index modulo the plane length in ranges)
    // and, within a strip, plane by plane, so that the range one plane below was evaluated just
    // before.  The plane length is the band's largest offset.  Rows are the same; only the order
    // in which blocks take them changes (the reductions sum their block partials in block order).
    int make_tiles(hipStream_t s)
    {
      tile_blocks = 0;
      if(!banded || ndiag == 0) return rt::kOk;
      std::vector<u32> ho(ndiag);
      F3_CHECK(hipMemcpy(ho.data(), boff.p, sizeof(u32) * ndiag, hipMemcpyDeviceToHost));
      const std::int64_t omax = std::int64_t(ho.back()) - std::int64_t(rows) + 1;
      const mg::RowGrid g = band_grid();
      const u64 nb = g.blocks;
      const u64 tp = std::max<u64>(1, u64((double(omax) + 0.5 * double(g.per)) / double(g.per)));
      if(omax <= 0 || tp < 2 || nb < 16) return rt::kOk;
      std::vector<u32> order(nb);
      const u64 q = nb / 8, r = nb % 8;
      for(u64 x = 0; x < 8; ++x)
      {
        const u64 start = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
        const u64 cnt = q + (x < r ? 1 : 0);
        std::vector<u32> t(cnt);
        for(u64 i = 0; i < cnt; ++i) t[i] = u32(start + i);
        std::stable_sort(t.begin(), t.end(), [&](u32 a, u32 b) { return a % tp < b % tp; });
        for(u64 i = 0; i < cnt; ++i) order[start + i] = t[i];
      }
      F3_RC(tile.alloc(sizeof(u32) * nb));
      F3_CHECK(hipMemcpyAsync(tile.p, order.data(), sizeof(u32) * nb, hipMemcpyHostToDevice, s));
      F3_CHECK(hipStreamSynchronize(s));
      tile_blocks = unsigned(nb);
      return rt::kOk;
    }

    // bytes of one row-evaluation pass over this matrix as stored (values + indices + row
    // pointers / chunk descriptors, SELL padding included), excluding the vector traffic
    u64 matrix_bytes() const
    {
      if(tensor)
      {
        u64 b = 0;
        for(int a = 0; a < 3; ++a) b += u64(tmo[a]) * (4 + u64(mg::kTensorW) * (4 + tw_bytes));
        return b;
      }
      if(banded && dict) return u64(dnw) * rows * 4 + u64(ndiag) * 4 + u64(dnw) * 8 * mg::kDictSlots * sizeof(VT);
      if(banded) return u64(bsym ? nds : ndiag) * rows * sizeof(VT) + u64(ndiag) * 4 * (bsym ? 2 : 1);
      if(sell)
      {
        const u64 nch = (rows + C - 1) / C;
        return sell_size * (sizeof(VT) + (sell_d16 ? 2 : 4)) + (nch + 1) * 4 + nch * 4 + (sell_perm ? rows * 4 : 0);
      }
      return nnz * (sizeof(VT) + 4) + (rows + 1) * 4;
    }

    // the TensorRows view of the factor tables (tw_bytes == sizeof(DT))
    template<typename DT>
    mg::TensorRows<DT> tensor_rows() const
    {
      mg::TensorRows<DT> T;
      for(int a = 0; a < 3; ++a)
      {
        T.ax[a] = mg::TensorAxis<DT>{tcnt[a].template as<u32>(), tidx[a].template as<u32>(), tw[a].template as<DT>()};
        T.mo[a] = tmo[a];
        T.mi[a] = tmi[a];
      }
      T.div0 = mg::udiv_magic_for(tmo[0]);
      T.div1 = mg::udiv_magic_for(tmo[1]);
      T.rows = rows;
      return T;
    }

    // factor tables (per axis: cnt[o], idx/w[o * kTensorW + k]) to the device, then every stored
    // entry checked against their product; *ok = the check passed
    template<typename DT>
    int set_tensor(const u32 mo[3], const u32 mi[3], const std::vector<u32> cnt[3], const std::vector<u32> idx[3],
                   const std::vector<DT> w[3], hipStream_t s, bool* ok)
    {
      *ok = false;
      if(mo[0] < 2 || mo[1] < 2) return rt::kOk;   // the row kernel divides by mo[0] and mo[1]
      // the row kernel's reciprocal division, checked at every quotient boundary it meets
      // (q < rows for mo[0], q < mo[1] mo[2] for mo[1])
      const u64 lim[2] = {rows, u64(mo[1]) * mo[2]};
      for(int a = 0; a < 2; ++a)
      {
        const u64 m = mg::udiv_magic_for(mo[a]);
        for(u64 k = 1; k * mo[a] <= lim[a]; ++k)
          if(mg::udiv_magic(u32(k * mo[a] - 1), m) != k - 1 ||
             (k * mo[a] < lim[a] && mg::udiv_magic(u32(k * mo[a]), m) != k))
            return rt::kOk;
      }
      std::vector<u32> sidx[3];
      for(int a = 0; a < 3; ++a)
      {
        tmo[a] = mo[a];
        tmi[a] = mi[a];
        // indices scaled by the axis's stride in the input lattice: a column is iz + iy + ix
        const u64 stride = a == 0 ? 1 : (a == 1 ? u64(mi[0]) : u64(mi[0]) * mi[1]);
        sidx[a].resize(idx[a].size());
        for(std::size_t k = 0; k < idx[a].size(); ++k) sidx[a][k] = u32(u64(idx[a][k]) * stride);
        F3_RC(tcnt[a].alloc(sizeof(u32) * cnt[a].size()));
        F3_RC(tidx[a].alloc(sizeof(u32) * idx[a].size()));
        F3_RC(tw[a].alloc(sizeof(DT) * w[a].size()));
        F3_CHECK(hipMemcpyAsync(tcnt[a].p, cnt[a].data(), sizeof(u32) * cnt[a].size(), hipMemcpyHostToDevice, s));
        F3_CHECK(hipMemcpyAsync(tidx[a].p, sidx[a].data(), sizeof(u32) * idx[a].size(), hipMemcpyHostToDevice, s));
        F3_CHECK(hipMemcpyAsync(tw[a].p, w[a].data(), sizeof(DT) * w[a].size(), hipMemcpyHostToDevice, s));
      }
      tw_bytes = sizeof(DT);
      tk = 0;
      for(int a = 0; a < 3; ++a)
        for(u32 c : cnt[a]) tk = std::max(tk, c);
      DevMem bad;
      F3_RC(bad.alloc(sizeof(int)));
      F3_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
      k_tensor_check<DT, VT><<<setup_blocks(rows), 256, 0, s>>>(tensor_rows<DT>(), rp.template as<u32>(), col.template as<u32>(),
                                                               val.template as<VT>(), bad.template as<int>());
      F3_RC(rt::check_launch());
      int hbad = 1;
      F3_CHECK(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
      F3_CHECK(hipStreamSynchronize(s));   // the host tables leave scope
      *ok = hbad == 0;
      return rt::kOk;
    }
  };

  // Tensor-product structure of a transfer pair on structured hierarchies.  P (fine x coarse)
  // of a Q1/Q2 hierarchy on the unit square or cube, numbered lexicographically on its node
  // lattice (the solver's lattice order), is Pz (x) Py (x) Px with 1D factors of 1-3 entries per
  // row, and R = P^T is the product of their transposes.  The lattices are taken as m^3 (then
  // m^2) nodes; P's factors are read from the rows of its first line along each axis (rows
  // (0,0,x), (0,y,0), (z,0,0), whose entries must lie on the matching coarse line), R's are
  // their transposes; then every stored entry of P and R is checked on the device against the
  // product in the arithmetic type, bit for bit.  Only if all of that holds are both evaluated
  // by mg::TensorRows (no matrix data read, identical sums); otherwise they keep their format.
  template<typename DT, typename VT>
  int build_tensor_pair(DevCsr<VT>& P, DevCsr<VT>& R, hipStream_t s)
  {
    constexpr u32 W = mg::kTensorW;
    P.tensor = R.tensor = false;
    if(P.rows == 0 || P.cols == 0 || P.rows >= (u64(1) << 32) || P.cols >= (u64(1) << 32) || R.rows != P.cols ||
       R.cols != P.rows || P.sell || R.sell || P.banded || R.banded)
      return rt::kOk;
    auto lattice = [](u64 n, int d, u32& m) {
      const double g = std::round(std::pow(double(n), 1.0 / d));
      for(u64 c = g > 1.0 ? u64(g) - 1 : 1; c <= u64(g) + 1; ++c)
      {
        u64 p = 1;
        for(int i = 0; i < d; ++i) p *= c;
        if(p == n) { m = u32(c); return true; }
      }
      return false;
    };
    for(int d = 3; d >= 2; --d)
    {
      u32 mf = 0, mc = 0;
      if(!lattice(P.rows, d, mf) || !lattice(P.cols, d, mc)) continue;
      const u32 mo[3] = {mf, mf, d == 3 ? mf : 1u}, mi[3] = {mc, mc, d == 3 ? mc : 1u};
      // the rows of the first line along each axis (a 2D z factor is the 1x1 identity)
      std::vector<u32> list;
      for(int a = 0; a < d; ++a)
        for(u32 i = 0; i < mo[a]; ++i) list.push_back(a == 0 ? i : (a == 1 ? i * mo[0] : i * mo[0] * mo[1]));
      const u32 nl = u32(list.size());
      DevMem dl, dc, dcol, dval;
      F3_RC(dl.alloc(sizeof(u32) * nl));
      F3_RC(dc.alloc(sizeof(u32) * nl));
      F3_RC(dcol.alloc(sizeof(u32) * nl * W));
      F3_RC(dval.alloc(sizeof(DT) * nl * W));
      F3_CHECK(hipMemcpyAsync(dl.p, list.data(), sizeof(u32) * nl, hipMemcpyHostToDevice, s));
      k_tensor_gather<DT, VT><<<setup_blocks(nl), 256, 0, s>>>(nl, dl.template as<u32>(), P.rp.template as<u32>(),
                                                             P.col.template as<u32>(), P.val.template as<VT>(),
                                                             dc.template as<u32>(), dcol.template as<u32>(), dval.template as<DT>());
      F3_RC(rt::check_launch());
      std::vector<u32> hc(nl), hcol(u64(nl) * W);
      std::vector<DT> hval(u64(nl) * W);
      F3_CHECK(hipMemcpyAsync(hc.data(), dc.p, sizeof(u32) * nl, hipMemcpyDeviceToHost, s));
      F3_CHECK(hipMemcpyAsync(hcol.data(), dcol.p, sizeof(u32) * nl * W, hipMemcpyDeviceToHost, s));
      F3_CHECK(hipMemcpyAsync(hval.data(), dval.p, sizeof(DT) * nl * W, hipMemcpyDeviceToHost, s));
      F3_CHECK(hipStreamSynchronize(s));
      std::vector<u32> cnt[3], idx[3];
      std::vector<DT> w[3];
      bool good = true;
      u32 li = 0;
      for(int a = 0; a < 3; ++a)
      {
        cnt[a].assign(mo[a], 0);
        idx[a].assign(u64(mo[a]) * W, 0);
        w[a].assign(u64(mo[a]) * W, DT(0));
        if(a >= d) { cnt[a][0] = 1; w[a][0] = DT(1); continue; }
        const u64 stride = a == 0 ? 1 : (a == 1 ? u64(mi[0]) : u64(mi[0]) * mi[1]);
        for(u32 i = 0; i < mo[a]; ++i, ++li)
        {
          if(hc[li] > W) { good = false; break; }
          cnt[a][i] = hc[li];
          for(u32 k = 0; k < hc[li]; ++k)
          {
            const u64 c = hcol[u64(li) * W + k];
            if(c % stride != 0 || c / stride >= mi[a]) { good = false; break; }
            idx[a][u64(i) * W + k] = u32(c / stride);
            w[a][u64(i) * W + k] = hval[u64(li) * W + k];
          }
          if(!good) break;
        }
        if(!good) break;
      }
      if(!good) continue;
      bool ok = false;
      F3_RC(P.template set_tensor<DT>(mo, mi, cnt, idx, w, s, &ok));
      if(!ok) continue;
      // R's factors: the transposes (coarse index j lists the fine indices i with a P entry
      // (i, j), ascending in i)
      std::vector<u32> rcnt[3], ridx[3];
      std::vector<DT> rw[3];
      for(int a = 0; a < 3 && good; ++a)
      {
        rcnt[a].assign(mi[a], 0);
        ridx[a].assign(u64(mi[a]) * W, 0);
        rw[a].assign(u64(mi[a]) * W, DT(0));
        for(u32 i = 0; i < mo[a] && good; ++i)
          for(u32 k = 0; k < cnt[a][i]; ++k)
          {
            const u32 j = idx[a][u64(i) * W + k];
            if(rcnt[a][j] >= W) { good = false; break; }
            ridx[a][u64(j) * W + rcnt[a][j]] = i;
            rw[a][u64(j) * W + rcnt[a][j]] = w[a][u64(i) * W + k];
            ++rcnt[a][j];
          }
      }
      if(!good) return rt::kOk;
      F3_RC(R.template set_tensor<DT>(mi, mo, rcnt, ridx, rw, s, &ok));
      if(!ok) return rt::kOk;
      P.tensor = R.tensor = true;
      return rt::kOk;
    }
    return rt::kOk;
  }

  // calls f(rows_evaluator, launch grid) for the matrix's format and lane count
  template<typename DT, typename VT, typename F>
  int with_rows(const DevCsr<VT>& m, F&& f)
  {
    auto grid = [&](int G) { return m.rows_per_block ? mg::row_grid_fixed(m.rows, m.rows_per_block) : mg::row_grid(m.rows, G); };
    if(m.tensor) return rt::kErrState;   // transfers only: with_transfer_rows
    if(m.banded && m.stage)
    {
      const mg::RowGrid g = m.band_grid();
      const u32* tile = m.tile_blocks == g.blocks ? m.tile.template as<u32>() : nullptr;
      if(m.bsym)
      {
        mg::BandedRows<DT, VT, true, true> A{};
        A.off = m.boff.template as<u32>(); A.val = m.bval.template as<VT>(); A.rows = m.rows; A.cols = m.cols;
        A.ndiag = m.ndiag; A.sx = m.bsx.template as<u32>(); A.nds = m.nds; A.tile = tile; A.st = m.stg;
        f(A, g);
      }
      else
      {
        mg::BandedRows<DT, VT, false, true> A{};
        A.off = m.boff.template as<u32>(); A.val = m.bval.template as<VT>(); A.rows = m.rows; A.cols = m.cols;
        A.ndiag = m.ndiag; A.tile = tile; A.st = m.stg;
        f(A, g);
      }
      return rt::kOk;
    }
    if(m.banded && m.dict)
    {
      // matrices of 2^20 rows or more: FP32 512 rows per block, two per thread in the interior
      // blocks; FP64 768 rows per block, three per thread one after another (every thread the same
      // count: 3D Q1 L7 outer product 40.0 us with row_grid's 320-row blocks, where a quarter of
      // the threads take a second row, 36.5 with 512, 35.3 with 768, 37.3 with 1024)
      const bool big = !m.rows_per_block && m.rows >= (u64(1) << 20);
      const mg::RowGrid g = big ? mg::row_grid_fixed(m.rows, sizeof(DT) == 4 ? 256 * F3_DICT_RT : 768) : grid(1);
      const u32* tile = m.tile_blocks == g.blocks ? m.tile.template as<u32>() : nullptr;
      // unrolled in full for the diagonal counts of Q1 operators (2D: 9; 3D: 27, 21 without the
      // face couplings below eps_f32^2) and of the 2D Q2 operator (25)
      auto go = [&](auto nd, auto trip) {
        mg::DictBandRows<DT, VT, decltype(nd)::value, decltype(trip)::value> A{
          m.doff.template as<u32>(), m.didx.template as<u32>(), m.dtab.template as<VT>(), m.rows, m.cols, m.ndiag, m.dnw,
          m.domin, m.domax};
        A.tile = tile;
        f(A, g);
      };
      const std::false_type plain;
      if(m.ndiag == 27)
      {
        // the FP64 3D Q1 stencil: one x load per run of three diagonals (DictBandRows TRIP_)
        if constexpr(sizeof(DT) == 8)
        {
          if(m.dtrip && m.allow_lane_shift) go(std::integral_constant<int, 27>(), std::true_type());
          else go(std::integral_constant<int, 27>(), plain);
        }
        else
          go(std::integral_constant<int, 27>(), plain);
      }
      else if(m.ndiag == 25) go(std::integral_constant<int, 25>(), plain);
      else if(m.ndiag == 21) go(std::integral_constant<int, 21>(), plain);
      else if(m.ndiag == 9) go(std::integral_constant<int, 9>(), plain);
      else go(std::integral_constant<int, 0>(), plain);
      return rt::kOk;
    }
    if(m.banded)
    {
      const mg::RowGrid g = grid(1);
      const u32* tile = m.tile_blocks == g.blocks ? m.tile.template as<u32>() : nullptr;
      if(m.bsym)
      {
        mg::BandedRows<DT, VT, true> A{m.boff.template as<u32>(), m.bval.template as<VT>(), m.rows, m.cols, m.ndiag,
                                       m.bsx.template as<u32>(), m.nds};
        A.tile = tile;
        f(A, g);
      }
      else
      {
        mg::BandedRows<DT, VT> A{m.boff.template as<u32>(), m.bval.template as<VT>(), m.rows, m.cols, m.ndiag};
        A.tile = tile;
        f(A, g);
      }
      return rt::kOk;
    }
    if(m.sell)
    {
      const mg::RowGrid g = grid(1);
      const u32* cs = m.cs.template as<u32>();
      const u32* cl = m.cl.template as<u32>();
      const u32* pm = m.perm.template as<u32>();
      const VT* v = m.sval.template as<VT>();
      if(m.sell_d16)
      {
        const std::int16_t* c = m.scol.template as<std::int16_t>();
        if(m.sell_perm) f(mg::SellRows<DT, VT, std::int16_t, true>{cs, cl, pm, c, v, m.rows, m.C}, g);
        else f(mg::SellRows<DT, VT, std::int16_t, false>{cs, cl, pm, c, v, m.rows, m.C}, g);
      }
      else
      {
        const u32* c = m.scol.template as<u32>();
        if(m.sell_perm) f(mg::SellRows<DT, VT, u32, true>{cs, cl, pm, c, v, m.rows, m.C}, g);
        else f(mg::SellRows<DT, VT, u32, false>{cs, cl, pm, c, v, m.rows, m.C}, g);
      }
      return rt::kOk;
    }
    const u32* rp = m.rp.template as<u32>();
    const u32* col = m.col.template as<u32>();
    const VT* val = m.val.template as<VT>();
    switch(m.lanes)
    {
      case 0:
      {
        // streaming evaluator: rounds of R rows filling about 85% of the LDS products
        const double mean = m.rows ? double(m.nnz) / double(m.rows) : 1.0;
        const double fit = 0.85 * double(mg::CsrStreamRows<DT, VT>::CAP) / (mean > 1.0 ? mean : 1.0);
        u64 R = fit < 1.0 ? 1 : (fit > 1024.0 ? 1024 : u64(fit));
        if(R >= 64) R = R / 64 * 64;
        // one round of R rows per block (a full LDS of products each) where rounds are short
        // (R <= 256 rows: one row per thread in the summing phase) and that still gives at least
        // 2048 blocks (8 per CU); otherwise the default grid.  Measured on the restriction of
        // 2D Q2 L9 (V-cycle 0.61 -> 0.59 ms); full 1024-row rounds of the 3D prolongation and
        // too few blocks on mid-size levels were slower.
        const bool full_rounds = R >= 64 && R <= 256 && m.rows >= R * 2048;
        const mg::RowGrid g = m.rows_per_block ? mg::row_grid_fixed(m.rows, m.rows_per_block)
                                               : (full_rounds ? mg::row_grid_fixed(m.rows, R) : grid(1));
        f(mg::CsrStreamRows<DT, VT>{rp, col, val, m.rows, R}, g);
        break;
      }
      case 2: f(mg::CsrRows<DT, VT, 2>{rp, col, val, m.rows}, grid(2)); break;
      case 4: f(mg::CsrRows<DT, VT, 4>{rp, col, val, m.rows}, grid(4)); break;
      case 8: f(mg::CsrRows<DT, VT, 8>{rp, col, val, m.rows}, grid(8)); break;
      case 16: f(mg::CsrRows<DT, VT, 16>{rp, col, val, m.rows}, grid(16)); break;
      case 32: f(mg::CsrRows<DT, VT, 32>{rp, col, val, m.rows}, grid(32)); break;
      case 64: f(mg::CsrRows<DT, VT, 64>{rp, col, val, m.rows}, grid(64)); break;
      default: return rt::kErrArg;
    }
    return rt::kOk;
  }

  // the same for a transfer matrix, which may also be evaluated as a tensor product: the row
  // evaluator is instantiated for the bound of its factors' entries per axis (2, 3, 5 or 8) and
  // for 2D (z factor the identity) or 3D
  template<typename DT, int K, bool D3, int RT>
  mg::TensorRows<DT, K, D3, RT> tensor_as(const mg::TensorRows<DT>& t)
  {
    mg::TensorRows<DT, K, D3, RT> r;
    for(int a = 0; a < 3; ++a) { r.ax[a] = t.ax[a]; r.mo[a] = t.mo[a]; r.mi[a] = t.mi[a]; }
    r.div0 = t.div0;
    r.div1 = t.div1;
    r.rows = t.rows;
    return r;
  }
  template<typename DT, typename VT, typename F>
  int with_transfer_rows(const DevCsr<VT>& m, F&& f)
  {
    if(!m.tensor) return with_rows<DT, VT>(m, f);
    if(m.tw_bytes != sizeof(DT)) return rt::kErrState;
    const mg::TensorRows<DT> t = m.template tensor_rows<DT>();
    // one row per thread in the transfers' block size (four rows per thread, 1024 per block, took
    // the 3D Q1 level-7 prolongation from 17.6 to 25.1 us)
    const mg::RowGrid g = m.rows_per_block ? mg::row_grid_fixed(m.rows, m.rows_per_block) : mg::row_grid(m.rows, 1);
    const bool d3 = m.tmo[2] > 1;
    auto go = [&](auto k) {
      constexpr int K = decltype(k)::value;
      if(d3) f(tensor_as<DT, K, true, 1>(t), g); else f(tensor_as<DT, K, false, 1>(t), g);
    };
    if(m.tk <= 2) go(std::integral_constant<int, 2>());
    else if(m.tk <= 3) go(std::integral_constant<int, 3>());
    else if(m.tk <= 5) go(std::integral_constant<int, 5>());
    else go(std::integral_constant<int, 8>());
    return rt::kOk;
  }

  inline unsigned nblocks(u64 threads)
  {
    const u64 b = (threads + kBlock - 1) / kBlock;
    return unsigned(std::max<u64>(b, 1));
  }
  inline unsigned stream_blocks(u64 n)
  {
    return unsigned(std::max<u64>(1, std::min<u64>((n + kBlock - 1) / kBlock, 4096)));
  }

  // Gauss-Jordan inverse with partial pivoting (FP64), for the coarse level
  bool invert_dense(std::vector<double>& a, u64 n)
  {
    std::vector<double> inv(n * n, 0.0);
    for(u64 i = 0; i < n; ++i) inv[i * n + i] = 1.0;
    for(u64 k = 0; k < n; ++k)
    {
      u64 pr = k;
      for(u64 i = k + 1; i < n; ++i)
        if(std::fabs(a[i * n + k]) > std::fabs(a[pr * n + k])) pr = i;
      if(a[pr * n + k] == 0.0) return false;
      if(pr != k)
        for(u64 j = 0; j < n; ++j) { std::swap(a[k * n + j], a[pr * n + j]); std::swap(inv[k * n + j], inv[pr * n + j]); }
      const double d = a[k * n + k];
      for(u64 j = 0; j < n; ++j) { a[k * n + j] /= d; inv[k * n + j] /= d; }
      for(u64 i = 0; i < n; ++i)
      {
        if(i == k) continue;
        const double f = a[i * n + k];
        if(f == 0.0) continue;
        for(u64 j = 0; j < n; ++j) { a[i * n + j] -= f * a[k * n + j]; inv[i * n + j] -= f * inv[k * n + j]; }
      }
    }
    a.swap(inv);
    return true;
  }

  std::vector<u32> make_mask(u64 n, const std::vector<u32>& fidx)
  {
    std::vector<u32> m((n + 31) / 32 + 1, 0u);
    for(u32 i : fidx) m[i >> 5] |= (1u << (i & 31u));
    return m;
  }

  template<typename T>
  __global__ void k_to_double(u64 n, const T* z, double* out)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      out[i] = double(z[i]);
  }

  // -----------------------------------------------------------------------------------------
  // the engine: type-erased interface + implementation per inner precision
  // -----------------------------------------------------------------------------------------
  struct Config
  {
    double omega_j = 0.5, omega_r = 1.0;
    int pre = 4, post = 4;
    int precond = 0;          // 0 multigrid cycle, 1 none (CG)
    int cycle = 0;            // 0 V, 1 F, 2 W (multigrid.hpp:1333-1519)
    bool use_graph = true;
    int format = 0;           // 0 CSR, 1 SELL-C-sigma, 2 banded where it fits (else SELL)
    u64 sell_C = 64, sell_sigma = 64;
    int lanes = 0;            // forced lanes per CSR row (0 = the streaming evaluator)
    bool index16 = true;      // SELL columns as 16-bit row offsets where they fit
    bool compact_transfer = true; // device P/R without the stored structural zeros
    // levels with at most this many rows (and whose vectors fit in one workgroup's LDS) run in
    // the single-workgroup coarse-tail kernel k_tail; 0 = every level keeps its own launches
    u64 tail_max_n = 1100;
    u64 rows_per_block = 0;   // row positions per block of the row kernels (0: by row_grid)
    u64 outer_rows_per_block = 0;   // the same for the FP64 outer matrix when it is a copy of its own
    // SELL sort window of the transfer matrices: prolongation rows have 1/2/4/8 entries in an
    // alternating pattern, so sorting by length inside windows removes most of the padding
    u64 transfer_sigma = 1;
    // transfer matrices (P, R) in CSR, evaluated by streaming, instead of the level format:
    // prolongation rows hold 1..8 entries (Q1), which SELL pads to the chunk's widest row
    // (measured 3D Q1 L7 FP32 V-cycle 0.603 -> 0.583 ms, 2D Q2 L9 0.643 -> 0.610 ms)
    bool transfer_csr = true;
    // transfers of structured hierarchies evaluated as tensor products of their 1D factors
    // (build_tensor_pair: detected and checked entry by entry; identical results, no matrix data)
    bool tensor_transfer = true;
    // the V-cycle kernels of a solve return at once when the solver status has left "progress"
    // (0: they run regardless; the PCG kernels behind them still check it, so the results are the
    // same, and a cycle launched after the last iteration does its work for nothing)
    bool vcycle_status = true;
    // launch shape of the transfer matrices: lanes per row (-1: as the level matrices) and row
    // positions per block (0: derived from the row count).  128 rows per block (more, smaller
    // blocks than the default grid on the mid-size levels) measured V-cycle 3D Q1 L7 584 -> 580 us,
    // 2D Q2 L9 590 -> 576 us; 64 rows, or lifting the 8192-block grid limit, were slower.
    int transfer_lanes = -1;
    u64 transfer_rows_per_block = 128;
    // FP64 bands that equal their mirror image bit for bit store only the center and upper
    // diagonals (0: always the full band; the results are identical either way)
    bool symmetric_band = true;
    bool band_dict = true;   // value-coded bands (mg::DictBandRows) where every diagonal has few values
    // FP64 value-coded rows of the 3D Q1 stencil read one x per run of three diagonals and take
    // the other two from the neighbouring lanes (identical results; 0: three loads)
    bool lane_shift = true;
    // processing order of the row ranges of banded matrices (DevCsr::make_tiles): 0 ascending,
    // 1 plane strips for the FP64 outer matrix, 2 plane strips for every banded matrix
    int plane_order = 1;
    // x of the banded row kernels staged in LDS per round of rows (mg::BandStage): 0 no, 1 the
    // FP64 outer matrix, 2 every banded matrix
    int band_stage_x = 0;
    double tol_rel = 1e-8, tol_abs = 1.0 / (2.220446049250313e-16 * 2.220446049250313e-16), tol_abs_low = 0.0;
    double div_rel = 1.0 / 2.220446049250313e-16, div_abs = 1.0 / (2.220446049250313e-16 * 2.220446049250313e-16);
    double stag_rate = 0.95;
    u64 min_iter = 0, max_iter = 1000, min_stag_iter = 0;
  };

  struct EngineBase
  {
    virtual ~EngineBase() {}
    virtual int init(const std::vector<HostLevel>& lv, const Config& cfg, hipStream_t s) = 0;
    virtual int vcycle(double* cor, const double* def) = 0;
    // epilogue (optional): enqueued on the solver's stream after the solve's last launch, before
    // the final wait
    virtual int solve(double* x, const double* b, bool correct, int* status, u64* iters,
                      const std::function<int()>* epilogue = nullptr) = 0;
    virtual int history(double* out, u64 cap, u64* len) = 0;
    virtual int time_kernel(int which, int reps, double* avg_ms, u64* bytes) = 0;
    virtual int info(u64* out, int cap) = 0;
    virtual int statistics(int reps, double* out, int cap) = 0;
    virtual bool set_tolerances(const Config& c) = 0;
  };

  template<typename DT, typename VT>
  struct Level
  {
    u64 n = 0;
    DevCsr<VT> A, P, R;
    DevMem mask, dinv, rhs, x, t, def;
    DT* xfinal = nullptr; // buffer holding this level's solution after its post-smoothing
  };

  template<typename DT, typename VT>
  struct Engine : public EngineBase
  {
    Config cfg;
    hipStream_t s = nullptr;
    std::vector<std::unique_ptr<Level<DT, VT>>> L;
    u64 ncoarse = 0;
    DevMem ainv;
    // outer FP64 PCG
    DevCsr<double> A64own;
    DevMem r, p, q, zcg, state, hist, partials;
    const DevCsr<double>* A64 = nullptr;
    u64 n0 = 0, hist_cap = 0;
    // an iteration is captured as one graph: the preconditioner half of iteration k (z = M r,
    // beta, p update; for k = 0 gamma and p from the first preconditioner application), the
    // outer half of iteration k + 1 (q = A p, alpha, x and r update, stopping criterion) and the
    // publication of the state
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    // published solver states (mg::k_publish): pinned, host-coherent; ring_dev is its device
    // address, pub_count the device's count of publications in the running solve
    static constexpr int kRing = 4;
    struct HostRing
    {
      mg::PcgState st[kRing];
      unsigned seq[kRing];
    };
    HostRing* ring = nullptr;
    HostRing* ring_dev = nullptr;
    DevMem pub_count;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    DevMem timing_state;   // solver state "in progress" for timing the outer product alone

    ~Engine() override
    {
      drop_graphs();
      if(ring) (void)hipHostFree(ring);
      if(t0) (void)hipEventDestroy(t0);
      if(t1) (void)hipEventDestroy(t1);
      for(hipEvent_t e : stat_ev) (void)hipEventDestroy(e);
    }

    const DevCsr<double>& outer_matrix() const { return *A64; }

    // tolerance changes after init; false if the history buffer is too small
    bool set_tolerances(const Config& c) override
    {
      if(c.max_iter + 2 > hist_cap) return false;
      cfg.tol_rel = c.tol_rel; cfg.tol_abs = c.tol_abs; cfg.tol_abs_low = c.tol_abs_low;
      cfg.min_iter = c.min_iter; cfg.max_iter = c.max_iter;
      return true;
    }

    // thin rule of the level matrices: FP64 keeps every entry, inner FP32 (and binary16 storage, which
    // computes in FP32) drops the couplings below eps_f32^2 of the row diagonal
    static double level_thin()
    {
      const double e = 1.1920928955078125e-07;
      return sizeof(VT) == 8 ? -1.0 : e * e;
    }

    int init(const std::vector<HostLevel>& lv, const Config& c, hipStream_t stream) override
    {
      cfg = c;
      s = stream;
      if(lv.empty()) return rt::kErrState;
      if(cfg.pre < 1 || cfg.post < 1) return rt::kErrArg;
      const bool want_sell = cfg.format >= 1;
      const bool want_banded = cfg.format == 2;
      L.clear();
      for(std::size_t i = 0; i < lv.size(); ++i)
      {
        std::unique_ptr<Level<DT, VT>> l(new Level<DT, VT>());
        const HostLevel& h = lv[i];
        l->n = h.A.rows;
        l->A.allow_sym = cfg.symmetric_band;
        l->A.allow_dict = cfg.band_dict;
        l->A.allow_lane_shift = cfg.lane_shift;
        if(h.A.on_device())
          F3_RC(l->A.upload_dev(h.A, s, want_sell, cfg.sell_C, cfg.sell_sigma, cfg.lanes, level_thin(), cfg.index16, want_banded));
        else
          F3_RC(l->A.upload(h.A, s, want_sell, cfg.sell_C, cfg.sell_sigma, cfg.lanes, level_thin(), cfg.index16, want_banded));
        l->A.rows_per_block = cfg.rows_per_block;
        if(i + 1 < lv.size())
        {
          if(h.P.rows != l->n || h.R.cols != l->n || h.P.cols != lv[i + 1].A.rows || h.R.rows != lv[i + 1].A.rows)
            return rt::kErrArg;
          const bool tsell = want_sell && !cfg.transfer_csr;
          const int tlanes = cfg.transfer_lanes >= 0 ? cfg.transfer_lanes : cfg.lanes;
          const double tthin = cfg.compact_transfer ? 0.0 : -1.0;
          if(h.P.on_device())
          {
            F3_RC(l->P.upload_dev(h.P, s, tsell, cfg.sell_C, cfg.transfer_sigma, tlanes, tthin, cfg.index16, false));
            F3_RC(l->R.upload_dev(h.R, s, tsell, cfg.sell_C, cfg.transfer_sigma, tlanes, tthin, cfg.index16, false));
          }
          else
          {
            F3_RC(l->P.upload(h.P, s, tsell, cfg.sell_C, cfg.transfer_sigma, tlanes, tthin, cfg.index16));
            F3_RC(l->R.upload(h.R, s, tsell, cfg.sell_C, cfg.transfer_sigma, tlanes, tthin, cfg.index16));
          }
          l->P.rows_per_block = cfg.transfer_rows_per_block;
          l->R.rows_per_block = cfg.transfer_rows_per_block;
          if(cfg.tensor_transfer) F3_RC((build_tensor_pair<DT, VT>(l->P, l->R, s)));
        }
        const std::vector<u32> m = make_mask(l->n, h.fidx);
        F3_RC(l->mask.alloc(sizeof(u32) * m.size()));
        F3_CHECK(hipMemcpyAsync(l->mask.p, m.data(), sizeof(u32) * m.size(), hipMemcpyHostToDevice, s));
        F3_RC(l->dinv.alloc(sizeof(DT) * l->n));
        F3_RC(l->rhs.alloc(sizeof(DT) * l->n));
        F3_RC(l->x.alloc(sizeof(DT) * l->n));
        F3_RC(l->t.alloc(sizeof(DT) * l->n));
        F3_RC(l->def.alloc(sizeof(DT) * l->n));
        F3_CHECK(hipMemsetAsync(l->x.p, 0, sizeof(DT) * l->n, s));
        F3_CHECK(hipMemsetAsync(l->t.p, 0, sizeof(DT) * l->n, s));
        // Jacobi setup on the device matrix in the inner storage precision
        if(l->n)
        {
          mg::k_make_dinv<DT, VT><<<nblocks(l->n), kBlock, 0, s>>>(
            l->n, l->A.rp.template as<u32>(), l->A.col.template as<u32>(), l->A.val.template as<VT>(),
            l->mask.template as<u32>(), DT(cfg.omega_j), DT(cfg.omega_r), l->dinv.template as<DT>());
          F3_RC(rt::check_launch());
        }
        L.push_back(std::move(l));
      }
      // FP64 coarse inverse of the unit-filtered coarsest matrix
      const HostLevel& hc = lv.back();
      ncoarse = hc.A.rows;
      if(ncoarse > 1024) return rt::kErrArg;
      {
        std::vector<double> a(ncoarse * ncoarse, 0.0);
        for(u64 i = 0; i < ncoarse; ++i)
          for(u32 k = hc.A.rp[i]; k < hc.A.rp[i + 1]; ++k) a[i * ncoarse + hc.A.col[k]] += hc.A.val[k];
        for(u32 i : hc.fidx)
        {
          for(u64 j = 0; j < ncoarse; ++j) a[u64(i) * ncoarse + j] = 0.0;
          a[u64(i) * ncoarse + i] = 1.0;
        }
        if(!invert_dense(a, ncoarse)) return rt::kErrNumeric;
        F3_RC(ainv.alloc(sizeof(double) * ncoarse * ncoarse));
        F3_CHECK(hipMemcpyAsync(ainv.p, a.data(), sizeof(double) * a.size(), hipMemcpyHostToDevice, s));
        F3_CHECK(hipStreamSynchronize(s));
      }
      // final-solution buffers: x, then sweeps alternate between x and t
      for(std::size_t i = 0; i + 1 < L.size(); ++i)
      {
        // pre: entry/restrict writes x, (pre-1) sweeps; post: prolong in place, post sweeps
        const int flips = (cfg.pre - 1) + cfg.post;
        L[i]->xfinal = (flips % 2 == 0) ? L[i]->x.template as<DT>() : L[i]->t.template as<DT>();
      }
      L.back()->xfinal = L.back()->x.template as<DT>();
      F3_RC(setup_tail(lv));
      F3_RC(build_cycle());

      // outer PCG
      const HostLevel& h0 = lv.front();
      n0 = h0.A.rows;
      // the outer FP64 matrix uses the hierarchy's format; with an FP64 hierarchy it is the
      // finest level's matrix itself
      bool shared = false;
      if constexpr(std::is_same<VT, double>::value)
      {
        A64 = &L.front()->A;
        shared = true;
      }
      if(!shared)
      {
        A64own.allow_sym = cfg.symmetric_band;
        A64own.allow_dict = cfg.band_dict;
        A64own.allow_lane_shift = cfg.lane_shift;
        if(h0.A.on_device())
          F3_RC(A64own.upload_dev(h0.A, s, want_sell, cfg.sell_C, cfg.sell_sigma, cfg.lanes, -1.0, cfg.index16, want_banded));
        else
          F3_RC(A64own.upload(h0.A, s, want_sell, cfg.sell_C, cfg.sell_sigma, cfg.lanes, -1.0, cfg.index16, want_banded));
        A64own.rows_per_block = cfg.outer_rows_per_block;
        A64 = &A64own;
      }
      // x staged in LDS by the banded kernels: the FP64 outer matrix (1), every banded matrix (2)
      if(cfg.band_stage_x >= 1)
      {
        if(shared) F3_RC(L.front()->A.build_stage(s));
        else F3_RC(A64own.build_stage(s));
      }
      if(cfg.band_stage_x == 2)
        for(std::size_t i = shared ? 1 : 0; i < L.size(); ++i) F3_RC(L[i]->A.build_stage(s));
      // plane-strip order of the banded row ranges: the FP64 outer matrix (1), every level (2)
      if(cfg.plane_order >= 1)
      {
        if(shared) F3_RC(L.front()->A.make_tiles(s));
        else F3_RC(A64own.make_tiles(s));
      }
      if(cfg.plane_order == 2)
        for(std::size_t i = shared ? 1 : 0; i < L.size(); ++i) F3_RC(L[i]->A.make_tiles(s));
      F3_RC(r.alloc(sizeof(double) * n0));
      F3_RC(p.alloc(sizeof(double) * n0));
      F3_RC(q.alloc(sizeof(double) * n0));
      F3_RC(zcg.alloc(sizeof(double) * n0));
      F3_RC(state.alloc(sizeof(mg::PcgState)));
      hist_cap = cfg.max_iter + 2;
      F3_RC(hist.alloc(sizeof(double) * hist_cap));
      // one partial per block of the largest reducing launch: row kernels launch at most
      // 8192 blocks (mg::row_grid), the streaming ones at most 4096
      const u64 max_blocks = 8192;
      F3_RC(partials.alloc(sizeof(double) * max_blocks));
      if(!ring)
      {
        F3_CHECK(hipHostMalloc(reinterpret_cast<void**>(&ring), sizeof(HostRing), hipHostMallocMapped | hipHostMallocCoherent));
        std::memset(ring, 0, sizeof(HostRing));
        F3_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&ring_dev), ring, 0));
      }
      F3_RC(pub_count.alloc(sizeof(unsigned)));
      if(!t0) F3_CHECK(hipEventCreate(&t0));
      if(!t1) F3_CHECK(hipEventCreate(&t1));
      F3_CHECK(hipStreamSynchronize(s));
      drop_graphs();
      return rt::kOk;
    }

    void drop_graphs()
    {
      if(gexec) { (void)hipGraphExecDestroy(gexec); gexec = nullptr; }
      if(graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    }

    mg::PcgState* st() const { return state.template as<mg::PcgState>(); }
    const int* status_ptr() const { return reinterpret_cast<const int*>(&st()->status); }

    // ---- V-cycle -------------------------------------------------------------------------
    // coarse tail: levels tail0..nl-1 run in one workgroup (k_tail) when tail0 < nl - 1;
    // tail_lds bytes of LDS hold their vectors
    std::size_t tail0 = 0;
    mg::TailDesc<DT, VT> tail{};
    u64 tail_lds = 0;
    std::vector<std::unique_ptr<DevMem>> tail_mem;

    // ELL copy for the tail kernel: 16-bit columns, the entries band_keeps(v, row diagonal, thin)
    // (thin = 0: exact zeros dropped; the level matrices take level_thin()), entry k of row r at
    // k * rows + r, padded with column 0 and value 0; with FP32 arithmetic a width of at most 32
    // other than 9, 21, 25 or 27 is padded to a multiple of 8 (the kernel's fixed-width row sums)
    int tail_ell(const HostCsr& h, mg::TailEll<VT>& e, double thin)
    {
      std::vector<double> dg(h.rows, 0.0);
      for(u64 r = 0; r < h.rows; ++r)
        for(u32 k = h.rp[r]; k < h.rp[r + 1]; ++k)
          if(u64(h.col[k]) == r) { dg[r] = h.val[k]; break; }
      u64 w = 0;
      for(u64 r = 0; r < h.rows; ++r)
      {
        u64 c = 0;
        for(u32 k = h.rp[r]; k < h.rp[r + 1]; ++k) c += band_keeps(h.val[k], dg[r], thin) ? 1 : 0;
        w = std::max(w, c);
      }
      if(sizeof(DT) == 4 && w <= 32 && w != 9 && w != 21 && w != 25 && w != 27) w = (w + 7) / 8 * 8;
      std::vector<std::uint16_t> col(w * h.rows, 0);
      std::vector<VT> val(w * h.rows, to_storage<VT>(0.0));
      for(u64 r = 0; r < h.rows; ++r)
      {
        u64 j = 0;
        for(u32 k = h.rp[r]; k < h.rp[r + 1]; ++k)
          if(band_keeps(h.val[k], dg[r], thin))
          {
            col[j * h.rows + r] = std::uint16_t(h.col[k]);
            val[j * h.rows + r] = to_storage<VT>(h.val[k]);
            ++j;
          }
      }
      std::unique_ptr<DevMem> mc(new DevMem()), mv(new DevMem());
      F3_RC(mc->alloc(sizeof(std::uint16_t) * col.size()));
      F3_RC(mv->alloc(sizeof(VT) * val.size()));
      if(!col.empty())
      {
        F3_CHECK(hipMemcpyAsync(mc->p, col.data(), sizeof(std::uint16_t) * col.size(), hipMemcpyHostToDevice, s));
        F3_CHECK(hipMemcpyAsync(mv->p, val.data(), sizeof(VT) * val.size(), hipMemcpyHostToDevice, s));
      }
      F3_CHECK(hipStreamSynchronize(s));   // host staging vectors go out of scope
      e = mg::TailEll<VT>{mc->template as<std::uint16_t>(), mv->template as<VT>(), u32(h.rows), u32(w)};
      tail_mem.push_back(std::move(mc));
      tail_mem.push_back(std::move(mv));
      return rt::kOk;
    }

    int setup_tail(const std::vector<HostLevel>& lv)
    {
      const std::size_t nl = L.size();
      tail_mem.clear();
      int dev = 0, cap = 65536;
      if(hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cap, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
      auto arena = [&](std::size_t t0) {
        u64 b = 0;
        for(std::size_t i = t0; i < nl; ++i) b += 5 * L[i]->n * sizeof(DT);
        return b;
      };
      tail0 = nl - 1;
      while(tail0 > 0 && L[tail0 - 1]->n <= cfg.tail_max_n && L[tail0 - 1]->n < 65536 &&
            nl - (tail0 - 1) <= std::size_t(mg::kTailMax) && arena(tail0 - 1) <= u64(cap))
        --tail0;
      tail = mg::TailDesc<DT, VT>{};
      tail_lds = 0;
      if(tail0 + 1 >= nl) return rt::kOk;
      tail.nlev = int(nl - tail0);
      tail.ncoarse = u32(ncoarse);
      tail.ainv = ainv.template as<double>();
      tail.rhs_in = L[tail0]->rhs.template as<DT>();
      tail.x_out = L[tail0]->xfinal;
      u64 off = 0;
      for(std::size_t i = tail0; i < nl; ++i)
      {
        Level<DT, VT>& l = *L[i];
        mg::TailLevel<DT, VT>& t = tail.lv[i - tail0];
        t.n = u32(l.n);
        t.off = u32(off);
        off += 5 * l.n;
        F3_RC(tail_ell(lv[i].A, t.A, std::max(0.0, level_thin())));
        if(i + 1 < nl)
        {
          F3_RC(tail_ell(lv[i].P, t.P, 0.0));
          F3_RC(tail_ell(lv[i].R, t.R, 0.0));
        }
        t.mask = l.mask.template as<u32>();
        t.dinv = l.dinv.template as<DT>();
      }
      tail_lds = off * sizeof(DT);
      if(tail_lds > 65536)
        F3_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&mg::k_tail<DT, VT>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(tail_lds)));
      return rt::kOk;
    }

    // ---- multigrid cycle ------------------------------------------------------------------
    // A cycle is a fixed sequence of level operations, built once at init from the cycle
    // kind (multigrid.hpp:1333-1519, _apply_cycle_v/_f/_w over _apply_rest, _apply_coarse,
    // _apply_prol and _apply_smooth_peak).  The operations on the levels of the coarse tail
    // (tail0 and below) are grouped into segments that one k_tail launch each runs; a bottom
    // operation stands for the coarse solve (k_coarse) or for such a segment.
    enum class Op : int { entry, sweep, defect, rest, rest_step, bottom, prol, copy_top };
    struct CycleOp { Op op; std::size_t lvl; u32 seg; };
    struct TailSeg { u32 first, count, out_buf; };
    std::vector<CycleOp> cycle_ops;
    std::vector<TailSeg> tail_segs;
    DevMem tail_ops;

    // the whole cycle over every level, the coarse solve at the coarsest
    std::vector<CycleOp> cycle_schedule() const
    {
      std::vector<CycleOp> o;
      const std::size_t b = L.size() - 1;
      auto push = [&](Op op, std::size_t l) { o.push_back(CycleOp{op, l, 0}); };
      auto sweeps = [&](std::size_t i, int k) {
        for(int j = 0; j < k; ++j) push(Op::sweep, i);
      };
      // _apply_rest(cur, smooth), multigrid.hpp:1711-1832: pre-smoothing from zero and defect
      // on every level below cur, and on cur itself when smooth (else cur holds the defect of
      // its peak smoothing); the restriction takes the next level's first smoothing step
      auto rest = [&](std::size_t cur, bool smooth) {
        for(std::size_t i = cur; i < b; ++i)
        {
          if(smooth || i > cur)
          {
            sweeps(i, cfg.pre - 1);
            push(Op::defect, i);
          }
          push(i + 1 < b ? Op::rest_step : Op::rest, i);
        }
      };
      // _apply_prol(cur, smooth), multigrid.hpp:1843-2000 (fixed coarse-grid correction)
      auto prol = [&](std::size_t cur, bool smooth) {
        for(std::size_t i = b; i-- > cur;)
        {
          push(Op::prol, i);
          if(smooth || i > cur) sweeps(i, cfg.post);
        }
      };
      // _apply_smooth_peak(lvl), multigrid.hpp:1630-1700: pre- then post-smoother (no peak
      // smoother is configured), then the defect
      auto peak = [&](std::size_t l) {
        sweeps(l, cfg.pre + cfg.post);
        push(Op::defect, l);
      };
      auto coarse = [&]() { push(Op::bottom, b); };

      push(Op::entry, 0);
      rest(0, true);
      if(cfg.cycle == 1)
      {
        // F: one peak level after the other, from the coarsest but one up to the level below
        // the top (multigrid.hpp:1366-1406)
        for(std::size_t pk = b > 0 ? b - 1 : 0; pk > 0; --pk)
        {
          coarse();
          prol(pk, false);
          peak(pk);
          rest(pk, false);
        }
        coarse();
      }
      else if(cfg.cycle == 2)
      {
        // W: 2^(levels below the top) coarse solves; the next peak level is the lowest level
        // whose counter is 0, the counters below it are reset (multigrid.hpp:1409-1519)
        coarse();
        std::vector<int> cnt(b + 1, 0);
        const std::size_t num = std::size_t(1) << b;
        for(std::size_t c = 1; c < num; ++c)
        {
          std::size_t pk = b;
          while(pk > 0)
            if(cnt[--pk] == 0) break;
          for(std::size_t i = b - 1; i > pk; --i) cnt[i] = 0;
          ++cnt[pk];
          prol(pk, false);
          peak(pk);
          rest(pk, false);
          coarse();
        }
      }
      else
        coarse();
      prol(0, true);
      return o;
    }

    int build_cycle()
    {
      const std::vector<CycleOp> full = cycle_schedule();
      const std::size_t nl = L.size();
      cycle_ops.clear();
      tail_segs.clear();
      if(tail0 + 1 >= nl)
        cycle_ops = full;
      else
      {
        // levels tail0.. run inside k_tail: a segment starts after the restriction into tail0
        // (the tail takes tail0's first smoothing step itself) and ends before the prolongation
        // out of it; every step records which buffer of its level's pair it reads
        std::vector<u32> ops;
        std::vector<u32> tin(nl, 0);
        bool open = false;
        auto close = [&]() {
          TailSeg& g = tail_segs.back();
          g.count = u32(ops.size()) - g.first;
          g.out_buf = tin[tail0];
          open = false;
        };
        for(const CycleOp& c : full)
        {
          const bool in_tail = c.op != Op::entry && c.lvl >= tail0;
          if(!in_tail)
          {
            if(open) close();
            CycleOp u = c;
            if(u.op == Op::rest_step && u.lvl + 1 == tail0) u.op = Op::rest;
            cycle_ops.push_back(u);
            continue;
          }
          if(!open)
          {
            cycle_ops.push_back(CycleOp{Op::bottom, tail0, u32(tail_segs.size())});
            tail_segs.push_back(TailSeg{u32(ops.size()), 0, 0});
            tin[tail0] = 0;
            open = true;
          }
          const u32 l = u32(c.lvl - tail0);
          switch(c.op)
          {
            case Op::sweep: ops.push_back(mg::tail_op(mg::kTailSweep, l, tin[c.lvl], 0, 0)); tin[c.lvl] ^= 1u; break;
            case Op::defect: ops.push_back(mg::tail_op(mg::kTailDefect, l, tin[c.lvl], 0, 0)); break;
            case Op::rest: ops.push_back(mg::tail_op(mg::kTailRest, l, 0, 0, 0)); break;
            case Op::rest_step: ops.push_back(mg::tail_op(mg::kTailRest, l, 0, 0, 1)); tin[c.lvl + 1] = 0; break;
            case Op::bottom: ops.push_back(mg::tail_op(mg::kTailCoarse, l, 0, 0, 0)); tin[c.lvl] = 0; break;
            case Op::prol: ops.push_back(mg::tail_op(mg::kTailProl, l, tin[c.lvl], tin[c.lvl + 1], 0)); break;
            default: return rt::kErrState;
          }
        }
        if(open) close();
        F3_RC(tail_ops.alloc(sizeof(u32) * std::max<std::size_t>(ops.size(), 1)));
        if(!ops.empty())
        {
          F3_CHECK(hipMemcpyAsync(tail_ops.p, ops.data(), sizeof(u32) * ops.size(), hipMemcpyHostToDevice, s));
          F3_CHECK(hipStreamSynchronize(s));
        }
      }
      // the top level's iterate must end in its xfinal buffer, which the PCG reads: follow it
      // through the upper operations (a bottom operation on level 0 writes xfinal itself)
      int where = 0;   // 0: x, 1: t, 2: xfinal
      for(const CycleOp& c : cycle_ops)
      {
        if(c.lvl != 0) continue;
        if(c.op == Op::entry) where = 0;
        else if(c.op == Op::sweep && where != 2) where ^= 1;
        else if(c.op == Op::bottom) where = 2;
      }
      if(where != 2 && (where == 1) != (L[0]->xfinal == L[0]->t.template as<DT>()))
        cycle_ops.push_back(CycleOp{Op::copy_top, 0, 0});
      return rt::kOk;
    }

    // evs (statistics only): an event is recorded before every operation k with span_start[k]
    // set (span_start[0] is) and one after the last operation, so that each span of operations
    // lies between two consecutive events
    int enqueue_vcycle(const double* def, const int* status, const std::vector<hipEvent_t>* evs = nullptr,
                       const std::vector<char>* span_start = nullptr)
    {
      const std::size_t nl = L.size();
      std::vector<DT*> cur(nl, nullptr), oth(nl, nullptr);
      std::size_t k = 0, e = 0;
      for(const CycleOp& c : cycle_ops)
      {
        const std::size_t i = c.lvl;
        Level<DT, VT>& f = *L[i];
        if(evs && (*span_start)[k]) F3_CHECK(hipEventRecord((*evs)[e++], s));
        ++k;
        switch(c.op)
        {
          case Op::entry:
            mg::k_mg_entry<DT><<<stream_blocks(f.n), kBlock, 0, s>>>(f.n, def, f.rhs.template as<DT>(), f.x.template as<DT>(),
                                                                    f.dinv.template as<DT>(), status);
            cur[0] = f.x.template as<DT>();
            oth[0] = f.t.template as<DT>();
            break;
          case Op::sweep:
            F3_RC(sweep(f, cur[i], oth[i], status));
            std::swap(cur[i], oth[i]);
            break;
          case Op::defect:
            F3_RC(with_rows<DT, VT>(f.A, [&](auto A, mg::RowGrid g) {
              mg::k_defect<<<g.blocks, kBlock, 0, s>>>(A, g.per, f.mask.template as<u32>(), f.rhs.template as<DT>(),
                                                       (const DT*)cur[i], f.def.template as<DT>(), status);
            }));
            break;
          case Op::rest:
          case Op::rest_step:
          {
            Level<DT, VT>& cl = *L[i + 1];
            DT* xc = c.op == Op::rest_step ? cl.x.template as<DT>() : nullptr;
            F3_RC(with_transfer_rows<DT, VT>(f.R, [&](auto R, mg::RowGrid g) {
              mg::k_restrict<<<g.blocks, kBlock, 0, s>>>(R, g.per, cl.mask.template as<u32>(), f.def.template as<DT>(),
                                                         cl.rhs.template as<DT>(), xc, cl.dinv.template as<DT>(), status);
            }));
            if(xc)
            {
              cur[i + 1] = xc;
              oth[i + 1] = cl.t.template as<DT>();
            }
            break;
          }
          case Op::bottom:
            if(tail0 + 1 < nl)
            {
              const TailSeg& g = tail_segs[c.seg];
              mg::k_tail<DT, VT><<<1, mg::kTailThreads, tail_lds, s>>>(tail, tail_ops.template as<u32>() + g.first, g.count,
                                                                      g.out_buf, status);
            }
            else
              mg::k_coarse<DT><<<1, kBlock, 0, s>>>(f.n, ainv.template as<double>(), f.mask.template as<u32>(),
                                                    f.rhs.template as<DT>(), f.x.template as<DT>(), status);
            cur[i] = f.xfinal;
            break;
          case Op::prol:
            F3_RC(with_transfer_rows<DT, VT>(f.P, [&](auto P, mg::RowGrid g) {
              mg::k_prolong<<<g.blocks, kBlock, 0, s>>>(P, g.per, f.mask.template as<u32>(), (const DT*)cur[i + 1], cur[i],
                                                        status);
            }));
            break;
          case Op::copy_top:
            mg::k_copy_vec<DT><<<stream_blocks(f.n), kBlock, 0, s>>>(f.n, cur[0], f.xfinal, status);
            std::swap(cur[0], oth[0]);
            break;
        }
      }
      if(evs) F3_CHECK(hipEventRecord((*evs)[e], s));
      if(cur[0] != L[0]->xfinal) return rt::kErrState;
      return rt::check_launch();
    }

    // ---- statistics (MultiGridHierarchy timings, multigrid.hpp:456-479,858-956) ------------
    // Per level: time of smoothing, defect, grid transfer (restriction and prolongation, counted
    // on the finer level as the reference does) and coarse solve per cycle, the number of each
    // operation per cycle, and the cycle's algorithmic bytes and flops on that level.  Fused
    // operations count where their main work is: the entry (conversion + first step) and the top
    // copy as smoothing on level 0, a restriction with the next level's first smoothing step as
    // transfer (its step is counted as a smoothing step of the coarser level), and a segment of
    // the single-workgroup coarse tail as the coarse solve of the tail's top level (one kernel:
    // the levels inside it report counts, bytes and flops, their time is that of the segment).
    // Cycles are launched operation by operation, with an event wherever the level or the
    // category changes (a run of smoothing steps on one level is one span): an event costs the
    // stream a few microseconds, as much as a small level's operation, so one per operation
    // would inflate the coarse levels' times.  The solve itself is not touched (its graphs
    // record no events).
    std::vector<hipEvent_t> stat_ev;
    static int op_category(Op op)
    {
      switch(op)
      {
        case Op::entry: case Op::sweep: case Op::copy_top: return 0;   // smoothing
        case Op::defect: return 1;
        case Op::rest: case Op::rest_step: case Op::prol: return 2;     // grid transfer
        case Op::bottom: return 3;                                       // coarse solve
      }
      return 0;
    }
    int statistics(int reps, double* out, int cap) override
    {
      const std::size_t nl = L.size(), nops = cycle_ops.size();
      const int need = 3 + 10 * int(nl);
      if(reps < 1 || !out || cap < need) return rt::kErrArg;
      // spans: maximal runs of operations with the same level and category
      std::vector<char> span_start(nops, 0);
      std::vector<std::size_t> span_slot;   // 4 * level + category of each span
      for(std::size_t k = 0; k < nops; ++k)
      {
        const std::size_t slot = 4 * cycle_ops[k].lvl + std::size_t(op_category(cycle_ops[k].op));
        if(k == 0 || slot != span_slot.back())
        {
          span_start[k] = 1;
          span_slot.push_back(slot);
        }
      }
      const std::size_t nspan = span_slot.size();
      // timing-only events: without the system-scope fence (a cache writeback and invalidation
      // at every record) an event costs the stream less; the host synchronises the stream
      // before it reads the times
      while(stat_ev.size() < nspan + 1)
      {
        hipEvent_t e = nullptr;
        F3_CHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        stat_ev.push_back(e);
      }
      std::vector<double> t(4 * nl, 0.0), cnt(4 * nl, 0.0), by(nl, 0.0), fl(nl, 0.0);
      F3_RC(enqueue_vcycle(r.template as<double>(), nullptr));   // warm-up
      F3_CHECK(hipStreamSynchronize(s));
      double total = 0.0;
      for(int rep = 0; rep < reps; ++rep)
      {
        F3_RC(enqueue_vcycle(r.template as<double>(), nullptr, &stat_ev, &span_start));
        F3_CHECK(hipStreamSynchronize(s));
        float ms = 0.f;
        F3_CHECK(hipEventElapsedTime(&ms, stat_ev[0], stat_ev[nspan]));
        total += ms;
        for(std::size_t k = 0; k < nspan; ++k)
        {
          F3_CHECK(hipEventElapsedTime(&ms, stat_ev[k], stat_ev[k + 1]));
          t[span_slot[k]] += double(ms);
        }
      }
      // counts, bytes and flops of the cycle's mathematical operations (the full schedule, tail
      // levels included)
      const u64 vb = sizeof(DT);
      for(const CycleOp& c : cycle_schedule())
      {
        const std::size_t i = c.lvl;
        const Level<DT, VT>& f = *L[i];
        const double n = double(f.n), nz = double(f.A.nnz);
        const double nc = i + 1 < nl ? double(L[i + 1]->n) : 0.0;
        switch(c.op)
        {
          case Op::entry: cnt[4 * i] += 1; by[i] += n * (8 + 3 * vb); fl[i] += n; break;
          case Op::sweep:
            cnt[4 * i] += 1; by[i] += double(f.A.matrix_bytes()) + n * vb * 4; fl[i] += 2 * nz + 3 * n; break;
          case Op::defect:
            cnt[4 * i + 1] += 1; by[i] += double(f.A.matrix_bytes()) + n * vb * 3; fl[i] += 2 * nz + n; break;
          case Op::rest: case Op::rest_step:
            cnt[4 * i + 2] += 1;
            by[i] += double(f.R.matrix_bytes()) + n * vb + nc * vb * (c.op == Op::rest ? 1 : 3);
            fl[i] += 2 * double(f.R.nnz);
            if(c.op == Op::rest_step) { cnt[4 * (i + 1)] += 1; fl[i + 1] += nc; }
            break;
          case Op::prol:
            cnt[4 * i + 2] += 1; by[i] += double(f.P.matrix_bytes()) + nc * vb + n * vb * 2; fl[i] += 2 * double(f.P.nnz) + n;
            break;
          case Op::bottom:
            cnt[4 * i + 3] += 1; by[i] += double(ncoarse * ncoarse * 8) + n * vb * 2; fl[i] += 2 * double(ncoarse * ncoarse);
            break;
          case Op::copy_top: break;
        }
      }
      out[0] = double(nl);
      out[1] = total / reps;
      out[2] = double(nops);
      for(std::size_t i = 0; i < nl; ++i)
      {
        double* o = out + 3 + 10 * i;
        for(int c = 0; c < 4; ++c) { o[c] = t[4 * i + c] / reps; o[4 + c] = cnt[4 * i + c]; }
        o[8] = by[i];
        o[9] = fl[i];
      }
      return rt::kOk;
    }

    int sweep(Level<DT, VT>& f, DT* xin, DT* xout, const int* status)
    {
      return with_rows<DT, VT>(f.A, [&](auto A, mg::RowGrid g) {
        mg::k_sweep<<<g.blocks, kBlock, 0, s>>>(A, g.per,f.dinv.template as<DT>(), f.rhs.template as<DT>(),
                                                    (const DT*)xin, xout, status);
      });
    }

    const DT* z_ptr() const { return L.front()->xfinal; }

    int vcycle(double* cor, const double* def) override
    {
      F3_RC(enqueue_vcycle(def, nullptr));
      // cor = (double) z
      k_to_double<<<stream_blocks(n0), kBlock, 0, s>>>(n0, z_ptr(), cor);
      F3_RC(rt::check_launch());
      F3_CHECK(hipStreamSynchronize(s));
      return rt::kOk;
    }

    // ---- PCG ------------------------------------------------------------------------------
    int enqueue_precond(const int* status)
    {
      if(cfg.precond == 0)
        return enqueue_vcycle(r.template as<double>(), status);
      mg::k_copy_filter<<<stream_blocks(n0), kBlock, 0, s>>>(n0, r.template as<double>(), mask0_ptr(), zcg.template as<double>(),
                                                            status);
      return rt::check_launch();
    }
    const u32* mask0_ptr() const { return L.front()->mask.template as<u32>(); }

    template<typename F>
    int with_z(F&& f)
    {
      if(cfg.precond == 0) f(z_ptr());
      else f((const double*)zcg.template as<double>());
      return rt::kOk;
    }

    // one-block finalisation of a reduction whose producing kernel ran with nb blocks
    template<int OP>
    void fin(unsigned nb)
    {
      mg::k_pcg_fin<OP><<<1, kBlock, 0, s>>>(partials.template as<double>(), nb, st(), hist.template as<double>());
    }

    // outer half of a PCG iteration (pcg.hpp:236-262): q = A p, alpha, x += alpha p,
    // r -= alpha q, the new defect and the stopping criterion
    int enqueue_outer_half()
    {
      mg::PcgState* S = st();
      double* part = partials.template as<double>();
      unsigned nb = 0;
      F3_RC(with_rows<double, double>(outer_matrix(), [&](auto A, mg::RowGrid g) {
        mg::k_pcg_spmv_dot<<<g.blocks, kBlock, 0, s>>>(A, g.per, mask0_ptr(), p.template as<double>(), q.template as<double>(),
                                                       S, part);
        nb = g.blocks;
      }));
      fin<mg::kFinAlpha>(nb);
      mg::k_pcg_update_norm<<<stream_blocks(n0), kBlock, 0, s>>>(n0, p.template as<double>(), q.template as<double>(),
                                                                x_cur, r.template as<double>(), S, part);
      fin<mg::kFinNewDefect>(stream_blocks(n0));
      return rt::check_launch();
    }

    // preconditioner half (pcg.hpp:264-285): z = M r, gamma, beta, p = z + beta p
    int enqueue_precond_half()
    {
      mg::PcgState* S = st();
      double* part = partials.template as<double>();
      F3_RC(enqueue_precond(cfg.vcycle_status ? status_ptr() : nullptr));
      F3_RC(with_z([&](auto z) {
        mg::k_pcg_gamma_beta<<<stream_blocks(n0), kBlock, 0, s>>>(n0, z, r.template as<double>(), S, part);
        fin<mg::kFinBeta>(stream_blocks(n0));
        mg::k_pcg_update_p<<<stream_blocks(n0), kBlock, 0, s>>>(n0, z, p.template as<double>(), S);
      }));
      return rt::check_launch();
    }

    double* x_cur = nullptr;   // the solution vector of the running solve
    double* x_graph = nullptr; // the one the captured graph was built for

    // the preconditioner half of iteration k, the outer half of iteration k + 1 and the
    // publication of its state: one graph
    int enqueue_iteration()
    {
      F3_RC(enqueue_precond_half());
      F3_RC(enqueue_outer_half());
      return enqueue_publish();
    }
    int enqueue_publish()
    {
      mg::k_publish<kRing><<<1, 64, 0, s>>>(st(), ring_dev->st, ring_dev->seq, pub_count.template as<unsigned>());
      return rt::check_launch();
    }

    int solve(double* x, const double* b, bool correct, int* status, u64* iters,
              const std::function<int()>* epilogue) override
    {
      if(!x || !b) return rt::kErrArg;
      mg::PcgState h{};
      h.tol_rel = cfg.tol_rel; h.tol_abs = cfg.tol_abs; h.tol_abs_low = cfg.tol_abs_low;
      h.div_rel = cfg.div_rel; h.div_abs = cfg.div_abs; h.stag_rate = cfg.stag_rate;
      h.min_iter = cfg.min_iter; h.max_iter = cfg.max_iter; h.min_stag_iter = cfg.min_stag_iter;
      h.hist_cap = hist_cap; h.status = mg::kUndefined;
      F3_CHECK(hipMemcpyAsync(state.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
      F3_CHECK(hipMemsetAsync(pub_count.p, 0, sizeof(unsigned), s));
      double* part = partials.template as<double>();
      x_cur = x;
      if(cfg.use_graph && (gexec == nullptr || x_graph != x))
      {
        drop_graphs();
        F3_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_iteration();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if(rc) { if(g) (void)hipGraphDestroy(g); return rc; }
        if(e != hipSuccess) return rt::kErrRuntime;
        graph = g;
        F3_CHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        x_graph = x;
      }
      if(correct)
      {
        unsigned nb = 0;
        F3_RC(with_rows<double, double>(outer_matrix(), [&](auto A, mg::RowGrid g) {
          mg::k_pcg_defect_norm<<<g.blocks, kBlock, 0, s>>>(A, g.per, mask0_ptr(), b, (const double*)x, r.template as<double>(), part);
          nb = g.blocks;
        }));
        fin<mg::kFinInitDefect>(nb);
      }
      else
      {
        // apply(cor, def): r = def, cor = 0
        F3_CHECK(hipMemsetAsync(x, 0, sizeof(double) * n0, s));
        mg::k_pcg_copy_norm<<<stream_blocks(n0), kBlock, 0, s>>>(n0, b, r.template as<double>(), part);
        fin<mg::kFinInitDefect>(stream_blocks(n0));
      }
      F3_RC(rt::check_launch());
      F3_RC(enqueue_publish());   // publication 1: the initial defect

      // publications launched and read (in order); the host reads publication j from slot
      // (j - 1) % kRing once its sequence number reads j
      unsigned launched = 1, read = 0;
      std::vector<double> seen_def;   // defects of iterations 0..read-1
      int seen = mg::kProgress;
      mg::PcgState last{};
      auto read_state = [&]() -> int {
        const unsigned want = read + 1, slot = (want - 1) % kRing;
        for(unsigned spin = 1; __atomic_load_n(&ring->seq[slot], __ATOMIC_ACQUIRE) != want; ++spin)
          if(spin % 256 == 0)
          {
            // the wait ends: once the stream has drained, the publication has arrived or is lost
            const hipError_t q = hipStreamQuery(s);
            if(q == hipSuccess)
            {
              if(__atomic_load_n(&ring->seq[slot], __ATOMIC_ACQUIRE) == want) break;
              return rt::kErrRuntime;
            }
            if(q != hipErrorNotReady) return rt::kErrRuntime;
          }
        last = ring->st[slot];
        seen = last.status;
        seen_def.push_back(last.def_cur);
        read = want;
        return rt::kOk;
      };
      auto launch_iteration = [&]() -> int {
        if(cfg.use_graph) F3_CHECK(hipGraphLaunch(gexec, s));
        else F3_RC(enqueue_iteration());
        ++launched;
        return rt::kOk;
      };
      // Host loop.  Far from convergence the next iteration is launched while the previous one
      // runs (the host reads states one publication behind).  When the defect history read so far
      // predicts that the outer half just launched may meet the criterion, its state is read
      // before the next graph (whose kernels would all return at once) is launched.
      const u64 max_launch = cfg.max_iter + 1;
      u64 k = 0;          // outer halves launched
      bool wait = false;  // the latest outer half may meet the criterion
      while(k < max_launch && seen == mg::kProgress)
      {
        if(wait)
        {
          while(read < launched && seen == mg::kProgress) F3_RC(read_state());
          if(seen != mg::kProgress) break;
        }
        F3_RC(launch_iteration());
        ++k;
        wait = k >= cfg.max_iter;
        if(seen_def.size() >= 2)
        {
          const std::size_t j = seen_def.size() - 1;   // latest iteration read
          const double d0 = seen_def[0], dj = seen_def[j], dp = seen_def[j - 1];
          const double rate = dp > 0.0 ? dj / dp : 1.0;
          const double goal = std::min(cfg.tol_abs, std::max(cfg.tol_rel * d0, cfg.tol_abs_low));
          const double pred = dj * std::pow(std::min(rate, 1.0), double(k - j));
          wait = wait || pred <= 10.0 * goal;
        }
        if(!wait)
          while(read + 1 < launched && seen == mg::kProgress) F3_RC(read_state());
      }
      if(epilogue) F3_RC((*epilogue)());
      while(read < launched) F3_RC(read_state());   // the last publication is the final state
      F3_CHECK(hipStreamSynchronize(s));
      if(status) *status = last.status;
      if(iters) *iters = last.num_iter;
      last_hist_len = std::min<u64>(last.num_iter + 1, hist_cap);
      return rt::kOk;
    }

    u64 last_hist_len = 0;

    int history(double* out, u64 cap, u64* len) override
    {
      const u64 m = std::min(cap, last_hist_len);
      if(m) F3_CHECK(hipMemcpy(out, hist.p, sizeof(double) * m, hipMemcpyDeviceToHost));
      if(len) *len = last_hist_len;
      return rt::kOk;
    }

    // times `reps` launches of one kernel of the hot path between two events on the engine's
    // stream: which = 0 fine-level smoothing sweep, 1 fine-level defect, 2 outer SpMV+dot,
    // 3 one full V-cycle, 4 one full PCG iteration (graph)
    int time_kernel(int which, int reps, double* avg_ms, u64* bytes) override
    {
      if(reps < 1) return rt::kErrArg;
      Level<DT, VT>& f = *L.front();
      // warm-up launch, then the timed region
      for(int pass = 0; pass < 2; ++pass)
      {
        const int nrep = pass == 0 ? 1 : reps;
        if(pass == 1) F3_CHECK(hipEventRecord(t0, s));
        for(int k = 0; k < nrep; ++k)
        {
          switch(which)
          {
            case 0: F3_RC(sweep(f, f.x.template as<DT>(), f.t.template as<DT>(), nullptr)); break;
            case 1:
              F3_RC(with_rows<DT, VT>(f.A, [&](auto A, mg::RowGrid g) {
                mg::k_defect<<<g.blocks, kBlock, 0, s>>>(A, g.per,f.mask.template as<u32>(), f.rhs.template as<DT>(),
                                                             (const DT*)f.x.template as<DT>(), f.def.template as<DT>(), nullptr);
              }));
              break;
            case 2:
            {
              // the outer FP64 product q = filter(A p) with its block partials of q.p, on a
              // state of its own that says "in progress"
              if(!timing_state.p)
              {
                mg::PcgState h{};
                h.status = mg::kProgress;
                F3_RC(timing_state.alloc(sizeof(mg::PcgState)));
                F3_CHECK(hipMemcpyAsync(timing_state.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
              }
              F3_RC(with_rows<double, double>(outer_matrix(), [&](auto A, mg::RowGrid g) {
                mg::k_pcg_spmv_dot<<<g.blocks, kBlock, 0, s>>>(A, g.per, mask0_ptr(), p.template as<double>(),
                                                               q.template as<double>(), timing_state.template as<mg::PcgState>(),
                                                               partials.template as<double>());
              }));
              break;
            }
            case 3: F3_RC(enqueue_vcycle(r.template as<double>(), nullptr)); break;
            default: return rt::kErrArg;
          }
        }
        if(pass == 1) F3_CHECK(hipEventRecord(t1, s));
        F3_CHECK(hipStreamSynchronize(s));
      }
      float ms = 0.f;
      F3_CHECK(hipEventElapsedTime(&ms, t0, t1));
      *avg_ms = double(ms) / reps;
      if(bytes)
      {
        const u64 vb = sizeof(DT);
        // sweep: matrix + x gathered once + rhs + dinv + x_out; defect: matrix + x + rhs + def (+mask)
        if(which == 0) *bytes = f.A.matrix_bytes() + f.n * vb * 4;
        else if(which == 1) *bytes = f.A.matrix_bytes() + f.n * vb * 3 + (f.n + 31) / 32 * 4;
        else if(which == 2) *bytes = outer_matrix().matrix_bytes() + n0 * 8 * 2 + (n0 + 31) / 32 * 4;
        else *bytes = 0;
      }
      return rt::kOk;
    }

    // info: [0] levels, [1] n finest, [2] nnz finest, [3] lanes finest, [4] sell size finest,
    // [5] kernels per V-cycle, [6] algorithmic bytes of one V-cycle, [7] bytes of one outer
    // iteration excluding the V-cycle, [8] sizeof inner DT, [9] sizeof VT
    int info(u64* out, int cap) override
    {
      std::vector<u64> v(17, 0);
      v[0] = L.size();
      v[1] = L.front()->n;
      v[2] = L.front()->A.nnz;
      v[3] = u64(L.front()->A.lanes);
      v[4] = L.front()->A.sell_size;
      u64 kern = 0, bytes = 0;
      const u64 vb = sizeof(DT);
      for(const CycleOp& c : cycle_ops)
      {
        const Level<DT, VT>& f = *L[c.lvl];
        const u64 nc = c.lvl + 1 < L.size() ? L[c.lvl + 1]->n : 0;
        ++kern;
        switch(c.op)
        {
          case Op::entry: bytes += f.n * (8 + vb * 3); break;
          case Op::sweep: bytes += f.A.matrix_bytes() + f.n * vb * 4; break;
          case Op::defect: bytes += f.A.matrix_bytes() + f.n * vb * 3; break;
          case Op::rest: bytes += f.R.matrix_bytes() + f.n * vb + nc * vb * 2; break;
          case Op::rest_step: bytes += f.R.matrix_bytes() + f.n * vb + nc * vb * 4; break;
          case Op::bottom: bytes += ncoarse * ncoarse * 8 + ncoarse * vb * 2; break;
          case Op::prol: bytes += f.P.matrix_bytes() + nc * vb + f.n * vb * 2; break;
          case Op::copy_top: bytes += f.n * vb * 2; break;
        }
      }
      v[5] = kern;
      v[6] = bytes;
      v[7] = outer_matrix().matrix_bytes() + n0 * 8 * 2 + n0 * 8 * 5 + n0 * (8 + vb) * 2 + n0 * (8 + vb * 2);
      v[8] = sizeof(DT);
      v[9] = sizeof(VT);
      v[10] = L.front()->A.sell_d16 ? 1 : 0;
      v[11] = L.front()->A.sell ? (L.front()->n + L.front()->A.C - 1) / L.front()->A.C : 0;
      v[12] = L.front()->A.banded ? (L.front()->A.dict ? 3 : (L.front()->A.bsym ? 2 : 1)) : 0;
      v[13] = L.front()->A.ndiag;
      v[14] = tail0 + 1 < L.size() ? L.size() - tail0 : 0;
      v[15] = tail_lds;
      for(const auto& l : L) v[16] += l->P.tensor ? 1 : 0;
      for(int i = 0; i < cap && i < 17; ++i) out[i] = v[i];
      return rt::kOk;
    }
  };
} // namespace

// =========================================================================================
// C ABI
// =========================================================================================
struct feat3_mgpcg
{
  int inner = 0;     // 0 FP64, 1 FP32, 2 FP16-stored values with FP32 arithmetic
  Config cfg;
  std::vector<HostLevel> levels;
  std::unique_ptr<EngineBase> eng;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // the solver's internal DOF order per level (feat3_mgpcg_set_level_order; empty: the caller's
  // numbering); applied to the level data at the first init after it was set
  std::vector<std::vector<u32>> orders;
  std::vector<char> pending;
  // level 0's order on the device and the internal FP64 vectors b, x of a permuted solve
  DevMem order0, xi, bi;
  u64 n0 = 0;
  bool permuted0 = false;
};

namespace
{
  // old index i -> new index order[i] (NULL: identity) of a host CSR matrix's rows and columns
  int permute_host_csr(HostCsr& m, const u32* row_order, const u32* col_order)
  {
    if(!row_order && !col_order) return rt::kOk;
    const u64 nnz = m.nnz();
    std::vector<u32> rp(m.rows + 1), col(nnz);
    std::vector<double> val(nnz);
    F3_RC(feat3_setup_permute_csr(m.rows, m.cols, m.rp.data(), m.col.data(), m.val.data(), row_order, col_order,
                                  rp.data(), col.data(), val.data()));
    m.rp.swap(rp); m.col.swap(col); m.val.swap(val);
    return rt::kOk;
  }

  // the same for a device-resident source: the permuted copy goes into store (kept until init
  // returns, like the caller's arrays it replaces)
  int permute_dev_csr(HostCsr& m, const u32* drow, const u32* dcol_order, hipStream_t s, std::deque<DevMem>& store)
  {
    if(!drow && !dcol_order) return rt::kOk;
    store.emplace_back(); DevMem& rp = store.back();
    store.emplace_back(); DevMem& col = store.back();
    store.emplace_back(); DevMem& val = store.back();
    F3_RC(rp.alloc(sizeof(u32) * (m.rows + 1)));
    F3_RC(col.alloc(sizeof(u32) * std::max<u64>(m.dnnz, 1)));
    F3_RC(val.alloc(sizeof(double) * std::max<u64>(m.dnnz, 1)));
    F3_RC(feat3_device_permute_csr(s, m.rows, m.cols, m.drp, m.dcol, m.dval, drow, dcol_order, rp.as<u32>(),
                                   col.as<u32>(), val.as<double>()));
    m.drp = rp.as<u32>(); m.dcol = col.as<u32>(); m.dval = val.as<double>();
    return rt::kOk;
  }

  // applies the pending level orders: A_l rows and columns by order_l, the Dirichlet list mapped
  // and sorted, P_l rows by order_l and columns by order_{l+1}, R_l the other way round
  int apply_orders(feat3_mgpcg* h, std::deque<DevMem>& store)
  {
    const std::size_t nl = h->levels.size();
    h->orders.resize(nl);
    h->pending.resize(nl, 0);
    bool any = false;
    for(std::size_t l = 0; l < nl; ++l) any = any || h->pending[l];
    if(!any) return rt::kOk;
    // device copies of the orders that device-resident levels need
    std::deque<DevMem> dord(nl);
    auto need_dev = [&](std::size_t l) {
      return h->levels[l].A.on_device() || (l > 0 && h->levels[l - 1].P.on_device());
    };
    for(std::size_t l = 0; l < nl; ++l)
      if(h->pending[l] && need_dev(l))
      {
        F3_RC(dord[l].alloc(sizeof(u32) * h->orders[l].size()));
        F3_CHECK(hipMemcpyAsync(dord[l].p, h->orders[l].data(), sizeof(u32) * h->orders[l].size(),
                                hipMemcpyHostToDevice, h->stream));
      }
    auto hord = [&](std::size_t l) -> const u32* { return l < nl && h->pending[l] ? h->orders[l].data() : nullptr; };
    auto dor = [&](std::size_t l) -> const u32* { return l < nl && h->pending[l] ? dord[l].as<u32>() : nullptr; };
    for(std::size_t l = 0; l < nl; ++l)
    {
      HostLevel& L = h->levels[l];
      if(L.A.on_device()) F3_RC(permute_dev_csr(L.A, dor(l), dor(l), h->stream, store));
      else F3_RC(permute_host_csr(L.A, hord(l), hord(l)));
      if(hord(l))
      {
        for(u32& i : L.fidx) i = h->orders[l][i];
        std::sort(L.fidx.begin(), L.fidx.end());
      }
      if(l + 1 < nl && L.P.rows)
      {
        if(L.P.on_device())
        {
          F3_RC(permute_dev_csr(L.P, dor(l), dor(l + 1), h->stream, store));
          F3_RC(permute_dev_csr(L.R, dor(l + 1), dor(l), h->stream, store));
        }
        else
        {
          F3_RC(permute_host_csr(L.P, hord(l), hord(l + 1)));
          F3_RC(permute_host_csr(L.R, hord(l + 1), hord(l)));
        }
      }
    }
    F3_CHECK(hipStreamSynchronize(h->stream));
    if(h->pending[0])
    {
      h->n0 = h->levels[0].A.rows;
      F3_RC(h->order0.alloc(sizeof(u32) * h->n0));
      F3_RC(h->xi.alloc(sizeof(double) * h->n0));
      F3_RC(h->bi.alloc(sizeof(double) * h->n0));
      F3_CHECK(hipMemcpy(h->order0.p, h->orders[0].data(), sizeof(u32) * h->n0, hipMemcpyHostToDevice));
      h->permuted0 = true;
    }
    for(std::size_t l = 0; l < nl; ++l) h->pending[l] = 0;
    return rt::kOk;
  }

  // caller vector -> internal order (b, x at entry) and back (x at exit), on the solver's stream
  int to_internal(feat3_mgpcg* h, const double* src, double* dst)
  {
    return feat3_device_permute_vector(h->stream, h->n0, h->order0.as<u32>(), 0, src, dst);
  }
  int to_caller_async(feat3_mgpcg* h, const double* src, double* dst)
  {
    return feat3_device_permute_vector(h->stream, h->n0, h->order0.as<u32>(), 1, src, dst);
  }
  int to_caller(feat3_mgpcg* h, const double* src, double* dst)
  {
    F3_RC(feat3_device_permute_vector(h->stream, h->n0, h->order0.as<u32>(), 1, src, dst));
    F3_CHECK(hipStreamSynchronize(h->stream));
    return rt::kOk;
  }
} // namespace

extern "C" {

int feat3_mgpcg_create(feat3_mgpcg** h, int inner_precision, int format)
{
  if(!h || inner_precision < 0 || inner_precision > 2 || format < 0 || format > 2) return rt::kErrArg;
  std::unique_ptr<feat3_mgpcg> o(new feat3_mgpcg());
  o->inner = inner_precision;
  o->cfg.format = format;
  if(hipStreamCreateWithFlags(&o->stream, hipStreamNonBlocking) != hipSuccess) return rt::kErrRuntime;
  o->own_stream = true;
  *h = o.release();
  return rt::kOk;
}

int feat3_mgpcg_destroy(feat3_mgpcg* h)
{
  if(!h) return rt::kOk;
  if(h->stream) (void)hipStreamSynchronize(h->stream);
  h->eng.reset();
  if(h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return rt::kOk;
}

// push one level (finest first); host arrays are copied.  The transfer arrays describe the
// prolongation P (n x nc) and restriction R (nc x n) to the next coarser level and are
// ignored (may be NULL) for the coarsest level.
int feat3_mgpcg_push_level(feat3_mgpcg* h, u64 n, const u32* row_ptr, const u32* col_ind, const double* val,
                           u64 nfilt, const u32* filter_idx, u64 nc, const u32* p_row_ptr, const u32* p_col_ind,
                           const double* p_val, const u32* r_row_ptr, const u32* r_col_ind, const double* r_val)
{
  if(!h || !row_ptr || (n > 0 && (!col_ind || !val))) return rt::kErrArg;
  HostLevel L;
  L.A.rows = n; L.A.cols = n;
  L.A.rp.assign(row_ptr, row_ptr + n + 1);
  const u64 nnz = row_ptr[n];
  L.A.col.assign(col_ind, col_ind + nnz);
  L.A.val.assign(val, val + nnz);
  for(u64 k = 0; k < nnz; ++k) if(col_ind[k] >= n) return rt::kErrArg;
  if(nfilt && !filter_idx) return rt::kErrArg;
  L.fidx.assign(filter_idx, filter_idx + nfilt);
  for(u32 i : L.fidx) if(i >= n) return rt::kErrArg;
  if(nc > 0)
  {
    if(!p_row_ptr || !p_col_ind || !p_val || !r_row_ptr || !r_col_ind || !r_val) return rt::kErrArg;
    L.P.rows = n; L.P.cols = nc;
    L.P.rp.assign(p_row_ptr, p_row_ptr + n + 1);
    L.P.col.assign(p_col_ind, p_col_ind + p_row_ptr[n]);
    L.P.val.assign(p_val, p_val + p_row_ptr[n]);
    L.R.rows = nc; L.R.cols = n;
    L.R.rp.assign(r_row_ptr, r_row_ptr + nc + 1);
    L.R.col.assign(r_col_ind, r_col_ind + r_row_ptr[nc]);
    L.R.val.assign(r_val, r_val + r_row_ptr[nc]);
    for(u32 c : L.P.col) if(c >= nc) return rt::kErrArg;
    for(u32 c : L.R.col) if(c >= n) return rt::kErrArg;
  }
  h->levels.push_back(std::move(L));
  h->eng.reset();
  return rt::kOk;
}

// the same level from device arrays (kept by the caller until feat3_mgpcg_init returns); levels
// of at most 65536 rows are copied to the host here (the coarse solve and the single-workgroup
// tail are set up there), larger ones are converted on the device by init
int feat3_mgpcg_push_level_device(feat3_mgpcg* h, u64 n, const u32* row_ptr, const u32* col_ind, const double* val,
                                  u64 nfilt, const u32* filter_idx, u64 nc, const u32* p_row_ptr, const u32* p_col_ind,
                                  const double* p_val, const u32* r_row_ptr, const u32* r_col_ind, const double* r_val)
{
  if(!h || !row_ptr || n == 0 || !col_ind || !val) return rt::kErrArg;
  if(nfilt && !filter_idx) return rt::kErrArg;
  if(nc > 0 && (!p_row_ptr || !p_col_ind || !p_val || !r_row_ptr || !r_col_ind || !r_val)) return rt::kErrArg;
  auto last = [&](const u32* rp, u64 rows, u64& out) -> int {
    u32 v = 0;
    F3_CHECK(hipMemcpy(&v, rp + rows, sizeof(u32), hipMemcpyDeviceToHost));
    out = v;
    return rt::kOk;
  };
  u64 nnz = 0, pn = 0, rn = 0;
  F3_RC(last(row_ptr, n, nnz));
  if(nc > 0) { F3_RC(last(p_row_ptr, n, pn)); F3_RC(last(r_row_ptr, nc, rn)); }
  std::vector<u32> fidx(nfilt);
  if(nfilt) F3_CHECK(hipMemcpy(fidx.data(), filter_idx, sizeof(u32) * nfilt, hipMemcpyDeviceToHost));
  if(n <= 65536)
  {
    // small: through the host path
    std::vector<u32> rp(n + 1), ci(nnz), prp, pci, rrp, rci;
    std::vector<double> v(nnz), pv, rv;
    F3_CHECK(hipMemcpy(rp.data(), row_ptr, sizeof(u32) * (n + 1), hipMemcpyDeviceToHost));
    if(nnz)
    {
      F3_CHECK(hipMemcpy(ci.data(), col_ind, sizeof(u32) * nnz, hipMemcpyDeviceToHost));
      F3_CHECK(hipMemcpy(v.data(), val, sizeof(double) * nnz, hipMemcpyDeviceToHost));
    }
    if(nc > 0)
    {
      prp.resize(n + 1); pci.resize(pn); pv.resize(pn); rrp.resize(nc + 1); rci.resize(rn); rv.resize(rn);
      F3_CHECK(hipMemcpy(prp.data(), p_row_ptr, sizeof(u32) * (n + 1), hipMemcpyDeviceToHost));
      F3_CHECK(hipMemcpy(rrp.data(), r_row_ptr, sizeof(u32) * (nc + 1), hipMemcpyDeviceToHost));
      if(pn)
      {
        F3_CHECK(hipMemcpy(pci.data(), p_col_ind, sizeof(u32) * pn, hipMemcpyDeviceToHost));
        F3_CHECK(hipMemcpy(pv.data(), p_val, sizeof(double) * pn, hipMemcpyDeviceToHost));
      }
      if(rn)
      {
        F3_CHECK(hipMemcpy(rci.data(), r_col_ind, sizeof(u32) * rn, hipMemcpyDeviceToHost));
        F3_CHECK(hipMemcpy(rv.data(), r_val, sizeof(double) * rn, hipMemcpyDeviceToHost));
      }
    }
    if(rp[0] != 0 || rp[n] != nnz) return rt::kErrArg;
    for(u64 r = 0; r < n; ++r) if(rp[r] > rp[r + 1]) return rt::kErrArg;
    if(nc > 0)
    {
      if(prp[0] != 0 || rrp[0] != 0) return rt::kErrArg;
      for(u64 r = 0; r < n; ++r) if(prp[r] > prp[r + 1]) return rt::kErrArg;
      for(u64 r = 0; r < nc; ++r) if(rrp[r] > rrp[r + 1]) return rt::kErrArg;
    }
    return feat3_mgpcg_push_level(h, n, rp.data(), ci.data(), v.data(), nfilt, fidx.data(), nc, prp.data(), pci.data(),
                                  pv.data(), rrp.data(), rci.data(), rv.data());
  }
  for(u32 i : fidx) if(i >= n) return rt::kErrArg;
  HostLevel L;
  L.A.rows = n; L.A.cols = n;
  L.A.drp = row_ptr; L.A.dcol = col_ind; L.A.dval = val; L.A.dnnz = nnz;
  L.fidx = std::move(fidx);
  if(nc > 0)
  {
    L.P.rows = n; L.P.cols = nc;
    L.P.drp = p_row_ptr; L.P.dcol = p_col_ind; L.P.dval = p_val; L.P.dnnz = pn;
    L.R.rows = nc; L.R.cols = n;
    L.R.drp = r_row_ptr; L.R.dcol = r_col_ind; L.R.dval = r_val; L.R.dnnz = rn;
  }
  h->levels.push_back(std::move(L));
  h->eng.reset();
  return rt::kOk;
}

int feat3_mgpcg_set_smoother(feat3_mgpcg* h, double omega_jacobi, double omega_richardson, int pre, int post)
{
  if(!h || pre < 1 || post < 1 || !(omega_jacobi > 0.0)) return rt::kErrArg;
  h->cfg.omega_j = omega_jacobi; h->cfg.omega_r = omega_richardson; h->cfg.pre = pre; h->cfg.post = post;
  h->eng.reset();
  return rt::kOk;
}

int feat3_mgpcg_set_tolerances(feat3_mgpcg* h, double tol_rel, double tol_abs, double tol_abs_low, u64 min_iter, u64 max_iter)
{
  if(!h || max_iter == 0) return rt::kErrArg;
  h->cfg.tol_rel = tol_rel;
  if(tol_abs > 0) h->cfg.tol_abs = tol_abs;
  h->cfg.tol_abs_low = tol_abs_low;
  h->cfg.min_iter = min_iter;
  h->cfg.max_iter = max_iter;
  // an initialised solver keeps running unless its history buffer is too small
  if(h->eng && !h->eng->set_tolerances(h->cfg)) h->eng.reset();
  return rt::kOk;
}

// 0: MG V-cycle preconditioner, 1: none (plain CG)
int feat3_mgpcg_set_precond(feat3_mgpcg* h, int kind)
{
  if(!h || kind < 0 || kind > 1) return rt::kErrArg;
  h->cfg.precond = kind;
  h->eng.reset();
  return rt::kOk;
}

// graph: replay the captured iteration (1) or launch kernel by kernel (0);
// lanes: lanes per CSR row for all levels (0 = the streaming evaluator);
// sell_C / sell_sigma: chunk height and sort window of the SELL format (format 1)
int feat3_mgpcg_set_launch(feat3_mgpcg* h, int graph, int lanes, u64 sell_C, u64 sell_sigma)
{
  if(!h) return rt::kErrArg;
  if(lanes != 0 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64) return rt::kErrArg;
  if(sell_C == 0 || sell_C > 1024 || sell_sigma == 0) return rt::kErrArg;
  h->cfg.use_graph = graph != 0;
  h->cfg.lanes = lanes;
  h->cfg.sell_C = sell_C;
  h->cfg.sell_sigma = sell_sigma;
  h->eng.reset();
  return rt::kOk;
}

// named options: "index16" (0/1), "compact_transfer" (0/1), "sell_C", "sell_sigma", "lanes",
// "tail_max_n", "rows_per_block" (level matrices; 0 = derived from the row count),
// "outer_rows_per_block" (the FP64 outer matrix's copy; 0 = derived), "lane_shift" (0/1),
// "transfer_sigma" (SELL sort window of P and R), "transfer_csr" (P and R in CSR, default 1)
int feat3_mgpcg_set_option(feat3_mgpcg* h, const char* name, double value)
{
  if(!h || !name) return rt::kErrArg;
  const std::string k(name);
  if(k == "index16") h->cfg.index16 = value != 0.0;
  else if(k == "compact_transfer") h->cfg.compact_transfer = value != 0.0;
  else if(k == "sell_C" && value >= 1 && value <= 1024) h->cfg.sell_C = u64(value);
  else if(k == "tail_max_n" && value >= 0) h->cfg.tail_max_n = u64(value);
  else if(k == "transfer_csr") h->cfg.transfer_csr = value != 0.0;
  else if(k == "tensor_transfer") h->cfg.tensor_transfer = value != 0.0;
  else if(k == "vcycle_status") h->cfg.vcycle_status = value != 0.0;
  else if(k == "transfer_lanes" && (value == -1 || value == 0 || value == 2 || value == 4 || value == 8 || value == 16 ||
                                    value == 32 || value == 64))
    h->cfg.transfer_lanes = int(value);
  else if(k == "transfer_rows_per_block" && value >= 0) h->cfg.transfer_rows_per_block = u64(value);
  else if(k == "symmetric_band") h->cfg.symmetric_band = value != 0.0;
  else if(k == "band_dict") h->cfg.band_dict = value != 0.0;
  else if(k == "lane_shift") h->cfg.lane_shift = value != 0.0;
  else if(k == "plane_order" && (value == 0 || value == 1 || value == 2)) h->cfg.plane_order = int(value);
  else if(k == "band_stage_x" && (value == 0 || value == 1 || value == 2)) h->cfg.band_stage_x = int(value);
  else if(k == "cycle" && (value == 0 || value == 1 || value == 2)) h->cfg.cycle = int(value);
  else if(k == "rows_per_block" && value >= 0) h->cfg.rows_per_block = u64(value);
  else if(k == "outer_rows_per_block" && value >= 0) h->cfg.outer_rows_per_block = u64(value);
  else if(k == "transfer_sigma" && value >= 1) h->cfg.transfer_sigma = u64(value);
  else if(k == "sell_sigma" && value >= 1) h->cfg.sell_sigma = u64(value);
  else if(k == "lanes" && (value == 0 || value == 2 || value == 4 || value == 8 || value == 16 || value == 32 || value == 64))
    h->cfg.lanes = int(value);
  else return rt::kErrArg;
  h->eng.reset();
  return rt::kOk;
}

// init_symbolic + init_numeric: uploads, converts to the inner precision, Jacobi setup,
// coarse inverse (solver/base.hpp:227-260, multigrid.hpp:814-818)
int feat3_mgpcg_set_level_order(feat3_mgpcg* h, int level, const uint32_t* order)
{
  if(!h || level < 0 || std::size_t(level) >= h->levels.size()) return rt::kErrState;
  if(!order) return rt::kErrArg;
  h->orders.resize(h->levels.size());
  h->pending.resize(h->levels.size(), 0);
  // a level's numbering is fixed once its data has been permuted
  if(!h->orders[level].empty()) return rt::kErrState;
  const u64 n = h->levels[level].A.rows;
  std::vector<u32> o(order, order + n);
  std::vector<char> hit(n, 0);
  for(u32 v : o)
  {
    if(u64(v) >= n || hit[v]) return rt::kErrArg;
    hit[v] = 1;
  }
  h->orders[level].swap(o);
  h->pending[level] = 1;
  h->eng.reset();
  return rt::kOk;
}

int feat3_mgpcg_init(feat3_mgpcg* h)
{
  if(!h || h->levels.empty()) return rt::kErrState;
  std::deque<DevMem> store;    // permuted copies of device-resident levels, until init returns
  F3_RC(apply_orders(h, store));
  std::unique_ptr<EngineBase> e;
  if(h->inner == 0) e.reset(new Engine<double, double>());
  else if(h->inner == 1) e.reset(new Engine<float, float>());
  else e.reset(new Engine<float, half>());
  const int rc = e->init(h->levels, h->cfg, h->stream);
  if(rc) return rc;
  h->eng = std::move(e);
  return rt::kOk;
}

// Solver::solve / IterativeSolver::correct(x, b): x and b are FP64 device vectors
int feat3_mgpcg_correct(feat3_mgpcg* h, void* caller_stream, double* x, const double* b, int* status, u64* iters)
{
  if(!h || !h->eng) return rt::kErrState;
  // the caller's stream (NULL: the null stream) has produced the inputs: wait for it (the
  // solver's own stream does not synchronise with it implicitly)
  if(hipStreamSynchronize(static_cast<hipStream_t>(caller_stream)) != hipSuccess) return rt::kErrRuntime;
  if(!h->permuted0) return h->eng->solve(x, b, true, status, iters);
  if(!x || !b) return rt::kErrArg;
  F3_RC(to_internal(h, b, h->bi.as<double>()));
  F3_RC(to_internal(h, x, h->xi.as<double>()));
  // the solution goes back to the caller's numbering right behind the solve's last launch
  const std::function<int()> back = [&]() { return to_caller_async(h, h->xi.as<double>(), x); };
  return h->eng->solve(h->xi.as<double>(), h->bi.as<double>(), true, status, iters, &back);
}

// SolverBase::apply(cor, def): cor = PCG(def) from a zero start
int feat3_mgpcg_apply(feat3_mgpcg* h, void* caller_stream, double* cor, const double* def, int* status, u64* iters)
{
  if(!h || !h->eng) return rt::kErrState;
  // the caller's stream (NULL: the null stream) has produced the inputs: wait for it (the
  // solver's own stream does not synchronise with it implicitly)
  if(hipStreamSynchronize(static_cast<hipStream_t>(caller_stream)) != hipSuccess) return rt::kErrRuntime;
  if(!h->permuted0) return h->eng->solve(cor, def, false, status, iters);
  if(!cor || !def) return rt::kErrArg;
  F3_RC(to_internal(h, def, h->bi.as<double>()));
  const std::function<int()> back = [&]() { return to_caller_async(h, h->xi.as<double>(), cor); };
  return h->eng->solve(h->xi.as<double>(), h->bi.as<double>(), false, status, iters, &back);
}

// the MultiGrid preconditioner alone (one V-cycle): cor = MG(def), FP64 in and out
int feat3_mgpcg_vcycle(feat3_mgpcg* h, void* caller_stream, double* cor, const double* def)
{
  if(!h || !h->eng) return rt::kErrState;
  // the caller's stream (NULL: the null stream) has produced the inputs: wait for it (the
  // solver's own stream does not synchronise with it implicitly)
  if(hipStreamSynchronize(static_cast<hipStream_t>(caller_stream)) != hipSuccess) return rt::kErrRuntime;
  if(!h->permuted0) return h->eng->vcycle(cor, def);
  if(!cor || !def) return rt::kErrArg;
  F3_RC(to_internal(h, def, h->bi.as<double>()));
  F3_RC(h->eng->vcycle(h->xi.as<double>(), h->bi.as<double>()));
  return to_caller(h, h->xi.as<double>(), cor);
}

int feat3_mgpcg_history(feat3_mgpcg* h, double* out, u64 cap, u64* len)
{
  if(!h || !h->eng) return rt::kErrState;
  return h->eng->history(out, cap, len);
}

int feat3_mgpcg_time_kernel(feat3_mgpcg* h, int which, int reps, double* avg_ms, u64* bytes)
{
  if(!h || !h->eng || !avg_ms) return rt::kErrState;
  return h->eng->time_kernel(which, reps, avg_ms, bytes);
}

int feat3_mgpcg_info(feat3_mgpcg* h, u64* out, int cap)
{
  if(!h || !h->eng || !out) return rt::kErrState;
  return h->eng->info(out, cap);
}

int feat3_mgpcg_statistics(feat3_mgpcg* h, int reps, double* out, int cap)
{
  if(!h || !h->eng) return rt::kErrState;
  return h->eng->statistics(reps, out, cap);
}

} // extern "C"
