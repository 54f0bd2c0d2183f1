// mg_kernels.hpp -- fused kernels of the device multigrid cycles and the device-resident PCG.
//
// The V-cycle follows MultiGrid::_apply_cycle_v (solver/multigrid.hpp:1333-1351,1711-2000)
// with Richardson(omega_R, Jacobi(omega_J)) smoothers (richardson.hpp:187-228,
// jacobi_precond.hpp:137-165).  Instead of FEAT's op-by-op sequence each smoothing step is
// one kernel:  x_out = x_in + dinv .* (rhs - A x_in),  dinv = omega_R*omega_J/diag(A) with
// dinv = 0 on Dirichlet rows (which is the filter_cor of the Jacobi correction).
// Post-smoothing runs directly on the level solution (Richardson on A c = rhs - A x from
// c = 0 is the same iteration as Richardson on A x = rhs started from x).  Every other
// SpMV carries the unit-filter as a row bitmask in its epilogue.
//
// Matrix formats are "row evaluators": CsrRows<G> (G lanes per row, sub-wave reduction) and
// SellRows (one lane per row of a sigma-sorted SELL-C chunk).
#pragma once

#include "device_common.hpp"

namespace feat3_amd
{
namespace mg
{
  typedef std::uint32_t u32;
  typedef std::uint64_t u64;

  enum : int { kUndefined = 0, kProgress = 1, kSuccess = 2, kAborted = 3, kDiverged = 4, kMaxIter = 5, kStagnated = 6 };

  // Launch-shape knobs of the row kernels for measurements (make EXTRA=-D...): F3_BAND_UNROLL
  // diagonals per batch of loads in the banded row sums, F3_ROW_WAVES an upper bound on the waves
  // per SIMD of the matrix row kernels (0: the compiler's choice)
#ifndef F3_BAND_UNROLL
#define F3_BAND_UNROLL 9
#endif
#ifndef F3_ROW_WAVES
#define F3_ROW_WAVES 0
#endif
#define F3_STR2(x) #x
#define F3_STR(x) F3_STR2(x)
#if F3_ROW_WAVES > 0
#define F3_ROW_ATTR __attribute__((amdgpu_waves_per_eu(1, F3_ROW_WAVES)))
#else
#define F3_ROW_ATTR
#endif

  // Row evaluators.  row_sum(p, lane, x) returns sum_j a_pj x_j for the row at position p,
  // reduced over the G lanes of the calling group (valid in every lane); row(p) maps the
  // position to the vector index (identity for CSR, the sigma permutation for SELL).
  template<typename DT_, typename VT_, int G_>
  struct CsrRows
  {
    static constexpr int G = G_;
    const u32* rp;
    const u32* col;
    const VT_* val;
    u64 rows;
    __device__ __forceinline__ u64 row(u64 p) const { return p; }
    __device__ __forceinline__ u64 block_of(u64 b) const { return b; }
    __device__ __forceinline__ DT_ row_sum(u64 p, int lane, const DT_* __restrict__ x) const
    {
      const u32 end = rp[p + 1];
      DT_ s = DT_(0);
      for(u32 i = rp[p] + u32(lane); i < end; i += u32(G_))
        s += dev::ld<DT_>(val, i) * x[col[i]];
      return dev::group_sum<G_>(s);
    }
  };

  // CSR evaluated by streaming: the rows of a block are processed in rounds of R consecutive
  // rows whose entries [rp[r0], rp[r1]) are contiguous; the 256 threads read values and columns
  // with unit stride (8 independent loads per thread in flight) and stage the products
  // val * x[col] in LDS, then thread t adds up row r0 + t in entry order.  A round whose
  // entries exceed the LDS capacity is summed with 8 lanes per row.  R is chosen on the host
  // so that a round's entries fill about 85% of the capacity.  for_rows has the block-level
  // loop; row_sum is not used.
  template<typename DT_, typename VT_>
  struct CsrStreamRows
  {
    static constexpr int G = 1;
    static constexpr int kCapBytes = 16 * 1024;
    static constexpr int CAP = kCapBytes / int(sizeof(DT_));
    const u32* rp;
    const u32* col;
    const VT_* val;
    u64 rows;
    u64 R;
    __device__ __forceinline__ u64 row(u64 p) const { return p; }
    __device__ __forceinline__ u64 block_of(u64 b) const { return b; }
  };

  template<typename T> struct IsStream { static constexpr bool value = false; };
  template<typename DT_, typename VT_> struct IsStream<CsrStreamRows<DT_, VT_>> { static constexpr bool value = true; };
  template<typename T> struct IsTensor { static constexpr bool value = false; };

  // column storage of the derived SELL format: absolute u32 indices, or 16-bit signed offsets
  // from the row index (usable when every |col - row| < 2^15, e.g. lexicographic 3D grids up
  // to 129^3 nodes); a padding slot holds offset 0 / value 0
  template<typename CI_> struct ColDecode;
  template<> struct ColDecode<u32>
  {
    __device__ static __forceinline__ u64 get(u32 c, u64) { return c; }
  };
  template<> struct ColDecode<std::int16_t>
  {
    __device__ static __forceinline__ u64 get(std::int16_t d, u64 r) { return u64(std::int64_t(r) + std::int64_t(d)); }
  };

  // SELL-C-sigma: one lane per row; a row's k-th entry of chunk c sits at cs[c] + (p % C) + k*C,
  // so a wave's loads are contiguous.  The loop runs to the chunk width cl[c] (uniform over the
  // chunk, padded with zeros), which keeps the trip count wave-uniform and lets it unroll.
  template<typename DT_, typename VT_, typename CI_, bool PERM_>
  struct SellRows
  {
    static constexpr int G = 1;
    const u32* cs;
    const u32* cl;
    const u32* perm;
    const CI_* col;
    const VT_* val;
    u64 rows, C;
    __device__ __forceinline__ u64 row(u64 p) const { return PERM_ ? u64(perm[p]) : p; }
    __device__ __forceinline__ u64 block_of(u64 b) const { return b; }
    __device__ __forceinline__ DT_ row_sum(u64 p, int, const DT_* __restrict__ x) const
    {
      const u64 ch = p / C;
      const u64 base = u64(cs[ch]) + (p - ch * C);
      const u32 len = cl[ch];
      const u64 r = row(p);
      DT_ s = DT_(0);
#pragma unroll 4
      for(u32 k = 0; k < len; ++k)
      {
        const u64 e = base + u64(k) * C;
        s += dev::ld<DT_>(val, e) * x[ColDecode<CI_>::get(col[e], r)];
      }
      return s;
    }
  };

  // Banded storage (the diagonals of FEAT's SparseMatrixBanded, apply_generic.hpp:519-598) of a
  // square matrix:
  // diagonal d of row p holds the entry at column p + off[d] + 1 - rows; a column outside
  // [0, cols) does not contribute.  Values are chunk-major, val[(p/64)*ndiag*64 + d*64 + p%64],
  // so one wave (64 consecutive rows) reads one contiguous region -- FEAT's diagonal-major
  // order would make every wave read ndiag streams one matrix-length apart.  One lane per row,
  // 64 consecutive x per diagonal, no column indices; the offsets are wave-uniform (scalar
  // loads).
  // SYM_ (FP64 bands): symmetric storage of a band that equals its mirror image bit for bit (a_pc = a_cp):
  // only the center and the upper diagonals are stored (nds of them, chunk-major as above), and
  // sx[d] is the stored diagonal that holds diagonal d.  A lower diagonal's value for row p is
  // the upper one stored at row c (its column), the same number, so the row sum takes the same
  // terms in the same order as the full band; the half of the values read a second time come
  // from the L2 (row c lies at most one node plane before p).
  // x of a banded matrix staged in LDS (BandedRows with STAGE_): a block evaluates its rows in
  // rounds of kStageRows consecutive rows r0.., and the columns those rows reach form a few
  // windows (3D Q1 in lattice order: 516 columns each in the node plane below, its own and the one
  // above), which all threads copy into LDS with contiguous loads before the round's row sums read
  // x from there: one global load per staged value instead of one gather per matrix entry (3D Q1:
  // 1548 loads instead of 6912 per round of 256 rows).  Window w holds the columns
  // r0 + wstart[w] .. r0 + wstart[w] + wlen[w] - 1 at LDS offset wbase[w] (a column outside the
  // matrix is staged as 0, and its product is selected away as in the unstaged sum); diagonal d
  // reads the column of row p at xoff[d] + (p - r0).
  constexpr u32 kStageRows = 256;
  constexpr u32 kStageWin = 8;
  constexpr u32 kStageBytes = 32 * 1024;
  struct BandStage
  {
    u32 nwin;
    int wstart[kStageWin];
    u32 wlen[kStageWin], wbase[kStageWin];
    const u32* xoff;   // per diagonal, device memory
  };

  // the banded row sums accumulate s = a * x + s with one rounding (fused multiply-add)
  __device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
  __device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

  template<typename DT_, typename VT_, bool SYM_ = false, bool STAGE_ = false>
  struct BandedRows
  {
    static constexpr int G = 1;
    static constexpr bool kStaged = STAGE_;
    // FP64 values are read with nontemporal loads: at 3D Q1 level 7 (464 MB, larger than the
    // 256 MB Infinity Cache) the FP64 sweep went 117 -> 107 us and the outer PCG product
    // 112 -> 106 us (FP32-inner solve 3.31 -> 3.26 ms, same GPU, alternating builds).  FP32
    // values (232 MB) stay ordinary loads: consecutive sweeps reuse them from the Infinity Cache,
    // which nontemporal loads give up (39.7 -> 45.5 us).
    static constexpr bool kNonTemporal = sizeof(VT_) == 8 && !SYM_;
    const u32* off;
    const VT_* val;
    u64 rows, cols;
    u32 ndiag;
    const u32* sx = nullptr;   // SYM_: stored diagonal of each diagonal
    u32 nds = 0;               // SYM_: stored diagonals
    // optional processing order of the row ranges (see DevCsr::make_tiles): logical block b of
    // an XCD works on range tile[b]
    const u32* tile = nullptr;
    BandStage st{};   // STAGE_: the LDS windows of x
    __device__ __forceinline__ u64 row(u64 p) const { return p; }
    __device__ __forceinline__ u64 block_of(u64 b) const { return tile ? u64(tile[b]) : b; }
    // Index arithmetic in 32 bits (the banded format is built for fewer than 2^31 rows): the
    // column c = p + off[d] + 1 - rows wraps past 2^32 below 0, so one unsigned compare with cols
    // tests both ends, and the chunk index (< 2^26) times the chunk stride uses a 24-bit multiply
    // (64-bit index arithmetic cost about 36 VALU per diagonal, the FP64 outer product's bound).
    __device__ __forceinline__ DT_ row_sum(u64 p64, int, const DT_* __restrict__ x) const
    {
      const dev::cptr<u32> co = dev::to_const(off);
      const u32 p = u32(p64), shift = 1u - u32(rows), nc = u32(cols);
      DT_ s = DT_(0);
      if constexpr(SYM_)
      {
        const dev::cptr<u32> cs = dev::to_const(sx);
        const u32 stride = nds * 64u;
_Pragma(F3_STR(unroll F3_BAND_UNROLL))
        for(u32 d = 0; d < ndiag; ++d)
        {
          const u32 c = p + co[d] + shift;
          const bool in = c < nc;
          const u32 vr = (in && c < p) ? c : p;   // lower diagonal: the value stored at row c
          const u32 at = __umul24(vr >> 6, stride) + cs[d] * 64u + (vr & 63u);
          const DT_ a = kNonTemporal ? dev::ld_nt<DT_>(val, at) : dev::ld<DT_>(val, at);
          const DT_ t = fmadd(a, x[in ? c : p], s);
          s = in ? t : s;
        }
        return s;
      }
      const u32 base = __umul24(p >> 6, ndiag * 64u) + (p & 63u);
_Pragma(F3_STR(unroll F3_BAND_UNROLL))
      for(u32 d = 0; d < ndiag; ++d)
      {
        // The load is unconditional (an in-range stand-in column) and the term selected away,
        // so that no branch splits the batch of loads
        const u32 c = p + co[d] + shift;
        const bool in = c < nc;
        const DT_ a = kNonTemporal ? dev::ld_nt<DT_>(val, base + d * 64u) : dev::ld<DT_>(val, base + d * 64u);
        const DT_ t = fmadd(a, x[in ? c : p], s);
        s = in ? t : s;
      }
      return s;
    }

    // the same row sum with x from the LDS windows of the round that starts at row r0 (the same
    // terms in the same order)
    __device__ __forceinline__ DT_ row_sum_staged(u32 p, u32 r0, const DT_* xs) const
    {
      const dev::cptr<u32> co = dev::to_const(off);
      const dev::cptr<u32> xo = dev::to_const(st.xoff);
      const u32 shift = 1u - u32(rows), nc = u32(cols), q = p - r0;
      DT_ s = DT_(0);
      if constexpr(SYM_)
      {
        const dev::cptr<u32> cs = dev::to_const(sx);
        const u32 stride = nds * 64u;
_Pragma(F3_STR(unroll F3_BAND_UNROLL))
        for(u32 d = 0; d < ndiag; ++d)
        {
          const u32 c = p + co[d] + shift;
          const bool in = c < nc;
          const u32 vr = (in && c < p) ? c : p;
          const u32 at = __umul24(vr >> 6, stride) + cs[d] * 64u + (vr & 63u);
          const DT_ a = dev::ld<DT_>(val, at);
          const DT_ t = fmadd(a, xs[xo[d] + q], s);
          s = in ? t : s;
        }
        return s;
      }
      const u32 base = __umul24(p >> 6, ndiag * 64u) + (p & 63u);
_Pragma(F3_STR(unroll F3_BAND_UNROLL))
      for(u32 d = 0; d < ndiag; ++d)
      {
        const u32 c = p + co[d] + shift;
        const bool in = c < nc;
        const DT_ a = kNonTemporal ? dev::ld_nt<DT_>(val, base + d * 64u) : dev::ld<DT_>(val, base + d * 64u);
        const DT_ t = fmadd(a, xs[xo[d] + q], s);
        s = in ? t : s;
      }
      return s;
    }
  };
  template<typename T> struct IsStagedBand { static constexpr bool value = false; };
  template<typename DT_, typename VT_, bool SYM_> struct IsStagedBand<BandedRows<DT_, VT_, SYM_, true>>
  {
    static constexpr bool value = true;
  };

  // Value-coded band: a banded matrix whose every diagonal holds at most kDictSlots distinct
  // values (bit patterns) stores, per row, one 4-bit code per diagonal into its diagonal's table
  // of values instead of the values themselves: the codes of 8 diagonals in one 32-bit word,
  // chunk-major (word w of row p at idx[(p / 64) * nw * 64 + w * 64 + p % 64]), the tables
  // [nw * 8][kDictSlots] in the storage type.  A block copies the tables into LDS (DT, as
  // BandedRows converts the stored values) before its rows; the row sum then takes the same
  // terms, in the same order and with the same arithmetic (fused multiply-adds) as BandedRows,
  // so the results are bit for bit those of the band it was built from (3D Q1:
  // 4 bytes of codes per 8 diagonals instead of 32 of FP32 or 64 of FP64 values).
  // ND_ > 0: exactly ND_ diagonals, every loop unrolled in full (all loads of a row in flight
  // together), and blocks whose rows reach only columns inside the matrix take the row sums
  // without the range tests.  ND_ = 0: any count, the offsets padded to nw * 8 with stand-ins
  // whose column is outside the matrix for every row (their terms are selected away).
#ifndef F3_DICT_RT
#define F3_DICT_RT 2   // FP32 rows per thread in the interior blocks of value-coded levels (3, 4: the same time)
#endif
  constexpr u32 kDictSlots = 16;
  constexpr u32 kDictDiag = 32;   // diagonals at most (4 code words per row)
  template<typename DT_, typename VT_, int ND_ = 0, bool TRIP_ = false>
  struct DictBandRows
  {
    static constexpr int G = 1;
    static constexpr int kND = ND_;
    const u32* off;    // offsets (off + rows - 1), ascending, padded to nw * 8
    const u32* idx;    // code words, chunk-major
    const VT_* tab;    // [nw * 8][kDictSlots] values
    u64 rows, cols;
    u32 ndiag, nw;     // diagonals, code words per row
    int omin, omax;    // the smallest and largest offset col - row
    const u32* tile = nullptr;
    const DT_* lt = nullptr;   // the tables in LDS (set by for_rows)
    // TRIP_: 27 diagonals in nine runs of three consecutive offsets (off_{3g} + 1 = off_{3g+1} =
    // off_{3g+2} - 1: the 3D Q1 stencil in lattice order), FP64
    static_assert(!TRIP_ || (ND_ == 27 && sizeof(DT_) == 8), "runs of three: 27 FP64 diagonals");
    __device__ __forceinline__ u64 row(u64 p) const { return p; }
    __device__ __forceinline__ u64 block_of(u64 b) const { return tile ? u64(tile[b]) : b; }
    // x at column c through a 32-bit byte offset (the band's rows are fewer than 2^26)
    __device__ __forceinline__ static DT_ xat(const DT_* __restrict__ x, u32 c)
    {
      return *reinterpret_cast<const DT_*>(reinterpret_cast<const char*>(x) + c * u32(sizeof(DT_)));
    }
    // a lane's value from the lane before it (DPP control 0x138, wave_shr:1) or after it (0x130,
    // wave_shl:1); the first / last lane keeps its own (its caller loads that one itself)
    template<int CTRL>
    __device__ __forceinline__ static DT_ lane_shift(DT_ v)
    {
      if constexpr(sizeof(DT_) == 8)
      {
        const u64 b = __builtin_bit_cast(u64, v);
        const u32 lo = u32(b), hi = u32(b >> 32);
        const u32 slo = u32(__builtin_amdgcn_update_dpp(int(lo), int(lo), CTRL, 0xF, 0xF, false));
        const u32 shi = u32(__builtin_amdgcn_update_dpp(int(hi), int(hi), CTRL, 0xF, 0xF, false));
        return __builtin_bit_cast(DT_, (u64(shi) << 32) | u64(slo));
      }
      else
        return __builtin_bit_cast(DT_, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, v), __builtin_bit_cast(int, v), CTRL,
                                                                   0xF, 0xF, false));
    }
    __device__ __forceinline__ DT_ term_value(u32 d, u32 code, u32 j) const
    {
      return lt[d * kDictSlots + __builtin_amdgcn_ubfe(code, 4u * j, 4u)];
    }
    // BandedRows' range test and selection per term
    __device__ __forceinline__ DT_ row_sum(u64 p64, int, const DT_* __restrict__ x) const
    {
      const dev::cptr<u32> co = dev::to_const(off);
      const u32 p = u32(p64), shift = 1u - u32(rows), nc = u32(cols);
      DT_ s = DT_(0);
      if constexpr(ND_ > 0)
      {
        // every load of the row first (code words, then x), then the terms in order
        constexpr u32 NW = (u32(ND_) + 7u) / 8u;
        const u32 base = __umul24(p >> 6, NW * 64u) + (p & 63u);
        u32 code[NW];
#pragma unroll
        for(u32 w = 0; w < NW; ++w) code[w] = idx[base + w * 64u];
        DT_ xv[ND_];
        bool in[ND_];
#pragma unroll
        for(u32 d = 0; d < u32(ND_); ++d)
        {
          const u32 c = p + co[d] + shift;
          in[d] = c < nc;
          xv[d] = xat(x, in[d] ? c : p);
        }
#pragma unroll
        for(u32 d = 0; d < u32(ND_); ++d)
        {
          const DT_ t = fmadd(term_value(d, code[d >> 3], d & 7u), xv[d], s);
          s = in[d] ? t : s;
        }
      }
      else
      {
        const u32 base = __umul24(p >> 6, nw * 64u) + (p & 63u);
        for(u32 w = 0; w < nw; ++w)
        {
          const u32 code = idx[base + w * 64u];
#pragma unroll
          for(u32 j = 0; j < 8u; ++j)
          {
            const u32 d = w * 8u + j;
            const DT_ a = term_value(d, code, j);
            const u32 c = p + co[d] + shift;
            const bool in = c < nc;
            const DT_ t = fmadd(a, xat(x, in ? c : p), s);
            s = in ? t : s;
          }
        }
      }
      return s;
    }
    // ND_ > 0, rows whose every column is inside the matrix (the caller checks the block's range
    // against omin / omax): the same terms without the range test, for R rows at once (every
    // load of the R rows issued before the first term is formed).  FP64 x is read with buffer
    // loads whose per-lane offset (p + omin) * 8 is the same for every diagonal and whose
    // diagonal offset (off_d - omin) * 8 is wave-uniform (a scalar register), so the addresses
    // cost no vector instructions (3D Q1 L7 outer product 50.6 -> 39.8 us); FP32 x keeps global
    // loads at p + off_d (buffer loads, all issued together, took the FP32 sweep 19.7 -> 21.1 us)
    template<int R>
    __device__ __forceinline__ void row_sums_interior(const u32 (&p)[R], const DT_* __restrict__ x, DT_ (&s)[R]) const
    {
      const dev::cptr<u32> co = dev::to_const(off);
      constexpr u32 NW = (u32(ND_ > 0 ? ND_ : 1) + 7u) / 8u;
      constexpr u32 NDX = u32(ND_ > 0 ? ND_ : 1);
      u32 code[R][NW];
      DT_ xv[R][NDX];
#pragma unroll
      for(int r = 0; r < R; ++r)
      {
        const u32 base = __umul24(p[r] >> 6, NW * 64u) + (p[r] & 63u);
#pragma unroll
        for(u32 w = 0; w < NW; ++w) code[r][w] = idx[base + w * 64u];
      }
      if constexpr(sizeof(DT_) == 8)
      {
        const u32 shift = 1u - u32(rows) - u32(omin);   // co[d] + shift = off_d - omin >= 0
        const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc(const_cast<DT_*>(x), 0, -1, 0x00020000);
        auto ldx = [&](u32 vo, u32 d) {
          return __builtin_bit_cast(DT_, __builtin_amdgcn_raw_buffer_load_b64(xr, vo, (co[d] + shift) * 8u, 0));
        };
        if constexpr(TRIP_ && R == 1)
        {
          {
            // the lanes of a wave hold consecutive rows, so x at off - 1 and off + 1 of a run is
            // the middle diagonal's x of the lane before and after: one load per run and two
            // lane shifts (DPP wave_shr:1 / wave_shl:1) instead of three loads; the first and
            // last lane, and a lane next to an inactive one, load their own (two branch regions,
            // issued with the middle loads)
            const u32 vo = (p[0] + u32(omin)) * 8u;
            const u32 lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
            const u64 ex = __builtin_amdgcn_read_exec();
            const bool own_l = lane == 0u || ((ex >> (lane - 1u)) & 1u) == 0u;
            const bool own_r = lane == 63u || ((ex >> (lane + 1u)) & 1u) == 0u;
            DT_ xc[9], bl[9], br[9];
#pragma unroll
            for(u32 g = 0; g < 9u; ++g) { xc[g] = ldx(vo, 3u * g + 1u); bl[g] = DT_(0); br[g] = DT_(0); }
            if(own_l)
            {
#pragma unroll
              for(u32 g = 0; g < 9u; ++g) bl[g] = ldx(vo, 3u * g);
            }
            if(own_r)
            {
#pragma unroll
              for(u32 g = 0; g < 9u; ++g) br[g] = ldx(vo, 3u * g + 2u);
            }
#pragma unroll
            for(u32 g = 0; g < 9u; ++g)
            {
              const DT_ sl = lane_shift<0x138>(xc[g]), sn = lane_shift<0x130>(xc[g]);
              xv[0][3u * g] = own_l ? bl[g] : sl;
              xv[0][3u * g + 1u] = xc[g];
              xv[0][3u * g + 2u] = own_r ? br[g] : sn;
            }
          }
        }
        else
        {
#pragma unroll
          for(int r = 0; r < R; ++r)
          {
            const u32 vo = (p[r] + u32(omin)) * 8u;
#pragma unroll
            for(u32 d = 0; d < u32(ND_); ++d) xv[r][d] = ldx(vo, d);
          }
        }
      }
      else
      {
        const u32 shift = 1u - u32(rows);
#pragma unroll
        for(int r = 0; r < R; ++r)
#pragma unroll
          for(u32 d = 0; d < u32(ND_); ++d) xv[r][d] = xat(x, p[r] + (co[d] + shift));
      }
#pragma unroll
      for(int r = 0; r < R; ++r)
      {
        s[r] = DT_(0);
#pragma unroll
        for(u32 d = 0; d < u32(ND_); ++d) s[r] = fmadd(term_value(d, code[r][d >> 3], d & 7u), xv[r][d], s[r]);
      }
    }
    __device__ __forceinline__ DT_ row_sum_interior(u32 p, const DT_* __restrict__ x) const
    {
      const u32 pp[1] = {p};
      DT_ s[1];
      row_sums_interior<1>(pp, x, s);
      return s[0];
    }
  };
  template<typename T> struct IsDictBand { static constexpr bool value = false; };
  template<typename DT_, typename VT_, int ND_, bool T_> struct IsDictBand<DictBandRows<DT_, VT_, ND_, T_>>
  {
    static constexpr bool value = true;
  };

  // Tensor-product transfer matrix on lattices numbered lexicographically (x fastest):
  // M[(oz,oy,ox),(iz,iy,ix)] = (wz * wy) * wx with the 1D factors of the three axes (a 2D
  // matrix has the 1x1 identity as its z factor).  Entry k of output index o along an axis is
  // (idx[o*kTensorW + k], w[o*kTensorW + k]), k < cnt[o], ascending; the indices are stored
  // scaled by the axis's stride in the input lattice (x: 1, y: mi[0], z: mi[0]*mi[1]), so a
  // column is iz + iy + ix, and the slots from cnt[o] to kTensorW hold index 0 and weight 0.  The
  // solver uses it for the grid transfers of structured hierarchies only after checking every
  // stored entry of the CSR matrix against it (column, count and the value in the arithmetic
  // type bit for bit), so a row sum reads no matrix data; the terms are the CSR row's, in its
  // column order, each product rounded before it is added as in the streaming CSR evaluation
  // (CsrStreamRows): the results are identical to it.
  constexpr u32 kTensorW = 8;
  template<typename DT_>
  struct TensorAxis
  {
    const u32* cnt;
    const u32* idx;
    const DT_* w;
  };

  // q / d for q < 2^32 with m = ceil(2^64 / d), 2 <= d < 2^32: floor(q m / 2^64) is exact (the
  // error q (m - 2^64/d) / 2^64 < 2^-32 <= 1/d cannot carry a quotient past an integer)
  __host__ __device__ __forceinline__ u32 udiv_magic(u32 q, u64 m)
  {
    const u64 lo = u64(q) * (m & 0xffffffffull);
    const u64 hi = u64(q) * (m >> 32) + (lo >> 32);
    return u32(hi >> 32);
  }
  inline u64 udiv_magic_for(u32 d) { return ~u64(0) / d + 1; }

  // the first K_ entries of a factor row (16-byte loads; the row has kTensorW = 8 slots)
  template<int K_, typename T>
  __device__ __forceinline__ void load_entries(const T* row, T (&out)[K_])
  {
    constexpr int per = int(16 / sizeof(T));
    typedef T vt __attribute__((ext_vector_type(per)));
#pragma unroll
    for(int k0 = 0; k0 < K_; k0 += per)
    {
      const vt v = *reinterpret_cast<const vt*>(row + k0);
#pragma unroll
      for(int j = 0; j < per; ++j)
        if(k0 + j < K_) out[k0 + j] = v[j];
    }
  }

  // K_: bound of the entries per axis (every cnt <= K_), D3_: the z factor is not the 1x1
  // identity.  One lane per row; the factor rows are read whole and the x gathers issued for all
  // K_^3 (K_^2) terms before the additions, an absent term reading x[0] and being selected away.
  template<typename DT_, int K_ = int(kTensorW), bool D3_ = true, int RT_ = 1>
  struct TensorRows
  {
    static constexpr int G = 1;
    static constexpr int RT = RT_;   // rows per thread of mg::for_rows
    TensorAxis<DT_> ax[3];
    u32 mo[3], mi[3];
    u64 div0, div1;   // udiv_magic multipliers of mo[0] and mo[1]
    u64 rows;
    __device__ __forceinline__ u64 row(u64 p) const { return p; }
    __device__ __forceinline__ u64 block_of(u64 b) const { return b; }
    __device__ __forceinline__ DT_ row_sum(u64 p, int, const DT_* __restrict__ x) const
    {
#pragma clang fp contract(off)
      constexpr int KZ = D3_ ? K_ : 1;
      const u32 q = u32(p);
      const u32 t = udiv_magic(q, div0), ox = q - t * mo[0];
      const u32 oz = D3_ ? udiv_magic(t, div1) : 0u, oy = t - oz * mo[1];
      const u32 nz = D3_ ? ax[2].cnt[oz] : 1u, ny = ax[1].cnt[oy], nx = ax[0].cnt[ox];
      u32 iz[KZ], iy[K_], ix[K_];
      DT_ wz[KZ], wy[K_], wx[K_];
      if constexpr(D3_)
      {
        load_entries<KZ>(ax[2].idx + oz * kTensorW, iz);
        load_entries<KZ>(ax[2].w + oz * kTensorW, wz);
      }
      else
      {
        iz[0] = 0u;
        wz[0] = DT_(1);
      }
      load_entries<K_>(ax[1].idx + oy * kTensorW, iy);
      load_entries<K_>(ax[1].w + oy * kTensorW, wy);
      load_entries<K_>(ax[0].idx + ox * kTensorW, ix);
      load_entries<K_>(ax[0].w + ox * kTensorW, wx);
      DT_ xv[KZ][K_][K_];
#pragma unroll
      for(int a = 0; a < KZ; ++a)
#pragma unroll
        for(int b = 0; b < K_; ++b)
        {
          const u32 base = iz[a] + iy[b];
#pragma unroll
          for(int c = 0; c < K_; ++c)
          {
            const bool in = u32(a) < nz && u32(b) < ny && u32(c) < nx;
            xv[a][b][c] = x[in ? base + ix[c] : 0u];
          }
        }
      DT_ s = DT_(0);
#pragma unroll
      for(int a = 0; a < KZ; ++a)
#pragma unroll
        for(int b = 0; b < K_; ++b)
        {
          const DT_ wzy = D3_ ? wz[a] * wy[b] : wy[b];
#pragma unroll
          for(int c = 0; c < K_; ++c)
          {
            const DT_ v = wzy * wx[c];
            const DT_ pr = v * xv[a][b][c];
            const DT_ s2 = s + pr;
            s = (u32(a) < nz && u32(b) < ny && u32(c) < nx) ? s2 : s;
          }
        }
      return s;
    }
  };

  template<typename DT_, int K_, bool D3_, int RT_> struct IsTensor<TensorRows<DT_, K_, D3_, RT_>>
  {
    static constexpr bool value = true;
  };

  // Launch shape of the row kernels: every block owns a contiguous range of `per` row
  // positions (a multiple of 64), its 256/G lane groups stride through that range, and the
  // blocks are XCD-remapped so that neighbouring ranges share an L2.
  struct RowGrid
  {
    u64 per;
    unsigned blocks;
  };

  inline RowGrid row_grid(u64 rows, int G, unsigned max_blocks = 8192)
  {
    const u64 groups_per_block = 256 / u64(G);
    u64 nb = (rows + groups_per_block - 1) / groups_per_block;
    nb = nb < 1 ? 1 : (nb > max_blocks ? max_blocks : nb);
    u64 per = (rows + nb - 1) / nb;
    per = (per + 63) / 64 * 64;
    if(per == 0) per = 64;
    RowGrid g;
    g.per = per;
    g.blocks = unsigned((rows + per - 1) / per);
    if(g.blocks == 0) g.blocks = 1;
    return g;
  }

  // a fixed number of row positions per block (rounded up to a multiple of 64, and raised where
  // the grid would exceed max_blocks, the size of the reduction partials)
  inline RowGrid row_grid_fixed(u64 rows, u64 per, unsigned max_blocks = 8192)
  {
    per = (per < 64 ? 64 : (per + 63) / 64 * 64);
    if((rows + per - 1) / per > max_blocks) per = ((rows + max_blocks - 1) / max_blocks + 63) / 64 * 64;
    RowGrid g;
    g.per = per;
    g.blocks = unsigned((rows + per - 1) / per);
    if(g.blocks == 0) g.blocks = 1;
    return g;
  }

  // the block-level loop of CsrStreamRows (all threads of the block take part)
  template<typename DT_, typename VT_, typename F>
  __device__ __forceinline__ void for_rows_stream(const CsrStreamRows<DT_, VT_>& A, u64 beg, u64 end,
                                                  const DT_* __restrict__ x, F&& f)
  {
    constexpr int CAP = CsrStreamRows<DT_, VT_>::CAP;
    constexpr unsigned NT = 256;
    __shared__ DT_ prod[CAP];
    for(u64 r0 = beg; r0 < end; r0 += A.R)
    {
      const u64 r1 = r0 + A.R < end ? r0 + A.R : end;
      const u64 e0 = u64(A.rp[r0]), ne = u64(A.rp[r1]) - e0;
      __syncthreads();   // the previous round's readers are done with prod
      if(ne <= u64(CAP))
      {
        const u64 my = r0 + threadIdx.x;
        const u64 mk0 = my < r1 ? u64(A.rp[my]) : 0, mk1 = my < r1 ? u64(A.rp[my + 1]) : 0;
        dev::stage_products<8, NT>(A.val, A.col, e0, ne, x, prod);
        __syncthreads();
        for(u64 row = my; row < r1; row += NT)
        {
          const bool first = row == my;
          const u64 k0 = (first ? mk0 : u64(A.rp[row])) - e0, k1 = (first ? mk1 : u64(A.rp[row + 1])) - e0;
          DT_ s = DT_(0);
          for(u64 k = k0; k < k1; ++k) s += prod[k];
          f(row, s, true);
        }
      }
      else
      {
        constexpr int G = 8;
        const int lane = int(threadIdx.x % G);
        for(u64 row = r0 + threadIdx.x / G; row < r1; row += NT / G)
        {
          DT_ s = DT_(0);
          for(u64 i = u64(A.rp[row]) + u64(lane); i < u64(A.rp[row + 1]); i += G)
            s += dev::ld<DT_>(A.val, i) * x[A.col[i]];
          s = dev::group_sum<G>(s);
          f(row, s, lane == 0);
        }
      }
    }
  }

  // calls f(position, row_sum, is_group_leader) for every row position of this block
  template<typename Rows_, typename DT_, typename F>
  __device__ __forceinline__ void for_rows(const Rows_& A, u64 per, const DT_* __restrict__ x, F&& f)
  {
    constexpr int G = Rows_::G;
    const u64 b = A.block_of(dev::xcd_remap(blockIdx.x, gridDim.x));
    const u64 beg = b * per;
    const u64 end = beg + per < A.rows ? beg + per : A.rows;
    if constexpr(IsStream<Rows_>::value)
      for_rows_stream(A, beg, end, x, f);
    else if constexpr(IsStagedBand<Rows_>::value)
    {
      constexpr u32 CAP = kStageBytes / u32(sizeof(DT_));
      __shared__ DT_ xs[CAP];
      const BandStage& S = A.st;
      for(u64 r0 = beg; r0 < end; r0 += kStageRows)
      {
        __syncthreads();   // the previous round's readers are done with xs
        for(u32 w = 0; w < S.nwin; ++w)
          for(u32 t = threadIdx.x; t < S.wlen[w]; t += blockDim.x)
          {
            const std::int64_t c = std::int64_t(r0) + S.wstart[w] + std::int64_t(t);
            xs[S.wbase[w] + t] = (c >= 0 && u64(c) < A.cols) ? x[c] : DT_(0);
          }
        __syncthreads();
        const u64 p = r0 + threadIdx.x;
        if(threadIdx.x < kStageRows && p < end) f(p, A.row_sum_staged(u32(p), u32(r0), xs), true);
      }
    }
    else if constexpr(IsDictBand<Rows_>::value)
    {
      __shared__ DT_ lt[kDictDiag * kDictSlots];
      for(u32 i = threadIdx.x; i < A.nw * 8u * kDictSlots; i += blockDim.x) lt[i] = dev::ld<DT_>(A.tab, i);
      __syncthreads();
      Rows_ B = A;
      B.lt = lt;
      // a block whose rows reach only columns inside the matrix skips the range tests; there an
      // FP32 thread takes two rows blockDim.x apart at once, their loads in flight together (3D
      // Q1 L7 sweep 19.2 -> 15.4 us; FP64, whose 54 registers of x per row then cost occupancy,
      // one row: outer product 39.4 us against 42.4 with two)
      constexpr int RT = sizeof(DT_) == 4 ? F3_DICT_RT : 1;
      if(Rows_::kND > 0 && std::int64_t(beg) + A.omin >= 0 && std::int64_t(end) - 1 + A.omax < std::int64_t(A.cols))
        for(u64 p0 = beg + threadIdx.x; p0 < end; p0 += RT * u64(blockDim.x))
        {
          u32 pp[RT];
          bool has[RT];
#pragma unroll
          for(int r = 0; r < RT; ++r)
          {
            const u64 q = p0 + u64(r) * blockDim.x;
            has[r] = q < end;
            pp[r] = u32(has[r] ? q : p0);
          }
          DT_ sr[RT];
          B.template row_sums_interior<RT>(pp, x, sr);
#pragma unroll
          for(int r = 0; r < RT; ++r)
            if(has[r]) f(pp[r], sr[r], true);
        }
      else
        for(u64 p = beg + threadIdx.x; p < end; p += blockDim.x) f(p, B.row_sum(p, 0, x), true);
    }
    else if constexpr(IsTensor<Rows_>::value)
    {
      // RT rows per thread, 256 apart (each wave's accesses stay contiguous), their row sums formed
      // before the RT epilogues so that the loads of all of them are in flight together; a
      // position past the range evaluates the range's first row and is not written (so RT > 1 is
      // used only where every block has RT full rounds of rows)
      constexpr int RT = Rows_::RT;
      for(u64 p0 = beg + threadIdx.x; p0 < end; p0 += u64(blockDim.x) * RT)
      {
        DT_ s[RT];
#pragma unroll
        for(int j = 0; j < RT; ++j)
        {
          const u64 p = p0 + u64(j) * blockDim.x;
          s[j] = A.row_sum(p < end ? p : beg, 0, x);
        }
#pragma unroll
        for(int j = 0; j < RT; ++j)
        {
          const u64 p = p0 + u64(j) * blockDim.x;
          if(p < end) f(p, s[j], true);
        }
      }
    }
    else
    {
      const int lane = int(threadIdx.x % G);
      const u64 ngrp = blockDim.x / G;
      for(u64 p = beg + threadIdx.x / G; p < end; p += ngrp)
      {
        const DT_ s = A.row_sum(p, lane, x);
        f(p, s, lane == 0);
      }
    }
  }

  __device__ __forceinline__ bool stopped(const int* status) { return status != nullptr && *status != kProgress; }

  // ---- V-cycle kernels -------------------------------------------------------------------

  // MG entry: rhs = def (converted to the inner precision, convert_precond.hpp:104),
  // first Richardson step from zero: x = dinv .* rhs
  template<typename DT_>
  __global__ void k_mg_entry(u64 n, const double* __restrict__ def, DT_* __restrict__ rhs, DT_* __restrict__ x,
                             const DT_* __restrict__ dinv, const int* status)
  {
    if(stopped(status)) return;
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
    {
      const DT_ r = DT_(def[i]);
      rhs[i] = r;
      x[i] = dinv[i] * r;
    }
  }

  // The row loop of value-coded bands with the epilogue's operands loaded ahead: pre(row) issues
  // the loads a row's epilogue needs (they do not depend on the row sum) before its code words
  // and x, and for a thread's first rows before the block stages the value tables, so they are in
  // flight during the staging and the row sums; f(row, s, pre(row)) then finishes the row.  The
  // rows and sums are those of for_rows.
  template<typename Rows_, typename DT_, typename Pre, typename F>
  __device__ __forceinline__ void for_dict_rows_pre(const Rows_& A, u64 per, const DT_* __restrict__ x, Pre&& pre, F&& f)
  {
    static_assert(IsDictBand<Rows_>::value, "value-coded bands only");
    const u64 b = A.block_of(dev::xcd_remap(blockIdx.x, gridDim.x));
    const u64 beg = b * per;
    const u64 end = beg + per < A.rows ? beg + per : A.rows;
    constexpr int RT = sizeof(DT_) == 4 ? F3_DICT_RT : 1;
    const bool interior =
      Rows_::kND > 0 && std::int64_t(beg) + A.omin >= 0 && std::int64_t(end) - 1 + A.omax < std::int64_t(A.cols);
    const int rt = interior ? RT : 1;   // rows per thread and round
    typedef decltype(pre(u32(0))) P;
    P pv[RT];
    u32 pp[RT];
    bool has[RT];
    auto load = [&](u64 p0) {
#pragma unroll
      for(int r = 0; r < RT; ++r)
      {
        const u64 q = p0 + u64(r) * blockDim.x;
        has[r] = r < rt && q < end;
        pp[r] = u32(has[r] ? q : p0);
        pv[r] = pre(pp[r]);
      }
    };
    u64 p0 = beg + threadIdx.x;
    if(p0 < end) load(p0);
    __shared__ DT_ lt[kDictDiag * kDictSlots];
    for(u32 i = threadIdx.x; i < A.nw * 8u * kDictSlots; i += blockDim.x) lt[i] = dev::ld<DT_>(A.tab, i);
    __syncthreads();
    Rows_ B = A;
    B.lt = lt;
    for(bool first = true; p0 < end; p0 += u64(rt) * blockDim.x, first = false)
    {
      if(!first) load(p0);
      if(interior)
      {
        DT_ sr[RT];
        B.template row_sums_interior<RT>(pp, x, sr);
#pragma unroll
        for(int r = 0; r < RT; ++r)
          if(has[r]) f(pp[r], sr[r], pv[r]);
      }
      else
        f(pp[0], B.row_sum(pp[0], 0, x), pv[0]);
    }
  }

  // damped Jacobi step  x_out = x_in + dinv .* (rhs - A x_in)
  template<typename Rows_, typename DT_>
  __global__ __launch_bounds__(256) F3_ROW_ATTR void k_sweep(Rows_ A, u64 per, const DT_* __restrict__ dinv, const DT_* __restrict__ rhs,
                                                 const DT_* __restrict__ xin, DT_* __restrict__ xout, const int* status)
  {
    if(stopped(status)) return;
    if constexpr(IsDictBand<Rows_>::value)
    {
      struct Op { DT_ x, d, b; };
      for_dict_rows_pre(A, per, xin, [&](u32 i) { return Op{xin[i], dinv[i], rhs[i]}; },
                        [&](u32 i, DT_ s, const Op& o) { xout[i] = o.x + o.d * (o.b - s); });
    }
    else
      for_rows(A, per, xin, [&](u64 p, DT_ s, bool w) {
        const u32 i = u32(A.row(p));
        if(w) xout[i] = xin[i] + dinv[i] * (rhs[i] - s);
      });
  }

  // defect  def = filter(rhs - A x)
  template<typename Rows_, typename DT_>
  __global__ __launch_bounds__(256) F3_ROW_ATTR void k_defect(Rows_ A, u64 per, const u32* __restrict__ mask, const DT_* __restrict__ rhs,
                                                  const DT_* __restrict__ x, DT_* __restrict__ def, const int* status)
  {
    if(stopped(status)) return;
    if constexpr(IsDictBand<Rows_>::value)
    {
      struct Op { DT_ b; bool m; };
      for_dict_rows_pre(A, per, x, [&](u32 i) { return Op{rhs[i], dev::masked_nn(mask, i)}; },
                        [&](u32 i, DT_ s, const Op& o) { def[i] = o.m ? DT_(0) : o.b - s; });
    }
    else
      for_rows(A, per, x, [&](u64 p, DT_ s, bool w) {
        const u64 row = A.row(p);
        if(w)
        {
          // rhs read whatever the mask says: a load in one arm of the selection became a branch
          // around it, which doubled the live registers of the two-row loop (117 against 76 VGPRs)
          const u32 i = u32(row);
          const DT_ d = rhs[i] - s;
          def[i] = dev::masked_nn(mask, i) ? DT_(0) : d;
        }
      });
  }

  // restriction  rhs_c = filter_c(R def_f);  optional first smoothing step x_c = dinv_c .* rhs_c
  template<typename Rows_, typename DT_>
  __global__ __launch_bounds__(256) void k_restrict(Rows_ R, u64 per, const u32* __restrict__ mask_c, const DT_* __restrict__ def_f,
                                                    DT_* __restrict__ rhs_c, DT_* __restrict__ x_c,
                                                    const DT_* __restrict__ dinv_c, const int* status)
  {
    if(stopped(status)) return;
    for_rows(R, per, def_f, [&](u64 p, DT_ s, bool w) {
      const u64 row = R.row(p);
      if(w)
      {
        const DT_ r = dev::masked(mask_c, row) ? DT_(0) : s;
        rhs_c[row] = r;
        if(x_c) x_c[row] = dinv_c[row] * r;
      }
    });
  }

  // prolongation and coarse-grid correction  x_f += filter_f(P x_c)
  template<typename Rows_, typename DT_>
  __global__ __launch_bounds__(256) void k_prolong(Rows_ P, u64 per, const u32* __restrict__ mask_f, const DT_* __restrict__ x_c,
                                                   DT_* __restrict__ x_f, const int* status)
  {
    if(stopped(status)) return;
    for_rows(P, per, x_c, [&](u64 p, DT_ s, bool w) {
      const u64 row = P.row(p);
      if(w && !dev::masked(mask_f, row)) x_f[row] = x_f[row] + s;
    });
  }

  // dst = src (a level solution that ended in the other buffer of its pair)
  template<typename DT_>
  __global__ void k_copy_vec(u64 n, const DT_* __restrict__ src, DT_* __restrict__ dst, const int* status)
  {
    if(stopped(status)) return;
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      dst[i] = src[i];
  }

  // exact coarse solve with the FP64 inverse of the unit-filtered coarse matrix, one
  // workgroup: x = filter(Ainv * rhs)
  template<typename DT_>
  __global__ __launch_bounds__(256) void k_coarse(u64 n, const double* __restrict__ ainv, const u32* __restrict__ mask,
                                                  const DT_* __restrict__ rhs, DT_* __restrict__ x, const int* status)
  {
    if(stopped(status)) return;
    __shared__ double sr[1024];
    for(u64 j = threadIdx.x; j < n; j += blockDim.x) sr[j] = double(rhs[j]);
    __syncthreads();
    for(u64 i = threadIdx.x; i < n; i += blockDim.x)
    {
      double s = 0.0;
      // unrolled so that a group of the row's loads is in flight at once (the sum keeps its order)
#pragma unroll 8
      for(u64 j = 0; j < n; ++j) s += ainv[i * n + j] * sr[j];
      x[i] = dev::masked(mask, i) ? DT_(0) : DT_(s);
    }
  }

  // ---- coarse tail: all small levels in one workgroup ----------------------------------------
  //
  // Below a few thousand rows a level step is a launch of ~4 us that moves a few KB.  k_tail
  // runs a segment of a multigrid cycle -- every step on the levels tail[0] (the finest small
  // level) down to the coarsest and back, and for F- and W-cycles the peak smoothing and repeated
  // coarse solves between them -- in one 1024-thread workgroup with barriers between the steps.
  // A segment starts when the restriction above has written tail[0]'s right-hand side, takes
  // tail[0]'s first smoothing step from x = 0, runs its list of steps and ends by writing
  // tail[0]'s iterate for the prolongation above.  All vectors of these levels live in LDS (x,
  // its ping-pong partner, rhs, defect, dinv: 5 n values per level); the matrices are ELL with
  // 16-bit columns (entry k of row r at k * n + r, padded with column 0 and value 0), so a step
  // issues one round of independent global loads (L2-resident after the first cycle) and gathers
  // from LDS.  The arithmetic of every step is that of k_sweep / k_defect / k_restrict /
  // k_prolong / k_coarse; only the order of the terms of a row sum may differ from the level
  // format's.
  constexpr int kTailMax = 12;
  constexpr int kTailThreads = 1024;

  // one step of a tail segment: kind | level << 4 | a << 8 | b << 9 | step << 10, where a is the
  // buffer (0: x0, 1: x1) of the level's iterate that the step reads, b that of the next
  // coarser level (prolongation), step the fused first smoothing step of a restriction
  enum TailKind : u32 { kTailSweep = 0, kTailDefect = 1, kTailRest = 2, kTailCoarse = 3, kTailProl = 4 };
  __host__ __device__ constexpr u32 tail_op(u32 kind, u32 lvl, u32 a, u32 b, u32 step)
  {
    return kind | (lvl << 4) | (a << 8) | (b << 9) | (step << 10);
  }

  template<typename VT_>
  struct TailEll
  {
    const std::uint16_t* col;
    const VT_* val;
    u32 n, w;   // rows, entries per row (padded)
  };

  template<typename DT_, typename VT_>
  struct TailLevel
  {
    u32 n, off;             // rows; offset of the level's vectors in the LDS arena (DT_ elements)
    TailEll<VT_> A, P, R;   // P: this level <- next coarser; R: next coarser <- this level
    const u32* mask;
    const DT_* dinv;
  };

  template<typename DT_, typename VT_>
  struct TailDesc
  {
    int nlev;
    u32 ncoarse;
    const double* ainv;
    const DT_* rhs_in;   // right-hand side of level tail[0], written by the restriction above
    DT_* x_out;          // its iterate at the end of a segment, read by the prolongation above
    TailLevel<DT_, VT_> lv[kTailMax];
  };

  // FP32 rows of 8, 9, 16, 21, 24, 25, 27 or 32 entries (the tail ELL keeps the widths of the Q1
  // and Q2 stencils, 9, 21, 25 and 27, and pads others to a multiple of 8 up to 32) issue every
  // load of the row before the first use: one round trip to the L2 per step instead of one per
  // group of unrolled entries (the steps of the tail are latency-bound); the terms are summed in
  // the order of the loop below
  template<u32 W_, typename DT_, typename VT_>
  __device__ __forceinline__ DT_ tail_row_w(const TailEll<VT_>& M, u32 r, const DT_* xs)
  {
    const u64 n = M.n;
    u32 c[W_];
    DT_ v[W_];
#pragma unroll
    for(u32 k = 0; k < W_; ++k)
    {
      c[k] = M.col[u64(k) * n + r];
      v[k] = dev::ld<DT_>(M.val, u64(k) * n + r);
    }
    DT_ s = DT_(0);
#pragma unroll
    for(u32 k = 0; k < W_; ++k) s += v[k] * xs[c[k]];
    return s;
  }

  template<typename DT_, typename VT_>
  __device__ __forceinline__ DT_ tail_row(const TailEll<VT_>& M, u32 r, const DT_* xs)
  {
    if constexpr(sizeof(DT_) == 4)
    {
      switch(M.w)
      {
        case 8: return tail_row_w<8, DT_, VT_>(M, r, xs);
        case 9: return tail_row_w<9, DT_, VT_>(M, r, xs);
        case 16: return tail_row_w<16, DT_, VT_>(M, r, xs);
        case 21: return tail_row_w<21, DT_, VT_>(M, r, xs);
        case 24: return tail_row_w<24, DT_, VT_>(M, r, xs);
        case 25: return tail_row_w<25, DT_, VT_>(M, r, xs);
        case 27: return tail_row_w<27, DT_, VT_>(M, r, xs);
        case 32: return tail_row_w<32, DT_, VT_>(M, r, xs);
        default: break;
      }
    }
    DT_ s = DT_(0);
    const u64 n = M.n;
#pragma unroll 9
    for(u32 k = 0; k < M.w; ++k)
    {
      const u64 at = u64(k) * n + r;
      s += dev::ld<DT_>(M.val, at) * xs[M.col[at]];
    }
    return s;
  }

  // ops[0, nops): the segment's steps; out_buf: the buffer of tail[0]'s iterate at the end
  template<typename DT_, typename VT_>
  __global__ __launch_bounds__(kTailThreads) void k_tail(TailDesc<DT_, VT_> d, const u32* __restrict__ ops, u32 nops,
                                                         u32 out_buf, const int* status)
  {
    if(stopped(status)) return;
    extern __shared__ double tail_arena_[];
    DT_* arena = reinterpret_cast<DT_*>(tail_arena_);
    const u32 tid = threadIdx.x;
    // vectors of level l at arena + off: [0,n) x0, [n,2n) x1, [2n,3n) rhs, [3n,4n) def, [4n,5n) dinv
    for(int i = 0; i < d.nlev; ++i)
    {
      const TailLevel<DT_, VT_>& l = d.lv[i];
      DT_* v = arena + l.off;
      for(u32 r = tid; r < l.n; r += kTailThreads)
      {
        const DT_ di = l.dinv[r];
        v[4 * l.n + r] = di;
        if(i == 0)
        {
          // the first smoothing step from x = 0 (as the restriction above writes it)
          const DT_ b = d.rhs_in[r];
          v[2 * l.n + r] = b;
          v[r] = di * b;
        }
      }
    }
    __syncthreads();
    for(u32 o = 0; o < nops; ++o)
    {
      const u32 op = ops[o];
      const u32 kind = op & 15u, li = (op >> 4) & 15u, a = (op >> 8) & 1u, b = (op >> 9) & 1u;
      const TailLevel<DT_, VT_>& f = d.lv[li];
      DT_* fv = arena + f.off;
      const u32 n = f.n;
      if(kind == kTailSweep)
      {
        const DT_* xi = fv + a * n;
        DT_* xo = fv + (a ^ 1u) * n;
        for(u32 r = tid; r < n; r += kTailThreads)
          xo[r] = xi[r] + fv[4 * n + r] * (fv[2 * n + r] - tail_row<DT_, VT_>(f.A, r, xi));
      }
      else if(kind == kTailDefect)
      {
        const DT_* xc = fv + a * n;
        for(u32 r = tid; r < n; r += kTailThreads)
        {
          const DT_ s = tail_row<DT_, VT_>(f.A, r, xc);
          fv[3 * n + r] = dev::masked(f.mask, r) ? DT_(0) : fv[2 * n + r] - s;
        }
      }
      else if(kind == kTailRest)
      {
        const TailLevel<DT_, VT_>& c = d.lv[li + 1];
        DT_* cv = arena + c.off;
        const bool first_step = ((op >> 10) & 1u) != 0u;
        for(u32 r = tid; r < c.n; r += kTailThreads)
        {
          const DT_ s = tail_row<DT_, VT_>(f.R, r, fv + 3 * n);
          const DT_ v = dev::masked(c.mask, r) ? DT_(0) : s;
          cv[2 * c.n + r] = v;
          if(first_step) cv[r] = cv[4 * c.n + r] * v;
        }
      }
      else if(kind == kTailCoarse)
      {
        // exact coarse solve with the FP64 inverse of the filtered coarsest matrix
        for(u32 i = tid; i < d.ncoarse; i += kTailThreads)
        {
          double s = 0.0;
#pragma unroll 8
          for(u32 j = 0; j < d.ncoarse; ++j) s += d.ainv[u64(i) * d.ncoarse + j] * double(fv[2 * n + j]);
          fv[i] = dev::masked(f.mask, i) ? DT_(0) : DT_(s);
        }
      }
      else
      {
        const TailLevel<DT_, VT_>& c = d.lv[li + 1];
        const DT_* xc = arena + c.off + b * c.n;
        DT_* xcur = fv + a * n;
        for(u32 r = tid; r < n; r += kTailThreads)
        {
          const DT_ s = tail_row<DT_, VT_>(f.P, r, xc);
          if(!dev::masked(f.mask, r)) xcur[r] = xcur[r] + s;
        }
      }
      __syncthreads();
    }
    const TailLevel<DT_, VT_>& t = d.lv[0];
    const DT_* xf = arena + t.off + out_buf * t.n;
    for(u32 r = tid; r < t.n; r += kTailThreads) d.x_out[r] = xf[r];
  }

  // ---- setup kernels -----------------------------------------------------------------------

  // dinv = omega_R * (omega_J / diag), 0 on Dirichlet rows (diagonal_generic.hpp:20-37,
  // component_invert_generic.hpp:21-36)
  template<typename DT_, typename VT_>
  __global__ void k_make_dinv(u64 rows, const u32* rp, const u32* col, const VT_* val, const u32* mask,
                              DT_ omega_j, DT_ omega_r, DT_* dinv)
  {
    const u64 row = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if(row >= rows) return;
    DT_ d = DT_(0);
    for(u32 i = rp[row]; i < rp[row + 1]; ++i)
      if(u64(col[i]) == row) { d = dev::ld<DT_>(val, i); break; }
    const DT_ inv = omega_j / d;
    dinv[row] = dev::masked(mask, row) ? DT_(0) : omega_r * inv;
  }

  // ---- PCG state and kernels (FP64) ----------------------------------------------------------

  // device-resident PCG scalars and the criteria of IterativeSolver (iterative.hpp:213-237)
  struct PcgState
  {
    double gamma, alpha, beta, def_init, def_cur, def_prev;
    double tol_rel, tol_abs, tol_abs_low, div_rel, div_abs, stag_rate;
    u64 num_iter, min_iter, max_iter, min_stag_iter, num_stag_iter, hist_cap;
    int status;
    int first;   // set with the initial defect: the next beta finalisation starts the recurrence
  };

  // _set_initial_defect (iterative.hpp:744-766)
  __device__ inline void set_initial_defect(PcgState* st, double* hist, double d)
  {
    st->def_init = st->def_cur = st->def_prev = d;
    st->num_iter = 0;
    st->num_stag_iter = 0;
    st->first = 1;
    if(st->hist_cap > 0) hist[0] = d;
    const double eps = 2.220446049250313e-16;
    if(!isfinite(d)) st->status = kAborted;
    else if(d <= eps * eps) st->status = kSuccess;
    else st->status = kProgress;
  }

  // _set_new_defect / _analyse_defect (iterative.hpp:791-879)
  __device__ inline void set_new_defect(PcgState* st, double* hist, double d)
  {
    ++st->num_iter;
    st->def_prev = st->def_cur;
    // with min_iter == max_iter and no stagnation test the reference skips the norm
    if((st->min_iter < st->max_iter) || (st->min_stag_iter > 0))
      st->def_cur = d;
    d = st->def_cur;
    if(st->num_iter < st->hist_cap) hist[st->num_iter] = d;
    int s = kProgress;
    if(!isfinite(d)) s = kAborted;
    else if((d > st->div_abs) || (d > st->div_rel * st->def_init)) s = kDiverged;
    else if(st->num_iter < st->min_iter) s = kProgress;
    else if((d <= st->tol_abs) && ((d <= st->tol_rel * st->def_init) || (d <= st->tol_abs_low))) s = kSuccess;
    else if(st->num_iter >= st->max_iter) s = kMaxIter;
    else if(st->min_stag_iter > 0)
    {
      if(d >= st->stag_rate * st->def_prev)
      {
        if(++st->num_stag_iter >= st->min_stag_iter) s = kStagnated;
      }
      else
        st->num_stag_iter = 0;
    }
    st->status = s;
  }

  // The reducing PCG kernels store one partial per block; k_pcg_fin (one block, launched next
  // on the stream) sums them in block order and updates the device scalars.

  // r = filter(b - A x)  (pcg.hpp:159-173 correct), partial sums of r.r
  template<typename Rows_>
  __global__ __launch_bounds__(256) F3_ROW_ATTR void k_pcg_defect_norm(Rows_ A, u64 per, const u32* __restrict__ mask, const double* __restrict__ b,
                                                           const double* __restrict__ x, double* __restrict__ r, double* partials)
  {
    double v = 0.0;
    for_rows(A, per, x, [&](u64 p, double s, bool w) {
      const u32 i = u32(A.row(p));
      if(w)
      {
        const double bs = b[i] - s;
        const double d = dev::masked_nn(mask, i) ? 0.0 : bs;
        r[i] = d;
        v += d * d;
      }
    });
    dev::block_partial(v, partials);
  }

  // r = def  (pcg.hpp:141-156 apply), partial sums of r.r
  static __global__ __launch_bounds__(256) void k_pcg_copy_norm(u64 n, const double* __restrict__ def, double* __restrict__ r,
                                                         double* partials)
  {
    double v = 0.0;
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
    {
      const double d = def[i];
      r[i] = d;
      v += d * d;
    }
    dev::block_partial(v, partials);
  }

  // p = z (correction converted back to FP64, convert_precond.hpp:115), partial sums of r.p
  template<typename DT_>
  __global__ __launch_bounds__(256) void k_pcg_p_gamma(u64 n, const DT_* __restrict__ z, const double* __restrict__ r,
                                                       double* __restrict__ p, const PcgState* st, double* partials)
  {
    if(st->status != kProgress) return;
    double v = 0.0;
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
    {
      const double zi = double(z[i]);
      p[i] = zi;
      v += r[i] * zi;
    }
    dev::block_partial(v, partials);
  }

  // q = filter(A p), partial sums of q.p   (pcg.hpp:236-242)
  template<typename Rows_>
  __global__ __launch_bounds__(256) F3_ROW_ATTR void k_pcg_spmv_dot(Rows_ A, u64 per, const u32* __restrict__ mask, const double* __restrict__ p,
                                                        double* __restrict__ q, const PcgState* st, double* partials)
  {
    if(st->status != kProgress) return;
    double v = 0.0;
    for_rows(A, per, p, [&](u64 pos, double s, bool w) {
      const u64 row = A.row(pos);
      if(w)
      {
        const u32 i = u32(row);
        const double qi = dev::masked_nn(mask, i) ? 0.0 : s;
        q[i] = qi;
        v += qi * p[i];
      }
    });
    dev::block_partial(v, partials);
  }

  // x += alpha p, r -= alpha q, partial sums of r.r   (pcg.hpp:244-259)
  static __global__ __launch_bounds__(256) void k_pcg_update_norm(u64 n, const double* __restrict__ p, const double* __restrict__ q,
                                                           double* __restrict__ x, double* __restrict__ r, const PcgState* st,
                                                           double* partials)
  {
    if(st->status != kProgress) return;
    const double alpha = st->alpha;
    // axpy with |alpha| < eps is a copy; written as a negation so that a NaN alpha updates (and the
    // NaN reaches the defect norm) as in dense_vector.hpp:891
    const bool upd = !(fabs(alpha) < 2.220446049250313e-16);
    double v = 0.0;
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
    {
      double ri = r[i];
      if(upd)
      {
        x[i] = x[i] + alpha * p[i];
        ri = ri + (-alpha) * q[i];
        r[i] = ri;
      }
      v += ri * ri;
    }
    dev::block_partial(v, partials);
  }

  // partial sums of r.z   (pcg.hpp:270-279)
  template<typename DT_>
  __global__ __launch_bounds__(256) void k_pcg_gamma_beta(u64 n, const DT_* __restrict__ z, const double* __restrict__ r,
                                                          const PcgState* st, double* partials)
  {
    if(st->status != kProgress) return;
    double v = 0.0;
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      v += r[i] * double(z[i]);
    dev::block_partial(v, partials);
  }

  enum : int { kFinInitDefect = 0, kFinAlpha = 1, kFinNewDefect = 2, kFinGamma = 3, kFinBeta = 4 };

  // finalisation of one PCG reduction (one block of 256 threads)
  template<int OP>
  __global__ __launch_bounds__(256) void k_pcg_fin(const double* __restrict__ partials, unsigned nb, PcgState* st, double* hist)
  {
    if(OP != kFinInitDefect && st->status != kProgress) return;
    const double t = dev::sum_partials(partials, nb);
    if(threadIdx.x != 0) return;
    if(OP == kFinInitDefect) set_initial_defect(st, hist, sqrt(t));        // iterative.hpp:744-766
    else if(OP == kFinAlpha) st->alpha = st->gamma / t;                     // pcg.hpp:240
    else if(OP == kFinNewDefect) set_new_defect(st, hist, sqrt(t));         // iterative.hpp:848-879
    else if(OP == kFinGamma) st->gamma = t;                                 // pcg.hpp:229
    else if(st->first)
    {
      // the initial preconditioner application: gamma = r.z, p = z (pcg.hpp:225-231)
      st->gamma = t;
      st->beta = 0.0;
      st->first = 0;
    }
    else { const double g2 = st->gamma; st->gamma = t; st->beta = t / g2; } // pcg.hpp:273-279
  }

  // Publication of the solver state to the host (one thread): publication c of a solve (c counted
  // on the device in *count) goes to slot c % R of a ring in pinned host memory, and after a
  // system-scope fence its sequence number c + 1 to seq[c % R], which the host polls.  It runs
  // whatever the status, so every publication the host launched arrives.
  template<int R>
  __global__ void k_publish(const PcgState* __restrict__ st, PcgState* ring, unsigned* seq, unsigned* count)
  {
    if(blockIdx.x != 0 || threadIdx.x != 0) return;
    const unsigned c = *count;
    ring[c % R] = *st;
    __threadfence_system();
    reinterpret_cast<volatile unsigned*>(seq)[c % R] = c + 1;
    *count = c + 1;
  }

  // p = beta p + z  (axpy with r == x: r = r*a + y; |beta| < eps: copy)   (pcg.hpp:281-283)
  template<typename DT_>
  __global__ void k_pcg_update_p(u64 n, const DT_* __restrict__ z, double* __restrict__ p, const PcgState* st)
  {
    if(st->status != kProgress) return;
    const double beta = st->beta;
    const bool copy = fabs(beta) < 2.220446049250313e-16;
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      p[i] = copy ? double(z[i]) : p[i] * beta + double(z[i]);
  }

  // unpreconditioned CG: the "preconditioner" copies and filters (iterative.hpp:1070-1075)
  static __global__ void k_copy_filter(u64 n, const double* __restrict__ r, const u32* __restrict__ mask, double* __restrict__ z,
                                const int* status)
  {
    if(stopped(status)) return;
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      z[i] = dev::masked(mask, i) ? 0.0 : r[i];
  }
} // namespace mg
} // namespace feat3_amd
