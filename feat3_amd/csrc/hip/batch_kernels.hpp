// batch_kernels.hpp -- kernels of the batched device MG-PCG: K systems in lockstep that share
// one sparsity graph, with B x B blocks per graph entry (B = 1 scalar, B = 3 the vector
// Laplacian of SparseMatrixBCSR<3,3>, sparse_matrix_bcsr.hpp:1295-1600), and values either
// shared by all K systems (K right-hand sides) or one set per system (K coefficient fields).
//
// Vectors are stored [node][component][system] (system fastest).  The row kernels give every
// lane S consecutive systems (one 16-byte vector access: 4 floats or 2 doubles) and a row
// G = K/S lanes, so a wave keeps 64/G rows and their x gathers in flight; a block's values are
// read once per lane group for all of its S systems per lane.  Every system keeps its own PCG
// scalars and status; a converged system's vectors are frozen, so each system performs
// exactly the iterations of its independent solve (pcg.hpp:188-288, iterative.hpp:744-879).
#pragma once

#include "device_common.hpp"
#include "mg_kernels.hpp"

#include <type_traits>

namespace feat3_amd
{
namespace batch
{
  typedef std::uint32_t u32;
  typedef std::uint64_t u64;
  using mg::RowGrid;

  // ---- S-wide accesses ---------------------------------------------------------------------
  template<typename T, int S_>
  __device__ __forceinline__ void ldv(const T* p, T (&o)[S_])
  {
    if constexpr(S_ == 1)
      o[0] = p[0];
    else
    {
      typedef T V __attribute__((ext_vector_type(S_)));
      const V v = *reinterpret_cast<const V*>(p);
#pragma unroll
      for(int i = 0; i < S_; ++i) o[i] = v[i];
    }
  }

  template<typename T, int S_>
  __device__ __forceinline__ void stv(T* p, const T (&o)[S_])
  {
    if constexpr(S_ == 1)
      p[0] = o[0];
    else
    {
      typedef T V __attribute__((ext_vector_type(S_)));
      V v;
#pragma unroll
      for(int i = 0; i < S_; ++i) v[i] = o[i];
      *reinterpret_cast<V*>(p) = v;
    }
  }

  template<typename DT_, typename VT_>
  __device__ __forceinline__ DT_ cvt(VT_ v) { return DT_(v); }
  template<>
  __device__ __forceinline__ float cvt<float, dev::half>(dev::half v) { return float(v); }

  template<typename DT_, typename VT_, int S_>
  __device__ __forceinline__ void ldvc(const VT_* p, DT_ (&o)[S_])
  {
    VT_ t[S_];
    ldv<VT_, S_>(p, t);
#pragma unroll
    for(int i = 0; i < S_; ++i) o[i] = cvt<DT_, VT_>(t[i]);
  }

  // ---- row evaluators: s[h][i] = sum_e sum_w A_e[h][w] x[col_e][w][k0 + i] -------------------
  // Entries are taken in batches of kU: the column loads of a batch, then its x gathers, are in
  // flight together instead of one dependent col -> x round trip per entry.
#ifndef F3_BATCH_UNROLL
#define F3_BATCH_UNROLL 4
#endif
  constexpr int kU = F3_BATCH_UNROLL;

  // DIAG_ (shared values, B > 1): every block is diagonal, and only its diagonal is stored,
  // [entry][h]; the off-diagonal entries of the blocks are exact zeros, so the row sums take the
  // terms of the full blocks that can be nonzero, in the same order
  template<typename DT_, typename VT_, int B_, bool SHARED_, int S_, bool DIAG_ = false>
  struct BlockRows
  {
    const u32* rp;
    const u32* col;
    const VT_* val;   // [entry][h][w] (shared), [entry][h] (shared, DIAG_) or [entry][k] (per system, B = 1)
    u64 rows;
    int K;
    const u32* ord;   // processing order of the rows (null: ascending)
    static constexpr bool kShared = SHARED_;
    static constexpr int kVpe = DIAG_ ? B_ : B_ * B_;   // stored values per entry (shared values)

    __device__ __forceinline__ void acc(u32 e, int k0, const DT_ (&xv)[B_][S_], DT_ (&s)[B_][S_]) const
    {
      if constexpr(SHARED_ && DIAG_)
      {
#pragma unroll
        for(int h = 0; h < B_; ++h)
        {
          const DT_ a = dev::ld<DT_>(val, u64(e) * B_ + u64(h));
#pragma unroll
          for(int i = 0; i < S_; ++i) s[h][i] += a * xv[h][i];
        }
      }
      else if constexpr(SHARED_)
      {
#pragma unroll
        for(int h = 0; h < B_; ++h)
#pragma unroll
          for(int w = 0; w < B_; ++w)
          {
            const DT_ a = dev::ld<DT_>(val, (u64(e) * B_ + u64(h)) * B_ + u64(w));
#pragma unroll
            for(int i = 0; i < S_; ++i) s[h][i] += a * xv[w][i];
          }
      }
      else
      {
        static_assert(B_ == 1, "per-system values are scalar");
        DT_ a[S_];
        ldvc<DT_, VT_, S_>(val + u64(e) * u64(K) + u64(k0), a);
#pragma unroll
        for(int i = 0; i < S_; ++i) s[0][i] += a[i] * xv[0][i];
      }
    }

    // shared values of entry e from a block-local copy (entries e0.. staged in LDS)
    __device__ __forceinline__ void acc_lds(u32 e, u32 e0, const VT_* lds, const DT_ (&xv)[B_][S_], DT_ (&s)[B_][S_]) const
    {
      if constexpr(DIAG_)
      {
#pragma unroll
        for(int h = 0; h < B_; ++h)
        {
          const DT_ a = cvt<DT_, VT_>(lds[(e - e0) * u32(B_) + u32(h)]);
#pragma unroll
          for(int i = 0; i < S_; ++i) s[h][i] += a * xv[h][i];
        }
        return;
      }
#pragma unroll
      for(int h = 0; h < B_; ++h)
#pragma unroll
        for(int w = 0; w < B_; ++w)
        {
          const DT_ a = cvt<DT_, VT_>(lds[(e - e0) * u32(B_ * B_) + u32(h * B_ + w)]);
#pragma unroll
          for(int i = 0; i < S_; ++i) s[h][i] += a * xv[w][i];
        }
    }

    // LDS != nullptr: the values of entries [e0, ...) are staged there (shared values only)
    __device__ __forceinline__ void row_sum(u64 p, int k0, const DT_* __restrict__ x, DT_ (&s)[B_][S_],
                                            const VT_* lds = nullptr, u32 e0 = 0) const
    {
#pragma unroll
      for(int h = 0; h < B_; ++h)
#pragma unroll
        for(int i = 0; i < S_; ++i) s[h][i] = DT_(0);
      const u32 end = rp[p + 1];
      u32 e = rp[p];
      for(; e + kU <= end; e += kU)
      {
        u64 j[kU];
#pragma unroll
        for(int u = 0; u < kU; ++u) j[u] = col[e + u];
        DT_ xv[kU][B_][S_];
#pragma unroll
        for(int u = 0; u < kU; ++u)
#pragma unroll
          for(int w = 0; w < B_; ++w) ldv<DT_, S_>(x + (j[u] * B_ + u64(w)) * u64(K) + u64(k0), xv[u][w]);
        if constexpr(SHARED_)
        {
          if(lds)
          {
#pragma unroll
            for(int u = 0; u < kU; ++u) acc_lds(e + u, e0, lds, xv[u], s);
            continue;
          }
        }
#pragma unroll
        for(int u = 0; u < kU; ++u) acc(e + u, k0, xv[u], s);
      }
      for(; e < end; ++e)
      {
        const u64 j = col[e];
        DT_ xv[B_][S_];
#pragma unroll
        for(int w = 0; w < B_; ++w) ldv<DT_, S_>(x + (j * B_ + u64(w)) * u64(K) + u64(k0), xv[w]);
        if constexpr(SHARED_)
        {
          if(lds) { acc_lds(e, e0, lds, xv, s); continue; }
        }
        acc(e, k0, xv, s);
      }
    }
  };

  // scalar graph applied per component (the transfers: csrsb, apply_generic.hpp:156-181)
  template<typename DT_, typename VT_, int B_, int S_>
  struct CompRows
  {
    const u32* rp;
    const u32* col;
    const VT_* val;
    u64 rows;
    int K;
    const u32* ord;   // processing order of the rows (null: ascending)
    static constexpr bool kShared = false;

    __device__ __forceinline__ void row_sum(u64 p, int k0, const DT_* __restrict__ x, DT_ (&s)[B_][S_]) const
    {
#pragma unroll
      for(int h = 0; h < B_; ++h)
#pragma unroll
        for(int i = 0; i < S_; ++i) s[h][i] = DT_(0);
      const u32 end = rp[p + 1];
      u32 e = rp[p];
      for(; e + kU <= end; e += kU)
      {
        u64 j[kU];
        DT_ v[kU];
#pragma unroll
        for(int u = 0; u < kU; ++u) { j[u] = col[e + u]; v[u] = dev::ld<DT_>(val, e + u); }
        DT_ xv[kU][B_][S_];
#pragma unroll
        for(int u = 0; u < kU; ++u)
#pragma unroll
          for(int h = 0; h < B_; ++h) ldv<DT_, S_>(x + (j[u] * B_ + u64(h)) * u64(K) + u64(k0), xv[u][h]);
#pragma unroll
        for(int u = 0; u < kU; ++u)
#pragma unroll
          for(int h = 0; h < B_; ++h)
#pragma unroll
            for(int i = 0; i < S_; ++i) s[h][i] += v[u] * xv[u][h][i];
      }
      for(; e < end; ++e)
      {
        const u64 j = col[e];
        const DT_ v = dev::ld<DT_>(val, e);
#pragma unroll
        for(int h = 0; h < B_; ++h)
        {
          DT_ xv[S_];
          ldv<DT_, S_>(x + (j * B_ + u64(h)) * u64(K) + u64(k0), xv);
#pragma unroll
          for(int i = 0; i < S_; ++i) s[h][i] += v * xv[i];
        }
      }
    }
  };

  // calls f(row, k0, s[B][S]) for the rows of this block: G lanes per row, lane l handles the
  // systems k0 = l*S .. k0+S-1; lanes with k0 >= K are idle.  Block b takes positions
  // [b*per, (b+1)*per) of the processing order (XCD-remapped, so one XCD works through one
  // contiguous part of it).
  template<int G_, int S_, int B_, typename Rows_, typename DT_, typename F>
  __device__ __forceinline__ void for_block_rows(const Rows_& A, u64 per, const DT_* __restrict__ x, F&& f)
  {
    const u64 b = dev::xcd_remap(blockIdx.x, gridDim.x);
    const u64 beg = b * per;
    const u64 end = beg + per < A.rows ? beg + per : A.rows;
    const int k0 = int(threadIdx.x % G_) * S_;
    if(k0 >= A.K) return;
    for(u64 q = beg + threadIdx.x / G_; q < end; q += blockDim.x / G_)
    {
      const u64 p = A.ord ? u64(A.ord[q]) : q;
      DT_ s[B_][S_];
      A.row_sum(p, k0, x, s);
      f(p, k0, s);
    }
  }

  // The same loop with the shared block values of each round of rows staged in LDS: a block
  // runs 256/G rows at a time; their entries [rp[q0], rp[q1]) are contiguous, so the block copies
  // their values with coalesced loads (each value once) instead of every lane of a row's group
  // loading all B*B values of every entry.  A round with more than CAP entries, a permuted row
  // order or per-system values use the direct loop.  All threads of the block take part in the
  // copies and barriers; lanes with k0 >= K only help copying.
  template<int G_, int S_, int B_, int CAP_, typename Rows_, typename DT_, typename VT_, typename F>
  __device__ __forceinline__ void for_block_rows_staged(const Rows_& A, u64 per, const DT_* __restrict__ x, VT_* lds, F&& f)
  {
    constexpr u32 BB = u32(Rows_::kVpe);
    constexpr unsigned R = 256 / G_;
    const u64 b = dev::xcd_remap(blockIdx.x, gridDim.x);
    const u64 beg = b * per;
    const u64 end = beg + per < A.rows ? beg + per : A.rows;
    const int k0 = int(threadIdx.x % G_) * S_;
    const unsigned g = threadIdx.x / G_;
    for(u64 q0 = beg; q0 < end; q0 += R)
    {
      const u64 q1 = q0 + R < end ? q0 + R : end;
      const u32 e0 = A.rp[q0], e1 = A.rp[q1];
      const bool staged = (e1 - e0) <= u32(CAP_);
      __syncthreads();   // the previous round's readers are done with lds
      if(staged)
        for(u32 i = threadIdx.x; i < (e1 - e0) * BB; i += blockDim.x) lds[i] = A.val[u64(e0) * BB + i];
      __syncthreads();
      const u64 p = q0 + g;
      if(p < q1 && k0 < A.K)
      {
        DT_ s[B_][S_];
        A.row_sum(p, k0, x, s, staged ? lds : nullptr, e0);
        f(p, k0, s);
      }
    }
  }

  __device__ __forceinline__ bool masked(const u32* mask, u64 node) { return dev::masked(mask, node); }
  __device__ __forceinline__ bool idle(const int* any) { return any != nullptr && *any == 0; }

  // dinv layout: [node][h] (shared values) or [node][h][k] (one matrix per system)
  template<int B_, bool SHARED_, int S_, typename DT_>
  __device__ __forceinline__ void ld_dinv(const DT_* dinv, u64 node, int h, int k0, int K, DT_ (&d)[S_])
  {
    if constexpr(SHARED_)
    {
      const DT_ v = dinv[node * B_ + u64(h)];
#pragma unroll
      for(int i = 0; i < S_; ++i) d[i] = v;
    }
    else
      ldv<DT_, S_>(dinv + (node * B_ + u64(h)) * u64(K) + u64(k0), d);
  }

  // Row kernels keep kU entries' gathers in flight per lane.  Left to itself the compiler keeps
  // 8 waves per SIMD by issuing those loads one after another (fewer live registers); the
  // bound below lets it spend registers on memory-level parallelism instead.
#ifndef F3_BATCH_WAVES
#define F3_BATCH_WAVES 4
#endif
#define F3_ROW_KERNEL __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, F3_BATCH_WAVES)))

  // ---- V-cycle -------------------------------------------------------------------------------
  // rhs = (inner) def, x = dinv * rhs (the first smoothing step from zero)
  template<typename DT_, int B_, bool SHARED_>
  __global__ void k_entry(u64 nodes, int K, const double* __restrict__ def, DT_* __restrict__ rhs, DT_* __restrict__ x,
                          const DT_* __restrict__ dinv, const int* any)
  {
    if(idle(any)) return;
    const u64 n = nodes * B_ * u64(K);
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
    {
      const u64 k = i % u64(K);
      const u64 nh = i / u64(K);
      const DT_ r = DT_(def[i]);
      rhs[i] = r;
      x[i] = dinv[SHARED_ ? nh : nh * u64(K) + k] * r;
    }
  }

  // entries of shared block values staged per round (LDS bytes: CAP * B * B * sizeof(value)).
  // Measured on the 32-RHS 3x3 sweep (3D Q1 L6): staging FP64 values 725 -> 507 us (outer FP64
  // SpMV 643 -> 490 us); staging FP32 values 260 -> 307 us (the 36 KB of LDS per block cost more
  // occupancy than the broadcast loads they replace), so only 8-byte values are staged.
  template<int VPE_, typename VT_>
  struct Stage
  {
    static constexpr bool enabled = sizeof(VT_) == 8;
    static constexpr int kBytes = 36 * 1024;
    static constexpr int CAP = kBytes / int(VPE_ * sizeof(VT_));   // VPE_: stored values per entry
  };

  // runs the row loop of a kernel, staged when the values are shared and the order ascending
  template<int G_, int S_, int B_, typename Rows_, typename DT_, typename F>
  __device__ __forceinline__ void row_loop(const Rows_& A, u64 per, const DT_* __restrict__ x, F&& f)
  {
    typedef typename std::remove_const<typename std::remove_pointer<decltype(A.val)>::type>::type VT;
    if constexpr(Rows_::kShared && Stage<Rows_::kVpe, VT>::enabled)
    {
      constexpr int CAP = Stage<Rows_::kVpe, VT>::CAP;
      __shared__ VT lds[CAP * Rows_::kVpe];
      if(A.ord == nullptr)
      {
        for_block_rows_staged<G_, S_, B_, CAP>(A, per, x, lds, f);
        return;
      }
    }
    for_block_rows<G_, S_, B_>(A, per, x, f);
  }

  // x_out = x_in + dinv (rhs - A x_in)
  template<int G_, int S_, int B_, bool SHARED_, typename Rows_, typename DT_>
  F3_ROW_KERNEL void k_sweep(Rows_ A, u64 per, const DT_* __restrict__ dinv, const DT_* __restrict__ rhs,
                                                 const DT_* __restrict__ xin, DT_* __restrict__ xout, const int* any)
  {
    if(idle(any)) return;
    const int K = A.K;
    row_loop<G_, S_, B_>(A, per, xin, [&](u64 p, int k0, const DT_ (&s)[B_][S_]) {
#pragma unroll
      for(int h = 0; h < B_; ++h)
      {
        const u64 i0 = (p * B_ + u64(h)) * u64(K) + u64(k0);
        DT_ xi[S_], ri[S_], di[S_], xo[S_];
        ldv<DT_, S_>(xin + i0, xi);
        ldv<DT_, S_>(rhs + i0, ri);
        ld_dinv<B_, SHARED_, S_>(dinv, p, h, k0, K, di);
#pragma unroll
        for(int i = 0; i < S_; ++i) xo[i] = xi[i] + di[i] * (ri[i] - s[h][i]);
        stv<DT_, S_>(xout + i0, xo);
      }
    });
  }

  // def = filter(rhs - A x)
  template<int G_, int S_, int B_, typename Rows_, typename DT_>
  F3_ROW_KERNEL void k_defect(Rows_ A, u64 per, const u32* __restrict__ mask, const DT_* __restrict__ rhs,
                                                  const DT_* __restrict__ x, DT_* __restrict__ def, const int* any)
  {
    if(idle(any)) return;
    const int K = A.K;
    row_loop<G_, S_, B_>(A, per, x, [&](u64 p, int k0, const DT_ (&s)[B_][S_]) {
      const bool m = masked(mask, p);
#pragma unroll
      for(int h = 0; h < B_; ++h)
      {
        const u64 i0 = (p * B_ + u64(h)) * u64(K) + u64(k0);
        DT_ ri[S_], d[S_];
        ldv<DT_, S_>(rhs + i0, ri);
#pragma unroll
        for(int i = 0; i < S_; ++i) d[i] = m ? DT_(0) : ri[i] - s[h][i];
        stv<DT_, S_>(def + i0, d);
      }
    });
  }

  // rhs_c = filter(R def_f); x_c = dinv_c rhs_c (first coarse smoothing step) unless x_c is null
  template<int G_, int S_, int B_, bool SHARED_, typename Rows_, typename DT_>
  F3_ROW_KERNEL void k_restrict(Rows_ R, u64 per, const u32* __restrict__ mask_c, const DT_* __restrict__ def_f,
                                                    DT_* __restrict__ rhs_c, DT_* __restrict__ x_c, const DT_* __restrict__ dinv_c,
                                                    const int* any)
  {
    if(idle(any)) return;
    const int K = R.K;
    for_block_rows<G_, S_, B_>(R, per, def_f, [&](u64 p, int k0, const DT_ (&s)[B_][S_]) {
      const bool m = masked(mask_c, p);
#pragma unroll
      for(int h = 0; h < B_; ++h)
      {
        const u64 i0 = (p * B_ + u64(h)) * u64(K) + u64(k0);
        DT_ r[S_];
#pragma unroll
        for(int i = 0; i < S_; ++i) r[i] = m ? DT_(0) : s[h][i];
        stv<DT_, S_>(rhs_c + i0, r);
        if(x_c)
        {
          DT_ d[S_];
          ld_dinv<B_, SHARED_, S_>(dinv_c, p, h, k0, K, d);
#pragma unroll
          for(int i = 0; i < S_; ++i) d[i] *= r[i];
          stv<DT_, S_>(x_c + i0, d);
        }
      }
    });
  }

  // x_f += filter(P x_c)
  template<int G_, int S_, int B_, typename Rows_, typename DT_>
  F3_ROW_KERNEL void k_prolong(Rows_ P, u64 per, const u32* __restrict__ mask_f, const DT_* __restrict__ x_c,
                                                   DT_* __restrict__ x_f, const int* any)
  {
    if(idle(any)) return;
    const int K = P.K;
    for_block_rows<G_, S_, B_>(P, per, x_c, [&](u64 p, int k0, const DT_ (&s)[B_][S_]) {
      if(masked(mask_f, p)) return;
#pragma unroll
      for(int h = 0; h < B_; ++h)
      {
        const u64 i0 = (p * B_ + u64(h)) * u64(K) + u64(k0);
        DT_ xf[S_];
        ldv<DT_, S_>(x_f + i0, xf);
#pragma unroll
        for(int i = 0; i < S_; ++i) xf[i] = xf[i] + s[h][i];
        stv<DT_, S_>(x_f + i0, xf);
      }
    });
  }

  // exact coarse solve: x[:, k] = Ainv_k * rhs[:, k] with FP64 inverses of the unit-filtered
  // coarse matrices (one shared inverse, or one per system); m = nodes * B rows
  template<typename DT_>
  __global__ __launch_bounds__(256) void k_coarse(u64 m, int K, int per_system, const double* __restrict__ ainv,
                                                  const u32* __restrict__ mask, int B, const DT_* __restrict__ rhs,
                                                  DT_* __restrict__ x, const int* any)
  {
    if(idle(any)) return;
    for(u64 t = u64(blockIdx.x) * blockDim.x + threadIdx.x; t < m * u64(K); t += u64(gridDim.x) * blockDim.x)
    {
      const u64 i = t / u64(K);
      const int k = int(t % u64(K));
      const double* a = ainv + (per_system ? u64(k) * m * m : 0) + i * m;
      double s = 0.0;
      // unrolled so that a group of the row's loads is in flight at once (the sum keeps its order)
#pragma unroll 8
      for(u64 j = 0; j < m; ++j) s += a[j] * double(rhs[j * u64(K) + u64(k)]);
      x[i * u64(K) + u64(k)] = masked(mask, i / u64(B)) ? DT_(0) : DT_(s);
    }
  }

  // ---- coarse tail of the shared-values solver: one workgroup per system ----------------------
  //
  // Below about a thousand nodes a level step is a launch of 5-7 us that moves a few hundred KB.
  // k_tail_sys runs the V-cycle's part on the levels tail[0] (entered from the restriction above,
  // which wrote tail[0]'s right-hand side) down to the coarsest and back in one workgroup of
  // kTailSysThreads per system, with barriers between the steps.  The system's vectors on these
  // levels live in LDS ([node][h]: x0, x1, rhs, def, dinv; 5 * nodes * B values per level); the
  // shared matrices are read from global memory in ELL order (TailSysEll: the entries of a wave's
  // rows coalesced), where the K workgroups keep them L2-resident.
  // A thread takes a node and its B components.  Every step sums a row's terms in the order of
  // k_sweep / k_defect / k_restrict / k_prolong / k_coarse with the same expressions, so the
  // results are bit for bit those of the per-level launches.
  constexpr int kTailSysMax = 8;
  constexpr int kTailSysThreads = 1024;

  // A matrix of the tail in ELL order: entry k of row r at col[k * rows + r] and values
  // val[(k * V + t) * rows + r] (V values per entry as stored: [h] or [h][w]), k < w; rp gives
  // the row lengths (entries past a row's end are padding, column 0 and value 0, left out)
  template<typename VT_>
  struct TailSysEll
  {
    const u32* rp;
    const u32* col;
    const VT_* val;
    u32 rows, w;
  };

  template<typename DT_, typename VT_>
  struct TailSysLevel
  {
    u32 nodes, off;                // nodes; offset of the level's vectors in LDS (DT_ elements)
    TailSysEll<VT_> A, P, R;       // P: rows = this level's nodes; R: rows = the next coarser level's
    const u32* mask;
    const DT_* dinv;               // [node][h]
  };

  // ELL copy of a CSR matrix with V values per entry (TailSysEll), one thread per (row, k)
  template<typename VT_>
  __global__ void k_tail_ell(u32 rows, u32 w, u32 V, const u32* __restrict__ rp, const u32* __restrict__ col,
                             const VT_* __restrict__ val, u32* __restrict__ ecol, VT_* __restrict__ eval)
  {
    const u64 n = u64(rows) * w;
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
    {
      const u32 k = u32(i / rows), r = u32(i % rows);
      const u32 e = rp[r] + k;
      const bool in = e < rp[r + 1];
      ecol[i] = in ? col[e] : 0u;
      for(u32 t = 0; t < V; ++t) eval[(u64(k) * V + t) * rows + r] = in ? val[u64(e) * V + t] : VT_(0.0f);
    }
  }

  template<typename DT_, typename VT_>
  struct TailSysDesc
  {
    int nlev, pre, post;
    u32 ncoarse;                   // rows of the coarsest level (nodes * B)
    const double* ainv;            // FP64 inverse of the filtered coarsest matrix (shared)
    const DT_* rhs_in;             // tail[0]'s right-hand side, [node][h][k]
    DT_* x_out;                    // tail[0]'s final iterate, [node][h][k]
    TailSysLevel<DT_, VT_> lv[kTailSysMax];
  };

  // s[h] = sum_e A_e[h][.] x[col_e][.] over the row's entries in ascending order (BlockRows with
  // shared values); x in LDS, [node][w].  The entries are taken in batches of NB whose loads (one
  // coalesced access per entry and value across the rows of a wave) are all issued before the
  // first use; loads past the ELL width read its last column, and terms past the row's end are
  // left out.
  template<typename DT_, typename VT_, int B_, bool DIAG_>
  __device__ __forceinline__ void tail_sys_block_row(const TailSysEll<VT_>& M, u32 p, const DT_* x, DT_ (&s)[B_])
  {
    constexpr u32 V = DIAG_ ? u32(B_) : u32(B_ * B_);
    constexpr u32 NB = V * sizeof(DT_) <= 12 ? 16 : V * sizeof(DT_) <= 36 ? 8 : 4;   // registers of a batch
#pragma unroll
    for(int h = 0; h < B_; ++h) s[h] = DT_(0);
    const u32 len = M.rp[p + 1] - M.rp[p];
    const u64 n = M.rows;
    for(u32 k0 = 0; k0 < M.w; k0 += NB)
    {
      u32 j[NB];
      DT_ a[NB][V];
#pragma unroll
      for(u32 u = 0; u < NB; ++u)
      {
        const u64 k = k0 + u < M.w ? k0 + u : M.w - 1;
        j[u] = M.col[k * n + p];
#pragma unroll
        for(u32 t = 0; t < V; ++t) a[u][t] = dev::ld<DT_>(M.val, (k * V + t) * n + p);
      }
#pragma unroll
      for(u32 u = 0; u < NB; ++u)
      {
        if(k0 + u < len)
        {
          DT_ xv[B_];
#pragma unroll
          for(int w = 0; w < B_; ++w) xv[w] = x[j[u] * u32(B_) + u32(w)];
          if constexpr(DIAG_)
          {
#pragma unroll
            for(int h = 0; h < B_; ++h) s[h] += a[u][h] * xv[h];
          }
          else
          {
#pragma unroll
            for(int h = 0; h < B_; ++h)
#pragma unroll
              for(int w = 0; w < B_; ++w) s[h] += a[u][h * B_ + w] * xv[w];
          }
        }
      }
    }
  }

  // s[h] = sum_e v_e x[col_e][h] (a scalar transfer graph applied per component, CompRows), in
  // batches of 16 entries as above
  template<typename DT_, typename VT_, int B_>
  __device__ __forceinline__ void tail_sys_comp_row(const TailSysEll<VT_>& M, u32 p, const DT_* x, DT_ (&s)[B_])
  {
    constexpr u32 NB = 16;
#pragma unroll
    for(int h = 0; h < B_; ++h) s[h] = DT_(0);
    const u32 len = M.rp[p + 1] - M.rp[p];
    const u64 n = M.rows;
    for(u32 k0 = 0; k0 < M.w; k0 += NB)
    {
      u32 j[NB];
      DT_ v[NB];
#pragma unroll
      for(u32 u = 0; u < NB; ++u)
      {
        const u64 k = k0 + u < M.w ? k0 + u : M.w - 1;
        j[u] = M.col[k * n + p];
        v[u] = dev::ld<DT_>(M.val, k * n + p);
      }
#pragma unroll
      for(u32 u = 0; u < NB; ++u)
      {
        if(k0 + u < len)
        {
#pragma unroll
          for(int h = 0; h < B_; ++h) s[h] += v[u] * x[j[u] * u32(B_) + u32(h)];
        }
      }
    }
  }

  // grid: K workgroups of kTailSysThreads (workgroup k = system k); dynamic LDS: the levels'
  // vectors.  Level li's iterate alternates between its buffers x0 and x1; the coarsest level's
  // solution is x0.
  template<typename DT_, typename VT_, int B_, bool DIAG_>
  __global__ __launch_bounds__(kTailSysThreads) void k_tail_sys(TailSysDesc<DT_, VT_> d, int K, const int* any)
  {
    if(idle(any)) return;
    extern __shared__ double tail_sys_arena_[];
    DT_* arena = reinterpret_cast<DT_*>(tail_sys_arena_);
    const u32 tid = threadIdx.x;
    const u64 k = blockIdx.x;
    // dinv of every level; tail[0]'s right-hand side and first smoothing step x0 = dinv rhs
    for(int i = 0; i < d.nlev; ++i)
    {
      const TailSysLevel<DT_, VT_>& l = d.lv[i];
      const u32 n = l.nodes * u32(B_);
      DT_* v = arena + l.off;
      for(u32 r = tid; r < n; r += kTailSysThreads)
      {
        const DT_ di = l.dinv[r];
        v[4 * n + r] = di;
        if(i == 0)
        {
          const DT_ b = d.rhs_in[u64(r) * u64(K) + k];
          v[2 * n + r] = b;
          v[r] = di * b;
        }
      }
    }
    __syncthreads();
    // bit li: the buffer (0: x0, 1: x1) that holds level li's iterate
    u32 curm = 0u;
    auto cur = [&](int li) { return (curm >> li) & 1u; };
    // x_out = x_in + dinv (rhs - A x_in) on level li, from its current buffer into the other one
    auto sweep = [&](int li) {
      const TailSysLevel<DT_, VT_>& l = d.lv[li];
      const u32 n = l.nodes * u32(B_);
      DT_* v = arena + l.off;
      const DT_* xi = v + cur(li) * n;
      DT_* xo = v + (cur(li) ^ 1u) * n;
      for(u32 p = tid; p < l.nodes; p += kTailSysThreads)
      {
        DT_ s[B_];
        tail_sys_block_row<DT_, VT_, B_, DIAG_>(l.A, p, xi, s);
#pragma unroll
        for(int h = 0; h < B_; ++h)
        {
          const u32 i = p * u32(B_) + u32(h);
          xo[i] = xi[i] + v[4 * n + i] * (v[2 * n + i] - s[h]);
        }
      }
      curm ^= 1u << li;
      __syncthreads();
    };
    for(int li = 0; li + 1 < d.nlev; ++li)
    {
      for(int j = 1; j < d.pre; ++j) sweep(li);
      const TailSysLevel<DT_, VT_>& f = d.lv[li];
      const TailSysLevel<DT_, VT_>& c = d.lv[li + 1];
      const u32 n = f.nodes * u32(B_), nc = c.nodes * u32(B_);
      DT_* fv = arena + f.off;
      DT_* cv = arena + c.off;
      // def = filter(rhs - A x)
      for(u32 p = tid; p < f.nodes; p += kTailSysThreads)
      {
        DT_ s[B_];
        tail_sys_block_row<DT_, VT_, B_, DIAG_>(f.A, p, fv + cur(li) * n, s);
        const bool m = masked(f.mask, p);
#pragma unroll
        for(int h = 0; h < B_; ++h)
        {
          const u32 i = p * u32(B_) + u32(h);
          fv[3 * n + i] = m ? DT_(0) : fv[2 * n + i] - s[h];
        }
      }
      __syncthreads();
      // rhs_c = filter(R def); x_c = dinv_c rhs_c unless c is the coarsest level
      const bool first_step = li + 2 < d.nlev;
      for(u32 p = tid; p < c.nodes; p += kTailSysThreads)
      {
        DT_ s[B_];
        tail_sys_comp_row<DT_, VT_, B_>(f.R, p, fv + 3 * n, s);
        const bool m = masked(c.mask, p);
#pragma unroll
        for(int h = 0; h < B_; ++h)
        {
          const u32 i = p * u32(B_) + u32(h);
          const DT_ r = m ? DT_(0) : s[h];
          cv[2 * nc + i] = r;
          if(first_step) cv[i] = cv[4 * nc + i] * r;
        }
      }
      __syncthreads();
    }
    {
      // exact coarse solve x0 = Ainv rhs
      const TailSysLevel<DT_, VT_>& c = d.lv[d.nlev - 1];
      DT_* cv = arena + c.off;
      const u32 m = d.ncoarse;
      for(u32 i = tid; i < m; i += kTailSysThreads)
      {
        const double* a = d.ainv + u64(i) * m;
        double s = 0.0;
#pragma unroll 8
        for(u32 j = 0; j < m; ++j) s += a[j] * double(cv[2 * m + j]);
        cv[i] = masked(c.mask, i / u32(B_)) ? DT_(0) : DT_(s);
      }
      __syncthreads();
    }
    for(int li = d.nlev - 2; li >= 0; --li)
    {
      // x += filter(P x_c)
      const TailSysLevel<DT_, VT_>& f = d.lv[li];
      const TailSysLevel<DT_, VT_>& c = d.lv[li + 1];
      const u32 n = f.nodes * u32(B_), nc = c.nodes * u32(B_);
      DT_* xf = arena + f.off + cur(li) * n;
      const DT_* xc = arena + c.off + cur(li + 1) * nc;
      for(u32 p = tid; p < f.nodes; p += kTailSysThreads)
      {
        DT_ s[B_];
        tail_sys_comp_row<DT_, VT_, B_>(f.P, p, xc, s);
        if(masked(f.mask, p)) continue;
#pragma unroll
        for(int h = 0; h < B_; ++h)
        {
          const u32 i = p * u32(B_) + u32(h);
          xf[i] = xf[i] + s[h];
        }
      }
      __syncthreads();
      for(int j = 0; j < d.post; ++j) sweep(li);
    }
    const TailSysLevel<DT_, VT_>& t = d.lv[0];
    const u32 n = t.nodes * u32(B_);
    const DT_* xf = arena + t.off + cur(0) * n;
    for(u32 r = tid; r < n; r += kTailSysThreads) d.x_out[u64(r) * u64(K) + k] = xf[r];
  }

  // ---- PCG -----------------------------------------------------------------------------------
  struct SysState
  {
    double gamma, alpha, beta, def_init, def_cur, def_prev;
    u64 num_iter, num_stag_iter;
    int status, pad;
  };

  // Publication of the K system states to the host (one block of 64 threads): publication c of a
  // solve (counted on the device in *count) goes to slot c % R of a ring in pinned host memory
  // (64 states per slot), each thread's copy fenced at system scope, then, after the barrier, the
  // sequence number c + 1 to seq[c % R], which the host polls.  Runs whatever the statuses.
  template<int R>
  __global__ __launch_bounds__(64) void k_publish_states(const SysState* __restrict__ st, int K, SysState* ring,
                                                         unsigned* seq, unsigned* count)
  {
    const unsigned c = *count;   // read by every thread before the barrier below
    if(int(threadIdx.x) < K) ring[u64(c % R) * 64 + threadIdx.x] = st[threadIdx.x];
    __threadfence_system();
    __syncthreads();
    if(threadIdx.x == 0)
    {
      reinterpret_cast<volatile unsigned*>(seq)[c % R] = c + 1;
      *count = c + 1;
    }
  }

  struct Criteria
  {
    double tol_rel, tol_abs, tol_abs_low, div_rel, div_abs, stag_rate;
    u64 min_iter, max_iter, min_stag_iter, hist_cap;
  };

  // per-system partial sums of a block, partials[block][k]: lane groups of G lanes with S
  // systems per lane (S = 1 for the streaming kernels); K <= G*S, at most 256*S values
  template<int G_, int S_>
  __device__ __forceinline__ void block_partials(const double (&v)[S_], double* partials, int K)
  {
    __shared__ double red[256 * S_];
    const int k0 = int(threadIdx.x % G_) * S_;
    const unsigned g = threadIdx.x / G_;
    if(k0 < K)
#pragma unroll
      for(int i = 0; i < S_; ++i) red[g * unsigned(K) + unsigned(k0 + i)] = v[i];
    __syncthreads();
    if(int(threadIdx.x) < K)
    {
      double s = 0.0;
      for(unsigned gg = 0; gg < blockDim.x / G_; ++gg) s += red[gg * unsigned(K) + threadIdx.x];
      partials[u64(blockIdx.x) * u64(K) + threadIdx.x] = s;
    }
  }

  __device__ __forceinline__ bool sys_active(const SysState* st, int k) { return st[k].status == mg::kProgress; }

  // r = filter(b - A x); partial sums of r.r per system
  template<int G_, int S_, int B_, typename Rows_>
  F3_ROW_KERNEL void k_pcg_defect(Rows_ A, u64 per, const u32* __restrict__ mask, const double* __restrict__ b,
                                                      const double* __restrict__ x, double* __restrict__ r, double* partials)
  {
    const int K = A.K;
    double v[S_];
#pragma unroll
    for(int i = 0; i < S_; ++i) v[i] = 0.0;
    row_loop<G_, S_, B_>(A, per, x, [&](u64 p, int k0, const double (&s)[B_][S_]) {
      const bool m = masked(mask, p);
#pragma unroll
      for(int h = 0; h < B_; ++h)
      {
        const u64 i0 = (p * B_ + u64(h)) * u64(K) + u64(k0);
        double bi[S_], d[S_];
        ldv<double, S_>(b + i0, bi);
#pragma unroll
        for(int i = 0; i < S_; ++i)
        {
          d[i] = m ? 0.0 : bi[i] - s[h][i];
          v[i] += d[i] * d[i];
        }
        stv<double, S_>(r + i0, d);
      }
    });
    block_partials<G_, S_>(v, partials, K);
  }

  // q = filter(A p); partial sums of q.p for the systems in progress
  template<int G_, int S_, int B_, typename Rows_>
  F3_ROW_KERNEL void k_pcg_spmv_dot(Rows_ A, u64 per, const u32* __restrict__ mask, const double* __restrict__ p,
                                                        double* __restrict__ q, const SysState* st, const int* any, double* partials)
  {
    if(idle(any)) return;
    const int K = A.K;
    double v[S_];
#pragma unroll
    for(int i = 0; i < S_; ++i) v[i] = 0.0;
    row_loop<G_, S_, B_>(A, per, p, [&](u64 row, int k0, const double (&s)[B_][S_]) {
      const bool m = masked(mask, row);
#pragma unroll
      for(int h = 0; h < B_; ++h)
      {
        const u64 i0 = (row * B_ + u64(h)) * u64(K) + u64(k0);
        double pi[S_], qi[S_];
        ldv<double, S_>(p + i0, pi);
#pragma unroll
        for(int i = 0; i < S_; ++i)
        {
          qi[i] = m ? 0.0 : s[h][i];
          if(sys_active(st, k0 + i)) v[i] += qi[i] * pi[i];
        }
        stv<double, S_>(q + i0, qi);
      }
    });
    block_partials<G_, S_>(v, partials, K);
  }

  // p = z; partial sums of r.p
  template<int G_, int B_, typename DT_>
  __global__ __launch_bounds__(256) void k_pcg_p_gamma(u64 nodes, int K, const DT_* __restrict__ z, const double* __restrict__ r,
                                                       double* __restrict__ p, const SysState* st, double* partials)
  {
    double v[1] = {0.0};
    const int k = int(threadIdx.x % G_);
    const u64 ngrp = blockDim.x / G_;
    const u64 m = nodes * B_;
    if(k < K && sys_active(st, k))
      for(u64 i = u64(blockIdx.x) * ngrp + threadIdx.x / G_; i < m; i += u64(gridDim.x) * ngrp)
      {
        const u64 t = i * u64(K) + u64(k);
        const double zi = double(z[t]);
        p[t] = zi;
        v[0] += r[t] * zi;
      }
    block_partials<G_, 1>(v, partials, K);
  }

  // x += alpha p, r -= alpha q (active systems); partial sums of r.r; and the next V-cycle's
  // entry for every system, as k_entry computes it from the new r: rhs = (inner) r, x0 = dinv rhs
  // (the V-cycle that follows then starts without k_entry, which would read r again)
  template<int G_, int B_, bool SHARED_, typename DT_>
  __global__ __launch_bounds__(256) void k_pcg_update(u64 nodes, int K, const double* __restrict__ p, const double* __restrict__ q,
                                                      double* __restrict__ x, double* __restrict__ r, const SysState* st,
                                                      const int* any, double* partials, DT_* __restrict__ rhs,
                                                      DT_* __restrict__ x0, const DT_* __restrict__ dinv)
  {
    if(idle(any)) return;
    double v[1] = {0.0};
    const int k = int(threadIdx.x % G_);
    const u64 ngrp = blockDim.x / G_;
    const u64 m = nodes * B_;
    if(k < K)
    {
      const bool act = sys_active(st, k);
      const double alpha = st[k].alpha;
      const bool upd = act && !(fabs(alpha) < 2.220446049250313e-16);   // a NaN alpha updates (dense_vector.hpp:891)
      for(u64 i = u64(blockIdx.x) * ngrp + threadIdx.x / G_; i < m; i += u64(gridDim.x) * ngrp)
      {
        const u64 t = i * u64(K) + u64(k);
        double ri = r[t];
        if(upd)
        {
          x[t] = x[t] + alpha * p[t];
          ri = ri + (-alpha) * q[t];
          r[t] = ri;
        }
        if(act) v[0] += ri * ri;
        const DT_ rv = DT_(ri);
        rhs[t] = rv;
        x0[t] = dinv[SHARED_ ? i : t] * rv;
      }
    }
    block_partials<G_, 1>(v, partials, K);
  }

  // partial sums of r.z
  template<int G_, int B_, typename DT_>
  __global__ __launch_bounds__(256) void k_pcg_rz(u64 nodes, int K, const DT_* __restrict__ z, const double* __restrict__ r,
                                                  const SysState* st, const int* any, double* partials)
  {
    if(idle(any)) return;
    double v[1] = {0.0};
    const int k = int(threadIdx.x % G_);
    const u64 ngrp = blockDim.x / G_;
    const u64 m = nodes * B_;
    if(k < K && sys_active(st, k))
      for(u64 i = u64(blockIdx.x) * ngrp + threadIdx.x / G_; i < m; i += u64(gridDim.x) * ngrp)
      {
        const u64 t = i * u64(K) + u64(k);
        v[0] += r[t] * double(z[t]);
      }
    block_partials<G_, 1>(v, partials, K);
  }

  // p = beta p + z (active systems)
  template<int G_, int B_, typename DT_>
  __global__ void k_pcg_update_p(u64 nodes, int K, const DT_* __restrict__ z, double* __restrict__ p, const SysState* st,
                                 const int* any)
  {
    if(idle(any)) return;
    const int k = int(threadIdx.x % G_);
    if(k >= K || !sys_active(st, k)) return;
    const double beta = st[k].beta;
    const bool copy = fabs(beta) < 2.220446049250313e-16;
    const u64 ngrp = blockDim.x / G_;
    const u64 m = nodes * B_;
    for(u64 i = u64(blockIdx.x) * ngrp + threadIdx.x / G_; i < m; i += u64(gridDim.x) * ngrp)
    {
      const u64 t = i * u64(K) + u64(k);
      p[t] = copy ? double(z[t]) : p[t] * beta + double(z[t]);
    }
  }

  // finalisation of one per-system reduction, first stage: block k sums system k's nb block
  // partials (stride K) in a fixed order, eight loads in flight per thread, into sums[k] (one
  // block for all K systems would read nb * K partials alone: 58-68 us at 8192 x 32)
  __global__ __launch_bounds__(256) void k_fin_sum(const double* __restrict__ partials, unsigned nb, int K,
                                                   double* __restrict__ sums)
  {
    const u64 k = blockIdx.x;
    double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const unsigned step = blockDim.x;
    unsigned b = threadIdx.x;
    for(; b + 7 * step < nb; b += 8 * step)
#pragma unroll
      for(unsigned j = 0; j < 8; ++j) acc[j] += partials[u64(b + j * step) * u64(K) + k];
    for(; b < nb; b += step) acc[0] += partials[u64(b) * u64(K) + k];
    const double t = dev::block_sum(((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7])));
    if(threadIdx.x == 0) sums[k] = t;
  }

  // second stage (one block): the sums into every system's PCG scalars and criteria; also
  // refreshes the "any system in progress" flag after a status change
  template<int OP>
  __global__ __launch_bounds__(256) void k_fin(const double* __restrict__ sums, int K, SysState* st,
                                               const Criteria* cr, double* hist, int* any)
  {
    if(int(threadIdx.x) < K)
    {
      const double t = sums[threadIdx.x];
      SysState& S = st[threadIdx.x];
      double* h = hist + u64(threadIdx.x) * cr->hist_cap;
      if(OP == mg::kFinInitDefect)
      {
        const double d = sqrt(t);
        S.def_init = S.def_cur = S.def_prev = d;
        S.num_iter = 0;
        S.num_stag_iter = 0;
        if(cr->hist_cap > 0) h[0] = d;
        const double eps = 2.220446049250313e-16;
        S.status = !isfinite(d) ? mg::kAborted : (d <= eps * eps ? mg::kSuccess : mg::kProgress);
      }
      else if(S.status == mg::kProgress)
      {
        if(OP == mg::kFinAlpha) S.alpha = S.gamma / t;
        else if(OP == mg::kFinGamma) S.gamma = t;
        else if(OP == mg::kFinBeta) { const double g2 = S.gamma; S.gamma = t; S.beta = t / g2; }
        else if(OP == mg::kFinNewDefect)
        {
          ++S.num_iter;
          S.def_prev = S.def_cur;
          if((cr->min_iter < cr->max_iter) || (cr->min_stag_iter > 0)) S.def_cur = sqrt(t);
          const double d = S.def_cur;
          if(S.num_iter < cr->hist_cap) h[S.num_iter] = d;
          int sv = mg::kProgress;
          if(!isfinite(d)) sv = mg::kAborted;
          else if((d > cr->div_abs) || (d > cr->div_rel * S.def_init)) sv = mg::kDiverged;
          else if(S.num_iter < cr->min_iter) sv = mg::kProgress;
          else if((d <= cr->tol_abs) && ((d <= cr->tol_rel * S.def_init) || (d <= cr->tol_abs_low))) sv = mg::kSuccess;
          else if(S.num_iter >= cr->max_iter) sv = mg::kMaxIter;
          else if(cr->min_stag_iter > 0)
          {
            if(d >= cr->stag_rate * S.def_prev) { if(++S.num_stag_iter >= cr->min_stag_iter) sv = mg::kStagnated; }
            else S.num_stag_iter = 0;
          }
          S.status = sv;
        }
      }
    }
    __syncthreads();
    if(threadIdx.x == 0 && (OP == mg::kFinInitDefect || OP == mg::kFinNewDefect))
    {
      int a = 0;
      for(int j = 0; j < K; ++j) a |= (st[j].status == mg::kProgress) ? 1 : 0;
      *any = a;
    }
  }
} // namespace batch
} // namespace feat3_amd
