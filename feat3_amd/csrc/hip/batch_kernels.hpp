// batch_kernels.hpp -- kernels of the batched device MG-PCG: K systems in lockstep that share
// one sparsity graph, with B x B blocks per graph entry (B = 1 scalar, B = 3 the vector
// Laplacian of SparseMatrixBCSR<3,3>, sparse_matrix_bcsr.hpp:1295-1600), and values either
// shared by all K systems (K right-hand sides) or one set per system (K coefficient fields).
//
// Vectors are stored [node][component][system] (system fastest): lane k of a G-lane group
// handles system k of one block row, so a block's values are read once for all systems and the
// gathers x[col][w][0..K) are contiguous.  Every system keeps its own PCG scalars and status;
// a converged system's vectors are frozen, so each system performs exactly the iterations of its
// independent solve (pcg.hpp:188-288, iterative.hpp:744-879).
#pragma once

#include "device_common.hpp"
#include "mg_kernels.hpp"

namespace feat3_amd
{
namespace batch
{
  typedef std::uint32_t u32;
  typedef std::uint64_t u64;
  using mg::RowGrid;

  // values of graph entry e, block position (h, w), system k
  template<typename DT_, typename VT_, int B_, bool SHARED_>
  __device__ __forceinline__ DT_ bval(const VT_* val, u64 e, int h, int w, int k, int K)
  {
    const u64 q = (e * B_ + u64(h)) * B_ + u64(w);
    return SHARED_ ? dev::ld<DT_>(val, q) : dev::ld<DT_>(val, q * u64(K) + u64(k));
  }

  // block row sums s[h] = sum_e sum_w A_e[h][w] x[col_e][w][k]
  template<typename DT_, typename VT_, int B_, bool SHARED_>
  struct BlockRows
  {
    const u32* rp;
    const u32* col;
    const VT_* val;
    u64 rows;
    int K;
    __device__ __forceinline__ void row_sum(u64 p, int k, const DT_* __restrict__ x, DT_ (&s)[B_]) const
    {
#pragma unroll
      for(int h = 0; h < B_; ++h) s[h] = DT_(0);
      const u32 end = rp[p + 1];
      for(u32 e = rp[p]; e < end; ++e)
      {
        const u64 j = col[e];
        DT_ xv[B_];
#pragma unroll
        for(int w = 0; w < B_; ++w) xv[w] = x[(j * B_ + u64(w)) * u64(K) + u64(k)];
#pragma unroll
        for(int h = 0; h < B_; ++h)
#pragma unroll
          for(int w = 0; w < B_; ++w)
            s[h] += bval<DT_, VT_, B_, SHARED_>(val, e, h, w, k, K) * xv[w];
      }
    }
  };

  // scalar graph applied per component (the transfers: csrsb, apply_generic.hpp:156-181)
  template<typename DT_, typename VT_, int B_>
  struct CompRows
  {
    const u32* rp;
    const u32* col;
    const VT_* val;
    u64 rows;
    int K;
    __device__ __forceinline__ void row_sum(u64 p, int k, const DT_* __restrict__ x, DT_ (&s)[B_]) const
    {
#pragma unroll
      for(int h = 0; h < B_; ++h) s[h] = DT_(0);
      const u32 end = rp[p + 1];
      for(u32 e = rp[p]; e < end; ++e)
      {
        const u64 j = col[e];
        const DT_ v = dev::ld<DT_>(val, e);
#pragma unroll
        for(int h = 0; h < B_; ++h) s[h] += v * x[(j * B_ + u64(h)) * u64(K) + u64(k)];
      }
    }
  };

  // calls f(row, k, s[B]) for the rows of this block; lanes k >= K of a group are idle
  template<int G_, int B_, typename Rows_, typename DT_, typename F>
  __device__ __forceinline__ void for_block_rows(const Rows_& A, u64 per, const DT_* __restrict__ x, F&& f)
  {
    const u64 b = dev::xcd_remap(blockIdx.x, gridDim.x);
    const u64 beg = b * per;
    const u64 end = beg + per < A.rows ? beg + per : A.rows;
    const int k = int(threadIdx.x % G_);
    if(k >= A.K) return;
    for(u64 p = beg + threadIdx.x / G_; p < end; p += blockDim.x / G_)
    {
      DT_ s[B_];
      A.row_sum(p, k, x, s);
      f(p, k, s);
    }
  }

  __device__ __forceinline__ bool masked(const u32* mask, u64 node) { return dev::masked(mask, node); }
  __device__ __forceinline__ bool idle(const int* any) { return any != nullptr && *any == 0; }

  // ---- V-cycle -------------------------------------------------------------------------------
  // dinv layout: [node][h] (shared values) or [node][h][k] (one matrix per system)
  template<int B_, bool SHARED_>
  __device__ __forceinline__ u64 dinv_index(u64 node, int h, int k, int K)
  {
    return SHARED_ ? node * B_ + u64(h) : (node * B_ + u64(h)) * u64(K) + u64(k);
  }

  template<typename DT_, int B_, bool SHARED_>
  __global__ void k_entry(u64 nodes, int K, const double* __restrict__ def, DT_* __restrict__ rhs, DT_* __restrict__ x,
                          const DT_* __restrict__ dinv, const int* any)
  {
    if(idle(any)) return;
    const u64 n = nodes * B_ * u64(K);
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
    {
      const int k = int(i % u64(K));
      const u64 nh = i / u64(K);
      const DT_ r = DT_(def[i]);
      rhs[i] = r;
      x[i] = dinv[dinv_index<B_, SHARED_>(nh / B_, int(nh % B_), k, K)] * r;
    }
  }

  template<int G_, int B_, bool SHARED_, typename Rows_, typename DT_>
  __global__ __launch_bounds__(256) void k_sweep(Rows_ A, u64 per, const DT_* __restrict__ dinv, const DT_* __restrict__ rhs,
                                                 const DT_* __restrict__ xin, DT_* __restrict__ xout, const int* any)
  {
    if(idle(any)) return;
    const int K = A.K;
    for_block_rows<G_, B_>(A, per, xin, [&](u64 p, int k, const DT_ (&s)[B_]) {
#pragma unroll
      for(int h = 0; h < B_; ++h)
      {
        const u64 i = (p * B_ + u64(h)) * u64(K) + u64(k);
        xout[i] = xin[i] + dinv[dinv_index<B_, SHARED_>(p, h, k, K)] * (rhs[i] - s[h]);
      }
    });
  }

  template<int G_, int B_, typename Rows_, typename DT_>
  __global__ __launch_bounds__(256) void k_defect(Rows_ A, u64 per, const u32* __restrict__ mask, const DT_* __restrict__ rhs,
                                                  const DT_* __restrict__ x, DT_* __restrict__ def, const int* any)
  {
    if(idle(any)) return;
    const int K = A.K;
    for_block_rows<G_, B_>(A, per, x, [&](u64 p, int k, const DT_ (&s)[B_]) {
      const bool m = masked(mask, p);
#pragma unroll
      for(int h = 0; h < B_; ++h)
      {
        const u64 i = (p * B_ + u64(h)) * u64(K) + u64(k);
        def[i] = m ? DT_(0) : rhs[i] - s[h];
      }
    });
  }

  template<int G_, int B_, bool SHARED_, typename Rows_, typename DT_>
  __global__ __launch_bounds__(256) void k_restrict(Rows_ R, u64 per, const u32* __restrict__ mask_c, const DT_* __restrict__ def_f,
                                                    DT_* __restrict__ rhs_c, DT_* __restrict__ x_c, const DT_* __restrict__ dinv_c,
                                                    const int* any)
  {
    if(idle(any)) return;
    const int K = R.K;
    for_block_rows<G_, B_>(R, per, def_f, [&](u64 p, int k, const DT_ (&s)[B_]) {
      const bool m = masked(mask_c, p);
#pragma unroll
      for(int h = 0; h < B_; ++h)
      {
        const u64 i = (p * B_ + u64(h)) * u64(K) + u64(k);
        const DT_ r = m ? DT_(0) : s[h];
        rhs_c[i] = r;
        if(x_c) x_c[i] = dinv_c[dinv_index<B_, SHARED_>(p, h, k, K)] * r;
      }
    });
  }

  template<int G_, int B_, typename Rows_, typename DT_>
  __global__ __launch_bounds__(256) void k_prolong(Rows_ P, u64 per, const u32* __restrict__ mask_f, const DT_* __restrict__ x_c,
                                                   DT_* __restrict__ x_f, const int* any)
  {
    if(idle(any)) return;
    const int K = P.K;
    for_block_rows<G_, B_>(P, per, x_c, [&](u64 p, int k, const DT_ (&s)[B_]) {
      if(masked(mask_f, p)) return;
#pragma unroll
      for(int h = 0; h < B_; ++h)
      {
        const u64 i = (p * B_ + u64(h)) * u64(K) + u64(k);
        x_f[i] = x_f[i] + s[h];
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
      for(u64 j = 0; j < m; ++j) s += a[j] * double(rhs[j * u64(K) + u64(k)]);
      x[i * u64(K) + u64(k)] = masked(mask, i / u64(B)) ? DT_(0) : DT_(s);
    }
  }

  // ---- PCG -----------------------------------------------------------------------------------
  struct SysState
  {
    double gamma, alpha, beta, def_init, def_cur, def_prev;
    u64 num_iter, num_stag_iter;
    int status, pad;
  };

  struct Criteria
  {
    double tol_rel, tol_abs, tol_abs_low, div_rel, div_abs, stag_rate;
    u64 min_iter, max_iter, min_stag_iter, hist_cap;
  };

  // per-system partial sums of a block: partials[block][k]
  template<int G_>
  __device__ __forceinline__ void block_partials_k(double v, double* partials, int K)
  {
    __shared__ double red[256];
    red[threadIdx.x] = v;
    __syncthreads();
    if(int(threadIdx.x) < K)
    {
      double s = 0.0;
      for(unsigned g = 0; g < blockDim.x / G_; ++g) s += red[g * G_ + threadIdx.x];
      partials[u64(blockIdx.x) * u64(K) + threadIdx.x] = s;
    }
  }

  __device__ __forceinline__ bool sys_active(const SysState* st, int k) { return st[k].status == mg::kProgress; }

  // r = filter(b - A x); partial sums of r.r per system
  template<int G_, int B_, typename Rows_>
  __global__ __launch_bounds__(256) void k_pcg_defect(Rows_ A, u64 per, const u32* __restrict__ mask, const double* __restrict__ b,
                                                      const double* __restrict__ x, double* __restrict__ r, double* partials)
  {
    const int K = A.K;
    double v = 0.0;
    for_block_rows<G_, B_>(A, per, x, [&](u64 p, int k, const double (&s)[B_]) {
      const bool m = masked(mask, p);
#pragma unroll
      for(int h = 0; h < B_; ++h)
      {
        const u64 i = (p * B_ + u64(h)) * u64(K) + u64(k);
        const double d = m ? 0.0 : b[i] - s[h];
        r[i] = d;
        v += d * d;
      }
    });
    block_partials_k<G_>(v, partials, K);
  }

  // p = z; partial sums of r.p
  template<int G_, int B_, typename DT_>
  __global__ __launch_bounds__(256) void k_pcg_p_gamma(u64 nodes, int K, const DT_* __restrict__ z, const double* __restrict__ r,
                                                       double* __restrict__ p, const SysState* st, double* partials)
  {
    double v = 0.0;
    const int k = int(threadIdx.x % G_);
    const u64 ngrp = blockDim.x / G_;
    const u64 m = nodes * B_;
    if(k < K && sys_active(st, k))
      for(u64 i = u64(blockIdx.x) * ngrp + threadIdx.x / G_; i < m; i += u64(gridDim.x) * ngrp)
      {
        const u64 q = i * u64(K) + u64(k);
        const double zi = double(z[q]);
        p[q] = zi;
        v += r[q] * zi;
      }
    block_partials_k<G_>(v, partials, K);
  }

  // q = filter(A p); partial sums of q.p
  template<int G_, int B_, typename Rows_>
  __global__ __launch_bounds__(256) void k_pcg_spmv_dot(Rows_ A, u64 per, const u32* __restrict__ mask, const double* __restrict__ p,
                                                        double* __restrict__ q, const SysState* st, const int* any, double* partials)
  {
    if(idle(any)) return;
    const int K = A.K;
    double v = 0.0;
    for_block_rows<G_, B_>(A, per, p, [&](u64 row, int k, const double (&s)[B_]) {
      if(!sys_active(st, k)) return;
      const bool m = masked(mask, row);
#pragma unroll
      for(int h = 0; h < B_; ++h)
      {
        const u64 i = (row * B_ + u64(h)) * u64(K) + u64(k);
        const double qi = m ? 0.0 : s[h];
        q[i] = qi;
        v += qi * p[i];
      }
    });
    block_partials_k<G_>(v, partials, K);
  }

  // x += alpha p, r -= alpha q (active systems); partial sums of r.r
  template<int G_, int B_>
  __global__ __launch_bounds__(256) void k_pcg_update(u64 nodes, int K, const double* __restrict__ p, const double* __restrict__ q,
                                                      double* __restrict__ x, double* __restrict__ r, const SysState* st,
                                                      const int* any, double* partials)
  {
    if(idle(any)) return;
    double v = 0.0;
    const int k = int(threadIdx.x % G_);
    const u64 ngrp = blockDim.x / G_;
    const u64 m = nodes * B_;
    if(k < K && sys_active(st, k))
    {
      const double alpha = st[k].alpha;
      const bool upd = fabs(alpha) >= 2.220446049250313e-16;
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
        v += ri * ri;
      }
    }
    block_partials_k<G_>(v, partials, K);
  }

  // partial sums of r.z
  template<int G_, int B_, typename DT_>
  __global__ __launch_bounds__(256) void k_pcg_rz(u64 nodes, int K, const DT_* __restrict__ z, const double* __restrict__ r,
                                                  const SysState* st, const int* any, double* partials)
  {
    if(idle(any)) return;
    double v = 0.0;
    const int k = int(threadIdx.x % G_);
    const u64 ngrp = blockDim.x / G_;
    const u64 m = nodes * B_;
    if(k < K && sys_active(st, k))
      for(u64 i = u64(blockIdx.x) * ngrp + threadIdx.x / G_; i < m; i += u64(gridDim.x) * ngrp)
      {
        const u64 t = i * u64(K) + u64(k);
        v += r[t] * double(z[t]);
      }
    block_partials_k<G_>(v, partials, K);
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

  // finalisation of one per-system reduction (one block of 256 threads); also refreshes the
  // "any system in progress" flag after a status change
  template<int OP>
  __global__ __launch_bounds__(256) void k_fin(const double* __restrict__ partials, unsigned nb, int K, SysState* st,
                                               const Criteria* cr, double* hist, int* any)
  {
    __shared__ double red[256];
    const int k = int(threadIdx.x % 64);
    const unsigned slice = threadIdx.x / 64;
    double s = 0.0;
    if(k < K)
      for(unsigned b = slice; b < nb; b += blockDim.x / 64) s += partials[u64(b) * u64(K) + u64(k)];
    red[threadIdx.x] = s;
    __syncthreads();
    if(int(threadIdx.x) < K)
    {
      double t = 0.0;
      for(unsigned sl = 0; sl < blockDim.x / 64; ++sl) t += red[sl * 64 + threadIdx.x];
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
