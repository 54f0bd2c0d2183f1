// feat_oracle.cpp -- CPU restatement of FEAT3's LAFEM generic arch kernels and of the
// PCG / Richardson / Jacobi / MultiGrid / ConvertPrecond solver semantics.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP product path is checked
// against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
// load the library built from it; the product never links, calls or falls back to it.
//
// Every function cites the reference file:line it restates (paths relative to the FEAT3
// source tree).  The restatement keeps FEAT's loop order and alias/epsilon special cases so
// that a single-threaded build reproduces the reference's CPU arithmetic.  The optional
// OpenMP build (-DORC_OMP) parallelises the row loops of the SpMV and the BLAS-1 loops for
// the multi-core CPU baseline; its reductions are then no longer in sequential order.
//
// Pinning (what fixes each function independently of this text):
//   banded apply              -- dense brute force (square/rectangular, diagonals leaving the
//                                matrix), bit-identity with the CSR apply on the Q1 matrix
//   csr/ell/csrb/csrsb apply  -- dense brute force on tiny matrices, closed-form eigenpairs of
//                                the pointstar FD matrix (pointstar_factory.hpp:388-411)
//   pointstar_fd              -- nnz formula, symmetric 5/7-point pattern, eigenpair KAT
//   PCG/Jacobi/Richardson     -- iteration counts of kernel/solver/basic_solver-test.cpp:136-263
//                                (CG 28, PCG-JAC 28, Richardson-JAC 1175, each +-2)
//   MultiGrid V-cycle + PCG   -- applications/poisson_dirichlet.cpp:266-276 (4 +- 1 iterations
//                                at level 5) and exact two-grid/coarse-solve identities
//   MultiGrid F-/W-cycle      -- recursive definitions evaluated independently in numpy
//                                (F = V on two levels, W on two levels = two V-cycles in
//                                defect correction, recursion on 3-5 levels)
//   transpose                 -- R = P^T identity on random patterns
//   half conversion           -- exhaustive round trip of all 65536 binary16 patterns
//   cell_gradient (adjoint)   -- no reference counterpart; pinned by central finite differences
//                                of the loss through two oracle solves and by the identity
//                                sum_e k_e dL/dk_e = -lambda^T b (A is linear in k)
// See tests/test_oracle_golden.py and tests/test_oracle_adjoint.py.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

namespace orc
{
  typedef std::uint64_t Index;

  template<typename T> inline T eps() { return std::numeric_limits<T>::epsilon(); }

  // ---------------------------------------------------------------------------------------
  // IEEE binary16 storage type (round to nearest even), used for the FP16-stored matrix
  // values of the mixed-precision smoother.  gcc 11 has no _Float16 on x86, so the
  // conversion is spelled out on the bit patterns.
  // ---------------------------------------------------------------------------------------
  struct half_t { std::uint16_t bits; };

  inline float half_to_float(std::uint16_t h)
  {
    const std::uint32_t sign = std::uint32_t(h & 0x8000u) << 16;
    const std::uint32_t expo = (h >> 10) & 0x1fu;
    const std::uint32_t mant = h & 0x3ffu;
    std::uint32_t f;
    if(expo == 0u)
    {
      if(mant == 0u)
        f = sign;
      else
      {
        // subnormal half -> normal float
        int e = -1;
        std::uint32_t m = mant;
        do { m <<= 1; ++e; } while((m & 0x400u) == 0u);
        f = sign | std::uint32_t(127 - 15 - e) << 23 | ((m & 0x3ffu) << 13);
      }
    }
    else if(expo == 0x1fu)
      f = sign | 0x7f800000u | (mant << 13);
    else
      f = sign | ((expo + 112u) << 23) | (mant << 13);
    float out;
    std::memcpy(&out, &f, 4);
    return out;
  }

  inline std::uint16_t float_to_half(float x)
  {
    std::uint32_t f;
    std::memcpy(&f, &x, 4);
    const std::uint16_t sign = std::uint16_t((f >> 16) & 0x8000u);
    const std::uint32_t expo = (f >> 23) & 0xffu;
    std::uint32_t mant = f & 0x7fffffu;
    if(expo == 0xffu) // inf / nan
      return std::uint16_t(sign | 0x7c00u | (mant ? 0x200u | (mant >> 13) : 0u));
    int e = int(expo) - 127 + 15;
    if(e >= 0x1f) // overflow -> inf
      return std::uint16_t(sign | 0x7c00u);
    if(e <= 0)
    {
      // subnormal or zero in half
      if(e < -10)
        return sign;
      mant |= 0x800000u;
      const int shift = 14 - e;
      std::uint32_t hm = mant >> shift;
      const std::uint32_t rem = mant & ((1u << shift) - 1u);
      const std::uint32_t halfway = 1u << (shift - 1);
      if(rem > halfway || (rem == halfway && (hm & 1u)))
        ++hm;
      return std::uint16_t(sign | hm);
    }
    std::uint32_t hm = mant >> 13;
    const std::uint32_t rem = mant & 0x1fffu;
    std::uint32_t he = std::uint32_t(e);
    if(rem > 0x1000u || (rem == 0x1000u && (hm & 1u)))
    {
      ++hm;
      if(hm == 0x400u) { hm = 0; ++he; }
      if(he >= 0x1fu) return std::uint16_t(sign | 0x7c00u);
    }
    return std::uint16_t(sign | (he << 10) | hm);
  }

  // value-storage traits: DT_ is the arithmetic type, VT_ the stored matrix value type
  template<typename DT_, typename VT_> struct Load { static DT_ get(const VT_& v) { return DT_(v); } };
  template<typename DT_> struct Load<DT_, half_t> { static DT_ get(const half_t& v) { return DT_(half_to_float(v.bits)); } };

  // ---------------------------------------------------------------------------------------
  // Arch kernels (lafem/arch/*_generic.hpp)
  // ---------------------------------------------------------------------------------------

  // apply_generic.hpp:24-69  (r = a*A*x + b*y; transposed variant included)
  template<typename DT_, typename IT_, typename VT_ = DT_>
  void csr_apply(DT_* r, DT_ a, const DT_* x, DT_ b, const DT_* y, const VT_* val, const IT_* col,
                 const IT_* rp, Index rows, Index cols, bool transposed)
  {
    const Index nr = transposed ? cols : rows;
    if(std::abs(b) < eps<DT_>())
      std::fill(r, r + nr, DT_(0));
    else if(r != y)
      std::copy(y, y + nr, r);
    if(transposed)
    {
      const DT_ ba = b / a;
      for(Index c(0); c < cols; ++c)
        r[c] = ba * r[c];
      for(Index row(0); row < rows; ++row)
        for(Index i(rp[row]); i < Index(rp[row + 1]); ++i)
          r[col[i]] += Load<DT_, VT_>::get(val[i]) * x[row];
      for(Index c(0); c < cols; ++c)
        r[c] = a * r[c];
    }
    else
    {
#ifdef ORC_OMP
#pragma omp parallel for schedule(static)
#endif
      for(std::int64_t srow = 0; srow < std::int64_t(rows); ++srow)
      {
        const Index row = Index(srow);
        DT_ sum(0);
        const IT_ end(rp[row + 1]);
        for(IT_ i(rp[row]); i < end; ++i)
          sum += Load<DT_, VT_>::get(val[i]) * x[col[i]];
        r[row] = (sum * a) + (b * r[row]);
      }
    }
  }

  // apply_generic.hpp:121-154  (block rows of BH x BW row-major blocks)
  template<typename DT_, typename IT_>
  void csrb_apply(DT_* r, DT_ a, const DT_* x, DT_ b, const DT_* y, const DT_* val, const IT_* col,
                  const IT_* rp, Index rows, int bh, int bw)
  {
    if(std::abs(b) < eps<DT_>())
      std::fill(r, r + rows * Index(bh), DT_(0));
    else if(r != y)
      std::copy(y, y + rows * Index(bh), r);
#ifdef ORC_OMP
#pragma omp parallel for schedule(static)
#endif
    for(std::int64_t srow = 0; srow < std::int64_t(rows); ++srow)
    {
      const Index row = Index(srow);
      DT_ bsum[16];
      for(int h(0); h < bh; ++h)
        bsum[h] = DT_(0);
      for(IT_ i(rp[row]); i < rp[row + 1]; ++i)
        for(int h(0); h < bh; ++h)
          for(int w(0); w < bw; ++w)
            bsum[h] += val[Index(i) * Index(bh * bw) + Index(h * bw + w)] * x[Index(col[i]) * Index(bw) + Index(w)];
      for(int h(0); h < bh; ++h)
        r[row * Index(bh) + Index(h)] = (bsum[h] * a) + (b * r[row * Index(bh) + Index(h)]);
    }
  }

  // apply_generic.hpp:156-181  (scalar CSR times a vector of bs-blocks)
  template<typename DT_, typename IT_, typename VT_ = DT_>
  void csrsb_apply(DT_* r, DT_ a, const DT_* x, DT_ b, const DT_* y, const VT_* val, const IT_* col,
                   const IT_* rp, Index rows, int bs)
  {
    if(std::abs(b) < eps<DT_>())
      std::fill(r, r + rows * Index(bs), DT_(0));
    else if(r != y)
      std::copy(y, y + rows * Index(bs), r);
#ifdef ORC_OMP
#pragma omp parallel for schedule(static)
#endif
    for(std::int64_t srow = 0; srow < std::int64_t(rows); ++srow)
    {
      const Index row = Index(srow);
      DT_ bsum[64];
      for(int h(0); h < bs; ++h)
        bsum[h] = DT_(0);
      for(IT_ i(rp[row]); i < rp[row + 1]; ++i)
      {
        const DT_ v = Load<DT_, VT_>::get(val[i]);
        for(int h(0); h < bs; ++h)
          bsum[h] += v * x[Index(col[i]) * Index(bs) + Index(h)];
      }
      for(int h(0); h < bs; ++h)
        r[row * Index(bs) + Index(h)] = (bsum[h] * a) + (b * r[row * Index(bs) + Index(h)]);
    }
  }

  // sparse_matrix_ell.hpp:865-960 (CSR -> SELL-C with sigma = 1, zero padded with column 0)
  template<typename IT_>
  Index ell_val_size(Index C, Index rows, const IT_* rp)
  {
    const Index nch = (rows + C - 1) / C;
    Index total = 0;
    for(Index c(0); c < nch; ++c)
    {
      Index cl = 0;
      for(Index i(c * C); i < std::min(rows, (c + 1) * C); ++i)
        cl = std::max(cl, Index(rp[i + 1] - rp[i]));
      total += C * cl;
    }
    return total;
  }

  template<typename DT_, typename IT_>
  void ell_from_csr(Index C, Index rows, const IT_* rp, const IT_* col, const DT_* val,
                    IT_* cs, IT_* cl, IT_* rl, DT_* eval, IT_* ecol)
  {
    const Index nch = (rows + C - 1) / C;
    for(Index c(0); c < nch; ++c)
      cl[c] = 0;
    for(Index i(0); i < rows; ++i)
    {
      rl[i] = rp[i + 1] - rp[i];
      if(rl[i] > cl[i / C])
        cl[i / C] = rl[i];
    }
    cs[0] = 0;
    for(Index c(0); c < nch; ++c)
      cs[c + 1] = cs[c] + IT_(C) * cl[c];
    // zero everything first: covers row padding and the rows of an incomplete last chunk
    std::fill(eval, eval + Index(cs[nch]), DT_(0));
    std::fill(ecol, ecol + Index(cs[nch]), IT_(0));
    for(Index i(0); i < rows; ++i)
    {
      Index k = 0;
      for(IT_ j(rp[i]); j < rp[i + 1]; ++j, ++k)
      {
        ecol[Index(cs[i / C]) + i % C + k * C] = col[j];
        eval[Index(cs[i / C]) + i % C + k * C] = val[j];
      }
    }
  }

  // apply_generic.hpp:230-382 (the C in {2,4,8,16,32} specialisations: r = a*sum + b*y)
  template<typename DT_, typename IT_>
  void ell_apply(DT_* r, DT_ a, const DT_* x, DT_ b, const DT_* y, const DT_* val, const IT_* col,
                 const IT_* cs, const IT_* cl, Index C, Index rows)
  {
    if(std::abs(b) < eps<DT_>())
      std::fill(r, r + rows, DT_(0));
    else if(r != y)
      std::copy(y, y + rows, r);
    const Index nch = (rows + C - 1) / C;
    for(Index c(0); c < nch; ++c)
    {
      for(Index k(0); k < C && c * C + k < rows; ++k)
      {
        DT_ sum(0);
        for(Index j(0); j < Index(cl[c]); ++j)
          sum += val[Index(cs[c]) + j * C + k] * x[col[Index(cs[c]) + j * C + k]];
        r[c * C + k] = sum * a + b * r[c * C + k];
      }
    }
  }

  // apply_generic.hpp:519-598 (SparseMatrixBanded): offsets[d] = diagonal offset + rows - 1,
  // val[d * rows + l] on diagonal d.  The reference splits the rows into segments on which the
  // same contiguous range [i, j) of (ascending) diagonals has its column inside [0, columns)
  // (start_offset/end_offset, :415-448) and sums a = i .. j-1 in order; per row that is the sum
  // over the diagonals with a column in range, in storage order, restated here row by row.
  template<typename DT_, typename IT_>
  void banded_apply(DT_* r, DT_ a, const DT_* x, DT_ b, const DT_* y, const DT_* val, const IT_* offsets,
                    Index noo, Index rows, Index columns)
  {
    if(std::abs(b) < eps<DT_>())
      std::fill(r, r + rows, DT_(0));
    else if(r != y)
      std::copy(y, y + rows, r);
    for(Index l(0); l < rows; ++l)
    {
      DT_ s(0);
      for(Index d(0); d < noo; ++d)
      {
        const Index c = l + Index(offsets[d]) + 1 - rows;   // wraps for columns below 0
        if(c < columns)
          s += val[d * rows + l] * x[c];
      }
      r[l] = b * r[l] + a * s;
    }
  }

  // axpy_generic.hpp:24-50
  template<typename DT_>
  void axpy(DT_* r, DT_ a, const DT_* x, const DT_* y, Index n)
  {
    if(r == y)
    {
#ifdef ORC_OMP
#pragma omp parallel for schedule(static)
#endif
      for(std::int64_t i = 0; i < std::int64_t(n); ++i)
        r[i] += a * x[i];
    }
    else if(r == x)
    {
#ifdef ORC_OMP
#pragma omp parallel for schedule(static)
#endif
      for(std::int64_t i = 0; i < std::int64_t(n); ++i)
      {
        r[i] *= a;
        r[i] += y[i];
      }
    }
    else
    {
#ifdef ORC_OMP
#pragma omp parallel for schedule(static)
#endif
      for(std::int64_t i = 0; i < std::int64_t(n); ++i)
        r[i] = (a * x[i]) + y[i];
    }
  }

  // dot_product_generic.hpp:20-41 (sequential sum)
  template<typename DT_>
  DT_ dot(const DT_* x, const DT_* y, Index n)
  {
    DT_ r(0);
#ifdef ORC_OMP
#pragma omp parallel for schedule(static) reduction(+:r)
#endif
    for(std::int64_t i = 0; i < std::int64_t(n); ++i)
      r += x[i] * y[i];
    return r;
  }

  // norm_generic.hpp:22-32
  template<typename DT_>
  DT_ norm2(const DT_* x, Index n)
  {
    DT_ r(0);
#ifdef ORC_OMP
#pragma omp parallel for schedule(static) reduction(+:r)
#endif
    for(std::int64_t i = 0; i < std::int64_t(n); ++i)
      r += x[i] * x[i];
    return DT_(std::sqrt(r));
  }

  // component_product_generic.hpp:21-51
  template<typename DT_>
  void component_product(DT_* r, const DT_* x, const DT_* y, Index n)
  {
#ifdef ORC_OMP
#pragma omp parallel for schedule(static)
#endif
    for(std::int64_t i = 0; i < std::int64_t(n); ++i)
      r[i] = x[i] * y[i];
  }

  // component_invert_generic.hpp:21-36
  template<typename DT_>
  void component_invert(DT_* r, const DT_* x, DT_ s, Index n)
  {
    for(Index i(0); i < n; ++i)
      r[i] = s / x[i];
  }

  // scale_generic.hpp (r = s*x)
  template<typename DT_>
  void scale(DT_* r, const DT_* x, DT_ s, Index n)
  {
    for(Index i(0); i < n; ++i)
      r[i] = x[i] * s;
  }

  // diagonal_generic.hpp:20-37 (first entry with col == row, else 0)
  template<typename DT_, typename IT_, typename VT_ = DT_>
  void diag_csr(DT_* d, const VT_* val, const IT_* col, const IT_* rp, Index rows)
  {
    for(Index row(0); row < rows; ++row)
    {
      d[row] = DT_(0);
      for(Index i(rp[row]); i < Index(rp[row + 1]); ++i)
        if(Index(col[i]) == row)
        {
          d[row] = Load<DT_, VT_>::get(val[i]);
          break;
        }
    }
  }

  // diagonal_generic.hpp:39-58 (diagonal of the diagonal block)
  template<typename DT_, typename IT_>
  void diag_csrb(DT_* d, const DT_* val, const IT_* col, const IT_* rp, Index rows, int bh, int bw)
  {
    for(Index row(0); row < rows; ++row)
    {
      for(int i(0); i < bh; ++i)
        d[row * Index(bh) + Index(i)] = DT_(0);
      for(Index c(rp[row]); c < Index(rp[row + 1]); ++c)
        if(Index(col[c]) == row)
          for(int i(0); i < bh; ++i)
            d[row * Index(bh) + Index(i)] = val[c * Index(bh * bw) + Index(i + i * bw)];
    }
  }

  // unit_filter_generic.hpp:21-37
  template<typename DT_, typename IT_>
  void unit_filter_rhs(DT_* v, const DT_* sv, const IT_* idx, Index ue)
  {
    for(Index i(0); i < ue; ++i)
      v[idx[i]] = sv[i];
  }
  template<typename DT_, typename IT_>
  void unit_filter_def(DT_* v, const IT_* idx, Index ue)
  {
    for(Index i(0); i < ue; ++i)
      v[idx[i]] = DT_(0);
  }

  // sparse_matrix_csr.hpp:1781-1844 (counting-sort transpose; ascending columns per row)
  template<typename DT_, typename IT_>
  void transpose_csr(Index rows, Index cols, const IT_* rp, const IT_* col, const DT_* val,
                     IT_* trp, IT_* tcol, DT_* tval)
  {
    std::vector<Index> cnt(cols + 1, 0);
    const Index nnz = rp[rows];
    for(Index i(0); i < nnz; ++i)
      ++cnt[Index(col[i]) + 1];
    for(Index c(0); c < cols; ++c)
      cnt[c + 1] += cnt[c];
    for(Index c(0); c <= cols; ++c)
      trp[c] = IT_(cnt[c]);
    for(Index row(0); row < rows; ++row)
      for(Index k(rp[row]); k < Index(rp[row + 1]); ++k)
      {
        const Index j = cnt[col[k]]++;
        tval[j] = val[k];
        tcol[j] = IT_(row);
      }
  }

  // ---------------------------------------------------------------------------------------
  // PointstarFactoryFD (pointstar_factory.hpp:168-411): the (2d+1)-point finite-difference
  // Laplacian on an m^d lexicographic grid, diagonal 2d, off-diagonals -1, ascending columns.
  // The construction below is a direct one (per row, neighbours in ascending order); the
  // reference's recursive block construction yields the same arrays.
  // ---------------------------------------------------------------------------------------
  inline Index pointstar_fd_nnz(Index m, Index d)
  {
    Index neq = m, nnz = 3 * m - 2;
    for(Index i(1); i < d; ++i)
    {
      nnz *= m;
      nnz += 2 * (m - 1) * neq;
      neq *= m;
    }
    return nnz;
  }

  template<typename DT_, typename IT_>
  void pointstar_fd(Index m, Index d, IT_* rp, IT_* col, DT_* val)
  {
    Index neq = 1;
    for(Index i(0); i < d; ++i)
      neq *= m;
    Index k = 0;
    rp[0] = 0;
    for(Index row(0); row < neq; ++row)
    {
      // neighbours at -stride (descending stride), diagonal, +stride (ascending stride)
      std::vector<Index> strides;
      for(Index q(1); q < neq; q *= m)
        strides.push_back(q);
      for(Index s(strides.size()); s > 0; --s)
      {
        const Index q = strides[s - 1];
        if((row / q) % m > 0) { col[k] = IT_(row - q); val[k] = DT_(-1); ++k; }
      }
      col[k] = IT_(row); val[k] = DT_(2 * d); ++k;
      for(Index s(0); s < strides.size(); ++s)
      {
        const Index q = strides[s];
        if((row / q) % m < m - 1) { col[k] = IT_(row + q); val[k] = DT_(-1); ++k; }
      }
      rp[row + 1] = IT_(k);
    }
  }

  // pointstar_factory.hpp:125-150
  template<typename DT_>
  void pointstar_q2_bubble(Index m, Index d, DT_* v)
  {
    Index neq = 1;
    for(Index i(0); i < d; ++i)
      neq *= m;
    const DT_ xsf = DT_(1) / DT_(m + 1);
    for(Index i(0); i < neq; ++i)
    {
      v[i] = DT_(1);
      for(Index quo(1); quo < neq; quo *= m)
      {
        const DT_ x(xsf * DT_(((i / quo) % m) + 1));
        v[i] *= DT_(4) * x * (DT_(1) - x);
      }
    }
  }

  // ---------------------------------------------------------------------------------------
  // Matrix operators used by the solvers (the container-level apply() semantics of
  // sparse_matrix_csr.hpp:1899-2010 / sparse_matrix_bcsr.hpp:1295-1600 on top of the kernels)
  // ---------------------------------------------------------------------------------------
  template<typename DT_>
  struct Operator
  {
    virtual ~Operator() {}
    virtual Index size_l() const = 0; // scalar length of A*x
    virtual Index size_r() const = 0; // scalar length of x
    virtual Index nnz() const = 0;
    // r = A*x  (sparse_matrix_csr.hpp:1899-1928)
    virtual void kernel(DT_* r, DT_ a, const DT_* x, DT_ b, const DT_* y) const = 0;
    void apply(DT_* r, const DT_* x) const
    {
      if(nnz() == 0) { std::fill(r, r + size_l(), DT_(0)); return; }
      kernel(r, DT_(1), x, DT_(0), r);
    }
    // r = y + alpha*A*x  (sparse_matrix_csr.hpp:1973-2010)
    void apply(DT_* r, const DT_* x, const DT_* y, DT_ alpha) const
    {
      if(nnz() == 0 || std::abs(alpha) < eps<DT_>())
      {
        if(r != y) std::copy(y, y + size_l(), r);
        return;
      }
      kernel(r, alpha, x, DT_(1), y);
    }
    virtual void diag(DT_* d) const = 0;
  };

  // scalar CSR (bs == 1) or scalar CSR applied to bs-blocked vectors (csrsb)
  template<typename DT_, typename VT_>
  struct CsrOp : public Operator<DT_>
  {
    Index rows, cols;
    int bs;
    std::vector<std::uint32_t> rp, col;
    std::vector<VT_> val;
    Index size_l() const override { return rows * Index(bs); }
    Index size_r() const override { return cols * Index(bs); }
    Index nnz() const override { return rp[rows]; }
    void kernel(DT_* r, DT_ a, const DT_* x, DT_ b, const DT_* y) const override
    {
      if(bs == 1)
        csr_apply<DT_, std::uint32_t, VT_>(r, a, x, b, y, val.data(), col.data(), rp.data(), rows, cols, false);
      else
        csrsb_apply<DT_, std::uint32_t, VT_>(r, a, x, b, y, val.data(), col.data(), rp.data(), rows, bs);
    }
    void diag(DT_* d) const override
    {
      std::vector<DT_> sd(rows);
      diag_csr<DT_, std::uint32_t, VT_>(sd.data(), val.data(), col.data(), rp.data(), rows);
      for(Index i(0); i < rows; ++i)
        for(int h(0); h < bs; ++h)
          d[i * Index(bs) + Index(h)] = sd[i];
    }
  };

  template<typename DT_>
  struct BcsrOp : public Operator<DT_>
  {
    Index rows, cols;
    int b;
    std::vector<std::uint32_t> rp, col;
    std::vector<DT_> val;
    Index size_l() const override { return rows * Index(b); }
    Index size_r() const override { return cols * Index(b); }
    Index nnz() const override { return rp[rows]; }
    void kernel(DT_* r, DT_ a, const DT_* x, DT_ bb, const DT_* y) const override
    {
      csrb_apply<DT_, std::uint32_t>(r, a, x, bb, y, val.data(), col.data(), rp.data(), rows, b, b);
    }
    void diag(DT_* d) const override
    {
      diag_csrb<DT_, std::uint32_t>(d, val.data(), col.data(), rp.data(), rows, b, b);
    }
  };

  // UnitFilter / UnitFilterBlocked with homogeneous values (unit_filter.hpp:399-441); the
  // index list is expanded to scalar entries for blocked vectors.
  template<typename DT_>
  struct Filter
  {
    std::vector<Index> idx;
    void filter_def(DT_* v) const { for(Index i : idx) v[i] = DT_(0); }
    void filter_cor(DT_* v) const { for(Index i : idx) v[i] = DT_(0); }
  };

  // ---------------------------------------------------------------------------------------
  // Solvers (kernel/solver/*.hpp)
  // ---------------------------------------------------------------------------------------
  enum class Status { undefined = 0, progress, success, aborted, diverged, max_iter, stagnated };

  // base.hpp:110-123
  inline bool status_success(Status s)
  {
    return s == Status::success || s == Status::max_iter || s == Status::stagnated;
  }

  template<typename DT_>
  struct SolverBase
  {
    virtual ~SolverBase() {}
    virtual void init() {}
    virtual Status apply(DT_* cor, const DT_* def) = 0;
  };

  template<typename DT_>
  struct Vec
  {
    std::vector<DT_> v;
    void resize(Index n) { v.assign(n, DT_(0)); }
    DT_* p() { return v.data(); }
    const DT_* p() const { return v.data(); }
  };

  // iterative.hpp:55-1077 (criteria, defect bookkeeping, preconditioner call)
  template<typename DT_>
  struct IterativeSolver : public SolverBase<DT_>
  {
    const Operator<DT_>* A;
    const Filter<DT_>* F;
    std::shared_ptr<SolverBase<DT_>> precond;
    DT_ tol_rel = std::sqrt(eps<DT_>());
    DT_ tol_abs = DT_(1) / (eps<DT_>() * eps<DT_>());
    DT_ tol_abs_low = DT_(0);
    DT_ div_rel = DT_(1) / eps<DT_>();
    DT_ div_abs = DT_(1) / (eps<DT_>() * eps<DT_>());
    DT_ stag_rate = DT_(0.95);
    Index min_iter = 0, max_iter = 100, num_iter = 0, min_stag_iter = 0, num_stag_iter = 0;
    DT_ def_init = 0, def_cur = 0, def_prev = 0;
    bool skip_def_calc = true;
    std::vector<double>* history = nullptr;

    Index n() const { return A->size_l(); }

    // iterative.hpp:476-505
    bool is_converged(DT_ d) const { return (d <= tol_abs) && ((d <= tol_rel * def_init) || (d <= tol_abs_low)); }
    bool is_diverged(DT_ d) const { return (d > div_abs) || (d > div_rel * def_init); }

    // iterative.hpp:744-766
    Status set_initial_defect(const DT_* def)
    {
      def_init = def_cur = def_prev = norm2(def, n());
      num_iter = 0;
      num_stag_iter = 0;
      if(history) history->push_back(double(def_init));
      if(!std::isfinite(double(def_init)))
        return Status::aborted;
      if(def_init <= eps<DT_>() * eps<DT_>())
        return Status::success;
      return Status::progress;
    }

    // iterative.hpp:791-832
    Status analyse_defect(Index it, DT_ dc, DT_ dp, bool check_stag)
    {
      if(!std::isfinite(double(dc)))
        return Status::aborted;
      if(is_diverged(dc))
        return Status::diverged;
      if(it < min_iter)
        return Status::progress;
      if(is_converged(dc))
        return Status::success;
      if(it >= max_iter)
        return Status::max_iter;
      if(check_stag && (min_stag_iter > 0))
      {
        if(dc >= stag_rate * dp)
        {
          if(++num_stag_iter >= min_stag_iter)
            return Status::stagnated;
        }
        else
          num_stag_iter = 0;
      }
      return Status::progress;
    }

    // iterative.hpp:848-879
    Status set_new_defect(const DT_* def)
    {
      ++num_iter;
      def_prev = def_cur;
      const bool calc_def = !skip_def_calc || (min_iter < max_iter) || (min_stag_iter > 0);
      if(calc_def)
      {
        def_cur = norm2(def, n());
        if(history) history->push_back(double(def_cur));
      }
      return analyse_defect(num_iter, def_cur, def_prev, true);
    }

    // iterative.hpp:1063-1076
    bool apply_precond(DT_* cor, const DT_* def)
    {
      if(precond)
        return status_success(precond->apply(cor, def));
      std::copy(def, def + n(), cor);
      F->filter_cor(cor);
      return true;
    }
  };

  // jacobi_precond.hpp:137-165 (inv_diag = omega / diag; cor = inv_diag .* def; filter_cor)
  template<typename DT_>
  struct Jacobi : public SolverBase<DT_>
  {
    const Operator<DT_>* A;
    const Filter<DT_>* F;
    DT_ omega;
    Vec<DT_> inv_diag;
    Jacobi(const Operator<DT_>* a, const Filter<DT_>* f, DT_ w) : A(a), F(f), omega(w) {}
    void init() override
    {
      inv_diag.resize(A->size_l());
      A->diag(inv_diag.p());
      component_invert(inv_diag.p(), inv_diag.p(), omega, A->size_l());
    }
    Status apply(DT_* cor, const DT_* def) override
    {
      component_product(cor, inv_diag.p(), def, A->size_l());
      F->filter_cor(cor);
      return Status::success;
    }
  };

  // richardson.hpp:160-228
  template<typename DT_>
  struct Richardson : public IterativeSolver<DT_>
  {
    DT_ omega;
    Vec<DT_> vdef, vcor;
    void init() override
    {
      if(this->precond) this->precond->init();
      vdef.resize(this->n());
      vcor.resize(this->n());
    }
    Status intern(DT_* sol, const DT_* rhs)
    {
      Status st = this->set_initial_defect(vdef.p());
      while(st == Status::progress)
      {
        if(!this->apply_precond(vcor.p(), vdef.p()))
          return Status::aborted;
        // sol = omega*cor + sol  (dense_vector.hpp:883-896 with |omega| < eps -> copy)
        if(std::abs(omega) < eps<DT_>()) {}
        else axpy(sol, omega, vcor.p(), sol, this->n());
        this->A->apply(vdef.p(), sol, rhs, DT_(-1));
        this->F->filter_def(vdef.p());
        st = this->set_new_defect(vdef.p());
      }
      return st;
    }
    Status apply(DT_* cor, const DT_* def) override
    {
      std::copy(def, def + this->n(), vdef.p());
      std::fill(cor, cor + this->n(), DT_(0));
      return intern(cor, def);
    }
    Status correct(DT_* sol, const DT_* rhs)
    {
      this->A->apply(vdef.p(), sol, rhs, DT_(-1));
      this->F->filter_def(vdef.p());
      return intern(sol, rhs);
    }
  };

  // pcg.hpp:140-288
  template<typename DT_>
  struct PCG : public IterativeSolver<DT_>
  {
    Vec<DT_> vr, vp, vt;
    void init() override
    {
      if(this->precond) this->precond->init();
      vr.resize(this->n());
      vp.resize(this->n());
      vt.resize(this->n());
    }
    Status intern(DT_* sol)
    {
      const Index n = this->n();
      DT_* r = vr.p();
      DT_* p = vp.p();
      DT_* q = vt.p();
      DT_* z = vt.p();
      Status st = this->set_initial_defect(r);
      if(st != Status::progress)
        return st;
      if(!this->apply_precond(p, r))
        return Status::aborted;
      DT_ gamma = dot(r, p, n);
      while(st == Status::progress)
      {
        this->A->apply(q, p);
        this->F->filter_def(q);
        const DT_ alpha = gamma / dot(q, p, n);
        // dense_vector.hpp:891: |alpha| < eps copies, anything else (a NaN alpha included) updates
        if(!(std::abs(alpha) < eps<DT_>())) axpy(sol, alpha, p, sol, n);
        if(!(std::abs(alpha) < eps<DT_>())) axpy(r, -alpha, q, r, n);
        st = this->set_new_defect(r);
        if(st != Status::progress)
          return st;
        if(!this->apply_precond(z, r))
          return Status::aborted;
        const DT_ gamma2 = gamma;
        gamma = dot(r, z, n);
        const DT_ beta = gamma / gamma2;
        // p = beta*p + z (axpy with r == x)
        if(std::abs(beta) < eps<DT_>()) std::copy(z, z + n, p);
        else axpy(p, beta, p, z, n);
      }
      return Status::undefined;
    }
    Status apply(DT_* cor, const DT_* def) override
    {
      std::copy(def, def + this->n(), vr.p());
      std::fill(cor, cor + this->n(), DT_(0));
      return intern(cor);
    }
    Status correct(DT_* sol, const DT_* rhs)
    {
      this->A->apply(vr.p(), sol, rhs, DT_(-1));
      this->F->filter_def(vr.p());
      return intern(sol);
    }
  };

  // Exact coarse solve: dense LU with partial pivoting of the filtered coarse matrix in FP64
  // (Dirichlet rows replaced by unit rows, as a direct solver sees a unit-filtered matrix);
  // the result is correction-filtered.
  template<typename DT_>
  struct DenseCoarse : public SolverBase<DT_>
  {
    const Operator<DT_>* A;
    const Filter<DT_>* F;
    Index n = 0;
    std::vector<double> lu;
    std::vector<Index> piv;
    std::vector<double> dense_in; // optional FP64 dense matrix supplied by the caller
    void init() override
    {
      n = A->size_l();
      lu.assign(n * n, 0.0);
      if(!dense_in.empty())
        lu = dense_in;
      else
      {
        // assemble dense by applying A to unit vectors
        std::vector<DT_> e(n, DT_(0)), c(n);
        for(Index j(0); j < n; ++j)
        {
          e[j] = DT_(1);
          A->apply(c.data(), e.data());
          for(Index i(0); i < n; ++i)
            lu[i * n + j] = double(c[i]);
          e[j] = DT_(0);
        }
      }
      for(Index i : F->idx)
      {
        for(Index j(0); j < n; ++j)
          lu[i * n + j] = 0.0;
        lu[i * n + i] = 1.0;
      }
      piv.resize(n);
      for(Index k(0); k < n; ++k)
      {
        Index pr = k;
        for(Index i(k + 1); i < n; ++i)
          if(std::abs(lu[i * n + k]) > std::abs(lu[pr * n + k]))
            pr = i;
        piv[k] = pr;
        if(pr != k)
          for(Index j(0); j < n; ++j)
            std::swap(lu[k * n + j], lu[pr * n + j]);
        for(Index i(k + 1); i < n; ++i)
        {
          lu[i * n + k] /= lu[k * n + k];
          for(Index j(k + 1); j < n; ++j)
            lu[i * n + j] -= lu[i * n + k] * lu[k * n + j];
        }
      }
    }
    Status apply(DT_* cor, const DT_* def) override
    {
      std::vector<double> y(n);
      for(Index i(0); i < n; ++i)
        y[i] = double(def[i]);
      for(Index i : F->idx)
        y[i] = 0.0;
      for(Index k(0); k < n; ++k)
        std::swap(y[k], y[piv[k]]);
      for(Index i(0); i < n; ++i)
        for(Index j(0); j < i; ++j)
          y[i] -= lu[i * n + j] * y[j];
      for(Index i(n); i > 0; --i)
      {
        for(Index j(i); j < n; ++j)
          y[i - 1] -= lu[(i - 1) * n + j] * y[j];
        y[i - 1] /= lu[(i - 1) * n + (i - 1)];
      }
      for(Index i(0); i < n; ++i)
        cor[i] = DT_(y[i]);
      F->filter_cor(cor);
      return Status::success;
    }
  };

  // multigrid.hpp:419-957 (hierarchy, 5 vectors per level) and 1268-2000 (V-cycle)
  template<typename DT_>
  struct MGLevel
  {
    std::unique_ptr<Operator<DT_>> A, P, R; // P: this(fine) <- next(coarse); R = P^T
    Filter<DT_> F;
    std::shared_ptr<SolverBase<DT_>> smoother_pre, smoother_post, coarse;
    Vec<DT_> rhs, sol, def, cor, tmp;
  };

  template<typename DT_>
  struct MultiGrid : public SolverBase<DT_>
  {
    std::vector<std::unique_ptr<MGLevel<DT_>>> lv; // 0 = finest
    void init() override
    {
      for(auto& l : lv)
      {
        const Index n = l->A->size_l();
        l->rhs.resize(n); l->sol.resize(n); l->def.resize(n); l->cor.resize(n); l->tmp.resize(n);
        if(l->smoother_pre) l->smoother_pre->init();
        if(l->smoother_post && l->smoother_post != l->smoother_pre) l->smoother_post->init();
        if(l->coarse && l->coarse != l->smoother_pre) l->coarse->init();
      }
    }
    int cycle = 0;  // 0 V, 1 F, 2 W (MultiGridCycle, multigrid.hpp:26-34)

    // multigrid.hpp:1711-1832: restriction from level cur down to the coarse level; the
    // pre-smoother runs on every level below cur, and on cur itself only when smooth
    void rest(std::size_t cur, bool smooth)
    {
      for(std::size_t i(cur); i + 1 < lv.size(); ++i)
      {
        MGLevel<DT_>& f = *lv[i];
        if(smooth || i > cur)
        {
          if(f.smoother_pre)
          {
            f.smoother_pre->apply(f.sol.p(), f.rhs.p());
            f.A->apply(f.def.p(), f.sol.p(), f.rhs.p(), DT_(-1));
          }
          else
          {
            std::fill(f.sol.v.begin(), f.sol.v.end(), DT_(0));
            f.def.v = f.rhs.v;
          }
        }
        f.F.filter_def(f.def.p());
        MGLevel<DT_>& c = *lv[i + 1];
        f.R->apply(c.rhs.p(), f.def.p());
        c.F.filter_def(c.rhs.p());
      }
    }
    // multigrid.hpp:1524-1587
    Status coarse()
    {
      MGLevel<DT_>& c = *lv.back();
      if(c.coarse)
      {
        if(!status_success(c.coarse->apply(c.sol.p(), c.rhs.p())))
          return Status::aborted;
      }
      else
      {
        c.sol.v = c.rhs.v;
        c.F.filter_cor(c.sol.p());
      }
      return Status::success;
    }
    // multigrid.hpp:1843-2000 (fixed coarse-grid correction, omega_cgc = 1): prolongation from
    // the coarse level up to level cur; the post-smoother runs on every level below cur, and
    // on cur itself only when smooth
    void prol(std::size_t cur, bool smooth)
    {
      for(std::size_t i(lv.size() - 1); i > cur;)
      {
        --i;
        MGLevel<DT_>& f = *lv[i];
        MGLevel<DT_>& c = *lv[i + 1];
        const Index n = f.A->size_l();
        f.P->apply(f.cor.p(), c.sol.p());
        f.F.filter_cor(f.cor.p());
        axpy(f.sol.p(), DT_(1), f.cor.p(), f.sol.p(), n);
        if(f.smoother_post && (smooth || i > cur))
        {
          f.A->apply(f.def.p(), f.sol.p(), f.rhs.p(), DT_(-1));
          f.F.filter_def(f.def.p());
          f.smoother_post->apply(f.cor.p(), f.def.p());
          axpy(f.sol.p(), DT_(1), f.cor.p(), f.sol.p(), n);
        }
      }
    }
    // multigrid.hpp:1589-1621: one smoother on the defect, correction added, defect updated
    void smooth_def(MGLevel<DT_>& l, SolverBase<DT_>& sm)
    {
      const Index n = l.A->size_l();
      sm.apply(l.cor.p(), l.def.p());
      l.F.filter_cor(l.cor.p());
      axpy(l.sol.p(), DT_(1), l.cor.p(), l.sol.p(), n);
      l.A->apply(l.def.p(), l.sol.p(), l.rhs.p(), DT_(-1));
      l.F.filter_def(l.def.p());
    }
    // multigrid.hpp:1630-1700 with no peak smoother: the pre- and then the post-smoother
    void smooth_peak(std::size_t i)
    {
      MGLevel<DT_>& l = *lv[i];
      l.A->apply(l.def.p(), l.sol.p(), l.rhs.p(), DT_(-1));
      l.F.filter_def(l.def.p());
      if(l.smoother_pre) smooth_def(l, *l.smoother_pre);
      if(l.smoother_post) smooth_def(l, *l.smoother_post);
    }
    Status apply(DT_* cor, const DT_* def) override
    {
      MGLevel<DT_>& top = *lv.front();
      std::copy(def, def + top.A->size_l(), top.rhs.p());
      const std::size_t last = lv.size() - 1;
      Status st = Status::success;
      rest(0, true);
      if(cycle == 1)
      {
        // multigrid.hpp:1366-1406: intermediate peak levels from the coarsest but one upwards
        for(std::size_t pk = last > 0 ? last - 1 : 0; pk > 0; --pk)
        {
          coarse();
          prol(pk, false);
          smooth_peak(pk);
          rest(pk, false);
        }
        st = coarse();
      }
      else if(cycle == 2)
      {
        // multigrid.hpp:1409-1519: 2^last coarse solves; each round's peak level is the lowest
        // level above the coarse one whose counter is 0, the counters below it are reset
        st = coarse();
        std::vector<int> counter(last + 1, 0);
        const std::size_t num_cgs = std::size_t(1) << last;
        for(std::size_t cgs(1); cgs < num_cgs; ++cgs)
        {
          std::size_t pk = last;
          while(pk > 0)
          {
            if(counter[--pk] == 0)
              break;
          }
          for(std::size_t i = last - 1; i > pk; --i)
            counter[i] = 0;
          ++counter[pk];
          prol(pk, false);
          smooth_peak(pk);
          rest(pk, false);
          st = coarse();
        }
      }
      else
        st = coarse();
      prol(0, true);
      std::copy(top.sol.v.begin(), top.sol.v.end(), cor);
      return st;
    }
  };

  // convert_precond.hpp:100-119 (outer FP64 defect -> inner precision, correction back)
  template<typename DTO_, typename DTI_>
  struct ConvertPrecond : public SolverBase<DTO_>
  {
    std::shared_ptr<SolverBase<DTI_>> inner;
    Index n;
    std::vector<DTI_> di, ci;
    void init() override { inner->init(); di.resize(n); ci.resize(n); }
    Status apply(DTO_* cor, const DTO_* def) override
    {
      for(Index i(0); i < n; ++i)
        di[i] = DTI_(def[i]);
      std::fill(ci.begin(), ci.end(), DTI_(0));
      Status st = inner->apply(ci.data(), di.data());
      if(!status_success(st))
        return st;
      for(Index i(0); i < n; ++i)
        cor[i] = DTO_(ci[i]);
      return Status::success;
    }
  };
} // namespace orc

// =========================================================================================
// C ABI for the test harness (ctypes)
// =========================================================================================
using orc::Index;
typedef std::uint32_t u32;
typedef std::uint64_t u64;

extern "C" {

// ---- arch kernels -----------------------------------------------------------------------
void orc_csr_apply_f64_u32(double* r, double a, const double* x, double b, const double* y, const double* val,
                           const u32* col, const u32* rp, u64 rows, u64 cols, int tr)
{ orc::csr_apply<double, u32>(r, a, x, b, y, val, col, rp, rows, cols, tr != 0); }
void orc_csr_apply_f64_u64(double* r, double a, const double* x, double b, const double* y, const double* val,
                           const u64* col, const u64* rp, u64 rows, u64 cols, int tr)
{ orc::csr_apply<double, u64>(r, a, x, b, y, val, col, rp, rows, cols, tr != 0); }
void orc_csr_apply_f32_u32(float* r, float a, const float* x, float b, const float* y, const float* val,
                           const u32* col, const u32* rp, u64 rows, u64 cols, int tr)
{ orc::csr_apply<float, u32>(r, a, x, b, y, val, col, rp, rows, cols, tr != 0); }
void orc_csr_apply_h16_u32(float* r, float a, const float* x, float b, const float* y, const std::uint16_t* val,
                           const u32* col, const u32* rp, u64 rows, u64 cols)
{
  orc::csr_apply<float, u32, orc::half_t>(r, a, x, b, y, reinterpret_cast<const orc::half_t*>(val), col, rp, rows, cols, false);
}
void orc_csrb_apply_f64_u32(double* r, double a, const double* x, double b, const double* y, const double* val,
                            const u32* col, const u32* rp, u64 rows, int bh, int bw)
{ orc::csrb_apply<double, u32>(r, a, x, b, y, val, col, rp, rows, bh, bw); }
void orc_csrb_apply_f32_u32(float* r, float a, const float* x, float b, const float* y, const float* val,
                            const u32* col, const u32* rp, u64 rows, int bh, int bw)
{ orc::csrb_apply<float, u32>(r, a, x, b, y, val, col, rp, rows, bh, bw); }
void orc_csrsb_apply_f64_u32(double* r, double a, const double* x, double b, const double* y, const double* val,
                             const u32* col, const u32* rp, u64 rows, int bs)
{ orc::csrsb_apply<double, u32>(r, a, x, b, y, val, col, rp, rows, bs); }
u64 orc_ell_val_size_u32(u64 C, u64 rows, const u32* rp) { return orc::ell_val_size<u32>(C, rows, rp); }
void orc_ell_from_csr_f64_u32(u64 C, u64 rows, const u32* rp, const u32* col, const double* val,
                              u32* cs, u32* cl, u32* rl, double* ev, u32* ec)
{ orc::ell_from_csr<double, u32>(C, rows, rp, col, val, cs, cl, rl, ev, ec); }
void orc_banded_apply_f64_u32(double* r, double a, const double* x, double b, const double* y, const double* val,
                              const u32* offsets, u64 noo, u64 rows, u64 columns)
{
  orc::banded_apply(r, a, x, b, y, val, offsets, noo, rows, columns);
}

void orc_banded_apply_f32_u32(float* r, float a, const float* x, float b, const float* y, const float* val,
                              const u32* offsets, u64 noo, u64 rows, u64 columns)
{
  orc::banded_apply(r, a, x, b, y, val, offsets, noo, rows, columns);
}

void orc_ell_apply_f64_u32(double* r, double a, const double* x, double b, const double* y, const double* val,
                           const u32* col, const u32* cs, const u32* cl, u64 C, u64 rows)
{ orc::ell_apply<double, u32>(r, a, x, b, y, val, col, cs, cl, C, rows); }
void orc_axpy_f64(double* r, double a, const double* x, const double* y, u64 n) { orc::axpy(r, a, x, y, n); }
void orc_axpy_f32(float* r, float a, const float* x, const float* y, u64 n) { orc::axpy(r, a, x, y, n); }
double orc_dot_f64(const double* x, const double* y, u64 n) { return orc::dot(x, y, n); }
float orc_dot_f32(const float* x, const float* y, u64 n) { return orc::dot(x, y, n); }
double orc_norm2_f64(const double* x, u64 n) { return orc::norm2(x, n); }
float orc_norm2_f32(const float* x, u64 n) { return orc::norm2(x, n); }
void orc_component_product_f64(double* r, const double* x, const double* y, u64 n) { orc::component_product(r, x, y, n); }
void orc_component_invert_f64(double* r, const double* x, double s, u64 n) { orc::component_invert(r, x, s, n); }
void orc_scale_f64(double* r, const double* x, double s, u64 n) { orc::scale(r, x, s, n); }
void orc_diag_csr_f64_u32(double* d, const double* val, const u32* col, const u32* rp, u64 rows)
{ orc::diag_csr<double, u32>(d, val, col, rp, rows); }
void orc_diag_csrb_f64_u32(double* d, const double* val, const u32* col, const u32* rp, u64 rows, int bh, int bw)
{ orc::diag_csrb<double, u32>(d, val, col, rp, rows, bh, bw); }
void orc_unit_filter_rhs_f64_u32(double* v, const double* sv, const u32* idx, u64 ue) { orc::unit_filter_rhs(v, sv, idx, ue); }
void orc_unit_filter_def_f64_u32(double* v, const u32* idx, u64 ue) { orc::unit_filter_def(v, idx, ue); }
void orc_transpose_csr_f64_u32(u64 rows, u64 cols, const u32* rp, const u32* col, const double* val,
                               u32* trp, u32* tcol, double* tval)
{ orc::transpose_csr<double, u32>(rows, cols, rp, col, val, trp, tcol, tval); }
std::uint16_t orc_float_to_half(float x) { return orc::float_to_half(x); }
float orc_half_to_float(std::uint16_t h) { return orc::half_to_float(h); }

// ---- pointstar factory --------------------------------------------------------------------
u64 orc_pointstar_fd_nnz(u64 m, u64 d) { return orc::pointstar_fd_nnz(m, d); }
void orc_pointstar_fd_f64_u32(u64 m, u64 d, u32* rp, u32* col, double* val) { orc::pointstar_fd<double, u32>(m, d, rp, col, val); }
void orc_pointstar_q2_bubble_f64(u64 m, u64 d, double* v) { orc::pointstar_q2_bubble<double>(m, d, v); }

} // extern "C"

// ---- solver hierarchy ---------------------------------------------------------------------
// A system is a list of levels (0 = finest).  Each level holds its matrix (scalar CSR with
// block == 1, or BCSR with block x block blocks), the Dirichlet node list, and the transfer
// to the next coarser level as scalar CSR P (fine x coarse) and R (coarse x fine); for
// block > 1 the transfers act node-wise on blocked vectors (csrsb).
struct OrcLevelData
{
  u64 n; int block;
  std::vector<u32> rp, col; std::vector<double> val;
  std::vector<u32> fidx;
  u64 nc = 0;
  std::vector<u32> prp, pcol, rrp, rcol; std::vector<double> pval, rval;
};

struct OrcSystem
{
  std::vector<OrcLevelData> levels;
};

namespace
{
  // value type converters for the three precisions
  template<typename DT_, typename VT_> VT_ store(double v) { return VT_(DT_(v)); }
  template<> orc::half_t store<float, orc::half_t>(double v) { orc::half_t h; h.bits = orc::float_to_half(float(v)); return h; }

  template<typename DT_, typename VT_>
  std::unique_ptr<orc::Operator<DT_>> make_csr(u64 rows, u64 cols, int bs, const std::vector<u32>& rp,
                                               const std::vector<u32>& col, const std::vector<double>& val)
  {
    auto op = std::unique_ptr<orc::CsrOp<DT_, VT_>>(new orc::CsrOp<DT_, VT_>());
    op->rows = rows; op->cols = cols; op->bs = bs; op->rp = rp; op->col = col;
    op->val.resize(val.size());
    for(std::size_t i(0); i < val.size(); ++i)
      op->val[i] = store<DT_, VT_>(val[i]);
    return std::unique_ptr<orc::Operator<DT_>>(op.release());
  }

  template<typename DT_, typename VT_>
  std::unique_ptr<orc::Operator<DT_>> make_matrix(const OrcLevelData& L)
  {
    if(L.block == 1)
      return make_csr<DT_, VT_>(L.n, L.n, 1, L.rp, L.col, L.val);
    auto op = std::unique_ptr<orc::BcsrOp<DT_>>(new orc::BcsrOp<DT_>());
    op->rows = L.n; op->cols = L.n; op->b = L.block; op->rp = L.rp; op->col = L.col;
    op->val.resize(L.val.size());
    for(std::size_t i(0); i < L.val.size(); ++i)
      op->val[i] = DT_(L.val[i]);
    return std::unique_ptr<orc::Operator<DT_>>(op.release());
  }

  // solver parameters (a flat double array so that ctypes can pass it)
  struct Params
  {
    int kind;          // 0 PCG-MG, 1 PCG-Jacobi, 2 CG, 3 Richardson-Jacobi, 4 MG as Richardson preconditioner
    double tol_rel;
    u64 max_iter;
    double omega_jac, omega_rich;
    u64 pre, post;
    int coarse;        // 0 dense exact (FP64), 1 smoother as coarse solver, 2 none (identity)
    int inner;         // 0 FP64, 1 FP32, 2 FP16-stored values with FP32 arithmetic
    double tol_abs;
    u64 min_iter;
    int use_def;       // 1: solve via correct(x, b) (Solver::solve); 0: apply(cor, def)
    int cycle;         // multigrid cycle: 0 V, 1 F, 2 W
  };

  Params parse(const double* p)
  {
    Params q;
    q.kind = int(p[0]); q.tol_rel = p[1]; q.max_iter = u64(p[2]);
    q.omega_jac = p[3]; q.omega_rich = p[4]; q.pre = u64(p[5]); q.post = u64(p[6]);
    q.coarse = int(p[7]); q.inner = int(p[8]); q.tol_abs = p[9]; q.min_iter = u64(p[10]);
    q.use_def = int(p[11]); q.cycle = int(p[12]);
    return q;
  }

  template<typename DT_>
  std::shared_ptr<orc::Richardson<DT_>> make_smoother(const orc::Operator<DT_>* A, const orc::Filter<DT_>* F,
                                                      double wj, double wr, u64 steps)
  {
    auto jac = std::make_shared<orc::Jacobi<DT_>>(A, F, DT_(wj));
    auto sm = std::make_shared<orc::Richardson<DT_>>();
    sm->A = A; sm->F = F; sm->precond = jac; sm->omega = DT_(wr);
    sm->min_iter = steps; sm->max_iter = steps;
    return sm;
  }

  // builds a MultiGrid<DT_> with value storage VT_ over all levels of the system
  template<typename DT_, typename VT_>
  std::shared_ptr<orc::MultiGrid<DT_>> build_mg(const OrcSystem& s, const Params& q)
  {
    auto mg = std::make_shared<orc::MultiGrid<DT_>>();
    mg->cycle = q.cycle;
    for(std::size_t i(0); i < s.levels.size(); ++i)
    {
      const OrcLevelData& L = s.levels[i];
      std::unique_ptr<orc::MGLevel<DT_>> lv(new orc::MGLevel<DT_>());
      lv->A = make_matrix<DT_, VT_>(L);
      for(u32 k : L.fidx)
        for(int b(0); b < L.block; ++b)
          lv->F.idx.push_back(Index(k) * Index(L.block) + Index(b));
      if(i + 1 < s.levels.size())
      {
        lv->P = make_csr<DT_, VT_>(L.n, L.nc, L.block, L.prp, L.pcol, L.pval);
        lv->R = make_csr<DT_, VT_>(L.nc, L.n, L.block, L.rrp, L.rcol, L.rval);
        if(q.pre > 0) lv->smoother_pre = make_smoother<DT_>(lv->A.get(), &lv->F, q.omega_jac, q.omega_rich, q.pre);
        if(q.post > 0)
          lv->smoother_post = (q.pre == q.post) ? lv->smoother_pre
                            : make_smoother<DT_>(lv->A.get(), &lv->F, q.omega_jac, q.omega_rich, q.post);
      }
      else
      {
        if(q.coarse == 0)
        {
          auto cs = std::make_shared<orc::DenseCoarse<DT_>>();
          cs->A = lv->A.get(); cs->F = &lv->F;
          // the coarse matrix in FP64 regardless of the hierarchy precision
          auto a64 = make_matrix<double, double>(L);
          const Index n = a64->size_l();
          cs->dense_in.assign(n * n, 0.0);
          std::vector<double> e(n, 0.0), c(n);
          for(Index j(0); j < n; ++j)
          {
            e[j] = 1.0;
            a64->apply(c.data(), e.data());
            for(Index r(0); r < n; ++r)
              cs->dense_in[r * n + j] = c[r];
            e[j] = 0.0;
          }
          lv->coarse = cs;
        }
        else if(q.coarse == 1)
          lv->coarse = make_smoother<DT_>(lv->A.get(), &lv->F, q.omega_jac, q.omega_rich, std::max<u64>(q.pre, 1));
      }
      mg->lv.push_back(std::move(lv));
    }
    return mg;
  }
} // namespace

extern "C" {

// wall time of the solve phase of the last orc_system_solve (matrix copies, hierarchy
// construction and solver init excluded)
static double g_last_solve_seconds = 0.0;
double orc_last_solve_seconds() { return g_last_solve_seconds; }

void* orc_system_create() { return new OrcSystem(); }
void orc_system_destroy(void* h) { delete static_cast<OrcSystem*>(h); }

int orc_system_add_level(void* h, u64 n, int block, const u32* rp, const u32* col, const double* val,
                         u64 nfilt, const u32* fidx, u64 nc, const u32* prp, const u32* pcol, const double* pval,
                         const u32* rrp, const u32* rcol, const double* rval)
{
  OrcSystem* s = static_cast<OrcSystem*>(h);
  OrcLevelData L;
  L.n = n; L.block = block;
  L.rp.assign(rp, rp + n + 1);
  const u64 nnz = rp[n];
  L.col.assign(col, col + nnz);
  L.val.assign(val, val + nnz * u64(block) * u64(block));
  L.fidx.assign(fidx, fidx + nfilt);
  L.nc = nc;
  if(nc > 0)
  {
    L.prp.assign(prp, prp + n + 1);
    L.pcol.assign(pcol, pcol + prp[n]);
    L.pval.assign(pval, pval + prp[n]);
    L.rrp.assign(rrp, rrp + nc + 1);
    L.rcol.assign(rcol, rcol + rrp[nc]);
    L.rval.assign(rval, rval + rrp[nc]);
  }
  s->levels.push_back(std::move(L));
  return 0;
}


// Runs one solve in FP64 on the finest level; the preconditioner hierarchy runs in the inner
// precision.  Returns the number of iterations; status, defect history (norms, one per
// computed defect, at most hist_cap) are written to the outputs.
u64 orc_system_solve(void* h, const double* praw, double* x, const double* b, int* status,
                     double* hist, u64 hist_cap, u64* hist_len)
{
  const OrcSystem& s = *static_cast<OrcSystem*>(h);
  const Params q = parse(praw);
  const OrcLevelData& L0 = s.levels.front();
  auto A = make_matrix<double, double>(L0);
  orc::Filter<double> F;
  for(u32 k : L0.fidx)
    for(int bb(0); bb < L0.block; ++bb)
      F.idx.push_back(Index(k) * Index(L0.block) + Index(bb));
  const Index n = A->size_l();

  std::shared_ptr<orc::SolverBase<double>> precond;
  if(q.kind == 0 || q.kind == 4)
  {
    if(q.inner == 0)
      precond = build_mg<double, double>(s, q);
    else
    {
      auto cp = std::make_shared<orc::ConvertPrecond<double, float>>();
      cp->n = n;
      if(q.inner == 1) cp->inner = build_mg<float, float>(s, q);
      else cp->inner = build_mg<float, orc::half_t>(s, q);
      precond = cp;
    }
  }
  else if(q.kind == 1 || q.kind == 3)
    precond = std::make_shared<orc::Jacobi<double>>(A.get(), &F, q.omega_jac);

  std::vector<double> history;
  orc::Status st;
  Index iters;
  if(q.kind == 3 || q.kind == 4)
  {
    orc::Richardson<double> sol;
    sol.A = A.get(); sol.F = &F; sol.precond = precond; sol.omega = q.omega_rich;
    sol.tol_rel = q.tol_rel; sol.max_iter = q.max_iter; sol.min_iter = q.min_iter;
    if(q.tol_abs > 0) sol.tol_abs = q.tol_abs;
    sol.history = &history;
    sol.init();
    const auto t0 = std::chrono::steady_clock::now();
    if(q.use_def) st = sol.correct(x, b);
    else st = sol.apply(x, b);
    g_last_solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    iters = sol.num_iter;
  }
  else
  {
    orc::PCG<double> sol;
    sol.A = A.get(); sol.F = &F; sol.precond = precond;
    sol.tol_rel = q.tol_rel; sol.max_iter = q.max_iter; sol.min_iter = q.min_iter;
    if(q.tol_abs > 0) sol.tol_abs = q.tol_abs;
    sol.history = &history;
    sol.init();
    const auto t0 = std::chrono::steady_clock::now();
    if(q.use_def) st = sol.correct(x, b);
    else st = sol.apply(x, b);
    g_last_solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    iters = sol.num_iter;
  }
  *status = int(st);
  const u64 m = std::min<u64>(hist_cap, history.size());
  for(u64 i(0); i < m; ++i)
    hist[i] = history[i];
  *hist_len = history.size();
  (void)n;
  return iters;
}

// One application of the MG preconditioner (one cycle of the kind the parameters name) in the
// requested precision: cor = MG(def).  Used to check the device cycle in isolation.
int orc_system_vcycle(void* h, const double* praw, double* cor, const double* def)
{
  const OrcSystem& s = *static_cast<OrcSystem*>(h);
  const Params q = parse(praw);
  const Index n = s.levels.front().n * Index(s.levels.front().block);
  orc::Status st;
  if(q.inner == 0)
  {
    auto mg = build_mg<double, double>(s, q);
    mg->init();
    st = mg->apply(cor, def);
  }
  else
  {
    orc::ConvertPrecond<double, float> cp;
    cp.n = n;
    if(q.inner == 1) cp.inner = build_mg<float, float>(s, q);
    else cp.inner = build_mg<float, orc::half_t>(s, q);
    cp.init();
    st = cp.apply(cor, def);
  }
  return int(st);
}

// dL/dk_e = -lambda_e^T E u_e for one system: E = element stiffness (nloc x nloc, lexicographic
// local nodes), u and lambda nodal vectors of the level-`level` Q`degree` space on the unit
// square/cube (lexicographic nodes, x fastest), grad per cell (x fastest).
int orc_cell_gradient(int dim, int degree, int level, const double* E, const double* u, const double* lam, double* grad)
{
  if(dim < 2 || dim > 3 || degree < 1 || degree > 2) return 1;
  const u64 m = u64(1) << level;
  const int q = degree + 1;
  const int nloc = dim == 2 ? q * q : q * q * q;
  const u64 np1 = u64(degree) * m + 1;
  const u64 ncell = dim == 2 ? m * m : m * m * m;
  std::vector<u64> nodes(nloc);
  for(u64 c = 0; c < ncell; ++c)
  {
    u64 e[3] = {c % m, (c / m) % m, c / (m * m)};
    for(int a = 0; a < nloc; ++a)
    {
      int s = a;
      u64 node = 0, stride = 1;
      for(int d = 0; d < dim; ++d)
      {
        node += (u64(degree) * e[d] + u64(s % q)) * stride;
        s /= q;
        stride *= np1;
      }
      nodes[a] = node;
    }
    double acc = 0.0;
    for(int a = 0; a < nloc; ++a)
      for(int b = 0; b < nloc; ++b)
        acc += lam[nodes[a]] * E[a * nloc + b] * u[nodes[b]];
    grad[c] = -acc;
  }
  return 0;
}

} // extern "C"
