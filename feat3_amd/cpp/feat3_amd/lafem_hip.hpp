// lafem_hip.hpp -- the LAFEM side of the reference interface on the HIP backend: the Mem::HIP
// tag, MemoryPool<Mem::HIP> and the containers DenseVector, SparseMatrixCSR,
// SparseMatrixELL, SparseMatrixBCSR, UnitFilter and Transfer over libfeat3_hip.so.
//
// Header-only C++14 on top of the C ABI in feat3_hip.h (no HIP headers needed: host code that
// includes this compiles with any C++ compiler and links -lfeat3_hip).  Member names, argument
// order and special cases follow the reference containers:
//   Mem tags                kernel/archs.hpp:82-118
//   MemoryPool              kernel/util/memory_pool.hpp:53-377
//   DenseVector             kernel/lafem/dense_vector.hpp:57-1181 (axpy/dot/norm2 ... :883-1026)
//   SparseMatrixCSR         kernel/lafem/sparse_matrix_csr.hpp:1899-2051,2321-2327,2491
//   SparseMatrixELL         kernel/lafem/sparse_matrix_ell.hpp:865-960,1983-2074
//   SparseMatrixBCSR        kernel/lafem/sparse_matrix_bcsr.hpp:1295-1600,1895-1901
//   UnitFilter              kernel/lafem/unit_filter.hpp:399-441
//   Transfer                kernel/lafem/transfer.hpp:249-285
// Errors are thrown as InternalError (kernel/util/exception.hpp), as the reference's CUDA
// backend does for library failures.
#pragma once

#include "feat3_hip.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <algorithm>
#include <vector>

namespace feat3_amd
{
  typedef std::uint64_t Index;

  namespace Mem
  {
    struct Main {};
    struct HIP {};   // the device tag this backend adds next to Main (and CUDA in the reference)
  }

  class InternalError : public std::runtime_error
  {
  public:
    explicit InternalError(const std::string& msg) : std::runtime_error(msg) {}
  };

  inline void hip_check(int rc, const char* what)
  {
    if(rc != FEAT3_HIP_OK)
      throw InternalError(std::string(what) + ": " + feat3_hip_error_string(rc));
  }

  namespace Math
  {
    template<typename T> inline T eps() { return std::numeric_limits<T>::epsilon(); }
    template<typename T> inline T abs(T x) { return x < T(0) ? -x : x; }
  }

  // MemoryPool<Mem::HIP>: refcounted device allocations, synchronous transfers (the reference's
  // upload/download are synchronous cudaMemcpy calls, memory_pool.cu:158-175)
  template<typename Mem_> class MemoryPool;

  template<>
  class MemoryPool<Mem::HIP>
  {
  public:
    static void initialise(int device = 0) { hip_check(feat3_hip_initialise(device), "MemoryPool<HIP>::initialise"); }
    static Index finalise()
    {
      std::uint64_t leaked = 0;
      feat3_hip_finalise(&leaked);
      return leaked;
    }
    template<typename T> static T* allocate_memory(Index count)
    {
      void* p = nullptr;
      hip_check(feat3_hip_malloc(&p, sizeof(T) * count), "MemoryPool<HIP>::allocate_memory");
      return static_cast<T*>(p);
    }
    static void increase_memory(void* p) { if(p) hip_check(feat3_hip_increase(p), "MemoryPool<HIP>::increase_memory"); }
    static void release_memory(void* p) { if(p) hip_check(feat3_hip_free(p), "MemoryPool<HIP>::release_memory"); }
    template<typename T> static void upload(T* dst, const T* src, Index count)
    {
      hip_check(feat3_hip_upload(nullptr, dst, src, sizeof(T) * count), "MemoryPool<HIP>::upload");
      synchronise();
    }
    template<typename T> static void download(T* dst, const T* src, Index count)
    {
      hip_check(feat3_hip_download(nullptr, dst, src, sizeof(T) * count), "MemoryPool<HIP>::download");
      synchronise();
    }
    template<typename T> static void copy(T* dst, const T* src, Index count)
    {
      if(dst != src) hip_check(feat3_hip_copy(nullptr, dst, src, sizeof(T) * count), "MemoryPool<HIP>::copy");
    }
    static void set_memory(double* p, double v, Index n) { hip_check(lafem_hip_set_memory_f64(nullptr, p, v, n), "set_memory"); }
    static void set_memory(float* p, float v, Index n) { hip_check(lafem_hip_set_memory_f32(nullptr, p, v, n), "set_memory"); }
    static void set_memory(std::uint32_t* p, std::uint32_t v, Index n) { hip_check(lafem_hip_set_memory_u32(nullptr, p, v, n), "set_memory"); }
    static void set_memory(std::uint64_t* p, std::uint64_t v, Index n) { hip_check(lafem_hip_set_memory_u64(nullptr, p, v, n), "set_memory"); }
    static void convert(float* d, const double* s, Index n) { hip_check(lafem_hip_convert_f32_f64(nullptr, d, s, n), "convert"); }
    static void convert(double* d, const float* s, Index n) { hip_check(lafem_hip_convert_f64_f32(nullptr, d, s, n), "convert"); }
    static void convert(double* d, const double* s, Index n) { copy(d, s, n); }
    static void convert(float* d, const float* s, Index n) { copy(d, s, n); }
    static void synchronise() { hip_check(feat3_hip_synchronise(nullptr), "MemoryPool<HIP>::synchronise"); }
  };

  namespace LAFEM
  {
    // the file modes of lafem/base.hpp:31-46 (same values); this backend reads and writes fm_mtx
    // and fm_csr
    enum class FileMode { fm_exp = 0, fm_dv, fm_mtx, fm_ell, fm_csr, fm_coo, fm_bm, fm_dm, fm_sv, fm_dvb, fm_bcsr, fm_cscr, fm_binary };

    // ---- typed dispatch onto the C ABI ------------------------------------------------------
    namespace Arch
    {
      template<typename DT_, typename IT_> struct Ops;

#define F3_OPS(DT, IT, D, I)                                                                                       \
      template<> struct Ops<DT, IT>                                                                                \
      {                                                                                                            \
        static int csr(DT* r, DT a, const DT* x, DT b, const DT* y, const DT* v, const IT* c, const IT* rp,       \
                       Index rows, Index cols, Index nnz, bool tr)                                                 \
        { return lafem_hip_apply_csr_##D##_##I(nullptr, r, a, x, b, y, v, c, rp, rows, cols, nnz, tr ? 1 : 0); }  \
        static int ell(DT* r, DT a, const DT* x, DT b, const DT* y, const DT* v, const IT* c, const IT* cs,       \
                       const IT* cl, Index C, Index rows)                                                          \
        { return lafem_hip_apply_ell_##D##_##I(nullptr, r, a, x, b, y, v, c, cs, cl, C, rows); }                  \
        static int banded(DT* r, DT a, const DT* x, DT b, const DT* y, const DT* v, const IT* offs, Index noo,    \
                          Index rows, Index cols)                                                                  \
        { return lafem_hip_apply_banded_##D##_##I(nullptr, r, a, x, b, y, v, offs, noo, rows, cols); }            \
        static int csrb(DT* r, DT a, const DT* x, DT b, const DT* y, const DT* v, const IT* c, const IT* rp,      \
                        Index rows, Index cols, Index nnz, int bh, int bw)                                         \
        { return lafem_hip_apply_csrb_##D##_##I(nullptr, r, a, x, b, y, v, c, rp, rows, cols, nnz, bh, bw); }     \
        static int diag_csr(DT* d, const DT* v, const IT* c, const IT* rp, Index rows)                            \
        { return lafem_hip_diagonal_csr_##D##_##I(nullptr, d, v, c, rp, rows); }                                   \
        static int diag_ell(DT* d, const DT* v, const IT* c, const IT* cs, const IT* cl, Index C, Index rows)     \
        { return lafem_hip_diagonal_ell_##D##_##I(nullptr, d, v, c, cs, cl, C, rows); }                           \
        static int diag_csrb(DT* d, const DT* v, const IT* c, const IT* rp, Index rows, int bh, int bw)           \
        { return lafem_hip_diagonal_csrb_##D##_##I(nullptr, d, v, c, rp, rows, bh, bw); }                          \
        static int filter_rhs(DT* v, const DT* sv, const IT* idx, Index ue, int bs)                               \
        { return lafem_hip_unit_filter_rhs_##D##_##I(nullptr, v, sv, idx, ue, bs); }                               \
        static int filter_def(DT* v, const IT* idx, Index ue, int bs)                                              \
        { return lafem_hip_unit_filter_def_##D##_##I(nullptr, v, idx, ue, bs); }                                   \
      };
      F3_OPS(double, std::uint32_t, f64, u32)
      F3_OPS(double, std::uint64_t, f64, u64)
      F3_OPS(float, std::uint32_t, f32, u32)
      F3_OPS(float, std::uint64_t, f32, u64)
#undef F3_OPS

      template<typename DT_> struct Blas;
#define F3_BLAS(DT, D)                                                                                             \
      template<> struct Blas<DT>                                                                                   \
      {                                                                                                            \
        static int axpy(DT* r, DT a, const DT* x, const DT* y, Index n) { return lafem_hip_axpy_##D(nullptr, r, a, x, y, n); } \
        static int dot(const DT* x, const DT* y, Index n, DT* h) { return lafem_hip_dot_value_##D(nullptr, x, y, n, h); } \
        static int norm2(const DT* x, Index n, DT* h) { return lafem_hip_norm2_value_##D(nullptr, x, n, h); }       \
        static int cprod(DT* r, const DT* x, const DT* y, Index n) { return lafem_hip_component_product_##D(nullptr, r, x, y, n); } \
        static int cinv(DT* r, const DT* x, DT s, Index n) { return lafem_hip_component_invert_##D(nullptr, r, x, s, n); } \
        static int scale(DT* r, const DT* x, DT s, Index n) { return lafem_hip_scale_##D(nullptr, r, x, s, n); }    \
      };
      F3_BLAS(double, f64)
      F3_BLAS(float, f32)
#undef F3_BLAS
    } // namespace Arch

    // a refcounted device array (Container::_elements, container.hpp:86-956)
    template<typename T_>
    class DeviceArray
    {
      T_* _p = nullptr;
      Index _n = 0;
    public:
      DeviceArray() = default;
      explicit DeviceArray(Index n) : _p(n ? MemoryPool<Mem::HIP>::allocate_memory<T_>(n) : nullptr), _n(n) {}
      DeviceArray(const DeviceArray& o) : _p(o._p), _n(o._n) { MemoryPool<Mem::HIP>::increase_memory(_p); }
      DeviceArray(DeviceArray&& o) noexcept : _p(o._p), _n(o._n) { o._p = nullptr; o._n = 0; }
      DeviceArray& operator=(DeviceArray o) { std::swap(_p, o._p); std::swap(_n, o._n); return *this; }
      ~DeviceArray() { if(_p) feat3_hip_free(_p); }
      T_* data() { return _p; }
      const T_* data() const { return _p; }
      Index size() const { return _n; }
      void upload(const T_* h) { if(_n) MemoryPool<Mem::HIP>::upload(_p, h, _n); }
      std::vector<T_> download() const
      {
        std::vector<T_> h(_n);
        if(_n) MemoryPool<Mem::HIP>::download(h.data(), _p, _n);
        return h;
      }
    };

    enum class CloneMode { Shallow, Layout, Weak, Deep, Allocate };

    template<typename Mem_, typename DT_, typename IT_ = Index> class DenseVector;

    // DenseVector<Mem::HIP, DT_, IT_> (dense_vector.hpp:57-1181)
    template<typename DT_, typename IT_>
    class DenseVector<Mem::HIP, DT_, IT_>
    {
      DeviceArray<DT_> _a;
      // the FileMode value the library's vector files take ("Filemode not supported!" else)
      static int file_mode(FileMode mode)
      {
        if(mode == FileMode::fm_exp) return 0;
        if(mode == FileMode::fm_dv || mode == FileMode::fm_binary) return 1;
        if(mode == FileMode::fm_mtx) return 2;
        throw InternalError("DenseVector: Filemode not supported!");
      }
    public:
      typedef Mem::HIP MemType;
      typedef DT_ DataType;
      typedef IT_ IndexType;

      DenseVector() = default;
      explicit DenseVector(Index size) : _a(size) {}
      DenseVector(Index size, DT_ value) : _a(size) { format(value); }
      explicit DenseVector(const std::vector<DT_>& host) : _a(host.size()) { _a.upload(host.data()); }

      // from a file in the reference's formats (DenseVector(FileMode, String), read_from,
      // dense_vector.hpp:587-718): fm_exp, fm_dv (or fm_binary), fm_mtx
      DenseVector(FileMode mode, const std::string& filename) { read_from(mode, filename); }

      Index size() const { return _a.size(); }
      DT_* elements() { return _a.data(); }
      const DT_* elements() const { return _a.data(); }
      std::vector<DT_> to_host() const { return _a.download(); }

      void read_from(FileMode mode, const std::string& filename)
      {
        const int m = file_mode(mode);
        std::uint64_t n = 0;
        hip_check(feat3_io_vector_size(filename.c_str(), m, &n), "DenseVector::read_from");
        std::vector<double> v(n);
        hip_check(feat3_io_read_vector(filename.c_str(), m, v.data(), n), "DenseVector::read_from");
        *this = DenseVector(std::vector<DT_>(v.begin(), v.end()));
      }
      // write_out(FileMode, String) (dense_vector.hpp:721-780); text modes keep 6 significant digits
      void write_out(FileMode mode, const std::string& filename) const
      {
        const int m = file_mode(mode);
        const std::vector<DT_> h = to_host();
        const std::vector<double> v(h.begin(), h.end());
        hip_check(feat3_io_write_vector(filename.c_str(), m, v.size(), v.data()), "DenseVector::write_out");
      }

      DenseVector clone(CloneMode mode = CloneMode::Deep) const
      {
        if(mode == CloneMode::Shallow || mode == CloneMode::Weak) { DenseVector v; v._a = _a; return v; }
        DenseVector v(size());
        if(mode == CloneMode::Deep) v.copy(*this);
        return v;
      }
      // container.hpp:632-636
      void format(DT_ value = DT_(0)) { if(size()) MemoryPool<Mem::HIP>::set_memory(elements(), value, size()); }
      void copy(const DenseVector& x)
      {
        if(x.size() != size()) throw InternalError("DenseVector::copy: size mismatch");
        MemoryPool<Mem::HIP>::copy(elements(), x.elements(), size());
      }
      // dense_vector.hpp:424-427 (convert between value types on the device)
      template<typename DT2_>
      void convert(const DenseVector<Mem::HIP, DT2_, IT_>& x)
      {
        if(size() != x.size()) _a = DeviceArray<DT_>(x.size());
        if(size()) MemoryPool<Mem::HIP>::convert(elements(), x.elements(), size());
      }
      // dense_vector.hpp:883-896: |alpha| < eps -> copy(y)
      void axpy(const DenseVector& x, const DenseVector& y, DT_ alpha = DT_(1))
      {
        if(x.size() != size() || y.size() != size()) throw InternalError("DenseVector::axpy: size mismatch");
        if(Math::abs(alpha) < Math::eps<DT_>()) { copy(y); return; }
        hip_check(Arch::Blas<DT_>::axpy(elements(), alpha, x.elements(), y.elements(), size()), "DenseVector::axpy");
      }
      void component_product(const DenseVector& x, const DenseVector& y)
      {
        hip_check(Arch::Blas<DT_>::cprod(elements(), x.elements(), y.elements(), size()), "DenseVector::component_product");
      }
      void component_invert(const DenseVector& x, DT_ alpha = DT_(1))
      {
        hip_check(Arch::Blas<DT_>::cinv(elements(), x.elements(), alpha, size()), "DenseVector::component_invert");
      }
      void scale(const DenseVector& x, DT_ alpha)
      {
        hip_check(Arch::Blas<DT_>::scale(elements(), x.elements(), alpha, size()), "DenseVector::scale");
      }
      DT_ dot(const DenseVector& x) const
      {
        DT_ r(0);
        hip_check(Arch::Blas<DT_>::dot(elements(), x.elements(), size(), &r), "DenseVector::dot");
        return r;
      }
      DT_ norm2() const
      {
        DT_ r(0);
        hip_check(Arch::Blas<DT_>::norm2(elements(), size(), &r), "DenseVector::norm2");
        return r;
      }
      DT_ norm2sqr() const { const DT_ n = norm2(); return n * n; }
    };

    // SparseMatrixCSR<Mem::HIP, DT_, IT_> (sparse_matrix_csr.hpp)
    template<typename Mem_, typename DT_, typename IT_ = Index> class SparseMatrixCSR;

    template<typename DT_, typename IT_>
    class SparseMatrixCSR<Mem::HIP, DT_, IT_>
    {
      Index _rows = 0, _cols = 0, _nnz = 0;
      DeviceArray<IT_> _rp, _col;
      DeviceArray<DT_> _val;
    public:
      typedef Mem::HIP MemType;
      typedef DT_ DataType;
      typedef IT_ IndexType;
      typedef DenseVector<Mem::HIP, DT_, IT_> VectorTypeL;
      typedef DenseVector<Mem::HIP, DT_, IT_> VectorTypeR;

      SparseMatrixCSR() = default;
      // from host arrays (the reference's convert() from a Mem::Main matrix, container.hpp:185-310)
      template<typename DTH_, typename ITH_>
      SparseMatrixCSR(Index rows, Index cols, const std::vector<ITH_>& row_ptr, const std::vector<ITH_>& col_ind,
                      const std::vector<DTH_>& val)
        : _rows(rows), _cols(cols), _nnz(Index(row_ptr.at(rows))), _rp(rows + 1), _col(_nnz), _val(_nnz)
      {
        std::vector<IT_> rp(row_ptr.begin(), row_ptr.end()), ci(col_ind.begin(), col_ind.end());
        std::vector<DT_> v(val.begin(), val.end());
        _rp.upload(rp.data());
        _col.upload(ci.data());
        _val.upload(v.data());
      }
      // from a file in the reference's formats (SparseMatrixCSR(FileMode, String),
      // sparse_matrix_csr.hpp:1125-1240): fm_mtx (Matrix Market) or fm_csr (binary container)
      SparseMatrixCSR(FileMode mode, const std::string& filename)
      {
        if(mode != FileMode::fm_mtx && mode != FileMode::fm_csr)
          throw InternalError("SparseMatrixCSR(file): only fm_mtx and fm_csr are supported");
        std::uint64_t r = 0, c = 0, z = 0;
        const bool mtx = mode == FileMode::fm_mtx;
        hip_check(mtx ? feat3_io_mtx_sizes(filename.c_str(), &r, &c, &z)
                      : feat3_io_csr_binary_sizes(filename.c_str(), &r, &c, &z), "SparseMatrixCSR(file)");
        std::vector<std::uint32_t> rp(r + 1), ci(z);
        std::vector<double> v(z);
        hip_check(mtx ? feat3_io_read_mtx(filename.c_str(), r, z, rp.data(), ci.data(), v.data())
                      : feat3_io_read_csr_binary(filename.c_str(), r, z, rp.data(), ci.data(), v.data()),
                  "SparseMatrixCSR(file)");
        *this = SparseMatrixCSR(Index(r), Index(c), rp, ci, v);
      }
      // write_out(FileMode, String) (sparse_matrix_csr.hpp:1266-1340)
      void write_out(FileMode mode, const std::string& filename, bool symmetric = false) const
      {
        if(mode != FileMode::fm_mtx && mode != FileMode::fm_csr)
          throw InternalError("SparseMatrixCSR::write_out: only fm_mtx and fm_csr are supported");
        const std::vector<IT_> rp = _rp.download(), ci = _col.download();
        const std::vector<DT_> v = _val.download();
        std::vector<std::uint32_t> rp32(rp.begin(), rp.end()), ci32(ci.begin(), ci.end());
        std::vector<double> v64(v.begin(), v.end());
        hip_check(mode == FileMode::fm_mtx
                    ? feat3_io_write_mtx(filename.c_str(), _rows, _cols, rp32.data(), ci32.data(), v64.data(), symmetric ? 1 : 0)
                    : feat3_io_write_csr_binary(filename.c_str(), _rows, _cols, rp32.data(), ci32.data(), v64.data()),
                  "SparseMatrixCSR::write_out");
      }
      Index rows() const { return _rows; }
      Index columns() const { return _cols; }
      Index used_elements() const { return _nnz; }
      const IT_* row_ptr() const { return _rp.data(); }
      const IT_* col_ind() const { return _col.data(); }
      const DT_* val() const { return _val.data(); }

      VectorTypeL create_vector_l() const { return VectorTypeL(_rows); }
      VectorTypeR create_vector_r() const { return VectorTypeR(_cols); }

      // r = A x  (sparse_matrix_csr.hpp:1899-1928)
      void apply(VectorTypeL& r, const VectorTypeR& x, bool transposed = false) const
      {
        if(_nnz == 0) { r.format(); return; }
        if(r.elements() == x.elements()) throw InternalError("SparseMatrixCSR::apply: r and x must not share memory");
        hip_check(Arch::Ops<DT_, IT_>::csr(r.elements(), DT_(1), x.elements(), DT_(0), r.elements(), _val.data(), _col.data(),
                                           _rp.data(), _rows, _cols, _nnz, transposed), "SparseMatrixCSR::apply");
      }
      // r = y + alpha A x  (sparse_matrix_csr.hpp:1973-2010)
      void apply(VectorTypeL& r, const VectorTypeR& x, const VectorTypeL& y, DT_ alpha = DT_(1), bool transposed = false) const
      {
        if(_nnz == 0 || Math::abs(alpha) < Math::eps<DT_>()) { r.copy(y); return; }
        if(r.elements() == x.elements()) throw InternalError("SparseMatrixCSR::apply: r and x must not share memory");
        hip_check(Arch::Ops<DT_, IT_>::csr(r.elements(), alpha, x.elements(), DT_(1), y.elements(), _val.data(), _col.data(),
                                           _rp.data(), _rows, _cols, _nnz, transposed), "SparseMatrixCSR::apply");
      }
      // sparse_matrix_csr.hpp:2321-2327
      void extract_diag(VectorTypeL& diag) const
      {
        hip_check(Arch::Ops<DT_, IT_>::diag_csr(diag.elements(), _val.data(), _col.data(), _rp.data(), _rows),
                  "SparseMatrixCSR::extract_diag");
      }
      VectorTypeL extract_diag() const { VectorTypeL d(_rows); extract_diag(d); return d; }
    };

    // SparseMatrixELL<Mem::HIP, DT_, IT_>: the reference's SELL-C (sigma = 1) layout, built on
    // the host from CSR (sparse_matrix_ell.hpp:865-960), C = 32 by default (:387)
    template<typename Mem_, typename DT_, typename IT_ = Index> class SparseMatrixELL;

    template<typename DT_, typename IT_>
    class SparseMatrixELL<Mem::HIP, DT_, IT_>
    {
      Index _rows = 0, _cols = 0, _C = 32;
      DeviceArray<IT_> _cs, _cl, _col;
      DeviceArray<DT_> _val;
    public:
      typedef DenseVector<Mem::HIP, DT_, IT_> VectorTypeL;
      typedef DenseVector<Mem::HIP, DT_, IT_> VectorTypeR;

      template<typename DTH_, typename ITH_>
      SparseMatrixELL(Index rows, Index cols, const std::vector<ITH_>& rp, const std::vector<ITH_>& ci,
                      const std::vector<DTH_>& v, Index C = 32)
        : _rows(rows), _cols(cols), _C(C)
      {
        const Index nch = (rows + C - 1) / C;
        std::vector<IT_> cl(nch, 0), cs(nch + 1, 0);
        for(Index i = 0; i < rows; ++i)
          cl[i / C] = std::max<IT_>(cl[i / C], IT_(rp[i + 1] - rp[i]));
        for(Index c = 0; c < nch; ++c) cs[c + 1] = cs[c] + IT_(C) * cl[c];
        std::vector<IT_> col(cs[nch], 0);
        std::vector<DT_> val(cs[nch], DT_(0));
        for(Index i = 0; i < rows; ++i)
          for(Index j = rp[i], k = 0; j < Index(rp[i + 1]); ++j, ++k)
          {
            col[cs[i / C] + i % C + k * C] = IT_(ci[j]);
            val[cs[i / C] + i % C + k * C] = DT_(v[j]);
          }
        _cs = DeviceArray<IT_>(nch + 1); _cs.upload(cs.data());
        _cl = DeviceArray<IT_>(nch); _cl.upload(cl.data());
        _col = DeviceArray<IT_>(col.size()); _col.upload(col.data());
        _val = DeviceArray<DT_>(val.size()); _val.upload(val.data());
      }
      Index rows() const { return _rows; }
      Index columns() const { return _cols; }
      VectorTypeL create_vector_l() const { return VectorTypeL(_rows); }
      VectorTypeR create_vector_r() const { return VectorTypeR(_cols); }
      // sparse_matrix_ell.hpp:1983-2041
      void apply(VectorTypeL& r, const VectorTypeR& x) const
      {
        hip_check(Arch::Ops<DT_, IT_>::ell(r.elements(), DT_(1), x.elements(), DT_(0), r.elements(), _val.data(), _col.data(),
                                           _cs.data(), _cl.data(), _C, _rows), "SparseMatrixELL::apply");
      }
      void apply(VectorTypeL& r, const VectorTypeR& x, const VectorTypeL& y, DT_ alpha = DT_(1)) const
      {
        if(Math::abs(alpha) < Math::eps<DT_>()) { r.copy(y); return; }
        hip_check(Arch::Ops<DT_, IT_>::ell(r.elements(), alpha, x.elements(), DT_(1), y.elements(), _val.data(), _col.data(),
                                           _cs.data(), _cl.data(), _C, _rows), "SparseMatrixELL::apply");
      }
      void extract_diag(VectorTypeL& d) const
      {
        hip_check(Arch::Ops<DT_, IT_>::diag_ell(d.elements(), _val.data(), _col.data(), _cs.data(), _cl.data(), _C, _rows),
                  "SparseMatrixELL::extract_diag");
      }
    };

    // SparseMatrixBanded<Mem::HIP, DT_, IT_> (sparse_matrix_banded.hpp): sorted diagonal offsets
    // stored as offset + rows - 1, values val[d * rows + row]
    template<typename Mem_, typename DT_, typename IT_ = Index> class SparseMatrixBanded;

    template<typename DT_, typename IT_>
    class SparseMatrixBanded<Mem::HIP, DT_, IT_>
    {
      Index _rows = 0, _cols = 0, _noo = 0;
      DeviceArray<IT_> _offsets;
      DeviceArray<DT_> _val;
    public:
      typedef DenseVector<Mem::HIP, DT_, IT_> VectorTypeL;
      typedef DenseVector<Mem::HIP, DT_, IT_> VectorTypeR;

      // from CSR arrays (throws if a row has more distinct diagonals than max_diag in total)
      template<typename DTH_, typename ITH_>
      SparseMatrixBanded(Index rows, Index cols, const std::vector<ITH_>& rp, const std::vector<ITH_>& ci,
                         const std::vector<DTH_>& v, Index max_diag = 64)
        : _rows(rows), _cols(cols)
      {
        std::vector<long long> offs;
        for(Index i = 0; i < rows; ++i)
          for(Index j = rp[i]; j < Index(rp[i + 1]); ++j)
          {
            const long long o = (long long)(ci[j]) - (long long)(i);
            auto it = std::lower_bound(offs.begin(), offs.end(), o);
            if(it == offs.end() || *it != o)
            {
              if(offs.size() >= max_diag) throw InternalError("SparseMatrixBanded: more than max_diag diagonals");
              offs.insert(it, o);
            }
          }
        _noo = offs.size();
        std::vector<IT_> ho(_noo);
        for(Index d = 0; d < _noo; ++d) ho[d] = IT_(offs[d] + (long long)(rows) - 1);
        std::vector<DT_> val(_noo * rows, DT_(0));
        for(Index i = 0; i < rows; ++i)
          for(Index j = rp[i]; j < Index(rp[i + 1]); ++j)
          {
            const long long o = (long long)(ci[j]) - (long long)(i);
            const Index d = Index(std::lower_bound(offs.begin(), offs.end(), o) - offs.begin());
            val[d * rows + i] = DT_(v[j]);
          }
        _offsets = DeviceArray<IT_>(_noo); _offsets.upload(ho.data());
        _val = DeviceArray<DT_>(val.size()); _val.upload(val.data());
      }
      Index rows() const { return _rows; }
      Index columns() const { return _cols; }
      Index num_of_offsets() const { return _noo; }
      VectorTypeL create_vector_l() const { return VectorTypeL(_rows); }
      VectorTypeR create_vector_r() const { return VectorTypeR(_cols); }
      // sparse_matrix_banded.hpp apply -> Arch::Apply::banded (apply.hpp:162)
      void apply(VectorTypeL& r, const VectorTypeR& x) const
      {
        hip_check(Arch::Ops<DT_, IT_>::banded(r.elements(), DT_(1), x.elements(), DT_(0), r.elements(), _val.data(),
                                              _offsets.data(), _noo, _rows, _cols), "SparseMatrixBanded::apply");
      }
      void apply(VectorTypeL& r, const VectorTypeR& x, const VectorTypeL& y, DT_ alpha = DT_(1)) const
      {
        if(Math::abs(alpha) < Math::eps<DT_>()) { r.copy(y); return; }
        hip_check(Arch::Ops<DT_, IT_>::banded(r.elements(), alpha, x.elements(), DT_(1), y.elements(), _val.data(),
                                              _offsets.data(), _noo, _rows, _cols), "SparseMatrixBanded::apply");
      }
    };

    // SparseMatrixBCSR<Mem::HIP, DT_, IT_, BH_, BW_> with row-major blocks (Tiny::Matrix)
    template<typename Mem_, typename DT_, typename IT_, int BH_, int BW_> class SparseMatrixBCSR;

    template<typename DT_, typename IT_, int BH_, int BW_>
    class SparseMatrixBCSR<Mem::HIP, DT_, IT_, BH_, BW_>
    {
      Index _rows = 0, _cols = 0, _nnz = 0;
      DeviceArray<IT_> _rp, _col;
      DeviceArray<DT_> _val;
    public:
      typedef DenseVector<Mem::HIP, DT_, IT_> VectorTypeL;   // blocked vectors stored flat
      typedef DenseVector<Mem::HIP, DT_, IT_> VectorTypeR;
      template<typename DTH_, typename ITH_>
      SparseMatrixBCSR(Index rows, Index cols, const std::vector<ITH_>& rp, const std::vector<ITH_>& ci,
                       const std::vector<DTH_>& v)
        : _rows(rows), _cols(cols), _nnz(Index(rp.at(rows))), _rp(rows + 1), _col(_nnz), _val(_nnz * BH_ * BW_)
      {
        std::vector<IT_> r(rp.begin(), rp.end()), c(ci.begin(), ci.end());
        std::vector<DT_> w(v.begin(), v.end());
        _rp.upload(r.data()); _col.upload(c.data()); _val.upload(w.data());
      }
      VectorTypeL create_vector_l() const { return VectorTypeL(_rows * BH_); }
      VectorTypeR create_vector_r() const { return VectorTypeR(_cols * BW_); }
      void apply(VectorTypeL& r, const VectorTypeR& x) const
      {
        hip_check(Arch::Ops<DT_, IT_>::csrb(r.elements(), DT_(1), x.elements(), DT_(0), r.elements(), _val.data(), _col.data(),
                                            _rp.data(), _rows, _cols, _nnz, BH_, BW_), "SparseMatrixBCSR::apply");
      }
      void apply(VectorTypeL& r, const VectorTypeR& x, const VectorTypeL& y, DT_ alpha = DT_(1)) const
      {
        if(_nnz == 0 || Math::abs(alpha) < Math::eps<DT_>()) { r.copy(y); return; }
        hip_check(Arch::Ops<DT_, IT_>::csrb(r.elements(), alpha, x.elements(), DT_(1), y.elements(), _val.data(), _col.data(),
                                            _rp.data(), _rows, _cols, _nnz, BH_, BW_), "SparseMatrixBCSR::apply");
      }
      void extract_diag(VectorTypeL& d) const
      {
        hip_check(Arch::Ops<DT_, IT_>::diag_csrb(d.elements(), _val.data(), _col.data(), _rp.data(), _rows, BH_, BW_),
                  "SparseMatrixBCSR::extract_diag");
      }
    };

    // UnitFilter<Mem::HIP, DT_, IT_> (homogeneous or given Dirichlet values), block_size > 1 is
    // the UnitFilterBlocked of the reference (unit_filter_blocked.hpp:344-393)
    template<typename Mem_, typename DT_, typename IT_ = Index> class UnitFilter;

    template<typename DT_, typename IT_>
    class UnitFilter<Mem::HIP, DT_, IT_>
    {
      DeviceArray<IT_> _idx;
      DeviceArray<DT_> _sv;
      int _bs = 1;
    public:
      UnitFilter() = default;
      template<typename ITH_>
      UnitFilter(const std::vector<ITH_>& idx, const std::vector<DT_>& values = {}, int block_size = 1)
        : _idx(idx.size()), _sv(idx.size() * Index(block_size)), _bs(block_size)
      {
        std::vector<IT_> i(idx.begin(), idx.end());
        std::vector<DT_> v(values);
        v.resize(idx.size() * Index(block_size), DT_(0));
        _idx.upload(i.data());
        _sv.upload(v.data());
      }
      Index used_elements() const { return _idx.size(); }
      template<typename V_> void filter_rhs(V_& v) const
      {
        hip_check(Arch::Ops<DT_, IT_>::filter_rhs(v.elements(), _sv.data(), _idx.data(), _idx.size(), _bs), "UnitFilter::filter_rhs");
      }
      template<typename V_> void filter_sol(V_& v) const { filter_rhs(v); }
      template<typename V_> void filter_def(V_& v) const
      {
        hip_check(Arch::Ops<DT_, IT_>::filter_def(v.elements(), _idx.data(), _idx.size(), _bs), "UnitFilter::filter_def");
      }
      template<typename V_> void filter_cor(V_& v) const { filter_def(v); }
    };

    // Transfer (transfer.hpp:249-285): prol = P x_c, rest = R x_f with R = P^T stored
    template<typename Matrix_>
    class Transfer
    {
      Matrix_ _prol, _rest;
    public:
      typedef typename Matrix_::VectorTypeL VectorType;
      Transfer(Matrix_ prol, Matrix_ rest) : _prol(std::move(prol)), _rest(std::move(rest)) {}
      bool is_ghost() const { return false; }
      bool prol(VectorType& vec_fine, const VectorType& vec_coarse) const { _prol.apply(vec_fine, vec_coarse); return true; }
      bool rest(const VectorType& vec_fine, VectorType& vec_coarse) const { _rest.apply(vec_coarse, vec_fine); return true; }
    };
  } // namespace LAFEM
} // namespace feat3_amd
