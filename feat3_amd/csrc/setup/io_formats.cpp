// io_formats.cpp -- the reference's on-disk formats of SparseMatrixCSR and DenseVector, so that
// matrices and vectors written by a FEAT build (its solvers' inputs, matrix_info, fixtures) feed
// the HIP path and back:
//
//  * binary fm_csr (sparse_matrix_csr.hpp:1266-1296 -> Container::_serialise<double, uint64_t>,
//    container.hpp:320-447): 8-byte words [total bytes, FileMode fm_csr = 4, feature hash of the
//    value type, feature hash of the index type, #elements = 1, #indices = 2, #elements_size = 1,
//    #indices_size = 2, #scalar_index = 4, #scalar_dt = 1], elements_size {nnz}, indices_size
//    {nnz, rows + 1}, scalar_index {rows * cols, rows, cols, nnz}, scalar_dt {0.0}, then val (FP64),
//    col_ind (u64), row_ptr (u64) and 16 bytes of padding;
//  * Matrix Market coordinate real general / symmetric (sparse_matrix_csr.hpp:1125-1240 reader,
//    1296-1340 writer): 1-based "row col value" lines after the "rows cols nnz" line, '%' comments;
//    a symmetric file lists the lower triangle and is expanded.  Entries are sorted by row and
//    column; of repeated (row, col) pairs the first is kept (the reference inserts into a map).
//
// and of DenseVector (dense_vector.hpp:587-780), mode = the FileMode value:
//
//  * fm_exp (0): one value per line in C++'s std::scientific (6 digits); the reader skips every
//    line holding a '#', reads the first number of every other line (an empty line counts as 0)
//    and, as the reference's getline loop does, drops a last line that has no newline;
//  * fm_dv (1): Container::_serialise<double, uint64_t>(fm_dv): 8-byte words [total bytes, 1,
//    feature hash of the value type, feature hash of the index type, #elements = 1, #indices = 0,
//    #elements_size = 1, #indices_size = 0, #scalar_index = 1, #scalar_dt = 0], elements_size {n},
//    scalar_index {n}, the n values (FP64) and 16 bytes of padding;
//  * fm_mtx (2): "%%MatrixMarket matrix array real general", '%' comment lines, "n 1", then one
//    value per line (std::scientific); a file whose value count differs from n is rejected (the
//    reference would read past its buffer).
//
// Indices are u32 in this library: files with more than 2^32 - 1 rows, columns or entries are
// rejected.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

typedef std::uint32_t u32;
typedef std::uint64_t u64;

namespace
{
  constexpr u64 kFmCsr = 4;                                            // FileMode::fm_csr (lafem/base.hpp:31-46)
  constexpr u64 kHashDouble = 8ull | (1ull << 33) | (1ull << 34);       // Type::Traits<double>::feature_hash
  constexpr u64 kHashIndex = 8ull | (1ull << 32);                       // Type::Traits<unsigned long>::feature_hash

  struct Entries
  {
    u64 rows = 0, cols = 0;
    std::map<u64, std::map<u64, double>> rc;
  };

  bool read_mtx(const char* path, Entries& E)
  {
    std::ifstream f(path);
    if(!f.is_open()) return false;
    std::string line;
    if(!std::getline(f, line)) return false;
    const bool general = line.find("%%MatrixMarket matrix coordinate real general") != std::string::npos;
    const bool symmetric = line.find("%%MatrixMarket matrix coordinate real symmetric") != std::string::npos;
    if(!general && !symmetric) return false;
    bool have_size = false;
    while(std::getline(f, line))
    {
      const std::size_t b = line.find_first_not_of(" \t\r");
      if(b == std::string::npos || line[b] == '%') continue;
      std::istringstream is(line);
      if(!have_size)
      {
        u64 nz = 0;
        if(!(is >> E.rows >> E.cols >> nz)) return false;
        have_size = true;
        continue;
      }
      u64 r, c;
      double v;
      if(!(is >> r >> c >> v) || r < 1 || c < 1 || r > E.rows || c > E.cols) return false;
      E.rc[r - 1].insert(std::make_pair(c - 1, v));
      if(symmetric && r != c) E.rc[c - 1].insert(std::make_pair(r - 1, v));
    }
    return have_size;
  }

  bool read_binary(const char* path, std::vector<u64>& words)
  {
    std::ifstream f(path, std::ios::binary);
    if(!f.is_open()) return false;
    u64 total = 0;
    f.read(reinterpret_cast<char*>(&total), sizeof(u64));
    if(!f.good() || total < 18 * 8 || total % 8 != 0) return false;
    words.assign(total / 8, 0);
    f.seekg(0);
    f.read(reinterpret_cast<char*>(words.data()), std::streamsize(total));
    if(!f.good()) return false;
    // layout checks of a CSR container serialised as <double, uint64>
    if(words[1] != kFmCsr || words[4] != 1 || words[5] != 2 || words[6] != 1 || words[7] != 2 || words[8] != 4)
      return false;
    if((words[2] >> 33 & 1) == 0 || (words[2] & 0xFFFFFFFFull) == 0) return false;   // floating value type
    if((words[3] >> 32 & 1) == 0) return false;                                       // integral index type
    return true;
  }

  // the reference converts its running word count between the value and index arrays through
  // float arithmetic (container.hpp:422, 527), which rounds counts above 2^24; the same rounding
  // is applied here so that its files are read at the offsets it wrote them
  u64 feat_offset(u64 gi)
  {
    return u64(std::ceil((float(gi) * float(sizeof(double))) / float(sizeof(u64))));
  }

  // positions (in words) of the arrays of a fm_csr file
  struct BinLayout
  {
    u64 rows, cols, nnz, val, col, rp;
  };

  // Every size comes from the file: the counts are bounded before any sum is formed (no sum of
  // them can wrap), and every array must lie inside the words read, in the order val < col < rp.
  // The header words are those of SparseMatrixCSR's container (sparse_matrix_csr.hpp:320-343):
  // one elements array, two index arrays, three scalar indices (rows, columns, nnz) and one
  // scalar value (the zero element).
  bool layout(const std::vector<u64>& w, BinLayout& L)
  {
    if(w.size() < 18 || w[6] != 1 || w[7] != 2 || w[8] != 4 || w[9] != 1) return false;
    const u64 hdr = 10 + w[6] + w[7] + w[8];
    const u64 limit = u64(1) << 40;                 // far above any u32-indexed matrix
    L.nnz = w[10];
    L.rows = w[13 + 1];
    L.cols = w[13 + 2];
    if(L.nnz >= limit || L.rows >= limit || L.cols >= limit) return false;
    if(w[11] != L.nnz || w[12] != L.rows + 1 || w[13 + 3] != L.nnz) return false;
    L.val = feat_offset(hdr) + w[9];   // DT2 = double: one word each
    if(L.val + L.nnz > w.size()) return false;
    L.col = feat_offset(L.val + L.nnz);
    if(L.col < L.val + L.nnz || L.col + L.nnz > w.size()) return false;
    L.rp = L.col + L.nnz;
    return L.rp + L.rows + 1 <= w.size();
  }

  double as_double(u64 w)
  {
    double d;
    std::memcpy(&d, &w, sizeof(d));
    return d;
  }
  u64 as_word(double d)
  {
    u64 w;
    std::memcpy(&w, &d, sizeof(w));
    return w;
  }
  constexpr int kFmExp = 0, kFmDv = 1, kFmMtx = 2;   // FileMode values (lafem/base.hpp:31-46)

  // dense_vector.hpp:669-705 (fm_exp) and 609-668 (fm_mtx, the values after the size line)
  bool read_vector_text(const char* path, int mode, std::vector<double>& v)
  {
    std::ifstream f(path);
    if(!f.is_open()) return false;
    v.clear();
    std::string line;
    u64 rows = 0;
    if(mode == kFmMtx)
    {
      if(!std::getline(f, line) || line.find("%%MatrixMarket matrix array real general") == std::string::npos)
        return false;
      bool have = false;
      while(std::getline(f, line))
      {
        if(f.eof()) return false;   // the reference: "Input-file is empty"
        const std::size_t b = line.find_first_not_of(' ');
        if(b != std::string::npos && line[b] == '%') continue;
        std::istringstream is(line);
        u64 cols = 0;
        if(!(is >> rows >> cols) || cols != 1) return false;
        have = true;
        break;
      }
      if(!have) return false;
    }
    while(true)
    {
      std::getline(f, line);
      if(mode == kFmExp && line.find('#') != std::string::npos)
      {
        if(f.eof()) break;
        continue;
      }
      if(f.eof() || f.fail()) break;
      const std::size_t b = line.find_first_not_of(' ');
      const std::string t = b == std::string::npos ? std::string() : line.substr(b);
      v.push_back(std::atof(t.c_str()));
    }
    return mode != kFmMtx || v.size() == rows;
  }

  bool read_vector_dv(const char* path, std::vector<double>& v)
  {
    std::ifstream f(path, std::ios::binary);
    if(!f.is_open()) return false;
    u64 total = 0;
    f.read(reinterpret_cast<char*>(&total), sizeof(u64));
    if(!f.good() || total < 14 * 8 || total % 8 != 0) return false;
    std::vector<u64> w(total / 8, 0);
    f.seekg(0);
    f.read(reinterpret_cast<char*>(w.data()), std::streamsize(total));
    if(!f.good()) return false;
    if(w[1] != u64(kFmDv) || w[4] != 1 || w[5] != 0 || w[6] != 1 || w[7] != 0) return false;
    if((w[2] >> 33 & 1) == 0 || (w[2] & 0xFFFFFFFFull) == 0) return false;   // floating value type
    if((w[3] >> 32 & 1) == 0) return false;                                   // integral index type
    const u64 n = w[10];
    const u64 at = feat_offset(10 + w[6] + w[7] + w[8]) + w[9];
    if(at + n > w.size()) return false;
    v.resize(n);
    for(u64 i = 0; i < n; ++i) v[i] = as_double(w[at + i]);
    return true;
  }

  bool read_vector(const char* path, int mode, std::vector<double>& v)
  {
    if(!path) return false;
    if(mode == kFmDv) return read_vector_dv(path, v);
    if(mode == kFmExp || mode == kFmMtx) return read_vector_text(path, mode, v);
    return false;
  }
} // namespace

extern "C" {

// DenseVector files: number of values, then the values (host memory, n = the size returned)
int feat3_io_vector_size(const char* path, int mode, uint64_t* n)
{
  std::vector<double> v;
  if(!n || !read_vector(path, mode, v)) return 1;
  *n = v.size();
  return 0;
}

int feat3_io_read_vector(const char* path, int mode, double* out, uint64_t n)
{
  std::vector<double> v;
  if(!read_vector(path, mode, v) || v.size() != n || (n > 0 && !out)) return 1;
  std::copy(v.begin(), v.end(), out);
  return 0;
}

int feat3_io_write_vector(const char* path, int mode, uint64_t n, const double* v)
{
  if(!path || (n > 0 && !v)) return 1;
  if(mode == kFmDv)
  {
    // _serialised_size: 10 header words, elements_size and scalar_index, the values, 16 bytes
    const u64 at = feat_offset(12);
    const u64 words = std::max<u64>(12 + n + 2, at + n);
    std::vector<u64> w(words, 0);
    w[0] = words * 8;
    w[1] = u64(kFmDv);
    w[2] = kHashDouble;
    w[3] = kHashIndex;
    w[4] = 1; w[5] = 0; w[6] = 1; w[7] = 0; w[8] = 1; w[9] = 0;
    w[10] = n;
    w[11] = n;
    for(u64 i = 0; i < n; ++i) w[at + i] = as_word(v[i]);
    std::ofstream f(path, std::ios::binary);
    if(!f.is_open()) return 1;
    f.write(reinterpret_cast<const char*>(w.data()), std::streamsize(words * 8));
    return f.good() ? 0 : 1;
  }
  if(mode != kFmExp && mode != kFmMtx) return 1;
  FILE* f = std::fopen(path, "w");
  if(!f) return 1;
  if(mode == kFmMtx)
  {
    std::fprintf(f, "%%%%MatrixMarket matrix array real general\n");
    std::fprintf(f, "%llu 1\n", (unsigned long long)n);
  }
  for(u64 i = 0; i < n; ++i) std::fprintf(f, "%e\n", v[i]);   // std::scientific, precision 6
  return std::fclose(f) == 0 ? 0 : 1;
}


// Matrix Market: sizes (nnz after symmetric expansion and removal of repeated entries)
int feat3_io_mtx_sizes(const char* path, uint64_t* rows, uint64_t* cols, uint64_t* nnz)
{
  Entries E;
  if(!path || !rows || !cols || !nnz || !read_mtx(path, E)) return 1;
  u64 n = 0;
  for(const auto& r : E.rc) n += r.second.size();
  if(E.rows >= (u64(1) << 32) || E.cols >= (u64(1) << 32) || n >= (u64(1) << 32)) return 1;
  *rows = E.rows; *cols = E.cols; *nnz = n;
  return 0;
}

int feat3_io_read_mtx(const char* path, uint64_t rows, uint64_t nnz, uint32_t* row_ptr, uint32_t* col_ind, double* val)
{
  Entries E;
  if(!path || !row_ptr || !read_mtx(path, E)) return 1;
  // the caller's arrays hold rows + 1 row pointers and nnz entries: the file must (still) match
  u64 z = 0;
  for(const auto& r : E.rc) z += r.second.size();
  if(E.rows != rows || z != nnz || E.rows >= (u64(1) << 32) || E.cols >= (u64(1) << 32) || z >= (u64(1) << 32)) return 1;
  if(z && (!col_ind || !val)) return 1;
  u64 k = 0;
  row_ptr[0] = 0;
  for(u64 r = 0; r < E.rows; ++r)
  {
    auto it = E.rc.find(r);
    if(it != E.rc.end())
      for(const auto& cv : it->second)
      {
        col_ind[k] = u32(cv.first);
        val[k] = cv.second;
        ++k;
      }
    row_ptr[r + 1] = u32(k);
  }
  return 0;
}

// general: every entry; symmetric != 0: the lower triangle (row >= col) with the symmetric header
int feat3_io_write_mtx(const char* path, uint64_t rows, uint64_t cols, const uint32_t* row_ptr, const uint32_t* col_ind,
                       const double* val, int symmetric)
{
  if(!path || !row_ptr) return 1;
  FILE* f = std::fopen(path, "w");
  if(!f) return 1;
  u64 nz = 0;
  for(u64 r = 0; r < rows; ++r)
    for(u32 k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
      if(!symmetric || r >= col_ind[k]) ++nz;
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate real %s\n", symmetric ? "symmetric" : "general");
  std::fprintf(f, "%llu %llu %llu\n", (unsigned long long)rows, (unsigned long long)cols, (unsigned long long)nz);
  for(u64 r = 0; r < rows; ++r)
    for(u32 k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
      if(!symmetric || r >= col_ind[k])
        std::fprintf(f, "%llu %llu %.17e\n", (unsigned long long)(r + 1), (unsigned long long)(col_ind[k] + 1), val[k]);
  const bool ok = std::fclose(f) == 0;
  return ok ? 0 : 1;
}

// binary fm_csr (Container::_serialise<double, uint64_t>)
int feat3_io_csr_binary_sizes(const char* path, uint64_t* rows, uint64_t* cols, uint64_t* nnz)
{
  std::vector<u64> w;
  BinLayout L;
  if(!path || !rows || !cols || !nnz || !read_binary(path, w) || !layout(w, L)) return 1;
  if(L.rows >= (u64(1) << 32) || L.cols >= (u64(1) << 32) || L.nnz >= (u64(1) << 32)) return 1;
  *rows = L.rows; *cols = L.cols; *nnz = L.nnz;
  return 0;
}

int feat3_io_read_csr_binary(const char* path, uint64_t rows, uint64_t nnz, uint32_t* row_ptr, uint32_t* col_ind,
                             double* val)
{
  std::vector<u64> w;
  BinLayout L;
  if(!path || !row_ptr || !read_binary(path, w) || !layout(w, L)) return 1;
  // the file is read again here: it must (still) have the sizes the caller's arrays were made
  // for, within the limits of feat3_io_csr_binary_sizes
  if(L.rows != rows || L.nnz != nnz) return 1;
  if(L.rows >= (u64(1) << 32) || L.cols >= (u64(1) << 32) || L.nnz >= (u64(1) << 32)) return 1;
  if(L.nnz && (!col_ind || !val)) return 1;
  for(u64 i = 0; i < L.nnz; ++i)
  {
    const u64 c = w[L.col + i];
    if(c >= L.cols) return 1;
    col_ind[i] = u32(c);
    val[i] = as_double(w[L.val + i]);
  }
  for(u64 i = 0; i <= L.rows; ++i)
  {
    const u64 p = w[L.rp + i];
    if(p > L.nnz || (i > 0 && p < w[L.rp + i - 1])) return 1;
    row_ptr[i] = u32(p);
  }
  return row_ptr[0] == 0 && row_ptr[L.rows] == u32(L.nnz) ? 0 : 1;
}

int feat3_io_write_csr_binary(const char* path, uint64_t rows, uint64_t cols, const uint32_t* row_ptr,
                              const uint32_t* col_ind, const double* val)
{
  if(!path || !row_ptr || rows == 0 || cols == 0) return 1;
  const u64 nnz = row_ptr[rows];
  // _serialised_size: 10 header words, 1 + 2 + 4 size/scalar words, 1 scalar_dt, the arrays, 16 bytes
  const u64 gwords = 10 + 1 + 2 + 4 + 1 + nnz + nnz + (rows + 1) + 2;
  const u64 val_at = feat_offset(17) + 1, col_at = feat_offset(val_at + nnz), rp_at = col_at + nnz;
  const u64 words = std::max(gwords, rp_at + rows + 1);
  std::vector<u64> w(words, 0);
  w[0] = words * 8;
  w[1] = kFmCsr;
  w[2] = kHashDouble;
  w[3] = kHashIndex;
  w[4] = 1; w[5] = 2; w[6] = 1; w[7] = 2; w[8] = 4; w[9] = 1;
  w[10] = nnz;
  w[11] = nnz; w[12] = rows + 1;
  w[13] = rows * cols; w[14] = rows; w[15] = cols; w[16] = nnz;
  w[17] = as_word(0.0);
  for(u64 i = 0; i < nnz; ++i) { w[val_at + i] = as_word(val[i]); w[col_at + i] = col_ind[i]; }
  for(u64 i = 0; i <= rows; ++i) w[rp_at + i] = row_ptr[i];
  std::ofstream f(path, std::ios::binary);
  if(!f.is_open()) return 1;
  f.write(reinterpret_cast<const char*>(w.data()), std::streamsize(words * 8));
  return f.good() ? 0 : 1;
}

} // extern "C"
