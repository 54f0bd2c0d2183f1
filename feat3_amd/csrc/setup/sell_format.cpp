// sell_format.cpp -- host-side CSR -> SELL-C-sigma conversion (the derived kernel format of
// the device SpMV; the CSR container of record stays unchanged).
//
// Generalises SparseMatrixELL::convert (lafem/sparse_matrix_ell.hpp:865-960, SELL-C with
// sigma = 1 and col-0 padding) by (a) sorting rows by descending length inside windows of
// sigma rows (a permutation is returned for sigma > 1), (b) padding each row of a chunk to the
// chunk width with a column that the row already holds (value 0), so that every slot the
// kernel reads addresses x inside [0, cols) -- also for rectangular transfer matrices -- and
// (c) optionally storing columns as 16-bit signed offsets from the row index when all fit.

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

#include "feat3_hip.h"

namespace
{
  typedef std::uint32_t u32;
  typedef std::uint64_t u64;

  std::vector<u32> sorted_positions(u64 rows, const u32* rp, u64 sigma)
  {
    std::vector<u32> pm(rows);
    std::iota(pm.begin(), pm.end(), 0u);
    if(sigma > 1)
      for(u64 w = 0; w < rows; w += sigma)
      {
        const u64 e = std::min(rows, w + sigma);
        std::stable_sort(pm.begin() + w, pm.begin() + e,
                         [&](u32 a, u32 b) { return rp[a + 1] - rp[a] > rp[b + 1] - rp[b]; });
      }
    return pm;
  }
}

extern "C" {

// number of chunks and stored slots (values incl. padding) of the SELL-C-sigma form
int feat3_sell_sizes(u64 rows, const u32* row_ptr, u64 C, u64 sigma, u64* nchunks, u64* val_size)
{
  if(!row_ptr || !nchunks || !val_size || C == 0 || sigma == 0) return 1;
  const std::vector<u32> pm = sorted_positions(rows, row_ptr, sigma);
  const u64 nch = (rows + C - 1) / C;
  u64 total = 0;
  for(u64 c = 0; c < nch; ++c)
  {
    u32 w = 0;
    for(u64 k = c * C; k < std::min(rows, (c + 1) * C); ++k) w = std::max(w, row_ptr[pm[k] + 1] - row_ptr[pm[k]]);
    total += C * w;
  }
  *nchunks = nch;
  *val_size = total;
  return 0;
}

// Fills cs[nchunks+1], cl[nchunks], perm[rows] (position -> row; identity for sigma = 1),
// col[val_size] (absolute columns) and val[val_size]; if d16 != NULL and every offset
// |col - row| fits in 16 bits, also d16[val_size] and *used_d16 = 1.
// Returns 1 (argument error) if a column index is >= cols.
int feat3_sell_build(u64 rows, u64 cols, const u32* row_ptr, const u32* col_ind, const double* val, u64 C, u64 sigma,
                     u32* cs, u32* cl, u32* perm, u32* col, double* sval, std::int16_t* d16, int* used_d16)
{
  if(!row_ptr || (row_ptr[rows] > 0 && (!col_ind || !val)) || !cs || !cl || !perm || !col || !sval || C == 0 || sigma == 0)
    return 1;
  for(u64 k = 0; k < row_ptr[rows]; ++k)
    if(col_ind[k] >= cols) return 1;
  const std::vector<u32> pm = sorted_positions(rows, row_ptr, sigma);
  auto len = [&](u32 r) { return row_ptr[r + 1] - row_ptr[r]; };
  auto pad_col = [&](u32 r) -> u32 {
    return len(r) > 0 ? col_ind[row_ptr[r]] : u32(std::min<u64>(r, cols > 0 ? cols - 1 : 0));
  };
  const u64 nch = (rows + C - 1) / C;
  cs[0] = 0;
  for(u64 c = 0; c < nch; ++c)
  {
    u32 w = 0;
    for(u64 k = c * C; k < std::min(rows, (c + 1) * C); ++k) w = std::max(w, len(pm[k]));
    cl[c] = w;
    cs[c + 1] = cs[c] + u32(C) * w;
  }
  const u64 vs = cs[nch];
  std::fill(col, col + vs, 0u);
  std::fill(sval, sval + vs, 0.0);
  for(u64 p = 0; p < rows; ++p)
  {
    const u32 r = pm[p];
    perm[p] = r;
    const u64 base = cs[p / C] + p % C;
    const u32 pc = pad_col(r);
    for(u64 k = 0; k < cl[p / C]; ++k) col[base + k * C] = pc;
    u64 k = 0;
    for(u32 j = row_ptr[r]; j < row_ptr[r + 1]; ++j, ++k)
    {
      col[base + k * C] = col_ind[j];
      sval[base + k * C] = val[j];
    }
  }
  // slots of the missing rows of an incomplete last chunk are never read; give them column 0
  bool fits = d16 != nullptr && cols > 0;
  for(u64 p = 0; p < rows && fits; ++p)
  {
    const u64 base = cs[p / C] + p % C;
    for(u64 k = 0; k < cl[p / C]; ++k)
    {
      const std::int64_t d = std::int64_t(col[base + k * C]) - std::int64_t(pm[p]);
      if(d < -32768 || d > 32767) { fits = false; break; }
    }
  }
  if(used_d16) *used_d16 = fits ? 1 : 0;
  if(fits)
  {
    std::fill(d16, d16 + vs, std::int16_t(0));
    for(u64 p = 0; p < rows; ++p)
    {
      const u64 base = cs[p / C] + p % C;
      for(u64 k = 0; k < cl[p / C]; ++k)
        d16[base + k * C] = std::int16_t(std::int64_t(col[base + k * C]) - std::int64_t(pm[p]));
    }
  }
  return 0;
}

} // extern "C"
