// assembly.hip -- numeric assembly of the coefficient-weighted Laplace values on the device, for
// many coefficient fields at once (the step before the hot path; its host form is
// feat3_setup_laplace in setup/structured_fe.cpp, the reference's BilinearOperatorAssembler with
// the Laplace operator of assembly/common_operators.hpp:36-97 scaled per cell).
//
// Structured Q1/Q2 Lagrange space on the refined unit square/cube, lexicographic node and cell
// numbering.  Entry (i, j) of A(k) is h^(d-2) * sum over the cells c holding both nodes of
// k_c * K_ref[loc_c(i)][loc_c(j)].  The graph (row_ptr, col_ind) is the host setup's; one thread
// per (row, field) walks its row.  The cells are summed in the host's order (z, y, x ascending)
// with separate multiply and add (no contraction), so the values are bit-identical to the host
// assembly.  Fields are [field][cell]; values are written [entry][field], the layout of the
// batched solver's per-system values.

#include "device_common.hpp"
#include "runtime.hpp"
#include "feat3_hip.h"

#include <hip/hip_runtime.h>

using namespace feat3_amd;
typedef std::uint32_t u32;
typedef std::uint64_t u64;

namespace
{
  constexpr int kMaxLoc = 9;   // Q1 2D (4), Q1 3D (8), Q2 2D (9)

  struct RefMatrix
  {
    double k[kMaxLoc * kMaxLoc];
  };

  struct GridDev
  {
    int dim, deg;
    u64 cells, m;
    __device__ void cell_range(u64 x, u64& lo, u64& hi) const
    {
      const u64 p = u64(deg);
      lo = (x >= p) ? (x - 1) / p : 0;
      if(x == 0) lo = 0;
      hi = x / p < cells - 1 ? x / p : cells - 1;
      if(lo > hi) lo = hi;
    }
  };

#pragma clang fp contract(off)
  __global__ __launch_bounds__(256) void k_assemble(GridDev G, RefMatrix R, double hs, int nfields, u64 n,
                                                    const double* __restrict__ kf, const u32* __restrict__ rp,
                                                    const u32* __restrict__ col, double* __restrict__ val)
  {
    const int q = G.deg + 1;
    int nloc = 1;
    for(int d = 0; d < G.dim; ++d) nloc *= q;
    u64 ncell = 1;
    for(int d = 0; d < G.dim; ++d) ncell *= G.cells;
    const u64 total = n * u64(nfields);
    for(u64 t = u64(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += u64(gridDim.x) * blockDim.x)
    {
      const u64 r = t / u64(nfields);
      const u64 f = t % u64(nfields);
      const double* kc_f = kf + f * ncell;
      u64 c[3] = {0, 0, 0}, clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
      u64 tmp = r;
      for(int d = 0; d < G.dim; ++d) { c[d] = tmp % G.m; tmp /= G.m; G.cell_range(c[d], clo[d], chi[d]); }
      const u64 p = u64(G.deg);
      for(u32 e = rp[r]; e < rp[r + 1]; ++e)
      {
        u64 j[3] = {0, 0, 0};
        u64 jt = col[e];
        for(int d = 0; d < G.dim; ++d) { j[d] = jt % G.m; jt /= G.m; }
        u64 elo[3] = {0, 0, 0}, ehi[3] = {0, 0, 0};
        bool any = true;
        for(int d = 0; d < G.dim; ++d)
        {
          u64 lo2, hi2;
          G.cell_range(j[d], lo2, hi2);
          elo[d] = clo[d] > lo2 ? clo[d] : lo2;
          ehi[d] = chi[d] < hi2 ? chi[d] : hi2;
          if(elo[d] > ehi[d]) any = false;
        }
        double s = 0.0;
        if(any)
        {
          const u64 ez0 = G.dim == 3 ? elo[2] : 0, ez1 = G.dim == 3 ? ehi[2] : 0;
          u64 ec[3];
          for(ec[2] = ez0; ec[2] <= ez1; ++ec[2])
            for(ec[1] = elo[1]; ec[1] <= ehi[1]; ++ec[1])
              for(ec[0] = elo[0]; ec[0] <= ehi[0]; ++ec[0])
              {
                int la = 0, lb = 0, mul = 1;
                for(int d = 0; d < G.dim; ++d)
                {
                  la += int(c[d] - p * ec[d]) * mul;
                  lb += int(j[d] - p * ec[d]) * mul;
                  mul *= q;
                }
                u64 ci = 0;
                for(int d = G.dim - 1; d >= 0; --d) ci = ci * G.cells + ec[d];
                const double prod = kc_f[ci] * R.k[la * nloc + lb];
                s = s + prod;
              }
        }
        val[u64(e) * u64(nfields) + f] = s * hs;
      }
    }
  }

  // kc[f][cc] = mean of the 2^dim children (feat3_setup_coarsen_cellfield's order)
  __global__ __launch_bounds__(256) void k_coarsen(int dim, u64 cf, int nfields, const double* __restrict__ kf,
                                                   double* __restrict__ kc)
  {
    const u64 cc = cf / 2;
    u64 ncc = 1, ncf = 1;
    for(int d = 0; d < dim; ++d) { ncc *= cc; ncf *= cf; }
    const u64 total = ncc * u64(nfields);
    const int nch = 1 << dim;
    for(u64 t = u64(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += u64(gridDim.x) * blockDim.x)
    {
      const u64 f = t / ncc;
      const u64 ci = t % ncc;
      u64 e[3] = {0, 0, 0};
      u64 tt = ci;
      for(int d = 0; d < dim; ++d) { e[d] = tt % cc; tt /= cc; }
      double s = 0.0;
      for(int ch = 0; ch < nch; ++ch)
      {
        u64 fi = 0;
        for(int d = dim - 1; d >= 0; --d) fi = fi * cf + 2 * e[d] + u64((ch >> d) & 1);
        s = s + kf[f * ncf + fi];
      }
      kc[f * ncc + ci] = s / double(nch);
    }
  }
#pragma clang fp contract(on)

  inline unsigned blocks_for(u64 items)
  {
    const u64 b = (items + 255) / 256;
    return unsigned(b < 1 ? 1 : (b > 65536 ? 65536 : b));
  }
} // namespace

extern "C" {

int feat3_assemble_laplace_fields(void* stream, int dim, int degree, int level, int nfields, const double* ref_stiffness,
                                  const double* kfields, const uint32_t* row_ptr, const uint32_t* col_ind, uint64_t n,
                                  double* val)
{
  if(!(dim == 2 || dim == 3) || !(degree == 1 || degree == 2) || (dim == 3 && degree == 2)) return rt::kErrArg;
  if(level < 0 || level > 12 || nfields < 1 || !ref_stiffness || !kfields || !row_ptr || !col_ind || !val) return rt::kErrArg;
  GridDev G{dim, degree, u64(1) << level, u64(degree) * (u64(1) << level) + 1};
  u64 nn = 1;
  for(int d = 0; d < dim; ++d) nn *= G.m;
  if(n != nn) return rt::kErrArg;
  const int q = degree + 1;
  const int nloc = dim == 2 ? q * q : q * q * q;
  RefMatrix R{};
  for(int i = 0; i < nloc * nloc; ++i) R.k[i] = ref_stiffness[i];
  // h^(d-2) as the host computes it
  const double h = 1.0 / double(G.cells);
  double hs = 1.0;
  for(int d = 2; d < dim; ++d) hs *= h;
  k_assemble<<<blocks_for(n * u64(nfields)), 256, 0, static_cast<hipStream_t>(stream)>>>(G, R, hs, nfields, n, kfields,
                                                                                          row_ptr, col_ind, val);
  return rt::check_launch();
}

int feat3_coarsen_cellfields(void* stream, int dim, int level_fine, int nfields, const double* kf, double* kc)
{
  if(!(dim == 2 || dim == 3) || level_fine < 1 || level_fine > 12 || nfields < 1 || !kf || !kc) return rt::kErrArg;
  const u64 cf = u64(1) << level_fine;
  u64 ncc = 1;
  for(int d = 0; d < dim; ++d) ncc *= cf / 2;
  k_coarsen<<<blocks_for(ncc * u64(nfields)), 256, 0, static_cast<hipStream_t>(stream)>>>(dim, cf, nfields, kf, kc);
  return rt::check_launch();
}

} // extern "C"
