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
    __host__ __device__ void cell_range(u64 x, u64& lo, u64& hi) const
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

#pragma clang fp contract(off)
  // ---- symbolic assembly: graphs, transfers and the Dirichlet list ------------------------
  //
  // All index arrays follow from closed forms per row, without scans or atomics:
  //  * a row pointer of a tensor-product pattern (entries of a row = product over the
  //    directions of a 1D count of its coordinate) is a separable prefix sum: with x fastest,
  //    rp[r] = sum_d S(c_d) * prod_{e<d} T * prod_{e>d} cnt(c_e), S the 1D prefix of cnt and T
  //    its total; S is summed per thread (at most a few thousand terms);
  //  * the position of a boundary node in the ascending Dirichlet list is r minus the number of
  //    interior nodes before it, again separable (interior coordinates are 1 .. m-2);
  //  * R = P^T: for a coarse coordinate j the fine coordinates whose prolongation row holds j
  //    form one interval per direction, so R's rows are tensor products too.
  // Columns come in ascending order (z, y, x loops) and values use the host setup's operations
  // in the host's order (no contraction), so every array equals the host's bit for bit.

  // 1D Lagrange basis of degree p on [0,1] (the host setup's lag)
  __host__ __device__ inline double lag1(int p, int a, double t)
  {
    if(p == 1) return a == 0 ? 1.0 - t : t;
    if(a == 0) return 2.0 * (t - 0.5) * (t - 1.0);
    if(a == 1) return 4.0 * t * (1.0 - t);
    return 2.0 * t * (t - 0.5);
  }

  // the prolongation rows of fine coordinate x (1D): coarse cells [lo, hi] holding x
  struct Prol1
  {
    u64 p, ccells;
    __host__ __device__ void cells(u64 x, u64& lo, u64& hi) const
    {
      lo = x >= 2 * p ? (x - 1) / (2 * p) : 0;
      hi = x / (2 * p) < ccells - 1 ? x / (2 * p) : ccells - 1;
      if(lo > hi) lo = hi;
    }
    __host__ __device__ u64 count(u64 x) const
    {
      u64 lo, hi;
      cells(x, lo, hi);
      return p * (hi - lo) + p + 1;
    }
    // 1D weight of coarse coordinate j at fine coordinate x: the first coarse cell (from lo)
    // holding both, the coarse basis function there (host: feat3_setup_prol)
    __host__ __device__ double weight(u64 x, u64 j) const
    {
      u64 lo, hi;
      cells(x, lo, hi);
      u64 ec = lo;
      while(ec < hi && !(j >= p * ec && j <= p * ec + p && x >= 2 * p * ec && x <= 2 * p * ec + 2 * p)) ++ec;
      const double t = double(x - 2 * p * ec) / double(2 * p);
      return lag1(int(p), int(j - p * ec), t);
    }
    // fine coordinates [xlo, xhi] whose prolongation row holds coarse coordinate j
    __host__ __device__ void fine_range(u64 j, u64 mf, u64& xlo, u64& xhi) const
    {
      const u64 w = 4 * p + 2;
      const u64 a = 2 * j > w ? 2 * j - w : 0, b = 2 * j + w < mf - 1 ? 2 * j + w : mf - 1;
      xlo = b + 1;
      xhi = 0;
      for(u64 x = a; x <= b; ++x)
      {
        u64 lo, hi;
        cells(x, lo, hi);
        if(j >= p * lo && j <= p * hi + p)
        {
          if(xlo > b) xlo = x;
          xhi = x;
        }
      }
    }
  };

  enum RowKind { kLaplace = 0, kProl = 1, kRest = 2 };

  // the 1D count of a coordinate for each pattern
  struct Counts
  {
    int kind;
    GridDev G;     // the grid whose nodes index the rows (fine for A and P, coarse for R)
    Prol1 T;       // prolongation geometry (P, R)
    u64 mf;        // fine nodes per direction (R)
    __host__ __device__ u64 cnt(u64 c) const
    {
      if(kind == kLaplace)
      {
        u64 lo, hi;
        G.cell_range(c, lo, hi);
        const u64 p = u64(G.deg);
        return (p * hi + p) - p * lo + 1;
      }
      if(kind == kProl) return T.count(c);
      u64 xlo, xhi;
      T.fine_range(c, mf, xlo, xhi);
      return xhi >= xlo ? xhi - xlo + 1 : 0;
    }
    __host__ __device__ u64 prefix(u64 c) const
    {
      u64 s = 0;
      for(u64 i = 0; i < c; ++i) s += cnt(i);
      return s;
    }
  };

  __global__ __launch_bounds__(256) void k_rowptr(Counts K, u64 n, u32* __restrict__ rp)
  {
    const int dim = K.G.dim;
    const u64 m = K.G.m;
    const u64 total1 = K.prefix(m);
    for(u64 r = u64(blockIdx.x) * blockDim.x + threadIdx.x; r <= n; r += u64(gridDim.x) * blockDim.x)
    {
      if(r == n)
      {
        u64 t = 1;
        for(int d = 0; d < dim; ++d) t *= total1;
        rp[n] = u32(t);
        continue;
      }
      u64 c[3] = {0, 0, 0};
      u64 tmp = r;
      for(int d = 0; d < dim; ++d) { c[d] = tmp % m; tmp /= m; }
      u64 s = 0;
      for(int d = 0; d < dim; ++d)
      {
        u64 term = K.prefix(c[d]);
        for(int e = 0; e < d; ++e) term *= total1;
        for(int e = d + 1; e < dim; ++e) term *= K.cnt(c[e]);
        s += term;
      }
      rp[r] = u32(s);
    }
  }

  // columns and values of the (coefficient-weighted) Laplace rows, and the Dirichlet list
  __global__ __launch_bounds__(256) void k_laplace_rows(GridDev G, RefMatrix R, double hs, const double* __restrict__ kcell,
                                                        u64 n, const u32* __restrict__ rp, u32* __restrict__ col,
                                                        double* __restrict__ val, u32* __restrict__ bnd)
  {
    const int q = G.deg + 1;
    int nloc = 1;
    for(int d = 0; d < G.dim; ++d) nloc *= q;
    const u64 p = u64(G.deg);
    for(u64 r = u64(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += u64(gridDim.x) * blockDim.x)
    {
      u64 c[3] = {0, 0, 0}, clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0}, nlo[3] = {0, 0, 0}, nhi[3] = {0, 0, 0};
      u64 tmp = r;
      bool boundary = false;
      u64 interior_before = 0;
      for(int d = 0; d < G.dim; ++d)
      {
        c[d] = tmp % G.m;
        tmp /= G.m;
        G.cell_range(c[d], clo[d], chi[d]);
        nlo[d] = p * clo[d];
        nhi[d] = p * chi[d] + p;
        if(c[d] == 0 || c[d] == G.m - 1) boundary = true;
      }
      if(boundary)
      {
        // interior nodes before r: sum_d IS(c_d) prod_{e<d} (m-2) prod_{e>d} [c_e interior]
        for(int d = 0; d < G.dim; ++d)
        {
          u64 term = c[d] == 0 ? 0 : (c[d] - 1 < G.m - 2 ? c[d] - 1 : G.m - 2);
          for(int e = 0; e < d; ++e) term *= G.m - 2;
          for(int e = d + 1; e < G.dim; ++e) term *= (c[e] >= 1 && c[e] <= G.m - 2) ? 1 : 0;
          interior_before += term;
        }
        bnd[r - interior_before] = u32(r);
      }
      u64 k = rp[r];
      u64 j[3] = {0, 0, 0};
      const u64 jz0 = G.dim == 3 ? nlo[2] : 0, jz1 = G.dim == 3 ? nhi[2] : 0;
      for(j[2] = jz0; j[2] <= jz1; ++j[2])
        for(j[1] = nlo[1]; j[1] <= nhi[1]; ++j[1])
          for(j[0] = nlo[0]; j[0] <= nhi[0]; ++j[0])
          {
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
              u64 e[3] = {0, 0, 0};
              const u64 ez0 = G.dim == 3 ? elo[2] : 0, ez1 = G.dim == 3 ? ehi[2] : 0;
              for(e[2] = ez0; e[2] <= ez1; ++e[2])
                for(e[1] = elo[1]; e[1] <= ehi[1]; ++e[1])
                  for(e[0] = elo[0]; e[0] <= ehi[0]; ++e[0])
                  {
                    int la = 0, lb = 0, mul = 1;
                    for(int d = 0; d < G.dim; ++d)
                    {
                      la += int(c[d] - p * e[d]) * mul;
                      lb += int(j[d] - p * e[d]) * mul;
                      mul *= q;
                    }
                    double kc = 1.0;
                    if(kcell)
                    {
                      u64 ci = 0;
                      for(int d = G.dim - 1; d >= 0; --d) ci = ci * G.cells + e[d];
                      kc = kcell[ci];
                    }
                    s += kc * R.k[la * nloc + lb];
                  }
            }
            u64 ji = 0;
            for(int d = G.dim - 1; d >= 0; --d) ji = ji * G.m + j[d];
            col[k] = u32(ji);
            val[k] = s * hs;
            ++k;
          }
    }
  }

  // prolongation rows (fine rows, coarse columns) or restriction rows (coarse rows, fine columns)
  __global__ __launch_bounds__(256) void k_transfer_rows(int kind, int dim, Prol1 T, u64 mf, u64 mc, u64 nrows,
                                                         const u32* __restrict__ rp, u32* __restrict__ col,
                                                         double* __restrict__ val)
  {
    const u64 p = T.p;
    const u64 mrow = kind == kProl ? mf : mc, mcol = kind == kProl ? mc : mf;
    for(u64 r = u64(blockIdx.x) * blockDim.x + threadIdx.x; r < nrows; r += u64(gridDim.x) * blockDim.x)
    {
      u64 c[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
      u64 tmp = r;
      for(int d = 0; d < dim; ++d)
      {
        c[d] = tmp % mrow;
        tmp /= mrow;
        if(kind == kProl)
        {
          u64 a, b;
          T.cells(c[d], a, b);
          lo[d] = p * a;
          hi[d] = p * b + p;
        }
        else
          T.fine_range(c[d], mf, lo[d], hi[d]);
      }
      u64 k = rp[r];
      u64 j[3] = {0, 0, 0};
      const u64 jz0 = dim == 3 ? lo[2] : 0, jz1 = dim == 3 ? hi[2] : 0;
      for(j[2] = jz0; j[2] <= jz1; ++j[2])
        for(j[1] = lo[1]; j[1] <= hi[1]; ++j[1])
          for(j[0] = lo[0]; j[0] <= hi[0]; ++j[0])
          {
            // the value is P(fine, coarse) in either orientation, multiplied as the host does
            double w = 1.0;
            for(int d = 0; d < dim; ++d)
              w *= kind == kProl ? T.weight(c[d], j[d]) : T.weight(j[d], c[d]);
            u64 ji = 0;
            for(int d = dim - 1; d >= 0; --d) ji = ji * mcol + j[d];
            col[k] = u32(ji);
            val[k] = w;
            ++k;
          }
    }
  }

#pragma clang fp contract(on)

  inline unsigned blocks_for(u64 items)
  {
    const u64 b = (items + 255) / 256;
    return unsigned(b < 1 ? 1 : (b > 65536 ? 65536 : b));
  }
} // namespace

namespace
{
  bool valid_space(int dim, int degree) { return (dim == 2 && (degree == 1 || degree == 2)) || (dim == 3 && degree == 1); }
  GridDev grid_of(int dim, int degree, int level)
  {
    return GridDev{dim, degree, u64(1) << level, u64(degree) * (u64(1) << level) + 1};
  }
  u64 ipow(u64 b, int e)
  {
    u64 r = 1;
    for(int i = 0; i < e; ++i) r *= b;
    return r;
  }
  Counts counts_for(int kind, int dim, int degree, int level_fine)
  {
    Counts K{};
    K.kind = kind;
    const GridDev F = grid_of(dim, degree, level_fine);
    const GridDev C = grid_of(dim, degree, level_fine > 0 ? level_fine - 1 : 0);
    K.G = kind == kRest ? C : F;
    K.T = Prol1{u64(degree), C.cells};
    K.mf = F.m;
    return K;
  }
  // rows and entries of a pattern (entries = (1D total)^dim)
  void pattern_sizes(const Counts& K, u64& rows, u64& nnz)
  {
    rows = ipow(K.G.m, K.G.dim);
    nnz = ipow(K.prefix(K.G.m), K.G.dim);
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


int feat3_device_laplace_sizes(int dim, int degree, int level, uint64_t* n, uint64_t* nnz, uint64_t* nbnd)
{
  if(!valid_space(dim, degree) || level < 0 || level > 12 || !n || !nnz || !nbnd) return rt::kErrArg;
  const Counts K = counts_for(kLaplace, dim, degree, level);
  u64 r = 0, z = 0;
  pattern_sizes(K, r, z);
  if(z >= (u64(1) << 32)) return rt::kErrArg;
  *n = r;
  *nnz = z;
  *nbnd = r - ipow(K.G.m - 2, dim);
  return 0;
}

int feat3_device_laplace(void* stream, int dim, int degree, int level, const double* ref_stiffness, const double* kcell,
                         uint32_t* row_ptr, uint32_t* col_ind, double* val, uint32_t* bnd)
{
  u64 n = 0, nnz = 0, nb = 0;
  if(feat3_device_laplace_sizes(dim, degree, level, &n, &nnz, &nb)) return rt::kErrArg;
  if(!ref_stiffness || !row_ptr || !col_ind || !val || !bnd) return rt::kErrArg;
  const Counts K = counts_for(kLaplace, dim, degree, level);
  const int q = degree + 1;
  const int nloc = dim == 2 ? q * q : q * q * q;
  RefMatrix R{};
  for(int i = 0; i < nloc * nloc; ++i) R.k[i] = ref_stiffness[i];
  const double h = 1.0 / double(K.G.cells);
  double hs = 1.0;
  for(int d = 2; d < dim; ++d) hs *= h;
  hipStream_t s = static_cast<hipStream_t>(stream);
  k_rowptr<<<blocks_for(n + 1), 256, 0, s>>>(K, n, row_ptr);
  k_laplace_rows<<<blocks_for(n), 256, 0, s>>>(K.G, R, hs, kcell, n, row_ptr, col_ind, val, bnd);
  return rt::check_launch();
}

int feat3_device_prol_sizes(int dim, int degree, int level_fine, uint64_t* nf, uint64_t* nc, uint64_t* nnz)
{
  if(!valid_space(dim, degree) || level_fine < 1 || level_fine > 12 || !nf || !nc || !nnz) return rt::kErrArg;
  const Counts K = counts_for(kProl, dim, degree, level_fine);
  u64 r = 0, z = 0;
  pattern_sizes(K, r, z);
  if(z >= (u64(1) << 32)) return rt::kErrArg;
  *nf = r;
  *nc = ipow(grid_of(dim, degree, level_fine - 1).m, dim);
  *nnz = z;
  return 0;
}

int feat3_device_transfer(void* stream, int dim, int degree, int level_fine, int transpose, uint32_t* row_ptr,
                          uint32_t* col_ind, double* val)
{
  u64 nf = 0, nc = 0, nnz = 0;
  if(feat3_device_prol_sizes(dim, degree, level_fine, &nf, &nc, &nnz)) return rt::kErrArg;
  if(!row_ptr || !col_ind || !val) return rt::kErrArg;
  const int kind = transpose ? kRest : kProl;
  const Counts K = counts_for(kind, dim, degree, level_fine);
  const u64 rows = transpose ? nc : nf;
  hipStream_t s = static_cast<hipStream_t>(stream);
  k_rowptr<<<blocks_for(rows + 1), 256, 0, s>>>(K, rows, row_ptr);
  k_transfer_rows<<<blocks_for(rows), 256, 0, s>>>(kind, dim, K.T, K.mf, grid_of(dim, degree, level_fine - 1).m, rows,
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
