// structured_fe.cpp -- host-side problem setup for the LAFEM/Solver hot path:
// Q1/Q2 Lagrange spaces on the refined unit square / unit cube, the (coefficient-weighted)
// Laplace matrix, the two-level prolongation matrix, the Dirichlet (unit filter) index set
// and load vectors.
//
// This is the step *before* the hot path (FEAT: kernel/assembly + kernel/geometry).  It runs
// once on the host and produces the CSR arrays, filter lists and vectors that the device
// containers take over.  DOFs are numbered lexicographically on the tensor grid
// (x fastest); the reference numbers them hierarchically through its refinery
// (geometry/intern/standard_vertex_refiner.hpp:65-250), which is a row/column permutation
// of the same matrices.
//
// Semantics restated from the reference:
//  * sparsity = "all DOFs sharing a cell" with ascending column indices
//    (assembly/symbolic_assembler.hpp:59-110, assemble_matrix_std1)
//  * matrix entries = sum over cells of k_c * int grad(phi_j).grad(phi_i), 3-point
//    Gauss-Legendre per direction (assembly/common_operators.hpp:36-97 LaplaceOperator,
//    cubature/auto_alias.hpp:187-205 "auto-degree:5")
//  * prolongation pattern = "fine DOF in a child of a coarse cell that holds the coarse DOF"
//    (symbolic_assembler.hpp:319-343,483-489 assemble_matrix_2lvl) including the structural
//    zeros that pattern implies; entries = the coarse basis function at the fine node, which
//    is what the row-averaged local L2 projection of grid_transfer.hpp:58-338 evaluates to
//    for nested Lagrange spaces
//  * Dirichlet set = all DOFs on the boundary in ascending order
//    (assembly/unit_filter_assembler.hpp:101-130)
//  * load vector = int f phi_i with the same cubature (LinearFunctionalAssembler)

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "feat3_hip.h"

namespace
{
  typedef std::uint32_t u32;
  typedef std::uint64_t u64;

  const double kPi = 3.14159265358979323846;

  // 3-point Gauss-Legendre rule mapped to [0,1]
  const double gq_t[3] = {0.5 - 0.5 * 0.77459666924148337704, 0.5, 0.5 + 0.5 * 0.77459666924148337704};
  const double gq_w[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};

  // 1D Lagrange basis of degree p (1 or 2) on [0,1] with equidistant nodes, value and slope
  double lag(int p, int a, double t)
  {
    if(p == 1)
      return a == 0 ? 1.0 - t : t;
    if(a == 0) return 2.0 * (t - 0.5) * (t - 1.0);
    if(a == 1) return 4.0 * t * (1.0 - t);
    return 2.0 * t * (t - 0.5);
  }
  double dlag(int p, int a, double t)
  {
    if(p == 1)
      return a == 0 ? -1.0 : 1.0;
    if(a == 0) return 4.0 * t - 3.0;
    if(a == 1) return 4.0 - 8.0 * t;
    return 4.0 * t - 1.0;
  }

  struct Grid
  {
    int dim, deg, level;
    u64 cells;   // cells per direction
    u64 m;       // nodes per direction
    u64 n;       // total nodes
    Grid(int d, int p, int l) : dim(d), deg(p), level(l)
    {
      cells = u64(1) << l;
      m = u64(p) * cells + 1;
      n = 1;
      for(int i = 0; i < d; ++i) n *= m;
    }
    void coords(u64 idx, u64* c) const
    {
      for(int d = 0; d < dim; ++d) { c[d] = idx % m; idx /= m; }
    }
    u64 index(const u64* c) const
    {
      u64 r = 0;
      for(int d = dim - 1; d >= 0; --d) r = r * m + c[d];
      return r;
    }
    // range of cells containing node coordinate x (inclusive)
    void cell_range(u64 x, u64& lo, u64& hi) const
    {
      const u64 p = u64(deg);
      lo = (x >= p) ? (x - 1) / p : 0;   // smallest e with p*e + p >= x
      if(x == 0) lo = 0;
      hi = std::min<u64>(x / p, cells - 1);
      if(lo > hi) lo = hi;
    }
  };

  // reference element stiffness on the unit cell (h = 1), local lexicographic node order;
  // scaled by h^(dim-2) for a cell of width h
  std::vector<double> ref_stiffness(int dim, int p)
  {
    const int q = p + 1;
    int nloc = 1;
    for(int d = 0; d < dim; ++d) nloc *= q;
    std::vector<double> K(std::size_t(nloc) * nloc, 0.0);
    int npt = 1;
    for(int d = 0; d < dim; ++d) npt *= 3;
    for(int ip = 0; ip < npt; ++ip)
    {
      int qi[3] = {0, 0, 0};
      int t = ip;
      double w = 1.0;
      for(int d = 0; d < dim; ++d) { qi[d] = t % 3; t /= 3; w *= gq_w[qi[d]]; }
      std::vector<double> grad(std::size_t(nloc) * dim);
      for(int a = 0; a < nloc; ++a)
      {
        int ai[3] = {0, 0, 0};
        int s = a;
        for(int d = 0; d < dim; ++d) { ai[d] = s % q; s /= q; }
        for(int g = 0; g < dim; ++g)
        {
          double v = 1.0;
          for(int d = 0; d < dim; ++d)
            v *= (d == g) ? dlag(p, ai[d], gq_t[qi[d]]) : lag(p, ai[d], gq_t[qi[d]]);
          grad[std::size_t(a) * dim + g] = v;
        }
      }
      for(int a = 0; a < nloc; ++a)
        for(int b = 0; b < nloc; ++b)
        {
          double s = 0.0;
          for(int g = 0; g < dim; ++g) s += grad[std::size_t(a) * dim + g] * grad[std::size_t(b) * dim + g];
          K[std::size_t(a) * nloc + b] += w * s;
        }
    }
    return K;
  }

  u64 box_count(const Grid& G, u64 row)
  {
    u64 c[3];
    G.coords(row, c);
    u64 cnt = 1;
    for(int d = 0; d < G.dim; ++d)
    {
      u64 lo, hi;
      G.cell_range(c[d], lo, hi);
      cnt *= (u64(G.deg) * hi + u64(G.deg)) - u64(G.deg) * lo + 1;
    }
    return cnt;
  }

  bool on_boundary(const Grid& G, u64 row)
  {
    u64 c[3];
    G.coords(row, c);
    for(int d = 0; d < G.dim; ++d)
      if(c[d] == 0 || c[d] == G.m - 1) return true;
    return false;
  }

  double h_pow(const Grid& G)
  {
    const double h = 1.0 / double(G.cells);
    double s = 1.0;
    for(int d = 2; d < G.dim; ++d) s *= h;
    for(int d = G.dim; d < 2; ++d) s /= h;
    return s;
  }

  // analytic right-hand sides f = -Laplace(u)
  double rhs_func(int func, int dim, const double* x)
  {
    if(func == 0) // sine bubble: f = dim * pi^2 * prod sin(pi x_i)
    {
      double v = double(dim) * kPi * kPi;
      for(int d = 0; d < dim; ++d) v *= std::sin(kPi * x[d]);
      return v;
    }
    if(func == 1) // constant one
      return 1.0;
    // exp bubble (analytic/common.hpp:580-622): u = prod ub(x_i)
    const double em1 = std::exp(-1.0);
    double u[3], uxx[3];
    for(int d = 0; d < dim; ++d)
    {
      const double t = x[d];
      const double e = std::exp(-(2.0 * t - 1.0) * (2.0 * t - 1.0));
      u[d] = (e - em1) / (1.0 - em1);
      uxx[d] = (64.0 * t * (1.0 - t) - 8.0) * e / (em1 - 1.0);
    }
    double lap = 0.0;
    for(int g = 0; g < dim; ++g)
    {
      double v = uxx[g];
      for(int d = 0; d < dim; ++d) if(d != g) v *= u[d];
      lap += v;
    }
    return -lap;
  }
} // namespace

extern "C" {

// Sizes of the Laplace system on level `level` (cells per direction 2^level) for a
// Lagrange space of degree 1 or 2 in 2 or 3 dimensions.
int feat3_setup_laplace_sizes(int dim, int degree, int level, u64* n, u64* nnz, u64* nbnd)
{
  if(dim < 2 || dim > 3 || degree < 1 || degree > 2 || level < 0 || level > 12) return 1;
  Grid G(dim, degree, level);
  u64 z = 0, b = 0;
  for(u64 r = 0; r < G.n; ++r)
  {
    z += box_count(G, r);
    if(on_boundary(G, r)) ++b;
  }
  *n = G.n; *nnz = z; *nbnd = b;
  return 0;
}

// Assembles the (coefficient-weighted) Laplace matrix.  kcell: one coefficient per cell in
// lexicographic cell order, or NULL for k = 1.  bnd receives the ascending Dirichlet list.
int feat3_setup_laplace(int dim, int degree, int level, const double* kcell, u32* rp, u32* col, double* val, u32* bnd)
{
  if(dim < 2 || dim > 3 || degree < 1 || degree > 2 || level < 0 || level > 12) return 1;
  Grid G(dim, degree, level);
  const std::vector<double> K = ref_stiffness(dim, degree);
  const int q = degree + 1;
  int nloc = 1;
  for(int d = 0; d < dim; ++d) nloc *= q;
  const double hs = h_pow(G);
  const u64 p = u64(degree);
  u64 k = 0, nb = 0;
  rp[0] = 0;
  for(u64 r = 0; r < G.n; ++r)
  {
    u64 c[3] = {0, 0, 0};
    G.coords(r, c);
    u64 clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0}, nlo[3] = {0, 0, 0}, nhi[3] = {0, 0, 0};
    for(int d = 0; d < dim; ++d)
    {
      G.cell_range(c[d], clo[d], chi[d]);
      nlo[d] = p * clo[d];
      nhi[d] = p * chi[d] + p;
    }
    // columns in ascending order: z slowest, x fastest
    u64 j[3] = {0, 0, 0};
    const u64 jz0 = dim == 3 ? nlo[2] : 0, jz1 = dim == 3 ? nhi[2] : 0;
    for(j[2] = jz0; j[2] <= jz1; ++j[2])
      for(j[1] = nlo[1]; j[1] <= nhi[1]; ++j[1])
        for(j[0] = nlo[0]; j[0] <= nhi[0]; ++j[0])
        {
          // cells containing both nodes: per direction intersection of cell ranges
          u64 elo[3] = {0, 0, 0}, ehi[3] = {0, 0, 0};
          bool any = true;
          for(int d = 0; d < dim; ++d)
          {
            u64 lo2, hi2;
            G.cell_range(j[d], lo2, hi2);
            elo[d] = std::max(clo[d], lo2);
            ehi[d] = std::min(chi[d], hi2);
            if(elo[d] > ehi[d]) any = false;
          }
          double s = 0.0;
          if(any)
          {
            u64 e[3] = {0, 0, 0};
            const u64 ez0 = dim == 3 ? elo[2] : 0, ez1 = dim == 3 ? ehi[2] : 0;
            for(e[2] = ez0; e[2] <= ez1; ++e[2])
              for(e[1] = elo[1]; e[1] <= ehi[1]; ++e[1])
                for(e[0] = elo[0]; e[0] <= ehi[0]; ++e[0])
                {
                  int la = 0, lb = 0, mul = 1;
                  for(int d = 0; d < dim; ++d)
                  {
                    la += int(c[d] - p * e[d]) * mul;
                    lb += int(j[d] - p * e[d]) * mul;
                    mul *= q;
                  }
                  double kc = 1.0;
                  if(kcell)
                  {
                    u64 ci = 0;
                    for(int d = dim - 1; d >= 0; --d) ci = ci * G.cells + e[d];
                    kc = kcell[ci];
                  }
                  s += kc * K[std::size_t(la) * nloc + lb];
                }
          }
          col[k] = u32(G.index(j));
          val[k] = s * hs;
          ++k;
        }
    rp[r + 1] = u32(k);
    if(on_boundary(G, r)) bnd[nb++] = u32(r);
  }
  return 0;
}

// Load vector b_i = int f phi_i (func: 0 sine bubble, 1 constant one, 2 exp bubble).
int feat3_setup_rhs(int dim, int degree, int level, int func, double* rhs)
{
  if(dim < 2 || dim > 3 || degree < 1 || degree > 2 || level < 0 || level > 12) return 1;
  Grid G(dim, degree, level);
  const int q = degree + 1;
  int nloc = 1, npt = 1;
  for(int d = 0; d < dim; ++d) { nloc *= q; npt *= 3; }
  const double h = 1.0 / double(G.cells);
  double vol = 1.0;
  for(int d = 0; d < dim; ++d) vol *= h;
  std::fill(rhs, rhs + G.n, 0.0);
  u64 ncell = 1;
  for(int d = 0; d < dim; ++d) ncell *= G.cells;
  std::vector<double> loc(nloc);
  for(u64 ci = 0; ci < ncell; ++ci)
  {
    u64 e[3] = {0, 0, 0};
    u64 t = ci;
    for(int d = 0; d < dim; ++d) { e[d] = t % G.cells; t /= G.cells; }
    std::fill(loc.begin(), loc.end(), 0.0);
    for(int ip = 0; ip < npt; ++ip)
    {
      int qi[3] = {0, 0, 0};
      int s = ip;
      double w = vol;
      double x[3] = {0, 0, 0};
      for(int d = 0; d < dim; ++d)
      {
        qi[d] = s % 3; s /= 3;
        w *= gq_w[qi[d]];
        x[d] = (double(e[d]) + gq_t[qi[d]]) * h;
      }
      const double f = rhs_func(func, dim, x);
      for(int a = 0; a < nloc; ++a)
      {
        int ai = a;
        double phi = 1.0;
        for(int d = 0; d < dim; ++d) { phi *= lag(degree, ai % q, gq_t[qi[d]]); ai /= q; }
        loc[a] += w * f * phi;
      }
    }
    for(int a = 0; a < nloc; ++a)
    {
      int ai = a;
      u64 nc[3] = {0, 0, 0};
      for(int d = 0; d < dim; ++d) { nc[d] = u64(degree) * e[d] + u64(ai % q); ai /= q; }
      rhs[G.index(nc)] += loc[a];
    }
  }
  return 0;
}

// Prolongation from level_fine-1 to level_fine: sizes, then arrays (fine rows x coarse cols).
int feat3_setup_prol_sizes(int dim, int degree, int level_fine, u64* nf, u64* nc, u64* nnz)
{
  if(dim < 2 || dim > 3 || degree < 1 || degree > 2 || level_fine < 1 || level_fine > 12) return 1;
  Grid F(dim, degree, level_fine), C(dim, degree, level_fine - 1);
  const u64 p = u64(degree);
  u64 z = 0;
  for(u64 r = 0; r < F.n; ++r)
  {
    u64 c[3];
    F.coords(r, c);
    u64 cnt = 1;
    for(int d = 0; d < dim; ++d)
    {
      // coarse cells ec with 2p*ec <= x <= 2p*ec + 2p
      const u64 x = c[d];
      const u64 lo = x >= 2 * p ? (x - 1) / (2 * p) : 0;
      const u64 hi = std::min<u64>(x / (2 * p), C.cells - 1);
      cnt *= p * (hi - std::min(lo, hi)) + p + 1;
    }
    z += cnt;
  }
  *nf = F.n; *nc = C.n; *nnz = z;
  return 0;
}

int feat3_setup_prol(int dim, int degree, int level_fine, u32* rp, u32* col, double* val)
{
  if(dim < 2 || dim > 3 || degree < 1 || degree > 2 || level_fine < 1 || level_fine > 12) return 1;
  Grid F(dim, degree, level_fine), C(dim, degree, level_fine - 1);
  const u64 p = u64(degree);
  u64 k = 0;
  rp[0] = 0;
  for(u64 r = 0; r < F.n; ++r)
  {
    u64 c[3] = {0, 0, 0};
    F.coords(r, c);
    u64 jlo[3] = {0, 0, 0}, jhi[3] = {0, 0, 0}, elo[3] = {0, 0, 0}, ehi[3] = {0, 0, 0};
    for(int d = 0; d < dim; ++d)
    {
      const u64 x = c[d];
      u64 lo = x >= 2 * p ? (x - 1) / (2 * p) : 0;
      const u64 hi = std::min<u64>(x / (2 * p), C.cells - 1);
      lo = std::min(lo, hi);
      elo[d] = lo; ehi[d] = hi;
      jlo[d] = p * lo; jhi[d] = p * hi + p;
    }
    u64 j[3] = {0, 0, 0};
    const u64 jz0 = dim == 3 ? jlo[2] : 0, jz1 = dim == 3 ? jhi[2] : 0;
    for(j[2] = jz0; j[2] <= jz1; ++j[2])
      for(j[1] = jlo[1]; j[1] <= jhi[1]; ++j[1])
        for(j[0] = jlo[0]; j[0] <= jhi[0]; ++j[0])
        {
          double w = 1.0;
          for(int d = 0; d < dim; ++d)
          {
            // a coarse cell containing both the fine node and the coarse node
            u64 ec = elo[d];
            while(ec < ehi[d] && !(j[d] >= p * ec && j[d] <= p * ec + p && c[d] >= 2 * p * ec && c[d] <= 2 * p * ec + 2 * p))
              ++ec;
            const double t = double(c[d] - 2 * p * ec) / double(2 * p);
            w *= lag(int(p), int(j[d] - p * ec), t);
          }
          col[k] = u32(C.index(j));
          val[k] = w;
          ++k;
        }
    rp[r + 1] = u32(k);
  }
  return 0;
}

// CSR transpose (ascending columns per row), as LAFEM's SparseMatrixCSR::transpose.
int feat3_setup_transpose(u64 rows, u64 cols, const u32* rp, const u32* col, const double* val,
                          u32* trp, u32* tcol, double* tval)
{
  std::vector<u64> cnt(cols + 1, 0);
  const u64 nnz = rp[rows];
  for(u64 i = 0; i < nnz; ++i) ++cnt[col[i] + 1];
  for(u64 c = 0; c < cols; ++c) cnt[c + 1] += cnt[c];
  for(u64 c = 0; c <= cols; ++c) trp[c] = u32(cnt[c]);
  for(u64 r = 0; r < rows; ++r)
    for(u64 k = rp[r]; k < rp[r + 1]; ++k)
    {
      const u64 j = cnt[col[k]]++;
      tval[j] = val[k];
      tcol[j] = u32(r);
    }
  return 0;
}

// Cell coefficient on the next coarser level: arithmetic mean of the 2^dim children.
int feat3_setup_coarsen_cellfield(int dim, int level_fine, const double* kf, double* kc)
{
  if(dim < 2 || dim > 3 || level_fine < 1) return 1;
  const u64 cf = u64(1) << level_fine, cc = cf / 2;
  u64 ncc = 1;
  for(int d = 0; d < dim; ++d) ncc *= cc;
  for(u64 ci = 0; ci < ncc; ++ci)
  {
    u64 e[3] = {0, 0, 0};
    u64 t = ci;
    for(int d = 0; d < dim; ++d) { e[d] = t % cc; t /= cc; }
    double s = 0.0;
    const int nch = 1 << dim;
    for(int ch = 0; ch < nch; ++ch)
    {
      u64 fi = 0;
      for(int d = dim - 1; d >= 0; --d) fi = fi * cf + 2 * e[d] + u64((ch >> d) & 1);
      s += kf[fi];
    }
    kc[ci] = s / double(nch);
  }
  return 0;
}

// The reference Q1/Q2 element stiffness on the unit cell (lexicographic local nodes),
// nloc x nloc, for the adjoint gradient kernel and for tests.
int feat3_setup_ref_stiffness(int dim, int degree, double* K)
{
  if(dim < 2 || dim > 3 || degree < 1 || degree > 2) return 1;
  const std::vector<double> k = ref_stiffness(dim, degree);
  std::copy(k.begin(), k.end(), K);
  return 0;
}

// n doubles uniform in [a, b] from the reference's xorshift64* stream (util/random.hpp:62-110:
// shifts 12/25/27, multiplier 2685821657736338717, a zero seed replaced by 28054777172512;
// doubles (next & 0xFFFFFFFF) / 0xFFFFFFFF mapped to a + x (b - a), random.hpp:259-280)
int feat3_setup_random_uniform(std::uint64_t seed, std::uint64_t n, double a, double b, double* out)
{
  if(n && !out) return 1;
  std::uint64_t x = seed != 0 ? seed : std::uint64_t(28054777172512ull);
  const double d = 1.0 / double(std::uint64_t(0xFFFFFFFFull));
  const bool ranged = a + 2.220446049250313e-16 < b;
  for(std::uint64_t i = 0; i < n; ++i)
  {
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    const double u = double((x * std::uint64_t(2685821657736338717ull)) & std::uint64_t(0xFFFFFFFFull)) * d;
    out[i] = ranged ? a + u * (b - a) : a;
  }
  return 0;
}

} // extern "C"
