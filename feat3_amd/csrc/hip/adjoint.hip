// adjoint.hip -- the per-element gradient of a loss through the variable-coefficient diffusion
// solve, for K coefficient fields at once.
//
// The system matrix is A(k) = sum_e k_e h^(d-2) K_ref (Q1/Q2 Lagrange on the refined unit
// square/cube, lexicographic local nodes), so for L(u(k)) with A(k) u = b and adjoint
// A lambda = dL/du (both with the homogeneous Dirichlet filter),
//     dL/dk_e = -lambda_e^T (h^(d-2) K_ref) u_e .
// The reference has no adjoint or coefficient operator; its element assembly
// (assembly/bilinear_operator_assembler.hpp, common_operators.hpp:36-97) defines K_ref.
//
// Layout: u, lambda [node][system], gradient [cell][system] (cells x fastest).  One thread per
// (cell, system): the K lanes of one cell read the K contiguous values of each local node, and
// neighbouring cells share nodes through L2, so HBM sees u and lambda about once.

#include "device_common.hpp"
#include "runtime.hpp"
#include "feat3_hip.h"

#include <hip/hip_runtime.h>

using namespace feat3_amd;
typedef std::uint64_t u64;

namespace
{
  constexpr int kMaxLoc = 9;   // Q1 2D (4), Q1 3D (8), Q2 2D (9)

  struct ElemMatrix
  {
    double k[kMaxLoc * kMaxLoc];
  };

  template<int DIM, int DEG>
  __global__ __launch_bounds__(256) void k_cell_gradient(u64 cells_per_dim, int K, ElemMatrix E, const double* __restrict__ u,
                                                         const double* __restrict__ lam, double* __restrict__ g)
  {
    constexpr int Q = DEG + 1;
    constexpr int NLOC = DIM == 2 ? Q * Q : Q * Q * Q;
    const u64 np1 = u64(DEG) * cells_per_dim + 1;
    u64 ncell = cells_per_dim * cells_per_dim;
    if(DIM == 3) ncell *= cells_per_dim;
    const u64 total = ncell * u64(K);
    for(u64 t = u64(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += u64(gridDim.x) * blockDim.x)
    {
      const u64 c = t / u64(K);
      const u64 k = t % u64(K);
      u64 e[3] = {0, 0, 0};
      u64 r = c;
#pragma unroll
      for(int d = 0; d < DIM; ++d) { e[d] = r % cells_per_dim; r /= cells_per_dim; }
      double uu[NLOC], ll[NLOC];
#pragma unroll
      for(int a = 0; a < NLOC; ++a)
      {
        int s = a;
        u64 node = 0, stride = 1;
#pragma unroll
        for(int d = 0; d < DIM; ++d)
        {
          node += (u64(DEG) * e[d] + u64(s % Q)) * stride;
          s /= Q;
          stride *= np1;
        }
        uu[a] = u[node * u64(K) + k];
        ll[a] = lam[node * u64(K) + k];
      }
      double acc = 0.0;
#pragma unroll
      for(int a = 0; a < NLOC; ++a)
      {
        double s = 0.0;
#pragma unroll
        for(int b = 0; b < NLOC; ++b) s += E.k[a * NLOC + b] * uu[b];
        acc += ll[a] * s;
      }
      g[t] = -acc;
    }
  }
} // namespace

extern "C" {

int feat3_adjoint_cell_gradient(void* stream, int dim, int degree, int level, int nsys, const double* elem_matrix,
                                const double* u, const double* lam, double* grad)
{
  if(!(dim == 2 || dim == 3) || !(degree == 1 || degree == 2) || (dim == 3 && degree == 2)) return rt::kErrArg;
  if(level < 0 || level > 12 || nsys < 1 || !elem_matrix || !u || !lam || !grad) return rt::kErrArg;
  const int q = degree + 1;
  const int nloc = dim == 2 ? q * q : q * q * q;
  ElemMatrix E{};
  for(int i = 0; i < nloc * nloc; ++i) E.k[i] = elem_matrix[i];
  const u64 m = u64(1) << level;
  u64 ncell = m * m * (dim == 3 ? m : 1);
  const u64 total = ncell * u64(nsys);
  const unsigned blocks = unsigned(total / 256 + 1 < 65536 ? total / 256 + 1 : 65536);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if(dim == 2 && degree == 1) k_cell_gradient<2, 1><<<blocks, 256, 0, s>>>(m, nsys, E, u, lam, grad);
  else if(dim == 2) k_cell_gradient<2, 2><<<blocks, 256, 0, s>>>(m, nsys, E, u, lam, grad);
  else k_cell_gradient<3, 1><<<blocks, 256, 0, s>>>(m, nsys, E, u, lam, grad);
  return rt::check_launch();
}

} // extern "C"
