/*
 * feat3_hip.h -- C ABI of the MI355X-native LAFEM + Solver hot path (libfeat3_hip.so).
 *
 * Conventions
 *  - Every function returns int: 0 on success, otherwise one of FEAT3_HIP_ERR_*; the
 *    message is feat3_hip_error_string(code).  A reference-side shim turns a non-zero code
 *    into FEAT::InternalError (util/exception.hpp), as the reference's CUDA backend does for
 *    cuBLAS/cuSPARSE errors (lafem/arch/apply.cu:104-105).
 *  - `stream` is a hipStream_t passed as void* (NULL = the legacy default stream).  Kernels
 *    are launched asynchronously on it; only the *_value reductions synchronise it, like the
 *    reference's host-returning cuBLAS reductions (lafem/arch/dot_product.cu:22-41).
 *  - All array arguments are device pointers unless stated otherwise; sizes are element
 *    counts.  Suffixes: f64/f32 = value type, u32/u64 = index type (FEAT's Index is u64,
 *    kernel/base_header.hpp:94-98), h16f32 = binary16 stored values with FP32 vectors.
 *  - No function allocates device memory per call: feat3_hip_initialise creates each device's
 *    reduction workspace (the dot/norm entry points return FEAT3_HIP_ERR_STATE on a device
 *    that was not initialised); the solver objects allocate in their init only.
 *  - The dot/norm entry points of one device share its workspace; they may be issued on any
 *    streams: each waits for the previous one (an event recorded by every reduction), so
 *    reductions on different streams are serialised against each other (not inside a stream
 *    capture, whose graph orders its own nodes).
 *  - Not thread-safe by contract (one host thread drives one device), as the reference.
 *
 * The first group mirrors LAFEM::Arch::<Op><Mem::CUDA> (lafem/arch/apply.hpp:267-303 and the
 * sibling headers): a reference build adds Mem::HIP and forwards Arch::<Op><Mem::HIP> to these
 * functions with the same raw-pointer arguments.  The second group mirrors
 * MemoryPool<Mem::CUDA> (util/memory_pool.hpp:53-377, util/memory_pool.cu:56-320).  The third
 * group is the fused device-resident solver with the Solver lifecycle of solver/base.hpp.
 */
#ifndef FEAT3_HIP_H
#define FEAT3_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FEAT3_HIP_OK 0
#define FEAT3_HIP_ERR_ARG 1     /* invalid argument: null pointer, unsupported size/format */
#define FEAT3_HIP_ERR_LAUNCH 2  /* kernel launch failed */
#define FEAT3_HIP_ERR_RUNTIME 3 /* HIP runtime error (allocation, copy, synchronisation) */
#define FEAT3_HIP_ERR_STATE 4   /* object used in the wrong lifecycle state */
#define FEAT3_HIP_ERR_NUMERIC 5 /* numerical breakdown during setup (singular coarse matrix) */

/* solver status, as FEAT::Solver::Status (solver/base.hpp:48-64) */
#define FEAT3_STATUS_UNDEFINED 0
#define FEAT3_STATUS_PROGRESS 1
#define FEAT3_STATUS_SUCCESS 2
#define FEAT3_STATUS_ABORTED 3
#define FEAT3_STATUS_DIVERGED 4
#define FEAT3_STATUS_MAX_ITER 5
#define FEAT3_STATUS_STAGNATED 6

const char* feat3_hip_error_string(int code);

/* ===================================================================================== */
/* Arch::Apply<Mem::HIP>                                                                  */
/* ===================================================================================== */

/* Apply::csr (apply.hpp:271-274, CUDA apply.cu:242-302): r = a*A*x + b*y; |b| < eps ignores
 * y; r may alias y but not x; transposed != 0 computes r = a*A^T*x + b*y. */
int lafem_hip_apply_csr_f64_u32(void* stream, double* r, double a, const double* x, double b, const double* y,
                                const double* val, const uint32_t* col_ind, const uint32_t* row_ptr, uint64_t rows,
                                uint64_t columns, uint64_t used_elements, int transposed);
int lafem_hip_apply_csr_f64_u64(void* stream, double* r, double a, const double* x, double b, const double* y,
                                const double* val, const uint64_t* col_ind, const uint64_t* row_ptr, uint64_t rows,
                                uint64_t columns, uint64_t used_elements, int transposed);
int lafem_hip_apply_csr_f32_u32(void* stream, float* r, float a, const float* x, float b, const float* y,
                                const float* val, const uint32_t* col_ind, const uint32_t* row_ptr, uint64_t rows,
                                uint64_t columns, uint64_t used_elements, int transposed);
int lafem_hip_apply_csr_f32_u64(void* stream, float* r, float a, const float* x, float b, const float* y,
                                const float* val, const uint64_t* col_ind, const uint64_t* row_ptr, uint64_t rows,
                                uint64_t columns, uint64_t used_elements, int transposed);
/* binary16 stored values, FP32 arithmetic (no reference counterpart: storage variant of
 * Apply::csr for the mixed-precision smoother) */
int lafem_hip_apply_csr_h16f32_u32(void* stream, float* r, float a, const float* x, float b, const float* y,
                                   const void* val, const uint32_t* col_ind, const uint32_t* row_ptr, uint64_t rows,
                                   uint64_t columns, uint64_t used_elements);

/* Apply::ell (apply.hpp:296, CUDA apply.cu:42-61,433-457): FEAT's SELL-C layout with
 * sigma = 1: val/col_ind[cs[row/C] + row%C + k*C], cl = chunk lengths. */
int lafem_hip_apply_ell_f64_u32(void* stream, double* r, double a, const double* x, double b, const double* y,
                                const double* val, const uint32_t* col_ind, const uint32_t* cs, const uint32_t* cl,
                                uint64_t C, uint64_t rows);
int lafem_hip_apply_ell_f64_u64(void* stream, double* r, double a, const double* x, double b, const double* y,
                                const double* val, const uint64_t* col_ind, const uint64_t* cs, const uint64_t* cl,
                                uint64_t C, uint64_t rows);
int lafem_hip_apply_ell_f32_u32(void* stream, float* r, float a, const float* x, float b, const float* y,
                                const float* val, const uint32_t* col_ind, const uint32_t* cs, const uint32_t* cl,
                                uint64_t C, uint64_t rows);
int lafem_hip_apply_ell_f32_u64(void* stream, float* r, float a, const float* x, float b, const float* y,
                                const float* val, const uint64_t* col_ind, const uint64_t* cs, const uint64_t* cl,
                                uint64_t C, uint64_t rows);
/* SELL-C-sigma (extension of Apply::ell, same chunk layout over rows permuted inside windows
 * of sigma rows): position p holds row perm[p] with true length rl[p]. */
int lafem_hip_apply_sell_f64_u32(void* stream, double* r, double a, const double* x, double b, const double* y,
                                 const double* val, const uint32_t* col_ind, const uint32_t* cs, const uint32_t* cl,
                                 const uint32_t* rl, const uint32_t* perm, uint64_t C, uint64_t rows);
int lafem_hip_apply_sell_f64_u64(void* stream, double* r, double a, const double* x, double b, const double* y,
                                 const double* val, const uint64_t* col_ind, const uint64_t* cs, const uint64_t* cl,
                                 const uint64_t* rl, const uint64_t* perm, uint64_t C, uint64_t rows);
int lafem_hip_apply_sell_f32_u32(void* stream, float* r, float a, const float* x, float b, const float* y,
                                 const float* val, const uint32_t* col_ind, const uint32_t* cs, const uint32_t* cl,
                                 const uint32_t* rl, const uint32_t* perm, uint64_t C, uint64_t rows);
int lafem_hip_apply_sell_f32_u64(void* stream, float* r, float a, const float* x, float b, const float* y,
                                 const float* val, const uint64_t* col_ind, const uint64_t* cs, const uint64_t* cl,
                                 const uint64_t* rl, const uint64_t* perm, uint64_t C, uint64_t rows);
int lafem_hip_apply_sell_h16f32_u32(void* stream, float* r, float a, const float* x, float b, const float* y,
                                    const void* val, const uint32_t* col_ind, const uint32_t* cs, const uint32_t* cl,
                                    const uint32_t* rl, const uint32_t* perm, uint64_t C, uint64_t rows);

/* Apply::banded (apply.hpp:162,299; generic apply_generic.hpp:519-598): SparseMatrixBanded
 * layout, val[d*rows + row] on diagonal d whose column is row + offsets[d] + 1 - rows;
 * diagonals with a column outside [0, columns) do not contribute; r = b*y + a*A*x. */
int lafem_hip_apply_banded_f64_u32(void* stream, double* r, double a, const double* x, double b, const double* y,
                                   const double* val, const uint32_t* offsets, uint64_t num_of_offsets, uint64_t rows,
                                   uint64_t columns);
int lafem_hip_apply_banded_f64_u64(void* stream, double* r, double a, const double* x, double b, const double* y,
                                   const double* val, const uint64_t* offsets, uint64_t num_of_offsets, uint64_t rows,
                                   uint64_t columns);
int lafem_hip_apply_banded_f32_u32(void* stream, float* r, float a, const float* x, float b, const float* y,
                                   const float* val, const uint32_t* offsets, uint64_t num_of_offsets, uint64_t rows,
                                   uint64_t columns);
int lafem_hip_apply_banded_f32_u64(void* stream, float* r, float a, const float* x, float b, const float* y,
                                   const float* val, const uint64_t* offsets, uint64_t num_of_offsets, uint64_t rows,
                                   uint64_t columns);

/* Apply::csrb (apply.hpp:277-281, CUDA apply.cu:203-231): BCSR with row-major blocks of
 * block_height x block_width, each in 1..9 (the reference's CUDA kernel takes BlockHeight < 10);
 * square blocks of size 1..4 run shape-specialised kernels, the other shapes a generic one. */
int lafem_hip_apply_csrb_f64_u32(void* stream, double* r, double a, const double* x, double b, const double* y,
                                 const double* val, const uint32_t* col_ind, const uint32_t* row_ptr, uint64_t rows,
                                 uint64_t columns, uint64_t used_elements, int block_height, int block_width);
int lafem_hip_apply_csrb_f64_u64(void* stream, double* r, double a, const double* x, double b, const double* y,
                                 const double* val, const uint64_t* col_ind, const uint64_t* row_ptr, uint64_t rows,
                                 uint64_t columns, uint64_t used_elements, int block_height, int block_width);
int lafem_hip_apply_csrb_f32_u32(void* stream, float* r, float a, const float* x, float b, const float* y,
                                 const float* val, const uint32_t* col_ind, const uint32_t* row_ptr, uint64_t rows,
                                 uint64_t columns, uint64_t used_elements, int block_height, int block_width);
int lafem_hip_apply_csrb_f32_u64(void* stream, float* r, float a, const float* x, float b, const float* y,
                                 const float* val, const uint64_t* col_ind, const uint64_t* row_ptr, uint64_t rows,
                                 uint64_t columns, uint64_t used_elements, int block_height, int block_width);

/* Apply::csrsb (apply.hpp:283, CUDA apply.cu:174-200): scalar CSR times a vector of
 * block_size-blocks (block_size in {1,2,3}). */
int lafem_hip_apply_csrsb_f64_u32(void* stream, double* r, double a, const double* x, double b, const double* y,
                                  const double* val, const uint32_t* col_ind, const uint32_t* row_ptr, uint64_t rows,
                                  uint64_t columns, uint64_t used_elements, int block_size);
int lafem_hip_apply_csrsb_f64_u64(void* stream, double* r, double a, const double* x, double b, const double* y,
                                  const double* val, const uint64_t* col_ind, const uint64_t* row_ptr, uint64_t rows,
                                  uint64_t columns, uint64_t used_elements, int block_size);
int lafem_hip_apply_csrsb_f32_u32(void* stream, float* r, float a, const float* x, float b, const float* y,
                                  const float* val, const uint32_t* col_ind, const uint32_t* row_ptr, uint64_t rows,
                                  uint64_t columns, uint64_t used_elements, int block_size);
int lafem_hip_apply_csrsb_f32_u64(void* stream, float* r, float a, const float* x, float b, const float* y,
                                  const float* val, const uint64_t* col_ind, const uint64_t* row_ptr, uint64_t rows,
                                  uint64_t columns, uint64_t used_elements, int block_size);

/* ===================================================================================== */
/* BLAS-1: Axpy::dv (axpy.hpp), DotProduct/Norm2::value (dot_product.hpp, norm.hpp),      */
/* ComponentProduct/ComponentInvert/Scale::value (component_*.hpp, scale.hpp)             */
/* ===================================================================================== */

/* r = a*x + y (r may alias x or y) */
int lafem_hip_axpy_f64(void* stream, double* r, double a, const double* x, const double* y, uint64_t n);
int lafem_hip_axpy_f32(void* stream, float* r, float a, const float* x, const float* y, uint64_t n);
/* <x,y> into a device scalar (deterministic order), or returned to the host (synchronises) */
int lafem_hip_dot_f64(void* stream, const double* x, const double* y, uint64_t n, double* result_dev);
int lafem_hip_dot_f32(void* stream, const float* x, const float* y, uint64_t n, float* result_dev);
int lafem_hip_dot_value_f64(void* stream, const double* x, const double* y, uint64_t n, double* result_host);
int lafem_hip_dot_value_f32(void* stream, const float* x, const float* y, uint64_t n, float* result_host);
/* <x,y> into a device scalar without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same kernels, grid and summation order as
 * lafem_hip_dot_f64, so the same result bit for bit; the block partials go to a static device
 * array of 2048 doubles (16 KiB per GPU, held until the process ends) instead of the
 * initialised workspace.  Every n is accepted: the grid has at most 2048 blocks, one partial
 * each.  Nothing needs clearing between calls (each call writes every partial it reads).  All
 * calls on one GPU share the array: calls that overlap on one GPU (on different streams, or
 * from several threads) race and must be ordered by the caller.  feat3_hip_initialise looks up
 * the array's device address; without it the first call on a device does. */
int lafem_hip_dot_noalloc_f64(void* stream, const double* x, const double* y, uint64_t n, double* result_dev);
/* ||x||_2 into a device scalar, or returned to the host (synchronises) */
int lafem_hip_norm2_f64(void* stream, const double* x, uint64_t n, double* result_dev);
int lafem_hip_norm2_f32(void* stream, const float* x, uint64_t n, float* result_dev);
int lafem_hip_norm2_value_f64(void* stream, const double* x, uint64_t n, double* result_host);
int lafem_hip_norm2_value_f32(void* stream, const float* x, uint64_t n, float* result_host);
/* r = x .* y */
int lafem_hip_component_product_f64(void* stream, double* r, const double* x, const double* y, uint64_t n);
int lafem_hip_component_product_f32(void* stream, float* r, const float* x, const float* y, uint64_t n);
/* r = s ./ x */
int lafem_hip_component_invert_f64(void* stream, double* r, const double* x, double s, uint64_t n);
int lafem_hip_component_invert_f32(void* stream, float* r, const float* x, float s, uint64_t n);
/* r = s * x */
int lafem_hip_scale_f64(void* stream, double* r, const double* x, double s, uint64_t n);
int lafem_hip_scale_f32(void* stream, float* r, const float* x, float s, uint64_t n);

/* ===================================================================================== */
/* Diagonal::{csr,ell,csrb} (diagonal.hpp; CUDA diagonal.cu:20-161)                       */
/* ===================================================================================== */
/* a missing diagonal entry reads 0; csrb: the diagonals of the square diagonal blocks
 * (block_height == block_width, 1..16) */
int lafem_hip_diagonal_csr_f64_u32(void* stream, double* diag, const double* val, const uint32_t* col_ind,
                                   const uint32_t* row_ptr, uint64_t rows);
int lafem_hip_diagonal_csr_f64_u64(void* stream, double* diag, const double* val, const uint64_t* col_ind,
                                   const uint64_t* row_ptr, uint64_t rows);
int lafem_hip_diagonal_csr_f32_u32(void* stream, float* diag, const float* val, const uint32_t* col_ind,
                                   const uint32_t* row_ptr, uint64_t rows);
int lafem_hip_diagonal_csr_f32_u64(void* stream, float* diag, const float* val, const uint64_t* col_ind,
                                   const uint64_t* row_ptr, uint64_t rows);
int lafem_hip_diagonal_ell_f64_u32(void* stream, double* diag, const double* val, const uint32_t* col_ind,
                                   const uint32_t* cs, const uint32_t* cl, uint64_t C, uint64_t rows);
int lafem_hip_diagonal_ell_f64_u64(void* stream, double* diag, const double* val, const uint64_t* col_ind,
                                   const uint64_t* cs, const uint64_t* cl, uint64_t C, uint64_t rows);
int lafem_hip_diagonal_ell_f32_u32(void* stream, float* diag, const float* val, const uint32_t* col_ind,
                                   const uint32_t* cs, const uint32_t* cl, uint64_t C, uint64_t rows);
int lafem_hip_diagonal_ell_f32_u64(void* stream, float* diag, const float* val, const uint64_t* col_ind,
                                   const uint64_t* cs, const uint64_t* cl, uint64_t C, uint64_t rows);
int lafem_hip_diagonal_csrb_f64_u32(void* stream, double* diag, const double* val, const uint32_t* col_ind,
                                    const uint32_t* row_ptr, uint64_t rows, int block_height, int block_width);
int lafem_hip_diagonal_csrb_f64_u64(void* stream, double* diag, const double* val, const uint64_t* col_ind,
                                    const uint64_t* row_ptr, uint64_t rows, int block_height, int block_width);
int lafem_hip_diagonal_csrb_f32_u32(void* stream, float* diag, const float* val, const uint32_t* col_ind,
                                    const uint32_t* row_ptr, uint64_t rows, int block_height, int block_width);
int lafem_hip_diagonal_csrb_f32_u64(void* stream, float* diag, const float* val, const uint64_t* col_ind,
                                    const uint64_t* row_ptr, uint64_t rows, int block_height, int block_width);

/* ===================================================================================== */
/* UnitFilter::{filter_rhs,filter_def} and the blocked variants (unit_filter.hpp,         */
/* unit_filter_blocked.hpp; CUDA unit_filter.cu:21-78, unit_filter_blocked.cu:21-96);      */
/* block_size 1 is the scalar filter.  filter_sol = filter_rhs, filter_cor = filter_def.  */
/* ===================================================================================== */
int lafem_hip_unit_filter_rhs_f64_u32(void* stream, double* v, const double* sv_elements, const uint32_t* sv_indices,
                                      uint64_t ue, int block_size);
int lafem_hip_unit_filter_rhs_f64_u64(void* stream, double* v, const double* sv_elements, const uint64_t* sv_indices,
                                      uint64_t ue, int block_size);
int lafem_hip_unit_filter_rhs_f32_u32(void* stream, float* v, const float* sv_elements, const uint32_t* sv_indices,
                                      uint64_t ue, int block_size);
int lafem_hip_unit_filter_rhs_f32_u64(void* stream, float* v, const float* sv_elements, const uint64_t* sv_indices,
                                      uint64_t ue, int block_size);
int lafem_hip_unit_filter_def_f64_u32(void* stream, double* v, const uint32_t* sv_indices, uint64_t ue, int block_size);
int lafem_hip_unit_filter_def_f64_u64(void* stream, double* v, const uint64_t* sv_indices, uint64_t ue, int block_size);
int lafem_hip_unit_filter_def_f32_u32(void* stream, float* v, const uint32_t* sv_indices, uint64_t ue, int block_size);
int lafem_hip_unit_filter_def_f32_u64(void* stream, float* v, const uint64_t* sv_indices, uint64_t ue, int block_size);

/* ===================================================================================== */
/* MemoryPool<Mem::HIP>::set_memory / convert (memory_pool.cu:26-41,176-227)              */
/* ===================================================================================== */
int lafem_hip_set_memory_f64(void* stream, double* dst, double value, uint64_t n);
int lafem_hip_set_memory_f32(void* stream, float* dst, float value, uint64_t n);
int lafem_hip_set_memory_u32(void* stream, uint32_t* dst, uint32_t value, uint64_t n);
int lafem_hip_set_memory_u64(void* stream, uint64_t* dst, uint64_t value, uint64_t n);
int lafem_hip_convert_f32_f64(void* stream, float* dst, const double* src, uint64_t n);
int lafem_hip_convert_f64_f32(void* stream, double* dst, const float* src, uint64_t n);
int lafem_hip_convert_u32_u64(void* stream, uint32_t* dst, const uint64_t* src, uint64_t n);
int lafem_hip_convert_u64_u32(void* stream, uint64_t* dst, const uint32_t* src, uint64_t n);
/* binary16 <-> FP32 (round to nearest even) */
int lafem_hip_convert_h16_f32(void* stream, void* dst, const float* src, uint64_t n);
int lafem_hip_convert_f32_h16(void* stream, float* dst, const void* src, uint64_t n);

/* ===================================================================================== */
/* MemoryPool<Mem::HIP> / Runtime (memory_pool.hpp:53-377, runtime.cpp:33-144)             */
/* ===================================================================================== */
/* initialise: select the device and create its reduction workspace (memory_pool.cu:56-72);
 * required once per device before the dot/norm entry points */
int feat3_hip_initialise(int device);
/* finalise: reports the number of allocations still held (memory_pool.cu:74-95) */
int feat3_hip_finalise(uint64_t* leaked);
int feat3_hip_device_count(int* n);
/* allocate_memory / increase_memory / release_memory (refcounted, memory_pool.cu:97-149) */
int feat3_hip_malloc(void** p, uint64_t bytes);
int feat3_hip_increase(void* p);
int feat3_hip_free(void* p);
/* upload / download / copy (memory_pool.cu:158-175), asynchronous on stream */
int feat3_hip_upload(void* stream, void* dst, const void* src, uint64_t bytes);
int feat3_hip_download(void* stream, void* dst, const void* src, uint64_t bytes);
int feat3_hip_copy(void* stream, void* dst, const void* src, uint64_t bytes);
int feat3_hip_synchronise(void* stream);
/* HBM read-rate calibration, no counterpart in the reference (for measuring the other kernels
 * against what the GPU sustains): reads `bytes` (a multiple of 16, 16-byte aligned) once with
 * nontemporal 16-byte loads; `out` is a device double that finite data never causes to be written */
int feat3_hip_read_probe(void* stream, const void* p, uint64_t bytes, double* out);
/* lanes per row of the CSR/BCSR apply kernels (0 = by mean row length); replaces the block
 * size globals that lafem/tuning.hpp:72-163 sets */
int feat3_hip_set_csr_lanes(int lanes);

/* ===================================================================================== */
/* Fused device solver: PCG( MultiGrid-V( Richardson(omega_R, Jacobi(omega_J)) ) ) with   */
/* an exact FP64 coarse solve.  Lifecycle of SolverBase (solver/base.hpp:184-316):        */
/*   create -> push_level (finest first) -> set_* -> init (init_symbolic+init_numeric)    */
/*   -> correct / apply (any number) -> destroy (done).                                   */
/* Reference objects replaced: PCG (solver/pcg.hpp:33-328), MultiGrid + MultiGridHierarchy */
/* (solver/multigrid.hpp:419-2035), Richardson (richardson.hpp:32-230), JacobiPrecond      */
/* (jacobi_precond.hpp:36-213), ConvertPrecond (convert_precond.hpp:41-237).               */
/* ===================================================================================== */
typedef struct feat3_mgpcg feat3_mgpcg;

/* inner_precision: 0 FP64, 1 FP32 inside FP64 PCG, 2 binary16 matrix values with FP32
 * arithmetic inside FP64 PCG; format: 0 CSR, 1 SELL-C-sigma for the level matrices, 2 banded
 * (SparseMatrixBanded layout, no column indices) for every square matrix with at most 64
 * distinct diagonals and at most 25% zero fill, SELL-C-sigma for the others.  The inner-precision
 * copies (1 and 2) of the level matrices, in every format, leave out the entries with
 * |a_ij| <= eps_f32^2 |a_ii|, which FP32 sums cannot resolve; FP64 copies keep every nonzero
 * (the banded FP64 layout has no slot for a stored zero).  NaN values are kept in every copy. */
int feat3_mgpcg_create(feat3_mgpcg** h, int inner_precision, int format);
int feat3_mgpcg_destroy(feat3_mgpcg* h);
/* one level (finest first), host arrays, copied: CSR matrix (n x n, u32 indices, FP64
 * values), Dirichlet index list, and for all but the coarsest level the prolongation
 * P (n x nc) and restriction R = P^T (nc x n) to the next coarser level
 * (MultiGridHierarchy::push_level, multigrid.hpp:612-700) */
int feat3_mgpcg_push_level(feat3_mgpcg* h, uint64_t n, const uint32_t* row_ptr, const uint32_t* col_ind,
                           const double* val, uint64_t nfilt, const uint32_t* filter_idx, uint64_t nc,
                           const uint32_t* p_row_ptr, const uint32_t* p_col_ind, const double* p_val,
                           const uint32_t* r_row_ptr, const uint32_t* r_col_ind, const double* r_val);
/* the same with device arrays (e.g. from feat3_device_laplace / feat3_device_transfer), valid until
 * feat3_mgpcg_init returns: init validates them on the device (rows ascending, columns strictly
 * ascending and in range) and builds the level formats there; levels of at most 65536 rows are
 * copied to the host at this call.  No reference counterpart (device-resident setup). */
int feat3_mgpcg_push_level_device(feat3_mgpcg* h, uint64_t n, const uint32_t* row_ptr, const uint32_t* col_ind,
                                  const double* val, uint64_t nfilt, const uint32_t* filter_idx, uint64_t nc,
                                  const uint32_t* p_row_ptr, const uint32_t* p_col_ind, const double* p_val,
                                  const uint32_t* r_row_ptr, const uint32_t* r_col_ind, const double* r_val);
/* The solver's internal order of the DOFs of one level (0 = finest; no reference counterpart):
 * order[i] = internal position of DOF i, a permutation of [0, n); host array, copied; set after
 * that level's push_level and before init, at most once per level.  At init the level matrix and
 * Dirichlet list and the transfers on both sides of the level are renumbered (host or device
 * arrays alike); correct / apply / vcycle still take and return vectors in the caller's
 * numbering (level 0's order is applied to b, x and def at entry and to x and cor at exit).
 * For a structured mesh in FEAT's hierarchical numbering, the lexicographic lattice order
 * (feat3_setup_lattice_order of the DOF nodes' coordinates) makes the level matrices banded. */
int feat3_mgpcg_set_level_order(feat3_mgpcg* h, int level, const uint32_t* order);
/* Jacobi damping, Richardson damping, pre/post smoothing steps (>= 1) */
int feat3_mgpcg_set_smoother(feat3_mgpcg* h, double omega_jacobi, double omega_richardson, int pre, int post);
/* IterativeSolver::set_tol_rel/set_tol_abs/set_min_iter/set_max_iter (iterative.hpp);
 * tol_abs <= 0 keeps the default 1/eps^2 */
int feat3_mgpcg_set_tolerances(feat3_mgpcg* h, double tol_rel, double tol_abs, double tol_abs_low, uint64_t min_iter,
                               uint64_t max_iter);
/* 0: multigrid preconditioner (V-cycle unless the option "cycle" says otherwise), 1: none (plain CG) */
int feat3_mgpcg_set_precond(feat3_mgpcg* h, int kind);
/* graph replay on/off, lanes per CSR row (0 auto), SELL chunk height and sort window */
int feat3_mgpcg_set_launch(feat3_mgpcg* h, int graph, int lanes, uint64_t sell_C, uint64_t sell_sigma);
/* named launch/format options (replaces the globals of lafem/tuning.hpp:72-163 that the
 * reference's block-size tuner sets): "index16" 0/1 (SELL columns as 16-bit row offsets where
 * they fit), "compact_transfer" 0/1 (device P/R without stored structural zeros), "sell_C",
 * "sell_sigma", "lanes", "tail_max_n" (levels of at most this many rows run in one LDS-resident
 * workgroup; 0 = none; default 1100), "rows_per_block" (row positions per workgroup of the
 * level-matrix kernels, a multiple of 64; 0 = derived from the row count),
 * "outer_rows_per_block" (the same for the outer FP64 matrix when it is a copy of its own),
 * "transfer_csr" (0/1:
 * transfer matrices in CSR evaluated by streaming, default 1, or in the level format),
 * "transfer_sigma" (SELL sort window of the transfer matrices when they are SELL, default 1),
 * "tensor_transfer" (1, default: transfer pairs of a structured hierarchy in lattice order whose
 * every stored entry equals the product of 1D factors are evaluated from those factors, with
 * the same sums; 0: always from the stored matrices),
 * "transfer_lanes" (lanes per row of CSR transfer matrices: -1 as "lanes", default; 0 streaming),
 * "transfer_rows_per_block" (row positions per workgroup of the transfer kernels, default 128;
 * 0 = derived from the row count), "symmetric_band" (1, default: an FP64 band equal to its mirror
 * image bit for bit stores only its center and upper diagonals; 0: the full band; same results),
 * "cycle" (the multigrid cycle, MultiGridCycle of multigrid.hpp:26-34: 0 V, default, 1 F,
 * 2 W; the coarse-tail kernel runs the cycle's steps on the small levels for each of them);
 * "plane_order" (order in which the blocks of a banded row kernel take their row ranges: 0
 * ascending, 1 plane strips for the FP64 outer matrix (default), 2 for every banded matrix);
 * "band_stage_x" (x of the banded row kernels staged in LDS per round of 256 rows, the columns
 * the round reaches copied once instead of gathered per entry: 0 no (default), 1 the FP64 outer
 * matrix, 2 every banded matrix; the same row sums, slower on gfx950; not used with
 * value-coded bands); "band_dict" (1, default: a band whose every diagonal holds at most 16
 * distinct values, bit patterns, is stored as 4-bit codes per row and diagonal into per-diagonal
 * tables of those values, which the row kernels read from LDS: the same values and row sums,
 * bit for bit, with 0.5 instead of 4 (FP32) or 8 (FP64) bytes per entry; 0: the values);
 * "lane_shift" (1, default: FP64 value-coded rows of 27 diagonals in nine runs of three
 * consecutive offsets load one x per run and take the other two from the neighbouring lanes,
 * the same sums bit for bit; 0: three loads); takes effect at the next init */
int feat3_mgpcg_set_option(feat3_mgpcg* h, const char* name, double value);
/* init_symbolic + init_numeric (solver/base.hpp:227-260) */
int feat3_mgpcg_init(feat3_mgpcg* h);
/* Solver::solve / IterativeSolver::correct (iterative.hpp:600,926-935): x, b FP64 device
 * vectors; caller_stream (NULL: the null stream) is synchronised first, since the solver runs
 * on a stream of its own; blocks until the solve has ended */
int feat3_mgpcg_correct(feat3_mgpcg* h, void* caller_stream, double* x, const double* b, int* status, uint64_t* iters);
/* SolverBase::apply(cor, def) (pcg.hpp:141-156) */
int feat3_mgpcg_apply(feat3_mgpcg* h, void* caller_stream, double* cor, const double* def, int* status, uint64_t* iters);
/* the MultiGrid preconditioner alone: cor = one cycle (V, F or W, option "cycle") applied to
 * def (multigrid.hpp:1268-1519) */
int feat3_mgpcg_vcycle(feat3_mgpcg* h, void* caller_stream, double* cor, const double* def);
/* defect norms of the last solve: [0] initial, [k] after iteration k */
int feat3_mgpcg_history(feat3_mgpcg* h, double* out_host, uint64_t cap, uint64_t* len);
/* average duration of `reps` launches of one hot-path kernel on the solver's stream
 * (which: 0 finest smoothing sweep, 1 finest defect, 2 outer FP64 product q = A p with its dot
 * partials, 3 full cycle) and its algorithmic bytes */
int feat3_mgpcg_time_kernel(feat3_mgpcg* h, int which, int reps, double* avg_ms, uint64_t* bytes);
/* [levels, n, nnz, lanes, sell size, kernels per cycle, cycle bytes, outer bytes,
 *  sizeof inner, sizeof stored value, 16-bit SELL columns on the finest level, SELL chunks
 *  of the finest level, finest level banded (0 no, 1 yes, 2 with symmetric storage: center and
 *  upper diagonals only, 3 value-coded), its number of diagonals, levels run by the
 *  single-workgroup coarse tail (0: none), LDS bytes of the tail, levels whose transfers are
 *  evaluated as tensor products of their 1D factors] */
int feat3_mgpcg_info(feat3_mgpcg* h, uint64_t* out, int cap);
/* Per-level statistics of the multigrid cycle (MultiGridHierarchy::get_time_smooth/_defect/
 * _transfer/_coarse, multigrid.hpp:858-956, and the flop counter of Statistics,
 * util/statistics.hpp:206-230): `reps` cycles launched operation by operation on the solver's
 * stream with an event wherever the level or the category changes (the solve's own graphs are
 * untouched).  out[0] levels L, out[1] ms per cycle (first to last event), out[2] operations
 * per cycle; then for each level
 * (finest first) 10 values: ms per cycle in smoothing, defect, grid transfer (restriction and
 * prolongation, on the finer level as the reference counts them), coarse solve; the number of
 * smoothing steps, defects, transfers and coarse solves per cycle; algorithmic bytes and flops
 * per cycle.  A segment of the single-workgroup coarse tail counts as the coarse solve of its
 * top level (its levels report counts, bytes and flops; their time is the segment's).
 * cap >= 3 + 10 L.  Creates its events at the first call. */
int feat3_mgpcg_statistics(feat3_mgpcg* h, int reps, double* out, int cap);

/* ===================================================================================== */
/* Batched fused solver: K systems that share one sparsity graph, solved in lockstep by   */
/* the same PCG(MG-V(Richardson(Jacobi))) as feat3_mgpcg; every system keeps its own PCG  */
/* scalars, status and history and performs exactly the iterations of its own solve.      */
/*  - shared values, block 1 or 3: K right-hand sides of one matrix (block 3 = the 3x3     */
/*    BCSR vector Laplacian, SparseMatrixBCSR<3,3>, sparse_matrix_bcsr.hpp:1295-1600; the */
/*    reference applies it once per right-hand side, apply.hpp:71-117)                     */
/*  - per-system values, block 1: K coefficient fields on one mesh                         */
/* The reference has no multi-system solver; the per-system semantics are those of        */
/* PCG (pcg.hpp:188-288) with MultiGrid (multigrid.hpp:1268-2000) and UnitFilterBlocked.   */
/* Vectors are [node][component][system] (system fastest), FP64, on the device.           */
/* ===================================================================================== */
typedef struct feat3_batch feat3_batch;

/* inner_precision as feat3_mgpcg_create; block 1 or 3 (3 needs shared_values);
 * 1 <= nsys <= 64 */
int feat3_batch_create(feat3_batch** h, int inner_precision, int block, int nsys, int shared_values);
int feat3_batch_destroy(feat3_batch* h);
/* one level (finest first), host arrays, copied: node graph (u32 CSR), values
 * [entry][bh][bw] (shared) or [entry][system] (per-system, block 1); Dirichlet node list
 * (all components filtered); scalar P (nodes x nc) and R = P^T applied per component */
int feat3_batch_push_level(feat3_batch* h, uint64_t nodes, const uint32_t* row_ptr, const uint32_t* col_ind,
                           const double* val, uint64_t nfilt, const uint32_t* filter_nodes, uint64_t nc,
                           const uint32_t* p_row_ptr, const uint32_t* p_col_ind, const double* p_val,
                           const uint32_t* r_row_ptr, const uint32_t* r_col_ind, const double* r_val);
int feat3_batch_set_smoother(feat3_batch* h, double omega_jacobi, double omega_richardson, int pre, int post);
int feat3_batch_set_tolerances(feat3_batch* h, double tol_rel, double tol_abs, double tol_abs_low, uint64_t min_iter,
                               uint64_t max_iter);
/* graph replay of one PCG iteration on/off */
int feat3_batch_set_graph(feat3_batch* h, int graph);
/* named options, applied at the next init: "tile_rows" (rows per locality tile of the row
 * kernels, feat3_setup_tile_order; default 0 = ascending row order), "diag_blocks" (1, default:
 * shared 3x3 values whose every block has exact zeros off its diagonal, a vector Laplacian's, are
 * stored as block diagonals on the device, the FP64 outer copy included; such a solver then
 * refuses feat3_batch_set_values; 0: full blocks), "tail_max_n" (shared values: the deepest
 * levels of at most this many nodes each, the coarsest included, run as one LDS-resident
 * workgroup per system, bit for bit the per-level launches; default 1000, 0 = off) */
int feat3_batch_set_option(feat3_batch* h, const char* name, double value);
int feat3_batch_init(feat3_batch* h);
/* x, b: FP64 device arrays [node][component][system], 16-byte aligned; status, iters: host
 * arrays of nsys */
int feat3_batch_correct(feat3_batch* h, void* caller_stream, double* x, const double* b, int* status, uint64_t* iters);
/* replaces the values of one level (0 = finest) from a device array in the layout of push_level
 * (FP64; [entry][h][w] or [entry][system]): converts to the storage precision, recomputes the
 * Jacobi diagonal, the FP64 outer copy (level 0) and the coarse inverses (coarsest level; a
 * small device-to-host copy); graph, transfers and Dirichlet nodes are kept.  The null stream is
 * synchronised first (values produced on another stream must be complete); blocks until done.
 * Refused (invalid object state) for a level whose FP32 inner graph was thinned at init (shared
 * values: blocks below eps_f32^2 of the diagonal left out, chosen for the values given then). */
int feat3_batch_set_values(feat3_batch* h, int level, const double* values);
/* defect norms of system `system` in the last solve */
int feat3_batch_history(feat3_batch* h, int system, double* out_host, uint64_t cap, uint64_t* len);
/* which: 0 finest smoothing sweep (all systems), 2 outer FP64 q = A p with q.p partials,
 * 3 one V-cycle; algorithmic bytes of one launch (0, 2) */
int feat3_batch_time_kernel(feat3_batch* h, int which, int reps, double* avg_ms, uint64_t* bytes);

/* ===================================================================================== */
/* Adjoint gradient of the variable-coefficient diffusion solve (no counterpart in the    */
/* reference, which has neither a coefficient operator nor an adjoint): for A(k) =        */
/* sum_e k_e E (E = h^(d-2) K_ref, the element stiffness of one cell), A u = b and          */
/* A lambda = dL/du, grad[e] = dL/dk_e = -lambda_e^T E u_e for nsys systems at once.        */
/* ===================================================================================== */
/* dim 2 (degree 1, 2) or 3 (degree 1); level: 2^level cells per direction; elem_matrix:   */
/* host array, nloc x nloc row-major, lexicographic local nodes; u, lam: device FP64       */
/* [node][system]; grad: device FP64 [cell][system], cells x fastest; enqueued on stream   */
int feat3_adjoint_cell_gradient(void* stream, int dim, int degree, int level, int nsys, const double* elem_matrix,
                                const double* u, const double* lam, double* grad);

/* ===================================================================================== */
/* Device-side assembly (the step before the hot path, for many coefficient fields): the   */
/* coefficient-weighted Laplace values of feat3_setup_laplace, bit-identical, and the      */
/* coarsening of cell fields of feat3_setup_coarsen_cellfield                              */
/* ===================================================================================== */
/* dim 2 (degree 1, 2) or 3 (degree 1); ref_stiffness: host array (feat3_setup_ref_stiffness);
 * kfields: device [field][cell] (cells x fastest); row_ptr/col_ind: device graph of
 * feat3_setup_laplace on this level (n rows); val: device [entry][field] */
int feat3_assemble_laplace_fields(void* stream, int dim, int degree, int level, int nfields, const double* ref_stiffness,
                                  const double* kfields, const uint32_t* row_ptr, const uint32_t* col_ind, uint64_t n,
                                  double* val);
/* kc[field][coarse cell] = mean of the 2^dim children on level level_fine (device arrays) */
int feat3_coarsen_cellfields(void* stream, int dim, int level_fine, int nfields, const double* kf, double* kc);

/* Symbolic and numeric assembly of a level on the device (the host setup's feat3_setup_laplace,
 * feat3_setup_prol and feat3_setup_transpose, bit for bit: the reference's
 * SymbolicAssembler::assemble_matrix_std1 / assemble_matrix_2lvl, symbolic_assembler.hpp:59-110,
 * 319-343, BilinearOperatorAssembler with the Laplace operator, UnitFilterAssembler,
 * unit_filter_assembler.hpp:101-130, and the transposed prolongation,
 * sparse_matrix_csr.hpp:1781-1844).  Sizes are closed forms computed on the host. */
int feat3_device_laplace_sizes(int dim, int degree, int level, uint64_t* n, uint64_t* nnz, uint64_t* nbnd);
/* row_ptr[n+1], col_ind[nnz], val[nnz], bnd[nbnd]: device; ref_stiffness: host
 * (feat3_setup_ref_stiffness); kcell: device coefficient per cell (x fastest) or NULL for k = 1 */
int feat3_device_laplace(void* stream, int dim, int degree, int level, const double* ref_stiffness, const double* kcell,
                         uint32_t* row_ptr, uint32_t* col_ind, double* val, uint32_t* bnd);
/* prolongation level_fine-1 -> level_fine: fine rows nf, coarse columns nc, entries nnz (R = P^T
 * has nc rows and the same entries) */
int feat3_device_prol_sizes(int dim, int degree, int level_fine, uint64_t* nf, uint64_t* nc, uint64_t* nnz);
/* transpose = 0: P (row_ptr[nf+1]); transpose = 1: R = P^T (row_ptr[nc+1]); device arrays */
int feat3_device_transfer(void* stream, int dim, int degree, int level_fine, int transpose, uint32_t* row_ptr,
                          uint32_t* col_ind, double* val);

/* ===================================================================================== */
/* Host-side problem setup (the step before the hot path: assembly, geometry)         */
/* ===================================================================================== */
int feat3_setup_laplace_sizes(int dim, int degree, int level, uint64_t* n, uint64_t* nnz, uint64_t* nbnd);
int feat3_setup_laplace(int dim, int degree, int level, const double* kcell, uint32_t* row_ptr, uint32_t* col_ind,
                        double* val, uint32_t* bnd);
int feat3_setup_rhs(int dim, int degree, int level, int func, double* rhs);
int feat3_setup_prol_sizes(int dim, int degree, int level_fine, uint64_t* nf, uint64_t* nc, uint64_t* nnz);
int feat3_setup_prol(int dim, int degree, int level_fine, uint32_t* row_ptr, uint32_t* col_ind, double* val);
int feat3_setup_transpose(uint64_t rows, uint64_t cols, const uint32_t* row_ptr, const uint32_t* col_ind,
                          const double* val, uint32_t* t_row_ptr, uint32_t* t_col_ind, double* t_val);
int feat3_setup_coarsen_cellfield(int dim, int level_fine, const double* kf, double* kc);

/* The reference's on-disk formats of SparseMatrixCSR (host, u32 indices): Matrix Market
 * coordinate real general/symmetric (sparse_matrix_csr.hpp:1125-1240, 1296-1340) and the binary
 * fm_csr container (Container::_serialise<double, uint64_t>, container.hpp:320-447).  *_sizes
 * reads the dimensions, read fills caller arrays (row_ptr rows + 1, col_ind and val nnz) and
 * fails unless the file still has those rows and nnz; header words and offsets of a binary file
 * are checked before they are used (a malformed file is an error, never an out-of-range read). */
int feat3_io_mtx_sizes(const char* path, uint64_t* rows, uint64_t* cols, uint64_t* nnz);
int feat3_io_read_mtx(const char* path, uint64_t rows, uint64_t nnz, uint32_t* row_ptr, uint32_t* col_ind, double* val);
int feat3_io_write_mtx(const char* path, uint64_t rows, uint64_t cols, const uint32_t* row_ptr, const uint32_t* col_ind,
                       const double* val, int symmetric);
int feat3_io_csr_binary_sizes(const char* path, uint64_t* rows, uint64_t* cols, uint64_t* nnz);
int feat3_io_read_csr_binary(const char* path, uint64_t rows, uint64_t nnz, uint32_t* row_ptr, uint32_t* col_ind,
                             double* val);
int feat3_io_write_csr_binary(const char* path, uint64_t rows, uint64_t cols, const uint32_t* row_ptr,
                              const uint32_t* col_ind, const double* val);
/* DenseVector files (DenseVector::read_from / write_out, dense_vector.hpp:587-780), host memory;
 * mode = the FileMode value: 0 fm_exp (one std::scientific value per line, '#' lines skipped),
 * 1 fm_dv (binary Container::_serialise<double, uint64_t>, container.hpp:320-447), 2 fm_mtx
 * (Matrix Market "array real general", n x 1).  Text files are written with 6 significant
 * digits, as the reference writes them.  Return 0, or 1 for an unreadable or malformed file. */
int feat3_io_vector_size(const char* path, int mode, uint64_t* n);
int feat3_io_read_vector(const char* path, int mode, double* out, uint64_t n);
int feat3_io_write_vector(const char* path, int mode, uint64_t n, const double* v);

/* CSR -> SELL-C-sigma on the host (generalises SparseMatrixELL::convert,
 * lafem/sparse_matrix_ell.hpp:865-960): rows sorted by length inside windows of sigma rows,
 * perm[p] = row at position p, padding slots repeat a column of their row with value 0, and
 * optional 16-bit offsets d16 = col - row (*used_d16 = 1 when all fit).  Host arrays. */
int feat3_sell_sizes(uint64_t rows, const uint32_t* row_ptr, uint64_t C, uint64_t sigma, uint64_t* nchunks,
                     uint64_t* val_size);
int feat3_sell_build(uint64_t rows, uint64_t cols, const uint32_t* row_ptr, const uint32_t* col_ind, const double* val,
                     uint64_t C, uint64_t sigma, uint32_t* cs, uint32_t* cl, uint32_t* perm, uint32_t* col, double* sval,
                     int16_t* d16, int* used_d16);
int feat3_setup_ref_stiffness(int dim, int degree, double* K);
/* a locality order of the rows of a graph: compact tiles of about `tile` rows grown by
 * breadth-first search from the lowest unvisited row, ascending inside a tile; tile <= 1 gives
 * the identity (no reference counterpart: a processing order for the batched row kernels) */
int feat3_setup_tile_order(uint64_t rows, const uint32_t* row_ptr, const uint32_t* col_ind, uint64_t tile,
                           uint32_t* order);
/* n doubles uniform in [a, b] from the xorshift64* stream of Random (util/random.hpp:62-110,
 * 259-280); seed 0 selects the reference's default seed */
int feat3_setup_random_uniform(uint64_t seed, uint64_t n, double a, double b, double* out);

/* Renumbered copy of a device CSR matrix (setup; no reference counterpart): old row i -> row
 * row_order[i] (a permutation), old column j -> column col_order[j] (a permutation of the
 * columns), NULL = identity; columns ascending in every row, values moved unchanged.  Device
 * arrays; out_* sized like the input; synchronises the stream (the row pointers are summed on the
 * host).  A row order that is not a permutation or a column out of range is an argument error. */
int feat3_device_permute_csr(void* stream, uint64_t rows, uint64_t cols, const uint32_t* row_ptr,
                             const uint32_t* col_ind, const double* val, const uint32_t* row_order,
                             const uint32_t* col_order, uint32_t* out_row_ptr, uint32_t* out_col_ind, double* out_val);
/* dst[order[i]] = src[i] (inverse = 0) or dst[i] = src[order[i]] (inverse = 1); device arrays,
 * src != dst, enqueued on stream */
int feat3_device_permute_vector(void* stream, uint64_t n, const uint32_t* order, int inverse, const double* src,
                                double* dst);

/* ===================================================================================== */
/* FEAT's mesh refinement and DOF numbering (host)                                        */
/* ===================================================================================== */
/* One StandardRefinery step of a conformal quad (dim 2) or hexa (dim 3) mesh
 * (geometry/standard_refinery.hpp, intern/standard_vertex_refiner.hpp:65-250,
 * intern/standard_index_refiner.hpp:1341-2700).  counts = {vertices, edges, quads, hexas}
 * (hexas 0 for dim 2); index sets as ConformalMesh's get_index_set<cell, face>: v_e (2 per
 * edge), v_q and e_q (4 per quad), v_c (8), e_c (12) and q_c (6) per hexa (dim 3 only, else
 * NULL); out_counts / the fine arrays are sized by feat3_mesh_refine_sizes.  Host arrays. */
int feat3_mesh_refine_sizes(int dim, const uint64_t* counts, uint64_t* out_counts);
int feat3_mesh_refine(int dim, const uint64_t* counts, const double* vtx, const uint32_t* v_e, const uint32_t* v_q,
                      const uint32_t* e_q, const uint32_t* v_c, const uint32_t* e_c, const uint32_t* q_c, double* f_vtx,
                      uint32_t* f_v_e, uint32_t* f_v_q, uint32_t* f_e_q, uint32_t* f_v_c, uint32_t* f_e_c,
                      uint32_t* f_q_c);
/* FEAT's DOF numbering of the Lagrange-1 or -2 space (space/lagrange{1,2}, DofMappingUniform of
 * space/dof_mapping_common.hpp:193-280) on UnitCubeFactory's square / cube
 * (geometry/common_factories.hpp:70-238) refined `level` times: lex_of_feat[i] = the
 * lexicographic lattice index (x fastest; the numbering of feat3_setup_laplace) of DOF i */
int feat3_setup_feat_order(int dim, int degree, int level, uint32_t* lex_of_feat);
/* order[i] = rank of node i when the nodes are sorted by their coordinates (last coordinate
 * slowest); coords: n x dim, distinct.  For the nodes of a tensor lattice (mesh vertices for
 * Lagrange-1; vertices, edge midpoints and cell centres for Lagrange-2) this is the
 * lexicographic lattice index, the order feat3_mgpcg_set_level_order takes. */
int feat3_setup_lattice_order(int dim, uint64_t n, const double* coords, uint32_t* order);
/* permuted copy of a CSR matrix: old row i -> row row_order[i], old column j -> column
 * col_order[j] (NULL: identity); columns ascending in every row; values moved unchanged */
int feat3_setup_permute_csr(uint64_t rows, uint64_t cols, const uint32_t* row_ptr, const uint32_t* col_ind,
                            const double* val, const uint32_t* row_order, const uint32_t* col_order,
                            uint32_t* out_row_ptr, uint32_t* out_col_ind, double* out_val);

#ifdef __cplusplus
}
#endif

#endif /* FEAT3_HIP_H */
