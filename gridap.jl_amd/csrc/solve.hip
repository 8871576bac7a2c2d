// FP64 SpMV and preconditioned CG on the assembled matrix.
//
// Gridap solves AffineFEOperators with a direct LU (src/Algebra/LinearSolvers.jl:
// 249-274); this engine keeps the matrix in HBM and runs CG instead.  The matrix
// is symmetric (the assembled forms are), so the CSC arrays are read as CSR rows.
// Reductions are two-level with a fixed order (block partials, then every
// consumer sums the partials in index order), so runs are bit-reproducible.
#include <algorithm>
#include <cmath>

#include "gh_internal.h"

namespace gh {
namespace {

constexpr int kBlock = 512;
constexpr int kParts = 512;  // blocks of the vector kernels = partials per reduction
constexpr int kLanes = 8;    // lanes per matrix row in the SpMV

__device__ __forceinline__ double block_sum(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) r += sh[k];
    sh[0] = r;
  }
  __syncthreads();
  r = sh[0];
  __syncthreads();
  return r;
}

// Sum of kParts partials in index order, identical in every block: one partial
// per thread (kParts == kBlock), then the fixed shuffle/LDS tree of block_sum.
static_assert(kParts == kBlock, "sum_parts assumes one partial per thread");
__device__ __forceinline__ double sum_parts(const double *parts, double *sh) {
  return block_sum(parts[threadIdx.x], sh);
}

// y = A x (rows from the CSC of a symmetric A); optional partial of x.y.
__global__ __launch_bounds__(kBlock) void spmv_kernel(int64_t n, const int32_t *__restrict__ cp,
                                                      const int32_t *__restrict__ rv,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ y, double *part) {
  __shared__ double sh[kBlock / 64];
  const int sub = threadIdx.x % kLanes;
  const int64_t rows_per_iter = (int64_t)gridDim.x * (kBlock / kLanes);
  double dot = 0.0;
  for (int64_t row = (int64_t)blockIdx.x * (kBlock / kLanes) + threadIdx.x / kLanes; row < n;
       row += rows_per_iter) {
    const int32_t p0 = cp[row], p1 = cp[row + 1];
    double s = 0.0;
    for (int32_t p = p0 + sub; p < p1; p += kLanes) s += val[p] * x[rv[p]];
#pragma unroll
    for (int o = kLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (sub == 0) {
      y[row] = s;
      dot += s * x[row];
    }
  }
  if (part) {
    double t = block_sum(dot, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

__global__ void diag_inv_kernel(int64_t n, const int32_t *__restrict__ cp,
                                const int32_t *__restrict__ rv, const double *__restrict__ val,
                                double *__restrict__ dinv) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double d = 0.0;
  for (int32_t p = cp[j]; p < cp[j + 1]; ++p)
    if (rv[p] == j) d = val[p];
  dinv[j] = d != 0.0 ? 1.0 / d : 1.0;
}

// r = b - q; with a diagonal (or no) preconditioner also z = M r, p = z.  With a
// factored preconditioner (`factored`) z, p and the r.z partials come from the
// apply kernels instead.  Partials of r.z, r.r, b.b.
__global__ __launch_bounds__(kBlock) void init_kernel(int64_t n, const double *__restrict__ b,
                                                      const double *__restrict__ q,
                                                      const double *__restrict__ dinv,
                                                      int factored, double *r, double *z,
                                                      double *p, double *part) {
  __shared__ double sh[kBlock / 64];
  double rz = 0.0, rr = 0.0, bb = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    double ri = b[i] - q[i];
    r[i] = ri;
    if (!factored) {
      double zi = dinv ? ri * dinv[i] : ri;
      z[i] = zi;
      p[i] = zi;
      rz += ri * zi;
    }
    rr += ri * ri;
    bb += b[i] * b[i];
  }
  rz = block_sum(rz, sh);
  rr = block_sum(rr, sh);
  bb = block_sum(bb, sh);
  if (threadIdx.x == 0) {
    if (!factored) part[blockIdx.x] = rz;
    part[kParts + blockIdx.x] = rr;
    part[2 * kParts + blockIdx.x] = bb;
  }
}

__global__ void copy_kernel(int64_t n, const double *__restrict__ a, double *__restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// scal layout: [0..1] rz (double-buffered by iteration parity), [2] rr, [3] bb.
__global__ __launch_bounds__(kBlock) void init_scalars(const double *part, double *scal) {
  __shared__ double sh[kBlock / 64];
  double rz = sum_parts(part, sh);
  double rr = sum_parts(part + kParts, sh);
  double bb = sum_parts(part + 2 * kParts, sh);
  if (threadIdx.x == 0) {
    scal[0] = rz;
    scal[1] = rz;
    scal[2] = rr;
    scal[3] = bb;
  }
}

// alpha = rz / (p.q); x += alpha p; r -= alpha q; z = M r (diagonal M; with a
// factored M the apply kernels follow); partials of r.z, r.r.
__global__ __launch_bounds__(kBlock) void update_kernel(int64_t n, int parity,
                                                        const double *__restrict__ pq_part,
                                                        const double *__restrict__ scal,
                                                        const double *__restrict__ p,
                                                        const double *__restrict__ q,
                                                        const double *__restrict__ dinv,
                                                        int factored, double *x, double *r,
                                                        double *z, double *part) {
  __shared__ double sh[kBlock / 64];
  const double pq = sum_parts(pq_part, sh);
  const double alpha = scal[parity] / pq;
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    x[i] += alpha * p[i];
    double ri = r[i] - alpha * q[i];
    r[i] = ri;
    if (!factored) {
      double zi = dinv ? ri * dinv[i] : ri;
      z[i] = zi;
      rz += ri * zi;
    }
    rr += ri * ri;
  }
  rz = block_sum(rz, sh);
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    if (!factored) part[blockIdx.x] = rz;
    part[kParts + blockIdx.x] = rr;
  }
}

// beta = rz_new / rz_old; p = z + beta p; block 0 publishes rz_new and r.r.
__global__ __launch_bounds__(kBlock) void direction_kernel(int64_t n, int parity,
                                                           const double *__restrict__ part,
                                                           double *scal,
                                                           const double *__restrict__ z,
                                                           double *p) {
  __shared__ double sh[kBlock / 64];
  const double rz_new = sum_parts(part, sh);
  const double rr = sum_parts(part + kParts, sh);
  const double beta = rz_new / scal[parity];
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    p[i] = z[i] + beta * p[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[parity ^ 1] = rz_new;
    scal[2] = rr;
  }
}

unsigned spmv_grid(int64_t n) {
  int64_t rows_per_block = kBlock / kLanes;
  int64_t g = (n + rows_per_block - 1) / rows_per_block;
  return (unsigned)std::min<int64_t>(std::max<int64_t>(g, 1), kParts);
}

}  // namespace

void spmv_device(Ctx &c, const double *x, double *y) {
  hipLaunchKernelGGL(spmv_kernel, dim3(spmv_grid(c.nfree)), dim3(kBlock), 0, c.stream,
                     (int64_t)c.nfree, c.colptr.get(), c.rowval.get(), c.nzval.get(), x, y,
                     (double *)nullptr);
  GH_CHECK_HIP(hipGetLastError());
}

void pcg_device(Ctx &c, int precond, double rtol, int maxit, gh_solve_stats *st) {
  const int64_t n = c.nfree;
  hipStream_t s = c.stream;
  for (auto *bf : {&c.s_r, &c.s_z, &c.s_p, &c.s_q, &c.s_dinv}) bf->resize(std::max<int64_t>(n, 1));
  c.s_part.resize(4 * kParts);
  c.s_scal.resize(4);
  // Partials must be zero where a grid has fewer than kParts blocks (the SpMV).
  GH_CHECK_HIP(hipMemsetAsync(c.s_part.get(), 0, sizeof(double) * 4 * kParts, s));
  double *dinv = nullptr;
  const int factored = precond == GH_PRECOND_FSAI;
  if (precond == GH_PRECOND_JACOBI) {
    hipLaunchKernelGGL(diag_inv_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                       c.colptr.get(), c.rowval.get(), c.nzval.get(), c.s_dinv.get());
    dinv = c.s_dinv.get();
  }
  if (factored) {
    GH_REQUIRE(c.fsai_built, GH_ERR_STATE, "no FSAI factor for this matrix: call gh_fsai_build");
    c.s_y.resize(std::max<int64_t>(n, 1));
  }
  // z = Gᵀ (G r) with the r.z partials in part[0, kParts)
  auto apply_factored = [&]() {
    fsai_apply_lower(c, c.s_r.get(), c.s_y.get());
    fsai_apply_upper(c, c.s_y.get(), c.s_z.get(), c.s_r.get(), c.s_part.get());
  };
  hipEvent_t e0, e1;
  GH_CHECK_HIP(hipEventCreate(&e0));
  GH_CHECK_HIP(hipEventCreate(&e1));
  GH_CHECK_HIP(hipEventRecord(e0, s));

  double *pq_part = c.s_part.get() + 3 * kParts;  // written only by the SpMV grid
  const unsigned vgrid = kParts;
  const unsigned mgrid = spmv_grid(n);
  double h[4] = {0, 0, 0, 0};
  double bnorm = 1.0, relres = INFINITY;
  int it = 0, restarts = 0;
  bool converged = false;
  while (true) {
    // (re)start from the true residual r = b - A x
    spmv_device(c, c.s_x.get(), c.s_q.get());
    hipLaunchKernelGGL(init_kernel, dim3(vgrid), dim3(kBlock), 0, s, n, c.s_b.get(),
                       c.s_q.get(), dinv, factored, c.s_r.get(), c.s_z.get(), c.s_p.get(),
                       c.s_part.get());
    if (factored) {
      apply_factored();
      hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                         c.s_z.get(), c.s_p.get());
    }
    hipLaunchKernelGGL(init_scalars, dim3(1), dim3(kBlock), 0, s, c.s_part.get(),
                       c.s_scal.get());
    GH_CHECK_HIP(hipGetLastError());
    GH_CHECK_HIP(hipMemcpyAsync(h, c.s_scal.get(), sizeof(h), hipMemcpyDeviceToHost, s));
    GH_CHECK_HIP(hipStreamSynchronize(s));
    bnorm = std::sqrt(h[3]) > 0 ? std::sqrt(h[3]) : 1.0;
    relres = std::sqrt(h[2]) / bnorm;
    if (relres <= rtol) {
      converged = true;
      break;
    }
    if (it >= maxit || restarts > 3) break;
    ++restarts;
    int parity = 0;  // scal[0] holds rz after init
    const int check = 8;
    while (it < maxit) {
      for (int k = 0; k < check && it < maxit; ++k, ++it) {
        hipLaunchKernelGGL(spmv_kernel, dim3(mgrid), dim3(kBlock), 0, s, n, c.colptr.get(),
                           c.rowval.get(), c.nzval.get(), c.s_p.get(), c.s_q.get(), pq_part);
        hipLaunchKernelGGL(update_kernel, dim3(vgrid), dim3(kBlock), 0, s, n, parity, pq_part,
                           c.s_scal.get(), c.s_p.get(), c.s_q.get(), dinv, factored,
                           c.s_x.get(), c.s_r.get(), c.s_z.get(), c.s_part.get());
        if (factored) apply_factored();
        hipLaunchKernelGGL(direction_kernel, dim3(vgrid), dim3(kBlock), 0, s, n, parity,
                           c.s_part.get(), c.s_scal.get(), c.s_z.get(), c.s_p.get());
        parity ^= 1;
      }
      GH_CHECK_HIP(hipGetLastError());
      GH_CHECK_HIP(hipMemcpyAsync(h, c.s_scal.get(), sizeof(h), hipMemcpyDeviceToHost, s));
      GH_CHECK_HIP(hipStreamSynchronize(s));
      if (!std::isfinite(h[2])) break;
      if (std::sqrt(h[2]) / bnorm <= rtol) break;
    }
  }
  GH_CHECK_HIP(hipEventRecord(e1, s));
  GH_CHECK_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  GH_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  st->iterations = it;
  st->converged = converged ? 1 : 0;
  st->rel_residual = relres;
  st->device_ms = ms;
}

}  // namespace gh
