// FP64 SpMV and preconditioned CG on the assembled matrix.
//
// Gridap solves AffineFEOperators with a direct LU (src/Algebra/LinearSolvers.jl:
// 249-274); this engine keeps the matrix in HBM and runs CG instead.  The matrix
// is symmetric (the assembled forms are), so the CSC arrays are read as CSR rows.
//
// One iteration is two kernels with a diagonal (or no) preconditioner and three
// with the factored FSAI preconditioner M = GᵀG:
//   spmv_dir      p' = z + β p and q = A p' in one pass (p' of the neighbours is
//                 formed on the fly from z and p), partials of p'·q;
//   update        x += α p', r -= α q, z = M r, partials of r·z and r·r  (diagonal M);
//   update_lower  x += α p', r' = r - α q, y = G r' (r' of the neighbours formed on
//                 the fly from r and q; r double-buffered), partials of r'·r';
//   upper_rz      z = Gᵀ y, partials of r'·z;
// G and Gᵀ are read from the solver's 16-slot row copy (weights with D^-1/2
// folded in, column indices), written with the factor.  When the free dofs form a
// full box of Q1 lattice nodes, the same iteration runs on stencil kernels with no
// column indices: A as 27 offset arrays, G as 14, Gᵀ read from G's arrays at the
// mirrored offsets.
// Reductions are deterministic and need no grid-wide synchronisation: a kernel
// writes one partial per block, and every block of the consuming kernel sums all
// of them in index order at its start (the same value everywhere); block 0 of
// spmv_dir publishes r·z for the next iteration's β.  The host sums the r·r
// partials itself at the convergence checks.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <mutex>

#include "gh_internal.h"

namespace gh {
namespace {

constexpr int kBlock = 512;
constexpr int kMaxParts = 8192;  // partial slots per reduced quantity (max blocks per grid)
constexpr int kSpmvLanes = 8;    // lanes per matrix row in the SpMV
constexpr int kApplyLanes = 4;   // lanes per row in the FSAI applications

// scalar slots
constexpr int kRz = 0;  // [0..1]: r·z, double-buffered by iteration parity
// convergence decided on the device: kStop = 0 (running), 1 (r·r met the threshold),
// 2 (r·r not finite); kStopIt = the iteration at whose start it was decided
constexpr int kStop = 5, kStopIt = 6;

__device__ __forceinline__ double block_sum(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) r += sh[k];
  return r;
}

// One partial per block: part[blockIdx.x] = Σ_block v.
__device__ __forceinline__ void put_partial(double v, double *part, double *sh) {
  v = block_sum(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// Σ part[0..np) in a fixed order, the same in every block.
__device__ __forceinline__ double sum_partials(const double *__restrict__ part, int np,
                                               double *sh) {
  constexpr int kUnroll = 8;
  const int T = (int)blockDim.x;
  double a = 0.0;
  for (int b0 = threadIdx.x; b0 < np; b0 += T * kUnroll) {
    double w[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int b = b0 + u * T;
      w[u] = b < np ? part[b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) a += w[u];
  }
  return block_sum(a, sh);
}

// Grids of the iteration kernels are capped at kPcgGrid blocks (grid-stride loops):
// every block of a consuming kernel re-reads all partials of its producer, so the
// partial count, not only the row count, sets the cost of a reduction.
constexpr int kPcgGrid = 1024;
static_assert(kPcgGrid <= kMaxParts, "partial slots");

// Σ a[0..np) and Σ b[0..np) in one pass (fixed order, the same in every block).
__device__ __forceinline__ void sum_partials2(const double *__restrict__ a,
                                              const double *__restrict__ b, int np, double *sh,
                                              double &ra, double &rb) {
  double x = 0.0, y = 0.0;
  for (int k = threadIdx.x; k < np; k += (int)blockDim.x) {
    x += a[k];
    y += b[k];
  }
  for (int o = 32; o > 0; o >>= 1) {
    x += __shfl_xor(x, o);
    y += __shfl_xor(y, o);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    sh[2 * w] = x;
    sh[2 * w + 1] = y;
  }
  __syncthreads();
  ra = rb = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
    ra += sh[2 * k];
    rb += sh[2 * k + 1];
  }
}

// True when the solve has stopped (scal[kStop] set by an earlier kernel).  Read once
// per workgroup and shared, so the whole workgroup returns or none of it does.
__device__ __forceinline__ bool block_stopped(const double *scal) {
  __shared__ int flag;
  if (threadIdx.x == 0) flag = scal[kStop] != 0.0;
  __syncthreads();
  return flag != 0;
}

// Start of an iteration: r·z (for β) and, after the first iteration of a start,
// r·r of the previous update against thr = (rtol ‖b‖)²; true when the solve stops
// here (block 0 records why and at which iteration).  Every block sums the same
// partials in the same order, so every block takes the same decision.
__device__ __forceinline__ bool iteration_start(double *scal, const double *__restrict__ part_rz,
                                                const double *__restrict__ part_rr, int np,
                                                int first, double thr, int it, double *sh,
                                                double &rz) {
  double rr = 0.0;
  if (first) {
    rz = sum_partials(part_rz, np, sh);
    return false;
  }
  sum_partials2(part_rz, part_rr, np, sh, rz, rr);
  const bool stop = !(rr > thr) || !isfinite(rr);
  if (stop && blockIdx.x == 0 && threadIdx.x == 0) {
    scal[kStop] = isfinite(rr) ? 1.0 : 2.0;
    scal[kStopIt] = (double)it;
  }
  return stop;
}

inline unsigned grid_rows(int64_t rows, int rows_per_block) {
  return (unsigned)std::min<int64_t>(
      std::max<int64_t>((rows + rows_per_block - 1) / rows_per_block, 1), kPcgGrid);
}

// y = A x (rows from the CSC of a symmetric A).
__global__ __launch_bounds__(kBlock) void spmv_kernel(int64_t n, const int32_t *__restrict__ cp,
                                                      const int32_t *__restrict__ rv,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ y) {
  const int sub = threadIdx.x % kSpmvLanes;
  const int64_t stride = (int64_t)gridDim.x * (kBlock / kSpmvLanes);
  for (int64_t row = (int64_t)blockIdx.x * (kBlock / kSpmvLanes) + threadIdx.x / kSpmvLanes;
       row < n; row += stride) {
    double s = 0.0;
    for (int32_t p = cp[row] + sub; p < cp[row + 1]; p += kSpmvLanes) s += val[p] * x[rv[p]];
#pragma unroll
    for (int o = kSpmvLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (sub == 0) y[row] = s;
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

// r = b - q; with a diagonal (or no) preconditioner also z = M r; partials of r·z,
// r·r, b·b.
__global__ __launch_bounds__(kBlock) void init_kernel(int64_t n, const double *__restrict__ b,
                                                      const double *__restrict__ q,
                                                      const double *__restrict__ dinv,
                                                      int factored, double *__restrict__ r,
                                                      double *__restrict__ z,
                                                      double *__restrict__ part_rz,
                                                      double *__restrict__ part_rr,
                                                      double *__restrict__ part_bb) {
  __shared__ double sh[kBlock / 64];
  double rz = 0.0, rr = 0.0, bb = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const double ri = b[i] - q[i];
    r[i] = ri;
    if (!factored) {
      const double zi = dinv ? ri * dinv[i] : ri;
      z[i] = zi;
      rz += ri * zi;
    }
    rr += ri * ri;
    bb += b[i] * b[i];
  }
  if (!factored) put_partial(rz, part_rz, sh);
  put_partial(rr, part_rr, sh);
  put_partial(bb, part_bb, sh);
}

// r·z of this iteration from its partials (published to scal[rz slot] by block 0);
// p' = z + β p, q = A p', partials of p'·q  (β = 0 on the first iteration of a start)
__global__ __launch_bounds__(kBlock) void spmv_dir_kernel(
    int64_t n, const int32_t *__restrict__ cp, const int32_t *__restrict__ rv,
    const double *__restrict__ val, const double *__restrict__ z,
    const double *__restrict__ pold, double *__restrict__ pnew, double *__restrict__ q,
    double *scal, int par, int first, const double *__restrict__ part_rz,
    const double *__restrict__ part_rr, int np_rz, double thr, int it,
    double *__restrict__ part_pq) {
  __shared__ double sh[2 * (kBlock / 64)];
  if (block_stopped(scal)) return;
  double rz;
  if (iteration_start(scal, part_rz, part_rr, np_rz, first, thr, it, sh, rz)) return;
  const double beta = first ? 0.0 : rz / scal[kRz + (par ^ 1)];
  if (blockIdx.x == 0 && threadIdx.x == 0) scal[kRz + par] = rz;
  const int sub = threadIdx.x % kSpmvLanes;
  const int64_t stride = (int64_t)gridDim.x * (kBlock / kSpmvLanes);
  double dot = 0.0;
  for (int64_t row = (int64_t)blockIdx.x * (kBlock / kSpmvLanes) + threadIdx.x / kSpmvLanes;
       row < n; row += stride) {
    double s = 0.0;
    if (first) {
      for (int32_t p = cp[row] + sub; p < cp[row + 1]; p += kSpmvLanes) s += val[p] * z[rv[p]];
    } else {
      for (int32_t p = cp[row] + sub; p < cp[row + 1]; p += kSpmvLanes) {
        const int32_t j = rv[p];
        s += val[p] * (z[j] + beta * pold[j]);
      }
    }
#pragma unroll
    for (int o = kSpmvLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (sub == 0) {
      const double pn = first ? z[row] : z[row] + beta * pold[row];
      pnew[row] = pn;
      q[row] = s;
      dot += pn * s;
    }
  }
  put_partial(dot, part_pq, sh);
}

// α = r·z / p'·q; x += α p'; r -= α q; z = M r (diagonal M or none); partials of
// r·z and r·r
__global__ __launch_bounds__(kBlock) void update_kernel(
    int64_t n, int par, const double *__restrict__ dinv, double *__restrict__ x,
    double *__restrict__ r, double *__restrict__ z, const double *__restrict__ p,
    const double *__restrict__ q, const double *scal, const double *__restrict__ part_pq,
    int np_pq, double *__restrict__ part_rz, double *__restrict__ part_rr) {
  __shared__ double sh[kBlock / 64];
  if (block_stopped(scal)) return;
  const double alpha = scal[kRz + par] / sum_partials(part_pq, np_pq, sh);
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    const double zi = dinv ? ri * dinv[i] : ri;
    z[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  }
  put_partial(rz, part_rz, sh);
  put_partial(rr, part_rr, sh);
}

// G and Gᵀ are 16-slot rows (kFsaiEllSlots): 4 lanes per row, 4 slots per lane, one
// 16-byte load of weights and one of column indices per lane.
constexpr int kSlots = kFsaiEllSlots;
static_assert(kSlots == 4 * kApplyLanes, "4 slots per lane");

// α = r·z / p'·q; x += α p'; r' = r - α q; y = G r'; partials of r'·r'
__global__ __launch_bounds__(kBlock) void update_lower_kernel(
    int64_t n, int par, const float *__restrict__ ge, const int32_t *__restrict__ gi,
    double *__restrict__ x, const double *__restrict__ rold, double *__restrict__ rnew,
    const double *__restrict__ p, const double *__restrict__ q, double *__restrict__ y,
    const double *scal, const double *__restrict__ part_pq, int np_pq,
    double *__restrict__ part_rr) {
  __shared__ double sh[kBlock / 64];
  if (block_stopped(scal)) return;
  const double alpha = scal[kRz + par] / sum_partials(part_pq, np_pq, sh);
  const int sub = threadIdx.x % kApplyLanes;
  const int64_t stride = (int64_t)gridDim.x * (kBlock / kApplyLanes);
  double rr = 0.0;
  for (int64_t row = (int64_t)blockIdx.x * (kBlock / kApplyLanes) + threadIdx.x / kApplyLanes;
       row < n; row += stride) {
    const float4 g = reinterpret_cast<const float4 *>(ge + row * kSlots)[sub];
    const int4 j = reinterpret_cast<const int4 *>(gi + row * kSlots)[sub];
    double s = (double)g.x * (rold[j.x] - alpha * q[j.x]);
    s += (double)g.y * (rold[j.y] - alpha * q[j.y]);
    s += (double)g.z * (rold[j.z] - alpha * q[j.z]);
    s += (double)g.w * (rold[j.w] - alpha * q[j.w]);
#pragma unroll
    for (int o = kApplyLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (sub == 0) {
      y[row] = s;
      const double ri = rold[row] - alpha * q[row];
      rnew[row] = ri;
      x[row] += alpha * p[row];
      rr += ri * ri;
    }
  }
  put_partial(rr, part_rr, sh);
}

// y = G r (start of a solve)
__global__ __launch_bounds__(kBlock) void lower_kernel(int64_t n, const float *__restrict__ ge,
                                                       const int32_t *__restrict__ gi,
                                                       const double *__restrict__ r,
                                                       double *__restrict__ y) {
  const int sub = threadIdx.x % kApplyLanes;
  const int64_t stride = (int64_t)gridDim.x * (kBlock / kApplyLanes);
  for (int64_t row = (int64_t)blockIdx.x * (kBlock / kApplyLanes) + threadIdx.x / kApplyLanes;
       row < n; row += stride) {
    const float4 g = reinterpret_cast<const float4 *>(ge + row * kSlots)[sub];
    const int4 j = reinterpret_cast<const int4 *>(gi + row * kSlots)[sub];
    double s = (double)g.x * r[j.x] + (double)g.y * r[j.y] + (double)g.z * r[j.z] +
               (double)g.w * r[j.w];
#pragma unroll
    for (int o = kApplyLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (sub == 0) y[row] = s;
  }
}

// z = Gᵀ y; partials of r·z
__global__ __launch_bounds__(kBlock) void upper_rz_kernel(
    int64_t n, const float *__restrict__ ge, const int32_t *__restrict__ gi,
    const double *__restrict__ y, const double *__restrict__ r, double *__restrict__ z,
    const double *scal, double *__restrict__ part_rz) {
  __shared__ double sh[kBlock / 64];
  if (block_stopped(scal)) return;
  const int sub = threadIdx.x % kApplyLanes;
  const int64_t stride = (int64_t)gridDim.x * (kBlock / kApplyLanes);
  double rz = 0.0;
  for (int64_t row = (int64_t)blockIdx.x * (kBlock / kApplyLanes) + threadIdx.x / kApplyLanes;
       row < n; row += stride) {
    const float4 g = reinterpret_cast<const float4 *>(ge + row * kSlots)[sub];
    const int4 j = reinterpret_cast<const int4 *>(gi + row * kSlots)[sub];
    double s = (double)g.x * y[j.x] + (double)g.y * y[j.y] + (double)g.z * y[j.z] +
               (double)g.w * y[j.w];
#pragma unroll
    for (int o = kApplyLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (sub == 0) {
      z[row] = s;
      rz += s * r[row];
    }
  }
  put_partial(rz, part_rz, sh);
}

// ---- box lattice: stencil kernels ---------------------------------------------------
// Row i ↔ box node (x, y, z); neighbour offset o = (dx+1) + 3 (dy+1) + 9 (dz+1) has dof
// offset dx + m0 dy + m0 m1 dz, ascending in o.  The assembled lattice matrix is
// symmetric bit for bit (both entries of a pair sum the same cells in the same order),
// so A is held as its lower half a27[o·N + i], o ≤ 13 (zero where the neighbour is
// outside the box), and A(i, i + off(o)) for o > 13 is read as row i + off(o)'s entry
// at offset 26 − o.  G is g14[o·N + i] for o ≤ 13 (lower offsets); Gᵀ row i collects
// g14[o·N + i − off(o)] from the mirrored neighbours.

struct Box {
  int64_t n, m0, m1, m2;
};

__device__ __forceinline__ uint32_t box_valid(const Box &B, int64_t row) {
  // bit o set when neighbour o of row i lies in the box (32-bit: N < 2^31)
  const int i = (int)row, m0 = (int)B.m0, m1 = (int)B.m1;
  const int x = i % m0, y = (i / m0) % m1, z = i / (m0 * m1);
  const uint32_t vx = (x > 0 ? 1u : 0u) | 2u | (x < B.m0 - 1 ? 4u : 0u);
  const uint32_t vy = (y > 0 ? 1u : 0u) | 2u | (y < B.m1 - 1 ? 4u : 0u);
  const uint32_t vz = (z > 0 ? 1u : 0u) | 2u | (z < B.m2 - 1 ? 4u : 0u);
  uint32_t m = 0;
#pragma unroll
  for (int o = 0; o < 27; ++o)
    if (((vx >> (o % 3)) & 1u) && ((vy >> ((o / 3) % 3)) & 1u) && ((vz >> (o / 9)) & 1u))
      m |= 1u << o;
  return m;
}

__device__ __forceinline__ int64_t box_off(const Box &B, int o) {
  return (int64_t)(o % 3 - 1) + B.m0 * ((o / 3) % 3 - 1) + B.m0 * B.m1 * (o / 9 - 1);
}

// a27 (lower half, o ≤ 13) from the CSR rows of the box-lattice matrix.  A lattice
// row holds exactly the neighbours that lie in the box, in ascending column order (the
// symbolic phase has verified the pattern), which is ascending offset order: the
// row's k-th entry belongs to the k-th in-box offset.  A block of kConvRows rows first stages its
// contiguous run of values in LDS with coalesced loads; then all lanes write array o
// together (coalesced, zeros outside the box written explicitly).  A row whose
// length differs from its in-box neighbour count is flagged.  32-bit indices
// (N < 2^31, checked by the caller).
constexpr int kConvRows = 128;

__global__ __launch_bounds__(kConvRows) void stencil_values_kernel(
    Box B, const int32_t *__restrict__ cp, const double *__restrict__ val,
    double *__restrict__ a27, int *__restrict__ bad) {
  __shared__ double sv[kConvRows * 27];
  const int n = (int)B.n;
  const int i0 = blockIdx.x * kConvRows;
  const int i1 = min(i0 + kConvRows, n);
  const int q0 = cp[i0], nq = cp[i1] - q0;
  if (nq > kConvRows * 27) {  // not a 27-point lattice matrix
    if (threadIdx.x == 0) atomicOr(bad, 1);
    return;
  }
  for (int k = threadIdx.x; k < nq; k += kConvRows) sv[k] = val[q0 + k];
  __syncthreads();
  const int i = i0 + (int)threadIdx.x;
  if (i >= n) return;
  const uint32_t vm = box_valid(B, i);
  int p = cp[i] - q0;
  const bool ok = cp[i + 1] - cp[i] == __popc(vm);
#pragma unroll
  for (int o = 0; o < 27; ++o) {
    double v = 0.0;
    if (((vm >> o) & 1u) && ok) v = sv[p++];
    if (o <= 13) a27[(size_t)o * n + i] = v;  // the upper half is read from the mirror
  }
  if (!ok) atomicOr(bad, 1);
}

// D^-1 of the box-lattice matrix from the centre array
__global__ void diag_inv_box_kernel(int64_t n, const double *__restrict__ a27,
                                    double *__restrict__ dinv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = a27[13 * n + i];
  dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
}

// stencil form of spmv_dir_kernel
__global__ __launch_bounds__(kBlock) void spmv_dir_box_kernel(
    Box B, const double *__restrict__ a27, const double *__restrict__ z,
    const double *__restrict__ pold, double *__restrict__ pnew, double *__restrict__ q,
    double *scal, int par, int first, const double *__restrict__ part_rz,
    const double *__restrict__ part_rr, int np_rz, double thr, int it,
    double *__restrict__ part_pq) {
  __shared__ double sh[2 * (kBlock / 64)];
  if (block_stopped(scal)) return;
  double rz;
  if (iteration_start(scal, part_rz, part_rr, np_rz, first, thr, it, sh, rz)) return;
  const double beta = first ? 0.0 : rz / scal[kRz + (par ^ 1)];
  if (blockIdx.x == 0 && threadIdx.x == 0) scal[kRz + par] = rz;
  double dot = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < B.n;
       i += (int64_t)gridDim.x * kBlock) {
    const uint32_t vm = box_valid(B, i);
    double s = 0.0;
    // (a branch-free form, reading row i for neighbours outside the box, needs 149
    // VGPRs and ran 28 µs instead of 23 µs at N = 250,047)
#pragma unroll
    for (int o = 0; o < 27; ++o) {
      if (!((vm >> o) & 1u)) continue;
      const int64_t j = i + box_off(B, o);
      const double pj = first ? z[j] : z[j] + beta * pold[j];
      // A(i, j) for an upper offset is A(j, i), row j's lower entry at offset 26 - o
      const double a = o <= 13 ? a27[(int64_t)o * B.n + i] : a27[(int64_t)(26 - o) * B.n + j];
      s += a * pj;
    }
    const double pn = first ? z[i] : z[i] + beta * pold[i];
    pnew[i] = pn;
    q[i] = s;
    dot += pn * s;
  }
  put_partial(dot, part_pq, sh);
}

// stencil form of update_lower_kernel
__global__ __launch_bounds__(kBlock) void update_lower_box_kernel(
    Box B, int par, const float *__restrict__ g14, double *__restrict__ x,
    const double *__restrict__ rold, double *__restrict__ rnew, const double *__restrict__ p,
    const double *__restrict__ q, double *__restrict__ y, const double *scal,
    const double *__restrict__ part_pq, int np_pq, double *__restrict__ part_rr) {
  __shared__ double sh[kBlock / 64];
  if (block_stopped(scal)) return;
  const double alpha = scal[kRz + par] / sum_partials(part_pq, np_pq, sh);
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < B.n;
       i += (int64_t)gridDim.x * kBlock) {
    const uint32_t vm = box_valid(B, i);
    double s = 0.0;
#pragma unroll
    for (int o = 0; o < 14; ++o) {
      if (!((vm >> o) & 1u)) continue;
      const int64_t j = i + box_off(B, o);
      s += (double)g14[(int64_t)o * B.n + i] * (rold[j] - alpha * q[j]);
    }
    y[i] = s;
    const double ri = rold[i] - alpha * q[i];
    rnew[i] = ri;
    x[i] += alpha * p[i];
    rr += ri * ri;
  }
  put_partial(rr, part_rr, sh);
}

// y = G r (start of a solve), stencil form
__global__ __launch_bounds__(kBlock) void lower_box_kernel(Box B, const float *__restrict__ g14,
                                                           const double *__restrict__ r,
                                                           double *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < B.n;
       i += (int64_t)gridDim.x * kBlock) {
    const uint32_t vm = box_valid(B, i);
    double s = 0.0;
#pragma unroll
    for (int o = 0; o < 14; ++o) {
      const bool in = (vm >> o) & 1u;
      const int64_t j = in ? i + box_off(B, o) : i;
      const double g = (double)g14[(int64_t)o * B.n + i];
      const double rj = r[j];
      s += in ? g * rj : 0.0;
    }
    y[i] = s;
  }
}

// z = Gᵀ y, partials of r·z; stencil form: entry (j, i) of G with j = i − off(o),
// o ≤ 13, sits at g14[o·N + j], and j is in the box when the mirrored neighbour
// 26 − o of i is
__global__ __launch_bounds__(kBlock) void upper_rz_box_kernel(
    Box B, const float *__restrict__ g14, const double *__restrict__ y,
    const double *__restrict__ r, double *__restrict__ z, const double *scal,
    double *__restrict__ part_rz) {
  __shared__ double sh[kBlock / 64];
  if (block_stopped(scal)) return;
  double rz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < B.n;
       i += (int64_t)gridDim.x * kBlock) {
    const uint32_t vm = box_valid(B, i);
    double s = 0.0;
#pragma unroll
    for (int o = 13; o >= 0; --o) {
      // branch-free: a neighbour outside the box reads row i itself (in bounds) and
      // its product is replaced by 0, so all loads of the row are issued together
      // (6.4 µs instead of 8.1 µs at N = 250,047)
      const bool in = (vm >> (26 - o)) & 1u;
      const int64_t j = in ? i - box_off(B, o) : i;
      const double g = (double)g14[(int64_t)o * B.n + j];
      const double yj = y[j];
      s += in ? g * yj : 0.0;
    }
    z[i] = s;
    rz += s * r[i];
  }
  put_partial(rz, part_rz, sh);
}

}  // namespace

void spmv_device(Ctx &c, const double *x, double *y) {
  hipLaunchKernelGGL(spmv_kernel, dim3(grid_rows(c.nfree, kBlock / kSpmvLanes)), dim3(kBlock), 0,
                     c.stream, (int64_t)c.nfree, c.colptr.get(), c.rowval.get(), c.nzval.get(),
                     x, y);
  GH_CHECK_HIP(hipGetLastError());
}

namespace {
// Σ of np partials on the host (convergence checks only; fixed order).
double host_sum(const double *v, int np) {
  double a = 0.0;
  for (int k = 0; k < np; ++k) a += v[k];
  return a;
}
}  // namespace

namespace {

// Device vectors of one solve: the solution x (initial guess on entry) and b, the CG
// vectors, the diagonal inverse, partial sums (4 x kMaxParts) and scalars (8).
struct PcgBufs {
  double *x, *b, *r, *r2, *z, *p, *p2, *q, *y, *dinv, *part, *scal;
};

__global__ void clear_stop_kernel(double *scal) {
  if (threadIdx.x < 2) scal[kStop + threadIdx.x] = 0.0;
}

// The solve on the given buffers.  box: the stencil form may be used (needs c.a27
// and, for FSAI, c.g14); timed: HIP events around the solve (device_ms), else 0.
void pcg_run(Ctx &c, const PcgBufs &V, int precond, double rtol, int maxit, bool allow_box,
             bool timed, gh_solve_stats *st) {
  const int64_t n = c.nfree;
  hipStream_t s = c.stream;
  const size_t nn = (size_t)std::max<int64_t>(n, 1);
  // the solve is timed from here: per-matrix preparation (D^-1, the stencil arrays)
  // included
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (timed) {
    GH_CHECK_HIP(hipEventCreate(&e0));
    GH_CHECK_HIP(hipEventCreate(&e1));
    GH_CHECK_HIP(hipEventRecord(e0, s));
  }
  double *dinv = nullptr;
  const int factored = precond == GH_PRECOND_FSAI;
  // box lattice: stencil kernels (A as 27 offset arrays, G as 14), one thread per row
  // GRIDAP_HIP_STENCIL_PCG=0 runs the general (CSR) kernels on a box lattice too, so
  // the two forms can be compared on the same system
  const char *stencil_env = std::getenv("GRIDAP_HIP_STENCIL_PCG");
  bool box = allow_box && c.box && n > 0 && n < (int64_t(1) << 31) &&
             (!factored || c.g14.get() != nullptr) && !(stencil_env && stencil_env[0] == '0');
  const Box B{n, c.box_m[0], c.box_m[1], c.box_m[2]};
  if (box) {
    c.a27.resize((size_t)14 * nn);
    DBuf<int> bad;
    bad.resize(1);
    GH_CHECK_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int), s));
    hipLaunchKernelGGL(stencil_values_kernel, dim3((unsigned)((n + kConvRows - 1) / kConvRows)),
                       dim3(kConvRows), 0, s, B, c.colptr.get(), c.nzval.get(), c.a27.get(),
                       bad.get());
    GH_CHECK_HIP(hipGetLastError());
    int hbad = 0;
    GH_CHECK_HIP(hipMemcpyAsync(&hbad, bad.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    GH_CHECK_HIP(hipStreamSynchronize(s));
    box = hbad == 0;
  }
  if (precond == GH_PRECOND_JACOBI) {
    if (box)
      hipLaunchKernelGGL(diag_inv_box_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                         n, c.a27.get(), V.dinv);
    else
      hipLaunchKernelGGL(diag_inv_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                         c.colptr.get(), c.rowval.get(), c.nzval.get(), V.dinv);
    dinv = V.dinv;
  }
  if (factored)
    GH_REQUIRE(c.fsai_built, GH_ERR_STATE, "no FSAI factor for this matrix: call gh_fsai_build");
  double *scal = V.scal;
  double *part_pq = V.part, *part_rz = part_pq + kMaxParts;
  double *part_rr = part_rz + kMaxParts, *part_bb = part_rr + kMaxParts;
  double *rbuf[2] = {V.r, factored ? V.r2 : V.r};
  double *pbuf[2] = {V.p, V.p2};
  const unsigned g_vec = grid_rows(n, kBlock);
  const unsigned g_spmv = box ? g_vec : grid_rows(n, kBlock / kSpmvLanes);
  const unsigned g_apply = box ? g_vec : grid_rows(n, kBlock / kApplyLanes);
  // partial counts of each quantity, by producer
  const int np_rz = factored ? (int)g_apply : (int)g_vec;
  double hrr[kPcgGrid], hbb[kPcgGrid];  // g_vec <= kPcgGrid partials

  double bnorm = 1.0, relres = INFINITY;
  int it = 0, restarts = 0;
  bool converged = false;
  while (true) {
    // (re)start from the true residual r = b - A x
    spmv_device(c, V.x, V.q);
    hipLaunchKernelGGL(init_kernel, dim3(g_vec), dim3(kBlock), 0, s, n, V.b, V.q,
                       dinv, factored, rbuf[0], V.z, part_rz, part_rr, part_bb);
    hipLaunchKernelGGL(clear_stop_kernel, dim3(1), dim3(64), 0, s, scal);
    if (factored && box) {
      hipLaunchKernelGGL(lower_box_kernel, dim3(g_apply), dim3(kBlock), 0, s, B, c.g14.get(),
                         rbuf[0], V.y);
      hipLaunchKernelGGL(upper_rz_box_kernel, dim3(g_apply), dim3(kBlock), 0, s, B,
                         c.g14.get(), V.y, rbuf[0], V.z, scal, part_rz);
    } else if (factored) {
      hipLaunchKernelGGL(lower_kernel, dim3(g_apply), dim3(kBlock), 0, s, n, c.gle.get(),
                         c.gli.get(), rbuf[0], V.y);
      hipLaunchKernelGGL(upper_rz_kernel, dim3(g_apply), dim3(kBlock), 0, s, n, c.gue.get(),
                         c.gui.get(), V.y, rbuf[0], V.z, scal, part_rz);
    }
    GH_CHECK_HIP(hipGetLastError());
    GH_CHECK_HIP(hipMemcpyAsync(hrr, part_rr, sizeof(double) * g_vec,
                                hipMemcpyDeviceToHost, s));
    GH_CHECK_HIP(hipMemcpyAsync(hbb, part_bb, sizeof(double) * g_vec,
                                hipMemcpyDeviceToHost, s));
    GH_CHECK_HIP(hipStreamSynchronize(s));
    const double bb = host_sum(hbb, (int)g_vec);
    bnorm = std::sqrt(bb) > 0 ? std::sqrt(bb) : 1.0;
    relres = std::sqrt(host_sum(hrr, (int)g_vec)) / bnorm;
    if (relres <= rtol) {
      converged = true;
      break;
    }
    if (it >= maxit || restarts > 3) break;
    ++restarts;
    // iterations are enqueued in chunks; each spmv_dir decides on the device whether
    // the previous update met the tolerance, and once it has, every later kernel of
    // the chunk returns at once.  The host reads the decision after each chunk.
    const int chunk = 16;
    const double thr = (rtol * bnorm) * (rtol * bnorm);
    int k = 0;  // iteration since this start: selects the buffers and scalar slots
    bool stopped = false;
    while (it < maxit && !stopped) {
      for (int m = 0; m < chunk && it + m < maxit; ++m, ++k) {
        const int par = k & 1;
        const int itk = it + m;
        double *pold = pbuf[par], *pnew = pbuf[par ^ 1];
        if (box)
          hipLaunchKernelGGL(spmv_dir_box_kernel, dim3(g_spmv), dim3(kBlock), 0, s, B,
                             c.a27.get(), V.z, pold, pnew, V.q, scal, par,
                             k == 0 ? 1 : 0, part_rz, part_rr, np_rz, thr, itk, part_pq);
        else
          hipLaunchKernelGGL(spmv_dir_kernel, dim3(g_spmv), dim3(kBlock), 0, s, n,
                             c.colptr.get(), c.rowval.get(), c.nzval.get(), V.z, pold,
                             pnew, V.q, scal, par, k == 0 ? 1 : 0, part_rz, part_rr,
                             np_rz, thr, itk, part_pq);
        if (factored && box) {
          double *rold = rbuf[par], *rnew = rbuf[par ^ 1];
          hipLaunchKernelGGL(update_lower_box_kernel, dim3(g_apply), dim3(kBlock), 0, s, B, par,
                             c.g14.get(), V.x, rold, rnew, pnew, V.q,
                             V.y, scal, part_pq, (int)g_spmv, part_rr);
          hipLaunchKernelGGL(upper_rz_box_kernel, dim3(g_apply), dim3(kBlock), 0, s, B,
                             c.g14.get(), V.y, rnew, V.z, scal, part_rz);
        } else if (factored) {
          double *rold = rbuf[par], *rnew = rbuf[par ^ 1];
          hipLaunchKernelGGL(update_lower_kernel, dim3(g_apply), dim3(kBlock), 0, s, n, par,
                             c.gle.get(), c.gli.get(), V.x, rold, rnew, pnew, V.q,
                             V.y, scal, part_pq, (int)g_spmv, part_rr);
          hipLaunchKernelGGL(upper_rz_kernel, dim3(g_apply), dim3(kBlock), 0, s, n, c.gue.get(),
                             c.gui.get(), V.y, rnew, V.z, scal, part_rz);
        } else {
          hipLaunchKernelGGL(update_kernel, dim3(g_vec), dim3(kBlock), 0, s, n, par, dinv,
                             V.x, rbuf[0], V.z, pnew, V.q, scal, part_pq,
                             (int)g_spmv, part_rz, part_rr);
        }
      }
      GH_CHECK_HIP(hipGetLastError());
      double st2[2] = {0.0, 0.0};
      GH_CHECK_HIP(hipMemcpyAsync(st2, scal + kStop, sizeof(st2), hipMemcpyDeviceToHost, s));
      GH_CHECK_HIP(hipStreamSynchronize(s));
      if (st2[0] != 0.0) {
        it = (int)st2[1];  // iterations completed before the stop
        stopped = true;
      } else {
        it = std::min(it + chunk, maxit);
      }
    }
    // the true residual of the next start decides convergence (and, if the stopped
    // r·r was not finite, the restart count ends the solve)
  }
  float ms = 0.f;
  if (timed) {
    GH_CHECK_HIP(hipEventRecord(e1, s));
    GH_CHECK_HIP(hipEventSynchronize(e1));
    GH_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  } else {
    GH_CHECK_HIP(hipStreamSynchronize(s));
  }
  st->iterations = it;
  st->converged = converged ? 1 : 0;
  st->rel_residual = relres;
  st->device_ms = ms;
}

}  // namespace

void pcg_device(Ctx &c, int precond, double rtol, int maxit, gh_solve_stats *st) {
  const size_t nn = (size_t)std::max<int64_t>(c.nfree, 1);
  for (auto *bf : {&c.s_r, &c.s_z, &c.s_p, &c.s_p2, &c.s_q}) bf->resize(nn);
  c.s_part.resize(4 * kMaxParts);
  c.s_scal.resize(8);
  if (precond == GH_PRECOND_JACOBI) c.s_dinv.resize(nn);
  if (precond == GH_PRECOND_FSAI) {
    c.s_y.resize(nn);
    c.s_r2.resize(nn);
  }
  const PcgBufs V{c.s_x.get(), c.s_b.get(), c.s_r.get(), c.s_r2.get(), c.s_z.get(),
                  c.s_p.get(), c.s_p2.get(), c.s_q.get(), c.s_y.get(), c.s_dinv.get(),
                  c.s_part.get(), c.s_scal.get()};
  pcg_run(c, V, precond, rtol, maxit, true, true, st);
}

// ---- the solve that allocates nothing (gh_pcg_noalloc) ------------------------------
// Its vectors live in one fixed device array, `scratch`, which every GPU the library
// runs on holds until the process ends: x, b, r, z, p, p', q, D^-1 of at most
// GH_NOALLOC_MAX_N entries each, then the partial sums and the scalars.  Every call on
// one GPU shares that GPU's copy (calls that overlap on one GPU race), and nothing in
// it is read before the call writes it.
namespace {
constexpr size_t kNoAllocVec = (size_t)GH_NOALLOC_MAX_N;
constexpr size_t kScratchDoubles = 8 * kNoAllocVec + 4 * (size_t)kMaxParts + 8;
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};  // device address of `scratch` on each GPU
std::mutex g_scratch_mu;
}  // namespace

// Look up this GPU's address of `scratch` once (the calling thread has the device
// current).  Host code uses that address, never the array's name.
void scratch_init(int device) {
  GH_REQUIRE(device >= 0 && device < kMaxDevices, GH_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr)
    throw Error(GH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = static_cast<double *>(p);
}

void pcg_noalloc(Ctx &c, int precond, const double *b, double *x, double rtol, int maxit,
                 gh_solve_stats *st) {
  GH_REQUIRE(precond == GH_PRECOND_NONE || precond == GH_PRECOND_JACOBI, GH_ERR_UNSUPPORTED,
             "the allocation-free solve provides Jacobi or no preconditioner");
  GH_REQUIRE(c.nfree <= GH_NOALLOC_MAX_N, GH_ERR_ARG,
             "num_free_dofs exceeds GH_NOALLOC_MAX_N, the allocation-free solve's limit");
  double *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = c.device >= 0 && c.device < kMaxDevices ? g_scratch[c.device] : nullptr;
  }
  GH_REQUIRE(base, GH_ERR_STATE, "scratch address not set up for this device");
  const size_t n = (size_t)c.nfree;
  const size_t L = kNoAllocVec;
  const PcgBufs V{base, base + L, base + 2 * L, nullptr, base + 3 * L, base + 4 * L,
                  base + 5 * L, base + 6 * L, nullptr, base + 7 * L, base + 8 * L,
                  base + 8 * L + 4 * (size_t)kMaxParts};
  const auto t0 = std::chrono::steady_clock::now();
  GH_CHECK_HIP(hipMemcpyAsync(V.x, x, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
  if (b)
    GH_CHECK_HIP(hipMemcpyAsync(V.b, b, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
  else
    GH_CHECK_HIP(hipMemcpyAsync(V.b, c.b.get(), sizeof(double) * n, hipMemcpyDeviceToDevice,
                                c.stream));
  pcg_run(c, V, precond, rtol, maxit, false, false, st);
  GH_CHECK_HIP(hipMemcpyAsync(x, V.x, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
  GH_CHECK_HIP(hipStreamSynchronize(c.stream));
  st->device_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  st->precond_used = precond;
}

}  // namespace gh
