// Assembly on affine simplices (P1/P2 on tetrahedra): Gridap's cell loop
// (src/FESpaces/SparseMatrixAssemblers.jl:214-248) for ∫κ∇v·∇u and ∫f v over a
// simplexified mesh, owner-computes.
//
// On an affine tetrahedron with DF_ij = ∂x_i/∂ξ_j the physical gradients are
// DF⁻ᵀ∇̂φ, so the element matrix is a contraction of a 6-entry cell metric with a
// reference tensor:
//     K_ab = κ|det DF| Σ_q w_q ∇̂φ_a(q)ᵀ (DF⁻¹DF⁻ᵀ) ∇̂φ_b(q) = Σ_k g_k Ĝ^k_ab,
//     g = κ|det DF| (M00, M11, M22, M01, M02, M12),  M = DF⁻¹DF⁻ᵀ,
//     Ĝ^dd_ab = Σ_q w_q ∂_dφ̂_a ∂_dφ̂_b,  Ĝ^de_ab = Σ_q w_q (∂_dφ̂_a ∂_eφ̂_b + ∂_eφ̂_a ∂_dφ̂_b),
// and the cell vector is f|det DF| m_a with m_a = Σ_q w_q φ̂_a(q).  The geometry
// (DF, its inverse and determinant) is FP64; the contraction is FP32 or FP64 by
// precision mode; every sum into nzval and b is FP64.
//
// Each CSC column (free dof j) is owned by a group of 32 lanes.  The symbolic
// phase lists the cells holding j in ascending cell order (the order of Gridap's
// cell loop); for each, lane b computes K_ab (a = j's local index) and adds it
// into the LDS slot of row dof_b, found by binary search in the column's rows.
// The lanes of one step hit distinct slots and steps are ordered, so the sums are
// deterministic; every nonzero is written to HBM once, with coalesced stores.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "gh_internal.h"

namespace gh {
namespace {

constexpr int kBlock = 256;
constexpr int kGroup = 32;                 // lanes per column
constexpr int kGroups = kBlock / kGroup;   // columns per block
constexpr int kGeoStride = 8;              // g[6], f|det|, pad

inline unsigned grid_for(int64_t n, int block = kBlock) {
  return (unsigned)std::max<int64_t>(1, (n + block - 1) / block);
}

__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  constexpr int kXcds = 8;
  const int q = nblocks / kXcds, r = nblocks % kXcds, x = b % kXcds;
  return x * q + min(x, r) + b / kXcds;
}

// ---- symbolic: dof -> (cell, local index), ascending cells ------------------------

__global__ void dof_cell_keys(int64_t ncells, int nloc, const int32_t *__restrict__ cell_dofs,
                              uint64_t invalid, uint64_t *__restrict__ keys) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells * nloc) return;
  const int64_t cell = t / nloc;
  const int a = (int)(t - cell * nloc);
  const int d = cell_dofs[t];
  keys[t] = d > 0 ? ((uint64_t)(d - 1) << 32) | (uint64_t)((cell << 4) | a) : invalid;
}

// adj[i] = (cell << 4 | a) of sorted entry i; adjptr[x] = first sorted position of
// free dof x, adjptr[nfree] = end of the free entries (non-free entries sort last).
__global__ void adjacency_fill(int64_t total, int64_t nfree, const uint64_t *__restrict__ sorted,
                               int32_t *__restrict__ adjptr, int32_t *__restrict__ adj) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int64_t di = (int64_t)(sorted[i] >> 32);
  const int64_t d = di < nfree ? di : nfree;
  int64_t prev = -1;
  if (i) {
    const int64_t dp = (int64_t)(sorted[i - 1] >> 32);
    prev = dp < nfree ? dp : nfree;
  }
  if (d < nfree) adj[i] = (int32_t)(sorted[i] & 0xffffffffu);
  for (int64_t x = prev + 1; x <= d; ++x) adjptr[x] = (int32_t)i;
  if (i == total - 1)
    for (int64_t x = d + 1; x <= nfree; ++x) adjptr[x] = (int32_t)total;
}

__global__ void max_column(int64_t n, const int32_t *__restrict__ colptr, int *__restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) atomicMax(out, colptr[j + 1] - colptr[j]);
}

// ---- numeric -----------------------------------------------------------------------

__global__ void simplex_geom_kernel(int64_t ncells, const int32_t *__restrict__ cell_nodes,
                                    const double *__restrict__ xyz,
                                    const double *__restrict__ kappa,
                                    const double *__restrict__ fsrc, double *__restrict__ geo) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int32_t *cn = cell_nodes + c * 4;
  double x[4][3];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int d = 0; d < 3; ++d) x[v][d] = xyz[(size_t)cn[v] * 3 + d];
  double e[3][3];  // e[j] = column j of DF = x_{j+1} - x_0
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int d = 0; d < 3; ++d) e[j][d] = x[j + 1][d] - x[0][d];
  auto cross = [](const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  double r[3][3];  // rows of DF⁻¹ times det
  cross(e[1], e[2], r[0]);
  cross(e[2], e[0], r[1]);
  cross(e[0], e[1], r[2]);
  const double det = e[0][0] * r[0][0] + e[0][1] * r[0][1] + e[0][2] * r[0][2];
  const double ad = fabs(det);
  const double s = (kappa ? kappa[c] : 1.0) * ad / (det * det);
  auto dot = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  double *g = geo + c * kGeoStride;
  g[0] = s * dot(r[0], r[0]);
  g[1] = s * dot(r[1], r[1]);
  g[2] = s * dot(r[2], r[2]);
  g[3] = s * dot(r[0], r[1]);
  g[4] = s * dot(r[0], r[2]);
  g[5] = s * dot(r[1], r[2]);
  g[6] = (fsrc ? fsrc[c] : 1.0) * ad;
  g[7] = 0.0;
}

struct GatherArgs {
  int64_t nfree;
  const int32_t *colptr, *rowval, *adjptr, *adj, *cell_dofs;
  const double *geo, *gref, *mref, *dvals;
  double *nzval, *b;
};

template <int NLOC, class T, int MAXNZ>
__global__ __launch_bounds__(kBlock) void simplex_gather_kernel(GatherArgs A) {
  __shared__ T ref[6 * NLOC * NLOC];
  __shared__ double mref[NLOC];
  __shared__ int rows[kGroups][MAXNZ];
  __shared__ double acc[kGroups][MAXNZ];
  for (int t = threadIdx.x; t < 6 * NLOC * NLOC; t += kBlock) ref[t] = (T)A.gref[t];
  if (threadIdx.x < NLOC) mref[threadIdx.x] = A.mref[threadIdx.x];
  const int g = threadIdx.x / kGroup, lane = threadIdx.x % kGroup;
  const int64_t j = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * kGroups + g;
  const bool active = j < A.nfree;
  int p0 = 0, n = 0;
  if (active) {
    p0 = A.colptr[j];
    n = A.colptr[j + 1] - p0;  // 1 <= n <= MAXNZ (checked on the host)
  }
  for (int k = lane; k < n; k += kGroup) {
    rows[g][k] = A.rowval[p0 + k];
    acc[g][k] = 0.0;
  }
  __syncthreads();
  double bl = 0.0;
  if (active && lane < NLOC) {
    const int q1 = A.adjptr[j + 1];
    for (int q = A.adjptr[j]; q < q1; ++q) {
      const int e = A.adj[q];
      const int cell = e >> 4, a = e & 15;
      const double *gc = A.geo + (size_t)cell * kGeoStride;
      const int d = A.cell_dofs[(size_t)cell * NLOC + lane];
      T kab = (T)0;
#pragma unroll
      for (int k = 0; k < 6; ++k) kab += (T)gc[k] * ref[(k * NLOC + a) * NLOC + lane];
      if (d > 0) {
        const int key = d - 1;
        int lo = 0, hi = n - 1;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (rows[g][mid] < key) lo = mid + 1;
          else hi = mid;
        }
        atomicAdd(&acc[g][lo], (double)kab);
      } else if (A.dvals) {
        bl -= (double)kab * A.dvals[-d - 1];
      }
      if (lane == 0) bl += gc[6] * mref[a];
    }
  }
  for (int o = kGroup / 2; o > 0; o >>= 1) bl += __shfl_xor(bl, o);
  __syncthreads();
  for (int k = lane; k < n; k += kGroup) A.nzval[p0 + k] = acc[g][k];
  if (active && lane == 0) A.b[j] = bl;
}

template <int NLOC, class T>
void launch_gather(Ctx &c, const GatherArgs &A, unsigned grid) {
  if (c.max_col <= 64)
    hipLaunchKernelGGL((simplex_gather_kernel<NLOC, T, 64>), dim3(grid), dim3(kBlock), 0,
                       c.stream, A);
  else if (c.max_col <= 128)
    hipLaunchKernelGGL((simplex_gather_kernel<NLOC, T, 128>), dim3(grid), dim3(kBlock), 0,
                       c.stream, A);
  else
    hipLaunchKernelGGL((simplex_gather_kernel<NLOC, T, 256>), dim3(grid), dim3(kBlock), 0,
                       c.stream, A);
}

}  // namespace

bool simplex_eligible(const Ctx &c) {
  return c.dim == 3 && !c.cartesian && c.nv == 4 && c.reffe_kind == GH_REFFE_PK &&
         c.ncomp == 1 && (c.nloc == 4 || c.nloc == 10);
}

void simplex_symbolic(Ctx &c) {
  c.simplex = false;
  c.adjptr.release();
  c.adj.release();
  if (!simplex_eligible(c) || c.nfree == 0) return;
  hipStream_t s = c.stream;
  const int64_t total = c.ncells * c.nloc;
  GH_REQUIRE(total < (int64_t(1) << 31) && c.ncells < (int64_t(1) << 27), GH_ERR_UNSUPPORTED,
             "mesh too large for the simplex adjacency");
  int rb = 1;
  while ((int64_t(1) << rb) <= c.nfree) ++rb;
  const uint64_t invalid = (uint64_t)c.nfree << 32;
  DBuf<uint64_t> keys, sorted;
  keys.resize(total);
  sorted.resize(total);
  hipLaunchKernelGGL(dof_cell_keys, dim3(grid_for(total)), dim3(kBlock), 0, s, c.ncells, c.nloc,
                     c.cell_dofs.get(), invalid, keys.get());
  GH_CHECK_HIP(hipGetLastError());
  {
    size_t bytes = 0;
    GH_CHECK_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys.get(), sorted.get(),
                                                   (int)total, 0, 32 + rb, s));
    DBuf<unsigned char> tmp;
    tmp.resize(std::max<size_t>(bytes, 1));
    GH_CHECK_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.get(), bytes, keys.get(), sorted.get(),
                                                   (int)total, 0, 32 + rb, s));
    GH_CHECK_HIP(hipStreamSynchronize(s));
  }
  keys.release();
  c.adjptr.resize(c.nfree + 1);
  c.adj.resize(total);
  hipLaunchKernelGGL(adjacency_fill, dim3(grid_for(total)), dim3(kBlock), 0, s, total, c.nfree,
                     sorted.get(), c.adjptr.get(), c.adj.get());
  DBuf<int> mx;
  mx.resize(1);
  GH_CHECK_HIP(hipMemsetAsync(mx.get(), 0, sizeof(int), s));
  hipLaunchKernelGGL(max_column, dim3(grid_for(c.nfree)), dim3(kBlock), 0, s, c.nfree,
                     c.colptr.get(), mx.get());
  GH_CHECK_HIP(hipGetLastError());
  int m = 0;
  GH_CHECK_HIP(hipMemcpyAsync(&m, mx.get(), sizeof(int), hipMemcpyDeviceToHost, s));
  GH_CHECK_HIP(hipStreamSynchronize(s));
  c.max_col = m;
  GH_REQUIRE(m <= 256, GH_ERR_UNSUPPORTED, "a column holds more than 256 entries");
  c.geo.resize((size_t)c.ncells * kGeoStride);
  c.simplex = true;
}

void assemble_simplex(Ctx &c, int precision) {
  GH_REQUIRE(c.simplex, GH_ERR_STATE, "no simplex adjacency: call gh_symbolic");
  const int NL = c.nloc;
  GH_REQUIRE((int)c.qw.size() == c.nq && (int)c.shp.size() == c.nq * NL &&
                 (int)c.grad.size() == c.nq * NL * 3,
             GH_ERR_STATE, "tabulation does not match the element (dofs per cell)");
  // reference tensor Ĝ^k and m from the uploaded tabulation (FP64)
  static const int pd[6] = {0, 1, 2, 0, 0, 1}, pe[6] = {0, 1, 2, 1, 2, 2};
  std::vector<double> gref(6 * NL * NL, 0.0), mref(NL, 0.0);
  for (int q = 0; q < c.nq; ++q) {
    const double w = c.qw[q];
    const double *G = &c.grad[(size_t)q * NL * 3];
    for (int k = 0; k < 6; ++k)
      for (int a = 0; a < NL; ++a)
        for (int b = 0; b < NL; ++b) {
          double v = G[a * 3 + pd[k]] * G[b * 3 + pe[k]];
          if (pd[k] != pe[k]) v += G[a * 3 + pe[k]] * G[b * 3 + pd[k]];
          gref[(k * NL + a) * NL + b] += w * v;
        }
    for (int a = 0; a < NL; ++a) mref[a] += w * c.shp[(size_t)q * NL + a];
  }
  c.sref.resize(gref.size() + mref.size());
  GH_CHECK_HIP(hipMemcpyAsync(c.sref.get(), gref.data(), sizeof(double) * gref.size(),
                              hipMemcpyHostToDevice, c.stream));
  GH_CHECK_HIP(hipMemcpyAsync(c.sref.get() + gref.size(), mref.data(),
                              sizeof(double) * mref.size(), hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(simplex_geom_kernel, dim3(grid_for(c.ncells)), dim3(kBlock), 0, c.stream,
                     c.ncells, c.cell_nodes.get(), c.xyz.get(),
                     c.have_kappa ? c.kappa.get() : nullptr, c.have_fsrc ? c.fsrc.get() : nullptr,
                     c.geo.get());
  GH_CHECK_HIP(hipGetLastError());
  GatherArgs A{};
  A.nfree = c.nfree;
  A.colptr = c.colptr.get();
  A.rowval = c.rowval.get();
  A.adjptr = c.adjptr.get();
  A.adj = c.adj.get();
  A.cell_dofs = c.cell_dofs.get();
  A.geo = c.geo.get();
  A.gref = c.sref.get();
  A.mref = c.sref.get() + gref.size();
  A.dvals = c.have_dvals ? c.dvals.get() : nullptr;
  A.nzval = c.nzval.get();
  A.b = c.b.get();
  const unsigned grid = (unsigned)((c.nfree + kGroups - 1) / kGroups);
  const bool f32 = precision == GH_PREC_FP32_ELEM;
  if (NL == 10) {
    if (f32) launch_gather<10, float>(c, A, grid);
    else launch_gather<10, double>(c, A, grid);
  } else {
    if (f32) launch_gather<4, float>(c, A, grid);
    else launch_gather<4, double>(c, A, grid);
  }
  GH_CHECK_HIP(hipGetLastError());
  // the host vectors above are read by the copies on the stream: wait for them
  GH_CHECK_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace gh
