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
// Three passes, each fully parallel and deterministic:
//  * element: one thread per cell computes the metric (FP64); the block then writes
//    every cell's element matrix (row-major, nloc x nloc) with coalesced stores, so
//    the entries a column needs from a cell form one contiguous row;
//  * sum: one thread per nonzero adds its entries in ascending cell order (the
//    order of Gridap's cell loop), from the contribution list that the symbolic
//    phase keeps (a stable sort of the (column, row) keys carries the cell index),
//    in FP64, and writes nzval once;
//  * rhs: one thread per free dof walks its cells (dof -> cell adjacency, ascending)
//    and adds f|det| m_a minus the lifting of the cell's Dirichlet dofs.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "gh_internal.h"

namespace gh {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int block = kBlock) {
  return (unsigned)std::max<int64_t>(1, (n + block - 1) / block);
}

// ---- symbolic: dof -> (cell, local index), ascending cells ------------------------

__global__ void dof_cell_keys(int64_t ncells, int nloc, int lbits,
                              const int32_t *__restrict__ cell_dofs, uint64_t invalid,
                              uint64_t *__restrict__ keys) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells * nloc) return;
  const int64_t cell = t / nloc;
  const int a = (int)(t - cell * nloc);
  const int d = cell_dofs[t];
  keys[t] = d > 0 ? ((uint64_t)(d - 1) << 32) | (uint64_t)((cell << lbits) | a) : invalid;
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

__global__ void dirichlet_mask_kernel(int64_t ncells, int nloc,
                                      const int32_t *__restrict__ cell_dofs,
                                      uint16_t *__restrict__ dmask) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  // bit a: local dof a is Dirichlet (nloc <= 16); else 1 when any is
  uint32_t m = 0;
  for (int a = 0; a < nloc; ++a)
    if (cell_dofs[c * nloc + a] < 0) m |= nloc <= 16 ? 1u << a : 1u;
  dmask[c] = (uint16_t)m;
}

// ---- numeric -----------------------------------------------------------------------

__device__ __forceinline__ int sym_index(int nloc, int a, int b) {
  const int lo = min(a, b), hi = max(a, b);
  return lo * nloc - lo * (lo - 1) / 2 + (hi - lo);
}

// Element matrices of kBlock cells: E[c][a][b] = Σ_k g_k(c) Ĝ^k_ab, g = κ|det| DF⁻¹DF⁻ᵀ.
template <int NLOC, class T>
__global__ __launch_bounds__(kBlock) void simplex_element_kernel(
    int64_t ncells, const int32_t *__restrict__ cell_nodes, const double *__restrict__ xyz,
    const double *__restrict__ kappa, const double *__restrict__ fsrc,
    const double *__restrict__ gref, T *__restrict__ E, double *__restrict__ fdet) {
  constexpr int NS = NLOC * (NLOC + 1) / 2;
  __shared__ T ref[6][NS];
  __shared__ double sg[kBlock][6];
  for (int t = threadIdx.x; t < 6 * NS; t += kBlock) {
    const int k = t / NS, e = t - k * NS;
    int a = 0, rem = e;
    while (rem >= NLOC - a) {
      rem -= NLOC - a;
      ++a;
    }
    ref[k][e] = (T)gref[(k * NLOC + a) * NLOC + a + rem];
  }
  const int64_t c0 = (int64_t)blockIdx.x * kBlock;
  const int64_t c = c0 + threadIdx.x;
  if (c < ncells) {
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
    auto cross = [](const double *u, const double *w, double *o) {
      o[0] = u[1] * w[2] - u[2] * w[1];
      o[1] = u[2] * w[0] - u[0] * w[2];
      o[2] = u[0] * w[1] - u[1] * w[0];
    };
    double r[3][3];  // rows of DF⁻¹ times det
    cross(e[1], e[2], r[0]);
    cross(e[2], e[0], r[1]);
    cross(e[0], e[1], r[2]);
    const double det = e[0][0] * r[0][0] + e[0][1] * r[0][1] + e[0][2] * r[0][2];
    const double ad = fabs(det);
    const double s = (kappa ? kappa[c] : 1.0) * ad / (det * det);
    auto dot = [](const double *u, const double *w) {
      return u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
    };
    sg[threadIdx.x][0] = s * dot(r[0], r[0]);
    sg[threadIdx.x][1] = s * dot(r[1], r[1]);
    sg[threadIdx.x][2] = s * dot(r[2], r[2]);
    sg[threadIdx.x][3] = s * dot(r[0], r[1]);
    sg[threadIdx.x][4] = s * dot(r[0], r[2]);
    sg[threadIdx.x][5] = s * dot(r[1], r[2]);
    fdet[c] = (fsrc ? fsrc[c] : 1.0) * ad;
  }
  __syncthreads();
  const int64_t ncb = min<int64_t>(kBlock, ncells - c0);
  constexpr int NN = NLOC * NLOC;
  T *out = E + c0 * NN;
  for (int idx = threadIdx.x; idx < ncb * NN; idx += kBlock) {
    const int cl = idx / NN, ab = idx - cl * NN;
    const int ee = sym_index(NLOC, ab / NLOC, ab % NLOC);
    T v = (T)0;
#pragma unroll
    for (int k = 0; k < 6; ++k) v += (T)sg[cl][k] * ref[k][ee];
    out[idx] = v;
  }
}

template <class T>
__global__ void simplex_sum_kernel(int64_t nnz, const int32_t *__restrict__ cstart,
                                   const int32_t *__restrict__ perm, const T *__restrict__ E,
                                   double *__restrict__ nzval) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nnz) return;
  // batches of 8: all index loads, then all value loads, then the ordered sum
  constexpr int U = 8;
  double s = 0.0;
  const int32_t e = cstart[p + 1];
  for (int32_t i0 = cstart[p]; i0 < e; i0 += U) {
    int32_t id[U];
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) id[u] = i0 + u < e ? perm[i0 + u] : -1;
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = id[u] >= 0 ? E[id[u]] : (T)0;
#pragma unroll
    for (int u = 0; u < U; ++u) s += (double)v[u];
  }
  nzval[p] = s;
}

template <int NLOC, class T>
__global__ void simplex_rhs_kernel(int64_t nfree, const int32_t *__restrict__ adjptr,
                                   const int32_t *__restrict__ adj,
                                   const double *__restrict__ fdet,
                                   const double *__restrict__ mref,
                                   const uint16_t *__restrict__ dmask,
                                   const int32_t *__restrict__ cell_dofs,
                                   const T *__restrict__ E, const double *__restrict__ dvals,
                                   double *__restrict__ b) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nfree) return;
  double s = 0.0;
  for (int q = adjptr[j]; q < adjptr[j + 1]; ++q) {
    const int e = adj[q];
    const int cell = e >> 4, a = e & 15;
    double v = fdet[cell] * mref[a];
    uint32_t m = dvals ? dmask[cell] : 0u;
    while (m) {
      const int bb = __ffs(m) - 1;
      m &= m - 1;
      const int d = cell_dofs[(size_t)cell * NLOC + bb];
      v -= (double)E[((size_t)cell * NLOC + a) * NLOC + bb] * dvals[-d - 1];
    }
    s += v;
  }
  b[j] = s;
}

template <int NLOC, class T>
void launch_simplex(Ctx &c) {
  c.elem.resize((size_t)c.ncells * NLOC * NLOC * sizeof(T));
  T *E = reinterpret_cast<T *>(c.elem.get());
  hipLaunchKernelGGL((simplex_element_kernel<NLOC, T>), dim3(grid_for(c.ncells)), dim3(kBlock),
                     0, c.stream, c.ncells, c.cell_nodes.get(), c.xyz.get(),
                     c.have_kappa ? c.kappa.get() : nullptr,
                     c.have_fsrc ? c.fsrc.get() : nullptr, c.sref.get(), E, c.fdet.get());
  elemlist_sum(c, sizeof(T) == 4);
  hipLaunchKernelGGL((simplex_rhs_kernel<NLOC, T>), dim3(grid_for(c.nfree)), dim3(kBlock), 0,
                     c.stream, c.nfree, c.adjptr.get(), c.adj.get(), c.fdet.get(),
                     c.sref.get() + 6 * NLOC * NLOC, c.dmask.get(), c.cell_dofs.get(),
                     (const T *)E, c.have_dvals ? c.dvals.get() : nullptr, c.b.get());
  GH_CHECK_HIP(hipGetLastError());
}

}  // namespace

bool simplex_eligible(const Ctx &c) {
  return c.dim == 3 && !c.cartesian && c.nv == 4 && c.reffe_kind == GH_REFFE_PK &&
         c.ncomp == 1 && (c.nloc == 4 || c.nloc == 10);
}

void elemlist_symbolic(Ctx &c) {
  c.elemlist = false;
  c.adjptr.release();
  c.adj.release();
  c.dmask.release();
  c.elem.release();
  if (!(simplex_eligible(c) || hexq2_eligible(c)) || c.nfree == 0) return;
  hipStream_t s = c.stream;
  const int64_t total = c.ncells * c.nloc;
  const int lbits = c.nloc <= 16 ? 4 : 7;
  GH_REQUIRE(total < (int64_t(1) << 31) && c.ncells < (int64_t(1) << (31 - lbits)),
             GH_ERR_UNSUPPORTED, "mesh too large for the element-list adjacency");
  int rb = 1;
  while ((int64_t(1) << rb) <= c.nfree) ++rb;
  const uint64_t invalid = (uint64_t)c.nfree << 32;
  DBuf<uint64_t> keys, sorted;
  keys.resize(total);
  sorted.resize(total);
  hipLaunchKernelGGL(dof_cell_keys, dim3(grid_for(total)), dim3(kBlock), 0, s, c.ncells, c.nloc,
                     lbits, c.cell_dofs.get(), invalid, keys.get());
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
  c.dmask.resize(c.ncells);
  hipLaunchKernelGGL(dirichlet_mask_kernel, dim3(grid_for(c.ncells)), dim3(kBlock), 0, s,
                     c.ncells, c.nloc, c.cell_dofs.get(), c.dmask.get());
  GH_CHECK_HIP(hipGetLastError());
  GH_CHECK_HIP(hipStreamSynchronize(s));
  c.fdet.resize(c.ncells);
  c.elemlist = true;
}

void elemlist_sum(Ctx &c, bool f32) {
  if (!c.nnz) return;
  if (f32)
    hipLaunchKernelGGL((simplex_sum_kernel<float>), dim3(grid_for(c.nnz)), dim3(kBlock), 0,
                       c.stream, c.nnz, c.cstart.get(), c.perm.get(),
                       (const float *)c.elem.get(), c.nzval.get());
  else
    hipLaunchKernelGGL((simplex_sum_kernel<double>), dim3(grid_for(c.nnz)), dim3(kBlock), 0,
                       c.stream, c.nnz, c.cstart.get(), c.perm.get(),
                       (const double *)c.elem.get(), c.nzval.get());
  GH_CHECK_HIP(hipGetLastError());
}

void assemble_simplex(Ctx &c, int precision) {
  GH_REQUIRE(c.elemlist && simplex_eligible(c), GH_ERR_STATE,
             "no simplex adjacency: call gh_symbolic");
  const int NL = c.nloc;
  GH_REQUIRE((int)c.qw.size() == c.nq && (int)c.shp.size() == c.nq * NL &&
                 (int)c.grad.size() == c.nq * NL * 3,
             GH_ERR_STATE, "tabulation does not match the element (dofs per cell)");
  // reference tensor Ĝ^k and m from the uploaded tabulation (FP64), once per tabulation
  const size_t ng = (size_t)6 * NL * NL;
  if (!c.sref_ready) {
    static const int pd[6] = {0, 1, 2, 0, 0, 1}, pe[6] = {0, 1, 2, 1, 2, 2};
    std::vector<double> ref(ng + NL, 0.0);
    for (int q = 0; q < c.nq; ++q) {
      const double w = c.qw[q];
      const double *G = &c.grad[(size_t)q * NL * 3];
      for (int k = 0; k < 6; ++k)
        for (int a = 0; a < NL; ++a)
          for (int b = 0; b < NL; ++b) {
            double v = G[a * 3 + pd[k]] * G[b * 3 + pe[k]];
            if (pd[k] != pe[k]) v += G[a * 3 + pe[k]] * G[b * 3 + pd[k]];
            ref[(k * NL + a) * NL + b] += w * v;
          }
      for (int a = 0; a < NL; ++a) ref[ng + a] += w * c.shp[(size_t)q * NL + a];
    }
    c.sref.resize(ref.size());
    GH_CHECK_HIP(hipMemcpy(c.sref.get(), ref.data(), sizeof(double) * ref.size(),
                           hipMemcpyHostToDevice));
    c.sref_ready = true;
  }
  const bool f32 = precision == GH_PREC_FP32_ELEM;
  if (NL == 10) {
    if (f32) launch_simplex<10, float>(c);
    else launch_simplex<10, double>(c);
  } else {
    if (f32) launch_simplex<4, float>(c);
    else launch_simplex<4, double>(c);
  }
  // enqueued only: gh_values_get (or a wait on gh_stream) orders the results
}

}  // namespace gh
