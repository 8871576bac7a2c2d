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
// Two passes, each fully parallel and deterministic:
//  * element (simplex_element_kernel): one thread per cell computes DF, det DF and
//    the metric in FP64 and writes the cell's metric record (g_0..g_5 in the element
//    precision and f|det DF| in FP64); block 0 also writes the reference tensor.  No
//    element matrix is stored;
//  * column sum (column_sum_kernel): one workgroup per range of consecutive CSC
//    columns stages the reference tensor and the metric records of the cells its
//    columns touch in LDS, forms each contribution Σ_k g_k Ĝ^k_ab there, and one
//    thread per nonzero adds its contributions in ascending cell order (the order of
//    Gridap's cell loop) in FP64 and writes nzval once.  The right-hand side
//    (f|det DF| m_a minus the lifting of the cell's Dirichlet dofs) is fused into the
//    same pass, one thread per column.
// A column too heavy for one range keeps the int32 contribution lists of the
// symbolic phase and the one-thread-per-nonzero sum (simplex_sum_kernel) over full
// element rows, with simplex_rhs_kernel for b; both forms give the same bits.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>

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

// ---- compact contribution lists ---------------------------------------------------
// The column sum runs one block per range of consecutive columns (bstart: the first
// column of each range; colblk: the range of each column).  A block's nonzeros, its
// contributions (sorted by (column, row), so in nonzero order) and its columns'
// adjacency entries are then three contiguous ranges, and 16 bits address each of
// them relative to the range start:
//   slot of a contribution  = (adjacency position − block's first) << obits | offset,
//                             offset: the entry within the cell (full column row, or
//                             packed upper-triangle entry for simplices);
//   noff of a nonzero       = its first contribution − block's first.

// C[j] = first contribution of column j (j <= nfree).
__global__ void column_counts_kernel(int64_t nfree, const int32_t *__restrict__ colptr,
                                     const int32_t *__restrict__ cstart,
                                     int32_t *__restrict__ C) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j <= nfree) C[j] = cstart[colptr[j]];
}

// Contribution q (emit index t = cell·nloc² + lj·nloc + li) of column J: the position
// of (cell, lj) in the adjacency of J's block and the entry's offset from the
// cell-column base, (cell·nloc + lj)·nloc (full rows, offset li) or, for simplices
// (ns > 0), the packed upper-triangle entry of (min, max) of li, lj.
// With ipos (full rows): instead of the slot, the contribution's position in its
// block's run, stored at the entry's place (adjacency position · nloc + li).
__global__ void contrib_slot_kernel(int64_t ncontrib, const int32_t *__restrict__ bstart,
                                    const int32_t *__restrict__ colblk, int nloc, int lbits,
                                    int obits, int ns, const int32_t *__restrict__ perm,
                                    const int32_t *__restrict__ cell_dofs,
                                    const int32_t *__restrict__ adjptr,
                                    const int32_t *__restrict__ adj,
                                    const int32_t *__restrict__ colptr,
                                    const int32_t *__restrict__ cstart,
                                    uint16_t *__restrict__ cslot, uint16_t *__restrict__ ipos,
                                    int32_t *__restrict__ bad) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncontrib) return;
  const int64_t t = perm[q];
  const int64_t nn = (int64_t)nloc * nloc;
  const int64_t cell = t / nn;
  const int r = (int)(t - cell * nn);
  const int lj = r / nloc, li = r - lj * nloc;
  const int J = cell_dofs[cell * nloc + lj] - 1;
  int lo = adjptr[J], hi = adjptr[J + 1] - 1;
  const int32_t want = (int32_t)((cell << lbits) | lj);
  while (lo < hi) {  // first entry >= want
    const int mid = (lo + hi) >> 1;
    if (adj[mid] < want) lo = mid + 1;
    else hi = mid;
  }
  if (adj[lo] != want) {
    atomicOr(bad, 1);
    return;
  }
  int off = li;
  if (ns) {
    const int a = min(li, lj), b = max(li, lj);
    off = a * nloc - a * (a - 1) / 2 + (b - a);
  }
  const int J0 = bstart[colblk[J]];
  if (ipos) {
    ipos[(int64_t)lo * nloc + li] = (uint16_t)(q - cstart[colptr[J0]]);
    return;
  }
  const int k = lo - adjptr[J0];
  if (((int64_t)k << obits) + off > 65535) atomicOr(bad, 2);
  cslot[q] = (uint16_t)((k << obits) | off);
}

// One wave per column: noff of its nonzeros, relative to its block's first entry.
__global__ void nonzero_offset_kernel(int64_t nfree, const int32_t *__restrict__ bstart,
                                      const int32_t *__restrict__ colblk,
                                      const int32_t *__restrict__ colptr,
                                      const int32_t *__restrict__ cstart,
                                      uint16_t *__restrict__ noff) {
  const int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (j >= nfree) return;
  const int32_t q0 = cstart[colptr[bstart[colblk[j]]]];
  for (int32_t p = colptr[j] + lane; p < colptr[j + 1]; p += 64)
    noff[p] = (uint16_t)(cstart[p] - q0);
}

// ---- numeric -----------------------------------------------------------------------

__device__ __forceinline__ int sym_index(int nloc, int a, int b) {
  const int lo = min(a, b), hi = max(a, b);
  return lo * nloc - lo * (lo - 1) / 2 + (hi - lo);
}

// Simplex element storage for the column sum (metric records): the element matrix of an
// affine simplex is K_ab = Σ_k g_k Ĝ^k_ab, so a cell is stored as its metric
// g_0..g_5 and f|det DF| (one record of kMetricRec entries of T; in FP32 f|det| is a
// double in entries 6-7, in FP64 entry 6, entry 7 zero), after a table of the
// reference tensor as T, Ĝ^k at packed upper-triangle entry e in R[6e + k], padded
// to metric_ref_pad entries.  The column sum forms each entry from the record and the
// table with the same operations, in the same order, as the element kernel's
// matrix output, so both give the same bits.
constexpr int kMetricRec = 8;
__host__ __device__ constexpr int metric_ref_pad(int ns) { return (6 * ns + 7) & ~7; }

template <class T>
using Vec2 = T __attribute__((ext_vector_type(2)));

// Element matrices of kBlock cells: E[c][a][b] = Σ_k g_k(c) Ĝ^k_ab, g = κ|det| DF⁻¹DF⁻ᵀ,
// as full rows, or (METRIC) the cell's metric record and, from block 0, the table.
template <int NLOC, class T, bool METRIC>
__global__ __launch_bounds__(kBlock) void simplex_element_kernel(
    int64_t ncells, const int32_t *__restrict__ cell_nodes, const double *__restrict__ xyz,
    const double *__restrict__ kappa, const double *__restrict__ fsrc,
    const double *__restrict__ gref, T *__restrict__ E, double *__restrict__ fdet) {
  constexpr int NS = NLOC * (NLOC + 1) / 2;
  __shared__ T ref[6][NS];
  __shared__ double sg[METRIC ? 1 : kBlock][6];
  if (!METRIC || blockIdx.x == 0)
    for (int t = threadIdx.x; t < 6 * NS; t += kBlock) {
      const int k = t / NS, e = t - k * NS;
      int a = 0, rem = e;
      while (rem >= NLOC - a) {
        rem -= NLOC - a;
        ++a;
      }
      const T v = (T)gref[(k * NLOC + a) * NLOC + a + rem];
      if (METRIC) E[e * 6 + k] = v;
      else ref[k][e] = v;
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
    const double g[6] = {s * dot(r[0], r[0]), s * dot(r[1], r[1]), s * dot(r[2], r[2]),
                         s * dot(r[0], r[1]), s * dot(r[0], r[2]), s * dot(r[1], r[2])};
    const double fd = (fsrc ? fsrc[c] : 1.0) * ad;
    if constexpr (METRIC) {
      T *rec = E + metric_ref_pad(NS) + c * kMetricRec;
      Vec2<T> *r2 = reinterpret_cast<Vec2<T> *>(rec);
      r2[0] = Vec2<T>{(T)g[0], (T)g[1]};
      r2[1] = Vec2<T>{(T)g[2], (T)g[3]};
      r2[2] = Vec2<T>{(T)g[4], (T)g[5]};
      if constexpr (sizeof(T) == 4) {
        *reinterpret_cast<double *>(rec + 6) = fd;
      } else {
        r2[3] = Vec2<T>{(T)fd, (T)0};
      }
    } else {
#pragma unroll
      for (int k = 0; k < 6; ++k) sg[threadIdx.x][k] = g[k];
      fdet[c] = fd;
    }
  }
  if constexpr (!METRIC) {
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
}

template <class T>
__global__ void simplex_sum_kernel(int64_t nnz, const int32_t *__restrict__ cstart,
                                   const int32_t *__restrict__ perm, const T *__restrict__ E,
                                   bool hexcols, double *__restrict__ nzval) {
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
    for (int u = 0; u < U; ++u) {
      // emit index (cell·81 + lj)·81 + li → the hexahedron's padded column layout
      const int64_t at = hexcols && id[u] >= 0
                             ? (int64_t)(id[u] / 81) * kHexColLd + hex_row_slot(id[u] % 81)
                             : id[u];
      v[u] = id[u] >= 0 ? E[at] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) s += (double)v[u];
  }
  nzval[p] = s;
}

// Right-hand side fused into the column sum of simplices (b = NULL: not fused): per
// adjacency entry (cell, a) the term f|det| m_a (f|det| from the cell's metric record)
// minus the lifting of the cell's Dirichlet dofs, and per column the sum of its terms
// in ascending cell order.
struct FusedRhs {
  const double *mref;
  const uint16_t *dmask;
  const int32_t *cell_dofs;
  const double *dvals;
  double *b;
};

// One block per column range [bstart[b], bstart[b+1]), in two phases:
//  * stage: every contribution of the range (a contiguous run) is formed or read
//    into LDS at its run position;
//  * sum: one thread per nonzero of the range adds its entries, noff[p] up to the
//    next nonzero's, in ascending cell order in FP64 and writes nzval once.
// Bit-identical to simplex_sum_kernel (same entries, same order).  The loads that
// do not depend on each other (slots or entry positions, adjacency, the first R
// nonzero offsets per thread) are all issued before the first wait, so a block
// costs about three memory round trips.
// Simplices (FULL_NLOC = 0, metric records): the metric of each of the range's
// cell-columns (one 32- or 64-byte record read per adjacency entry) and the
// reference table go to LDS, and each contribution forms its entry Σ_k g_k Ĝ^k_e
// from them through its slot (adjacency position, packed entry e).
// Full columns (FULL_NLOC = 81, the vector-Q2 hexahedron layout, kHexColLd entries
// per column; FULL_NLOC = -1, nloc rows per column of the general n-cube element
// matrices): the block walks its cell-columns entry by entry with contiguous reads
// and writes each entry that is a contribution to its run position (cslot holds
// those positions, 0xffff for rows that are not free).
// LDS (maxbase: the most adjacency entries of a block, even; maxq: the most
// contributions), full columns: maxq entries; metric records: maxbase right-hand
// side terms (when fused), the table (metric_ref_pad(ns) entries), 6 metric entries
// per adjacency entry, then maxq entries.
// Entry (row li, column lj) of a cell's element matrix in the vector-Q2 hexahedron's
// padded columns.
__device__ __forceinline__ uint32_t hex_elem_addr(uint32_t cell, uint32_t lj, int li) {
  return (cell * 81u + lj) * (uint32_t)kHexColLd + hex_row_slot(li);
}

// Σ_k g_k Ĝ^k_e over 6 entries of T at g and r, in the element kernel's order.
template <class T>
__device__ __forceinline__ T metric_entry(const T *g, const T *r) {
  T v = (T)0;
#pragma unroll
  for (int k = 0; k < 6; ++k) v += g[k] * r[k];
  return v;
}

// metric_entry over 6 entries at g and r that are 2-entry aligned (3 paired LDS reads each).
template <class T>
__device__ __forceinline__ T metric_entry_lds(const T *g, const T *r) {
  const Vec2<T> *g2 = reinterpret_cast<const Vec2<T> *>(g);
  const Vec2<T> *r2 = reinterpret_cast<const Vec2<T> *>(r);
  const Vec2<T> g0 = g2[0], g1 = g2[1], g2v = g2[2], r0 = r2[0], r1 = r2[1], r2v = r2[2];
  const T gg[6] = {g0.x, g0.y, g1.x, g1.y, g2v.x, g2v.y};
  const T rr[6] = {r0.x, r0.y, r1.x, r1.y, r2v.x, r2v.y};
  return metric_entry(gg, rr);
}

template <class T, int FULL_NLOC>
__global__ __launch_bounds__(kBlock) void column_sum_kernel(
    const int32_t *__restrict__ bstart, int nloc, int lbits, int obits, int ns, int maxbase,
    const int32_t *__restrict__ colptr, const int32_t *__restrict__ bbase,
    const int32_t *__restrict__ adjptr, const int32_t *__restrict__ adj,
    const uint16_t *__restrict__ noff, const uint16_t *__restrict__ cslot,
    const T *__restrict__ E, double *__restrict__ nzval, FusedRhs F) {
  extern __shared__ __align__(16) uint32_t shm[];
  // staged entries per thread and batch (simplex ranges hold at most kBlock·8 contributions)
  constexpr int U = FULL_NLOC != 0 ? 12 : 8;
  // nonzeros per thread whose offsets are prefetched (a simplex range of kBlock·8
  // contributions holds about 3·kBlock nonzeros)
  constexpr int R = FULL_NLOC != 0 ? 8 : 4;
  const int rpad = metric_ref_pad(ns);
  unsigned char *lds = reinterpret_cast<unsigned char *>(shm);
  double *rv = reinterpret_cast<double *>(lds);
  T *refl = reinterpret_cast<T *>(lds + (F.b ? 8 * maxbase : 0));
  T *gl = refl + rpad;
  T *vals = FULL_NLOC != 0 ? reinterpret_cast<T *>(lds) : gl + 6 * maxbase;
  const int64_t j0 = bstart[blockIdx.x], j1 = bstart[blockIdx.x + 1];
  const int a0 = adjptr[j0], na = adjptr[j1] - a0;
  const int32_t q0 = bbase[blockIdx.x], nq = bbase[blockIdx.x + 1] - q0;
  const int32_t p0 = colptr[j0], np = colptr[j1] - p0;
  const int tid = threadIdx.x;
  const int lmask = (1 << lbits) - 1;
  // nonzero offsets of the first R rounds (start and end of each)
  uint32_t ost[R], oen[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = tid + r * kBlock;
    ost[r] = i < np ? noff[p0 + i] : 0u;
    oen[r] = i + 1 < np ? noff[p0 + i + 1] : (uint32_t)nq;
  }
  if constexpr (FULL_NLOC != 0) {
    // full columns: FULL_NLOC > 0, the vector-Q2 hexahedron's padded columns; FULL_NLOC
    // < 0, nloc rows per column of the general n-cube element matrices
    const int fl = FULL_NLOC > 0 ? FULL_NLOC : nloc;
    const uint16_t *ip = cslot + (int64_t)a0 * fl;
    const int ne = na * fl;
    for (int eb = 0; eb < ne; eb += kBlock * U) {
      uint32_t pos[U];
      int32_t ae[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = eb + tid + u * kBlock;
        pos[u] = e < ne ? ip[e] : 0xffffu;
        ae[u] = e < ne ? adj[a0 + e / fl] : 0;
      }
      T v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = eb + tid + u * kBlock;
        const int li = e - (e / fl) * fl;
        const uint32_t cell = (uint32_t)(ae[u] >> lbits), lj = (uint32_t)(ae[u] & lmask);
        const uint32_t at = FULL_NLOC > 0 ? hex_elem_addr(cell, lj, li)
                                          : (cell * (uint32_t)fl + lj) * (uint32_t)fl + li;
        v[u] = pos[u] != 0xffffu ? E[at] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (pos[u] != 0xffffu) vals[pos[u]] = v[u];
    }
  } else {
    const int omask = (1 << obits) - 1;
    const T *recs = E + rpad;
    uint32_t sl[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = tid + u * kBlock;
      sl[u] = q < nq ? cslot[q0 + q] : 0xffffffffu;
    }
    for (int t = tid; t < 6 * ns; t += kBlock) refl[t] = E[t];
    __syncthreads();  // the table is read by the liftings below
    for (int k = tid; k < na; k += kBlock) {
      const int e = adj[a0 + k];
      const uint32_t cell = (uint32_t)(e >> lbits), lj = (uint32_t)(e & lmask);
      const T *rec = recs + (size_t)cell * kMetricRec;
      const Vec2<T> *r2 = reinterpret_cast<const Vec2<T> *>(rec);
      const Vec2<T> x0 = r2[0], x1 = r2[1], x2 = r2[2];
      Vec2<T> *dst = reinterpret_cast<Vec2<T> *>(gl + 6 * k);
      dst[0] = x0;
      dst[1] = x1;
      dst[2] = x2;
      if (F.b) {
        const T g[6] = {x0.x, x0.y, x1.x, x1.y, x2.x, x2.y};
        double v = (sizeof(T) == 4 ? *reinterpret_cast<const double *>(rec + 6)
                                   : (double)rec[6]) * F.mref[lj];
        uint32_t m = F.dvals ? F.dmask[cell] : 0u;
        while (m) {
          const int bb = __ffs(m) - 1;
          m &= m - 1;
          const int d = F.cell_dofs[(size_t)cell * nloc + bb];
          const int lo = min((int)lj, bb), hi = max((int)lj, bb);
          const T ev = metric_entry(g, refl + 6 * (lo * nloc - lo * (lo - 1) / 2 + (hi - lo)));
          v -= (double)ev * F.dvals[-d - 1];
        }
        rv[k] = v;
      }
    }
    __syncthreads();
    for (int qb = 0; qb < nq; qb += kBlock * U) {
      if (qb) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int q = qb + tid + u * kBlock;
          sl[u] = q < nq ? cslot[q0 + q] : 0xffffffffu;
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (sl[u] != 0xffffffffu)
          vals[qb + tid + u * kBlock] =
              metric_entry_lds(gl + 6 * (sl[u] >> obits), refl + 6 * (sl[u] & omask));
    }
  }
  __syncthreads();
  // the range's column sums of b on the last threads, which hold the fewest nonzeros
  if (F.b && tid >= kBlock - (j1 - j0)) {
    const int64_t J = j0 + (tid - (kBlock - (j1 - j0)));
    double bs = 0.0;
    for (int k = adjptr[J] - a0, ke = adjptr[J + 1] - a0; k < ke; ++k) bs += rv[k];
    F.b[J] = bs;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = tid + r * kBlock;
    if (i < np) {
      double s = 0.0;
      for (uint32_t k = ost[r]; k < oen[r]; ++k) s += (double)vals[k];
      nzval[p0 + i] = s;
    }
  }
  for (int i = tid + R * kBlock; i < np; i += kBlock) {
    const int st = noff[p0 + i];
    const int en = i + 1 < np ? (int)noff[p0 + i + 1] : nq;
    double s = 0.0;
    for (int k = st; k < en; ++k) s += (double)vals[k];
    nzval[p0 + i] = s;
  }
}

// b from full-row element matrices, one thread per column (when the column sum does
// not fuse it).
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

}  // namespace

bool elemlist_sum(Ctx &c, bool f32, const FusedRhs *rhs);

namespace {

template <int NLOC, class T, bool METRIC>
void launch_simplex(Ctx &c) {
  constexpr int NS = NLOC * (NLOC + 1) / 2;
  c.elem.resize(METRIC ? ((size_t)metric_ref_pad(NS) + (size_t)c.ncells * kMetricRec) * sizeof(T)
                       : (size_t)c.ncells * NLOC * NLOC * sizeof(T));
  T *E = reinterpret_cast<T *>(c.elem.get());
  hipLaunchKernelGGL((simplex_element_kernel<NLOC, T, METRIC>), dim3(grid_for(c.ncells)),
                     dim3(kBlock), 0, c.stream, c.ncells, c.cell_nodes.get(), c.xyz.get(),
                     c.have_kappa ? c.kappa.get() : nullptr,
                     c.have_fsrc ? c.fsrc.get() : nullptr, c.sref.get(), E, c.fdet.get());
  const FusedRhs F{c.sref.get() + 6 * NLOC * NLOC, c.dmask.get(), c.cell_dofs.get(),
                   c.have_dvals ? c.dvals.get() : nullptr, c.b.get()};
  const bool fused = elemlist_sum(c, sizeof(T) == 4, &F);
  GH_REQUIRE(fused == METRIC, GH_ERR_STATE, "simplex storage does not match the sum");
  if (!fused)
    hipLaunchKernelGGL((simplex_rhs_kernel<NLOC, T>), dim3(grid_for(c.nfree)),
                       dim3(kBlock), 0, c.stream, c.nfree, c.adjptr.get(), c.adj.get(),
                       c.fdet.get(), c.sref.get() + 6 * NLOC * NLOC, c.dmask.get(),
                       c.cell_dofs.get(), (const T *)E, c.have_dvals ? c.dvals.get() : nullptr,
                       c.b.get());
  GH_CHECK_HIP(hipGetLastError());
}

}  // namespace

bool simplex_eligible(const Ctx &c) {
  return c.dim == 3 && !c.cartesian && c.nv == 4 && c.reffe_kind == GH_REFFE_PK &&
         c.ncomp == 1 && (c.nloc == 4 || c.nloc == 10);
}

// Replace cstart/perm by the compact lists when they fit: 16-bit slots and offsets
// within a block's ranges, and hexahedron element storage addressable in 32 bits.
// Simplices then store metric records.  The column ranges are cut greedily in column
// order: a range grows while it holds at most kBlock columns and a budget of
// contributions (kBlock·8 for simplices, kBlock·12 for full columns) and an FP32 LDS
// target, and its extents fit 16 bits and the LDS (FP32 within kMaxColumnLds; FP64
// within the device's per-workgroup limit, at most kMaxColumnLds64).  Budgeted ranges keep the LDS of a
// block near the mean: the numbering puts the vertex dofs, which touch many more cells
// than edge or face dofs, first, and a fixed column count per range sized the LDS of
// every block for the heaviest.
constexpr size_t kMaxColumnLds = 48 * 1024;
constexpr size_t kMaxColumnLds64 = 64 * 1024;
// A range also stops growing once its FP32 LDS would pass kColumnLdsTarget, so that 8
// workgroups (32 waves) fit the 160 KB of a CU; a single column may still exceed it.
constexpr size_t kColumnLdsTarget = 18 * 1024;

// maxbase as the kernel takes it: at least 2 and even (metric entries stay 16-byte aligned)
int lds_maxbase(int maxbase) { return (std::max(maxbase, 1) + 1) & ~1; }

size_t column_sum_lds(bool metric, int64_t maxbase, int64_t maxq, size_t esz, bool rhs, int ns) {
  const size_t mb = (size_t)lds_maxbase((int)std::min<int64_t>(maxbase, 1 << 30));
  const size_t mq = (size_t)std::max<int64_t>(maxq, 1);
  if (!metric) return esz * mq;
  return (rhs ? sizeof(double) * mb : 0) + esz * ((size_t)metric_ref_pad(ns) + 6 * mb + mq);
}

void compact_contributions(Ctx &c, int lbits) {
  c.slots = false;
  c.metric = false;
  c.inverse = false;
  c.cslot.release();
  c.noff.release();
  c.bbase.release();
  c.bstart.release();
  if (!c.nnz || !c.cstart.get() || !c.perm.get()) return;
  // GRIDAP_HIP_COMPACT_LISTS=0 keeps the int32 lists (the fallback sum), so the two
  // sums can be compared on the same system
  if (const char *env = std::getenv("GRIDAP_HIP_COMPACT_LISTS"))
    if (env[0] == '0') return;
  hipStream_t s = c.stream;
  const int64_t N = c.nfree;
  // simplices: metric records; otherwise full columns (the padded vector-Q2 layout of
  // the Cartesian hexahedron, or nloc rows per column of the general n-cube element
  // matrices), addressed in 32 bits
  const bool metric = simplex_eligible(c);
  const bool padded = hexq2_eligible(c);
  // general n-cube elements: full columns for the large ones (vector Q2: the mapped 32³
  // sum 0.95 -> 0.31 ms); the lists for the small ones, whose gathers stay near each
  // other (scalar Q2 48³: 0.37 ms with the lists, 0.79 ms with full columns)
  if (!metric && !padded && !(hexgen_eligible(c) && c.nloc >= 48)) return;
  const int ns = metric ? c.nloc * (c.nloc + 1) / 2 : 0;
  int obits = 1;
  while ((1 << obits) < (metric ? ns : c.nloc)) ++obits;
  if (!metric && c.ncells * (int64_t)c.nloc * (padded ? kHexColLd : c.nloc) >= (int64_t(1) << 32))
    return;
  int lds64 = 0;
  GH_CHECK_HIP(hipDeviceGetAttribute(&lds64, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                     c.device));
  const size_t max64 = std::min(kMaxColumnLds64, (size_t)lds64);
  // first contribution (C) and first adjacency entry (A) of every column, on the host
  std::vector<int32_t> C(N + 1), A(N + 1);
  {
    DBuf<int32_t> cd;
    cd.resize(N + 1);
    hipLaunchKernelGGL(column_counts_kernel, dim3(grid_for(N + 1)), dim3(kBlock), 0, s, N,
                       c.colptr.get(), c.cstart.get(), cd.get());
    GH_CHECK_HIP(hipGetLastError());
    GH_CHECK_HIP(hipMemcpyAsync(C.data(), cd.get(), sizeof(int32_t) * (N + 1),
                                hipMemcpyDeviceToHost, s));
    GH_CHECK_HIP(hipMemcpyAsync(A.data(), c.adjptr.get(), sizeof(int32_t) * (N + 1),
                                hipMemcpyDeviceToHost, s));
    GH_CHECK_HIP(hipStreamSynchronize(s));
  }
  const int64_t ncontrib = C[N];
  const int64_t budget = (int64_t)kBlock * (metric ? 8 : 12);
  auto fits = [&](int64_t a, int64_t q) {
    return (a << obits) <= 65536 && q <= 65535 &&
           column_sum_lds(metric, a, q, 4, metric, ns) <= kMaxColumnLds &&
           column_sum_lds(metric, a, q, 8, metric, ns) <= (metric ? max64 : kMaxColumnLds);
  };
  std::vector<int32_t> bs{0}, bq;
  bs.reserve(N / 4 + 2);
  int64_t maxa = 0, maxq = 0, jb = 0;
  for (int64_t j = 1; j <= N; ++j) {
    if (j < N) {  // extend [jb, j) to [jb, j + 1)?
      const int64_t a = A[j + 1] - A[jb], q = C[j + 1] - C[jb];
      if (j - jb < kBlock && q <= budget && fits(a, q) &&
          column_sum_lds(metric, a, q, 4, metric, ns) <= kColumnLdsTarget)
        continue;
    }
    const int64_t a = A[j] - A[jb], q = C[j] - C[jb];
    if (!fits(a, q)) return;  // one column too large: keep cstart/perm and full rows
    maxa = std::max(maxa, a);
    maxq = std::max(maxq, q);
    bs.push_back((int32_t)j);
    jb = j;
  }
  const int64_t nblocks = (int64_t)bs.size() - 1;
  std::vector<int32_t> colblk(N), hb(nblocks + 1);
  for (int64_t b = 0; b < nblocks; ++b) {
    hb[b] = C[bs[b]];
    for (int64_t j = bs[b]; j < bs[b + 1]; ++j) colblk[j] = (int32_t)b;
  }
  hb[nblocks] = C[N];
  DBuf<uint16_t> cslot, noff;
  DBuf<int32_t> bbase, bstart, cblk, dev;
  bstart.resize(nblocks + 1);
  bbase.resize(nblocks + 1);
  cblk.resize(N);
  dev.resize(1);
  GH_CHECK_HIP(hipMemcpyAsync(bstart.get(), bs.data(), sizeof(int32_t) * (nblocks + 1),
                              hipMemcpyHostToDevice, s));
  GH_CHECK_HIP(hipMemcpyAsync(bbase.get(), hb.data(), sizeof(int32_t) * (nblocks + 1),
                              hipMemcpyHostToDevice, s));
  GH_CHECK_HIP(hipMemcpyAsync(cblk.get(), colblk.data(), sizeof(int32_t) * N,
                              hipMemcpyHostToDevice, s));
  // full columns: the inverse form (ipos: one run position per (free adjacency
  // entry, row)); simplices: per-contribution slots
  const bool inverse = !metric;
  const int64_t nslot = inverse ? (int64_t)A[N] * c.nloc : ncontrib;
  cslot.resize(std::max<int64_t>(nslot, 1));
  if (inverse)
    GH_CHECK_HIP(hipMemsetAsync(cslot.get(), 0xff, sizeof(uint16_t) * cslot.n, s));
  noff.resize(c.nnz);
  GH_CHECK_HIP(hipMemsetAsync(dev.get(), 0, sizeof(int32_t), s));
  if (ncontrib)
    hipLaunchKernelGGL(contrib_slot_kernel, dim3(grid_for(ncontrib)), dim3(kBlock), 0, s,
                       ncontrib, bstart.get(), cblk.get(), c.nloc, lbits, obits, ns,
                       c.perm.get(), c.cell_dofs.get(), c.adjptr.get(), c.adj.get(),
                       c.colptr.get(), c.cstart.get(), inverse ? nullptr : cslot.get(),
                       inverse ? cslot.get() : nullptr, dev.get());
  hipLaunchKernelGGL(nonzero_offset_kernel, dim3(grid_for(N * 64)), dim3(kBlock), 0, s, N,
                     bstart.get(), cblk.get(), c.colptr.get(), c.cstart.get(), noff.get());
  GH_CHECK_HIP(hipGetLastError());
  int32_t bad = 0;
  GH_CHECK_HIP(hipMemcpyAsync(&bad, dev.get(), 4, hipMemcpyDeviceToHost, s));
  GH_CHECK_HIP(hipStreamSynchronize(s));
  if (bad) return;
  std::swap(c.cslot.p, cslot.p);
  std::swap(c.cslot.n, cslot.n);
  std::swap(c.noff.p, noff.p);
  std::swap(c.noff.n, noff.n);
  std::swap(c.bbase.p, bbase.p);
  std::swap(c.bbase.n, bbase.n);
  std::swap(c.bstart.p, bstart.p);
  std::swap(c.bstart.n, bstart.n);
  c.cstart.release();
  c.perm.release();
  c.maxbase = (int)maxa;
  c.maxq = (int)maxq;
  c.obits = obits;
  c.metric = metric;
  c.inverse = inverse;
  c.slots = true;
  c.elem.release();
}

void elemlist_symbolic(Ctx &c) {
  c.elemlist = false;
  c.adjptr.release();
  c.adj.release();
  c.dmask.release();
  c.elem.release();
  if (!(simplex_eligible(c) || hexq2_eligible(c) || hexgen_eligible(c)) || c.nfree == 0) return;
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
  compact_contributions(c, lbits);
}

template <class T>
void launch_column_sum(Ctx &c, const FusedRhs &F) {
  const int lbits = c.nloc <= 16 ? 4 : 7;
  const int ns = c.metric ? c.nloc * (c.nloc + 1) / 2 : 0;
  const size_t shm =
      column_sum_lds(c.metric, c.maxbase, c.maxq, sizeof(T), F.b != nullptr, ns);
  const int64_t nblocks = (int64_t)c.bstart.n - 1;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(kBlock), shm, c.stream, c.bstart.get(),
                       c.nloc, lbits, c.obits, ns, lds_maxbase(c.maxbase),
                       c.colptr.get(), c.bbase.get(), c.adjptr.get(), c.adj.get(),
                       c.noff.get(), c.cslot.get(), (const T *)c.elem.get(), c.nzval.get(), F);
  };
  if (!c.inverse) go(column_sum_kernel<T, 0>);
  else if (hexq2_eligible(c)) go(column_sum_kernel<T, 81>);
  else go(column_sum_kernel<T, -1>);
}

// nzval from the element entries; with rhs (metric records only) also b, fused
bool elemlist_sum(Ctx &c, bool f32, const FusedRhs *rhs) {
  if (!c.nnz) return false;
  const bool fuse = rhs && c.slots && c.metric;
  FusedRhs F{};
  if (fuse) F = *rhs;
  c.last_kernel = c.slots ? GH_KERNEL_COLUMN_SUM : GH_KERNEL_LIST_SUM;
  if (c.slots) {
    if (f32) launch_column_sum<float>(c, F);
    else launch_column_sum<double>(c, F);
  } else if (f32)
    hipLaunchKernelGGL((simplex_sum_kernel<float>), dim3(grid_for(c.nnz)), dim3(kBlock), 0,
                       c.stream, c.nnz, c.cstart.get(), c.perm.get(),
                       (const float *)c.elem.get(), hexq2_eligible(c), c.nzval.get());
  else
    hipLaunchKernelGGL((simplex_sum_kernel<double>), dim3(grid_for(c.nnz)), dim3(kBlock), 0,
                       c.stream, c.nnz, c.cstart.get(), c.perm.get(),
                       (const double *)c.elem.get(), hexq2_eligible(c), c.nzval.get());
  GH_CHECK_HIP(hipGetLastError());
  return fuse;
}

void elemlist_sum(Ctx &c, bool f32) { (void)elemlist_sum(c, f32, nullptr); }

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
  if (c.metric) {
    GH_REQUIRE(c.slots, GH_ERR_STATE, "metric element storage without compact lists");
    if (NL == 10) {
      if (f32) launch_simplex<10, float, true>(c);
      else launch_simplex<10, double, true>(c);
    } else {
      if (f32) launch_simplex<4, float, true>(c);
      else launch_simplex<4, double, true>(c);
    }
  } else if (NL == 10) {
    if (f32) launch_simplex<10, float, false>(c);
    else launch_simplex<10, double, false>(c);
  } else {
    if (f32) launch_simplex<4, float, false>(c);
    else launch_simplex<4, double, false>(c);
  }
  // enqueued only: gh_values_get (or a wait on gh_stream) orders the results
}

}  // namespace gh
