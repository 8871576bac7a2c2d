// Symbolic phase on the device: the CSC pattern of create_from_nz.
//
// Gridap (src/Algebra/SparseMatrixCSC.jl:72-283) counts, allocates, inserts each
// (i,j) touched by a cell into column j by binary search and compacts: the result
// is, per column, the sorted unique rows of every (i>0, j>0) pair of every cell,
// explicit zeros kept.  Here every cell emits its pairs as packed (col,row) keys,
// one radix sort orders them, and a unique-flag scan writes rowval/colptr.  The
// order is total, so the pattern is bit-identical whatever the launch order.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <vector>

#include "gh_internal.h"

namespace gh {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)std::max<int64_t>(1, g);
}

int bits_for(int64_t v) {  // bits needed to hold values 0..v
  int b = 1;
  while ((int64_t(1) << b) <= v) ++b;
  return b;
}

// One thread per (cell, lj, li): the pairs of _add_entries!
// (src/Algebra/AlgebraInterfaces.jl:175-187), Dirichlet ids skipped.
__global__ void emit_pairs(int64_t ncells, int nloc, const int32_t *__restrict__ cell_dofs,
                           int rb, uint64_t invalid, uint64_t *__restrict__ keys) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t per = (int64_t)nloc * nloc;
  if (t >= ncells * per) return;
  int64_t c = t / per;
  int r = (int)(t - c * per);
  int lj = r / nloc, li = r - lj * nloc;
  int32_t j = cell_dofs[c * nloc + lj];
  int32_t i = cell_dofs[c * nloc + li];
  keys[t] = (i > 0 && j > 0) ? ((uint64_t)(j - 1) << rb) | (uint64_t)(i - 1) : invalid;
}

__global__ void unique_flags(int64_t n, const uint64_t *__restrict__ keys, uint64_t invalid,
                             int32_t *__restrict__ flags) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  uint64_t key = keys[k];
  flags[k] = key < invalid && (k == 0 || keys[k - 1] != key);
}

// colptr from the sorted unique keys: column j starts at the unique position of its
// first key; colptr[ncols] = nnz.  A column with no entry is flagged (every free
// dof holds at least its diagonal entry).
__global__ void column_starts(int64_t n, const uint64_t *__restrict__ keys, uint64_t invalid,
                              int rb, const int32_t *__restrict__ flags,
                              const int32_t *__restrict__ pos, int32_t *__restrict__ colptr) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || !flags[k]) return;
  const uint64_t col = keys[k] >> rb;
  if (k == 0 || (keys[k - 1] >> rb) != col) colptr[col] = pos[k];
}

__global__ void column_check(int64_t ncols, int64_t nnz, int32_t *__restrict__ colptr,
                             int32_t *__restrict__ bad) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j > ncols) return;
  if (j == ncols) {
    colptr[j] = (int32_t)nnz;
    return;
  }
  if (colptr[j] < 0) atomicOr(bad, 1);
}

// Contribution ranges of the nonzeros (CSC order): cstart[pos[k]] = k for the first
// entry of each nonzero; cstart[nnz] = number of entries with a free row and column.
__global__ void contrib_ranges(int64_t n, const uint64_t *__restrict__ keys, uint64_t invalid,
                               const int32_t *__restrict__ flags, const int32_t *__restrict__ pos,
                               int64_t nnz, int32_t *__restrict__ cstart) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (flags[k]) cstart[pos[k]] = (int32_t)k;
  if (keys[k] < invalid && (k + 1 == n || keys[k + 1] >= invalid)) cstart[nnz] = (int32_t)(k + 1);
}

__global__ void iota_kernel(int64_t n, int32_t *__restrict__ v) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) v[k] = (int32_t)k;
}

__global__ void scatter_rows(int64_t n, const uint64_t *__restrict__ keys,
                             const int32_t *__restrict__ flags, const int32_t *__restrict__ pos,
                             uint64_t rowmask, int32_t *__restrict__ rowval) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || !flags[k]) return;
  rowval[pos[k]] = (int32_t)(keys[k] & rowmask);
}

// ---- lattice eligibility (Cartesian mesh, scalar Q1) ----

__device__ __forceinline__ int64_t cart_node(int D, const int64_t *np, int64_t cell,
                                             const int64_t *nc, int a) {
  int64_t node = 0, stride = 1, r = cell;
  for (int d = 0; d < D; ++d) {
    int64_t ci = r % nc[d];
    r /= nc[d];
    node += (ci + ((a >> d) & 1)) * stride;
    stride *= np[d];
  }
  return node;
}

struct Lattice {
  int D;
  int64_t nc[3], np[3];
};

__global__ void lattice_scatter(Lattice L, int64_t ncells, int nloc,
                                const int32_t *__restrict__ cell_dofs,
                                int32_t *__restrict__ node_dof) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells * nloc) return;
  int64_t c = t / nloc;
  int a = (int)(t - c * nloc);
  node_dof[cart_node(L.D, L.np, c, L.nc, a)] = cell_dofs[t];
}

__global__ void lattice_verify(Lattice L, int64_t ncells, int nloc,
                               const int32_t *__restrict__ cell_dofs,
                               const int32_t *__restrict__ node_dof, int32_t *bad) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells * nloc) return;
  int64_t c = t / nloc;
  int a = (int)(t - c * nloc);
  if (node_dof[cart_node(L.D, L.np, c, L.nc, a)] != cell_dofs[t]) atomicOr(bad, 1);
}

__global__ void free_flags(int64_t nnodes, const int32_t *__restrict__ node_dof,
                           int32_t *__restrict__ f, int32_t *bad) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes) return;
  int32_t d = node_dof[n];
  if (d == 0) atomicOr(bad, 2);
  f[n] = d > 0;
}

// Free dofs must be numbered in node order (CLagrangian numbering) and the
// pattern column of node n must be the sorted free lattice neighbours of n.
// Also records, per free node, the bit mask of free stencil neighbours (bit o for
// offset o = Σ_k (δ_k+1) 3^k) that the numeric lattice kernel uses in place of
// the 3^D neighbour loads.
__global__ void lattice_check_pattern(Lattice L, int64_t nnodes,
                                      const int32_t *__restrict__ node_dof,
                                      const int32_t *__restrict__ scan,
                                      const int32_t *__restrict__ colptr,
                                      const int32_t *__restrict__ rowval,
                                      uint32_t *__restrict__ fmask, int32_t *bad) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes) return;
  fmask[n] = 0u;
  int32_t d = node_dof[n];
  if (d <= 0) return;
  if (d != scan[n] + 1) {
    atomicOr(bad, 4);
    return;
  }
  int64_t xi[3] = {0, 0, 0}, r = n;
  for (int k = 0; k < L.D; ++k) {
    xi[k] = r % L.np[k];
    r /= L.np[k];
  }
  int nb = 1;
  for (int k = 0; k < L.D; ++k) nb *= 3;
  int32_t p = colptr[d - 1], pend = colptr[d];
  uint32_t m = 0u;
  for (int o = 0; o < nb; ++o) {
    int oo = o;
    int64_t node = 0, stride = 1;
    bool in = true;
    for (int k = 0; k < L.D; ++k) {
      int64_t v = xi[k] + (oo % 3) - 1;
      oo /= 3;
      in = in && v >= 0 && v < L.np[k];
      node += v * stride;
      stride *= L.np[k];
    }
    if (!in) continue;
    int32_t nd = node_dof[node];
    if (nd <= 0) continue;
    if (p >= pend || rowval[p] != nd - 1) {
      atomicOr(bad, 8);
      return;
    }
    ++p;
    m |= 1u << o;
  }
  if (p != pend) atomicOr(bad, 8);
  fmask[n] = m;
}

// ---- lattice symbolic phase (no sort) ----
// With the numbering verified to be node order (free dof = 1 + number of free nodes
// before it), column n of the pattern is the free lattice neighbours of node n in
// ascending node order, which is ascending dof order: two nodes of a Q1 lattice share
// a cell exactly when every coordinate differs by at most one.

// Per node: the free-neighbour mask (bit o, o = Σ_k (δ_k+1) 3^k), its column length
// (0 for a node that owns no column), and the numbering checks.
__global__ void lattice_masks(Lattice L, int64_t nnodes, const int32_t *__restrict__ node_dof,
                              const int32_t *__restrict__ scan, uint32_t *__restrict__ fmask,
                              int32_t *__restrict__ count, int32_t *bad) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes) return;
  fmask[n] = 0u;
  count[n] = 0;
  const int32_t d = node_dof[n];
  if (d == 0) {
    atomicOr(bad, 2);
    return;
  }
  if (d < 0) return;
  if (d != scan[n] + 1) {
    atomicOr(bad, 4);
    return;
  }
  int64_t xi[3] = {0, 0, 0}, r = n;
  for (int k = 0; k < L.D; ++k) {
    xi[k] = r % L.np[k];
    r /= L.np[k];
  }
  int nb = 1;
  for (int k = 0; k < L.D; ++k) nb *= 3;
  uint32_t m = 0u;
  for (int o = 0; o < nb; ++o) {
    int oo = o;
    int64_t node = 0, stride = 1;
    bool in = true;
    for (int k = 0; k < L.D; ++k) {
      const int64_t v = xi[k] + (oo % 3) - 1;
      oo /= 3;
      in = in && v >= 0 && v < L.np[k];
      node += v * stride;
      stride *= L.np[k];
    }
    if (in && node_dof[node] > 0) m |= 1u << o;
  }
  fmask[n] = m;
  count[n] = __popc(m);
}

// colptr and rowval of every free node's column
__global__ void lattice_columns(Lattice L, int64_t nnodes, const int32_t *__restrict__ node_dof,
                                const uint32_t *__restrict__ fmask,
                                const int32_t *__restrict__ start, int32_t *__restrict__ colptr,
                                int32_t *__restrict__ rowval) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes) return;
  const int32_t d = node_dof[n];
  if (d <= 0) return;
  int32_t p = start[n];
  colptr[d - 1] = p;
  const uint32_t m = fmask[n];
  int64_t stride[3] = {1, L.np[0], L.np[0] * L.np[1]};
  int nb = 1;
  for (int k = 0; k < L.D; ++k) nb *= 3;
  for (int o = 0; o < nb; ++o) {
    if (!((m >> o) & 1u)) continue;
    int oo = o;
    int64_t node = n;
    for (int k = 0; k < L.D; ++k) {
      node += ((oo % 3) - 1) * stride[k];
      oo /= 3;
    }
    rowval[p++] = node_dof[node] - 1;
  }
}

// nzval start of each tile's columns: colptr of the first free node's dof, -1
// when the tile has no free node.  The numeric kernel reads it with one scalar
// load instead of deriving it from the loaded dofs.
__global__ void lattice_tile_p0(int64_t nnodes, int64_t ntiles,
                                const int32_t *__restrict__ node_dof,
                                const int32_t *__restrict__ colptr,
                                int32_t *__restrict__ tile_p0, int32_t *__restrict__ tile_d0) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  int32_t p0 = -1, d0 = -1;
  for (int64_t n = t * kTileNodes; n < std::min<int64_t>((t + 1) * kTileNodes, nnodes); ++n) {
    const int32_t d = node_dof[n];
    if (d > 0) {
      p0 = colptr[d - 1];
      d0 = d;
      break;
    }
  }
  tile_p0[t] = p0;
  tile_d0[t] = d0;
}

template <class T>
void exclusive_sum(const T *in, T *out, int64_t n, hipStream_t s) {
  size_t bytes = 0;
  GH_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s));
  DBuf<unsigned char> tmp;
  tmp.resize(std::max<size_t>(bytes, 1));
  GH_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.get(), bytes, in, out, (int)n, s));
  GH_CHECK_HIP(hipStreamSynchronize(s));
}

// The same scan with a kept temporary buffer: enqueued only.
template <class T>
void exclusive_sum_ws(const T *in, T *out, int64_t n, hipStream_t s, DBuf<unsigned char> &ws) {
  size_t bytes = 0;
  GH_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s));
  ws.reserve(std::max<size_t>(bytes, 1));
  GH_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(ws.get(), bytes, in, out, (int)n, s));
}

}  // namespace

void symbolic_general(Ctx &c, bool contributions) {
  hipStream_t s = c.stream;
  const int64_t per = (int64_t)c.nloc * c.nloc;
  const int64_t total = c.ncells * per;
  GH_REQUIRE(total < (int64_t(1) << 31), GH_ERR_UNSUPPORTED,
             "cell pair count exceeds 2^31 in one symbolic pass");
  const int rb = bits_for(std::max<int64_t>(c.nfree, 1));
  GH_REQUIRE(2 * rb <= 64, GH_ERR_UNSUPPORTED, "dof count too large for packed keys");
  const uint64_t invalid = (uint64_t)c.nfree << rb;

  DBuf<uint64_t> keys, sorted;
  DBuf<int32_t> vals, svals;  // contribution ids (emit index), carried by a stable sort
  keys.resize(std::max<int64_t>(total, 1));
  sorted.resize(std::max<int64_t>(total, 1));
  if (total)
    hipLaunchKernelGGL(emit_pairs, dim3(grid_for(total)), dim3(kBlock), 0, s, c.ncells,
                       c.nloc, c.cell_dofs.get(), rb, invalid, keys.get());
  GH_CHECK_HIP(hipGetLastError());
  if (contributions) {
    vals.resize(std::max<int64_t>(total, 1));
    svals.resize(std::max<int64_t>(total, 1));
    if (total)
      hipLaunchKernelGGL(iota_kernel, dim3(grid_for(total)), dim3(kBlock), 0, s, total,
                         vals.get());
    size_t bytes = 0;
    GH_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys.get(), sorted.get(),
                                                    vals.get(), svals.get(), (int)total, 0,
                                                    2 * rb, s));
    DBuf<unsigned char> tmp;
    tmp.resize(std::max<size_t>(bytes, 1));
    GH_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.get(), bytes, keys.get(), sorted.get(),
                                                    vals.get(), svals.get(), (int)total, 0,
                                                    2 * rb, s));
    GH_CHECK_HIP(hipStreamSynchronize(s));
    vals.release();
  } else {
    size_t bytes = 0;
    GH_CHECK_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys.get(), sorted.get(),
                                                   (int)total, 0, 2 * rb, s));
    DBuf<unsigned char> tmp;
    tmp.resize(std::max<size_t>(bytes, 1));
    GH_CHECK_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.get(), bytes, keys.get(), sorted.get(),
                                                   (int)total, 0, 2 * rb, s));
    GH_CHECK_HIP(hipStreamSynchronize(s));
  }
  keys.release();

  DBuf<int32_t> flags, pos;
  flags.resize(std::max<int64_t>(total, 1));
  pos.resize(std::max<int64_t>(total, 1));
  if (total)
    hipLaunchKernelGGL(unique_flags, dim3(grid_for(total)), dim3(kBlock), 0, s, total,
                       sorted.get(), invalid, flags.get());
  GH_CHECK_HIP(hipGetLastError());
  int64_t nnz = 0;
  if (total) {
    exclusive_sum(flags.get(), pos.get(), total, s);
    int32_t last_pos = 0, last_flag = 0;
    GH_CHECK_HIP(hipMemcpy(&last_pos, pos.get() + total - 1, 4, hipMemcpyDeviceToHost));
    GH_CHECK_HIP(hipMemcpy(&last_flag, flags.get() + total - 1, 4, hipMemcpyDeviceToHost));
    nnz = (int64_t)last_pos + last_flag;
  }
  c.nnz = nnz;
  c.colptr.resize(c.nfree + 1);
  {
    GH_CHECK_HIP(hipMemsetAsync(c.colptr.get(), 0xff, sizeof(int32_t) * (c.nfree + 1), s));
    DBuf<int32_t> bad;
    bad.resize(1);
    GH_CHECK_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), s));
    if (total)
      hipLaunchKernelGGL(column_starts, dim3(grid_for(total)), dim3(kBlock), 0, s, total,
                         sorted.get(), invalid, rb, flags.get(), pos.get(), c.colptr.get());
    hipLaunchKernelGGL(column_check, dim3(grid_for(c.nfree + 1)), dim3(kBlock), 0, s, c.nfree,
                       nnz, c.colptr.get(), bad.get());
    GH_CHECK_HIP(hipGetLastError());
    int32_t hbad = 0;
    GH_CHECK_HIP(hipMemcpyAsync(&hbad, bad.get(), 4, hipMemcpyDeviceToHost, s));
    GH_CHECK_HIP(hipStreamSynchronize(s));
    GH_REQUIRE(hbad == 0, GH_ERR_UNSUPPORTED, "a free dof with no matrix entry");
  }
  c.rowval.resize(std::max<int64_t>(nnz, 1));
  if (total)
    hipLaunchKernelGGL(scatter_rows, dim3(grid_for(total)), dim3(kBlock), 0, s, total,
                       sorted.get(), flags.get(), pos.get(), (uint64_t(1) << rb) - 1,
                       c.rowval.get());
  GH_CHECK_HIP(hipGetLastError());
  c.cstart.release();
  c.perm.release();
  if (contributions) {
    c.cstart.resize(nnz + 1);
    GH_CHECK_HIP(hipMemsetAsync(c.cstart.get(), 0, sizeof(int32_t) * (nnz + 1), s));
    if (total)
      hipLaunchKernelGGL(contrib_ranges, dim3(grid_for(total)), dim3(kBlock), 0, s, total,
                         sorted.get(), invalid, flags.get(), pos.get(), nnz, c.cstart.get());
    GH_CHECK_HIP(hipGetLastError());
    // the emit index cell * nloc^2 + (column local) * nloc + (row local) addresses the
    // row-major element matrices directly
    std::swap(c.perm.p, svals.p);
    std::swap(c.perm.n, svals.n);
  }
  GH_CHECK_HIP(hipStreamSynchronize(s));
  c.have_pattern = true;
  c.nzval.resize(std::max<int64_t>(nnz, 1));
  c.b.resize(std::max<int64_t>(c.nfree, 1));
}

// Free dofs forming a full box of lattice nodes (numbered in node order, which the
// structured check has verified): the dof offsets of the 27 neighbours are constant.
// In node order a box's corners hold dofs 1 and N; it is a box when the corners span
// exactly N nodes and no free node lies outside them.
__global__ void box_corners(int64_t nnodes, int64_t nfree, const int32_t *__restrict__ node_dof,
                            int64_t *__restrict__ corner) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes) return;
  const int32_t d = node_dof[n];
  if (d == 1) corner[0] = n;
  if (d == nfree) corner[1] = n;
}

__global__ void box_outside(int64_t nnodes, int64_t np0, int64_t np1, const int64_t *lo,
                            const int64_t *hi, const int32_t *__restrict__ node_dof,
                            int32_t *bad) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes || node_dof[n] <= 0) return;
  const int64_t x[3] = {n % np0, (n / np0) % np1, n / (np0 * np1)};
  for (int d = 0; d < 3; ++d)
    if (x[d] < lo[d] || x[d] > hi[d]) atomicOr(bad, 1);
}

void detect_box(Ctx &c) {
  c.box = false;
  c.a27.release();
  c.g14.release();
  if (!c.structured || c.dim != 3 || c.nfree <= 0) return;
  hipStream_t s = c.stream;
  DBuf<int64_t> dv;
  dv.resize(8);  // corners, then lo[3], hi[3]
  hipLaunchKernelGGL(box_corners, dim3(grid_for(c.nnodes)), dim3(kBlock), 0, s, c.nnodes,
                     c.nfree, c.node_dof.get(), dv.get());
  int64_t corner[2] = {0, 0};
  GH_CHECK_HIP(hipMemcpyAsync(corner, dv.get(), sizeof(corner), hipMemcpyDeviceToHost, s));
  GH_CHECK_HIP(hipStreamSynchronize(s));
  const int64_t np0 = c.part[0] + 1, np1 = c.part[1] + 1;
  int64_t lh[6];
  for (int k = 0; k < 2; ++k) {
    const int64_t n = corner[k];
    lh[3 * k + 0] = n % np0;
    lh[3 * k + 1] = (n / np0) % np1;
    lh[3 * k + 2] = n / (np0 * np1);
  }
  int64_t vol = 1;
  for (int d = 0; d < 3; ++d) {
    if (lh[3 + d] < lh[d]) return;
    vol *= lh[3 + d] - lh[d] + 1;
  }
  if (vol != c.nfree) return;
  DBuf<int32_t> bad;
  bad.resize(1);
  GH_CHECK_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), s));
  GH_CHECK_HIP(hipMemcpyAsync(dv.get() + 2, lh, sizeof(lh), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(box_outside, dim3(grid_for(c.nnodes)), dim3(kBlock), 0, s, c.nnodes, np0,
                     np1, dv.get() + 2, dv.get() + 5, c.node_dof.get(), bad.get());
  GH_CHECK_HIP(hipGetLastError());
  int32_t hbad = 0;
  GH_CHECK_HIP(hipMemcpyAsync(&hbad, bad.get(), 4, hipMemcpyDeviceToHost, s));
  GH_CHECK_HIP(hipStreamSynchronize(s));
  if (hbad) return;
  for (int d = 0; d < 3; ++d) c.box_m[d] = lh[3 + d] - lh[d] + 1;
  c.box = true;
}

// The pattern of a Cartesian scalar-Q1 space without the sort: lattice node ↔ dof map
// from cell_dofs (consistency checked), numbering checked to be node order, then each
// column written directly.  False (nothing set) when a check fails; the caller then
// runs symbolic_general.  On success the lattice kernel's inputs (node_dof,
// node_fmask, tile starts) are set as check_structured sets them.
// 32-bit forms of the lattice kernels (every index of the lattice fits): the 64-bit
// divisions of the general forms dominate their time.
template <class I>
__device__ __forceinline__ I cart_node_i(int D, const int64_t *np, I cell, const int64_t *nc,
                                         int a) {
  I node = 0, stride = 1, r = cell;
  for (int d = 0; d < D; ++d) {
    const I ci = r % (I)nc[d];
    r /= (I)nc[d];
    node += (ci + ((a >> d) & 1)) * stride;
    stride *= (I)np[d];
  }
  return node;
}

// One thread per cell: the cell's first node from one index split, its 2^D vertices
// by adding the lattice strides (no divisions per vertex); the cell's dof ids are one
// contiguous run.
__device__ __forceinline__ int32_t cell_base_node(const Lattice &L, int32_t c, int32_t *st) {
  int32_t node = 0, stride = 1, r = c;
  for (int d = 0; d < L.D; ++d) {
    const int32_t ci = r % (int32_t)L.nc[d];
    r /= (int32_t)L.nc[d];
    node += ci * stride;
    st[d] = stride;
    stride *= (int32_t)L.np[d];
  }
  return node;
}

__global__ void lattice_scatter32(Lattice L, int32_t ncells, int nloc,
                                  const int32_t *__restrict__ cell_dofs,
                                  int32_t *__restrict__ node_dof) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  int32_t st[3] = {0, 0, 0};
  const int32_t base = cell_base_node(L, c, st);
  for (int a = 0; a < nloc; ++a)
    node_dof[base + (a & 1) * st[0] + ((a >> 1) & 1) * st[1] + ((a >> 2) & 1) * st[2]] =
        cell_dofs[c * nloc + a];
}

__global__ void lattice_verify32(Lattice L, int32_t ncells, int nloc,
                                 const int32_t *__restrict__ cell_dofs,
                                 const int32_t *__restrict__ node_dof, int32_t *bad) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  int32_t st[3] = {0, 0, 0};
  const int32_t base = cell_base_node(L, c, st);
  bool ok = true;
  for (int a = 0; a < nloc; ++a)
    ok = ok && node_dof[base + (a & 1) * st[0] + ((a >> 1) & 1) * st[1] +
                        ((a >> 2) & 1) * st[2]] == cell_dofs[c * nloc + a];
  if (!ok) atomicOr(bad, 1);
}

// lattice_masks with 32-bit indices and the stencil unrolled (D known at compile time)
template <int D>
__global__ __launch_bounds__(256) void lattice_masks32(Lattice L, int32_t nnodes,
                                                       const int32_t *__restrict__ node_dof,
                                                       const int32_t *__restrict__ scan,
                                                       uint32_t *__restrict__ fmask,
                                                       int32_t *__restrict__ count, int32_t *bad) {
  const int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes) return;
  const int32_t d = node_dof[n];
  uint32_t m = 0u;
  if (d == 0) atomicOr(bad, 2);
  if (d > 0 && d != scan[n] + 1) atomicOr(bad, 4);
  if (d > 0 && d == scan[n] + 1) {
    int32_t xi[3] = {0, 0, 0}, np[3] = {1, 1, 1}, r = n;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      np[k] = (int32_t)L.np[k];
      xi[k] = r % np[k];
      r /= np[k];
    }
    constexpr int NB = D == 3 ? 27 : 9;
#pragma unroll
    for (int o = 0; o < NB; ++o) {
      int oo = o;
      int32_t node = 0, stride = 1;
      bool in = true;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const int32_t v = xi[k] + (oo % 3) - 1;
        oo /= 3;
        in = in && v >= 0 && v < np[k];
        node += v * stride;
        stride *= np[k];
      }
      if (in && node_dof[node] > 0) m |= 1u << o;
    }
  }
  fmask[n] = m;
  count[n] = __popc(m);
}

// colptr and rowval: one thread per node writes its column's rows into the
// workgroup's LDS run (the columns of 256 consecutive nodes are one contiguous run of
// rowval), which the workgroup then stores coalesced.
template <int D>
__global__ __launch_bounds__(256) void lattice_rows32(Lattice L, int32_t nnodes, int64_t nnz,
                                                      const int32_t *__restrict__ node_dof,
                                                      const uint32_t *__restrict__ fmask,
                                                      const int32_t *__restrict__ start,
                                                      int32_t *__restrict__ colptr,
                                                      int32_t *__restrict__ rowval) {
  constexpr int NB = D == 3 ? 27 : 9;
  __shared__ int32_t buf[256 * NB];
  const int32_t n0 = blockIdx.x * 256, n = n0 + (int32_t)threadIdx.x;
  const int32_t base = start[n0];
  const int64_t endp = n0 + 256 < nnodes ? (int64_t)start[n0 + 256] : nnz;
  if (n < nnodes) {
    const int32_t d = node_dof[n];
    if (d > 0) {
      const int32_t p = start[n];
      colptr[d - 1] = p;
      const uint32_t m = fmask[n];
      int k = p - base;
      int32_t off[3] = {1, (int32_t)L.np[0], (int32_t)(L.np[0] * L.np[1])};
#pragma unroll
      for (int o = 0; o < NB; ++o) {
        if (!((m >> o) & 1u)) continue;
        int oo = o;
        int32_t nb = n;
#pragma unroll
        for (int j = 0; j < D; ++j) {
          nb += ((oo % 3) - 1) * off[j];
          oo /= 3;
        }
        buf[k++] = node_dof[nb] - 1;
      }
    }
  }
  __syncthreads();
  const int len = (int)(endp - base);
  for (int t = threadIdx.x; t < len; t += 256) rowval[(int64_t)base + t] = buf[t];
}

// Clears the flags and the box record: box[0], box[1] = node of dof 1 and of the last
// dof, box[2..4] / box[5..7] = their lattice coordinates, box[8] = 1 when the free nodes
// span exactly that box of the lattice.
__global__ void lattice_init(int32_t *flag, int64_t *box) {
  if (threadIdx.x == 0) flag[0] = 0;
  if (threadIdx.x < 9) box[threadIdx.x] = threadIdx.x < 2 ? -1 : 0;
}

// out = {flags, free count, nnz}; colptr[nfree] = nnz.
__global__ void lattice_summary(int64_t nnodes, int64_t nfree, const int32_t *flag,
                                const int32_t *f, const int32_t *scan, const int32_t *count,
                                const int32_t *start, int32_t *colptr, int64_t *out) {
  if (threadIdx.x != 0) return;
  const int64_t nnz = (int64_t)start[nnodes - 1] + count[nnodes - 1];
  out[0] = flag[0];
  out[1] = (int64_t)scan[nnodes - 1] + f[nnodes - 1];
  out[2] = nnz;
  colptr[nfree] = (int32_t)nnz;
}

__global__ void box_bounds(int64_t np0, int64_t np1, int64_t nfree, int64_t *box) {
  if (threadIdx.x != 0) return;
  if (box[0] < 0 || box[1] < 0) return;
  int64_t vol = 1;
  for (int k = 0; k < 2; ++k) {
    const int64_t n = box[k];
    box[2 + 3 * k] = n % np0;
    box[3 + 3 * k] = (n / np0) % np1;
    box[4 + 3 * k] = n / (np0 * np1);
  }
  for (int d = 0; d < 3; ++d) {
    if (box[5 + d] < box[2 + d]) return;
    vol *= box[5 + d] - box[2 + d] + 1;
  }
  box[8] = vol == nfree;
}

__global__ void box_outside32(int32_t nnodes, int32_t np0, int32_t np1, int64_t *box,
                              const int32_t *__restrict__ node_dof) {
  const int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes || box[8] == 0 || node_dof[n] <= 0) return;
  const int32_t x[3] = {n % np0, (n / np0) % np1, n / (np0 * np1)};
  for (int d = 0; d < 3; ++d)
    if (x[d] < box[2 + d] || x[d] > box[5 + d]) box[8] = 0;
}

// Lattice pattern with 32-bit indices: a fixed number of launches, the workspace
// kept in the context, and two reads back (the counts, then the box).
bool symbolic_lattice32(Ctx &c) {
  hipStream_t s = c.stream;
  Lattice L;
  L.D = c.dim;
  for (int d = 0; d < 3; ++d) {
    L.nc[d] = d < c.dim ? c.part[d] : 1;
    L.np[d] = d < c.dim ? c.part[d] + 1 : 1;
  }
  const int32_t nn = (int32_t)c.nnodes;
  c.node_dof.resize(c.nnodes);
  c.node_fmask.resize(c.nnodes);
  c.ws_f.resize(c.nnodes);
  c.ws_scan.resize(c.nnodes);
  c.ws_count.resize(c.nnodes);
  c.ws_start.resize(c.nnodes);
  c.ws_flag.resize(1);
  c.ws_box.resize(9);
  c.colptr.resize(c.nfree + 1);
  int64_t *h = c.h_small.get(16);
  hipLaunchKernelGGL(lattice_init, dim3(1), dim3(64), 0, s, c.ws_flag.get(), c.ws_box.get());
  const int32_t nc32 = (int32_t)c.ncells;
  hipLaunchKernelGGL(lattice_scatter32, dim3(grid_for(nc32)), dim3(kBlock), 0, s, L, nc32,
                     c.nloc, c.cell_dofs.get(), c.node_dof.get());
  hipLaunchKernelGGL(lattice_verify32, dim3(grid_for(nc32)), dim3(kBlock), 0, s, L, nc32,
                     c.nloc, c.cell_dofs.get(), c.node_dof.get(), c.ws_flag.get());
  hipLaunchKernelGGL(free_flags, dim3(grid_for(nn)), dim3(kBlock), 0, s, c.nnodes,
                     c.node_dof.get(), c.ws_f.get(), c.ws_flag.get());
  GH_CHECK_HIP(hipGetLastError());
  exclusive_sum_ws(c.ws_f.get(), c.ws_scan.get(), c.nnodes, s, c.ws_cub);
  if (c.dim == 3)
    hipLaunchKernelGGL(lattice_masks32<3>, dim3(grid_for(nn)), dim3(kBlock), 0, s, L, nn,
                       c.node_dof.get(), c.ws_scan.get(), c.node_fmask.get(), c.ws_count.get(),
                       c.ws_flag.get());
  else
    hipLaunchKernelGGL(lattice_masks32<2>, dim3(grid_for(nn)), dim3(kBlock), 0, s, L, nn,
                       c.node_dof.get(), c.ws_scan.get(), c.node_fmask.get(), c.ws_count.get(),
                       c.ws_flag.get());
  GH_CHECK_HIP(hipGetLastError());
  exclusive_sum_ws(c.ws_count.get(), c.ws_start.get(), c.nnodes, s, c.ws_cub);
  hipLaunchKernelGGL(lattice_summary, dim3(1), dim3(64), 0, s, c.nnodes, c.nfree,
                     c.ws_flag.get(), c.ws_f.get(), c.ws_scan.get(), c.ws_count.get(),
                     c.ws_start.get(), c.colptr.get(), c.ws_box.get());
  GH_CHECK_HIP(hipGetLastError());
  GH_CHECK_HIP(hipMemcpyAsync(h, c.ws_box.get(), 3 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  GH_CHECK_HIP(hipStreamSynchronize(s));
  if (h[0] != 0 || h[1] != c.nfree) return false;
  const int64_t nnz = h[2];
  c.nnz = nnz;
  c.rowval.resize(std::max<int64_t>(nnz, 1));
  if (c.dim == 3)
    hipLaunchKernelGGL(lattice_rows32<3>, dim3(grid_for(nn)), dim3(256), 0, s, L, nn, nnz,
                       c.node_dof.get(), c.node_fmask.get(), c.ws_start.get(), c.colptr.get(),
                       c.rowval.get());
  else
    hipLaunchKernelGGL(lattice_rows32<2>, dim3(grid_for(nn)), dim3(256), 0, s, L, nn, nnz,
                       c.node_dof.get(), c.node_fmask.get(), c.ws_start.get(), c.colptr.get(),
                       c.rowval.get());
  const int64_t ntiles = (c.nnodes + kTileNodes - 1) / kTileNodes;
  c.tile_p0.resize(ntiles);
  c.tile_d0.resize(ntiles);
  hipLaunchKernelGGL(lattice_tile_p0, dim3(grid_for(ntiles)), dim3(kBlock), 0, s, c.nnodes,
                     ntiles, c.node_dof.get(), c.colptr.get(), c.tile_p0.get(), c.tile_d0.get());
  GH_CHECK_HIP(hipGetLastError());
  // box of free nodes (the stencil PCG), decided on the device
  c.box = false;
  c.a27.release();
  c.g14.release();
  if (c.dim == 3) {
    hipLaunchKernelGGL(lattice_init, dim3(1), dim3(64), 0, s, c.ws_flag.get(), c.ws_box.get());
    hipLaunchKernelGGL(box_corners, dim3(grid_for(nn)), dim3(kBlock), 0, s, c.nnodes, c.nfree,
                       c.node_dof.get(), c.ws_box.get());
    hipLaunchKernelGGL(box_bounds, dim3(1), dim3(64), 0, s, L.np[0], L.np[1], c.nfree,
                       c.ws_box.get());
    hipLaunchKernelGGL(box_outside32, dim3(grid_for(nn)), dim3(kBlock), 0, s, nn,
                       (int32_t)L.np[0], (int32_t)L.np[1], c.ws_box.get(), c.node_dof.get());
    GH_CHECK_HIP(hipGetLastError());
    GH_CHECK_HIP(hipMemcpyAsync(h + 3, c.ws_box.get(), 9 * sizeof(int64_t),
                                hipMemcpyDeviceToHost, s));
  }
  GH_CHECK_HIP(hipStreamSynchronize(s));
  if (c.dim == 3 && h[3 + 8] == 1) {
    for (int d = 0; d < 3; ++d) c.box_m[d] = h[3 + 5 + d] - h[3 + 2 + d] + 1;
    c.box = true;
  }
  c.cstart.release();
  c.perm.release();
  c.have_pattern = true;
  c.structured = true;
  c.nzval.resize(std::max<int64_t>(nnz, 1));
  c.b.resize(std::max<int64_t>(c.nfree, 1));
  return true;
}

bool symbolic_lattice(Ctx &c) {
  c.structured = false;
  if (!(c.lattice_topo && c.reffe_kind == GH_REFFE_QK && c.order == 1 && c.ncomp == 1 &&
        c.nloc == (1 << c.dim) && c.nfree > 0 && c.nnodes < (int64_t(1) << 31)))
    return false;
  if (c.ncells * c.nloc < (int64_t(1) << 31) && c.nnodes * 27 < (int64_t(1) << 31))
    return symbolic_lattice32(c);
  hipStream_t s = c.stream;
  Lattice L;
  L.D = c.dim;
  for (int d = 0; d < 3; ++d) {
    L.nc[d] = d < c.dim ? c.part[d] : 1;
    L.np[d] = d < c.dim ? c.part[d] + 1 : 1;
  }
  DBuf<int32_t> node_dof, bad, f, scan, count, start;
  DBuf<uint32_t> fmask;
  node_dof.resize(c.nnodes);
  fmask.resize(c.nnodes);
  bad.resize(1);
  f.resize(c.nnodes);
  scan.resize(c.nnodes);
  count.resize(c.nnodes);
  start.resize(c.nnodes);
  GH_CHECK_HIP(hipMemsetAsync(node_dof.get(), 0, sizeof(int32_t) * c.nnodes, s));
  GH_CHECK_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), s));
  const int64_t t = c.ncells * c.nloc;
  hipLaunchKernelGGL(lattice_scatter, dim3(grid_for(t)), dim3(kBlock), 0, s, L, c.ncells, c.nloc,
                     c.cell_dofs.get(), node_dof.get());
  hipLaunchKernelGGL(lattice_verify, dim3(grid_for(t)), dim3(kBlock), 0, s, L, c.ncells, c.nloc,
                     c.cell_dofs.get(), node_dof.get(), bad.get());
  hipLaunchKernelGGL(free_flags, dim3(grid_for(c.nnodes)), dim3(kBlock), 0, s, c.nnodes,
                     node_dof.get(), f.get(), bad.get());
  GH_CHECK_HIP(hipGetLastError());
  exclusive_sum(f.get(), scan.get(), c.nnodes, s);
  hipLaunchKernelGGL(lattice_masks, dim3(grid_for(c.nnodes)), dim3(kBlock), 0, s, L, c.nnodes,
                     node_dof.get(), scan.get(), fmask.get(), count.get(), bad.get());
  GH_CHECK_HIP(hipGetLastError());
  exclusive_sum(count.get(), start.get(), c.nnodes, s);
  int32_t hbad = 0, last_scan = 0, last_f = 0, last_start = 0, last_count = 0;
  GH_CHECK_HIP(hipMemcpy(&hbad, bad.get(), 4, hipMemcpyDeviceToHost));
  GH_CHECK_HIP(hipMemcpy(&last_scan, scan.get() + c.nnodes - 1, 4, hipMemcpyDeviceToHost));
  GH_CHECK_HIP(hipMemcpy(&last_f, f.get() + c.nnodes - 1, 4, hipMemcpyDeviceToHost));
  GH_CHECK_HIP(hipMemcpy(&last_start, start.get() + c.nnodes - 1, 4, hipMemcpyDeviceToHost));
  GH_CHECK_HIP(hipMemcpy(&last_count, count.get() + c.nnodes - 1, 4, hipMemcpyDeviceToHost));
  if (hbad != 0 || (int64_t)last_scan + last_f != c.nfree) return false;
  const int64_t nnz = (int64_t)last_start + last_count;
  c.nnz = nnz;
  c.colptr.resize(c.nfree + 1);
  c.rowval.resize(std::max<int64_t>(nnz, 1));
  hipLaunchKernelGGL(lattice_columns, dim3(grid_for(c.nnodes)), dim3(kBlock), 0, s, L, c.nnodes,
                     node_dof.get(), fmask.get(), start.get(), c.colptr.get(), c.rowval.get());
  const int32_t nnz32 = (int32_t)nnz;
  GH_CHECK_HIP(hipMemcpyAsync(c.colptr.get() + c.nfree, &nnz32, 4, hipMemcpyHostToDevice, s));
  GH_CHECK_HIP(hipGetLastError());
  // the lattice kernel's inputs
  std::swap(c.node_dof.p, node_dof.p);
  std::swap(c.node_dof.n, node_dof.n);
  std::swap(c.node_fmask.p, fmask.p);
  std::swap(c.node_fmask.n, fmask.n);
  const int64_t ntiles = (c.nnodes + kTileNodes - 1) / kTileNodes;
  c.tile_p0.resize(ntiles);
  c.tile_d0.resize(ntiles);
  hipLaunchKernelGGL(lattice_tile_p0, dim3(grid_for(ntiles)), dim3(kBlock), 0, s, c.nnodes,
                     ntiles, c.node_dof.get(), c.colptr.get(), c.tile_p0.get(), c.tile_d0.get());
  GH_CHECK_HIP(hipGetLastError());
  GH_CHECK_HIP(hipStreamSynchronize(s));
  c.cstart.release();
  c.perm.release();
  c.have_pattern = true;
  c.structured = true;
  c.nzval.resize(std::max<int64_t>(nnz, 1));
  c.b.resize(std::max<int64_t>(c.nfree, 1));
  detect_box(c);
  return true;
}

void check_structured(Ctx &c) {
  c.structured = false;
  c.node_dof.release();
  c.node_fmask.release();
  c.tile_p0.release();
  c.tile_d0.release();
  if (!(c.lattice_topo && c.reffe_kind == GH_REFFE_QK && c.order == 1 && c.ncomp == 1 &&
        c.nloc == (1 << c.dim) && c.nfree > 0))
    return;
  hipStream_t s = c.stream;
  Lattice L;
  L.D = c.dim;
  for (int d = 0; d < 3; ++d) {
    L.nc[d] = c.part[d];
    L.np[d] = c.part[d] + 1;
  }
  c.node_dof.resize(c.nnodes);
  c.node_fmask.resize(c.nnodes);
  DBuf<int32_t> bad, f, scan;
  bad.resize(1);
  f.resize(c.nnodes);
  scan.resize(c.nnodes);
  GH_CHECK_HIP(hipMemsetAsync(c.node_dof.get(), 0, sizeof(int32_t) * c.nnodes, s));
  GH_CHECK_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), s));
  int64_t t = c.ncells * c.nloc;
  hipLaunchKernelGGL(lattice_scatter, dim3(grid_for(t)), dim3(kBlock), 0, s, L, c.ncells,
                     c.nloc, c.cell_dofs.get(), c.node_dof.get());
  hipLaunchKernelGGL(lattice_verify, dim3(grid_for(t)), dim3(kBlock), 0, s, L, c.ncells,
                     c.nloc, c.cell_dofs.get(), c.node_dof.get(), bad.get());
  hipLaunchKernelGGL(free_flags, dim3(grid_for(c.nnodes)), dim3(kBlock), 0, s, c.nnodes,
                     c.node_dof.get(), f.get(), bad.get());
  GH_CHECK_HIP(hipGetLastError());
  exclusive_sum(f.get(), scan.get(), c.nnodes, s);
  hipLaunchKernelGGL(lattice_check_pattern, dim3(grid_for(c.nnodes)), dim3(kBlock), 0, s, L,
                     c.nnodes, c.node_dof.get(), scan.get(), c.colptr.get(), c.rowval.get(),
                     c.node_fmask.get(), bad.get());
  const int64_t ntiles = (c.nnodes + kTileNodes - 1) / kTileNodes;
  c.tile_p0.resize(ntiles);
  c.tile_d0.resize(ntiles);
  hipLaunchKernelGGL(lattice_tile_p0, dim3(grid_for(ntiles)), dim3(kBlock), 0, s, c.nnodes,
                     ntiles, c.node_dof.get(), c.colptr.get(), c.tile_p0.get(),
                     c.tile_d0.get());
  GH_CHECK_HIP(hipGetLastError());
  int32_t hbad = 0;
  GH_CHECK_HIP(hipMemcpyAsync(&hbad, bad.get(), 4, hipMemcpyDeviceToHost, s));
  GH_CHECK_HIP(hipStreamSynchronize(s));
  c.structured = (hbad == 0);
  if (!c.structured) {
    c.node_dof.release();
    c.node_fmask.release();
    c.tile_p0.release();
    c.tile_d0.release();
  }
  detect_box(c);
}

// Greedy colouring of cells so that no two cells of one colour share a free dof:
// each colour's scatter is race-free and the colour sequence fixes the summation
// order (deterministic).  Cells keep ascending order inside a colour.
void build_coloring(Ctx &c) {
  if (c.ncolors > 0) return;
  const int64_t C = c.ncells;
  const int nloc = c.nloc;
  const std::vector<int32_t> &cd = c.h_cell_dofs;
  std::vector<int64_t> dptr(c.nfree + 1, 0);
  for (int64_t k = 0; k < C * nloc; ++k)
    if (cd[k] > 0) dptr[cd[k]]++;
  for (int64_t j = 0; j < c.nfree; ++j) dptr[j + 1] += dptr[j];
  std::vector<int32_t> dcells(dptr[c.nfree]);
  {
    std::vector<int64_t> cur(dptr.begin(), dptr.end() - 1);
    for (int64_t cell = 0; cell < C; ++cell)
      for (int a = 0; a < nloc; ++a) {
        int32_t d = cd[cell * nloc + a];
        if (d > 0) dcells[cur[d - 1]++] = (int32_t)cell;
      }
  }
  std::vector<int32_t> color(C, -1);
  std::vector<int64_t> stamp(64, -1);
  int ncol = 0;
  for (int64_t cell = 0; cell < C; ++cell) {
    for (int a = 0; a < nloc; ++a) {
      int32_t d = cd[cell * nloc + a];
      if (d <= 0) continue;
      for (int64_t k = dptr[d - 1]; k < dptr[d]; ++k) {
        int32_t other = dcells[k];
        int32_t col = color[other];
        if (col >= 0) {
          if ((size_t)col >= stamp.size()) stamp.resize(col * 2 + 1, -1);
          stamp[col] = cell;
        }
      }
    }
    int chosen = 0;
    while ((size_t)chosen < stamp.size() && stamp[chosen] == cell) ++chosen;
    if ((size_t)chosen >= stamp.size()) stamp.resize(chosen * 2 + 1, -1);
    color[cell] = chosen;
    ncol = std::max(ncol, chosen + 1);
  }
  c.color_ptr.assign(ncol + 1, 0);
  for (int64_t cell = 0; cell < C; ++cell) c.color_ptr[color[cell] + 1]++;
  for (int k = 0; k < ncol; ++k) c.color_ptr[k + 1] += c.color_ptr[k];
  std::vector<int32_t> cells(C);
  {
    std::vector<int64_t> cur(c.color_ptr.begin(), c.color_ptr.end() - 1);
    for (int64_t cell = 0; cell < C; ++cell) cells[cur[color[cell]]++] = (int32_t)cell;
  }
  c.color_cells.resize(std::max<int64_t>(C, 1));
  GH_CHECK_HIP(hipMemcpy(c.color_cells.get(), cells.data(), sizeof(int32_t) * C,
                         hipMemcpyHostToDevice));
  c.ncolors = ncol;
}

}  // namespace gh
