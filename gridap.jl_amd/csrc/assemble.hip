// Numeric phase: element integration + scatter into the CSC nzval and the RHS b,
// for a(u,v)=∫κ∇v⋅∇u dΩ and l(v)=∫f v dΩ with Dirichlet lifting.
//
// Two device paths produce the same matrix:
//  * lattice ("owner computes"): Cartesian mesh + scalar Q1 with CLagrangian
//    numbering (checked at symbolic time).  One thread owns one CSC column (one
//    free node): it gathers the element contributions of its 2^D incident cells
//    into its 3^D stencil entries in registers, in ascending cell order (the
//    order of Gridap's cell loop), stages the column in LDS and the workgroup
//    writes its contiguous nzval range once, coalesced.  No atomics, no
//    read-modify-write of nzval, deterministic.
//  * general: one thread per cell computes the element matrix with the
//    isoparametric (or Cartesian) Jacobian at every quadrature point and adds it
//    into nzval through a binary search of the column (nz_index,
//    src/Algebra/SparseMatrixCSC.jl:14-22).  Cells are launched colour by colour
//    (no two cells of a colour share a free dof), so the scatter is race-free and
//    deterministic.
//
// Element math (src/CellData/CellQuadratures.jl:138-160,
// src/Fields/FieldsInterfaces.jl:737-760): K_ij = Σ_q κ (J⁻ᵀ∇̂φ_i)⋅(J⁻ᵀ∇̂φ_j) |det J| w_q,
// b_i = Σ_q f φ_i |det J| w_q, then b -= K u_d on cells with Dirichlet dofs
// (src/CellData/AttachDirichlet.jl:59-84).
#include <algorithm>
#include <climits>
#include <cmath>
#include <type_traits>
#include <vector>

#include "gh_internal.h"

namespace gh {
namespace {

// ---------------------------------------------------------------------------
// Lattice (owner-computes) kernel
// ---------------------------------------------------------------------------

struct LatticeArgs {
  int nn[3];       // lattice nodes per axis
  int nc[3];       // cells per axis
  int nnodes;
  double inv_nn0, inv_nn1;  // reciprocals for the node-index split
  const int32_t *node_dof;
  const uint32_t *fmask;    // free stencil neighbours per node (symbolic phase)
  const int32_t *tile_p0;   // nzval start of each 64-node tile (symbolic phase)
  const int32_t *tile_d0;   // dof of each tile's first free node (symbolic phase)
  const int32_t *colptr;
  const double *kappa;      // per cell or nullptr (κ = 1)
  const double *fsrc;       // per cell or nullptr (f = 1)
  const double *dvals;      // Dirichlet values or nullptr (homogeneous)
  double *nzval;
  double *b;
  double Kc[8];  // reference stiffness (κ=1) by class a⊕b of the node pair
  double b0;     // reference load ∫φ_a (f=1), equal for every a on a box cell
  // mapped lattice: per cell, the upper triangle of the element matrix (row-major)
  // then the element load, entry-major ([k][cell]), FP32 or FP64
  const void *elem;
  // element loads in FP64 ([a][cell]) in place of elem's load slots, or nullptr
  const double *eload;
  int ncells;
};

template <int D>
struct Pow3 {
  static constexpr int v = 3 * Pow3<D - 1>::v;
};
template <>
struct Pow3<0> {
  static constexpr int v = 1;
};

// Staging rounds of the lattice kernel's tile: LDS per wave is 64/R columns, so
// more waves fit a CU (measured on MI355X, 128^3 Q1).
constexpr int kLatticeStageRounds = 2;
// Tiles (waves) per workgroup of the lattice kernel; each wave stages through its
// own LDS and synchronises only with itself.
constexpr int kLatticeWaves = 4;



// floor(n / d) for 0 <= n < 2^31 from a double reciprocal, corrected exactly.
__device__ __forceinline__ int div_by(int n, int d, double inv) {
  int q = (int)((double)n * inv);
  const int r = n - q * d;
  q += (r >= d) - (r < 0);
  return q;
}

// Wave-wide exclusive prefix sum of v (64 lanes).
__device__ __forceinline__ int wave_exclusive_sum(int v, int lane) {
  int incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  return incl - v;
}

// Inputs of one 64-node tile for one lane, loaded ahead of their use.
template <int D>
struct TileIn {
  int xi[3];
  uint32_t fm;
  double kc[1 << D];
  int cid[1 << D];  // mapped lattice: the incident cells (-1 outside)
  double fsum;
};

template <int D, bool MAPPED>
__device__ __forceinline__ void load_tile(const LatticeArgs &P, int tile, int lane,
                                          TileIn<D> &T) {
  constexpr int NV = 1 << D;
  const int n = tile * kTileNodes + lane;
  T.xi[0] = T.xi[1] = T.xi[2] = 0;
  if (n < P.nnodes) {
    const int t = div_by(n, P.nn[0], P.inv_nn0);
    T.xi[0] = n - t * P.nn[0];
    if (D == 3) {
      T.xi[2] = div_by(t, P.nn[1], P.inv_nn1);
      T.xi[1] = t - T.xi[2] * P.nn[1];
    } else {
      T.xi[1] = t;
    }
    T.fm = P.fmask[n];  // 0 exactly for nodes that own no column (Dirichlet)
  } else {
    T.fm = 0u;
  }
  // κ and f of the incident cells, in ascending cell id (bit k of e: cell
  // x_k - 1 + e_k); cells outside the lattice contribute nothing.
  T.fsum = 0.0;
  const bool valid = n < P.nnodes;
#pragma unroll
  for (int e = 0; e < NV; ++e) {
    bool in = valid;
    int cid = 0, cstride = 1;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const int ck = T.xi[k] - 1 + ((e >> k) & 1);
      in = in && ck >= 0 && ck < P.nc[k];
      cid += ck * cstride;
      cstride *= P.nc[k];
    }
    if (MAPPED) {
      T.cid[e] = in ? cid : -1;
    } else {
      T.kc[e] = in ? (P.kappa ? P.kappa[cid] : 1.0) : 0.0;
      T.fsum += in ? (P.fsrc ? P.fsrc[cid] : 1.0) : 0.0;
    }
  }
}

// One wave owns one tile of 64 consecutive lattice nodes (kLatticeWaves tiles per
// workgroup); the tile's nodes own a contiguous range of CSC columns, hence of nzval.  On a box
// cell with a tensor-product Q1 basis and a symmetric rule, K̂[a][b] depends only
// on a⊕b, so the entry of column n at stencil offset δ is K̂[class(δ)] times the
// sum of κ over the cells that hold both nodes.  Those sums and products are FP64
// (a few dozen flops per column; the kernel is bound by the nzval stores).
// Column offsets come from the free-neighbour masks (one load per node instead
// of 3^D); the tile is staged in LDS and stored in 16-byte chunks.
//
// EK = 0: Cartesian cells (K̂ and κ); EK = 4 / 8: mapped cells, whose element matrices
// and loads the element kernel below stored in FP32 / FP64: the entry of column n at
// offset δ is Σ over the incident cells e, ascending, of K_e[a_e(n)][a_e(n+δ)].
template <int D, int EK>
__global__ __launch_bounds__(64 * kLatticeWaves) void lattice_q1_kernel(const LatticeArgs P,
                                                                        int ntiles) {
  using ET = typename std::conditional<EK == 4, float, double>::type;
  constexpr int NB = Pow3<D>::v;  // stencil entries per column
  constexpr int NV = 1 << D;      // incident cells per node
  constexpr int R = kLatticeStageRounds;
  constexpr int LPR = 64 / R;     // lanes staged per round
  __shared__ __align__(16) double stages[kLatticeWaves][LPR * NB + 2];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double *stage = stages[wave];
  // XCD-aware order at workgroup granularity (workgroups are dealt round-robin)
  const int tile = xcd_remap(blockIdx.x, gridDim.x) * kLatticeWaves + wave;
  if (tile >= ntiles) return;
  const int P0 = P.tile_p0[tile];  // uniform: every load below is independent of it
  const int dbase = P.tile_d0[tile];
  TileIn<D> cur;
  load_tile<D, EK != 0>(P, tile, lane, cur);
  if (P0 < 0) return;  // wave-uniform: no free node in the tile
  const uint32_t fm = cur.fm;
  // Free dofs are numbered in node order (checked in the symbolic phase): the
  // lane's dof is the tile's first free dof plus the free lanes below it.
  const unsigned long long freem = __ballot(fm != 0u);
  const int d0 = fm != 0u ? dbase + __popcll(freem & ((1ull << lane) - 1ull)) : 0;

  // acc[o] = K̂[class(o)] Σ_{cells holding node and neighbour o} κ_c
  double acc[NB];
  double bsum = 0.0;
  if (EK != 0) {
    constexpr int NT = NV * (NV + 1) / 2;
    const ET *E = static_cast<const ET *>(P.elem);
    const int64_t nce = P.ncells;
#pragma unroll
    for (int o = 0; o < NB; ++o) acc[o] = 0.0;
#pragma unroll
    for (int e = 0; e < NV; ++e) {
      const int cid = cur.cid[e];
      if (cid < 0) continue;
      const int a = (~e) & (NV - 1);  // the node's local index in cell e
      bsum += P.eload ? P.eload[(int64_t)a * nce + cid] : (double)E[(int64_t)(NT + a) * nce + cid];
#pragma unroll
      for (int o = 0; o < NB; ++o) {
        int oo = o, bl = 0;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < D; ++k) {
          const int bk = ((a >> k) & 1) + oo % 3 - 1;
          oo /= 3;
          ok = ok && bk >= 0 && bk <= 1;
          bl |= (bk & 1) << k;
        }
        if (!ok) continue;
        const int lo = a < bl ? a : bl, hi = a < bl ? bl : a;
        const int t = lo * NV - lo * (lo - 1) / 2 + (hi - lo);
        acc[o] += (double)E[(int64_t)t * nce + cid];
      }
    }
  } else
#pragma unroll
  for (int o = 0; o < NB; ++o) {
    double sum = 0.0;
    int cls = 0;
#pragma unroll
    for (int e = 0; e < NV; ++e) {
      bool ok = true;
      int oo = o;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const int dk = oo % 3 - 1;  // neighbour displacement on axis k
        oo /= 3;
        const int ek = (e >> k) & 1;
        ok = ok && !(dk == -1 && ek == 1) && !(dk == 1 && ek == 0);
      }
      if (ok) sum += cur.kc[e];
    }
    {
      int oo = o;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        cls |= (oo % 3 != 1) << k;
        oo /= 3;
      }
    }
    acc[o] = P.Kc[cls] * sum;
  }

  double bacc = EK != 0 ? bsum : P.b0 * cur.fsum;
  if (P.dvals && d0 > 0) {
    // Dirichlet lifting b -= Σ_δ A(n, n+δ) u_d(n+δ) over Dirichlet neighbours.
#pragma unroll
    for (int o = 0; o < NB; ++o) {
      if ((fm >> o) & 1u) continue;
      int oo = o, node = 0, stride = 1;
      bool in = true;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const int v = cur.xi[k] + oo % 3 - 1;
        oo /= 3;
        in = in && v >= 0 && v < P.nn[k];
        node += v * stride;
        stride *= P.nn[k];
      }
      if (!in) continue;
      const int nd = P.node_dof[node];
      if (nd < 0) bacc -= acc[o] * P.dvals[-nd - 1];
    }
  }

  // Column offsets inside the tile's nzval range from the masks.
  const int cnt = __popc(fm);
  const int off = wave_exclusive_sum(cnt, lane);
  const int total = __shfl(off + cnt, 63);
  if (d0 > 0) P.b[d0 - 1] = bacc;

  // R rounds of LPR lanes each: stage that lane group's columns, store them.
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int beg = r == 0 ? 0 : __shfl(off, r * LPR);
    const int end = r == R - 1 ? total : __shfl(off, (r + 1) * LPR);
    if (end == beg) continue;  // wave-uniform
    const int Pr = P0 + beg;
    const int sh = Pr & 1;  // stage[sh + k] <-> nzval[Pr + k]; 16-byte aligned chunks
    if (r > 0) wave_sync();  // the previous round's stage has been read out
    if (d0 > 0 && lane / LPR == r) {
      int w = sh + off - beg;
#pragma unroll
      for (int o = 0; o < NB; ++o)
        if ((fm >> o) & 1u) stage[w++] = acc[o];
    }
    wave_sync();
    const int lo = sh, hi = sh + (end - beg);
    double2 *dst2 = reinterpret_cast<double2 *>(P.nzval + (Pr - sh));
    const double2 *st2 = reinterpret_cast<const double2 *>(stage);
    for (int c = (lo + 1) / 2 + lane; c < hi / 2; c += 64) dst2[c] = st2[c];
    if (lane == 0 && (lo & 1)) P.nzval[Pr] = stage[lo];
    if (lane == 1 && (hi & 1)) P.nzval[Pr + (end - beg) - 1] = stage[hi - 1];
  }
}

// Reference-cell stiffness and load on the Cartesian cell (Jacobian diag(h)).
void lattice_reference(const Ctx &c, double *K, double *bv) {
  const int D = c.dim, nv = 1 << D;
  double det = 1.0;
  for (int d = 0; d < D; ++d) det *= c.sizes[d];
  for (int i = 0; i < nv * nv; ++i) K[i] = 0.0;
  for (int i = 0; i < nv; ++i) bv[i] = 0.0;
  for (int q = 0; q < c.nq; ++q) {
    double dV = std::fabs(det) * c.qw[q];
    for (int j = 0; j < nv; ++j)
      for (int i = 0; i < nv; ++i) {
        double dot = 0.0;
        for (int d = 0; d < D; ++d)
          dot += (c.grad[(q * nv + i) * D + d] / c.sizes[d]) *
                 (c.grad[(q * nv + j) * D + d] / c.sizes[d]);
        K[i * nv + j] += dot * dV;
      }
    for (int i = 0; i < nv; ++i) bv[i] += c.shp[q * nv + i] * dV;
  }
}

// The lattice kernel needs K̂[a][b] = K̂[a⊕b][0] and equal load entries (true for
// a tensor-product basis with a symmetric rule on a box); checked to 1e-13.
bool lattice_reference_symmetric(const Ctx &c, double *Kc, double *b0) {
  const int nv = 1 << c.dim;
  double K[64], bv[8];
  lattice_reference(c, K, bv);
  double kmax = 0.0, bmax = 0.0;
  for (int i = 0; i < nv * nv; ++i) kmax = std::max(kmax, std::fabs(K[i]));
  for (int i = 0; i < nv; ++i) bmax = std::max(bmax, std::fabs(bv[i]));
  for (int a = 0; a < nv; ++a)
    for (int b = 0; b < nv; ++b)
      if (std::fabs(K[a * nv + b] - K[(a ^ b) * nv]) > 1e-13 * kmax) return false;
  for (int a = 0; a < nv; ++a)
    if (std::fabs(bv[a] - bv[0]) > 1e-13 * bmax) return false;
  for (int k = 0; k < 8; ++k) Kc[k] = k < nv ? K[k * nv] : 0.0;
  *b0 = bv[0];
  return true;
}

template <int D>
void launch_lattice(Ctx &c, const double *Kc, double b0) {
  c.last_kernel = GH_KERNEL_Q1_LATTICE;
  LatticeArgs P{};
  for (int k = 0; k < 3; ++k) {
    P.nn[k] = (int)(c.part[k] + 1);
    P.nc[k] = (int)c.part[k];
  }
  P.nnodes = (int)c.nnodes;
  P.inv_nn0 = 1.0 / P.nn[0];
  P.inv_nn1 = 1.0 / P.nn[1];
  P.node_dof = c.node_dof.get();
  P.fmask = c.node_fmask.get();
  P.tile_p0 = c.tile_p0.get();
  P.tile_d0 = c.tile_d0.get();
  P.colptr = c.colptr.get();
  P.kappa = c.have_kappa ? c.kappa.get() : nullptr;
  P.fsrc = c.have_fsrc ? c.fsrc.get() : nullptr;
  P.dvals = c.have_dvals ? c.dvals.get() : nullptr;
  P.nzval = c.nzval.get();
  P.b = c.b.get();
  for (int k = 0; k < 8; ++k) P.Kc[k] = Kc[k];
  P.b0 = b0;
  const int ntiles = (int)((c.nnodes + kTileNodes - 1) / kTileNodes);
  const unsigned grid = (unsigned)((ntiles + kLatticeWaves - 1) / kLatticeWaves);
  hipLaunchKernelGGL((lattice_q1_kernel<D, 0>), dim3(grid), dim3(kTileNodes * kLatticeWaves), 0,
                     c.stream, P, ntiles);
  GH_CHECK_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Mapped lattice: element kernel (per cell, isoparametric Jacobian per point) then
// the owner-computes column kernel over the stored element matrices
// ---------------------------------------------------------------------------

struct QElemArgs {
  int nc[3], nn[3];
  int ncells;
  int nq;
  const double *xyz;  // node coordinates, or nullptr: the Cartesian cells of sizes h
  double h[3];
  const double *tab;  // per q: w, shp[NV], grad[NV*D]
  const double *kappa, *fsrc;
  const double *sigma;  // mass coefficient per cell, or nullptr
  const double *fq;     // source per cell and point (replaces fsrc), or nullptr
};

// The element matrix and load of every cell, with the arithmetic of cell_q1_kernel
// (FP64 Jacobian, FP32 or FP64 element math); stored as the upper triangle of the
// symmetric matrix (row-major) and the load, entry-major.
template <int D, bool F32>
__global__ __launch_bounds__(256) void q1_lattice_elem_kernel(const QElemArgs A,
                                                              typename std::conditional<F32, float, double>::type *E) {
  using T = typename std::conditional<F32, float, double>::type;
  constexpr int NL = 1 << D;
  constexpr int NT = NL * (NL + 1) / 2;
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= A.ncells) return;
  int ci[3] = {0, 0, 0};
  {
    int r = cell;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      ci[k] = r % A.nc[k];
      r /= A.nc[k];
    }
  }
  double X[NL][D];
#pragma unroll
  for (int a = 0; a < NL; ++a) {
    int node = 0, stride = 1;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      node += (ci[k] + ((a >> k) & 1)) * stride;
      stride *= A.nn[k];
    }
#pragma unroll
    for (int k = 0; k < D; ++k)
      X[a][k] = A.xyz ? A.xyz[(int64_t)node * D + k] : (double)(ci[k] + ((a >> k) & 1)) * A.h[k];
  }
  const double kap = A.kappa ? A.kappa[cell] : 1.0;
  const double fc = A.fsrc ? A.fsrc[cell] : 1.0;
  T K[NT], F[NL];
#pragma unroll
  for (int t = 0; t < NT; ++t) K[t] = 0;
#pragma unroll
  for (int a = 0; a < NL; ++a) F[a] = 0;
  const int stride = 1 + NL + NL * D;
  for (int q = 0; q < A.nq; ++q) {
    const double *tq = A.tab + q * stride;
    double J[D * D], Ji[D * D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < NL; ++a) s += tq[1 + NL + a * D + i] * X[a][j];
        J[i * D + j] = s;
      }
    const double det = det_inv<D>(J, Ji);
    const T dV = (T)(fabs(det) * tq[0]);
    T g[NL][D];
#pragma unroll
    for (int a = 0; a < NL; ++a)
#pragma unroll
      for (int r = 0; r < D; ++r) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) s += Ji[r * D + k] * tq[1 + NL + a * D + k];
        g[a][r] = (T)s;
      }
    const T kd = (T)kap * dV;
    const T sd = A.sigma ? (T)A.sigma[cell] * dV : (T)0;
    const T fd = (T)(A.fq ? A.fq[(int64_t)cell * A.nq + q] : fc) * dV;
    int t = 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
#pragma unroll
      for (int j = i; j < NL; ++j, ++t) {
        T dot = 0;
#pragma unroll
        for (int r = 0; r < D; ++r) dot += g[i][r] * g[j][r];
        K[t] += dot * kd;
        if (A.sigma) K[t] += (T)tq[1 + i] * (T)tq[1 + j] * sd;
      }
      F[i] += (T)tq[1 + i] * fd;
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) E[(int64_t)t * A.ncells + cell] = K[t];
#pragma unroll
  for (int a = 0; a < NL; ++a) E[(int64_t)(NT + a) * A.ncells + cell] = F[a];
}

// The owner-computes column kernel over the element matrices stored in c.qelem
// (FP32 when f32, else FP64); eload: FP64 element loads in place of qelem's load
// slots, or nullptr; lift: subtract the Dirichlet lifting from b.
template <int D>
void lattice_columns(Ctx &c, bool f32, const double *eload, bool lift) {
  LatticeArgs P{};
  for (int k = 0; k < 3; ++k) {
    P.nn[k] = (int)(c.part[k] + 1);
    P.nc[k] = (int)c.part[k];
  }
  P.nnodes = (int)c.nnodes;
  P.inv_nn0 = 1.0 / P.nn[0];
  P.inv_nn1 = 1.0 / P.nn[1];
  P.node_dof = c.node_dof.get();
  P.fmask = c.node_fmask.get();
  P.tile_p0 = c.tile_p0.get();
  P.tile_d0 = c.tile_d0.get();
  P.colptr = c.colptr.get();
  P.dvals = lift && c.have_dvals ? c.dvals.get() : nullptr;
  P.nzval = c.nzval.get();
  P.b = c.b.get();
  P.elem = c.qelem.get();
  P.eload = eload;
  P.ncells = (int)c.ncells;
  const int ntiles = (int)((c.nnodes + kTileNodes - 1) / kTileNodes);
  const unsigned grid = (unsigned)((ntiles + kLatticeWaves - 1) / kLatticeWaves);
  const dim3 block(kTileNodes * kLatticeWaves);
  if (f32)
    hipLaunchKernelGGL((lattice_q1_kernel<D, 4>), dim3(grid), block, 0, c.stream, P, ntiles);
  else
    hipLaunchKernelGGL((lattice_q1_kernel<D, 8>), dim3(grid), block, 0, c.stream, P, ntiles);
  GH_CHECK_HIP(hipGetLastError());
}

template <int D>
void launch_lattice_mapped(Ctx &c, int precision) {
  c.last_kernel = GH_KERNEL_Q1_LATTICE;
  constexpr int NL = 1 << D;
  constexpr int NT = NL * (NL + 1) / 2;
  GH_REQUIRE(c.cartesian || (c.nv == NL && c.xyz.n == (size_t)c.nnodes * D), GH_ERR_STATE,
             "the mapped lattice needs the node coordinates");
  const bool f32 = precision == GH_PREC_FP32_ELEM;
  const size_t esz = f32 ? sizeof(float) : sizeof(double);
  c.qelem.resize((size_t)(NT + NL) * c.ncells * esz);
  const int stride = 1 + NL + NL * D;
  std::vector<double> tab(size_t(c.nq) * stride);
  for (int q = 0; q < c.nq; ++q) {
    tab[q * stride] = c.qw[q];
    for (int a = 0; a < NL; ++a) tab[q * stride + 1 + a] = c.shp[q * NL + a];
    for (int a = 0; a < NL * D; ++a) tab[q * stride + 1 + NL + a] = c.grad[q * NL * D + a];
  }
  c.nl_tab.resize(tab.size());
  GH_CHECK_HIP(hipMemcpyAsync(c.nl_tab.get(), tab.data(), sizeof(double) * tab.size(),
                              hipMemcpyHostToDevice, c.stream));
  QElemArgs A{};
  for (int k = 0; k < 3; ++k) {
    A.nc[k] = (int)c.part[k];
    A.nn[k] = (int)(c.part[k] + 1);
  }
  A.ncells = (int)c.ncells;
  A.nq = c.nq;
  A.xyz = c.cartesian ? nullptr : c.xyz.get();
  for (int k = 0; k < 3; ++k) A.h[k] = c.sizes[k];
  A.tab = c.nl_tab.get();
  A.kappa = c.have_kappa ? c.kappa.get() : nullptr;
  A.fsrc = c.have_fsrc ? c.fsrc.get() : nullptr;
  A.sigma = c.have_sigma ? c.sigma.get() : nullptr;
  A.fq = c.have_fq ? c.fq.get() : nullptr;
  const unsigned eg = (unsigned)((c.ncells + 255) / 256);
  if (f32)
    hipLaunchKernelGGL((q1_lattice_elem_kernel<D, true>), dim3(eg), dim3(256), 0, c.stream, A,
                       (float *)c.qelem.get());
  else
    hipLaunchKernelGGL((q1_lattice_elem_kernel<D, false>), dim3(eg), dim3(256), 0, c.stream, A,
                       (double *)c.qelem.get());
  GH_CHECK_HIP(hipGetLastError());
  lattice_columns<D>(c, f32, nullptr, true);
  // the packed tabulation's host copy lives until here
  GH_CHECK_HIP(hipStreamSynchronize(c.stream));
}

// ---------------------------------------------------------------------------
// General per-cell kernel (Q1 on quads/hexes, any geometry), coloured scatter
// ---------------------------------------------------------------------------

struct CellArgs {
  const int32_t *cells;  // cells of this colour
  int64_t ncells;
  const int32_t *cell_dofs;
  const int32_t *cell_nodes;  // nullptr: Cartesian analytic
  const double *xyz;
  double h[3];
  int nq;
  const double *tab;  // per q: w, shp[NLOC], grad[NLOC*D]
  const double *kappa, *fsrc, *dvals;
  const double *sigma;  // mass coefficient per cell, or nullptr (no mass term)
  const double *fq;  // f at the quadrature points (cells x nq) or nullptr
  const int32_t *colptr, *rowval;
  double *nzval, *b;
};

template <int D, bool F32>
__global__ __launch_bounds__(128) void cell_q1_kernel(const CellArgs A) {
  constexpr int NL = 1 << D;
  using T = typename std::conditional<F32, float, double>::type;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= A.ncells) return;
  const int64_t cell = A.cells[t];
  int32_t dofs[NL];
#pragma unroll
  for (int a = 0; a < NL; ++a) dofs[a] = A.cell_dofs[cell * NL + a];
  double X[NL][D];
  if (A.cell_nodes) {
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      int64_t nd = A.cell_nodes[cell * NL + a];
#pragma unroll
      for (int k = 0; k < D; ++k) X[a][k] = A.xyz[nd * D + k];
    }
  }
  const double kap = A.kappa ? A.kappa[cell] : 1.0;
  const double fc = A.fsrc ? A.fsrc[cell] : 1.0;
  T K[NL][NL];
  T bv[NL];
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    bv[i] = 0;
#pragma unroll
    for (int j = 0; j < NL; ++j) K[i][j] = 0;
  }
  const int stride = 1 + NL + NL * D;
  for (int q = 0; q < A.nq; ++q) {
    const double *tq = A.tab + q * stride;
    double J[D * D], Ji[D * D];
    if (A.cell_nodes) {
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
          double s = 0.0;
#pragma unroll
          for (int a = 0; a < NL; ++a) s += tq[1 + NL + a * D + i] * X[a][j];
          J[i * D + j] = s;
        }
    } else {
#pragma unroll
      for (int i = 0; i < D * D; ++i) J[i] = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) J[k * D + k] = A.h[k];
    }
    const double det = det_inv<D>(J, Ji);
    const T dV = (T)(fabs(det) * tq[0]);
    T g[NL][D];
#pragma unroll
    for (int a = 0; a < NL; ++a)
#pragma unroll
      for (int r = 0; r < D; ++r) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) s += Ji[r * D + k] * tq[1 + NL + a * D + k];
        g[a][r] = (T)s;
      }
    const T kd = (T)kap * dV;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
#pragma unroll
      for (int j = 0; j < NL; ++j) {
        T dot = 0;
#pragma unroll
        for (int r = 0; r < D; ++r) dot += g[i][r] * g[j][r];
        K[i][j] += dot * kd;
      }
      bv[i] += (T)tq[1 + i] * ((T)(A.fq ? A.fq[cell * A.nq + q] : fc) * dV);
    }
    if (A.sigma) {  // + σ φ_i φ_j at this point
      const T sd = (T)A.sigma[cell] * dV;
#pragma unroll
      for (int i = 0; i < NL; ++i)
#pragma unroll
        for (int j = 0; j < NL; ++j) K[i][j] += (T)tq[1 + i] * (T)tq[1 + j] * sd;
    }
  }
  // Dirichlet lifting b -= K u_d on the cell.
  if (A.dvals) {
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < NL; ++j)
        if (dofs[j] < 0) s += (double)K[i][j] * A.dvals[-dofs[j] - 1];
      bv[i] = (T)((double)bv[i] - s);
    }
  }
  // Scatter: columns lj, rows li, Dirichlet ids skipped.
#pragma unroll
  for (int lj = 0; lj < NL; ++lj) {
    const int32_t j = dofs[lj];
    if (j <= 0) continue;
    const int32_t p0 = A.colptr[j - 1], p1 = A.colptr[j];
#pragma unroll
    for (int li = 0; li < NL; ++li) {
      const int32_t i = dofs[li];
      if (i <= 0) continue;
      int lo = p0, hi = p1;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (A.rowval[mid] < i - 1) lo = mid + 1;
        else hi = mid;
      }
      if (lo < p1 && A.rowval[lo] == i - 1) A.nzval[lo] += (double)K[li][lj];
    }
  }
#pragma unroll
  for (int li = 0; li < NL; ++li)
    if (dofs[li] > 0) A.b[dofs[li] - 1] += (double)bv[li];
}

void launch_general(Ctx &c, int precision) {
  c.last_kernel = GH_KERNEL_CELL;
  GH_REQUIRE(c.reffe_kind == GH_REFFE_QK && c.order == 1 && c.ncomp == 1 &&
                 c.nloc == (1 << c.dim),
             GH_ERR_UNSUPPORTED, "general path supports scalar Q1 on quads/hexes");
  GH_REQUIRE(c.cartesian || c.nv == c.nloc, GH_ERR_UNSUPPORTED,
             "geometry must be Q1 (2^D nodes per cell)");
  GH_REQUIRE((int)c.qw.size() == c.nq && (int)c.shp.size() == c.nq * c.nloc, GH_ERR_STATE,
             "tabulation does not match the element (dofs per cell)");
  build_coloring(c);
  const int NL = c.nloc, D = c.dim;
  const int stride = 1 + NL + NL * D;
  std::vector<double> tab(size_t(c.nq) * stride);
  for (int q = 0; q < c.nq; ++q) {
    tab[q * stride] = c.qw[q];
    for (int a = 0; a < NL; ++a) tab[q * stride + 1 + a] = c.shp[q * NL + a];
    for (int a = 0; a < NL * D; ++a) tab[q * stride + 1 + NL + a] = c.grad[q * NL * D + a];
  }
  DBuf<double> dtab;
  dtab.resize(tab.size());
  GH_CHECK_HIP(hipMemcpyAsync(dtab.get(), tab.data(), sizeof(double) * tab.size(),
                              hipMemcpyHostToDevice, c.stream));
  GH_CHECK_HIP(hipMemsetAsync(c.nzval.get(), 0, sizeof(double) * std::max<int64_t>(c.nnz, 1),
                              c.stream));
  GH_CHECK_HIP(hipMemsetAsync(c.b.get(), 0, sizeof(double) * std::max<int64_t>(c.nfree, 1),
                              c.stream));
  CellArgs A{};
  A.cell_dofs = c.cell_dofs.get();
  A.cell_nodes = c.cartesian ? nullptr : c.cell_nodes.get();
  A.xyz = c.cartesian ? nullptr : c.xyz.get();
  for (int k = 0; k < 3; ++k) A.h[k] = c.sizes[k];
  A.nq = c.nq;
  A.tab = dtab.get();
  A.kappa = c.have_kappa ? c.kappa.get() : nullptr;
  A.fsrc = c.have_fsrc ? c.fsrc.get() : nullptr;
  A.sigma = c.have_sigma ? c.sigma.get() : nullptr;
  A.dvals = c.have_dvals ? c.dvals.get() : nullptr;
  A.fq = c.have_fq ? c.fq.get() : nullptr;
  A.colptr = c.colptr.get();
  A.rowval = c.rowval.get();
  A.nzval = c.nzval.get();
  A.b = c.b.get();
  for (int col = 0; col < c.ncolors; ++col) {
    A.cells = c.color_cells.get() + c.color_ptr[col];
    A.ncells = c.color_ptr[col + 1] - c.color_ptr[col];
    if (A.ncells == 0) continue;
    unsigned grid = (unsigned)((A.ncells + 127) / 128);
    bool f32 = precision == GH_PREC_FP32_ELEM;
    if (D == 2) {
      if (f32) hipLaunchKernelGGL((cell_q1_kernel<2, true>), dim3(grid), dim3(128), 0, c.stream, A);
      else hipLaunchKernelGGL((cell_q1_kernel<2, false>), dim3(grid), dim3(128), 0, c.stream, A);
    } else {
      if (f32) hipLaunchKernelGGL((cell_q1_kernel<3, true>), dim3(grid), dim3(128), 0, c.stream, A);
      else hipLaunchKernelGGL((cell_q1_kernel<3, false>), dim3(grid), dim3(128), 0, c.stream, A);
    }
    GH_CHECK_HIP(hipGetLastError());
  }
  GH_CHECK_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace

void launch_lattice_columns(Ctx &c, bool f32, const double *eload, bool lift) {
  if (c.dim == 3)
    lattice_columns<3>(c, f32, eload, lift);
  else
    lattice_columns<2>(c, f32, eload, lift);
}

void assemble_diffusion(Ctx &c, int precision, int path) {
  GH_REQUIRE(!c.have_kappa || c.kappa.n == (size_t)c.ncells, GH_ERR_STATE,
             "kappa was uploaded for a different mesh");
  GH_REQUIRE(!c.have_dvals || c.dvals.n == (size_t)c.ndiri, GH_ERR_STATE,
             "Dirichlet values were uploaded for a different space");
  GH_REQUIRE(!c.have_sigma || (c.reffe_kind == GH_REFFE_QK && c.order == 1 && c.ncomp == 1),
             GH_ERR_UNSUPPORTED, "the cell mass term is provided for scalar Q1 spaces");
  if (c.elemlist && simplex_eligible(c)) {
    GH_REQUIRE(path != GH_PATH_STRUCTURED, GH_ERR_UNSUPPORTED,
               "the lattice path needs a Cartesian mesh, a scalar Q1 CLagrangian space and a "
               "symmetric quadrature rule");
    assemble_simplex(c, precision);
    return;
  }
  GH_REQUIRE(!c.have_fq || c.fq.n == (size_t)c.ncells * c.nq, GH_ERR_STATE,
             "the quadrature-point source was uploaded for another mesh or quadrature");
  double Kc[8] = {0}, b0 = 0.0;
  // the lattice kernel takes Cartesian cells with f constant per cell and no mass term;
  // mapped cells, f at the quadrature points or a mass term store the element matrices
  // per cell first (the same owner-computes column kernel then sums them)
  bool lattice = c.structured && path != GH_PATH_GENERAL;
  if (lattice && (!c.cartesian || c.have_fq || c.have_sigma)) {
    // element matrices per cell, owner-computes columns
    GH_REQUIRE((int)c.shp.size() == c.nq * c.nloc && (int)c.grad.size() == c.nq * c.nloc * c.dim,
               GH_ERR_STATE, "tabulation does not match the element (dofs per cell)");
    if (c.dim == 3)
      launch_lattice_mapped<3>(c, precision);
    else
      launch_lattice_mapped<2>(c, precision);
    return;
  }
  if (lattice) {
    GH_REQUIRE((int)c.shp.size() == c.nq * c.nloc, GH_ERR_STATE,
               "tabulation does not match the element (dofs per cell)");
    lattice = lattice_reference_symmetric(c, Kc, &b0);
  }
  GH_REQUIRE(path != GH_PATH_STRUCTURED || lattice, GH_ERR_UNSUPPORTED,
             "the lattice path needs a Cartesian mesh, a scalar Q1 CLagrangian space and a "
             "symmetric quadrature rule");
  if (lattice) {
    // Enqueued only: gh_values_get (or a wait on gh_stream) orders the results.
    // The lattice arithmetic is FP64 in both precision modes (see the kernel).
    if (c.dim == 3)
      launch_lattice<3>(c, Kc, b0);
    else
      launch_lattice<2>(c, Kc, b0);
  } else {
    launch_general(c, precision);
  }
}

}  // namespace gh
