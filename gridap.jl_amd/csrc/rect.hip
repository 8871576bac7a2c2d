// Rectangular field pairs: a test space (rows) and a different trial space (columns) on
// one mesh, the off-diagonal blocks of a multi-field form whose fields live in
// different FE spaces (src/MultiField/BlockSparseMatrixAssemblers.jl:19-36; the block
// (i, j) of a MultiFieldFESpace pairs the test basis of field i with the trial basis of
// field j, src/MultiField/MultiFieldCellFields.jl).
//
// Symbolic: every cell emits the (trial column, test row) pairs of its free dofs as
// packed keys (_add_entries! with rows and columns from two dof maps,
// src/Algebra/AlgebraInterfaces.jl:175-187), one radix sort orders them with their
// emit index, and a unique-flag scan writes rowval, colptr (columns may be empty: a
// trial dof whose cells hold only Dirichlet test dofs) and the contribution lists
// (cstart, perm).  Numeric: one thread per cell forms the block's element matrix, for
// scalar test ψ and trial φ
//   B_c[lj][li] = Σ_q w|J| (κ_c ∇ψ_li⋅∇φ_lj + σ_c ψ_li φ_lj),
// or the divergence coupling between a scalar and a vector space (local dofs
// component-blocked, node + nn·component)
//   scalar test, vector trial: B_c[(b,j)][a] = κ_c Σ_q w|J| ψ_a ∂_j φ_b,
//   vector test, scalar trial: B_c[b][(a,i)] = κ_c Σ_q w|J| ∂_i ψ_a φ_b
// (the geometry of the cell: Cartesian sizes, or the Q1 map of its vertices) and its
// lifting load -Σ_{lj Dirichlet} B_c[lj][li] u_D(lj); one thread per
// nonzero sums its entries in ascending cell order and writes nzval once; one thread
// per test row sums the lifting loads of its cells (ascending).  Deterministic, no
// atomics on values.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "gh_internal.h"

namespace gh {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, (n + kBlock - 1) / kBlock);
}

int bits_for(int64_t v) {  // bits needed to hold values 0..v
  int b = 1;
  while ((int64_t(1) << b) <= v) ++b;
  return b;
}

// t = cell * ns * nt + lj * ns + li: the emit index addresses B_c[lj][li] directly.
__global__ void rect_emit(int64_t ncells, int ns, int nt, const int32_t *__restrict__ rows,
                          const int32_t *__restrict__ cols, int rb, uint64_t invalid,
                          uint64_t *__restrict__ keys, int32_t *__restrict__ ids) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)ns * nt;
  if (t >= ncells * per) return;
  const int64_t c = t / per;
  const int r = (int)(t - c * per);
  const int lj = r / ns, li = r - lj * ns;
  const int32_t j = cols[c * nt + lj];
  const int32_t i = rows[c * ns + li];
  keys[t] = (i > 0 && j > 0) ? ((uint64_t)(j - 1) << rb) | (uint64_t)(i - 1) : invalid;
  ids[t] = (int32_t)t;
}

__global__ void rect_flags(int64_t n, const uint64_t *__restrict__ keys, uint64_t invalid,
                           int32_t *__restrict__ flags) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t key = keys[k];
  flags[k] = key < invalid && (k == 0 || keys[k - 1] != key);
}

// Per unique key: its row into rowval, one count into its column; contribution starts.
__global__ void rect_rows(int64_t n, const uint64_t *__restrict__ keys, uint64_t invalid,
                          int rb, const int32_t *__restrict__ flags,
                          const int32_t *__restrict__ pos, int64_t nnz,
                          int32_t *__restrict__ rowval, int32_t *__restrict__ colcount,
                          int32_t *__restrict__ cstart) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t key = keys[k];
  if (flags[k]) {
    rowval[pos[k]] = (int32_t)(key & ((uint64_t(1) << rb) - 1));
    atomicAdd(&colcount[key >> rb], 1);  // a count: the same for any order
    cstart[pos[k]] = (int32_t)k;
  }
  if (key < invalid && (k + 1 == n || keys[k + 1] >= invalid)) cstart[nnz] = (int32_t)(k + 1);
}

struct RectArgs {
  int64_t ncells;
  int ns, nt, nq, nv;        // local dofs of the test and trial spaces
  int nns, ntn;              // their scalar shape functions (nodes)
  bool div;                  // the divergence coupling (else κ ∇ψ⋅∇φ + σ ψ φ)
  const double *ttab;        // per q: w, ψ[nns], ∇̂ψ[nns*D], φ[ntn], ∇̂φ[ntn*D]
  const double *gtab;        // Q1 map gradients per q ([nv*D]), or nullptr (Cartesian)
  const int32_t *cell_nodes; // 0-based node ids of the cell's vertices
  const double *xyz;
  double h[3];
  const double *kappa, *sigma;  // per cell, or nullptr (no such term)
  const int32_t *tcell_dofs;    // trial dof ids (lifting)
  const double *tdvals;         // trial Dirichlet values, or nullptr
};

template <class T, int D>
__global__ __launch_bounds__(kBlock) void rect_elem_kernel(const RectArgs A, T *__restrict__ E,
                                                           double *__restrict__ Ev) {
  const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= A.ncells) return;
  const int ns = A.ns, nt = A.nt, nns = A.nns, ntn = A.ntn;
  const int stride = 1 + (nns + ntn) * (1 + D);
  const double kap = A.kappa ? A.kappa[cell] : (A.div ? 1.0 : 0.0);
  const double sig = A.sigma ? A.sigma[cell] : 0.0;
  T *Ec = E + cell * ns * nt;
  double *Evc = Ev + cell * ns;
  if (!A.gtab) {
    // Cartesian cells: J = diag(h) at every point, so each entry is summed over the
    // points in a register and stored once
    double ih[D], det = 1.0;
    for (int d = 0; d < D; ++d) {
      ih[d] = 1.0 / A.h[d];
      det *= A.h[d];
    }
    det = fabs(det);
    for (int lj = 0; lj < nt; ++lj)
      for (int li = 0; li < ns; ++li) {
        double s = 0.0;
        for (int q = 0; q < A.nq; ++q) {
          const double *tq = A.ttab + (size_t)q * stride;
          const double *psi = tq + 1, *gpsi = psi + nns, *phi = gpsi + nns * D,
                       *gphi = phi + ntn;
          const double w = tq[0] * det;
          double v;
          if (A.div && nt > ntn) {  // (b, j) x a: ψ_a ∂_j φ_b
            const int b = lj % ntn, j = lj / ntn;
            v = kap * psi[li] * (gphi[b * D + j] * ih[j]);
          } else if (A.div) {       // b x (a, i): ∂_i ψ_a φ_b
            const int a = li % nns, i = li / nns;
            v = kap * (gpsi[a * D + i] * ih[i]) * phi[lj];
          } else {
            double dot = 0.0;
            for (int d = 0; d < D; ++d)
              dot += (gpsi[li * D + d] * ih[d]) * (gphi[lj * D + d] * ih[d]);
            v = kap * dot + sig * (psi[li] * phi[lj]);
          }
          s += w * v;
        }
        Ec[lj * ns + li] = (T)s;
      }
  } else {
    // mapped cells: the Jacobian of the Q1 map at each point
    for (int k = 0; k < ns * nt; ++k) Ec[k] = 0;
    double X[8][D];
    if (A.gtab)
      for (int v = 0; v < A.nv; ++v)
        for (int d = 0; d < D; ++d) X[v][d] = A.xyz[(int64_t)A.cell_nodes[cell * A.nv + v] * D + d];
    for (int q = 0; q < A.nq; ++q) {
      const double *tq = A.ttab + (size_t)q * stride;
      double Ji[D][D], det = 1.0;
      if (A.gtab) {
        // J_de = ∂x_d/∂ξ_e of the Q1 map; ∇φ = J⁻ᵀ∇̂φ
        double J[D * D], Jinv[D * D];
        for (int k = 0; k < D * D; ++k) J[k] = 0.0;
        const double *gq = A.gtab + (size_t)q * A.nv * D;
        for (int v = 0; v < A.nv; ++v)
          for (int d = 0; d < D; ++d)
            for (int e = 0; e < D; ++e) J[d * D + e] += X[v][d] * gq[v * D + e];
        det = det_inv<D>(J, Jinv);
        for (int d = 0; d < D; ++d)
          for (int e = 0; e < D; ++e) Ji[d][e] = Jinv[d * D + e];
      } else {
        for (int d = 0; d < D; ++d) {
          det *= A.h[d];
          for (int e = 0; e < D; ++e) Ji[d][e] = d == e ? 1.0 / A.h[d] : 0.0;
        }
      }
      const double w = tq[0] * fabs(det);
      const double *psi = tq + 1, *gpsi = psi + nns, *phi = gpsi + nns * D, *gphi = phi + ntn;
      if (A.div && nt > ntn) {  // scalar test, vector trial: ψ_a ∂_j φ_b
        for (int b = 0; b < ntn; ++b)
          for (int j = 0; j < D; ++j) {
            double g = 0.0;
            for (int e = 0; e < D; ++e) g += Ji[e][j] * gphi[b * D + e];
            const double wg = w * kap * g;
            for (int a = 0; a < nns; ++a) Ec[(b + ntn * j) * ns + a] += (T)(wg * psi[a]);
          }
        continue;
      }
      if (A.div) {  // vector test, scalar trial: ∂_i ψ_a φ_b
        for (int a = 0; a < nns; ++a)
          for (int i = 0; i < D; ++i) {
            double g = 0.0;
            for (int e = 0; e < D; ++e) g += Ji[e][i] * gpsi[a * D + e];
            const double wg = w * kap * g;
            for (int b = 0; b < ntn; ++b) Ec[b * ns + a + nns * i] += (T)(wg * phi[b]);
          }
        continue;
      }
      for (int lj = 0; lj < nt; ++lj) {
        double gj[D];
        for (int d = 0; d < D; ++d) {
          double s = 0.0;
          for (int e = 0; e < D; ++e) s += Ji[e][d] * gphi[lj * D + e];
          gj[d] = s;
        }
        for (int li = 0; li < ns; ++li) {
          double dot = 0.0;
          for (int d = 0; d < D; ++d) {
            double s = 0.0;
            for (int e = 0; e < D; ++e) s += Ji[e][d] * gpsi[li * D + e];
            dot += s * gj[d];
          }
          Ec[lj * ns + li] += (T)(w * (kap * dot + sig * (psi[li] * phi[lj])));
        }
      }
    }
  }
  for (int li = 0; li < ns; ++li) {
    double v = 0.0;
    if (A.tdvals)
      for (int lj = 0; lj < nt; ++lj) {
        const int32_t d = A.tcell_dofs[cell * nt + lj];
        if (d < 0) v -= (double)Ec[lj * ns + li] * A.tdvals[-d - 1];
      }
    Evc[li] = v;
  }
}

template <class T>
__global__ void rect_sum_kernel(int64_t nnz, const int32_t *__restrict__ cstart,
                                const int32_t *__restrict__ perm, const T *__restrict__ E,
                                double *__restrict__ nzval) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  // batches of 4: the index loads, then the value loads, then the ordered sum
  constexpr int U = 4;
  const int32_t e = cstart[k + 1];
  double s = 0.0;
  for (int32_t m0 = cstart[k]; m0 < e; m0 += U) {
    int32_t id[U];
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) id[u] = m0 + u < e ? perm[m0 + u] : -1;
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = id[u] >= 0 ? E[id[u]] : (T)0;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (id[u] >= 0) s += (double)v[u];
  }
  nzval[k] = s;
}

__global__ void rect_rhs_kernel(int64_t nfree, int ns, const int32_t *__restrict__ adjptr,
                                const int32_t *__restrict__ adj, const double *__restrict__ Ev,
                                double *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfree) return;
  double s = 0.0;
  for (int32_t k = adjptr[i]; k < adjptr[i + 1]; ++k) s += Ev[adj[k]];
  b[i] = s;
}

// out[dst[k]] = v[k]: one block's stored values to their composite positions.  A block
// column's values are one consecutive run of composite positions, so both sides of the
// move stay coalesced.
__global__ void compose_move_kernel(int64_t n, const double *__restrict__ v,
                                    const int32_t *__restrict__ dst, double *__restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    out[dst[k]] = v[k];
}

}  // namespace

void compose_move(Ctx &c, int block) {
  GH_REQUIRE(block >= 0 && block < (int)c.comp.size(), GH_ERR_ARG, "no such composite block");
  const auto &blk = c.comp[(size_t)block];
  Ctx &s = *blk->src;
  GH_REQUIRE(s.have_values, GH_ERR_STATE, "the composite block holds no assembled values");
  GH_REQUIRE(s.have_pattern && s.nnz == blk->n, GH_ERR_STATE,
             "the composite block's pattern changed since it was registered");
  if (blk->n == 0) return;
  // on the block's stream: after the assembly that wrote its values, before the next
  const int64_t blocks = std::min<int64_t>((blk->n + 255) / 256, 8192);
  compose_move_kernel<<<(unsigned)blocks, 256, 0, s.stream>>>(blk->n, s.nzval.get(),
                                                             blk->dst.get(), c.comp_nz.get());
  GH_CHECK_HIP(hipGetLastError());
}

void rect_symbolic(Ctx &c) {
  hipStream_t s = c.stream;
  const int ns = c.nloc, nt = c.tnloc;
  const int64_t total = c.ncells * ns * nt;
  GH_REQUIRE(total < (int64_t(1) << 31), GH_ERR_UNSUPPORTED,
             "cell pair count exceeds 2^31 in one symbolic pass");
  const int rb = bits_for(std::max<int64_t>({c.nfree, c.tnfree, 1}));
  GH_REQUIRE(2 * rb <= 64, GH_ERR_UNSUPPORTED, "dof count too large for packed keys");
  const uint64_t invalid = (uint64_t)std::max<int64_t>(c.tnfree, 1) << rb;
  DBuf<uint64_t> keys, sorted;
  DBuf<int32_t> ids, sids;
  const int64_t n = std::max<int64_t>(total, 1);
  keys.resize(n);
  sorted.resize(n);
  ids.resize(n);
  sids.resize(n);
  if (total)
    hipLaunchKernelGGL(rect_emit, dim3(grid_for(total)), dim3(kBlock), 0, s, c.ncells, ns, nt,
                       c.cell_dofs.get(), c.tcell_dofs.get(), rb, invalid, keys.get(),
                       ids.get());
  GH_CHECK_HIP(hipGetLastError());
  size_t bytes = 0;
  GH_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys.get(), sorted.get(),
                                                  ids.get(), sids.get(), (int)total, 0, 2 * rb,
                                                  s));
  {
    DBuf<unsigned char> tmp;
    tmp.resize(std::max<size_t>(bytes, 1));
    GH_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.get(), bytes, keys.get(), sorted.get(),
                                                    ids.get(), sids.get(), (int)total, 0,
                                                    2 * rb, s));
    GH_CHECK_HIP(hipStreamSynchronize(s));
  }
  keys.release();
  ids.release();
  DBuf<int32_t> flags, pos;
  flags.resize(n);
  pos.resize(n);
  int64_t nnz = 0;
  if (total) {
    hipLaunchKernelGGL(rect_flags, dim3(grid_for(total)), dim3(kBlock), 0, s, total,
                       sorted.get(), invalid, flags.get());
    GH_CHECK_HIP(hipGetLastError());
    size_t sb = 0;
    GH_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, flags.get(), pos.get(),
                                                  (int)total, s));
    DBuf<unsigned char> tmp;
    tmp.resize(std::max<size_t>(sb, 1));
    GH_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.get(), sb, flags.get(), pos.get(),
                                                  (int)total, s));
    int32_t last[2] = {0, 0};
    GH_CHECK_HIP(hipMemcpyAsync(&last[0], pos.get() + total - 1, 4, hipMemcpyDeviceToHost, s));
    GH_CHECK_HIP(hipMemcpyAsync(&last[1], flags.get() + total - 1, 4, hipMemcpyDeviceToHost, s));
    GH_CHECK_HIP(hipStreamSynchronize(s));
    nnz = (int64_t)last[0] + last[1];
  }
  c.nnz = nnz;
  c.rowval.resize(std::max<int64_t>(nnz, 1));
  c.cstart.resize(nnz + 1);
  DBuf<int32_t> cnt;
  cnt.resize(c.tnfree + 1);
  GH_CHECK_HIP(hipMemsetAsync(cnt.get(), 0, sizeof(int32_t) * (c.tnfree + 1), s));
  GH_CHECK_HIP(hipMemsetAsync(c.cstart.get(), 0, sizeof(int32_t) * (nnz + 1), s));
  if (total)
    hipLaunchKernelGGL(rect_rows, dim3(grid_for(total)), dim3(kBlock), 0, s, total, sorted.get(),
                       invalid, rb, flags.get(), pos.get(), nnz, c.rowval.get(), cnt.get(),
                       c.cstart.get());
  GH_CHECK_HIP(hipGetLastError());
  // colptr = exclusive scan of the column counts (the last count is 0, so the last
  // entry is nnz)
  c.colptr.resize(c.tnfree + 1);
  {
    size_t sb = 0;
    GH_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, cnt.get(), c.colptr.get(),
                                                  (int)(c.tnfree + 1), s));
    DBuf<unsigned char> tmp;
    tmp.resize(std::max<size_t>(sb, 1));
    GH_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.get(), sb, cnt.get(), c.colptr.get(),
                                                  (int)(c.tnfree + 1), s));
    GH_CHECK_HIP(hipStreamSynchronize(s));
  }
  c.perm.resize(n);
  std::swap(c.perm.p, sids.p);
  std::swap(c.perm.n, sids.n);
  // test adjacency for the lifting loads: per free test dof, its (cell, local) entries
  // in ascending cell order, as indices cell * ns + li into the per-cell loads
  std::vector<int32_t> aptr(c.nfree + 1, 0), adj;
  for (int64_t k = 0; k < c.ncells * ns; ++k) {
    const int32_t d = c.h_cell_dofs[k];
    if (d > 0) aptr[d]++;
  }
  for (int64_t i = 0; i < c.nfree; ++i) aptr[i + 1] += aptr[i];
  adj.resize(std::max<int32_t>(aptr[c.nfree], 1));
  std::vector<int32_t> cur(aptr.begin(), aptr.end() - 1);
  for (int64_t k = 0; k < c.ncells * ns; ++k) {
    const int32_t d = c.h_cell_dofs[k];
    if (d > 0) adj[cur[d - 1]++] = (int32_t)k;
  }
  c.adjptr.resize(aptr.size());
  c.adj.resize(adj.size());
  GH_CHECK_HIP(hipMemcpy(c.adjptr.get(), aptr.data(), sizeof(int32_t) * aptr.size(),
                         hipMemcpyHostToDevice));
  GH_CHECK_HIP(hipMemcpy(c.adj.get(), adj.data(), sizeof(int32_t) * adj.size(),
                         hipMemcpyHostToDevice));
  c.have_pattern = true;
  c.nzval.resize(std::max<int64_t>(nnz, 1));
  c.b.resize(std::max<int64_t>(c.nfree, 1));
}

void rect_assemble(Ctx &c, int form, int precision) {
  const int D = c.dim, ns = c.nloc, nt = c.tnloc;
  const int nns = ns / c.ncomp, ntn = nt / c.tncomp;
  const bool div = form == GH_FORM_DIVERGENCE;
  GH_REQUIRE(D == 2 || D == 3, GH_ERR_UNSUPPORTED, "rectangular pairs on 2-D and 3-D meshes");
  GH_REQUIRE(div ? (c.ncomp == 1) != (c.tncomp == 1) : (c.ncomp == 1 && c.tncomp == 1),
             GH_ERR_UNSUPPORTED,
             div ? "the divergence coupling pairs a scalar space with a vector space"
                 : "κ ∇ψ⋅∇φ + σ ψ φ pairs two scalar spaces");
  GH_REQUIRE(c.nq > 0 && c.tnq == c.nq && (int)c.shp.size() == c.nq * nns &&
                 (int)c.grad.size() == c.nq * nns * D && (int)c.tshp.size() == c.nq * ntn &&
                 (int)c.tgrad.size() == c.nq * ntn * D,
             GH_ERR_STATE, "upload the test and trial tabulations at the same points");
  GH_REQUIRE(c.cartesian || (c.have_gtab && c.gtab_nq == c.nq && c.nv <= 8 &&
                             c.nv == (1 << D)),
             GH_ERR_STATE,
             "a mapped mesh needs the geometry tabulation (gh_geometry_tabulation) at the "
             "points of the tabulations");
  GH_REQUIRE(div || c.have_kappa || c.have_sigma, GH_ERR_STATE,
             "a rectangular pair needs κ (gh_coef_upload) or σ (gh_mass_upload)");
  GH_REQUIRE(!div || !c.have_sigma, GH_ERR_UNSUPPORTED,
             "the divergence coupling has no mass term");
  const int stride = 1 + (nns + ntn) * (1 + D);
  std::vector<double> tab((size_t)c.nq * stride);
  for (int q = 0; q < c.nq; ++q) {
    double *tq = &tab[(size_t)q * stride];
    tq[0] = c.qw[q];
    double *psi = tq + 1, *gpsi = psi + nns, *phi = gpsi + nns * D, *gphi = phi + ntn;
    for (int a = 0; a < nns; ++a) psi[a] = c.shp[(size_t)q * nns + a];
    for (int a = 0; a < nns * D; ++a) gpsi[a] = c.grad[(size_t)q * nns * D + a];
    for (int a = 0; a < ntn; ++a) phi[a] = c.tshp[(size_t)q * ntn + a];
    for (int a = 0; a < ntn * D; ++a) gphi[a] = c.tgrad[(size_t)q * ntn * D + a];
  }
  DBuf<double> dtab;
  dtab.resize(tab.size());
  GH_CHECK_HIP(hipMemcpyAsync(dtab.get(), tab.data(), sizeof(double) * tab.size(),
                              hipMemcpyHostToDevice, c.stream));
  const bool f32 = precision == GH_PREC_FP32_ELEM;
  c.elem.resize((size_t)c.ncells * ns * nt * (f32 ? sizeof(float) : sizeof(double)));
  c.evec.resize((size_t)c.ncells * ns);
  RectArgs A{};
  A.ncells = c.ncells;
  A.ns = ns;
  A.nt = nt;
  A.nns = nns;
  A.ntn = ntn;
  A.div = div;
  A.nq = c.nq;
  A.nv = c.nv;
  A.ttab = dtab.get();
  A.gtab = c.cartesian ? nullptr : c.gtab.get();
  A.cell_nodes = c.cartesian ? nullptr : c.cell_nodes.get();
  A.xyz = c.cartesian ? nullptr : c.xyz.get();
  for (int k = 0; k < 3; ++k) A.h[k] = c.sizes[k];
  A.kappa = c.have_kappa ? c.kappa.get() : nullptr;
  A.sigma = c.have_sigma ? c.sigma.get() : nullptr;
  A.tcell_dofs = c.tcell_dofs.get();
  A.tdvals = c.have_dvals ? c.dvals.get() : nullptr;
  const unsigned eg = grid_for(c.ncells);
  if (D == 2) {
    if (f32) hipLaunchKernelGGL((rect_elem_kernel<float, 2>), dim3(eg), dim3(kBlock), 0, c.stream, A, (float *)c.elem.get(), c.evec.get());
    else hipLaunchKernelGGL((rect_elem_kernel<double, 2>), dim3(eg), dim3(kBlock), 0, c.stream, A, (double *)c.elem.get(), c.evec.get());
  } else {
    if (f32) hipLaunchKernelGGL((rect_elem_kernel<float, 3>), dim3(eg), dim3(kBlock), 0, c.stream, A, (float *)c.elem.get(), c.evec.get());
    else hipLaunchKernelGGL((rect_elem_kernel<double, 3>), dim3(eg), dim3(kBlock), 0, c.stream, A, (double *)c.elem.get(), c.evec.get());
  }
  GH_CHECK_HIP(hipGetLastError());
  if (c.nnz) {
    if (f32)
      hipLaunchKernelGGL(rect_sum_kernel<float>, dim3(grid_for(c.nnz)), dim3(kBlock), 0, c.stream,
                         c.nnz, c.cstart.get(), c.perm.get(), (const float *)c.elem.get(),
                         c.nzval.get());
    else
      hipLaunchKernelGGL(rect_sum_kernel<double>, dim3(grid_for(c.nnz)), dim3(kBlock), 0, c.stream,
                         c.nnz, c.cstart.get(), c.perm.get(), (const double *)c.elem.get(),
                         c.nzval.get());
    GH_CHECK_HIP(hipGetLastError());
  }
  if (c.nfree) {
    hipLaunchKernelGGL(rect_rhs_kernel, dim3(grid_for(c.nfree)), dim3(kBlock), 0, c.stream,
                       c.nfree, ns, c.adjptr.get(), c.adj.get(), c.evec.get(), c.b.get());
    GH_CHECK_HIP(hipGetLastError());
  }
  c.last_kernel = GH_KERNEL_LIST_SUM;
  // the tabulation buffer is freed on return: wait for the kernel that reads it
  GH_CHECK_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace gh
