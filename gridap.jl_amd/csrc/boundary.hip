// Boundary integrals into the same CSC and right-hand side: the terms of a weak form
// integrated over a BoundaryTriangulation (src/Geometry/BoundaryTriangulations.jl:
// 122-183) of a Cartesian mesh, as in the Neumann and Nitsche drivers of
// test/GridapTests/PoissonTests.jl:101-108 (scalar Q1, and the driver's own order 2 on
// quadrilaterals; vector spaces component by component, as in the driver's vector case
// and the Neumann loads of test/GridapTests/StokesTaylorHoodTests.jl:56).
//
// Per boundary face F (a facet `side` of cell c, outward unit normal n) with the face
// quadrature (reference weights w_q on the facet, the cell's reference tabulation at
// the facet points):
//   A_ab += Σ_q w_q |J_F| (α φ_a φ_b + β (φ_a n⋅∇φ_b + n⋅∇φ_a φ_b))
//   b_a  += Σ_q w_q |J_F| (φ_a g_q + n⋅∇φ_a h_q)
// with α, β per face and g_q, h_q per face quadrature point (Neumann data, Nitsche
// penalty and consistency data, evaluated by the host), then the Dirichlet lifting
// b_a -= Σ_{b Dirichlet} A_ab u_b of the face matrix (src/CellData/AttachDirichlet.jl:
// 59-84).  On a Cartesian mesh ∇φ = ∇̂φ / h per axis, |J_F| is the product of the cell
// sizes along the facet and n = ±e_axis.
//
// Two kernels after the cell kernels, on lists built once per face upload and pattern:
//  * face kernel, one thread per face: the face matrix (upper triangle; K_ab and K_ba are
//    the same products summed in the same order) and the lifted face load, stored
//    entry-major ([k][face], coalesced stores);
//  * sum kernel, one thread per destination: a nonzero (row i, column j) of two free dofs
//    of a face's cell sums its face contributions (faces ascending) and adds the sum
//    into nzval once; one thread per free dof of a boundary cell adds its face loads
//    into b.  No two threads write one value, so the additions are race-free and
//    deterministic, and the assembled matrix keeps the symmetry of the cell terms
//    bitwise.
// Vector spaces (one component per dimension) use one scalar face matrix for every
// component: the lists pair rows and columns of the same component only.
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "gh_internal.h"

namespace gh {
namespace {

// bdst_pos value of a destination with no entry in the pattern (not met when the pattern
// is the cell pattern of the same space)
constexpr int32_t kNoEntry = INT32_MIN;

struct FaceArgs {
  int64_t nfaces;
  const int32_t *face_cell;  // 0-based cell of each face
  const int32_t *face_side;  // 0-based n-cube facet of each face
  const int32_t *cell_dofs;
  const double *tab;         // per side, per q: w, shp[NL], grad[NL*D]
  int nqf;
  const double *alpha, *beta, *gq, *hq;  // per face / per face and point (or nullptr)
  const double *dvals;
  double h[3];
  double *K, *r;             // out: [t][face] upper triangle, [a][face] load
};

// The matrix and load of every boundary face.
template <int D, int NL>
__global__ __launch_bounds__(128) void boundary_face_kernel(const FaceArgs A) {
  constexpr int NT = NL * (NL + 1) / 2;
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= A.nfaces) return;
  const int64_t cell = A.face_cell[f];
  const int side = A.face_side[f];
  // n-cube facet order (src/ReferenceFEs/ExtrusionPolytopes.jl): facets come in pairs
  // (coordinate 0, coordinate 1) of one normal axis, the axes from the last to the first
  const int axis = D - 1 - side / 2;
  const double sgn = (side & 1) ? 1.0 : -1.0;
  double jf = 1.0;
#pragma unroll
  for (int k = 0; k < D; ++k)
    if (k != axis) jf *= A.h[k];
  const double inv_h = 1.0 / A.h[axis];
  const double al = A.alpha ? A.alpha[f] : 0.0;
  const double be = A.beta ? A.beta[f] : 0.0;
  double K[NT], r[NL];
#pragma unroll
  for (int t = 0; t < NT; ++t) K[t] = 0.0;
#pragma unroll
  for (int a = 0; a < NL; ++a) r[a] = 0.0;
  const int stride = 1 + NL + NL * D;
  for (int q = 0; q < A.nqf; ++q) {
    const double *tq = A.tab + ((size_t)side * A.nqf + q) * stride;
    const double dS = tq[0] * jf;
    const double g = A.gq ? A.gq[f * A.nqf + q] : 0.0;
    const double hh = A.hq ? A.hq[f * A.nqf + q] : 0.0;
    double phi[NL], dn[NL];
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      phi[a] = tq[1 + a];
      dn[a] = sgn * tq[1 + NL + a * D + axis] * inv_h;
      r[a] += dS * (phi[a] * g + dn[a] * hh);
    }
    int t = 0;
#pragma unroll
    for (int a = 0; a < NL; ++a)
#pragma unroll
      for (int b = a; b < NL; ++b, ++t)
        K[t] += dS * (al * (phi[a] * phi[b]) + be * (phi[a] * dn[b] + dn[a] * phi[b]));
  }
  if (A.dvals) {
    // Dirichlet lifting r_a -= Σ_b K_ab u_b over the face's Dirichlet dofs
    // (src/CellData/AttachDirichlet.jl:59-84)
    double ud[NL];
#pragma unroll
    for (int b = 0; b < NL; ++b) {
      const int32_t d = A.cell_dofs[cell * NL + b];
      ud[b] = d < 0 ? A.dvals[-d - 1] : 0.0;
    }
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      double s = 0.0;
#pragma unroll
      for (int b = 0; b < NL; ++b) {
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        s += K[lo * NL - lo * (lo - 1) / 2 + (hi - lo)] * ud[b];
      }
      r[a] -= s;
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) A.K[t * A.nfaces + f] = K[t];
#pragma unroll
  for (int a = 0; a < NL; ++a) A.r[a * A.nfaces + f] = r[a];
}

// Vector spaces (NC components sharing the NL scalar shape functions, local dofs
// node + NL·component), component-wise terms (v⊙u, v⊙(n⋅∇u), (n⋅∇v)⊙u): the scalar face
// matrix (the same for every component) and the loads ∫_F φ_a g_i + ∂nφ_a h_i minus the
// lifting of the component's Dirichlet dofs; g_i, h_i at the face points
// ([(f*nqf + q)*NC + i]).
template <int NL, int NC>
__global__ __launch_bounds__(128) void boundary_face_vec_kernel(const FaceArgs A) {
  constexpr int D = NC;
  constexpr int NT = NL * (NL + 1) / 2;
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= A.nfaces) return;
  const int64_t cell = A.face_cell[f];
  const int side = A.face_side[f];
  const int axis = D - 1 - side / 2;
  const double sgn = (side & 1) ? 1.0 : -1.0;
  double jf = 1.0;
#pragma unroll
  for (int k = 0; k < D; ++k)
    if (k != axis) jf *= A.h[k];
  const double inv_h = 1.0 / A.h[axis];
  const double al = A.alpha ? A.alpha[f] : 0.0;
  const double be = A.beta ? A.beta[f] : 0.0;
  double K[NT], r[NL * NC];
#pragma unroll
  for (int t = 0; t < NT; ++t) K[t] = 0.0;
#pragma unroll
  for (int k = 0; k < NL * NC; ++k) r[k] = 0.0;
  const int stride = 1 + NL + NL * D;
  for (int q = 0; q < A.nqf; ++q) {
    const double *tq = A.tab + ((size_t)side * A.nqf + q) * stride;
    const double dS = tq[0] * jf;
    double phi[NL], dn[NL];
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      phi[a] = tq[1 + a];
      dn[a] = sgn * tq[1 + NL + a * D + axis] * inv_h;
    }
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      const double g = A.gq ? A.gq[(f * A.nqf + q) * NC + i] : 0.0;
      const double hh = A.hq ? A.hq[(f * A.nqf + q) * NC + i] : 0.0;
#pragma unroll
      for (int a = 0; a < NL; ++a) r[a + NL * i] += dS * (phi[a] * g + dn[a] * hh);
    }
    int t = 0;
#pragma unroll
    for (int a = 0; a < NL; ++a)
#pragma unroll
      for (int b = a; b < NL; ++b, ++t)
        K[t] += dS * (al * (phi[a] * phi[b]) + be * (phi[a] * dn[b] + dn[a] * phi[b]));
  }
  if (A.dvals) {
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int a = 0; a < NL; ++a) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < NL; ++b) {
          const int32_t d = A.cell_dofs[cell * NL * NC + b + NL * i];
          const int lo = a < b ? a : b, hi = a < b ? b : a;
          if (d < 0) s += K[lo * NL - lo * (lo - 1) / 2 + (hi - lo)] * A.dvals[-d - 1];
        }
        r[a + NL * i] -= s;
      }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) A.K[t * A.nfaces + f] = K[t];
#pragma unroll
  for (int k = 0; k < NL * NC; ++k) A.r[k * A.nfaces + f] = r[k];
}

struct SumArgs {
  int64_t ndst, nown;
  const int32_t *dst_ptr, *dst_ci, *dst_pos;  // contributions: indices into K
  const int32_t *own_dof, *own_ptr, *own_fa;  // loads: indices into r
  const double *K, *r;
  double *nzval, *b;
};

// Threads [0, ndst): one nonzero each; [ndst, ndst + nown): one load row each.
__global__ __launch_bounds__(256) void boundary_sum_kernel(const SumArgs A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // batches of 4: the index loads, then the value loads, then the ordered sum
  constexpr int U = 4;
  if (t < A.ndst) {
    const int32_t pos = A.dst_pos[t];
    const int32_t e = A.dst_ptr[t + 1];
    double s = 0.0;
    for (int32_t k0 = A.dst_ptr[t]; k0 < e; k0 += U) {
      int32_t id[U];
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) id[u] = k0 + u < e ? A.dst_ci[k0 + u] : -1;
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = id[u] >= 0 ? A.K[id[u]] : 0.0;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (id[u] >= 0) s += v[u];
    }
    if (pos >= 0) A.nzval[pos] += s;
  } else if (t < A.ndst + A.nown) {
    const int64_t o = t - A.ndst;
    const int32_t e = A.own_ptr[o + 1];
    double s = 0.0;
    for (int32_t k0 = A.own_ptr[o]; k0 < e; k0 += U) {
      int32_t id[U];
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) id[u] = k0 + u < e ? A.own_fa[k0 + u] : -1;
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = id[u] >= 0 ? A.r[id[u]] : 0.0;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (id[u] >= 0) s += v[u];
    }
    A.b[A.own_dof[o] - 1] += s;
  }
}

// The nzval position of each destination: a binary search of its row in its column
// (nz_index, src/Algebra/SparseMatrixCSC.jl:14-22), once per pattern.
__global__ __launch_bounds__(256) void boundary_positions_kernel(
    int64_t n, const int32_t *row, const int32_t *col, const int32_t *colptr,
    const int32_t *rowval, int32_t *pos) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int32_t i = row[t] - 1, j = col[t];
  const int32_t p1 = colptr[j];
  int lo = colptr[j - 1], hi = p1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (rowval[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  pos[t] = lo < p1 && rowval[lo] == i ? lo : kNoEntry;
}

template <class T>
void upload(DBuf<T> &d, const std::vector<T> &h) {
  d.resize(std::max<size_t>(h.size(), 1));
  if (!h.empty())
    GH_CHECK_HIP(hipMemcpy(d.get(), h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
}

// Destination and owner lists of the uploaded faces (free dofs ascending; each
// destination's rows ascending inside its column; contributions faces ascending).
void boundary_lists(Ctx &c) {
  const int64_t F = c.nbfaces;
  const int nl = c.nloc;
  const int nn = nl / c.ncomp;        // the face matrix acts on each component alike
  const int nt = nn * (nn + 1) / 2;
  GH_REQUIRE((int64_t)nt * F < INT32_MAX, GH_ERR_UNSUPPORTED,
             "too many boundary faces for the 32-bit face lists");
  const std::vector<int32_t> &cd = c.h_cell_dofs;
  std::vector<int32_t> cnt(c.nfree + 2, 0);
  for (int64_t f = 0; f < F; ++f) {
    const int64_t cell = c.h_face_cell[f];
    for (int a = 0; a < nl; ++a) {
      const int32_t d = cd[cell * nl + a];
      if (d > 0) cnt[d]++;
    }
  }
  // owner entries (face, local index), faces ascending
  std::vector<int32_t> own, optr(1, 0), slot(c.nfree + 1, -1);
  for (int64_t d = 1; d <= c.nfree; ++d)
    if (cnt[d] > 0) {
      slot[d] = optr.back();
      own.push_back((int32_t)d);
      optr.push_back(optr.back() + cnt[d]);
    }
  std::vector<int32_t> ef(optr.back()), ea(optr.back());
  for (int64_t f = 0; f < F; ++f) {
    const int64_t cell = c.h_face_cell[f];
    for (int a = 0; a < nl; ++a) {
      const int32_t d = cd[cell * nl + a];
      if (d > 0) {
        ef[slot[d]] = (int32_t)f;
        ea[slot[d]++] = a;
      }
    }
  }
  std::vector<int32_t> ofa(ef.size());
  for (size_t k = 0; k < ef.size(); ++k) ofa[k] = (int32_t)(ea[k] * F + ef[k]);
  // destinations of each owner column j: rows i of the faces' cells (none when the
  // faces carry loads only)
  const bool matrix = c.have_balpha || c.have_bbeta;
  std::vector<int32_t> dptr(1, 0), dci, drow, dcol;
  std::vector<std::pair<int32_t, int32_t>> tmp;  // (row, contribution)
  for (size_t o = 0; matrix && o < own.size(); ++o) {
    tmp.clear();
    for (int32_t k = optr[o]; k < optr[o + 1]; ++k) {
      const int64_t f = ef[k];
      const int a = ea[k];
      const int64_t cell = c.h_face_cell[f];
      const int an = a % nn, comp = a / nn;
      for (int bn = 0; bn < nn; ++bn) {  // rows of the same component
        const int32_t i = cd[cell * nl + bn + nn * comp];
        if (i <= 0) continue;
        const int lo = std::min(an, bn), hi = std::max(an, bn);
        const int t = lo * nn - lo * (lo - 1) / 2 + (hi - lo);
        tmp.emplace_back(i, (int32_t)(t * F + f));
      }
    }
    // rows ascending; stable, so each row keeps its faces ascending
    std::stable_sort(tmp.begin(), tmp.end(),
                     [](const std::pair<int32_t, int32_t> &x,
                        const std::pair<int32_t, int32_t> &y) { return x.first < y.first; });
    for (size_t k = 0; k < tmp.size(); ++k) {
      if (k == 0 || tmp[k].first != tmp[k - 1].first) {
        if (k > 0) dptr.push_back((int32_t)dci.size());
        drow.push_back(tmp[k].first);
        dcol.push_back(own[o]);
      }
      dci.push_back(tmp[k].second);
    }
    if (!tmp.empty()) dptr.push_back((int32_t)dci.size());
  }
  upload(c.bdst_ptr, dptr);
  upload(c.bdst_ci, dci);
  upload(c.bdst_row, drow);
  upload(c.bdst_col, dcol);
  upload(c.bown_dof, own);
  upload(c.bown_ptr, optr);
  upload(c.bown_fa, ofa);
  const int64_t nd = (int64_t)drow.size();
  c.bdst_pos.resize(std::max<int64_t>(nd, 1));
  if (nd > 0) {
    hipLaunchKernelGGL(boundary_positions_kernel, dim3((unsigned)((nd + 255) / 256)), dim3(256),
                       0, c.stream, nd, c.bdst_row.get(), c.bdst_col.get(), c.colptr.get(),
                       c.rowval.get(), c.bdst_pos.get());
    GH_CHECK_HIP(hipGetLastError());
  }
  c.bface_K.resize(std::max<int64_t>((int64_t)nt * F, 1));  // the face kernels always write it
  c.bface_r.resize(std::max<int64_t>((int64_t)nl * F, 1));
  c.bown_n = (int64_t)own.size();
  c.bdst_n = nd;  // built: kept until the faces, the space or the pattern change
}

}  // namespace

void assemble_boundary(Ctx &c) {
  if (c.nbfaces == 0) return;
  const int nn = c.nloc / c.ncomp;  // scalar shape functions
  const bool q1 = c.order == 1 && nn == (1 << c.dim);
  const bool q2quad = c.order == 2 && c.dim == 2 && nn == 9;
  const bool vec = c.ncomp > 1;
  GH_REQUIRE(c.cartesian && c.reffe_kind == GH_REFFE_QK && (c.ncomp == 1 || c.ncomp == c.dim) &&
                 (q1 || q2quad),
             GH_ERR_UNSUPPORTED,
             "boundary terms are provided for Q1, and Q2 on quadrilaterals, on Cartesian "
             "meshes (vector spaces: component-wise terms)");
  GH_REQUIRE(c.bface_tab_ok, GH_ERR_STATE, "no face tabulation: call gh_face_tabulation");
  if (c.bdst_n < 0) boundary_lists(c);
  FaceArgs A{};
  A.nfaces = c.nbfaces;
  A.face_cell = c.bface_cell.get();
  A.face_side = c.bface_side.get();
  A.cell_dofs = c.cell_dofs.get();
  A.tab = c.bface_tab.get();
  A.nqf = c.nqf;
  A.alpha = c.have_balpha ? c.balpha.get() : nullptr;
  A.beta = c.have_bbeta ? c.bbeta.get() : nullptr;
  A.gq = c.have_bg ? c.bg.get() : nullptr;
  A.hq = c.have_bh ? c.bh.get() : nullptr;
  A.dvals = c.have_dvals ? c.dvals.get() : nullptr;
  for (int k = 0; k < 3; ++k) A.h[k] = c.sizes[k];
  A.K = c.bface_K.get();
  A.r = c.bface_r.get();
  const unsigned fgrid = (unsigned)((c.nbfaces + 127) / 128);
  if (vec && q2quad)
    hipLaunchKernelGGL((boundary_face_vec_kernel<9, 2>), dim3(fgrid), dim3(128), 0, c.stream, A);
  else if (vec && c.dim == 2)
    hipLaunchKernelGGL((boundary_face_vec_kernel<4, 2>), dim3(fgrid), dim3(128), 0, c.stream, A);
  else if (vec)
    hipLaunchKernelGGL((boundary_face_vec_kernel<8, 3>), dim3(fgrid), dim3(128), 0, c.stream, A);
  else if (q2quad)
    hipLaunchKernelGGL((boundary_face_kernel<2, 9>), dim3(fgrid), dim3(128), 0, c.stream, A);
  else if (c.dim == 2)
    hipLaunchKernelGGL((boundary_face_kernel<2, 4>), dim3(fgrid), dim3(128), 0, c.stream, A);
  else
    hipLaunchKernelGGL((boundary_face_kernel<3, 8>), dim3(fgrid), dim3(128), 0, c.stream, A);
  GH_CHECK_HIP(hipGetLastError());
  SumArgs S{};
  S.ndst = c.bdst_n;
  S.nown = c.bown_n;
  S.dst_ptr = c.bdst_ptr.get();
  S.dst_ci = c.bdst_ci.get();
  S.dst_pos = c.bdst_pos.get();
  S.own_dof = c.bown_dof.get();
  S.own_ptr = c.bown_ptr.get();
  S.own_fa = c.bown_fa.get();
  S.K = c.bface_K.get();
  S.r = c.bface_r.get();
  S.nzval = c.nzval.get();
  S.b = c.b.get();
  const int64_t nt = S.ndst + S.nown;
  if (nt > 0) {
    hipLaunchKernelGGL(boundary_sum_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0,
                       c.stream, S);
    GH_CHECK_HIP(hipGetLastError());
  }
}

}  // namespace gh
