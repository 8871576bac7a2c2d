// Boundary integrals into the same CSC and right-hand side: the terms of a weak form
// integrated over a BoundaryTriangulation (src/Geometry/BoundaryTriangulations.jl:
// 122-183) of a Cartesian mesh, as in the Neumann and Nitsche drivers of
// test/GridapTests/PoissonTests.jl:101-108 (scalar Q1, and the driver's own order 2 on
// quadrilaterals).
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
// sizes along the facet and n = ±e_axis.  One thread per face; faces are launched colour
// by colour (no two faces of one colour share a free dof) after the cell kernels, so
// the additions are race-free and deterministic.
#include <algorithm>
#include <cmath>

#include "gh_internal.h"

namespace gh {
namespace {

struct FaceArgs {
  const int32_t *faces;      // face indices of this colour
  int64_t nfaces;
  const int32_t *face_cell;  // 0-based cell of each face
  const int32_t *face_side;  // 0-based n-cube facet of each face
  const int32_t *cell_dofs;
  const double *tab;         // per side, per q: w, shp[NL], grad[NL*D]
  int nqf;
  const double *alpha, *beta, *gq, *hq;  // per face / per face and point (or nullptr)
  const double *dvals;
  double h[3];
  const int32_t *colptr, *rowval;
  double *nzval, *b;
};

// NL: the cell's scalar shape functions (4, 8: Q1; 9: Q2 quadrilateral)
template <int D, int NL>
__global__ __launch_bounds__(128) void boundary_face_kernel(const FaceArgs A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= A.nfaces) return;
  const int64_t f = A.faces[t];
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
  int32_t dofs[NL];
#pragma unroll
  for (int a = 0; a < NL; ++a) dofs[a] = A.cell_dofs[cell * NL + a];
  double K[NL][NL], r[NL];
#pragma unroll
  for (int a = 0; a < NL; ++a) {
    r[a] = 0.0;
#pragma unroll
    for (int bb = 0; bb < NL; ++bb) K[a][bb] = 0.0;
  }
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
#pragma unroll
    for (int a = 0; a < NL; ++a)
#pragma unroll
      for (int bb = 0; bb < NL; ++bb)
        K[a][bb] += dS * (al * phi[a] * phi[bb] + be * (phi[a] * dn[bb] + dn[a] * phi[bb]));
  }
  if (A.dvals) {
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      double s = 0.0;
#pragma unroll
      for (int bb = 0; bb < NL; ++bb)
        if (dofs[bb] < 0) s += K[a][bb] * A.dvals[-dofs[bb] - 1];
      r[a] -= s;
    }
  }
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
        const int mid = (lo + hi) >> 1;
        if (A.rowval[mid] < i - 1) lo = mid + 1;
        else hi = mid;
      }
      if (lo < p1 && A.rowval[lo] == i - 1) A.nzval[lo] += K[li][lj];
    }
  }
#pragma unroll
  for (int li = 0; li < NL; ++li)
    if (dofs[li] > 0) A.b[dofs[li] - 1] += r[li];
}

}  // namespace

// Greedy colouring of the boundary faces on shared free dofs (faces ascending inside a
// colour); faces of one cell always differ in colour.
void boundary_coloring(Ctx &c) {
  const int64_t F = c.nbfaces;
  const int nl = c.nloc;
  const std::vector<int32_t> &cd = c.h_cell_dofs;
  std::vector<int32_t> color(F, -1);
  // free dof -> faces already coloured that hold it
  std::vector<std::vector<int32_t>> dof_faces;
  std::vector<int64_t> stamp(8, -1);
  int ncol = 0;
  std::vector<int32_t> slot(c.nfree + 1, -1);  // free dof -> its list in dof_faces
  for (int64_t f = 0; f < F; ++f) {
    const int64_t cell = c.h_face_cell[f];
    for (int a = 0; a < nl; ++a) {
      const int32_t d = cd[cell * nl + a];
      if (d <= 0) continue;
      if (slot[d] < 0) {
        slot[d] = (int32_t)dof_faces.size();
        dof_faces.emplace_back();
      }
      for (int32_t other : dof_faces[slot[d]]) {
        const int col = color[other];
        if ((size_t)col >= stamp.size()) stamp.resize(2 * col + 2, -1);
        stamp[col] = f;
      }
    }
    int chosen = 0;
    while ((size_t)chosen < stamp.size() && stamp[chosen] == f) ++chosen;
    if ((size_t)chosen >= stamp.size()) stamp.resize(2 * chosen + 2, -1);
    color[f] = chosen;
    ncol = std::max(ncol, chosen + 1);
    for (int a = 0; a < nl; ++a) {
      const int32_t d = cd[cell * nl + a];
      if (d > 0) dof_faces[slot[d]].push_back((int32_t)f);
    }
  }
  c.bface_color_ptr.assign(ncol + 1, 0);
  for (int64_t f = 0; f < F; ++f) c.bface_color_ptr[color[f] + 1]++;
  for (int k = 0; k < ncol; ++k) c.bface_color_ptr[k + 1] += c.bface_color_ptr[k];
  std::vector<int32_t> order(std::max<int64_t>(F, 1));
  std::vector<int64_t> cur(c.bface_color_ptr.begin(), c.bface_color_ptr.end() - 1);
  for (int64_t f = 0; f < F; ++f) order[cur[color[f]]++] = (int32_t)f;
  c.bface_order.resize(order.size());
  GH_CHECK_HIP(hipMemcpy(c.bface_order.get(), order.data(), sizeof(int32_t) * order.size(),
                         hipMemcpyHostToDevice));
  c.bface_ncolors = ncol;
}

void assemble_boundary(Ctx &c) {
  if (c.nbfaces == 0) return;
  const bool q1 = c.order == 1 && c.nloc == (1 << c.dim);
  const bool q2quad = c.order == 2 && c.dim == 2 && c.nloc == 9;
  GH_REQUIRE(c.cartesian && c.reffe_kind == GH_REFFE_QK && c.ncomp == 1 && (q1 || q2quad),
             GH_ERR_UNSUPPORTED,
             "boundary terms are provided for scalar Q1, and Q2 on quadrilaterals, on "
             "Cartesian meshes");
  GH_REQUIRE(c.bface_tab_ok, GH_ERR_STATE, "no face tabulation: call gh_face_tabulation");
  if (c.bface_ncolors == 0) boundary_coloring(c);
  FaceArgs A{};
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
  A.colptr = c.colptr.get();
  A.rowval = c.rowval.get();
  A.nzval = c.nzval.get();
  A.b = c.b.get();
  for (int col = 0; col < c.bface_ncolors; ++col) {
    A.faces = c.bface_order.get() + c.bface_color_ptr[col];
    A.nfaces = c.bface_color_ptr[col + 1] - c.bface_color_ptr[col];
    if (A.nfaces == 0) continue;
    const unsigned grid = (unsigned)((A.nfaces + 127) / 128);
    if (q2quad)
      hipLaunchKernelGGL((boundary_face_kernel<2, 9>), dim3(grid), dim3(128), 0, c.stream, A);
    else if (c.dim == 2)
      hipLaunchKernelGGL((boundary_face_kernel<2, 4>), dim3(grid), dim3(128), 0, c.stream, A);
    else
      hipLaunchKernelGGL((boundary_face_kernel<3, 8>), dim3(grid), dim3(128), 0, c.stream, A);
    GH_CHECK_HIP(hipGetLastError());
  }
}

}  // namespace gh
