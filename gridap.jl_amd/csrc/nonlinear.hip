// Residual and Jacobian of a nonlinear form at a given FE function uh, on the device:
// Gridap's residual!/jacobian!/residual_and_jacobian! of an FEOperatorFromWeakForm
// (src/FESpaces/FEOperatorsFromWeakForm.jl:57-95), whose cell loops are the ones of
// assemble_vector!/assemble_matrix! into a fixed pattern
// (src/FESpaces/SparseMatrixAssemblers.jl:42-68).
//
// Form (the p-Laplacian of test/GridapTests/PLaplacianTests.jl:15-17,31-32, with a
// per-cell factor κ and a per-cell source f):
//   res(u,v)    = ∫ κ ∇v⋅(|∇u|^(p-2) ∇u) - f v
//   jac(u,du,v) = ∫ κ ∇v⋅((p-2)|∇u|^(p-4)(∇u⋅∇du)∇u + |∇u|^(p-2)∇du)
// r_a  = Σ_q w|J| (κ |∇u|^(p-2) ∇φ_a⋅∇u - f φ_a),
// J_ab = Σ_q w|J| κ (|∇u|^(p-2) ∇φ_a⋅∇φ_b + (p-2)|∇u|^(p-4)(∇u⋅∇φ_a)(∇u⋅∇φ_b)),
// with ∇u = Σ_b u_b ∇φ_b and u_b the cell's values of uh (free values, or Dirichlet
// values at Dirichlet dofs).  Rows and columns of Dirichlet dofs are skipped (the
// residual has no lifting: uh already carries its Dirichlet values).  For p = 2 this is
// the linear diffusion operator, and J is symmetric for every p.
//
// One thread per cell (Q1 on quads or hexahedra, Cartesian or isoparametric map), the
// element Jacobian in FP64 (or FP32 element math), added into nzval by a binary search
// of its column; cells are launched colour by colour (no two cells of one colour share
// a free dof), so the scatter is race-free and deterministic.
#include <algorithm>
#include <cmath>

#include "gh_internal.h"

namespace gh {
namespace {

struct NlArgs {
  const int32_t *cells;  // cells of this colour
  int64_t ncells;
  const int32_t *cell_dofs;
  const int32_t *cell_nodes;  // nullptr: Cartesian analytic
  const double *xyz;
  double h[3];
  int nq;
  const double *tab;  // per q: w, shp[NLOC], grad[NLOC*D]
  const double *kappa, *fsrc;
  const double *ufree, *udiri;  // the FE function uh: free and Dirichlet values
  double p;
  const int32_t *colptr, *rowval;
  double *nzval, *b;
};

template <int D, bool F32>
__global__ __launch_bounds__(128) void plaplacian_q1_kernel(const NlArgs A) {
  constexpr int NL = 1 << D;
  using T = typename std::conditional<F32, float, double>::type;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= A.ncells) return;
  const int64_t cell = A.cells[t];
  int32_t dofs[NL];
  double u[NL];
#pragma unroll
  for (int a = 0; a < NL; ++a) {
    const int32_t d = A.cell_dofs[cell * NL + a];
    dofs[a] = d;
    u[a] = d > 0 ? A.ufree[d - 1] : (A.udiri ? A.udiri[-d - 1] : 0.0);
  }
  double X[NL][D];
  if (A.cell_nodes) {
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      const int64_t nd = A.cell_nodes[cell * NL + a];
#pragma unroll
      for (int k = 0; k < D; ++k) X[a][k] = A.xyz[nd * D + k];
    }
  }
  const double kap = A.kappa ? A.kappa[cell] : 1.0;
  const double fc = A.fsrc ? A.fsrc[cell] : 0.0;
  const double pm2 = A.p - 2.0;
  T K[NL][NL];
  double r[NL];
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    r[i] = 0.0;
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
    const double dV = fabs(det_inv<D>(J, Ji)) * tq[0];
    double g[NL][D], gu[D];
#pragma unroll
    for (int k = 0; k < D; ++k) gu[k] = 0.0;
#pragma unroll
    for (int a = 0; a < NL; ++a)
#pragma unroll
      for (int rr = 0; rr < D; ++rr) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) s += Ji[rr * D + k] * tq[1 + NL + a * D + k];
        g[a][rr] = s;
        gu[rr] += u[a] * s;
      }
    double n2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) n2 += gu[k] * gu[k];
    const double nrm = sqrt(n2);
    // |∇u|^(p-2) and (p-2)|∇u|^(p-4) (the latter 0 where ∇u = 0 or p = 2)
    const double s0 = pm2 == 0.0 ? 1.0 : (nrm > 0.0 ? pow(nrm, pm2) : 0.0);
    const double s1 = (pm2 != 0.0 && nrm > 0.0) ? pm2 * s0 / n2 : 0.0;
    double ga[NL];  // ∇u⋅∇φ_a
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) s += gu[k] * g[a][k];
      ga[a] = s;
      r[a] += dV * (kap * s0 * s - fc * tq[1 + a]);
    }
    const T w0 = (T)(kap * dV * s0), w1 = (T)(kap * dV * s1);
#pragma unroll
    for (int i = 0; i < NL; ++i)
#pragma unroll
      for (int j = 0; j < NL; ++j) {
        T dot = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) dot += (T)g[i][k] * (T)g[j][k];
        K[i][j] += w0 * dot + w1 * ((T)ga[i] * (T)ga[j]);
      }
  }
  // scatter: columns lj, rows li, Dirichlet ids skipped
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
      if (lo < p1 && A.rowval[lo] == i - 1) A.nzval[lo] += (double)K[li][lj];
    }
  }
#pragma unroll
  for (int li = 0; li < NL; ++li)
    if (dofs[li] > 0) A.b[dofs[li] - 1] += r[li];
}

__global__ void zero_kernel(int64_t n, double *__restrict__ v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] = 0.0;
}

void zero(Ctx &c, double *v, int64_t n) {
  if (n <= 0) return;
  const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(zero_kernel, dim3(grid), dim3(256), 0, c.stream, n, v);
}

}  // namespace

void assemble_plaplacian(Ctx &c, int precision) {
  GH_REQUIRE(c.reffe_kind == GH_REFFE_QK && c.order == 1 && c.ncomp == 1 &&
                 c.nloc == (1 << c.dim),
             GH_ERR_UNSUPPORTED, "the p-Laplacian form is provided for scalar Q1 on quads/hexes");
  GH_REQUIRE(c.cartesian || c.nv == c.nloc, GH_ERR_UNSUPPORTED,
             "geometry must be Q1 (2^D nodes per cell)");
  GH_REQUIRE((int)c.qw.size() == c.nq && (int)c.shp.size() == c.nq * c.nloc, GH_ERR_STATE,
             "tabulation does not match the element (dofs per cell)");
  GH_REQUIRE(c.have_state && c.ustate.n >= (size_t)std::max<int64_t>(c.nfree, 1),
             GH_ERR_STATE, "no FE function: call gh_fefunction_upload");
  GH_REQUIRE(!c.have_kappa || c.kappa.n == (size_t)c.ncells, GH_ERR_STATE,
             "kappa was uploaded for a different mesh");
  build_coloring(c);
  c.last_kernel = GH_KERNEL_CELL;
  const int NL = c.nloc, D = c.dim;
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
  zero(c, c.nzval.get(), c.nnz);
  zero(c, c.b.get(), c.nfree);
  NlArgs A{};
  A.cell_dofs = c.cell_dofs.get();
  A.cell_nodes = c.cartesian ? nullptr : c.cell_nodes.get();
  A.xyz = c.cartesian ? nullptr : c.xyz.get();
  for (int k = 0; k < 3; ++k) A.h[k] = c.sizes[k];
  A.nq = c.nq;
  A.tab = c.nl_tab.get();
  A.kappa = c.have_kappa ? c.kappa.get() : nullptr;
  A.fsrc = c.have_fsrc ? c.fsrc.get() : nullptr;
  A.ufree = c.ustate.get();
  A.udiri = c.have_dvals ? c.dvals.get() : nullptr;
  A.p = c.plap_p;
  A.colptr = c.colptr.get();
  A.rowval = c.rowval.get();
  A.nzval = c.nzval.get();
  A.b = c.b.get();
  const bool f32 = precision == GH_PREC_FP32_ELEM;
  for (int col = 0; col < c.ncolors; ++col) {
    A.cells = c.color_cells.get() + c.color_ptr[col];
    A.ncells = c.color_ptr[col + 1] - c.color_ptr[col];
    if (A.ncells == 0) continue;
    const unsigned grid = (unsigned)((A.ncells + 127) / 128);
    if (D == 2) {
      if (f32) hipLaunchKernelGGL((plaplacian_q1_kernel<2, true>), dim3(grid), dim3(128), 0, c.stream, A);
      else hipLaunchKernelGGL((plaplacian_q1_kernel<2, false>), dim3(grid), dim3(128), 0, c.stream, A);
    } else {
      if (f32) hipLaunchKernelGGL((plaplacian_q1_kernel<3, true>), dim3(grid), dim3(128), 0, c.stream, A);
      else hipLaunchKernelGGL((plaplacian_q1_kernel<3, false>), dim3(grid), dim3(128), 0, c.stream, A);
    }
    GH_CHECK_HIP(hipGetLastError());
  }
  // the packed tabulation's host copy lives until here
  GH_CHECK_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace gh
