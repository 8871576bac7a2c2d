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
// the linear diffusion operator, and J is symmetric for every p (bitwise: J_ab and J_ba
// are the same products, summed in the same order).
//
// Q1 on quads or hexahedra, Cartesian or isoparametric map, one thread per cell, the
// element Jacobian in FP64 (or FP32 element math), the residual in FP64 always.  Two
// ways into nzval:
//  * lattice meshes (Cartesian topology, nodes numbered lexicographically): the
//    element Jacobians and residuals are stored per cell and the owner-computes column
//    kernel of the linear lattice path sums them (each free node writes its column
//    once, coalesced; incident cells in ascending order);
//  * quadrilateral / hexahedral meshes without the lattice topology: the element
//    Jacobians (full columns) and residuals are stored per cell and the element-list
//    sums of the general n-cube path add them per nonzero and per b row, cells
//    ascending;
//  * on request (GH_PATH_GENERAL), or where neither applies: each thread adds its
//    element into nzval by a binary search of the column; cells are launched colour by
//    colour (no two cells of one colour share a free dof), so the scatter is race-free
//    and deterministic.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "gh_internal.h"

namespace gh {
namespace {

// The upper triangle Kt (row-major: (i, j >= i)) of the element Jacobian and the
// residual r of one Q1 cell with nodes X (mapped) or box sizes h (Cartesian), at the
// cell values u of uh.  tab: per q, w, shp[NL], grad[NL*D].
template <int D, typename T>
__device__ __forceinline__ void plap_element(const double (&X)[1 << D][D], bool mapped,
                                             const double *h, const double (&u)[1 << D],
                                             double kap, double fc, double pm2,
                                             const double *tab, int nq,
                                             T (&Kt)[(1 << D) * ((1 << D) + 1) / 2],
                                             double (&r)[1 << D]) {
  constexpr int NL = 1 << D;
  constexpr int NT = NL * (NL + 1) / 2;
#pragma unroll
  for (int t = 0; t < NT; ++t) Kt[t] = 0;
#pragma unroll
  for (int a = 0; a < NL; ++a) r[a] = 0.0;
  const int stride = 1 + NL + NL * D;
  for (int q = 0; q < nq; ++q) {
    const double *tq = tab + q * stride;
    double J[D * D], Ji[D * D];
    if (mapped) {
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
      for (int k = 0; k < D; ++k) J[k * D + k] = h[k];
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
    int t = 0;
#pragma unroll
    for (int i = 0; i < NL; ++i)
#pragma unroll
      for (int j = i; j < NL; ++j, ++t) {
        T dot = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) dot += (T)g[i][k] * (T)g[j][k];
        Kt[t] += w0 * dot + w1 * ((T)ga[i] * (T)ga[j]);
      }
  }
}

// Index of (i, j) in the row-major upper triangle of an NL x NL matrix.
template <int NL>
__device__ __forceinline__ int tri_index(int i, int j) {
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  return lo * NL - lo * (lo - 1) / 2 + (hi - lo);
}

__device__ __forceinline__ double uh_value(int32_t d, const double *ufree,
                                           const double *udiri) {
  return d > 0 ? ufree[d - 1] : (udiri ? udiri[-d - 1] : 0.0);
}

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
  constexpr int NT = NL * (NL + 1) / 2;
  using T = typename std::conditional<F32, float, double>::type;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= A.ncells) return;
  const int64_t cell = A.cells[t];
  int32_t dofs[NL];
  double u[NL];
#pragma unroll
  for (int a = 0; a < NL; ++a) {
    dofs[a] = A.cell_dofs[cell * NL + a];
    u[a] = uh_value(dofs[a], A.ufree, A.udiri);
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
  T K[NT];
  double r[NL];
  plap_element<D, T>(X, A.cell_nodes != nullptr, A.h, u, kap, fc, A.p - 2.0, A.tab, A.nq, K, r);
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
      if (lo < p1 && A.rowval[lo] == i - 1) A.nzval[lo] += (double)K[tri_index<NL>(li, lj)];
    }
  }
#pragma unroll
  for (int li = 0; li < NL; ++li)
    if (dofs[li] > 0) A.b[dofs[li] - 1] += r[li];
}

struct NlLatArgs {
  int nc[3], nn[3];
  int ncells;
  int nq;
  const int32_t *cell_dofs;
  const double *xyz;  // node coordinates, or nullptr: the Cartesian cells of sizes h
  double h[3];
  const double *tab;
  const double *kappa, *fsrc;
  const double *ufree, *udiri;
  double p;
};

// Lattice meshes: the element Jacobian's upper triangle (E, [t][cell]) and the FP64
// residual (R, [a][cell]) of every cell, for the owner-computes column kernel.
template <int D, bool F32>
__global__ __launch_bounds__(256) void plaplacian_lattice_elem_kernel(
    const NlLatArgs A, typename std::conditional<F32, float, double>::type *E, double *R) {
  constexpr int NL = 1 << D;
  constexpr int NT = NL * (NL + 1) / 2;
  using T = typename std::conditional<F32, float, double>::type;
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= A.ncells) return;
  double X[NL][D], u[NL];
  if (A.xyz) {
    int ci[3] = {0, 0, 0};
    int rem = cell;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      ci[k] = rem % A.nc[k];
      rem /= A.nc[k];
    }
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      int node = 0, stride = 1;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        node += (ci[k] + ((a >> k) & 1)) * stride;
        stride *= A.nn[k];
      }
#pragma unroll
      for (int k = 0; k < D; ++k) X[a][k] = A.xyz[(int64_t)node * D + k];
    }
  }
#pragma unroll
  for (int a = 0; a < NL; ++a)
    u[a] = uh_value(A.cell_dofs[(int64_t)cell * NL + a], A.ufree, A.udiri);
  const double kap = A.kappa ? A.kappa[cell] : 1.0;
  const double fc = A.fsrc ? A.fsrc[cell] : 0.0;
  T K[NT];
  double r[NL];
  plap_element<D, T>(X, A.xyz != nullptr, A.h, u, kap, fc, A.p - 2.0, A.tab, A.nq, K, r);
#pragma unroll
  for (int t = 0; t < NT; ++t) E[(int64_t)t * A.ncells + cell] = K[t];
#pragma unroll
  for (int a = 0; a < NL; ++a) R[(int64_t)a * A.ncells + cell] = r[a];
}

// Cartesian lattice cells: the geometry is the same box for every cell, so the
// physical gradients g_a(q) = ∇̂φ_a(q)/h and their products P_q[ab] = g_a(q)⋅g_b(q)
// come from the table (wave-uniform loads) and only uh's values live in registers:
// J_ab += w0 P_q[ab] + w1 (∇u⋅g_a)(∇u⋅g_b), which keeps the kernel at three waves per
// SIMD instead of one.  Table per q: w|det|, φ[NL], g[NL*D], P[NT].
template <int D, bool F32>
__global__ __launch_bounds__(256) void plaplacian_cart_elem_kernel(
    const NlLatArgs A, typename std::conditional<F32, float, double>::type *E, double *R) {
  constexpr int NL = 1 << D;
  constexpr int NT = NL * (NL + 1) / 2;
  constexpr int stride = 1 + NL + NL * D + NT;
  using T = typename std::conditional<F32, float, double>::type;
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= A.ncells) return;
  double u[NL];
#pragma unroll
  for (int a = 0; a < NL; ++a)
    u[a] = uh_value(A.cell_dofs[(int64_t)cell * NL + a], A.ufree, A.udiri);
  const double kap = A.kappa ? A.kappa[cell] : 1.0;
  const double fc = A.fsrc ? A.fsrc[cell] : 0.0;
  const double pm2 = A.p - 2.0;
  T K[NT];
  double r[NL];
#pragma unroll
  for (int t = 0; t < NT; ++t) K[t] = 0;
#pragma unroll
  for (int a = 0; a < NL; ++a) r[a] = 0.0;
  for (int q = 0; q < A.nq; ++q) {
    const double *tq = A.tab + q * stride;
    const double *gq = tq + 1 + NL;
    const double dV = tq[0];
    double gu[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < NL; ++a) s += u[a] * gq[a * D + k];
      gu[k] = s;
    }
    double n2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) n2 += gu[k] * gu[k];
    const double nrm = sqrt(n2);
    const double s0 = pm2 == 0.0 ? 1.0 : (nrm > 0.0 ? pow(nrm, pm2) : 0.0);
    const double s1 = (pm2 != 0.0 && nrm > 0.0) ? pm2 * s0 / n2 : 0.0;
    double ga[NL];
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) s += gu[k] * gq[a * D + k];
      ga[a] = s;
      r[a] += dV * (kap * s0 * s - fc * tq[1 + a]);
    }
    const T w0 = (T)(kap * dV * s0), w1 = (T)(kap * dV * s1);
    const double *P = gq + NL * D;
    int t = 0;
#pragma unroll
    for (int i = 0; i < NL; ++i)
#pragma unroll
      for (int j = i; j < NL; ++j, ++t) K[t] += w0 * (T)P[t] + w1 * ((T)ga[i] * (T)ga[j]);
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) E[(int64_t)t * A.ncells + cell] = K[t];
#pragma unroll
  for (int a = 0; a < NL; ++a) R[(int64_t)a * A.ncells + cell] = r[a];
}

// Meshes without the lattice topology: the element Jacobian as full columns
// (E[cell][b][a], both halves from one value) and the FP64 residual (R[cell][a]) of
// every cell, for the element-list sums.
template <int D, bool F32>
__global__ __launch_bounds__(256) void plaplacian_list_elem_kernel(
    const NlArgs A, typename std::conditional<F32, float, double>::type *E, double *R) {
  constexpr int NL = 1 << D;
  constexpr int NT = NL * (NL + 1) / 2;
  using T = typename std::conditional<F32, float, double>::type;
  const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= A.ncells) return;
  double X[NL][D], u[NL];
#pragma unroll
  for (int a = 0; a < NL; ++a) {
    u[a] = uh_value(A.cell_dofs[cell * NL + a], A.ufree, A.udiri);
    if (A.cell_nodes) {
      const int64_t nd = A.cell_nodes[cell * NL + a];
#pragma unroll
      for (int k = 0; k < D; ++k) X[a][k] = A.xyz[nd * D + k];
    }
  }
  const double kap = A.kappa ? A.kappa[cell] : 1.0;
  const double fc = A.fsrc ? A.fsrc[cell] : 0.0;
  T K[NT];
  double r[NL];
  plap_element<D, T>(X, A.cell_nodes != nullptr, A.h, u, kap, fc, A.p - 2.0, A.tab, A.nq, K, r);
  T *Ec = E + cell * NL * NL;
  int t = 0;
#pragma unroll
  for (int a = 0; a < NL; ++a)
#pragma unroll
    for (int b = a; b < NL; ++b, ++t) {
      Ec[b * NL + a] = K[t];
      Ec[a * NL + b] = K[t];
    }
#pragma unroll
  for (int a = 0; a < NL; ++a) R[cell * NL + a] = r[a];
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

void assemble_plaplacian(Ctx &c, int precision, int path) {
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
  const bool lattice = c.structured && path != GH_PATH_GENERAL;
  GH_REQUIRE(path != GH_PATH_STRUCTURED || lattice, GH_ERR_UNSUPPORTED,
             "the lattice path needs a Cartesian-topology mesh with a scalar Q1 CLagrangian space");
  const int NL = c.nloc, D = c.dim;
  const int NT = NL * (NL + 1) / 2;
  // Cartesian lattice cells take the table of plaplacian_cart_elem_kernel (per q: w|det|,
  // φ, g = ∇̂φ/h, P_ab = g_a⋅g_b for a <= b), the other kernels per q: w, φ, ∇̂φ
  const bool cart_lattice = lattice && c.cartesian;
  const int stride = cart_lattice ? 1 + NL + NL * D + NT : 1 + NL + NL * D;
  std::vector<double> tab(size_t(c.nq) * stride, 0.0);
  double det = 1.0;
  for (int k = 0; k < D; ++k) det *= c.sizes[k];
  for (int q = 0; q < c.nq; ++q) {
    double *tq = &tab[(size_t)q * stride];
    tq[0] = cart_lattice ? c.qw[q] * std::fabs(det) : c.qw[q];
    for (int a = 0; a < NL; ++a) tq[1 + a] = c.shp[q * NL + a];
    double *g = tq + 1 + NL;
    for (int a = 0; a < NL; ++a)
      for (int k = 0; k < D; ++k)
        g[a * D + k] = c.grad[(q * NL + a) * D + k] / (cart_lattice ? c.sizes[k] : 1.0);
    if (cart_lattice) {
      int t = 0;
      for (int i = 0; i < NL; ++i)
        for (int j = i; j < NL; ++j, ++t) {
          double dot = 0.0;
          for (int k = 0; k < D; ++k) dot += g[i * D + k] * g[j * D + k];
          g[NL * D + t] = dot;
        }
    }
  }
  c.nl_tab.resize(tab.size());
  GH_CHECK_HIP(hipMemcpyAsync(c.nl_tab.get(), tab.data(), sizeof(double) * tab.size(),
                              hipMemcpyHostToDevice, c.stream));
  const bool f32 = precision == GH_PREC_FP32_ELEM;
  if (lattice) {
    GH_REQUIRE(c.cartesian || c.xyz.n == (size_t)c.nnodes * D, GH_ERR_STATE,
               "the mapped lattice needs the node coordinates");
    c.last_kernel = GH_KERNEL_Q1_LATTICE;
    // Jacobian triangles (FP32 or FP64) then FP64 residuals; NT is even, so the
    // residual block starts 8-byte aligned
    const size_t ebytes = (size_t)NT * c.ncells * (f32 ? sizeof(float) : sizeof(double));
    c.qelem.resize(ebytes + (size_t)NL * c.ncells * sizeof(double));
    double *R = reinterpret_cast<double *>(c.qelem.get() + ebytes);
    NlLatArgs A{};
    for (int k = 0; k < 3; ++k) {
      A.nc[k] = (int)c.part[k];
      A.nn[k] = (int)(c.part[k] + 1);
      A.h[k] = c.sizes[k];
    }
    A.ncells = (int)c.ncells;
    A.nq = c.nq;
    A.cell_dofs = c.cell_dofs.get();
    A.xyz = c.cartesian ? nullptr : c.xyz.get();
    A.tab = c.nl_tab.get();
    A.kappa = c.have_kappa ? c.kappa.get() : nullptr;
    A.fsrc = c.have_fsrc ? c.fsrc.get() : nullptr;
    A.ufree = c.ustate.get();
    A.udiri = c.have_dvals ? c.dvals.get() : nullptr;
    A.p = c.plap_p;
    const unsigned eg = (unsigned)((c.ncells + 255) / 256);
    void *E = c.qelem.get();
    if (c.cartesian) {
      if (D == 2) {
        if (f32) hipLaunchKernelGGL((plaplacian_cart_elem_kernel<2, true>), dim3(eg), dim3(256), 0, c.stream, A, (float *)E, R);
        else hipLaunchKernelGGL((plaplacian_cart_elem_kernel<2, false>), dim3(eg), dim3(256), 0, c.stream, A, (double *)E, R);
      } else {
        if (f32) hipLaunchKernelGGL((plaplacian_cart_elem_kernel<3, true>), dim3(eg), dim3(256), 0, c.stream, A, (float *)E, R);
        else hipLaunchKernelGGL((plaplacian_cart_elem_kernel<3, false>), dim3(eg), dim3(256), 0, c.stream, A, (double *)E, R);
      }
    } else if (D == 2) {
      if (f32) hipLaunchKernelGGL((plaplacian_lattice_elem_kernel<2, true>), dim3(eg), dim3(256), 0, c.stream, A, (float *)E, R);
      else hipLaunchKernelGGL((plaplacian_lattice_elem_kernel<2, false>), dim3(eg), dim3(256), 0, c.stream, A, (double *)E, R);
    } else {
      if (f32) hipLaunchKernelGGL((plaplacian_lattice_elem_kernel<3, true>), dim3(eg), dim3(256), 0, c.stream, A, (float *)E, R);
      else hipLaunchKernelGGL((plaplacian_lattice_elem_kernel<3, false>), dim3(eg), dim3(256), 0, c.stream, A, (double *)E, R);
    }
    GH_CHECK_HIP(hipGetLastError());
    // the residual carries no lifting: uh already holds its Dirichlet values
    launch_lattice_columns(c, f32, R, false);
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
    return;
  }
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
  if (path != GH_PATH_GENERAL && c.elemlist && hexgen_eligible(c)) {
    // element-list sums (sets last_kernel)
    c.elem.resize((size_t)c.ncells * NL * NL * (f32 ? sizeof(float) : sizeof(double)));
    c.evec.resize((size_t)c.ncells * NL);
    A.ncells = c.ncells;
    const unsigned eg = (unsigned)((c.ncells + 255) / 256);
    void *E = c.elem.get();
    double *R = c.evec.get();
    if (D == 2) {
      if (f32) hipLaunchKernelGGL((plaplacian_list_elem_kernel<2, true>), dim3(eg), dim3(256), 0, c.stream, A, (float *)E, R);
      else hipLaunchKernelGGL((plaplacian_list_elem_kernel<2, false>), dim3(eg), dim3(256), 0, c.stream, A, (double *)E, R);
    } else {
      if (f32) hipLaunchKernelGGL((plaplacian_list_elem_kernel<3, true>), dim3(eg), dim3(256), 0, c.stream, A, (float *)E, R);
      else hipLaunchKernelGGL((plaplacian_list_elem_kernel<3, false>), dim3(eg), dim3(256), 0, c.stream, A, (double *)E, R);
    }
    GH_CHECK_HIP(hipGetLastError());
    elemlist_sum(c, f32);
    hexgen_rhs(c);
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
    return;
  }
  build_coloring(c);
  c.last_kernel = GH_KERNEL_CELL;
  zero(c, c.nzval.get(), c.nnz);
  zero(c, c.b.get(), c.nfree);
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
