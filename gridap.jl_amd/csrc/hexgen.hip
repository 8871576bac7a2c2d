// Element-list assembly of the n-cube Lagrangian spaces the specialised kernels do not
// take: scalar Q2 diffusion ∫ κ ∇v·∇u + σ u v, f v (test/GridapTests/PoissonTests.jl,
// order 2), scalar Q1 diffusion on meshes without the lattice topology,
// vector Q1 and Q2 isotropic elasticity ∫ ε(v) : σ(ε(u)), b·v
// (test/GridapTests/IsotropicDamageTests.jl:76-85), on Cartesian cells or on the
// isoparametric Q1 map of each cell's vertices (a CartesianDiscreteModel with a map,
// src/Geometry/CartesianGrids.jl:107-124).  The cell integrals follow IntegrationMap
// (src/Fields/FieldsInterfaces.jl:737-760): per quadrature point the Jacobian
// J = Σ_v X_v ⊗ ∇̂N_v of the geometry map, w|det J| and the physical gradients
// J⁻ᵀ∇̂φ_a (push_∇, src/Fields/FieldsInterfaces.jl:94-98), then the integrand summed
// over the points.
//
// Two passes over element matrices stored as full columns:
//  * hexgen_element_kernel: one workgroup per cell; the points' weights and physical
//    gradients go to LDS (FP64), each thread then forms the block of its ordered pair
//    of scalar nodes (a, b) (the products w (g_a g_b) commute, so the blocks of (a, b)
//    and (b, a) are transposes bit for bit), and the cell's load minus the lifting of
//    its Dirichlet dofs (attach_dirichlet, src/CellData/AttachDirichlet.jl:59-84).  Element
//    arithmetic in FP64; the stored matrices in the precision mode's type;
//  * the column-range sum of the element-list path in its full-column form (each range
//    walks its cells' columns with contiguous reads and stages the entries in LDS at
//    their run positions, one thread per nonzero then adds them in ascending cell
//    order), or the contribution-list sum when a column does not fit the 16-bit
//    ranges, for nzval; hexgen_rhs_kernel (one thread per free dof, ascending cells)
//    for b.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gh_internal.h"

namespace gh {
namespace {

constexpr int kHgThreads = 256;
constexpr int kHgMaxNn = 27;   // scalar shape functions (Q2 hexahedron)
constexpr int kHgMaxNq = 64;   // quadrature points

struct HgArgs {
  int D, nn, nc, nloc, nq, nv;
  int cpb;                     // cells per workgroup (kHgThreads / cpb threads each)
  int64_t ncells;
  int elasticity;
  double lam, mu;              // Lamé parameters for E = 1 (scaled by the cell's E)
  double h[3];                 // Cartesian cell sizes (cell_nodes == nullptr)
  const int32_t *cell_nodes;   // 0-based geometry nodes (nv per cell), or nullptr
  const double *xyz;           // node coordinates, D per node
  const double *tab;           // per point: w, φ_a (nn), ∇̂φ_a (nn·D)
  const double *gtab;          // per point: ∇̂N_v (nv·D) of the geometry map
  const double *coef;          // κ (diffusion) or E (elasticity) per cell, or nullptr: 1
  const double *fsrc;          // f per cell (diffusion), or nullptr: 1
  const double *fq;            // f per cell and point (diffusion; replaces fsrc), or nullptr
  const double *sigma;         // mass coefficient σ per cell (diffusion), or nullptr
  double body[3];              // body force (elasticity)
  const int32_t *cell_dofs;
  const double *dvals;         // Dirichlet values, or nullptr
  void *E;                     // per cell: nloc columns of nloc rows
  double *Ev;                  // per cell: nloc load entries
};

template <int D>
__device__ __forceinline__ double det_inv(const double J[3][3], double Ji[3][3]) {
  if constexpr (D == 2) {
    const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double r = 1.0 / det;
    Ji[0][0] = J[1][1] * r;
    Ji[0][1] = -J[0][1] * r;
    Ji[1][0] = -J[1][0] * r;
    Ji[1][1] = J[0][0] * r;
    return det;
  }
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double r = 1.0 / det;
  Ji[0][0] = c00 * r;
  Ji[1][0] = c01 * r;
  Ji[2][0] = c02 * r;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
  return det;
}

// LDS of one cell of the element kernel, in doubles: weights, gradients, dofs.
__host__ __device__ inline size_t hexgen_slot_doubles(int nq, int nn, int D, int nloc) {
  return (size_t)nq + (size_t)nq * nn * D + ((size_t)nloc + 1) / 2;
}

template <class T, int D>
__global__ __launch_bounds__(kHgThreads) void hexgen_element_kernel(HgArgs A) {
  extern __shared__ double sh[];
  const int nn = A.nn, nq = A.nq, nloc = A.nloc;
  // cpb cells per workgroup, tpc threads each (small elements share a workgroup)
  const int tpc = kHgThreads / A.cpb;
  const int slot = threadIdx.x / tpc, tid = threadIdx.x - slot * tpc;
  const int64_t cell = (int64_t)blockIdx.x * A.cpb + slot;
  const bool live = cell < A.ncells;                // no early return: barriers follow
  double *wq = sh + (size_t)slot * hexgen_slot_doubles(nq, nn, D, nloc);  // nq: w |det J|
  double *gp = wq + nq;                             // nq x nn x D: J⁻ᵀ∇̂φ
  int32_t *cd = reinterpret_cast<int32_t *>(gp + (size_t)nq * nn * D);  // nloc dofs
  const int stride = 1 + nn + nn * D;
  for (int a = tid; live && a < nloc; a += tpc) cd[a] = A.cell_dofs[cell * nloc + a];
  for (int q = tid; live && q < nq; q += tpc) {
    const double *tq = A.tab + (size_t)q * stride;
    double J[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, Ji[3][3];
    double det;
    if (A.cell_nodes) {
      for (int d = 0; d < D; ++d)
        for (int e = 0; e < D; ++e) J[d][e] = 0.0;
      const double *gg = A.gtab + (size_t)q * A.nv * D;
      for (int v = 0; v < A.nv; ++v) {
        const double *X = A.xyz + (size_t)A.cell_nodes[cell * A.nv + v] * D;
        for (int d = 0; d < D; ++d)
          for (int e = 0; e < D; ++e) J[d][e] += X[d] * gg[v * D + e];
      }
      det = det_inv<D>(J, Ji);
    } else {
      det = 1.0;
      for (int d = 0; d < 3; ++d)
        for (int e = 0; e < 3; ++e) Ji[d][e] = 0.0;
      for (int d = 0; d < D; ++d) {
        det *= A.h[d];
        Ji[d][d] = 1.0 / A.h[d];
      }
    }
    wq[q] = tq[0] * fabs(det);
    const double *gr = tq + 1 + nn;
    for (int a = 0; a < nn; ++a)
      for (int d = 0; d < D; ++d) {
        double g = 0.0;
        for (int e = 0; e < D; ++e) g += Ji[e][d] * gr[a * D + e];
        gp[((size_t)q * nn + a) * D + d] = g;
      }
  }
  __syncthreads();
  T *E = static_cast<T *>(A.E) + (size_t)cell * nloc * nloc;
  const double cf = live && A.coef ? A.coef[cell] : 1.0;
  // one thread per ordered pair (a, b), consecutive threads on consecutive rows a, so
  // the column stores are contiguous runs; every product is formed as w (g_a g_b),
  // which is the same number for (b, a), so the matrix is symmetric bit for bit
  for (int p = tid; live && p < nn * nn; p += tpc) {
    const int a = p % nn, b = p / nn;
    if (!A.elasticity) {
      double s = 0.0;
      for (int q = 0; q < nq; ++q) {
        const double *ga = gp + ((size_t)q * nn + a) * D, *gb = gp + ((size_t)q * nn + b) * D;
        double dot = 0.0;
        for (int d = 0; d < D; ++d) dot += ga[d] * gb[d];
        s += wq[q] * dot;
      }
      double v = cf * s;
      if (A.sigma) {  // + σ Σ_q w|J| φ_a φ_b (the products commute: symmetric bit for bit)
        double m = 0.0;
        for (int q = 0; q < nq; ++q)
          m += wq[q] * (A.tab[(size_t)q * stride + 1 + a] * A.tab[(size_t)q * stride + 1 + b]);
        v += A.sigma[cell] * m;
      }
      // vector spaces: the component-wise (vector) Laplacian ∇v⊙∇u, δ_ij blocks
      for (int j = 0; j < A.nc; ++j)
        for (int i = 0; i < A.nc; ++i)
          E[(b + nn * j) * nloc + a + nn * i] = (T)(i == j ? v : 0.0);
      continue;
    }
    // K[(a,i),(b,j)] = Σ_q w [λ g_ai g_bj + μ (g_aj g_bi + δ_ij g_a·g_b)]
    //                = λ G_ij + μ (G_ji + δ_ij tr G),  G_ij = Σ_q w (g_ai g_bj)
    double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int q = 0; q < nq; ++q) {
      const double *ga = gp + ((size_t)q * nn + a) * D, *gb = gp + ((size_t)q * nn + b) * D;
      const double w = wq[q];
      double gav[3], gbv[3];
      for (int d = 0; d < D; ++d) {
        gav[d] = ga[d];
        gbv[d] = gb[d];
      }
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) G[i][j] += w * (gav[i] * gbv[j]);
    }
    double tr = 0.0;
    for (int d = 0; d < D; ++d) tr += G[d][d];
    for (int j = 0; j < D; ++j)
      for (int i = 0; i < D; ++i) {
        const double k = A.lam * G[i][j] + A.mu * (G[j][i] + (i == j ? tr : 0.0));
        E[(b + nn * j) * nloc + a + nn * i] = (T)(cf * k);
      }
  }
  __syncthreads();  // the cell's matrix is read below
  // load minus the lifting of the cell's Dirichlet dofs
  const double f = live && A.fsrc ? A.fsrc[cell] : 1.0;
  for (int li = tid; live && li < nloc; li += tpc) {
    const int a = li % nn, i = li / nn;
    double m = 0.0;
    double v;
    if (A.fq) {  // Σ_q w|J| f_q φ_a
      for (int q = 0; q < nq; ++q)
        m += wq[q] * A.fq[(cell * nq + q) * A.nc + i] * A.tab[(size_t)q * stride + 1 + a];
      v = m;
    } else {
      for (int q = 0; q < nq; ++q) m += wq[q] * A.tab[(size_t)q * stride + 1 + a];
      v = (A.nc > 1 ? A.body[i] : f) * m;
    }
    if (A.dvals)
      for (int lk = 0; lk < nloc; ++lk)
        if (cd[lk] < 0) v -= (double)E[lk * nloc + li] * A.dvals[-cd[lk] - 1];
    A.Ev[cell * nloc + li] = v;
  }
}

// Scalar Q1 (NL = 2^D nodes, the FE basis is the cell map's): one thread per cell,
// the upper triangle of the element matrix accumulated in registers point by point
// (FP64), then stored as full columns (both halves from one value) with the load
// minus the lifting.  A workgroup per cell leaves these small cells latency-bound.
template <class T, int D>
__global__ __launch_bounds__(256) void hexgen_q1_kernel(HgArgs A) {
  constexpr int NL = 1 << D;
  constexpr int NT = NL * (NL + 1) / 2;
  const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= A.ncells) return;
  double X[NL][D];
#pragma unroll
  for (int v = 0; v < NL; ++v)
#pragma unroll
    for (int d = 0; d < D; ++d)
      X[v][d] = A.cell_nodes ? A.xyz[(size_t)A.cell_nodes[cell * NL + v] * D + d] : 0.0;
  const double kap = A.coef ? A.coef[cell] : 1.0;
  const double sig = A.sigma ? A.sigma[cell] : 0.0;
  const double fc = A.fsrc ? A.fsrc[cell] : 1.0;
  double K[NT], F[NL];
#pragma unroll
  for (int t = 0; t < NT; ++t) K[t] = 0.0;
#pragma unroll
  for (int a = 0; a < NL; ++a) F[a] = 0.0;
  const int stride = 1 + NL + NL * D;
  for (int q = 0; q < A.nq; ++q) {
    const double *tq = A.tab + (size_t)q * stride;
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Ji[3][3];
    double det;
    if (A.cell_nodes) {
#pragma unroll
      for (int v = 0; v < NL; ++v)
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
          for (int e = 0; e < D; ++e) J[d][e] += X[v][d] * tq[1 + NL + v * D + e];
      det = det_inv<D>(J, Ji);
    } else {
      det = 1.0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        det *= A.h[d];
#pragma unroll
        for (int e = 0; e < D; ++e) Ji[d][e] = d == e ? 1.0 / A.h[d] : 0.0;
      }
    }
    const double w = tq[0] * fabs(det);
    double g[NL][D];
#pragma unroll
    for (int a = 0; a < NL; ++a)
#pragma unroll
      for (int d = 0; d < D; ++d) {
        double s = 0.0;
#pragma unroll
        for (int e = 0; e < D; ++e) s += Ji[e][d] * tq[1 + NL + a * D + e];
        g[a][d] = s;
      }
    const double fq = A.fq ? A.fq[cell * A.nq + q] : fc;
    int t = 0;
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      F[a] += w * fq * tq[1 + a];
#pragma unroll
      for (int b = a; b < NL; ++b, ++t) {
        double dot = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) dot += g[a][d] * g[b][d];
        K[t] += w * (kap * dot + sig * (tq[1 + a] * tq[1 + b]));
      }
    }
  }
  T *E = static_cast<T *>(A.E) + (size_t)cell * NL * NL;
  int t = 0;
#pragma unroll
  for (int a = 0; a < NL; ++a)
#pragma unroll
    for (int b = a; b < NL; ++b, ++t) {
      E[b * NL + a] = (T)K[t];
      E[a * NL + b] = (T)K[t];
    }
  const int32_t *cd = A.cell_dofs + cell * NL;
#pragma unroll
  for (int a = 0; a < NL; ++a) {
    double v = F[a];
    if (A.dvals)
#pragma unroll
      for (int b = 0; b < NL; ++b) {
        const int32_t d = cd[b];
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        if (d < 0) v -= (double)(T)K[lo * NL - lo * (lo - 1) / 2 + (hi - lo)] * A.dvals[-d - 1];
      }
    A.Ev[cell * NL + a] = v;
  }
}

// b[j] = Σ over j's cell-local entries (ascending cells) of the cell's load entry.
__global__ void hexgen_rhs_kernel(int64_t nfree, int nloc, int lbits,
                                  const int32_t *__restrict__ adjptr,
                                  const int32_t *__restrict__ adj,
                                  const double *__restrict__ Ev, double *__restrict__ b) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nfree) return;
  const int lmask = (1 << lbits) - 1;
  double s = 0.0;
  for (int k = adjptr[j]; k < adjptr[j + 1]; ++k) {
    const int e = adj[k];
    s += Ev[(int64_t)(e >> lbits) * nloc + (e & lmask)];
  }
  b[j] = s;
}

}  // namespace

bool hexgen_eligible(const Ctx &c) {
  if (c.reffe_kind != GH_REFFE_QK || (c.dim != 2 && c.dim != 3) || hexq2_eligible(c))
    return false;
  if (!(c.cartesian || c.nv == (1 << c.dim))) return false;
  const bool scalar_q2 = c.ncomp == 1 && c.order == 2;
  const bool vector = c.ncomp == c.dim && (c.order == 1 || c.order == 2);
  // scalar Q1 without the lattice topology (unstructured hexahedra or quadrilaterals):
  // element lists instead of the coloured per-cell scatter
  const bool scalar_q1 = c.ncomp == 1 && c.order == 1 && !c.cartesian && !c.lattice_topo;
  int nn = 1;
  for (int d = 0; d < c.dim; ++d) nn *= c.order + 1;
  return (scalar_q2 || vector || scalar_q1) && c.nloc == nn * c.ncomp && nn <= kHgMaxNn;
}

void hexgen_rhs(Ctx &c) {
  const int lbits = c.nloc <= 16 ? 4 : 7;
  hipLaunchKernelGGL(hexgen_rhs_kernel, dim3((unsigned)std::max<int64_t>(1, (c.nfree + 255) / 256)),
                     dim3(256), 0, c.stream, c.nfree, c.nloc, lbits, c.adjptr.get(), c.adj.get(),
                     c.evec.get(), c.b.get());
  GH_CHECK_HIP(hipGetLastError());
}

void hexgen_assemble(Ctx &c, int precision, bool elasticity) {
  GH_REQUIRE(hexgen_eligible(c) && (elasticity ? c.ncomp == c.dim
                                                : (c.ncomp == 1 || c.ncomp == c.dim)),
             GH_ERR_UNSUPPORTED,
             elasticity ? "elasticity needs a vector space with one component per dimension"
                        : "diffusion needs a scalar space or one component per dimension");
  GH_REQUIRE(c.elemlist, GH_ERR_STATE, "no element lists for this space");
  GH_REQUIRE(!c.have_sigma || (!elasticity && c.sigma.n == (size_t)c.ncells), GH_ERR_STATE,
             "the cell mass term belongs to the diffusion form of this mesh");
  GH_REQUIRE(!c.have_fq || (!elasticity && c.fq.n == (size_t)c.ncells * c.nq * c.ncomp),
             GH_ERR_STATE,
             "the quadrature-point source belongs to another mesh or quadrature");
  const int D = c.dim, nloc = c.nloc, nn = nloc / c.ncomp;
  GH_REQUIRE(c.nq >= 1 && c.nq <= kHgMaxNq && (int)c.qw.size() == c.nq &&
                 (int)c.shp.size() == c.nq * nn && (int)c.grad.size() == c.nq * nn * D,
             GH_ERR_STATE, "the tabulation does not match the element (scalar shape functions)");
  // per point: w, φ, ∇̂φ
  const int stride = 1 + nn + nn * D;
  std::vector<double> tab((size_t)c.nq * stride);
  for (int q = 0; q < c.nq; ++q) {
    tab[(size_t)q * stride] = c.qw[q];
    for (int a = 0; a < nn; ++a) tab[(size_t)q * stride + 1 + a] = c.shp[(size_t)q * nn + a];
    for (int a = 0; a < nn * D; ++a)
      tab[(size_t)q * stride + 1 + nn + a] = c.grad[(size_t)q * nn * D + a];
  }
  DBuf<double> dtab, dgeo;
  dtab.resize(tab.size());
  GH_CHECK_HIP(hipMemcpyAsync(dtab.get(), tab.data(), sizeof(double) * tab.size(),
                              hipMemcpyHostToDevice, c.stream));
  const double *gtab = nullptr;
  if (!c.cartesian) {
    if (nn == c.nv) {  // Q1: the geometry map is the FE basis
      dgeo.resize((size_t)c.nq * nn * D);
      GH_CHECK_HIP(hipMemcpyAsync(dgeo.get(), c.grad.data(), sizeof(double) * dgeo.n,
                                  hipMemcpyHostToDevice, c.stream));
      gtab = dgeo.get();
    } else {
      GH_REQUIRE(c.have_gtab && c.gtab_nq == c.nq, GH_ERR_STATE,
                 "a mapped mesh with a Q2 space needs the geometry tabulation "
                 "(gh_geometry_tabulation) at the points of the FE tabulation");
      gtab = c.gtab.get();
    }
  }
  const bool f32 = precision == GH_PREC_FP32_ELEM;
  const size_t esz = f32 ? sizeof(float) : sizeof(double);
  c.elem.resize((size_t)c.ncells * nloc * nloc * esz);
  c.evec.resize((size_t)c.ncells * nloc);
  HgArgs A{};
  A.D = D;
  A.nn = nn;
  A.nc = c.ncomp;
  A.nloc = nloc;
  A.nq = c.nq;
  A.nv = c.nv;
  A.elasticity = elasticity;
  A.lam = c.nu / ((1.0 + c.nu) * (1.0 - 2.0 * c.nu));
  A.mu = 1.0 / (2.0 * (1.0 + c.nu));
  for (int d = 0; d < 3; ++d) {
    A.h[d] = c.sizes[d];
    A.body[d] = c.body[d];
  }
  A.cell_nodes = c.cartesian ? nullptr : c.cell_nodes.get();
  A.xyz = c.cartesian ? nullptr : c.xyz.get();
  A.tab = dtab.get();
  A.gtab = gtab;
  A.coef = c.have_kappa ? c.kappa.get() : nullptr;
  A.fsrc = c.have_fsrc ? c.fsrc.get() : nullptr;
  A.fq = c.have_fq ? c.fq.get() : nullptr;
  A.sigma = c.have_sigma ? c.sigma.get() : nullptr;
  A.cell_dofs = c.cell_dofs.get();
  A.dvals = c.have_dvals ? c.dvals.get() : nullptr;
  A.E = c.elem.get();
  A.Ev = c.evec.get();
  // cells per workgroup: the most (a power of two, at most 16) whose node pairs still
  // give every thread one
  int cpb = 1;
  while (cpb < 16 && 2 * cpb * nn * nn <= kHgThreads) cpb *= 2;
  A.cpb = cpb;
  A.ncells = c.ncells;
  const size_t shm = sizeof(double) * hexgen_slot_doubles(c.nq, nn, D, nloc) * cpb;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)((c.ncells + cpb - 1) / cpb)), dim3(kHgThreads),
                       shm, c.stream, A);
  };
  auto go1 = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)((c.ncells + 255) / 256)), dim3(256), 0, c.stream, A);
  };
  if (c.order == 1 && c.ncomp == 1) {
    if (D == 2) {
      if (f32) go1(hexgen_q1_kernel<float, 2>);
      else go1(hexgen_q1_kernel<double, 2>);
    } else {
      if (f32) go1(hexgen_q1_kernel<float, 3>);
      else go1(hexgen_q1_kernel<double, 3>);
    }
  } else if (D == 2) {
    if (f32) go(hexgen_element_kernel<float, 2>);
    else go(hexgen_element_kernel<double, 2>);
  } else {
    if (f32) go(hexgen_element_kernel<float, 3>);
    else go(hexgen_element_kernel<double, 3>);
  }
  GH_CHECK_HIP(hipGetLastError());
  elemlist_sum(c, f32);
  hexgen_rhs(c);
  // the tabulation buffers are freed on return: wait for the kernels that read them
  GH_CHECK_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace gh
