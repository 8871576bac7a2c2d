// Linear elasticity with vector Q2 on hexahedra: the cell loop of Gridap's
// assemble_matrix_and_vector (src/FESpaces/SparseMatrixAssemblers.jl:214-248) for
//     a(u,v) = ∫ ε(v) : σ(ε(u)),  σ = λ tr(ε) I + 2 μ ε,  λ, μ from E_c and ν,
//     l(v)   = ∫ v · f  (f constant),
// with Gridap's component-blocked local dofs (node + 27 comp).  For isotropic σ the
// 81x81 element matrix is built from the 27x27 gradient products
//     P_ij[a][b] = Σ_q w_q |J_q| g_ai(q) g_bj(q),   g = ∂φ/∂x = J⁻ᵀ ∇̂φ,
//     K[(a,i),(b,j)] = E_c (λ̂ P_ij[a][b] + μ̂ P_ji[a][b] + μ̂ δ_ij Σ_k P_kk[a][b]),
// λ̂ = ν/((1+ν)(1-2ν)), μ̂ = 1/(2(1+ν)).  One wave per cell: the CartesianMap
// Jacobian is diagonal, so g = ĝ/h and w|J| = w·h0·h1·h2 are formed in FP64 from
// the reference tabulation as the operands are loaded; the nine products run on
// the matrix cores (v_mfma_f32_32x32x2f32, FP32) and the blocks are formed and
// stored from the accumulators, or the six distinct ones in FP64 on the vector
// units; the 81x81 matrix is written row-major (a column's
// entries from one cell are a contiguous row), and the sum/rhs passes of the
// element-list path assemble them.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gh_internal.h"

namespace gh {
namespace {

constexpr int kNn = 27;                      // Q2 nodes
constexpr int kNq = 27;                      // degree-4 tensor rule
constexpr int kNloc = 3 * kNn;               // 81

typedef float v16f __attribute__((ext_vector_type(16)));

struct HexArgs {
  int64_t ncells;
  double h[3];
  const double *E;            // per cell (NULL: 1)
  const double *ref;          // ĝ[q][a][d] (27*27*3), w[q] (27), φ̂[q][a] (27*27)
  double lam, mu;             // λ̂, μ̂
  void *E_out;                // ncells x 81 x 81 (T)
  double *mvec;               // ncells x 27: Σ_q w |J| φ_a
};

template <class T>
__global__ __launch_bounds__(64) void hex_elasticity_element_kernel(HexArgs A) {
  const int lane = threadIdx.x;
  const int64_t cell = blockIdx.x;
  const double *rg = A.ref, *rw = A.ref + kNq * kNn * 3, *rs = rw + kNq;
  // CartesianMap: J = diag(h) at every point (src/Geometry/CartesianGrids.jl:259-269),
  // so g = ∂φ/∂x = ĝ / h and w|J| = w h0 h1 h2 (FP64)
  const double det = A.h[0] * A.h[1] * A.h[2];
  const double ih[3] = {1.0 / A.h[0], 1.0 / A.h[1], 1.0 / A.h[2]};
  // m_a = Σ_q w |J| φ̂_a (FP64)
  if (lane < kNn) {
    double m = 0.0;
    for (int q = 0; q < kNq; ++q) m += rw[q] * det * rs[q * kNn + lane];
    A.mvec[cell * kNn + lane] = m;
  }
  const double Ec = A.E ? A.E[cell] : 1.0;
  T *out = reinterpret_cast<T *>(A.E_out) + cell * (int64_t)(kNloc * kNloc);
  if constexpr (sizeof(T) == 4) {
    // All nine products P_ij = Aw_i G_jᵀ on the matrix cores (A[row a][k q] =
    // g_i(q, a) w|J|_q, B[k q][col b] = g_j(q, b); lane l: row/col l % 32,
    // k = 2 s + l / 32), so that P_ji comes out in the same lane layout as P_ij and
    // each block K_ij = λ̂ P_ij + μ̂ P_ji (+ μ̂ Σ_k P_kk on the diagonal) is formed in
    // registers and stored straight from the accumulator layout.
    const int r = lane & 31, kh = lane >> 5;
    constexpr int S = (kNq + 1) / 2;
    float av[S][3], bv[S][3];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int q = 2 * s + kh;
      const bool ok = q < kNq && r < kNn;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double gq = ok ? rg[(q * kNn + r) * 3 + k] * ih[k] : 0.0;
        bv[s][k] = (float)gq;
        av[s][k] = ok ? (float)(gq * rw[q] * det) : 0.0f;
      }
    }
    auto prod = [&](int i, int j) {
      v16f acc = v16f{0};
#pragma unroll
      for (int s = 0; s < S; ++s)
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[s][i], bv[s][j], acc, 0, 0, 0);
      return acc;
    };
    const float lam = (float)(Ec * A.lam), mu = (float)(Ec * A.mu);
    auto store = [&](int i, int j, const v16f &x, const v16f &y, const v16f *tr) {
      if (r >= kNn) return;
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const int a = (t / 4) * 8 + kh * 4 + (t % 4);
        if (a >= kNn) continue;
        float v = lam * x[t] + mu * y[t];
        if (tr) v += mu * (*tr)[t];
        out[(a + kNn * i) * kNloc + r + kNn * j] = v;
      }
    };
    {  // diagonal blocks: λ̂ P_ii + μ̂ P_ii + μ̂ Σ_k P_kk
      const v16f d0 = prod(0, 0), d1 = prod(1, 1), d2 = prod(2, 2);
      v16f tr;
#pragma unroll
      for (int t = 0; t < 16; ++t) tr[t] = d0[t] + d1[t] + d2[t];
      store(0, 0, d0, d0, &tr);
      store(1, 1, d1, d1, &tr);
      store(2, 2, d2, d2, &tr);
    }
#pragma unroll
    for (int pr = 0; pr < 3; ++pr) {  // off-diagonal pairs (i, j), i < j
      const int i = pr == 2 ? 1 : 0, j = pr == 0 ? 1 : 2;
      const v16f x = prod(i, j), y = prod(j, i);
      store(i, j, x, y, nullptr);
      store(j, i, y, x, nullptr);
    }
  } else {
    __shared__ double P[6][kNn][kNn + 1];
    static const int pi[6] = {0, 1, 2, 0, 0, 1}, pj[6] = {0, 1, 2, 1, 2, 2};
    for (int e = lane; e < 6 * kNn * kNn; e += 64) {
      const int pp = e / (kNn * kNn), ab = e % (kNn * kNn), a = ab / kNn, b = ab % kNn;
      const int i = pi[pp], j = pj[pp];
      double s = 0.0;
      for (int q = 0; q < kNq; ++q)
        s += rw[q] * rg[(q * kNn + a) * 3 + i] * rg[(q * kNn + b) * 3 + j];
      P[pp][a][b] = s * det * ih[i] * ih[j];
    }
    __syncthreads();
    const double lam = Ec * A.lam, mu = Ec * A.mu;
    auto pidx = [](int i, int j) {  // slot of the product P_ij with i <= j
      return i == j ? i : (i == 0 ? (j == 1 ? 3 : 4) : 5);
    };
    for (int e = lane; e < kNloc * kNloc; e += 64) {
      const int p = e / kNloc, qd = e - p * kNloc;
      const int a = p % kNn, i = p / kNn, b = qd % kNn, j = qd / kNn;
      double pij, pji;  // P_ji[a][b] = P_ij[b][a]; stored products have i <= j
      if (i <= j) {
        pij = P[pidx(i, j)][a][b];
        pji = P[pidx(i, j)][b][a];
      } else {
        pij = P[pidx(j, i)][b][a];
        pji = P[pidx(j, i)][a][b];
      }
      double v = lam * pij + mu * pji;
      if (i == j) v += mu * (P[0][a][b] + P[1][a][b] + P[2][a][b]);
      out[e] = (T)v;
    }
  }
}

template <class T>
__global__ void hex_rhs_kernel(int64_t nfree, const int32_t *__restrict__ adjptr,
                               const int32_t *__restrict__ adj, const double *__restrict__ mvec,
                               double f0, double f1, double f2,
                               const uint16_t *__restrict__ dflag,
                               const int32_t *__restrict__ cell_dofs,
                               const T *__restrict__ E,
                               const double *__restrict__ dvals, double *__restrict__ b) {
  const int64_t jd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (jd >= nfree) return;
  const double f[3] = {f0, f1, f2};
  double s = 0.0;
  for (int q = adjptr[jd]; q < adjptr[jd + 1]; ++q) {
    const int e = adj[q];
    const int cell = e >> 7, a = e & 127;
    double v = f[a / kNn] * mvec[(size_t)cell * kNn + a % kNn];
    if (dvals && dflag[cell]) {
      for (int bb = 0; bb < kNloc; ++bb) {
        const int d = cell_dofs[(size_t)cell * kNloc + bb];
        if (d < 0) {
          v -= (double)E[((size_t)cell * kNloc + a) * kNloc + bb] * dvals[-d - 1];
        }
      }
    }
    s += v;
  }
  b[jd] = s;
}

}  // namespace

bool hexq2_eligible(const Ctx &c) {
  return c.dim == 3 && c.reffe_kind == GH_REFFE_QK && c.order == 2 && c.ncomp == 3 &&
         c.nloc == kNloc && c.cartesian;
}

void assemble_elasticity(Ctx &c, int precision) {
  GH_REQUIRE(hexq2_eligible(c) && c.elemlist, GH_ERR_UNSUPPORTED,
             "elasticity is provided for vector Q2 on Cartesian hexahedra");
  GH_REQUIRE(c.nq == kNq && (int)c.shp.size() == kNq * kNn &&
                 (int)c.grad.size() == kNq * kNn * 3,
             GH_ERR_STATE, "the elasticity path needs the scalar Q2 tabulation at the 27-point rule");
  if (!c.sref_ready) {
    // ĝ[q][a][d], w[q], φ̂[q][a] of the uploaded (scalar Q2) tabulation
    std::vector<double> ref((size_t)kNq * kNn * 3 + kNq + (size_t)kNq * kNn, 0.0);
    std::copy(c.grad.begin(), c.grad.end(), ref.begin());
    std::copy(c.qw.begin(), c.qw.end(), ref.begin() + kNq * kNn * 3);
    std::copy(c.shp.begin(), c.shp.end(), ref.begin() + kNq * kNn * 3 + kNq);
    c.sref.resize(ref.size());
    GH_CHECK_HIP(hipMemcpy(c.sref.get(), ref.data(), sizeof(double) * ref.size(),
                           hipMemcpyHostToDevice));
    c.sref_ready = true;
  }
  const bool f32 = precision == GH_PREC_FP32_ELEM;
  const size_t esz = f32 ? sizeof(float) : sizeof(double);
  c.elem.resize((size_t)c.ncells * kNloc * kNloc * esz);
  c.fdet.resize((size_t)c.ncells * kNn);
  HexArgs A{};
  A.ncells = c.ncells;
  for (int d = 0; d < 3; ++d) A.h[d] = c.sizes[d];
  A.E = c.have_kappa ? c.kappa.get() : nullptr;
  A.ref = c.sref.get();
  A.lam = c.nu / ((1.0 + c.nu) * (1.0 - 2.0 * c.nu));
  A.mu = 1.0 / (2.0 * (1.0 + c.nu));
  A.E_out = c.elem.get();
  A.mvec = c.fdet.get();
  if (c.ncells) {
    if (f32)
      hipLaunchKernelGGL(hex_elasticity_element_kernel<float>, dim3((unsigned)c.ncells), dim3(64),
                         0, c.stream, A);
    else
      hipLaunchKernelGGL(hex_elasticity_element_kernel<double>, dim3((unsigned)c.ncells),
                         dim3(64), 0, c.stream, A);
  }
  GH_CHECK_HIP(hipGetLastError());
  elemlist_sum(c, f32);
  const unsigned grid = (unsigned)std::max<int64_t>(1, (c.nfree + 255) / 256);
  const double *dv = c.have_dvals ? c.dvals.get() : nullptr;
  if (f32)
    hipLaunchKernelGGL(hex_rhs_kernel<float>, dim3(grid), dim3(256), 0, c.stream, c.nfree,
                       c.adjptr.get(), c.adj.get(), c.fdet.get(), c.body[0], c.body[1],
                       c.body[2], c.dmask.get(), c.cell_dofs.get(),
                       (const float *)c.elem.get(), dv, c.b.get());
  else
    hipLaunchKernelGGL(hex_rhs_kernel<double>, dim3(grid), dim3(256), 0, c.stream, c.nfree,
                       c.adjptr.get(), c.adj.get(), c.fdet.get(), c.body[0], c.body[1],
                       c.body[2], c.dmask.get(), c.cell_dofs.get(),
                       (const double *)c.elem.get(), dv, c.b.get());
  GH_CHECK_HIP(hipGetLastError());
}

}  // namespace gh
