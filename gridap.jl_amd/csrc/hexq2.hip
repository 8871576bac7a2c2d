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
// units; the 81x81 matrix is written column by column (kHexColLd entries per
// column, so a column's entries from one cell are contiguous), and the sum/rhs
// passes of the element-list path assemble them.
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

// FP64 products on the vector units, one wave per cell: the six distinct P_ij
// (i <= j) go to LDS and the 81x81 matrix is written column by column.
__global__ __launch_bounds__(64) void hex_elasticity_fp64_kernel(HexArgs A) {
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
  double *out = reinterpret_cast<double *>(A.E_out) + cell * (int64_t)(kNloc * kHexColLd);
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
  for (int e = lane; e < kNloc * kNloc; e += 64) {  // column qd, row p
    const int qd = e / kNloc, p = e - qd * kNloc;
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
    out[qd * kHexColLd + hex_row_slot(p)] = v;
  }
}

// FP32 products on the matrix cores, persistent: kHexWaves waves per block, each wave
// walks cells blockIdx·kHexWaves + wave, stepping by the grid's wave count.  On the
// Cartesian mesh every cell has the same J = diag(h), so the MFMA operands
//     A[row a][k q] = g_i(q, a) w_q |J|,  B[k q][col b] = g_j(q, b)   (g = ĝ/h, FP64 → FP32)
// are formed once per block into LDS in the lane layout of v_mfma_f32_32x32x2f32
// (lane l: row/col l % 32, k = 2 s + l / 32, zero padding beyond 27), and only E_c
// changes from cell to cell.  All nine products P_ij are run so that P_ji comes out
// in the lane layout of P_ij; each pair of blocks K_ij, K_ji is formed in registers
// and stored straight from the accumulators.
constexpr int kHexWaves = 4;
constexpr int64_t kHexMfmaBlocks = 1024;  // persistent grid: 4 blocks per CU
constexpr int kHexS = (kNq + 1) / 2;  // k-steps of 2 quadrature points

__global__ __launch_bounds__(64 * kHexWaves)
__attribute__((amdgpu_waves_per_eu(2, 2))) void hex_elasticity_mfma_kernel(
    HexArgs A, const double *__restrict__ Ecell, const double *__restrict__ ref,
    float *__restrict__ Eout, double *__restrict__ mvec) {
  __shared__ float opa[kHexS][3][64], opb[kHexS][3][64];
  __shared__ __align__(16) float tiles[kHexWaves][kNn][kHexColLd];
  const double *rg = ref, *rw = ref + kNq * kNn * 3, *rs = rw + kNq;
  const double det = A.h[0] * A.h[1] * A.h[2];
  const double ih[3] = {1.0 / A.h[0], 1.0 / A.h[1], 1.0 / A.h[2]};
  for (int e = threadIdx.x; e < kHexS * 3 * 64; e += 64 * kHexWaves) {
    const int s = e / (3 * 64), k = (e / 64) % 3, l = e % 64;
    const int r = l & 31, q = 2 * s + (l >> 5);
    float av = 0.0f, bv = 0.0f;
    if (q < kNq && r < kNn) {
      const double gq = rg[(q * kNn + r) * 3 + k] * ih[k];
      bv = (float)gq;
      av = (float)(gq * rw[q] * det);
    }
    opa[s][k][l] = av;
    opb[s][k][l] = bv;
  }
  // wave index in a scalar register, so that the cell index and the E_c load are
  // scalar: a vector load would make each cell wait for the previous cell's stores
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // m_a = Σ_q w |J| φ̂_a (FP64), the same for every cell
  double m = 0.0;
  if (lane < kNn)
    for (int q = 0; q < kNq; ++q) m += rw[q] * det * rs[q * kNn + lane];
  __syncthreads();
  const int r = lane & 31, kh = lane >> 5;
  for (int64_t cell = (int64_t)blockIdx.x * kHexWaves + wave; cell < A.ncells;
       cell += (int64_t)gridDim.x * kHexWaves) {
    if (lane < kNn) mvec[cell * kNn + lane] = m;
    const double Ec = Ecell ? Ecell[cell] : 1.0;
    float *out = Eout + cell * (int64_t)(kNloc * kHexColLd);
    auto prod = [&](int i, int j) {
      v16f acc = v16f{0};
#pragma unroll
      for (int s = 0; s < kHexS; ++s)
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(opa[s][i][lane], opb[s][j][lane], acc, 0,
                                                   0, 0);
      return acc;
    };
    const float lam = (float)(Ec * A.lam), mu = (float)(Ec * A.mu);
    // lane (r, kh) holds column b = r and rows a = 8 g + 4 kh + 0..3 of each
    // accumulator.  The three blocks (i, j) of column block j are placed in the wave's
    // LDS tile [27 columns][84 row slots] (slot 28 i + a; slot 27 + 28 i is padding
    // and receives the zero rows of the padded operands), and the finished tile — 27
    // whole columns, 9 KB contiguous in the element matrix — leaves with nine 1 KB
    // float4 stores, so every cache line is written whole by one instruction.
    float4 *tile = reinterpret_cast<float4 *>(&tiles[wave][0][0]);
    auto put = [&](int i, const v16f &x, const v16f &y, const v16f *tr) {
      if (r < kNn) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int a0 = g * 8 + kh * 4;
          if (a0 > kNn) continue;
          float4 v;
          float *vp = reinterpret_cast<float *>(&v);
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int t = 4 * g + u;
            vp[u] = lam * x[t] + mu * y[t];
            if (tr) vp[u] += mu * (*tr)[t];
          }
          tile[r * 21 + 7 * i + a0 / 4] = v;
        }
      }
    };
    auto flush = [&](int j) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      float4 *dst = reinterpret_cast<float4 *>(out + kNn * j * kHexColLd);
#pragma unroll
      for (int pass = 0; pass < 9; ++pass) {
        const int it = pass * 64 + lane;
        if (it < kNn * 21) dst[it] = tile[it];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    // column block j needs λ̂ P_ij + μ̂ P_ji for i != j and (λ̂ + μ̂) P_jj + μ̂ Σ_k P_kk;
    // each product is run once and kept until its last block is placed
    const v16f d0 = prod(0, 0), d1 = prod(1, 1), d2 = prod(2, 2);
    v16f tr;
#pragma unroll
    for (int t = 0; t < 16; ++t) tr[t] = d0[t] + d1[t] + d2[t];
    const v16f p10 = prod(1, 0), p01 = prod(0, 1);
    put(0, d0, d0, &tr);
    put(1, p10, p01, nullptr);
    const v16f p20 = prod(2, 0), p02 = prod(0, 2);
    put(2, p20, p02, nullptr);
    flush(0);
    put(0, p01, p10, nullptr);
    put(1, d1, d1, &tr);
    const v16f p21 = prod(2, 1), p12 = prod(1, 2);
    put(2, p21, p12, nullptr);
    flush(1);
    put(0, p02, p20, nullptr);
    put(1, p12, p21, nullptr);
    put(2, d2, d2, &tr);
    flush(2);
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
          v -= (double)E[((size_t)cell * kNloc + bb) * kHexColLd + hex_row_slot(a)] *
               dvals[-d - 1];
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
  c.elem.resize((size_t)c.ncells * kNloc * kHexColLd * esz);
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
    if (f32) {
      const int64_t want = (c.ncells + kHexWaves - 1) / kHexWaves;
      hipLaunchKernelGGL(hex_elasticity_mfma_kernel,
                         dim3((unsigned)std::min<int64_t>(want, kHexMfmaBlocks)),
                         dim3(64 * kHexWaves), 0, c.stream, A, A.E, A.ref,
                         (float *)A.E_out, A.mvec);
    } else
      hipLaunchKernelGGL(hex_elasticity_fp64_kernel, dim3((unsigned)c.ncells), dim3(64), 0,
                         c.stream, A);
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
