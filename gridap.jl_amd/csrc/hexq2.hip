// Linear elasticity with vector Q2 on hexahedra: the cell loop of Gridap's
// assemble_matrix_and_vector (src/FESpaces/SparseMatrixAssemblers.jl:214-248) for
//     a(u,v) = ∫ ε(v) : σ(ε(u)),  σ = λ tr(ε) I + 2 μ ε,  λ, μ from E_c and ν,
//     l(v)   = ∫ v · f  (f constant),
// with Gridap's component-blocked local dofs (node + 27 comp).  For isotropic σ the
// 81x81 element matrix is built from the 27x27 gradient products
//     P_ij[a][b] = Σ_q w_q |J_q| g_ai(q) g_bj(q),   g = ∂φ/∂x = J⁻ᵀ ∇̂φ,
//     K[(a,i),(b,j)] = E_c (λ̂ P_ij[a][b] + μ̂ P_ji[a][b] + μ̂ δ_ij Σ_k P_kk[a][b]),
// λ̂ = ν/((1+ν)(1-2ν)), μ̂ = 1/(2(1+ν)).  Persistent waves, each walking cells: the
// CartesianMap Jacobian is diagonal, so g = ĝ/h and w|J| = w·h0·h1·h2 are formed in
// FP64 from the reference tabulation once per workgroup; the products run on
// the matrix cores (v_mfma_f32_32x32x2f32 in FP32 mode, v_mfma_f64_16x16x4f64 in
// FP64 mode) and the blocks are formed and stored from the accumulators; the 81x81
// matrix is written column by column (kHexColLd entries per column, so a column's
// entries from one cell are contiguous), and the sum/rhs passes of the element-list
// path assemble them.
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

// FP64 products on the matrix cores (v_mfma_f64_16x16x4f64), persistent, kHex64Waves
// waves per workgroup.  Lane layout of the instruction (checked on MI355X): A lane l =
// (row l % 16, k l / 16), B lane l = (column l % 16, k l / 16), result value t of lane
// l = (row 4 t + l / 16, column l % 16).  A 27x27 product is 2x2 tiles of 16x16 over
// 7 k-steps of 4 quadrature points (zero padding beyond 27): 28 instructions.  The
// operands (A = g_i w|J|, B = g_j, formed in FP64 as in the FP32 kernel) sit in LDS
// for the workgroup.  Σ_k P_kk accumulates in one accumulator set; column block j
// needs P_jj again (recomputed) and the pairs P_ij, P_ji, i != j, which are kept in
// registers until both of their blocks are placed.  Blocks go through the wave's LDS
// tile [27 columns][84 row slots] and leave as one contiguous 18 KB column block.
constexpr int kHex64Waves = 2;
constexpr int64_t kHex64Blocks = 512;  // persistent grid: 2 workgroups per CU
constexpr int kHex64S = 7;             // k-steps of 4 quadrature points

typedef double v4d __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(64 * kHex64Waves) void hex_elasticity_mfma64_kernel(
    HexArgs A, const double *__restrict__ Ecell, const double *__restrict__ ref,
    double *__restrict__ Eout, double *__restrict__ mvec) {
  __shared__ double opa[2][kHex64S][3][64], opb[2][kHex64S][3][64];
  __shared__ __align__(16) double tiles[kHex64Waves][kNn][kHexColLd];
  const double *rg = ref, *rw = ref + kNq * kNn * 3, *rs = rw + kNq;
  const double det = A.h[0] * A.h[1] * A.h[2];
  const double ih[3] = {1.0 / A.h[0], 1.0 / A.h[1], 1.0 / A.h[2]};
  for (int e = threadIdx.x; e < 2 * kHex64S * 3 * 64; e += 64 * kHex64Waves) {
    const int t = e / (kHex64S * 3 * 64), s = (e / (3 * 64)) % kHex64S, k = (e / 64) % 3;
    const int l = e % 64;
    const int node = 16 * t + (l & 15), q = 4 * s + (l >> 4);
    double av = 0.0, bv = 0.0;
    if (q < kNq && node < kNn) {
      bv = rg[(q * kNn + node) * 3 + k] * ih[k];
      av = bv * rw[q] * det;
    }
    opa[t][s][k][l] = av;
    opb[t][s][k][l] = bv;
  }
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double m = 0.0;
  if (lane < kNn)
    for (int q = 0; q < kNq; ++q) m += rw[q] * det * rs[q * kNn + lane];
  __syncthreads();
  struct Prod {
    v4d c[2][2];
  };
  auto zero = [](Prod &P) {
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
#pragma unroll
      for (int tb = 0; tb < 2; ++tb) P.c[ta][tb] = v4d{0.0, 0.0, 0.0, 0.0};
  };
  auto accum = [&](Prod &P, int i, int j) {
#pragma unroll
    for (int s = 0; s < kHex64S; ++s)
#pragma unroll
      for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb)
          P.c[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(opa[ta][s][i][lane],
                                                             opb[tb][s][j][lane],
                                                             P.c[ta][tb], 0, 0, 0);
  };
  auto prod = [&](int i, int j) {
    Prod P;
    zero(P);
    accum(P, i, j);
    return P;
  };
  double *tile = &tiles[wave][0][0];
  const int r4 = lane >> 4, cl = lane & 15;
  for (int64_t cell = (int64_t)blockIdx.x * kHex64Waves + wave; cell < A.ncells;
       cell += (int64_t)gridDim.x * kHex64Waves) {
    if (lane < kNn) mvec[cell * kNn + lane] = m;
    const double Ec = Ecell ? Ecell[cell] : 1.0;
    const double lam = Ec * A.lam, mu = Ec * A.mu;
    double *out = Eout + cell * (int64_t)(kNloc * kHexColLd);
    // block (i, j) = λ̂ x + μ̂ y (+ μ̂ tr): row a = 16 ta + 4 t + l/16, column b = 16 tb + l%16
    auto put = [&](int i, const Prod &x, const Prod &y, const Prod *tr) {
#pragma unroll
      for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb)
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int a = 16 * ta + 4 * t + r4, b = 16 * tb + cl;
            if (a > kNn || b >= kNn) continue;  // row 27 is the zero padding row
            double v = lam * x.c[ta][tb][t] + mu * y.c[ta][tb][t];
            if (tr) v += mu * tr->c[ta][tb][t];
            tile[b * kHexColLd + 28 * i + a] = v;
          }
    };
    auto flush = [&](int j) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      double2 *dst = reinterpret_cast<double2 *>(out + kNn * j * kHexColLd);
      const double2 *src = reinterpret_cast<const double2 *>(tile);
      for (int it = lane; it < kNn * kHexColLd / 2; it += 64) dst[it] = src[it];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    Prod tr;
    zero(tr);
    accum(tr, 0, 0);
    accum(tr, 1, 1);
    accum(tr, 2, 2);
    const Prod p10 = prod(1, 0), p01 = prod(0, 1);
    {
      const Prod d = prod(0, 0);
      put(0, d, d, &tr);
    }
    put(1, p10, p01, nullptr);
    const Prod p20 = prod(2, 0), p02 = prod(0, 2);
    put(2, p20, p02, nullptr);
    flush(0);
    put(0, p01, p10, nullptr);
    {
      const Prod d = prod(1, 1);
      put(1, d, d, &tr);
    }
    const Prod p21 = prod(2, 1), p12 = prod(1, 2);
    put(2, p21, p12, nullptr);
    flush(1);
    put(0, p02, p20, nullptr);
    put(1, p12, p21, nullptr);
    {
      const Prod d = prod(2, 2);
      put(2, d, d, &tr);
    }
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

void assemble_elasticity(Ctx &c, int precision, int path) {
  GH_REQUIRE(hexq2_eligible(c) && (c.elemlist || c.q2lattice), GH_ERR_UNSUPPORTED,
             "elasticity is provided for vector Q2 on Cartesian hexahedra");
  GH_REQUIRE(c.nq == kNq && (int)c.shp.size() == kNq * kNn &&
                 (int)c.grad.size() == kNq * kNn * 3,
             GH_ERR_STATE, "the elasticity path needs the scalar Q2 tabulation at the 27-point rule");
  if (c.q2lattice && path != GH_PATH_GENERAL) {
    // owner-computes lattice kernel: no element matrices
    q2lattice_assemble(c, precision == GH_PREC_FP32_ELEM);
    return;
  }
  GH_REQUIRE(path != GH_PATH_STRUCTURED, GH_ERR_UNSUPPORTED,
             "the vector-Q2 lattice path is not available for this space");
  GH_REQUIRE(c.elemlist, GH_ERR_STATE, "no element lists for this space");
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
    {
      const int64_t want = (c.ncells + kHex64Waves - 1) / kHex64Waves;
      hipLaunchKernelGGL(hex_elasticity_mfma64_kernel,
                         dim3((unsigned)std::min<int64_t>(want, kHex64Blocks)),
                         dim3(64 * kHex64Waves), 0, c.stream, A, A.E, A.ref,
                         (double *)A.E_out, A.mvec);
    }
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
