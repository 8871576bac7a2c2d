// Learned factorized sparse approximate inverse (FSAI) for the lattice Q1 path.
//
// Gridap has no iterative solver (src/Algebra/LinearSolvers.jl:249-303 ships LU
// only); this is the preconditioner of the engine's PCG.  G is lower triangular
// on the pattern of tril(A) and M = GᵀG.  With D = diag(A),
//     G = Ĝ D^(-1/2),   Ĝ_ij = Σ_{cells c ∋ i,j} y_c[pair(i,j)]  (i > j),
//     Ĝ_ii = exp(Σ_{cells c ∋ i} y_c[i]) > 0,
// where y_c (36 numbers: 8 diagonal log-terms, 28 strictly-lower pairs in local
// numbering) is the output of a per-cell MLP 10 -> 64 -> 64 -> 64 -> 36 with GELU,
// BF16 weights and FP32 accumulation.  M is SPD by construction; y = 0 gives
// Ĝ = I, i.e. Jacobi.  Inputs per cell: the cell's share κ_c K̂_aa / A_ii of each of
// its 8 nodes' diagonal (0 at Dirichlet nodes), log κ_c and log h.
//
// Training minimises the probe estimate of ‖I - G A Gᵀ‖²_F / N: for Rademacher z,
// r = z - G A Gᵀ z, L = ‖r‖² / (m N), and on the pattern
//     ∂L/∂G_ij = -2 (r_i v_j + z_i q_j) / (m N),  v = A Gᵀ z,  q = A Gᵀ r,
// then ∂L/∂Ĝ_ij = ∂L/∂G_ij D_jj^(-1/2), gathered to the cells and back-propagated
// through the network.  Every kernel here is deterministic (no float atomics).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "gh_internal.h"

namespace gh {
namespace {

constexpr int kBlock = 256;
constexpr int kMlpCellsPerBlock = 256;  // backward / weight-gradient partials
constexpr int kMlpFwdCells = 32;        // forward: small blocks, many waves in flight
constexpr int kEllSlots = kFsaiEllSlots;

inline unsigned grid_for(int64_t n, int block = kBlock) {
  return (unsigned)std::max<int64_t>(1, (n + block - 1) / block);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}

// Round to nearest even (finite inputs).
__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float gelu(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_d(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) +
         x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// splitmix64 of (seed, probe, index): the Rademacher probe sign.
__host__ __device__ __forceinline__ double rademacher(uint64_t seed, uint64_t probe,
                                                      uint64_t i) {
  uint64_t x = seed + 0x9E3779B97F4A7C15ull * (probe * 0x100000001ull + i + 1ull);
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return (x & 1ull) ? 1.0 : -1.0;
}

struct Lat3 {
  int nn[3], nc[3];
  double inv0, inv1;
};

__device__ __forceinline__ void node_xyz(const Lat3 &L, int n, int *xi) {
  xi[0] = n % L.nn[0];
  const int t = n / L.nn[0];
  xi[1] = t % L.nn[1];
  xi[2] = t / L.nn[1];
}

// ---- set-up --------------------------------------------------------------

__global__ void diag_setup_kernel(int64_t n, const int32_t *__restrict__ cp,
                                  const int32_t *__restrict__ rv,
                                  const double *__restrict__ val, int32_t *__restrict__ dpos,
                                  double *__restrict__ fs) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int lo = cp[j], hi = cp[j + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (rv[mid] < j) lo = mid + 1;
    else hi = mid;
  }
  dpos[j] = lo;  // the diagonal is in the pattern of every free column
  const double d = val[lo];
  fs[j] = d > 0.0 ? 1.0 / sqrt(d) : 1.0;
}

__global__ void grf_kappa_kernel(Lat3 L, int64_t ncells, double ox, double oy, double oz,
                                 double hx, double hy, double hz, int nmodes,
                                 const double *__restrict__ omega,
                                 const double *__restrict__ phase, double amp,
                                 double *__restrict__ kappa) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int cx = (int)(c % L.nc[0]);
  const int cy = (int)((c / L.nc[0]) % L.nc[1]);
  const int cz = (int)(c / ((int64_t)L.nc[0] * L.nc[1]));
  const double x = ox + (cx + 0.5) * hx, y = oy + (cy + 0.5) * hy, z = oz + (cz + 0.5) * hz;
  double g = 0.0;
  for (int k = 0; k < nmodes; ++k)
    g += cos(omega[3 * k] * x + omega[3 * k + 1] * y + omega[3 * k + 2] * z + phase[k]);
  kappa[c] = exp(amp * g);
}

// ---- network inputs --------------------------------------------------------

__global__ void features_kernel(Lat3 L, int64_t ncells, const double *__restrict__ kappa,
                                const int32_t *__restrict__ node_dof,
                                const int32_t *__restrict__ dpos,
                                const double *__restrict__ nzval, double kdiag, float logh,
                                float *__restrict__ feat) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int cx = (int)(c % L.nc[0]);
  const int cy = (int)((c / L.nc[0]) % L.nc[1]);
  const int cz = (int)(c / ((int64_t)L.nc[0] * L.nc[1]));
  const double k = kappa ? kappa[c] : 1.0;
  float *f = feat + c * kMlpInPad;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int node = (cx + (a & 1)) + L.nn[0] * ((cy + ((a >> 1) & 1)) + L.nn[1] * (cz + (a >> 2)));
    const int d = node_dof[node];
    f[a] = d > 0 ? (float)(k * kdiag / nzval[dpos[d - 1]]) : 0.0f;
  }
  f[8] = (float)log(k);
  f[9] = logh;
#pragma unroll
  for (int a = 10; a < kMlpInPad; ++a) f[a] = 0.0f;
}

// ---- MLP: lane = output unit, one cell row (uniform) at a time ---------------

template <int NIN, int NINP, int NOUT, int NOUTP, bool GELU, bool SOA = false>
__global__ __launch_bounds__(64) void mlp_fwd(const float *__restrict__ in,
                                              const uint16_t *__restrict__ wq, int woff,
                                              int boff, float *__restrict__ zout,
                                              float *__restrict__ hout, int C, int cpb) {
  const int j = threadIdx.x;
  float w[NIN];
#pragma unroll
  for (int k = 0; k < NIN; ++k) w[k] = j < NOUT ? bf16_to_f32(wq[woff + j * NIN + k]) : 0.0f;
  const float bias = j < NOUT ? bf16_to_f32(wq[boff + j]) : 0.0f;
  const int c0 = blockIdx.x * cpb, c1 = min(c0 + cpb, C);
  for (int c = c0; c < c1; ++c) {
    const float *row = in + (size_t)c * NINP;
    float acc = bias;
#pragma unroll
    for (int k = 0; k < NIN; ++k) acc = fmaf(w[k], row[k], acc);
    if (j < NOUT) {
      zout[SOA ? (size_t)j * C + c : (size_t)c * NOUTP + j] = acc;
      if (GELU) hout[(size_t)c * NOUTP + j] = gelu(acc);
    }
  }
}

// Inference forward of the whole network on the matrix cores
// (v_mfma_f32_16x16x32_bf16): a wave owns 16-cell tiles; layer inputs are rounded
// to BF16 (A operand: lane l holds cell l%16, inputs 8(l/16)..+7), weights are the
// BF16 values (B operand: output unit l%16, inputs 8(l/16)..+7, kept in registers),
// accumulation is FP32 (D: output unit l%16, cells 4(l/16)..+3).  Bias + GELU are
// applied in FP32 and the activations go through LDS to the next layer's A layout.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kFwdWaves = 4;

__device__ __forceinline__ __bf16 to_bf(float f) {
  const uint16_t h = f32_to_bf16(f);
  __bf16 b;
  memcpy(&b, &h, 2);
  return b;
}
__device__ __forceinline__ __bf16 bf_bits(uint16_t h) {
  __bf16 b;
  memcpy(&b, &h, 2);
  return b;
}

__global__ __launch_bounds__(kFwdWaves * 64) void mlp_fwd_mfma(const float *__restrict__ feat,
                                                               const uint16_t *__restrict__ wq,
                                                               int C, float *__restrict__ yo) {
  __shared__ float act[kFwdWaves][16][kMlpHidden + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r16 = lane & 15, kg = lane >> 4;
  const __bf16 zero = to_bf(0.0f);
  v8bf w1[4], w2[4][2], w3[4][2], w4[3][2];
  float b1[4], b2[4], b3[4], b4[3];
  // 8 consecutive BF16 of a 64-wide weight row: one 16-byte load (row offsets are
  // multiples of 8 elements)
  auto row8 = [&](int off) {
    const uint4 u = *reinterpret_cast<const uint4 *>(wq + off);
    v8bf v;
    const uint32_t d[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[2 * i] = bf_bits((uint16_t)(d[i] & 0xffffu));
      v[2 * i + 1] = bf_bits((uint16_t)(d[i] >> 16));
    }
    return v;
  };
  static_assert(kMlpW2 % 8 == 0 && kMlpW3 % 8 == 0 && kMlpW4 % 8 == 0, "aligned rows");
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int n = t * 16 + r16;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = kg * 8 + i;
      w1[t][i] = k < kMlpIn ? bf_bits(wq[kMlpW1 + n * kMlpIn + k]) : zero;
    }
#pragma unroll
    for (int sk = 0; sk < 2; ++sk) {
      const int kk = sk * 32 + kg * 8;
      w2[t][sk] = row8(kMlpW2 + n * kMlpHidden + kk);
      w3[t][sk] = row8(kMlpW3 + n * kMlpHidden + kk);
      if (t < 3) {
        if (n < kMlpOut) {
          w4[t][sk] = row8(kMlpW4 + n * kMlpHidden + kk);
        } else {
#pragma unroll
          for (int i = 0; i < 8; ++i) w4[t][sk][i] = zero;
        }
      }
    }
    b1[t] = bf16_to_f32(wq[kMlpB1 + n]);
    b2[t] = bf16_to_f32(wq[kMlpB2 + n]);
    b3[t] = bf16_to_f32(wq[kMlpB3 + n]);
    if (t < 3) b4[t] = n < kMlpOut ? bf16_to_f32(wq[kMlpB4 + n]) : 0.0f;
  }
  const int ntiles = (C + 15) / 16;
  for (int t0 = blockIdx.x * kFwdWaves; t0 < ntiles; t0 += gridDim.x * kFwdWaves) {
    const int tile = t0 + w;
    const int cell = tile * 16 + r16;
    const bool valid = tile < ntiles && cell < C;
    v8bf a;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = kg * 8 + i;
      a[i] = (valid && k < kMlpInPad) ? to_bf(feat[(size_t)cell * kMlpInPad + k]) : zero;
    }
    v4f acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      acc[t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w1[t], acc[t], 0, 0, 0);
    }
    // hidden layers 2 and 3
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
      const float *bias = layer == 0 ? b1 : b2;
      __syncthreads();
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) act[w][kg * 4 + j][t * 16 + r16] = gelu(acc[t][j] + bias[t]);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int sk = 0; sk < 2; ++sk) {
        v8bf x;
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = to_bf(act[w][r16][sk * 32 + kg * 8 + i]);
#pragma unroll
        for (int t = 0; t < 4; ++t)
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, layer == 0 ? w2[t][sk] : w3[t][sk],
                                                          acc[t], 0, 0, 0);
      }
    }
    // output layer
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) act[w][kg * 4 + j][t * 16 + r16] = gelu(acc[t][j] + b3[t]);
    __syncthreads();
    v4f out[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) out[t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int sk = 0; sk < 2; ++sk) {
      v8bf x;
#pragma unroll
      for (int i = 0; i < 8; ++i) x[i] = to_bf(act[w][r16][sk * 32 + kg * 8 + i]);
#pragma unroll
      for (int t = 0; t < 3; ++t)
        out[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, w4[t][sk], out[t], 0, 0, 0);
    }
    // output-major store yo[n][cell] through LDS: 16 consecutive cells per row
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) act[w][kg * 4 + j][t * 16 + r16] = out[t][j] + b4[t];
    __syncthreads();
    if (tile < ntiles) {
      const int c = tile * 16 + r16;
#pragma unroll
      for (int nb = 0; nb < kMlpOut; nb += 4) {
        const int n = nb + kg;
        if (n < kMlpOut && c < C) yo[(size_t)n * C + c] = act[w][r16][n];
      }
    }
  }
}

// dz_in = (Wᵀ dz_out) ⊙ gelu'(z_in); lane = input unit
template <int NOUT, int NOUTP, int NIN, int NINP>
__global__ __launch_bounds__(64) void mlp_bwd(const float *__restrict__ dzout,
                                              const uint16_t *__restrict__ wq, int woff,
                                              const float *__restrict__ zin,
                                              float *__restrict__ dzin, int C, int cpb) {
  const int k = threadIdx.x;
  float w[NOUT];
#pragma unroll
  for (int j = 0; j < NOUT; ++j) w[j] = k < NIN ? bf16_to_f32(wq[woff + j * NIN + k]) : 0.0f;
  const int c0 = blockIdx.x * cpb, c1 = min(c0 + cpb, C);
  for (int c = c0; c < c1; ++c) {
    const float *row = dzout + (size_t)c * NOUTP;
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < NOUT; ++j) s = fmaf(w[j], row[j], s);
    if (k < NIN) dzin[(size_t)c * NINP + k] = s * gelu_d(zin[(size_t)c * NINP + k]);
  }
}

// Per-block partial of dW = Σ_c dz_c h_cᵀ and db = Σ_c dz_c; lane = output unit.
template <int NIN, int NINP, int NOUT, int NOUTP>
__global__ __launch_bounds__(64) void mlp_wgrad(const float *__restrict__ dz,
                                                const float *__restrict__ hin,
                                                float *__restrict__ part, int C, int cpb) {
  const int j = threadIdx.x;
  float acc[NIN];
#pragma unroll
  for (int k = 0; k < NIN; ++k) acc[k] = 0.0f;
  float accb = 0.0f;
  const int c0 = blockIdx.x * cpb, c1 = min(c0 + cpb, C);
  for (int c = c0; c < c1; ++c) {
    const float d = j < NOUT ? dz[(size_t)c * NOUTP + j] : 0.0f;
    const float *row = hin + (size_t)c * NINP;
#pragma unroll
    for (int k = 0; k < NIN; ++k) acc[k] = fmaf(d, row[k], acc[k]);
    accb += d;
  }
  float *pp = part + (size_t)blockIdx.x * (NOUT * NIN + NOUT);
  if (j < NOUT) {
#pragma unroll
    for (int k = 0; k < NIN; ++k) pp[j * NIN + k] = acc[k];
    pp[NOUT * NIN + j] = accb;
  }
}

__global__ void reduce_parts(const float *__restrict__ part, int nblocks, int len,
                             float *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= len) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; ++b) s += part[(size_t)b * len + t];
  out[t] = (float)s;
}

__global__ void adam_kernel(int n, float *__restrict__ w, const float *__restrict__ g,
                            float *__restrict__ m, float *__restrict__ v, float lr, float b1t,
                            float b2t, uint16_t *__restrict__ wq) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gi = g[t];
  const float mi = 0.9f * m[t] + 0.1f * gi;
  const float vi = 0.999f * v[t] + 0.001f * gi * gi;
  m[t] = mi;
  v[t] = vi;
  const float wi = w[t] - lr * (mi / (1.0f - b1t)) / (sqrtf(vi / (1.0f - b2t)) + 1e-8f);
  w[t] = wi;
  wq[t] = f32_to_bf16(wi);
}

__global__ void to_bf16_kernel(int n, const float *__restrict__ w, uint16_t *__restrict__ wq) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) wq[t] = f32_to_bf16(w[t]);
}

// ---- Ĝ from the cell outputs (owner computes: one thread per lattice node) ----

__device__ __forceinline__ int pair_index(int a, int b) {  // a > b, local numbering
  return 8 + a * (a - 1) / 2 + b;
}

// Ĝ entry of node n and its stencil neighbour m = n + (DX, DY, DZ) (the diagonal for
// 0, 0, 0): Σ over the cells holding both, in ascending cell order, so that both
// endpoints compute the same bits.  Along an axis with offset 0 the cell index is
// x-1 or x, else min(x, x+D); cells outside the lattice are skipped.
template <int DX, int DY, int DZ>
__device__ __forceinline__ float pair_sum(const Lat3 &L, const int *xn,
                                          const float *__restrict__ yo, size_t C) {
  constexpr int NX = DX == 0 ? 2 : 1, NY = DY == 0 ? 2 : 1, NZ = DZ == 0 ? 2 : 1;
  float s = 0.0f;
#pragma unroll
  for (int kz = 0; kz < NZ; ++kz)
#pragma unroll
    for (int ky = 0; ky < NY; ++ky)
#pragma unroll
      for (int kx = 0; kx < NX; ++kx) {
        const int cx = DX == 0 ? xn[0] - 1 + kx : xn[0] + min(DX, 0);
        const int cy = DY == 0 ? xn[1] - 1 + ky : xn[1] + min(DY, 0);
        const int cz = DZ == 0 ? xn[2] - 1 + kz : xn[2] + min(DZ, 0);
        if (cx < 0 || cx >= L.nc[0] || cy < 0 || cy >= L.nc[1] || cz < 0 || cz >= L.nc[2])
          continue;
        const size_t cell = (size_t)cx + (size_t)L.nc[0] * (cy + (size_t)L.nc[1] * cz);
        const int ln = (xn[0] - cx) | ((xn[1] - cy) << 1) | ((xn[2] - cz) << 2);
        const int lm = (xn[0] + DX - cx) | ((xn[1] + DY - cy) << 1) | ((xn[2] + DZ - cz) << 2);
        const int idx = ln == lm ? ln : (ln > lm ? pair_index(ln, lm) : pair_index(lm, ln));
        s += yo[(size_t)idx * C + cell];
      }
  return s;
}

template <int O>
__device__ __forceinline__ float pair_sum_o(const Lat3 &L, const int *xn,
                                            const float *__restrict__ yo, size_t C) {
  return pair_sum<O % 3 - 1, (O / 3) % 3 - 1, O / 9 - 1>(L, xn, yo, C);
}

// One thread per lattice node n (row i of G): the column-i prefix (gl, rows of G)
// and suffix (gu, rows of Gᵀ) of Ĝ in CSC positions, and the solver's copy of
// G = Ĝ D^-1/2 in 16-slot rows: gle[i] = G(i, j ≤ i), gue[i] = G(j ≥ i, i).
template <bool CSC>
__global__ void g_assemble_kernel(Lat3 L, int nnodes, const int32_t *__restrict__ node_dof,
                                  const uint32_t *__restrict__ fmask,
                                  const int32_t *__restrict__ colptr,
                                  const float *__restrict__ yo, const double *__restrict__ fs,
                                  float *__restrict__ gl, float *__restrict__ gu,
                                  float *__restrict__ gle, float *__restrict__ gue,
                                  int64_t nfree, float *__restrict__ g14) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes) return;
  const int d0 = node_dof[n];
  if (d0 <= 0) return;
  int xi[3];
  node_xyz(L, n, xi);
  const uint32_t fm = fmask[n];
  const int p0 = colptr[d0 - 1];
  const double sn = fs[d0 - 1];
  float *el = gle + (size_t)(d0 - 1) * kEllSlots, *eu = gue + (size_t)(d0 - 1) * kEllSlots;
  int kl = 0, ku = 0;
  float sums[27];
  const size_t C = (size_t)L.nc[0] * L.nc[1] * L.nc[2];
#define GH_PAIR(O) sums[O] = ((fm >> (O)) & 1u) ? pair_sum_o<O>(L, xi, yo, C) : 0.0f;
  GH_PAIR(0) GH_PAIR(1) GH_PAIR(2) GH_PAIR(3) GH_PAIR(4) GH_PAIR(5) GH_PAIR(6) GH_PAIR(7)
  GH_PAIR(8) GH_PAIR(9) GH_PAIR(10) GH_PAIR(11) GH_PAIR(12) GH_PAIR(13) GH_PAIR(14)
  GH_PAIR(15) GH_PAIR(16) GH_PAIR(17) GH_PAIR(18) GH_PAIR(19) GH_PAIR(20) GH_PAIR(21)
  GH_PAIR(22) GH_PAIR(23) GH_PAIR(24) GH_PAIR(25) GH_PAIR(26)
#undef GH_PAIR
#pragma unroll
  for (int o = 0; o < 27; ++o) {
    if (!((fm >> o) & 1u)) continue;
    const int pos = p0 + __popc(fm & ((1u << o) - 1u));
    const float sum = sums[o];
    if (o == 13) {
      const float g = expf(sum);
      if (CSC) {
        gl[pos] = g;
        gu[pos] = g;
      }
      const float ge = (float)((double)g * sn);
      el[kl++] = ge;
      eu[ku++] = ge;
      if (g14) g14[13 * nfree + (d0 - 1)] = ge;
    } else if (o < 13) {
      if (CSC) gl[pos] = sum;  // Ĝ(n, m), m < n
      const int m = (xi[0] + o % 3 - 1) + L.nn[0] * ((xi[1] + (o / 3) % 3 - 1) +
                                                     L.nn[1] * (xi[2] + o / 9 - 1));
      el[kl++] = (float)((double)sum * fs[node_dof[m] - 1]);
      if (g14) g14[o * nfree + (d0 - 1)] = el[kl - 1];
    } else {
      if (CSC) gu[pos] = sum;  // Ĝ(m, n), m > n
      eu[ku++] = (float)((double)sum * sn);
    }
  }
  for (; kl < kEllSlots; ++kl) el[kl] = 0.0f;
  for (; ku < kEllSlots; ++ku) eu[ku] = 0.0f;
  if (g14)  // lower offsets whose neighbour is not free (outside the box): zero
    for (int o = 0; o < 13; ++o)
      if (!((fm >> o) & 1u)) g14[o * nfree + (d0 - 1)] = 0.0f;
}

// Column indices of the 16-slot rows (from the pattern; padding points at the row
// itself, with zero weight); flags rows whose halves exceed 16 entries.
__global__ void ell_index_kernel(int64_t n, const int32_t *__restrict__ cp,
                                 const int32_t *__restrict__ rv, const int32_t *__restrict__ dpos,
                                 int32_t *__restrict__ gli, int32_t *__restrict__ gui,
                                 int *__restrict__ overflow) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int lo = cp[i], d = dpos[i], hi = cp[i + 1];
  if (d - lo + 1 > kEllSlots || hi - d > kEllSlots) atomicOr(overflow, 1);
  for (int k = 0; k < kEllSlots; ++k) {
    gli[i * kEllSlots + k] = lo + k <= d ? rv[lo + k] : (int32_t)i;
    gui[i * kEllSlots + k] = d + k < hi ? rv[d + k] : (int32_t)i;
  }
}

// ∂L/∂y_c from the pattern gradient (the transpose of g_assemble).
__global__ void cell_grad_kernel(Lat3 L, int64_t ncells, const int32_t *__restrict__ node_dof,
                                 const uint32_t *__restrict__ fmask,
                                 const int32_t *__restrict__ colptr,
                                 const int32_t *__restrict__ dpos,
                                 const float *__restrict__ gG, const float *__restrict__ gl,
                                 float *__restrict__ dy) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int cx = (int)(c % L.nc[0]);
  const int cy = (int)((c / L.nc[0]) % L.nc[1]);
  const int cz = (int)(c / ((int64_t)L.nc[0] * L.nc[1]));
  int dof[8];
  uint32_t fm[8];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int node = (cx + (a & 1)) + L.nn[0] * ((cy + ((a >> 1) & 1)) + L.nn[1] * (cz + (a >> 2)));
    dof[a] = node_dof[node];
    fm[a] = dof[a] > 0 ? fmask[node] : 0u;
  }
  float *out = dy + c * kMlpOutPad;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    float g = 0.0f;
    if (dof[a] > 0) {
      const int pd = dpos[dof[a] - 1];
      g = gG[pd] * gl[pd];
    }
    out[a] = g;
  }
#pragma unroll
  for (int a = 1; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < a; ++b) {
      float g = 0.0f;
      if (dof[a] > 0 && dof[b] > 0) {
        const int dx = (b & 1) - (a & 1), dyy = ((b >> 1) & 1) - ((a >> 1) & 1),
                  dz = (b >> 2) - (a >> 2);
        const int o = (dx + 1) + 3 * (dyy + 1) + 9 * (dz + 1);
        const int pos = colptr[dof[a] - 1] + __popc(fm[a] & ((1u << o) - 1u));
        g = gG[pos];
      }
      out[pair_index(a, b)] = g;
    }
#pragma unroll
  for (int k = kMlpOut; k < kMlpOutPad; ++k) out[k] = 0.0f;
}

// ---- applications of G and Gᵀ ------------------------------------------------

constexpr int kApplyLanes = 4;

// y_i = Σ_{j <= i} Ĝ_ij s_j r_j  (column-i prefix)
__global__ __launch_bounds__(256) void apply_lower_kernel(int64_t n, const int32_t *__restrict__ cp,
                                                          const int32_t *__restrict__ rv,
                                                          const int32_t *__restrict__ dpos,
                                                          const float *__restrict__ gl,
                                                          const double *__restrict__ fs,
                                                          const double *__restrict__ r,
                                                          double *__restrict__ y) {
  const int sub = threadIdx.x % kApplyLanes;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kApplyLanes;
  double s = 0.0;
  if (row < n) {
    for (int p = cp[row] + sub; p <= dpos[row]; p += kApplyLanes) {
      const int j = rv[p];
      s += (double)gl[p] * fs[j] * r[j];
    }
  }
#pragma unroll
  for (int o = kApplyLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (row < n && sub == 0) y[row] = s;
}

// z_j = s_j Σ_{i >= j} Ĝ_ij y_i  (column-j suffix); optional block partials of z·d
__global__ __launch_bounds__(256) void apply_upper_kernel(int64_t n, const int32_t *__restrict__ cp,
                                                          const int32_t *__restrict__ rv,
                                                          const int32_t *__restrict__ dpos,
                                                          const float *__restrict__ gu,
                                                          const double *__restrict__ fs,
                                                          const double *__restrict__ y,
                                                          double *__restrict__ z,
                                                          const double *__restrict__ d,
                                                          double *__restrict__ part) {
  __shared__ double sh[4];
  const int sub = threadIdx.x % kApplyLanes;
  double dot = 0.0;
  const int64_t rows_per_iter = (int64_t)gridDim.x * (blockDim.x / kApplyLanes);
  for (int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kApplyLanes; row < n;
       row += rows_per_iter) {
    double s = 0.0;
    for (int p = dpos[row] + sub; p < cp[row + 1]; p += kApplyLanes) s += (double)gu[p] * y[rv[p]];
#pragma unroll
    for (int o = kApplyLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (sub == 0) {
      const double zj = fs[row] * s;
      z[row] = zj;
      if (d) dot += zj * d[row];
    }
  }
  if (part) {
    for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = dot;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
  }
}

// ---- training pieces -------------------------------------------------------------

__global__ void rademacher_kernel(int64_t n, uint64_t seed, uint64_t probe,
                                  double *__restrict__ z) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = rademacher(seed, probe, (uint64_t)i);
}

// r = z - u; block partials of r·r added to lpart (one slot per block)
__global__ __launch_bounds__(256) void residual_kernel(int64_t n, const double *__restrict__ z,
                                                       const double *__restrict__ u,
                                                       double *__restrict__ r,
                                                       double *__restrict__ lpart) {
  __shared__ double sh[4];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double ri = z[i] - u[i];
    r[i] = ri;
    acc += ri * ri;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = acc;
  __syncthreads();
  if (threadIdx.x == 0) lpart[blockIdx.x] += sh[0] + sh[1] + sh[2] + sh[3];
}

// gĜ_ij += coef (r_i v_j + z_i q_j) s_j on the lower pattern (column-i prefix)
__global__ void pattern_grad_kernel(int64_t n, const int32_t *__restrict__ cp,
                                    const int32_t *__restrict__ rv,
                                    const int32_t *__restrict__ dpos,
                                    const double *__restrict__ fs, const double *__restrict__ r,
                                    const double *__restrict__ v, const double *__restrict__ z,
                                    const double *__restrict__ q, double coef,
                                    float *__restrict__ gG) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ri = r[i], zi = z[i];
  for (int p = cp[i]; p <= dpos[i]; ++p) {
    const int j = rv[p];
    gG[p] += (float)(coef * (ri * v[j] + zi * q[j]) * fs[j]);
  }
}

Lat3 lattice_of(const Ctx &c) {
  Lat3 L;
  for (int k = 0; k < 3; ++k) {
    L.nn[k] = (int)(c.part[k] + 1);
    L.nc[k] = (int)c.part[k];
  }
  L.inv0 = 1.0 / L.nn[0];
  L.inv1 = 1.0 / L.nn[1];
  return L;
}

constexpr int kResidualBlocks = 512;

}  // namespace

void fsai_prepare(Ctx &c) {
  GH_REQUIRE(c.structured && c.cartesian && c.dim == 3, GH_ERR_UNSUPPORTED,
             "the learned FSAI needs a 3-D Cartesian mesh with a scalar Q1 CLagrangian space");
  GH_REQUIRE(c.have_values, GH_ERR_STATE, "assemble the matrix before the FSAI set-up");
  const int64_t n = c.nfree, C = c.ncells;
  if (!c.fsai_ready) {
    c.diagpos.resize(std::max<int64_t>(n, 1));
    c.fs.resize(std::max<int64_t>(n, 1));
    c.gl.resize(std::max<int64_t>(c.nnz, 1));
    c.gu.resize(std::max<int64_t>(c.nnz, 1));
    c.gG.resize(std::max<int64_t>(c.nnz, 1));
    c.feat.resize(C * kMlpInPad);
    for (auto *b : {&c.z1, &c.h1, &c.z2, &c.h2, &c.z3, &c.h3, &c.dz1, &c.dz2, &c.dz3})
      b->resize(C * kMlpHidden);
    c.yo.resize(C * kMlpOut);  // output-major: yo[n * C + cell]
    c.dy.resize(C * kMlpOutPad);
    const int nb = (int)((C + kMlpCellsPerBlock - 1) / kMlpCellsPerBlock);
    c.wpart.resize((size_t)nb * (kMlpHidden * kMlpHidden + kMlpHidden));
    for (auto *b : {&c.pz, &c.pw, &c.pv, &c.pu, &c.pr, &c.pt, &c.pq})
      b->resize(std::max<int64_t>(n, 1));
    c.lpart.resize(kResidualBlocks);
    for (auto *b : {&c.gle, &c.gue}) b->resize((size_t)std::max<int64_t>(n, 1) * kEllSlots);
    for (auto *b : {&c.gli, &c.gui}) b->resize((size_t)std::max<int64_t>(n, 1) * kEllSlots);
  }
  hipLaunchKernelGGL(diag_setup_kernel, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n,
                     c.colptr.get(), c.rowval.get(), c.nzval.get(), c.diagpos.get(),
                     c.fs.get());
  GH_CHECK_HIP(hipGetLastError());
  if (!c.fsai_ready) {
    DBuf<int> ovf;
    ovf.resize(1);
    GH_CHECK_HIP(hipMemsetAsync(ovf.get(), 0, sizeof(int), c.stream));
    hipLaunchKernelGGL(ell_index_kernel, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n,
                       c.colptr.get(), c.rowval.get(), c.diagpos.get(), c.gli.get(), c.gui.get(),
                       ovf.get());
    GH_CHECK_HIP(hipGetLastError());
    int h = 0;
    GH_CHECK_HIP(hipMemcpyAsync(&h, ovf.get(), sizeof(int), hipMemcpyDeviceToHost, c.stream));
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
    GH_REQUIRE(h == 0, GH_ERR_UNSUPPORTED, "FSAI rows longer than the 16-slot layout");
    c.fsai_ready = true;
  }
}

void mlp_sync_bf16(Ctx &c) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for(kMlpParams)), dim3(kBlock), 0, c.stream,
                     kMlpParams, c.params.get(), c.params_bf16.get());
  GH_CHECK_HIP(hipGetLastError());
}

void fsai_forward_build(Ctx &c, bool inference) {
  GH_REQUIRE(c.have_mlp, GH_ERR_STATE, "no network weights: call gh_mlp_init or gh_mlp_set");
  fsai_prepare(c);
  const Lat3 L = lattice_of(c);
  const int C = (int)c.ncells;
  // K̂_aa of the box cell (every a), from the uploaded tabulation
  double kdiag = 0.0;
  {
    double det = c.sizes[0] * c.sizes[1] * c.sizes[2];
    for (int q = 0; q < c.nq; ++q) {
      double dot = 0.0;
      for (int d = 0; d < 3; ++d) {
        const double g = c.grad[(q * 8 + 0) * 3 + d] / c.sizes[d];
        dot += g * g;
      }
      kdiag += dot * std::fabs(det) * c.qw[q];
    }
  }
  const float logh = (float)(std::log(c.sizes[0] * c.sizes[1] * c.sizes[2]) / 3.0);
  hipStream_t s = c.stream;
  hipLaunchKernelGGL(features_kernel, dim3(grid_for(C)), dim3(kBlock), 0, s, L, (int64_t)C,
                     c.have_kappa ? c.kappa.get() : nullptr, c.node_dof.get(), c.diagpos.get(),
                     c.nzval.get(), kdiag, logh, c.feat.get());
  if (inference) {
    const int ntiles = (C + 15) / 16;
    const unsigned grid =
        (unsigned)std::max(1, std::min((ntiles + kFwdWaves - 1) / kFwdWaves, 512));
    hipLaunchKernelGGL(mlp_fwd_mfma, dim3(grid), dim3(kFwdWaves * 64), 0, s, c.feat.get(),
                       c.params_bf16.get(), C, c.yo.get());
  } else {
    const int nb = (C + kMlpFwdCells - 1) / kMlpFwdCells;
    const uint16_t *wq = c.params_bf16.get();
    hipLaunchKernelGGL((mlp_fwd<kMlpIn, kMlpInPad, kMlpHidden, kMlpHidden, true>), dim3(nb),
                       dim3(64), 0, s, c.feat.get(), wq, kMlpW1, kMlpB1, c.z1.get(), c.h1.get(), C,
                       kMlpFwdCells);
    hipLaunchKernelGGL((mlp_fwd<kMlpHidden, kMlpHidden, kMlpHidden, kMlpHidden, true>), dim3(nb),
                       dim3(64), 0, s, c.h1.get(), wq, kMlpW2, kMlpB2, c.z2.get(), c.h2.get(), C,
                       kMlpFwdCells);
    hipLaunchKernelGGL((mlp_fwd<kMlpHidden, kMlpHidden, kMlpHidden, kMlpHidden, true>), dim3(nb),
                       dim3(64), 0, s, c.h2.get(), wq, kMlpW3, kMlpB3, c.z3.get(), c.h3.get(), C,
                       kMlpFwdCells);
    hipLaunchKernelGGL((mlp_fwd<kMlpHidden, kMlpHidden, kMlpOut, kMlpOutPad, false, true>), dim3(nb),
                       dim3(64), 0, s, c.h3.get(), wq, kMlpW4, kMlpB4, c.yo.get(),
                       (float *)nullptr, C, kMlpFwdCells);
  }
  // box lattice: also the 14 lower-offset arrays of G for the stencil PCG kernels
  float *g14 = nullptr;
  if (c.box) {
    c.g14.resize((size_t)14 * std::max<int64_t>(c.nfree, 1));
    g14 = c.g14.get();
  }
  if (inference)
    hipLaunchKernelGGL(g_assemble_kernel<false>, dim3(grid_for(c.nnodes)), dim3(kBlock), 0, s, L,
                       (int)c.nnodes, c.node_dof.get(), c.node_fmask.get(), c.colptr.get(),
                       c.yo.get(), c.fs.get(), c.gl.get(), c.gu.get(), c.gle.get(), c.gue.get(),
                       c.nfree, g14);
  else
    hipLaunchKernelGGL(g_assemble_kernel<true>, dim3(grid_for(c.nnodes)), dim3(kBlock), 0, s, L,
                       (int)c.nnodes, c.node_dof.get(), c.node_fmask.get(), c.colptr.get(),
                       c.yo.get(), c.fs.get(), c.gl.get(), c.gu.get(), c.gle.get(), c.gue.get(),
                       c.nfree, g14);
  GH_CHECK_HIP(hipGetLastError());
  c.fsai_built = true;
}

void fsai_apply_lower(Ctx &c, const double *r, double *y) {
  const int64_t n = c.nfree;
  const unsigned grid = (unsigned)((n * kApplyLanes + 255) / 256);
  hipLaunchKernelGGL(apply_lower_kernel, dim3(std::max(grid, 1u)), dim3(256), 0, c.stream, n,
                     c.colptr.get(), c.rowval.get(), c.diagpos.get(), c.gl.get(), c.fs.get(),
                     r, y);
  GH_CHECK_HIP(hipGetLastError());
}

// grid of the upper apply: `parts` blocks when partials are requested (the PCG
// reduction expects exactly that many), else one row group per thread group.
void fsai_apply_upper_grid(Ctx &c, const double *y, double *z, const double *d, double *part,
                           int parts) {
  const int64_t n = c.nfree;
  unsigned grid = part ? (unsigned)parts : (unsigned)((n * kApplyLanes + 255) / 256);
  hipLaunchKernelGGL(apply_upper_kernel, dim3(std::max(grid, 1u)), dim3(256), 0, c.stream, n,
                     c.colptr.get(), c.rowval.get(), c.diagpos.get(), c.gu.get(), c.fs.get(),
                     y, z, d, part);
  GH_CHECK_HIP(hipGetLastError());
}

double fsai_train_step(Ctx &c, int nprobes, uint64_t seed, float lr) {
  GH_REQUIRE(nprobes >= 1 && nprobes <= 1024, GH_ERR_ARG, "nprobes must be in 1..1024");
  fsai_forward_build(c, false);
  hipStream_t s = c.stream;
  const int64_t n = c.nfree;
  const int C = (int)c.ncells;
  const Lat3 L = lattice_of(c);
  GH_CHECK_HIP(hipMemsetAsync(c.gG.get(), 0, sizeof(float) * std::max<int64_t>(c.nnz, 1), s));
  GH_CHECK_HIP(hipMemsetAsync(c.lpart.get(), 0, sizeof(double) * kResidualBlocks, s));
  const double coef = -2.0 / ((double)nprobes * (double)std::max<int64_t>(n, 1));
  for (int k = 0; k < nprobes; ++k) {
    hipLaunchKernelGGL(rademacher_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, n, seed,
                       (uint64_t)k, c.pz.get());
    fsai_apply_upper_grid(c, c.pz.get(), c.pw.get(), nullptr, nullptr, 0);   // w = Gᵀ z
    spmv_device(c, c.pw.get(), c.pv.get());                                  // v = A w
    fsai_apply_lower(c, c.pv.get(), c.pu.get());                             // u = G v
    hipLaunchKernelGGL(residual_kernel, dim3(kResidualBlocks), dim3(256), 0, s, n, c.pz.get(),
                       c.pu.get(), c.pr.get(), c.lpart.get());               // r = z - u
    fsai_apply_upper_grid(c, c.pr.get(), c.pt.get(), nullptr, nullptr, 0);   // t = Gᵀ r
    spmv_device(c, c.pt.get(), c.pq.get());                                  // q = A t
    hipLaunchKernelGGL(pattern_grad_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, n,
                       c.colptr.get(), c.rowval.get(), c.diagpos.get(), c.fs.get(), c.pr.get(),
                       c.pv.get(), c.pz.get(), c.pq.get(), coef, c.gG.get());
  }
  hipLaunchKernelGGL(cell_grad_kernel, dim3(grid_for(C)), dim3(kBlock), 0, s, L, (int64_t)C,
                     c.node_dof.get(), c.node_fmask.get(), c.colptr.get(), c.diagpos.get(),
                     c.gG.get(), c.gl.get(), c.dy.get());
  GH_CHECK_HIP(hipGetLastError());
  // back-propagation (weights are the BF16 values the forward pass used)
  const int nb = (C + kMlpCellsPerBlock - 1) / kMlpCellsPerBlock;
  const uint16_t *wq = c.params_bf16.get();
  float *gr = c.grads.get();
  auto wgrad_reduce = [&](int off, int len) {
    hipLaunchKernelGGL(reduce_parts, dim3(grid_for(len)), dim3(kBlock), 0, s, c.wpart.get(), nb,
                       len, gr + off);
  };
  hipLaunchKernelGGL((mlp_wgrad<kMlpHidden, kMlpHidden, kMlpOut, kMlpOutPad>), dim3(nb),
                     dim3(64), 0, s, c.dy.get(), c.h3.get(), c.wpart.get(), C, kMlpCellsPerBlock);
  wgrad_reduce(kMlpW4, kMlpOut * kMlpHidden + kMlpOut);
  hipLaunchKernelGGL((mlp_bwd<kMlpOut, kMlpOutPad, kMlpHidden, kMlpHidden>), dim3(nb), dim3(64),
                     0, s, c.dy.get(), wq, kMlpW4, c.z3.get(), c.dz3.get(), C, kMlpCellsPerBlock);
  hipLaunchKernelGGL((mlp_wgrad<kMlpHidden, kMlpHidden, kMlpHidden, kMlpHidden>), dim3(nb),
                     dim3(64), 0, s, c.dz3.get(), c.h2.get(), c.wpart.get(), C,
                     kMlpCellsPerBlock);
  wgrad_reduce(kMlpW3, kMlpHidden * kMlpHidden + kMlpHidden);
  hipLaunchKernelGGL((mlp_bwd<kMlpHidden, kMlpHidden, kMlpHidden, kMlpHidden>), dim3(nb),
                     dim3(64), 0, s, c.dz3.get(), wq, kMlpW3, c.z2.get(), c.dz2.get(), C,
                     kMlpCellsPerBlock);
  hipLaunchKernelGGL((mlp_wgrad<kMlpHidden, kMlpHidden, kMlpHidden, kMlpHidden>), dim3(nb),
                     dim3(64), 0, s, c.dz2.get(), c.h1.get(), c.wpart.get(), C,
                     kMlpCellsPerBlock);
  wgrad_reduce(kMlpW2, kMlpHidden * kMlpHidden + kMlpHidden);
  hipLaunchKernelGGL((mlp_bwd<kMlpHidden, kMlpHidden, kMlpHidden, kMlpHidden>), dim3(nb),
                     dim3(64), 0, s, c.dz2.get(), wq, kMlpW2, c.z1.get(), c.dz1.get(), C,
                     kMlpCellsPerBlock);
  hipLaunchKernelGGL((mlp_wgrad<kMlpIn, kMlpInPad, kMlpHidden, kMlpHidden>), dim3(nb), dim3(64),
                     0, s, c.dz1.get(), c.feat.get(), c.wpart.get(), C, kMlpCellsPerBlock);
  wgrad_reduce(kMlpW1, kMlpHidden * kMlpIn + kMlpHidden);
  GH_CHECK_HIP(hipGetLastError());
  if (lr > 0.0f) {
    ++c.adam_t;
    const float b1t = (float)std::pow(0.9, c.adam_t), b2t = (float)std::pow(0.999, c.adam_t);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(kMlpParams)), dim3(kBlock), 0, s, kMlpParams,
                       c.params.get(), c.grads.get(), c.adam_m.get(), c.adam_v.get(), lr, b1t,
                       b2t, c.params_bf16.get());
    GH_CHECK_HIP(hipGetLastError());
    c.fsai_built = false;  // G still holds the factor of the weights before the step
  }
  double lp[kResidualBlocks];
  GH_CHECK_HIP(hipMemcpyAsync(lp, c.lpart.get(), sizeof(lp), hipMemcpyDeviceToHost, s));
  GH_CHECK_HIP(hipStreamSynchronize(s));
  double loss = 0.0;
  for (int b = 0; b < kResidualBlocks; ++b) loss += lp[b];
  return loss / ((double)nprobes * (double)std::max<int64_t>(n, 1));
}

void grf_kappa(Ctx &c, int nmodes, const double *omega, const double *phase, double amp) {
  GH_REQUIRE(c.cartesian && c.dim == 3, GH_ERR_UNSUPPORTED,
             "the random-field coefficient is defined on a 3-D Cartesian mesh");
  GH_REQUIRE(nmodes >= 1 && nmodes <= 4096 && omega && phase, GH_ERR_ARG, "bad mode arrays");
  DBuf<double> dom, dph;
  dom.resize(3 * nmodes);
  dph.resize(nmodes);
  GH_CHECK_HIP(hipMemcpyAsync(dom.get(), omega, sizeof(double) * 3 * nmodes,
                              hipMemcpyHostToDevice, c.stream));
  GH_CHECK_HIP(hipMemcpyAsync(dph.get(), phase, sizeof(double) * nmodes, hipMemcpyHostToDevice,
                              c.stream));
  c.kappa.resize(c.ncells);
  const Lat3 L = lattice_of(c);
  hipLaunchKernelGGL(grf_kappa_kernel, dim3(grid_for(c.ncells)), dim3(kBlock), 0, c.stream, L,
                     c.ncells, c.origin[0], c.origin[1], c.origin[2], c.sizes[0], c.sizes[1],
                     c.sizes[2], nmodes, dom.get(), dph.get(), amp, c.kappa.get());
  GH_CHECK_HIP(hipGetLastError());
  GH_CHECK_HIP(hipStreamSynchronize(c.stream));
  c.have_kappa = true;
}

}  // namespace gh
