// Vector-Q2 elasticity on a Cartesian mesh, owner-computes: the assembled matrix
// without element matrices in memory.
//
// On the CartesianMap every cell has the Jacobian diag(h), so the element matrix of
// cell c is K_c = E_c K̂ with one 81x81 reference matrix
//     K̂[(a,i),(b,j)] = λ̂ P_ij[a][b] + μ̂ P_ji[a][b] + μ̂ δ_ij Σ_k P_kk[a][b],
//     P_ij[a][b] = Σ_q w_q |J| g_ai(q) g_bj(q),   g = ĝ/h
// (λ̂ = ν/((1+ν)(1-2ν)), μ̂ = 1/(2(1+ν)); the weak form ε(v) : σ(ε(u)) of
// test/GridapTests/IsotropicDamageTests.jl:13-17), formed once in FP64 and made
// exactly symmetric.  The Q2 nodes form a lattice of (2n+1)^3 points and the
// conforming numbering gives each node's free components consecutive dof ids, so a
// column (node n, component i) holds, in ascending row order, the free components of
// the nodes of the cells around n sorted by their first free dof.  One wavefront
// owns a node: it gathers the (at most 125) neighbour nodes, ranks them by dof id
// (the sorted row order of create_from_nz, src/Algebra/SparseMatrixCSC.jl:264-283),
// forms every entry as Σ_c E_c K̂[(a_c(m),k),(a_c(n),i)] over the cells shared by n and
// m in ascending cell order (Gridap's cell loop), stages the column in LDS and writes
// its nzval run once with coalesced stores.  The Dirichlet neighbours' entries give
// the lifting, and b = f_i Σ_c ∫φ_a - lifting is written by the same wave.  No element
// matrix, no contribution list, no atomics; deterministic.
//
// The symbolic phase builds the node -> dof map of the lattice from cell_dof_ids,
// checks that every cell agrees with it, and checks once that the rows this kernel
// produces are exactly the stored pattern (rowval); otherwise the element-list path
// assembles the space.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "gh_internal.h"

namespace gh {
namespace {

constexpr int kQ2Waves = 8;      // waves per workgroup (they share K̂ in LDS)
constexpr int kQ2MaxNb = 125;    // neighbour nodes of a vertex node (5^3)
constexpr int kQ2MaxRun = 3 * kQ2MaxNb;

// Gridap's local Q2 node order on the HEX (8 vertices, 12 edges, 6 faces, the
// interior; src/ReferenceFEs/CLagrangianRefFEs.jl:626-764) as lattice offsets
// tx + 3 ty + 9 tz inside the cell, and back.
__constant__ int8_t kLexOfLoc[27] = {
    0, 2, 6, 8, 18, 20, 24, 26,                  // vertices (x fastest)
    1, 7, 19, 25, 3, 5, 21, 23, 9, 11, 15, 17,   // edges: x-, y-, z-directed
    4, 22, 10, 16, 12, 14,                       // faces z=0, z=1, y=0, y=1, x=0, x=1
    13};                                         // interior
__constant__ int8_t kLocOfLex[27] = {
    0, 8, 1, 12, 20, 13, 2, 9, 3, 16, 22, 17, 24, 26, 25, 18, 23, 19,
    4, 10, 5, 14, 21, 15, 6, 11, 7};

struct Q2Args {
  int nc[3];                 // cells per axis
  int nn[3];                 // lattice nodes per axis (2 nc + 1)
  int64_t nnodes;
  const int32_t *node_dof;   // nnodes x 3, signed Gridap dof ids
  const int32_t *colptr;
  const int32_t *rowval;     // check mode only
  const double *E;           // per cell, or nullptr (E = 1)
  const void *Kref;          // 81 x 81 (T), [col (b,j)][row (a,i)]
  const double *mref;        // 27: Σ_q w |J| φ_a
  double body[3];
  const double *dvals;       // Dirichlet values or nullptr
  double *nzval;
  double *b;
  int *bad;                  // check mode: set when the rows differ from rowval
  const int32_t *nodes;      // the nodes to assemble, or nullptr: every node
  int64_t nlist;             // their count (with nodes)
};


template <class T, bool CHECK>
__global__ __launch_bounds__(64 * kQ2Waves) void q2_lattice_kernel(const Q2Args P) {
  __shared__ T K[81 * 81];
  __shared__ int keys[kQ2Waves][128];
  __shared__ int nfcs[kQ2Waves][128];
  __shared__ __align__(16) double stages[kQ2Waves][kQ2MaxRun + 2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (!CHECK) {
    const T *Kg = static_cast<const T *>(P.Kref);
    for (int t = threadIdx.x; t < 81 * 81; t += blockDim.x) K[t] = Kg[t];
    __syncthreads();
  }
  int *key = keys[wave];
  int *nfc = nfcs[wave];
  double *stage = stages[wave];
  const int N0 = P.nn[0], N1 = P.nn[1];
  const int64_t nwork = P.nodes ? P.nlist : P.nnodes;
  for (int64_t w = (int64_t)blockIdx.x * kQ2Waves + wave; w < nwork;
       w += (int64_t)gridDim.x * kQ2Waves) {
    const int64_t node = P.nodes ? (int64_t)P.nodes[w] : w;
    int own[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) own[k] = P.node_dof[node * 3 + k];
    if (own[0] <= 0 && own[1] <= 0 && own[2] <= 0) continue;  // wave-uniform
    const int X = (int)(node % N0), Y = (int)((node / N0) % N1), Z = (int)(node / ((int64_t)N0 * N1));
    const int c[3] = {X, Y, Z};
    // the cells around the node (per axis: one if the coordinate is odd, else two)
    int clo[3], chi[3], lo[3], nb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (c[k] & 1) {
        clo[k] = chi[k] = c[k] >> 1;
      } else {
        clo[k] = max((c[k] >> 1) - 1, 0);
        chi[k] = min(c[k] >> 1, P.nc[k] - 1);
      }
      lo[k] = 2 * clo[k];
      nb[k] = 2 * (chi[k] - clo[k]) + 3;
    }
    const int nbox = nb[0] * nb[1] * nb[2];
    // neighbour slots s = lane and lane + 64
    int md[2][3], mx[2][3];
    bool valid[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int s = lane + 64 * h;
      valid[h] = s < nbox;
      int kk = 0, nf = 0;
      if (valid[h]) {
        mx[h][0] = lo[0] + s % nb[0];
        mx[h][1] = lo[1] + (s / nb[0]) % nb[1];
        mx[h][2] = lo[2] + s / (nb[0] * nb[1]);
        const int64_t m = mx[h][0] + (int64_t)N0 * (mx[h][1] + (int64_t)N1 * mx[h][2]);
        kk = INT_MAX;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          md[h][k] = P.node_dof[m * 3 + k];
          if (md[h][k] > 0) {
            kk = min(kk, md[h][k]);
            ++nf;
          }
        }
      } else {
        kk = INT_MAX;
#pragma unroll
        for (int k = 0; k < 3; ++k) md[h][k] = 0, mx[h][k] = 0;
      }
      key[s] = kk;
      nfc[s] = nf;
    }
    wave_sync();
    // position of each neighbour's first free component in the column: the free
    // components of every neighbour with a smaller dof id come first
    int pos[2] = {0, 0}, total = 0;
    const int mine[2] = {key[lane], key[lane + 64]};
    for (int t = 0; t < nbox; ++t) {
      const int kt = key[t], ft = nfc[t];
      total += ft;
#pragma unroll
      for (int h = 0; h < 2; ++h) pos[h] += kt < mine[h] ? ft : 0;
    }
    double ev[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) ev[e] = 0.0;
    {
      int e = 0;
      for (int cz = clo[2]; cz <= chi[2]; ++cz)
        for (int cy = clo[1]; cy <= chi[1]; ++cy)
          for (int cx = clo[0]; cx <= chi[0]; ++cx, ++e)
            ev[e] = P.E ? P.E[cx + (int64_t)P.nc[0] * (cy + (int64_t)P.nc[1] * cz)] : 1.0;
    }
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
      const int j = own[i];
      if (j <= 0) continue;  // wave-uniform
      const int p0 = P.colptr[j - 1];
      const int sh = p0 & 1;  // stage[sh + t] <-> nzval[p0 + t]: 16-byte aligned chunks
      double lift = 0.0;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (!valid[h]) continue;
        double v[3] = {0.0, 0.0, 0.0};
        int e = 0;
        for (int cz = clo[2]; cz <= chi[2]; ++cz)
          for (int cy = clo[1]; cy <= chi[1]; ++cy)
            for (int cx = clo[0]; cx <= chi[0]; ++cx, ++e) {
              const int tx = mx[h][0] - 2 * cx, ty = mx[h][1] - 2 * cy, tz = mx[h][2] - 2 * cz;
              if (tx < 0 || tx > 2 || ty < 0 || ty > 2 || tz < 0 || tz > 2) continue;
              if (CHECK) continue;
              const int am = kLocOfLex[tx + 3 * ty + 9 * tz];
              const int an = kLocOfLex[(X - 2 * cx) + 3 * (Y - 2 * cy) + 9 * (Z - 2 * cz)];
              const T *col = &K[(an + 27 * i) * 81];
              const T Ec = (T)ev[e];
#pragma unroll
              for (int k = 0; k < 3; ++k) v[k] += (double)(Ec * col[am + 27 * k]);
            }
        int r = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int d = md[h][k];
          if (d > 0) {
            if (CHECK) {
              const int at = p0 + pos[h] + r;
              if (at >= P.colptr[j] || P.rowval[at] != d - 1) *P.bad = 1;
            } else {
              stage[sh + pos[h] + r] = v[k];
            }
            ++r;
          } else if (d < 0 && P.dvals) {
            lift += v[k] * P.dvals[-d - 1];
          }
        }
      }
      if (CHECK) {
        if (lane == 0 && P.colptr[j] - p0 != total) *P.bad = 1;
        continue;
      }
      wave_sync();
      double2 *dst2 = reinterpret_cast<double2 *>(P.nzval + (p0 - sh));
      const double2 *st2 = reinterpret_cast<const double2 *>(stage);
      const int lo2 = sh, hi2 = sh + total;
      for (int t = (lo2 + 1) / 2 + lane; t < hi2 / 2; t += 64) dst2[t] = st2[t];
      if (lane == 0 && (lo2 & 1)) P.nzval[p0] = stage[lo2];
      if (lane == 1 && (hi2 & 1)) P.nzval[p0 + total - 1] = stage[hi2 - 1];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) lift += __shfl_xor(lift, o);
      if (lane == 0) {
        double load = 0.0;
        int e = 0;
        for (int cz = clo[2]; cz <= chi[2]; ++cz)
          for (int cy = clo[1]; cy <= chi[1]; ++cy)
            for (int cx = clo[0]; cx <= chi[0]; ++cx, ++e)
              load += P.mref[kLocOfLex[(X - 2 * cx) + 3 * (Y - 2 * cy) + 9 * (Z - 2 * cz)]];
        P.b[j - 1] = P.body[i] * load - lift;
      }
      wave_sync();  // the stage is free for the next component
    }
    wave_sync();    // key/nfc are free for the next node
  }
}

// node -> dof map of the lattice from cell_dof_ids (cells agree on shared nodes; the
// check kernel below verifies it)
__global__ void q2_node_scatter(int64_t ncells, int nc0, int nc1, int N0, int N1,
                                const int32_t *__restrict__ cell_dofs,
                                int32_t *__restrict__ node_dof) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells * 27) return;
  const int64_t cell = t / 27;
  const int a = (int)(t % 27);
  const int cx = (int)(cell % nc0), cy = (int)((cell / nc0) % nc1), cz = (int)(cell / ((int64_t)nc0 * nc1));
  const int lx = kLexOfLoc[a];
  const int64_t node = (2 * cx + lx % 3) + (int64_t)N0 * ((2 * cy + (lx / 3) % 3) +
                                                         (int64_t)N1 * (2 * cz + lx / 9));
#pragma unroll
  for (int k = 0; k < 3; ++k) node_dof[node * 3 + k] = cell_dofs[cell * 81 + a + 27 * k];
}

__global__ void q2_node_verify(int64_t ncells, int nc0, int nc1, int N0, int N1,
                               const int32_t *__restrict__ cell_dofs,
                               const int32_t *__restrict__ node_dof, int *bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells * 27) return;
  const int64_t cell = t / 27;
  const int a = (int)(t % 27);
  const int cx = (int)(cell % nc0), cy = (int)((cell / nc0) % nc1), cz = (int)(cell / ((int64_t)nc0 * nc1));
  const int lx = kLexOfLoc[a];
  const int64_t node = (2 * cx + lx % 3) + (int64_t)N0 * ((2 * cy + (lx / 3) % 3) +
                                                         (int64_t)N1 * (2 * cz + lx / 9));
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (node_dof[node * 3 + k] != cell_dofs[cell * 81 + a + 27 * k]) *bad = 1;
}

// ---- column templates ---------------------------------------------------------------
// Away from the boundary every node of one parity type (x, y, z even or odd) sees the
// same column: the same cells around it, the same neighbours, the same sorted row
// order (the conforming numbering is translation invariant there) and the same
// Dirichlet neighbours.  The symbolic phase groups the nodes by a position class per
// axis (the first and last four lattice positions each form a class, the rest split
// by parity), builds each class's template from one node and checks every node
// against it (rows equal to rowval, Dirichlet neighbours where the template has them):
//   slot s < len: the s-th row of the column (neighbour offset and component),
//   slot len <= s < len + nd: a Dirichlet neighbour (its entries give the lifting),
//   pos[e][t]: the slot of local row t = a + 27 k of the template's e-th cell.
// The column of a templated node is then Σ_e E_e K̂[:, (a_e(n), i)] scattered by pos
// into an LDS accumulator, cell by cell in ascending cell order (Gridap's cell loop);
// the products and sums are those of the ranking kernel above, so both give the same
// bits.  Nodes that match no template keep the ranking kernel.
constexpr int kQ2Slots = 3 * kQ2MaxNb;  // neighbour nodes x components
constexpr int kQ2AxisClasses = 10;
constexpr int kQ2Classes = kQ2AxisClasses * kQ2AxisClasses * kQ2AxisClasses;

struct Q2ClassHdr {
  int32_t ne, len, nd, pad;
  int8_t an[8];     // local Q2 node of the column's node in cell e
  int8_t dc[3][8];  // cell e = (X >> 1, Y >> 1, Z >> 1) + dc[.][e]
};

__host__ __device__ inline int q2_axis_class(int x, int n) {
  if (x < 4) return x;
  if (x > n - 5) return 4 + (n - 1 - x);
  return 8 + (x & 1);
}

// neighbour of `node` addressed by a slot code (offset code * 3 + component), or -1
// outside the lattice
__device__ __forceinline__ int64_t q2_neighbour(const Q2Args &P, int X, int Y, int Z, int code,
                                                int *comp) {
  *comp = code % 3;
  const int o = code / 3;
  const int mx = X + o % 5 - 2, my = Y + (o / 5) % 5 - 2, mz = Z + o / 25 - 2;
  if (mx < 0 || my < 0 || mz < 0 || mx >= P.nn[0] || my >= P.nn[1] || mz >= P.nn[2]) return -1;
  return mx + (int64_t)P.nn[0] * (my + (int64_t)P.nn[1] * mz);
}

// One wave per node: the node's class when its column equals the template's, else -1.
__global__ __launch_bounds__(256) void q2_class_check(const Q2Args P,
                                                      const int16_t *__restrict__ clsmap,
                                                      const Q2ClassHdr *__restrict__ hdr,
                                                      const int16_t *__restrict__ code,
                                                      int16_t *__restrict__ node_class) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int64_t node = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); node < P.nnodes; node += nw) {
    const int X = (int)(node % P.nn[0]), Y = (int)((node / P.nn[0]) % P.nn[1]);
    const int Z = (int)(node / ((int64_t)P.nn[0] * P.nn[1]));
    const int cls = clsmap[q2_axis_class(X, P.nn[0]) +
                           kQ2AxisClasses * (q2_axis_class(Y, P.nn[1]) +
                                             kQ2AxisClasses * q2_axis_class(Z, P.nn[2]))];
    bool ok = cls >= 0;
    if (ok) {
      const Q2ClassHdr &H = hdr[cls];
      // the cells around the node are the template's, in the same order
      const int c[3] = {X, Y, Z};
      int clo[3], chi[3];
      for (int k = 0; k < 3; ++k) {
        if (c[k] & 1) {
          clo[k] = chi[k] = c[k] >> 1;
        } else {
          clo[k] = max((c[k] >> 1) - 1, 0);
          chi[k] = min(c[k] >> 1, P.nc[k] - 1);
        }
      }
      int e = 0;
      for (int cz = clo[2]; cz <= chi[2]; ++cz)
        for (int cy = clo[1]; cy <= chi[1]; ++cy)
          for (int cx = clo[0]; cx <= chi[0]; ++cx, ++e)
            if (e >= 8 || cx - (X >> 1) != H.dc[0][e] || cy - (Y >> 1) != H.dc[1][e] ||
                cz - (Z >> 1) != H.dc[2][e])
              ok = false;
      if (e != H.ne) ok = false;
      const int16_t *cc = code + (int64_t)cls * kQ2Slots;
      for (int i = 0; i < 3 && ok; ++i) {
        const int j = P.node_dof[node * 3 + i];
        if (j <= 0) continue;
        const int p0 = P.colptr[j - 1];
        if (P.colptr[j] - p0 != H.len) {
          ok = false;
          break;
        }
        for (int s = lane; s < H.len; s += 64) {
          int k;
          const int64_t m = q2_neighbour(P, X, Y, Z, cc[s], &k);
          if (m < 0 || P.node_dof[m * 3 + k] != P.rowval[p0 + s] + 1) ok = false;
        }
      }
      for (int s = lane; s < H.nd && ok; s += 64) {
        int k;
        const int64_t m = q2_neighbour(P, X, Y, Z, cc[H.len + s], &k);
        if (m < 0 || P.node_dof[m * 3 + k] >= 0) ok = false;
      }
    }
    ok = __all(ok);
    if (lane == 0) node_class[node] = ok ? (int16_t)cls : (int16_t)-1;
  }
}

// One record per templated node with a free component, written by the symbolic phase:
// {lattice position x | y << 10 | z << 20, template, dof of component 0, 1, 2 (signed),
// first nonzero of each free component's column (-1: Dirichlet)}.
struct Q2NodeRec {
  int32_t xyz, cls, own[3], p0[3];
};
constexpr int kQ2AccLd = 376;  // LDS accumulator of one column (>= kQ2Slots + 1, even)

// Read-only data through the constant address space: a wave-uniform address becomes a
// scalar load, whose wait (lgkmcnt) is not held up by the wave's outstanding nzval
// stores the way a vector load's wait (vmcnt) is.
template <class T>
__device__ __forceinline__ T ldc(const T *p) {
  typedef const __attribute__((address_space(4))) T ConstT;
  return *(ConstT *)p;
}

__device__ __forceinline__ int4 ldc4(const int4 *p) {
  const int *q = reinterpret_cast<const int *>(p);
  return make_int4(ldc(q), ldc(q + 1), ldc(q + 2), ldc(q + 3));
}

__device__ __forceinline__ int sbyte(int w, int k) { return (int)(int8_t)(w >> (8 * k)); }

// One wave per node, one component's column at a time: lane L takes the cell's local
// rows t = L and t = 64 + L (L < 17) of the column; pos2 holds, per template, cell and
// lane, the slots of those two rows (16 bits each).  The records are sorted by
// template, so the per-lane slot tables are loaded once per template change; every
// other read is a wave-uniform scalar load, so no wait in the node loop is held up by
// the wave's outstanding nzval stores.
template <class T, int W>
__global__ __launch_bounds__(64 * W) void q2_class_kernel(
    const Q2Args P, const int4 *__restrict__ hdr, const uint32_t *__restrict__ pos2,
    const int16_t *__restrict__ code, const int4 *__restrict__ rec, int64_t nrec) {
  __shared__ T K[81 * 81];
  // per wave: one column's accumulator; one more region, shared by the waves, takes
  // the entries of inactive lanes (their values are never read)
  __shared__ __align__(16) double accs[(W + 1) * kQ2AccLd];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const T *Kg = static_cast<const T *>(P.Kref);
  for (int t = threadIdx.x; t < 81 * 81; t += blockDim.x) K[t] = Kg[t];
  __syncthreads();
  double *acc = accs + wave * kQ2AccLd;
  const int dummy = (W - wave) * kQ2AccLd + lane;  // the shared region, from acc
  const bool hi = lane < 81 - 64;                  // the lane also takes row 64 + lane
  const int64_t nw = (int64_t)gridDim.x * W, gw = (int64_t)blockIdx.x * W + wave;
  const int nc0 = P.nc[0], nc01 = P.nc[0] * P.nc[1];
  // the current template
  int cur = -1, ne = 0, len = 0, nd = 0;
  int an[8], coff[8];
  uint32_t pp[8];
  double msum = 0.0;
#pragma unroll
  for (int e = 0; e < 8; ++e) an[e] = coff[e] = 0, pp[e] = 0;
  // strided over the sorted records: every wave sees each template's share of the
  // nodes (balanced work) and still changes template rarely
  for (int64_t r = gw; r < nrec; r += nw) {
    const int4 ra = ldc4(rec + 2 * r), rb = ldc4(rec + 2 * r + 1);
    const int xyz = ra.x, cls = ra.y;
    const int own[3] = {ra.z, ra.w, rb.x};
    const int p0s[3] = {rb.y, rb.z, rb.w};
    if (cls != cur) {  // wave-uniform; rare: the records are sorted by template
      cur = cls;
      const int4 h0 = ldc4(hdr + 3 * cls), h1 = ldc4(hdr + 3 * cls + 1), h2 = ldc4(hdr + 3 * cls + 2);
      ne = h0.x;
      len = h0.y;
      nd = h0.z;
      msum = 0.0;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = e & 3;
        an[e] = sbyte(e < 4 ? h1.x : h1.y, k);
        const int dx = sbyte(e < 4 ? h1.z : h1.w, k), dy = sbyte(e < 4 ? h2.x : h2.y, k),
                  dz = sbyte(e < 4 ? h2.z : h2.w, k);
        coff[e] = dx + nc0 * dy + nc01 * dz;
        if (e < ne) {
          pp[e] = pos2[((int64_t)cls * 8 + e) * 64 + lane];
          msum += ldc(P.mref + an[e]);
        }
      }
      // wait for the slot tables here, on this rare path, so that the common path
      // carries no wait on vector loads (which would also wait for the stores)
#pragma unroll
      for (int e = 0; e < 8; ++e) asm volatile("" ::"v"(pp[e]));
    }
    const int X = xyz & 1023, Y = (xyz >> 10) & 1023, Z = xyz >> 20;
    const int64_t cbase = (X >> 1) + (int64_t)nc0 * (Y >> 1) + (int64_t)nc01 * (Z >> 1);
    T Ec[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) Ec[e] = e < ne ? (T)(P.E ? ldc(P.E + cbase + coff[e]) : 1.0) : (T)0;
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
      const int j = own[i];
      if (j <= 0) continue;  // wave-uniform
      const int p0 = p0s[i];
      const int sh = p0 & 1;  // acc[sh + slot] <-> nzval[p0 + slot]: 16-byte aligned chunks
      double2 *acc2 = reinterpret_cast<double2 *>(acc);
      for (int t = lane; t < kQ2AccLd / 2; t += 64) acc2[t] = make_double2(0.0, 0.0);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        if (e < ne) {
          const T *kc = &K[(an[e] + 27 * i) * 81];
          // the lane's two rows of this cell have distinct slots: both reads, then both
          // writes; the next cell's reads follow these writes in the wave's LDS order
          const int a0 = sh + (int)(pp[e] & 0xffffu);
          const int a1 = hi ? sh + (int)(pp[e] >> 16) : dummy;
          const T k0 = kc[lane], k1 = kc[hi ? 64 + lane : lane];
          const double v0 = acc[a0], v1 = acc[a1];
          acc[a0] = v0 + (double)(Ec[e] * k0);
          acc[a1] = v1 + (double)(Ec[e] * k1);
        }
      }
      wave_sync();
      double2 *dst2 = reinterpret_cast<double2 *>(P.nzval + (p0 - sh));
      const double2 *st2 = reinterpret_cast<const double2 *>(acc);
      const int lo2 = sh, hi2 = sh + len;
      for (int t = (lo2 + 1) / 2 + lane; t < hi2 / 2; t += 64) dst2[t] = st2[t];
      if (lane == 0 && (lo2 & 1)) P.nzval[p0] = acc[lo2];
      if (lane == 1 && (hi2 & 1)) P.nzval[p0 + len - 1] = acc[hi2 - 1];
      double lift = 0.0;
      if (P.dvals) {
        const int16_t *cc = code + (int64_t)cls * kQ2Slots + len;
        for (int t = lane; t < nd; t += 64) {
          int k;
          const int64_t m = q2_neighbour(P, X, Y, Z, cc[t], &k);
          lift += acc[sh + len + t] * P.dvals[-P.node_dof[m * 3 + k] - 1];
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) lift += __shfl_xor(lift, o);
      if (lane == 0) P.b[j - 1] = P.body[i] * msum - lift;
      wave_sync();  // the accumulator is free for the next component
    }
  }
}

Q2Args make_args(const Ctx &c) {
  Q2Args A{};
  for (int k = 0; k < 3; ++k) {
    A.nc[k] = (int)c.part[k];
    A.nn[k] = 2 * (int)c.part[k] + 1;
  }
  A.nnodes = (int64_t)A.nn[0] * A.nn[1] * A.nn[2];
  A.node_dof = c.q2_node_dof.get();
  A.colptr = c.colptr.get();
  A.rowval = c.rowval.get();
  return A;
}

// Host copy of kLexOfLoc (the template builder runs on the host).
const int8_t kLexOfLocHost[27] = {0, 2, 6, 8, 18, 20, 24, 26, 1, 7, 19, 25, 3, 5,
                                  21, 23, 9, 11, 15, 17, 4, 22, 10, 16, 12, 14, 13};

// The column templates of a lattice space (the node map must be on the device): one per
// position class, built from the first node of the class; every node is then checked
// against its class's template.  Nodes without one are listed for the ranking kernel.
void q2_build_classes(Ctx &c, const Q2Args &A) {
  const int64_t nn = A.nnodes;
  hipStream_t s = c.stream;
  std::vector<int32_t> nd((size_t)nn * 3);
  GH_CHECK_HIP(hipMemcpyAsync(nd.data(), c.q2_node_dof.get(), sizeof(int32_t) * nd.size(),
                              hipMemcpyDeviceToHost, s));
  GH_CHECK_HIP(hipStreamSynchronize(s));
  int first[3][kQ2AxisClasses];
  for (int k = 0; k < 3; ++k) {
    std::fill(first[k], first[k] + kQ2AxisClasses, -1);
    for (int x = A.nn[k] - 1; x >= 0; --x) first[k][q2_axis_class(x, A.nn[k])] = x;
  }
  std::vector<int16_t> clsmap(kQ2Classes, -1);
  std::vector<Q2ClassHdr> hdr;
  std::vector<uint32_t> pos;
  std::vector<int16_t> codes;
  std::vector<int> slot_of(kQ2Slots);
  for (int az = 0; az < kQ2AxisClasses; ++az)
    for (int ay = 0; ay < kQ2AxisClasses; ++ay)
      for (int ax = 0; ax < kQ2AxisClasses; ++ax) {
        const int rx = first[0][ax], ry = first[1][ay], rz = first[2][az];
        if (rx < 0 || ry < 0 || rz < 0) continue;
        const int r[3] = {rx, ry, rz};
        int clo[3], chi[3];
        for (int k = 0; k < 3; ++k) {
          if (r[k] & 1) {
            clo[k] = chi[k] = r[k] >> 1;
          } else {
            clo[k] = std::max((r[k] >> 1) - 1, 0);
            chi[k] = std::min(r[k] >> 1, A.nc[k] - 1);
          }
        }
        // the window: every node of the cells around the representative
        std::vector<std::pair<int32_t, int>> freev;
        std::vector<int> dirv;
        bool valid = true;
        for (int mz = 2 * clo[2]; mz <= 2 * chi[2] + 2; ++mz)
          for (int my = 2 * clo[1]; my <= 2 * chi[1] + 2; ++my)
            for (int mx = 2 * clo[0]; mx <= 2 * chi[0] + 2; ++mx) {
              const int64_t m = mx + (int64_t)A.nn[0] * (my + (int64_t)A.nn[1] * mz);
              const int oc = (mx - rx + 2) + 5 * (my - ry + 2) + 25 * (mz - rz + 2);
              for (int k = 0; k < 3; ++k) {
                const int32_t d = nd[(size_t)m * 3 + k];
                if (d > 0) freev.push_back({d, oc * 3 + k});
                else if (d < 0) dirv.push_back(oc * 3 + k);
                else valid = false;
              }
            }
        if (!valid) continue;
        std::sort(freev.begin(), freev.end());
        std::fill(slot_of.begin(), slot_of.end(), -1);
        Q2ClassHdr H{};
        H.len = (int)freev.size();
        H.nd = (int)dirv.size();
        std::vector<int16_t> cc(kQ2Slots, 0);
        for (int t = 0; t < H.len; ++t) {
          slot_of[freev[t].second] = t;
          cc[t] = (int16_t)freev[t].second;
        }
        for (int t = 0; t < H.nd; ++t) {
          slot_of[dirv[t]] = H.len + t;
          cc[H.len + t] = (int16_t)dirv[t];
        }
        std::vector<uint32_t> pc((size_t)8 * 64, 0);  // [e][lane]: rows lane, 64 + lane
        int e = 0;
        for (int cz = clo[2]; cz <= chi[2]; ++cz)
          for (int cy = clo[1]; cy <= chi[1]; ++cy)
            for (int cx = clo[0]; cx <= chi[0]; ++cx, ++e) {
              H.dc[0][e] = (int8_t)(cx - (rx >> 1));
              H.dc[1][e] = (int8_t)(cy - (ry >> 1));
              H.dc[2][e] = (int8_t)(cz - (rz >> 1));
              const int an_lex = (rx - 2 * cx) + 3 * (ry - 2 * cy) + 9 * (rz - 2 * cz);
              H.an[e] = -1;
              for (int a = 0; a < 27; ++a)
                if (kLexOfLocHost[a] == an_lex) H.an[e] = (int8_t)a;
              for (int t = 0; t < 81; ++t) {
                const int a = t % 27, k = t / 27, lx = kLexOfLocHost[a];
                const int mx = 2 * cx + lx % 3, my = 2 * cy + (lx / 3) % 3, mz = 2 * cz + lx / 9;
                const int oc = (mx - rx + 2) + 5 * (my - ry + 2) + 25 * (mz - rz + 2);
                const int sl = slot_of[oc * 3 + k];
                if (sl < 0 || H.an[e] < 0) valid = false;
                pc[(size_t)e * 64 + t % 64] |= (uint32_t)(sl & 0xffff) << (t < 64 ? 0 : 16);
              }
            }
        H.ne = e;
        if (!valid) continue;
        clsmap[ax + kQ2AxisClasses * (ay + kQ2AxisClasses * az)] = (int16_t)hdr.size();
        hdr.push_back(H);
        pos.insert(pos.end(), pc.begin(), pc.end());
        codes.insert(codes.end(), cc.begin(), cc.end());
      }
  DBuf<int16_t> node_class;
  node_class.resize(nn);
  if (hdr.empty()) {
    GH_CHECK_HIP(hipMemsetAsync(node_class.get(), 0xff, sizeof(int16_t) * nn, s));
  } else {
    c.q2_cls_hdr.resize(hdr.size() * sizeof(Q2ClassHdr));
    c.q2_cls_pos.resize(pos.size());
    c.q2_cls_code.resize(codes.size());
    DBuf<int16_t> dmap;
    dmap.resize(clsmap.size());
    GH_CHECK_HIP(hipMemcpyAsync(c.q2_cls_hdr.get(), hdr.data(), hdr.size() * sizeof(Q2ClassHdr),
                                hipMemcpyHostToDevice, s));
    GH_CHECK_HIP(hipMemcpyAsync(c.q2_cls_pos.get(), pos.data(), pos.size() * sizeof(uint32_t),
                                hipMemcpyHostToDevice, s));
    GH_CHECK_HIP(hipMemcpyAsync(c.q2_cls_code.get(), codes.data(), codes.size() * sizeof(int16_t),
                                hipMemcpyHostToDevice, s));
    GH_CHECK_HIP(hipMemcpyAsync(dmap.get(), clsmap.data(), clsmap.size() * sizeof(int16_t),
                                hipMemcpyHostToDevice, s));
    const unsigned g = (unsigned)std::min<int64_t>((nn + 3) / 4, 8192);
    hipLaunchKernelGGL(q2_class_check, dim3(g), dim3(256), 0, s, A, dmap.get(),
                       (const Q2ClassHdr *)c.q2_cls_hdr.get(), c.q2_cls_code.get(),
                       node_class.get());
    GH_CHECK_HIP(hipGetLastError());
    GH_CHECK_HIP(hipStreamSynchronize(s));  // dmap is released at scope end
  }
  std::vector<int16_t> ncls((size_t)nn);
  std::vector<int32_t> cp((size_t)c.nfree + 1);
  GH_CHECK_HIP(hipMemcpyAsync(ncls.data(), node_class.get(), sizeof(int16_t) * nn,
                              hipMemcpyDeviceToHost, s));
  GH_CHECK_HIP(hipMemcpyAsync(cp.data(), c.colptr.get(), sizeof(int32_t) * cp.size(),
                              hipMemcpyDeviceToHost, s));
  GH_CHECK_HIP(hipStreamSynchronize(s));
  std::vector<int32_t> irr;
  std::vector<Q2NodeRec> recs;
  for (int64_t m = 0; m < nn; ++m) {
    const int32_t *o = &nd[(size_t)m * 3];
    if (o[0] <= 0 && o[1] <= 0 && o[2] <= 0) continue;
    if (ncls[m] < 0) {
      irr.push_back((int32_t)m);
      continue;
    }
    Q2NodeRec R{};
    const int64_t X = m % A.nn[0], Y = (m / A.nn[0]) % A.nn[1], Z = m / ((int64_t)A.nn[0] * A.nn[1]);
    R.xyz = (int32_t)(X | (Y << 10) | (Z << 20));
    R.cls = ncls[m];
    for (int k = 0; k < 3; ++k) {
      R.own[k] = o[k];
      R.p0[k] = o[k] > 0 ? cp[o[k] - 1] : -1;
    }
    recs.push_back(R);
  }
  // by template, then lattice order: a wave's run of records rarely changes template
  std::stable_sort(recs.begin(), recs.end(),
                   [](const Q2NodeRec &x, const Q2NodeRec &y) { return x.cls < y.cls; });
  c.q2_nrec = (int64_t)recs.size();
  c.q2_rec.resize(recs.size() * 8);
  if (!recs.empty())
    GH_CHECK_HIP(hipMemcpy(c.q2_rec.get(), recs.data(), sizeof(Q2NodeRec) * recs.size(),
                           hipMemcpyHostToDevice));
  c.q2_nirregular = (int64_t)irr.size();
  c.q2_irregular.resize(irr.size());
  if (!irr.empty())
    GH_CHECK_HIP(hipMemcpy(c.q2_irregular.get(), irr.data(), sizeof(int32_t) * irr.size(),
                           hipMemcpyHostToDevice));
}

unsigned q2_class_grid(int64_t nnodes, bool f64, int waves) {
  const int64_t want = (nnodes + waves - 1) / waves;
  const char *env = std::getenv("GRIDAP_HIP_Q2_GRID");
  // measured at 48^3: FP32 8 waves per workgroup (3 workgroups per CU), 2048 workgroups;
  // FP64 16 waves (one workgroup per CU by LDS), 512 workgroups
  const int64_t cap = env && std::atoi(env) > 0 ? std::atoi(env) : (f64 ? 512 : 2048);
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, cap));
}

unsigned q2_grid(int64_t nnodes, bool f64) {
  const int64_t want = (nnodes + kQ2Waves - 1) / kQ2Waves;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, f64 ? 256 : 512));
}

}  // namespace

// Symbolic phase of the lattice path (a Cartesian vector-Q2 space with its pattern):
// the node -> dof map, its agreement with every cell, and the produced rows against
// rowval.  Sets c.q2lattice.
void q2lattice_prepare(Ctx &c) {
  c.q2lattice = false;
  c.q2_node_dof.release();
  const char *env = std::getenv("GRIDAP_HIP_Q2_LATTICE");
  if (env && env[0] == '0') return;
  if (!(c.cartesian && c.dim == 3 && c.reffe_kind == GH_REFFE_QK && c.order == 2 &&
        c.ncomp == 3 && c.nloc == 81 && c.have_pattern && c.nfree > 0))
    return;
  const int64_t nn = (2 * c.part[0] + 1) * (2 * c.part[1] + 1) * (2 * c.part[2] + 1);
  if (nn * 3 >= (int64_t(1) << 31)) return;
  hipStream_t s = c.stream;
  c.q2_node_dof.resize(nn * 3);
  GH_CHECK_HIP(hipMemsetAsync(c.q2_node_dof.get(), 0, sizeof(int32_t) * nn * 3, s));
  Q2Args A = make_args(c);
  const int64_t work = c.ncells * 27;
  const unsigned g = (unsigned)((work + 255) / 256);
  hipLaunchKernelGGL(q2_node_scatter, dim3(g), dim3(256), 0, s, c.ncells, A.nc[0], A.nc[1],
                     A.nn[0], A.nn[1], c.cell_dofs.get(), c.q2_node_dof.get());
  DBuf<int> bad;
  bad.resize(1);
  GH_CHECK_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int), s));
  hipLaunchKernelGGL(q2_node_verify, dim3(g), dim3(256), 0, s, c.ncells, A.nc[0], A.nc[1],
                     A.nn[0], A.nn[1], c.cell_dofs.get(), c.q2_node_dof.get(), bad.get());
  A.bad = bad.get();
  hipLaunchKernelGGL((q2_lattice_kernel<float, true>), dim3(q2_grid(A.nnodes, false)),
                     dim3(64 * kQ2Waves), 0, s, A);
  GH_CHECK_HIP(hipGetLastError());
  int hbad = 1;
  GH_CHECK_HIP(hipMemcpyAsync(&hbad, bad.get(), sizeof(int), hipMemcpyDeviceToHost, s));
  GH_CHECK_HIP(hipStreamSynchronize(s));
  if (hbad) {
    c.q2_node_dof.release();
    return;
  }
  const char *tenv = std::getenv("GRIDAP_HIP_Q2_TEMPLATES");
  if ((tenv && tenv[0] == '0') || A.nn[0] > 1024 || A.nn[1] > 1024 || A.nn[2] > 1024) {
    // every node through the ranking kernel
    c.q2_rec.release();
    c.q2_nrec = 0;
    c.q2_irregular.release();
    c.q2_nirregular = -1;
  } else {
    q2_build_classes(c, A);
  }
  c.q2lattice = true;
}

// The numeric phase on the lattice: K̂ and m̂ from the uploaded tabulation (FP64, once
// per tabulation, cell sizes and ν), then one launch writes nzval and b.
void q2lattice_assemble(Ctx &c, bool f32, bool vlap) {
  GH_REQUIRE(c.q2lattice, GH_ERR_STATE, "no vector-Q2 lattice for this space");
  c.last_kernel = GH_KERNEL_Q2_LATTICE;
  // the form is part of the key: ν lies in (-1, 0.5), -2 marks the vector Laplacian
  const double key[4] = {vlap ? -2.0 : c.nu, c.sizes[0], c.sizes[1], c.sizes[2]};
  const bool fresh = !(c.q2_ref_ready && std::equal(key, key + 4, c.q2_ref_key));
  if (fresh) {
    const int nq = c.nq, NN = 27;
    const double vol = c.sizes[0] * c.sizes[1] * c.sizes[2];
    std::vector<double> g((size_t)nq * NN * 3);
    for (int q = 0; q < nq; ++q)
      for (int a = 0; a < NN; ++a)
        for (int d = 0; d < 3; ++d)
          g[((size_t)q * NN + a) * 3 + d] = c.grad[((size_t)q * NN + a) * 3 + d] / c.sizes[d];
    std::vector<double> P((size_t)9 * NN * NN, 0.0);  // P[(i*3+j)][a][b]
    for (int q = 0; q < nq; ++q) {
      const double w = c.qw[q] * vol;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int a = 0; a < NN; ++a) {
            const double ga = w * g[((size_t)q * NN + a) * 3 + i];
            for (int b = 0; b < NN; ++b)
              P[((size_t)(i * 3 + j) * NN + a) * NN + b] += ga * g[((size_t)q * NN + b) * 3 + j];
          }
    }
    const double lam = c.nu / ((1.0 + c.nu) * (1.0 - 2.0 * c.nu));
    const double mu = 1.0 / (2.0 * (1.0 + c.nu));
    std::vector<double> K((size_t)81 * 81);
    auto Pij = [&](int i, int j, int a, int b) { return P[((size_t)(i * 3 + j) * NN + a) * NN + b]; };
    for (int i = 0; i < 3; ++i)
      for (int a = 0; a < NN; ++a)
        for (int j = 0; j < 3; ++j)
          for (int b = 0; b < NN; ++b) {
            double v = vlap ? 0.0 : lam * Pij(i, j, a, b) + mu * Pij(j, i, a, b);
            if (i == j)  // vector Laplacian: δ_ij ∇φ_a⋅∇φ_b
              v += (vlap ? 1.0 : mu) * (Pij(0, 0, a, b) + Pij(1, 1, a, b) + Pij(2, 2, a, b));
            K[(size_t)(a + 27 * i) * 81 + (b + 27 * j)] = v;  // [row][col]
          }
    // exactly symmetric; stored by column ([col][row]) for the kernel
    std::vector<double> Kc((size_t)81 * 81);
    std::vector<float> Kf((size_t)81 * 81);
    for (int r = 0; r < 81; ++r)
      for (int cc = 0; cc < 81; ++cc) {
        const double v = 0.5 * (K[(size_t)r * 81 + cc] + K[(size_t)cc * 81 + r]);
        Kc[(size_t)cc * 81 + r] = v;
        Kf[(size_t)cc * 81 + r] = (float)v;
      }
    std::vector<double> m(NN, 0.0);
    for (int q = 0; q < nq; ++q)
      for (int a = 0; a < NN; ++a) m[a] += c.qw[q] * vol * c.shp[(size_t)q * NN + a];
    c.q2_ref.resize(81 * 81 + NN);
    c.q2_ref32.resize(81 * 81);
    GH_CHECK_HIP(hipMemcpy(c.q2_ref.get(), Kc.data(), sizeof(double) * 81 * 81,
                           hipMemcpyHostToDevice));
    GH_CHECK_HIP(hipMemcpy(c.q2_ref.get() + 81 * 81, m.data(), sizeof(double) * NN,
                           hipMemcpyHostToDevice));
    GH_CHECK_HIP(hipMemcpy(c.q2_ref32.get(), Kf.data(), sizeof(float) * 81 * 81,
                           hipMemcpyHostToDevice));
    std::copy(key, key + 4, c.q2_ref_key);
    c.q2_ref_ready = true;
  }
  Q2Args A = make_args(c);
  A.E = c.have_kappa ? c.kappa.get() : nullptr;
  A.Kref = f32 ? (const void *)c.q2_ref32.get() : (const void *)c.q2_ref.get();
  A.mref = c.q2_ref.get() + 81 * 81;
  for (int k = 0; k < 3; ++k) A.body[k] = c.body[k];
  A.dvals = c.have_dvals ? c.dvals.get() : nullptr;
  A.nzval = c.nzval.get();
  A.b = c.b.get();
  // templated nodes
  if (c.q2_nirregular >= 0) {
    static_assert(sizeof(Q2ClassHdr) == 48 && sizeof(Q2NodeRec) == 32, "record layouts");
    const int4 *hdr = (const int4 *)c.q2_cls_hdr.get();
    const int4 *rec = (const int4 *)c.q2_rec.get();
    const char *wenv = std::getenv("GRIDAP_HIP_Q2_WAVES");
    const bool w8 = wenv ? std::atoi(wenv) == 8 : f32;
    if (c.q2_nrec > 0) {
      const unsigned g = q2_class_grid(c.q2_nrec, !f32, w8 ? 8 : 16);
      auto go = [&](auto kern, int w) {
        hipLaunchKernelGGL(kern, dim3(g), dim3(64 * w), 0, c.stream, A, hdr,
                           (const uint32_t *)c.q2_cls_pos.get(), c.q2_cls_code.get(), rec,
                           c.q2_nrec);
      };
      if (f32) {
        if (w8) go(q2_class_kernel<float, 8>, 8);
        else go(q2_class_kernel<float, 16>, 16);
      } else {
        if (w8) go(q2_class_kernel<double, 8>, 8);
        else go(q2_class_kernel<double, 16>, 16);
      }
    }
    GH_CHECK_HIP(hipGetLastError());
    if (c.q2_nirregular == 0) return;
    A.nodes = c.q2_irregular.get();
    A.nlist = c.q2_nirregular;
  }
  // the rest through the ranking kernel
  const int64_t work = A.nodes ? A.nlist : A.nnodes;
  if (f32)
    hipLaunchKernelGGL((q2_lattice_kernel<float, false>), dim3(q2_grid(work, false)),
                       dim3(64 * kQ2Waves), 0, c.stream, A);
  else
    hipLaunchKernelGGL((q2_lattice_kernel<double, false>), dim3(q2_grid(work, true)),
                       dim3(64 * kQ2Waves), 0, c.stream, A);
  GH_CHECK_HIP(hipGetLastError());
}

}  // namespace gh
