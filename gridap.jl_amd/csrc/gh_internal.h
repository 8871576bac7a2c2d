// Internal declarations of the gfx950 assembly engine (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <memory>
#include <vector>

#include "gridap_hip.h"

namespace gh {

// Lattice nodes per tile of the lattice kernel (one wavefront).
constexpr int kTileNodes = 64;

// Workgroups are dealt round-robin to the 8 XCDs; remap so XCD x takes the contiguous
// x-th slice of the work and the data that neighbouring work items share stays in
// that XCD's L2.  Bijective on [0, nblocks).
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  constexpr int kXcds = 8;
  const int q = nblocks / kXcds, r = nblocks % kXcds, x = b % kXcds;
  return x * q + min(x, r) + b / kXcds;
}

// LDS written by some lanes of a wavefront is visible to the wavefront's other lanes.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Error carried from the point of failure to the ABI boundary, where it becomes a
// status code plus gh_last_error() text.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define GH_CHECK_HIP(expr)                                                              \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      int code_ = (e_ == hipErrorOutOfMemory) ? GH_ERR_OOM : GH_ERR_HIP;                \
      throw ::gh::Error(code_, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    }                                                                                   \
  } while (0)

#define GH_REQUIRE(cond, code, msg)                                                     \
  do {                                                                                  \
    if (!(cond)) throw ::gh::Error((code), (msg));                                      \
  } while (0)

// Owning device buffer (hipMalloc/hipFree), sized in elements.
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void resize(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    GH_CHECK_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  // grow-only: keeps the buffer (and its contents) when it is large enough
  void reserve(size_t count) {
    if (count > n) resize(count);
  }
  T *get() const { return p; }
};

// Element-matrix storage of the vector-Q2 hexahedron (81 local dofs, node + 27 comp):
// column-major, one column of kHexColLd entries per local dof, the rows of component
// i at 28 i + node, so that four consecutive nodes of one component are one aligned
// 16-byte group.  Entry (row li, column lj) of cell c: (c·81 + lj)·kHexColLd + slot(li).
constexpr int kHexColLd = 84;
__host__ __device__ inline int hex_row_slot(int li) { return li + li / 27; }

// det J and J⁻¹ of a D x D matrix (row-major), FP64.
template <int D>
__host__ __device__ inline double det_inv(const double *J, double *Ji) {
  if (D == 2) {
    double det = J[0] * J[3] - J[1] * J[2];
    double r = 1.0 / det;
    Ji[0] = J[3] * r; Ji[1] = -J[1] * r; Ji[2] = -J[2] * r; Ji[3] = J[0] * r;
    return det;
  } else {
    double a = J[0], b = J[1], cc = J[2], d = J[3], e = J[4], f = J[5], g = J[6], h = J[7],
           i = J[8];
    double c0 = e * i - f * h, c1 = f * g - d * i, c2 = d * h - e * g;
    double det = a * c0 + b * c1 + cc * c2;
    double r = 1.0 / det;
    Ji[0] = c0 * r; Ji[1] = (cc * h - b * i) * r; Ji[2] = (b * f - cc * e) * r;
    Ji[3] = c1 * r; Ji[4] = (a * i - cc * g) * r; Ji[5] = (cc * d - a * f) * r;
    Ji[6] = c2 * r; Ji[7] = (b * g - a * h) * r; Ji[8] = (a * e - b * d) * r;
    return det;
  }
}

// Pinned host memory (small read-backs of the symbolic phase).
template <class T>
struct HBuf {
  T *p = nullptr;
  size_t n = 0;
  HBuf() = default;
  HBuf(const HBuf &) = delete;
  HBuf &operator=(const HBuf &) = delete;
  ~HBuf() {
    if (p) (void)hipHostFree(p);
  }
  T *get(size_t count) {
    if (count > n) {
      if (p) (void)hipHostFree(p);
      p = nullptr;
      n = 0;
      GH_CHECK_HIP(hipHostMalloc((void **)&p, count * sizeof(T), hipHostMallocDefault));
      n = count;
    }
    return p;
  }
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // workspace of the lattice symbolic phase, kept between calls (it is rebuilt for
  // every pattern, so reusing it saves the allocations, not any result)
  DBuf<int32_t> ws_f, ws_scan, ws_count, ws_start, ws_flag;
  DBuf<int64_t> ws_box;
  DBuf<unsigned char> ws_cub;
  HBuf<int64_t> h_small;

  // ---- mesh ----
  int dim = 0;
  bool cartesian = false;
  bool lattice_topo = false;        // cells and nodes of a Cartesian grid (part), any coordinates
  double origin[3] = {0, 0, 0}, sizes[3] = {1, 1, 1};
  int64_t part[3] = {1, 1, 1};
  int64_t nnodes = 0, ncells = 0;
  int nv = 0;                       // geometric nodes per cell
  DBuf<double> xyz;                 // nnodes x dim
  DBuf<int32_t> cell_nodes;         // ncells x nv, 0-based

  // ---- FE space ----
  bool have_space = false;
  int reffe_kind = 0, order = 0, ncomp = 0, nloc = 0;
  int64_t nfree = 0, ndiri = 0;
  DBuf<int32_t> cell_dofs;          // ncells x nloc, Gridap signed 1-based ids
  std::vector<int32_t> h_cell_dofs; // host copy (colouring)

  // ---- trial space of a rectangular field pair (gh_trial_space_upload): the columns,
  //      while the space above gives the rows ----
  bool have_trial = false;
  int torder = 0, tnloc = 0, tncomp = 1;
  int64_t tnfree = 0, tndiri = 0;
  DBuf<int32_t> tcell_dofs;         // ncells x tnloc, Gridap signed 1-based ids
  int tnq = 0;                      // points of the trial tabulation
  std::vector<double> tshp, tgrad;  // trial values and reference gradients there

  // ---- composite matrix of a multi-field form (gh_compose_block): the stored values of
  //      other contexts' matrices, each moved to its composite positions ----
  struct ComposeBlock {
    Ctx *src = nullptr;        // the context whose nzval is the block
    int64_t n = 0;             // its stored values
    DBuf<int32_t> dst;         // composite position of each, 0-based
  };
  std::vector<std::unique_ptr<ComposeBlock>> comp;
  int64_t comp_nnz = -1;     // composite stored values (-1: none registered)
  DBuf<double> comp_nz;
  std::vector<bool> comp_taken;  // composite positions some block writes

  // ---- reference tabulation ----
  int nq = 0;
  std::vector<double> qw, shp, grad;

  // ---- pattern ----
  bool have_pattern = false;
  int64_t nnz = 0;
  DBuf<int32_t> colptr;             // nfree+1, 0-based
  DBuf<int32_t> rowval;             // nnz, 0-based

  // ---- structured (lattice) eligibility ----
  bool structured = false;
  DBuf<int32_t> node_dof;           // nnodes: Gridap dof of each lattice node
  DBuf<uint32_t> node_fmask;        // nnodes: bit o set when stencil neighbour o is free
  DBuf<int32_t> tile_p0;            // per 64-node tile: nzval start of its columns, or -1
  DBuf<int32_t> tile_d0;            // per 64-node tile: dof of its first free node, or -1

  // ---- element-list path (P1/P2 tetrahedra, vector Q2 hexahedra) ----
  bool elemlist = false;            // adjacency, contribution lists and masks ready
  DBuf<int32_t> adjptr;             // nfree+1: cells of each free dof
  DBuf<int32_t> adj;                // (cell << 4 | local index), cells ascending per dof
  DBuf<int32_t> cstart;             // nnz+1: contribution range of each nonzero
  DBuf<int32_t> perm;               // contributions: emit index cell·nloc² + lj·nloc + li
  // compact contribution lists (replace cstart/perm when they fit): the column sum
  // takes one range of consecutive columns per block (bstart: first column of each);
  // per contribution a 16-bit slot (block adjacency position << obits | entry offset),
  // per nonzero its first contribution relative to its block (noff), per block its
  // first contribution (bbase)
  bool slots = false;
  bool metric = false;              // simplices: per-cell metric records (entries formed in the sum)
  bool inverse = false;             // cslot holds run positions per (adjacency entry, row)
  int obits = 0;                    // offset bits of a slot
  int maxbase = 0;                  // most adjacency entries of one block
  int maxq = 0;                     // most contributions of one block
  DBuf<uint16_t> cslot;
  DBuf<uint16_t> noff;
  DBuf<int32_t> bbase;
  DBuf<int32_t> bstart;
  DBuf<uint16_t> dmask;             // per cell: bit b set when local dof b is Dirichlet
  DBuf<unsigned char> elem;         // per cell: the element matrix or metric record (FP32/FP64)
  DBuf<double> fdet;                // full-row simplices: f |det DF| per cell; hexes: Σ w|J| φ_a (C x 27)
  double nu = 0.3;                  // elasticity: Poisson ratio
  double body[3] = {0, 0, 0};       // elasticity: body force
  DBuf<double> sref;                // reference tensor Ĝ (6 x nloc x nloc) and m (nloc)
  bool sref_ready = false;          // sref matches the uploaded tabulation

  // ---- vector Q2 on a Cartesian mesh, owner-computes (no element matrices) ----
  int last_kernel = 0;              // GH_KERNEL_* of the last numeric phase
  bool q2lattice = false;           // the lattice kernel assembles this space
  DBuf<int32_t> q2_node_dof;        // (2n+1)^3 lattice nodes x 3: signed dof ids
  DBuf<double> q2_ref;              // K̂ (81 x 81, by column) and m̂ (27), FP64
  DBuf<int32_t> q2_rec;             // per templated node: {node, template, dofs, column starts}
  int64_t q2_nrec = 0;
  DBuf<unsigned char> q2_cls_hdr;   // per template: Q2ClassHdr
  DBuf<int32_t> q2_cls_pos;         // per template, cell, lane: slots of rows lane, 64 + lane
  DBuf<int16_t> q2_cls_code;        // per template, slot: neighbour offset code * 3 + comp
  DBuf<int32_t> q2_irregular;       // lattice nodes without a template (ranking kernel)
  int64_t q2_nirregular = 0;
  DBuf<float> q2_ref32;             // K̂ in FP32
  bool q2_ref_ready = false;
  double q2_ref_key[4] = {0, 0, 0, 0};  // ν and cell sizes of q2_ref
  DBuf<unsigned char> qelem;        // mapped lattice: per-cell stiffness triangle and load (SoA)
  DBuf<double> evec;                // general n-cube element lists: per cell its load entries
  DBuf<double> gtab;                // geometry map gradients ∇̂N_v at the tabulation's points
  bool have_gtab = false;
  int gtab_nq = 0;

  // ---- colouring for the general path ----
  int ncolors = 0;
  std::vector<int64_t> color_ptr;   // ncolors+1
  DBuf<int32_t> color_cells;        // cells grouped by colour, ascending inside a colour

  // ---- nonlinear forms: the FE function uh (free values; Dirichlet values in dvals) ----
  DBuf<double> ustate;
  bool have_state = false;
  double plap_p = 2.0;              // p of GH_FORM_PLAPLACIAN
  DBuf<double> nl_tab;              // packed tabulation of the nonlinear kernels

  // ---- boundary faces (BoundaryTriangulation) and their terms ----
  int64_t nbfaces = 0;
  std::vector<int32_t> h_face_cell; // 0-based cell of each face (host copy, owner lists)
  DBuf<int32_t> bface_cell, bface_side;  // 0-based cell and n-cube facet of each face
  int nqf = 0;                      // quadrature points per face
  bool bface_tab_ok = false;        // face tabulation matches the element
  DBuf<double> bface_tab;           // per side, per point: w, shp[nloc], grad[nloc*dim]
  DBuf<double> balpha, bbeta, bg, bh;
  bool have_balpha = false, have_bbeta = false, have_bg = false, have_bh = false;
  // face assembly lists (boundary_lists; built once per face upload and pattern):
  // destinations = (column j, row i) entries of free dofs shared by a face's cell, each
  // with its face-matrix contributions; owners = free dofs of boundary cells, each with
  // its face loads
  int64_t bdst_n = -1;              // destinations (-1: lists not built)
  int64_t bown_n = 0;               // owners
  DBuf<int32_t> bdst_ptr, bdst_ci, bdst_row, bdst_col, bdst_pos;
  DBuf<int32_t> bown_dof, bown_ptr, bown_fa;
  DBuf<double> bface_K, bface_r;    // face matrices (upper triangle) and loads, [k][face]

  // ---- coefficients / values ----
  DBuf<double> kappa, fsrc, dvals;
  bool have_kappa = false, have_fsrc = false, have_dvals = false;
  DBuf<double> sigma;               // per cell: mass (reaction) coefficient of ∫ σ u v
  bool have_sigma = false;
  DBuf<double> fq;                  // f at the quadrature points, cells x nq (general path)
  bool have_fq = false;
  bool have_values = false;
  DBuf<double> nzval, b;

  // ---- solver scratch ----
  DBuf<double> s_r, s_z, s_p, s_q, s_x, s_dinv, s_b, s_part, s_scal, s_y;
  DBuf<double> s_r2, s_p2;          // second buffers of r (factored M) and p
  // box lattice: the free dofs are all nodes of a box of box_m[0] x box_m[1] x box_m[2]
  // lattice nodes, numbered in node order (Q1, every boundary node outside the box
  // Dirichlet), so A is a 27-point stencil in dof space and PCG runs the stencil
  // kernels: A as 27 offset arrays of N values (a27, zero outside the box) and G as
  // 14 offset arrays (g14, lower offsets in ascending order), no column indices
  bool box = false;
  int64_t box_m[3] = {0, 0, 0};
  DBuf<double> a27;
  DBuf<float> g14;

  // ---- learned FSAI (lattice Q1) ----
  bool fsai_ready = false;   // structures allocated for the current pattern
  bool fsai_built = false;   // G holds the factor of the current matrix
  DBuf<int32_t> diagpos;     // N: nzval position of each column's diagonal
  DBuf<double> fs;           // N: diag(A)^(-1/2)
  DBuf<float> gl, gu;        // NNZ: Ĝ in column-prefix (rows of G) and suffix (rows of Gᵀ) layout
  DBuf<float> gG;            // NNZ: loss gradient w.r.t. Ĝ, prefix layout
  DBuf<float> gle, gue;      // N x 16: rows of G = Ĝ D^-1/2 and of Gᵀ (solver copy)
  DBuf<int32_t> gli, gui;    // N x 16: their column indices (padding: the row itself)
  DBuf<float> feat;          // C x 16 (10 used)
  DBuf<float> z1, h1, z2, h2, z3, h3;  // C x 64 pre-/post-activations
  DBuf<float> yo, dy;        // C x 40 (36 used): network output and its gradient
  DBuf<float> dz1, dz2, dz3; // C x 64
  DBuf<float> params, grads, adam_m, adam_v;  // NPARAM, FP32 master copy and Adam state
  DBuf<uint16_t> params_bf16;                 // the weights the kernels use
  DBuf<float> wpart;         // per-block weight-gradient partials
  DBuf<double> pz, pw, pv, pu, pr, pt, pq, lpart;  // probe vectors, loss partials
  int adam_t = 0;
  bool have_mlp = false;
};

// pattern.hip
// contributions: also keep, per nonzero, the ascending list of the (cell, a, b)
// element entries that sum into it (cstart, perm)
void symbolic_general(Ctx &c, bool contributions);
void check_structured(Ctx &c);
// the pattern of a Cartesian scalar-Q1 space written directly (no sort); false when the
// numbering is not CLagrangian node order (then symbolic_general + check_structured)
bool symbolic_lattice(Ctx &c);
void build_coloring(Ctx &c);
// assemble.hip
void assemble_diffusion(Ctx &c, int precision, int path);
// simplex.hip
bool simplex_eligible(const Ctx &c);
bool hexgen_eligible(const Ctx &c);      // scalar Q2 / vector Q1, Q2 n-cubes for hexgen
void hexgen_assemble(Ctx &c, int precision, bool elasticity);
// b[j] = Σ over j's adjacency (ascending cells) of the cell loads in c.evec
// ([cell][local], FP64); enqueued on c.stream.
void hexgen_rhs(Ctx &c);
void rect_symbolic(Ctx &c);                  // pattern of a rectangular field pair
void rect_assemble(Ctx &c, int form, int precision);  // its values and lifting loads
// one composite block's current nzval moved to its positions in c.comp_nz, enqueued on
// the block's stream
void compose_move(Ctx &c, int block);
void elemlist_symbolic(Ctx &c);         // dof -> cell adjacency, Dirichlet masks
void elemlist_sum(Ctx &c, bool f32);     // nzval from the element entries
void assemble_simplex(Ctx &c, int precision);
// boundary.hip
void assemble_boundary(Ctx &c);  // adds the boundary-face terms to nzval and b
// nonlinear.hip
void assemble_plaplacian(Ctx &c, int precision, int path);  // residual -> b, Jacobian -> nzval
// Owner-computes column kernel of the Q1 lattice over the element matrices in c.qelem
// (upper triangle, FP32 when f32, else FP64; loads in qelem's load slots, or FP64 in
// eload); lift: subtract the Dirichlet lifting from b.  Enqueued on c.stream.
void launch_lattice_columns(Ctx &c, bool f32, const double *eload, bool lift);

// hexq2.hip
bool hexq2_eligible(const Ctx &c);
void assemble_elasticity(Ctx &c, int precision, int path);
// hexq2_lattice.hip
void q2lattice_prepare(Ctx &c);                 // symbolic phase: node map, pattern check
// numeric phase: nzval and b; vlap: the vector Laplacian ∫ κ ∇v⊙∇u instead of elasticity
void q2lattice_assemble(Ctx &c, bool f32, bool vlap = false);
// solve.hip
void spmv_device(Ctx &c, const double *x, double *y);
void pcg_device(Ctx &c, int precond, double rtol, int maxit, gh_solve_stats *st);
// the allocation-free solve: vectors in the library's fixed device array
void scratch_init(int device);
void pcg_noalloc(Ctx &c, int precond, const double *b, double *x, double rtol, int maxit,
                 gh_solve_stats *st);
// fsai.hip
constexpr int kFsaiEllSlots = 16;  // slots per row of the solver's G and Gᵀ
constexpr int kMlpIn = 10, kMlpInPad = 16, kMlpHidden = 64, kMlpOut = 36, kMlpOutPad = 40;
constexpr int kMlpW1 = 0, kMlpB1 = 640, kMlpW2 = 704, kMlpB2 = 4800, kMlpW3 = 4864,
              kMlpB3 = 8960, kMlpW4 = 9024, kMlpB4 = 11328, kMlpParams = 11364;
void fsai_prepare(Ctx &c);                       // diag positions, D^-1/2, buffers
// features -> MLP -> Ĝ; inference: the BF16 matrix-core forward (no stored
// activations), else the FP32 forward that keeps the activations for training
void fsai_forward_build(Ctx &c, bool inference);
double fsai_train_step(Ctx &c, int nprobes, uint64_t seed, float lr);
void fsai_apply_lower(Ctx &c, const double *r, double *y);   // y = Ĝ (D^-1/2 r)
void mlp_sync_bf16(Ctx &c);
void grf_kappa(Ctx &c, int nmodes, const double *omega, const double *phase, double amp);

}  // namespace gh
