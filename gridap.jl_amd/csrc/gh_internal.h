// Internal declarations of the gfx950 assembly engine (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "gridap_hip.h"

namespace gh {

// Lattice nodes per tile of the lattice kernel (one wavefront).
constexpr int kTileNodes = 64;

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
  T *get() const { return p; }
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // ---- mesh ----
  int dim = 0;
  bool cartesian = false;
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

  // ---- colouring for the general path ----
  int ncolors = 0;
  std::vector<int64_t> color_ptr;   // ncolors+1
  DBuf<int32_t> color_cells;        // cells grouped by colour, ascending inside a colour

  // ---- coefficients / values ----
  DBuf<double> kappa, fsrc, dvals;
  bool have_kappa = false, have_fsrc = false, have_dvals = false;
  bool have_values = false;
  DBuf<double> nzval, b;

  // ---- solver scratch ----
  DBuf<double> s_r, s_z, s_p, s_q, s_x, s_dinv, s_b, s_part, s_scal;
};

// pattern.hip
void symbolic_general(Ctx &c);
void check_structured(Ctx &c);
void build_coloring(Ctx &c);
// assemble.hip
void assemble_diffusion(Ctx &c, int precision, int path);
// solve.hip
void spmv_device(Ctx &c, const double *x, double *y);
void pcg_device(Ctx &c, int precond, double rtol, int maxit, gh_solve_stats *st);

}  // namespace gh
