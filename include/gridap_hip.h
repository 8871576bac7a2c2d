/*
 * gridap_hip.h -- C ABI of the MI355X finite-element assembly engine.
 *
 * One opaque context per device holds the mesh, the FE space (cell dof ids), the
 * CSC pattern and the assembled values in HBM.  The host (Gridap's Julia side
 * through `ccall`, or any FFI) hands over Gridap's own data and gets back a
 * SparseMatrixCSC{Float64,Int64} (1-based colptr/rowval) and a Vector{Float64}.
 *
 * Conventions (Gridap's): cell tables are Table{Int32} (data + 1-based ptrs,
 * src/Arrays/Tables.jl:23-30); dof ids are 1-based, free > 0, Dirichlet < 0
 * (src/FESpaces/CLagrangianFESpaces.jl:155-207); matrices are CSC with 1-based
 * Int64 colptr/rowval (src/FESpaces/SparseMatrixAssemblers.jl:155-160).
 *
 * Every entry returns GH_OK (0) or an error code; gh_last_error() gives the
 * message.  No C++ exception crosses the ABI.  Host pointers are read or written
 * only during the call.  Calls on one context are serialised by the caller
 * (Gridap assembles from one task); the library uses one HIP stream per context.
 */
#ifndef GRIDAP_HIP_H
#define GRIDAP_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define GH_OK 0
#define GH_ERR_ARG 1         /* bad argument / size mismatch (the @assert checks of
                                src/FESpaces/SparseMatrixAssemblers.jl:219-220)       */
#define GH_ERR_RANGE 2       /* dof id out of range                                   */
#define GH_ERR_HIP 3         /* HIP runtime error                                     */
#define GH_ERR_OOM 4         /* device allocation failed                              */
#define GH_ERR_STATE 5       /* call out of order (e.g. numeric before symbolic)      */
#define GH_ERR_NOCONV 6      /* iterative solver did not reach the tolerance          */
#define GH_ERR_UNSUPPORTED 7 /* reference element / form not provided by the engine  */

/* ---- reference elements (Lagrangian, src/ReferenceFEs/CLagrangianRefFEs.jl) ---- */
#define GH_REFFE_QK 1 /* n-cube Q_k (QUAD/HEX), k = order */
#define GH_REFFE_PK 2 /* simplex P_k (TRI/TET)             */

/* ---- weak forms ---- */
#define GH_FORM_DIFFUSION 1 /* a(u,v)=∫ κ ∇v⋅∇u dΩ,  l(v)=∫ f v dΩ, κ and f constant per cell
                             * (or f at the quadrature points, gh_source_qp_upload); scalar
                             * Q1, P1, P2, and Q2 on quads/hexes; on vector Q1/Q2 spaces of
                             * the general n-cube path the vector Laplacian ∫ κ ∇v⊙∇u with
                             * l(v)=∫ v⋅f, f the constant body force of gh_elasticity_params */
#define GH_FORM_ELASTICITY 2 /* a(u,v)=∫ ε(v) : σ(ε(u)) dΩ, σ = λ tr(ε) I + 2 μ ε with λ, μ from a
                              * per-cell Young's modulus (the κ array) and ν; l(v)=∫ v⋅f dΩ with f
                              * constant (gh_elasticity_params); vector Q1 on quads/hexes and
                              * vector Q2 on hexes, Cartesian or mapped */
#define GH_FORM_PLAPLACIAN 3 /* residual and Jacobian at uh (gh_fefunction_upload) of
                              * res(u,v) = ∫ κ ∇v⋅(|∇u|^(p-2)∇u) - f v dΩ, jac(u,du,v) = ∫ κ ∇v⋅
                              * ((p-2)|∇u|^(p-4)(∇u⋅∇du)∇u + |∇u|^(p-2)∇du) dΩ, κ (NULL = 1) and
                              * f (NULL = 0) per cell from gh_coef_upload, p >= 2
                              * (gh_plaplacian_params); gh_assemble writes the residual to b and
                              * the Jacobian to nzval; scalar Q1 on quads/hexes.  Lattice
                              * meshes take the owner-computes column route unless path is
                              * GH_PATH_GENERAL (the coloured per-cell scatter) */
#define GH_FORM_DIVERGENCE 4 /* rectangular pairs (gh_trial_space_upload) between a scalar
                              * and a vector space (one component per dimension):
                              * scalar test ψ, vector trial u: ∫ κ ψ (∇⋅u) dΩ; vector test v,
                              * scalar trial φ: ∫ κ (∇⋅v) φ dΩ (the q (∇⋅u) and -(∇⋅v) p
                              * blocks of test/GridapTests/StokesTaylorHoodTests.jl:54, with
                              * κ = 1 and -1); κ per cell from gh_coef_upload (NULL = 1) */

/* ---- element precision ---- */
#define GH_PREC_FP64 0      /* element arithmetic and accumulation in FP64           */
#define GH_PREC_FP32_ELEM 1 /* element arithmetic in FP32, nzval/b accumulated FP64  */

/* ---- numeric kernel choice ---- */
#define GH_PATH_AUTO 0       /* owner-computes lattice kernel when the space allows it */
#define GH_PATH_GENERAL 1    /* per-cell element kernel + coloured deterministic scatter */
#define GH_PATH_STRUCTURED 2 /* owner-computes lattice kernel (error if not eligible)  */

/* ---- preconditioners ---- */
/* Numeric kernels (gh_pattern_info.last_kernel): which form assembled the values. */
#define GH_KERNEL_NONE 0       /* no numeric phase since the pattern was built          */
#define GH_KERNEL_Q1_LATTICE 1 /* owner-computes scalar Q1 lattice kernel               */
#define GH_KERNEL_CELL 2       /* per-cell kernel + coloured deterministic scatter      */
#define GH_KERNEL_COLUMN_SUM 3 /* element lists: column ranges with 16-bit lists        */
#define GH_KERNEL_LIST_SUM 4   /* element lists: int32 lists, one thread per nonzero (with
                                  the general n-cube element kernel: scalar Q2, vector Q1,
                                  mapped vector Q2)                                     */
#define GH_KERNEL_Q2_LATTICE 5 /* owner-computes vector-Q2 lattice kernel               */

#define GH_PRECOND_NONE 0
#define GH_PRECOND_JACOBI 1
#define GH_PRECOND_FSAI 2 /* learned factorized sparse approximate inverse, M = GᵀG;
                             Jacobi-PCG is rerun from x0 if it fails to converge  */

typedef struct gh_ctx gh_ctx;

typedef struct gh_solve_stats {
  int32_t iterations;
  int32_t converged;   /* 1 when ||b - A x|| / ||b|| <= rtol (true residual, FP64) */
  double rel_residual; /* true residual at exit                                  */
  double device_ms;    /* solver wall time on the device stream                  */
  int32_t precond_used; /* GH_PRECOND_* that produced x (FSAI falls back to JACOBI) */
  int32_t reserved;
} gh_solve_stats;

/* Context.  Replaces nothing in Gridap (it is the state a GPU assembler holds). */
int gh_create(gh_ctx **ctx, int device);
int gh_destroy(gh_ctx *ctx);
const char *gh_last_error(const gh_ctx *ctx);
/* The HIP stream (hipStream_t) every kernel of this context runs on. */
int gh_stream(gh_ctx *ctx, void **stream);

/* Mesh from a CartesianDescriptor (origin, sizes, partition): analytic node
 * coordinates, CartesianCellNodes and CartesianMap Jacobian.
 * Replaces CartesianGrid/CartesianMap evaluation, src/Geometry/CartesianGrids.jl:107-283. */
int gh_mesh_cartesian(gh_ctx *ctx, int dim, const double *origin, const double *sizes,
                      const int64_t *partition);

/* Unstructured mesh: node coordinates (dim x nnodes, Julia Point layout, i.e. the
 * coordinates of node n at xyz[dim*n ..]) and cell_node_ids Table (1-based).
 * Replaces UnstructuredGrid's cell map, src/Geometry/UnstructuredGrids.jl:52-100. */
int gh_mesh_upload(gh_ctx *ctx, int dim, int64_t nnodes, const double *xyz, int64_t ncells,
                   const int32_t *cell_nodes_data, const int32_t *cell_nodes_ptrs);

/* FE space: get_cell_dof_ids(V) as a Table{Int32} plus num_free_dofs and
 * num_dirichlet_dofs.  Replaces the rows/cols inputs of the cell loops,
 * src/FESpaces/SparseMatrixAssemblers.jl:174-248 (cell_rows = cell_cols, U = V). */
int gh_space_upload(gh_ctx *ctx, int reffe_kind, int order, int ncomp,
                    const int32_t *cell_dofs_data, const int32_t *cell_dofs_ptrs,
                    int64_t ncells, int64_t nfree, int64_t ndiri);

/* Trial space of a rectangular field pair: the columns, while gh_space_upload's space
 * gives the rows (the block (i, j) of a multi-field form whose test field i and trial
 * field j live in different spaces on one mesh; the cols of
 * SparseMatrixAssemblers.jl:174-248 when cell_cols differ from cell_rows, and
 * src/MultiField/BlockSparseMatrixAssemblers.jl:19-36).  Scalar Q1/Q2 spaces on
 * quadrilaterals and hexahedra.  gh_symbolic then builds the (test nfree) x (trial
 * nfree) pattern (empty columns allowed), gh_coef_upload's Dirichlet values are the
 * trial space's, and gh_assemble (GH_FORM_DIFFUSION) writes
 *   B_c[lj][li] = ∫ κ ∇ψ_li⋅∇φ_lj + σ ψ_li φ_lj   (κ: gh_coef_upload, σ: gh_mass_upload)
 * into nzval and the lifting -B u_D into b.  Uploading a test space again ends the pair. */
int gh_trial_space_upload(gh_ctx *ctx, int reffe_kind, int order, int ncomp,
                          const int32_t *cell_dofs_data, const int32_t *cell_dofs_ptrs,
                          int64_t ncells, int64_t nfree, int64_t ndiri);

/* The trial basis at the points of the test tabulation (gh_reffe_tabulation): values
 * shp[q*nloc + a] and reference gradients grad[(q*nloc + a)*dim + d] (the trial half
 * of the cell field pair that IntegrationMap evaluates, src/Fields/FieldsInterfaces.jl:
 * 737-760). */
int gh_trial_tabulation(gh_ctx *ctx, int nq, int nloc, const double *shp, const double *grad);

/* Reference tabulation, computed once per reference cell by the host:
 * quadrature weights qw[nq], shape values shp[nq*nloc] and reference gradients
 * grad[nq*nloc*dim] (point-major).  Replaces the per-cell-type evaluation of
 * src/CellData/CellQuadratures.jl:138-160 with src/Arrays/CompressedArrays.jl:66-94. */
int gh_reffe_tabulation(gh_ctx *ctx, int nq, int nloc, const double *qw, const double *shp,
                        const double *grad);

/* Gradients of the geometry map's shape functions (the Q1 basis of the cell's nv
 * vertices) at the points of the reference tabulation, grad[nq*nv*dim] (point-major),
 * for meshes uploaded with their nodes when the FE basis is not the geometry basis
 * (Q2 spaces on mapped hexahedra).  Replaces the evaluation of the cell map's gradient,
 * get_cell_map (src/Geometry/Grids.jl:234-241), at the quadrature points; cleared by the
 * next gh_reffe_tabulation. */
int gh_geometry_tabulation(gh_ctx *ctx, int nq, int nv, const double *grad);

/* Boundary integrals (Cartesian meshes; scalar Q1, and scalar Q2 on quadrilaterals).
 * The reference tabulation at the facet quadrature points, per facet (side) of the
 * n-cube in Gridap's face order
 * (QUAD: y=0, y=1, x=0, x=1; HEX: z=0, z=1, y=0, y=1, x=0, x=1): facet weights qw[nqf]
 * (of the reference facet), cell shape values shp[nsides*nqf*nloc] and reference
 * gradients grad[nsides*nqf*nloc*dim].  Replaces the face-to-cell evaluation of a
 * BoundaryTriangulation measure, src/Geometry/BoundaryTriangulations.jl:122-183. */
int gh_face_tabulation(gh_ctx *ctx, int nsides, int nqf, const double *qw, const double *shp,
                       const double *grad);

/* The boundary faces and the terms integrated over them, added by gh_assemble
 * (GH_FORM_DIFFUSION) into the same nzval and b: face f is facet face_sides[f]
 * (1-based) of cell face_cells[f] (1-based) with outward normal n, and
 *   A += ∫_F α φ_a φ_b + β (φ_a n⋅∇φ_b + n⋅∇φ_a φ_b),  b += ∫_F φ_a g + n⋅∇φ_a h
 * with α[f], β[f] per face (NULL = 0) and g, h at the face quadrature points,
 * gq[f*nqf + q], hq[f*nqf + q] (NULL = 0), followed by the Dirichlet lifting of the face
 * matrix.  Neumann: g = n⋅∇u; Nitsche (test/GridapTests/PoissonTests.jl:101-108):
 * α = γ/h, β = -1, g = (γ/h) u_D, h = -u_D.  nfaces = 0 removes the boundary terms.
 * Vector spaces (one component per dimension) take the terms component by component
 * (v⊙u, v⊙(n⋅∇u), (n⋅∇v)⊙u: the scalar face matrix on every component) with g and h
 * per component, gq[(f*nqf + q)*ncomp + i] and hq alike (PoissonTests.jl's vector case;
 * the loads v⋅(n⋅∇u) - (n⋅v) p of test/GridapTests/StokesTaylorHoodTests.jl:56); they
 * are added after the cell terms of GH_FORM_DIFFUSION or GH_FORM_ELASTICITY.
 * The first gh_assemble after an upload (or a new pattern) builds the face lists: for
 * each nonzero the faces touch, its face contributions (faces ascending). */
int gh_boundary_upload(gh_ctx *ctx, int64_t nfaces, const int32_t *face_cells,
                       const int32_t *face_sides, const double *alpha, const double *beta,
                       const double *gq, const double *hq);

/* Symbolic phase: counter -> allocation -> sorted unique rows per column, built on
 * the device.  nnz_out = stored entries of create_from_nz.
 * Replaces symbolic_loop_matrix! + nz_allocation + create_from_nz pattern,
 * src/FESpaces/SparseMatrixAssemblers.jl:174-210, src/Algebra/SparseMatrixCSC.jl:214-283. */
int gh_symbolic(gh_ctx *ctx, int64_t *nnz_out);

/* 1-based Int64 colptr[nfree+1] and rowval[nnz], bit-identical to the
 * SparseMatrixCSC that create_from_nz returns. */
int gh_pattern_get(gh_ctx *ctx, int64_t *colptr, int64_t *rowval);

/* Per-cell coefficients resident on the device: kappa[ncells] (NULL = 1) and
 * f[ncells] (NULL = 1) for GH_FORM_DIFFUSION; Dirichlet values[ndiri] of uhd
 * (NULL = homogeneous).  These are the CellField inputs of the weak form. */
int gh_coef_upload(gh_ctx *ctx, const double *kappa, const double *fsrc,
                   const double *dirichlet_vals);

/* Per-cell mass (reaction) coefficient σ of the diffusion form: a(u,v) += ∫ σ u v dΩ
 * (NULL: no mass term).  The u⋅v terms of the multi-field forms of
 * test/MultiFieldTests/BlockSparseMatrixAssemblersTests.jl:101-104 (∫ u2⋅v2 - u1⋅v2).
 * Scalar Q1; the assembly then runs the per-cell kernel. */
int gh_mass_upload(gh_ctx *ctx, const double *sigma);

/* A source f(x) given at the quadrature points of every cell, fq[cell*nq + q] (the
 * rule of the uploaded tabulation), in place of the per-cell f of gh_coef_upload
 * (NULL: back to the per-cell f).  The x-dependent load of
 * test/GridapTests/PoissonTests.jl:101-108 (l(v) = ∫ v f dΩ with f a function).
 * Scalar Q1 and Q2, and vector spaces of the general n-cube element kernel with
 * fq[(cell*nq + q)*ncomp + i] (l(v) = ∫ v⋅f dΩ). */
int gh_source_qp_upload(gh_ctx *ctx, const double *fq);

/* Numeric phase on device-resident inputs: nzval and b in HBM, the matrix part of
 * assemble_matrix_and_vector! with attach_dirichlet lifting.  The work is enqueued
 * on the context stream; gh_values_get / gh_spmv / gh_pcg wait for it.
 * Replaces numeric_loop_matrix_and_vector! + InserterCSC.add_entry!,
 * src/FESpaces/SparseMatrixAssemblers.jl:365-405, src/Algebra/SparseMatrixCSC.jl:124-150,
 * src/CellData/AttachDirichlet.jl:59-84. */
int gh_assemble(gh_ctx *ctx, int form_kind, int precision, int path);

/* Material and load of GH_FORM_ELASTICITY: Poisson ratio ν in (-1, 0.5) and the body
 * force f[dim] (NULL = 0).  The weak form of Gridap's elasticity drivers
 * (ε(v) ⊙ (σ∘ε(u)), src/TensorValues/ and test/GridapTests/IsotropicDamageTests.jl:13-17). */
int gh_elasticity_params(gh_ctx *ctx, double nu, const double *body_force);

/* The FE function uh at which a nonlinear form's residual and Jacobian are assembled:
 * its free values[nfree] and Dirichlet values[ndiri] (NULL = 0), as
 * get_free_dof_values(uh) / get_dirichlet_dof_values(uh) of a Gridap FEFunction.
 * Replaces the uh argument of residual!/jacobian!(…, op, uh),
 * src/FESpaces/FEOperatorsFromWeakForm.jl:57-95. */
int gh_fefunction_upload(gh_ctx *ctx, const double *free_vals, const double *dirichlet_vals);

/* Exponent p >= 2 of GH_FORM_PLAPLACIAN (test/GridapTests/PLaplacianTests.jl:15-17). */
int gh_plaplacian_params(gh_ctx *ctx, double p);

/* Copy the assembled nzval[nnz] and b[nfree] to the host (either may be NULL). */
int gh_values_get(gh_ctx *ctx, double *nzval, double *b);

/* Composite matrix of a multi-field form on the device: the matrix that
 * src/MultiField/BlockSparseMatrixAssemblers.jl:19-36 (and the consecutive multi-field
 * layout of src/MultiField/MultiFieldFESpaces.jl) builds from the field-pair blocks.
 * ctx holds the composite's nnz values; gh_compose_block registers another context
 * src (same device, pattern built) whose k-th stored value goes to composite position
 * dst[k] (0-based, k < n = src's nnz).  Positions must be distinct across all blocks;
 * those no block covers read as stored zeros.  nnz <= 2^31 - 1.  src must stay alive
 * while it is registered.
 * Blocks are numbered 0, 1, ... in registration order.  gh_compose_move(block) moves
 * the block's current assembled values to their positions on the device, ordered
 * after the work its context has enqueued and before what it enqueues next (so one
 * context may serve several blocks, reassembled between their moves).
 * gh_compose_values_get waits for the moves and copies the nnz composite values to the
 * host.  gh_compose_clear drops the blocks. */
int gh_compose_block(gh_ctx *ctx, gh_ctx *src, const int64_t *dst, int64_t n, int64_t nnz);
int gh_compose_move(gh_ctx *ctx, int block);
int gh_compose_values_get(gh_ctx *ctx, double *nzval);
int gh_compose_clear(gh_ctx *ctx);

/* Load a matrix with this context's pattern (nzval[nnz], the stored values of a
 * SparseMatrixCSC whose colptr/rowval are gh_pattern_get's) and/or a right-hand side
 * b[nfree] into the device, in place of the assembled ones; either may be NULL.  The
 * matrix argument of solve!(x, solver, A, b) / numerical_setup(ss, A),
 * src/Algebra/LinearSolvers.jl:202-274: gh_spmv and gh_pcg then work on this A. */
int gh_values_set(gh_ctx *ctx, const double *nzval, const double *b);

/* Which numeric form the symbolic phase chose for the uploaded space (diagnostic;
 * no counterpart in Gridap). */
typedef struct gh_pattern_info {
  int64_t nnz;
  int64_t nranges;     /* column ranges of the element-list column sum (0: not used)   */
  int32_t lattice;     /* 1: owner-computes lattice kernel (Cartesian scalar Q1, or
                          vector Q2 elasticity without element matrices)            */
  int32_t box;         /* 1: free dofs form a box of lattice nodes (stencil PCG)       */
  int32_t elemlist;    /* 1: element-list path (simplices, vector Q2 hexahedra)        */
  int32_t compact;     /* 1: compact 16-bit contribution lists (column-range sum);
                          0 on the element-list path: int32 lists, one thread per nonzero */
  int32_t metric;      /* 1: simplex metric records (entries formed in the column sum) */
  int32_t inverse;     /* 1: full-column staging (vector Q2 element matrices)          */
  int32_t last_kernel; /* GH_KERNEL_* of the last numeric phase on this pattern          */
  int32_t fallback_nodes; /* vector-Q2 lattice: nodes without a column template, which the
                          ranking kernel assembles (-1: templates switched off)      */
} gh_pattern_info;
int gh_pattern_info_get(gh_ctx *ctx, gh_pattern_info *info);

/* One-call numeric phase from host arrays: gh_coef_upload(cell_coef, cell_source,
 * dirichlet_vals) + gh_assemble(form_kind, precision, GH_PATH_AUTO) + gh_values_get.
 * The quadrature is the one of the uploaded tabulation (gh_reffe_tabulation). */
int gh_numeric(gh_ctx *ctx, int form_kind, const double *cell_coef, const double *cell_source,
               int precision, double *nzval, double *b, const double *dirichlet_vals);

/* y = A x with the assembled matrix (host vectors of length nfree).  The kernels read
 * the CSC columns as rows, i.e. they compute Aᵀx, which is A x for the symmetric
 * matrices of these forms: every numeric path except the vector-Q2 element path
 * (GH_PATH_GENERAL, whose FP32/FP64 matrix-core products round the two entries of a
 * pair differently in the last bits) stores A symmetric bit for bit.
 * Replaces the SpMV of residual!, src/Algebra/LinearSolvers.jl:36-58. */
int gh_spmv(gh_ctx *ctx, const double *x, double *y);

/* Preconditioned CG in FP64 on the assembled (SPD, see gh_spmv) matrix; b NULL = the assembled
 * b.  x holds the initial guess on entry and the solution on exit.
 * Stands in for solve!(x, LUSolver(), op), src/Algebra/LinearSolvers.jl:202-274. */
int gh_pcg(gh_ctx *ctx, int precond, const double *b, double *x, double rtol, int maxit,
           gh_solve_stats *stats);

/* gh_pcg without any device allocation.  No counterpart in Gridap: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Jacobi (GH_PRECOND_JACOBI) or no preconditioner,
 * the general CSR kernels, nfree <= GH_NOALLOC_MAX_N.  Its vectors live in one fixed
 * device array of the library (8 x GH_NOALLOC_MAX_N doubles plus partial sums, about
 * 17 MB), which every GPU the library runs on holds until the process ends; all calls
 * on one GPU share it, so calls that overlap on one GPU race.  Same results as gh_pcg
 * with the same preconditioner on the CSR kernels (GRIDAP_HIP_STENCIL_PCG=0);
 * stats->device_ms is the host wall time of the call. */
#define GH_NOALLOC_MAX_N 262144
int gh_pcg_noalloc(gh_ctx *ctx, int precond, const double *b, double *x, double rtol,
                   int maxit, gh_solve_stats *stats);

/* ---- learned FSAI preconditioner (3-D Cartesian mesh, scalar Q1) ----
 * No counterpart in Gridap (it ships LU only, src/Algebra/LinearSolvers.jl:249-303).
 * G = Ĝ D^(-1/2) on the pattern of tril(A); Ĝ is assembled from the outputs of a
 * per-cell MLP (10 -> 64 -> 64 -> 64 -> 36, GELU, BF16 weights, FP32 accumulate). */

/* Number of network parameters (flat FP32 layout W1,b1,W2,b2,W3,b3,W4,b4). */
int gh_mlp_nparams(int64_t *n);
/* Seeded initialisation (Xavier-uniform hidden layers, zero last layer: G = Jacobi). */
int gh_mlp_init(gh_ctx *ctx, uint64_t seed);
/* Set / get the FP32 master weights; setting resets the optimiser state. */
int gh_mlp_set(gh_ctx *ctx, const float *params, int64_t n);
int gh_mlp_get(gh_ctx *ctx, float *params, int64_t n);
/* Per-cell κ = exp(amp Σ_k cos(ω_k·x_c + φ_k)) at the cell centres x_c, written to the
 * device coefficient (a Gaussian random field by random Fourier features). */
int gh_grf_kappa(gh_ctx *ctx, int nmodes, const double *omega, const double *phase,
                 double amp);
/* Features -> network -> Ĝ for the currently assembled matrix. */
int gh_fsai_build(gh_ctx *ctx);
/* One training step on the currently assembled matrix: nprobes Rademacher probes of
 * ||I - G A Gᵀ||_F^2 / N, back-propagation and one Adam step (lr = 0: gradient only).
 * loss receives the probe estimate before the step. */
int gh_fsai_train_step(gh_ctx *ctx, int nprobes, uint64_t seed, float lr, double *loss);
/* Diagnostic copies: features [C x 16], network output [C x 40], Ĝ in the column-prefix
 * and column-suffix layouts [nnz] (written by gh_fsai_train_step), last parameter
 * gradient [nparams] (any may be NULL). */
int gh_fsai_get(gh_ctx *ctx, float *feat, float *yout, float *gl, float *gu, float *grad);
/* The factor the solver applies, 16 slots per row [N x 16]: gle/gli = weights and
 * 0-based column indices of row i of G = Ĝ D^(-1/2) (columns j <= i), gue/gui = row i
 * of Gᵀ (entries G(j, i), j >= i); unused slots have weight 0 (any may be NULL). */
int gh_fsai_factor_get(gh_ctx *ctx, float *gle, int32_t *gli, float *gue, int32_t *gui);

#ifdef __cplusplus
}
#endif
#endif
