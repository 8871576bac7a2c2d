/*
 * gridap_oracle.h -- CPU FP64 restatement of Gridap.jl's assembly path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP assembly
 * engine: only tests, the smoke check and the CPU-baseline leg of the benchmark
 * may load it.  The product path never calls it and has no CPU fallback.
 *
 * Conventions follow Gridap: node/cell/dof ids are 1-based, free dofs are
 * positive, Dirichlet dofs negative (-1..-ndiri), CSC output is 1-based Int64.
 * Per-function reference citations are in gridap_oracle.c.
 */
#ifndef GRIDAP_ORACLE_H
#define GRIDAP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- quadrature (src/ReferenceFEs/TensorProductQuadratures.jl:68-79) ---- */
int gor_gauss_legendre01(int npts, double *x, double *w);
int gor_tensor_quadrature(int D, int degree, double *coords, double *weights);

/* ---- Lagrangian n-cube Q1/Q2 and simplex P1/P2 tabulation ---- */
int gor_qk_num_nodes(int D, int order);
int gor_qk_nodes(int D, int order, double *nodes);
int gor_qk_tabulate(int D, int order, int npts, const double *pts, double *shp,
                    double *grad);

/* ---- Cartesian model (src/Geometry/CartesianGrids.jl, CartesianDiscreteModels.jl) ---- */
int64_t gor_cart_num_nodes(int D, const int64_t *part);
int64_t gor_cart_num_cells(int D, const int64_t *part);
void gor_cart_node_coords(int D, const double *origin, const double *sizes,
                          const int64_t *part, double *xyz);
void gor_cart_cell_nodes(int D, const int64_t *part, int32_t *cell_nodes);
int gor_ncube_face_id(int D, int free_mask, int anchor_bits);
void gor_cart_vertex_labels(int D, const int64_t *part, int32_t *labels);

/* ---- CLagrangian numbering (src/FESpaces/CLagrangianFESpaces.jl:155-312) ---- */
void gor_face_tag_index(int64_t nfaces, const int32_t *face_entity, int ntags,
                        const int32_t *tag_entity_ptrs, const int32_t *tag_entities,
                        int8_t *face_to_tag);
int gor_clagrangian_numbering(int64_t nnodes, const int8_t *node_to_tag, int ncomp,
                              const uint8_t *tag_comp_mask, int32_t *node_comp_to_dof,
                              int64_t *nfree, int64_t *ndiri);
void gor_clagrangian_cell_dofs(int64_t ncells, int nnodes_cell, const int32_t *cell_nodes,
                               int ncomp, const int32_t *node_comp_to_dof,
                               int32_t *cell_dofs);

/* ---- element integration (src/CellData/CellQuadratures.jl:138-160,
 *      src/Fields/FieldsInterfaces.jl:737-760) ---- */
int gor_cell_matrices_diffusion(int D, int64_t ncells, int nloc, const int32_t *cell_nodes,
                                const double *xyz, int npts, const double *qw,
                                const double *shp_grad, const double *kappa,
                                double *cellmats);
int gor_cell_matrices_diffusion_cartesian(int D, const double *sizes, int64_t ncells,
                                          int nloc, int npts, const double *qw,
                                          const double *shp_grad, const double *kappa,
                                          double *cellmats);
int gor_cell_vectors_source(int D, int64_t ncells, int nloc, const int32_t *cell_nodes,
                            const double *xyz, int npts, const double *qw,
                            const double *shp, const double *shp_grad,
                            const double *fsrc, double *cellvecs);
int gor_cell_vectors_source_cartesian(int D, const double *sizes, int64_t ncells, int nloc,
                                      int npts, const double *qw, const double *shp,
                                      const double *fsrc, double *cellvecs);
int gor_cell_vectors_source_qp(int D, int64_t ncells, int nloc, const int32_t *cell_nodes,
                               const double *xyz, const double *cartesian_sizes, int npts,
                               const double *qw, const double *shp, const double *shp_grad,
                               const double *fq, double *cellvecs);
void gor_attach_dirichlet(int64_t ncells, int nloc, const int32_t *cell_dofs,
                          const double *cellmats, const double *dirichlet_vals,
                          double *cellvecs);

/* ---- MinMemory CSC builder (src/Algebra/SparseMatrixCSC.jl:72-283) ---- */
int64_t gor_csc_count(int64_t ncells, int nloc, const int32_t *cell_dofs, int64_t ncols,
                      int64_t *colnnzmax);
int64_t gor_csc_assemble(int64_t ncells, int nloc, const int32_t *cell_dofs,
                         const double *cellmats, int64_t nrows, int64_t ncols,
                         int64_t *colptr, int64_t *rowval, double *nzval);
int64_t gor_csc_from_entries(int64_t nent, const int64_t *I, const int64_t *J, const double *V,
                             const uint8_t *has_value, int64_t ncols, int64_t *colnnzmax_out,
                             int64_t *colptr, int64_t *rowval, double *nzval);
int gor_numeric_refill_cartesian(int D, const double *sizes, int64_t ncells, int nloc,
                                 const int32_t *cell_dofs, int npts, const double *qw,
                                 const double *shp, const double *shp_grad,
                                 const double *kappa, const double *fsrc, int64_t n,
                                 const int64_t *colptr, const int64_t *rowval, double *nzval,
                                 double *b);
/* gor_numeric_refill_cartesian on nthreads OpenMP threads (parity colouring of the
 * Cartesian cells; agrees with it to rounding).  For timing the CPU baseline. */
int gor_numeric_refill_cartesian_par(int D, const double *sizes, const int64_t *part,
                                     int nloc, const int32_t *cell_dofs, int npts,
                                     const double *qw, const double *shp,
                                     const double *shp_grad, const double *kappa,
                                     const double *fsrc, int64_t n, const int64_t *colptr,
                                     const int64_t *rowval, double *nzval, double *b,
                                     int nthreads);
void gor_assemble_vector(int64_t ncells, int nloc, const int32_t *cell_dofs,
                         const double *cellvecs, int64_t nrows, double *b);

/* ---- solver checks (no reference counterpart: Gridap ships only LU) ---- */
void gor_csc_spmv(int64_t n, const int64_t *colptr, const int64_t *rowval,
                  const double *nzval, const double *x, double *y);
int gor_pcg_jacobi(int64_t n, const int64_t *colptr, const int64_t *rowval,
                   const double *nzval, const double *b, double *x, double rtol,
                   int maxit, int *iters, double *relres);

#ifdef __cplusplus
}
#endif
#endif
