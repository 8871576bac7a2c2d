/*
 * gridap_oracle.c -- CPU FP64 restatement of Gridap.jl's matrix/vector assembly path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests, the
 * smoke check and the benchmark's CPU-baseline leg load it; nothing in the product
 * path links or calls it.
 *
 * Every function cites the Gridap source it restates (paths relative to the
 * Gridap.jl v0.20.8 tree).  Pinning of each function (fixtures, closed forms,
 * invariants) lives in tests/test_oracle_*.py.
 */
#include "gridap_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Quadrature                                                                */
/* ------------------------------------------------------------------------- */

/* Gauss-Legendre nodes/weights on [-1,1] by Newton iteration on P_n (the
 * classical method; QuadGK's gauss(T,n) returns the same rule, ascending), then
 * the affine map to [0,1] done by _tensor_product_legendre:
 * src/ReferenceFEs/TensorProductQuadratures.jl:68-79. */
int gor_gauss_legendre01(int n, double *x, double *w) {
  if (n < 1) return 1;
  /* Newton in extended precision so the rounded nodes/weights are (nearly always)
   * the correctly rounded values. */
  for (int i = 0; i < n; ++i) {
    long double z = -cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
    long double pn = 0, pnm1 = 0, pp = 0;
    for (int it = 0; it < 200; ++it) {
      long double pk = z, pkm1 = 1.0L;
      for (int k = 2; k <= n; ++k) {
        long double pkp1 = ((2.0L * k - 1.0L) * z * pk - (k - 1.0L) * pkm1) / k;
        pkm1 = pk;
        pk = pkp1;
      }
      pn = (n == 1) ? z : pk;
      pnm1 = (n == 1) ? 1.0L : pkm1;
      pp = n * (z * pn - pnm1) / (z * z - 1.0L);
      long double dz = pn / pp;
      z -= dz;
      if (fabsl(dz) < 1e-21L) break;
    }
    {
      long double pk = z, pkm1 = 1.0L;
      for (int k = 2; k <= n; ++k) {
        long double pkp1 = ((2.0L * k - 1.0L) * z * pk - (k - 1.0L) * pkm1) / k;
        pkm1 = pk;
        pk = pkp1;
      }
      pn = (n == 1) ? z : pk;
      pnm1 = (n == 1) ? 1.0L : pkm1;
      pp = n * (z * pn - pnm1) / (z * z - 1.0L);
    }
    long double wi = 2.0L / ((1.0L - z * z) * pp * pp);
    x[i] = (double)((z + 1.0L) / 2.0L);
    w[i] = (double)(wi / 2.0L);
  }
  /* symmetric rule: exact symmetry and the exact centre node */
  for (int i = 0; i < n / 2; ++i) {
    x[n - 1 - i] = 1.0 - x[i];
    w[n - 1 - i] = w[i];
  }
  if (n % 2 == 1) x[n / 2] = 0.5;
  return 0;
}

/* Tensor-product rule, x fastest (CartesianIndices order), weights multiplied
 * in axis order: src/ReferenceFEs/Quadratures.jl:191,203-222.
 * n points per axis = div(degree,2)+1 (src/ReferenceFEs/Quadratures.jl:191). */
int gor_tensor_quadrature(int D, int degree, double *coords, double *weights) {
  int n = degree / 2 + 1;
  double x1[64], w1[64];
  if (n > 64 || D < 1 || D > 3) return 1;
  gor_gauss_legendre01(n, x1, w1);
  int total = 1;
  for (int d = 0; d < D; ++d) total *= n;
  for (int k = 0; k < total; ++k) {
    int r = k;
    double w = 1.0;
    for (int d = 0; d < D; ++d) {
      int i = r % n;
      r /= n;
      coords[k * D + d] = x1[i];
      w *= w1[i];
    }
    weights[k] = w;
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* n-cube faces and Lagrangian nodes                                         */
/* ------------------------------------------------------------------------- */

static int binom(int n, int k) {
  int r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

/* 1-based id of the n-cube face with free axes `free_mask` (bit d = axis d) and
 * fixed coordinates `anchor_bits` (bit d = 1 when fixed axis d sits at 1).
 * Order of src/ReferenceFEs/ExtrusionPolytopes.jl:460-478: stable sorts by
 * anchor then extrusion, axis D most significant, then by dimension. */
int gor_ncube_face_id(int D, int free_mask, int anchor_bits) {
  int d = __builtin_popcount(free_mask);
  int offset = 0;
  for (int k = 0; k < d; ++k) offset += binom(D, k) << (D - k);
  int rank_ext = 0;
  for (int m = 0; m < free_mask; ++m)
    if (__builtin_popcount(m) == d) ++rank_ext;
  int a = 0, bit = 0;
  for (int ax = 0; ax < D; ++ax) {
    if (free_mask & (1 << ax)) continue;
    if (anchor_bits & (1 << ax)) a |= 1 << bit;
    ++bit;
  }
  return offset + rank_ext * (1 << (D - d)) + a + 1;
}

int gor_qk_num_nodes(int D, int order) {
  int n = 1;
  for (int d = 0; d < D; ++d) n *= order + 1;
  return n;
}

/* Lagrangian nodes of Q_order on the n-cube in Gridap's order: faces in face
 * order (vertices, edges, ..., interior), each face's own interior nodes with
 * x fastest.  src/ReferenceFEs/CLagrangianRefFEs.jl:626-764 (_compute_nodes,
 * _interior_nodes, _add_terms!). nodes: nnodes x D. */
int gor_qk_nodes(int D, int order, double *nodes) {
  int k = 0;
  for (int d = 0; d <= D; ++d) {
    for (int mask = 0; mask < (1 << D); ++mask) {
      if (__builtin_popcount(mask) != d) continue;
      for (int anchor = 0; anchor < (1 << (D - d)); ++anchor) {
        /* expand anchor bits onto fixed axes */
        int fixed_val[3] = {0, 0, 0};
        int bit = 0;
        for (int ax = 0; ax < D; ++ax) {
          if (mask & (1 << ax)) continue;
          fixed_val[ax] = (anchor >> bit) & 1;
          ++bit;
        }
        /* interior lattice points on free axes: indices 1..order-1, x fastest */
        int nin = 1;
        for (int ax = 0; ax < D; ++ax)
          if (mask & (1 << ax)) nin *= (order - 1);
        if (order == 1 && d > 0) nin = 0;
        for (int t = 0; t < nin; ++t) {
          int r = t;
          for (int ax = 0; ax < D; ++ax) {
            if (mask & (1 << ax)) {
              int i = r % (order - 1) + 1;
              r /= (order - 1);
              nodes[k * D + ax] = (double)i / order;
            } else {
              nodes[k * D + ax] = (double)fixed_val[ax];
            }
          }
          ++k;
        }
      }
    }
  }
  return k;
}

/* 1-D Lagrange basis on equispaced points j/order, value and derivative. */
static void lagrange1d(int order, int j, double t, double *val, double *der) {
  double v = 1.0, dv = 0.0;
  double xj = (double)j / order;
  for (int m = 0; m <= order; ++m) {
    if (m == j) continue;
    double xm = (double)m / order;
    double f = (t - xm) / (xj - xm);
    dv = dv * f + v / (xj - xm);
    v *= f;
  }
  *val = v;
  *der = dv;
}

/* Shape functions of the Lagrangian Q_order element at points: the nodal basis
 * inv(V)*monomials of src/ReferenceFEs/ReferenceFEInterfaces.jl:563-583 equals
 * the tensor product of 1-D Lagrange polynomials through the nodes.
 * shp: npts x nnodes; grad: npts x nnodes x D. */
int gor_qk_tabulate(int D, int order, int npts, const double *pts, double *shp,
                    double *grad) {
  int nn = gor_qk_num_nodes(D, order);
  double *nodes = (double *)malloc(sizeof(double) * nn * D);
  gor_qk_nodes(D, order, nodes);
  for (int p = 0; p < npts; ++p) {
    for (int a = 0; a < nn; ++a) {
      double v[3], dv[3];
      for (int ax = 0; ax < D; ++ax) {
        int j = (int)lround(nodes[a * D + ax] * order);
        lagrange1d(order, j, pts[p * D + ax], &v[ax], &dv[ax]);
      }
      double s = 1.0;
      for (int ax = 0; ax < D; ++ax) s *= v[ax];
      if (shp) shp[p * nn + a] = s;
      if (grad) {
        for (int g = 0; g < D; ++g) {
          double s2 = 1.0;
          for (int ax = 0; ax < D; ++ax) s2 *= (ax == g) ? dv[ax] : v[ax];
          grad[(p * nn + a) * D + g] = s2;
        }
      }
    }
  }
  free(nodes);
  return nn;
}

/* ------------------------------------------------------------------------- */
/* Cartesian model                                                           */
/* ------------------------------------------------------------------------- */

int64_t gor_cart_num_nodes(int D, const int64_t *part) {
  int64_t n = 1;
  for (int d = 0; d < D; ++d) n *= part[d] + 1;
  return n;
}

int64_t gor_cart_num_cells(int D, const int64_t *part) {
  int64_t n = 1;
  for (int d = 0; d < D; ++d) n *= part[d];
  return n;
}

/* Node coordinates x0 + (I-1)*dx over LinearIndices(partition.+1), x fastest:
 * src/Geometry/CartesianGrids.jl:107-124.  xyz: nnodes x D. */
void gor_cart_node_coords(int D, const double *origin, const double *sizes,
                          const int64_t *part, double *xyz) {
  int64_t nn = gor_cart_num_nodes(D, part);
  for (int64_t n = 0; n < nn; ++n) {
    int64_t r = n;
    for (int d = 0; d < D; ++d) {
      int64_t i = r % (part[d] + 1);
      r /= (part[d] + 1);
      xyz[n * D + d] = origin[d] + (double)i * sizes[d];
    }
  }
}

/* Cell nodes: for cell I (x fastest) the 2^D nodes at I-1 + lnode over
 * CartesianIndices((2,..,2)), 1-based: src/Geometry/CartesianGrids.jl:152-161. */
void gor_cart_cell_nodes(int D, const int64_t *part, int32_t *cell_nodes) {
  int64_t nc = gor_cart_num_cells(D, part);
  int nv = 1 << D;
  for (int64_t c = 0; c < nc; ++c) {
    int64_t ci[3] = {0, 0, 0}, r = c;
    for (int d = 0; d < D; ++d) {
      ci[d] = r % part[d];
      r /= part[d];
    }
    for (int a = 0; a < nv; ++a) {
      int64_t node = 0, stride = 1;
      for (int d = 0; d < D; ++d) {
        node += (ci[d] + ((a >> d) & 1)) * stride;
        stride *= part[d] + 1;
      }
      cell_nodes[c * nv + a] = (int32_t)(node + 1);
    }
  }
}

/* Vertex geolabels of a CartesianDiscreteModel: the result of
 * _generate_pre_geolabel! + _fix_geolabels (src/Geometry/CartesianDiscreteModels.jl:
 * 137-266) on a Cartesian grid is the id of the reference n-cube face whose free
 * axes are the axes where the vertex is interior and whose fixed coordinates are
 * 0 at index 0 and 1 at index n; interior vertices get num_faces = 3^D. */
void gor_cart_vertex_labels(int D, const int64_t *part, int32_t *labels) {
  int64_t nn = gor_cart_num_nodes(D, part);
  int interior = 1;
  for (int d = 0; d < D; ++d) interior *= 3;
  for (int64_t n = 0; n < nn; ++n) {
    int64_t r = n;
    int free_mask = 0, anchor = 0;
    for (int d = 0; d < D; ++d) {
      int64_t i = r % (part[d] + 1);
      r /= (part[d] + 1);
      if (i > 0 && i < part[d]) free_mask |= 1 << d;
      else if (i == part[d]) anchor |= 1 << d;
    }
    labels[n] = (free_mask == (1 << D) - 1) ? interior
                                            : gor_ncube_face_id(D, free_mask, anchor);
  }
}

/* ------------------------------------------------------------------------- */
/* CLagrangian numbering                                                     */
/* ------------------------------------------------------------------------- */

/* get_face_tag_index: first tag (1-based position in `tags`) whose entity list
 * holds the face's entity, else UNSET=0.  src/Geometry/FaceLabelings.jl:411-427. */
void gor_face_tag_index(int64_t nfaces, const int32_t *face_entity, int ntags,
                        const int32_t *tag_entity_ptrs, const int32_t *tag_entities,
                        int8_t *face_to_tag) {
  for (int64_t f = 0; f < nfaces; ++f) {
    face_to_tag[f] = 0;
    for (int t = 0; t < ntags; ++t) {
      int found = 0;
      for (int k = tag_entity_ptrs[t]; k < tag_entity_ptrs[t + 1]; ++k)
        if (tag_entities[k] == face_entity[f]) { found = 1; break; }
      if (found) { face_to_tag[f] = (int8_t)(t + 1); break; }
    }
  }
}

/* Node-major numbering, components interleaved per node, free ids 1.. and
 * Dirichlet ids -1.. in node order.  tag_comp_mask: ntags x ncomp (1 = Dirichlet).
 * src/FESpaces/CLagrangianFESpaces.jl:155-275. node_comp_to_dof: nnodes x ncomp. */
int gor_clagrangian_numbering(int64_t nnodes, const int8_t *node_to_tag, int ncomp,
                              const uint8_t *tag_comp_mask, int32_t *node_comp_to_dof,
                              int64_t *nfree, int64_t *ndiri) {
  int64_t nf = 0, nd = 0;
  for (int64_t n = 0; n < nnodes; ++n) {
    int tag = node_to_tag[n];
    for (int c = 0; c < ncomp; ++c) {
      int diri = (tag != 0) && tag_comp_mask[(tag - 1) * ncomp + c];
      if (diri) node_comp_to_dof[n * ncomp + c] = (int32_t)(-(++nd));
      else node_comp_to_dof[n * ncomp + c] = (int32_t)(++nf);
    }
  }
  *nfree = nf;
  *ndiri = nd;
  return 0;
}

/* Component-blocked local dofs: ldof = lnode + nnodes_cell*(comp-1)
 * (src/ReferenceFEs/LagrangianDofBases.jl:77-96), filled as in
 * _fill_cell_dofs_clagrangian! (src/FESpaces/CLagrangianFESpaces.jl:355-379). */
void gor_clagrangian_cell_dofs(int64_t ncells, int nnodes_cell, const int32_t *cell_nodes,
                               int ncomp, const int32_t *node_comp_to_dof,
                               int32_t *cell_dofs) {
  int nloc = nnodes_cell * ncomp;
  for (int64_t c = 0; c < ncells; ++c)
    for (int a = 0; a < nnodes_cell; ++a) {
      int64_t node = cell_nodes[c * nnodes_cell + a] - 1;
      for (int k = 0; k < ncomp; ++k)
        cell_dofs[c * nloc + a + nnodes_cell * k] = node_comp_to_dof[node * ncomp + k];
    }
}

/* ------------------------------------------------------------------------- */
/* Element integration                                                       */
/* ------------------------------------------------------------------------- */

static double det_inv(int D, const double *J, double *Ji) {
  if (D == 2) {
    double det = J[0] * J[3] - J[1] * J[2];
    Ji[0] = J[3] / det; Ji[1] = -J[1] / det;
    Ji[2] = -J[2] / det; Ji[3] = J[0] / det;
    return det;
  }
  double a = J[0], b = J[1], c = J[2], d = J[3], e = J[4], f = J[5], g = J[6], h = J[7],
         i = J[8];
  double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
  Ji[0] = (e * i - f * h) / det; Ji[1] = (c * h - b * i) / det; Ji[2] = (b * f - c * e) / det;
  Ji[3] = (f * g - d * i) / det; Ji[4] = (a * i - c * g) / det; Ji[5] = (c * d - a * f) / det;
  Ji[6] = (d * h - e * g) / det; Ji[7] = (b * g - a * h) / det; Ji[8] = (a * e - b * d) / det;
  return det;
}

/* Shared body: given Jt per point (Jt[i][j] = dx_j/dxi_i, the gradient of the
 * cell map, src/Fields/AffineMaps.jl / src/Geometry/CartesianGrids.jl:259-276),
 * push gradients forward with inv(Jt) (push_∇, src/Fields/FieldsInterfaces.jl:
 * 98-110, src/Fields/ApplyOptimizations.jl:302-310), form aq[p,i,j] =
 * kappa*grad_i.grad_j and integrate with dV = |det Jt| w (IntegrationMap for a
 * 3-D aq, src/Fields/FieldsInterfaces.jl:737-760: dV per point, then j outer,
 * i, p inner).  Output column-major (i fastest) per cell. */
static void diffusion_cell(int D, int nloc, int npts, const double *qw,
                           const double *shp_grad, const double *Jt_all, double kappa,
                           double *K, double *gphys, double *dV) {
  for (int p = 0; p < npts; ++p) {
    const double *Jt = Jt_all + p * D * D;
    double Ji[9];
    double det = det_inv(D, Jt, Ji);
    dV[p] = fabs(det) * qw[p];
    for (int a = 0; a < nloc; ++a)
      for (int r = 0; r < D; ++r) {
        double s = 0.0;
        for (int k = 0; k < D; ++k) s += Ji[r * D + k] * shp_grad[(p * nloc + a) * D + k];
        gphys[(p * nloc + a) * D + r] = s;
      }
  }
  for (int j = 0; j < nloc; ++j)
    for (int i = 0; i < nloc; ++i) {
      double rij = 0.0;
      for (int p = 0; p < npts; ++p) {
        double dot = 0.0;
        for (int r = 0; r < D; ++r)
          dot += gphys[(p * nloc + i) * D + r] * gphys[(p * nloc + j) * D + r];
        rij += (kappa * dot) * dV[p];
      }
      K[j * nloc + i] = rij;
    }
}

/* Isoparametric cell map x(xi) = sum_a x_a phi_a(xi) (UnstructuredGrid cell map,
 * src/Geometry/UnstructuredGrids.jl:52-100): Jt = sum_a grad phi_a (x) x_a. */
int gor_cell_matrices_diffusion(int D, int64_t ncells, int nloc, const int32_t *cell_nodes,
                                const double *xyz, int npts, const double *qw,
                                const double *shp_grad, const double *kappa,
                                double *cellmats) {
  double *gphys = (double *)malloc(sizeof(double) * npts * nloc * D);
  double *dV = (double *)malloc(sizeof(double) * npts);
  double *Jt = (double *)malloc(sizeof(double) * npts * D * D);
  for (int64_t c = 0; c < ncells; ++c) {
    for (int p = 0; p < npts; ++p)
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
          double s = 0.0;
          for (int a = 0; a < nloc; ++a) {
            int64_t node = cell_nodes[c * nloc + a] - 1;
            s += shp_grad[(p * nloc + a) * D + i] * xyz[node * D + j];
          }
          Jt[(p * D + i) * D + j] = s;
        }
    diffusion_cell(D, nloc, npts, qw, shp_grad, Jt, kappa ? kappa[c] : 1.0,
                   cellmats + c * nloc * nloc, gphys, dV);
  }
  free(gphys); free(dV); free(Jt);
  return 0;
}

/* CartesianMap: Jt = diagonal(sizes) for every cell (a Fill,
 * src/Geometry/CartesianGrids.jl:271-276). */
int gor_cell_matrices_diffusion_cartesian(int D, const double *sizes, int64_t ncells,
                                          int nloc, int npts, const double *qw,
                                          const double *shp_grad, const double *kappa,
                                          double *cellmats) {
  double *gphys = (double *)malloc(sizeof(double) * npts * nloc * D);
  double *dV = (double *)malloc(sizeof(double) * npts);
  double *Jt = (double *)calloc(npts * D * D, sizeof(double));
  for (int p = 0; p < npts; ++p)
    for (int d = 0; d < D; ++d) Jt[(p * D + d) * D + d] = sizes[d];
  for (int64_t c = 0; c < ncells; ++c)
    diffusion_cell(D, nloc, npts, qw, shp_grad, Jt, kappa ? kappa[c] : 1.0,
                   cellmats + c * nloc * nloc, gphys, dV);
  free(gphys); free(dV); free(Jt);
  return 0;
}

/* Element vector of l(v) = ∫ v f dΩ with f constant per cell: IntegrationMap for a
 * 2-D aq (src/Fields/FieldsInterfaces.jl:762-775): r[i] += aq[p,i]*dV[p], p outer. */
static void source_cell(int nloc, int npts, const double *shp, const double *dV, double f,
                        double *r) {
  for (int i = 0; i < nloc; ++i) r[i] = 0.0;
  for (int p = 0; p < npts; ++p)
    for (int i = 0; i < nloc; ++i) r[i] += (shp[p * nloc + i] * f) * dV[p];
}

int gor_cell_vectors_source(int D, int64_t ncells, int nloc, const int32_t *cell_nodes,
                            const double *xyz, int npts, const double *qw,
                            const double *shp, const double *shp_grad,
                            const double *fsrc, double *cellvecs) {
  double *dV = (double *)malloc(sizeof(double) * npts);
  for (int64_t c = 0; c < ncells; ++c) {
    for (int p = 0; p < npts; ++p) {
      double Jt[9], Ji[9];
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
          double s = 0.0;
          for (int a = 0; a < nloc; ++a) {
            int64_t node = cell_nodes[c * nloc + a] - 1;
            s += shp_grad[(p * nloc + a) * D + i] * xyz[node * D + j];
          }
          Jt[i * D + j] = s;
        }
      dV[p] = fabs(det_inv(D, Jt, Ji)) * qw[p];
    }
    source_cell(nloc, npts, shp, dV, fsrc ? fsrc[c] : 1.0, cellvecs + c * nloc);
  }
  free(dV);
  return 0;
}

/* Element vector of l(v) = ∫ v f dΩ with f given at the physical quadrature
 * points (fq: ncells x npts), same IntegrationMap loop order; cartesian_sizes
 * NULL selects the isoparametric map from node coordinates. */
int gor_cell_vectors_source_qp(int D, int64_t ncells, int nloc, const int32_t *cell_nodes,
                               const double *xyz, const double *cartesian_sizes, int npts,
                               const double *qw, const double *shp, const double *shp_grad,
                               const double *fq, double *cellvecs) {
  double *dV = (double *)malloc(sizeof(double) * npts);
  for (int64_t c = 0; c < ncells; ++c) {
    for (int p = 0; p < npts; ++p) {
      double Jt[9] = {0}, Ji[9];
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
          double s = 0.0;
          if (cartesian_sizes) {
            s = (i == j) ? cartesian_sizes[i] : 0.0;
          } else {
            for (int a = 0; a < nloc; ++a) {
              int64_t node = cell_nodes[c * nloc + a] - 1;
              s += shp_grad[(p * nloc + a) * D + i] * xyz[node * D + j];
            }
          }
          Jt[i * D + j] = s;
        }
      dV[p] = fabs(det_inv(D, Jt, Ji)) * qw[p];
    }
    double *r = cellvecs + c * nloc;
    for (int i = 0; i < nloc; ++i) r[i] = 0.0;
    for (int p = 0; p < npts; ++p)
      for (int i = 0; i < nloc; ++i) r[i] += (shp[p * nloc + i] * fq[c * npts + p]) * dV[p];
  }
  free(dV);
  return 0;
}

int gor_cell_vectors_source_cartesian(int D, const double *sizes, int64_t ncells, int nloc,
                                      int npts, const double *qw, const double *shp,
                                      const double *fsrc, double *cellvecs) {
  double *dV = (double *)malloc(sizeof(double) * npts);
  double det = 1.0;
  for (int d = 0; d < D; ++d) det *= sizes[d];
  for (int p = 0; p < npts; ++p) dV[p] = fabs(det) * qw[p];
  for (int64_t c = 0; c < ncells; ++c)
    source_cell(nloc, npts, shp, dV, fsrc ? fsrc[c] : 1.0, cellvecs + c * nloc);
  free(dV);
  return 0;
}

/* AttachDirichletMap: on cells with a Dirichlet dof, vec := vec - mat*vals where
 * vals are the cell dof values of uhd (0 on free dofs, g on Dirichlet dofs).
 * src/CellData/AttachDirichlet.jl:59-84; cell_is_dirichlet from
 * src/FESpaces/CLagrangianFESpaces.jl:143-152. */
void gor_attach_dirichlet(int64_t ncells, int nloc, const int32_t *cell_dofs,
                          const double *cellmats, const double *dirichlet_vals,
                          double *cellvecs) {
  for (int64_t c = 0; c < ncells; ++c) {
    const int32_t *dofs = cell_dofs + c * nloc;
    int isd = 0;
    for (int a = 0; a < nloc; ++a) isd |= dofs[a] < 0;
    if (!isd) continue;
    const double *K = cellmats + c * nloc * nloc;
    for (int i = 0; i < nloc; ++i) {
      double s = 0.0;
      for (int j = 0; j < nloc; ++j) {
        double v = dofs[j] < 0 ? dirichlet_vals[-dofs[j] - 1] : 0.0;
        s += K[j * nloc + i] * v;
      }
      cellvecs[c * nloc + i] -= s;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* MinMemory CSC builder                                                     */
/* ------------------------------------------------------------------------- */

/* Symbolic loop with CounterCSC: colnnzmax[j] += 1 for every (i>0,j>0) pair,
 * duplicates included (src/Algebra/SparseMatrixCSC.jl:82-85, loop order of
 * _add_entries!, src/Algebra/AlgebraInterfaces.jl:162-173). Returns the sum. */
int64_t gor_csc_count(int64_t ncells, int nloc, const int32_t *cell_dofs, int64_t ncols,
                      int64_t *colnnzmax) {
  memset(colnnzmax, 0, sizeof(int64_t) * ncols);
  for (int64_t c = 0; c < ncells; ++c) {
    const int32_t *dofs = cell_dofs + c * nloc;
    for (int lj = 0; lj < nloc; ++lj) {
      int32_t j = dofs[lj];
      if (j <= 0) continue;
      for (int li = 0; li < nloc; ++li)
        if (dofs[li] > 0) colnnzmax[j - 1] += 1;
    }
  }
  int64_t total = 0;
  for (int64_t j = 0; j < ncols; ++j) total += colnnzmax[j];
  return total;
}

/* InserterCSC.add_entry! (src/Algebra/SparseMatrixCSC.jl:102-150): sorted insert
 * of row i into column j by searchsortedfirst, shifting the tail one slot, or
 * accumulation into an existing entry.  has_value = 0 is the `v::Nothing`
 * method (structure only).  1-based i, j.  Returns -1 on column overflow
 * (the @check at SparseMatrixCSC.jl:135). */
static int inserter_add(int64_t *colptr, int64_t *colnnz, int64_t *rowval, double *nzval,
                        int64_t i, int64_t j, double v, int has_value) {
  int64_t pini = colptr[j - 1];
  int64_t pend = pini + colnnz[j - 1] - 1;
  int64_t lo = pini, hi = pend + 1;
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (rowval[mid - 1] < i) lo = mid + 1;
    else hi = mid;
  }
  int64_t p = lo;
  if (p > pend) {
    colnnz[j - 1] += 1;
    rowval[p - 1] = i;
    if (has_value) nzval[p - 1] = v;
  } else if (rowval[p - 1] != i) {
    if (!(pend + 1 < colptr[j])) return -1;
    for (int64_t k = pend; k >= p; --k) {
      rowval[k] = rowval[k - 1];
      nzval[k] = nzval[k - 1];
    }
    colnnz[j - 1] += 1;
    rowval[p - 1] = i;
    if (has_value) nzval[p - 1] = v;
  } else if (has_value) {
    nzval[p - 1] += v;
  }
  return 0;
}

/* nz_allocation: colptr from the counts by length_to_ptrs!, zeroed nzval and
 * colnnz (src/Algebra/SparseMatrixCSC.jl:250-262). */
static void nz_allocation(int64_t ncols, const int64_t *colnnzmax, int64_t *colptr,
                          int64_t *colnnz, double *nzval) {
  colptr[0] = 1;
  for (int64_t j = 0; j < ncols; ++j) colptr[j + 1] = colptr[j] + colnnzmax[j];
  for (int64_t k = 0; k < colptr[ncols] - 1; ++k) nzval[k] = 0.0;
  memset(colnnz, 0, sizeof(int64_t) * ncols);
}

/* create_from_nz: compact every column to its used prefix and rebuild colptr
 * (src/Algebra/SparseMatrixCSC.jl:264-283).  Returns nnz. */
static int64_t create_from_nz(int64_t ncols, int64_t *colptr, const int64_t *colnnz,
                              int64_t *rowval, double *nzval) {
  int64_t k = 1;
  for (int64_t j = 0; j < ncols; ++j) {
    int64_t pini = colptr[j];
    int64_t pend = pini + colnnz[j] - 1;
    for (int64_t p = pini; p <= pend; ++p) {
      nzval[k - 1] = nzval[p - 1];
      rowval[k - 1] = rowval[p - 1];
      ++k;
    }
  }
  colptr[0] = 1;
  for (int64_t j = 0; j < ncols; ++j) colptr[j + 1] = colptr[j] + colnnz[j];
  return colptr[ncols] - 1;
}

/* Full assemble_matrix with the default MinMemory builder
 * (src/FESpaces/SparseMatrixAssemblers.jl:70-77): counter + symbolic loop,
 * nz_allocation, numeric loop in cell order with lj outer / li inner
 * (src/Algebra/AlgebraInterfaces.jl:175-187) through the inserter, then
 * create_from_nz.  rowval/nzval are sized by gor_csc_count's total; on return
 * they hold the compacted 1-based CSC.  Returns nnz, or -1 on overflow. */
int64_t gor_csc_assemble(int64_t ncells, int nloc, const int32_t *cell_dofs,
                         const double *cellmats, int64_t nrows, int64_t ncols,
                         int64_t *colptr, int64_t *rowval, double *nzval) {
  (void)nrows;
  int64_t *colnnz = (int64_t *)calloc(ncols > 0 ? ncols : 1, sizeof(int64_t));
  gor_csc_count(ncells, nloc, cell_dofs, ncols, colnnz);
  nz_allocation(ncols, colnnz, colptr, colnnz, nzval);
  for (int64_t c = 0; c < ncells; ++c) {
    const int32_t *dofs = cell_dofs + c * nloc;
    const double *K = cellmats + c * nloc * nloc;
    for (int lj = 0; lj < nloc; ++lj) {
      int64_t j = dofs[lj];
      if (j <= 0) continue;
      for (int li = 0; li < nloc; ++li) {
        int64_t i = dofs[li];
        if (i <= 0) continue;
        if (inserter_add(colptr, colnnz, rowval, nzval, i, j, K[lj * nloc + li], 1)) {
          free(colnnz);
          return -1;
        }
      }
    }
  }
  int64_t nnz = create_from_nz(ncols, colptr, colnnz, rowval, nzval);
  free(colnnz);
  return nnz;
}

/* The builder protocol driven entry by entry, as test_assembler and
 * AlgebraInterfacesTests do: the counter sees every entry (valued or not), the
 * inserter adds them in order, create_from_nz compacts.  colnnzmax_out (ncols)
 * receives the counter state.  Returns nnz or -1. */
int64_t gor_csc_from_entries(int64_t nent, const int64_t *I, const int64_t *J, const double *V,
                             const uint8_t *has_value, int64_t ncols, int64_t *colnnzmax_out,
                             int64_t *colptr, int64_t *rowval, double *nzval) {
  int64_t *colnnz = (int64_t *)calloc(ncols > 0 ? ncols : 1, sizeof(int64_t));
  for (int64_t k = 0; k < nent; ++k) colnnz[J[k] - 1] += 1;
  if (colnnzmax_out) memcpy(colnnzmax_out, colnnz, sizeof(int64_t) * ncols);
  nz_allocation(ncols, colnnz, colptr, colnnz, nzval);
  for (int64_t k = 0; k < nent; ++k)
    if (inserter_add(colptr, colnnz, rowval, nzval, I[k], J[k], V[k], has_value[k])) {
      free(colnnz);
      return -1;
    }
  int64_t nnz = create_from_nz(ncols, colptr, colnnz, rowval, nzval);
  free(colnnz);
  return nnz;
}

/* nz_index: binary search of row i in column j of an existing CSC
 * (src/Algebra/SparseMatrixCSC.jl:14-22); 1-based, -1 when absent. */
static int64_t nz_index(const int64_t *colptr, const int64_t *rowval, int64_t i, int64_t j) {
  int64_t lo = colptr[j - 1], hi = colptr[j];  /* positions lo..hi-1 */
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (rowval[mid - 1] < i) lo = mid + 1;
    else hi = mid;
  }
  return (lo < colptr[j] && rowval[lo - 1] == i) ? lo : -1;
}

/* In-place numeric re-assembly of A and b on a Cartesian model, the steady-state
 * path of assemble_matrix_and_vector! (src/FESpaces/SparseMatrixAssemblers.jl:
 * 86-97): fillstored!(A,0) and fill!(b,0), then for each cell in order evaluate
 * the element matrix and vector (the lazy cell arrays, as in
 * gor_cell_matrices_diffusion_cartesian / gor_cell_vectors_source_cartesian) and
 * add them with add_entry!(A::AbstractSparseMatrix) = nz_index + += 
 * (src/Algebra/AlgebraInterfaces.jl:591-597), lj outer / li inner.
 * Homogeneous Dirichlet data (no lifting).  Returns -1 if an entry is missing. */
int gor_numeric_refill_cartesian(int D, const double *sizes, int64_t ncells, int nloc,
                                 const int32_t *cell_dofs, int npts, const double *qw,
                                 const double *shp, const double *shp_grad,
                                 const double *kappa, const double *fsrc, int64_t n,
                                 const int64_t *colptr, const int64_t *rowval, double *nzval,
                                 double *b) {
  double *gphys = (double *)malloc(sizeof(double) * npts * nloc * D);
  double *dV = (double *)malloc(sizeof(double) * npts);
  double *Jt = (double *)calloc(npts * D * D, sizeof(double));
  double *K = (double *)malloc(sizeof(double) * nloc * nloc);
  double *r = (double *)malloc(sizeof(double) * nloc);
  for (int p = 0; p < npts; ++p)
    for (int d = 0; d < D; ++d) Jt[(p * D + d) * D + d] = sizes[d];
  for (int64_t k = 0; k < colptr[n] - 1; ++k) nzval[k] = 0.0;
  for (int64_t k = 0; k < n; ++k) b[k] = 0.0;
  int status = 0;
  for (int64_t c = 0; c < ncells && status == 0; ++c) {
    diffusion_cell(D, nloc, npts, qw, shp_grad, Jt, kappa ? kappa[c] : 1.0, K, gphys, dV);
    source_cell(nloc, npts, shp, dV, fsrc ? fsrc[c] : 1.0, r);
    const int32_t *dofs = cell_dofs + c * nloc;
    for (int lj = 0; lj < nloc; ++lj) {
      int64_t j = dofs[lj];
      if (j <= 0) continue;
      for (int li = 0; li < nloc; ++li) {
        int64_t i = dofs[li];
        if (i <= 0) continue;
        int64_t p = nz_index(colptr, rowval, i, j);
        if (p < 0) { status = -1; break; }
        nzval[p - 1] += K[lj * nloc + li];
      }
    }
    for (int li = 0; li < nloc; ++li)
      if (dofs[li] > 0) b[dofs[li] - 1] += r[li];
  }
  free(gphys); free(dV); free(Jt); free(K); free(r);
  return status;
}

/* The same numeric loop on `nthreads` OpenMP threads, for timing only.  Gridap's
 * loop is sequential; here the cells of a Cartesian partition part[0..D-1] (cell
 * c = i0 + part0·(i1 + part1·i2)) are split into 2^D colours by the parity of their
 * coordinates.  Two cells of one colour share no node, so the threads of one colour
 * never add into the same nonzero or rhs entry.  Within a nonzero the additions come
 * in colour order, not cell order, so the result agrees with the sequential loop to
 * rounding only.  Returns -1 if an entry is missing. */
int gor_numeric_refill_cartesian_par(int D, const double *sizes, const int64_t *part,
                                     int nloc, const int32_t *cell_dofs, int npts,
                                     const double *qw, const double *shp,
                                     const double *shp_grad, const double *kappa,
                                     const double *fsrc, int64_t n, const int64_t *colptr,
                                     const int64_t *rowval, double *nzval, double *b,
                                     int nthreads) {
  int64_t p0 = part[0], p1 = D > 1 ? part[1] : 1, p2 = D > 2 ? part[2] : 1;
  int status = 0;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t k = 0; k < colptr[n] - 1; ++k) nzval[k] = 0.0;
  for (int64_t k = 0; k < n; ++k) b[k] = 0.0;
  for (int col = 0; col < (1 << D); ++col) {
    int64_t o0 = col & 1, o1 = (col >> 1) & 1, o2 = (col >> 2) & 1;
    int64_t m0 = (p0 - o0 + 1) / 2, m1 = (p1 - o1 + 1) / 2, m2 = (p2 - o2 + 1) / 2;
    int64_t mc = m0 * m1 * m2;
#pragma omp parallel num_threads(nthreads) reduction(| : status)
    {
      double *gphys = (double *)malloc(sizeof(double) * npts * nloc * D);
      double *dV = (double *)malloc(sizeof(double) * npts);
      double *Jt = (double *)calloc(npts * D * D, sizeof(double));
      double *K = (double *)malloc(sizeof(double) * nloc * nloc);
      double *r = (double *)malloc(sizeof(double) * nloc);
      for (int p = 0; p < npts; ++p)
        for (int d = 0; d < D; ++d) Jt[(p * D + d) * D + d] = sizes[d];
#pragma omp for schedule(static)
      for (int64_t t = 0; t < mc; ++t) {
        int64_t i0 = o0 + 2 * (t % m0), i1 = o1 + 2 * ((t / m0) % m1),
                i2 = o2 + 2 * (t / (m0 * m1));
        int64_t c = i0 + p0 * (i1 + p1 * i2);
        diffusion_cell(D, nloc, npts, qw, shp_grad, Jt, kappa ? kappa[c] : 1.0, K, gphys, dV);
        source_cell(nloc, npts, shp, dV, fsrc ? fsrc[c] : 1.0, r);
        const int32_t *dofs = cell_dofs + c * nloc;
        for (int lj = 0; lj < nloc; ++lj) {
          int64_t j = dofs[lj];
          if (j <= 0) continue;
          for (int li = 0; li < nloc; ++li) {
            int64_t i = dofs[li];
            if (i <= 0) continue;
            int64_t p = nz_index(colptr, rowval, i, j);
            if (p < 0) { status = -1; continue; }
            nzval[p - 1] += K[lj * nloc + li];
          }
        }
        for (int li = 0; li < nloc; ++li)
          if (dofs[li] > 0) b[dofs[li] - 1] += r[li];
      }
      free(gphys); free(dV); free(Jt); free(K); free(r);
    }
  }
  return status ? -1 : 0;
}

/* assemble_vector with ArrayBuilder: dense zeros, then b[i] += v for i > 0 in
 * cell order (src/Algebra/AlgebraInterfaces.jl:202-210,380-435). */
void gor_assemble_vector(int64_t ncells, int nloc, const int32_t *cell_dofs,
                         const double *cellvecs, int64_t nrows, double *b) {
  memset(b, 0, sizeof(double) * nrows);
  for (int64_t c = 0; c < ncells; ++c)
    for (int li = 0; li < nloc; ++li) {
      int32_t i = cell_dofs[c * nloc + li];
      if (i > 0) b[i - 1] += cellvecs[c * nloc + li];
    }
}

/* ------------------------------------------------------------------------- */
/* Solver checks                                                             */
/* ------------------------------------------------------------------------- */

void gor_csc_spmv(int64_t n, const int64_t *colptr, const int64_t *rowval,
                  const double *nzval, const double *x, double *y) {
  memset(y, 0, sizeof(double) * n);
  for (int64_t j = 0; j < n; ++j)
    for (int64_t p = colptr[j] - 1; p < colptr[j + 1] - 1; ++p)
      y[rowval[p] - 1] += nzval[p] * x[j];
}

/* Textbook Jacobi-preconditioned CG in FP64, stopping on ||b - A x|| / ||b|| <= rtol
 * with the residual recomputed from scratch at the end. */
int gor_pcg_jacobi(int64_t n, const int64_t *colptr, const int64_t *rowval,
                   const double *nzval, const double *b, double *x, double rtol,
                   int maxit, int *iters, double *relres) {
  double *r = (double *)malloc(sizeof(double) * n), *z = (double *)malloc(sizeof(double) * n);
  double *p = (double *)malloc(sizeof(double) * n), *q = (double *)malloc(sizeof(double) * n);
  double *dinv = (double *)malloc(sizeof(double) * n);
  for (int64_t j = 0; j < n; ++j) {
    double d = 0.0;
    for (int64_t k = colptr[j] - 1; k < colptr[j + 1] - 1; ++k)
      if (rowval[k] - 1 == j) d = nzval[k];
    dinv[j] = 1.0 / d;
  }
  gor_csc_spmv(n, colptr, rowval, nzval, x, q);
  double bb = 0.0, rz = 0.0, rr = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    r[i] = b[i] - q[i];
    z[i] = r[i] * dinv[i];
    p[i] = z[i];
    bb += b[i] * b[i];
    rz += r[i] * z[i];
    rr += r[i] * r[i];
  }
  double bn = sqrt(bb);
  if (bn == 0.0) bn = 1.0;
  int it = 0;
  while (sqrt(rr) / bn > rtol && it < maxit) {
    gor_csc_spmv(n, colptr, rowval, nzval, p, q);
    double pq = 0.0;
    for (int64_t i = 0; i < n; ++i) pq += p[i] * q[i];
    double alpha = rz / pq;
    double rz_new = 0.0;
    rr = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      z[i] = r[i] * dinv[i];
      rz_new += r[i] * z[i];
      rr += r[i] * r[i];
    }
    double beta = rz_new / rz;
    rz = rz_new;
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    ++it;
  }
  gor_csc_spmv(n, colptr, rowval, nzval, x, q);
  double rt = 0.0;
  for (int64_t i = 0; i < n; ++i) rt += (b[i] - q[i]) * (b[i] - q[i]);
  *iters = it;
  *relres = sqrt(rt) / bn;
  free(r); free(z); free(p); free(q); free(dinv);
  return (*relres <= rtol) ? 0 : 1;
}
