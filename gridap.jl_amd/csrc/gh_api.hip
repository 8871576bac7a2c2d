// C ABI of the assembly engine: argument checks, host<->device copies and the
// status-code boundary.  Compute lives in pattern.hip, assemble.hip, solve.hip.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "gh_internal.h"

struct gh_ctx {
  gh::Ctx c;
};

namespace {

template <class F>
int guarded(gh_ctx *ctx, F &&f) {
  if (!ctx) return GH_ERR_ARG;
  try {
    GH_CHECK_HIP(hipSetDevice(ctx->c.device));
    f(ctx->c);
    ctx->c.err.clear();
    return GH_OK;
  } catch (const gh::Error &e) {
    ctx->c.err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    ctx->c.err = "host allocation failed";
    return GH_ERR_OOM;
  } catch (const std::exception &e) {
    ctx->c.err = e.what();
    return GH_ERR_HIP;
  }
}

void clear_boundary(gh::Ctx &c) {
  c.nbfaces = 0;
  c.h_face_cell.clear();
  c.bface_cell.release();
  c.bface_side.release();
  c.balpha.release();
  c.bbeta.release();
  c.bg.release();
  c.bh.release();
  c.have_balpha = c.have_bbeta = c.have_bg = c.have_bh = false;
  c.bdst_n = -1;
  c.bown_n = 0;
  for (gh::DBuf<int32_t> *v : {&c.bdst_ptr, &c.bdst_ci, &c.bdst_row, &c.bdst_col, &c.bdst_pos,
                               &c.bown_dof, &c.bown_ptr, &c.bown_fa})
    v->release();
  c.bface_K.release();
  c.bface_r.release();
}

// An uploaded quadrilateral/hexahedral mesh whose cells are those of a Cartesian grid:
// lexicographic node ids, lexicographic cells, Gridap's local node order (x fastest),
// as a CartesianDiscreteModel with a map has (src/Geometry/CartesianGrids.jl:107-124).
// The lattice pattern and the owner-computes kernels then apply, with the node
// coordinates for the geometry.  Sets part and returns true when it is one.
bool lattice_topology(int dim, int64_t nnodes, int64_t ncells, int nv,
                      const std::vector<int32_t> &cn, int64_t *part) {
  if (nv != (1 << dim) || ncells < 1 || cn[0] != 0 || cn[1] != 1) return false;
  const int64_t n0 = (int64_t)cn[2] - cn[0];
  if (n0 < 2) return false;
  int64_t p[3] = {n0 - 1, 1, 1};
  int64_t n01 = 0;
  if (dim == 3) {
    n01 = (int64_t)cn[4] - cn[0];
    if (n01 < 2 * n0 || n01 % n0) return false;
    p[1] = n01 / n0 - 1;
    if (ncells % (p[0] * p[1])) return false;
    p[2] = ncells / (p[0] * p[1]);
    if (n01 * (p[2] + 1) != nnodes) return false;
  } else {
    if (ncells % p[0]) return false;
    p[1] = ncells / p[0];
    if (n0 * (p[1] + 1) != nnodes) return false;
  }
  for (int64_t k = 0; k < ncells; ++k) {
    const int64_t cx = k % p[0], cy = (k / p[0]) % p[1], cz = dim == 3 ? k / (p[0] * p[1]) : 0;
    const int64_t base = cx + n0 * (cy + (p[1] + 1) * cz);
    for (int a = 0; a < nv; ++a) {
      const int64_t node = base + (a & 1) + n0 * ((a >> 1) & 1) + n01 * ((a >> 2) & 1);
      if (cn[(size_t)k * nv + a] != node) return false;
    }
  }
  for (int d = 0; d < 3; ++d) part[d] = d < dim ? p[d] : 1;
  return true;
}

// A new mesh voids the FE space uploaded for the previous one: its cell_dof_ids were
// numbered on the old cells (gh_space_upload must follow gh_mesh_*).
void invalidate_space(gh::Ctx &c) {
  c.have_space = false;
  c.nloc = 0;
  c.nfree = c.ndiri = 0;
  c.cell_dofs.release();
  c.h_cell_dofs.clear();
  c.h_cell_dofs.shrink_to_fit();
  c.have_kappa = c.have_fsrc = c.have_dvals = false;
  c.have_sigma = false;
  c.sigma.release();
  c.have_fq = false;
  c.fq.release();
  c.have_state = false;
  c.nq = 0;
  c.nqf = 0;
  c.bface_tab_ok = false;
  clear_boundary(c);
}

void invalidate_pattern(gh::Ctx &c) {
  c.have_pattern = false;
  c.have_values = false;
  c.last_kernel = GH_KERNEL_NONE;
  c.structured = false;
  c.ncolors = 0;
  c.color_ptr.clear();
  c.fsai_ready = false;
  c.fsai_built = false;
  c.elemlist = false;
  c.sref_ready = false;
  c.q2lattice = false;
  c.q2_ref_ready = false;
  c.q2_node_dof.release();
  c.q2_rec.release();
  c.q2_nrec = 0;
  c.q2_irregular.release();
  c.q2_nirregular = 0;
  // state derived from the previous pattern (set again by the symbolic phase when it
  // applies to the new one)
  c.box = false;
  c.a27.release();
  c.g14.release();
  c.slots = false;
  c.metric = false;
  c.inverse = false;
  c.cslot.release();
  c.noff.release();
  c.bbase.release();
  c.bstart.release();
  c.bdst_n = -1;  // the face lists hold nzval positions of the old pattern
}

}  // namespace

extern "C" {

int gh_create(gh_ctx **out, int device) {
  if (!out) return GH_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return GH_ERR_HIP;
  if (device < 0 || device >= ndev) return GH_ERR_ARG;
  gh_ctx *ctx = new (std::nothrow) gh_ctx();
  if (!ctx) return GH_ERR_OOM;
  ctx->c.device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return GH_ERR_HIP;
  }
  try {
    gh::scratch_init(device);  // once per device: the allocation-free solve's array
  } catch (const gh::Error &e) {
    (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
    return e.code;
  }
  *out = ctx;
  return GH_OK;
}

int gh_destroy(gh_ctx *ctx) {
  if (!ctx) return GH_OK;
  (void)hipSetDevice(ctx->c.device);
  (void)hipStreamSynchronize(ctx->c.stream);
  hipStream_t s = ctx->c.stream;
  delete ctx;  // frees device buffers
  (void)hipStreamDestroy(s);
  return GH_OK;
}

const char *gh_last_error(const gh_ctx *ctx) {
  return ctx ? ctx->c.err.c_str() : "null context";
}

int gh_stream(gh_ctx *ctx, void **stream) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(stream, GH_ERR_ARG, "stream pointer is NULL");
    *stream = (void *)c.stream;
  });
}

int gh_mesh_cartesian(gh_ctx *ctx, int dim, const double *origin, const double *sizes,
                      const int64_t *partition) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(dim == 2 || dim == 3, GH_ERR_ARG, "Cartesian mesh dimension must be 2 or 3");
    GH_REQUIRE(origin && sizes && partition, GH_ERR_ARG, "NULL descriptor array");
    int64_t nn = 1, nc = 1;
    for (int d = 0; d < dim; ++d) {
      GH_REQUIRE(partition[d] >= 1, GH_ERR_ARG, "partition entries must be >= 1");
      GH_REQUIRE(sizes[d] > 0, GH_ERR_ARG, "cell sizes must be positive");
      nn *= partition[d] + 1;
      nc *= partition[d];
    }
    GH_REQUIRE(nn < (int64_t(1) << 31), GH_ERR_ARG, "too many nodes for 32-bit ids");
    c.dim = dim;
    c.cartesian = true;
    c.lattice_topo = true;
    for (int d = 0; d < 3; ++d) {
      c.origin[d] = d < dim ? origin[d] : 0.0;
      c.sizes[d] = d < dim ? sizes[d] : 1.0;
      c.part[d] = d < dim ? partition[d] : 1;
    }
    c.nnodes = nn;
    c.ncells = nc;
    c.nv = 1 << dim;
    c.xyz.release();
    c.cell_nodes.release();
    invalidate_space(c);
    invalidate_pattern(c);
  });
}

int gh_mesh_upload(gh_ctx *ctx, int dim, int64_t nnodes, const double *xyz, int64_t ncells,
                   const int32_t *data, const int32_t *ptrs) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(dim == 2 || dim == 3, GH_ERR_ARG, "mesh dimension must be 2 or 3");
    GH_REQUIRE(nnodes > 0 && ncells > 0 && xyz && data && ptrs, GH_ERR_ARG,
               "empty mesh or NULL array");
    GH_REQUIRE(ptrs[0] == 1, GH_ERR_ARG, "cell_node_ids ptrs must start at 1");
    int nv = ptrs[1] - ptrs[0];
    GH_REQUIRE(nv > 0 && nv <= 32, GH_ERR_ARG, "unsupported nodes per cell");
    std::vector<int32_t> cn(size_t(ncells) * nv);
    for (int64_t k = 0; k < ncells; ++k) {
      GH_REQUIRE(ptrs[k + 1] - ptrs[k] == nv, GH_ERR_UNSUPPORTED,
                 "mixed cell types are not supported");
      for (int a = 0; a < nv; ++a) {
        int32_t n = data[ptrs[k] - 1 + a];
        GH_REQUIRE(n >= 1 && n <= nnodes, GH_ERR_RANGE, "cell node id out of range");
        cn[size_t(k) * nv + a] = n - 1;
      }
    }
    c.dim = dim;
    c.cartesian = false;
    c.lattice_topo = lattice_topology(dim, nnodes, ncells, nv, cn, c.part);
    c.nnodes = nnodes;
    c.ncells = ncells;
    c.nv = nv;
    c.xyz.resize(size_t(nnodes) * dim);
    c.cell_nodes.resize(cn.size());
    GH_CHECK_HIP(hipMemcpy(c.xyz.get(), xyz, sizeof(double) * nnodes * dim,
                           hipMemcpyHostToDevice));
    GH_CHECK_HIP(hipMemcpy(c.cell_nodes.get(), cn.data(), sizeof(int32_t) * cn.size(),
                           hipMemcpyHostToDevice));
    invalidate_space(c);
    invalidate_pattern(c);
  });
}

int gh_space_upload(gh_ctx *ctx, int reffe_kind, int order, int ncomp, const int32_t *data,
                    const int32_t *ptrs, int64_t ncells, int64_t nfree, int64_t ndiri) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.dim > 0, GH_ERR_STATE, "upload a mesh before the FE space");
    GH_REQUIRE(ncells == c.ncells, GH_ERR_ARG, "cell count differs from the mesh");
    GH_REQUIRE(data && ptrs, GH_ERR_ARG, "NULL cell_dof_ids array");
    GH_REQUIRE(reffe_kind == GH_REFFE_QK || reffe_kind == GH_REFFE_PK, GH_ERR_UNSUPPORTED,
               "unknown reference element kind");
    GH_REQUIRE(order >= 1 && ncomp >= 1, GH_ERR_ARG, "order and ncomp must be >= 1");
    GH_REQUIRE(nfree >= 0 && ndiri >= 0 && nfree < (int64_t(1) << 31) &&
                   ndiri < (int64_t(1) << 31),
               GH_ERR_ARG, "dof counts out of range");
    GH_REQUIRE(ptrs[0] == 1, GH_ERR_ARG, "cell_dof_ids ptrs must start at 1");
    int nloc = ptrs[1] - ptrs[0];
    GH_REQUIRE(nloc >= 1 && nloc <= 256, GH_ERR_UNSUPPORTED, "unsupported dofs per cell");
    std::vector<int32_t> cd(size_t(ncells) * nloc);
    for (int64_t k = 0; k < ncells; ++k) {
      GH_REQUIRE(ptrs[k + 1] - ptrs[k] == nloc, GH_ERR_UNSUPPORTED,
                 "cells with different numbers of dofs are not supported");
      const int32_t *src = data + (ptrs[k] - 1);
      for (int a = 0; a < nloc; ++a) {
        int32_t d = src[a];
        GH_REQUIRE(d != 0 && d <= nfree && -d <= ndiri, GH_ERR_RANGE,
                   "cell dof id out of range");
      }
      std::memcpy(&cd[size_t(k) * nloc], src, sizeof(int32_t) * nloc);
    }
    c.reffe_kind = reffe_kind;
    c.order = order;
    c.ncomp = ncomp;
    c.nloc = nloc;
    c.nfree = nfree;
    c.ndiri = ndiri;
    c.cell_dofs.resize(cd.size());
    GH_CHECK_HIP(hipMemcpy(c.cell_dofs.get(), cd.data(), sizeof(int32_t) * cd.size(),
                           hipMemcpyHostToDevice));
    c.h_cell_dofs.swap(cd);
    c.have_space = true;
    c.have_trial = false;  // a new test space ends the previous rectangular pair
    c.tcell_dofs.release();
    c.tshp.clear();
    c.tgrad.clear();
    c.tnq = 0;
    invalidate_pattern(c);
    c.have_dvals = false;
    c.have_state = false;
    clear_boundary(c);
  });
}

int gh_trial_space_upload(gh_ctx *ctx, int reffe_kind, int order, int ncomp,
                          const int32_t *data, const int32_t *ptrs, int64_t ncells,
                          int64_t nfree, int64_t ndiri) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.have_space, GH_ERR_STATE, "upload the test space before the trial space");
    GH_REQUIRE(ncells == c.ncells, GH_ERR_ARG, "cell count differs from the mesh");
    GH_REQUIRE(data && ptrs, GH_ERR_ARG, "NULL cell_dof_ids array");
    GH_REQUIRE(reffe_kind == GH_REFFE_QK && c.reffe_kind == GH_REFFE_QK && order >= 1 &&
                   order <= 2 && c.order >= 1 && c.order <= 2 &&
                   (ncomp == 1 || ncomp == c.dim) && (c.ncomp == 1 || c.ncomp == c.dim) &&
                   (ncomp == 1 || c.ncomp == 1),
               GH_ERR_UNSUPPORTED,
               "rectangular pairs are provided between Q1/Q2 spaces on quadrilaterals and "
               "hexahedra, scalar or with one component per dimension, at most one of them "
               "vector-valued");
    GH_REQUIRE(nfree >= 0 && ndiri >= 0 && nfree < (int64_t(1) << 31) &&
                   ndiri < (int64_t(1) << 31),
               GH_ERR_ARG, "dof counts out of range");
    GH_REQUIRE(ptrs[0] == 1, GH_ERR_ARG, "cell_dof_ids ptrs must start at 1");
    const int nloc = ptrs[1] - ptrs[0];
    GH_REQUIRE(nloc >= 1 && nloc <= 256, GH_ERR_UNSUPPORTED, "unsupported dofs per cell");
    std::vector<int32_t> cd(size_t(ncells) * nloc);
    for (int64_t k = 0; k < ncells; ++k) {
      GH_REQUIRE(ptrs[k + 1] - ptrs[k] == nloc, GH_ERR_UNSUPPORTED,
                 "cells with different numbers of dofs are not supported");
      const int32_t *src = data + (ptrs[k] - 1);
      for (int a = 0; a < nloc; ++a)
        GH_REQUIRE(src[a] != 0 && src[a] <= nfree && -src[a] <= ndiri, GH_ERR_RANGE,
                   "cell dof id out of range");
      std::memcpy(&cd[size_t(k) * nloc], src, sizeof(int32_t) * nloc);
    }
    GH_REQUIRE(nloc % ncomp == 0, GH_ERR_ARG, "dofs per cell not a multiple of ncomp");
    c.torder = order;
    c.tnloc = nloc;
    c.tncomp = ncomp;
    c.tnfree = nfree;
    c.tndiri = ndiri;
    c.tcell_dofs.resize(cd.size());
    GH_CHECK_HIP(hipMemcpy(c.tcell_dofs.get(), cd.data(), sizeof(int32_t) * cd.size(),
                           hipMemcpyHostToDevice));
    c.tshp.clear();
    c.tgrad.clear();
    c.tnq = 0;
    c.have_trial = true;
    invalidate_pattern(c);
    c.have_dvals = false;
    c.have_state = false;
    clear_boundary(c);
  });
}

int gh_trial_tabulation(gh_ctx *ctx, int nq, int nloc, const double *shp, const double *grad) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.have_trial, GH_ERR_STATE, "upload the trial space first");
    GH_REQUIRE(nq >= 1 && nloc == c.tnloc / c.tncomp && shp && grad, GH_ERR_ARG,
               "the trial tabulation needs nq >= 1 points and the trial space's scalar shape "
               "functions");
    c.tshp.assign(shp, shp + (size_t)nq * nloc);
    c.tgrad.assign(grad, grad + (size_t)nq * nloc * c.dim);
    c.tnq = nq;
    c.have_values = false;
  });
}

int gh_reffe_tabulation(gh_ctx *ctx, int nq, int nloc, const double *qw, const double *shp,
                        const double *grad) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.dim > 0, GH_ERR_STATE, "upload a mesh before the tabulation");
    GH_REQUIRE(nq >= 1 && nq <= 512 && nloc >= 1 && nloc <= 256, GH_ERR_ARG,
               "tabulation sizes out of range");
    GH_REQUIRE(qw && shp && grad, GH_ERR_ARG, "NULL tabulation array");
    if (nq != c.nq) {
      c.have_fq = false;
      c.fq.release();
    }
    c.nq = nq;
    c.qw.assign(qw, qw + nq);
    c.shp.assign(shp, shp + size_t(nq) * nloc);
    c.grad.assign(grad, grad + size_t(nq) * nloc * c.dim);
    c.have_values = false;
    c.sref_ready = false;
    c.q2_ref_ready = false;
    c.have_gtab = false;  // the geometry tabulation belongs to the previous points
    c.gtab.release();
  });
}

int gh_geometry_tabulation(gh_ctx *ctx, int nq, int nv, const double *grad) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.dim > 0 && !c.cartesian, GH_ERR_STATE,
               "the geometry tabulation belongs to a mesh uploaded with its nodes");
    GH_REQUIRE(nq == c.nq, GH_ERR_ARG,
               "the geometry tabulation must use the points of the FE tabulation");
    GH_REQUIRE(nv == c.nv, GH_ERR_ARG, "one geometry shape function per cell node");
    GH_REQUIRE(grad, GH_ERR_ARG, "NULL geometry tabulation");
    c.gtab.resize((size_t)nq * nv * c.dim);
    GH_CHECK_HIP(hipMemcpy(c.gtab.get(), grad, sizeof(double) * c.gtab.n,
                           hipMemcpyHostToDevice));
    c.have_gtab = true;
    c.gtab_nq = nq;
    c.have_values = false;
  });
}

int gh_face_tabulation(gh_ctx *ctx, int nsides, int nqf, const double *qw, const double *shp,
                       const double *grad) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.have_space, GH_ERR_STATE, "upload the FE space before the face tabulation");
    GH_REQUIRE(nsides == 2 * c.dim, GH_ERR_ARG, "one tabulation per facet of the n-cube");
    GH_REQUIRE(nqf >= 1 && nqf <= 64, GH_ERR_ARG, "face quadrature size out of range");
    GH_REQUIRE(qw && shp && grad, GH_ERR_ARG, "NULL face tabulation array");
    // the scalar shape functions (a vector space's components share them)
    const int NL = c.nloc / c.ncomp, D = c.dim, stride = 1 + NL + NL * D;
    std::vector<double> tab((size_t)nsides * nqf * stride);
    for (int s = 0; s < nsides; ++s)
      for (int q = 0; q < nqf; ++q) {
        double *t = &tab[((size_t)s * nqf + q) * stride];
        t[0] = qw[q];
        for (int a = 0; a < NL; ++a) t[1 + a] = shp[((size_t)s * nqf + q) * NL + a];
        for (int a = 0; a < NL * D; ++a)
          t[1 + NL + a] = grad[((size_t)s * nqf + q) * NL * D + a];
      }
    c.bface_tab.resize(tab.size());
    GH_CHECK_HIP(hipMemcpy(c.bface_tab.get(), tab.data(), sizeof(double) * tab.size(),
                           hipMemcpyHostToDevice));
    c.nqf = nqf;
    c.bface_tab_ok = true;
    c.have_values = false;
  });
}

int gh_boundary_upload(gh_ctx *ctx, int64_t nfaces, const int32_t *face_cells,
                       const int32_t *face_sides, const double *alpha, const double *beta,
                       const double *gq, const double *hq) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(!c.have_trial, GH_ERR_UNSUPPORTED,
               "boundary terms belong to square (same-space) pairs");
    GH_REQUIRE(c.have_space, GH_ERR_STATE, "upload the FE space before boundary terms");
    GH_REQUIRE(nfaces >= 0 && nfaces < (int64_t(1) << 31), GH_ERR_ARG, "bad face count");
    clear_boundary(c);
    c.have_values = false;
    if (nfaces == 0) return;
    GH_REQUIRE(face_cells && face_sides, GH_ERR_ARG, "NULL face arrays");
    GH_REQUIRE(c.bface_tab_ok && c.nqf > 0, GH_ERR_STATE,
               "upload the face tabulation (gh_face_tabulation) first");
    std::vector<int32_t> fc(nfaces), fs(nfaces);
    for (int64_t f = 0; f < nfaces; ++f) {
      GH_REQUIRE(face_cells[f] >= 1 && face_cells[f] <= c.ncells, GH_ERR_RANGE,
                 "face cell out of range");
      GH_REQUIRE(face_sides[f] >= 1 && face_sides[f] <= 2 * c.dim, GH_ERR_RANGE,
                 "face side out of range");
      fc[f] = face_cells[f] - 1;
      fs[f] = face_sides[f] - 1;
    }
    c.bface_cell.resize(nfaces);
    c.bface_side.resize(nfaces);
    GH_CHECK_HIP(hipMemcpy(c.bface_cell.get(), fc.data(), sizeof(int32_t) * nfaces,
                           hipMemcpyHostToDevice));
    GH_CHECK_HIP(hipMemcpy(c.bface_side.get(), fs.data(), sizeof(int32_t) * nfaces,
                           hipMemcpyHostToDevice));
    auto up = [&](gh::DBuf<double> &buf, bool &flag, const double *src, size_t n) {
      flag = src != nullptr;
      if (!flag) return;
      buf.resize(n);
      GH_CHECK_HIP(hipMemcpy(buf.get(), src, sizeof(double) * n, hipMemcpyHostToDevice));
    };
    up(c.balpha, c.have_balpha, alpha, nfaces);
    up(c.bbeta, c.have_bbeta, beta, nfaces);
    up(c.bg, c.have_bg, gq, (size_t)nfaces * c.nqf * c.ncomp);
    up(c.bh, c.have_bh, hq, (size_t)nfaces * c.nqf * c.ncomp);
    c.h_face_cell.swap(fc);
    c.nbfaces = nfaces;
  });
}

int gh_symbolic(gh_ctx *ctx, int64_t *nnz_out) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.have_space, GH_ERR_STATE, "upload the FE space before gh_symbolic");
    GH_REQUIRE(c.nloc > 0 && c.cell_dofs.n == size_t(c.ncells) * c.nloc &&
                   c.h_cell_dofs.size() == c.cell_dofs.n,
               GH_ERR_STATE, "the FE space does not match the mesh: upload it again");
    invalidate_pattern(c);
    if (c.have_trial) {
      gh::rect_symbolic(c);
      if (nnz_out) *nnz_out = c.nnz;
      return;
    }
    if (!gh::symbolic_lattice(c)) {
      gh::symbolic_general(c, gh::simplex_eligible(c) || gh::hexq2_eligible(c) ||
                                  gh::hexgen_eligible(c));
      gh::check_structured(c);
    }
    gh::elemlist_symbolic(c);
    gh::q2lattice_prepare(c);
    if (nnz_out) *nnz_out = c.nnz;
  });
}

int gh_pattern_get(gh_ctx *ctx, int64_t *colptr, int64_t *rowval) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.have_pattern, GH_ERR_STATE, "no pattern: call gh_symbolic first");
    std::vector<int32_t> cp((c.have_trial ? c.tnfree : c.nfree) + 1), rv(c.nnz);
    GH_CHECK_HIP(hipMemcpyAsync(cp.data(), c.colptr.get(), sizeof(int32_t) * cp.size(),
                                hipMemcpyDeviceToHost, c.stream));
    if (c.nnz)
      GH_CHECK_HIP(hipMemcpyAsync(rv.data(), c.rowval.get(), sizeof(int32_t) * rv.size(),
                                  hipMemcpyDeviceToHost, c.stream));
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
    if (colptr)
      for (size_t k = 0; k < cp.size(); ++k) colptr[k] = int64_t(cp[k]) + 1;
    if (rowval)
      for (size_t k = 0; k < rv.size(); ++k) rowval[k] = int64_t(rv[k]) + 1;
  });
}

int gh_coef_upload(gh_ctx *ctx, const double *kappa, const double *fsrc,
                   const double *dirichlet_vals) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.have_space, GH_ERR_STATE, "upload the FE space before coefficients");
    auto up = [&](gh::DBuf<double> &buf, bool &flag, const double *src, int64_t n) {
      flag = src != nullptr && n > 0;
      if (!flag) return;
      buf.resize(n);
      GH_CHECK_HIP(hipMemcpyAsync(buf.get(), src, sizeof(double) * n, hipMemcpyHostToDevice,
                                  c.stream));
    };
    up(c.kappa, c.have_kappa, kappa, c.ncells);
    up(c.fsrc, c.have_fsrc, fsrc, c.ncells);
    // Dirichlet values of the trial space (the test space's own for a square pair)
    up(c.dvals, c.have_dvals, dirichlet_vals, c.have_trial ? c.tndiri : c.ndiri);
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
  });
}

int gh_mass_upload(gh_ctx *ctx, const double *sigma) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.have_space, GH_ERR_STATE, "upload the FE space before coefficients");
    c.have_values = false;
    c.have_sigma = sigma != nullptr && c.ncells > 0;
    if (!c.have_sigma) {
      c.sigma.release();
      return;
    }
    GH_REQUIRE(c.reffe_kind == GH_REFFE_QK && c.ncomp == 1 && (c.order == 1 || c.order == 2),
               GH_ERR_UNSUPPORTED, "the cell mass term is provided for scalar Q1 and Q2 spaces");
    c.sigma.resize(c.ncells);
    GH_CHECK_HIP(hipMemcpyAsync(c.sigma.get(), sigma, sizeof(double) * c.ncells,
                                hipMemcpyHostToDevice, c.stream));
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
  });
}

int gh_source_qp_upload(gh_ctx *ctx, const double *fq) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(!c.have_trial || !fq, GH_ERR_UNSUPPORTED,
               "sources belong to the test field's own (square) pair");
    GH_REQUIRE(c.have_space && c.nq > 0, GH_ERR_STATE,
               "upload the FE space and the tabulation before the source");
    c.have_fq = fq != nullptr;
    c.have_values = false;
    if (!fq) {
      c.fq.release();
      return;
    }
    GH_REQUIRE(c.reffe_kind == GH_REFFE_QK &&
                   ((c.ncomp == 1 && (c.order == 1 || gh::hexgen_eligible(c))) ||
                    (c.ncomp == c.dim && gh::hexgen_eligible(c))),
               GH_ERR_UNSUPPORTED,
               "sources at quadrature points are provided for scalar Q1 and Q2 spaces and for "
               "vector spaces of the general n-cube path (fq[(cell*nq + q)*ncomp + i])");
    const size_t n = (size_t)c.ncells * c.nq * c.ncomp;
    c.fq.resize(n);
    GH_CHECK_HIP(hipMemcpy(c.fq.get(), fq, sizeof(double) * n, hipMemcpyHostToDevice));
  });
}

int gh_assemble(gh_ctx *ctx, int form_kind, int precision, int path) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(form_kind == GH_FORM_DIFFUSION || form_kind == GH_FORM_ELASTICITY ||
                   form_kind == GH_FORM_PLAPLACIAN ||
                   (form_kind == GH_FORM_DIVERGENCE && c.have_trial),
               GH_ERR_UNSUPPORTED, "unknown form kind (the divergence form needs a trial space)");
    GH_REQUIRE(precision == GH_PREC_FP64 || precision == GH_PREC_FP32_ELEM, GH_ERR_ARG,
               "unknown precision mode");
    GH_REQUIRE(path == GH_PATH_AUTO || path == GH_PATH_GENERAL || path == GH_PATH_STRUCTURED,
               GH_ERR_ARG, "unknown numeric path");
    GH_REQUIRE(c.have_pattern, GH_ERR_STATE, "no pattern: call gh_symbolic first");
    GH_REQUIRE(c.nq > 0, GH_ERR_STATE, "no reference tabulation uploaded");
    if (c.have_trial) {
      GH_REQUIRE((form_kind == GH_FORM_DIFFUSION || form_kind == GH_FORM_DIVERGENCE) &&
                     path != GH_PATH_STRUCTURED,
                 GH_ERR_UNSUPPORTED,
                 "a rectangular pair assembles κ ∇ψ⋅∇φ + σ ψ φ or the divergence coupling on "
                 "the general path");
      GH_REQUIRE(!c.have_fsrc && !c.have_fq, GH_ERR_UNSUPPORTED,
                 "sources belong to the test field's own (square) pair");
      gh::rect_assemble(c, form_kind, precision);
    } else if (form_kind == GH_FORM_PLAPLACIAN) {
      gh::assemble_plaplacian(c, precision, path);
    } else if (form_kind == GH_FORM_ELASTICITY) {
      if (gh::hexgen_eligible(c)) {
        GH_REQUIRE(path != GH_PATH_STRUCTURED, GH_ERR_UNSUPPORTED,
                   "the lattice path assembles vector Q2 on Cartesian hexahedra");
        gh::hexgen_assemble(c, precision, true);
      } else {
        gh::assemble_elasticity(c, precision, path);
      }
      gh::assemble_boundary(c);  // vector spaces: boundary loads
    } else {
      if (gh::hexq2_eligible(c)) {
        // vector Q2 on Cartesian hexahedra: the vector Laplacian on the elasticity lattice
        GH_REQUIRE(c.q2lattice && path != GH_PATH_GENERAL, GH_ERR_UNSUPPORTED,
                   "the vector Laplacian on Cartesian vector-Q2 hexahedra takes the lattice "
                   "path");
        GH_REQUIRE((int)c.shp.size() == c.nq * 27 && (int)c.grad.size() == c.nq * 27 * 3,
                   GH_ERR_STATE, "the vector Laplacian needs the scalar Q2 tabulation");
        GH_REQUIRE(!c.have_sigma && !c.have_fq, GH_ERR_UNSUPPORTED,
                   "the vector-Q2 lattice takes κ per cell and a constant body force");
        gh::q2lattice_assemble(c, precision == GH_PREC_FP32_ELEM, true);
      } else if (gh::hexgen_eligible(c)) {
        GH_REQUIRE(path != GH_PATH_STRUCTURED, GH_ERR_UNSUPPORTED,
                   "the lattice path assembles scalar Q1 spaces");
        gh::hexgen_assemble(c, precision, false);
      } else {
        gh::assemble_diffusion(c, precision, path);
      }
      gh::assemble_boundary(c);
    }
    c.have_values = true;
    c.fsai_built = false;
  });
}

int gh_elasticity_params(gh_ctx *ctx, double nu, const double *body_force) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(nu > -1.0 && nu < 0.5, GH_ERR_ARG, "Poisson ratio must lie in (-1, 0.5)");
    c.nu = nu;
    for (int d = 0; d < 3; ++d) c.body[d] = body_force && d < c.dim ? body_force[d] : 0.0;
    c.have_values = false;
  });
}

int gh_fefunction_upload(gh_ctx *ctx, const double *free_vals, const double *dirichlet_vals) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(!c.have_trial, GH_ERR_UNSUPPORTED, "nonlinear forms need a square pair");
    GH_REQUIRE(c.have_space, GH_ERR_STATE, "upload the FE space first");
    GH_REQUIRE(free_vals || c.nfree == 0, GH_ERR_ARG, "NULL free values");
    c.ustate.resize(std::max<int64_t>(c.nfree, 1));
    if (c.nfree)
      GH_CHECK_HIP(hipMemcpyAsync(c.ustate.get(), free_vals, sizeof(double) * c.nfree,
                                  hipMemcpyHostToDevice, c.stream));
    c.have_dvals = dirichlet_vals != nullptr && c.ndiri > 0;
    if (c.have_dvals) {
      c.dvals.resize(c.ndiri);
      GH_CHECK_HIP(hipMemcpyAsync(c.dvals.get(), dirichlet_vals, sizeof(double) * c.ndiri,
                                  hipMemcpyHostToDevice, c.stream));
    }
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
    c.have_state = true;
  });
}

int gh_plaplacian_params(gh_ctx *ctx, double p) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(std::isfinite(p) && p >= 2.0, GH_ERR_ARG, "p must be a finite number >= 2");
    c.plap_p = p;
  });
}

int gh_values_get(gh_ctx *ctx, double *nzval, double *b) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.have_values, GH_ERR_STATE, "no assembled values: call gh_assemble first");
    if (nzval && c.nnz)
      GH_CHECK_HIP(hipMemcpyAsync(nzval, c.nzval.get(), sizeof(double) * c.nnz,
                                  hipMemcpyDeviceToHost, c.stream));
    if (b && c.nfree)
      GH_CHECK_HIP(hipMemcpyAsync(b, c.b.get(), sizeof(double) * c.nfree,
                                  hipMemcpyDeviceToHost, c.stream));
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
  });
}

int gh_compose_block(gh_ctx *ctx, gh_ctx *src, const int64_t *dst, int64_t n, int64_t nnz) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(src && src != ctx, GH_ERR_ARG, "a composite block is another context");
    GH_REQUIRE(src->c.device == c.device, GH_ERR_ARG, "a composite block is on another device");
    GH_REQUIRE(src->c.have_pattern, GH_ERR_STATE, "the block has no pattern: call gh_symbolic");
    GH_REQUIRE(n == src->c.nnz, GH_ERR_ARG, "n is not the block's number of stored values");
    GH_REQUIRE(n == 0 || dst, GH_ERR_ARG, "NULL dst");
    GH_REQUIRE(nnz >= 0, GH_ERR_ARG, "negative nnz");
    GH_REQUIRE(nnz <= INT32_MAX, GH_ERR_UNSUPPORTED, "composite above 2^31 - 1 stored values");
    GH_REQUIRE(c.comp_nnz < 0 || c.comp_nnz == nnz, GH_ERR_ARG,
               "nnz differs from the composite's registered blocks");
    // marked on a copy, kept only when the whole block is valid
    std::vector<bool> taken = c.comp_nnz < 0 ? std::vector<bool>((size_t)nnz, false)
                                             : c.comp_taken;
    std::vector<int32_t> d((size_t)n);
    for (int64_t k = 0; k < n; ++k) {
      GH_REQUIRE(dst[k] >= 0 && dst[k] < nnz, GH_ERR_ARG, "composite position out of range");
      GH_REQUIRE(!taken[(size_t)dst[k]], GH_ERR_ARG,
                 "composite position written by two stored values");
      taken[(size_t)dst[k]] = true;
      d[(size_t)k] = (int32_t)dst[k];
    }
    auto blk = std::make_unique<gh::Ctx::ComposeBlock>();
    blk->src = &src->c;
    blk->n = n;
    blk->dst.resize((size_t)n);
    if (n)
      GH_CHECK_HIP(hipMemcpyAsync(blk->dst.get(), d.data(), sizeof(int32_t) * n,
                                  hipMemcpyHostToDevice, c.stream));
    if (c.comp_nnz < 0) {
      // positions no block covers read as stored zeros
      c.comp_nz.resize((size_t)nnz);
      if (nnz)
        GH_CHECK_HIP(hipMemsetAsync(c.comp_nz.get(), 0, sizeof(double) * nnz, c.stream));
    }
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
    c.comp.push_back(std::move(blk));
    c.comp_taken.swap(taken);
    c.comp_nnz = nnz;
  });
}

int gh_compose_move(gh_ctx *ctx, int block) {
  return guarded(ctx, [&](gh::Ctx &c) { gh::compose_move(c, block); });
}

int gh_compose_values_get(gh_ctx *ctx, double *nzval) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.comp_nnz >= 0, GH_ERR_STATE, "no composite blocks: call gh_compose_block");
    GH_REQUIRE(nzval || c.comp_nnz == 0, GH_ERR_ARG, "NULL nzval");
    for (const auto &blk : c.comp) GH_CHECK_HIP(hipStreamSynchronize(blk->src->stream));
    if (c.comp_nnz)
      GH_CHECK_HIP(hipMemcpyAsync(nzval, c.comp_nz.get(), sizeof(double) * c.comp_nnz,
                                  hipMemcpyDeviceToHost, c.stream));
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
  });
}

int gh_compose_clear(gh_ctx *ctx) {
  return guarded(ctx, [&](gh::Ctx &c) {
    c.comp.clear();
    c.comp_nz.release();
    c.comp_taken.clear();
    c.comp_nnz = -1;
  });
}

int gh_values_set(gh_ctx *ctx, const double *nzval, const double *b) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.have_pattern, GH_ERR_STATE, "no pattern: call gh_symbolic first");
    GH_REQUIRE(nzval || c.have_values, GH_ERR_STATE,
               "no matrix on the device: pass nzval or assemble first");
    if (nzval && c.nnz)
      GH_CHECK_HIP(hipMemcpyAsync(c.nzval.get(), nzval, sizeof(double) * c.nnz,
                                  hipMemcpyHostToDevice, c.stream));
    if (b && c.nfree)
      GH_CHECK_HIP(hipMemcpyAsync(c.b.get(), b, sizeof(double) * c.nfree,
                                  hipMemcpyHostToDevice, c.stream));
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
    c.have_values = true;
    if (nzval) c.fsai_built = false;
  });
}

int gh_pattern_info_get(gh_ctx *ctx, gh_pattern_info *info) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(info, GH_ERR_ARG, "NULL info");
    GH_REQUIRE(c.have_pattern, GH_ERR_STATE, "no pattern: call gh_symbolic first");
    gh_pattern_info r{};
    r.nnz = c.nnz;
    r.nranges = c.slots && c.bstart.n ? (int64_t)c.bstart.n - 1 : 0;
    r.lattice = c.structured || c.q2lattice;
    r.box = c.box;
    r.elemlist = c.elemlist;
    r.compact = c.slots;
    r.metric = c.metric;
    r.inverse = c.inverse;
    r.last_kernel = c.last_kernel;
    r.fallback_nodes = c.q2lattice ? (int32_t)c.q2_nirregular : 0;
    *info = r;
  });
}

int gh_numeric(gh_ctx *ctx, int form_kind, const double *cell_coef, const double *cell_source,
               int precision, double *nzval, double *b, const double *dirichlet_vals) {
  int st = gh_coef_upload(ctx, cell_coef, cell_source, dirichlet_vals);
  if (st) return st;
  st = gh_assemble(ctx, form_kind, precision, GH_PATH_AUTO);
  if (st) return st;
  return gh_values_get(ctx, nzval, b);
}

int gh_spmv(gh_ctx *ctx, const double *x, double *y) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(!c.have_trial, GH_ERR_UNSUPPORTED, "SpMV needs a square matrix (no trial space)");
    GH_REQUIRE(c.have_values, GH_ERR_STATE, "no assembled matrix");
    GH_REQUIRE(x && y, GH_ERR_ARG, "NULL vector");
    c.s_x.resize(c.nfree);
    c.s_q.resize(c.nfree);
    GH_CHECK_HIP(hipMemcpyAsync(c.s_x.get(), x, sizeof(double) * c.nfree,
                                hipMemcpyHostToDevice, c.stream));
    gh::spmv_device(c, c.s_x.get(), c.s_q.get());
    GH_CHECK_HIP(hipMemcpyAsync(y, c.s_q.get(), sizeof(double) * c.nfree,
                                hipMemcpyDeviceToHost, c.stream));
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
  });
}

int gh_pcg(gh_ctx *ctx, int precond, const double *b, double *x, double rtol, int maxit,
           gh_solve_stats *stats) {
  int st = guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(!c.have_trial, GH_ERR_UNSUPPORTED, "PCG needs a square matrix (no trial space)");
    GH_REQUIRE(c.have_values, GH_ERR_STATE, "no assembled matrix");
    GH_REQUIRE(x, GH_ERR_ARG, "NULL solution vector");
    GH_REQUIRE(precond == GH_PRECOND_NONE || precond == GH_PRECOND_JACOBI ||
                   precond == GH_PRECOND_FSAI,
               GH_ERR_UNSUPPORTED, "unknown preconditioner");
    GH_REQUIRE(rtol > 0 && maxit >= 0, GH_ERR_ARG, "bad tolerance or iteration cap");
    size_t n = c.nfree;
    c.s_x.resize(std::max<size_t>(n, 1));
    c.s_b.resize(std::max<size_t>(n, 1));
    GH_CHECK_HIP(hipMemcpyAsync(c.s_x.get(), x, sizeof(double) * n, hipMemcpyHostToDevice,
                                c.stream));
    if (b)
      GH_CHECK_HIP(hipMemcpyAsync(c.s_b.get(), b, sizeof(double) * n, hipMemcpyHostToDevice,
                                  c.stream));
    else
      GH_CHECK_HIP(hipMemcpyAsync(c.s_b.get(), c.b.get(), sizeof(double) * n,
                                  hipMemcpyDeviceToDevice, c.stream));
    if (precond == GH_PRECOND_FSAI && !c.fsai_built) gh::fsai_forward_build(c, true);
    gh_solve_stats local{};
    gh::pcg_device(c, precond, rtol, maxit, &local);
    local.precond_used = precond;
    if (!local.converged && precond == GH_PRECOND_FSAI) {
      // Jacobi-PCG from the caller's initial guess is the fallback of the learned
      // preconditioner; the reported time covers both attempts.
      const double first_ms = local.device_ms;
      const int first_it = local.iterations;
      GH_CHECK_HIP(hipMemcpyAsync(c.s_x.get(), x, sizeof(double) * n, hipMemcpyHostToDevice,
                                  c.stream));
      gh::pcg_device(c, GH_PRECOND_JACOBI, rtol, maxit, &local);
      local.precond_used = GH_PRECOND_JACOBI;
      local.device_ms += first_ms;
      local.iterations += first_it;
    }
    GH_CHECK_HIP(hipMemcpyAsync(x, c.s_x.get(), sizeof(double) * n, hipMemcpyDeviceToHost,
                                c.stream));
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
    if (stats) *stats = local;
    if (!local.converged)
      throw gh::Error(GH_ERR_NOCONV, "PCG did not reach the requested tolerance");
  });
  return st;
}

int gh_pcg_noalloc(gh_ctx *ctx, int precond, const double *b, double *x, double rtol,
                   int maxit, gh_solve_stats *stats) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(!c.have_trial, GH_ERR_UNSUPPORTED, "PCG needs a square matrix (no trial space)");
    GH_REQUIRE(c.have_values, GH_ERR_STATE, "no assembled matrix");
    GH_REQUIRE(x, GH_ERR_ARG, "NULL solution vector");
    GH_REQUIRE(rtol > 0 && maxit >= 0, GH_ERR_ARG, "bad tolerance or iteration cap");
    gh_solve_stats local{};
    gh::pcg_noalloc(c, precond, b, x, rtol, maxit, &local);
    if (stats) *stats = local;
    if (!local.converged)
      throw gh::Error(GH_ERR_NOCONV, "PCG did not reach the requested tolerance");
  });
}

int gh_mlp_nparams(int64_t *n) {
  if (!n) return GH_ERR_ARG;
  *n = gh::kMlpParams;
  return GH_OK;
}

namespace {
void mlp_alloc(gh::Ctx &c) {
  c.params.resize(gh::kMlpParams);
  c.grads.resize(gh::kMlpParams);
  c.adam_m.resize(gh::kMlpParams);
  c.adam_v.resize(gh::kMlpParams);
  c.params_bf16.resize(gh::kMlpParams);
}
void mlp_reset_optimizer(gh::Ctx &c) {
  GH_CHECK_HIP(hipMemsetAsync(c.adam_m.get(), 0, sizeof(float) * gh::kMlpParams, c.stream));
  GH_CHECK_HIP(hipMemsetAsync(c.adam_v.get(), 0, sizeof(float) * gh::kMlpParams, c.stream));
  GH_CHECK_HIP(hipMemsetAsync(c.grads.get(), 0, sizeof(float) * gh::kMlpParams, c.stream));
  c.adam_t = 0;
}
}  // namespace

int gh_mlp_init(gh_ctx *ctx, uint64_t seed) {
  return guarded(ctx, [&](gh::Ctx &c) {
    std::vector<float> w(gh::kMlpParams, 0.0f);
    uint64_t state = seed;
    auto uniform = [&]() {  // splitmix64 -> [-1, 1)
      uint64_t x = (state += 0x9E3779B97F4A7C15ull);
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      x ^= x >> 31;
      return (float)((double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0);
    };
    auto xavier = [&](int off, int nout, int nin) {
      const float a = std::sqrt(6.0f / (float)(nin + nout));
      for (int k = 0; k < nout * nin; ++k) w[off + k] = a * uniform();
    };
    xavier(gh::kMlpW1, gh::kMlpHidden, gh::kMlpIn);
    xavier(gh::kMlpW2, gh::kMlpHidden, gh::kMlpHidden);
    xavier(gh::kMlpW3, gh::kMlpHidden, gh::kMlpHidden);
    // last layer zero: Ĝ = I, the preconditioner starts as Jacobi
    mlp_alloc(c);
    GH_CHECK_HIP(hipMemcpyAsync(c.params.get(), w.data(), sizeof(float) * w.size(),
                                hipMemcpyHostToDevice, c.stream));
    mlp_reset_optimizer(c);
    gh::mlp_sync_bf16(c);
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
    c.have_mlp = true;
    c.fsai_built = false;
  });
}

int gh_mlp_set(gh_ctx *ctx, const float *params, int64_t n) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(params && n == gh::kMlpParams, GH_ERR_ARG, "expected gh_mlp_nparams() weights");
    for (int64_t k = 0; k < n; ++k)
      GH_REQUIRE(std::isfinite(params[k]), GH_ERR_ARG, "non-finite network weight");
    mlp_alloc(c);
    GH_CHECK_HIP(hipMemcpyAsync(c.params.get(), params, sizeof(float) * n, hipMemcpyHostToDevice,
                                c.stream));
    mlp_reset_optimizer(c);
    gh::mlp_sync_bf16(c);
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
    c.have_mlp = true;
    c.fsai_built = false;
  });
}

int gh_mlp_get(gh_ctx *ctx, float *params, int64_t n) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.have_mlp, GH_ERR_STATE, "no network weights");
    GH_REQUIRE(params && n == gh::kMlpParams, GH_ERR_ARG, "expected gh_mlp_nparams() weights");
    GH_CHECK_HIP(hipMemcpyAsync(params, c.params.get(), sizeof(float) * n, hipMemcpyDeviceToHost,
                                c.stream));
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
  });
}

int gh_grf_kappa(gh_ctx *ctx, int nmodes, const double *omega, const double *phase,
                 double amp) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.dim > 0, GH_ERR_STATE, "upload a mesh first");
    gh::grf_kappa(c, nmodes, omega, phase, amp);
    c.have_values = false;
    c.fsai_built = false;
  });
}

int gh_fsai_build(gh_ctx *ctx) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(!c.have_trial, GH_ERR_UNSUPPORTED, "FSAI needs a square matrix (no trial space)");
    GH_REQUIRE(c.have_values, GH_ERR_STATE, "assemble the matrix first");
    gh::fsai_forward_build(c, true);
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
  });
}

int gh_fsai_train_step(gh_ctx *ctx, int nprobes, uint64_t seed, float lr, double *loss) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(!c.have_trial, GH_ERR_UNSUPPORTED, "FSAI needs a square matrix (no trial space)");
    GH_REQUIRE(c.have_values, GH_ERR_STATE, "assemble the matrix first");
    GH_REQUIRE(lr >= 0.0f && std::isfinite(lr), GH_ERR_ARG, "bad learning rate");
    const double l = gh::fsai_train_step(c, nprobes, seed, lr);
    if (loss) *loss = l;
  });
}

int gh_fsai_get(gh_ctx *ctx, float *feat, float *yout, float *gl, float *gu, float *grad) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.fsai_ready && c.have_mlp, GH_ERR_STATE, "no FSAI state: build or train first");
    auto cp = [&](void *dst, const void *src, size_t bytes) {
      if (dst) GH_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
    };
    const size_t C = (size_t)c.ncells;
    std::vector<float> yt(yout ? C * gh::kMlpOut : 0);
    cp(feat, c.feat.get(), sizeof(float) * C * gh::kMlpInPad);
    cp(yout ? yt.data() : nullptr, c.yo.get(), sizeof(float) * C * gh::kMlpOut);
    cp(gl, c.gl.get(), sizeof(float) * c.nnz);
    cp(gu, c.gu.get(), sizeof(float) * c.nnz);
    cp(grad, c.grads.get(), sizeof(float) * gh::kMlpParams);
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
    if (yout)  // device layout is output-major; hand back cells x kMlpOutPad
      for (size_t cell = 0; cell < C; ++cell)
        for (int n = 0; n < gh::kMlpOutPad; ++n)
          yout[cell * gh::kMlpOutPad + n] = n < gh::kMlpOut ? yt[(size_t)n * C + cell] : 0.0f;
  });
}

int gh_fsai_factor_get(gh_ctx *ctx, float *gle, int32_t *gli, float *gue, int32_t *gui) {
  return guarded(ctx, [&](gh::Ctx &c) {
    GH_REQUIRE(c.fsai_built, GH_ERR_STATE, "no FSAI factor: call gh_fsai_build");
    const size_t n = (size_t)c.nfree * gh::kFsaiEllSlots;
    auto cp = [&](void *dst, const void *src, size_t bytes) {
      if (dst) GH_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
    };
    cp(gle, c.gle.get(), sizeof(float) * n);
    cp(gli, c.gli.get(), sizeof(int32_t) * n);
    cp(gue, c.gue.get(), sizeof(float) * n);
    cp(gui, c.gui.get(), sizeof(int32_t) * n);
    GH_CHECK_HIP(hipStreamSynchronize(c.stream));
  });
}

}  // extern "C"
