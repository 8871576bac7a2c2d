"""ctypes front end of the CPU FP64 oracle (libgridap_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests, the smoke check and the
benchmark's CPU-baseline leg, never by the product package.  Each wrapper
names the restated Gridap routine; the C file cites the reference lines.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgridap_oracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle with its Makefile (gcc only)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        i64, i32, dbl = C.c_int64, C.c_int, C.c_double
        vp = C.c_void_p
        L.gor_gauss_legendre01.argtypes = [i32, _f64p, _f64p]
        L.gor_tensor_quadrature.argtypes = [i32, i32, _f64p, _f64p]
        L.gor_qk_num_nodes.argtypes = [i32, i32]
        L.gor_qk_nodes.argtypes = [i32, i32, _f64p]
        L.gor_qk_tabulate.argtypes = [i32, i32, i32, _f64p, _f64p, _f64p]
        L.gor_cart_num_nodes.argtypes = [i32, _i64p]
        L.gor_cart_num_nodes.restype = i64
        L.gor_cart_num_cells.argtypes = [i32, _i64p]
        L.gor_cart_num_cells.restype = i64
        L.gor_cart_node_coords.argtypes = [i32, _f64p, _f64p, _i64p, _f64p]
        L.gor_cart_cell_nodes.argtypes = [i32, _i64p, _i32p]
        L.gor_ncube_face_id.argtypes = [i32, i32, i32]
        L.gor_cart_vertex_labels.argtypes = [i32, _i64p, _i32p]
        L.gor_face_tag_index.argtypes = [i64, _i32p, i32, _i32p, _i32p, _i8p]
        L.gor_clagrangian_numbering.argtypes = [i64, _i8p, i32, _u8p, _i32p,
                                                C.POINTER(i64), C.POINTER(i64)]
        L.gor_clagrangian_cell_dofs.argtypes = [i64, i32, _i32p, i32, _i32p, _i32p]
        L.gor_cell_matrices_diffusion.argtypes = [i32, i64, i32, _i32p, _f64p, i32,
                                                  _f64p, _f64p, vp, _f64p]
        L.gor_cell_matrices_diffusion_cartesian.argtypes = [i32, _f64p, i64, i32, i32,
                                                            _f64p, _f64p, vp, _f64p]
        L.gor_cell_vectors_source.argtypes = [i32, i64, i32, _i32p, _f64p, i32, _f64p,
                                              _f64p, _f64p, vp, _f64p]
        L.gor_cell_vectors_source_cartesian.argtypes = [i32, _f64p, i64, i32, i32, _f64p,
                                                        _f64p, vp, _f64p]
        L.gor_cell_vectors_source_qp.argtypes = [i32, i64, i32, _i32p, _f64p, vp, i32, _f64p,
                                                 _f64p, _f64p, _f64p, _f64p]
        L.gor_attach_dirichlet.argtypes = [i64, i32, _i32p, _f64p, _f64p, _f64p]
        L.gor_csc_from_entries.argtypes = [i64, _i64p, _i64p, _f64p, _u8p, i64, _i64p, _i64p,
                                           _i64p, _f64p]
        L.gor_csc_from_entries.restype = i64
        L.gor_csc_count.argtypes = [i64, i32, _i32p, i64, _i64p]
        L.gor_csc_count.restype = i64
        L.gor_csc_assemble.argtypes = [i64, i32, _i32p, _f64p, i64, i64, _i64p, _i64p,
                                       _f64p]
        L.gor_csc_assemble.restype = i64
        L.gor_assemble_vector.argtypes = [i64, i32, _i32p, _f64p, i64, _f64p]
        L.gor_numeric_refill_cartesian.argtypes = [i32, _f64p, i64, i32, _i32p, i32, _f64p,
                                                   _f64p, _f64p, vp, vp, i64, _i64p, _i64p,
                                                   _f64p, _f64p]
        L.gor_numeric_refill_cartesian_par.argtypes = [i32, _f64p, _i64p, i32, _i32p, i32,
                                                       _f64p, _f64p, _f64p, vp, vp, i64,
                                                       _i64p, _i64p, _f64p, _f64p, i32]
        L.gor_csc_spmv.argtypes = [i64, _i64p, _i64p, _f64p, _f64p, _f64p]
        L.gor_pcg_jacobi.argtypes = [i64, _i64p, _i64p, _f64p, _f64p, _f64p, dbl, i32,
                                     C.POINTER(C.c_int), C.POINTER(dbl)]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- quadrature / reference FE -------------------------------------------

def gauss_legendre01(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().gor_gauss_legendre01(n, x, w)
    return x, w


def tensor_quadrature(D, degree):
    n = (degree // 2 + 1) ** D
    x = np.zeros(n * D)
    w = np.zeros(n)
    assert lib().gor_tensor_quadrature(D, degree, x, w) == 0
    return x.reshape(n, D), w


def qk_nodes(D, order):
    nn = lib().gor_qk_num_nodes(D, order)
    x = np.zeros(nn * D)
    lib().gor_qk_nodes(D, order, x)
    return x.reshape(nn, D)


def qk_tabulate(D, order, pts):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    npts = pts.shape[0]
    nn = lib().gor_qk_num_nodes(D, order)
    shp = np.zeros(npts * nn)
    grad = np.zeros(npts * nn * D)
    lib().gor_qk_tabulate(D, order, npts, pts.ravel(), shp, grad)
    return shp.reshape(npts, nn), grad.reshape(npts, nn, D)


# ---- Cartesian model / numbering -----------------------------------------

def cartesian_model(domain, partition):
    D = len(partition)
    part = np.asarray(partition, dtype=np.int64)
    origin = np.array([domain[2 * d] for d in range(D)], dtype=np.float64)
    sizes = np.array([(domain[2 * d + 1] - domain[2 * d]) / partition[d] for d in range(D)])
    L = lib()
    nn = L.gor_cart_num_nodes(D, part)
    nc = L.gor_cart_num_cells(D, part)
    xyz = np.zeros(nn * D)
    L.gor_cart_node_coords(D, origin, sizes, part, xyz)
    cn = np.zeros(nc * (1 << D), dtype=np.int32)
    L.gor_cart_cell_nodes(D, part, cn)
    lab = np.zeros(nn, dtype=np.int32)
    L.gor_cart_vertex_labels(D, part, lab)
    return dict(D=D, origin=origin, sizes=sizes, partition=part,
                xyz=xyz.reshape(nn, D), cell_nodes=cn.reshape(nc, 1 << D),
                vertex_labels=lab)


def ncube_face_id(D, free_mask, anchor_bits):
    return lib().gor_ncube_face_id(D, free_mask, anchor_bits)


def cartesian_tag_entities(D, tag):
    """Entities of a tag of a CartesianDiscreteModel (_add_cartesian_tags!,
    src/Geometry/CartesianDiscreteModels.jl:158-170): tag ids 1..3^D-1 hold the
    entity of the same id, then "interior" and "boundary"."""
    nf = 3 ** D
    if isinstance(tag, str):
        if tag == "boundary":
            return list(range(1, nf))
        if tag == "interior":
            return [nf]
        if tag.startswith("tag_"):
            return [int(tag[4:])]
        raise KeyError(tag)
    t = int(tag)
    if 1 <= t < nf:
        return [t]
    if t == nf:
        return [nf]
    if t == nf + 1:
        return list(range(1, nf))
    raise KeyError(tag)


def face_tag_index(face_entity, tag_entities):
    ptrs = np.zeros(len(tag_entities) + 1, dtype=np.int32)
    for t, e in enumerate(tag_entities):
        ptrs[t + 1] = ptrs[t] + len(e)
    ents = np.array([x for e in tag_entities for x in e] or [0], dtype=np.int32)
    fe = np.ascontiguousarray(face_entity, dtype=np.int32)
    out = np.zeros(len(fe), dtype=np.int8)
    lib().gor_face_tag_index(len(fe), fe, len(tag_entities), ptrs, ents, out)
    return out


def clagrangian(node_to_tag, ncomp=1, masks=None):
    node_to_tag = np.ascontiguousarray(node_to_tag, dtype=np.int8)
    ntags = int(node_to_tag.max()) if len(node_to_tag) else 0
    if masks is None:
        m = np.ones(max(ntags, 1) * ncomp, dtype=np.uint8)
    else:
        m = np.array(masks, dtype=np.uint8).reshape(-1)
    out = np.zeros(len(node_to_tag) * ncomp, dtype=np.int32)
    nf, nd = C.c_int64(), C.c_int64()
    lib().gor_clagrangian_numbering(len(node_to_tag), node_to_tag, ncomp, m, out,
                                    C.byref(nf), C.byref(nd))
    return out.reshape(-1, ncomp), nf.value, nd.value


def clagrangian_cell_dofs(cell_nodes, node_comp_to_dof):
    cn = np.ascontiguousarray(cell_nodes, dtype=np.int32)
    ncomp = node_comp_to_dof.shape[1]
    out = np.zeros(cn.shape[0] * cn.shape[1] * ncomp, dtype=np.int32)
    lib().gor_clagrangian_cell_dofs(cn.shape[0], cn.shape[1], cn.ravel(), ncomp,
                                    np.ascontiguousarray(node_comp_to_dof.ravel()), out)
    return out.reshape(cn.shape[0], cn.shape[1] * ncomp)


# ---- element integration / assembly --------------------------------------

def cell_matrices_diffusion(D, cell_nodes, xyz, qw, grad, kappa=None, cartesian_sizes=None):
    cn = np.ascontiguousarray(cell_nodes, dtype=np.int32)
    nc, nloc = cn.shape
    out = np.zeros(nc * nloc * nloc)
    kap = None if kappa is None else np.ascontiguousarray(kappa, dtype=np.float64)
    g = np.ascontiguousarray(grad, dtype=np.float64).ravel()
    qw = np.ascontiguousarray(qw, dtype=np.float64)
    if cartesian_sizes is not None:
        lib().gor_cell_matrices_diffusion_cartesian(
            D, np.asarray(cartesian_sizes, dtype=np.float64), nc, nloc, len(qw), qw, g,
            _ptr(kap), out)
    else:
        lib().gor_cell_matrices_diffusion(D, nc, nloc, cn.ravel(),
                                          np.ascontiguousarray(xyz).ravel(), len(qw), qw,
                                          g, _ptr(kap), out)
    return out.reshape(nc, nloc, nloc)   # [c, lj, li] (column-major per cell)


def cell_vectors_source(D, cell_nodes, xyz, qw, shp, grad, fsrc=None, cartesian_sizes=None):
    cn = np.ascontiguousarray(cell_nodes, dtype=np.int32)
    nc, nloc = cn.shape
    out = np.zeros(nc * nloc)
    f = None if fsrc is None else np.ascontiguousarray(fsrc, dtype=np.float64)
    qw = np.ascontiguousarray(qw, dtype=np.float64)
    s = np.ascontiguousarray(shp, dtype=np.float64).ravel()
    if cartesian_sizes is not None:
        lib().gor_cell_vectors_source_cartesian(
            D, np.asarray(cartesian_sizes, dtype=np.float64), nc, nloc, len(qw), qw, s,
            _ptr(f), out)
    else:
        lib().gor_cell_vectors_source(D, nc, nloc, cn.ravel(),
                                      np.ascontiguousarray(xyz).ravel(), len(qw), qw, s,
                                      np.ascontiguousarray(grad).ravel(), _ptr(f), out)
    return out.reshape(nc, nloc)


def physical_quadrature_points(D, cell_nodes, xyz, shp):
    """x_q = Σ_a φ_a(ξ_q) x_a per cell (the isoparametric map); [ncells, npts, D]."""
    X = np.asarray(xyz)[np.asarray(cell_nodes) - 1]          # [c, a, D]
    return np.einsum("qa,cad->cqd", np.asarray(shp), X)


def cell_vectors_source_qp(D, cell_nodes, xyz, qw, shp, grad, fq, cartesian_sizes=None):
    cn = np.ascontiguousarray(cell_nodes, dtype=np.int32)
    nc, nloc = cn.shape
    out = np.zeros(nc * nloc)
    cs = None if cartesian_sizes is None else np.asarray(cartesian_sizes, dtype=np.float64)
    lib().gor_cell_vectors_source_qp(D, nc, nloc, cn.ravel(), np.ascontiguousarray(xyz).ravel(),
                                     _ptr(cs), len(qw), np.ascontiguousarray(qw),
                                     np.ascontiguousarray(shp).ravel(),
                                     np.ascontiguousarray(grad).ravel(),
                                     np.ascontiguousarray(fq, dtype=np.float64).ravel(), out)
    return out.reshape(nc, nloc)


def csc_from_entries(entries, ncols):
    """Drive counter -> allocation -> inserter -> create_from_nz with a list of
    (i, j, v) entries, v=None meaning the `nothing` (structure-only) method."""
    n = len(entries)
    I = np.array([e[0] for e in entries], dtype=np.int64)
    J = np.array([e[1] for e in entries], dtype=np.int64)
    V = np.array([0.0 if e[2] is None else e[2] for e in entries], dtype=np.float64)
    H = np.array([e[2] is not None for e in entries], dtype=np.uint8)
    cnt = np.zeros(ncols, dtype=np.int64)
    colptr = np.zeros(ncols + 1, dtype=np.int64)
    rowval = np.zeros(max(n, 1), dtype=np.int64)
    nzval = np.zeros(max(n, 1))
    nnz = lib().gor_csc_from_entries(n, I, J, V, H, ncols, cnt, colptr, rowval, nzval)
    if nnz < 0:
        raise RuntimeError("CSC inserter overflow")
    return cnt, colptr, rowval[:nnz].copy(), nzval[:nnz].copy()


def attach_dirichlet(cell_dofs, cellmats, dirichlet_vals, cellvecs):
    cd = np.ascontiguousarray(cell_dofs, dtype=np.int32)
    vecs = np.ascontiguousarray(cellvecs, dtype=np.float64).copy()
    dv = np.ascontiguousarray(dirichlet_vals, dtype=np.float64)
    if dv.size == 0:
        dv = np.zeros(1)
    lib().gor_attach_dirichlet(cd.shape[0], cd.shape[1], cd.ravel(),
                               np.ascontiguousarray(cellmats).ravel(), dv, vecs.ravel())
    return vecs


def csc_count(cell_dofs, ncols):
    cd = np.ascontiguousarray(cell_dofs, dtype=np.int32)
    cnt = np.zeros(ncols, dtype=np.int64)
    tot = lib().gor_csc_count(cd.shape[0], cd.shape[1], cd.ravel(), ncols, cnt)
    return cnt, tot


def assemble_csc(cell_dofs, cellmats, nrows, ncols):
    """MinMemory SparseMatrixCSC assembly; returns 1-based (colptr, rowval, nzval)."""
    cd = np.ascontiguousarray(cell_dofs, dtype=np.int32)
    _, tot = csc_count(cd, ncols)
    colptr = np.zeros(ncols + 1, dtype=np.int64)
    rowval = np.zeros(max(tot, 1), dtype=np.int64)
    nzval = np.zeros(max(tot, 1))
    nnz = lib().gor_csc_assemble(cd.shape[0], cd.shape[1], cd.ravel(),
                                 np.ascontiguousarray(cellmats, dtype=np.float64).ravel(),
                                 nrows, ncols, colptr, rowval, nzval)
    if nnz < 0:
        raise RuntimeError("CSC inserter overflow")
    return colptr, rowval[:nnz].copy(), nzval[:nnz].copy()


def assemble_vector(cell_dofs, cellvecs, nrows):
    cd = np.ascontiguousarray(cell_dofs, dtype=np.int32)
    b = np.zeros(nrows)
    lib().gor_assemble_vector(cd.shape[0], cd.shape[1], cd.ravel(),
                              np.ascontiguousarray(cellvecs, dtype=np.float64).ravel(),
                              nrows, b)
    return b


def numeric_refill_cartesian(D, sizes, cell_dofs, qw, shp, grad, kappa, fsrc, colptr, rowval):
    """assemble_matrix_and_vector! into an existing pattern (numeric phase only)."""
    cd = np.ascontiguousarray(cell_dofs, dtype=np.int32)
    n = len(colptr) - 1
    nzval = np.zeros(max(len(rowval), 1))
    b = np.zeros(max(n, 1))
    k = None if kappa is None else np.ascontiguousarray(kappa, dtype=np.float64)
    f = None if fsrc is None else np.ascontiguousarray(fsrc, dtype=np.float64)
    st = lib().gor_numeric_refill_cartesian(
        D, np.asarray(sizes, dtype=np.float64), cd.shape[0], cd.shape[1], cd.ravel(), len(qw),
        np.ascontiguousarray(qw), np.ascontiguousarray(shp).ravel(),
        np.ascontiguousarray(grad).ravel(), _ptr(k), _ptr(f), n, colptr, rowval, nzval, b)
    if st != 0:
        raise RuntimeError("entry missing from the pattern")
    return nzval[:len(rowval)], b[:n]


def numeric_refill_cartesian_par(D, sizes, part, cell_dofs, qw, shp, grad, kappa, fsrc,
                                 colptr, rowval, nthreads):
    """numeric_refill_cartesian on nthreads OpenMP threads (parity-coloured cells); agrees
    with the sequential loop to rounding.  Timing baseline only."""
    cd = np.ascontiguousarray(cell_dofs, dtype=np.int32)
    assert cd.shape[0] == int(np.prod(part)) and len(part) == D
    n = len(colptr) - 1
    nzval = np.zeros(max(len(rowval), 1))
    b = np.zeros(max(n, 1))
    k = None if kappa is None else np.ascontiguousarray(kappa, dtype=np.float64)
    f = None if fsrc is None else np.ascontiguousarray(fsrc, dtype=np.float64)
    st = lib().gor_numeric_refill_cartesian_par(
        D, np.asarray(sizes, dtype=np.float64), np.asarray(part, dtype=np.int64), cd.shape[1],
        cd.ravel(), len(qw), np.ascontiguousarray(qw), np.ascontiguousarray(shp).ravel(),
        np.ascontiguousarray(grad).ravel(), _ptr(k), _ptr(f), n, colptr, rowval, nzval, b,
        int(nthreads))
    if st != 0:
        raise RuntimeError("entry missing from the pattern")
    return nzval[:len(rowval)], b[:n]


def csc_spmv(colptr, rowval, nzval, x):
    y = np.zeros(len(colptr) - 1)
    lib().gor_csc_spmv(len(colptr) - 1, colptr, rowval, nzval,
                       np.ascontiguousarray(x, dtype=np.float64), y)
    return y


def pcg_jacobi(colptr, rowval, nzval, b, rtol=1e-8, maxit=10000):
    x = np.zeros(len(b))
    it, rr = C.c_int(), C.c_double()
    lib().gor_pcg_jacobi(len(b), colptr, rowval, nzval, np.ascontiguousarray(b), x, rtol,
                         maxit, C.byref(it), C.byref(rr))
    return x, it.value, rr.value


def to_scipy(colptr, rowval, nzval, n, m=None):
    import scipy.sparse as sp
    m = n if m is None else m
    return sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(m, n))


# ---- complete driver: the Poisson configurations -------------------------

def poisson_q1(domain, partition, kappa=None, fsrc=None, dirichlet_tags=("boundary",),
               g=None, degree=2, mesh="cartesian"):
    """Q1 Poisson a(u,v)=∫κ∇v⋅∇u, l(v)=∫f v with Dirichlet tags, exactly as an
    AffineFEOperator over a CartesianDiscreteModel assembles it
    (src/FESpaces/AffineFEOperators.jl:23-54 -> Assemblers.jl:496-541 ->
    SparseMatrixAssemblers.jl:99-106).  mesh="cartesian" uses the CartesianMap
    Jacobian; mesh="unstructured" the isoparametric map from node coordinates.
    g: callable x -> value for the Dirichlet data (None = homogeneous)."""
    D = len(partition)
    model = cartesian_model(domain, partition)
    ents = [cartesian_tag_entities(D, t) for t in dirichlet_tags]
    node_to_tag = face_tag_index(model["vertex_labels"], ents)
    n2d, nfree, ndiri = clagrangian(node_to_tag)
    cell_dofs = clagrangian_cell_dofs(model["cell_nodes"], n2d)
    qx, qw = tensor_quadrature(D, degree)
    shp, grad = qk_tabulate(D, 1, qx)
    cs = model["sizes"] if mesh == "cartesian" else None
    mats = cell_matrices_diffusion(D, model["cell_nodes"], model["xyz"], qw, grad, kappa,
                                   cartesian_sizes=cs)
    vecs = cell_vectors_source(D, model["cell_nodes"], model["xyz"], qw, shp, grad, fsrc,
                               cartesian_sizes=cs)
    diri_vals = np.zeros(ndiri)
    if g is not None:
        dnodes = np.nonzero(n2d[:, 0] < 0)[0]
        ids = -n2d[dnodes, 0] - 1
        diri_vals[ids] = [g(model["xyz"][n]) for n in dnodes]
        vecs = attach_dirichlet(cell_dofs, mats, diri_vals, vecs)
    colptr, rowval, nzval = assemble_csc(cell_dofs, mats, nfree, nfree)
    b = assemble_vector(cell_dofs, vecs, nfree)
    return dict(model=model, node_to_dof=n2d, nfree=nfree, ndiri=ndiri,
                cell_dofs=cell_dofs, qx=qx, qw=qw, shp=shp, grad=grad, cellmats=mats,
                cellvecs=vecs, dirichlet_vals=diri_vals, colptr=colptr, rowval=rowval,
                nzval=nzval, b=b)


# ---- nonlinear forms: residual and Jacobian at an FE function --------------------------

def cell_values(cell_dofs, free_vals, dirichlet_vals):
    """The cell-wise dof values of an FE function: free values at ids > 0, Dirichlet
    values at ids < 0 (scatter_free_and_dirichlet_values,
    src/FESpaces/FESpaceInterface.jl, with Gridap's signed dof ids)."""
    cd = np.asarray(cell_dofs, dtype=np.int64)
    fv = np.asarray(free_vals, dtype=np.float64)
    dv = np.asarray(dirichlet_vals, dtype=np.float64) if dirichlet_vals is not None else None
    out = np.zeros(cd.shape)
    pos = cd > 0
    out[pos] = fv[cd[pos] - 1]
    if dv is not None and dv.size:
        out[~pos] = dv[-cd[~pos] - 1]
    return out


def plaplacian_cells(D, cell_nodes, xyz, qw, shp, grad, u_cell, p, kappa=None, fsrc=None,
                     cartesian_sizes=None):
    """Cell residuals and Jacobians of the p-Laplacian of
    test/GridapTests/PLaplacianTests.jl:15-17,31-32 (flux(∇u) = |∇u|^(p-2) ∇u,
    dflux(∇du,∇u) = (p-2)|∇u|^(p-4)(∇u⋅∇du)∇u + |∇u|^(p-2)∇du), scaled by κ per cell,
    with the source term -∫ f v, at the FE function whose cell values are u_cell:
        r_a  = Σ_q w|J| (κ ∇φ_a⋅flux(∇u) - f φ_a)
        J_ab = Σ_q w|J| κ ∇φ_a⋅dflux(∇φ_b, ∇u)
    (residual!/jacobian! of src/FESpaces/FEOperatorsFromWeakForm.jl:57-95, integrated
    as IntegrationMap, src/Fields/FieldsInterfaces.jl:737-775; the geometry as in
    cell_matrices_diffusion: the isoparametric Jt, or diag(sizes) for CartesianMap).
    Returns (cellmats [c, lj, li], cellvecs [c, a])."""
    cn = np.asarray(cell_nodes, dtype=np.int64) - 1
    nc, nl = cn.shape
    qw = np.asarray(qw, dtype=np.float64)
    shp = np.asarray(shp, dtype=np.float64).reshape(len(qw), nl)
    grad = np.asarray(grad, dtype=np.float64).reshape(len(qw), nl, D)
    if cartesian_sizes is not None:
        Jt = np.broadcast_to(np.diag(np.asarray(cartesian_sizes, dtype=np.float64)[:D]),
                             (nc, len(qw), D, D))
    else:
        X = np.asarray(xyz, dtype=np.float64)[cn]                       # [c, a, D]
        Jt = np.einsum("qai,caj->cqij", grad, X)
    Ji = np.linalg.inv(Jt)
    dV = np.abs(np.linalg.det(Jt)) * qw[None, :]                         # [c, q]
    g = np.einsum("cqrk,qak->cqar", Ji, grad)                            # [c, q, a, r]
    gu = np.einsum("ca,cqar->cqr", np.asarray(u_cell, dtype=np.float64), g)
    n2 = np.sum(gu * gu, axis=-1)
    nrm = np.sqrt(n2)
    with np.errstate(divide="ignore", invalid="ignore"):
        s0 = np.ones_like(nrm) if p == 2 else np.where(nrm > 0, nrm ** (p - 2), 0.0)
        s1 = np.zeros_like(nrm) if p == 2 else np.where(nrm > 0, (p - 2) * nrm ** (p - 4), 0.0)
    kap = np.ones(nc) if kappa is None else np.asarray(kappa, dtype=np.float64)
    f = np.zeros(nc) if fsrc is None else np.asarray(fsrc, dtype=np.float64)
    ga = np.einsum("cqr,cqar->cqa", gu, g)                                # ∇u⋅∇φ_a
    r = np.einsum("cq,cqa->ca", dV * kap[:, None] * s0, ga) - \
        np.einsum("cq,qa->ca", dV * f[:, None], shp)
    K = np.einsum("cq,cqar,cqbr->cab", dV * kap[:, None] * s0, g, g) + \
        np.einsum("cq,cqa,cqb->cab", dV * kap[:, None] * s1, ga, ga)
    return np.ascontiguousarray(K.transpose(0, 2, 1)), r                 # [c, lj, li]


def plaplacian_system(cell_dofs, nfree, cellmats, cellvecs):
    """Global Jacobian (CSC) and residual from the cell arrays (assemble_matrix! and
    assemble_vector!, Dirichlet rows and columns skipped; no lifting)."""
    colptr, rowval, nzval = assemble_csc(cell_dofs, cellmats, nfree, nfree)
    r = assemble_vector(cell_dofs, cellvecs, nfree)
    return colptr, rowval, nzval, r


# ---- boundary integrals ------------------------------------------------------------------

def boundary_faces(model, tag_entities):
    """The boundary facets of a Cartesian model whose entity is in tag_entities, found
    from the vertex labels: a cell facet (the 2^(D-1) local vertices with one local
    coordinate fixed) whose vertices all lie on one box side; its entity is that box
    side's face id (BoundaryTriangulation(model, tags), src/Geometry/
    BoundaryTriangulations.jl:122-183, with FaceLabeling entities,
    src/Geometry/CartesianDiscreteModels.jl:137-170).  Returns [(cell, local vertex
    list, fixed axis, fixed value)] in cell order, facets of one cell in axis order."""
    D = model["D"]
    xyz, cn = model["xyz"], np.asarray(model["cell_nodes"]) - 1
    lo = xyz.min(0)
    hi = xyz.max(0)
    out = []
    ents = set(tag_entities)
    for c in range(cn.shape[0]):
        for axis in range(D):
            for val in (0, 1):
                loc = [a for a in range(1 << D) if ((a >> axis) & 1) == val]
                X = xyz[cn[c, loc]]
                side = lo[axis] if val == 0 else hi[axis]
                if not np.all(np.abs(X[:, axis] - side) < 1e-12):
                    continue
                free = ((1 << D) - 1) & ~(1 << axis)
                ent = ncube_face_id(D, free, val << axis)
                if ent in ents:
                    out.append((c, loc, axis, val))
    return out


def boundary_face_arrays(model, faces, degree, alpha=0.0, beta=0.0, g=None, h=None, order=1):
    """Face matrices [f, lj, li] and vectors [f, a] (cell-local numbering) of
        ∫_F α φ_a φ_b + β (φ_a n⋅∇φ_b + n⋅∇φ_a φ_b)   and   ∫_F φ_a g + n⋅∇φ_a h
    for Q_order cells on the Q1 map of their vertices: the facet quadrature (tensor
    Gauss of `degree` on the (D-1)-cube) is mapped into the cell's reference
    coordinates, the cell basis and its gradients pushed forward by the Q1 map's
    Jacobian are evaluated there, the facet measure is sqrt(det(TᵀT)) of the facet
    tangents T, and the normal is the unit normal of the facet oriented away from the
    cell centre (the geometry of src/Geometry/BoundaryTriangulations.jl:122-183 and
    get_normal_vector).  g, h: callables (x, n) -> value (None = 0)."""
    D = model["D"]
    xyz, cn = model["xyz"], np.asarray(model["cell_nodes"]) - 1
    x1, w1 = gauss_legendre01(degree // 2 + 1)
    if D == 2:
        eta, we = x1[:, None], w1
    else:                                                  # first facet coordinate fastest
        eta = np.array([[x1[i], x1[j]] for j in range(len(x1)) for i in range(len(x1))])
        we = np.array([w1[i] * w1[j] for j in range(len(x1)) for i in range(len(x1))])
    nl = (order + 1) ** D
    mats = np.zeros((len(faces), nl, nl))
    vecs = np.zeros((len(faces), nl))
    for k, (c, loc, axis, val) in enumerate(faces):
        others = [d for d in range(D) if d != axis]
        xi = np.zeros((len(we), D))
        xi[:, others] = eta
        xi[:, axis] = val
        shp1, grad1 = qk_tabulate(D, 1, xi)               # the Q1 cell map
        shp, grad = qk_tabulate(D, order, xi)             # [q, a], [q, a, D]
        X = xyz[cn[c]]                                     # [v, D]
        centre = X.mean(0)
        for q in range(len(we)):
            Jt = grad1[q].T @ X                            # Jt[i, j] = ∂x_j/∂ξ_i
            gphys = np.linalg.solve(Jt, grad[q].T).T       # [a, D]
            T = Jt[others]                                 # facet tangents (rows)
            dS = np.sqrt(np.linalg.det(T @ T.T)) * we[q]
            if D == 2:
                n = np.array([T[0, 1], -T[0, 0]])
            else:
                n = np.cross(T[0], T[1])
            n = n / np.linalg.norm(n)
            x = shp1[q] @ X
            if np.dot(n, x - centre) < 0:
                n = -n
            dn = gphys @ n
            phi = shp[q]
            mats[k] += dS * (alpha * np.outer(phi, phi) +
                             beta * (np.outer(phi, dn) + np.outer(dn, phi)))
            gv = 0.0 if g is None else g(x, n)
            hv = 0.0 if h is None else h(x, n)
            vecs[k] += dS * (phi * gv + dn * hv)
    # [f, li, lj] -> [f, lj, li]
    return np.ascontiguousarray(mats.transpose(0, 2, 1)), vecs


def poisson_q1_with_boundary(domain, partition, dirichlet_entities, face_terms, degree=2,
                             face_degree=2, g=None, fsrc=None, kappa=None):
    """Q1 Poisson with boundary terms, assembled as Gridap's AffineFEOperator does with
    a weak form over Ω and boundary triangulations: cell arrays then face arrays in one
    assembly (the cell contributions first), Dirichlet lifting of both
    (src/FESpaces/Assemblers.jl:485-541, src/CellData/AttachDirichlet.jl).
    dirichlet_entities: the entity ids of the Dirichlet tag; face_terms: [(entity ids,
    alpha, beta, g(x, n), h(x, n))]; kappa: the diffusion coefficient per cell (1)."""
    D = len(partition)
    model = cartesian_model(domain, partition)
    ents = sorted(set(dirichlet_entities))
    node_to_tag = face_tag_index(model["vertex_labels"], [ents] if ents else [])
    n2d, nfree, ndiri = clagrangian(node_to_tag)
    cd = clagrangian_cell_dofs(model["cell_nodes"], n2d)
    qx, qw = tensor_quadrature(D, degree)
    shp, grad = qk_tabulate(D, 1, qx)
    mats = cell_matrices_diffusion(D, model["cell_nodes"], model["xyz"], qw, grad, kappa=kappa,
                                   cartesian_sizes=model["sizes"])
    vecs = cell_vectors_source(D, model["cell_nodes"], model["xyz"], qw, shp, grad, fsrc,
                               cartesian_sizes=model["sizes"])
    all_cd, all_m, all_v = [cd], [mats], [vecs]
    for tents, al, be, gf, hf in face_terms:
        faces = boundary_faces(model, tents)
        fm, fv = boundary_face_arrays(model, faces, face_degree, al, be, gf, hf)
        all_cd.append(cd[[f[0] for f in faces]])
        all_m.append(fm)
        all_v.append(fv)
    cd_all = np.ascontiguousarray(np.concatenate(all_cd), dtype=np.int32)
    m_all = np.ascontiguousarray(np.concatenate(all_m))
    v_all = np.ascontiguousarray(np.concatenate(all_v))
    dvals = np.zeros(ndiri)
    if g is not None and ndiri:
        dnodes = np.nonzero(n2d[:, 0] < 0)[0]
        dvals[-n2d[dnodes, 0] - 1] = [g(model["xyz"][n]) for n in dnodes]
        v_all = attach_dirichlet(cd_all, m_all, dvals, v_all)
    colptr, rowval, nzval = assemble_csc(cd_all, m_all, nfree, nfree)
    b = assemble_vector(cd_all, v_all, nfree)
    return dict(model=model, node_to_dof=n2d, nfree=nfree, ndiri=ndiri, cell_dofs=cd,
                colptr=colptr, rowval=rowval, nzval=nzval, b=b, dirichlet_vals=dvals)
