"""Parity of the HIP assembly engine with the FP64 oracle (needs an MI355X).

Bar: CSC pattern bit-exact (1-based Int64 colptr/rowval equal to the oracle's
create_from_nz output); FP64 element path: nzval and b within 1e-12 relative
(max-abs error / max-abs value); FP32 element path: relative Frobenius error
<= 1e-6.  At the full 128^3 size the checks are size-independent properties.
"""
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import gridap_hip as G
from gridap_hip import _lib

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_FROB_TOL = 1e-6


def announce(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    assert a.shape == b.shape
    if a.size == 0:
        return 0.0
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def relfrob(a, b):
    assert a.shape == b.shape
    if a.size == 0:
        return 0.0
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def build(domain, part, kappa=None, fsrc=None, tags=("boundary",), g=None, map=None,
          precision="fp64", path="auto", degree=2):
    model = G.CartesianDiscreteModel(domain, part, map=map)
    reffe = G.ReferenceFE(G.lagrangian, G.Float64, 1)
    V = G.TestFESpace(model, reffe, dirichlet_tags=list(tags))
    U = G.TrialFESpace(V, g)
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, degree)
    k = 1.0 if kappa is None else G.CellField(kappa, Ω)
    f = 1.0 if fsrc is None else G.CellField(fsrc, Ω)
    a = lambda u, v: G.integrate(k * G.inner(G.gradient(v), G.gradient(u)), dΩ)
    l = lambda v: G.integrate(f * v, dΩ)
    assem = G.SparseMatrixAssembler(U, V, precision=precision, path=path)
    op = G.AffineFEOperator(a, l, U, V, assem)
    return op, V, U, model


def oracle_problem(oracle, domain, part, kappa=None, fsrc=None, tags=("boundary",), g=None,
                   mesh="cartesian", xyz=None):
    if xyz is None:
        return oracle.poisson_q1(domain, part, kappa=kappa, fsrc=fsrc, dirichlet_tags=tags,
                                 g=g, mesh=mesh)
    # mapped grid: same topology, moved nodes, isoparametric map
    D = len(part)
    m = oracle.cartesian_model(domain, part)
    m["xyz"] = xyz
    ents = [oracle.cartesian_tag_entities(D, t) for t in tags]
    n2d, nfree, ndiri = oracle.clagrangian(oracle.face_tag_index(m["vertex_labels"], ents))
    cd = oracle.clagrangian_cell_dofs(m["cell_nodes"], n2d)
    qx, qw = oracle.tensor_quadrature(D, 2)
    shp, grad = oracle.qk_tabulate(D, 1, qx)
    mats = oracle.cell_matrices_diffusion(D, m["cell_nodes"], xyz, qw, grad, kappa)
    vecs = oracle.cell_vectors_source(D, m["cell_nodes"], xyz, qw, shp, grad, fsrc)
    if g is not None and ndiri:
        dv = np.zeros(ndiri)
        dnodes = np.nonzero(n2d[:, 0] < 0)[0]
        dv[-n2d[dnodes, 0] - 1] = [g(xyz[n]) for n in dnodes]
        vecs = oracle.attach_dirichlet(cd, mats, dv, vecs)
    colptr, rowval, nzval = oracle.assemble_csc(cd, mats, nfree, nfree)
    b = oracle.assemble_vector(cd, vecs, nfree)
    return dict(colptr=colptr, rowval=rowval, nzval=nzval, b=b, nfree=nfree)


def check_against(oracle_res, op, precision):
    A, b = op.get_matrix(), op.get_vector()
    np.testing.assert_array_equal(A.colptr, oracle_res["colptr"])
    np.testing.assert_array_equal(A.rowval, oracle_res["rowval"])
    if precision == "fp64":
        assert relmax(A.nzval, oracle_res["nzval"]) <= FP64_TOL
        assert relmax(b, oracle_res["b"]) <= FP64_TOL
    else:
        assert relfrob(A.nzval, oracle_res["nzval"]) <= FP32_FROB_TOL
        assert relfrob(b, oracle_res["b"]) <= FP32_FROB_TOL


# --------------------------------------------------------------------------

@pytest.mark.parametrize("path", ["auto", "general"])
def test_config1_2d_q1_fp64(oracle, path):
    """2-D Poisson, 64x64 QUAD, Q1, degree 2, κ=1, f=1, Dirichlet "boundary", FP64."""
    announce("config1 64x64 fp64 path=%s" % path)
    dom, part = (0, 1, 0, 1), (64, 64)
    op, V, U, _ = build(dom, part, path=path)
    ref = oracle_problem(oracle, dom, part)
    assert op.get_matrix().nnz() == 34969
    check_against(ref, op, "fp64")
    announce("config1 solve")
    uh = G.solve(op, G.PCGSolver(rtol=1e-12))
    A = oracle.to_scipy(ref["colptr"], ref["rowval"], ref["nzval"], ref["nfree"])
    xd = spla.spsolve(A.tocsc(), ref["b"])
    assert relmax(uh.free_values, xd) <= 1e-9


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("path", ["auto", "general"])
def test_3d_lognormal_kappa(oracle, precision, path):
    announce("3d 16x12x10 kappa precision=%s path=%s" % (precision, path))
    rng = np.random.default_rng(11)
    dom, part = (0, 1, 0, 1.5, 0, 0.8), (16, 12, 10)
    kappa = np.exp(rng.normal(size=int(np.prod(part))))
    fsrc = rng.uniform(0.5, 2.0, size=int(np.prod(part)))
    op, *_ = build(dom, part, kappa=kappa, fsrc=fsrc, precision=precision, path=path)
    assert op.assem.ctx is not None
    ref = oracle_problem(oracle, dom, part, kappa=kappa, fsrc=fsrc)
    check_against(ref, op, precision)


def test_structured_and_general_paths_agree(oracle):
    announce("3d 20^3 structured vs general fp64")
    rng = np.random.default_rng(12)
    part = (20, 20, 20)
    kappa = np.exp(rng.normal(size=8000))
    op1, *_ = build((0, 1) * 3, part, kappa=kappa, path="structured")
    op2, *_ = build((0, 1) * 3, part, kappa=kappa, path="general")
    A1, A2 = op1.get_matrix(), op2.get_matrix()
    np.testing.assert_array_equal(A1.rowval, A2.rowval)
    assert relmax(A1.nzval, A2.nzval) <= 1e-14
    assert relmax(op1.get_vector(), op2.get_vector()) <= 1e-14


def test_nonhomogeneous_dirichlet_reproduces_linear_solution(oracle):
    """AffineFEOperatorsTests.jl:52-86 in 3-D: u linear, f = 0, κ constant (a linear u
    solves -∇⋅(κ∇u) = 0 only for constant κ) -> uh = u at the nodes."""
    announce("3d 9x7x8 linear u dirichlet lifting")
    u = lambda x: 1 + 2 * x[:, 0] - x[:, 1] + 0.5 * x[:, 2]
    part = (9, 7, 8)
    kappa = np.full(int(np.prod(part)), 2.5)
    for path in ("auto", "general"):
        op, V, U, model = build((0, 1, 0, 2, 0, 1), part, kappa=kappa,
                                fsrc=np.zeros(int(np.prod(part))), g=u, path=path)
        ref = oracle_problem(oracle, (0, 1, 0, 2, 0, 1), part, kappa=kappa,
                             fsrc=np.zeros(int(np.prod(part))),
                             g=lambda x: 1 + 2 * x[0] - x[1] + 0.5 * x[2])
        check_against(ref, op, "fp64")
        uh = G.solve(op, G.PCGSolver(rtol=1e-13))
        xyz = model.node_coordinates()
        exact = u(xyz[V.free_dof_to_node - 1])
        assert np.max(np.abs(uh.free_values - exact)) <= 1e-9


def test_partial_dirichlet_tags(oracle):
    announce("3d 10x9x8 dirichlet tag_25 (x=0 face) + tag_23")
    rng = np.random.default_rng(14)
    part = (10, 9, 8)
    kappa = np.exp(rng.normal(size=int(np.prod(part))))
    tags = ("tag_25", "tag_23")
    op, *_ = build((0, 1) * 3, part, kappa=kappa, tags=tags)
    assert op.assem.ctx is not None
    ref = oracle_problem(oracle, (0, 1) * 3, part, kappa=kappa, tags=tags)
    check_against(ref, op, "fp64")


@pytest.mark.parametrize("path", ["auto", "general"])
def test_mapped_grid_general_path(oracle, path):
    """Non-affine geometry (jittered interior nodes): isoparametric Jacobian per
    quadrature point.  auto: the mesh keeps the lattice topology, so the element
    matrices are stored per cell and the owner-computes lattice kernel sums them;
    general: the per-cell kernel with the coloured scatter."""
    announce("3d 8x7x6 jittered mapped grid path=%s" % path)
    part = (8, 7, 6)
    rng = np.random.default_rng(15)
    h = np.array([1 / 8, 1 / 7, 1 / 6])

    def jitter(xyz):
        out = xyz.copy()
        interior = np.all((xyz > 1e-12) & (xyz < 1 - 1e-12), axis=1)
        out[interior] += rng.uniform(-0.2, 0.2, size=(interior.sum(), 3)) * h
        return out

    kappa = np.exp(rng.normal(size=int(np.prod(part))))
    op, V, U, model = build((0, 1) * 3, part, kappa=kappa, map=jitter, path=path)
    kern = _lib.GH_KERNEL_Q1_LATTICE if path == "auto" else _lib.GH_KERNEL_CELL
    assert op.assem.ctx.pattern_info()["last_kernel"] == kern
    ref = oracle_problem(oracle, (0, 1) * 3, part, kappa=kappa, xyz=model.node_coordinates())
    check_against(ref, op, "fp64")
    op32, *_ = build((0, 1) * 3, part, kappa=kappa, map=lambda x: model.node_coordinates(),
                     precision="fp32", path=path)
    check_against(ref, op32, "fp32")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mapped_lattice_matches_cell_kernel(oracle, precision):
    """A smoothly mapped 2-D and 3-D lattice with κ lognormal, Dirichlet data g on part
    of the boundary (lifting): the mapped lattice kernel against the oracle and against
    the per-cell kernel (same element arithmetic, another summation order)."""
    for dom, part in (((0, 1, 0, 1), (17, 13)), ((0, 1) * 3, (11, 9, 10))):
        announce("mapped lattice %s %s" % (part, precision))
        D = len(part)
        warp = lambda x: x + 0.04 * np.prod(np.sin(np.pi * x), axis=-1)[..., None]
        kappa = np.exp(np.random.default_rng(D).normal(size=int(np.prod(part))))
        g = lambda x: 1.0 + x[..., 0] - 0.5 * x[..., 1]
        tags = ("tag_5", "tag_7") if D == 2 else ("tag_21", "tag_23")
        op, V, U, model = build(dom, part, kappa=kappa, map=warp, g=g, tags=tags,
                                precision=precision)
        assert op.assem.ctx.pattern_info()["last_kernel"] == _lib.GH_KERNEL_Q1_LATTICE
        ref = oracle_problem(oracle, dom, part, kappa=kappa, xyz=model.node_coordinates(),
                             tags=tags, g=lambda x: 1.0 + x[0] - 0.5 * x[1])
        check_against(ref, op, precision)
        op2, *_ = build(dom, part, kappa=kappa, map=warp, g=g, tags=tags, precision=precision,
                        path="general")
        A1, A2 = op.get_matrix(), op2.get_matrix()
        tol = 1e-13 if precision == "fp64" else 1e-6
        assert relmax(A1.nzval, A2.nzval) <= tol
        assert relmax(op.get_vector(), op2.get_vector()) <= tol


@pytest.mark.parametrize("path", ["auto", "general"])
def test_repeat_is_bit_identical(path):
    announce("determinism 24^3 path=%s" % path)
    rng = np.random.default_rng(16)
    kappa = np.exp(rng.normal(size=24 ** 3))
    op, V, U, model = build((0, 1) * 3, (24, 24, 24), kappa=kappa, path=path,
                            precision="fp32")
    A0, b0 = op.get_matrix().nzval.copy(), op.get_vector().copy()
    for _ in range(3):
        A1, b1 = op.assem.assemble(*_forms(op, V, U, kappa), lift=True)
        assert np.array_equal(A1.nzval, A0) and np.array_equal(b1, b0)


def _forms(op, V, U, kappa):
    Ω = G.Triangulation(V.model)
    dΩ = G.Measure(Ω, 2)
    u, v = G.get_trial_fe_basis(U), G.get_fe_basis(V)
    return (G.integrate(G.CellField(kappa, Ω) * G.inner(G.gradient(v), G.gradient(u)), dΩ),
            G.integrate(1.0 * v, dΩ))


def test_edge_cases(oracle):
    announce("edge: 1x1 all-Dirichlet, 1x1 Neumann, 1x7x1 thin, 2x1x1")
    op, V, *_ = build((0, 1, 0, 1), (1, 1))
    assert V.num_free_dofs() == 0 and op.get_matrix().nnz() == 0
    assert op.get_vector().size == 0
    op, *_ = build((0, 1, 0, 1), (1, 1), tags=())
    ref = oracle_problem(oracle, (0, 1, 0, 1), (1, 1), tags=())
    check_against(ref, op, "fp64")
    for part, tags in [((1, 7, 1), ("boundary",)), ((1, 7, 1), ()), ((2, 1, 1), ("tag_25",))]:
        op, *_ = build((0, 1) * 3, part, tags=tags)
        ref = oracle_problem(oracle, (0, 1) * 3, part, tags=tags)
        check_against(ref, op, "fp64")


def test_error_paths():
    announce("error paths through the C ABI")
    ctx = _lib.Context(0)
    with pytest.raises(G.GridapHipError) as e:
        ctx.symbolic()
    assert e.value.code == _lib.GH_ERR_STATE
    ctx.mesh_cartesian([0, 0], [0.5, 0.5], [2, 2])
    ptrs = np.arange(5, dtype=np.int32) * 4 + 1
    bad = np.array([1, 2, 3, 99] * 4, dtype=np.int32)
    with pytest.raises(G.GridapHipError) as e:
        ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, bad, ptrs, 9, 0)
    assert e.value.code == _lib.GH_ERR_RANGE
    with pytest.raises(G.GridapHipError) as e:
        ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, bad[:12], ptrs[:4], 9, 0)
    assert e.value.code == _lib.GH_ERR_ARG
    with pytest.raises(G.GridapHipError) as e:
        ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, np.array([1, 2, 0, 4] * 4, np.int32), ptrs, 9, 0)
    assert e.value.code == _lib.GH_ERR_RANGE
    # a permuted (non-CLagrangian) numbering is valid but not lattice-eligible
    cd = np.array([[1, 2, 4, 5], [2, 3, 5, 6], [4, 5, 7, 8], [5, 6, 8, 9]], np.int32)
    perm = np.array([9, 8, 7, 6, 5, 4, 3, 2, 1], np.int32)
    ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, perm[cd - 1].ravel(), ptrs, 9, 0)
    nnz = ctx.symbolic()
    assert nnz == 49  # 3x3 nodes, all free: 4 corners x 4 + 4 edges x 6 + centre 9
    q = G.Quadrature(2, 2)
    shp, grad = G.ReferenceFE(G.lagrangian, G.Float64, 1).tabulate(2, q.coords)
    ctx.reffe_tabulation(q.weights, shp, grad)
    ctx.coef_upload()
    with pytest.raises(G.GridapHipError) as e:
        ctx.assemble(path=_lib.GH_PATH_STRUCTURED)
    assert e.value.code == _lib.GH_ERR_UNSUPPORTED
    ctx.assemble(path=_lib.GH_PATH_AUTO)  # falls to the general path
    with pytest.raises(G.GridapHipError) as e:
        ctx.pcg(b=np.ones(9), precond=_lib.GH_PRECOND_JACOBI, rtol=1e-12, maxit=0)
    assert e.value.code == _lib.GH_ERR_NOCONV
    ctx.close()


def test_spmv_and_pcg_match_scipy(oracle):
    announce("spmv + jacobi pcg 24x20x16 lognormal kappa")
    rng = np.random.default_rng(17)
    part = (24, 20, 16)
    kappa = np.exp(rng.normal(size=int(np.prod(part))))
    op, V, *_ = build((0, 1) * 3, part, kappa=kappa)
    A = op.get_matrix().to_scipy()
    x = rng.normal(size=V.num_free_dofs())
    y = op.assem.ctx.spmv(x)
    with pytest.raises(ValueError):
        op.assem.ctx.spmv(x[:-1])
    assert relmax(y, A @ x) <= 1e-13
    xs, st = op.assem.ctx.pcg(b=op.get_vector(), rtol=1e-10)
    assert st.converged == 1 and st.rel_residual <= 1e-10
    r = op.get_vector() - A @ xs
    assert np.linalg.norm(r) / np.linalg.norm(op.get_vector()) <= 1e-10
    xd = spla.spsolve(A.tocsc(), op.get_vector())
    assert relmax(xs, xd) <= 1e-7
    x2, st2 = op.assem.ctx.pcg(b=op.get_vector(), rtol=1e-10)
    assert np.array_equal(xs, x2) and st2.iterations == st.iterations


def test_64cube_against_oracle(oracle):
    announce("64^3 lognormal kappa fp32 + fp64 vs oracle")
    rng = np.random.default_rng(18)
    part = (64, 64, 64)
    kappa = np.exp(rng.normal(size=64 ** 3))
    ref = oracle_problem(oracle, (0, 1) * 3, part, kappa=kappa)
    for precision in ("fp64", "fp32"):
        op, *_ = build((0, 1) * 3, part, kappa=kappa, precision=precision)
        check_against(ref, op, precision)


def test_config2_128cube_properties():
    """3-D Q1 128^3, κ=exp(z), z~N(0,1): closed-form NNZ, symmetry, zero row sums
    of rows with a complete stencil; the lattice kernel (FP64 sums) against the
    per-cell general kernel in FP32 element math (relative Frobenius <= 1e-6) and in
    FP64 (1e-12), both at the full size."""
    announce("config2 128^3 properties")
    rng = np.random.default_rng(19)
    n = 128
    kappa = np.exp(rng.normal(size=n ** 3))
    op64, V, U, model = build((0, 1) * 3, (n, n, n), kappa=kappa, precision="fp64")
    assert op64.assem.ctx.pattern_info()["lattice"] == 1
    A = op64.get_matrix()
    assert A.nnz() == 54439939 == (3 * (n - 1) - 2) ** 3
    assert V.num_free_dofs() == 2048383
    ctx = op64.assem.ctx
    for prec, tol, err in (("fp32", FP32_FROB_TOL, relfrob), ("fp64", FP64_TOL, relmax)):
        announce("config2 128^3 general kernel %s vs lattice" % prec)
        ctx.assemble(_lib.GH_FORM_DIFFUSION, op64.assem._PREC[prec], _lib.GH_PATH_GENERAL)
        nz, bg = ctx.values_get()
        assert err(nz, A.nzval) <= tol, (prec, err(nz, A.nzval))
        assert err(bg, op64.get_vector()) <= tol
        if prec == "fp32":
            assert not np.array_equal(nz, A.nzval)  # the FP32 element math did run
    S = A.to_scipy()
    assert abs(S - S.T).max() <= 1e-15 * abs(S).max()
    # rows whose node is at least 2 nodes away from the boundary: zero row sum
    idx = np.indices((n + 1,) * 3).reshape(3, -1, order="F")
    far = np.all((idx >= 2) & (idx <= n - 2), axis=0)
    free_far = V.node_and_comp_to_dof[far, 0] - 1
    rs = S @ np.ones(S.shape[0])
    diag = S.diagonal()
    assert np.max(np.abs(rs[free_far]) / diag[free_far]) <= 1e-13


@pytest.mark.parametrize("part,mapped", [((13, 11), False), ((9, 8, 7), False),
                                         ((8, 7, 6), True)])
def test_source_at_quadrature_points(oracle, part, mapped):
    """l(v) = ∫ v f dΩ with f a function of x (PoissonTests.jl:101-108), f evaluated at
    the physical quadrature points: nzval and b against the oracle's
    cell_vectors_source_qp (FP64, 1e-12), Dirichlet data lifted."""
    D = len(part)
    announce("source f(x) at quadrature points %s mapped=%s" % (part, mapped))
    f = lambda x: np.sin(3 * x[..., 0]) + x[..., 1] * x[..., -1]
    g = lambda x: 1.0 + x[..., 0] - x[..., 1]
    mp = (lambda x: x + 0.03 * np.sin(4 * x[..., ::-1])) if mapped else None
    model = G.CartesianDiscreteModel((0, 1) * D, part, map=mp)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1),
                      dirichlet_tags="boundary")
    U = G.TrialFESpace(V, g)
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 2)
    kappa = np.exp(np.random.default_rng(D).normal(size=model.num_cells()))
    k = G.CellField(kappa, Ω)
    op = G.AffineFEOperator(lambda u, v: G.integrate(k * G.inner(G.gradient(v), G.gradient(u)), dΩ),
                            lambda v: G.integrate(v * f, dΩ), U, V)
    cn, xyz = model.cell_node_ids(), model.node_coordinates()
    qx, qw = oracle.tensor_quadrature(D, 2)
    shp, grad = oracle.qk_tabulate(D, 1, qx)
    xq = oracle.physical_quadrature_points(D, cn, xyz, shp)
    fq = f(xq)
    cs = None if mapped else model.sizes
    mats = (oracle.cell_matrices_diffusion(D, cn, xyz, qw, grad, kappa, cartesian_sizes=cs))
    vecs = oracle.cell_vectors_source_qp(D, cn, xyz, qw, shp, grad, fq, cartesian_sizes=cs)
    cd = V.cell_dof_ids
    dv = U.dirichlet_values()
    vecs = oracle.attach_dirichlet(cd, mats, dv, vecs)
    cp, rv, nz = oracle.assemble_csc(cd, mats, V.nfree, V.nfree)
    b = oracle.assemble_vector(cd, vecs, V.nfree)
    A = op.get_matrix()
    np.testing.assert_array_equal(A.rowval, rv)
    assert relmax(A.nzval, nz) <= FP64_TOL
    assert relmax(op.get_vector(), b) <= FP64_TOL
