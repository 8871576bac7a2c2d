"""Newton FEOperator on the device: residual and Jacobian of the p-Laplacian at an FE
function against the oracle, and Newton's method on PLaplacianTests.jl's driver
(needs an MI355X).

Bar: FP64 residual and Jacobian within 1e-12 relative (max-abs error / max-abs value)
of the oracle, pattern bit-exact; FP32 element math within 1e-6 relative Frobenius;
Newton iterates within 1e-10 of the oracle's Newton (direct solves) and the
manufactured solution u = x + y reproduced to 1e-9.
"""
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import gridap_hip as G

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300) if a.size else 0.0


def relfrob(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _operator(part, p, kappa=None, fsrc=None, g=None, map=None, tags=("boundary",),
              precision="fp64", path="auto"):
    D = len(part)
    model = G.CartesianDiscreteModel((0, 1) * D, part, map=map)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1),
                      dirichlet_tags=list(tags))
    U = G.TrialFESpace(V, g)
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 2)
    law = G.PLaplacian(p)
    k = 1.0 if kappa is None else G.CellField(kappa, Ω)
    f = 0.0 if fsrc is None else G.CellField(fsrc, Ω)
    res = lambda u, v: G.integrate(k * G.inner(G.gradient(v), law.flux(G.gradient(u))) -
                                   f * v, dΩ)
    jac = lambda u, du, v: G.integrate(
        k * G.inner(G.gradient(v), law.dflux(G.gradient(du), G.gradient(u))), dΩ)
    assem = G.SparseMatrixAssembler(U, V, precision=precision, path=path)
    return G.FEOperator(res, jac, U, V, assem), V, U, model


def _oracle_system(oracle, model, V, uh, p, kappa, fsrc):
    D = model.D
    cd = V.cell_dof_ids
    xyz = model.node_coordinates()
    cn = model.cell_node_ids()
    qx, qw = oracle.tensor_quadrature(D, 2)
    shp, grad = oracle.qk_tabulate(D, 1, qx)
    uc = oracle.cell_values(cd, uh.free_values, uh.dirichlet_values)
    cs = model.sizes if model.is_affine_cartesian() else None
    K, r = oracle.plaplacian_cells(D, cn, xyz, qw, shp, grad, uc, p, kappa, fsrc,
                                   cartesian_sizes=cs)
    return oracle.plaplacian_system(cd, V.nfree, K, r)


@pytest.mark.parametrize("part,mapped,p", [((9, 7), False, 3.0), ((9, 7), True, 4.0),
                                           ((7, 6, 5), False, 3.0), ((6, 5, 4), True, 3.5)])
def test_residual_and_jacobian_match_oracle(oracle, part, mapped, p):
    announce("p-Laplacian residual+Jacobian %s mapped=%s p=%g" % (part, mapped, p))
    D = len(part)
    rng = np.random.default_rng(len(part) * 10 + int(mapped))
    C = int(np.prod(part))
    kappa = np.exp(rng.normal(size=C))
    fsrc = rng.uniform(-1, 1, C)
    mp = (lambda x: x + 0.04 * np.sin(3 * x[..., ::-1])) if mapped else None
    g = lambda x: 1.0 + x[..., 0] - 0.5 * x[..., 1]
    op, V, U, model = _operator(part, p, kappa, fsrc, g=g, map=mp)
    uh = G.FEFunction(U, rng.normal(size=V.nfree))
    b, A = G.residual_and_jacobian(op, uh)
    cp, rv, nz, r = _oracle_system(oracle, model, V, uh, p, kappa, fsrc)
    np.testing.assert_array_equal(A.colptr, cp)
    np.testing.assert_array_equal(A.rowval, rv)
    assert relmax(A.nzval, nz) <= 1e-12
    assert relmax(b, r) <= 1e-12
    # in-place re-assembly at another uh, into the same A and b
    uh2 = G.FEFunction(U, rng.normal(size=V.nfree))
    b2, A2 = op.residual_and_jacobian_inplace(b, A, uh2)
    assert A2 is A and b2 is b
    cp, rv, nz, r = _oracle_system(oracle, model, V, uh2, p, kappa, fsrc)
    assert relmax(A.nzval, nz) <= 1e-12 and relmax(b, r) <= 1e-12
    # FP32 element math
    op32, *_ = _operator(part, p, kappa, fsrc, g=g, map=mp, precision="fp32")
    b32, A32 = G.residual_and_jacobian(op32, uh2)
    assert relfrob(A32.nzval, nz) <= 1e-6 and relfrob(b32, r) <= 1e-6


@pytest.mark.parametrize("part,mapped", [((9, 7), False), ((8, 6, 5), True)])
def test_lattice_route_matches_cell_kernel(part, mapped):
    """Lattice meshes take the owner-computes column route (element Jacobians and FP64
    residuals stored per cell, summed per column); the coloured per-cell scatter
    (path "general") computes the same system in another summation order."""
    announce("p-Laplacian lattice vs per-cell kernel %s mapped=%s" % (part, mapped))
    rng = np.random.default_rng(3 + len(part))
    C = int(np.prod(part))
    kappa = np.exp(rng.normal(size=C))
    fsrc = rng.uniform(-1, 1, C)
    mp = (lambda x: x + 0.04 * np.sin(3 * x[..., ::-1])) if mapped else None
    g = lambda x: x[..., 0] - x[..., 1]
    for precision, tol in (("fp64", 1e-13), ("fp32", 1e-6)):
        op, V, U, model = _operator(part, 3.0, kappa, fsrc, g=g, map=mp, precision=precision)
        opc, *_ = _operator(part, 3.0, kappa, fsrc, g=g, map=mp, precision=precision,
                            path="general")
        uh = G.FEFunction(U, rng.normal(size=V.nfree))
        b, A = G.residual_and_jacobian(op, uh)
        assert op.assem.ctx.pattern_info()["last_kernel"] == G._lib.GH_KERNEL_Q1_LATTICE
        bc, Ac = G.residual_and_jacobian(opc, uh)
        assert opc.assem.ctx.pattern_info()["last_kernel"] == G._lib.GH_KERNEL_CELL
        assert relmax(A.nzval, Ac.nzval) <= tol and relmax(b, bc) <= 1e-13
        # the Jacobian is symmetric bitwise on both routes
        for M in (A, Ac):
            S = M.to_scipy()
            assert (S - S.T).count_nonzero() == 0


def test_newton_plaplacian_driver():
    """PLaplacianTests.jl:8-47 with Q1 (u = x + y is in the space): p = 3, f = 0, random
    initial guess, NewtonRaphsonSolver with device PCG; the solution is the interpolant
    of u and the last steps converge quadratically."""
    announce("p-Laplacian Newton driver 4x4 and 24x20")
    g = lambda x: x[..., 0] + x[..., 1]
    for part in ((4, 4), (24, 20)):
        op, V, U, model = _operator(part, 3.0, g=g)
        uh0 = G.FEFunction(U, np.random.default_rng(0).uniform(0, 1, V.nfree))
        nls = G.NewtonRaphsonSolver(G.PCGSolver(rtol=1e-13), 1e-12, 20)
        uh, _ = G.solve_inplace(uh0, G.FESolver(nls), op)
        uI = g(model.node_coordinates()[V.free_dof_to_node - 1])
        assert np.max(np.abs(uh.free_values - uI)) <= 1e-9
        h = nls.history
        assert h[-1] < 1e-12 * h[0]
        # quadratic convergence above the rounding floor: r_{k+1}/r_0 <= C (r_k/r_0)^2
        tail = [x / h[0] for x in h if x > 1e-13 * h[0]]
        assert len(tail) >= 2 and tail[-1] <= 1e3 * tail[-2] ** 2, h


def test_newton_iterates_match_oracle(oracle):
    """3-D 12x10x8, κ lognormal, f per cell, p = 3: every Newton iterate of the device
    (residual/Jacobian assembled on the device, PCG to 1e-13) against the oracle's
    Newton with direct solves, from the same initial guess."""
    announce("p-Laplacian Newton iterates vs oracle 12x10x8")
    part = (12, 10, 8)
    rng = np.random.default_rng(7)
    C = int(np.prod(part))
    kappa = np.exp(0.5 * rng.normal(size=C))
    fsrc = rng.uniform(0.5, 1.5, C)
    g = lambda x: 0.2 * x[..., 0] - x[..., 2]
    op, V, U, model = _operator(part, 3.0, kappa, fsrc, g=g)
    x = rng.uniform(-0.2, 0.2, V.nfree)
    xo = x.copy()
    pcg = G.PCGSolver(rtol=1e-13)
    r0 = None
    for it in range(9):                # the oracle's Newton converges at iteration 8
        b, A = op.residual_and_jacobian(G.FEFunction(U, x))
        cp, rv, nz, r = _oracle_system(oracle, model, V, G.FEFunction(U, xo), 3.0, kappa, fsrc)
        r0 = np.abs(r).max() if r0 is None else r0
        # residuals against the initial residual's scale (late iterates are rounding noise)
        assert np.abs(b - r).max() <= 1e-10 * r0, it
        assert relmax(A.nzval, nz) <= 1e-10, it
        x = x + pcg.solve(op.assem, -b, A)
        xo = xo + spla.spsolve(oracle.to_scipy(cp, rv, nz, V.nfree).tocsc(), -r)
        assert relmax(x, xo) <= 1e-9, it
    assert np.abs(r).max() <= 1e-10 * r0   # converged


def _unstructured_q1(oracle, part, seed):
    """A Cartesian Q1 mesh with its nodes renumbered at random and moved (no lattice
    topology left): coordinates, 1-based cell nodes, CLagrangian cell dofs of the
    boundary-Dirichlet space in the new node order."""
    import conforming_oracle as CO
    D = len(part)
    m = oracle.cartesian_model((0.0, 1.0) * D, part)
    rng = np.random.default_rng(seed)
    nn = m["xyz"].shape[0]
    perm = rng.permutation(nn)                 # new id of old node k: perm[k]
    xyz = np.empty_like(m["xyz"])
    xyz[perm] = m["xyz"] + 0.05 * np.prod(np.sin(np.pi * m["xyz"]), axis=-1)[:, None]
    cn = (perm[m["cell_nodes"] - 1] + 1).astype(np.int32)
    lab = np.empty_like(m["vertex_labels"])
    lab[perm] = m["vertex_labels"]
    node_to_tag = oracle.face_tag_index(lab, [sorted(CO.boundary_entities(D))])
    n2d, nfree, ndiri = oracle.clagrangian(node_to_tag)
    cd = oracle.clagrangian_cell_dofs(cn, n2d)
    return xyz, cn, cd, nfree, ndiri


@pytest.mark.parametrize("part", [(9, 7), (6, 5, 4)])
def test_unstructured_mesh_element_lists_match_oracle(oracle, part):
    """A Q1 mesh without the lattice topology (renumbered, moved nodes), through the C
    ABI: the residual and Jacobian go to the element-list sums (no coloured scatter),
    within 1e-12 of the oracle (FP64) and 1e-6 (FP32 element math); the coloured
    scatter (path "general") agrees."""
    from gridap_hip import _lib
    D = len(part)
    announce("p-Laplacian element lists, unstructured %s" % (part,))
    xyz, cn, cd, nfree, ndiri = _unstructured_q1(oracle, part, 11)
    C = cn.shape[0]
    rng = np.random.default_rng(D + 20)
    kappa, fsrc = np.exp(rng.normal(size=C)), rng.uniform(-1, 1, C)
    u, dv = rng.normal(size=nfree), rng.normal(size=ndiri)
    p = 3.0
    qx, qw = oracle.tensor_quadrature(D, 2)
    shp, grad = oracle.qk_tabulate(D, 1, qx)
    ptrs = (np.arange(C + 1) * (1 << D) + 1).astype(np.int32)
    ctx = _lib.Context(0)
    ctx.mesh_upload(xyz, cn.ravel(), ptrs)
    ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, np.ascontiguousarray(cd, np.int32).ravel(), ptrs,
                     nfree, ndiri)
    ctx.reffe_tabulation(qw, shp, grad.reshape(len(qw), -1))
    ctx.symbolic()
    ctx.coef_upload(kappa, fsrc, None)
    ctx.fefunction_upload(u, dv)
    ctx.plaplacian_params(p)
    uc = oracle.cell_values(cd, u, dv)
    K, r = oracle.plaplacian_cells(D, cn, xyz, qw, shp, grad, uc, p, kappa, fsrc)
    cp, rv, nzr, br = oracle.plaplacian_system(cd, nfree, K, r)
    colptr, rowval = ctx.pattern_get()
    np.testing.assert_array_equal(colptr, cp)
    np.testing.assert_array_equal(rowval, rv)
    out = {}
    for prec, path in ((_lib.GH_PREC_FP64, _lib.GH_PATH_AUTO), (_lib.GH_PREC_FP32_ELEM, _lib.GH_PATH_AUTO),
                       (_lib.GH_PREC_FP64, _lib.GH_PATH_GENERAL)):
        ctx.assemble(_lib.GH_FORM_PLAPLACIAN, prec, path)
        out[(prec, path)] = ctx.values_get() + (ctx.pattern_info()["last_kernel"],)
    ctx.close()
    nz, b, kern = out[(_lib.GH_PREC_FP64, _lib.GH_PATH_AUTO)]
    assert kern in (_lib.GH_KERNEL_COLUMN_SUM, _lib.GH_KERNEL_LIST_SUM)
    assert relmax(nz, nzr) <= 1e-12 and relmax(b, br) <= 1e-12
    nz32, b32, _ = out[(_lib.GH_PREC_FP32_ELEM, _lib.GH_PATH_AUTO)]
    assert relfrob(nz32, nzr) <= 1e-6 and relmax(b32, br) <= 1e-12   # residual in FP64
    nzc, bc, kc = out[(_lib.GH_PREC_FP64, _lib.GH_PATH_GENERAL)]
    assert kc == _lib.GH_KERNEL_CELL
    assert relmax(nzc, nz) <= 1e-13 and relmax(bc, b) <= 1e-13
