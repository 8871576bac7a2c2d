"""Multi-field assembly on the device against the oracle (needs an MI355X).

Mirrors test/MultiFieldTests/BlockSparseMatrixAssemblersTests.jl: two and three copies
of a Q1 space on a 5x5 QUAD model with Dirichlet data sol(x) = x1 + x2 on the boundary,
biform2/liform2 and biform3/liform3 (:101-104), assembled with the consecutive style
(one matrix) and with the block styles BlockMultiFieldStyle() and
BlockMultiFieldStyle(2, (1, n-1)) for trial and test: every block layout equals the
monolithic matrix (:48-75), and the monolithic one equals the oracle (pattern bit-exact,
values and b within 1e-12).  The mass term ∫ σ u v alone is checked on a 3-D mesh.
"""
import sys

import numpy as np
import pytest

import gridap_hip as G
from gridap_hip import _lib
import multifield_oracle as MF

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300) if a.size else 0.0


sol = lambda x: x[..., 0] + x[..., 1]


def biform2(u, v, dΩ):
    (u1, u2), (v1, v2) = u, v
    return G.integrate(G.inner(G.gradient(u1), G.gradient(v1)) + G.inner(u2, v2) -
                       G.inner(u1, v2), dΩ)


def liform2(v, dΩ):
    v1, v2 = v
    return G.integrate(v1 * 1.0 - v2 * 1.0, dΩ)


def biform3(u, v, dΩ):
    (u1, u2, u3), (v1, v2, v3) = u, v
    return G.integrate(G.inner(G.gradient(u1), G.gradient(v1)) + G.inner(u2, v2) -
                       G.inner(u1, v2) - G.inner(u3, v2) - G.inner(u2, v3), dΩ)


def liform3(v, dΩ):
    v1, v2, v3 = v
    return G.integrate(v1 * 1.0 - v2 * 1.0 + v3 * 2.0, dΩ)


# the touched pairs (test field, trial field) of each form and their mass/stiffness
PAIRS2 = {(0, 0): ("stiff", 1.0), (1, 1): ("mass", 1.0), (1, 0): ("mass", -1.0)}
VECS2 = {0: 1.0, 1: -1.0}
PAIRS3 = {(0, 0): ("stiff", 1.0), (1, 1): ("mass", 1.0), (1, 0): ("mass", -1.0),
          (1, 2): ("mass", -1.0), (2, 1): ("mass", -1.0)}
VECS3 = {0: 1.0, 1: -1.0, 2: 2.0}


def oracle_system(n, pairs, vecs, lift=True):
    f = MF.q1_field((0, 1, 0, 1), (5, 5), g=lambda x: x[0] + x[1])
    cd = f["cell_dofs"]
    mats = {k: s * f[kind] for k, (kind, s) in pairs.items()}
    cv = {i: s * f["src"] for i, s in vecs.items()}
    return f, MF.assemble([cd] * n, [f["nfree"]] * n, [f["ndiri"]] * n, mats, cv,
                          dvals=[f["dvals"]] * n if lift else None)


def spaces(n, style=None):
    model = G.CartesianDiscreteModel((0, 1, 0, 1), (5, 5))
    reffe = G.ReferenceFE(G.lagrangian, G.Float64, 1)
    V = G.TestFESpace(model, reffe, dirichlet_tags="boundary")
    U = G.TrialFESpace(V, sol)
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 2)
    return V, U, dΩ


def slice_csc(A, rows, cols):
    """Rows and columns (0-based index arrays, each a union of field ranges) of a 1-based
    SparseMatrixCSC, keeping stored entries (explicit zeros included)."""
    rmap = -np.ones(A.m, dtype=np.int64)
    rmap[rows] = np.arange(len(rows))
    cp, rv, nz = [1], [], []
    for c in cols:
        lo, hi = A.colptr[c] - 1, A.colptr[c + 1] - 1
        r = rmap[A.rowval[lo:hi] - 1]
        keep = r >= 0
        rv.append(r[keep] + 1)
        nz.append(A.nzval[lo:hi][keep])
        cp.append(cp[-1] + int(keep.sum()))
    return (np.array(cp), np.concatenate(rv) if rv else np.zeros(0, np.int64),
            np.concatenate(nz) if nz else np.zeros(0))


@pytest.mark.parametrize("n", [2, 3])
def test_multifield_consecutive_matches_oracle(oracle, n):
    announce("multi-field consecutive, %d fields, 5x5 QUAD" % n)
    V, U, dΩ = spaces(n)
    Y = G.MultiFieldFESpace([V] * n)
    X = G.MultiFieldFESpace([U] * n)
    a = (lambda u, v: biform2(u, v, dΩ)) if n == 2 else (lambda u, v: biform3(u, v, dΩ))
    l = (lambda v: liform2(v, dΩ)) if n == 2 else (lambda v: liform3(v, dΩ))
    pairs, vecs = (PAIRS2, VECS2) if n == 2 else (PAIRS3, VECS3)
    _, (cp, rv, nz, b) = oracle_system(n, pairs, vecs)
    assem = G.SparseMatrixAssembler(X, Y)
    A1 = G.assemble_matrix(a, assem, X, Y)
    np.testing.assert_array_equal(A1.colptr, cp)
    np.testing.assert_array_equal(A1.rowval, rv)
    assert relmax(A1.nzval, nz) <= 1e-12
    A2, b2 = G.assemble_matrix_and_vector(a, l, X, Y, assem)
    np.testing.assert_array_equal(A2.rowval, rv)
    assert relmax(A2.nzval, nz) <= 1e-12
    assert relmax(b2, b) <= 1e-12
    # the vector alone: sources only, no lifting
    _, (_, _, _, bv) = oracle_system(n, pairs, vecs, lift=False)
    b1 = G.assemble_vector(l, assem, Y)
    assert relmax(b1, bv) <= 1e-12
    op = G.AffineFEOperator(a, l, X, Y)
    assert relmax(op.get_matrix().nzval, nz) <= 1e-12 and relmax(op.get_vector(), b) <= 1e-12


@pytest.mark.parametrize("n", [2, 3])
@pytest.mark.parametrize("trial_style", ["block", "block2"])
@pytest.mark.parametrize("test_style", ["block", "block2"])
def test_multifield_blocks_equal_monolithic(n, trial_style, test_style):
    announce("multi-field blocks %s x %s, %d fields" % (test_style, trial_style, n))
    V, U, dΩ = spaces(n)
    style = {"block": G.BlockMultiFieldStyle(), "block2": G.BlockMultiFieldStyle(2, (1, n - 1))}
    a = (lambda u, v: biform2(u, v, dΩ)) if n == 2 else (lambda u, v: biform3(u, v, dΩ))
    l = (lambda v: liform2(v, dΩ)) if n == 2 else (lambda v: liform3(v, dΩ))
    X, Y = G.MultiFieldFESpace([U] * n), G.MultiFieldFESpace([V] * n)
    A1, b1 = G.assemble_matrix_and_vector(a, l, X, Y)
    Xb = G.MultiFieldFESpace([U] * n, style=style[trial_style])
    Yb = G.MultiFieldFESpace([V] * n, style=style[test_style])
    Ab, bb = G.assemble_matrix_and_vector(a, l, Xb, Yb)
    assert isinstance(Ab, G.BlockSparseMatrix) and isinstance(bb, G.BlockVector)
    nf = V.nfree
    rgroups, cgroups = Yb.groups(), Xb.groups()
    assert len(Ab.blocks) == len(rgroups) and len(Ab.blocks[0]) == len(cgroups)
    # each block: the monolithic matrix on its fields, stored entries included
    for r, rg in enumerate(rgroups):
        rows = np.concatenate([np.arange(i * nf, (i + 1) * nf) for i in rg])
        for c, cg in enumerate(cgroups):
            cols = np.concatenate([np.arange(j * nf, (j + 1) * nf) for j in cg])
            cp, rv, nz = slice_csc(A1, rows, cols)
            B = Ab.blocks[r][c]
            assert B.shape == (len(rows), len(cols))
            np.testing.assert_array_equal(B.colptr, cp)
            np.testing.assert_array_equal(B.rowval, rv)
            np.testing.assert_array_equal(B.nzval, nz)
        np.testing.assert_array_equal(bb.blocks[r], b1[rows])
    # mul!: the block product equals the monolithic one when both layouts agree
    if rgroups == cgroups:
        x = np.random.default_rng(n).normal(size=A1.n)
        perm = np.concatenate([np.arange(j * nf, (j + 1) * nf) for g in cgroups for j in g])
        y1 = A1.to_scipy() @ x
        yb = Ab.to_scipy() @ x[perm]
        np.testing.assert_allclose(yb, y1[perm], rtol=1e-14, atol=1e-14)


def test_mass_term_3d_matches_oracle(oracle):
    """∫ κ ∇v⋅∇u + σ u v with per-cell κ and σ on a 3-D Cartesian mesh, on the lattice
    route (element matrices per cell, owner-computes columns) and on the per-cell
    kernel: FP64 against the oracle; FP32 element math within 1e-6."""
    part = (7, 6, 5)
    announce("mass + stiffness %s" % (part,))
    C = int(np.prod(part))
    rng = np.random.default_rng(5)
    kappa = np.exp(rng.normal(size=C))
    sigma = np.exp(rng.normal(size=C))
    f = MF.q1_field((0, 1, 0, 1, 0, 1), part, g=lambda x: 1.0 + x[0] * x[2])
    mats = kappa[:, None, None] * f["stiff"] + sigma[:, None, None] * f["mass"]
    cp, rv, nz, b = MF.assemble([f["cell_dofs"]], [f["nfree"]], [f["ndiri"]], {(0, 0): mats},
                                {0: f["src"]}, dvals=[f["dvals"]])
    model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), part)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1),
                      dirichlet_tags="boundary")
    U = G.TrialFESpace(V, lambda x: 1.0 + x[..., 0] * x[..., 2])
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 2)
    k, s = G.CellField(kappa, Ω), G.CellField(sigma, Ω)
    a = lambda u, v: G.integrate(k * G.inner(G.gradient(v), G.gradient(u)) + s * (u * v), dΩ)
    l = lambda v: G.integrate(1.0 * v, dΩ)
    for prec, tol, path, kern in (("fp64", 1e-12, "auto", _lib.GH_KERNEL_Q1_LATTICE),
                                  ("fp32", 1e-6, "auto", _lib.GH_KERNEL_Q1_LATTICE),
                                  ("fp64", 1e-12, "general", _lib.GH_KERNEL_CELL)):
        # auto: element matrices per cell, then the owner-computes lattice kernel;
        # general: the per-cell kernel with coloured scatter
        op = G.AffineFEOperator(a, l, U, V, G.SparseMatrixAssembler(U, V, precision=prec,
                                                                    path=path))
        A = op.get_matrix()
        assert op.assem.ctx.pattern_info()["last_kernel"] == kern
        np.testing.assert_array_equal(A.colptr, cp)
        np.testing.assert_array_equal(A.rowval, rv)
        assert relmax(A.nzval, nz) <= tol
        assert relmax(op.get_vector(), b) <= tol
    # the mass matrix alone is SPD: the device PCG solves it
    m = lambda u, v: G.integrate(s * (v * u), dΩ)
    opm = G.AffineFEOperator(m, l, U, V)
    uh = G.solve(opm)
    r = opm.get_vector() - opm.get_matrix().to_scipy() @ uh.free_values
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(opm.get_vector())


def _mixed_system(part, mapped, precision="fp64", style=None):
    """A Q2 field and a Q1 field on one mesh (Dirichlet data on both) coupled both ways:
    a = ∫ ∇u1⋅∇v1 + u2 v2 - u1 v2 + 0.5 ∇u2⋅∇v1, l = ∫ v1 - v2.  The blocks (2, 1) and
    (1, 2) pair two different spaces (rectangular)."""
    D = len(part)
    warp = lambda x: x + 0.04 * np.prod(np.sin(np.pi * x), axis=-1)[..., None]
    model = G.CartesianDiscreteModel((0, 1) * D, part, map=warp if mapped else None)
    V1 = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 2), dirichlet_tags="boundary")
    V2 = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1), dirichlet_tags="boundary")
    U1 = G.TrialFESpace(V1, lambda x: x[..., 0] + x[..., 1] ** 2)
    U2 = G.TrialFESpace(V2, lambda x: 1.0 - x[..., 0])
    dΩ = G.Measure(G.Triangulation(model), 4)
    kw = {} if style is None else {"style": style}
    X, Y = G.MultiFieldFESpace([U1, U2], **kw), G.MultiFieldFESpace([V1, V2], **kw)
    a = lambda u, v: G.integrate(G.inner(G.gradient(u[0]), G.gradient(v[0])) +
                                 G.inner(u[1], v[1]) - G.inner(u[0], v[1]) +
                                 0.5 * G.inner(G.gradient(u[1]), G.gradient(v[0])), dΩ)
    l = lambda v: G.integrate(v[0] * 1.0 - v[1] * 1.0, dΩ)
    return G.assemble_matrix_and_vector(a, l, X, Y, G.SparseMatrixAssembler(X, Y, precision=precision)), \
        (model, V1, V2, U1, U2)


def _mixed_oracle(oracle, model, V1, V2, U1, U2):
    import conforming_oracle as CO
    D = model.D
    X = model.node_coordinates()[model.cell_node_ids() - 1]
    qx, qw = oracle.tensor_quadrature(D, 4)
    s1, g1 = oracle.qk_tabulate(D, 1, qx)
    s2, g2 = oracle.qk_tabulate(D, 2, qx)
    wdet, gp1 = CO.physical_gradients(X, qw, g1, g1)
    _, gp2 = CO.physical_gradients(X, qw, g1, g2)
    C = X.shape[0]
    one = np.ones(C)
    mats = {(0, 0): CO.mixed_cells(wdet, gp2, gp2, s2, s2, kappa=one),
            (1, 1): CO.mixed_cells(wdet, gp1, gp1, s1, s1, sigma=one),
            (1, 0): -CO.mixed_cells(wdet, gp1, gp2, s1, s2, sigma=one),
            (0, 1): 0.5 * CO.mixed_cells(wdet, gp2, gp1, s2, s1, kappa=one)}
    vecs = {0: CO.ncube_loads(wdet, s2, one[:, None]), 1: -CO.ncube_loads(wdet, s1, one[:, None])}
    cds = [np.asarray(V1.cell_dof_ids), np.asarray(V2.cell_dof_ids)]
    return MF.assemble(cds, [V1.nfree, V2.nfree], [V1.ndirichlet, V2.ndirichlet], mats, vecs,
                       dvals=[U1.dirichlet_values(), U2.dirichlet_values()])


@pytest.mark.parametrize("part,mapped", [((5, 4), False), ((4, 3, 3), True)])
def test_mixed_order_fields_match_oracle(oracle, part, mapped):
    """Rectangular blocks between a Q2 and a Q1 field (consecutive layout): pattern
    bit-exact, values and b within 1e-12 of the oracle (FP64), 1e-6 (FP32 element
    math); the block layout equals the consecutive matrix entry for entry."""
    announce("mixed-order fields (Q2, Q1) %s mapped=%s" % (part, mapped))
    (A, b), spaces = _mixed_system(part, mapped)
    cp, rv, nz, br = _mixed_oracle(oracle, *spaces)
    np.testing.assert_array_equal(A.colptr, cp)
    np.testing.assert_array_equal(A.rowval, rv)
    assert relmax(A.nzval, nz) <= 1e-12 and relmax(b, br) <= 1e-12
    (A32, b32), _ = _mixed_system(part, mapped, precision="fp32")
    assert np.linalg.norm(A32.nzval - nz) <= 1e-6 * np.linalg.norm(nz)
    assert relmax(b32, br) <= 1e-6
    (Ab, bb), _ = _mixed_system(part, mapped, style=G.BlockMultiFieldStyle())
    n1 = spaces[1].nfree
    S = A.to_scipy().tocsc()
    for (i, j), (rs, cs) in {(0, 0): (slice(0, n1), slice(0, n1)),
                             (0, 1): (slice(0, n1), slice(n1, None)),
                             (1, 0): (slice(n1, None), slice(0, n1)),
                             (1, 1): (slice(n1, None), slice(n1, None))}.items():
        blk = Ab.blocks[i][j].to_scipy().tocsc()
        ref = S[rs, cs].tocsc()
        assert blk.shape == ref.shape and abs(blk - ref).max() == 0, (i, j)
    np.testing.assert_array_equal(np.concatenate(bb.blocks), b)


def _stokes(part, precision="fp64", calls=1):
    """test/GridapTests/StokesTaylorHoodTests.jl's Taylor–Hood form on (0,2)² with
    u = (x1² + 2 x2², -x1²) and p = x1 + 3 x2 (f = -Δu + ∇p = (-5, 5), g = ∇⋅u = 2 x1),
    velocity Dirichlet data on the whole boundary and the pressure fixed at the vertex
    (0, 0) (the driver's Neumann part fixes the pressure level instead)."""
    model = G.CartesianDiscreteModel((0, 2, 0, 2), part)
    ue = lambda x: np.stack([x[..., 0] ** 2 + 2 * x[..., 1] ** 2, -x[..., 0] ** 2], axis=-1)
    pe = lambda x: x[..., 0] + 3 * x[..., 1]
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.VectorValue[2, G.Float64], 2),
                      dirichlet_tags="boundary")
    Q = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1), dirichlet_tags="tag_1")
    U, P = G.TrialFESpace(V, ue), G.TrialFESpace(Q, pe)
    Y, X = G.MultiFieldFESpace([V, Q]), G.MultiFieldFESpace([U, P])
    dΩ = G.Measure(G.Triangulation(model), 2)
    g = lambda x: 2.0 * x[..., 0]
    a = lambda u, v: G.integrate(G.inner(G.gradient(v[0]), G.gradient(u[0])) -
                                 G.divergence(v[0]) * u[1] + v[1] * G.divergence(u[0]), dΩ)
    l = lambda v: G.integrate(G.dot(v[0], G.VectorValue(-5.0, 5.0)) + v[1] * g, dΩ)
    assem = G.SparseMatrixAssembler(X, Y, precision=precision)
    runs = [G.assemble_matrix_and_vector(a, l, X, Y, assem) for _ in range(calls)]
    A, b = runs[0]
    return A, b, (model, V, Q, U, P, ue, pe) + ((runs, assem) if calls > 1 else ())


def _stokes_oracle(oracle, model, V, Q, U, P):
    import conforming_oracle as CO
    X = model.node_coordinates()[model.cell_node_ids() - 1]
    qx, qw = oracle.tensor_quadrature(2, 2)
    s1, g1 = oracle.qk_tabulate(2, 1, qx)
    s2, g2 = oracle.qk_tabulate(2, 2, qx)
    wdet, gp2 = CO.physical_gradients(X, qw, g1, g2)
    C = X.shape[0]
    mats = {(0, 0): CO.vector_laplacian_cells(wdet, gp2),
            (1, 0): CO.divergence_cells(wdet, s1, gp2, scalar_test=True),
            (0, 1): -CO.divergence_cells(wdet, s1, gp2, scalar_test=False)}
    xq = CO.q1_map_points(X, s1)                                      # [C, nq, 2]
    vecs = {0: CO.ncube_loads(wdet, s2, np.tile([-5.0, 5.0], (C, 1))),
            1: np.einsum("cq,cq,qa->ca", wdet, 2.0 * xq[..., 0], s1)}
    cds = [np.asarray(V.cell_dof_ids), np.asarray(Q.cell_dof_ids)]
    return MF.assemble(cds, [V.nfree, Q.nfree], [V.ndirichlet, Q.ndirichlet], mats, vecs,
                       dvals=[U.dirichlet_values(), P.dirichlet_values()])


@pytest.mark.parametrize("part", [(3, 3), (8, 6)])
def test_stokes_taylor_hood_matches_oracle_and_solution(oracle, part):
    """Taylor–Hood Stokes (vector Q2 velocity, Q1 pressure): the vector Laplacian block,
    the q (∇⋅u) and -(∇⋅v) p rectangular blocks and both loads on the device, composed
    in the consecutive layout — pattern bit-exact and values within 1e-12 of the oracle
    (FP64; 1e-6 FP32 element math); the direct solve reproduces u and p (they lie in the
    spaces) to 1e-9, the driver's own bar (StokesTaylorHoodTests.jl:78-82)."""
    import scipy.sparse.linalg as spla
    announce("Stokes Taylor-Hood %s" % (part,))
    A, b, (model, V, Q, U, P, ue, pe) = _stokes(part)
    cp, rv, nz, br = _stokes_oracle(oracle, model, V, Q, U, P)
    np.testing.assert_array_equal(A.colptr, cp)
    np.testing.assert_array_equal(A.rowval, rv)
    assert relmax(A.nzval, nz) <= 1e-12 and relmax(b, br) <= 1e-12
    x = spla.spsolve(A.to_scipy().tocsc(), b)
    uI = G.interpolate(ue, U).free_values
    pI = G.interpolate(pe, P).free_values
    assert np.max(np.abs(x[:V.nfree] - uI)) <= 1e-9 * np.max(np.abs(uI))
    assert np.max(np.abs(x[V.nfree:] - pI)) <= 1e-9 * np.max(np.abs(pI))
    A32, b32, _ = _stokes(part, precision="fp32")
    assert np.linalg.norm(A32.nzval - nz) <= 1e-6 * np.linalg.norm(nz)
    assert relmax(b32, br) <= 1e-6


def test_stokes_taylor_hood_driver_as_written(oracle):
    """test/GridapTests/StokesTaylorHoodTests.jl as written: (0,2)², 3×3 cells,
    "dirichlet" = entities [1,2,5] for the velocity, "neumann" = [6,7,8] carrying
    ∫ v⋅(n⋅∇u) - (n⋅v) p dΓ, no pressure constraint.  The composite system matches the
    oracle (pattern bit-exact, 1e-12) and its solve reproduces u and p to 1e-9."""
    import scipy.sparse.linalg as spla
    announce("Stokes Taylor-Hood driver with Neumann terms")
    model = G.CartesianDiscreteModel((0, 2, 0, 2), (3, 3))
    labels = model.get_face_labeling()
    labels.add_tag_from_tags("dirichlet", [1, 2, 5])
    labels.add_tag_from_tags("neumann", [6, 7, 8])
    ue = lambda x: np.stack([x[..., 0] ** 2 + 2 * x[..., 1] ** 2, -x[..., 0] ** 2], axis=-1)
    pe = lambda x: x[..., 0] + 3 * x[..., 1]
    # ∇u[i, j] = ∂_i u_j
    gu = lambda x: np.stack([np.stack([2 * x[..., 0], -2 * x[..., 0]], -1),
                             np.stack([4 * x[..., 1], 0 * x[..., 1]], -1)], -2)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.VectorValue[2, G.Float64], 2),
                      labels=labels, dirichlet_tags="dirichlet")
    Q = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1), labels=labels)
    U, P = G.TrialFESpace(V, ue), G.TrialFESpace(Q)
    Y, X = G.MultiFieldFESpace([V, Q]), G.MultiFieldFESpace([U, P])
    Ω = G.Triangulation(model)
    Γ = G.BoundaryTriangulation(model, labels, tags="neumann")
    nΓ = G.get_normal_vector(Γ)
    dΩ, dΓ = G.Measure(Ω, 2), G.Measure(Γ, 2)
    g = lambda x: 2.0 * x[..., 0]
    a = lambda u, v: G.integrate(G.inner(G.gradient(v[0]), G.gradient(u[0])) -
                                 G.divergence(v[0]) * u[1] + v[1] * G.divergence(u[0]), dΩ)
    l = lambda v: (G.integrate(G.dot(v[0], G.VectorValue(-5.0, 5.0)) + v[1] * g, dΩ) +
                   G.integrate(G.dot(v[0], G.dot(nΓ, gu)) - G.dot(nΓ, v[0]) * pe, dΓ))
    A, b = G.assemble_matrix_and_vector(a, l, X, Y, G.SparseMatrixAssembler(X, Y))
    # oracle: the cell blocks and loads of _stokes_oracle plus the Neumann face loads
    import conforming_oracle as CO
    Xc = model.node_coordinates()[model.cell_node_ids() - 1]
    qx, qw = oracle.tensor_quadrature(2, 2)
    s1, g1 = oracle.qk_tabulate(2, 1, qx)
    s2, g2 = oracle.qk_tabulate(2, 2, qx)
    wdet, gp2 = CO.physical_gradients(Xc, qw, g1, g2)
    C = Xc.shape[0]
    mats = {(0, 0): CO.vector_laplacian_cells(wdet, gp2),
            (1, 0): CO.divergence_cells(wdet, s1, gp2, scalar_test=True),
            (0, 1): -CO.divergence_cells(wdet, s1, gp2, scalar_test=False)}
    xq = CO.q1_map_points(Xc, s1)
    v0 = CO.ncube_loads(wdet, s2, np.tile([-5.0, 5.0], (C, 1)))
    mo = oracle.cartesian_model((0, 2, 0, 2), (3, 3))
    faces = oracle.boundary_faces(mo, [6, 7, 8])
    for i in range(2):
        gi = lambda x, n, i=i: (n @ gu(np.asarray(x)))[i] - n[i] * pe(np.asarray(x))
        _, fv = oracle.boundary_face_arrays(mo, faces, 2, 0.0, 0.0, gi, None, order=2)
        for k, fc in enumerate(faces):
            v0[fc[0], 9 * i:9 * (i + 1)] += fv[k]
    vecs = {0: v0, 1: np.einsum("cq,cq,qa->ca", wdet, 2.0 * xq[..., 0], s1)}
    cds = [np.asarray(V.cell_dof_ids), np.asarray(Q.cell_dof_ids)]
    cp, rv, nz, br = MF.assemble(cds, [V.nfree, Q.nfree], [V.ndirichlet, Q.ndirichlet], mats,
                                 vecs, dvals=[U.dirichlet_values(), P.dirichlet_values()])
    np.testing.assert_array_equal(A.colptr, cp)
    np.testing.assert_array_equal(A.rowval, rv)
    assert relmax(A.nzval, nz) <= 1e-12 and relmax(b, br) <= 1e-12
    x = spla.spsolve(A.to_scipy().tocsc(), b)
    uI, pI = G.interpolate(ue, U).free_values, G.interpolate(pe, P).free_values
    assert np.max(np.abs(x[:V.nfree] - uI)) <= 1e-9 * np.max(np.abs(uI))
    assert np.max(np.abs(x[V.nfree:] - pI)) <= 1e-9 * np.max(np.abs(pI))


@pytest.mark.parametrize("mapped", [True, False])
def test_stokes_taylor_hood_3d(oracle, mapped):
    """Taylor–Hood Stokes on a 3-D mesh, Cartesian (the vector Laplacian on the
    vector-Q2 lattice kernel) or under a sheared affine map (the vector Laplacian of the
    general n-cube element kernel and the divergence couplings with the Q1 geometry map), u = (x1² + 2 x2², -x1², x3), p = x1 + 3 x2 + x3,
    velocity Dirichlet data on the boundary, pressure fixed at one vertex: against the
    oracle (pattern bit-exact, 1e-12) and the direct solve reproduces u and p to 1e-9
    (both lie in the spaces under an affine map)."""
    import scipy.sparse.linalg as spla
    import conforming_oracle as CO
    announce("Stokes Taylor-Hood 3-D 3x3x2 mapped=%s" % mapped)
    M = np.array([[1.0, 0.2, 0.1], [0.0, 0.9, 0.15], [0.05, 0.0, 1.1]])
    # Cartesian: the vector Laplacian takes the vector-Q2 lattice kernel
    model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), (3, 3, 2),
                                     map=(lambda x: x @ M.T) if mapped else None)
    ue = lambda x: np.stack([x[..., 0] ** 2 + 2 * x[..., 1] ** 2, -x[..., 0] ** 2, x[..., 2]],
                            axis=-1)
    pe = lambda x: x[..., 0] + 3 * x[..., 1] + x[..., 2]
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.VectorValue[3, G.Float64], 2),
                      dirichlet_tags="boundary")
    Q = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1), dirichlet_tags="tag_01")
    U, P = G.TrialFESpace(V, ue), G.TrialFESpace(Q, pe)
    Y, X = G.MultiFieldFESpace([V, Q]), G.MultiFieldFESpace([U, P])
    dΩ = G.Measure(G.Triangulation(model), 4)
    f = (-5.0, 5.0, 1.0)                              # -Δu + ∇p
    g = lambda x: 2.0 * x[..., 0] + 1.0                # ∇⋅u
    a = lambda u, v: G.integrate(G.inner(G.gradient(v[0]), G.gradient(u[0])) -
                                 G.divergence(v[0]) * u[1] + v[1] * G.divergence(u[0]), dΩ)
    l = lambda v: G.integrate(G.dot(v[0], G.VectorValue(*f)) + v[1] * g, dΩ)
    A, b = G.assemble_matrix_and_vector(a, l, X, Y, G.SparseMatrixAssembler(X, Y))
    Xc = model.node_coordinates()[model.cell_node_ids() - 1]
    qx, qw = oracle.tensor_quadrature(3, 4)
    s1, g1 = oracle.qk_tabulate(3, 1, qx)
    s2, g2 = oracle.qk_tabulate(3, 2, qx)
    wdet, gp2 = CO.physical_gradients(Xc, qw, g1, g2)
    C = Xc.shape[0]
    mats = {(0, 0): CO.vector_laplacian_cells(wdet, gp2),
            (1, 0): CO.divergence_cells(wdet, s1, gp2, scalar_test=True),
            (0, 1): -CO.divergence_cells(wdet, s1, gp2, scalar_test=False)}
    xq = CO.q1_map_points(Xc, s1)
    vecs = {0: CO.ncube_loads(wdet, s2, np.tile(f, (C, 1))),
            1: np.einsum("cq,cq,qa->ca", wdet, g(xq), s1)}
    cds = [np.asarray(V.cell_dof_ids), np.asarray(Q.cell_dof_ids)]
    cp, rv, nz, br = MF.assemble(cds, [V.nfree, Q.nfree], [V.ndirichlet, Q.ndirichlet], mats,
                                 vecs, dvals=[U.dirichlet_values(), P.dirichlet_values()])
    np.testing.assert_array_equal(A.colptr, cp)
    np.testing.assert_array_equal(A.rowval, rv)
    assert relmax(A.nzval, nz) <= 1e-12 and relmax(b, br) <= 1e-12
    x = spla.spsolve(A.to_scipy().tocsc(), b)
    uI, pI = G.interpolate(ue, U).free_values, G.interpolate(pe, P).free_values
    assert np.max(np.abs(x[:V.nfree] - uI)) <= 1e-9 * np.max(np.abs(uI))
    assert np.max(np.abs(x[V.nfree:] - pI)) <= 1e-9 * np.max(np.abs(pI))


def biform3b(u, v, dΩ):
    (u1, u2, u3), (v1, v2, v3) = u, v
    return G.integrate(G.inner(G.gradient(u1), G.gradient(v1)) + G.inner(u2, v2) +
                       G.inner(u1, v2) - G.inner(u3, v2) + G.inner(u2, v3), dΩ)


def _matrices(A):
    return [B for row in A.blocks for B in row] if isinstance(A, G.BlockSparseMatrix) else [A]


@pytest.mark.parametrize("blocked", [False, True])
def test_device_composition_equals_host(oracle, blocked):
    """The first assembly of a form on a multi-field assembler composes the field-pair
    values on the host and registers the composition on the device; later assemblies
    move each pair's values into the composite on the device, right after the pair's
    own assembly (the five pairs of the three-field form share one assembler, so each
    move precedes the next pair's reassembly).  Bitwise equal to the host composition;
    a changed form on the same pairs (first call of its own, already on the device)
    equals a fresh assembler's host composition bitwise and the oracle within 1e-12."""
    announce("device composition, three fields, blocked=%s" % blocked)
    V, U, dΩ = spaces(3)
    if blocked:
        st = G.BlockMultiFieldStyle(2, (1, 2))
        X, Y = G.MultiFieldFESpace([U] * 3, style=st), G.MultiFieldFESpace([V] * 3, style=st)
    else:
        X, Y = G.MultiFieldFESpace([U] * 3), G.MultiFieldFESpace([V] * 3)
    a = lambda u, v: biform3(u, v, dΩ)
    ab = lambda u, v: biform3b(u, v, dΩ)
    l = lambda v: liform3(v, dΩ)
    assem = G.SparseMatrixAssembler(X, Y)
    A1, b1 = G.assemble_matrix_and_vector(a, l, X, Y, assem)
    assert assem._composers, "the first assembly registers the device composition"
    A2, b2 = G.assemble_matrix_and_vector(a, l, X, Y, assem)
    for B1, B2 in zip(_matrices(A1), _matrices(A2)):
        np.testing.assert_array_equal(B2.colptr, B1.colptr)
        np.testing.assert_array_equal(B2.rowval, B1.rowval)
        np.testing.assert_array_equal(B2.nzval, B1.nzval)
    Ab, _ = G.assemble_matrix_and_vector(ab, l, X, Y, assem)
    Ar, _ = G.assemble_matrix_and_vector(ab, l, X, Y, G.SparseMatrixAssembler(X, Y))
    for B1, B2 in zip(_matrices(Ar), _matrices(Ab)):
        np.testing.assert_array_equal(B2.nzval, B1.nzval)
    if not blocked:
        pairs = dict(PAIRS3)
        pairs[(1, 0)] = ("mass", 1.0)
        pairs[(2, 1)] = ("mass", 1.0)
        _, (cp, rv, nz, _) = oracle_system(3, pairs, VECS3)
        np.testing.assert_array_equal(Ab.rowval, rv)
        assert relmax(Ab.nzval, nz) <= 1e-12
    # the matrix-free call in between leaves the composition in place
    G.assemble_vector(l, assem, Y)
    A3, _ = G.assemble_matrix_and_vector(a, l, X, Y, assem)
    for B1, B3 in zip(_matrices(A1), _matrices(A3)):
        np.testing.assert_array_equal(B3.nzval, B1.nzval)


def test_stokes_device_composition_equals_host(oracle):
    """Taylor–Hood Stokes assembled three times on one assembler: the first call composes
    on the host, the later ones move the vector Laplacian and both rectangular coupling
    blocks into the composite on the device; matrices and vectors bitwise equal."""
    announce("Stokes device composition (8, 6)")
    _, _, env = _stokes((8, 6), calls=3)
    runs, assem = env[-2], env[-1]
    assert assem._composers
    A1, b1 = runs[0]
    for A, b in runs[1:]:
        np.testing.assert_array_equal(A.colptr, A1.colptr)
        np.testing.assert_array_equal(A.rowval, A1.rowval)
        np.testing.assert_array_equal(A.nzval, A1.nzval)
        np.testing.assert_array_equal(b, b1)


@pytest.mark.parametrize("blocked", [False, True])
def test_multifield_inplace_equals_fresh(blocked):
    """assemble_matrix_and_vector! on a multi-field assembler: A's stored values and b
    are overwritten in place (the device composition copies straight into A's values),
    bitwise equal to a fresh assembler's result; the first in-place call of a fresh
    assembler (host composition) too; a matrix of another layout is rejected."""
    announce("multi-field in place, three fields, blocked=%s" % blocked)
    V, U, dΩ = spaces(3)
    if blocked:
        st = G.BlockMultiFieldStyle(2, (1, 2))
        X, Y = G.MultiFieldFESpace([U] * 3, style=st), G.MultiFieldFESpace([V] * 3, style=st)
    else:
        X, Y = G.MultiFieldFESpace([U] * 3), G.MultiFieldFESpace([V] * 3)
    a = lambda u, v: biform3(u, v, dΩ)
    ab = lambda u, v: biform3b(u, v, dΩ)
    l = lambda v: liform3(v, dΩ)
    arr = lambda x: x.to_array() if isinstance(x, G.BlockVector) else x
    assem = G.SparseMatrixAssembler(X, Y)
    A, b = G.assemble_matrix_and_vector(a, l, X, Y, assem)
    Ar, br = G.assemble_matrix_and_vector(ab, l, X, Y, G.SparseMatrixAssembler(X, Y))
    nzs = [B.nzval for B in _matrices(A)]
    A2, b2 = G.assemble_matrix_and_vector_inplace(ab, l, A, b, assem)
    assert A2 is A and b2 is b
    for B, Bref, nz in zip(_matrices(A), _matrices(Ar), nzs):
        assert B.nzval is nz
        np.testing.assert_array_equal(B.nzval, Bref.nzval)
    np.testing.assert_array_equal(arr(b), arr(br))
    A1, _ = G.assemble_matrix_and_vector(a, l, X, Y, G.SparseMatrixAssembler(X, Y))
    assert G.assemble_matrix_inplace(a, A, G.SparseMatrixAssembler(X, Y)) is A
    for B, Bref in zip(_matrices(A), _matrices(A1)):
        np.testing.assert_array_equal(B.nzval, Bref.nzval)
    X2, Y2 = G.MultiFieldFESpace([U] * 2), G.MultiFieldFESpace([V] * 2)
    Aother = G.assemble_matrix(lambda u, v: biform2(u, v, dΩ), G.SparseMatrixAssembler(X2, Y2),
                               X2, Y2)
    with pytest.raises(ValueError):
        G.assemble_matrix_inplace(a, Aother, assem)
