"""Re-assembly into a fixed pattern (assemble_matrix!, assemble_vector!,
assemble_matrix_and_vector!; src/FESpaces/Assemblers.jl:308-346 ->
SparseMatrixAssemblers.jl:60-63), the numeric path of a nonlinear solve (needs an MI355X).

A Picard loop for -∇·(κ(u)∇u) = 1, κ(u) = 1 + (10u)² per cell, re-assembles its system in
place each iteration: the matrix and vector objects, the pattern and the device context
stay the same, and every iterate's system equals the oracle's assembly with that κ
(FP64 bar: 1e-12 relative).  On the element-list path (P2 tetrahedra) the in-place
re-assembly is bit-identical to a fresh assembly with a new assembler.
"""
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import gridap_hip as G

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def announce(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def cell_means(V, u_free):
    """Per-cell mean of the nodal values (homogeneous Dirichlet dofs are 0)."""
    cd = np.asarray(V.get_cell_dof_ids())
    vals = np.where(cd > 0, u_free[np.maximum(cd, 1) - 1], 0.0)
    return vals.mean(axis=1)


def test_picard_reassembly_q1_matches_oracle(oracle):
    announce("picard Q1 12x10x8 in-place re-assembly")
    dom, part = (0, 1, 0, 1, 0, 1), (12, 10, 8)
    model = G.CartesianDiscreteModel(dom, part)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1),
                      dirichlet_tags="boundary")
    U = G.TrialFESpace(V)
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 2)
    assem = G.SparseMatrixAssembler(U, V)

    def forms(kappa):
        k = G.CellField(kappa, Ω)
        return (lambda u, v: G.integrate(k * G.inner(G.gradient(v), G.gradient(u)), dΩ),
                lambda v: G.integrate(1.0 * v, dΩ))

    ncells = int(np.prod(part))
    a, l = forms(np.ones(ncells))
    A, b = G.assemble_matrix_and_vector(a, l, U, V, assem)
    ctx, colptr, rowval, nzval = assem.ctx, A.colptr.copy(), A.rowval.copy(), A.nzval
    u = spla.spsolve(A.to_scipy().tocsc(), b)
    steps = []
    for it in range(4):
        kappa = 1.0 + (10.0 * cell_means(V, u)) ** 2
        a, l = forms(kappa)
        A2, b2 = G.assemble_matrix_and_vector_inplace(a, l, A, b, assem)
        assert A2 is A and b2 is b and A.nzval is nzval and assem.ctx is ctx
        np.testing.assert_array_equal(A.colptr, colptr)
        np.testing.assert_array_equal(A.rowval, rowval)
        ref = oracle.poisson_q1(dom, part, kappa=kappa)
        np.testing.assert_array_equal(A.colptr, ref["colptr"])
        assert relmax(A.nzval, ref["nzval"]) <= FP64_TOL
        assert relmax(b, ref["b"]) <= FP64_TOL
        u_new = spla.spsolve(A.to_scipy().tocsc(), b)
        steps.append(np.max(np.abs(u_new - u)))
        u = u_new
    # the Picard map contracts on this problem: the updates shrink
    assert steps[-1] < steps[0]


def test_inplace_matrix_and_vector_match_fresh_assembly():
    announce("in-place assemble_matrix! / assemble_vector! Q1 9x7x5")
    part = (9, 7, 5)
    model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), part)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1),
                      dirichlet_tags="boundary")
    U = G.TrialFESpace(V)
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 2)
    rng = np.random.default_rng(3)
    assem = G.SparseMatrixAssembler(U, V)
    k1 = G.CellField(np.exp(rng.normal(size=int(np.prod(part)))), Ω)
    k2 = G.CellField(np.exp(rng.normal(size=int(np.prod(part)))), Ω)
    f2 = G.CellField(rng.uniform(0.5, 2.0, size=int(np.prod(part))), Ω)
    A = G.assemble_matrix(lambda u, v: G.integrate(k1 * G.inner(G.gradient(v), G.gradient(u)), dΩ),
                          assem, U, V)
    b = G.assemble_vector(lambda v: G.integrate(1.0 * v, dΩ), assem, V)
    a2 = lambda u, v: G.integrate(k2 * G.inner(G.gradient(v), G.gradient(u)), dΩ)
    l2 = lambda v: G.integrate(f2 * v, dΩ)
    nz = A.nzval
    assert G.assemble_matrix_inplace(a2, A, assem) is A and A.nzval is nz
    assert G.assemble_vector_inplace(l2, b, assem) is b
    fresh = G.SparseMatrixAssembler(U, V)
    A_ref = G.assemble_matrix(a2, fresh, U, V)
    b_ref = G.assemble_vector(l2, fresh, V)
    np.testing.assert_array_equal(A.nzval, A_ref.nzval)
    np.testing.assert_array_equal(b, b_ref)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_inplace_p2_tets_bit_identical_to_fresh(precision):
    announce("in-place P2 6^3x6 jittered precision=%s" % precision)
    part = (6, 6, 6)
    model = G.simplexify(G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), part)).jittered(9)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 2),
                      dirichlet_tags="boundary")
    U = G.TrialFESpace(V)
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 4)
    rng = np.random.default_rng(8)
    C = model.num_cells()
    k1, k2 = (G.CellField(np.exp(rng.normal(size=C)), Ω) for _ in range(2))
    form = lambda k: (lambda u, v: G.integrate(k * G.inner(G.gradient(v), G.gradient(u)), dΩ))
    l = lambda v: G.integrate(1.0 * v, dΩ)
    assem = G.SparseMatrixAssembler(U, V, precision=precision)
    A, b = G.assemble_matrix_and_vector(form(k1), l, U, V, assem)
    G.assemble_matrix_and_vector_inplace(form(k2), l, A, b, assem)
    A_ref, b_ref = G.assemble_matrix_and_vector(form(k2), l, U, V,
                                                G.SparseMatrixAssembler(U, V, precision=precision))
    np.testing.assert_array_equal(A.colptr, A_ref.colptr)
    np.testing.assert_array_equal(A.nzval, A_ref.nzval)
    np.testing.assert_array_equal(b, b_ref)


def test_inplace_rejects_foreign_pattern():
    announce("in-place pattern mismatch error")
    m1 = G.CartesianDiscreteModel((0, 1, 0, 1), (6, 5))
    m2 = G.CartesianDiscreteModel((0, 1, 0, 1), (5, 6))
    spaces = []
    for m in (m1, m2):
        V = G.TestFESpace(m, G.ReferenceFE(G.lagrangian, G.Float64, 1), dirichlet_tags="boundary")
        spaces.append((G.TrialFESpace(V), V, G.Measure(G.Triangulation(m), 2)))
    (U1, V1, d1), (U2, V2, d2) = spaces
    A1 = G.assemble_matrix(lambda u, v: G.integrate(G.inner(G.gradient(v), G.gradient(u)), d1),
                           U1, V1)
    assem2 = G.SparseMatrixAssembler(U2, V2)
    with pytest.raises(ValueError):
        G.assemble_matrix_inplace(
            lambda u, v: G.integrate(G.inner(G.gradient(v), G.gradient(u)), d2), A1, assem2)
    with pytest.raises(ValueError):
        G.assemble_vector_inplace(lambda v: G.integrate(1.0 * v, d2), np.zeros(3), assem2)
