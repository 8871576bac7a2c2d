"""The matrices that reach the device SpMV and PCG are symmetric bit for bit on every
default numeric path (needs an MI355X): the SpMV reads the CSC columns as rows, so it
computes Aᵀx, which is A x exactly when A is symmetric.  Checked for the Q1 lattice
kernel (Cartesian and mapped), the per-cell kernel, P1/P2 tetrahedra and the vector-Q2
lattice kernel, with lognormal coefficients and Dirichlet data."""
import sys

import numpy as np
import pytest

import gridap_hip as G
from gridap_hip import _lib

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def assert_bitwise_symmetric(A):
    S = A.to_scipy().tocsr()
    T = S.T.tocsr()
    S.sort_indices()
    T.sort_indices()
    np.testing.assert_array_equal(S.indptr, T.indptr)
    np.testing.assert_array_equal(S.indices, T.indices)
    np.testing.assert_array_equal(S.data, T.data)


def q1_op(part, path="auto", precision="fp64", warp=None):
    model = G.CartesianDiscreteModel((0, 1) * len(part), part, map=warp)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1),
                      dirichlet_tags=["boundary"])
    U = G.TrialFESpace(V, lambda x: 1.0 + x[..., 0])
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 2)
    k = G.CellField(np.exp(np.random.default_rng(1).normal(size=int(np.prod(part)))), Ω)
    a = lambda u, v: G.integrate(k * G.inner(G.gradient(v), G.gradient(u)), dΩ)
    l = lambda v: G.integrate(1.0 * v, dΩ)
    return G.AffineFEOperator(a, l, U, V, G.SparseMatrixAssembler(U, V, precision=precision,
                                                                  path=path))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_q1_paths_symmetric(precision):
    warp = lambda x: x + 0.03 * np.prod(np.sin(np.pi * x), axis=-1)[..., None]
    for part, path, w, kern in (((12, 10, 9), "auto", None, _lib.GH_KERNEL_Q1_LATTICE),
                                ((12, 10, 9), "auto", warp, _lib.GH_KERNEL_Q1_LATTICE),
                                ((12, 10, 9), "general", warp, _lib.GH_KERNEL_CELL)):
        announce("symmetry Q1 %s path=%s mapped=%s %s" % (part, path, w is not None, precision))
        op = q1_op(part, path, precision, w)
        assert op.assem.ctx.pattern_info()["last_kernel"] == kern
        assert_bitwise_symmetric(op.get_matrix())


@pytest.mark.parametrize("order", [1, 2])
def test_tet_symmetric(order):
    announce("symmetry P%d tets" % order)
    model = G.simplexify(G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), (6, 5, 4))).jittered(3)
    reffe = G.ReferenceFE(G.lagrangian, G.Float64, order)
    V = G.TestFESpace(model, reffe, dirichlet_tags="boundary")
    U = G.TrialFESpace(V)
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 2 * order)
    k = G.CellField(np.exp(np.random.default_rng(2).normal(size=model.num_cells())), Ω)
    for prec in ("fp64", "fp32"):
        op = G.AffineFEOperator(lambda u, v: G.integrate(k * G.inner(G.gradient(v), G.gradient(u)), dΩ),
                                lambda v: G.integrate(1.0 * v, dΩ), U, V,
                                G.SparseMatrixAssembler(U, V, precision=prec))
        assert_bitwise_symmetric(op.get_matrix())


def test_q2_lattice_symmetric():
    announce("symmetry vector Q2 lattice 7x6x5")
    part = (7, 6, 5)
    model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), part)
    reffe = G.ReferenceFE(G.lagrangian, G.VectorValue[3, G.Float64], 2)
    V = G.TestFESpace(model, reffe, dirichlet_tags=["tag_25"])
    U = G.TrialFESpace(V)
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 4)
    E = G.CellField(np.exp(np.random.default_rng(3).normal(0, 0.5, int(np.prod(part)))), Ω)
    σ = G.IsotropicStress(E, 0.3)
    for prec in ("fp64", "fp32"):
        op = G.AffineFEOperator(lambda u, v: G.integrate(G.inner(G.ε(v), σ(G.ε(u))), dΩ),
                                lambda v: G.integrate(G.dot(v, G.VectorValue(0.0, 0.0, -1.0)), dΩ),
                                U, V, G.SparseMatrixAssembler(U, V, precision=prec))
        assert op.assem.ctx.pattern_info()["last_kernel"] == _lib.GH_KERNEL_Q2_LATTICE
        assert_bitwise_symmetric(op.get_matrix())
