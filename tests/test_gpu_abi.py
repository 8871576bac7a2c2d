"""The C ABI called the way a Julia `ccall` binding calls it (needs an MI355X).

gh_numeric is the one-call numeric entry of the binding: host κ, f and Dirichlet
values in, nzval and b out, on a pattern built from Gridap's own cell_dof_ids
(here: the oracle's CLagrangian numbering).  Bar: pattern bit-exact, FP64 values
within 1e-12 relative (max-abs error / max-abs value) of the oracle.
Also: state errors when a mesh is replaced under an uploaded space, and the solver
working on the matrix it is given (gh_values_set).
"""
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import gridap_hip as G
from gridap_hip import _lib

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300) if a.size else 0.0


def _table_ptrs(ncells, nloc):
    return (np.arange(ncells + 1, dtype=np.int64) * nloc + 1).astype(np.int32)


def _ctx_from_oracle(ref, D, part):
    """Context fed with the oracle's model, numbering and tabulation (Gridap's data)."""
    m = ref["model"]
    ctx = _lib.Context(0)
    ctx.mesh_cartesian(m["origin"], m["sizes"], m["partition"])
    cd = np.ascontiguousarray(ref["cell_dofs"], dtype=np.int32)
    ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, cd.ravel(), _table_ptrs(*cd.shape), ref["nfree"],
                     ref["ndiri"])
    ctx.reffe_tabulation(ref["qw"], ref["shp"], ref["grad"].reshape(len(ref["qw"]), -1))
    return ctx


@pytest.mark.parametrize("path", [_lib.GH_PATH_AUTO, _lib.GH_PATH_GENERAL])
def test_gh_numeric_matches_oracle(oracle, path):
    """gh_numeric with κ, f and nonzero Dirichlet data, 3-D Q1 (13x11x9)."""
    part = (13, 11, 9)
    announce("gh_numeric Q1 %s path %d" % (part, path))
    rng = np.random.default_rng(31)
    C = int(np.prod(part))
    kappa = np.exp(rng.normal(size=C))
    fsrc = rng.uniform(-1, 1, C)
    g = lambda x: 1.0 + x[0] - 2.0 * x[1] * x[2]
    ref = oracle.poisson_q1((0, 1) * 3, part, kappa=kappa, fsrc=fsrc, g=g)
    ctx = _ctx_from_oracle(ref, 3, part)
    nnz = ctx.symbolic()
    colptr, rowval = ctx.pattern_get()
    np.testing.assert_array_equal(colptr, ref["colptr"])
    np.testing.assert_array_equal(rowval, ref["rowval"])
    assert nnz == len(ref["rowval"])
    if path == _lib.GH_PATH_GENERAL:
        # the general kernel through gh_assemble on the same uploaded inputs
        ctx.coef_upload(kappa, fsrc, ref["dirichlet_vals"])
        ctx.assemble(_lib.GH_FORM_DIFFUSION, _lib.GH_PREC_FP64, path)
        nz, b = ctx.values_get()
    else:
        nz, b = ctx.numeric(_lib.GH_FORM_DIFFUSION, kappa, fsrc, _lib.GH_PREC_FP64,
                            ref["dirichlet_vals"])
    assert relmax(nz, ref["nzval"]) <= 1e-12
    assert relmax(b, ref["b"]) <= 1e-12
    # the same call again with other coefficients re-assembles into the same pattern
    nz2, b2 = ctx.numeric(_lib.GH_FORM_DIFFUSION, 2.0 * kappa, None, _lib.GH_PREC_FP64, None)
    ref2 = oracle.poisson_q1((0, 1) * 3, part, kappa=2.0 * kappa)
    assert relmax(nz2, ref2["nzval"]) <= 1e-12
    assert relmax(b2, ref2["b"]) <= 1e-12
    with pytest.raises(ValueError):
        ctx.numeric(_lib.GH_FORM_DIFFUSION, kappa[:-1])
    ctx.close()


def test_mesh_reupload_voids_the_space():
    """A new (larger) mesh under an uploaded space: gh_symbolic refuses with
    GH_ERR_STATE before any kernel runs; a space uploaded for the new mesh works."""
    announce("mesh re-upload voids the FE space")
    ctx = _lib.Context(0)
    ctx.mesh_cartesian([0, 0], [0.5, 0.5], [2, 2])
    cd = np.array([[1, 2, 4, 5], [2, 3, 5, 6], [4, 5, 7, 8], [5, 6, 8, 9]], np.int32)
    ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, cd.ravel(), _table_ptrs(4, 4), 9, 0)
    assert ctx.symbolic() == 49
    ctx.mesh_cartesian([0, 0], [0.25, 0.25], [4, 4])
    with pytest.raises(G.GridapHipError) as e:
        ctx.symbolic()
    assert e.value.code == _lib.GH_ERR_STATE
    with pytest.raises(G.GridapHipError) as e:
        ctx.assemble()
    assert e.value.code == _lib.GH_ERR_STATE
    # unstructured re-upload with the same cell count also voids it
    ctx.mesh_cartesian([0, 0], [0.5, 0.5], [2, 2])
    ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, cd.ravel(), _table_ptrs(4, 4), 9, 0)
    xyz = np.array([[x, y] for y in (0, .5, 1) for x in (0, .5, 1)], dtype=np.float64)
    ctx.mesh_upload(xyz, cd.ravel(), _table_ptrs(4, 4))
    with pytest.raises(G.GridapHipError) as e:
        ctx.symbolic()
    assert e.value.code == _lib.GH_ERR_STATE
    ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, cd.ravel(), _table_ptrs(4, 4), 9, 0)
    assert ctx.symbolic() == 49
    ctx.close()


def _poisson(assem, model, kappa, fval):
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 2)
    k = G.CellField(kappa, Ω)
    a = lambda u, v: G.integrate(k * G.inner(G.gradient(v), G.gradient(u)), dΩ)
    l = lambda v: G.integrate(fval * v, dΩ)
    return G.AffineFEOperator(a, l, assem.U, assem.V, assem), l


def test_solve_uses_the_operators_matrix():
    """solve(op) works on op's matrix, whatever the assembler assembled since: a
    vector-only assembly and a second operator on the same assembler in between."""
    part = (10, 9, 8)
    announce("solve uses op.matrix after other assemblies on one assembler")
    model = G.CartesianDiscreteModel((0, 1) * 3, part)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1),
                      dirichlet_tags="boundary")
    U = G.TrialFESpace(V)
    assem = G.SparseMatrixAssembler(U, V)
    rng = np.random.default_rng(5)
    C = int(np.prod(part))
    op1, l1 = _poisson(assem, model, np.exp(rng.normal(size=C)), 1.0)
    bvec = G.assemble_vector(l1, assem, V)
    np.testing.assert_array_equal(bvec, op1.get_vector())
    op2, _ = _poisson(assem, model, np.exp(rng.normal(size=C)) * 5.0, 2.0)
    for op in (op1, op2, op1):
        x = G.solve(op, G.PCGSolver(rtol=1e-12)).get_free_dof_values()
        xd = spla.spsolve(op.get_matrix().to_scipy().tocsc(), op.get_vector())
        assert relmax(x, xd) <= 1e-9
    # the device SpMV follows the loaded matrix too
    y = rng.normal(size=V.num_free_dofs())
    assem.load_matrix(op2.get_matrix())
    assert relmax(assem.ctx.spmv(y), op2.get_matrix().to_scipy() @ y) <= 1e-13
    with pytest.raises(ValueError):
        assem.ctx.values_set(np.zeros(3))


def test_pattern_info_lattice():
    announce("pattern info: lattice and box on Cartesian Q1")
    model = G.CartesianDiscreteModel((0, 1) * 3, (6, 5, 4))
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1),
                      dirichlet_tags="boundary")
    op, _ = _poisson(G.SparseMatrixAssembler(G.TrialFESpace(V), V), model,
                     np.ones(120), 1.0)
    info = op.assem.ctx.pattern_info()
    assert info["lattice"] == 1 and info["box"] == 1 and info["elemlist"] == 0
    assert info["nnz"] == op.get_matrix().nnz()
