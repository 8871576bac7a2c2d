"""Parity of the vector-Q2 elasticity path with the FP64 oracle (needs an MI355X).

Bar: CSC pattern bit-exact with the oracle's create_from_nz output on the same
conforming cell_dof_ids; FP64 mode: nzval and b within 1e-12 relative (max-abs error /
max-abs value); FP32 mode (element products on the matrix cores): relative Frobenius
error <= 1e-6.  Size-independent checks: the quadratic patch identity A u_I = b,
rigid-body modes in the kernel of the unconstrained matrix, FP32 vs FP64.
"""
import sys

import numpy as np
import pytest

import gridap_hip as G
from gridap_hip import _lib
import conforming_oracle as CO

pytestmark = pytest.mark.gpu

CLAMP_X0 = ["tag_01", "tag_03", "tag_05", "tag_07", "tag_13", "tag_15", "tag_17", "tag_19",
            "tag_25"]


def announce(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300) if a.size else 0.0


def relfrob(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def build(part, E, nu=0.3, force=(0.0, 0.0, -1.0), tags=CLAMP_X0, g=None, precision="fp64",
          path="auto"):
    model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), part)
    reffe = G.ReferenceFE(G.lagrangian, G.VectorValue[3, G.Float64], 2)
    V = G.TestFESpace(model, reffe, dirichlet_tags=list(tags))
    U = G.TrialFESpace(V, g)
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 4)
    σ = G.IsotropicStress(G.CellField(E, Ω) if not np.isscalar(E) else E, nu)
    a = lambda u, v: G.integrate(G.inner(G.ε(v), σ(G.ε(u))), dΩ)
    l = lambda v: G.integrate(G.dot(v, G.VectorValue(*force)), dΩ)
    assem = G.SparseMatrixAssembler(U, V, precision=precision, path=path)
    return G.AffineFEOperator(a, l, U, V, assem), V


def entities(tags):
    if tags == ["boundary"]:
        return CO.boundary_entities(3)
    return {int(t.split("_")[1]) for t in tags}


def oracle_system(oracle, part, E, nu, force, tags, g=None):
    r = CO.elasticity_q2_box(oracle, part, E, nu, np.asarray(force), [entities(list(tags))], g=g)
    cp, rv, nz = oracle.assemble_csc(r["cell_dofs"], r["cellmats"], r["nfree"], r["nfree"])
    b = oracle.assemble_vector(r["cell_dofs"], r["cellvecs"], r["nfree"])
    return r, cp, rv, nz, b


@pytest.mark.parametrize("path", ["auto", "general"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_elasticity_matches_oracle(oracle, precision, path):
    part = (3, 2, 2)
    announce("elasticity Q2 %s %s clamped x=0 path=%s" % (part, precision, path))
    E = np.exp(np.random.default_rng(3).normal(0, 0.5, 12))
    op, V = build(part, E, precision=precision, path=path)
    info = op.assem.ctx.pattern_info()
    assert info["lattice"] == 1 and info["elemlist"] == 1
    assert info["last_kernel"] == (_lib.GH_KERNEL_Q2_LATTICE if path == "auto"
                                   else _lib.GH_KERNEL_COLUMN_SUM)
    r, cp, rv, nz, b = oracle_system(oracle, part, E, 0.3, (0, 0, -1), CLAMP_X0)
    np.testing.assert_array_equal(V.cell_dof_ids, r["cell_dofs"])
    A = op.get_matrix()
    np.testing.assert_array_equal(A.colptr, cp)
    np.testing.assert_array_equal(A.rowval, rv)
    if precision == "fp64":
        assert relmax(A.nzval, nz) <= 1e-12
        assert relmax(op.get_vector(), b) <= 1e-12
    else:
        assert relfrob(A.nzval, nz) <= 1e-6
        assert relfrob(op.get_vector(), b) <= 1e-6


def test_elasticity_quadratic_patch_and_solve(oracle):
    part = (3, 3, 2)
    announce("elasticity Q2 quadratic patch + lifting + PCG %s" % (part,))
    nu = 0.3
    lam, mu = CO.lame(1.0, nu)
    f = -np.array([lam + mu, lam + mu, lam + 3 * mu])
    u = lambda x: np.stack([x[..., 0] * x[..., 1], x[..., 1] * x[..., 2],
                            x[..., 2] * x[..., 0] + x[..., 0] ** 2], axis=-1)
    op, V = build(part, 1.0, nu, f, ["boundary"], g=u)
    uI = u(V.free_dof_coords)[np.arange(V.nfree), V.free_dof_comp]
    A, b = op.get_matrix().to_scipy(), op.get_vector()
    assert np.max(np.abs(A @ uI - b)) <= 1e-12 * np.max(np.abs(b))
    r, cp, rv, nz, bo = oracle_system(oracle, part, np.ones(18), nu, f, ["boundary"],
                                      g=lambda x: u(np.asarray(x)))
    assert relmax(b, bo) <= 1e-12
    x = G.solve(op).get_free_dof_values()
    assert np.max(np.abs(x - uI)) <= 1e-8


def test_elasticity_is_deterministic():
    announce("elasticity Q2 repeat bit-identical (4^3)")
    E = np.exp(np.random.default_rng(4).normal(0, 0.5, 64))
    op, V = build((4, 4, 4), E, precision="fp32")
    ctx = op.assem.ctx
    nz1, b1 = ctx.values_get()
    ctx.assemble(_lib.GH_FORM_ELASTICITY, op.assem.precision, op.assem.path)
    nz2, b2 = ctx.values_get()
    np.testing.assert_array_equal(nz1, nz2)
    np.testing.assert_array_equal(b1, b2)


def test_elasticity_fp32_vs_fp64_and_rigid_modes():
    part = (12, 12, 12)
    announce("elasticity Q2 %s FP32 vs FP64, rigid modes" % (part,))
    E = np.exp(np.random.default_rng(12).normal(0, 0.5, 12 ** 3))
    op64, V = build(part, E, precision="fp64")
    op32, _ = build(part, E, precision="fp32")
    assert relfrob(op32.get_matrix().nzval, op64.get_matrix().nzval) <= 1e-6
    # without Dirichlet tags the six rigid motions span the kernel of A
    opn, Vn = build(part, E, tags=[], precision="fp32")
    xyz = Vn.free_dof_coords
    comp = Vn.free_dof_comp
    A = opn.get_matrix()
    Ad = opn.assem.ctx
    scale = np.abs(A.nzval).max()
    for k in range(3):
        t = (comp == k).astype(np.float64)
        assert np.abs(Ad.spmv(t)).max() <= 1e-5 * scale
    for a, b in ((0, 1), (1, 2), (2, 0)):
        rot = np.where(comp == a, -xyz[:, b], 0.0) + np.where(comp == b, xyz[:, a], 0.0)
        assert np.abs(Ad.spmv(rot)).max() <= 1e-5 * scale


def test_elasticity_column_sum_matches_list_sum(monkeypatch):
    """Column-range sum (inverse staging over padded columns) vs the fallback sum over
    the int32 lists: same entries, same order, nzval and b bit-identical."""
    part = (5, 4, 3)
    announce("elasticity Q2 column sum vs list sum %s" % (part,))
    E = np.exp(np.random.default_rng(7).normal(0, 0.5, 60))
    for precision in ("fp32", "fp64"):
        monkeypatch.setenv("GRIDAP_HIP_COMPACT_LISTS", "1")
        op1, _ = build(part, E, precision=precision, path="general")
        monkeypatch.setenv("GRIDAP_HIP_COMPACT_LISTS", "0")
        op2, _ = build(part, E, precision=precision, path="general")
        assert op1.assem.ctx.pattern_info()["compact"] == 1
        assert op2.assem.ctx.pattern_info()["compact"] == 0
        np.testing.assert_array_equal(op1.get_matrix().nzval, op2.get_matrix().nzval)
        np.testing.assert_array_equal(op1.get_vector(), op2.get_vector())


@pytest.mark.parametrize("path", ["auto", "general"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_elasticity_17cube_matches_oracle(oracle, precision, path):
    """4913 cells: more cells than waves in the FP32 element kernel's grid (1024
    workgroups x 4 waves), so waves walk several cells in both modes (general path);
    the lattice kernel walks 35^3 nodes with persistent waves (auto path); the whole
    system against the oracle."""
    part = (17, 17, 17)
    announce("elasticity Q2 %s %s vs oracle path=%s" % (part, precision, path))
    C = int(np.prod(part))
    E = np.exp(np.random.default_rng(17).normal(0, 0.5, C))
    op, V = build(part, E, precision=precision, path=path)
    info = op.assem.ctx.pattern_info()
    assert info["elemlist"] == 1 and info["compact"] == 1 and info["inverse"] == 1
    assert info["lattice"] == 1
    r, cp, rv, nz, b = oracle_system(oracle, part, E, 0.3, (0, 0, -1), CLAMP_X0)
    np.testing.assert_array_equal(V.cell_dof_ids, r["cell_dofs"])
    A = op.get_matrix()
    np.testing.assert_array_equal(A.colptr, cp)
    np.testing.assert_array_equal(A.rowval, rv)
    if precision == "fp64":
        assert relmax(A.nzval, nz) <= 1e-12
        assert relmax(op.get_vector(), b) <= 1e-12
    else:
        assert relfrob(A.nzval, nz) <= 1e-6
        assert relfrob(op.get_vector(), b) <= 1e-6


def _config3_E(part, seed=11):
    return np.exp(np.random.default_rng(seed).normal(0, 0.5, int(np.prod(part))))


def test_elasticity_48cube_fp32_vs_fp64():
    """The benchmarked workload at full size (48^3, E=exp(N(0,0.5^2)), ν=0.3,
    f=(0,0,-1), x=0 clamped): FP32 (matrix-core products) against FP64 within 1e-6
    relative Frobenius, nzval and b; NNZ and free dofs equal to the closed counts."""
    part = (48, 48, 48)
    announce("elasticity Q2 48^3 fp32 vs fp64 (config 3)")
    E = _config3_E(part)
    op64, V = build(part, E, precision="fp64")
    assert op64.assem.ctx.pattern_info()["lattice"] == 1
    A64, b64 = op64.get_matrix(), op64.get_vector()
    assert A64.nnz() == 506929500 and V.num_free_dofs() == 2709792
    op64.assem.ctx.close()
    for prec, path in (("fp32", "auto"), ("fp64", "general"), ("fp32", "general")):
        announce("elasticity Q2 48^3 %s path=%s against the FP64 lattice" % (prec, path))
        op, _ = build(part, E, precision=prec, path=path)
        A, b = op.get_matrix(), op.get_vector()
        op.assem.ctx.close()
        np.testing.assert_array_equal(A.rowval, A64.rowval)
        if prec == "fp64":
            assert relmax(A.nzval, A64.nzval) <= 1e-12   # the MFMA element path, FP64
            assert relmax(b, b64) <= 1e-12
        else:
            assert relfrob(A.nzval, A64.nzval) <= 1e-6
            assert relfrob(b, b64) <= 1e-6
            assert not np.array_equal(A.nzval, A64.nzval)
        del A, b


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_elasticity_48cube_rigid_modes(precision):
    """48^3 with the lognormal E and no Dirichlet tags: the three translations and
    three rotations are in the kernel of A (checked with the device SpMV)."""
    part = (48, 48, 48)
    announce("elasticity Q2 48^3 rigid modes %s" % precision)
    opn, Vn = build(part, _config3_E(part), tags=[], precision=precision)
    xyz, comp = Vn.free_dof_coords, Vn.free_dof_comp
    Ad = opn.assem.ctx
    scale = np.abs(opn.get_matrix().nzval).max()
    tol = 1e-5 if precision == "fp32" else 1e-10
    for k in range(3):
        t = (comp == k).astype(np.float64)
        assert np.abs(Ad.spmv(t)).max() <= tol * scale
    for a, b in ((0, 1), (1, 2), (2, 0)):
        rot = np.where(comp == a, -xyz[:, b], 0.0) + np.where(comp == b, xyz[:, a], 0.0)
        assert np.abs(Ad.spmv(rot)).max() <= tol * scale


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_elasticity_48cube_quadratic_patch(precision):
    """48^3, E = 1, Dirichlet data u on the whole boundary, f = -div σ(u) for a
    quadratic u: the Q2 space holds u, so A u_I = b up to rounding (max-abs residual
    relative to max|A| max|u_I|: 1e-12 FP64, 2e-5 FP32)."""
    part = (48, 48, 48)
    announce("elasticity Q2 48^3 quadratic patch %s" % precision)
    nu = 0.3
    lam, mu = CO.lame(1.0, nu)
    f = -np.array([lam + mu, lam + mu, lam + 3 * mu])
    u = lambda x: np.stack([x[..., 0] * x[..., 1], x[..., 1] * x[..., 2],
                            x[..., 2] * x[..., 0] + x[..., 0] ** 2], axis=-1)
    op, V = build(part, 1.0, nu, f, ["boundary"], g=u, precision=precision)
    uI = u(V.free_dof_coords)[np.arange(V.nfree), V.free_dof_comp]
    b = op.get_vector()
    r = op.assem.ctx.spmv(uI) - b
    # relative to max|A| max|u_I|: FP32 entries carry ~1e-7 relative rounding each
    rel = np.max(np.abs(r)) / (np.max(np.abs(op.get_matrix().nzval)) * np.max(np.abs(uI)))
    assert rel <= (2e-5 if precision == "fp32" else 1e-12), rel


def test_elasticity_lattice_symmetric_and_switch(monkeypatch):
    """The lattice kernel's matrix is symmetric bit for bit (K̂ exactly symmetric, both
    entries of a pair sum the same cells in the same order); GRIDAP_HIP_Q2_LATTICE=0
    leaves the element-list path, which agrees to rounding."""
    part = (6, 5, 4)
    announce("elasticity Q2 lattice symmetry and switch %s" % (part,))
    E = np.exp(np.random.default_rng(6).normal(0, 0.5, int(np.prod(part))))
    op, V = build(part, E, precision="fp32")
    S = op.get_matrix().to_scipy()
    assert (S != S.T).nnz == 0
    monkeypatch.setenv("GRIDAP_HIP_Q2_LATTICE", "0")
    op2, _ = build(part, E, precision="fp64")
    assert op2.assem.ctx.pattern_info()["lattice"] == 0
    monkeypatch.delenv("GRIDAP_HIP_Q2_LATTICE")
    op3, _ = build(part, E, precision="fp64")
    assert relmax(op3.get_matrix().nzval, op2.get_matrix().nzval) <= 1e-12
    assert relmax(op3.get_vector(), op2.get_vector()) <= 1e-12


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_elasticity_templates_match_ranking(monkeypatch, precision):
    """Column templates (one per position class, checked node by node against rowval)
    and the per-node ranking kernel add the same products in the same order: nzval
    bit-identical, b within 1e-14 (the lifting sums in another order).  15x13x11
    cells: every node has a template; Dirichlet data on the clamped face."""
    part = (15, 13, 11)
    announce("elasticity Q2 %s templates vs ranking %s" % (part, precision))
    E = _config3_E(part, seed=5)
    g = lambda x: np.stack([0.1 * x[..., 1], -0.2 * x[..., 2], 0.3 + x[..., 1] * x[..., 2]],
                           axis=-1)
    op1, _ = build(part, E, precision=precision, g=g)
    info1 = op1.assem.ctx.pattern_info()
    assert info1["last_kernel"] == _lib.GH_KERNEL_Q2_LATTICE and info1["fallback_nodes"] == 0
    monkeypatch.setenv("GRIDAP_HIP_Q2_TEMPLATES", "0")
    op2, _ = build(part, E, precision=precision, g=g)
    assert op2.assem.ctx.pattern_info()["fallback_nodes"] == -1
    np.testing.assert_array_equal(op1.get_matrix().nzval, op2.get_matrix().nzval)
    assert relmax(op1.get_vector(), op2.get_vector()) <= 1e-14


def test_elasticity_templates_with_irregular_nodes(oracle, monkeypatch):
    """40 single free components turned Dirichlet at random (the other free dofs
    renumbered in order): the nodes around them match no template and go through the
    ranking kernel, the rest through the templates.  FP64 against the oracle's
    assembly of the same cell matrices with the same ids (1e-12), and bitwise against
    the ranking kernel alone."""
    part = (9, 8, 7)
    announce("elasticity Q2 %s templates with irregular nodes" % (part,))
    C = int(np.prod(part))
    E = _config3_E(part, seed=8)
    model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), part)
    reffe = G.ReferenceFE(G.lagrangian, G.VectorValue[3, G.Float64], 2)
    V = G.TestFESpace(model, reffe, dirichlet_tags=CLAMP_X0)
    cd = V.cell_dof_ids.astype(np.int64)
    rng = np.random.default_rng(9)
    S = rng.choice(np.arange(1, V.nfree + 1), size=40, replace=False)
    keep = np.ones(V.nfree + 1, bool)
    keep[0] = False
    keep[S] = False
    newid = np.zeros(V.nfree + 1, np.int64)
    newid[keep] = np.arange(1, keep.sum() + 1)
    newid[S] = -(V.ndirichlet + 1 + np.arange(S.size))
    cd2 = np.where(cd > 0, newid[np.maximum(cd, 0)], cd).astype(np.int32)
    nfree2, ndiri2 = int(keep.sum()), V.ndirichlet + S.size
    q = G.Quadrature(3, 4)
    shp, grad = V.reffe.tabulate(3, q.coords, simplex=False)
    ptrs = (np.arange(C + 1) * 81 + 1).astype(np.int32)
    out = {}
    for env in ("1", "0"):
        monkeypatch.setenv("GRIDAP_HIP_Q2_TEMPLATES", env)
        ctx = _lib.Context(0)
        ctx.mesh_cartesian(model.origin, model.sizes, model.partition)
        ctx.space_upload(_lib.GH_REFFE_QK, 2, 3, cd2.ravel(), ptrs, nfree2, ndiri2)
        ctx.symbolic()
        ctx.reffe_tabulation(q.weights, shp, grad)
        ctx.coef_upload(E, None, None)
        ctx.elasticity_params(0.3, (0.0, 0.0, -1.0))
        ctx.assemble(_lib.GH_FORM_ELASTICITY, _lib.GH_PREC_FP64, _lib.GH_PATH_AUTO)
        info = ctx.pattern_info()
        assert info["last_kernel"] == _lib.GH_KERNEL_Q2_LATTICE, info
        if env == "1":
            assert 0 < info["fallback_nodes"] < (2 * part[0] + 1) * (2 * part[1] + 1) * (2 * part[2] + 1)
        nz, b = ctx.values_get()
        out[env] = (ctx.pattern_get(), nz, b)
        ctx.close()
    np.testing.assert_array_equal(out["1"][1], out["0"][1])
    r = CO.elasticity_q2_box(oracle, part, E, 0.3, np.array([0.0, 0.0, -1.0]),
                             [entities(CLAMP_X0)])
    np.testing.assert_array_equal(r["cell_dofs"], V.cell_dof_ids)
    cp, rv, nz = oracle.assemble_csc(cd2, r["cellmats"], nfree2, nfree2)
    bo = oracle.assemble_vector(cd2, r["cellvecs"], nfree2)
    (colptr, rowval), nzd, bd = out["1"]
    np.testing.assert_array_equal(colptr, cp)
    np.testing.assert_array_equal(rowval, rv)
    assert relmax(nzd, nz) <= 1e-12
    assert relmax(bd, bo) <= 1e-12
