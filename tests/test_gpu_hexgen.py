"""The general n-cube element-list path against the FP64 oracle (needs an MI355X):
scalar Q2 diffusion and vector Q1 / Q2 elasticity on Cartesian and on mapped
(CartesianDiscreteModel with a map) meshes.

Bar: CSC pattern bit-exact with the oracle's create_from_nz output on the oracle's own
conforming numbering; FP64 element arithmetic in both modes, nzval and b within 1e-12
relative (max-abs error / max-abs value) in FP64 and 1e-6 (relative Frobenius) with
FP32 element matrices.  Size-independent checks at larger sizes: the quadratic patch
identity A u_I = b on Cartesian cells, rigid motions in the kernel of the mapped
elasticity matrix, the mapped path with the identity map against the Cartesian
vector-Q2 path, symmetry bit for bit.
"""
import sys

import numpy as np
import pytest

import gridap_hip as G
from gridap_hip import _lib
import conforming_oracle as CO

pytestmark = pytest.mark.gpu

# the element-list sums: column ranges with full-column staging, or the int32 lists
SUMS = (_lib.GH_KERNEL_COLUMN_SUM, _lib.GH_KERNEL_LIST_SUM)


def announce(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300) if a.size else 0.0


def relfrob(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def warp(x):
    return x + 0.05 * np.prod(np.sin(np.pi * x), axis=-1)[..., None]


def entities_of(D, tags):
    out = set()
    for t in tags:
        out |= CO.boundary_entities(D) if t == "boundary" else {int(t.split("_")[1])}
    return out


def product(part, order, elasticity, tags, g, coef, src, nu=0.3, mapped=False,
            precision="fp64", identity_map=False):
    D = len(part)
    dom = tuple(v for _ in range(D) for v in (0.0, 1.0))
    mp = warp if mapped else ((lambda x: x) if identity_map else None)
    model = G.CartesianDiscreteModel(dom, part, map=mp)
    T = G.VectorValue[D, G.Float64] if elasticity else G.Float64
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, T, order), dirichlet_tags=list(tags))
    U = G.TrialFESpace(V, g)
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 2 * order)
    c = G.CellField(coef, Ω)
    if elasticity:
        σ = G.IsotropicStress(c, nu)
        a = lambda u, v: G.integrate(G.inner(G.ε(v), σ(G.ε(u))), dΩ)
        l = lambda v: G.integrate(G.dot(v, G.VectorValue(*src)), dΩ)
    else:
        f = G.CellField(src, Ω)
        a = lambda u, v: G.integrate(c * G.inner(G.gradient(v), G.gradient(u)), dΩ)
        l = lambda v: G.integrate(f * v, dΩ)
    op = G.AffineFEOperator(a, l, U, V, G.SparseMatrixAssembler(U, V, precision=precision))
    return op, V, model


def oracle_system(oracle, part, order, elasticity, tags, g, coef, src, nu=0.3, mapped=False):
    D = len(part)
    ncomp = D if elasticity else 1
    _, cd, nfree, ndiri, _ = CO.cartesian_numbering(part, order, ncomp, [entities_of(D, tags)])
    dom = tuple(v for _ in range(D) for v in (0.0, 1.0))
    m = oracle.cartesian_model(dom, part)
    xyz = warp(m["xyz"]) if mapped else m["xyz"]
    X = xyz[m["cell_nodes"] - 1]
    C = X.shape[0]
    qx, qw = oracle.tensor_quadrature(D, 2 * order)
    shp, grad = oracle.qk_tabulate(D, order, qx)
    _, g1 = oracle.qk_tabulate(D, 1, qx)
    wdet, gp = CO.physical_gradients(X, qw, g1, grad)
    if elasticity:
        K = CO.ncube_elasticity_cells(wdet, gp, nu, coef)
        F = CO.ncube_loads(wdet, shp, np.tile(np.asarray(src, float), (C, 1)))
    else:
        K = CO.ncube_diffusion_cells(wdet, gp, coef)
        F = CO.ncube_loads(wdet, shp, np.asarray(src, float)[:, None])
    n1, _ = oracle.qk_tabulate(D, 1, oracle.qk_nodes(D, order))
    P = CO.q1_map_points(X, n1)
    nn = P.shape[1]
    dv = np.zeros(ndiri)
    if g is not None:
        for c in range(C):
            for a in range(ncomp * nn):
                if cd[c, a] < 0:
                    val = np.asarray(g(P[c, a % nn][None]), float).reshape(-1)
                    dv[-cd[c, a] - 1] = val[a // nn] if elasticity else val[0]
    F = oracle.attach_dirichlet(cd, K, dv, F)
    cp, rv, nz = oracle.assemble_csc(cd, K, nfree, nfree)
    b = oracle.assemble_vector(cd, F, nfree)
    return cd, cp, rv, nz, b


def check(op, V, ref, precision):
    cd, cp, rv, nz, b = ref
    np.testing.assert_array_equal(V.cell_dof_ids, cd)
    A = op.get_matrix()
    np.testing.assert_array_equal(A.colptr, cp)
    np.testing.assert_array_equal(A.rowval, rv)
    assert op.assem.ctx.pattern_info()["last_kernel"] in SUMS
    if precision == "fp64":
        assert relmax(A.nzval, nz) <= 1e-12
        assert relmax(op.get_vector(), b) <= 1e-12
    else:
        assert relfrob(A.nzval, nz) <= 1e-6
        assert relfrob(op.get_vector(), b) <= 1e-6


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("mapped", [False, True])
def test_scalar_q2_diffusion_matches_oracle(oracle, mapped, precision):
    part = (5, 4, 3)
    announce("hexgen scalar Q2 %s mapped=%s %s" % (part, mapped, precision))
    C = int(np.prod(part))
    rng = np.random.default_rng(3)
    kappa, fsrc = np.exp(rng.normal(size=C)), rng.normal(size=C)
    g = lambda x: 1.0 + x[..., 0] - 2.0 * x[..., 1] * x[..., 2]
    tags = ("tag_21", "tag_23", "tag_25")
    op, V, _ = product(part, 2, False, tags, g, kappa, fsrc, mapped=mapped, precision=precision)
    check(op, V, oracle_system(oracle, part, 2, False, tags, g, kappa, fsrc, mapped=mapped),
          precision)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("D,mapped", [(2, False), (2, True), (3, False), (3, True)])
def test_vector_q1_elasticity_matches_oracle(oracle, D, mapped, precision):
    part = (9, 7) if D == 2 else (6, 5, 4)
    announce("hexgen vector Q1 elasticity %s mapped=%s %s" % (part, mapped, precision))
    C = int(np.prod(part))
    E = np.exp(np.random.default_rng(D).normal(size=C))
    body = (0.3, -1.0) if D == 2 else (0.3, 0.1, -1.0)
    g = lambda x: np.stack([0.01 * x[..., 1], -0.02 * x[..., 0]] +
                           ([0.005 * x[..., 0] * x[..., 1]] if D == 3 else []), axis=-1)
    tags = ("tag_7",) if D == 2 else ("tag_21",)
    op, V, _ = product(part, 1, True, tags, g, E, body, nu=0.28, mapped=mapped,
                       precision=precision)
    check(op, V, oracle_system(oracle, part, 1, True, tags, g, E, body, nu=0.28,
                               mapped=mapped), precision)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mapped_vector_q2_elasticity_matches_oracle(oracle, precision):
    part = (3, 3, 2)
    announce("hexgen mapped vector Q2 elasticity %s %s" % (part, precision))
    C = int(np.prod(part))
    E = np.exp(np.random.default_rng(11).normal(size=C))
    body = (0.0, 0.2, -1.0)
    g = lambda x: np.stack([0.01 * x[..., 2], 0.0 * x[..., 0], -0.01 * x[..., 0]], axis=-1)
    tags = ("tag_21", "tag_25")
    op, V, _ = product(part, 2, True, tags, g, E, body, mapped=True, precision=precision)
    check(op, V, oracle_system(oracle, part, 2, True, tags, g, E, body, mapped=True),
          precision)


def test_identity_map_matches_cartesian_vector_q2():
    """Vector Q2 elasticity with the identity map (nodes uploaded, general element
    lists) against the Cartesian vector-Q2 path at 10x9x8: same pattern, values within
    1e-12 (FP64)."""
    part = (10, 9, 8)
    announce("hexgen identity-mapped vs Cartesian vector Q2 %s" % (part,))
    C = int(np.prod(part))
    E = np.exp(np.random.default_rng(5).normal(size=C))
    body = (0.0, 0.0, -1.0)
    g = lambda x: np.stack([0.01 * x[..., 1], 0.0 * x[..., 0], 0.0 * x[..., 0]], axis=-1)
    tags = ("tag_21",)
    op1, V1, _ = product(part, 2, True, tags, g, E, body, identity_map=True)
    op2, V2, _ = product(part, 2, True, tags, g, E, body)
    assert op1.assem.ctx.pattern_info()["last_kernel"] in SUMS
    assert op2.assem.ctx.pattern_info()["last_kernel"] not in SUMS
    A1, A2 = op1.get_matrix(), op2.get_matrix()
    np.testing.assert_array_equal(A1.colptr, A2.colptr)
    np.testing.assert_array_equal(A1.rowval, A2.rowval)
    assert relmax(A1.nzval, A2.nzval) <= 1e-12
    assert relmax(op1.get_vector(), op2.get_vector()) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_scalar_q2_quadratic_patch_24cube(precision):
    """-Δu = f with u quadratic (in the Q2 space on affine cells, and the degree-4 rule
    integrates the products exactly): A u_I = b at 24^3 (14 M nonzeros), relative to
    max|A| max|u_I|: 1e-12 FP64, 2e-5 FP32; A symmetric bit for bit on sampled entries."""
    part = (24, 24, 24)
    announce("hexgen scalar Q2 quadratic patch %s %s" % (part, precision))
    u = lambda x: 1 + x[..., 0] - 2 * x[..., 1] + x[..., 0] ** 2 + 0.5 * x[..., 1] * x[..., 2]
    C = int(np.prod(part))
    op, V, _ = product(part, 2, False, ("boundary",), u, np.ones(C), np.full(C, -2.0),
                       precision=precision)
    A = op.get_matrix()
    uI = u(V.free_dof_coords)
    r = A.to_scipy() @ uI - op.get_vector()
    rel = np.max(np.abs(r)) / (np.max(np.abs(A.nzval)) * np.max(np.abs(uI)))
    assert rel <= (1e-12 if precision == "fp64" else 2e-5), rel
    S = A.to_scipy()
    idx = np.random.default_rng(1).integers(0, V.nfree, 2000)
    sub = S[idx][:, idx].toarray()
    np.testing.assert_array_equal(sub, sub.T)


@pytest.mark.parametrize("order", [1, 2])
def test_mapped_elasticity_rigid_motions(order):
    """No Dirichlet dofs: A is symmetric bit for bit, and the six rigid motions,
    interpolated at the dof nodes of the mapped mesh, lie in its kernel (1e-11 relative
    to max|A| max|r|), FP64."""
    part = (16, 14, 12) if order == 1 else (8, 7, 6)
    announce("hexgen mapped rigid motions order %d %s" % (order, part))
    C = int(np.prod(part))
    E = np.exp(np.random.default_rng(order).normal(size=C))
    op, V, model = product(part, order, True, (), None, E, (0.0, 0.0, -1.0), mapped=True)
    A = op.get_matrix().to_scipy()
    assert (A != A.T).nnz == 0          # symmetric bit for bit
    if order == 1:
        x = model.node_coordinates()[V.free_dof_to_node - 1]
        comp = V.free_dof_to_comp - 1
    else:
        x, comp = V.free_dof_coords, V.free_dof_comp
    modes = []
    for i in range(3):
        modes.append((comp == i).astype(float))
    for i in range(3):
        for j in range(i + 1, 3):
            r = np.zeros(V.nfree)
            r[comp == i] = -x[comp == i, j]
            r[comp == j] = x[comp == j, i]
            modes.append(r)
    amax = np.max(np.abs(A.data))
    for r in modes:
        assert np.max(np.abs(A @ r)) <= 1e-11 * amax * np.max(np.abs(r))


def _unstructured_q1(oracle, part, seed):
    """A Cartesian Q1 mesh with its nodes renumbered at random and moved (no lattice
    topology left): coordinates, 1-based cell nodes, CLagrangian cell dofs of the
    boundary-Dirichlet space in the new node order."""
    D = len(part)
    m = oracle.cartesian_model((0.0, 1.0) * D, part)
    rng = np.random.default_rng(seed)
    nn = m["xyz"].shape[0]
    perm = rng.permutation(nn)                 # new id of old node k: perm[k]
    xyz = np.empty_like(m["xyz"])
    xyz[perm] = warp(m["xyz"])
    cn = (perm[m["cell_nodes"] - 1] + 1).astype(np.int32)
    lab = np.empty_like(m["vertex_labels"])
    lab[perm] = m["vertex_labels"]
    ents = CO.boundary_entities(D)
    node_to_tag = oracle.face_tag_index(lab, [sorted(ents)])
    n2d, nfree, ndiri = oracle.clagrangian(node_to_tag)
    cd = oracle.clagrangian_cell_dofs(cn, n2d)
    return xyz, cn, cd, nfree, ndiri, n2d


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("part", [(17, 13), (9, 8, 7)])
def test_unstructured_q1_element_lists_match_oracle(oracle, part, precision):
    """Scalar Q1 on a mesh without the lattice topology (renumbered, moved nodes),
    through the C ABI: ∫ κ ∇v⋅∇u + σ u v with κ, σ, f per cell and Dirichlet data go to
    the general n-cube element lists (no coloured scatter); pattern bit-exact, values
    within 1e-12 (FP64) / 1e-6 (FP32) of the oracle."""
    D = len(part)
    announce("unstructured Q1 element lists %s %s" % (part, precision))
    xyz, cn, cd, nfree, ndiri, n2d = _unstructured_q1(oracle, part, 7)
    C = cn.shape[0]
    rng = np.random.default_rng(D)
    kappa, sigma, fsrc = np.exp(rng.normal(size=C)), np.exp(rng.normal(size=C)), rng.normal(size=C)
    dv = rng.normal(size=ndiri)
    qx, qw = oracle.tensor_quadrature(D, 2)
    shp, grad = oracle.qk_tabulate(D, 1, qx)
    ptrs = (np.arange(C + 1) * (1 << D) + 1).astype(np.int32)
    ctx = _lib.Context(0)
    ctx.mesh_upload(xyz, cn.ravel(), ptrs)
    ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, np.ascontiguousarray(cd, np.int32).ravel(), ptrs,
                     nfree, ndiri)
    ctx.reffe_tabulation(qw, shp, grad.reshape(len(qw), -1))
    ctx.symbolic()
    ctx.coef_upload(kappa, fsrc, dv)
    ctx.mass_upload(sigma)
    prec = _lib.GH_PREC_FP64 if precision == "fp64" else _lib.GH_PREC_FP32_ELEM
    ctx.assemble(_lib.GH_FORM_DIFFUSION, prec, _lib.GH_PATH_AUTO)
    info = ctx.pattern_info()
    assert info["lattice"] == 0 and info["last_kernel"] in SUMS, info
    nz, b = ctx.values_get()
    colptr, rowval = ctx.pattern_get()
    ctx.close()
    X = xyz[cn - 1]
    wdet, gp = CO.physical_gradients(X, qw, grad, grad)
    mats = CO.ncube_diffusion_cells(wdet, gp, kappa) + \
        sigma[:, None, None] * np.einsum("cq,qa,qb->cab", wdet, shp, shp)
    Kc = oracle.cell_matrices_diffusion(D, cn, xyz, qw, grad, kappa=kappa)
    assert np.abs(CO.ncube_diffusion_cells(wdet, gp, kappa) - Kc).max() <= 1e-12 * np.abs(Kc).max()
    vecs = CO.ncube_loads(wdet, shp, fsrc[:, None])
    vecs = oracle.attach_dirichlet(cd, mats, dv, vecs)
    cp, rv, nzr = oracle.assemble_csc(cd, mats, nfree, nfree)
    br = oracle.assemble_vector(cd, vecs, nfree)
    np.testing.assert_array_equal(colptr, cp)
    np.testing.assert_array_equal(rowval, rv)
    if precision == "fp64":
        assert relmax(nz, nzr) <= 1e-12 and relmax(b, br) <= 1e-12
    else:
        assert relfrob(nz, nzr) <= 1e-6 and relfrob(b, br) <= 1e-6


@pytest.mark.parametrize("case", ["mapped_vector_q2", "mapped_vector_q2_wide"])
def test_column_ranges_match_list_sum(monkeypatch, case):
    """Vector Q2 (81 entries per element column, the full-column form of the range sum;
    smaller elements keep the lists): the column-range sum and the int32-list sum
    (GRIDAP_HIP_COMPACT_LISTS=0) add the same element entries in the same order: nzval
    and b bit-identical, FP32 and FP64."""
    announce("hexgen column ranges vs list sum %s" % case)
    part, order, el, tags, mapped = {
        "mapped_vector_q2": ((4, 3, 3), 2, True, ("tag_21",), True),
        "mapped_vector_q2_wide": ((9, 7, 6), 2, True, ("tag_21", "tag_26"), True)}[case]
    D = len(part)
    C = int(np.prod(part))
    coef = np.exp(np.random.default_rng(C).normal(size=C))
    src = (0.1, -1.0, 0.3)[:D] if el else np.random.default_rng(1).normal(size=C)
    g = (lambda x: np.stack([0.01 * x[..., -1]] * D, axis=-1)) if el else \
        (lambda x: 1.0 + x[..., 0] * x[..., 1])
    for precision in ("fp64", "fp32"):
        out = {}
        for env in ("1", "0"):
            monkeypatch.setenv("GRIDAP_HIP_COMPACT_LISTS", env)
            op, V, _ = product(part, order, el, tags, g, coef, src, mapped=mapped,
                               precision=precision)
            info = op.assem.ctx.pattern_info()
            assert info["compact"] == (env == "1") and info["inverse"] == (env == "1"), info
            assert info["last_kernel"] == (_lib.GH_KERNEL_COLUMN_SUM if env == "1"
                                           else _lib.GH_KERNEL_LIST_SUM), info
            out[env] = (op.get_matrix().nzval, op.get_vector())
            op.assem.ctx.close()
        np.testing.assert_array_equal(out["1"][0], out["0"][0])
        np.testing.assert_array_equal(out["1"][1], out["0"][1])
