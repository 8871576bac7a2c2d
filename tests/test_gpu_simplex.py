"""Parity of the simplex (P1/P2 tetrahedra) assembly path with the FP64 oracle
(needs an MI355X).

Bar: CSC pattern bit-exact with the oracle's create_from_nz output on the same
cell_dof_ids; FP64 mode: nzval and b within 1e-12 relative (max-abs error / max-abs
value); FP32-contraction mode: relative Frobenius error <= 1e-6.  At 48^3 the check
is the size-independent polynomial identity A u_I = b (P2 holds quadratics).
"""
import sys

import numpy as np
import pytest

import gridap_hip as G
from gridap_hip import _lib
import conforming_oracle as CO

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    assert a.shape == b.shape
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300) if a.size else 0.0


def relfrob(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def build(part, order, seed=5, kappa=None, fsrc=None, tags=("boundary",), g=None,
          precision="fp64", jitter=0.2):
    cart = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), part)
    model = G.simplexify(cart)
    if jitter:
        model = model.jittered(seed, jitter)
    reffe = G.ReferenceFE(G.lagrangian, G.Float64, order)
    V = G.TestFESpace(model, reffe, dirichlet_tags=list(tags))
    U = G.TrialFESpace(V, g)
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 2 * order)
    k = 1.0 if kappa is None else G.CellField(kappa, Ω)
    f = 1.0 if fsrc is None else G.CellField(fsrc, Ω)
    a = lambda u, v: G.integrate(k * G.inner(G.gradient(v), G.gradient(u)), dΩ)
    l = lambda v: G.integrate(f * v, dΩ)
    assem = G.SparseMatrixAssembler(U, V, precision=precision)
    op = G.AffineFEOperator(a, l, U, V, assem)
    return op, V, model


def oracle_system(oracle, part, order, xyz, kappa=None, fsrc=None, ents=None, g=None):
    q = G.Quadrature(3, 2 * order, simplex=True)
    r = CO.tet_poisson(part, xyz, order, kappa=kappa, fsrc=fsrc, qx=q.coords, qw=q.weights,
                       dirichlet_entities=ents, g=g)
    cp, rv, nz = oracle.assemble_csc(r["cell_dofs"], r["cellmats"], r["nfree"], r["nfree"])
    b = oracle.assemble_vector(r["cell_dofs"], r["cellvecs"], r["nfree"])
    return r, cp, rv, nz, b


def check(op, V, r, cp, rv, nz, b, precision="fp64"):
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


@pytest.mark.parametrize("order,precision", [(2, "fp64"), (2, "fp32"), (1, "fp64")])
def test_tet_matches_oracle(oracle, order, precision):
    part = (4, 3, 3)
    announce("simplex P%d %s %s jittered" % (order, precision, part))
    rng = np.random.default_rng(order)
    C = 6 * int(np.prod(part))
    kappa = np.exp(rng.normal(0, 1, C))
    fsrc = rng.uniform(-1, 1, C)
    op, V, model = build(part, order, kappa=kappa, fsrc=fsrc, precision=precision)
    r, cp, rv, nz, b = oracle_system(oracle, part, order, model.node_coordinates(), kappa, fsrc)
    check(op, V, r, cp, rv, nz, b, precision)


def test_tet_partial_tags_and_lifting(oracle):
    part = (3, 4, 2)
    announce("simplex P2 partial tags + lifting %s" % (part,))
    g = lambda x: 1.0 + x[..., 0] - 2.0 * x[..., 1] * x[..., 2]
    op, V, model = build(part, 2, tags=("tag_21", "tag_01", "tag_25"), g=g)
    ents = {21, 1, 25}
    r, cp, rv, nz, b = oracle_system(oracle, part, 2, model.node_coordinates(), ents=ents,
                                     g=lambda p: 1.0 + p[0] - 2.0 * p[1] * p[2])
    check(op, V, r, cp, rv, nz, b)


def quadratic_case(part, seed):
    u = lambda x: (1 + x[..., 0] - 2 * x[..., 1] + 0.5 * x[..., 2] + x[..., 0] ** 2 +
                   x[..., 1] ** 2 + 2 * x[..., 1] * x[..., 2] - 0.5 * x[..., 2] ** 2)
    C = 6 * int(np.prod(part))
    op, V, model = build(part, 2, seed=seed, fsrc=np.full(C, -3.0), g=u)
    return op, V, u(V.free_dof_coords)


def test_tet_reproduces_quadratics_and_solves():
    announce("simplex P2 quadratic reproduction + PCG (6x5x4)")
    op, V, uI = quadratic_case((6, 5, 4), 2)
    A, b = op.get_matrix().to_scipy(), op.get_vector()
    assert np.max(np.abs(A @ uI - b)) <= 1e-12 * np.max(np.abs(b))
    x = G.solve(op).get_free_dof_values()
    assert np.max(np.abs(x - uI)) <= 1e-9


def test_tet_16cube_matches_oracle(oracle):
    part = (16, 16, 16)
    announce("simplex P2 16^3 vs oracle")
    rng = np.random.default_rng(16)
    C = 6 * int(np.prod(part))
    kappa = np.exp(rng.normal(0, 1, C))
    op, V, model = build(part, 2, seed=16, kappa=kappa)
    r, cp, rv, nz, b = oracle_system(oracle, part, 2, model.node_coordinates(), kappa)
    check(op, V, r, cp, rv, nz, b)


def test_tet_assembly_is_deterministic():
    announce("simplex P2 repeat bit-identical (8^3)")
    part = (8, 8, 8)
    rng = np.random.default_rng(8)
    kappa = np.exp(rng.normal(0, 1, 6 * 512))
    op, V, _ = build(part, 2, kappa=kappa)
    assem = op.assem
    nz1, b1 = assem.ctx.values_get()
    assem.ctx.assemble(_lib.GH_FORM_DIFFUSION, assem.precision, assem.path)
    nz2, b2 = assem.ctx.values_get()
    np.testing.assert_array_equal(nz1, nz2)
    np.testing.assert_array_equal(b1, b2)


def test_tet_48cube_polynomial_identity():
    announce("simplex P2 48^3 A u_I = b")
    op, V, uI = quadratic_case((48, 48, 48), 48)
    A = op.get_matrix()
    Ad = op.assem.ctx
    r = Ad.spmv(uI) - op.get_vector()
    assert np.max(np.abs(r)) <= 1e-11 * np.max(np.abs(op.get_vector()))
    # pattern: symmetric, sorted, diagonal present
    S = A.to_scipy()
    assert (S != S.T).nnz == 0 or abs(S - S.T).max() <= 1e-12 * abs(S).max()
    d = S.diagonal()
    assert np.all(d > 0)


def test_tet_column_sum_matches_list_sum(monkeypatch):
    """The column-range sum (16-bit lists, packed triangles) and the fallback sum over
    the symbolic phase's int32 lists add the same entries in the same order: nzval
    and b bit-identical."""
    announce("simplex P2 column sum vs list sum (10x9x8)")
    part = (10, 9, 8)
    rng = np.random.default_rng(21)
    C = 6 * int(np.prod(part))
    kappa = np.exp(rng.normal(0, 1, C))
    g = lambda x: 1.0 + x[..., 0] - x[..., 1] * x[..., 2]
    for precision in ("fp32", "fp64"):
        monkeypatch.setenv("GRIDAP_HIP_COMPACT_LISTS", "1")
        op1, _, _ = build(part, 2, kappa=kappa, tags=("tag_21", "tag_01"), g=g,
                          precision=precision)
        monkeypatch.setenv("GRIDAP_HIP_COMPACT_LISTS", "0")
        op2, _, _ = build(part, 2, kappa=kappa, tags=("tag_21", "tag_01"), g=g,
                          precision=precision)
        assert op1.assem.ctx.pattern_info()["compact"] == 1
        assert op2.assem.ctx.pattern_info()["compact"] == 0
        assert op1.assem.ctx.pattern_info()["last_kernel"] == _lib.GH_KERNEL_COLUMN_SUM
        assert op2.assem.ctx.pattern_info()["last_kernel"] == _lib.GH_KERNEL_LIST_SUM
        np.testing.assert_array_equal(op1.get_matrix().nzval, op2.get_matrix().nzval)
        np.testing.assert_array_equal(op1.get_vector(), op2.get_vector())


def _fan_mesh(m):
    """m ring vertices around a centre vertex, with an apex above and below: 2m
    tetrahedra share the centre (vertex 0)."""
    t = 2 * np.pi * np.arange(m) / m
    xyz = np.vstack([[0.0, 0.0, 0.0], np.c_[np.cos(t), np.sin(t), np.zeros(m)],
                     [0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])
    top, bot = m + 1, m + 2
    ring = 1 + np.arange(m)
    nxt = 1 + (np.arange(m) + 1) % m
    cells = np.vstack([np.c_[np.zeros(m, int), ring, nxt, np.full(m, top)],
                       np.c_[np.zeros(m, int), nxt, ring, np.full(m, bot)]])
    return xyz, cells


def _p1_reference(xyz, cells, kappa):
    """Dense P1 stiffness and load (f = 1): K_ab = κ|det DF|/6 ∇λ_a·∇λ_b, b_a = |det|/24."""
    n = xyz.shape[0]
    K, b = np.zeros((n, n)), np.zeros(n)
    for c, v in enumerate(cells):
        DF = (xyz[v[1:]] - xyz[v[0]]).T
        det = np.linalg.det(DF)
        g = np.linalg.inv(DF)                       # rows: ∇λ_1..3
        grads = np.vstack([-g.sum(axis=0), g])
        K[np.ix_(v, v)] += kappa[c] * abs(det) / 6.0 * grads @ grads.T
        b[v] += abs(det) / 24.0
    return K, b


def test_high_valence_column_gets_its_own_range(monkeypatch):
    """A vertex shared by 600 tetrahedra: its column alone holds 2400 contributions,
    more than a column range's budget, so it forms a range of its own.  P1, every dof
    free; the compact sum and the int32-list sum agree bitwise, and both match the
    dense P1 reference (FP64: 1e-12 relative)."""
    announce("simplex P1 fan mesh, 600 tets on one vertex")
    xyz, cells = _fan_mesh(300)
    C, n = cells.shape[0], xyz.shape[0]
    rng = np.random.default_rng(17)
    kappa = np.exp(rng.normal(0, 1, C))
    reffe = G.ReferenceFE(G.lagrangian, G.Float64, 1)
    q = G.Quadrature(3, 2, simplex=True)
    shp, grad = reffe.tabulate(3, q.coords, simplex=True)
    ptrs = (np.arange(C + 1) * 4 + 1).astype(np.int32)
    cd = (cells + 1).astype(np.int32)
    out = {}
    for env in ("1", "0"):
        monkeypatch.setenv("GRIDAP_HIP_COMPACT_LISTS", env)
        ctx = _lib.Context(0)
        ctx.mesh_upload(xyz, cd.ravel(), ptrs)
        ctx.space_upload(_lib.GH_REFFE_PK, 1, 1, cd.ravel(), ptrs, n, 0)
        ctx.reffe_tabulation(q.weights, shp, grad)
        ctx.symbolic()
        info = ctx.pattern_info()
        assert info["elemlist"] == 1 and info["compact"] == (env == "1"), info
        ctx.coef_upload(kappa, None, None)
        ctx.assemble(_lib.GH_FORM_DIFFUSION, _lib.GH_PREC_FP64, _lib.GH_PATH_AUTO)
        info = ctx.pattern_info()
        assert info["last_kernel"] == (_lib.GH_KERNEL_COLUMN_SUM if env == "1"
                                       else _lib.GH_KERNEL_LIST_SUM), info
        assert (info["nranges"] > 0) == (env == "1"), info
        nz, b = ctx.values_get()
        colptr, rowval = ctx.pattern_get()
        out[env] = (colptr, rowval, nz, b)
        ctx.close()
    for k in range(4):
        np.testing.assert_array_equal(out["1"][k], out["0"][k])
    colptr, rowval, nz, b = out["1"]
    A = np.zeros((n, n))
    for j in range(n):
        for p in range(colptr[j] - 1, colptr[j + 1] - 1):
            A[rowval[p] - 1, j] = nz[p]
    K, bref = _p1_reference(xyz, cells, kappa)
    assert colptr[1] - colptr[0] == n          # the centre couples to every vertex
    assert np.max(np.abs(A - K)) <= 1e-12 * np.max(np.abs(K))
    assert np.max(np.abs(b - bref)) <= 1e-12 * np.max(np.abs(bref))


def _entry_keys(colptr, rowval, N):
    """Key col·N + row (0-based) of every stored entry of a CSC pattern (1-based
    colptr/rowval), ascending because columns are stored in order, rows sorted."""
    return np.repeat(np.arange(N, dtype=np.int64), np.diff(colptr)) * N + (rowval - 1)


def test_tet_96cube_full_size():
    """The benchmarked P2 workload at full size (96^3 x 6 tets, jitter ±0.2h, seed 7),
    FP64 and FP32 contraction: the column-range cut (compact lists, ~1e5 ranges) runs;
    NNZ equals the closed count 230n^3 - 690n^2 + 732n - 273; the quadratic identity
    A u_I = b holds (relative to max|A| max|u_I|: 1e-12 FP64, 2e-5 FP32); the values
    are symmetric bit for bit on 200k sampled entries; every diagonal entry is
    positive; FP32 against FP64 within 1e-6 relative Frobenius (nzval and b)."""
    n = 96
    part = (n, n, n)
    u = lambda x: (1 + x[..., 0] - 2 * x[..., 1] + 0.5 * x[..., 2] + x[..., 0] ** 2 +
                   x[..., 1] ** 2 + 2 * x[..., 1] * x[..., 2] - 0.5 * x[..., 2] ** 2)
    C = 6 * n ** 3
    res = {}
    for precision in ("fp64", "fp32"):
        announce("simplex P2 96^3x6 full size %s" % precision)
        op, V, model = build(part, 2, seed=7, fsrc=np.full(C, -3.0), g=u, precision=precision)
        info = op.assem.ctx.pattern_info()
        assert info["compact"] == 1 and info["metric"] == 1 and info["nranges"] > 10000, info
        A = op.get_matrix()
        N = V.num_free_dofs()
        assert A.nnz() == 230 * n ** 3 - 690 * n ** 2 + 732 * n - 273 == 197200239
        assert N == (2 * n - 1) ** 3
        uI = u(V.free_dof_coords)
        b = op.get_vector()
        r = op.assem.ctx.spmv(uI) - b
        rel = np.max(np.abs(r)) / (np.max(np.abs(A.nzval)) * np.max(np.abs(uI)))
        assert rel <= (1e-12 if precision == "fp64" else 2e-5), rel
        key = _entry_keys(A.colptr, A.rowval, N)
        sample = np.random.default_rng(96).integers(0, A.nnz(), 200000)
        col, row = key[sample] // N, key[sample] % N
        mirror = np.searchsorted(key, row * N + col)
        assert np.all(key[mirror] == row * N + col)            # structurally symmetric
        np.testing.assert_array_equal(A.nzval[sample], A.nzval[mirror])
        dpos = np.searchsorted(key, np.arange(N, dtype=np.int64) * (N + 1))
        assert np.all(key[dpos] == np.arange(N, dtype=np.int64) * (N + 1))
        assert np.all(A.nzval[dpos] > 0)
        res[precision] = (A.nzval, b)
        del key
        op.assem.ctx.close()
    assert relfrob(res["fp32"][0], res["fp64"][0]) <= 1e-6
    assert relfrob(res["fp32"][1], res["fp64"][1]) <= 1e-6


def test_tet_model_from_json_assembles_identically(tmp_path):
    """A jittered P2 model written in Gridap's JSON format and read back (what the Julia
    side reads) assembles to the same pattern and the same bits on the device."""
    announce("simplex P2 model through Gridap JSON (8x7x6)")
    from gridap_hip import io as GIO
    part = (8, 7, 6)
    model = G.simplexify(G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), part)).jittered(3)
    path = tmp_path / "tets.json"
    GIO.to_json_file(model, str(path))
    m2 = GIO.DiscreteModelFromFile(str(path))
    rng = np.random.default_rng(8)
    kappa = np.exp(rng.normal(0, 1, m2.num_cells()))
    out = []
    for m in (model, m2):
        reffe = G.ReferenceFE(G.lagrangian, G.Float64, 2)
        V = G.TestFESpace(m, reffe, dirichlet_tags=["tag_21", "tag_01"])
        U = G.TrialFESpace(V, lambda x: 1.0 + x[..., 0] * x[..., 2])
        dΩ = G.Measure(G.Triangulation(m), 4)
        k = G.CellField(kappa, G.Triangulation(m))
        op = G.AffineFEOperator(
            lambda u, v: G.integrate(k * G.inner(G.gradient(v), G.gradient(u)), dΩ),
            lambda v: G.integrate(1.0 * v, dΩ), U, V)
        out.append((op.get_matrix(), op.get_vector()))
    (A1, b1), (A2, b2) = out
    np.testing.assert_array_equal(A1.colptr, A2.colptr)
    np.testing.assert_array_equal(A1.rowval, A2.rowval)
    np.testing.assert_array_equal(A1.nzval, A2.nzval)
    np.testing.assert_array_equal(b1, b2)
