"""Pins the CPU oracle to Gridap's own fixtures, closed forms and invariants.

Each test names the reference test or mathematical fact it checks; together they
make a plausible mistake in the restatement (a wrong node/face order, a dropped
Jacobian factor, a transposed element matrix, a mis-sorted CSC insert) fail.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla


# ---------------------------------------------------------------- quadrature

@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("degree", range(1, 11))
def test_tensor_gauss_exactness(oracle, D, degree):
    """test/ReferenceFEsTests/TensorProductQuadraturesTests.jl:29-38: every monomial
    of the degree's basis integrates to 1/prod(e+1) with error < 1e-15."""
    x, w = oracle.tensor_quadrature(D, degree)
    exps = np.indices((degree + 1,) * D).reshape(D, -1).T
    errs, exact = [], []
    for e in exps:
        q = np.sum(w * np.prod(x ** e, axis=1))
        ex = 1.0 / np.prod(e + 1.0)
        errs.append(abs(q - ex))
        exact.append(ex)
    assert max(errs) / max(exact) < 1e-15


def test_gauss_closed_forms(oracle):
    x, w = oracle.gauss_legendre01(2)
    np.testing.assert_allclose(x, [(1 - 3 ** -0.5) / 2, (1 + 3 ** -0.5) / 2], rtol=0, atol=1e-16)
    np.testing.assert_allclose(w, [0.5, 0.5], rtol=0, atol=1e-16)
    x, w = oracle.gauss_legendre01(3)
    s = np.sqrt(3 / 5)
    np.testing.assert_allclose(x, [(1 - s) / 2, 0.5, (1 + s) / 2], rtol=0, atol=1e-16)
    np.testing.assert_allclose(w, [5 / 18, 8 / 18, 5 / 18], rtol=0, atol=1e-16)


def test_gauss_matches_golub_welsch(oracle):
    """Independent library routine: numpy's leggauss (Golub-Welsch, itself good to a
    few ulps; the exactness test above is the sharp check)."""
    for n in range(1, 12):
        x, w = oracle.gauss_legendre01(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        np.testing.assert_allclose(x, (xr + 1) / 2, rtol=0, atol=2e-16)
        np.testing.assert_allclose(w, wr / 2, rtol=0, atol=2e-15)


# ---------------------------------------------------------- reference element

def test_q2_node_order(oracle):
    """Face-ownership order (vertices, edges, faces, interior; x fastest), as the
    QUAD Q2 fixtures of test/ReferenceFEsTests/CLagrangianRefFEsTests.jl:124-131
    imply (edge own nodes 5..8)."""
    x = oracle.qk_nodes(2, 2)
    np.testing.assert_array_equal(x, [[0, 0], [1, 0], [0, 1], [1, 1], [0.5, 0], [0.5, 1],
                                      [0, 0.5], [1, 0.5], [0.5, 0.5]])
    x3 = oracle.qk_nodes(3, 2)
    assert x3.shape == (27, 3)
    np.testing.assert_array_equal(x3[24], [0.0, 0.5, 0.5])   # face 25 (x=0) centre
    np.testing.assert_array_equal(x3[26], [0.5, 0.5, 0.5])


@pytest.mark.parametrize("D,order", [(2, 1), (3, 1), (2, 2), (3, 2)])
def test_shape_functions_nodal_and_partition_of_unity(oracle, D, order):
    nodes = oracle.qk_nodes(D, order)
    shp, grad = oracle.qk_tabulate(D, order, nodes)
    np.testing.assert_allclose(shp, np.eye(len(nodes)), atol=1e-14)
    pts = np.random.default_rng(0).random((7, D))
    shp, grad = oracle.qk_tabulate(D, order, pts)
    np.testing.assert_allclose(shp.sum(1), 1.0, atol=1e-14)
    np.testing.assert_allclose(grad.sum(1), 0.0, atol=1e-13)
    # reproduces linear functions: Σ x_a φ_a = x, Σ x_a ∇φ_a = I
    np.testing.assert_allclose(shp @ nodes, pts, atol=1e-14)
    np.testing.assert_allclose(np.einsum("pad,ae->ped", grad, nodes),
                               np.broadcast_to(np.eye(D), (7, D, D)), atol=1e-13)


def test_shape_gradient_by_finite_differences(oracle):
    pts = np.array([[0.3, 0.6, 0.2]])
    _, g = oracle.qk_tabulate(3, 2, pts)
    h = 1e-6
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        sp_, _ = oracle.qk_tabulate(3, 2, pts + e)
        sm, _ = oracle.qk_tabulate(3, 2, pts - e)
        np.testing.assert_allclose((sp_ - sm)[0] / (2 * h), g[0, :, d], atol=1e-8)


# ----------------------------------------------------------- faces and labels

def test_ncube_face_ids_hex_x0_face(oracle):
    """x=0 face of the unit cube = tags {1,3,5,7,13,15,17,19,25}
    (test/GridapTests/IsotropicDamageTests.jl:71-74)."""
    ids = set()
    # vertices with x=0
    for a in range(8):
        if not a & 1:
            ids.add(oracle.ncube_face_id(3, 0, a))
    # edges parallel to y or z lying on x=0
    for free in (2, 4):
        for anchor in range(8):
            if anchor & free or anchor & 1:
                continue
            ids.add(oracle.ncube_face_id(3, free, anchor))
    ids.add(oracle.ncube_face_id(3, 6, 0))
    assert ids == {1, 3, 5, 7, 13, 15, 17, 19, 25}
    assert oracle.ncube_face_id(3, 7, 0) == 27


def test_ncube_face_ids_quad(oracle):
    # vertices 1-4, edges y=0:5, y=1:6, x=0:7, x=1:8, interior 9
    assert [oracle.ncube_face_id(2, 0, a) for a in range(4)] == [1, 2, 3, 4]
    assert oracle.ncube_face_id(2, 1, 0) == 5
    assert oracle.ncube_face_id(2, 1, 2) == 6
    assert oracle.ncube_face_id(2, 2, 0) == 7
    assert oracle.ncube_face_id(2, 2, 1) == 8
    assert oracle.ncube_face_id(2, 3, 0) == 9


def _numbering(oracle, partition, tags, masks=None, ncomp=1):
    D = len(partition)
    m = oracle.cartesian_model((0, 1) * D, partition)
    ents = [oracle.cartesian_tag_entities(D, t) for t in tags]
    n2t = oracle.face_tag_index(m["vertex_labels"], ents)
    n2d, nf, nd = oracle.clagrangian(n2t, ncomp=ncomp, masks=masks)
    return m, n2d, nf, nd


def test_clagrangian_fixtures_scalar(oracle):
    """test/FESpacesTests/CLagrangianFESpacesTests.jl:58-72, 85-101."""
    tags = [1, 2, 4, 5, 8]
    m, n2d, nf, nd = _numbering(oracle, (2, 2), tags)
    assert n2d[:, 0].tolist() == [-1, -2, -3, 1, 2, -4, 3, 4, -5]
    m, n2d, nf, nd = _numbering(oracle, (2, 2), tags, masks=[1, 1, 0, 1, 1])
    assert n2d[:, 0].tolist() == [-1, -2, -3, 1, 2, -4, 3, 4, 5]
    cd = oracle.clagrangian_cell_dofs(m["cell_nodes"], n2d)
    assert cd.tolist() == [[-1, -2, 1, 2], [-2, -3, 2, -4], [1, 2, 3, 4], [2, -4, 4, 5]]
    m, n2d, nf, nd = _numbering(oracle, (2, 2), [])
    assert n2d[:, 0].tolist() == list(range(1, 10))
    assert m["cell_nodes"].tolist() == [[1, 2, 4, 5], [2, 3, 5, 6], [4, 5, 7, 8], [5, 6, 8, 9]]


def test_clagrangian_fixtures_vector(oracle):
    """test/FESpacesTests/CLagrangianFESpacesTests.jl:31-43, 74-80, 103-110."""
    m, n2d, nf, nd = _numbering(oracle, (2, 2), [], ncomp=2)
    cd = oracle.clagrangian_cell_dofs(m["cell_nodes"], n2d)
    assert cd.tolist() == [[1, 3, 7, 9, 2, 4, 8, 10], [3, 5, 9, 11, 4, 6, 10, 12],
                           [7, 9, 13, 15, 8, 10, 14, 16], [9, 11, 15, 17, 10, 12, 16, 18]]
    tags = [1, 2, 4, 5, 8]
    m, n2d, nf, nd = _numbering(oracle, (2, 2), tags, ncomp=2)
    assert [tuple(r) for r in n2d] == [(-1, -2), (-3, -4), (-5, -6), (1, 2), (3, 4), (-7, -8),
                                       (5, 6), (7, 8), (-9, -10)]
    masks2 = [(1, 1), (1, 0), (0, 0), (0, 1), (1, 1)]
    m, n2d, nf, nd = _numbering(oracle, (2, 2), tags, masks=masks2, ncomp=2)
    assert [tuple(r) for r in n2d] == [(-1, -2), (1, -3), (-4, 2), (3, 4), (5, 6), (-5, -6),
                                       (7, 8), (9, 10), (11, 12)]
    cd = oracle.clagrangian_cell_dofs(m["cell_nodes"], n2d)
    assert cd.tolist() == [[-1, 1, 3, 5, -2, -3, 4, 6], [1, -4, 5, -5, -3, 2, 6, -6],
                           [3, 5, 7, 9, 4, 6, 8, 10], [5, -5, 9, 11, 6, -6, 10, 12]]


def test_boundary_tag_closed_form(oracle):
    """"boundary" = every entity but the interior: the Dirichlet nodes are exactly
    the nodes with an index 0 or n on some axis."""
    for part in [(3, 4), (2, 3, 4)]:
        D = len(part)
        m, n2d, nf, nd = _numbering(oracle, part, ["boundary"])
        idx = np.indices([n + 1 for n in part]).reshape(D, -1, order="F")
        onb = np.zeros(idx.shape[1], bool)
        for d in range(D):
            onb |= (idx[d] == 0) | (idx[d] == part[d])
        assert np.array_equal(n2d[:, 0] < 0, onb)
        assert nf == np.prod([n - 1 for n in part])


# --------------------------------------------------------------- CSC builder

def test_builder_protocol_fixture(oracle):
    """test/AlgebraTests/AlgebraInterfacesTests.jl:106-138 (MinMemory CSC, 6x9)."""
    entries = [(1, 1, 1.0), (1, 1, None), (3, 1, 4.0), (3, 1, 2.0), (2, 1, 8.0), (4, 9, 3.0)]
    cnt, colptr, rowval, nzval = oracle.csc_from_entries(entries, 9)
    assert cnt.tolist() == [5, 0, 0, 0, 0, 0, 0, 0, 1]
    # the fixture's counter saw (1,1),(1,1),(3,1),(2,1),(4,9): drive exactly that
    cnt2, _, _, _ = oracle.csc_from_entries(
        [(1, 1, 1.0), (1, 1, None), (3, 1, None), (2, 1, None), (4, 9, 3.0)], 9)
    assert cnt2.tolist() == [4, 0, 0, 0, 0, 0, 0, 0, 1]
    assert rowval.tolist() == [1, 2, 3, 4]
    J = np.repeat(np.arange(1, 10), np.diff(colptr))
    assert J.tolist() == [1, 1, 1, 9]
    assert nzval.tolist() == [1.0, 8.0, 6.0, 3.0]


def test_builder_keeps_explicit_zeros_and_sorts(oracle):
    rng = np.random.default_rng(3)
    entries = [(int(i), int(j), float(v)) for i, j, v in
               zip(rng.integers(1, 20, 300), rng.integers(1, 15, 300), rng.normal(size=300))]
    entries += [(5, 5, 0.0), (7, 3, None)]
    cnt, colptr, rowval, nzval = oracle.csc_from_entries(entries, 15)
    dense = np.zeros((20, 15))
    touched = np.zeros((20, 15), bool)
    for i, j, v in entries:
        touched[i - 1, j - 1] = True
        if v is not None:
            dense[i - 1, j - 1] += v
    A = sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(20, 15))
    assert A.has_sorted_indices
    np.testing.assert_allclose(A.toarray(), dense, atol=1e-13)
    mask = np.zeros((20, 15), bool)
    mask[rowval - 1, np.repeat(np.arange(15), np.diff(colptr))] = True
    assert np.array_equal(mask, touched)


# ------------------------------------------------------ assembled systems

def test_sparse_matrix_assembler_known_values(oracle):
    """test/FESpacesTests/SparseMatrixAssemblersTests.jl:15-28,107-122: 2x2 QUAD Q1,
    dirichlet_tags=[1,2,3,4,6,5], ∫∇v⊙∇u and ∫v⊙b with b(x)=x[2], degree 2."""
    r = oracle.poisson_q1((0, 1, 0, 1), (2, 2), dirichlet_tags=[1, 2, 3, 4, 6, 5])
    A = oracle.to_scipy(r["colptr"], r["rowval"], r["nzval"], r["nfree"]).toarray()
    assert r["nfree"] == 3
    np.testing.assert_allclose(A[0, 0], 1.333333333333333, rtol=1e-14)
    np.testing.assert_allclose(A[1, 0], -0.33333333333333, rtol=1e-12)
    np.testing.assert_allclose(A[0, 1], -0.33333333333333, rtol=1e-12)
    np.testing.assert_allclose(A[1, 1], 2.666666666666666, rtol=1e-14)
    np.testing.assert_allclose(A[2, 1], -0.33333333333333, rtol=1e-12)
    np.testing.assert_allclose(A[1, 2], -0.33333333333333, rtol=1e-12)
    np.testing.assert_allclose(A[2, 2], 1.333333333333333, rtol=1e-14)
    m = r["model"]
    xq = oracle.physical_quadrature_points(2, m["cell_nodes"], m["xyz"], r["shp"])
    vecs = oracle.cell_vectors_source_qp(2, m["cell_nodes"], m["xyz"], r["qw"], r["shp"],
                                         r["grad"], xq[:, :, 1])
    b = oracle.assemble_vector(r["cell_dofs"], vecs, r["nfree"])
    np.testing.assert_allclose(b, [0.0625, 0.125, 0.0625], rtol=1e-14)


def test_affine_operator_reproduces_linear_solution(oracle):
    """test/FESpacesTests/AffineFEOperatorsTests.jl:52-86: u=x+y, f=0, Dirichlet on
    "boundary", degree 4: the FE solution equals u at the nodes."""
    u = lambda x: x[0] + x[1]
    r = oracle.poisson_q1((0, 1, 0, 1), (2, 2), fsrc=np.zeros(4), g=u, degree=4)
    A = oracle.to_scipy(r["colptr"], r["rowval"], r["nzval"], r["nfree"])
    x = spla.spsolve(A.tocsc(), r["b"])
    free_nodes = np.nonzero(r["node_to_dof"][:, 0] > 0)[0]
    np.testing.assert_allclose(x, [u(r["model"]["xyz"][n]) for n in free_nodes], atol=1e-14)
    # and on a larger, jittered-free 3-D case with the isoparametric map
    u3 = lambda x: 1 + 2 * x[0] - x[1] + 0.5 * x[2]
    r = oracle.poisson_q1((0, 1, 0, 2, 0, 1), (3, 4, 2), fsrc=np.zeros(24), g=u3,
                          mesh="unstructured")
    A = oracle.to_scipy(r["colptr"], r["rowval"], r["nzval"], r["nfree"])
    x = spla.spsolve(A.tocsc(), r["b"])
    free_nodes = np.nonzero(r["node_to_dof"][:, 0] > 0)[0]
    np.testing.assert_allclose(x, [u3(r["model"]["xyz"][n]) for n in free_nodes], atol=1e-13)


@pytest.mark.parametrize("part", [(5, 7), (4, 3, 5)])
def test_closed_form_nnz(oracle, part):
    """Cartesian Q1 + boundary Dirichlet: NNZ = Π(3 m_d - 2), m_d = n_d - 1."""
    r = oracle.poisson_q1((0, 1) * len(part), part)
    assert len(r["rowval"]) == np.prod([3 * (n - 1) - 2 for n in part])
    assert r["nfree"] == np.prod([n - 1 for n in part])


def test_pattern_brute_force(oracle):
    """Dense assembly of every touched (i,j): the CSC pattern is exactly that set
    (explicit zeros kept), values equal the dense sums."""
    rng = np.random.default_rng(1)
    kappa = np.exp(rng.normal(size=4 * 3 * 3))
    r = oracle.poisson_q1((0, 1, 0, 1, 0, 1), (4, 3, 3), kappa=kappa,
                          dirichlet_tags=["tag_25", "tag_21"])
    n = r["nfree"]
    dense = np.zeros((n, n))
    touched = np.zeros((n, n), bool)
    for c, dofs in enumerate(r["cell_dofs"]):
        K = r["cellmats"][c].T  # stored [lj, li]
        for li, i in enumerate(dofs):
            for lj, j in enumerate(dofs):
                if i > 0 and j > 0:
                    dense[i - 1, j - 1] += K[li, lj]
                    touched[i - 1, j - 1] = True
    A = oracle.to_scipy(r["colptr"], r["rowval"], r["nzval"], n)
    mask = np.zeros((n, n), bool)
    mask[r["rowval"] - 1, np.repeat(np.arange(n), np.diff(r["colptr"]))] = True
    assert np.array_equal(mask, touched)
    np.testing.assert_allclose(A.toarray(), dense, rtol=0, atol=1e-13)


def test_stencil_closed_forms(oracle):
    """Uniform unit-square grid, κ=1: the Q1 stiffness row of an interior node is
    8/3 on the diagonal and -1/3 for the 8 neighbours; in 3-D (unit cube, h) it is
    8h/3 and -h/6 (face-diagonal), -h/12 (body-diagonal), 0 (edge neighbour)."""
    r = oracle.poisson_q1((0, 1, 0, 1), (6, 6))
    A = oracle.to_scipy(r["colptr"], r["rowval"], r["nzval"], r["nfree"]).toarray()
    row = A[12]  # an interior free dof far from the boundary (node (3,3))
    assert np.isclose(row[12], 8 / 3, rtol=1e-14)
    nz = sorted(np.round(row[np.abs(row) > 1e-15], 14))
    assert nz == sorted([round(8 / 3, 14)] + [round(-1 / 3, 14)] * 8)
    h = 1 / 6
    r = oracle.poisson_q1((0, 1, 0, 1, 0, 1), (6, 6, 6))
    A = oracle.to_scipy(r["colptr"], r["rowval"], r["nzval"], r["nfree"]).toarray()
    k = 2 + 5 * (2 + 5 * 2)  # free dof of node (3,3,3)
    row = A[k]
    assert np.isclose(row[k], 8 * h / 3, rtol=1e-14)
    vals = sorted(np.round(row[np.abs(row) > 1e-15] / h, 12))
    expect = sorted([round(8 / 3, 12)] + [round(-1 / 6, 12)] * 12 + [round(-1 / 12, 12)] * 8)
    assert vals == expect


def test_invariants_neumann(oracle):
    """No Dirichlet tags: symmetric, zero row sums (constants in the kernel), and
    Σ b = ∫ f dΩ = |Ω| for f = 1; with per-cell κ and the isoparametric map on a
    mapped grid too."""
    rng = np.random.default_rng(2)
    part = (3, 4, 2)
    kappa = np.exp(rng.normal(size=24))
    for mesh in ("cartesian", "unstructured"):
        r = oracle.poisson_q1((0, 2, 0, 1, 0, 1), part, kappa=kappa, dirichlet_tags=[],
                              mesh=mesh)
        A = oracle.to_scipy(r["colptr"], r["rowval"], r["nzval"], r["nfree"])
        assert abs(A - A.T).max() < 1e-14
        np.testing.assert_allclose(A @ np.ones(r["nfree"]), 0.0, atol=1e-14)
        assert np.isclose(r["b"].sum(), 2.0, rtol=1e-14)


def test_cartesian_cell_matrix_equals_kappa_times_reference(oracle):
    """K_c = κ_c K̂ on a Cartesian grid: the CartesianMap path and the
    isoparametric path agree, and scale linearly in κ."""
    rng = np.random.default_rng(5)
    kappa = np.exp(rng.normal(size=8))
    r1 = oracle.poisson_q1((0, 1, 0, 1, 0, 1), (2, 2, 2), kappa=kappa, mesh="cartesian")
    r2 = oracle.poisson_q1((0, 1, 0, 1, 0, 1), (2, 2, 2), kappa=kappa, mesh="unstructured")
    np.testing.assert_allclose(r1["cellmats"], r2["cellmats"], rtol=1e-13, atol=1e-15)
    K0 = r1["cellmats"][0] / kappa[0]
    for c in range(8):
        scale = np.abs(r1["cellmats"][c]).max()
        np.testing.assert_allclose(r1["cellmats"][c], kappa[c] * K0, rtol=1e-14,
                                   atol=1e-15 * scale)


def test_pcg_oracle_matches_direct_solve(oracle):
    rng = np.random.default_rng(4)
    kappa = np.exp(rng.normal(size=6 * 5 * 4))
    r = oracle.poisson_q1((0, 1, 0, 1, 0, 1), (6, 5, 4), kappa=kappa)
    x, it, rr = oracle.pcg_jacobi(r["colptr"], r["rowval"], r["nzval"], r["b"], rtol=1e-12)
    A = oracle.to_scipy(r["colptr"], r["rowval"], r["nzval"], r["nfree"])
    xd = spla.spsolve(A.tocsc(), r["b"])
    assert rr <= 1e-12
    np.testing.assert_allclose(x, xd, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("part", [(7, 5), (5, 4, 3)])
def test_numeric_refill_sequential_and_coloured(oracle, part):
    """assemble_matrix_and_vector! into a fixed pattern (the numeric loop the CPU
    baseline times) reproduces the full assembly, and its parity-coloured OpenMP
    form agrees with it to rounding for any thread count."""
    D = len(part)
    rng = np.random.default_rng(11)
    kappa = np.exp(rng.normal(size=int(np.prod(part))))
    r = oracle.poisson_q1((0, 1) * D, part, kappa=kappa)
    colptr = np.asarray(r["colptr"], dtype=np.int64)
    rowval = np.asarray(r["rowval"], dtype=np.int64)
    sizes = np.full(D, 1.0) / np.asarray(part)
    args = (r["cell_dofs"], r["qw"], r["shp"], r["grad"], kappa, None, colptr, rowval)
    nz, b = oracle.numeric_refill_cartesian(D, sizes, *args)
    np.testing.assert_allclose(nz, r["nzval"], rtol=1e-14, atol=1e-15 * np.abs(nz).max())
    np.testing.assert_allclose(b, r["b"], rtol=1e-14, atol=1e-15 * np.abs(b).max())
    for nt in (1, 3, 4):
        nzp, bp = oracle.numeric_refill_cartesian_par(D, sizes, part, *args, nt)
        np.testing.assert_allclose(nzp, nz, rtol=1e-14, atol=1e-15 * np.abs(nz).max())
        np.testing.assert_allclose(bp, b, rtol=1e-14, atol=1e-15 * np.abs(b).max())
