"""Pins of the simplex / conforming-numbering oracle (CPU): reference fixtures for
the polytope face order and the conforming numbering, closed forms for the
quadrature and the P2 element, invariants of the simplexified mesh."""
import itertools

import numpy as np
import pytest

import conforming_oracle as O


def test_tet_polytope_matches_reference_fixture():
    # test/ReferenceFEsTests/ExtrusionPolytopesTests.jl:43-48: 15 faces, vertex coords
    assert len(O.TET.nfaces) == 15
    np.testing.assert_array_equal(O.TET.vertex_coords,
                                  [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    assert [len(O.TET.faces[d]) for d in range(4)] == [4, 6, 4, 1]
    assert len(O.TRI.nfaces) == 7 and len(O.HEX.nfaces) == 27 and len(O.QUAD.nfaces) == 9
    # every edge of the tet appears once; faces are the 4 triangles
    assert sorted(O.TET.faces[1]) == sorted(itertools.combinations(range(1, 5), 2))
    assert sorted(O.TET.faces[2]) == sorted(itertools.combinations(range(1, 5), 3))


def test_ncube_face_order_matches_the_q1_oracle(oracle):
    # the n-cube ids of the restated recursion agree with gor_ncube_face_id, which
    # is pinned by the reference's face-label fixtures
    for D, poly in ((2, O.QUAD), (3, O.HEX)):
        for k, (anchor, ext) in enumerate(poly.nfaces):
            fixed = [None if e else a for a, e in zip(anchor, ext)]
            assert O.box_entity(D, fixed) == k + 1
            free = sum(1 << d for d in range(D) if fixed[d] is None)
            anc = sum(1 << d for d in range(D) if fixed[d] == 1)
            assert oracle.ncube_face_id(D, free, anc) == k + 1


def quad_model_numbering(partition, dirichlet_entities, order):
    n0, n1 = partition
    cells = []
    for j in range(n1):
        for i in range(n0):
            cells.append([i + (n0 + 1) * j, i + 1 + (n0 + 1) * j,
                          i + (n0 + 1) * (j + 1), i + 1 + (n0 + 1) * (j + 1)])
    cells = np.array(cells)
    nn = (n0 + 1) * (n1 + 1)
    ijk = lambda n: (n % (n0 + 1), n // (n0 + 1))
    edges, ne = O.first_touch_faces(cells, O.QUAD.faces[1])
    vtag = [1 if O.simplex_face_entity(2, partition, [ijk(n)]) in dirichlet_entities else 0
            for n in range(nn)]
    etag = [0] * ne
    for c in range(len(cells)):
        for le, lv in enumerate(O.QUAD.faces[1]):
            ent = O.simplex_face_entity(2, partition, [ijk(cells[c][v - 1]) for v in lv])
            etag[edges[c, le] - 1] = 1 if ent in dirichlet_entities else 0
    own, nloc = O.lagrangian_own_dofs(O.QUAD, order)
    cf = np.arange(1, len(cells) + 1).reshape(-1, 1)
    return O.conforming_cell_dofs([cells + 1, edges, cf], [vtag, etag, [0] * len(cells)],
                                  own, nloc)


def test_conforming_q2_numbering_matches_reference_fixture():
    # test/FESpacesTests/ConformingFESpacesTests.jl:16-46: 3x3 QUAD, Q2, tags 1 and 6
    cd, nfree, ndiri, _ = quad_model_numbering((3, 3), {1, 6}, 2)
    r = [[-1, 1, 4, 5, 14, 15, 16, 17, 35], [1, 2, 5, 6, 18, 19, 17, 20, 36],
         [2, 3, 6, 7, 21, 22, 20, 23, 37], [4, 5, 8, 9, 15, 24, 25, 26, 38],
         [5, 6, 9, 10, 19, 27, 26, 28, 39], [6, 7, 10, 11, 22, 29, 28, 30, 40],
         [8, 9, 12, -2, 24, -4, 31, 32, 41], [9, 10, -2, -3, 27, -5, 32, 33, 42],
         [10, 11, -3, 13, 29, -6, 33, 34, 43]]
    np.testing.assert_array_equal(cd, r)
    assert (nfree, ndiri) == (43, 6)


def test_conforming_q1_numbering_without_tags():
    # order 1, no Dirichlet tags: vertex dofs only, numbered in vertex order
    # (CLagrangianFESpacesTests.jl:23, cell dofs == cell node ids)
    cd, nfree, ndiri, _ = quad_model_numbering((2, 2), set(), 1)
    np.testing.assert_array_equal(cd, [[1, 2, 4, 5], [2, 3, 5, 6], [4, 5, 7, 8], [5, 6, 8, 9]])
    assert (nfree, ndiri) == (9, 0)


def test_tet_rule_is_exact_to_degree_5():
    P, W = O.tet_degree5_rule()
    assert len(W) == 14 and np.all(W > 0)
    assert np.all(P >= 0) and np.all(P.sum(1) <= 1)
    for a, b, c in itertools.product(range(7), repeat=3):
        if a + b + c <= 5:
            got = np.sum(W * P[:, 0] ** a * P[:, 1] ** b * P[:, 2] ** c)
            assert abs(got - O.tet_monomial_integral(a, b, c)) < 1e-16
    # not exact at degree 6 (so the test above has teeth)
    errs = [abs(np.sum(W * P[:, 0] ** a * P[:, 1] ** b * P[:, 2] ** c) -
                O.tet_monomial_integral(a, b, c))
            for a, b, c in itertools.product(range(7), repeat=3) if a + b + c == 6]
    assert max(errs) > 1e-8


def test_p2_basis_nodal_and_complete():
    X = O.p2_tet_nodes()
    shp, grad = O.p2_tet_tabulate(X)
    np.testing.assert_allclose(shp, np.eye(10), atol=1e-15)
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 0.3, (20, 3))
    s, g = O.p2_tet_tabulate(x)
    np.testing.assert_allclose(s.sum(1), 1.0, atol=1e-14)
    np.testing.assert_allclose(g.sum(1), 0.0, atol=1e-13)
    # quadratic reproduction: u = 1 + 2x - y + 3z + xy - 2z^2 + yz
    u = lambda p: 1 + 2 * p[:, 0] - p[:, 1] + 3 * p[:, 2] + p[:, 0] * p[:, 1] - 2 * p[:, 2] ** 2 + p[:, 1] * p[:, 2]
    np.testing.assert_allclose(s @ u(X), u(x), atol=1e-13)
    # gradients by central differences
    h = 1e-6
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        fd = (O.p2_tet_tabulate(x + e)[0] - O.p2_tet_tabulate(x - e)[0]) / (2 * h)
        np.testing.assert_allclose(g[:, :, d], fd, atol=1e-8)


def test_p2_reference_stiffness_exact_vs_quadrature():
    K = O.p2_tet_stiffness_exact()
    P, W = O.tet_degree5_rule()
    _, g = O.p2_tet_tabulate(P)
    Kq = np.einsum("q,qad,qbd->ab", W, g, g)
    np.testing.assert_allclose(Kq, K, atol=1e-15)
    np.testing.assert_allclose(K, K.T, atol=1e-16)
    np.testing.assert_allclose(K.sum(1), 0.0, atol=1e-15)
    ev = np.linalg.eigvalsh(K)
    assert ev[0] > -1e-14 and ev[1] > 1e-3              # one constant null mode
    # the vertex-vertex entries of the reference P2 stiffness: K_11 = 3/10 · (3/6) ... as
    # the closed form ∫|∇φ_1|² = ∫(4λ-1)² |∇λ_1|², |∇λ_1|² = 3, ∫(4λ-1)² = 1/10
    assert abs(K[0, 0] - 3 * (16 * 2 / 120 - 8 / 24 + 1 / 6)) < 1e-15


def test_simplexified_mesh_invariants():
    part = (3, 2, 2)
    cells = O.simplexify_cartesian(part)
    assert len(cells) == 6 * np.prod(part)
    n0, n1, n2 = part
    xyz = np.array([[i / n0, j / n1, k / n2] for k in range(n2 + 1) for j in range(n1 + 1)
                    for i in range(n0 + 1)])
    vol = 0.0
    tri_count = {}
    for c in cells:
        J = (xyz[c[1:]] - xyz[c[0]]).T
        d = np.linalg.det(J)
        assert abs(d) > 1e-12
        vol += abs(d) / 6
        for f in itertools.combinations(sorted(c), 3):
            tri_count[f] = tri_count.get(f, 0) + 1
    assert abs(vol - 1.0) < 1e-14
    assert set(tri_count.values()) <= {1, 2}             # conforming: a face has <= 2 cells
    # boundary triangles: 2 per boundary quad face
    nb = sum(1 for v in tri_count.values() if v == 1)
    assert nb == 2 * 2 * (n0 * n1 + n1 * n2 + n0 * n2)


def test_p2_numbering_invariants():
    part = (3, 2, 2)
    cells, cd, nfree, ndiri, edges = O.p2_tet_numbering(part, O.boundary_entities(3))
    n0, n1, n2 = part
    nedges = (n0 * (n1 + 1) * (n2 + 1) + (n0 + 1) * n1 * (n2 + 1) + (n0 + 1) * (n1 + 1) * n2 +
              n0 * n1 * (n2 + 1) + n0 * (n1 + 1) * n2 + (n0 + 1) * n1 * n2 + n0 * n1 * n2)
    assert edges.max() == nedges
    nnodes = (n0 + 1) * (n1 + 1) * (n2 + 1)
    assert nfree + ndiri == nnodes + nedges
    # first touch: edge ids appear in increasing order of first occurrence
    first = []
    seen = set()
    for e in edges.ravel():
        if e not in seen:
            seen.add(e)
            first.append(e)
    assert first == list(range(1, nedges + 1))
    # conformity: the dof of an edge depends only on its two vertices
    by_pair = {}
    for c in range(len(cells)):
        for k, (a, b) in enumerate(O.TET.faces[1]):
            key = frozenset((cells[c][a - 1], cells[c][b - 1]))
            assert by_pair.setdefault(key, cd[c, 4 + k]) == cd[c, 4 + k]
    # vertex dofs: free interior vertices first, in vertex order
    vdofs = {}
    for c in range(len(cells)):
        for a in range(4):
            vdofs[cells[c][a]] = cd[c, a]
    interior = [n for n in range(nnodes) if 0 < n % (n0 + 1) < n0 and
                0 < (n // (n0 + 1)) % (n1 + 1) < n1 and 0 < n // ((n0 + 1) * (n1 + 1)) < n2]
    assert [vdofs[n] for n in interior] == list(range(1, len(interior) + 1))
    # a Dirichlet edge has both vertices on one boundary plane
    for key, dof in by_pair.items():
        u, v = [O.node_ijk(part, n) for n in key]
        on_plane = any((u[d] == v[d] == 0) or (u[d] == v[d] == part[d]) for d in range(3))
        assert (dof < 0) == on_plane


def jittered_xyz(part, seed, amp=0.2):
    n0, n1, n2 = part
    xyz = np.array([[i / n0, j / n1, k / n2] for k in range(n2 + 1) for j in range(n1 + 1)
                    for i in range(n0 + 1)])
    rng = np.random.default_rng(seed)
    ijk = np.array([O.node_ijk(part, n) for n in range(len(xyz))])
    inner = np.all((ijk > 0) & (ijk < np.array(part)), axis=1)
    h = 1.0 / np.array(part)
    xyz[inner] += rng.uniform(-amp, amp, (inner.sum(), 3)) * h
    return xyz


@pytest.mark.parametrize("order", [1, 2])
def test_tet_assembly_reproduces_polynomials(oracle, order):
    # P_k holds every polynomial of degree k, so with f = -Δu and Dirichlet data from
    # u the discrete system is solved exactly by the interpolant: A u_I = b.
    part = (3, 3, 2)
    xyz = jittered_xyz(part, 1)
    if order == 2:
        u = lambda p: (1 + p[0] - 2 * p[1] + 0.5 * p[2] + p[0] ** 2 + p[1] ** 2 +
                       2 * p[1] * p[2] - 0.5 * p[2] ** 2)
        lap = 2.0 + 2.0 - 1.0
    else:
        u = lambda p: 1 + p[0] - 2 * p[1] + 0.5 * p[2]
        lap = 0.0
    P, W = O.tet_degree5_rule()
    r = O.tet_poisson(part, xyz, order, fsrc=np.full(6 * 18, -lap), qx=P, qw=W, g=u)
    cp, rv, nz = oracle.assemble_csc(r["cell_dofs"], r["cellmats"], r["nfree"], r["nfree"])
    b = oracle.assemble_vector(r["cell_dofs"], r["cellvecs"], r["nfree"])
    # interpolant at the free dofs
    cells, cd = r["cells"], r["cell_dofs"]
    X = xyz[cells]
    nodes = np.concatenate([X, 0.5 * (X[:, [a - 1 for a, _ in O.TET.faces[1]]] +
                                      X[:, [b - 1 for _, b in O.TET.faces[1]]])], axis=1)
    uI = np.zeros(r["nfree"])
    for c in range(len(cells)):
        for a in range(cd.shape[1]):
            if cd[c, a] > 0:
                uI[cd[c, a] - 1] = u(nodes[c, a])
    A = oracle.to_scipy(cp, rv, nz, r["nfree"])
    assert np.max(np.abs(A @ uI - b)) < 1e-12 * max(1.0, np.max(np.abs(b)))
    # without the Dirichlet lifting the identity fails (the check has teeth)
    r0 = O.tet_poisson(part, xyz, order, fsrc=np.full(6 * 18, -lap), qx=P, qw=W)
    b0 = oracle.assemble_vector(r0["cell_dofs"], r0["cellvecs"], r0["nfree"])
    assert np.max(np.abs(A @ uI - b0)) > 1e-3


def test_vector_numbering_with_components_matches_reference_fixture():
    # test/FESpacesTests/ConformingFESpacesTests.jl:48-65: 3x3 QUAD, vector order 1,
    # tags 1 and 6 with masks (true, true) and (false, true)
    cells, cd, nfree, ndiri, dtag = O.cartesian_numbering((3, 3), 1, 2, [{1}, {6}],
                                                           comp_masks=[(1, 1), (0, 1)])
    r = [[-1, 1, 7, 9, -2, 2, 8, 10], [1, 3, 9, 11, 2, 4, 10, 12], [3, 5, 11, 13, 4, 6, 12, 14],
         [7, 9, 15, 17, 8, 10, 16, 18], [9, 11, 17, 19, 10, 12, 18, 20],
         [11, 13, 19, 21, 12, 14, 20, 22], [15, 17, 23, 25, 16, 18, 24, -3],
         [17, 19, 25, 26, 18, 20, -3, -4], [19, 21, 26, 27, 20, 22, -4, 28]]
    np.testing.assert_array_equal(cd, r)
    assert (nfree, ndiri) == (28, 4)
    assert dtag == [1, 1, 2, 2]
    # and the scalar Q2 fixture through the same Cartesian driver
    _, cd2, nf2, nd2, _ = O.cartesian_numbering((3, 3), 2, 1, [{1}, {6}])
    assert (nf2, nd2) == (43, 6) and cd2[0].tolist() == [-1, 1, 4, 5, 14, 15, 16, 17, 35]


def rigid_modes(nodes):
    nn = len(nodes)
    modes = []
    for k in range(3):
        u = np.zeros((3, nn))
        u[k] = 1.0
        modes.append(u.ravel())
    for a, b in ((0, 1), (1, 2), (2, 0)):            # infinitesimal rotations
        u = np.zeros((3, nn))
        u[a] = -nodes[:, b]
        u[b] = nodes[:, a]
        modes.append(u.ravel())
    return np.array(modes)


def test_elasticity_element_kernel_is_rigid_motions(oracle):
    h = (0.5, 0.25, 0.2)
    K, m = O.elasticity_reference(oracle, h, 0.3)
    assert K.shape == (81, 81)
    np.testing.assert_allclose(K, K.T, atol=1e-15 * np.abs(K).max())
    nodes = oracle.qk_nodes(3, 2) * np.array(h)
    R = rigid_modes(nodes)
    assert np.abs(K @ R.T).max() < 1e-13 * np.abs(K).max()
    ev = np.linalg.eigvalsh(K)
    assert np.all(ev[:6] < 1e-12 * ev[-1]) and ev[6] > 1e-4 * ev[-1]
    np.testing.assert_allclose(m.sum(), np.prod(h), rtol=1e-14)


def test_elasticity_assembly_reproduces_quadratics(oracle):
    # u = (xy, yz, zx + x^2) with E = 1: -div σ(u) = -((λ+μ)(1,1,1) + μ(0,0,2)), and the
    # Q2 interpolant solves the discrete system exactly: A u_I = b
    part = (3, 2, 2)
    nu = 0.3
    lam, mu = O.lame(1.0, nu)
    f = -np.array([lam + mu, lam + mu, lam + 3 * mu])
    u = lambda x: np.array([x[0] * x[1], x[1] * x[2], x[2] * x[0] + x[0] ** 2])
    r = O.elasticity_q2_box(oracle, part, np.ones(12), nu, f, [O.boundary_entities(3)], g=u)
    cp, rv, nz = oracle.assemble_csc(r["cell_dofs"], r["cellmats"], r["nfree"], r["nfree"])
    b = oracle.assemble_vector(r["cell_dofs"], r["cellvecs"], r["nfree"])
    nodes = oracle.qk_nodes(3, 2)
    uI = np.zeros(r["nfree"])
    for c, cell in enumerate(r["cells"]):
        x0 = np.array([O.lattice_ijk(part, cell[0])[d] * r["h"][d] for d in range(3)])
        for a in range(81):
            if r["cell_dofs"][c, a] > 0:
                uI[r["cell_dofs"][c, a] - 1] = u(x0 + nodes[a % 27] * np.array(r["h"]))[a // 27]
    A = oracle.to_scipy(cp, rv, nz, r["nfree"])
    assert np.max(np.abs(A @ uI - b)) < 1e-12 * max(1.0, np.abs(b).max())
    r0 = O.elasticity_q2_box(oracle, part, np.ones(12), nu, f, [O.boundary_entities(3)])
    b0 = oracle.assemble_vector(r0["cell_dofs"], r0["cellvecs"], r0["nfree"])
    assert np.max(np.abs(A @ uI - b0)) > 1e-3           # the lifting matters
