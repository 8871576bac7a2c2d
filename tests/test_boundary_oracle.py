"""Pins of the oracle's boundary integrals (no GPU).

boundary_faces / boundary_face_arrays restate integration over a BoundaryTriangulation
(src/Geometry/BoundaryTriangulations.jl:122-183) with generic facet geometry (facet
measure from the tangents, normal oriented away from the cell).  Fixed by:
* closed forms: the Q1 boundary mass matrix of an edge, h/6 [[2,1],[1,2]], and of a
  square face, h²/36 [[4,2,2,1],...]; ∫_Γ 1 = |∂Ω|;
* consistency: for a linear u, ∫_F φ_a (n⋅∇u) = (n⋅∇u) ∫_F φ_a, with ∫_F φ_a = |F|/2^(D-1)
  at the facet's vertices;
* the manufactured-solution bar of test/GridapTests/PoissonTests.jl:101-124: Dirichlet
  on tags 1, 2, 5, Neumann on 7, 8, Nitsche (γ = 10) on 6; a u that Q1 holds (bilinear,
  f = -Δu = 0) is reproduced to rounding, and a wrong normal sign breaks it.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

U = lambda x: 1 + x[0] + 2 * x[1] + 0.5 * x[0] * x[1]
GU = lambda x: np.array([1 + 0.5 * x[1], 2 + 0.5 * x[0]])


def test_boundary_mass_closed_forms(oracle):
    m = oracle.cartesian_model((0, 1, 0, 1), (4, 4))
    faces = oracle.boundary_faces(m, [5])
    assert len(faces) == 4
    M, _ = oracle.boundary_face_arrays(m, faces, 2, alpha=1.0)
    np.testing.assert_allclose(M[0][:2, :2] * 24.0, [[2, 1], [1, 2]], atol=1e-14)
    assert np.abs(M[0][2:, :]).max() == 0.0 and np.abs(M[0][:, 2:]).max() == 0.0
    m3 = oracle.cartesian_model((0, 1, 0, 1, 0, 1), (3, 3, 3))
    f3 = oracle.boundary_faces(m3, [21])
    M3, _ = oracle.boundary_face_arrays(m3, f3, 2, alpha=1.0)
    ref = np.array([[4, 2, 2, 1], [2, 4, 1, 2], [2, 1, 4, 2], [1, 2, 2, 4]]) / 36.0 / 9.0
    np.testing.assert_allclose(M3[0][:4, :4], ref, atol=1e-15)
    # ∫_Γ 1 over the whole boundary = perimeter / surface area
    for mod, ents, meas in ((m, [5, 6, 7, 8], 4.0), (m3, list(range(21, 27)), 6.0)):
        fa = oracle.boundary_faces(mod, ents)
        Mb, _ = oracle.boundary_face_arrays(mod, fa, 2, alpha=1.0)
        assert abs(Mb.sum() - meas) < 1e-13


@pytest.mark.parametrize("D", [2, 3])
def test_normal_derivative_consistency(oracle, D):
    m = oracle.cartesian_model((0, 1) * D, (3,) * D)
    a = np.arange(1, D + 1, dtype=float)
    ents = [5, 6, 7, 8] if D == 2 else list(range(21, 27))
    faces = oracle.boundary_faces(m, ents)
    _, V = oracle.boundary_face_arrays(m, faces, 2, g=lambda x, n: float(a @ n))
    _, V1 = oracle.boundary_face_arrays(m, faces, 2, g=lambda x, n: 1.0)
    # ∫_F φ_a = |F| / 2^(D-1) at the facet's vertices, 0 elsewhere
    area = (1.0 / 3.0) ** (D - 1)
    for k, (c, loc, axis, val) in enumerate(faces):
        n = np.zeros(D)
        n[axis] = 1.0 if val else -1.0
        exp = np.zeros(1 << D)
        exp[loc] = area / 2 ** (D - 1)
        np.testing.assert_allclose(V1[k], exp, atol=1e-15)
        np.testing.assert_allclose(V[k], exp * (a @ n), atol=1e-14)


def _driver(oracle, part, normal_sign=1.0):
    gam = 10.0
    h = 1.0 / part[0]
    terms = [([7, 8], 0.0, 0.0, lambda x, n: normal_sign * (GU(x) @ n), None),
             ([6], gam / h, -1.0, lambda x, n: gam / h * U(x), lambda x, n: -U(x))]
    r = oracle.poisson_q1_with_boundary((0, 1, 0, 1), part, [1, 2, 5], terms, g=U,
                                        fsrc=np.zeros(int(np.prod(part))))
    A = oracle.to_scipy(r["colptr"], r["rowval"], r["nzval"], r["nfree"])
    x = spla.spsolve(A.tocsc(), r["b"])
    n2d = r["node_to_dof"][:, 0]
    xyz = r["model"]["xyz"]
    free = np.nonzero(n2d > 0)[0]
    uI = np.array([U(p) for p in xyz[free]])
    return A, np.abs(x[n2d[free] - 1] - uI).max() / np.abs(uI).max()


@pytest.mark.parametrize("part", [(4, 4), (7, 5)])
def test_poisson_driver_manufactured_solution(oracle, part):
    A, err = _driver(oracle, part)
    assert err <= 1e-12                     # PoissonTests.jl:123-124 asks 1e-8
    assert abs(A - A.T).max() <= 1e-15 * abs(A).max()
    _, bad = _driver(oracle, part, normal_sign=-1.0)
    assert bad > 1e-3


# ---- the driver at its own order 2 (PoissonTests.jl:12, u_scal = x1² + x2, f = -Δu) ----

U2 = lambda x: x[..., 0] ** 2 + x[..., 1]
GU2 = lambda x: np.array([2 * x[0], 1.0])


def _terms(part, gam=10.0, normal_sign=1.0, u=U, gu=GU):
    h = 1.0 / part[0]
    return [([7, 8], 0.0, 0.0, lambda x, n: normal_sign * (gu(x) @ n), None),
            ([6], gam / h, -1.0, lambda x, n: gam / h * u(x), lambda x, n: -u(x))]


def test_qk_driver_order1_equals_q1_driver(oracle):
    """The Q_k restatement (conforming numbering, numpy cell integrals, Q_k face arrays)
    at order 1 equals the pinned Q1 driver."""
    import conforming_oracle as CO
    part = (5, 4)
    r1 = oracle.poisson_q1_with_boundary((0, 1, 0, 1), part, [1, 2, 5], _terms(part), g=U,
                                         fsrc=np.zeros(int(np.prod(part))))
    rk = CO.poisson_qk_with_boundary(oracle, (0, 1, 0, 1), part, 1, [1, 2, 5], _terms(part),
                                     2, 2, g=U)
    np.testing.assert_array_equal(rk["colptr"], r1["colptr"])
    np.testing.assert_array_equal(rk["rowval"], r1["rowval"])
    np.testing.assert_allclose(rk["nzval"], r1["nzval"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(rk["b"], r1["b"], rtol=0, atol=1e-13)


def _free_coords(oracle, r, order):
    import conforming_oracle as CO
    m, cd = r["model"], r["cell_dofs"]
    X = m["xyz"][m["cell_nodes"] - 1]
    n1, _ = oracle.qk_tabulate(2, 1, oracle.qk_nodes(2, order))
    P = CO.q1_map_points(X, n1)
    x = np.zeros((r["nfree"], 2))
    x[cd[cd > 0] - 1] = P[cd > 0]
    return x


@pytest.mark.parametrize("part", [(4, 4), (6, 5)])
def test_qk_driver_order2_manufactured_solution(oracle, part):
    """PoissonTests.jl at its own order 2 with u = x1² + x2 (in Q2) and f = -Δu = -2:
    the solution at the dof nodes is u to rounding (the test asks 1e-8 in L2); a wrong
    normal sign breaks it."""
    import conforming_oracle as CO
    u1 = lambda x: float(x[0] ** 2 + x[1])

    def solve(sign):
        r = CO.poisson_qk_with_boundary(oracle, (0, 1, 0, 1), part, 2, [1, 2, 5],
                                        _terms(part, normal_sign=sign, u=u1, gu=GU2),
                                        4, 4, g=u1, f=lambda x: np.full(len(x), -2.0))
        A = oracle.to_scipy(r["colptr"], r["rowval"], r["nzval"], r["nfree"])
        sol = spla.spsolve(A.tocsc(), r["b"])
        uI = U2(_free_coords(oracle, r, 2))
        return A, np.abs(sol - uI).max() / np.abs(uI).max()
    A, err = solve(1.0)
    assert err <= 1e-11
    assert abs(A - A.T).max() <= 1e-15 * abs(A).max()
    _, bad = solve(-1.0)
    assert bad > 1e-3
