"""Boundary integrals on the device (needs an MI355X): Neumann and Nitsche terms of
test/GridapTests/PoissonTests.jl:9-108 assembled into the same CSC and b as the cell
terms, against the oracle.

Bar: pattern bit-exact, nzval and b within 1e-12 relative (max-abs error / max-abs
value) on both numeric paths (lattice + face kernel, general + face kernel); the
driver's manufactured solution reproduced by the device solve (PoissonTests.jl:123-124
asks 1e-8 relative).
"""
import sys

import numpy as np
import pytest

import gridap_hip as G

pytestmark = pytest.mark.gpu

U_EX = lambda x: 1 + x[..., 0] + 2 * x[..., 1] + 0.5 * x[..., 0] * x[..., 1]
GU_EX = lambda x: np.array([1 + 0.5 * x[1], 2 + 0.5 * x[0]])


def announce(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300) if a.size else 0.0


def poisson_driver(part, path="auto", gam=10.0):
    """PoissonTests.jl:9-108 with Q1 and a u that Q1 holds."""
    model = G.CartesianDiscreteModel((0, 1, 0, 1), part)
    labels = model.get_face_labeling()
    labels.add_tag_from_tags("dirichlet", [1, 2, 5])
    labels.add_tag_from_tags("neumann", [7, 8])
    labels.add_tag_from_tags("nitsche", 6)
    h = 1.0 / part[0]
    Ω = G.Triangulation(model)
    Γn = G.BoundaryTriangulation(model, labels, tags="neumann")
    Γd = G.BoundaryTriangulation(model, labels, tags="nitsche")
    dΩ, dΓn, dΓd = G.Measure(Ω, 2), G.Measure(Γn, 2), G.Measure(Γd, 2)
    nn, nd = G.get_normal_vector(Γn), G.get_normal_vector(Γd)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1),
                      dirichlet_tags="dirichlet", labels=labels)
    U = G.TrialFESpace(V, U_EX)
    uh = G.interpolate(U_EX, U)
    a = lambda u, v: (G.integrate(G.inner(G.gradient(v), G.gradient(u)), dΩ) +
                      G.integrate((gam / h) * v * u - v * G.dot(nd, G.gradient(u)) -
                                  G.dot(nd, G.gradient(v)) * u, dΓd))
    l = lambda v: (G.integrate(0.0 * v, dΩ) +
                   G.integrate(v * G.dot(nn, G.gradient(uh)), dΓn) +
                   G.integrate((gam / h) * v * uh - G.dot(nd, G.gradient(v)) * uh, dΓd))
    assem = G.SparseMatrixAssembler(U, V, path=path)
    return G.AffineFEOperator(a, l, U, V, assem), V, model


def oracle_driver(oracle, part, gam=10.0):
    h = 1.0 / part[0]
    U1 = lambda x: U_EX(np.asarray(x))
    terms = [([7, 8], 0.0, 0.0, lambda x, n: GU_EX(x) @ n, None),
             ([6], gam / h, -1.0, lambda x, n: gam / h * U1(x), lambda x, n: -U1(x))]
    return oracle.poisson_q1_with_boundary((0, 1, 0, 1), part, [1, 2, 5], terms, g=U1,
                                           fsrc=np.zeros(int(np.prod(part))))


@pytest.mark.parametrize("path", ["auto", "general"])
@pytest.mark.parametrize("part", [(4, 4), (13, 9)])
def test_poisson_driver_matches_oracle(oracle, part, path):
    announce("boundary terms PoissonTests driver %s path=%s" % (part, path))
    op, V, model = poisson_driver(part, path)
    ref = oracle_driver(oracle, part)
    A, b = op.get_matrix(), op.get_vector()
    np.testing.assert_array_equal(A.colptr, ref["colptr"])
    np.testing.assert_array_equal(A.rowval, ref["rowval"])
    assert relmax(A.nzval, ref["nzval"]) <= 1e-12
    assert relmax(b, ref["b"]) <= 1e-12
    info = op.assem.ctx.pattern_info()
    assert info["lattice"] == 1          # the face kernel runs after either cell kernel


def test_poisson_driver_solution():
    announce("boundary terms PoissonTests driver 32x24 solve")
    op, V, model = poisson_driver((32, 24))
    uh = G.solve(op, G.PCGSolver(rtol=1e-12))
    x = model.node_coordinates()[V.free_dof_to_node - 1]
    uI = U_EX(x)
    assert np.max(np.abs(uh.get_free_dof_values() - uI)) / np.max(np.abs(uI)) <= 1e-8


def test_3d_robin_and_neumann_match_oracle(oracle):
    """3-D 9x8x7: Robin (α = 3) on z = 0 and z = 1, a Neumann load g(x) on x = 1,
    Dirichlet on x = 0, κ = 1, f = 1."""
    part = (9, 8, 7)
    announce("boundary terms 3-D Robin + Neumann %s" % (part,))
    model = G.CartesianDiscreteModel((0, 1) * 3, part)
    labels = model.get_face_labeling()
    Ω = G.Triangulation(model)
    Γr = G.BoundaryTriangulation(model, labels, tags=["tag_21", "tag_22"])
    Γn = G.BoundaryTriangulation(model, labels, tags="tag_26")
    dΩ, dΓr, dΓn = G.Measure(Ω, 2), G.Measure(Γr, 2), G.Measure(Γn, 2)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1), dirichlet_tags="tag_25")
    U = G.TrialFESpace(V, lambda x: 1.0 + x[..., 1])
    g = lambda x: np.sin(3 * x[..., 1]) + x[..., 2]
    a = lambda u, v: (G.integrate(G.inner(G.gradient(v), G.gradient(u)), dΩ) +
                      G.integrate(3.0 * v * u, dΓr))
    l = lambda v: G.integrate(1.0 * v, dΩ) + G.integrate(v * g, dΓn)
    op = G.AffineFEOperator(a, l, U, V)
    gx = lambda x: 1.0 + x[1]
    ref = oracle.poisson_q1_with_boundary(
        (0, 1) * 3, part, [25],
        [([21, 22], 3.0, 0.0, None, None),
         ([26], 0.0, 0.0, lambda x, n: np.sin(3 * x[1]) + x[2], None)], g=gx)
    A, b = op.get_matrix(), op.get_vector()
    np.testing.assert_array_equal(A.rowval, ref["rowval"])
    assert relmax(A.nzval, ref["nzval"]) <= 1e-12
    assert relmax(b, ref["b"]) <= 1e-12
    # the face kernel adds each face matrix (symmetric bitwise) into column j by the
    # owner of j, faces ascending: the assembled matrix stays symmetric bitwise
    S = A.to_scipy()
    assert (S - S.T).count_nonzero() == 0


# ---- the driver at its own order (PoissonTests.jl:12: order = 2, degree = order) -------

U2 = lambda x: x[..., 0] ** 2 + x[..., 1]          # u_scal (PoissonTests.jl:35)
GU2 = lambda x: np.array([2 * x[0], 1.0])


def poisson_driver_q2(part, gam=10.0, precision="fp64"):
    """PoissonTests.jl:9-108 as written for scalar data: Q2, degree 2, f = -Δu = -2 as a
    function of x, Neumann data n⋅∇uh of the interpolant, Nitsche on tag 6."""
    order = 2
    model = G.CartesianDiscreteModel((0, 1, 0, 1), part)
    labels = model.get_face_labeling()
    labels.add_tag_from_tags("dirichlet", [1, 2, 5])
    labels.add_tag_from_tags("neumann", [7, 8])
    labels.add_tag_from_tags("nitsche", 6)
    h = 1.0 / part[0]
    Ω = G.Triangulation(model)
    Γn = G.BoundaryTriangulation(model, labels, tags="neumann")
    Γd = G.BoundaryTriangulation(model, labels, tags="nitsche")
    dΩ, dΓn, dΓd = G.Measure(Ω, order), G.Measure(Γn, order), G.Measure(Γd, order)
    nn, nd = G.get_normal_vector(Γn), G.get_normal_vector(Γd)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, order),
                      dirichlet_tags="dirichlet", labels=labels)
    U = G.TrialFESpace(V, U2)
    uh = G.interpolate(U2, U)
    f = lambda x: np.full(x.shape[:-1], -2.0)
    a = lambda u, v: (G.integrate(G.inner(G.gradient(v), G.gradient(u)), dΩ) +
                      G.integrate((gam / h) * v * u - v * G.dot(nd, G.gradient(u)) -
                                  G.dot(nd, G.gradient(v)) * u, dΓd))
    l = lambda v: (G.integrate(v * f, dΩ) +
                   G.integrate(v * G.dot(nn, G.gradient(uh)), dΓn) +
                   G.integrate((gam / h) * v * uh - G.dot(nd, G.gradient(v)) * uh, dΓd))
    assem = G.SparseMatrixAssembler(U, V, precision=precision)
    return G.AffineFEOperator(a, l, U, V, assem), V


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("part", [(4, 4), (11, 7)])
def test_poisson_driver_order2_matches_oracle(oracle, part, precision):
    """The scalar case of the driver at its own order 2 (general n-cube element lists,
    source at the quadrature points, Q2 face kernel): pattern bit-exact, nzval and b
    against the oracle (1e-12 FP64, 1e-6 relative Frobenius FP32), and the device
    solve reproduces u = x1² + x2 (PoissonTests.jl:123 asks 1e-8 relative in L2)."""
    import conforming_oracle as CO
    announce("PoissonTests driver order 2 %s %s" % (part, precision))
    op, V = poisson_driver_q2(part, precision=precision)
    gam, h = 10.0, 1.0 / part[0]
    u1 = lambda x: float(x[0] ** 2 + x[1])
    terms = [([7, 8], 0.0, 0.0, lambda x, n: GU2(x) @ n, None),
             ([6], gam / h, -1.0, lambda x, n: gam / h * u1(x), lambda x, n: -u1(x))]
    ref = CO.poisson_qk_with_boundary(oracle, (0, 1, 0, 1), part, 2, [1, 2, 5], terms, 2, 2,
                                      g=u1, f=lambda x: np.full(len(x), -2.0))
    np.testing.assert_array_equal(V.cell_dof_ids, ref["cell_dofs"])
    A, b = op.get_matrix(), op.get_vector()
    np.testing.assert_array_equal(A.colptr, ref["colptr"])
    np.testing.assert_array_equal(A.rowval, ref["rowval"])
    if precision == "fp64":
        assert relmax(A.nzval, ref["nzval"]) <= 1e-12
        assert relmax(b, ref["b"]) <= 1e-12
    else:
        assert np.linalg.norm(A.nzval - ref["nzval"]) <= 1e-6 * np.linalg.norm(ref["nzval"])
        assert np.linalg.norm(b - ref["b"]) <= 1e-6 * np.linalg.norm(ref["b"])
    if precision == "fp64":
        x = G.solve(op).get_free_dof_values()
        uI = U2(V.free_dof_coords)
        assert np.abs(x - uI).max() <= 1e-8 * np.abs(uI).max()


@pytest.mark.parametrize("mass", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_3d_faces_folded_into_lattice_match_face_kernel(oracle, mass, precision):
    """3-D 10x9x8 with Nitsche terms (α = γ/h, β = -1) on z = 1, Robin (α = 3) on y = 0, a
    Neumann load on x = 1 and Dirichlet data on x = 0.  On the lattice the faces are
    folded into the owner-computes columns (with a cell mass term: the stored element
    matrices, FP32 or FP64); path "general" runs the per-cell kernel and the separate
    face kernel.  Both agree to 1e-12 (FP64), and without the mass term the FP64 matrix
    equals the oracle's."""
    part = (10, 9, 8)
    announce("boundary faces folded into the lattice %s mass=%s %s" % (part, mass, precision))
    h = 1.0 / part[2]
    gam = 10.0
    ue = lambda x: 1.0 + x[..., 0] - 0.5 * x[..., 2]
    model = G.CartesianDiscreteModel((0, 1) * 3, part)
    labels = model.get_face_labeling()
    Ω = G.Triangulation(model)
    Γd = G.BoundaryTriangulation(model, labels, tags="tag_22")
    Γr = G.BoundaryTriangulation(model, labels, tags="tag_23")
    Γn = G.BoundaryTriangulation(model, labels, tags="tag_26")
    dΩ, dΓd, dΓr, dΓn = (G.Measure(t, 2) for t in (Ω, Γd, Γr, Γn))
    nd = G.get_normal_vector(Γd)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1), dirichlet_tags="tag_25")
    U = G.TrialFESpace(V, ue)
    uh = G.interpolate(ue, U)
    kappa = np.exp(np.random.default_rng(4).normal(size=model.num_cells()))
    k = G.CellField(kappa, Ω)
    g = lambda x: np.cos(2 * x[..., 1]) + x[..., 2]
    sig = 2.5 if mass else 0.0

    def a(u, v):
        vol = k * G.inner(G.gradient(v), G.gradient(u))
        if mass:
            vol = vol + sig * v * u
        return (G.integrate(vol, dΩ) +
                G.integrate((gam / h) * v * u - v * G.dot(nd, G.gradient(u)) -
                            G.dot(nd, G.gradient(v)) * u, dΓd) +
                G.integrate(3.0 * v * u, dΓr))

    l = lambda v: (G.integrate(1.0 * v, dΩ) + G.integrate(v * g, dΓn) +
                   G.integrate((gam / h) * v * uh - G.dot(nd, G.gradient(v)) * uh, dΓd))
    ops = {path: G.AffineFEOperator(a, l, U, V, G.SparseMatrixAssembler(U, V, precision=precision,
                                                                        path=path))
           for path in ("auto", "general")}
    assert ops["auto"].assem.ctx.pattern_info()["last_kernel"] == G._lib.GH_KERNEL_Q1_LATTICE
    assert ops["general"].assem.ctx.pattern_info()["last_kernel"] == G._lib.GH_KERNEL_CELL
    A, b = ops["auto"].get_matrix(), ops["auto"].get_vector()
    Ag, bg = ops["general"].get_matrix(), ops["general"].get_vector()
    np.testing.assert_array_equal(A.rowval, Ag.rowval)
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert relmax(A.nzval, Ag.nzval) <= tol and relmax(b, bg) <= tol
    S = A.to_scipy()
    assert (S - S.T).count_nonzero() == 0
    if not mass and precision == "fp64":
        U1 = lambda x: ue(np.asarray(x))
        ref = oracle.poisson_q1_with_boundary(
            (0, 1) * 3, part, [25],
            [([22], gam / h, -1.0, lambda x, n: gam / h * U1(x), lambda x, n: -U1(x)),
             ([23], 3.0, 0.0, None, None),
             ([26], 0.0, 0.0, lambda x, n: np.cos(2 * x[1]) + x[2], None)],
            g=U1, kappa=kappa)
        np.testing.assert_array_equal(A.rowval, ref["rowval"])
        assert relmax(A.nzval, ref["nzval"]) <= 1e-12 and relmax(b, ref["b"]) <= 1e-12


UV_EX = lambda x: np.stack([x[..., 0] ** 2 + x[..., 1], 4 * x[..., 0] - x[..., 1] ** 2], axis=-1)
# ∇u[i, j] = ∂_i u_j
GUV_EX = lambda x: np.array([[2 * x[0], 4.0], [1.0, -2 * x[1]]])


@pytest.mark.parametrize("part", [(4, 4), (9, 7)])
def test_poisson_driver_vector_case(oracle, part):
    """PoissonTests.jl's vector case (VectorValue{2}, order 2): u = (x1² + x2, 4x1 - x2²),
    f = -Δu, Neumann and Nitsche terms acting component by component.  Against the
    oracle (pattern bit-exact, 1e-12) and the device PCG reproduces u (it lies in the
    space) to 1e-8 relative."""
    announce("PoissonTests vector case %s" % (part,))
    gam = 10.0
    model = G.CartesianDiscreteModel((0, 1, 0, 1), part)
    labels = model.get_face_labeling()
    labels.add_tag_from_tags("dirichlet", [1, 2, 5])
    labels.add_tag_from_tags("neumann", [7, 8])
    labels.add_tag_from_tags("nitsche", 6)
    h = 1.0 / part[0]
    Ω = G.Triangulation(model)
    Γn = G.BoundaryTriangulation(model, labels, tags="neumann")
    Γd = G.BoundaryTriangulation(model, labels, tags="nitsche")
    dΩ, dΓn, dΓd = G.Measure(Ω, 2), G.Measure(Γn, 2), G.Measure(Γd, 2)
    nn, nd = G.get_normal_vector(Γn), G.get_normal_vector(Γd)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.VectorValue[2, G.Float64], 2),
                      dirichlet_tags="dirichlet", labels=labels)
    U = G.TrialFESpace(V, UV_EX)
    uh = G.interpolate(UV_EX, U)
    f = lambda x: np.stack([np.full(x.shape[:-1], -2.0), np.full(x.shape[:-1], 2.0)], axis=-1)
    a = lambda u, v: (G.integrate(G.inner(G.gradient(v), G.gradient(u)), dΩ) +
                      G.integrate((gam / h) * v * u - v * G.dot(nd, G.gradient(u)) -
                                  G.dot(nd, G.gradient(v)) * u, dΓd))
    l = lambda v: (G.integrate(G.inner(v, f), dΩ) +
                   G.integrate(v * G.dot(nn, G.gradient(uh)), dΓn) +
                   G.integrate((gam / h) * v * uh - G.dot(nd, G.gradient(v)) * uh, dΓd))
    op = G.AffineFEOperator(a, l, U, V)
    import conforming_oracle as CO
    U1 = lambda x: UV_EX(np.asarray(x))
    ref = CO.poisson_qk_vector_with_boundary(
        oracle, (0, 1, 0, 1), part, 2, np.asarray(V.cell_dof_ids), V.nfree, V.ndirichlet,
        U.dirichlet_values(),
        [([7, 8], 0.0, 0.0, lambda x, n: n @ GUV_EX(x), None),
         ([6], gam / h, -1.0, lambda x, n: gam / h * U1(x), lambda x, n: -U1(x))],
        2, 2, f)
    A, b = op.get_matrix(), op.get_vector()
    np.testing.assert_array_equal(A.colptr, ref["colptr"])
    np.testing.assert_array_equal(A.rowval, ref["rowval"])
    assert relmax(A.nzval, ref["nzval"]) <= 1e-12
    assert relmax(b, ref["b"]) <= 1e-12
    u = G.solve(op, G.PCGSolver(rtol=1e-13))
    uI = G.interpolate(UV_EX, U).free_values
    assert np.max(np.abs(u.free_values - uI)) <= 1e-8 * np.max(np.abs(uI))
