"""Pins of the n-cube isoparametric restatement in the oracle
(conforming_oracle.physical_gradients, ncube_diffusion_cells, ncube_elasticity_cells,
ncube_loads), which checks the general n-cube element-list path:

* on Cartesian cells it equals the C oracle's Cartesian diffusion matrices and the box
  elasticity reference (itself pinned by rigid motions and quadratic reproduction);
* on mapped cells its diffusion matrices equal the C oracle's isoparametric ones (Q1
  directly; Q2 with the Q2 nodes placed by the Q1 map, the same geometry);
* on mapped cells the elasticity matrices hold the rigid motions in their kernel
  (they lie in every Q_k space on a Q1 map, and ε of them vanishes pointwise), and
  their block trace is (λ + (D+1)μ) times the diffusion matrix — a transposed gradient
  pair or a dropped μ term breaks one of the two;
* the loads sum to f times the cell volume (partition of unity).
"""
import numpy as np
import pytest

import conforming_oracle as CO


def warp(x):
    return x + 0.06 * np.prod(np.sin(np.pi * x), axis=-1)[..., None]


def cells_of(oracle, D, part, mapped):
    dom = tuple(v for _ in range(D) for v in (0.0, 1.0))
    m = oracle.cartesian_model(dom, part)
    xyz = warp(m["xyz"]) if mapped else m["xyz"]
    cn = m["cell_nodes"]
    return m, xyz, cn, xyz[cn - 1]


@pytest.mark.parametrize("D,order", [(2, 1), (3, 1), (2, 2), (3, 2)])
def test_cartesian_cells_match_references(oracle, D, order):
    h = np.array([0.3, 0.5, 0.7])[:D]
    X = np.array([[(a >> d) & 1 for d in range(D)] for a in range(2 ** D)], float) * h
    qx, qw = oracle.tensor_quadrature(D, 2 * order)
    shp, grad = oracle.qk_tabulate(D, order, qx)
    _, g1 = oracle.qk_tabulate(D, 1, qx)
    wdet, g = CO.physical_gradients(X[None], qw, g1, grad)
    Kd = CO.ncube_diffusion_cells(wdet, g, [2.5])
    Kc = oracle.cell_matrices_diffusion(D, np.zeros((1, grad.shape[1]), np.int32), None, qw,
                                        grad, kappa=np.array([2.5]), cartesian_sizes=h)
    np.testing.assert_allclose(Kd, Kc, rtol=0, atol=1e-13 * np.abs(Kc).max())
    Ke = CO.ncube_elasticity_cells(wdet, g, 0.3, [4.0])[0]
    Kr, m = CO.elasticity_reference(oracle, h, 0.3, qdeg=2 * order, order=order)
    np.testing.assert_allclose(Ke, 4.0 * Kr, rtol=0, atol=1e-13 * np.abs(Kr).max())
    F = CO.ncube_loads(wdet, shp, np.array([[1.5] * D]))[0]
    np.testing.assert_allclose(F, 1.5 * np.tile(m, D), rtol=0, atol=1e-15)


@pytest.mark.parametrize("D,order", [(2, 1), (3, 1), (2, 2), (3, 2)])
def test_mapped_diffusion_matches_c_oracle(oracle, D, order):
    part = (3, 2, 2)[:D]
    m, xyz, cn, X = cells_of(oracle, D, part, True)
    qx, qw = oracle.tensor_quadrature(D, 2 * order)
    shp, grad = oracle.qk_tabulate(D, order, qx)
    _, g1 = oracle.qk_tabulate(D, 1, qx)
    wdet, g = CO.physical_gradients(X, qw, g1, grad)
    kappa = np.exp(np.random.default_rng(D).normal(size=len(cn)))
    K = CO.ncube_diffusion_cells(wdet, g, kappa)
    # the C oracle maps each cell with its FE basis through the FE nodes: place the
    # nodes with the Q1 map, which the Q_k basis reproduces
    n1, _ = oracle.qk_tabulate(D, 1, oracle.qk_nodes(D, order))
    P = CO.q1_map_points(X, n1)                          # [C, nn, D]
    nn = P.shape[1]
    cnodes = (np.arange(len(cn) * nn, dtype=np.int32).reshape(len(cn), nn) + 1)
    Kc = oracle.cell_matrices_diffusion(D, cnodes, P.reshape(-1, D), qw, grad, kappa=kappa)
    np.testing.assert_allclose(K, Kc, rtol=0, atol=1e-12 * np.abs(Kc).max())
    # loads: Σ_a f ∫φ_a = f |cell|
    F = CO.ncube_loads(wdet, shp, np.full((len(cn), 1), 2.0))
    np.testing.assert_allclose(F.sum(1), 2.0 * wdet.sum(1), rtol=1e-13)


@pytest.mark.parametrize("D,order", [(2, 1), (3, 1), (2, 2), (3, 2)])
def test_mapped_elasticity_rigid_motions_and_trace(oracle, D, order):
    part = (3, 2, 2)[:D]
    m, xyz, cn, X = cells_of(oracle, D, part, True)
    qx, qw = oracle.tensor_quadrature(D, 2 * order)
    _, grad = oracle.qk_tabulate(D, order, qx)
    _, g1 = oracle.qk_tabulate(D, 1, qx)
    wdet, g = CO.physical_gradients(X, qw, g1, grad)
    nu = 0.27
    K = CO.ncube_elasticity_cells(wdet, g, nu)
    Kd = CO.ncube_diffusion_cells(wdet, g)
    lam, mu = CO.lame(1.0, nu)
    nn = g.shape[2]
    tr = sum(K[:, i * nn:(i + 1) * nn, i * nn:(i + 1) * nn] for i in range(D))
    np.testing.assert_allclose(tr, (lam + (D + 1) * mu) * Kd, rtol=0,
                               atol=1e-12 * np.abs(Kd).max())
    n1, _ = oracle.qk_tabulate(D, 1, oracle.qk_nodes(D, order))
    P = CO.q1_map_points(X, n1)                          # node coordinates per cell
    modes = []
    for i in range(D):                                   # translations
        r = np.zeros((len(cn), D * nn))
        r[:, i * nn:(i + 1) * nn] = 1.0
        modes.append(r)
    for i in range(D):                                   # rotations in the (i, j) planes
        for j in range(i + 1, D):
            r = np.zeros((len(cn), D * nn))
            r[:, i * nn:(i + 1) * nn] = -P[:, :, j]
            r[:, j * nn:(j + 1) * nn] = P[:, :, i]
            modes.append(r)
    for r in modes:
        Kr = np.einsum("cab,cb->ca", K, r)
        assert np.abs(Kr).max() <= 1e-12 * np.abs(K).max() * np.abs(r).max()
    # a non-rigid motion (stretch along x) is not in the kernel
    r = np.zeros((len(cn), D * nn))
    r[:, :nn] = P[:, :, 0]
    assert np.abs(np.einsum("cab,cb->ca", K, r)).max() > 1e-3 * np.abs(K).max()


@pytest.mark.parametrize("D", [2, 3])
def test_mixed_cells_pins(oracle, D):
    """mixed_cells (the cell matrices of a rectangular field pair): with one space for
    test and trial it is the square diffusion + mass matrix; with a Q2 test space and a
    Q1 trial space, every Q1 basis function is its Q2 interpolant (Q1 ⊂ Q2), so the
    block equals the square Q2 matrix times the interpolation matrix N[a, b] =
    φ^Q1_b(x^Q2_a) — a transposed operand or a swapped index fails it."""
    part = (3, 2, 2)[:D]
    m, xyz, cn, X = cells_of(oracle, D, part, True)
    qx, qw = oracle.tensor_quadrature(D, 4)
    s1, g1 = oracle.qk_tabulate(D, 1, qx)
    s2, g2 = oracle.qk_tabulate(D, 2, qx)
    wdet, gp1 = CO.physical_gradients(X, qw, g1, g1)
    _, gp2 = CO.physical_gradients(X, qw, g1, g2)
    rng = np.random.default_rng(D)
    kappa, sigma = np.exp(rng.normal(size=len(cn))), np.exp(rng.normal(size=len(cn)))
    sq = CO.mixed_cells(wdet, gp1, gp1, s1, s1, kappa, sigma)
    ref = CO.ncube_diffusion_cells(wdet, gp1, kappa) + \
        sigma[:, None, None] * np.einsum("cq,qa,qb->cab", wdet, s1, s1)
    np.testing.assert_allclose(sq, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    B = CO.mixed_cells(wdet, gp2, gp1, s2, s1, kappa, sigma)          # [C, n1, n2]
    K2 = CO.mixed_cells(wdet, gp2, gp2, s2, s2, kappa, sigma)         # [C, n2, n2]
    N, _ = oracle.qk_tabulate(D, 1, oracle.qk_nodes(D, 2))           # [n2 nodes, n1]
    # φ^Q1_b = Σ_a N[a, b] ψ^Q2_a, so B[c, b, li] = Σ_a N[a, b] K2[c, a, li]
    np.testing.assert_allclose(B, np.einsum("ab,cal->cbl", N, K2), rtol=0,
                               atol=1e-12 * np.abs(K2).max())
    assert B.shape == (len(cn), 2 ** D, 3 ** D)


@pytest.mark.parametrize("D", [2, 3])
def test_divergence_and_vector_laplacian_pins(oracle, D):
    """divergence_cells and vector_laplacian_cells (the Taylor–Hood blocks):
    * q (∇⋅u) applied to the Q2 interpolant of a linear vector field u = A x + c equals
      tr(A) ∫ ψ_a (∇⋅u is the constant tr A) — a transposed gradient or a wrong
      component offset fails it;
    * the (∇⋅v) p block is the transpose of the q (∇⋅u) block of the same two bases;
    * the vector Laplacian is δ_ij times the scalar diffusion matrix, and constant
      vector fields lie in its kernel."""
    part = (3, 2, 2)[:D]
    m, xyz, cn, X = cells_of(oracle, D, part, True)
    qx, qw = oracle.tensor_quadrature(D, 4)
    s1, g1 = oracle.qk_tabulate(D, 1, qx)
    s2, g2 = oracle.qk_tabulate(D, 2, qx)
    wdet, gp2 = CO.physical_gradients(X, qw, g1, g2)
    C, nn = len(cn), s2.shape[1]
    kappa = np.exp(np.random.default_rng(D).normal(size=C))
    B = CO.divergence_cells(wdet, s1, gp2, scalar_test=True, kappa=kappa)   # [C, D nn, n1]
    A = np.random.default_rng(D + 1).normal(size=(D, D))
    c0 = np.arange(1.0, D + 1)
    n1, _ = oracle.qk_tabulate(D, 1, oracle.qk_nodes(D, 2))
    P = CO.q1_map_points(X, n1)                                  # Q2 nodes [C, nn, D]
    U = np.einsum("ij,cbj->cib", A, P) + c0[None, :, None]       # [C, D, nn]
    Bu = np.einsum("cka,ck->ca", B, U.reshape(C, D * nn))
    ref = kappa[:, None] * np.trace(A) * np.einsum("cq,qa->ca", wdet, s1)
    np.testing.assert_allclose(Bu, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    Bt = CO.divergence_cells(wdet, s1, gp2, scalar_test=False, kappa=kappa)
    np.testing.assert_array_equal(Bt, B.transpose(0, 2, 1))
    L = CO.vector_laplacian_cells(wdet, gp2)
    K = CO.ncube_diffusion_cells(wdet, gp2)
    for i in range(D):
        for j in range(D):
            blk = L[:, i * nn:(i + 1) * nn, j * nn:(j + 1) * nn]
            np.testing.assert_array_equal(blk, K if i == j else 0 * K)
    e = np.tile(np.repeat(np.arange(1.0, D + 1), nn), (C, 1))
    assert np.abs(np.einsum("cab,cb->ca", L, e)).max() <= 1e-12 * np.abs(L).max()


def _solve(cp, rv, nz, b, n):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    return spla.spsolve(sp.csc_matrix((nz, np.asarray(rv) - 1, np.asarray(cp) - 1),
                                      shape=(n, n)), b)


def test_vector_poisson_oracle_reproduces_u(oracle):
    """poisson_qk_vector_with_boundary (PoissonTests.jl's vector case): its system's solve
    is the Q2 interpolant of u = (x1² + x2, 4x1 - x2²) (u lies in the space), with the
    Neumann and Nitsche terms as the driver writes them — a dropped face term, a wrong
    normal sign or a component mix-up breaks it."""
    import gridap_hip as G
    uv = lambda x: np.stack([x[..., 0] ** 2 + x[..., 1], 4 * x[..., 0] - x[..., 1] ** 2], -1)
    guv = lambda x: np.array([[2 * x[0], 4.0], [1.0, -2 * x[1]]])
    f = lambda x: np.stack([np.full(x.shape[:-1], -2.0), np.full(x.shape[:-1], 2.0)], -1)
    part, gam = (5, 4), 10.0
    h = 1.0 / part[0]
    model = G.CartesianDiscreteModel((0, 1, 0, 1), part)
    labels = model.get_face_labeling()
    labels.add_tag_from_tags("dirichlet", [1, 2, 5])
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.VectorValue[2, G.Float64], 2),
                      dirichlet_tags="dirichlet", labels=labels)
    U = G.TrialFESpace(V, uv)
    u1 = lambda x: uv(np.asarray(x))
    ref = CO.poisson_qk_vector_with_boundary(
        oracle, (0, 1, 0, 1), part, 2, np.asarray(V.cell_dof_ids), V.nfree, V.ndirichlet,
        U.dirichlet_values(),
        [([7, 8], 0.0, 0.0, lambda x, n: n @ guv(x), None),
         ([6], gam / h, -1.0, lambda x, n: gam / h * u1(x), lambda x, n: -u1(x))], 2, 2, f)
    x = _solve(ref["colptr"], ref["rowval"], ref["nzval"], ref["b"], V.nfree)
    uI = G.interpolate(uv, U).free_values
    assert np.abs(x - uI).max() <= 1e-12 * np.abs(uI).max()


def test_stokes_oracle_reproduces_u_and_p(oracle):
    """The Taylor–Hood blocks (vector_laplacian_cells, divergence_cells) assembled by the
    multi-field oracle for StokesTaylorHoodTests.jl's u and p (velocity Dirichlet data
    on the boundary, pressure fixed at a vertex): the solve is the interpolant of u and
    p — a sign flip in either coupling or a transposed block breaks it."""
    import gridap_hip as G
    import multifield_oracle as MF
    model = G.CartesianDiscreteModel((0, 2, 0, 2), (3, 3))
    ue = lambda x: np.stack([x[..., 0] ** 2 + 2 * x[..., 1] ** 2, -x[..., 0] ** 2], axis=-1)
    pe = lambda x: x[..., 0] + 3 * x[..., 1]
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.VectorValue[2, G.Float64], 2),
                      dirichlet_tags="boundary")
    Q = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1), dirichlet_tags="tag_1")
    U, P = G.TrialFESpace(V, ue), G.TrialFESpace(Q, pe)
    X = model.node_coordinates()[model.cell_node_ids() - 1]
    qx, qw = oracle.tensor_quadrature(2, 2)
    s1, g1 = oracle.qk_tabulate(2, 1, qx)
    s2, g2 = oracle.qk_tabulate(2, 2, qx)
    wdet, gp2 = CO.physical_gradients(X, qw, g1, g2)
    C = X.shape[0]
    mats = {(0, 0): CO.vector_laplacian_cells(wdet, gp2),
            (1, 0): CO.divergence_cells(wdet, s1, gp2, scalar_test=True),
            (0, 1): -CO.divergence_cells(wdet, s1, gp2, scalar_test=False)}
    xq = CO.q1_map_points(X, s1)
    vecs = {0: CO.ncube_loads(wdet, s2, np.tile([-5.0, 5.0], (C, 1))),
            1: np.einsum("cq,cq,qa->ca", wdet, 2.0 * xq[..., 0], s1)}
    cds = [np.asarray(V.cell_dof_ids), np.asarray(Q.cell_dof_ids)]
    cp, rv, nz, b = MF.assemble(cds, [V.nfree, Q.nfree], [V.ndirichlet, Q.ndirichlet], mats,
                                vecs, dvals=[U.dirichlet_values(), P.dirichlet_values()])
    x = _solve(cp, rv, nz, b, V.nfree + Q.nfree)
    uI, pI = G.interpolate(ue, U).free_values, G.interpolate(pe, P).free_values
    assert np.abs(x[:V.nfree] - uI).max() <= 1e-11 * np.abs(uI).max()
    assert np.abs(x[V.nfree:] - pI).max() <= 1e-10 * np.abs(pI).max()
