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
