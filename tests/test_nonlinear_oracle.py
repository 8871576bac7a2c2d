"""Pins of the oracle's p-Laplacian residual and Jacobian (no GPU).

plaplacian_cells restates res/jac of test/GridapTests/PLaplacianTests.jl:15-17,31-32
(κ-scaled, with a source).  What fixes it independently of its own formula:
* p = 2 is linear diffusion: the Jacobian equals the C oracle's diffusion cell matrices
  and the residual equals K u - b;
* the Jacobian is the derivative of the residual: central finite differences agree
  (and a Jacobian with the (p-2) term dropped does not, so the check has teeth);
* the Jacobian is symmetric;
* Newton with these arrays reproduces PLaplacianTests.jl's manufactured solution
  u = x + y (f = 0), which Q1 holds exactly, with quadratic convergence.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla


def _mapped(oracle, domain, part, seed, amp=0.15):
    m = oracle.cartesian_model(domain, part)
    xyz = m["xyz"].copy()
    D = len(part)
    lo = np.array(domain[0::2], dtype=float)
    hi = np.array(domain[1::2], dtype=float)
    h = (hi - lo) / np.array(part)
    inner = np.all((xyz > lo + 1e-12) & (xyz < hi - 1e-12), axis=1)
    rng = np.random.default_rng(seed)
    xyz[inner] += rng.uniform(-amp, amp, (inner.sum(), D)) * h
    m["xyz"] = xyz
    return m


def _setup(oracle, D, part, seed, mapped=True):
    domain = (0, 1) * D
    ref = oracle.poisson_q1(domain, part)
    m = _mapped(oracle, domain, part, seed) if mapped else ref["model"]
    return ref, m


@pytest.mark.parametrize("D,part", [(2, (5, 4)), (3, (3, 3, 2))])
def test_p2_is_linear_diffusion(oracle, D, part):
    ref, m = _setup(oracle, D, part, 1)
    rng = np.random.default_rng(2)
    C = int(np.prod(part))
    kappa = np.exp(rng.normal(size=C))
    fsrc = rng.uniform(-1, 1, C)
    uc = rng.normal(size=ref["cell_dofs"].shape)
    K, r = oracle.plaplacian_cells(D, m["cell_nodes"], m["xyz"], ref["qw"], ref["shp"],
                                   ref["grad"], uc, 2.0, kappa, fsrc)
    Kd = oracle.cell_matrices_diffusion(D, m["cell_nodes"], m["xyz"], ref["qw"], ref["grad"],
                                        kappa)
    bd = oracle.cell_vectors_source(D, m["cell_nodes"], m["xyz"], ref["qw"], ref["shp"],
                                    ref["grad"], fsrc)
    np.testing.assert_allclose(K, Kd, atol=1e-14 * np.abs(Kd).max())
    # r = K u - b (K stored [c, lj, li])
    np.testing.assert_allclose(r, np.einsum("cji,cj->ci", Kd, uc) - bd, atol=1e-13)


def _fd_jacobian(oracle, D, m, ref, uc, p, kappa, eps=1e-6):
    nl = uc.shape[1]
    Kfd = np.zeros((uc.shape[0], nl, nl))
    for b in range(nl):
        up, um = uc.copy(), uc.copy()
        up[:, b] += eps
        um[:, b] -= eps
        _, rp = oracle.plaplacian_cells(D, m["cell_nodes"], m["xyz"], ref["qw"], ref["shp"],
                                        ref["grad"], up, p, kappa)
        _, rm = oracle.plaplacian_cells(D, m["cell_nodes"], m["xyz"], ref["qw"], ref["shp"],
                                        ref["grad"], um, p, kappa)
        Kfd[:, b, :] = (rp - rm) / (2 * eps)          # [c, lj = b, li]
    return Kfd


@pytest.mark.parametrize("D,part,p", [(2, (4, 3), 3.0), (2, (4, 3), 4.5), (3, (2, 3, 2), 3.0)])
def test_jacobian_is_the_residual_derivative(oracle, D, part, p):
    ref, m = _setup(oracle, D, part, 3)
    rng = np.random.default_rng(4)
    C = int(np.prod(part))
    kappa = np.exp(rng.normal(size=C))
    uc = rng.normal(size=ref["cell_dofs"].shape)
    K, _ = oracle.plaplacian_cells(D, m["cell_nodes"], m["xyz"], ref["qw"], ref["shp"],
                                   ref["grad"], uc, p, kappa)
    Kfd = _fd_jacobian(oracle, D, m, ref, uc, p, kappa)
    scale = np.abs(K).max()
    assert np.abs(K - Kfd).max() <= 1e-7 * scale
    np.testing.assert_allclose(K, K.transpose(0, 2, 1), atol=1e-14 * scale)
    # a Jacobian without the (p-2)|∇u|^(p-4)(∇u⋅∇du)∇u term fails the same check
    K_missing = _first_term_only(D, m, ref, uc, p, kappa)
    assert np.abs(K_missing - Kfd).max() > 1e-3 * scale


def _first_term_only(D, m, ref, uc, p, kappa):
    """The planted mistake: Σ_q w|J| κ |∇u|^(p-2) ∇φ_a⋅∇φ_b, the second term of dflux
    dropped ([c, lj, li])."""
    cn = np.asarray(m["cell_nodes"]) - 1
    qw = np.asarray(ref["qw"])
    nl = cn.shape[1]
    grad = np.asarray(ref["grad"]).reshape(len(qw), nl, D)
    X = np.asarray(m["xyz"])[cn]
    Jt = np.einsum("qai,caj->cqij", grad, X)
    Ji = np.linalg.inv(Jt)
    dV = np.abs(np.linalg.det(Jt)) * qw[None, :]
    g = np.einsum("cqrk,qak->cqar", Ji, grad)
    gu = np.einsum("ca,cqar->cqr", uc, g)
    s0 = np.sqrt(np.sum(gu * gu, -1)) ** (p - 2)
    K1 = np.einsum("cq,cqar,cqbr->cab", dV * kappa[:, None] * s0, g, g)
    return K1.transpose(0, 2, 1)


def _newton_oracle(oracle, D, part, p, kappa=None, fsrc=None, g=None, x0=None, tol=1e-12,
                   maxit=30):
    """Newton (NewtonRaphsonSolver, src/Algebra/NLSolvers.jl:47-77) on the oracle's
    arrays with a direct solve; returns the iterates' max-residual history and x."""
    ref = oracle.poisson_q1((0, 1) * D, part, g=g)
    m = ref["model"]
    cd = ref["cell_dofs"]
    nfree = ref["nfree"]
    x = np.zeros(nfree) if x0 is None else x0.copy()
    hist = []
    for it in range(maxit + 1):
        uc = oracle.cell_values(cd, x, ref["dirichlet_vals"])
        K, r = oracle.plaplacian_cells(D, m["cell_nodes"], m["xyz"], ref["qw"], ref["shp"],
                                       ref["grad"], uc, p, kappa, fsrc,
                                       cartesian_sizes=m["sizes"])
        cp, rv, nz, b = oracle.plaplacian_system(cd, nfree, K, r)
        hist.append(np.abs(b).max())
        if it == 0 and hist[0] < tol:
            break
        if it > 0 and hist[-1] < tol * hist[0]:
            break
        J = oracle.to_scipy(cp, rv, nz, nfree)
        x = x + spla.spsolve(J.tocsc(), -b)
    return ref, x, hist


def test_newton_reproduces_the_plaplacian_driver(oracle):
    """PLaplacianTests.jl:8-47 with Q1: p = 3, u = x + y on the boundary, f = 0, random
    initial guess; the discrete solution is the interpolant of u."""
    p = 3.0
    g = lambda x: x[0] + x[1]
    ref0 = oracle.poisson_q1((0, 1, 0, 1), (4, 4))
    x0 = np.random.default_rng(0).uniform(0, 1, ref0["nfree"])
    ref, x, hist = _newton_oracle(oracle, 2, (4, 4), p, g=g, x0=x0, tol=1e-13)
    xyz = ref["model"]["xyz"]
    n2d = ref["node_to_dof"][:, 0]
    free_nodes = np.nonzero(n2d > 0)[0]
    uI = xyz[free_nodes, 0] + xyz[free_nodes, 1]
    assert np.max(np.abs(x[n2d[free_nodes] - 1] - uI)) <= 1e-10
    # quadratic convergence at the end: r_{k+1} <= C r_k^2
    tail = [h for h in hist if h > 1e-14]
    assert len(hist) <= 12
    assert tail[-1] <= 10.0 * tail[-2] ** 2 / max(tail[0], 1.0) + 1e-13
