"""Parity of the learned-FSAI kernels with the FP64 oracle (needs an MI355X).

Bar (BF16 weights, FP32 accumulation; the oracle in FP64 with the same
BF16-rounded weights): features within 1e-6 relative; training forward (FP32
activations): probe loss within 1e-5 relative, weight gradient within 1e-3
relative L2 per layer; inference forward (matrix cores, layer inputs rounded to
BF16, mirrored by the oracle): network outputs and Ĝ within 1e-4 of max |value| for
99.9% of the entries and within 1e-2 for all (an input that lands on a BF16
rounding boundary may round the other way in FP32 than in FP64).  Probes are the
same splitmix64 Rademacher vectors on both sides.
"""
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import gridap_hip as G
from gridap_hip import _lib
import fsai_oracle as F

pytestmark = pytest.mark.gpu

PART = (6, 5, 4)


def announce(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def make_context(part, kappa):
    model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), part)
    reffe = G.ReferenceFE(G.lagrangian, G.Float64, 1)
    V = G.TestFESpace(model, reffe, dirichlet_tags="boundary")
    q = G.Quadrature(3, 2)
    shp, grad = reffe.tabulate(3, q.coords)
    cd = V.get_cell_dof_ids()
    ptrs = (np.arange(cd.shape[0] + 1, dtype=np.int64) * cd.shape[1] + 1).astype(np.int32)
    ctx = _lib.Context(0)
    ctx.mesh_cartesian(model.origin, model.sizes, model.partition)
    ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, cd.ravel(), ptrs, V.nfree, V.ndirichlet)
    ctx.reffe_tabulation(q.weights, shp, grad)
    ctx.symbolic()
    ctx.coef_upload(kappa=np.asarray(kappa, dtype=np.float64), fsrc=np.ones(len(kappa)))
    ctx.assemble(_lib.GH_FORM_DIFFUSION, _lib.GH_PREC_FP64, _lib.GH_PATH_STRUCTURED)
    return ctx


def random_params(seed, scale4=0.05):
    p = F.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed)
    p[F.OFF["W4"]:] = rng.normal(0, scale4, F.NPARAMS - F.OFF["W4"])
    p[F.OFF["B1"]:F.OFF["W2"]] = rng.normal(0, 0.1, 64)
    return p.astype(np.float32)


@pytest.fixture(scope="module")
def case(oracle):
    announce("fsai fixture: %s cells, lognormal kappa" % (PART,))
    rng = np.random.default_rng(11)
    kappa = np.exp(rng.normal(0, 1, int(np.prod(PART))))
    prob = F.Problem(oracle, PART, kappa)
    ctx = make_context(PART, kappa)
    colptr, rowval = ctx.pattern_get()
    return dict(prob=prob, ctx=ctx, kappa=kappa, colptr=colptr, rowval=rowval)


def dense_factor(ctx, N):
    """Dense G (from the rows of G) and Gᵀ (from the rows of Gᵀ) of the solver's copy."""
    gle, gli, gue, gui = ctx.fsai_factor_get()
    lo = np.zeros((N, N))
    upT = np.zeros((N, N))
    for i in range(N):
        for k in range(16):
            lo[i, gli[i, k]] += gle[i, k]
            upT[i, gui[i, k]] += gue[i, k]
    return lo, upT.T


def test_mlp_init_matches_oracle(case):
    announce("fsai: mlp_init weights")
    ctx = case["ctx"]
    for seed in (0, 7, 2**63 + 5):
        ctx.mlp_init(seed)
        np.testing.assert_array_equal(ctx.mlp_get(), F.init_params(seed))


def test_forward_matches_oracle(case):
    announce("fsai: features, network, factor")
    ctx, prob = case["ctx"], case["prob"]
    p = random_params(3)
    ctx.mlp_set(p)
    ctx.fsai_build()
    got = ctx.fsai_get()
    X = prob.features()
    feat = got["feat"].astype(np.float64)
    np.testing.assert_allclose(feat[:, :10], X, rtol=1e-6, atol=1e-7)
    assert np.all(feat[:, 10:] == 0)
    # the oracle's input: the device's FP32 features (so BF16 rounding sees the same values)
    y, _ = F.mlp_forward(p, feat[:, :10], act_bf16=True)
    yd = got["yout"][:, :36].astype(np.float64)

    def close(a, b):
        e = np.abs(a - b) / np.max(np.abs(b))
        return np.quantile(e, 0.999) <= 1e-4 and e.max() <= 1e-2

    assert close(yd, y)
    S = 1.0 / np.sqrt(np.diag(prob.A))
    Gref = prob.ghat(y) * S[None, :]                  # G = Ĝ D^-1/2
    lo, up = dense_factor(ctx, prob.N)
    mask = prob.lower
    assert np.all(lo[~mask] == 0) and np.all(up[~mask] == 0)
    assert close(lo[mask], Gref[mask])
    np.testing.assert_array_equal(lo, up)             # G and Gᵀ copies are exact transposes


def test_train_step_matches_oracle(case):
    announce("fsai: loss and weight gradient")
    ctx, prob = case["ctx"], case["prob"]
    p = random_params(5, scale4=0.1)
    ctx.mlp_set(p)
    seed, m = 1234, 3
    loss = ctx.fsai_train_step(m, seed, 0.0)          # lr 0: gradient only
    np.testing.assert_array_equal(ctx.mlp_get(), p)   # weights untouched
    L, g, _ = prob.loss_and_grad(p, seed, m)
    assert abs(loss - L) <= 1e-5 * L
    gd = ctx.fsai_get()["grad"].astype(np.float64)
    for name, lo, hi in (("W1", 0, 704), ("W2", 704, 4864), ("W3", 4864, 9024),
                         ("W4", 9024, F.NPARAMS)):
        err = np.linalg.norm(gd[lo:hi] - g[lo:hi]) / np.linalg.norm(g[lo:hi])
        assert err <= 1e-3, (name, err)


def test_factor_is_deterministic(case):
    announce("fsai: determinism")
    ctx = case["ctx"]
    ctx.mlp_set(random_params(6))
    ctx.fsai_build()
    a = ctx.fsai_factor_get()
    l1 = ctx.fsai_train_step(2, 99, 0.0)
    g1 = ctx.fsai_get()
    ctx.fsai_build()
    b = ctx.fsai_factor_get()
    l2 = ctx.fsai_train_step(2, 99, 0.0)
    g2 = ctx.fsai_get()
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
    assert l1 == l2
    for k in ("grad", "gl", "gu", "yout"):
        np.testing.assert_array_equal(g1[k], g2[k])


def test_untrained_network_is_jacobi(case):
    announce("fsai: zero last layer vs Jacobi-PCG")
    ctx, prob = case["ctx"], case["prob"]
    ctx.mlp_init(1)
    ctx.fsai_build()
    lo, up = dense_factor(ctx, prob.N)
    S = (1.0 / np.sqrt(np.diag(prob.A))).astype(np.float32).astype(np.float64)
    np.testing.assert_array_equal(lo, np.diag(S))     # Ĝ = I: G = D^-1/2 (in FP32)
    np.testing.assert_array_equal(up, np.diag(S))
    xj, sj = ctx.pcg(precond=_lib.GH_PRECOND_JACOBI, rtol=1e-10)
    xf, sf = ctx.pcg(precond=_lib.GH_PRECOND_FSAI, rtol=1e-10)
    assert sf.precond_used == _lib.GH_PRECOND_FSAI
    assert abs(sf.iterations - sj.iterations) <= 1
    np.testing.assert_allclose(xf, xj, rtol=1e-8, atol=1e-12 * np.max(np.abs(xj)))


def test_fsai_pcg_solves_the_system(case, oracle):
    announce("fsai: FSAI-PCG solution vs scipy")
    ctx, prob = case["ctx"], case["prob"]
    ctx.mlp_set(random_params(8, scale4=0.05))
    ctx.fsai_build()
    x, st = ctx.pcg(precond=_lib.GH_PRECOND_FSAI, rtol=1e-10)
    assert st.converged == 1 and st.precond_used == _lib.GH_PRECOND_FSAI
    nz, b = ctx.values_get()
    A = oracle.to_scipy(case["colptr"], case["rowval"], nz, prob.N).tocsc()
    xs = spla.spsolve(A, b)
    assert np.max(np.abs(x - xs)) <= 1e-8 * np.max(np.abs(xs))


def test_diverging_factor_falls_back_to_jacobi(case):
    announce("fsai: overflowing factor -> Jacobi fallback")
    ctx = case["ctx"]
    p = F.init_params(2)
    p[F.OFF["B4"]:F.OFF["B4"] + 8] = 100.0            # Ĝ_ii = exp(800) = inf in FP32
    ctx.mlp_set(p)
    x, st = ctx.pcg(precond=_lib.GH_PRECOND_FSAI, rtol=1e-10)
    assert st.converged == 1 and st.precond_used == _lib.GH_PRECOND_JACOBI
    xj, _ = ctx.pcg(precond=_lib.GH_PRECOND_JACOBI, rtol=1e-10)
    np.testing.assert_allclose(x, xj, rtol=1e-9, atol=1e-14)


def test_training_lowers_the_loss(case):
    announce("fsai: 30 Adam steps")
    ctx = case["ctx"]
    ctx.mlp_init(4)
    first = ctx.fsai_train_step(4, 77, 0.0)
    for k in range(30):
        ctx.fsai_train_step(4, 1000 + k, 3e-3)
    last = ctx.fsai_train_step(4, 77, 0.0)
    assert last < 0.9 * first, (first, last)
    ctx.fsai_build()
    x, st = ctx.pcg(precond=_lib.GH_PRECOND_FSAI, rtol=1e-10)
    assert st.converged == 1 and st.precond_used == _lib.GH_PRECOND_FSAI


def test_grf_kappa_matches_formula(case):
    announce("fsai: random-field coefficient")
    from gridap_hip.fsai import grf_modes
    ctx = make_context(PART, np.ones(int(np.prod(PART))))
    omega, phase, amp = grf_modes(3, nmodes=16, corr_len=0.3)
    ctx.grf_kappa(omega, phase, amp)
    ctx.assemble(_lib.GH_FORM_DIFFUSION, _lib.GH_PREC_FP64, _lib.GH_PATH_STRUCTURED)
    ctx.mlp_init(0)
    ctx.fsai_build()
    logk = ctx.fsai_get()["feat"][:, 8].astype(np.float64)
    h = 1.0 / np.array(PART)
    c = np.arange(int(np.prod(PART)))
    xc = np.stack([(c % PART[0] + 0.5) * h[0], ((c // PART[0]) % PART[1] + 0.5) * h[1],
                   (c // (PART[0] * PART[1]) + 0.5) * h[2]], axis=1)
    want = amp * np.cos(xc @ omega.T + phase).sum(axis=1)
    np.testing.assert_allclose(logk, want, rtol=1e-6, atol=1e-6)


def test_trained_network_beats_jacobi_on_heldout():
    # The committed weights (python -m gridap_hip.fsai: 512 training instances) against
    # the same network untrained (zero last layer == Jacobi) on held-out instances of
    # the 64^3 κ = exp(GRF) problem: fewer PCG iterations to 1e-8, same solution.
    announce("fsai: committed weights vs Jacobi on held-out 64^3 instances")
    from gridap_hip import fsai as FS
    inst = FS.PoissonInstances(64)
    net = FS.LearnedFSAI.load(FS.DEFAULT_WEIGHTS)
    for seed in list(FS.HELDOUT_SEEDS)[:2]:
        inst.load(seed)
        xj, sj = inst.solve("jacobi")
        net.attach(inst.ctx)
        inst.ctx.fsai_build()
        xf, sf = inst.solve("fsai")
        assert sf.converged == 1 and sf.precond_used == _lib.GH_PRECOND_FSAI
        assert sf.iterations <= 0.7 * sj.iterations, (sf.iterations, sj.iterations)
        assert np.max(np.abs(xf - xj)) <= 1e-6 * np.max(np.abs(xj))
        inst.ctx.mlp_init(0)                          # untrained: exactly Jacobi
        inst.ctx.fsai_build()
        _, su = inst.solve("fsai")
        assert abs(su.iterations - sj.iterations) <= 8
    inst.ctx.close()


def test_stencil_pcg_matches_general_pcg(case, monkeypatch):
    """On a box lattice PCG runs stencil kernels (A as 27 offset arrays, G as 14, Gᵀ from
    mirrored offsets); GRIDAP_HIP_STENCIL_PCG=0 runs the general CSR / 16-slot kernels on
    the same system.  Both must reach the same solution with the same iteration count
    (up to rounding: the sums run in another order)."""
    announce("fsai: stencil PCG vs general PCG (Jacobi and FSAI)")
    ctx = case["ctx"]
    ctx.mlp_set(random_params(9, scale4=0.05))
    ctx.fsai_build()
    for precond in (_lib.GH_PRECOND_JACOBI, _lib.GH_PRECOND_FSAI):
        monkeypatch.setenv("GRIDAP_HIP_STENCIL_PCG", "1")
        xs, ss = ctx.pcg(precond=precond, rtol=1e-10)
        monkeypatch.setenv("GRIDAP_HIP_STENCIL_PCG", "0")
        xg, sg = ctx.pcg(precond=precond, rtol=1e-10)
        assert ss.converged == 1 and sg.converged == 1
        assert ss.precond_used == precond and sg.precond_used == precond
        assert abs(ss.iterations - sg.iterations) <= 1
        np.testing.assert_allclose(xs, xg, rtol=1e-8, atol=1e-12 * np.max(np.abs(xg)))


def test_context_reuse_drops_box_state(oracle):
    """One context, two spaces in turn: the first has its free dofs on a box of lattice
    nodes (stencil PCG), the second does not (a single Dirichlet vertex).  Nothing of the
    first pattern's box form may reach the second solve: both match scipy."""
    part = (5, 4, 3)
    announce("context reuse: box space, then a non-box space, PCG vs scipy")
    model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), part)
    reffe = G.ReferenceFE(G.lagrangian, G.Float64, 1)
    q = G.Quadrature(3, 2)
    shp, grad = reffe.tabulate(3, q.coords)
    C = int(np.prod(part))
    kappa = np.exp(np.random.default_rng(5).normal(0, 1, C))
    ctx = _lib.Context(0)
    ctx.mesh_cartesian(model.origin, model.sizes, model.partition)
    for tags in ("boundary", ["tag_01"]):
        V = G.TestFESpace(model, reffe, dirichlet_tags=tags)
        cd = V.get_cell_dof_ids()
        ptrs = (np.arange(cd.shape[0] + 1, dtype=np.int64) * cd.shape[1] + 1).astype(np.int32)
        ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, cd.ravel(), ptrs, V.nfree, V.ndirichlet)
        ctx.reffe_tabulation(q.weights, shp, grad)
        ctx.symbolic()
        ctx.coef_upload(kappa=kappa, fsrc=np.ones(C))
        ctx.assemble(_lib.GH_FORM_DIFFUSION, _lib.GH_PREC_FP64, _lib.GH_PATH_AUTO)
        x, st = ctx.pcg(precond=_lib.GH_PRECOND_JACOBI, rtol=1e-10)
        assert st.converged == 1
        colptr, rowval = ctx.pattern_get()
        nz, b = ctx.values_get()
        A = oracle.to_scipy(colptr, rowval, nz, V.nfree).tocsc()
        xs = spla.spsolve(A, b)
        assert np.max(np.abs(x - xs)) <= 1e-8 * np.max(np.abs(xs))
    ctx.close()
