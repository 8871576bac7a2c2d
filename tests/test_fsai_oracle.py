"""Pins of the learned-FSAI oracle (CPU): the probe generator against published
splitmix64 outputs, BF16 rounding and GELU against torch, the loss against closed
forms, and both gradients against central finite differences."""
import math

import numpy as np
import pytest
import torch

import fsai_oracle as F

GOLDEN = 0x9E3779B97F4A7C15


def test_splitmix64_known_answers():
    # splitmix64 seeded with 1234567: its first five outputs (Rosetta Code,
    # "Pseudo-random numbers/Splitmix64")
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    got = [F.splitmix64_mix((1234567 + GOLDEN * (k + 1)) % 2**64) for k in range(5)]
    assert got == want
    z = F.rademacher(1234567, 0, 5)
    assert list(z) == [1.0 if w & 1 else -1.0 for w in want]


def test_rademacher_probes_are_balanced_and_distinct():
    z0, z1 = F.rademacher(5, 0, 4000), F.rademacher(5, 1, 4000)
    assert set(np.unique(z0)) == {-1.0, 1.0}
    assert abs(z0.mean()) < 0.06 and abs(z1.mean()) < 0.06
    assert abs(z0 @ z1) / 4000 < 0.06


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(0)
    w = np.concatenate([rng.normal(0, 1, 2000), rng.normal(0, 1e-3, 2000),
                        [1.0, 1 + 2**-8, 1 + 3 * 2**-8, -1 - 2**-8, 0.0, 65504.0]]).astype(np.float32)
    ref = torch.from_numpy(w).to(torch.bfloat16).float().double().numpy()
    np.testing.assert_array_equal(F.bf16_round(w), ref)
    # ties go to even
    assert F.bf16_round(np.float32(1 + 2**-8)) == 1.0
    assert F.bf16_round(np.float32(1 + 3 * 2**-8)) == 1 + 2**-6


def test_gelu_and_derivative():
    x = np.linspace(-6, 6, 241)
    t = torch.tensor(x, requires_grad=True)
    y = torch.nn.functional.gelu(t)
    y.sum().backward()
    np.testing.assert_allclose(F.gelu(x), y.detach().numpy(), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(F.gelu_d(x), t.grad.numpy(), rtol=1e-13, atol=1e-15)


@pytest.fixture(scope="module")
def prob(oracle):
    rng = np.random.default_rng(3)
    part = (4, 4, 3)
    kappa = np.exp(rng.normal(0, 1, int(np.prod(part))))
    return F.Problem(oracle, part, kappa)


def random_params(seed, scale4=0.05):
    p = F.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed)
    p[F.OFF["W4"]:] = rng.normal(0, scale4, F.NPARAMS - F.OFF["W4"])
    p[F.OFF["B1"]:F.OFF["W2"]] = rng.normal(0, 0.1, 64)
    return p


def test_features_partition_the_diagonal(prob):
    # A_ii = Σ_{c ∋ i} κ_c K̂_aa and K̂_aa is the same for every corner of the box
    # cell, so each free node's features over its cells sum to one.
    X = prob.features()
    tot = np.zeros(prob.N)
    for c, dofs in enumerate(prob.cell_dofs):
        for a in range(8):
            if dofs[a] > 0:
                tot[dofs[a] - 1] += X[c, a]
    np.testing.assert_allclose(tot, 1.0, rtol=1e-13)
    np.testing.assert_allclose(X[:, 8], np.log(prob.kappa), rtol=1e-15)
    assert np.allclose(X[:, 9], math.log(1 / 4 * 1 / 4 * 1 / 3) / 3)
    Kd = prob.r["cellmats"][0] / prob.kappa[0]
    np.testing.assert_allclose(np.diag(Kd), prob.kdiag, rtol=1e-13)


def test_zero_last_layer_gives_jacobi(prob):
    p = F.init_params(1)
    y, _ = F.mlp_forward(p, prob.features())
    assert np.all(y == 0)
    np.testing.assert_array_equal(prob.ghat(y), np.eye(prob.N))


def test_ghat_structure(prob):
    y, _ = F.mlp_forward(random_params(2), prob.features())
    Gh = prob.ghat(y)
    assert np.all(np.triu(Gh, 1) == 0)
    assert np.all(np.diag(Gh) > 0)
    assert not np.any((Gh != 0) & ~prob.lower)          # inside the pattern of tril(A)


def test_loss_closed_forms(prob):
    S = np.diag(1 / np.sqrt(np.diag(prob.A)))
    B = S @ prob.A @ S
    Linv = np.linalg.inv(np.linalg.cholesky(B))
    L0, _ = prob.loss_and_ghat_grad(Linv, 9, 3)
    assert L0 < 1e-24                                   # G A Gᵀ = I exactly
    # Ĝ = I: E‖(I - B) z‖² = ‖I - B‖_F² for Rademacher z (Hutchinson)
    m = 600
    L1, _ = prob.loss_and_ghat_grad(np.eye(prob.N), 4, m)
    exact = np.linalg.norm(np.eye(prob.N) - B, "fro") ** 2 / prob.N
    assert abs(L1 - exact) < 0.08 * exact


def test_ghat_gradient_finite_difference(prob):
    rng = np.random.default_rng(5)
    Gh = np.where(prob.lower, rng.normal(0, 0.1, (prob.N, prob.N)), 0.0)
    Gh[np.diag_indices(prob.N)] = np.exp(rng.normal(0, 0.2, prob.N))
    L, gG = prob.loss_and_ghat_grad(Gh, 13, 3)
    assert np.all(gG[~prob.lower] == 0)
    for trial in range(3):
        E = np.where(prob.lower, rng.normal(0, 1, (prob.N, prob.N)), 0.0)
        eps = 1e-6
        fd = (prob.loss_and_ghat_grad(Gh + eps * E, 13, 3)[0] -
              prob.loss_and_ghat_grad(Gh - eps * E, 13, 3)[0]) / (2 * eps)
        assert abs(fd - np.sum(gG * E)) < 1e-7 * max(abs(fd), 1e-3)


def test_parameter_gradient_finite_difference(prob):
    p = random_params(4, scale4=0.1)
    L, g, _ = prob.loss_and_grad(p, 21, 2, bf16=False)
    rng = np.random.default_rng(6)
    for sl in ("W1", "W2", "W4", "all"):
        d = np.zeros_like(p)
        if sl == "all":
            d = rng.normal(0, 1, p.size)
        else:
            lo = F.OFF[sl]
            d[lo:lo + 64] = rng.normal(0, 1, 64)
        eps = 1e-6
        fd = (prob.loss(p + eps * d, 21, 2, bf16=False) -
              prob.loss(p - eps * d, 21, 2, bf16=False)) / (2 * eps)
        assert abs(fd - g @ d) < 1e-6 * max(abs(fd), 1e-4), sl


def test_init_params_layout():
    p = F.init_params(7)
    assert p.dtype == np.float32 and p.size == F.NPARAMS
    assert np.all(p[F.OFF["W4"]:] == 0)                 # last layer zero
    for lo, hi, nin, nout in ((0, 640, 10, 64), (704, 4800, 64, 64), (4864, 8960, 64, 64)):
        a = math.sqrt(6 / (nin + nout))
        assert np.all(np.abs(p[lo:hi]) <= a) and np.std(p[lo:hi]) > 0.5 * a / math.sqrt(3)
    for lo, hi in ((640, 704), (4800, 4864), (8960, 9024)):
        assert np.all(p[lo:hi] == 0)                    # biases start at zero
    first = F.splitmix64_mix((7 + GOLDEN) % 2**64)
    u = np.float32((first >> 11) * 2.0 ** -52 - 1.0)
    assert p[0] == np.sqrt(np.float32(6) / np.float32(74)) * u
