"""FP64 CPU restatement of the engine's learned FSAI (TEST INFRASTRUCTURE ONLY).

Gridap has no preconditioner or iterative solver (src/Algebra/LinearSolvers.jl:
249-303 ships LU only), so nothing here restates Gridap; it restates the
engine's own definition on small meshes with dense linear algebra, so the device
kernels can be checked against it:

    G = Ĝ D^(-1/2),  Ĝ_ij = Σ_{c ∋ i,j} y_c[pair],  Ĝ_ii = exp(Σ_{c ∋ i} y_c[i]),
    L = Σ_k ‖z_k - G A Gᵀ z_k‖² / (m N)  (Rademacher z_k from splitmix64),
    y_c = MLP(x_c) with BF16-rounded weights, GELU, 10 -> 64 -> 64 -> 64 -> 36.

The dense gradient and the network back-propagation are pinned by finite
differences in tests/test_fsai_oracle.py.
"""
import math

import numpy as np
from scipy.special import erf

IN, HID, OUT = 10, 64, 36
OFF = dict(W1=0, B1=640, W2=704, B2=4800, W3=4864, B3=8960, W4=9024, B4=11328)
NPARAMS = 11364
MASK64 = (1 << 64) - 1


def bf16_round(w):
    """Round FP32 values to BF16 (nearest even), returned as float64."""
    u = np.asarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    return dict(W1=p[0:640].reshape(64, 10), b1=p[640:704], W2=p[704:4800].reshape(64, 64),
                b2=p[4800:4864], W3=p[4864:8960].reshape(64, 64), b3=p[8960:9024],
                W4=p[9024:11328].reshape(36, 64), b4=p[11328:11364])


def pack(d):
    return np.concatenate([d["W1"].ravel(), d["b1"], d["W2"].ravel(), d["b2"], d["W3"].ravel(),
                           d["b3"], d["W4"].ravel(), d["b4"]])


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)


def mlp_forward(params, X, bf16=True, act_bf16=False):
    """bf16=False evaluates the same network with unrounded weights (for finite
    differences, which the rounding would make piecewise constant).  act_bf16=True
    also rounds every layer input to BF16, as the matrix-core inference forward does
    (its products are exact in FP32 and accumulated in FP32)."""
    P = unpack(bf16_round(params) if bf16 else params)
    r = bf16_round if act_bf16 else (lambda a: a)
    z1 = r(X) @ P["W1"].T + P["b1"]
    h1 = gelu(z1)
    z2 = r(h1) @ P["W2"].T + P["b2"]
    h2 = gelu(z2)
    z3 = r(h2) @ P["W3"].T + P["b3"]
    h3 = gelu(z3)
    y = r(h3) @ P["W4"].T + P["b4"]
    return y, dict(X=X, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, P=P)


def mlp_backward(cache, dy):
    P = cache["P"]
    g = {}
    g["W4"] = dy.T @ cache["h3"]
    g["b4"] = dy.sum(0)
    dz3 = (dy @ P["W4"]) * gelu_d(cache["z3"])
    g["W3"] = dz3.T @ cache["h2"]
    g["b3"] = dz3.sum(0)
    dz2 = (dz3 @ P["W3"]) * gelu_d(cache["z2"])
    g["W2"] = dz2.T @ cache["h1"]
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ P["W2"]) * gelu_d(cache["z1"])
    g["W1"] = dz1.T @ cache["X"]
    g["b1"] = dz1.sum(0)
    return pack(g)


def splitmix64_mix(x):
    """The output function of splitmix64 (Steele, Lea, Flood, OOPSLA 2014)."""
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK64
    return x ^ (x >> 31)


def rademacher(seed, probe, n):
    """splitmix64 of (seed, probe, i) -> ±1, identical to the device generator:
    element i of probe k is the (k*2^32 + k + i + 1)-th output of splitmix64 seeded
    with seed."""
    out = np.empty(n)
    for i in range(n):
        x = splitmix64_mix((seed + 0x9E3779B97F4A7C15 * (probe * 0x100000001 + i + 1)) & MASK64)
        out[i] = 1.0 if (x & 1) else -1.0
    return out


def pair_index(a, b):
    return 8 + a * (a - 1) // 2 + b


class Problem:
    """Q1 Poisson on a Cartesian n0 x n1 x n2 grid of (0,1)^3 with Dirichlet boundary,
    from the assembly oracle; dense A on the free dofs."""

    def __init__(self, oracle, part, kappa):
        self.part = tuple(part)
        r = oracle.poisson_q1((0, 1, 0, 1, 0, 1), part, kappa=kappa)
        self.r = r
        self.kappa = np.asarray(kappa, dtype=np.float64)
        self.N = r["nfree"]
        self.A = oracle.to_scipy(r["colptr"], r["rowval"], r["nzval"], self.N).toarray()
        self.cell_dofs = r["cell_dofs"]
        self.sizes = r["model"]["sizes"]
        # pattern of tril(A): the assembled pattern, explicit zeros kept
        pat = np.zeros((self.N, self.N), dtype=bool)
        cp, rv = r["colptr"], r["rowval"]
        for j in range(self.N):
            pat[rv[cp[j] - 1:cp[j + 1] - 1] - 1, j] = True
        self.lower = np.tril(pat)
        # reference κ=1 element diagonal of the box cell
        K0 = r["cellmats"][0] / self.kappa[0]
        self.kdiag = K0[0, 0]

    def features(self):
        d = np.diag(self.A)
        C = len(self.kappa)
        X = np.zeros((C, IN))
        for c in range(C):
            for a in range(8):
                dof = self.cell_dofs[c, a]
                X[c, a] = self.kappa[c] * self.kdiag / d[dof - 1] if dof > 0 else 0.0
        X[:, 8] = np.log(self.kappa)
        X[:, 9] = math.log(float(np.prod(self.sizes))) / 3.0
        return X

    def ghat(self, y):
        """Dense Ĝ from the cell outputs y [C x 36]."""
        N = self.N
        G = np.zeros((N, N))
        logdiag = np.zeros(N)
        for c, dofs in enumerate(self.cell_dofs):
            for a in range(8):
                if dofs[a] > 0:
                    logdiag[dofs[a] - 1] += y[c, a]
            for a in range(1, 8):
                for b in range(a):
                    i, j = dofs[a], dofs[b]
                    if i > 0 and j > 0:
                        G[i - 1, j - 1] += y[c, pair_index(a, b)]
        G[np.diag_indices(N)] = np.exp(logdiag)
        return G

    def ghat_grad_to_cells(self, gG, Gh):
        """∂L/∂y from ∂L/∂Ĝ (transpose of ghat)."""
        C = len(self.kappa)
        dy = np.zeros((C, OUT))
        for c, dofs in enumerate(self.cell_dofs):
            for a in range(8):
                if dofs[a] > 0:
                    i = dofs[a] - 1
                    dy[c, a] = gG[i, i] * Gh[i, i]
            for a in range(1, 8):
                for b in range(a):
                    i, j = dofs[a], dofs[b]
                    if i > 0 and j > 0:
                        dy[c, pair_index(a, b)] = gG[i - 1, j - 1]
        return dy

    def loss_and_ghat_grad(self, Gh, seed, nprobes):
        N = self.N
        S = np.diag(1.0 / np.sqrt(np.diag(self.A)))
        B = S @ self.A @ S
        L = 0.0
        gG = np.zeros((N, N))
        for k in range(nprobes):
            z = rademacher(seed, k, N)
            u = Gh @ (B @ (Gh.T @ z))
            r = z - u
            L += r @ r
            gG += np.outer(r, B @ (Gh.T @ z)) + np.outer(z, B @ (Gh.T @ r))
        scale = 1.0 / (nprobes * N)
        return L * scale, -2.0 * scale * gG * self.lower

    def loss(self, params, seed, nprobes, bf16=True):
        y, _ = mlp_forward(params, self.features(), bf16)
        return self.loss_and_ghat_grad(self.ghat(y), seed, nprobes)[0]

    def loss_and_grad(self, params, seed, nprobes, bf16=True):
        X = self.features()
        y, cache = mlp_forward(params, X, bf16)
        Gh = self.ghat(y)
        L, gG = self.loss_and_ghat_grad(Gh, seed, nprobes)
        dy = self.ghat_grad_to_cells(gG, Gh)
        return L, mlp_backward(cache, dy), dict(X=X, y=y, Gh=Gh, gG=gG, dy=dy)


def init_params(seed):
    """The device's gh_mlp_init: Xavier-uniform hidden layers from splitmix64,
    zero last layer."""
    w = np.zeros(NPARAMS, dtype=np.float64)
    state = seed

    def uniform():
        nonlocal state
        state = (state + 0x9E3779B97F4A7C15) & MASK64
        x = splitmix64_mix(state)
        return float((x >> 11) * (2.0 / 9007199254740992.0) - 1.0)

    for off, nout, nin in ((0, 64, 10), (704, 64, 64), (4864, 64, 64)):
        a = np.sqrt(np.float32(6.0) / np.float32(nin + nout))
        for k in range(nout * nin):
            w[off + k] = a * np.float32(uniform())
    return w.astype(np.float32)
