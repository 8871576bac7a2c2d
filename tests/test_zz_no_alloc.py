"""gh_pcg_noalloc, the solve that allocates no device memory, against gh_pcg on the
same systems (needs an MI355X).

Both run the same CSR kernels (GRIDAP_HIP_STENCIL_PCG=0 selects them for gh_pcg), so
the bar is bitwise: identical solutions and iteration counts.  Two calls in a row,
the second on a smaller system, share the library's fixed device array; a system
above GH_NOALLOC_MAX_N is refused with GH_ERR_ARG before anything runs.
"""
import sys

import numpy as np
import pytest

import gridap_hip as G
from gridap_hip import _lib

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def _system(part, seed, tags=("tag_21", "tag_22", "tag_23")):
    """3-D Q1 with κ lognormal and Dirichlet data on some faces only (not a box of
    free nodes, so gh_pcg runs its CSR kernels too)."""
    model = G.CartesianDiscreteModel((0, 1) * 3, part)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1),
                      dirichlet_tags=list(tags))
    U = G.TrialFESpace(V, lambda x: 1.0 + x[..., 0] - x[..., 2])
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 2)
    C = int(np.prod(part))
    k = G.CellField(np.exp(np.random.default_rng(seed).normal(size=C)), Ω)
    a = lambda u, v: G.integrate(k * G.inner(G.gradient(v), G.gradient(u)), dΩ)
    l = lambda v: G.integrate(1.0 * v, dΩ)
    return G.AffineFEOperator(a, l, U, V, G.SparseMatrixAssembler(U, V))


def test_noalloc_pcg_matches_pcg(monkeypatch):
    monkeypatch.setenv("GRIDAP_HIP_STENCIL_PCG", "0")
    for part, seed in (((40, 36, 32), 1), ((12, 10, 9), 2)):
        op = _system(part, seed)
        ctx = op.assem.ctx
        assert ctx.pattern_info()["box"] == 0
        for precond in (_lib.GH_PRECOND_JACOBI, _lib.GH_PRECOND_NONE):
            announce("noalloc pcg %s precond %d N=%d" % (part, precond, ctx.nfree))
            x0 = np.random.default_rng(seed).normal(size=ctx.nfree) * 0.1
            xa, sa = ctx.pcg(b=None, x0=x0, precond=precond, rtol=1e-10, maxit=5000)
            xb, sb = ctx.pcg_noalloc(b=None, x0=x0, precond=precond, rtol=1e-10, maxit=5000)
            assert sa.converged == sb.converged == 1
            assert sa.iterations == sb.iterations
            np.testing.assert_array_equal(xa, xb)
            assert sa.rel_residual == sb.rel_residual <= 1e-10
            # an explicit right-hand side
            b = np.random.default_rng(seed + 7).normal(size=ctx.nfree)
            xa, sa = ctx.pcg(b=b, precond=precond, rtol=1e-9, maxit=5000)
            xb, sb = ctx.pcg_noalloc(b=b, precond=precond, rtol=1e-9, maxit=5000)
            assert sa.iterations == sb.iterations
            np.testing.assert_array_equal(xa, xb)


def test_noalloc_pcg_limits():
    announce("noalloc pcg: size limit and FSAI refused")
    big = _system((66, 66, 66), 3, tags=("boundary",))   # 65^3 = 274,625 free dofs
    assert big.assem.ctx.nfree > _lib.GH_NOALLOC_MAX_N
    with pytest.raises(G.GridapHipError) as e:
        big.assem.ctx.pcg_noalloc()
    assert e.value.code == _lib.GH_ERR_ARG
    small = _system((6, 5, 4), 4)
    with pytest.raises(G.GridapHipError) as e:
        small.assem.ctx.pcg_noalloc(precond=_lib.GH_PRECOND_FSAI)
    assert e.value.code == _lib.GH_ERR_UNSUPPORTED
    x, st = small.assem.ctx.pcg_noalloc(rtol=1e-10)
    assert st.converged == 1
