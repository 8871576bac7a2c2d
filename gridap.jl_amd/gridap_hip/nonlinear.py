"""FEOperator (residual / Jacobian at an FE function) and Newton's method.

Gridap's FEOperatorFromWeakForm (src/FESpaces/FEOperatorsFromWeakForm.jl:1-95)
assembles res(uh, v) with assemble_vector! and jac(uh, du, v) with assemble_matrix!
into a pattern allocated once; NewtonRaphsonSolver (src/Algebra/NLSolvers.jl:11-88)
iterates dx = -J⁻¹ r, x += dx until max|r| < tol·max|r₀|.  Here the residual and the
Jacobian are assembled by one device pass (GH_FORM_PLAPLACIAN) into the assembler's
resident pattern, and the linearised systems are solved on the device (PCG: the
p-Laplacian Jacobian is symmetric positive definite for p >= 2).
"""
import numpy as np

from . import _lib
from .assemblers import PCGSolver, SparseMatrixAssembler, SparseMatrixCSC, _space
from .celldata import (PLaplacianJacobianTerm, PLaplacianResidualTerm, SourceTerm,
                       _cell_coef, get_fe_basis, get_trial_fe_basis)
from .fespaces import FEFunction, TrialFESpace
from .reference_fes import Quadrature


def _contrib_terms(out):
    from .celldata import DomainContribution
    if not isinstance(out, DomainContribution):
        raise TypeError("weak form must return integrate(...) contributions")
    return out.terms


class FEOperator:
    """FEOperator(res, jac, U, V[, assem]) or FEOperator(res, U, V[, assem]) (the Jacobian
    of the recognised residual is then the exact derivative, as Gridap's autodiff
    gives), src/FESpaces/FEOperatorsFromWeakForm.jl:24-46."""

    def __init__(self, res, *args):
        if len(args) >= 3 and callable(args[0]):
            jac, U, V = args[:3]
            rest = args[3:]
        else:
            jac = None
            U, V = args[:2]
            rest = args[2:]
        if isinstance(V, TrialFESpace):
            raise TypeError("the test space of an FEOperator must not be a TrialFESpace")
        self.res, self.jac = res, jac
        self.trial, self.test = U, V
        self.assem = rest[0] if rest else SparseMatrixAssembler(U, V)
        self._A = None

    def get_trial(self):
        return self.trial

    def get_test(self):
        return self.test

    # lowering ---------------------------------------------------------------
    def _lower(self, uh):
        V = self.test
        ncells = _space(V).model.num_cells()
        v = get_fe_basis(V)
        terms = _contrib_terms(self.res(uh, v))
        p = kappa = None
        fsrc = np.zeros(ncells)
        degs = set()
        for t, meas in terms:
            degs.add(meas.degree)
            if isinstance(t, PLaplacianResidualTerm):
                if t.state is not uh:
                    raise ValueError("the residual must be evaluated at the given uh")
                coef = _cell_coef(t.coef)
                k = coef.per_cell(ncells) if coef is not None else None
                k = np.ones(ncells) if k is None else k
                if p is not None and p != t.p:
                    raise NotImplementedError("one p-Laplacian exponent per operator")
                p = t.p
                kappa = k if kappa is None else kappa + k
            elif isinstance(t, SourceTerm):
                # res += ∫ c v  <=>  the kernel's -∫ f v with f = -c
                coef = _cell_coef(t.coef)
                c = coef.per_cell(ncells) if coef is not None else None
                fsrc = fsrc - (np.ones(ncells) if c is None else c)
            else:
                raise NotImplementedError("residuals: κ ∇v⋅flux(∇u) and source terms f v")
        if p is None:
            raise NotImplementedError("the residual has no p-Laplacian term")
        if self.jac is not None:
            du = get_trial_fe_basis(self.trial)
            jterms = _contrib_terms(self.jac(uh, du, v))
            kj = None
            for t, meas in jterms:
                degs.add(meas.degree)
                if not isinstance(t, PLaplacianJacobianTerm) or t.p != p:
                    raise NotImplementedError("the Jacobian must be κ ∇v⋅dflux(∇du, ∇u) of the "
                                              "residual's law")
                coef = _cell_coef(t.coef)
                k = coef.per_cell(ncells) if coef is not None else None
                k = np.ones(ncells) if k is None else k
                kj = k if kj is None else kj + k
            if kj is None or not np.array_equal(kj, kappa):
                raise NotImplementedError("the Jacobian's coefficient differs from the "
                                          "residual's")
        if len(degs) != 1:
            raise NotImplementedError("residual and Jacobian terms must share one Measure")
        return p, kappa, fsrc, degs.pop()

    def _assemble(self, uh, A=None, b=None):
        if not isinstance(uh, FEFunction):
            raise TypeError("uh must be an FEFunction of the trial space")
        p, kappa, fsrc, degree = self._lower(uh)
        assem = self.assem
        assem._setup(degree)
        ctx = assem.ctx
        ctx.coef_upload(None if np.all(kappa == 1.0) else kappa,
                        None if not np.any(fsrc) else fsrc, None)
        dv = np.asarray(uh.dirichlet_values, dtype=np.float64)
        ctx.fefunction_upload(np.ascontiguousarray(uh.free_values, dtype=np.float64),
                              dv if dv.size else None)
        ctx.plaplacian_params(p)
        if A is not None and not (np.array_equal(A.colptr, assem.colptr) and
                                  np.array_equal(A.rowval, assem.rowval)):
            raise ValueError("A does not have this assembler's sparsity pattern")
        ctx.assemble(_lib.GH_FORM_PLAPLACIAN, assem.precision, assem.path)
        nz, r = ctx.values_get(None if A is None else A.nzval, b)
        if A is None:
            n = assem.space.nfree
            A = SparseMatrixCSC(n, n, assem.colptr.copy(), assem.rowval.copy(), nz)
        assem.device_matrix = A
        return r, A

    # Gridap's API -------------------------------------------------------------
    def residual_and_jacobian(self, uh):
        """(b, A) = residual_and_jacobian(op, uh) (FEOperatorsFromWeakForm.jl:86-95)."""
        return self._assemble(uh)

    def residual(self, uh):
        return self._assemble(uh)[0]

    def jacobian(self, uh):
        return self._assemble(uh)[1]

    def residual_and_jacobian_inplace(self, b, A, uh):
        """residual_and_jacobian!(b, A, op, uh): b and A's stored values overwritten."""
        return self._assemble(uh, A=A, b=b)

    def residual_inplace(self, b, uh):
        return self._assemble(uh, b=b)[0]

    def jacobian_inplace(self, A, uh):
        return self._assemble(uh, A=A)[1]


def residual(op, uh):
    return op.residual(uh)


def jacobian(op, uh):
    return op.jacobian(uh)


def residual_and_jacobian(op, uh):
    return op.residual_and_jacobian(uh)


class NewtonRaphsonSolver:
    """NewtonRaphsonSolver(ls, tol, max_nliters) (src/Algebra/NLSolvers.jl:11-88): stops
    when max|r| < tol (at the start) or max|r| < tol·max|r₀|; raises after max_nliters
    iterations without convergence.  ls solves J dx = -r on the device."""

    def __init__(self, ls=None, tol=1e-10, max_nliters=20):
        self.ls = ls or PCGSolver(rtol=1e-12)
        self.tol = float(tol)
        self.max_nliters = int(max_nliters)
        self.history = []   # max|r| after each residual evaluation

    def solve(self, x, op):
        """x (free values, updated in place) -> x."""
        U = op.get_trial()
        self.history = []
        b, A = op.residual_and_jacobian(FEFunction(U, x))
        m0 = float(np.max(np.abs(b))) if b.size else 0.0
        self.history.append(m0)
        if m0 < self.tol:
            return x
        for it in range(1, self.max_nliters + 1):
            dx = self.ls.solve(op.assem, -b, A)
            x += dx
            b, A = op.residual_and_jacobian_inplace(b, A, FEFunction(U, x))
            m = float(np.max(np.abs(b)))
            self.history.append(m)
            if m < self.tol * m0:
                return x
        raise RuntimeError("NewtonRaphsonSolver failed to converge in %d iterations. Final "
                           "residual: %g" % (self.max_nliters, self.history[-1]))


class FESolver:
    """FESolver(nls) (src/FESpaces/FESolvers.jl): solve!(uh, solver, op)."""

    def __init__(self, nls):
        self.nls = nls


def solve_inplace(uh, solver, op):
    """uh, cache = solve!(uh, solver, op): Newton from uh's free values (updated)."""
    nls = solver.nls if isinstance(solver, FESolver) else solver
    x = np.array(uh.free_values, dtype=np.float64)
    nls.solve(x, op)
    uh.free_values = x
    return uh, nls
