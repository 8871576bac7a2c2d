"""Weak-form expressions: CellField, gradient, inner, Measure, integrate.

Gridap builds lazy cell arrays from `∫(κ*∇(v)⋅∇(u))dΩ`
(src/CellData/DomainContributions.jl:188-200, src/CellData/CellQuadratures.jl:
138-170).  Here the same expressions are recorded symbolically and lowered to a
form kind plus per-cell coefficients, which is what the device kernels consume:

    a(u,v) = integrate(κ * inner(gradient(v), gradient(u)), dΩ)   -> diffusion, κ
    a(u,v) = integrate(σ * u * v, dΩ)                             -> mass, σ (scalar Q1)
    l(v)   = integrate(f * v, dΩ)                                 -> source, f
    a(u,v) = integrate(inner(ε(v), σ(ε(u))), dΩ), σ = IsotropicStress(E, ν)
                                                                  -> elasticity, E, ν
    l(v)   = integrate(dot(v, VectorValue(f1, f2, f3)), dΩ)       -> body force

κ, f and E are numbers or per-cell arrays (CellField(values, Ω)).

Over a BoundaryTriangulation Γ (measure dΓ, normal n = get_normal_vector(Γ)):

    ∫( α*v*u + β*(v*(n⋅∇(u)) + (n⋅∇(v))*u) )dΓ   -> face matrix (Nitsche, Robin)
    ∫( v*g + (n⋅∇(v))*h )dΓ                        -> face vector (Neumann, Nitsche)

where α, β are numbers or per-cell arrays and g, h may also be functions of x, FE
functions uh, n⋅∇(uh), or products of these; they are evaluated at the face
quadrature points by the host.
"""
import numbers

import numpy as np

from .geometry import BoundaryTriangulation, Triangulation


class CellField:
    """A field constant on each cell: CellField(values_per_cell, Ω) or a number."""

    def __init__(self, values, trian=None):
        self.values = values
        self.trian = trian

    def per_cell(self, ncells):
        v = self.values
        if isinstance(v, numbers.Number):
            return None if v == 1 else np.full(ncells, float(v))
        a = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        if a.shape[0] != ncells:
            raise ValueError("CellField has %d values for %d cells" % (a.shape[0], ncells))
        return a

    def __mul__(self, other):
        return _scale(other, self)

    __rmul__ = __mul__


class PointField:
    """A coefficient g(x) evaluated at the quadrature points (boundary terms)."""

    def __init__(self, func):
        self.func = func


class FEValue:
    """The values of an FE function uh at the quadrature points."""

    def __init__(self, uh):
        self.uh = uh


class NormalGradValue:
    """n⋅∇uh of an FE function at the boundary quadrature points."""

    def __init__(self, uh):
        self.uh = uh


class CoefProduct:
    """Product of coefficients, evaluated factor by factor."""

    def __init__(self, factors):
        self.factors = factors


_POINT_COEFS = (PointField, FEValue, NormalGradValue, CoefProduct)


def _coef_product(a, b):
    if a is None:
        return b
    if b is None:
        return a
    if isinstance(a, _POINT_COEFS) or isinstance(b, _POINT_COEFS):
        fa = a.factors if isinstance(a, CoefProduct) else [a]
        fb = b.factors if isinstance(b, CoefProduct) else [b]
        return CoefProduct(fa + fb)
    va = a.values if isinstance(a, CellField) else a
    vb = b.values if isinstance(b, CellField) else b
    if isinstance(va, numbers.Number) and isinstance(vb, numbers.Number):
        return CellField(float(va) * float(vb))
    return CellField(np.asarray(va, dtype=np.float64) * np.asarray(vb, dtype=np.float64))


class _Expr:
    coef = None
    fields = None   # multi-field: (test field,) or (test field, trial field)

    def __mul__(self, other):
        return _scale(other, self)

    __rmul__ = __mul__

    def __add__(self, other):
        return _Sum(_terms(self) + _terms(other))

    __radd__ = __add__

    def __neg__(self):
        return _Sum([_scale(-1.0, t) for t in _terms(self)])

    def __sub__(self, other):
        return self + (-other)

    def __rsub__(self, other):
        return other + (-self)


def _terms(e):
    return list(e.terms) if isinstance(e, _Sum) else [e]


class FEBasis(_Expr):
    """get_fe_basis(V) (test, v) or get_trial_fe_basis(U) (trial, u); field: the index of
    the field in a MultiFieldFESpace (None for a single-field space)."""

    def __init__(self, space, trial, field=None):
        self.space = space
        self.trial = trial
        self.field = field


def _f(term, fields):
    """term with its multi-field (test[, trial]) field indices recorded."""
    term.fields = fields
    return term


class _Grad(_Expr):
    def __init__(self, basis):
        self.basis = basis


class DiffusionTerm(_Expr):
    """coef * ∇v⋅∇u."""

    def __init__(self, coef=None):
        self.coef = coef


class SourceTerm(_Expr):
    """coef * v."""

    def __init__(self, coef=None):
        self.coef = coef


class _SymGrad(_Expr):
    """ε(u) = (∇u + ∇uᵀ)/2 of an FE basis."""

    def __init__(self, basis):
        self.basis = basis


class IsotropicStress:
    """σ(ε) = λ tr(ε) I + 2 μ ε with λ = Eν/((1+ν)(1-2ν)), μ = E/(2(1+ν)), E a number or
    per-cell CellField (test/GridapTests/IsotropicDamageTests.jl:13-17 spells the same
    law); σ(ε(u)) is Gridap's σ∘ε(u)."""

    def __init__(self, E, nu):
        self.E = _as_coef(E)
        self.nu = float(nu)

    def __call__(self, eps):
        if not isinstance(eps, _SymGrad):
            raise NotImplementedError("σ applies to ε(u) of a trial basis")
        return _Stress(eps, self)


class _Stress(_Expr):
    def __init__(self, eps, material):
        self.eps = eps
        self.material = material


class ElasticityTerm(_Expr):
    """ε(v) : σ(ε(u)) with per-cell E (coef) and ν."""

    def __init__(self, coef=None, nu=0.3):
        self.coef = coef
        self.nu = nu


class BodyForceTerm(_Expr):
    """v · f with f constant."""

    def __init__(self, force):
        self.force = tuple(force)


class _Sum(_Expr):
    def __init__(self, terms):
        self.terms = terms


class _Normal:
    """get_normal_vector(Γ): the outward unit normal of a boundary triangulation."""

    def __init__(self, trian):
        self.trian = trian


class _NormalGrad(_Expr):
    """n⋅∇(v) or n⋅∇(u) of an FE basis."""

    def __init__(self, basis):
        self.basis = basis


class _Div(_Expr):
    """coef * (∇⋅u) of an FE basis of a vector-valued space."""

    def __init__(self, basis, coef=None):
        self.basis = basis
        self.coef = coef


class DivergenceTerm(_Expr):
    """coef * q (∇⋅u) (part "q_divu": scalar test, vector trial) or coef * (∇⋅v) p (part
    "divv_p": vector test, scalar trial), the coupling blocks of
    test/GridapTests/StokesTaylorHoodTests.jl:54."""

    def __init__(self, part, coef=None):
        self.part, self.coef = part, coef


class _NormalComp(_Expr):
    """n⋅v of the test basis of a vector-valued space on a boundary."""

    def __init__(self, basis):
        self.basis = basis


class NormalTensorField:
    """n⋅T(x) of a tensor-valued function T (T[..., i, j], Gridap's ∇u convention
    ∇u[i, j] = ∂_i u_j) at the boundary quadrature points: the vector Σ_i n_i T_ij."""

    def __init__(self, func):
        self.func = func


class VectorLoadTerm(_Expr):
    """A boundary load on a vector-valued test basis: v⋅g with g = coef (a
    NormalTensorField or a vector function of x) for kind "vec", or (n⋅v) coef with a
    scalar coef for kind "normal" (the terms v⋅(n⋅∇u) and -(n⋅v) p of
    test/GridapTests/StokesTaylorHoodTests.jl:56)."""

    def __init__(self, kind, coef=None):
        self.kind, self.coef = kind, coef


class MassTerm(_Expr):
    """coef * v u (test times trial)."""

    def __init__(self, coef=None):
        self.coef = coef


class NitscheTerm(_Expr):
    """coef * v (n⋅∇u) (part "v_dnu") or coef * (n⋅∇v) u (part "dnv_u")."""

    def __init__(self, part, coef=None):
        self.part, self.coef = part, coef


class NormalLoadTerm(_Expr):
    """coef * (n⋅∇v) (a boundary load on the normal derivative of the test function)."""

    def __init__(self, coef=None):
        self.coef = coef


class _GradState(_Expr):
    """∇uh of an FE function (the state of a nonlinear form)."""

    def __init__(self, uh):
        self.uh = uh


class _Flux(_Expr):
    """flux(∇uh) = |∇uh|^(p-2) ∇uh (PLaplacianTests.jl:15)."""

    def __init__(self, p, state):
        self.p, self.state = p, state


class _DFlux(_Expr):
    """dflux(∇du, ∇uh) = (p-2)|∇uh|^(p-4)(∇uh⋅∇du)∇uh + |∇uh|^(p-2)∇du
    (PLaplacianTests.jl:16)."""

    def __init__(self, p, state):
        self.p, self.state = p, state


class PLaplacianResidualTerm(_Expr):
    """coef * ∇v⋅flux(∇uh)."""

    def __init__(self, p, state, coef=None):
        self.p, self.state, self.coef = p, state, coef


class PLaplacianJacobianTerm(_Expr):
    """coef * ∇v⋅dflux(∇du, ∇uh)."""

    def __init__(self, p, state, coef=None):
        self.p, self.state, self.coef = p, state, coef


class PLaplacian:
    """The constitutive law of test/GridapTests/PLaplacianTests.jl:14-17:
    law = PLaplacian(p); law.flux(∇(u)) and law.dflux(∇(du), ∇(u)) stand for Gridap's
    flux∘∇(u) and dflux∘(∇(du),∇(u)) inside res(u,v) and jac(u,du,v)."""

    def __init__(self, p):
        self.p = float(p)
        if not self.p >= 2.0:
            raise ValueError("p must be >= 2")

    def flux(self, grad_u):
        if not isinstance(grad_u, _GradState):
            raise NotImplementedError("flux applies to ∇(uh) of an FE function")
        return _Flux(self.p, grad_u.uh)

    def dflux(self, grad_du, grad_u):
        if not (isinstance(grad_du, _Grad) and grad_du.basis.trial and
                isinstance(grad_u, _GradState)):
            raise NotImplementedError("dflux applies to (∇(du) of the trial basis, ∇(uh))")
        return _DFlux(self.p, grad_u.uh)


def compose(f, *args):
    """f∘args (Gridap's ∘ on CellFields)."""
    return f(*args)


def _is_coef(x):
    from .fespaces import FEFunction
    if isinstance(x, (_Expr, FEBasis)):
        return False
    return (isinstance(x, (numbers.Number, np.ndarray, CellField, FEFunction)) or
            isinstance(x, _POINT_COEFS) or callable(x))


def _product(x, y):
    """x * y for one ordering of the two factors, or None."""
    cy = _is_coef(y)
    if isinstance(x, FEBasis) and not x.trial:                       # v * y
        if isinstance(y, FEBasis) and y.trial:
            return _f(MassTerm(), (x.field, y.field))
        if isinstance(y, _Div) and y.basis.trial:                    # q * (∇⋅u)
            return _f(DivergenceTerm("q_divu", y.coef), (x.field, y.basis.field))
        if isinstance(y, _NormalGrad) and y.basis.trial:
            return NitscheTerm("v_dnu")
        if cy:
            return _f(SourceTerm(_as_coef(y)), (x.field,))
    if isinstance(x, _Div):
        if not x.basis.trial and isinstance(y, FEBasis) and y.trial:  # (∇⋅v) * p
            return _f(DivergenceTerm("divv_p", x.coef), (x.basis.field, y.field))
        if cy:
            return _Div(x.basis, _coef_product(x.coef, _as_coef(y)))
    if isinstance(x, _Sum) and isinstance(y, FEBasis):               # (Σ terms) * basis
        return _Sum([_scale(y, t) for t in x.terms])
    if isinstance(x, _NormalComp) and cy:                            # (n⋅v) * p
        return _f(VectorLoadTerm("normal", _as_coef(y)), (x.basis.field,))
    if isinstance(x, VectorLoadTerm) and cy and x.kind == "normal":
        return _f(VectorLoadTerm("normal", _coef_product(x.coef, _as_coef(y))), x.fields)
    if isinstance(x, _NormalGrad) and not x.basis.trial:            # (n⋅∇v) * y
        if isinstance(y, FEBasis) and y.trial:
            return NitscheTerm("dnv_u")
        if cy:
            return NormalLoadTerm(_as_coef(y))
    if isinstance(x, SourceTerm):                                    # (c v) * y
        tf = x.fields[0] if x.fields else None
        if isinstance(y, FEBasis) and y.trial:
            return _f(MassTerm(x.coef), (tf, y.field))
        if isinstance(y, _Div) and y.basis.trial:
            return _f(DivergenceTerm("q_divu", _coef_product(x.coef, y.coef)),
                      (tf, y.basis.field))
        if isinstance(y, _NormalGrad) and y.basis.trial:
            return NitscheTerm("v_dnu", x.coef)
        if cy:
            return _f(SourceTerm(_coef_product(x.coef, _as_coef(y))), x.fields)
    if isinstance(x, NormalLoadTerm):                                # (c n⋅∇v) * y
        if isinstance(y, FEBasis) and y.trial:
            return NitscheTerm("dnv_u", x.coef)
        if cy:
            return NormalLoadTerm(_coef_product(x.coef, _as_coef(y)))
    if cy:
        if isinstance(x, (DiffusionTerm, MassTerm)):
            return _f(type(x)(_coef_product(x.coef, _as_coef(y))), x.fields)
        if isinstance(x, DivergenceTerm):
            return _f(DivergenceTerm(x.part, _coef_product(x.coef, _as_coef(y))), x.fields)
        if isinstance(x, NitscheTerm):
            return NitscheTerm(x.part, _coef_product(x.coef, _as_coef(y)))
        if isinstance(x, (PLaplacianResidualTerm, PLaplacianJacobianTerm)):
            return type(x)(x.p, x.state, _coef_product(x.coef, _as_coef(y)))
        if isinstance(x, _Sum):
            return _Sum([_scale(y, t) for t in x.terms])
    return None


def _scale(factor, expr):
    """factor * expr, in either order of the two operands."""
    for x, y in ((expr, factor), (factor, expr)):
        r = _product(x, y)
        if r is not None:
            return r
    if isinstance(expr, FEBasis) and expr.trial and _is_coef(factor):
        raise NotImplementedError("only ∫ f v source terms are provided")
    raise NotImplementedError("unsupported weak-form product")


def _as_coef(x):
    from .fespaces import FEFunction
    if isinstance(x, (CellField,) + _POINT_COEFS):
        return x
    if isinstance(x, numbers.Number):
        return CellField(float(x))
    if isinstance(x, np.ndarray):
        return CellField(x)
    if isinstance(x, FEFunction):
        return FEValue(x)
    if callable(x):
        return PointField(x)
    raise TypeError("unsupported coefficient %r" % (x,))


def _cell_coef(coef):
    """Per-cell coefficient of a cell integral (constant on each cell)."""
    if coef is not None and not isinstance(coef, CellField):
        raise NotImplementedError("coefficients of cell integrals must be constant per cell "
                                  "(a number or a per-cell array)")
    return coef


def get_normal_vector(trian):
    if not isinstance(trian, BoundaryTriangulation):
        raise TypeError("get_normal_vector needs a BoundaryTriangulation")
    return _Normal(trian)


def gradient(x):
    from .fespaces import FEFunction
    if isinstance(x, FEFunction):
        return _GradState(x)
    if not isinstance(x, FEBasis):
        raise NotImplementedError("gradient of an FE basis or an FE function only")
    return _Grad(x)


grad = gradient


def divergence(x):
    """∇⋅u of an FE basis of a vector-valued space."""
    if not isinstance(x, FEBasis):
        raise NotImplementedError("divergence of an FE basis only")
    return _Div(x)


def symmetric_gradient(x):
    if not isinstance(x, FEBasis):
        raise NotImplementedError("ε of an FE basis only")
    return _SymGrad(x)


ε = symmetric_gradient


def inner(a, b):
    from .fespaces import FEFunction
    from .reference_fes import VectorValue
    if isinstance(b, _Normal):
        a, b = b, a
    if isinstance(a, _Normal):
        if isinstance(b, _Grad):
            return _NormalGrad(b.basis)
        if isinstance(b, _GradState):
            return NormalGradValue(b.uh)
        if isinstance(b, FEBasis) and not b.trial:
            return _NormalComp(b)
        if callable(b) and not isinstance(b, (_Expr, FEBasis)):
            return NormalTensorField(b)
        raise NotImplementedError("n⋅x is provided for gradients, test bases and tensor "
                                  "functions only")
    if isinstance(a, NormalTensorField):
        a, b = b, a
    if isinstance(a, FEBasis) and not a.trial and isinstance(b, NormalTensorField):
        return _f(VectorLoadTerm("vec", b), (a.field,))
    if isinstance(a, _Grad) and isinstance(b, _Grad) and a.basis.trial != b.basis.trial:
        v, u = (b, a) if a.basis.trial else (a, b)
        return _f(DiffusionTerm(), (v.basis.field, u.basis.field))
    if isinstance(a, _Grad) and not a.basis.trial and isinstance(b, _Flux):
        return PLaplacianResidualTerm(b.p, b.state)
    if isinstance(a, _Grad) and not a.basis.trial and isinstance(b, _DFlux):
        return PLaplacianJacobianTerm(b.p, b.state)
    if isinstance(a, _SymGrad) and isinstance(b, _Stress) and not a.basis.trial and \
            b.eps.basis.trial:
        return ElasticityTerm(b.material.E, b.material.nu)
    if isinstance(a, FEBasis) and isinstance(b, VectorValue) and not a.trial:
        return _f(BodyForceTerm(b.comps), (a.field,))
    if isinstance(b, FEBasis) and isinstance(a, VectorValue) and not b.trial:
        return _f(BodyForceTerm(a.comps), (b.field,))
    if isinstance(a, FEBasis) and not isinstance(b, _Expr):
        return _scale(b, a)
    if isinstance(b, FEBasis) and not isinstance(a, _Expr):
        return _scale(a, b)
    try:
        return _scale(a, b)
    except NotImplementedError:
        raise NotImplementedError("inner/dot of these operands is not provided") from None


dot = inner


class Measure:
    """Measure(Ω, degree): the tensor-product Gauss rule of that degree."""

    def __init__(self, trian, degree):
        if not isinstance(trian, (Triangulation, BoundaryTriangulation)):
            raise TypeError("Measure(Ω, degree) needs a Triangulation or BoundaryTriangulation")
        self.trian = trian
        self.degree = int(degree)

    def __rmul__(self, integrand):
        if isinstance(integrand, Integrand):
            return integrate(integrand.expr, self)
        return NotImplemented


class DomainContribution:
    """Terms integrated over measures (src/CellData/DomainContributions.jl:7-56)."""

    def __init__(self, terms):
        self.terms = terms  # list of (term, measure)

    def __add__(self, other):
        return DomainContribution(self.terms + other.terms)

    def __radd__(self, other):
        if other == 0:
            return self
        return NotImplemented


_CELL_TERMS = (DiffusionTerm, MassTerm, SourceTerm, ElasticityTerm, BodyForceTerm,
               PLaplacianResidualTerm, PLaplacianJacobianTerm, DivergenceTerm)
_FACE_TERMS = (MassTerm, NitscheTerm, SourceTerm, NormalLoadTerm, VectorLoadTerm)


def integrate(expr, measure):
    terms = expr.terms if isinstance(expr, _Sum) else [expr]
    boundary = isinstance(measure.trian, BoundaryTriangulation)
    for t in terms:
        if not isinstance(t, _FACE_TERMS if boundary else _CELL_TERMS):
            raise NotImplementedError("unsupported integrand %r over %s" %
                                      (t, "a boundary" if boundary else "the cells"))
    return DomainContribution([(t, measure) for t in terms])


class Integrand:
    """integral(expr) * dΩ — the `∫(expr)dΩ` spelling."""

    def __init__(self, expr):
        self.expr = expr


def integral(expr):
    return Integrand(expr)


def get_fe_basis(V):
    """v = get_fe_basis(V); for a MultiFieldFESpace the tuple (v1, v2, ...)."""
    from .multifield import MultiFieldFESpace
    if isinstance(V, MultiFieldFESpace):
        return tuple(FEBasis(Vi, trial=False, field=i) for i, Vi in enumerate(V.spaces))
    return FEBasis(V, trial=False)


def get_trial_fe_basis(U):
    """u = get_trial_fe_basis(U); for a MultiFieldFESpace the tuple (u1, u2, ...)."""
    from .multifield import MultiFieldFESpace
    if isinstance(U, MultiFieldFESpace):
        return tuple(FEBasis(Ui, trial=True, field=i) for i, Ui in enumerate(U.spaces))
    return FEBasis(U, trial=True)
