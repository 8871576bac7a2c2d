"""Host side of the drop-in (no GPU): the Gridap-API mirror produces the same
model, labels, numbering and tabulation as the oracle, and the C-ABI library
loads and exports every entry point of include/gridap_hip.h."""
import ctypes
import os
import re

import numpy as np
import pytest

import gridap_hip as G
from gridap_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("part", [(2, 2), (5, 3), (3, 4, 2), (6, 5, 7)])
@pytest.mark.parametrize("tags", [["boundary"], [], [1, 2, 4, 5, 8], ["tag_25", "tag_21"]])
def test_numbering_matches_oracle(oracle, part, tags):
    D = len(part)
    if D == 2 and any(isinstance(t, str) and t.startswith("tag_2") for t in tags):
        pytest.skip("3-D tag names")
    if D == 3 and tags == [1, 2, 4, 5, 8]:
        tags = [1, 2, 4, 5, 8, 13, 22]
    dom = tuple(v for d in range(D) for v in (0.0, 1.0 + d))
    model = G.CartesianDiscreteModel(dom, part)
    om = oracle.cartesian_model(dom, part)
    np.testing.assert_array_equal(model.cell_node_ids(), om["cell_nodes"])
    np.testing.assert_allclose(model.node_coordinates(), om["xyz"], rtol=0, atol=1e-15)
    np.testing.assert_array_equal(model.get_face_labeling().vertex_entity, om["vertex_labels"])
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1), dirichlet_tags=tags)
    ents = [oracle.cartesian_tag_entities(D, t) for t in tags]
    n2d, nf, nd = oracle.clagrangian(oracle.face_tag_index(om["vertex_labels"], ents))
    cd = oracle.clagrangian_cell_dofs(om["cell_nodes"], n2d)
    np.testing.assert_array_equal(V.get_cell_dof_ids(), cd)
    assert (V.num_free_dofs(), V.num_dirichlet_dofs()) == (nf, nd)


def test_vector_numbering_fixture():
    """test/FESpacesTests/CLagrangianFESpacesTests.jl:103-110 through the host API."""
    model = G.CartesianDiscreteModel((0, 1, 0, 1), (2, 2))
    reffe = G.ReferenceFE(G.lagrangian, G.VectorValue[2, G.Float64], 1)
    V = G.FESpace(model, reffe, dirichlet_tags=[1, 2, 4, 5, 8],
                  dirichlet_masks=[(True, True), (True, False), (False, False), (False, True),
                                   (True, True)])
    assert [tuple(r) for r in V.node_and_comp_to_dof] == [
        (-1, -2), (1, -3), (-4, 2), (3, 4), (5, 6), (-5, -6), (7, 8), (9, 10), (11, 12)]
    assert V.free_dof_to_node.tolist() == [2, 3, 4, 4, 5, 5, 7, 7, 8, 8, 9, 9]
    assert V.dirichlet_dof_to_comp.tolist() == [1, 2, 2, 1, 1, 2]


@pytest.mark.parametrize("D,degree", [(2, 2), (3, 2), (3, 4), (2, 5)])
def test_tabulation_matches_oracle(oracle, D, degree):
    q = G.Quadrature(D, degree)
    x, w = oracle.tensor_quadrature(D, degree)
    np.testing.assert_allclose(q.coords, x, rtol=0, atol=2e-16)
    np.testing.assert_allclose(q.weights, w, rtol=0, atol=4e-15)
    for order in (1, 2):
        reffe = G.ReferenceFE(G.lagrangian, G.Float64, order)
        np.testing.assert_array_equal(reffe.nodes(D), oracle.qk_nodes(D, order))
        shp, grad = reffe.tabulate(D, q.coords)
        oshp, ograd = oracle.qk_tabulate(D, order, q.coords)
        np.testing.assert_allclose(shp, oshp, rtol=0, atol=1e-15)
        np.testing.assert_allclose(grad, ograd, rtol=0, atol=1e-14)


def test_weak_form_lowering():
    model = G.CartesianDiscreteModel((0, 1, 0, 1), (2, 3))
    Ω = G.Triangulation(model)
    dΩ = G.Measure(Ω, 2)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1),
                      dirichlet_tags="boundary")
    u, v = G.get_trial_fe_basis(V), G.get_fe_basis(V)
    k = np.arange(1.0, 7.0)
    c = G.integrate(G.CellField(k, Ω) * G.inner(G.gradient(v), G.gradient(u)), dΩ)
    assert len(c.terms) == 1 and np.array_equal(c.terms[0][0].coef.values, k)
    c2 = G.integral(G.dot(G.gradient(v), G.gradient(u)) * 2.0) * dΩ
    assert c2.terms[0][0].coef.values == 2.0
    s = G.integrate(v * 3.0, dΩ) + G.integrate(1.0 * v, dΩ)
    assert len(s.terms) == 2
    # a source f(x) is evaluated at the quadrature points; a coefficient of a bilinear
    # term must stay constant per cell
    assem = G.SparseMatrixAssembler(G.TrialFESpace(V), V)
    *_, extra = assem._lower(None, G.integrate(v * (lambda x: x[..., 0]), dΩ))
    assert extra["fq"].shape == (6, 4)
    with pytest.raises(NotImplementedError):
        assem._lower(G.integrate((lambda x: x[..., 0]) * G.inner(G.gradient(v),
                                                                 G.gradient(u)), dΩ), None)
    with pytest.raises(NotImplementedError):
        G.FESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 3))


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "gridap_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*\*?\s*(gh_\w+)\(", text, re.M)))


def test_abi_library_exports_every_header_symbol():
    syms = _header_symbols()
    assert len(syms) >= 16
    L = ctypes.CDLL(_lib.LIB_PATH)
    for s in syms:
        assert hasattr(L, s), "missing export %s" % s
    # the Python binding declares exactly the header's entry points
    assert sorted(_lib.SIGNATURES) == syms


def test_product_does_not_reach_the_oracle():
    pkg = os.path.join(ROOT, "gridap.jl_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "gridap_oracle" not in text and "import oracle" not in text, f


def _tet_moment_error(P, W, degree):
    """Max |Σ w x^a y^b z^c - a!b!c!/(a+b+c+3)!| over monomials of total degree == degree."""
    from math import factorial
    err = 0.0
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            c = degree - a - b
            exact = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)
            err = max(err, abs(np.sum(W * P[:, 0] ** a * P[:, 1] ** b * P[:, 2] ** c) - exact))
    return err


@pytest.mark.parametrize("degree,npts,exact_to", [(1, 1, 1), (2, 4, 2), (3, 8, 3),
                                                  (4, 14, 5), (5, 14, 5)])
def test_tet_rules_are_witherden_vincent(degree, npts, exact_to):
    """Quadrature(TET, degree): Gridap's point count, interior points, positive weights
    summing to 1/6, exact for every monomial up to the rule's degree and not beyond
    (so a mistyped digit of the transcribed table fails)."""
    q = G.Quadrature(3, degree, simplex=True)
    P, W = q.coords, q.weights
    assert len(W) == npts and np.all(W > 0)
    assert np.all(P > 0) and np.all(P.sum(1) < 1)
    assert abs(W.sum() - 1.0 / 6.0) <= 6e-17
    for d in range(exact_to + 1):
        assert _tet_moment_error(P, W, d) <= 6e-17, d
    assert _tet_moment_error(P, W, exact_to + 1) > 1e-7


def test_tet_degree4_rule_equals_the_moment_equation_root():
    """The transcribed 14-point rule is the interior positive-weight root of the
    degree-5 moment equations for two S31 orbits and one S22 orbit, which the oracle
    derives independently by least squares; points listed in Gridap's orbit order
    (S31: (s,s,t), (s,t,s), (t,s,s), (s,s,s) in bi-unit coordinates, then S22)."""
    import conforming_oracle as CO
    q = G.Quadrature(3, 4, simplex=True)
    P2, W2 = CO.tet_degree5_rule()
    key = lambda P, W: np.lexsort(np.c_[P, W].T)
    k1, k2 = key(q.coords, q.weights), key(P2, W2)
    np.testing.assert_allclose(q.coords[k1], P2[k2], atol=1e-14)
    np.testing.assert_allclose(q.weights[k1], W2[k2], atol=1e-16)
    # orbit order of the table: the first S31 orbit's fourth point is (s,s,s) mapped
    s = q.coords[3]
    assert s[0] == s[1] == s[2] and np.all(q.coords[0][:2] == s[:2]) and q.coords[0][2] != s[2]


def test_taylor_hood_form_lowering():
    """StokesTaylorHoodTests.jl's a((u,p),(v,q)) = ∫ ∇v⊙∇u - (∇⋅v) p + q (∇⋅u) splits into
    the vector Laplacian on (v, u), the (∇⋅v) p coupling with coefficient -1 on (v, p)
    and the q (∇⋅u) coupling on (q, u); l = ∫ v⋅f + q g into a body force on v and a
    source on q."""
    model = G.CartesianDiscreteModel((0, 2, 0, 2), (3, 3))
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.VectorValue[2, G.Float64], 2),
                      dirichlet_tags="boundary")
    Q = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1), dirichlet_tags="tag_1")
    Y = G.MultiFieldFESpace([V, Q])
    X = G.MultiFieldFESpace([G.TrialFESpace(V), G.TrialFESpace(Q)])
    dΩ = G.Measure(G.Triangulation(model), 2)
    u, v = G.get_trial_fe_basis(X), G.get_fe_basis(Y)
    c = G.integrate(G.inner(G.gradient(v[0]), G.gradient(u[0])) - G.divergence(v[0]) * u[1] +
                    v[1] * G.divergence(u[0]), dΩ)
    got = [(type(t).__name__, t.fields, getattr(t, "part", None),
            None if t.coef is None else t.coef.values) for t, _ in c.terms]
    assert got == [("DiffusionTerm", (0, 0), None, None),
                   ("DivergenceTerm", (0, 1), "divv_p", -1.0),
                   ("DivergenceTerm", (1, 0), "q_divu", None)]
    l = G.integrate(G.dot(v[0], G.VectorValue(-5.0, 5.0)) + v[1] * (lambda x: x[..., 0]), dΩ)
    assert [(type(t).__name__, t.fields) for t, _ in l.terms] == \
        [("BodyForceTerm", (0,)), ("SourceTerm", (1,))]
    with pytest.raises(NotImplementedError):
        G.divergence(2.0)
