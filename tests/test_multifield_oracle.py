"""Pins of the multi-field oracle (CPU): the Q1 mass element matrix against its closed
form, and the multi-field assembly against the single-field restatement."""
import numpy as np

import multifield_oracle as MF


def test_q1_mass_closed_form(oracle):
    """Q1 mass matrix on a Cartesian cell: the tensor product of the 1-D matrices
    h/6 [[2, 1], [1, 2]] (Gridap's node order, x fastest), and Σ_ab M_ab = σ |cell|."""
    for D, h in ((2, (0.25, 0.5)), (3, (0.2, 0.5, 0.125))):
        qx, qw = oracle.tensor_quadrature(D, 2)
        shp, _ = oracle.qk_tabulate(D, 1, qx)
        M = MF.q1_cell_mass_cartesian(D, h, qw, shp, np.array([3.0]))[0]
        m1 = [hk / 6.0 * np.array([[2.0, 1.0], [1.0, 2.0]]) for hk in h]
        ref = m1[0]
        for k in range(1, D):
            ref = np.kron(m1[k], ref)
        np.testing.assert_allclose(M, 3.0 * ref, rtol=1e-14, atol=0)
        assert abs(M.sum() - 3.0 * np.prod(h)) <= 1e-14


def test_one_field_equals_single_field_assembly(oracle):
    """With one field the multi-field restatement is the single-field one: same pattern
    and values as oracle.assemble_csc / assemble_vector with attach_dirichlet."""
    g = lambda x: x[0] + 2.0 * x[1]
    f = MF.q1_field((0, 1, 0, 1), (4, 3), g=g)
    mats = f["stiff"] + 0.5 * f["mass"]
    cp, rv, nz, b = MF.assemble([f["cell_dofs"]], [f["nfree"]], [f["ndiri"]],
                                {(0, 0): mats}, {0: f["src"]}, dvals=[f["dvals"]])
    cp1, rv1, nz1 = oracle.assemble_csc(f["cell_dofs"], mats, f["nfree"], f["nfree"])
    vecs = oracle.attach_dirichlet(f["cell_dofs"], mats, f["dvals"], f["src"])
    b1 = oracle.assemble_vector(f["cell_dofs"], vecs, f["nfree"])
    np.testing.assert_array_equal(cp, cp1)
    np.testing.assert_array_equal(rv, rv1)
    np.testing.assert_allclose(nz, nz1, rtol=0, atol=1e-15)
    np.testing.assert_allclose(b, b1, rtol=0, atol=1e-14)


def test_two_fields_are_shifted_blocks(oracle):
    """Two copies of one field with pairs (1,1) stiffness, (2,2) mass, (2,1) -mass: the
    monolithic matrix is the block matrix of the single-field ones (scipy), and its
    pattern has no entry in the untouched block (1,2)."""
    import scipy.sparse as sp
    f = MF.q1_field((0, 1, 0, 1), (3, 4))
    n = f["nfree"]
    cd = f["cell_dofs"]
    cp, rv, nz, b = MF.assemble([cd, cd], [n, n], [f["ndiri"]] * 2,
                                {(0, 0): f["stiff"], (1, 1): f["mass"], (1, 0): -f["mass"]},
                                {0: f["src"], 1: -f["src"]})
    A = sp.csc_matrix((nz, rv - 1, cp - 1), shape=(2 * n, 2 * n))
    S = oracle.to_scipy(*oracle.assemble_csc(cd, f["stiff"], n, n), n)
    M = oracle.to_scipy(*oracle.assemble_csc(cd, f["mass"], n, n), n)
    ref = sp.bmat([[S, None], [-M, M]]).toarray()
    np.testing.assert_allclose(A.toarray(), ref, rtol=0, atol=1e-15)
    assert A[:n, n:].nnz == 0
    bs = oracle.assemble_vector(cd, f["src"], n)
    np.testing.assert_allclose(b, np.concatenate([bs, -bs]), rtol=0, atol=1e-15)
