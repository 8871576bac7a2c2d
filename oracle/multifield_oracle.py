"""Multi-field assembly oracle (FP64, CPU).

TEST INFRASTRUCTURE ONLY: imported by tests, never by the product package.

Restates Gridap's multi-field assembly for fields sharing one cell loop:
* free dofs of the ConsecutiveMultiFieldStyle: field 1's, then field 2's, ...
  (src/MultiField/MultiFieldFESpaces.jl, the consecutive cell dof ids);
* per cell, the touched blocks of the block cell matrix are added column block by
  column block, rows inside (AddEntriesMap on a MatrixBlock, src/Arrays/
  AlgebraMaps.jl:173-182), each block entry by entry with columns outer, rows inner
  and non-positive (Dirichlet) ids skipped (_add_entries!,
  src/Algebra/AlgebraInterfaces.jl:175-187);
* the builder is the MinMemory CSC one (counter, inserter, create_from_nz), driven
  through the C oracle (oracle.csc_from_entries);
* the lifting vec_i -= Σ_j mat_ij u_dj of every cell with Dirichlet dofs
  (attach_dirichlet, src/CellData/AttachDirichlet.jl:59-84) per field pair.
The mass element matrix of a Cartesian Q1 cell is σ |det J| Σ_q w_q φ_a φ_b
(IntegrationMap, src/Fields/FieldsInterfaces.jl:737-760).
"""
import numpy as np

import oracle as O


def q1_cell_mass_cartesian(D, sizes, qw, shp, sigma):
    """[c, lj, li] mass matrices σ_c |det J| Σ_q w_q φ_li(q) φ_lj(q) of a Cartesian Q1 mesh
    (|det J| = Π h_k)."""
    detj = float(np.prod(np.asarray(sizes, dtype=np.float64)[:D]))
    shp = np.asarray(shp, dtype=np.float64)
    ref = np.einsum("q,qa,qb->ab", np.asarray(qw, dtype=np.float64), shp, shp) * detj
    sigma = np.asarray(sigma, dtype=np.float64)
    return sigma[:, None, None] * ref[None, :, :]


def consecutive_ids(cell_dofs, nfree, ndiri):
    """Cell dof ids of every field in the consecutive multi-field numbering: free ids
    shifted by the free dofs of the fields before, Dirichlet ids by their Dirichlet
    dofs."""
    out, fo, do = [], 0, 0
    for cd, nf, nd in zip(cell_dofs, nfree, ndiri):
        cd = np.asarray(cd, dtype=np.int64)
        out.append(np.where(cd > 0, cd + fo, cd - do))
        fo += nf
        do += nd
    return out


def assemble(cell_dofs, nfree, ndiri, cellmats, cellvecs, dvals=None):
    """Monolithic (ConsecutiveMultiFieldStyle) A and b.

    cell_dofs[f]: [C, nloc_f] Gridap ids of field f; cellmats[(i, j)]: [C, nloc_j,
    nloc_i] (column-major per cell) for the touched pairs; cellvecs[i]: [C, nloc_i] for
    the touched test fields; dvals[f]: Dirichlet values of field f (None: zero).
    Returns 1-based (colptr, rowval, nzval) and b."""
    nf = len(cell_dofs)
    ids = consecutive_ids(cell_dofs, nfree, ndiri)
    N = int(sum(nfree))
    C = ids[0].shape[0]
    pairs = [(i, j) for j in range(nf) for i in range(nf) if (i, j) in cellmats]
    I, J, V = [], [], []
    for c in range(C):
        for (i, j) in pairs:
            K = cellmats[(i, j)][c]
            ri, cj = ids[i][c], ids[j][c]
            for lj in range(len(cj)):
                if cj[lj] <= 0:
                    continue
                for li in range(len(ri)):
                    if ri[li] <= 0:
                        continue
                    I.append(ri[li])
                    J.append(cj[lj])
                    V.append(K[lj, li])
    _, colptr, rowval, nzval = O.csc_from_entries(list(zip(I, J, V)), N)
    # right-hand side with the lifting of every touched pair
    b = np.zeros(N)
    vecs = {i: np.array(v, dtype=np.float64) for i, v in cellvecs.items()}
    for (i, j) in pairs:
        if dvals is None or dvals[j] is None or len(dvals[j]) == 0:
            continue
        dj = np.asarray(dvals[j], dtype=np.float64)
        cdj = np.asarray(cell_dofs[j], dtype=np.int64)
        ud = np.where(cdj < 0, dj[np.maximum(-cdj, 1) - 1], 0.0)       # [C, nloc_j]
        lift = np.einsum("cji,cj->ci", cellmats[(i, j)], ud)
        vecs[i] = vecs.get(i, np.zeros_like(lift)) - lift
    for i, v in vecs.items():
        ri = ids[i]
        for c in range(C):
            for a in range(ri.shape[1]):
                if ri[c, a] > 0:
                    b[ri[c, a] - 1] += v[c, a]
    return colptr, rowval, nzval, b


def q1_field(domain, partition, dirichlet_tags=("boundary",), g=None, degree=2):
    """One scalar Q1 field on a Cartesian model: its cell dofs and counts, Dirichlet
    values of g (a function of x, or None), and the per-cell stiffness (κ = 1), mass
    (σ = 1) and source (f = 1) arrays on the CartesianMap."""
    D = len(partition)
    model = O.cartesian_model(domain, partition)
    ents = [O.cartesian_tag_entities(D, t) for t in dirichlet_tags]
    node_to_tag = O.face_tag_index(model["vertex_labels"], ents)
    n2d, nfree, ndiri = O.clagrangian(node_to_tag)
    cd = O.clagrangian_cell_dofs(model["cell_nodes"], n2d)
    qx, qw = O.tensor_quadrature(D, degree)
    shp, grad = O.qk_tabulate(D, 1, qx)
    cs = model["sizes"]
    stiff = O.cell_matrices_diffusion(D, model["cell_nodes"], model["xyz"], qw, grad,
                                      cartesian_sizes=cs)
    C = cd.shape[0]
    mass = q1_cell_mass_cartesian(D, cs, qw, shp, np.ones(C))
    src = O.cell_vectors_source(D, model["cell_nodes"], model["xyz"], qw, shp, grad,
                                cartesian_sizes=cs)
    dv = np.zeros(ndiri)
    if g is not None and ndiri:
        dnodes = np.nonzero(n2d[:, 0] < 0)[0]
        dv[-n2d[dnodes, 0] - 1] = [g(model["xyz"][n]) for n in dnodes]
    return dict(model=model, cell_dofs=cd, nfree=nfree, ndiri=ndiri, dvals=dv, stiff=stiff,
                mass=mass, src=src, qw=qw, shp=shp, grad=grad)
