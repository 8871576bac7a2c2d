"""Lowering of boundary integrals to the device's per-face data.

Gridap integrates ∫(…)dΓ over a BoundaryTriangulation by restricting the cell's basis
to the facet quadrature points (src/Geometry/BoundaryTriangulations.jl:122-183,
src/CellData/CellQuadratures.jl:138-170).  The device kernel (gh_boundary_upload)
takes per face a matrix part α φ_a φ_b + β (φ_a n⋅∇φ_b + n⋅∇φ_a φ_b) and a vector part
φ_a g + n⋅∇φ_a h; this module turns the recorded weak-form terms into α, β (per face)
and g, h (per face quadrature point), evaluating every coefficient (numbers, per-cell
arrays, functions of x, FE functions uh and n⋅∇uh) at the face points.
"""
import numpy as np

from .celldata import (CellField, CoefProduct, FEValue, MassTerm, NitscheTerm,
                       NormalGradValue, NormalLoadTerm, NormalTensorField, PointField,
                       SourceTerm, VectorLoadTerm)
from .reference_fes import facet_points, facet_quadrature


class FaceSet:
    """The union of the boundary faces of several triangulations, with the reference
    tabulation at the facet quadrature points of every n-cube facet."""

    def __init__(self, model, reffe, trians, degree):
        D = model.D
        keys = {}
        for tr in trians:
            for c, sd in zip(tr.face_cells, tr.face_sides):
                keys.setdefault((int(c), int(sd)), len(keys))
        self.cells = np.array([k[0] for k in keys], dtype=np.int64)
        self.sides = np.array([k[1] for k in keys], dtype=np.int64)
        self.index = keys
        self.model, self.D = model, D
        qx, qw = facet_quadrature(D, degree)
        self.qw = qw
        self.nqf = len(qw)
        shp, grad = [], []
        for side in range(2 * D):
            s, g = reffe.tabulate(D, facet_points(D, side, qx))
            shp.append(s)
            grad.append(g)
        self.shp = np.ascontiguousarray(np.stack(shp))      # [side, q, a]
        self.grad = np.ascontiguousarray(np.stack(grad))    # [side, q, a, D]
        h = np.asarray(model.sizes, dtype=np.float64)
        axis = D - 1 - self.sides // 2
        sgn = np.where(self.sides & 1, 1.0, -1.0)
        self.axis, self.sgn = axis, sgn                      # outward normal sgn e_axis
        # n⋅∇φ_a at the face points: ±∂φ̂_a/∂ξ_axis / h_axis
        self.dn = (self.grad[self.sides, :, :, :][np.arange(len(axis)), :, :, axis] *
                   (sgn / h[axis])[:, None, None])          # [f, q, a]
        ijk = np.stack(np.unravel_index(self.cells, model.partition, order="F"), axis=1)
        xi = np.stack([facet_points(D, sd, qx) for sd in range(2 * D)])[self.sides]  # [f,q,D]
        self.x = np.asarray(model.origin)[None, None, :] + (ijk[:, None, :] + xi) * h

    def positions(self, trian):
        return np.array([self.index[(int(c), int(s))]
                         for c, s in zip(trian.face_cells, trian.face_sides)], dtype=np.int64)


def _cell_values(uh):
    """Cell-wise dof values of an FE function (free ids > 0, Dirichlet ids < 0)."""
    V = uh.space.space if hasattr(uh.space, "space") else uh.space
    cd = np.asarray(V.cell_dof_ids, dtype=np.int64)
    free = np.asarray(uh.free_values, dtype=np.float64)
    diri = np.asarray(uh.dirichlet_values, dtype=np.float64)
    out = np.zeros(cd.shape)
    pos = cd > 0
    out[pos] = free[cd[pos] - 1]
    if diri.size:
        out[~pos] = diri[-cd[~pos] - 1]
    return out


def eval_coef(coef, fs, pos, ncells):
    """coef at the quadrature points of the faces fs[pos] -> [len(pos), nqf]."""
    nq = fs.nqf
    if coef is None:
        return np.ones((len(pos), nq))
    if isinstance(coef, CellField):
        v = coef.values
        if np.isscalar(v):
            return np.full((len(pos), nq), float(v))
        pc = coef.per_cell(ncells)
        return np.repeat(pc[fs.cells[pos]][:, None], nq, axis=1)
    if isinstance(coef, PointField):
        X = fs.x[pos].reshape(-1, fs.D)
        val = np.asarray(coef.func(X), dtype=np.float64)
        return np.broadcast_to(val.reshape(-1), (len(pos) * nq,)).reshape(len(pos), nq).copy()
    if isinstance(coef, FEValue):
        uc = _cell_values(coef.uh)[fs.cells[pos]]                         # [f, a]
        return np.einsum("fqa,fa->fq", fs.shp[fs.sides[pos]], uc)
    if isinstance(coef, NormalGradValue):
        uc = _cell_values(coef.uh)[fs.cells[pos]]
        return np.einsum("fqa,fa->fq", fs.dn[pos], uc)
    if isinstance(coef, CoefProduct):
        out = np.ones((len(pos), nq))
        for c in coef.factors:
            out = out * eval_coef(c, fs, pos, ncells)
        return out
    raise NotImplementedError("unsupported boundary coefficient %r" % (coef,))


def _cell_values_vec(uh, nc):
    """[C, nc, nn] cell values of a (vector) FE function, components blocked."""
    uc = _cell_values(uh)
    return uc.reshape(uc.shape[0], nc, -1)


def eval_coef_vec(coef, fs, pos, ncells, nc):
    """coef at the face points, [len(pos), nqf, k] with k = 1 (scalar) or nc."""
    nq = fs.nqf
    if isinstance(coef, FEValue):
        uc = _cell_values_vec(coef.uh, nc)[fs.cells[pos]]                  # [f, k, a]
        return np.einsum("fqa,fka->fqk", fs.shp[fs.sides[pos]], uc)
    if isinstance(coef, NormalGradValue):
        uc = _cell_values_vec(coef.uh, nc)[fs.cells[pos]]
        return np.einsum("fqa,fka->fqk", fs.dn[pos], uc)
    if isinstance(coef, PointField):
        X = fs.x[pos].reshape(-1, fs.D)
        val = np.asarray(coef.func(X), dtype=np.float64)
        k = val.size // (len(pos) * nq)
        return val.reshape(len(pos), nq, k)
    if isinstance(coef, CoefProduct):
        out = np.ones((len(pos), nq, 1))
        for c in coef.factors:
            out = out * eval_coef_vec(c, fs, pos, ncells, nc)
        return out
    return eval_coef(coef, fs, pos, ncells)[:, :, None]


def _vector_loads(terms, fs, ncells, D):
    """g [nfaces, nqf, D] of the vector boundary loads (v⋅g): v⋅(n⋅T(x)) and (n⋅v) c."""
    F, nq = len(fs.cells), fs.nqf
    g = np.zeros((F, nq, D))
    h = np.zeros((F, nq, D))
    for t, m in terms:
        pos = fs.positions(m.trian)
        if isinstance(t, SourceTerm):             # v⊙c: component-wise g
            np.add.at(g, pos, np.broadcast_to(eval_coef_vec(t.coef, fs, pos, ncells, D),
                                              (len(pos), nq, D)))
            continue
        if isinstance(t, NormalLoadTerm):         # (n⋅∇v)⊙c: component-wise h
            np.add.at(h, pos, np.broadcast_to(eval_coef_vec(t.coef, fs, pos, ncells, D),
                                              (len(pos), nq, D)))
            continue
        if not isinstance(t, VectorLoadTerm):
            raise NotImplementedError("vector boundary loads: v⊙g, (n⋅∇v)⊙h, v⋅(n⋅∇u) and "
                                      "(n⋅v) p")
        n = np.zeros((len(pos), D))
        n[np.arange(len(pos)), fs.axis[pos]] = fs.sgn[pos]
        if t.kind == "vec":
            if not isinstance(t.coef, NormalTensorField):
                raise NotImplementedError("v⋅g on a boundary: g = n⋅T(x)")
            X = fs.x[pos].reshape(-1, D)
            T = np.asarray(t.coef.func(X), dtype=np.float64).reshape(len(pos), nq, D, D)
            val = np.einsum("fi,fqij->fqj", n, T)
        else:
            c = eval_coef(t.coef, fs, pos, ncells)                        # [f, q]
            val = c[:, :, None] * n[:, None, :]
        np.add.at(g, pos, val)
    return g, h


def lower_boundary(model, reffe, mat_terms, vec_terms):
    """(term, measure) lists over boundary measures -> FaceSet and per-face data
    (alpha, beta or None; gq, hq [nfaces, nqf] or None; for a vector-valued space gq is
    [nfaces, nqf, ncomp] and only loads are taken)."""
    terms = list(mat_terms) + list(vec_terms)
    if not terms:
        return None
    vector = getattr(reffe, "ncomp", 1) > 1
    degs = {m.degree for _, m in terms}
    if len(degs) > 1:
        raise NotImplementedError("boundary terms must share one quadrature degree")
    trians = []
    for _, m in terms:
        if all(m.trian is not t for t in trians):
            trians.append(m.trian)
    fs = FaceSet(model, reffe, trians, degs.pop())
    F, nq, ncells = len(fs.cells), fs.nqf, model.num_cells()
    alpha = np.zeros(F)
    beta = {"v_dnu": np.zeros(F), "dnv_u": np.zeros(F)}
    gq, hq = np.zeros((F, nq)), np.zeros((F, nq))
    have = dict(alpha=False, beta=False, g=False, h=False)
    for t, m in mat_terms:
        pos = fs.positions(m.trian)
        if isinstance(t.coef, (PointField, FEValue, NormalGradValue, CoefProduct)):
            raise NotImplementedError("boundary matrix coefficients must be constant per face")
        val = eval_coef(t.coef, fs, pos, ncells)[:, 0]
        if isinstance(t, MassTerm):
            np.add.at(alpha, pos, val)
            have["alpha"] = True
        elif isinstance(t, NitscheTerm):
            np.add.at(beta[t.part], pos, val)
            have["beta"] = True
        else:
            raise NotImplementedError("boundary bilinear terms: v u, v (n⋅∇u), (n⋅∇v) u")
    if not np.array_equal(beta["v_dnu"], beta["dnv_u"]):
        raise NotImplementedError("the terms v (n⋅∇u) and (n⋅∇v) u must carry the same "
                                  "coefficient (symmetric Nitsche form)")
    if vector:
        # component-wise terms: the scalar face matrix per component, loads per component
        g, h = _vector_loads(vec_terms, fs, ncells, model.D)
        use = bool(vec_terms)
        return dict(faces=fs, alpha=alpha if have["alpha"] else None,
                    beta=beta["v_dnu"] if have["beta"] else None,
                    gq=np.ascontiguousarray(g) if use else None,
                    hq=np.ascontiguousarray(h) if use and np.any(h) else None)
    for t, m in vec_terms:
        pos = fs.positions(m.trian)
        val = eval_coef(t.coef, fs, pos, ncells)
        if isinstance(t, SourceTerm):
            np.add.at(gq, pos, val)
            have["g"] = True
        elif isinstance(t, NormalLoadTerm):
            np.add.at(hq, pos, val)
            have["h"] = True
        else:
            raise NotImplementedError("boundary linear terms: v g, (n⋅∇v) h")
    return dict(faces=fs, alpha=alpha if have["alpha"] else None,
                beta=beta["v_dnu"] if have["beta"] else None,
                gq=gq if have["g"] else None, hq=hq if have["h"] else None)
