"""Plain-Python restatement of Gridap's simplex/conforming path (TEST INFRASTRUCTURE
ONLY): extrusion-polytope face order, first-touch grid topology, the conforming
dof numbering, hex -> tet simplexification and the P2 tetrahedral Poisson
assembly.  Loops are written for clarity on small meshes; each function cites the
Gridap lines it follows (Gridap.jl v0.20.8).
"""
import math

import numpy as np

HEX_AXIS, TET_AXIS = 1, 2


# ---- reference polytopes -------------------------------------------------------

def _nface_boundary(anchor, extrusion, extend, isanchor, out):
    """src/ReferenceFEs/ExtrusionPolytopes.jl:_nfaceboundary! — recursive
    enumeration of the n-faces (anchor, extrusion) of an extrusion polytope."""
    D = len(extrusion)
    out.append((tuple(anchor), tuple(extrusion)))
    ext = list(extend)
    for i in range(D):
        cur = ext[i]
        if cur <= 0:
            continue
        ext = [0 if j == i else ext[j] for j in range(D)]
        unit = [1 if j == i else 0 for j in range(D)]
        if cur == HEX_AXIS:
            _nface_boundary([a + u for a, u in zip(anchor, unit)], extrusion, ext, False, out)
        elif isanchor:
            base = [anchor[j] if j >= i else 0 for j in range(D)]
            _nface_boundary([a + u for a, u in zip(base, unit)], extrusion, ext, False, out)
        _nface_boundary(anchor, [x + u * cur for x, u in zip(extrusion, unit)], ext, False, out)
    return out


def _nface_dim(extrusion):
    return sum(1 for e in extrusion if e > 0)


def polytope_nfaces(anchor, extrusion):
    """The n-faces of the face (anchor, extrusion), in Gridap's order
    (src/ReferenceFEs/ExtrusionPolytopes.jl:_polytopenfaces): by dimension, then by
    the extrusion (last axis most significant), then by the anchor (last axis most
    significant); ties keep the recursion order."""
    D = len(extrusion)
    faces = _nface_boundary(list(anchor), [0] * D, list(extrusion), True, [])
    keyed = sorted(range(len(faces)), key=lambda k: (
        _nface_dim(faces[k][1]),
        tuple(faces[k][1][i] for i in reversed(range(D))),
        tuple(faces[k][0][i] for i in reversed(range(D))), k))
    return [faces[k] for k in keyed]


class Polytope:
    """Face lists of an extrusion polytope: faces[d] = list of vertex tuples
    (1-based local vertex ids), vertex coordinates."""

    def __init__(self, extrusion):
        self.extrusion = tuple(extrusion)
        self.D = len(extrusion)
        nf = polytope_nfaces([0] * self.D, extrusion)
        ids = {f: k for k, f in enumerate(nf)}
        self.nfaces = nf
        self.faces = [[] for _ in range(self.D + 1)]
        for f in nf:
            sub = polytope_nfaces(f[0], f[1])
            verts = tuple(ids[s] + 1 for s in sub if _nface_dim(s[1]) == 0)
            self.faces[_nface_dim(f[1])].append(verts)
        self.vertex_coords = np.array([f[0] for f in nf if _nface_dim(f[1]) == 0], dtype=float)


TET = Polytope((TET_AXIS,) * 3)
TRI = Polytope((TET_AXIS,) * 2)
HEX = Polytope((HEX_AXIS,) * 3)
QUAD = Polytope((HEX_AXIS,) * 2)


def oriented_simplexify_hex():
    """The six tetrahedra of a HEX, local vertex lists (1-based), from
    src/ReferenceFEs/ExtrusionPolytopes.jl:290-299 (oriented variant, the default of
    simplexify)."""
    return [(1, 2, 3, 7), (1, 2, 5, 7), (2, 3, 4, 7), (2, 4, 7, 8), (2, 5, 6, 7), (2, 6, 7, 8)]


# ---- grid topology ------------------------------------------------------------------

def first_touch_faces(cell_vertices, lface_vertices):
    """cell -> d-face ids (1-based) by first touch: cells in order, local faces in
    order; an unnumbered face gets the next id in every cell that holds it
    (src/Geometry/GridTopologies.jl:1184-1251, _cell_to_faces_fill!)."""
    ncells = len(cell_vertices)
    out = np.zeros((ncells, len(lface_vertices)), dtype=np.int64)
    seen = {}
    nxt = 1
    for c in range(ncells):
        for lf, lv in enumerate(lface_vertices):
            key = frozenset(int(cell_vertices[c][v - 1]) for v in lv)
            if key not in seen:
                seen[key] = nxt
                nxt += 1
            out[c, lf] = seen[key]
    return out, nxt - 1


# ---- conforming numbering -------------------------------------------------------------

def conforming_cell_dofs(d_cell_faces, d_face_tags, d_lface_own, nloc, ldof_comp=None,
                         comp_masks=None):
    """Gridap's generic conforming numbering (src/FESpaces/ConformingFESpaces.jl:
    367-575): every d-face owns the dofs its reference face owns; faces are visited
    dimension by dimension in face-id order and their own dofs get the next free
    (tag unset) or Dirichlet (tag set) id; cell dofs are gathered face by face.

    d_cell_faces[d]: cells x nlfaces_d array of 1-based d-face ids.
    d_face_tags[d]: per d-face, 0 = free, else the Dirichlet tag (1-based).
    d_lface_own[d][lf]: local dof ids (0-based) owned by local d-face lf, in order.
    comp_masks (optional, with ldof_comp): per tag, per component, whether that
    component is Dirichlet on the tag's faces
    (_split_face_own_dofs_into_free_and_dirichlet_with_components!, :577-636).
    Returns (cell_dofs [cells x nloc], nfree, ndiri, dirichlet_dof_tag)."""
    D = len(d_cell_faces) - 1
    ncells = d_cell_faces[0].shape[0]
    face_dofs = []
    nfree = ndiri = 0
    diri_tag = []
    for d in range(D + 1):
        nface = len(d_face_tags[d])
        # own dof count of each d-face: from any cell holding it (one cell type)
        nown = np.zeros(nface, dtype=np.int64)
        owner_lf = np.zeros(nface, dtype=np.int64)
        for c in range(ncells):
            for lf, f in enumerate(d_cell_faces[d][c]):
                nown[f - 1] = len(d_lface_own[d][lf])
                owner_lf[f - 1] = lf
        ids = []
        for f in range(nface):
            tag = d_face_tags[d][f]
            mine = []
            for k in range(nown[f]):
                diri = tag != 0
                if diri and comp_masks is not None:
                    ldof = d_lface_own[d][owner_lf[f]][k]
                    diri = bool(comp_masks[tag - 1][ldof_comp[ldof]])
                if not diri:
                    nfree += 1
                    mine.append(nfree)
                else:
                    ndiri += 1
                    mine.append(-ndiri)
                    diri_tag.append(tag)
            ids.append(mine)
        face_dofs.append(ids)
    cd = np.zeros((ncells, nloc), dtype=np.int64)
    for c in range(ncells):
        for d in range(D + 1):
            for lf, f in enumerate(d_cell_faces[d][c]):
                for k, ldof in enumerate(d_lface_own[d][lf]):
                    cd[c, ldof] = face_dofs[d][f - 1][k]
    return cd, nfree, ndiri, diri_tag


def lagrangian_own_dofs(poly, order):
    """Local dofs owned by each face for scalar Lagrangian order 1 or 2 with Gridap's
    node order (vertices, then one node per edge, then per quad face / hex interior
    for Q2), src/ReferenceFEs/LagrangianRefFEs.jl own-nodes ordering."""
    D = poly.D
    out = []
    k = 0
    for d in range(D + 1):
        per = []
        for lf, verts in enumerate(poly.faces[d]):
            if d == 0:
                per.append([k])
                k += 1
            elif order == 2 and (d == 1 or len(verts) == 2 ** d):
                per.append([k])   # edge midpoint; face/cell centre of an n-cube face
                k += 1
            else:
                per.append([])
        out.append(per)
    return out, k


# ---- Cartesian box entities (labels) ---------------------------------------------------

def box_entity(D, fixed):
    """1-based id of the face of the D-cube [0,1]^D where axis d is fixed at
    fixed[d] in {0, 1} or free (None): Gridap's n-cube face order, which is also the
    CartesianDiscreteModel entity numbering (src/Geometry/CartesianDiscreteModels.jl)."""
    poly = HEX if D == 3 else QUAD
    ext = tuple(0 if f is not None else HEX_AXIS for f in fixed)
    anc = tuple(f if f is not None else 0 for f in fixed)
    for k, f in enumerate(poly.nfaces):
        if f == (anc, ext):
            return k + 1
    raise ValueError(fixed)


def simplex_face_entity(D, partition, node_ijk_list):
    """Entity label of a face of the simplexified Cartesian model whose vertices
    have lattice coordinates node_ijk_list: the smallest box face containing them
    (the label transfer of src/Geometry/UnstructuredDiscreteModels.jl:85-311 for
    vertices, edges on hex edges/faces, face diagonals and body diagonals)."""
    fixed = []
    for d in range(D):
        vals = {int(ijk[d]) for ijk in node_ijk_list}
        if vals == {0}:
            fixed.append(0)
        elif vals == {partition[d]}:
            fixed.append(1)
        else:
            fixed.append(None)
    return box_entity(D, fixed)


def boundary_entities(D):
    return set(range(1, 3 ** D))        # every entity but the interior (the last)


# ---- the simplexified Cartesian model ------------------------------------------------

def simplexify_cartesian(partition):
    """Tet cells (0-based node ids, hex-major, the six local tets in order) of the
    simplexified CartesianDiscreteModel (src/Geometry/Grids.jl:487-568)."""
    n0, n1, n2 = partition
    tets = oriented_simplexify_hex()
    cells = []
    for k in range(n2):
        for j in range(n1):
            for i in range(n0):
                hexn = [(i + a) + (n0 + 1) * ((j + b) + (n1 + 1) * (k + c))
                        for c in (0, 1) for b in (0, 1) for a in (0, 1)]
                for t in tets:
                    cells.append([hexn[v - 1] for v in t])
    return np.array(cells, dtype=np.int64)


def node_ijk(partition, n):
    n0, n1 = partition[0] + 1, partition[1] + 1
    return (n % n0, (n // n0) % n1, n // (n0 * n1))


def tet_numbering(partition, dirichlet_entities, order=2):
    """P1/P2 conforming cell dofs on the simplexified Cartesian grid."""
    cells = simplexify_cartesian(partition)
    D = 3
    nnodes = int(np.prod([p + 1 for p in partition]))
    cell_vertices = cells + 1
    edges, nedges = first_touch_faces(cells, TET.faces[1])
    faces2, nf2 = first_touch_faces(cells, TET.faces[2])
    cellsd = np.arange(1, len(cells) + 1).reshape(-1, 1)
    vtag = [1 if simplex_face_entity(D, partition, [node_ijk(partition, n)]) in dirichlet_entities
            else 0 for n in range(nnodes)]
    etag = [0] * nedges
    for c in range(len(cells)):
        for le, lv in enumerate(TET.faces[1]):
            e = edges[c, le] - 1
            ent = simplex_face_entity(D, partition, [node_ijk(partition, cells[c][v - 1]) for v in lv])
            etag[e] = 1 if ent in dirichlet_entities else 0
    own, nloc = lagrangian_own_dofs(TET, order)
    cd, nfree, ndiri, _ = conforming_cell_dofs([cell_vertices, edges, faces2, cellsd],
                                               [vtag, etag, [0] * nf2, [0] * len(cells)],
                                               own, nloc)
    return cells, cd, nfree, ndiri, edges


def p2_tet_numbering(partition, dirichlet_entities):
    return tet_numbering(partition, dirichlet_entities, 2)


# ---- P2 on the tetrahedron ---------------------------------------------------------------

def p2_tet_nodes():
    """Gridap's P2 TET nodes: the 4 vertices, then the 6 edge midpoints in TET edge
    order."""
    V = TET.vertex_coords
    mids = [0.5 * (V[a - 1] + V[b - 1]) for a, b in TET.faces[1]]
    return np.vstack([V, np.array(mids)])


def p2_tet_tabulate(x):
    """Nodal P2 basis on the reference tet at points x [nq x 3]: values [nq x 10],
    gradients [nq x 10 x 3].  λ = (1-x-y-z, x, y, z) are the vertex barycentrics;
    vertex a: λ_a(2λ_a - 1); edge (a,b): 4 λ_a λ_b."""
    x = np.atleast_2d(x)
    lam = np.stack([1 - x.sum(1), x[:, 0], x[:, 1], x[:, 2]], 1)
    dlam = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
    nq = len(x)
    shp = np.zeros((nq, 10))
    grad = np.zeros((nq, 10, 3))
    for a in range(4):
        shp[:, a] = lam[:, a] * (2 * lam[:, a] - 1)
        grad[:, a, :] = (4 * lam[:, a] - 1)[:, None] * dlam[a]
    for k, (a, b) in enumerate(TET.faces[1]):
        a -= 1
        b -= 1
        shp[:, 4 + k] = 4 * lam[:, a] * lam[:, b]
        grad[:, 4 + k, :] = 4 * (lam[:, a][:, None] * dlam[b] + lam[:, b][:, None] * dlam[a])
    return shp, grad


def tet_monomial_integral(a, b, c):
    """∫ x^a y^b z^c over the reference tet = a! b! c! / (a+b+c+3)!"""
    return math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)


def p2_tet_stiffness_exact():
    """Reference-tet P2 stiffness ∫∇φ_a·∇φ_b by exact polynomial integration (no
    quadrature): the gradients are linear in (x, y, z)."""
    # each gradient component = c0 + c1 x + c2 y + c3 z; coefficients from the
    # tabulation at 4 points (exact for linear functions)
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
    _, g = p2_tet_tabulate(pts)                     # [4 x 10 x 3]
    # linear function values at the vertices -> monomial coefficients
    M = np.array([[1, 0, 0, 0], [1, 1, 0, 0], [1, 0, 1, 0], [1, 0, 0, 1]], dtype=float)
    coef = np.einsum("pv,vad->pad", np.linalg.inv(M), g)  # [4 coefs x 10 x 3]
    I = np.zeros((4, 4))
    mono = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    for i, mi in enumerate(mono):
        for j, mj in enumerate(mono):
            I[i, j] = tet_monomial_integral(*(np.add(mi, mj)))
    return np.einsum("iad,ij,jbd->ab", coef, I, coef)


def p1_tet_tabulate(x):
    x = np.atleast_2d(x)
    lam = np.stack([1 - x.sum(1), x[:, 0], x[:, 1], x[:, 2]], 1)
    dlam = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
    return lam, np.broadcast_to(dlam, (len(x), 4, 3)).copy()


def tet_poisson(partition, xyz, order=2, kappa=None, fsrc=None, qx=None, qw=None,
                dirichlet_entities=None, g=None):
    """Cell matrices and vectors of ∫κ∇v·∇u, ∫f v for P1/P2 on the affine tets of the
    simplexified Cartesian grid with node coordinates xyz: DF from the 4 vertices,
    physical gradients DF⁻ᵀ∇̂φ, measure |det DF| (the affine cell map of
    src/Geometry/UnstructuredGrids.jl with the integration of
    src/CellData/CellQuadratures.jl:138-160), FP64.  g: callable for the Dirichlet
    values at the dof nodes (vertices, edge midpoints); the cell vectors are then
    lifted (attach_dirichlet)."""
    D = 3
    if dirichlet_entities is None:
        dirichlet_entities = boundary_entities(D)
    cells, cd, nfree, ndiri, edges = tet_numbering(partition, dirichlet_entities, order)
    C = len(cells)
    kappa = np.ones(C) if kappa is None else np.asarray(kappa, dtype=float)
    fsrc = np.ones(C) if fsrc is None else np.asarray(fsrc, dtype=float)
    shp, grad = (p2_tet_tabulate if order == 2 else p1_tet_tabulate)(qx)
    X = xyz[cells]                                       # C x 4 x 3
    J = np.transpose(X[:, 1:, :] - X[:, :1, :], (0, 2, 1))   # J_ij = ∂x_i/∂ξ_j
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)
    G = np.einsum("qad,cdi->cqai", grad, Jinv)           # physical gradients
    dV = np.abs(det)[:, None] * qw[None, :]              # C x nq
    mats = np.einsum("c,cq,cqai,cqbi->cab", kappa, dV, G, G)
    vecs = np.einsum("c,cq,qa->ca", fsrc, dV, shp)
    dvals = np.zeros(ndiri)
    if g is not None:
        nloc = cd.shape[1]
        nodes = np.concatenate([X, 0.5 * (X[:, [a - 1 for a, b in TET.faces[1]]] +
                                          X[:, [b - 1 for a, b in TET.faces[1]]])], axis=1)
        for c in range(C):
            for a in range(nloc):
                if cd[c, a] < 0:
                    dvals[-cd[c, a] - 1] = g(nodes[c, a])
        for c in range(C):
            m = cd[c] < 0
            if m.any():
                vecs[c] -= mats[c][:, m] @ dvals[-cd[c, m] - 1]
    return dict(cells=cells, cell_dofs=cd, nfree=nfree, ndiri=ndiri, cellmats=mats,
                cellvecs=vecs, edges=edges, dirichlet_vals=dvals)


def p2_tet_poisson(partition, xyz, **kw):
    return tet_poisson(partition, xyz, 2, **kw)


# ---- quadrature -----------------------------------------------------------------------

def _tet_orbit_points(w1, a1, w2, a2, w3, b):
    pts, wts = [], []
    for w, a in ((w1, a1), (w2, a2)):                   # S31 orbits: (a, a, a, 1-3a)
        for k in range(4):
            lam = [a] * 4
            lam[k] = 1 - 3 * a
            pts.append(lam[1:])
            wts.append(w)
    for i in range(4):                                   # S22 orbit: (b, b, ½-b, ½-b)
        for j in range(i + 1, 4):
            lam = [0.5 - b] * 4
            lam[i] = lam[j] = b
            pts.append(lam[1:])
            wts.append(w3)
    return np.array(pts), np.array(wts)


def tet_degree5_rule():
    """The symmetric 14-point rule on the reference tet exact to degree 5 (two S31
    orbits and one S22 orbit, interior points, positive weights): Gridap's
    Quadrature(TET, 4) and (TET, 5) use a rule of this structure
    (src/ReferenceFEs/Quadratures.jl:157-175, Witherden & Vincent 2015).  Derived
    here from the moment equations by Newton's method; the interior,
    positive-weight root is unique for this orbit structure."""
    from scipy.optimize import least_squares
    exps = [(a, b, c) for a in range(6) for b in range(6) for c in range(6) if a + b + c <= 5]

    def res(p):
        P, W = _tet_orbit_points(*p)
        return [np.sum(W * P[:, 0] ** a * P[:, 1] ** b * P[:, 2] ** c) - tet_monomial_integral(a, b, c)
                for a, b, c in exps]

    r = least_squares(res, [0.012, 0.09, 0.019, 0.31, 0.007, 0.045], xtol=3e-16, ftol=3e-16,
                      gtol=3e-16)
    return _tet_orbit_points(*r.x)


# ---- Cartesian conforming (vector) Lagrangian spaces -------------------------------------

def cartesian_cells(partition):
    """Cell node ids (0-based, local x fastest) of a CartesianDiscreteModel."""
    D = len(partition)
    n = [p + 1 for p in partition]
    out = []
    for idx in np.ndindex(*reversed(partition)):
        ci = list(reversed(idx))
        nodes = []
        for a in range(2 ** D):
            off = [(a >> d) & 1 for d in range(D)]
            node, stride = 0, 1
            for d in range(D):
                node += (ci[d] + off[d]) * stride
                stride *= n[d]
            nodes.append(node)
        out.append(nodes)
    return np.array(out, dtype=np.int64)


def lattice_ijk(partition, node):
    n = [p + 1 for p in partition]
    out, r = [], node
    for d in range(len(partition)):
        out.append(r % n[d])
        r //= n[d]
    return tuple(out)


def qk_vector_own_dofs(poly, order, ncomp):
    """Own local dofs per face of the vector Lagrangian Q_k element: each own node
    contributes its ncomp dofs, node-major (local dof = node + nnodes * comp)."""
    own, nnodes = lagrangian_own_dofs(poly, order)
    vown = [[[n + nnodes * c for n in nodes for c in range(ncomp)] for nodes in per]
            for per in own]
    comp = [ld // nnodes for ld in range(nnodes * ncomp)]
    return vown, nnodes * ncomp, comp


def cartesian_numbering(partition, order, ncomp, tag_entities, comp_masks=None):
    """Conforming numbering of a (vector) Q_k space on a Cartesian grid with the
    first-touch grid topology; tag_entities: list of entity sets, one per Dirichlet
    tag (a face takes the first tag holding its entity)."""
    D = len(partition)
    poly = HEX if D == 3 else QUAD
    cells = cartesian_cells(partition)
    d_faces = [cells + 1]
    for d in range(1, D):
        d_faces.append(first_touch_faces(cells, poly.faces[d])[0])
    d_faces.append(np.arange(1, len(cells) + 1).reshape(-1, 1))
    d_tags = []
    for d in range(D + 1):
        nf = int(d_faces[d].max())
        tags = [0] * nf
        for c in range(len(cells)):
            for lf, lv in enumerate(poly.faces[d]):
                ent = simplex_face_entity(D, partition,
                                          [lattice_ijk(partition, cells[c][v - 1]) for v in lv])
                t = 0
                for k, ents in enumerate(tag_entities):
                    if ent in ents:
                        t = k + 1
                        break
                tags[d_faces[d][c, lf] - 1] = t
        d_tags.append(tags)
    own, nloc, comp = qk_vector_own_dofs(poly, order, ncomp)
    cd, nfree, ndiri, dtag = conforming_cell_dofs(d_faces, d_tags, own, nloc, comp, comp_masks)
    return cells, cd, nfree, ndiri, dtag


# ---- isotropic linear elasticity with vector Q2 on boxes -----------------------------------

def lame(E, nu):
    """λ, μ of an isotropic material (test/GridapTests/IsotropicDamageTests.jl:13-17)."""
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def elasticity_reference(oracle, h, nu, qdeg=4, order=2):
    """Element matrix of ∫ ε(v) : σ(ε(u)) for E = 1 on a box of sizes h, vector Q_k
    (component-blocked local dofs), and ∫ φ_a over the box: σ = λ tr(ε) I + 2 μ ε, so
    K[(a,i),(b,j)] = Σ_q w|J| [λ g_ai g_bj + μ (g_aj g_bi + δ_ij g_a·g_b)], g = ∂φ/∂x."""
    D = len(h)
    qx, qw = oracle.tensor_quadrature(D, qdeg)
    shp, grad = oracle.qk_tabulate(D, order, qx)
    nn = shp.shape[1]
    g = grad / np.asarray(h)[None, None, :]               # physical gradients (box)
    det = float(np.prod(h))
    lam, mu = lame(1.0, nu)
    K = np.zeros((D * nn, D * nn))
    for i in range(D):
        for j in range(D):
            blk = lam * np.einsum("q,qa,qb->ab", qw, g[:, :, i], g[:, :, j]) \
                + mu * np.einsum("q,qa,qb->ab", qw, g[:, :, j], g[:, :, i])
            if i == j:
                blk = blk + mu * np.einsum("q,qak,qbk->ab", qw, g, g)
            K[i * nn:(i + 1) * nn, j * nn:(j + 1) * nn] = det * blk
    m = det * (qw @ shp)                                  # ∫ φ_a
    return K, m


def elasticity_q2_box(oracle, partition, E, nu, body_force, tag_entities, g=None,
                      domain=None, qdeg=4):
    """Cell matrices / vectors of the Cartesian vector-Q2 elasticity problem:
    K_c = E_c K̂, F_c[(a,k)] = f_k ∫φ_a, Dirichlet data g(x) (callable -> D-vector) at
    the Q2 nodes, lifted into F (attach_dirichlet)."""
    D = len(partition)
    domain = domain or tuple(v for _ in range(D) for v in (0.0, 1.0))
    h = [(domain[2 * d + 1] - domain[2 * d]) / partition[d] for d in range(D)]
    cells, cd, nfree, ndiri, _ = cartesian_numbering(partition, 2, D, tag_entities)
    Khat, m = elasticity_reference(oracle, h, nu, qdeg)
    E = np.asarray(E, dtype=float)
    mats = E[:, None, None] * Khat[None]
    nn = len(m)
    F = np.concatenate([body_force[k] * m for k in range(D)])
    vecs = np.repeat(F[None], len(cells), axis=0)
    dvals = np.zeros(ndiri)
    if g is not None:
        nodes = oracle.qk_nodes(D, 2)                    # reference Q2 nodes
        for c in range(len(cells)):
            x0 = np.array([domain[2 * d] + lattice_ijk(partition, cells[c][0])[d] * h[d]
                           for d in range(D)])
            for a in range(D * nn):
                if cd[c, a] < 0:
                    x = x0 + nodes[a % nn] * np.asarray(h)
                    dvals[-cd[c, a] - 1] = g(x)[a // nn]
        for c in range(len(cells)):
            msk = cd[c] < 0
            if msk.any():
                vecs[c] = vecs[c] - mats[c][:, msk] @ dvals[-cd[c, msk] - 1]
    return dict(cells=cells, cell_dofs=cd, nfree=nfree, ndiri=ndiri, cellmats=mats,
                cellvecs=vecs, dirichlet_vals=dvals, Khat=Khat, m=m, h=h)


# ---- n-cube elements on the isoparametric Q1 map ---------------------------------------
# get_cell_map (src/Geometry/Grids.jl:234-241): x(ξ) = Σ_v X_v N_v(ξ) with the Q1 basis of
# the cell's vertices; IntegrationMap (src/Fields/FieldsInterfaces.jl:737-760) sums the
# integrand times w|det J| over the points; gradients are pushed forward with J⁻ᵀ
# (push_∇, src/Fields/FieldsInterfaces.jl:94-98).

def physical_gradients(X, qw, ggrad, grad):
    """w|det J| [C, nq] and the physical gradients J⁻ᵀ∇̂φ [C, nq, nn, D] of the FE basis,
    J[c, q] = Σ_v X[c, v] ⊗ ggrad[q, v] (X: [C, nv, D] vertex coordinates; ggrad:
    [nq, nv, D] Q1 gradients; grad: [nq, nn, D] FE gradients)."""
    X = np.asarray(X, dtype=np.float64)
    J = np.einsum("cvd,qve->cqde", X, np.asarray(ggrad, dtype=np.float64))
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)                                  # [c, q, e, d]
    g = np.einsum("cqed,qae->cqad", Jinv, np.asarray(grad, dtype=np.float64))
    return np.asarray(qw, dtype=np.float64)[None, :] * np.abs(det), g


def ncube_diffusion_cells(wdet, g, kappa=None):
    """[C, nn, nn]: κ_c Σ_q w|det J| ∇φ_a·∇φ_b."""
    K = np.einsum("cq,cqad,cqbd->cab", wdet, g, g)
    return K if kappa is None else np.asarray(kappa, dtype=np.float64)[:, None, None] * K


def ncube_elasticity_cells(wdet, g, nu, E=None):
    """[C, D·nn, D·nn] with component-blocked local dofs (node + nn·comp):
    K[(a,i),(b,j)] = E_c Σ_q w|det J| [λ g_ai g_bj + μ (g_aj g_bi + δ_ij g_a·g_b)],
    λ, μ of E = 1 and ν (lame)."""
    C, nq, nn, D = g.shape
    lam, mu = lame(1.0, nu)
    K = np.zeros((C, D * nn, D * nn))
    dot = np.einsum("cq,cqad,cqbd->cab", wdet, g, g)
    for i in range(D):
        for j in range(D):
            blk = lam * np.einsum("cq,cqa,cqb->cab", wdet, g[..., i], g[..., j]) \
                + mu * np.einsum("cq,cqa,cqb->cab", wdet, g[..., j], g[..., i])
            if i == j:
                blk = blk + mu * dot
            K[:, i * nn:(i + 1) * nn, j * nn:(j + 1) * nn] = blk
    return K if E is None else np.asarray(E, dtype=np.float64)[:, None, None] * K


def ncube_loads(wdet, shp, values):
    """[C, ncomp·nn]: values[c, k] Σ_q w|det J| φ_a for component k (component-blocked);
    values: [C, ncomp]."""
    m = np.einsum("cq,qa->ca", wdet, np.asarray(shp, dtype=np.float64))
    v = np.asarray(values, dtype=np.float64)
    return np.concatenate([v[:, k:k + 1] * m for k in range(v.shape[1])], axis=1)


def mixed_cells(wdet, g_test, g_trial, shp_test, shp_trial, kappa=None, sigma=None):
    """[C, nt, ns] cell matrices of a rectangular field pair (test basis ψ, trial basis φ
    on one cell, IntegrationMap of src/Fields/FieldsInterfaces.jl:737-760 with the two
    bases of a MultiFieldCellField block): entry [c, lj, li] =
    Σ_q w|det J| (κ_c ∇ψ_li·∇φ_lj + σ_c ψ_li φ_lj); g_*: [C, nq, n, D] physical gradients
    (physical_gradients), shp_*: [nq, n]; κ, σ per cell or None (term absent)."""
    C = wdet.shape[0]
    out = np.zeros((C, g_trial.shape[2], g_test.shape[2]))
    if kappa is not None:
        out += np.asarray(kappa, dtype=np.float64)[:, None, None] * \
            np.einsum("cq,cqbd,cqad->cba", wdet, g_trial, g_test)
    if sigma is not None:
        out += np.asarray(sigma, dtype=np.float64)[:, None, None] * \
            np.einsum("cq,qb,qa->cba", wdet, np.asarray(shp_trial, dtype=np.float64),
                      np.asarray(shp_test, dtype=np.float64))
    return out


def vector_laplacian_cells(wdet, g, kappa=None):
    """[C, D·nn, D·nn] component-blocked (node + nn·comp) cell matrices of the vector
    Laplacian ∫ κ ∇v⊙∇u: δ_ij times the scalar diffusion matrix."""
    K = ncube_diffusion_cells(wdet, g, kappa)
    D = g.shape[3]
    return np.einsum("ij,cab->ciajb", np.eye(D), K).reshape(K.shape[0], D * K.shape[1], -1)


def divergence_cells(wdet, shp_scalar, g_vector, scalar_test=True, kappa=None):
    """Cell matrices [C, nt, ns] of the divergence coupling between a scalar basis ψ
    (shp_scalar [nq, ns]) and the scalar basis of a vector space (physical gradients
    g_vector [C, nq, nn, D]; local dofs node + nn·comp), test/GridapTests/
    StokesTaylorHoodTests.jl:54 with κ per cell:
      scalar_test: entry [c, (b, j), a] = κ Σ_q w|det J| ψ_a ∂_j φ_b   (q (∇⋅u));
      otherwise:   entry [c, b, (a, i)] = κ Σ_q w|det J| ∂_i φ_a ψ_b   ((∇⋅v) p)."""
    shp = np.asarray(shp_scalar, dtype=np.float64)
    C, nq, nn, D = g_vector.shape
    B = np.einsum("cq,qa,cqbj->cjba", wdet, shp, g_vector).reshape(C, D * nn, shp.shape[1])
    if kappa is not None:
        B = np.asarray(kappa, dtype=np.float64)[:, None, None] * B
    return B if scalar_test else np.ascontiguousarray(B.transpose(0, 2, 1))


def q1_map_points(X, n1):
    """Physical coordinates [C, npts, D] of reference points under each cell's Q1 map
    (n1: [npts, nv] Q1 shape values at the points)."""
    return np.einsum("pv,cvd->cpd", np.asarray(n1, dtype=np.float64), X)


def poisson_qk_with_boundary(oracle, domain, partition, order, dirichlet_entities,
                             face_terms, degree, face_degree, g=None, f=None):
    """Scalar Q_order Poisson ∫ ∇v⋅∇u with a source f(x) and boundary terms on a
    Cartesian model, assembled as Gridap's AffineFEOperator does with a weak form over
    Ω and boundary triangulations (test/GridapTests/PoissonTests.jl:9-108): the
    conforming numbering, cell arrays then face arrays in one assembly
    (src/FESpaces/Assemblers.jl:485-541), the Dirichlet lifting of both
    (src/CellData/AttachDirichlet.jl:59-84) with g at the dof nodes.  face_terms:
    [(entity ids, alpha, beta, g(x, n), h(x, n))]; f: callable of x (None: 0)."""
    D = len(partition)
    model = oracle.cartesian_model(domain, partition)
    _, cd, nfree, ndiri, _ = cartesian_numbering(partition, order, 1,
                                                 [set(dirichlet_entities)])
    X = model["xyz"][model["cell_nodes"] - 1]
    C = X.shape[0]
    qx, qw = oracle.tensor_quadrature(D, degree)
    shp, grad = oracle.qk_tabulate(D, order, qx)
    n1q, g1 = oracle.qk_tabulate(D, 1, qx)
    wdet, gp = physical_gradients(X, qw, g1, grad)
    mats = ncube_diffusion_cells(wdet, gp)
    if f is None:
        vecs = np.zeros((C, shp.shape[1]))
    else:
        xq = q1_map_points(X, n1q)                                  # [C, nq, D]
        fq = np.asarray(f(xq.reshape(-1, D)), dtype=np.float64).reshape(C, -1)
        vecs = np.einsum("cq,cq,qa->ca", wdet, fq, shp)
    all_cd, all_m, all_v = [cd], [mats], [vecs]
    for tents, al, be, gf, hf in face_terms:
        faces = oracle.boundary_faces(model, tents)
        fm, fv = oracle.boundary_face_arrays(model, faces, face_degree, al, be, gf, hf,
                                             order=order)
        all_cd.append(cd[[fc[0] for fc in faces]])
        all_m.append(fm)
        all_v.append(fv)
    cd_all = np.ascontiguousarray(np.concatenate(all_cd), dtype=np.int32)
    m_all = np.ascontiguousarray(np.concatenate(all_m))
    v_all = np.ascontiguousarray(np.concatenate(all_v))
    dvals = np.zeros(ndiri)
    if g is not None and ndiri:
        n1, _ = oracle.qk_tabulate(D, 1, oracle.qk_nodes(D, order))
        P = q1_map_points(X, n1)                                    # [C, nn, D]
        neg = cd < 0
        dvals[-cd[neg] - 1] = [g(x) for x in P[neg]]
        v_all = oracle.attach_dirichlet(cd_all, m_all, dvals, v_all)
    colptr, rowval, nzval = oracle.assemble_csc(cd_all, m_all, nfree, nfree)
    b = oracle.assemble_vector(cd_all, v_all, nfree)
    return dict(model=model, cell_dofs=cd, nfree=nfree, ndiri=ndiri, colptr=colptr,
                rowval=rowval, nzval=nzval, b=b, dirichlet_vals=dvals)


def poisson_qk_vector_with_boundary(oracle, domain, partition, order, cell_dofs, nfree, ndiri,
                                    dvals, face_terms, degree, face_degree, f):
    """The vector-valued case of test/GridapTests/PoissonTests.jl (VectorValue{2}): every
    term acts component by component (∇v⊙∇u, v⊙u, v⊙(n⋅∇u), (n⋅∇v)⊙u), so each cell and
    face matrix is δ_ij times the scalar one and each load is the scalar load of its
    component.  cell_dofs [C, ncomp·nn] (components blocked), the Dirichlet values and
    counts are inputs (the numbering is pinned elsewhere); face_terms: [(entity ids,
    alpha, beta, g(x, n) -> vector, h(x, n) -> vector)]; f(x) -> vector.  Assembled as
    poisson_qk_with_boundary assembles the scalar case."""
    D = len(partition)
    model = oracle.cartesian_model(domain, partition)
    X = model["xyz"][model["cell_nodes"] - 1]
    C = X.shape[0]
    qx, qw = oracle.tensor_quadrature(D, degree)
    shp, grad = oracle.qk_tabulate(D, order, qx)
    n1q, g1 = oracle.qk_tabulate(D, 1, qx)
    wdet, gp = physical_gradients(X, qw, g1, grad)
    mats = vector_laplacian_cells(wdet, gp)
    xq = q1_map_points(X, n1q)                                      # [C, nq, D]
    fq = np.asarray(f(xq.reshape(-1, D)), dtype=np.float64).reshape(C, len(qw), D)
    vecs = np.concatenate([np.einsum("cq,cq,qa->ca", wdet, fq[..., i], shp) for i in range(D)],
                          axis=1)
    cd = np.asarray(cell_dofs)
    all_cd, all_m, all_v = [cd], [mats], [vecs]
    nn = shp.shape[1]
    for tents, al, be, gf, hf in face_terms:
        faces = oracle.boundary_faces(model, tents)
        fvs = []
        for i in range(D):
            gi = None if gf is None else (lambda x, n, i=i: gf(x, n)[i])
            hi = None if hf is None else (lambda x, n, i=i: hf(x, n)[i])
            fm, fv = oracle.boundary_face_arrays(model, faces, face_degree, al, be, gi, hi,
                                                 order=order)
            fvs.append(fv)
        blk = np.einsum("ij,fab->fiajb", np.eye(D), fm).reshape(len(faces), D * nn, D * nn)
        all_cd.append(cd[[fc[0] for fc in faces]])
        all_m.append(blk)
        all_v.append(np.concatenate(fvs, axis=1))
    cd_all = np.ascontiguousarray(np.concatenate(all_cd), dtype=np.int32)
    m_all = np.ascontiguousarray(np.concatenate(all_m))
    v_all = np.ascontiguousarray(np.concatenate(all_v))
    if ndiri:
        v_all = oracle.attach_dirichlet(cd_all, m_all, np.asarray(dvals, dtype=np.float64), v_all)
    colptr, rowval, nzval = oracle.assemble_csc(cd_all, m_all, nfree, nfree)
    b = oracle.assemble_vector(cd_all, v_all, nfree)
    return dict(colptr=colptr, rowval=rowval, nzval=nzval, b=b)
