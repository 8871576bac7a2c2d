"""Reference elements and quadratures on the host.

Gridap tabulates shape functions once per reference cell type
(src/Arrays/CompressedArrays.jl:66-94); the engine receives that tabulation
(weights, values, reference gradients at the quadrature points) through
gh_reffe_tabulation and keeps it on the device.
"""
import numpy as np


class _Lagrangian:
    def __repr__(self):
        return "lagrangian"


lagrangian = _Lagrangian()
Float64 = float


class VectorValue:
    """VectorValue[D, Float64] marks a vector-valued Lagrangian space;
    VectorValue(0.0, 0.0, -1.0) is a constant vector (a body force)."""

    def __class_getitem__(cls, args):
        D = args[0] if isinstance(args, tuple) else args
        return ("VectorValue", int(D))

    def __init__(self, *comps):
        self.comps = tuple(float(c) for c in comps)

    def __len__(self):
        return len(self.comps)

    def __iter__(self):
        return iter(self.comps)


class LagrangianRefFE:
    """ReferenceFE(lagrangian, T, order) on the n-cube of the model.

    Nodes follow src/ReferenceFEs/CLagrangianRefFEs.jl:626-764 (faces in face
    order, own nodes x fastest); the shape functions are the nodal Lagrange basis
    (inv(V)*monomials, src/ReferenceFEs/ReferenceFEInterfaces.jl:563-583), i.e. the
    tensor product of 1-D Lagrange polynomials through the nodes."""

    def __init__(self, value_type, order):
        self.order = int(order)
        if self.order < 1:
            raise ValueError("order must be >= 1")
        if isinstance(value_type, tuple) and value_type[0] == "VectorValue":
            self.ncomp = value_type[1]
        else:
            self.ncomp = 1

    def nodes(self, D):
        k = self.order
        out = []
        for d in range(D + 1):
            for mask in range(1 << D):
                if bin(mask).count("1") != d:
                    continue
                for anchor in range(1 << (D - d)):
                    fixed = {}
                    bit = 0
                    for ax in range(D):
                        if not mask >> ax & 1:
                            fixed[ax] = (anchor >> bit) & 1
                            bit += 1
                    free_axes = [ax for ax in range(D) if mask >> ax & 1]
                    nin = (k - 1) ** len(free_axes)
                    for t in range(nin):
                        r = t
                        x = np.zeros(D)
                        for ax in range(D):
                            if ax in fixed:
                                x[ax] = fixed[ax]
                            else:
                                x[ax] = (r % (k - 1) + 1) / k
                                r //= (k - 1)
                        out.append(x)
        return np.array(out)

    def tabulate(self, D, pts, simplex=False):
        """Shape values [npts, nnodes] and reference gradients [npts, nnodes, D]."""
        if simplex:
            return tabulate_simplex(D, self.order, pts)
        k = self.order
        nodes = self.nodes(D)
        pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
        grid = np.arange(k + 1) / k
        npts, nn = pts.shape[0], nodes.shape[0]
        shp = np.ones((npts, nn))
        vals = np.zeros((D, npts, nn))
        ders = np.zeros((D, npts, nn))
        for ax in range(D):
            j = np.rint(nodes[:, ax] * k).astype(int)
            t = pts[:, ax][:, None]
            v = np.ones((npts, nn))
            dv = np.zeros((npts, nn))
            xj = grid[j][None, :]
            for m in range(k + 1):
                xm = grid[m]
                use = (j != m)[None, :]
                f = np.where(use, (t - xm) / np.where(use, xj - xm, 1.0), 1.0)
                dv = np.where(use, dv * f + v / np.where(use, xj - xm, 1.0), dv)
                v = v * f
            vals[ax] = v
            ders[ax] = dv
        for ax in range(D):
            shp *= vals[ax]
        grad = np.ones((npts, nn, D))
        for g in range(D):
            for ax in range(D):
                grad[:, :, g] *= ders[ax] if ax == g else vals[ax]
        return shp, grad


# Local edges of TET in Gridap's face order (vertex pairs, 0-based).
_TET_EDGES = ((0, 1), (0, 2), (1, 2), (0, 3), (1, 3), (2, 3))


def tabulate_simplex(D, order, pts):
    """Nodal Lagrangian basis of order 1 or 2 on the reference TET (vertices (0,0,0),
    (1,0,0), (0,1,0), (0,0,1)); nodes: the vertices, then one per edge in TET edge
    order (src/ReferenceFEs/LagrangianRefFEs.jl).  With barycentrics λ = (1-x-y-z,
    x, y, z): P1 φ_a = λ_a; P2 φ_a = λ_a (2 λ_a - 1) at vertices, 4 λ_a λ_b on edges."""
    if D != 3 or order not in (1, 2):
        raise NotImplementedError("simplicial elements: P1 and P2 on TET")
    x = np.atleast_2d(np.asarray(pts, dtype=np.float64))
    lam = np.stack([1.0 - x.sum(1), x[:, 0], x[:, 1], x[:, 2]], axis=1)
    dlam = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    npts = x.shape[0]
    if order == 1:
        return lam.copy(), np.broadcast_to(dlam, (npts, 4, 3)).copy()
    shp = np.zeros((npts, 10))
    grad = np.zeros((npts, 10, 3))
    shp[:, :4] = lam * (2.0 * lam - 1.0)
    grad[:, :4, :] = (4.0 * lam - 1.0)[:, :, None] * dlam[None, :, :]
    for k, (a, b) in enumerate(_TET_EDGES):
        shp[:, 4 + k] = 4.0 * lam[:, a] * lam[:, b]
        grad[:, 4 + k, :] = 4.0 * (lam[:, a][:, None] * dlam[b] + lam[:, b][:, None] * dlam[a])
    return shp, grad


def ReferenceFE(name, value_type=Float64, order=1):
    if name is not lagrangian:
        raise NotImplementedError("only lagrangian reference elements are provided")
    return LagrangianRefFE(value_type, order)


def gauss_legendre01(n):
    """n-point Gauss-Legendre rule mapped from (-1,1) to (0,1), ascending
    (src/ReferenceFEs/TensorProductQuadratures.jl:68-79)."""
    x, w = np.polynomial.legendre.leggauss(n)
    return (x + 1.0) / 2.0, w / 2.0


# Witherden-Vincent symmetric TET rules, the ones Gridap's Quadrature(TET, degree)
# selects for degree <= 10 (src/ReferenceFEs/Quadratures.jl:157-165).  Published
# values (F.D. Witherden, P.E. Vincent, Comput. Math. Appl. 2015) on the bi-unit TET
# with vertices (-1,-1,-1), (1,-1,-1), (-1,1,-1), (-1,-1,1), as tabulated in
# src/ReferenceFEs/WitherdenVincentQuadratures.jl:367-400: a centre point of weight
# w0, orbits "S31" (w, s, t) -> points (s,s,t), (s,t,s), (t,s,s), (s,s,s), orbits
# "S22" (w, s, t) -> (s,t,t), (t,s,t), (s,s,t), (s,t,s), (t,s,s), (t,t,s), in this
# order (:471-513).  The map to the unit TET is x -> (x+1)/2, w -> w/2 (:1290-1299),
# then the weights are scaled to sum to the TET measure 1/6.
_WV_TET = {
    1: (4.0 / 3.0, (), ()),
    2: (None, ((0.33333333333333333333, -0.72360679774997896964, 0.17082039324993690892),),
        ()),
    3: (None, ((0.18162379004944980942, -0.34367339496723662642, -0.96897981509829012074),
               (0.15170954328388352391, -0.78390550020314279176, 0.35171650060942837529)),
        ()),
    4: (None, ((0.15025056762402113440, -0.37822816147339878041, -0.86531551557980365878),
               (0.097990724155149266058, -0.81452949937821754720, 0.44358849813465264159)),
        ((0.056728027702775288584, -0.90899259174870070102, -0.091007408251299298984),)),
}
_WV_TET[0] = _WV_TET[1]
_WV_TET[5] = _WV_TET[4]


def _tet_rule(degree):
    """Quadrature(TET, degree) for degrees 0-5: points (n x 3) and weights."""
    if degree not in _WV_TET:
        raise NotImplementedError("TET quadrature is provided for degrees 0 to 5")
    w0, s31, s22 = _WV_TET[degree]
    rows = []
    if w0 is not None:
        rows.append((w0, -0.5, -0.5, -0.5))
    for w, a, b in s31:
        for x in ((a, a, b), (a, b, a), (b, a, a), (a, a, a)):
            rows.append((w,) + x)
    for w, a, b in s22:
        for x in ((a, b, b), (b, a, b), (a, a, b), (a, b, a), (b, a, a), (b, b, a)):
            rows.append((w,) + x)
    wx = np.array(rows, dtype=np.float64)
    w = wx[:, 0] / 2.0
    pts = (wx[:, 1:] + 1.0) / 2.0
    tot = 0.0
    for v in w:                     # left-to-right sum, as a plain loop adds them
        tot += v
    return pts, w * ((1.0 / 6.0) / tot)


class Quadrature:
    """Quadrature(polytope, degree).  n-cube: tensor-product Gauss-Legendre rule,
    div(degree,2)+1 points per axis, x fastest (src/ReferenceFEs/Quadratures.jl:191,
    203-222).  simplex=True: the Witherden-Vincent TET rules above."""

    def __init__(self, D, degree, simplex=False):
        self.degree = degree
        self.D = D
        self.simplex = simplex
        if simplex:
            if D != 3:
                raise NotImplementedError("simplicial quadrature is provided on TET")
            self.coords, self.weights = _tet_rule(degree)
            return
        n = degree // 2 + 1
        x1, w1 = gauss_legendre01(n)
        idx = np.indices((n,) * D).reshape(D, -1, order="F")
        self.coords = np.stack([x1[idx[d]] for d in range(D)], axis=1)
        w = np.ones(idx.shape[1])
        for d in range(D):
            w = w * w1[idx[d]]
        self.weights = w


def facet_points(D, side, pts_f):
    """Points of the reference n-cube facet `side` (n-cube face order: pairs of
    opposite facets, normal axis from the last to the first) for facet-local points
    pts_f [n, D-1] (in the remaining axes, ascending)."""
    axis = D - 1 - side // 2
    pts = np.zeros((pts_f.shape[0], D))
    others = [k for k in range(D) if k != axis]
    for j, k in enumerate(others):
        pts[:, k] = pts_f[:, j]
    pts[:, axis] = float(side & 1)
    return pts


def facet_quadrature(D, degree):
    """The tensor Gauss-Legendre rule of `degree` on the reference facet (a (D-1)-cube),
    as Quadrature(facet polytope, degree): div(degree,2)+1 points per axis, x fastest."""
    q = Quadrature(D - 1, degree) if D > 2 else None
    if D == 2:
        x1, w1 = gauss_legendre01(degree // 2 + 1)
        return x1[:, None], w1
    return q.coords, q.weights
