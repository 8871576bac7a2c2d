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
    """Marker for vector-valued Lagrangian spaces: VectorValue[D, Float64]."""

    def __class_getitem__(cls, args):
        D = args[0] if isinstance(args, tuple) else args
        return ("VectorValue", int(D))


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

    def tabulate(self, D, pts):
        """Shape values [npts, nnodes] and reference gradients [npts, nnodes, D]."""
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


def ReferenceFE(name, value_type=Float64, order=1):
    if name is not lagrangian:
        raise NotImplementedError("only lagrangian reference elements are provided")
    return LagrangianRefFE(value_type, order)


def gauss_legendre01(n):
    """n-point Gauss-Legendre rule mapped from (-1,1) to (0,1), ascending
    (src/ReferenceFEs/TensorProductQuadratures.jl:68-79)."""
    x, w = np.polynomial.legendre.leggauss(n)
    return (x + 1.0) / 2.0, w / 2.0


class Quadrature:
    """Tensor-product Gauss-Legendre rule of a degree on the n-cube: div(degree,2)+1
    points per axis, x fastest (src/ReferenceFEs/Quadratures.jl:191,203-222)."""

    def __init__(self, D, degree):
        n = degree // 2 + 1
        x1, w1 = gauss_legendre01(n)
        idx = np.indices((n,) * D).reshape(D, -1, order="F")
        self.coords = np.stack([x1[idx[d]] for d in range(D)], axis=1)
        w = np.ones(idx.shape[1])
        for d in range(D):
            w = w * w1[idx[d]]
        self.weights = w
        self.degree = degree
        self.D = D
