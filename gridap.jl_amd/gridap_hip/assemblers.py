"""SparseMatrixAssembler, assemble_matrix/vector, AffineFEOperator, solve.

The Gridap entry points keep their names and meaning
(src/FESpaces/Assemblers.jl:302-400, src/FESpaces/SparseMatrixAssemblers.jl:
42-170, src/FESpaces/AffineFEOperators.jl:8-98, src/FESpaces/FESolvers.jl:65):
the result is a CSC matrix with 1-based Int64 colptr/rowval (Gridap's
SparseMatrixCSC{Float64,Int}) and a dense Float64 vector.  All numerics run on
the device through libgridap_hip; this module only moves Gridap's data across
the C ABI.
"""
import numpy as np

from . import _lib
from ._lib import Context
from .boundary import lower_boundary
from .celldata import CellField, CoefProduct, FEValue, PointField
from .celldata import (BodyForceTerm, DiffusionTerm, DomainContribution, ElasticityTerm,
                       MassTerm, SourceTerm, _cell_coef, get_fe_basis, get_trial_fe_basis)
from .geometry import BoundaryTriangulation
from .fespaces import FEFunction, TrialFESpace
from .reference_fes import Float64, Quadrature, ReferenceFE, lagrangian


class SparseMatrixCSC:
    """Julia SparseMatrixCSC layout: m x n, 1-based colptr[n+1], rowval[nnz], nzval."""

    def __init__(self, m, n, colptr, rowval, nzval):
        self.m, self.n = m, n
        self.colptr, self.rowval, self.nzval = colptr, rowval, nzval

    @property
    def shape(self):
        return (self.m, self.n)

    def nnz(self):
        return len(self.rowval)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.nzval, self.rowval - 1, self.colptr - 1),
                             shape=(self.m, self.n))

    def __getitem__(self, ij):
        i, j = ij  # 1-based, as in Julia
        lo, hi = self.colptr[j - 1] - 1, self.colptr[j] - 1
        k = lo + np.searchsorted(self.rowval[lo:hi], i)
        return self.nzval[k] if k < hi and self.rowval[k] == i else 0.0


def _space(U):
    return U.space if isinstance(U, TrialFESpace) else U


class SparseMatrixAssembler:
    """SparseMatrixAssembler(U, V) on one device.

    precision: "fp64" (element math in FP64) or "fp32" (element math in FP32,
    nzval/b accumulated in FP64); path: "auto", "general" or "structured"."""

    _PREC = {"fp64": _lib.GH_PREC_FP64, "fp32": _lib.GH_PREC_FP32_ELEM}
    _PATH = {"auto": _lib.GH_PATH_AUTO, "general": _lib.GH_PATH_GENERAL,
             "structured": _lib.GH_PATH_STRUCTURED}

    def __new__(cls, U, V, *args, **kwargs):
        from .multifield import MultiFieldFESpace, MultiFieldSparseMatrixAssembler
        if isinstance(U, MultiFieldFESpace) or isinstance(V, MultiFieldFESpace):
            return MultiFieldSparseMatrixAssembler(U, V, *args, **kwargs)
        return super().__new__(cls)

    def __init__(self, U, V, device=0, precision="fp64", path="auto"):
        Us, Vs = _space(U), _space(V)
        if Us is not Vs:
            raise NotImplementedError("trial and test spaces must share one numbering")
        self.U, self.V = U, V
        self.space = Vs
        self.precision = self._PREC[precision]
        self.path = self._PATH[path]
        self.device = device
        self.ctx = None
        self.nnz = None
        self.colptr = self.rowval = None
        self._degree = None
        # the SparseMatrixCSC whose stored values the device holds (None: the device
        # values are not those of any matrix handed out, e.g. after a vector-only
        # assembly); solvers load their matrix into the device when it is not this one
        self.device_matrix = None
        self._face_degree = None    # degree of the face tabulation on the device
        self._has_boundary = False  # boundary terms uploaded
        self._has_fq = False        # a source at the quadrature points uploaded
        self._has_sigma = False     # a cell mass coefficient uploaded

    def get_rows(self):
        return self.space.get_free_dof_ids()

    get_cols = get_rows

    # device set-up ---------------------------------------------------------
    def _setup(self, degree):
        V = self.space
        model = V.model
        self._simplex = bool(getattr(V, "simplex", False))
        if self.ctx is None:
            ctx = Context(self.device)
            D = model.D
            if model.is_affine_cartesian():
                ctx.mesh_cartesian(model.origin, model.sizes, model.partition)
            else:
                cn = model.cell_node_ids()
                ptrs = (np.arange(cn.shape[0] + 1, dtype=np.int64) * cn.shape[1] + 1)
                ctx.mesh_upload(model.node_coordinates(), cn.ravel(), ptrs.astype(np.int32))
            cd = V.cell_dof_ids
            ptrs = (np.arange(cd.shape[0] + 1, dtype=np.int64) * cd.shape[1] + 1)
            kind = _lib.GH_REFFE_PK if self._simplex else _lib.GH_REFFE_QK
            ctx.space_upload(kind, V.order, V.ncomp, cd.ravel(),
                             ptrs.astype(np.int32), V.nfree, V.ndirichlet)
            self.nnz = ctx.symbolic()
            self.colptr, self.rowval = ctx.pattern_get()
            self.ctx = ctx
            self._D = D
        if degree != self._degree:
            q = Quadrature(model.D, degree, simplex=self._simplex)
            shp, grad = V.reffe.tabulate(model.D, q.coords, simplex=self._simplex)
            self.ctx.reffe_tabulation(q.weights, shp, grad)
            if not self._simplex and V.order != 1 and not model.is_affine_cartesian():
                # the cell map is the Q1 map of the cell's vertices (get_cell_map)
                _, ggrad = ReferenceFE(lagrangian, Float64, 1).tabulate(model.D, q.coords)
                self.ctx.geometry_tabulation(ggrad)
            self._degree = degree

    @staticmethod
    def _split(contrib):
        """(cell terms, boundary terms) of a DomainContribution (or None)."""
        if contrib is None:
            return None, []
        cell = [(t, m) for t, m in contrib.terms if not isinstance(m.trian, BoundaryTriangulation)]
        face = [(t, m) for t, m in contrib.terms if isinstance(m.trian, BoundaryTriangulation)]
        return (DomainContribution(cell) if cell else None), face

    def _lower(self, matcontrib, veccontrib):
        """Weak form -> (form kind, per-cell coefficient, per-cell source, degree, extra);
        boundary terms go to extra["boundary"] (lower_boundary's per-face data)."""
        ncells = self.space.model.num_cells()
        had_mat = matcontrib is not None
        matcontrib, mface = self._split(matcontrib)
        veccontrib, vface = self._split(veccontrib)
        if had_mat and matcontrib is None:
            raise NotImplementedError("a bilinear form needs a cell (volume) term")
        kappa, fsrc, sigma, degs = None, None, None, set()
        have_stiff = False
        point_sources = []  # (coefficient, degree) of sources f(x) given as functions
        nu, force = None, None
        have_mat = have_vec = False
        kinds = set()
        if matcontrib is not None:
            for term, meas in matcontrib.terms:
                if isinstance(term, ElasticityTerm):
                    kinds.add(_lib.GH_FORM_ELASTICITY)
                    if nu is not None and nu != term.nu:
                        raise NotImplementedError("one Poisson ratio per form")
                    nu = term.nu
                elif isinstance(term, DiffusionTerm):
                    kinds.add(_lib.GH_FORM_DIFFUSION)
                elif isinstance(term, MassTerm):
                    # σ u v: the mass coefficient of the diffusion form
                    kinds.add(_lib.GH_FORM_DIFFUSION)
                    coef = _cell_coef(term.coef)
                    sg = coef.per_cell(ncells) if coef is not None else None
                    sg = np.ones(ncells) if sg is None else sg
                    sigma = sg if sigma is None else sigma + sg
                    degs.add(meas.degree)
                    have_mat = True
                    continue
                else:
                    raise NotImplementedError("bilinear forms: κ ∇v⋅∇u, σ u v or ε(v) : σ(ε(u))")
                coef = _cell_coef(term.coef)
                k = coef.per_cell(ncells) if coef is not None else None
                k = np.ones(ncells) if k is None else k
                kappa = k if kappa is None else kappa + k
                degs.add(meas.degree)
                have_mat = have_stiff = True
        if veccontrib is not None:
            for term, meas in veccontrib.terms:
                if isinstance(term, BodyForceTerm):
                    kinds.add(_lib.GH_FORM_ELASTICITY)
                    f = np.asarray(term.force, dtype=np.float64)
                    force = f if force is None else force + f
                elif isinstance(term, SourceTerm):
                    kinds.add(_lib.GH_FORM_DIFFUSION)
                    if term.coef is not None and not isinstance(term.coef, CellField):
                        point_sources.append((term.coef, meas.degree))
                        degs.add(meas.degree)
                        have_vec = True
                        continue
                    coef = _cell_coef(term.coef)
                    f = coef.per_cell(ncells) if coef is not None else None
                    f = np.ones(ncells) if f is None else f
                    fsrc = f if fsrc is None else fsrc + f
                else:
                    raise NotImplementedError("linear forms: f v or v⋅f")
                degs.add(meas.degree)
                have_vec = True
        if kinds == {_lib.GH_FORM_DIFFUSION, _lib.GH_FORM_ELASTICITY} and nu is None and \
                self.space.ncomp > 1 and fsrc is None and not point_sources:
            # ∇v⊙∇u with v⋅f on a vector space: the vector Laplacian and its body force
            kinds = {_lib.GH_FORM_DIFFUSION}
        if len(kinds) > 1:
            raise NotImplementedError("one form kind per operator")
        if len(degs) > 1:
            raise NotImplementedError("matrix and vector terms must share one Measure")
        if not degs and (mface or vface):
            # boundary terms only (e.g. a Neumann load): the cell part is empty
            form = _lib.GH_FORM_DIFFUSION
            kinds.add(form)
            degs.add(2 * self.space.order)
            fsrc = np.zeros(ncells)
        if not degs:
            raise ValueError("empty weak form")
        form = kinds.pop()
        if have_mat and kappa is not None and np.all(kappa == 1.0):
            kappa = None
        if form == _lib.GH_FORM_DIFFUSION:
            if have_vec and fsrc is not None and np.all(fsrc == 1.0):
                fsrc = None
            if not have_vec:
                fsrc = np.zeros(ncells)
        if not have_stiff:
            kappa = np.zeros(ncells)
        extra = dict(nu=0.3 if nu is None else nu,
                     force=np.zeros(3) if force is None else force, sigma=sigma)
        if point_sources:
            # f at the quadrature points: the per-cell sources join it, fsrc is unused
            deg = next(iter(degs))
            fq = sum(_eval_cell_points(cf, self.space, deg) for cf, _ in point_sources)
            if fsrc is not None:
                if fq.ndim != 2:
                    raise NotImplementedError("per-cell and point sources on a vector space")
                fq = fq + fsrc[:, None]
            extra["fq"] = np.ascontiguousarray(fq)
            fsrc = np.zeros(ncells)
        if mface or vface:
            if form != _lib.GH_FORM_DIFFUSION and not (self.space.ncomp > 1 and not mface):
                raise NotImplementedError("boundary terms are provided with the diffusion "
                                          "form (vector spaces: loads with either form)")
            extra["boundary"] = lower_boundary(self.space.model, self.space.reffe, mface, vface)
        return form, kappa, fsrc, degs.pop(), extra

    def _upload_boundary(self, bnd):
        if bnd is None:
            if self._has_boundary:
                self.ctx.boundary_upload(np.zeros(0, np.int32), np.zeros(0, np.int32))
                self._has_boundary = False
            return
        fs = bnd["faces"]
        key = (fs.nqf, fs.qw.tobytes())
        if self._face_degree != key:
            self.ctx.face_tabulation(fs.qw, fs.shp, fs.grad)
            self._face_degree = key
        self.ctx.boundary_upload(fs.cells + 1, fs.sides + 1, bnd["alpha"], bnd["beta"],
                                 bnd["gq"], bnd["hq"])
        self._has_boundary = True

    def assemble(self, matcontrib, veccontrib, lift, A=None, b=None, trial=None,
                 copy_values=True):
        """Device numeric phase; returns (SparseMatrixCSC, b).  With A (a matrix on this
        assembler's pattern) and/or b given, their values are overwritten in place and
        they are returned (fillstored! then the numeric loop,
        src/FESpaces/SparseMatrixAssemblers.jl:60-63).  trial: the space whose Dirichlet
        values are lifted (default: the assembler's U).  copy_values=False leaves the
        matrix values on the device and returns (None, b)."""
        form, kappa, fsrc, degree, extra = self._lower(matcontrib, veccontrib)
        self._setup(degree)
        dvals = None
        if lift:
            U = self.U if trial is None else trial
            dv = U.dirichlet_values() if isinstance(U, TrialFESpace) else None
            if dv is not None and dv.size and np.any(dv != 0):
                dvals = dv
        self.ctx.coef_upload(kappa, fsrc, dvals)
        sigma = extra.get("sigma")
        if sigma is not None or self._has_sigma:
            self.ctx.mass_upload(sigma)
            self._has_sigma = sigma is not None
        fq = extra.get("fq")
        if fq is not None or self._has_fq:
            self.ctx.source_qp_upload(fq)
            self._has_fq = fq is not None
        self._upload_boundary(extra.get("boundary"))
        if form == _lib.GH_FORM_ELASTICITY or self.space.ncomp > 1:
            # the body force loads elasticity and the vector Laplacian
            self.ctx.elasticity_params(extra["nu"], extra["force"])
        if A is not None and not (A.shape == (self.space.nfree, self.space.nfree) and
                                  np.array_equal(A.colptr, self.colptr) and
                                  np.array_equal(A.rowval, self.rowval)):
            # Gridap's add_entry! fails on an entry missing from the stored pattern
            raise ValueError("A does not have this assembler's sparsity pattern")
        self.ctx.assemble(form, self.precision, self.path)
        if (matcontrib is None or not copy_values) and A is None:
            # a vector-only assembly: the matrix values stay on the device
            _, bv = self.ctx.values_get(None, b, want_nz=False)
            self.device_matrix = None
            return None, bv
        nz, bv = self.ctx.values_get(None if A is None else A.nzval, b)
        if A is None:
            n = self.space.nfree
            A = SparseMatrixCSC(n, n, self.colptr.copy(), self.rowval.copy(), nz)
        self.device_matrix = A if matcontrib is not None else None
        return A, bv

    def load_matrix(self, A):
        """Make the device hold A's stored values (A carries this assembler's pattern),
        so that the device SpMV and PCG work on A."""
        if self.device_matrix is A:
            return
        if not (A.shape == (self.space.nfree, self.space.nfree) and
                np.array_equal(A.colptr, self.colptr) and np.array_equal(A.rowval, self.rowval)):
            raise ValueError("A does not have this assembler's sparsity pattern")
        self.ctx.values_set(np.ascontiguousarray(A.nzval, dtype=np.float64))
        self.device_matrix = A


def _eval_cell_points(coef, V, degree):
    """A coefficient at the quadrature points of every cell, [ncells, nq] (scalar
    spaces) or [ncells, nq, ncomp] (vector spaces): functions of x at the physical points
    (CartesianMap, or the isoparametric map of the nodes), FE functions by their cell
    values and the basis, per-cell values broadcast."""
    model = V.model
    if V.simplex or V.order not in (1, 2) or V.ncomp not in (1, model.D):
        raise NotImplementedError("sources f(x) are provided for Q1 and Q2 spaces (scalar, or "
                                  "one component per dimension) on n-cube meshes")
    out = _cell_points(coef, V, degree)                       # [c, q, k], k = 1 or ncomp
    if V.ncomp == 1:
        if out.shape[2] != 1:
            raise ValueError("a vector-valued source on a scalar space")
        return out[:, :, 0]
    return np.ascontiguousarray(np.broadcast_to(out, out.shape[:2] + (V.ncomp,)))


def _cell_points(coef, V, degree):
    model = V.model
    D = model.D
    q = Quadrature(D, degree)
    shp, _ = V.reffe.tabulate(D, q.coords)                    # [q, a]: the scalar basis
    # the cell map: the Q1 basis of the cell's vertices (get_cell_map)
    gshp = shp if V.order == 1 else ReferenceFE(lagrangian, Float64, 1).tabulate(D, q.coords)[0]
    C, nq = model.num_cells(), len(q.weights)
    if isinstance(coef, CoefProduct):
        out = np.ones((C, nq, 1))
        for c in coef.factors:
            out = out * _cell_points(c, V, degree)
        return out
    if isinstance(coef, CellField):
        pc = coef.per_cell(C)
        return np.ones((C, nq, 1)) * (1.0 if pc is None else pc[:, None, None])
    if isinstance(coef, PointField):
        X = model.node_coordinates()[model.cell_node_ids() - 1]   # [c, v, D]
        x = np.einsum("qv,cvd->cqd", gshp, X)
        val = np.asarray(coef.func(x.reshape(-1, D)), dtype=np.float64)
        k = val.size // (C * nq)
        return np.broadcast_to(val.reshape(-1), (C * nq * k,)).reshape(C, nq, k).copy()
    if isinstance(coef, FEValue):
        uh = coef.uh
        cd = np.asarray(V.cell_dof_ids, dtype=np.int64)
        fv = np.asarray(uh.free_values, dtype=np.float64)
        dv = np.asarray(uh.dirichlet_values, dtype=np.float64)
        uc = np.where(cd > 0, fv[np.maximum(cd, 1) - 1],
                      dv[np.maximum(-cd, 1) - 1] if dv.size else 0.0)
        nc = V.ncomp
        uc = uc.reshape(C, nc, -1)                             # component-blocked local dofs
        return np.einsum("cka,qa->cqk", uc, shp)
    raise NotImplementedError("unsupported source coefficient %r" % (coef,))


def _contrib(form, *args):
    out = form(*args) if callable(form) else form
    if not isinstance(out, DomainContribution):
        raise TypeError("weak form must return integrate(...) contributions")
    return out


def _is_assembler(x):
    from .multifield import MultiFieldSparseMatrixAssembler
    return isinstance(x, (SparseMatrixAssembler, MultiFieldSparseMatrixAssembler))


def assemble_matrix(a, U, V=None, assem=None):
    """assemble_matrix(a, U, V) or assemble_matrix(a, assem, U, V)
    (src/FESpaces/Assemblers.jl:372-375)."""
    if _is_assembler(U):
        assem, U, V = U, V, assem
    V = U if V is None else V
    assem = assem or SparseMatrixAssembler(U, V)
    m = _contrib(a, get_trial_fe_basis(U), get_fe_basis(V))
    return assem.assemble(m, None, lift=False)[0]


def assemble_vector(l, V, assem=None):
    """assemble_vector(l, V) (src/FESpaces/Assemblers.jl:382-385)."""
    if _is_assembler(V):
        assem, V = V, assem
    assem = assem or SparseMatrixAssembler(V, V)
    vcon = _contrib(l, get_fe_basis(V))
    return assem.assemble(None, vcon, lift=False)[1]


def assemble_matrix_and_vector(a, l, U, V, assem=None):
    """assemble_matrix_and_vector(a, l, U, V) with Dirichlet lifting
    (src/FESpaces/Assemblers.jl:392-400,496-541)."""
    assem = assem or SparseMatrixAssembler(U, V)
    m = _contrib(a, get_trial_fe_basis(U), get_fe_basis(V))
    vcon = _contrib(l, get_fe_basis(V))
    return assem.assemble(m, vcon, lift=True)


def assemble_matrix_inplace(a, A, assem, U=None, V=None):
    """assemble_matrix!(f, A, assem, U, V) (src/FESpaces/Assemblers.jl:308-312 ->
    SparseMatrixAssemblers.jl:60-63): overwrite the stored values of A, which must carry
    assem's pattern, with the assembly of a(u, v).  The pattern is not rebuilt, so a
    nonlinear (Picard or Newton) loop re-assembles its matrix at the cost of the numeric
    phase only."""
    U = assem.U if U is None else U
    V = assem.V if V is None else V
    m = _contrib(a, get_trial_fe_basis(U), get_fe_basis(V))
    return assem.assemble(m, None, lift=False, A=A)[0]


def assemble_vector_inplace(l, b, assem, V=None):
    """assemble_vector!(f, b, assem, V) (src/FESpaces/Assemblers.jl:319-322): overwrite b
    (a contiguous float64 vector of num_free_dofs entries) with the assembly of l(v)."""
    V = assem.V if V is None else V
    vcon = _contrib(l, get_fe_basis(V))
    return assem.assemble(None, vcon, lift=False, b=b)[1]


def assemble_matrix_and_vector_inplace(a, l, A, b, assem, U=None, V=None):
    """assemble_matrix_and_vector!(f, l, A, b, assem, U, V) with the Dirichlet lifting
    of U (src/FESpaces/Assemblers.jl:330-334,342-346): overwrite A's stored values and
    b in place."""
    U = assem.U if U is None else U
    V = assem.V if V is None else V
    m = _contrib(a, get_trial_fe_basis(U), get_fe_basis(V))
    vcon = _contrib(l, get_fe_basis(V))
    return assem.assemble(m, vcon, lift=True, A=A, b=b)


class AffineFEOperator:
    """AffineFEOperator(a, l, U, V[, assem]) or AffineFEOperator(weakform, U, V[, assem])
    with weakform(u, v) -> (a_contribution, l_contribution)
    (src/FESpaces/AffineFEOperators.jl:23-54)."""

    def __init__(self, *args, assem=None):
        if len(args) >= 4 and callable(args[0]) and callable(args[1]):
            a, l, U, V = args[:4]
            if len(args) > 4:
                assem = args[4]
            weakform = lambda u, v: (a(u, v), l(v))
        else:
            weakform, U, V = args[:3]
            if len(args) > 3:
                assem = args[3]
        self.trial, self.test = U, V
        self.assem = assem or SparseMatrixAssembler(U, V)
        u, v = get_trial_fe_basis(U), get_fe_basis(V)
        mcon, vcon = weakform(u, v)
        self.matrix, self.vector = self.assem.assemble(mcon, vcon, lift=True)

    def get_matrix(self):
        return self.matrix

    def get_vector(self):
        return self.vector

    def get_trial(self):
        return self.trial

    def get_test(self):
        return self.test


def get_matrix(op):
    return op.get_matrix()


def get_vector(op):
    return op.get_vector()


class PCGSolver:
    """FP64 preconditioned CG on the device (Jacobi by default); stands in for
    Gridap's LUSolver (src/Algebra/LinearSolvers.jl:249-274)."""

    def __init__(self, rtol=1e-12, maxit=100000, precond="jacobi"):
        self.rtol = rtol
        self.maxit = maxit
        self.precond = {"none": _lib.GH_PRECOND_NONE,
                        "jacobi": _lib.GH_PRECOND_JACOBI,
                        "fsai": _lib.GH_PRECOND_FSAI}[precond]
        self.stats = None

    def solve(self, assem, b, A=None):
        """x with A x = b on the device; A (default: the matrix the device holds) must
        carry assem's pattern and is loaded into the device when it is not the one there."""
        if A is not None:
            assem.load_matrix(A)
        elif assem.device_matrix is None:
            raise RuntimeError("no matrix on the device: pass A")
        x, st = assem.ctx.pcg(b=b, precond=self.precond, rtol=self.rtol, maxit=self.maxit)
        self.stats = st
        return x


LinearFESolver = PCGSolver


def solve(op, solver=None):
    """uh = solve(op) (src/FESpaces/FESolvers.jl:65-69)."""
    solver = solver or PCGSolver()
    x = solver.solve(op.assem, op.vector, op.matrix)
    return FEFunction(op.trial, x)
