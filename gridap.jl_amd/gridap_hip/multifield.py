"""MultiFieldFESpace and block assembly (src/MultiField/MultiFieldFESpaces.jl,
src/MultiField/BlockSparseMatrixAssemblers.jl).

A multi-field weak form is a sum of terms between one test field i and one trial field
j, e.g. ∫(∇(u1)⋅∇(v1) + u2⋅v2 - u1⋅v2)dΩ
(test/MultiFieldTests/BlockSparseMatrixAssemblersTests.jl:101-104).  Gridap assembles
only the touched field pairs (its cell matrices are ArrayBlocks with touched flags), so
the pattern of the pair (i, j) is the single-field pattern of the two spaces and
nothing else.  Each touched pair is assembled on the device by the single-field
engine (the same kernels; one device context per distinct field space, reused by every
pair on it), and the pairs are laid out as Gridap lays them out:

* ConsecutiveMultiFieldStyle (the default): one matrix over the free dofs of field 1,
  then field 2, ...; column (j, k) holds the rows of every touched pair (i, j) of
  column k, shifted by the free-dof offset of field i (the rows stay sorted because the
  offsets grow with i);
* BlockMultiFieldStyle(NB, SB, P): NB x NB blocks, block b holding the fields
  P[SB_1 + ... + SB_{b-1} + 1 : ...] consecutively (get_block_ranges,
  src/MultiField/MultiFieldFESpaces.jl); each block matrix is laid out like the
  consecutive one over its fields.

The right-hand side of test field i is its source terms minus the lifting
Σ_j A_ij(:, D_j) u_Dj of every touched pair (attach_dirichlet per field pair).
Pairs between two different spaces on one mesh (rectangular blocks, e.g. a Q2 and a Q1
scalar field coupled by ∫ σ u1 v2 or ∫ κ ∇u1⋅∇v2) are assembled on the device by a
pair context that holds the test space as rows and the trial space as columns
(gh_trial_space_upload), and so are the divergence couplings q (∇⋅u) and (∇⋅v) p
between a scalar and a vector field (the velocity-pressure blocks of a Taylor–Hood
Stokes form, test/GridapTests/StokesTaylorHoodTests.jl:54).
"""
import numpy as np

from .celldata import DomainContribution


class ConsecutiveMultiFieldStyle:
    """The free dofs of all fields in one range, field after field."""


class BlockMultiFieldStyle:
    """BlockMultiFieldStyle(NB, SB, P): NB blocks; SB[b] fields in block b; P the
    1-based permutation of the fields (src/MultiField/MultiFieldFESpaces.jl)."""

    def __init__(self, NB=None, SB=None, P=None):
        self.NB, self.SB, self.P = NB, SB, P

    def parameters(self, nfields):
        NB = nfields if self.NB is None else int(self.NB)
        SB = tuple([1] * NB) if self.SB is None else tuple(int(s) for s in self.SB)
        P = tuple(range(1, nfields + 1)) if self.P is None else tuple(int(p) for p in self.P)
        if len(SB) != NB or sum(SB) != nfields or sorted(P) != list(range(1, nfields + 1)):
            raise ValueError("BlockMultiFieldStyle(%r, %r, %r) does not fit %d fields"
                             % (self.NB, self.SB, self.P, nfields))
        return NB, SB, P

    def block_fields(self, nfields):
        """0-based field indices of each block (get_block_ranges)."""
        NB, SB, P = self.parameters(nfields)
        out, k = [], 0
        for b in range(NB):
            out.append([p - 1 for p in P[k:k + SB[b]]])
            k += SB[b]
        return out


class MultiFieldFESpace:
    """MultiFieldFESpace([V1, V2, ...]; style) of test or trial spaces."""

    def __init__(self, spaces, style=None):
        self.spaces = list(spaces)
        if not self.spaces:
            raise ValueError("a MultiFieldFESpace needs at least one field")
        self.style = ConsecutiveMultiFieldStyle() if style is None else style

    def __len__(self):
        return len(self.spaces)

    def __iter__(self):
        return iter(self.spaces)

    def __getitem__(self, i):
        return self.spaces[i]

    def num_fields(self):
        return len(self.spaces)

    def num_free_dofs(self):
        return int(sum(V.nfree for V in self.spaces))

    @property
    def nfree(self):
        return self.num_free_dofs()

    def field_offsets(self):
        """Free-dof offset of each field in the consecutive layout."""
        return np.concatenate([[0], np.cumsum([V.nfree for V in self.spaces])]).astype(np.int64)

    def groups(self):
        """Fields of each block (one block for the consecutive style)."""
        if isinstance(self.style, BlockMultiFieldStyle):
            return self.style.block_fields(len(self.spaces))
        return [list(range(len(self.spaces)))]


class BlockSparseMatrix:
    """NB x NB blocks of SparseMatrixCSC (Gridap's BlockMatrix of the block assembler)."""

    def __init__(self, blocks):
        self.blocks = blocks

    @property
    def shape(self):
        return (sum(r[0].m for r in self.blocks), sum(B.n for B in self.blocks[0]))

    def nnz(self):
        return sum(B.nnz() for r in self.blocks for B in r)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.bmat([[B.to_scipy() for B in r] for r in self.blocks], format="csc")


class BlockVector:
    """NB blocks of a vector (Gridap's BlockVector of the block assembler)."""

    def __init__(self, blocks):
        self.blocks = blocks

    def to_array(self):
        return np.concatenate(self.blocks)


def _underlying(U):
    return getattr(U, "space", U)


def _same_space(A, B):
    a, b = _underlying(A), _underlying(B)
    if a is b:
        return True
    return (a.nfree == b.nfree and a.ndirichlet == b.ndirichlet and
            np.array_equal(np.asarray(a.cell_dof_ids), np.asarray(b.cell_dof_ids)))


def compose_plan(sub, row_fields, col_fields, nrows, ncols):
    """The layout of compose_blocks for the patterns of sub: (m, n, colptr, rowval,
    [(pair, destination indices)]).  It depends on the patterns only, so an assembler
    reuses it for every assembly on its fixed patterns."""
    roff = np.concatenate([[0], np.cumsum([nrows[i] for i in row_fields])]).astype(np.int64)
    colptrs, rowvals, scatter = [np.ones(1, dtype=np.int64)], [], []
    base = 0
    for j in col_fields:
        nc = ncols[j]
        parts = [(r, (i, j)) for r, i in enumerate(row_fields) if (i, j) in sub]
        cnt = np.zeros(nc, dtype=np.int64)
        for _, key in parts:
            cnt += np.diff(sub[key].colptr)
        start = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
        total = int(start[-1])
        rv = np.empty(total, dtype=np.int64)
        done = np.zeros(nc, dtype=np.int64)   # entries already placed in each column
        for r, key in parts:
            B = sub[key]
            c = np.diff(B.colptr)
            first = np.repeat(start[:-1] + done, c)
            within = np.arange(B.nnz(), dtype=np.int64) - np.repeat(B.colptr[:-1] - 1, c)
            dst = first + within
            rv[dst] = B.rowval + roff[r]
            scatter.append((key, dst + base))
            done += c
        colptrs.append(start[1:] + 1 + base)
        rowvals.append(rv)
        base += total
    cp = np.concatenate(colptrs)
    rv = np.concatenate(rowvals) if rowvals else np.zeros(0, dtype=np.int64)
    return int(roff[-1]), int(sum(ncols[j] for j in col_fields)), cp, rv, scatter


def compose_blocks(sub, row_fields, col_fields, nrows, ncols, plan=None):
    """One matrix over the fields row_fields x col_fields from the field-pair matrices
    sub[(i, j)] (SparseMatrixCSC, 1-based; missing pairs are untouched): column
    (j, k) is the concatenation over i of the rows of sub[(i, j)] column k, each shifted
    by the offset of field i within row_fields.  plan: compose_plan of the same
    patterns (computed here when None)."""
    from .assemblers import SparseMatrixCSC
    m, n, cp, rv, scatter = plan or compose_plan(sub, row_fields, col_fields, nrows, ncols)
    nz = np.empty(len(rv), dtype=np.float64)
    for key, dst in scatter:
        nz[dst] = sub[key].nzval
    # the pattern arrays are the plan's, shared read-only by every matrix composed on it
    cp.flags.writeable = False
    rv.flags.writeable = False
    return SparseMatrixCSC(m, n, cp, rv, nz)


class RectangularPairAssembler:
    """The block (i, j) of a multi-field form whose test field i and trial field j live
    in different spaces on one mesh: rows are Vi's free dofs, columns Uj's
    (SparseMatrixAssemblers.jl:174-248 with cell_rows and cell_cols from two spaces).
    Cell terms κ ∇u⋅∇v and σ u v between scalar spaces, or the divergence coupling
    κ q (∇⋅u) / κ (∇⋅v) p between a scalar and a vector space (κ constant per cell);
    the lifting of Uj's Dirichlet values goes to the test field's right-hand side."""

    def __init__(self, U, V, device=0, precision="fp64"):
        from . import _lib
        self.U, self.V = U, V
        self.device = device
        self.precision = {"fp64": _lib.GH_PREC_FP64, "fp32": _lib.GH_PREC_FP32_ELEM}[precision] \
            if isinstance(precision, str) else precision
        self.ctx = None
        self._degree = None
        self.colptr = self.rowval = None

    def _setup(self, degree):
        from . import _lib
        from .reference_fes import Float64, Quadrature, ReferenceFE, lagrangian
        Vs, Us = _underlying(self.V), _underlying(self.U)
        if getattr(Vs, "simplex", False) or getattr(Us, "simplex", False):
            raise NotImplementedError("rectangular pairs are provided on quadrilaterals and "
                                      "hexahedra")
        model = Vs.model
        if self.ctx is None:
            ctx = _lib.Context(self.device)
            if model.is_affine_cartesian():
                ctx.mesh_cartesian(model.origin, model.sizes, model.partition)
            else:
                cn = model.cell_node_ids()
                ptrs = (np.arange(cn.shape[0] + 1, dtype=np.int64) * cn.shape[1] + 1)
                ctx.mesh_upload(model.node_coordinates(), cn.ravel(), ptrs.astype(np.int32))
            for up, S in ((ctx.space_upload, Vs), (ctx.trial_space_upload, Us)):
                cd = np.asarray(S.cell_dof_ids)
                ptrs = (np.arange(cd.shape[0] + 1, dtype=np.int64) * cd.shape[1] + 1)
                up(_lib.GH_REFFE_QK, S.order, S.ncomp, cd.ravel(), ptrs.astype(np.int32),
                   S.nfree, S.ndirichlet)
            ctx.symbolic()
            self.colptr, self.rowval = ctx.pattern_get()
            self.ctx = ctx
        if degree != self._degree:
            D = model.D
            q = Quadrature(D, degree)
            shp, grad = Vs.reffe.tabulate(D, q.coords)
            self.ctx.reffe_tabulation(q.weights, shp, grad)
            tshp, tgrad = Us.reffe.tabulate(D, q.coords)
            self.ctx.trial_tabulation(tshp, tgrad)
            if not model.is_affine_cartesian():
                _, ggrad = ReferenceFE(lagrangian, Float64, 1).tabulate(D, q.coords)
                self.ctx.geometry_tabulation(ggrad)
            self._degree = degree

    def assemble(self, matcontrib, veccontrib, lift, trial=None, copy_values=True):
        from . import _lib
        from .assemblers import SparseMatrixCSC
        from .celldata import DiffusionTerm, DivergenceTerm, MassTerm, _cell_coef
        from .fespaces import TrialFESpace
        if veccontrib is not None:
            raise NotImplementedError("sources belong to the test field's own pair")
        Vs, Us = _underlying(self.V), _underlying(self.U)
        ncells = Vs.model.num_cells()
        kappa = sigma = None
        degs, kinds = set(), set()
        for term, meas in (matcontrib.terms if matcontrib is not None else []):
            if isinstance(term, DivergenceTerm):
                want = (1, Vs.model.D) if term.part == "q_divu" else (Vs.model.D, 1)
                if (Vs.ncomp, Us.ncomp) != want:
                    raise NotImplementedError("the divergence coupling pairs a scalar and a "
                                              "vector space as q (∇⋅u) or (∇⋅v) p")
                kinds.add(_lib.GH_FORM_DIVERGENCE)
            elif isinstance(term, (DiffusionTerm, MassTerm)):
                kinds.add(_lib.GH_FORM_DIFFUSION)
            else:
                raise NotImplementedError("rectangular pairs: κ ∇u⋅∇v, σ u v and divergence "
                                          "cell terms")
            coef = _cell_coef(term.coef)
            k = coef.per_cell(ncells) if coef is not None else None
            k = np.ones(ncells) if k is None else k
            if isinstance(term, MassTerm):
                sigma = k if sigma is None else sigma + k
            else:
                kappa = k if kappa is None else kappa + k
            degs.add(meas.degree)
        if len(degs) != 1 or len(kinds) != 1:
            raise NotImplementedError("the terms of a block must share one Measure and kind")
        form = kinds.pop()
        self._setup(degs.pop())
        dvals = None
        U = self.U if trial is None else trial
        if lift and isinstance(U, TrialFESpace):
            dv = U.dirichlet_values()
            if dv is not None and dv.size and np.any(dv != 0):
                dvals = dv
        self.ctx.coef_upload(kappa, None, dvals)
        self.ctx.mass_upload(sigma)
        self.ctx.assemble(form, self.precision, _lib.GH_PATH_GENERAL)
        if not copy_values:
            return self.ctx.values_get(None, None, want_nz=False)
        nz, b = self.ctx.values_get()
        A = SparseMatrixCSC(Vs.nfree, Us.nfree, self.colptr.copy(), self.rowval.copy(), nz)
        return A, b


class MultiFieldSparseMatrixAssembler:
    """SparseMatrixAssembler(X, Y) for multi-field spaces: the touched field pairs on
    the device, laid out by the style of Y (and X)."""

    def __init__(self, U, V, device=0, precision="fp64", path="auto"):
        if not (isinstance(U, MultiFieldFESpace) and isinstance(V, MultiFieldFESpace)):
            raise TypeError("multi-field assembly needs MultiFieldFESpace trial and test spaces")
        if len(U) != len(V):
            raise ValueError("trial and test spaces have different numbers of fields")
        if type(U.style) is not type(V.style):
            raise ValueError("trial and test spaces must share one MultiFieldStyle")
        self.U, self.V = U, V
        self.device, self.precision, self.path = device, precision, path
        self._assems = []   # (test space, single-field assembler)
        self._rects = []    # ((test space, trial space), rectangular pair assembler)
        self._plans = {}    # compose_plan per (row fields, column fields, touched pairs)
        # per plan key: (context holding the composite on the device, {pair: block})
        self._composers = {}

    @staticmethod
    def _key(pairs, rg, cg):
        return (tuple(rg), tuple(cg), tuple(sorted(k for k in pairs if k[0] in rg and k[1] in cg)))

    def _compose(self, sub, rg, cg, nrows, ncols):
        # the pair patterns are the fixed patterns of the single-field assemblers
        key = self._key(sub, rg, cg)
        if key not in self._plans:
            self._plans[key] = compose_plan(sub, rg, cg, nrows, ncols)
        return compose_blocks(sub, rg, cg, nrows, ncols, plan=self._plans[key])

    def _register(self, key):
        """Compose the plan of key on the device from now on: one context holds the
        composite values and every touched pair's assembled values move there."""
        from ._lib import Context
        _, _, _, rv, scatter = self._plans[key]
        if not scatter or key in self._composers:
            return
        ctx = Context(self.device)
        blocks = {}
        for idx, (pair, dst) in enumerate(scatter):
            ctx.compose_block(self._pair(*pair).ctx, dst, len(rv))
            blocks[pair] = idx
        self._composers[key] = (ctx, blocks)

    def _compose_device(self, key, T=None):
        """The composite of key from the device, into T's stored values when given."""
        from .assemblers import SparseMatrixCSC
        m, n, cp, rv, _ = self._plans[key]
        if T is not None:
            self._check_target(T, m, n, cp, rv)
        if key in self._composers:
            nz = self._composers[key][0].compose_values_get(None if T is None else T.nzval)
        else:
            nz = np.empty(0)
        if T is not None:
            return T
        cp.flags.writeable = False
        rv.flags.writeable = False
        return SparseMatrixCSC(m, n, cp, rv, nz)

    def get_rows(self):
        return np.arange(1, self.V.num_free_dofs() + 1)

    def get_cols(self):
        return np.arange(1, self.U.num_free_dofs() + 1)

    def _pair(self, i, j):
        from .assemblers import SparseMatrixAssembler
        Vi, Uj = self.V.spaces[i], self.U.spaces[j]
        if not _same_space(Vi, Uj):
            for (test, trial), assem in self._rects:
                if _same_space(test, Vi) and _same_space(trial, Uj):
                    return assem
            assem = RectangularPairAssembler(Uj, Vi, device=self.device,
                                             precision=self.precision)
            self._rects.append(((Vi, Uj), assem))
            return assem
        for space, assem in self._assems:
            if _same_space(space, Vi):
                return assem
        assem = SparseMatrixAssembler(Uj, Vi, device=self.device, precision=self.precision,
                                      path=self.path)
        self._assems.append((Vi, assem))
        return assem

    @staticmethod
    def _split(contrib, kind):
        """{fields: DomainContribution} of a multi-field contribution."""
        out = {}
        if contrib is None:
            return out
        for term, meas in contrib.terms:
            f = term.fields
            if f is None or len(f) != kind or any(x is None for x in f):
                raise NotImplementedError("multi-field terms must involve one test field%s"
                                          % (" and one trial field" if kind == 2 else ""))
            out.setdefault(tuple(f), []).append((term, meas))
        return {k: DomainContribution(v) for k, v in out.items()}

    def _targets(self, A, nr, nc):
        """The blocks of an in-place matrix A, checked against this assembler's layout."""
        if isinstance(self.V.style, BlockMultiFieldStyle):
            if not (isinstance(A, BlockSparseMatrix) and len(A.blocks) == nr and
                    all(len(row) == nc for row in A.blocks)):
                raise ValueError("A must be a BlockSparseMatrix of %d x %d blocks" % (nr, nc))
            return A.blocks
        from .assemblers import SparseMatrixCSC
        if not isinstance(A, SparseMatrixCSC):
            raise ValueError("A must be a SparseMatrixCSC")
        return [[A]]

    @staticmethod
    def _check_target(T, m, n, cp, rv):
        # Gridap's add_entry! fails on an entry missing from the stored pattern
        if not (T.m == m and T.n == n and (T.colptr is cp or np.array_equal(T.colptr, cp)) and
                (T.rowval is rv or np.array_equal(T.rowval, rv))):
            raise ValueError("A does not have this assembler's sparsity pattern")
        if not (isinstance(T.nzval, np.ndarray) and T.nzval.dtype == np.float64 and
                T.nzval.flags.c_contiguous and T.nzval.flags.writeable):
            raise ValueError("A's stored values must be a writeable contiguous float64 array")

    def assemble(self, matcontrib, veccontrib, lift, A=None, b=None):
        """Returns (matrix, vector) in the layout of the spaces' style.  With A and/or b
        given (a matrix of this assembler's layout from an earlier assembly, a vector of
        num_free_dofs entries or a BlockVector), their values are overwritten in place
        and they are returned (assemble_matrix_and_vector!, src/FESpaces/Assemblers.jl:
        330-346)."""
        if A is not None and matcontrib is None:
            raise ValueError("an in-place matrix needs a bilinear form")
        mats = self._split(matcontrib, 2)
        vecs = self._split(veccontrib, 1)
        sub = {}
        rhs = [np.zeros(V.nfree) for V in self.V.spaces]
        groups, cgroups = self.V.groups(), self.U.groups()
        keys = {(tuple(rg), tuple(cg)): self._key(mats, rg, cg) for rg in groups for cg in cgroups}
        # composed on the device once every touched plan has been composed on the host
        # (the first assembly of a form on this assembler builds the pair patterns)
        ondev = matcontrib is not None and all(
            k in self._composers or (k in self._plans and not self._plans[k][4])
            for k in keys.values())
        for (i, j) in sorted(mats):
            pa = self._pair(i, j)
            Aij, lij = pa.assemble(mats[(i, j)], None, lift=lift, trial=self.U.spaces[j],
                                   copy_values=not ondev)
            if ondev:
                # move the pair's values now: a later pair may share its assembler
                for k in keys.values():
                    if (i, j) in k[2]:
                        ctx, blocks = self._composers[k]
                        ctx.compose_move(blocks[(i, j)])
            sub[(i, j)] = Aij
            if lift:
                rhs[i] = rhs[i] + lij
        for (i,) in sorted(vecs):
            _, bi = self._pair(i, i).assemble(None, vecs[(i,)], lift=False)
            rhs[i] = rhs[i] + bi
        nrows = [V.nfree for V in self.V.spaces]
        ncols = [U.nfree for U in self.U.spaces]
        mat = vec = None
        blocked = isinstance(self.V.style, BlockMultiFieldStyle)
        if matcontrib is not None:
            targets = None if A is None else self._targets(A, len(groups), len(cgroups))
            blocks = []
            for r, rg in enumerate(groups):
                row = []
                for c, cg in enumerate(cgroups):
                    T = None if targets is None else targets[r][c]
                    if ondev:
                        B = self._compose_device(keys[(tuple(rg), tuple(cg))], T)
                    else:
                        B = self._compose(sub, rg, cg, nrows, ncols)
                        if T is not None:
                            self._check_target(T, B.m, B.n, B.colptr, B.rowval)
                            T.nzval[...] = B.nzval
                            B = T
                    row.append(B)
                blocks.append(row)
            if not ondev:
                for k in keys.values():
                    self._register(k)
            mat = A if A is not None else (BlockSparseMatrix(blocks) if blocked else blocks[0][0])
        if veccontrib is not None or lift:
            parts = [np.concatenate([rhs[i] for i in g]) if g else np.zeros(0) for g in groups]
            if b is not None:
                outs = b.blocks if blocked and isinstance(b, BlockVector) else [b]
                if len(outs) != len(parts) or any(
                        not isinstance(o, np.ndarray) or o.shape != p.shape for o, p in
                        zip(outs, parts)):
                    raise ValueError("b does not have this assembler's layout")
                for o, p in zip(outs, parts):
                    o[...] = p
                vec = b
            else:
                vec = BlockVector(parts) if blocked else parts[0]
        return mat, vec
