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
Pairs between two different spaces (rectangular blocks) are not provided.
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
    return SparseMatrixCSC(m, n, cp.copy(), rv.copy(), nz)


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
        self._plans = {}    # compose_plan per (row fields, column fields, touched pairs)

    def _compose(self, sub, rg, cg, nrows, ncols):
        # the pair patterns are the fixed patterns of the single-field assemblers
        key = (tuple(rg), tuple(cg), tuple(sorted(k for k in sub if k[0] in rg and k[1] in cg)))
        if key not in self._plans:
            self._plans[key] = compose_plan(sub, rg, cg, nrows, ncols)
        return compose_blocks(sub, rg, cg, nrows, ncols, plan=self._plans[key])

    def get_rows(self):
        return np.arange(1, self.V.num_free_dofs() + 1)

    def get_cols(self):
        return np.arange(1, self.U.num_free_dofs() + 1)

    def _pair(self, i, j):
        from .assemblers import SparseMatrixAssembler
        Vi, Uj = self.V.spaces[i], self.U.spaces[j]
        if not _same_space(Vi, Uj):
            raise NotImplementedError("blocks between two different FE spaces (rectangular "
                                      "blocks) are not provided")
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

    def assemble(self, matcontrib, veccontrib, lift, A=None, b=None):
        if A is not None or b is not None:
            raise NotImplementedError("in-place multi-field assembly is not provided")
        mats = self._split(matcontrib, 2)
        vecs = self._split(veccontrib, 1)
        nf = len(self.V)
        sub = {}
        rhs = [np.zeros(V.nfree) for V in self.V.spaces]
        for (i, j) in sorted(mats):
            Aij, lij = self._pair(i, j).assemble(mats[(i, j)], None, lift=lift,
                                                 trial=self.U.spaces[j])
            sub[(i, j)] = Aij
            if lift:
                rhs[i] = rhs[i] + lij
        for (i,) in sorted(vecs):
            _, bi = self._pair(i, i).assemble(None, vecs[(i,)], lift=False)
            rhs[i] = rhs[i] + bi
        nrows = [V.nfree for V in self.V.spaces]
        ncols = [U.nfree for U in self.U.spaces]
        groups = self.V.groups()
        mat = vec = None
        if matcontrib is not None:
            blocks = [[self._compose(sub, rg, cg, nrows, ncols) for cg in self.U.groups()]
                      for rg in groups]
            mat = blocks[0][0] if not isinstance(self.V.style, BlockMultiFieldStyle) \
                else BlockSparseMatrix(blocks)
        if veccontrib is not None or lift:
            parts = [np.concatenate([rhs[i] for i in g]) if g else np.zeros(0) for g in groups]
            vec = parts[0] if not isinstance(self.V.style, BlockMultiFieldStyle) \
                else BlockVector(parts)
        return mat, vec
