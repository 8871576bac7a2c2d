"""ctypes binding of libgridap_hip.so (the C ABI in include/gridap_hip.h).

The HIP library is the only compute path: importing this module without the
built library raises, and a failed device call raises GridapHipError with the
library's message.  There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgridap_hip.so")

GH_OK = 0
GH_ERR_ARG = 1
GH_ERR_RANGE = 2
GH_ERR_HIP = 3
GH_ERR_OOM = 4
GH_ERR_STATE = 5
GH_ERR_NOCONV = 6
GH_ERR_UNSUPPORTED = 7

GH_REFFE_QK = 1
GH_REFFE_PK = 2
GH_FORM_DIFFUSION = 1
GH_PREC_FP64 = 0
GH_PREC_FP32_ELEM = 1
GH_FORM_ELASTICITY = 2
GH_FORM_PLAPLACIAN = 3
GH_FORM_DIVERGENCE = 4
GH_PATH_AUTO = 0
GH_PATH_GENERAL = 1
GH_PATH_STRUCTURED = 2
GH_KERNEL_NONE = 0
GH_KERNEL_Q1_LATTICE = 1
GH_KERNEL_CELL = 2
GH_KERNEL_COLUMN_SUM = 3
GH_KERNEL_LIST_SUM = 4
GH_KERNEL_Q2_LATTICE = 5
GH_PRECOND_NONE = 0
GH_PRECOND_JACOBI = 1
GH_PRECOND_FSAI = 2
GH_NOALLOC_MAX_N = 262144

# Every symbol the header declares, with (restype, argtypes).
_i32, _i64, _dbl, _vp = C.c_int32, C.c_int64, C.c_double, C.c_void_p
_pi64 = C.POINTER(C.c_int64)


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32),
                ("rel_residual", C.c_double), ("device_ms", C.c_double),
                ("precond_used", C.c_int32), ("reserved", C.c_int32)]


class PatternInfo(C.Structure):
    _fields_ = [("nnz", C.c_int64), ("nranges", C.c_int64), ("lattice", C.c_int32),
                ("box", C.c_int32), ("elemlist", C.c_int32), ("compact", C.c_int32),
                ("metric", C.c_int32), ("inverse", C.c_int32), ("last_kernel", C.c_int32),
                ("fallback_nodes", C.c_int32)]


SIGNATURES = {
    "gh_create": (_i32, [C.POINTER(_vp), _i32]),
    "gh_destroy": (_i32, [_vp]),
    "gh_last_error": (C.c_char_p, [_vp]),
    "gh_stream": (_i32, [_vp, C.POINTER(_vp)]),
    "gh_mesh_cartesian": (_i32, [_vp, _i32, _vp, _vp, _vp]),
    "gh_mesh_upload": (_i32, [_vp, _i32, _i64, _vp, _i64, _vp, _vp]),
    "gh_space_upload": (_i32, [_vp, _i32, _i32, _i32, _vp, _vp, _i64, _i64, _i64]),
    "gh_trial_space_upload": (_i32, [_vp, _i32, _i32, _i32, _vp, _vp, _i64, _i64, _i64]),
    "gh_trial_tabulation": (_i32, [_vp, _i32, _i32, _vp, _vp]),
    "gh_reffe_tabulation": (_i32, [_vp, _i32, _i32, _vp, _vp, _vp]),
    "gh_geometry_tabulation": (_i32, [_vp, _i32, _i32, _vp]),
    "gh_face_tabulation": (_i32, [_vp, _i32, _i32, _vp, _vp, _vp]),
    "gh_boundary_upload": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gh_symbolic": (_i32, [_vp, _pi64]),
    "gh_pattern_get": (_i32, [_vp, _vp, _vp]),
    "gh_coef_upload": (_i32, [_vp, _vp, _vp, _vp]),
    "gh_mass_upload": (_i32, [_vp, _vp]),
    "gh_assemble": (_i32, [_vp, _i32, _i32, _i32]),
    "gh_source_qp_upload": (_i32, [_vp, _vp]),
    "gh_values_get": (_i32, [_vp, _vp, _vp]),
    "gh_values_set": (_i32, [_vp, _vp, _vp]),
    "gh_compose_block": (_i32, [_vp, _vp, _vp, _i64, _i64]),
    "gh_compose_move": (_i32, [_vp, _i32]),
    "gh_compose_values_get": (_i32, [_vp, _vp]),
    "gh_compose_clear": (_i32, [_vp]),
    "gh_fefunction_upload": (_i32, [_vp, _vp, _vp]),
    "gh_plaplacian_params": (_i32, [_vp, _dbl]),
    "gh_pattern_info_get": (_i32, [_vp, C.POINTER(PatternInfo)]),
    "gh_numeric": (_i32, [_vp, _i32, _vp, _vp, _i32, _vp, _vp, _vp]),
    "gh_spmv": (_i32, [_vp, _vp, _vp]),
    "gh_pcg": (_i32, [_vp, _i32, _vp, _vp, _dbl, _i32, C.POINTER(SolveStats)]),
    "gh_pcg_noalloc": (_i32, [_vp, _i32, _vp, _vp, _dbl, _i32, C.POINTER(SolveStats)]),
    "gh_mlp_nparams": (_i32, [_pi64]),
    "gh_mlp_init": (_i32, [_vp, C.c_uint64]),
    "gh_mlp_set": (_i32, [_vp, _vp, _i64]),
    "gh_mlp_get": (_i32, [_vp, _vp, _i64]),
    "gh_grf_kappa": (_i32, [_vp, _i32, _vp, _vp, _dbl]),
    "gh_fsai_build": (_i32, [_vp]),
    "gh_fsai_train_step": (_i32, [_vp, _i32, C.c_uint64, C.c_float, C.POINTER(_dbl)]),
    "gh_fsai_get": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "gh_fsai_factor_get": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "gh_elasticity_params": (_i32, [_vp, _dbl, _vp]),
}

_lib = None


class GridapHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gridap_hip error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Load libgridap_hip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libgridap_hip.so is not built (expected at %s); run "
                "`python gridap.jl_amd/build.py` or __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    """Raw pointer of a contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One device context (gh_ctx): mesh, space, pattern and values in HBM."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        st = L.gh_create(C.byref(h), int(device))
        if st != GH_OK:
            raise GridapHipError(st, "gh_create failed (no usable HIP device %d)" % device)
        self.h = h
        self.device = device
        self.nfree = 0  # sizes of the arrays the library returns, tracked here so the
        self.nnz = 0    # host buffers always match what the library writes
        self.ncols = 0  # matrix columns: nfree, or the trial space's free dofs
        self.ncells = 0

    def check(self, st):
        if st != GH_OK:
            msg = lib().gh_last_error(self.h)
            raise GridapHipError(st, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            lib().gh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stream(self):
        s = C.c_void_p()
        self.check(lib().gh_stream(self.h, C.byref(s)))
        return s.value

    # thin wrappers ---------------------------------------------------------
    def mesh_cartesian(self, origin, sizes, partition):
        D = len(partition)
        o = np.ascontiguousarray(origin, dtype=np.float64)
        s = np.ascontiguousarray(sizes, dtype=np.float64)
        p = np.ascontiguousarray(partition, dtype=np.int64)
        self.check(lib().gh_mesh_cartesian(self.h, D, ptr(o), ptr(s), ptr(p)))
        self.ncells = int(np.prod(p))

    def mesh_upload(self, xyz, cell_nodes_data, cell_nodes_ptrs):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        d = np.ascontiguousarray(cell_nodes_data, dtype=np.int32)
        p = np.ascontiguousarray(cell_nodes_ptrs, dtype=np.int32)
        self.check(lib().gh_mesh_upload(self.h, xyz.shape[1], xyz.shape[0], ptr(xyz),
                                         len(p) - 1, ptr(d), ptr(p)))
        self.ncells = len(p) - 1

    def space_upload(self, reffe_kind, order, ncomp, data, ptrs, nfree, ndiri):
        d = np.ascontiguousarray(data, dtype=np.int32)
        p = np.ascontiguousarray(ptrs, dtype=np.int32)
        self.check(lib().gh_space_upload(self.h, reffe_kind, order, ncomp, ptr(d), ptr(p),
                                         len(p) - 1, int(nfree), int(ndiri)))
        self.nfree = int(nfree)
        self.ncols = int(nfree)
        self.nnz = 0

    def trial_space_upload(self, reffe_kind, order, ncomp, data, ptrs, nfree, ndiri):
        """gh_trial_space_upload: the columns of a rectangular field pair (the space of
        space_upload gives the rows)."""
        d = np.ascontiguousarray(data, dtype=np.int32)
        p = np.ascontiguousarray(ptrs, dtype=np.int32)
        self.check(lib().gh_trial_space_upload(self.h, reffe_kind, order, ncomp, ptr(d), ptr(p),
                                               len(p) - 1, int(nfree), int(ndiri)))
        self.ncols = int(nfree)
        self.nnz = 0

    def trial_tabulation(self, shp, grad):
        """The trial basis at the points of reffe_tabulation: shp [nq, nloc], grad
        [nq, nloc, dim] (or [nq, nloc*dim])."""
        shp = np.ascontiguousarray(shp, dtype=np.float64)
        grad = np.ascontiguousarray(grad, dtype=np.float64)
        self.check(lib().gh_trial_tabulation(self.h, shp.shape[0], shp.shape[1], ptr(shp),
                                             ptr(grad)))

    def reffe_tabulation(self, qw, shp, grad):
        qw = np.ascontiguousarray(qw, dtype=np.float64)
        shp = np.ascontiguousarray(shp, dtype=np.float64)
        grad = np.ascontiguousarray(grad, dtype=np.float64)
        self.check(lib().gh_reffe_tabulation(self.h, len(qw), shp.shape[1], ptr(qw),
                                              ptr(shp), ptr(grad)))

    def geometry_tabulation(self, grad):
        """Gradients of the geometry map's shape functions at the points of the
        reference tabulation, [nq, nv, dim]."""
        grad = np.ascontiguousarray(grad, dtype=np.float64)
        self.check(lib().gh_geometry_tabulation(self.h, grad.shape[0], grad.shape[1],
                                                 ptr(grad)))

    def face_tabulation(self, qw, shp, grad):
        """Per n-cube facet: the cell's reference tabulation at the facet points;
        shp [nsides, nqf, nloc], grad [nsides, nqf, nloc, dim]."""
        qw = np.ascontiguousarray(qw, dtype=np.float64)
        shp = np.ascontiguousarray(shp, dtype=np.float64)
        grad = np.ascontiguousarray(grad, dtype=np.float64)
        self.check(lib().gh_face_tabulation(self.h, shp.shape[0], len(qw), ptr(qw), ptr(shp),
                                             ptr(grad)))

    def boundary_upload(self, face_cells, face_sides, alpha=None, beta=None, gq=None, hq=None):
        fc = np.ascontiguousarray(face_cells, dtype=np.int32)
        fs = np.ascontiguousarray(face_sides, dtype=np.int32)
        if fc.shape != fs.shape:
            raise ValueError("face_cells and face_sides differ in length")
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (alpha, beta, gq, hq)]
        self.check(lib().gh_boundary_upload(self.h, len(fc), ptr(fc), ptr(fs),
                                             *[ptr(a) for a in arrs]))

    def symbolic(self):
        nnz = C.c_int64()
        self.check(lib().gh_symbolic(self.h, C.byref(nnz)))
        self.nnz = nnz.value
        return nnz.value

    def pattern_get(self):
        colptr = np.zeros(self.ncols + 1, dtype=np.int64)
        rowval = np.zeros(max(self.nnz, 1), dtype=np.int64)
        self.check(lib().gh_pattern_get(self.h, ptr(colptr), ptr(rowval)))
        return colptr, rowval[:self.nnz]

    def coef_upload(self, kappa=None, fsrc=None, dirichlet_vals=None):
        k = None if kappa is None else np.ascontiguousarray(kappa, dtype=np.float64)
        f = None if fsrc is None else np.ascontiguousarray(fsrc, dtype=np.float64)
        g = None if dirichlet_vals is None else np.ascontiguousarray(dirichlet_vals,
                                                                     dtype=np.float64)
        self.check(lib().gh_coef_upload(self.h, ptr(k), ptr(f), ptr(g)))

    def mass_upload(self, sigma=None):
        """Per-cell mass coefficient σ of ∫ σ u v (None: no mass term)."""
        a = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64)
        self.check(lib().gh_mass_upload(self.h, ptr(a)))

    def source_qp_upload(self, fq=None):
        """f at the quadrature points, [ncells, nq] (None: the per-cell f again)."""
        a = None if fq is None else np.ascontiguousarray(fq, dtype=np.float64)
        self.check(lib().gh_source_qp_upload(self.h, ptr(a)))

    def elasticity_params(self, nu, body_force=None):
        f = None if body_force is None else np.ascontiguousarray(body_force, dtype=np.float64)
        self.check(lib().gh_elasticity_params(self.h, float(nu), ptr(f)))

    def assemble(self, form=GH_FORM_DIFFUSION, precision=GH_PREC_FP64, path=GH_PATH_AUTO):
        self.check(lib().gh_assemble(self.h, form, precision, path))

    def values_get(self, nz=None, b=None, want_nz=True):
        """nzval and b of the last assembly; into nz / b (contiguous float64 arrays of
        nnz / nfree entries, written in place) when given, else into new arrays.
        want_nz=False copies b only and returns (None, b)."""
        if not want_nz:
            if b is not None and (not isinstance(b, np.ndarray) or b.dtype != np.float64 or
                                  not b.flags.c_contiguous or b.shape != (self.nfree,)):
                raise ValueError("b must be a contiguous float64 array of %d entries"
                                 % self.nfree)
            bo = b if b is not None and self.nfree > 0 else np.zeros(max(self.nfree, 1))
            self.check(lib().gh_values_get(self.h, None, ptr(bo)))
            return None, (b if b is not None else bo[:self.nfree])
        for name, a, n in (("nz", nz, self.nnz), ("b", b, self.nfree)):
            if a is not None and (not isinstance(a, np.ndarray) or a.dtype != np.float64 or
                                  not a.flags.c_contiguous or a.shape != (n,)):
                raise ValueError("%s must be a contiguous float64 array of %d entries" % (name, n))
        nzo = nz if nz is not None and self.nnz > 0 else np.zeros(max(self.nnz, 1))
        bo = b if b is not None and self.nfree > 0 else np.zeros(max(self.nfree, 1))
        self.check(lib().gh_values_get(self.h, ptr(nzo), ptr(bo)))
        return (nz if nz is not None else nzo[:self.nnz]), (b if b is not None else bo[:self.nfree])

    def compose_block(self, src, dst, nnz):
        """Register src's matrix (a Context with its pattern built) as a block of this
        context's composite of nnz stored values: its k-th stored value goes to
        composite position dst[k] (0-based)."""
        d = np.ascontiguousarray(dst, dtype=np.int64)
        if d.shape != (src.nnz,):
            raise ValueError("dst must have one entry per stored value of the block")
        self.check(lib().gh_compose_block(self.h, src.h, ptr(d) if d.size else None,
                                          d.size, int(nnz)))
        self._comp_nnz = int(nnz)

    def compose_move(self, block):
        """Move block's current assembled values to their composite positions (on the
        device, ordered with the block context's own work)."""
        self.check(lib().gh_compose_move(self.h, int(block)))

    def compose_values_get(self, nz=None):
        """The composite's values, once every move enqueued so far has finished."""
        n = getattr(self, "_comp_nnz", None)
        if n is None:
            raise RuntimeError("no composite blocks registered")
        if nz is not None and (not isinstance(nz, np.ndarray) or nz.dtype != np.float64 or
                               not nz.flags.c_contiguous or nz.shape != (n,)):
            raise ValueError("nz must be a contiguous float64 array of %d entries" % n)
        out = nz if nz is not None and n > 0 else np.empty(max(n, 1))
        self.check(lib().gh_compose_values_get(self.h, ptr(out)))
        return nz if nz is not None else out[:n]

    def compose_clear(self):
        self.check(lib().gh_compose_clear(self.h))
        self._comp_nnz = None

    def fefunction_upload(self, free_vals, dirichlet_vals=None):
        """The FE function uh of a nonlinear form (free and Dirichlet values)."""
        u = np.ascontiguousarray(free_vals, dtype=np.float64)
        if u.shape != (self.nfree,):
            raise ValueError("free values must have num_free_dofs entries")
        g = None if dirichlet_vals is None else np.ascontiguousarray(dirichlet_vals,
                                                                     dtype=np.float64)
        self.check(lib().gh_fefunction_upload(self.h, ptr(u), ptr(g)))

    def plaplacian_params(self, p):
        self.check(lib().gh_plaplacian_params(self.h, float(p)))

    def values_set(self, nz=None, b=None):
        """Load nzval (on this context's pattern) and/or b into the device."""
        for name, a, n in (("nz", nz, self.nnz), ("b", b, self.nfree)):
            if a is not None and (not isinstance(a, np.ndarray) or a.dtype != np.float64 or
                                  not a.flags.c_contiguous or a.shape != (n,)):
                raise ValueError("%s must be a contiguous float64 array of %d entries" % (name, n))
        self.check(lib().gh_values_set(self.h, ptr(nz), ptr(b)))

    def pattern_info(self):
        """The numeric form chosen for the uploaded space (dict of gh_pattern_info)."""
        info = PatternInfo()
        self.check(lib().gh_pattern_info_get(self.h, C.byref(info)))
        return {k: int(getattr(info, k)) for k, _ in PatternInfo._fields_}

    def numeric(self, form, cell_coef=None, cell_source=None, precision=GH_PREC_FP64,
                dirichlet_vals=None):
        """gh_numeric: the one-call numeric phase from host arrays -> (nzval, b)."""
        k = None if cell_coef is None else np.ascontiguousarray(cell_coef, dtype=np.float64)
        f = None if cell_source is None else np.ascontiguousarray(cell_source, dtype=np.float64)
        g = None if dirichlet_vals is None else np.ascontiguousarray(dirichlet_vals,
                                                                     dtype=np.float64)
        for name, a, n in (("cell_coef", k, self.ncells), ("cell_source", f, self.ncells)):
            if a is not None and a.shape != (n,):
                raise ValueError("%s must have one value per cell" % name)
        nz = np.zeros(max(self.nnz, 1))
        b = np.zeros(max(self.nfree, 1))
        self.check(lib().gh_numeric(self.h, int(form), ptr(k), ptr(f), int(precision), ptr(nz),
                                    ptr(b), ptr(g)))
        return nz[:self.nnz], b[:self.nfree]

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.nfree,):
            raise ValueError("x must have num_free_dofs entries")
        y = np.zeros_like(x)
        self.check(lib().gh_spmv(self.h, ptr(x), ptr(y)))
        return y

    def pcg(self, b=None, x0=None, precond=GH_PRECOND_JACOBI, rtol=1e-8, maxit=10000,
            raise_on_noconv=True):
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        n = self.nfree
        if bb is not None and bb.shape != (n,):
            raise ValueError("b must have num_free_dofs entries")
        x = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
        if x.shape != (n,):
            raise ValueError("x0 must have num_free_dofs entries")
        st = SolveStats()
        code = lib().gh_pcg(self.h, precond, ptr(bb), ptr(x), float(rtol), int(maxit),
                            C.byref(st))
        if code != GH_OK and (code != GH_ERR_NOCONV or raise_on_noconv):
            self.check(code)
        return x, st

    def pcg_noalloc(self, b=None, x0=None, precond=GH_PRECOND_JACOBI, rtol=1e-8, maxit=10000,
                    raise_on_noconv=True):
        """gh_pcg_noalloc: the same solve with no device allocation (vectors in the
        library's fixed device array; nfree <= GH_NOALLOC_MAX_N)."""
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        n = self.nfree
        if bb is not None and bb.shape != (n,):
            raise ValueError("b must have num_free_dofs entries")
        x = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
        if x.shape != (n,):
            raise ValueError("x0 must have num_free_dofs entries")
        st = SolveStats()
        code = lib().gh_pcg_noalloc(self.h, precond, ptr(bb), ptr(x), float(rtol), int(maxit),
                                    C.byref(st))
        if code != GH_OK and (code != GH_ERR_NOCONV or raise_on_noconv):
            self.check(code)
        return x, st

    # learned FSAI ------------------------------------------------------------
    @staticmethod
    def mlp_nparams():
        n = C.c_int64()
        st = lib().gh_mlp_nparams(C.byref(n))
        if st != GH_OK:
            raise GridapHipError(st, "gh_mlp_nparams")
        return n.value

    def mlp_init(self, seed):
        self.check(lib().gh_mlp_init(self.h, C.c_uint64(seed)))

    def mlp_set(self, params):
        w = np.ascontiguousarray(params, dtype=np.float32)
        self.check(lib().gh_mlp_set(self.h, ptr(w), w.size))

    def mlp_get(self):
        w = np.zeros(self.mlp_nparams(), dtype=np.float32)
        self.check(lib().gh_mlp_get(self.h, ptr(w), w.size))
        return w

    def grf_kappa(self, omega, phase, amp):
        om = np.ascontiguousarray(omega, dtype=np.float64).reshape(-1)
        ph = np.ascontiguousarray(phase, dtype=np.float64)
        self.check(lib().gh_grf_kappa(self.h, len(ph), ptr(om), ptr(ph), float(amp)))

    def fsai_build(self):
        self.check(lib().gh_fsai_build(self.h))

    def fsai_train_step(self, nprobes, seed, lr):
        loss = C.c_double()
        self.check(lib().gh_fsai_train_step(self.h, int(nprobes), C.c_uint64(seed),
                                            C.c_float(lr), C.byref(loss)))
        return loss.value

    def fsai_get(self):
        feat = np.zeros((self.ncells, 16), dtype=np.float32)
        yo = np.zeros((self.ncells, 40), dtype=np.float32)
        gl = np.zeros(max(self.nnz, 1), dtype=np.float32)
        gu = np.zeros(max(self.nnz, 1), dtype=np.float32)
        gr = np.zeros(self.mlp_nparams(), dtype=np.float32)
        self.check(lib().gh_fsai_get(self.h, ptr(feat), ptr(yo), ptr(gl), ptr(gu), ptr(gr)))
        return dict(feat=feat, yout=yo, gl=gl[:self.nnz], gu=gu[:self.nnz], grad=gr)

    def fsai_factor_get(self):
        """The solver's factor: rows of G and of Gᵀ, 16 slots each (weights, 0-based
        column indices)."""
        n = max(self.nfree, 1) * 16
        gle, gue = np.zeros(n, dtype=np.float32), np.zeros(n, dtype=np.float32)
        gli, gui = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
        self.check(lib().gh_fsai_factor_get(self.h, ptr(gle), ptr(gli), ptr(gue), ptr(gui)))
        k = self.nfree * 16
        return (gle[:k].reshape(-1, 16), gli[:k].reshape(-1, 16), gue[:k].reshape(-1, 16),
                gui[:k].reshape(-1, 16))
