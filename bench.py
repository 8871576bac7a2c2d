"""Benchmark: assembled NNZ/s of the 3-D Q1 Poisson numeric assembly on one MI355X.

Workload (the single-GPU configuration of the metric): (0,1)^3, 128^3 HEX cells,
scalar Q1, quadrature degree 2, κ_c = exp(z_c), z ~ N(0,1) i.i.d. (seeded),
f = 1, homogeneous Dirichlet on "boundary"; element math FP32, nzval/b
accumulated FP64.  One step = one numeric assembly of A (CSC nzval, 54,439,939
entries) and b, inputs resident in HBM (assemble_matrix_and_vector! on the fixed
pattern).  The pattern is built once beforehand (reported separately as the
full-assembly time).

Prints one JSON line.  `roofline` uses the algorithmic bytes per launch
8*NNZ + 32*C + 8*C + 8*N and the kernel's average duration from HIP events on the
library's stream.  `cpu_baseline` times the CPU oracle's numeric re-assembly of
the same system (1 thread); `cpu_baseline_omp` the same loop on every host core
the process may use (OpenMP, parity-coloured cells).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gridap.jl_amd"))

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s
METRIC = "assembled NNZ/s & %HBM roofline; PCG s-to-1e-8 learned-SPAI vs Jacobi, 1 GPU"


class HipEvents:
    """hipEvent timing on an existing stream (the library's)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.L.hipEventSynchronize.argtypes = [C.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.L.hipEventDestroy.argtypes = [C.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [C.c_void_p]

    def sync(self, stream):
        assert self.L.hipStreamSynchronize(C.c_void_p(stream)) == 0

    def create(self, n):
        evs = []
        for _ in range(n):
            e = C.c_void_p()
            assert self.L.hipEventCreate(C.byref(e)) == 0
            evs.append(e)
        return evs

    def record(self, e, stream):
        assert self.L.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed_ms(self, e0, e1):
        assert self.L.hipEventSynchronize(e1) == 0
        ms = C.c_float()
        assert self.L.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
        return ms.value

    def destroy(self, evs):
        for e in evs:
            self.L.hipEventDestroy(e)


def fill_ceiling(ev, nbytes, reps=20):
    """Measured store bandwidth of the device: hipMemsetAsync of nbytes (the runtime's
    fill kernel, stores only), the mean over reps launches timed with HIP events on the
    null stream; B/s."""
    L = ev.L
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipMemsetAsync.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
    L.hipFree.argtypes = [C.c_void_p]
    p = C.c_void_p()
    assert L.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
    try:
        for _ in range(3):
            assert L.hipMemsetAsync(p, 1, C.c_size_t(nbytes), None) == 0
        ev.sync(0)
        evs = ev.create(2)
        ev.record(evs[0], 0)
        for _ in range(reps):
            assert L.hipMemsetAsync(p, 1, C.c_size_t(nbytes), None) == 0
        ev.record(evs[1], 0)
        ms = ev.elapsed_ms(evs[0], evs[1]) / reps
        ev.destroy(evs)
    finally:
        L.hipFree(p)
    return nbytes / (ms * 1e-3)


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f)
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--path", default="auto", choices=["auto", "general", "structured"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-pcg", action="store_true", help="skip the Jacobi-PCG solve")
    ap.add_argument("--no-fsai", action="store_true",
                    help="skip the learned-FSAI vs Jacobi PCG comparison (64^3 GRF instances)")
    ap.add_argument("--fsai-instances", type=int, default=16,
                    help="held-out instances in the learned-FSAI comparison")
    ap.add_argument("--no-p2", action="store_true",
                    help="skip the P2 tetrahedral assembly leg (96^3 jittered, 6 tets/hex)")
    ap.add_argument("--p2-n", type=int, default=96)
    ap.add_argument("--no-ncube", action="store_true",
                    help="skip the general n-cube element-list leg (mapped vector Q2, scalar Q2)")
    ap.add_argument("--no-mapped", action="store_true",
                    help="skip the mapped-hexahedra Q1 leg (general per-cell path)")
    ap.add_argument("--no-elasticity", action="store_true",
                    help="skip the vector-Q2 elasticity leg (48^3 hexes)")
    ap.add_argument("--elast-n", type=int, default=48)
    ap.add_argument("--no-plaplacian", action="store_true",
                    help="skip the p-Laplacian residual + Jacobian leg")
    ap.add_argument("--no-boundary", action="store_true", help="skip the boundary-terms leg")
    ap.add_argument("--no-multifield", action="store_true", help="skip the two-field leg")
    ap.add_argument("--mf-n", type=int, default=48)
    ap.add_argument("--no-stokes", action="store_true", help="skip the Taylor-Hood Stokes leg")
    ap.add_argument("--seed", type=int, default=2024)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine runs on a single MI355X and no "
              "multi-GPU path is built (asked for --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import gridap_hip as G
    from gridap_hip import _lib

    n = args.n
    D = 3
    t_host = time.time()
    model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), (n, n, n))
    reffe = G.ReferenceFE(G.lagrangian, G.Float64, 1)
    V = G.TestFESpace(model, reffe, dirichlet_tags="boundary")
    ncells = model.num_cells()
    N = V.num_free_dofs()
    rng = np.random.default_rng(args.seed)
    kappa = np.exp(rng.standard_normal(ncells))
    q = G.Quadrature(D, 2)
    shp, grad = reffe.tabulate(D, q.coords)
    cd = V.get_cell_dof_ids()
    ptrs = (np.arange(ncells + 1, dtype=np.int64) * cd.shape[1] + 1).astype(np.int32)
    host_setup_s = time.time() - t_host

    ctx = _lib.Context(0)
    ctx.mesh_cartesian(model.origin, model.sizes, model.partition)
    ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, cd.ravel(), ptrs, N, V.num_dirichlet_dofs())
    ctx.reffe_tabulation(q.weights, shp, grad)
    t0 = time.time()
    nnz = ctx.symbolic()
    symbolic_s = time.time() - t0
    ctx.coef_upload(kappa, None, None)
    prec = {"fp32": _lib.GH_PREC_FP32_ELEM, "fp64": _lib.GH_PREC_FP64}[args.precision]
    path = {"auto": _lib.GH_PATH_AUTO, "general": _lib.GH_PATH_GENERAL,
            "structured": _lib.GH_PATH_STRUCTURED}[args.path]
    stream = ctx.stream()
    ev = HipEvents()

    for _ in range(args.warmup):
        ctx.assemble(_lib.GH_FORM_DIFFUSION, prec, path)
    ev.sync(stream)

    # Timed region: K back-to-back numeric assemblies (one kernel launch each on the
    # lattice path) on the library stream, HIP events on that stream around them.
    # ms_per_step includes the dispatch gap between launches, so the roofline
    # fraction below is a lower bound for the kernel itself.
    evs = ev.create(2)
    t_wall = time.time()
    ev.record(evs[0], stream)
    for k in range(args.steps):
        ctx.assemble(_lib.GH_FORM_DIFFUSION, prec, path)
    ev.record(evs[1], stream)
    total_ms = ev.elapsed_ms(evs[0], evs[1])
    wall_s = time.time() - t_wall
    ev.destroy(evs)
    ms_per_step = total_ms / args.steps
    kernel_ms = ms_per_step

    nzval, b = ctx.values_get()
    assert np.all(np.isfinite(nzval)) and np.all(np.isfinite(b))
    colptr, rowval = ctx.pattern_get()

    algo_bytes = 8 * nnz + 4 * 8 * ncells + 8 * ncells + 8 * N
    # what the lattice kernel must move: nzval and b written, the free-neighbour
    # mask (4 B/node), κ (8 B/cell) and the per-tile starts (8 B/64 nodes) read
    ntiles = (model.num_nodes() + 63) // 64
    compulsory_bytes = 8 * nnz + 8 * N + 4 * model.num_nodes() + 8 * ncells + 8 * ntiles
    achieved = algo_bytes / (kernel_ms * 1e-3)
    fill_bw = fill_ceiling(ev, 8 * nnz)
    traffic = load_traffic("q1_lattice_128cube_traffic.json") if n == 128 else None
    out = {
        "metric": METRIC,
        "value": nnz / (ms_per_step * 1e-3),
        "unit": "NNZ/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # the lattice kernel computes in FP64 in both precision modes
        "dtype": "f32 element math, f64 nzval/b"
                 if args.precision == "fp32" and path == _lib.GH_PATH_GENERAL else "f64",
        "data": "synthetic (κ=exp(N(0,1)) per cell, seed %d)" % args.seed,
        "config": {
            "workload": "3D Q1 Poisson numeric assembly (A and b), %d^3 HEX, degree 2, "
                        "Dirichlet boundary" % n,
            "cells": ncells, "free_dofs": N, "nnz": nnz,
            "path": "lattice owner-computes" if path != _lib.GH_PATH_GENERAL else "general",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": None if traffic is None else traffic.get("bytes_per_launch"),
            "algorithmic_bytes_per_launch": algo_bytes,
            "kernel_ms": kernel_ms,
            "compulsory_bytes_per_launch": compulsory_bytes,
            "frac_on_compulsory_bytes": compulsory_bytes / (kernel_ms * 1e-3) / HBM_PEAK,
            # context: what a plain store stream of the nzval bytes reaches on this
            # device (the kernel's traffic is 90 % nzval stores); peak stays the
            # datasheet 8 TB/s
            "measured_fill_GBps": fill_bw / 1e9,
            "achieved_over_measured_fill": achieved / fill_bw,
        },
        "full_assembly_s": symbolic_s + ms_per_step * 1e-3,
        "full_assembly": full_assembly(ctx, ev, stream, prec, path, nnz, N, algo_bytes,
                                       colptr, rowval, symbolic_s, kernel_ms),
        "symbolic_s": symbolic_s,
        "host_setup_s": host_setup_s,
        "timed_wall_s": wall_s,
    }

    if not args.no_pcg:
        x, st = ctx.pcg(b=None, precond=_lib.GH_PRECOND_JACOBI, rtol=1e-8, maxit=20000,
                        raise_on_noconv=False)
        out["pcg_jacobi"] = {"iterations": int(st.iterations), "converged": int(st.converged),
                             "rel_residual": float(st.rel_residual),
                             "seconds": float(st.device_ms) * 1e-3}

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        t = time.time()
        nz_cpu, b_cpu = O.numeric_refill_cartesian(D, model.sizes, cd, q.weights, shp, grad,
                                                   kappa, None, colptr, rowval)
        cpu_s = time.time() - t
        out["cpu_baseline"] = {
            "value": nnz / cpu_s, "unit": "NNZ/s", "cores": 1, "kind": "port",
            "sample": "the full %d^3 workload once: CPU restatement of Gridap's "
                      "assemble_matrix_and_vector! numeric loop (FP64 element integration + "
                      "nz_index insertion), 1 thread, %.1f s" % (n, cpu_s),
        }
        out["gpu_vs_cpu_rel_frobenius"] = float(np.linalg.norm(nzval - nz_cpu) /
                                                np.linalg.norm(nz_cpu))
        # the same loop on all host cores this process may use (OpenMP, cells split into
        # 2^D parity colours so that threads never add into one entry)
        nthreads = len(os.sched_getaffinity(0))
        if os.environ.get("OMP_NUM_THREADS", "").isdigit():
            nthreads = min(nthreads, int(os.environ["OMP_NUM_THREADS"]))
        t = time.time()
        nz_omp, _ = O.numeric_refill_cartesian_par(D, model.sizes, model.partition, cd,
                                                   q.weights, shp, grad, kappa, None, colptr,
                                                   rowval, nthreads)
        omp_s = time.time() - t
        out["cpu_baseline_omp"] = {
            "value": nnz / omp_s, "unit": "NNZ/s", "cores": nthreads, "kind": "port",
            "sample": "the same restatement on %d OpenMP threads (cells in 2^3 parity "
                      "colours), full %d^3 workload once, %.2f s" % (nthreads, n, omp_s),
            "rel_diff_vs_1_thread": float(np.abs(nz_omp - nz_cpu).max() /
                                          np.abs(nz_cpu).max()),
        }
    ctx.close()

    if not args.no_fsai:
        out["pcg_learned_fsai"] = fsai_leg(args.fsai_instances)
    if not args.no_mapped:
        out["q1_mapped_assembly"] = q1_mapped_leg(args.n, ev)
    if not args.no_p2:
        out["p2_tet_assembly"] = p2_tet_leg(args.p2_n, ev)
    if not args.no_ncube:
        out["ncube_general"] = ncube_general_leg(ev)
    if not args.no_elasticity:
        out["q2_elasticity_assembly"] = elasticity_leg(args.elast_n, ev)
    if not args.no_plaplacian:
        out["plaplacian_assembly"] = plaplacian_leg(args.n, ev)
    if not args.no_boundary:
        out["boundary_terms_assembly"] = boundary_leg(args.n, ev)
    if not args.no_multifield:
        out["multifield_assembly"] = multifield_leg(args.mf_n)
    if not args.no_stokes:
        out["stokes_taylor_hood_assembly"] = stokes_leg(ev)
    print(json.dumps(out), flush=True)


def full_assembly(ctx, ev, stream, prec, path, nnz, N, algo_bytes, colptr, rowval, symbolic_s,
                  kernel_ms, reps=7):
    """assemble_matrix without a cached pattern: the symbolic phase (colptr/rowval from
    the cell dofs, with its checks) plus the numeric phase, repeated on one context
    (its device buffers are reused, as a caching allocator would).  Median wall time
    of `reps` calls, each ending with a stream synchronisation.  Bytes: the numeric
    assembly's plus the Int64 rowval and colptr of Gridap's SparseMatrixCSC."""
    from gridap_hip import _lib
    walls = []
    for _ in range(reps):
        ev.sync(stream)
        t = time.perf_counter()
        ctx.symbolic()
        ctx.assemble(_lib.GH_FORM_DIFFUSION, prec, path)
        ev.sync(stream)
        walls.append(time.perf_counter() - t)
    cp, rv = ctx.pattern_get()
    assert np.array_equal(cp, colptr) and np.array_equal(rv, rowval)
    warm_ms = float(np.median(walls)) * 1e3
    full_bytes = algo_bytes + 8 * nnz + 8 * (N + 1)
    return {
        "cold_s": symbolic_s + kernel_ms * 1e-3,
        "warm_ms": warm_ms,
        "warm_ms_all": [w * 1e3 for w in walls],
        "roofline_warm": {"bound": "hbm", "achieved": full_bytes / (warm_ms * 1e-3) / 1e9,
                          "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                          "frac": full_bytes / (warm_ms * 1e-3) / HBM_PEAK,
                          "algorithmic_bytes": full_bytes},
    }


MFMA_FP32_PEAK = 157.3e12  # MI355X dense FP32 matrix, FLOP/s
MFMA_FP64_PEAK = 78.6e12   # MI355X dense FP64 matrix, FLOP/s


def elasticity_leg(n, ev, steps=5, warmup=1, seed=11):
    """Numeric assembly of 3-D linear elasticity with vector Q2 on n^3 hexes: E_c =
    exp(N(0, 0.5^2)) per cell, ν = 0.3, f = (0, 0, -1), x = 0 face clamped, degree-4
    rule (27 points).  Two numeric paths on one pattern, each in FP32 and FP64 mode:
    the owner-computes lattice kernel (the default on a Cartesian mesh: E_c K̂ gathered
    per column, no element matrices) and the element path (element gradient products
    on the matrix cores, element matrices summed by the column-range kernel).  The
    FP64 lattice values are the reference of every relative Frobenius error."""
    import gridap_hip as G
    from gridap_hip import _lib
    t = time.time()
    model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), (n, n, n))
    reffe = G.ReferenceFE(G.lagrangian, G.VectorValue[3, G.Float64], 2)
    clamp = ["tag_01", "tag_03", "tag_05", "tag_07", "tag_13", "tag_15", "tag_17", "tag_19",
             "tag_25"]
    V = G.TestFESpace(model, reffe, dirichlet_tags=clamp)
    C = model.num_cells()
    E = np.exp(np.random.default_rng(seed).normal(0.0, 0.5, C))
    q = G.Quadrature(3, 4)
    shp, grad = reffe.tabulate(3, q.coords)
    host_s = time.time() - t
    ctx = _lib.Context(0)
    ctx.mesh_cartesian(model.origin, model.sizes, model.partition)
    cd = V.cell_dof_ids
    ptr = (np.arange(cd.shape[0] + 1, dtype=np.int64) * cd.shape[1] + 1).astype(np.int32)
    ctx.space_upload(_lib.GH_REFFE_QK, 2, 3, cd.ravel(), ptr, V.nfree, V.ndirichlet)
    ctx.reffe_tabulation(q.weights, shp, grad)
    t = time.time()
    nnz = ctx.symbolic()
    symbolic_s = time.time() - t
    lattice = bool(ctx.pattern_info()["lattice"])
    ctx.coef_upload(E, None, None)
    ctx.elasticity_params(0.3, [0.0, 0.0, -1.0])
    stream = ctx.stream()

    def timed(prec, path):
        for _ in range(warmup):
            ctx.assemble(_lib.GH_FORM_ELASTICITY, prec, path)
        ev.sync(stream)
        evs = ev.create(2)
        ev.record(evs[0], stream)
        for _ in range(steps):
            ctx.assemble(_lib.GH_FORM_ELASTICITY, prec, path)
        ev.record(evs[1], stream)
        ms = ev.elapsed_ms(evs[0], evs[1]) / steps
        ev.destroy(evs)
        nz, b = ctx.values_get()
        return ms, nz, b

    F32, F64 = _lib.GH_PREC_FP32_ELEM, _lib.GH_PREC_FP64
    AUTO, GEN = _lib.GH_PATH_AUTO, _lib.GH_PATH_GENERAL
    ms_l64, nz_ref, b_ref = timed(F64, AUTO)
    ms_l32, nz, b = timed(F32, AUTO)
    err_l32 = float(np.linalg.norm(nz - nz_ref) / np.linalg.norm(nz_ref))
    del nz
    ms_e32, nz, b_e32 = timed(F32, GEN)
    err_e32 = float(np.linalg.norm(nz - nz_ref) / np.linalg.norm(nz_ref))
    del nz
    ms_e64, nz, _ = timed(F64, GEN)
    err_e64 = float(np.abs(nz - nz_ref).max() / np.abs(nz_ref).max())
    del nz
    N = V.nfree
    ctx.close()
    flops = C * 27 * 2 * (6 * 6 * 81 + 81 * 6 * 81)      # dense Voigt BᵀDB count
    algo = 8 * nnz + 4 * 81 * C + 8 * C + 8 * N

    def hbm(ms):
        return {"bound": "hbm", "achieved": algo / (ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": algo / (ms * 1e-3) / HBM_PEAK,
                "algorithmic_bytes_per_launch": algo, "kernel_ms": ms}

    def mfma(ms, peak):
        return {"bound": "mfma", "achieved": flops / (ms * 1e-3) / 1e12, "peak": peak / 1e12,
                "unit": "TFLOP/s", "frac": flops / (ms * 1e-3) / peak, "flops_per_launch": flops}

    return {
        "workload": "3-D linear elasticity, vector Q2 on %d^3 HEX, degree 4, E=exp(N(0,0.5^2)) "
                    "per cell, nu=0.3, f=(0,0,-1), x=0 face clamped; FP32 element math, FP64 "
                    "nzval/b" % n,
        "cells": C, "free_dofs": N, "nnz": int(nnz),
        "path": "lattice owner-computes (E_c K̂ per column, no element matrices)" if lattice
                else "element matrices on MFMA + column-range sum",
        "ms_per_assembly": ms_l32, "value": nnz / (ms_l32 * 1e-3), "unit": "NNZ/s",
        "roofline_hbm": hbm(ms_l32),
        "fp32_vs_fp64_rel_frobenius": err_l32,
        "fp64_ms_per_assembly": ms_l64, "fp64_roofline_hbm": hbm(ms_l64),
        "element_path_mfma": {
            "ms_per_assembly": ms_e32, "value": nnz / (ms_e32 * 1e-3),
            "roofline_mfma_dense_count": mfma(ms_e32, MFMA_FP32_PEAK),
            "roofline_hbm": hbm(ms_e32),
            "fp32_vs_fp64_lattice_rel_frobenius": err_e32,
            "fp64_ms_per_assembly": ms_e64,
            "fp64_roofline_mfma_dense_count": mfma(ms_e64, MFMA_FP64_PEAK),
            "fp64_vs_fp64_lattice_rel_max": err_e64,
            "rhs_rel_diff_fp32": float(np.linalg.norm(b_e32 - b_ref) / np.linalg.norm(b_ref)),
        },
        "symbolic_s": symbolic_s, "host_numbering_s": host_s,
    }


def q1_mapped_leg(n, ev, steps=5, warmup=1, seed=2024):
    """The 3-D Q1 Poisson workload on a mapped (non-affine) n^3 hexahedral mesh:
    x -> x + 0.03 sin(2πx1) sin(2πx2) sin(2πx3) (1, 1, 1), so every cell has its own
    Jacobian at each quadrature point; κ=exp(N(0,1)) per cell, Dirichlet boundary.
    The general path: per-cell element kernel (isoparametric Jacobian per point) and
    the coloured deterministic scatter.  FP64 and FP32 element math."""
    import gridap_hip as G
    from gridap_hip import _lib
    warp = lambda x: x + 0.03 * np.prod(np.sin(2 * np.pi * x), axis=-1)[..., None]
    model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), (n, n, n), map=warp)
    reffe = G.ReferenceFE(G.lagrangian, G.Float64, 1)
    V = G.TestFESpace(model, reffe, dirichlet_tags="boundary")
    C = model.num_cells()
    kappa = np.exp(np.random.default_rng(seed).standard_normal(C))
    q = G.Quadrature(3, 2)
    shp, grad = reffe.tabulate(3, q.coords)
    ctx = _lib.Context(0)
    cn = model.cell_node_ids()
    ptr = lambda a: (np.arange(a.shape[0] + 1, dtype=np.int64) * a.shape[1] + 1).astype(np.int32)
    ctx.mesh_upload(model.node_coordinates(), cn.ravel(), ptr(cn))
    cd = V.get_cell_dof_ids()
    ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, cd.ravel(), ptr(cd), V.num_free_dofs(),
                     V.num_dirichlet_dofs())
    ctx.reffe_tabulation(q.weights, shp, grad)
    t = time.time()
    nnz = ctx.symbolic()
    symbolic_s = time.time() - t
    ctx.coef_upload(kappa, None, None)
    stream = ctx.stream()
    N = V.num_free_dofs()
    nnodes = model.num_nodes()
    # numeric bytes of the lattice workload plus the mesh the map needs: 8 node ids
    # (4 B) per cell and 3 FP64 coordinates per node
    algo = 8 * nnz + 4 * 8 * C + 8 * C + 8 * N + 4 * 8 * C + 24 * nnodes
    out = {"workload": "3D Q1 Poisson numeric assembly on a mapped %d^3 HEX mesh "
                       "(x + 0.03 sin sin sin), κ=exp(N(0,1)), degree 2, Dirichlet boundary; "
                       "per-cell kernel + coloured scatter" % n,
           "cells": C, "free_dofs": N, "nnz": int(nnz), "symbolic_s": symbolic_s,
           "algorithmic_bytes_per_launch": algo}
    ref = None
    for name, prec in (("fp64", _lib.GH_PREC_FP64), ("fp32", _lib.GH_PREC_FP32_ELEM)):
        for _ in range(warmup):
            ctx.assemble(_lib.GH_FORM_DIFFUSION, prec, _lib.GH_PATH_AUTO)
        ev.sync(stream)
        evs = ev.create(2)
        ev.record(evs[0], stream)
        for _ in range(steps):
            ctx.assemble(_lib.GH_FORM_DIFFUSION, prec, _lib.GH_PATH_AUTO)
        ev.record(evs[1], stream)
        ms = ev.elapsed_ms(evs[0], evs[1]) / steps
        ev.destroy(evs)
        nz, b = ctx.values_get()
        assert np.all(np.isfinite(nz))
        info = ctx.pattern_info()
        out[name] = {"ms_per_assembly": ms, "value": nnz / (ms * 1e-3), "unit": "NNZ/s",
                     "kernel": info["last_kernel"],
                     "roofline": {"bound": "hbm", "achieved": algo / (ms * 1e-3) / 1e9,
                                  "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                                  "frac": algo / (ms * 1e-3) / HBM_PEAK}}
        if ref is None:
            ref = nz
        else:
            out[name]["rel_frobenius_vs_fp64"] = float(np.linalg.norm(nz - ref) /
                                                       np.linalg.norm(ref))
    # the per-cell kernel with coloured scatter on the same mesh (GH_PATH_GENERAL), and
    # the same mesh with its nodes shuffled inside runs of 64 consecutive ids (no lattice
    # topology left, the locality of a mesh generator's numbering kept): the general
    # n-cube element lists
    for name, prec in (("fp64", _lib.GH_PREC_FP64), ("fp32", _lib.GH_PREC_FP32_ELEM)):
        ctx.assemble(_lib.GH_FORM_DIFFUSION, prec, _lib.GH_PATH_GENERAL)
        ev.sync(stream)
        evs = ev.create(2)
        ev.record(evs[0], stream)
        for _ in range(steps):
            ctx.assemble(_lib.GH_FORM_DIFFUSION, prec, _lib.GH_PATH_GENERAL)
        ev.record(evs[1], stream)
        out[name]["cell_kernel_ms"] = ev.elapsed_ms(evs[0], evs[1]) / steps
        ev.destroy(evs)
    ctx.close()
    rng = np.random.default_rng(seed + 1)
    perm = np.arange(nnodes)
    for k0 in range(0, nnodes, 64):
        perm[k0:k0 + 64] = k0 + rng.permutation(min(64, nnodes - k0))
    xyz2 = np.empty_like(model.node_coordinates())
    xyz2[perm] = model.node_coordinates()
    cn2 = (perm[cn - 1] + 1).astype(np.int32)
    isd = np.zeros(nnodes, dtype=bool)
    isd[perm] = V.node_and_comp_to_dof[:, 0] < 0
    ids = np.empty(nnodes, dtype=np.int64)
    ids[~isd] = np.arange(1, np.count_nonzero(~isd) + 1)
    ids[isd] = -np.arange(1, np.count_nonzero(isd) + 1)
    cd2 = ids[cn2 - 1].astype(np.int32)
    ctx = _lib.Context(0)
    ctx.mesh_upload(xyz2, cn2.ravel(), ptr(cn2))
    ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, cd2.ravel(), ptr(cd2), N, int(np.count_nonzero(isd)))
    ctx.reffe_tabulation(q.weights, shp, grad)
    t = time.time()
    nnz2 = ctx.symbolic()
    sym2 = time.time() - t
    ctx.coef_upload(kappa, None, None)
    stream = ctx.stream()
    un = {"workload": "the same mesh with its nodes shuffled inside runs of 64 ids (no "
                      "lattice topology): general n-cube element lists", "nnz": int(nnz2),
          "symbolic_s": sym2}
    for name, prec in (("fp64", _lib.GH_PREC_FP64), ("fp32", _lib.GH_PREC_FP32_ELEM)):
        ctx.assemble(_lib.GH_FORM_DIFFUSION, prec, _lib.GH_PATH_AUTO)
        ev.sync(stream)
        evs = ev.create(2)
        ev.record(evs[0], stream)
        for _ in range(steps):
            ctx.assemble(_lib.GH_FORM_DIFFUSION, prec, _lib.GH_PATH_AUTO)
        ev.record(evs[1], stream)
        ms = ev.elapsed_ms(evs[0], evs[1]) / steps
        ev.destroy(evs)
        un[name] = {"ms_per_assembly": ms, "value": nnz2 / (ms * 1e-3), "unit": "NNZ/s",
                    "kernel": ctx.pattern_info()["last_kernel"]}
    # the p-Laplacian (p = 3) residual + Jacobian at a random FE function on the same
    # mesh: element lists, and the coloured per-cell scatter (path "general")
    ctx.fefunction_upload(rng.standard_normal(N), np.zeros(int(np.count_nonzero(isd))))
    ctx.plaplacian_params(3.0)
    pl = {}
    for name, prec in (("fp64", _lib.GH_PREC_FP64), ("fp32", _lib.GH_PREC_FP32_ELEM)):
        for key, path in (("ms_per_assembly", _lib.GH_PATH_AUTO),
                          ("cell_kernel_ms", _lib.GH_PATH_GENERAL)):
            ctx.assemble(_lib.GH_FORM_PLAPLACIAN, prec, path)
            ev.sync(stream)
            evs = ev.create(2)
            ev.record(evs[0], stream)
            for _ in range(steps):
                ctx.assemble(_lib.GH_FORM_PLAPLACIAN, prec, path)
            ev.record(evs[1], stream)
            pl.setdefault(name, {})[key] = ev.elapsed_ms(evs[0], evs[1]) / steps
            ev.destroy(evs)
            if path == _lib.GH_PATH_AUTO:
                pl[name]["kernel"] = ctx.pattern_info()["last_kernel"]
    un["plaplacian_p3"] = pl
    out["unstructured"] = un
    ctx.close()
    return out


def ncube_general_leg(ev, nq2=32, ns=48, steps=3, warmup=1, seed=5):
    """The general n-cube element-list path: vector Q2 elasticity on a mapped nq2^3
    hexahedral mesh (x + 0.03 sin sin sin) and scalar Q2 Poisson on a Cartesian ns^3
    mesh, E or κ lognormal per cell, degree 4 (27 points), FP64 and FP32 element
    matrices.  Per assembly: the element kernel writes every cell's full matrix, the
    contribution-list sum reads one entry, one int32 list entry per contribution and one
    start per nonzero, and writes nzval; "model bytes" counts exactly that traffic."""
    import gridap_hip as G
    from gridap_hip import _lib
    warp = lambda x: x + 0.03 * np.prod(np.sin(2 * np.pi * x), axis=-1)[..., None]
    out = {}
    for key, n, vector in (("mapped_vector_q2_elasticity", nq2, True),
                           ("scalar_q2_poisson", ns, False)):
        model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), (n, n, n),
                                         map=warp if vector else None)
        T = G.VectorValue[3, G.Float64] if vector else G.Float64
        V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, T, 2), dirichlet_tags="boundary")
        U = G.TrialFESpace(V)
        Ω = G.Triangulation(model)
        dΩ = G.Measure(Ω, 4)
        C = model.num_cells()
        c = G.CellField(np.exp(np.random.default_rng(seed).standard_normal(C)), Ω)
        if vector:
            σ = G.IsotropicStress(c, 0.3)
            a = lambda u, v: G.integrate(G.inner(G.ε(v), σ(G.ε(u))), dΩ)
            l = lambda v: G.integrate(G.dot(v, G.VectorValue(0.0, 0.0, -1.0)), dΩ)
            form = _lib.GH_FORM_ELASTICITY
        else:
            a = lambda u, v: G.integrate(c * G.inner(G.gradient(v), G.gradient(u)), dΩ)
            l = lambda v: G.integrate(1.0 * v, dΩ)
            form = _lib.GH_FORM_DIFFUSION
        res = {}
        ref = None
        for name in ("fp64", "fp32"):
            assem = G.SparseMatrixAssembler(U, V, precision=name)
            t = time.time()
            op = G.AffineFEOperator(a, l, U, V, assem)
            first_s = time.time() - t
            ctx = assem.ctx
            prec = assem.precision
            stream = ctx.stream()
            for _ in range(warmup):
                ctx.assemble(form, prec, _lib.GH_PATH_AUTO)
            ev.sync(stream)
            evs = ev.create(2)
            ev.record(evs[0], stream)
            for _ in range(steps):
                ctx.assemble(form, prec, _lib.GH_PATH_AUTO)
            ev.record(evs[1], stream)
            ms = ev.elapsed_ms(evs[0], evs[1]) / steps
            ev.destroy(evs)
            nz, _ = ctx.values_get()
            nnz, N, nloc = int(assem.nnz), V.num_free_dofs(), V.cell_dof_ids.shape[1]
            esz = 8 if name == "fp64" else 4
            ncontrib = int(np.count_nonzero(V.cell_dof_ids > 0, axis=1).astype(np.int64) ** 2 @
                           np.ones(C, dtype=np.int64))
            model_bytes = C * nloc * nloc * esz + ncontrib * (esz + 4) + 12 * nnz + 8 * N
            res[name] = {"ms_per_assembly": ms, "value": nnz / (ms * 1e-3), "unit": "NNZ/s",
                         "kernel": ctx.pattern_info()["last_kernel"],
                         "first_assembly_s": first_s,
                         "model_bytes": model_bytes,
                         "model_GBps": model_bytes / (ms * 1e-3) / 1e9,
                         "model_frac_of_hbm": model_bytes / (ms * 1e-3) / HBM_PEAK}
            if ref is None:
                ref = nz
            else:
                res[name]["rel_frobenius_vs_fp64"] = float(np.linalg.norm(nz - ref) /
                                                           np.linalg.norm(ref))
            ctx.close()
        res.update({"cells": C, "free_dofs": int(V.num_free_dofs()), "nnz": int(nnz),
                    "workload": ("vector Q2 elasticity, mapped %d^3 HEX (x + 0.03 sin sin sin), "
                                 "E lognormal, ν=0.3" % n) if vector else
                                ("scalar Q2 Poisson, Cartesian %d^3 HEX, κ lognormal" % n)})
        out[key] = res
    return out


def p2_tet_leg(n, ev, steps=10, warmup=2, seed=7):
    """Numeric assembly of scalar P2 Poisson on the simplexified n^3 grid (6 tets per
    hex, interior vertices jittered by U(-0.2h, 0.2h), κ lognormal σ=1, Dirichlet
    boundary, degree-4 rule), FP64 geometry and FP32 contraction, K launches timed
    with HIP events on the library stream."""
    import gridap_hip as G
    from gridap_hip import _lib
    t = time.time()
    model = G.simplexify(G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), (n, n, n))).jittered(seed)
    reffe = G.ReferenceFE(G.lagrangian, G.Float64, 2)
    V = G.TestFESpace(model, reffe, dirichlet_tags="boundary")
    C = model.num_cells()
    kappa = np.exp(np.random.default_rng(seed).standard_normal(C))
    q = G.Quadrature(3, 4, simplex=True)
    shp, grad = reffe.tabulate(3, q.coords, simplex=True)
    host_s = time.time() - t
    ctx = _lib.Context(0)
    cn = model.cell_node_ids()
    ptr = lambda a: (np.arange(a.shape[0] + 1, dtype=np.int64) * a.shape[1] + 1).astype(np.int32)
    ctx.mesh_upload(model.node_coordinates(), cn.ravel(), ptr(cn))
    cd = V.cell_dof_ids
    ctx.space_upload(_lib.GH_REFFE_PK, 2, 1, cd.ravel(), ptr(cd), V.nfree, V.ndirichlet)
    ctx.reffe_tabulation(q.weights, shp, grad)
    t = time.time()
    nnz = ctx.symbolic()
    symbolic_s = time.time() - t
    ctx.coef_upload(kappa, None, None)
    stream = ctx.stream()
    for _ in range(warmup):
        ctx.assemble(_lib.GH_FORM_DIFFUSION, _lib.GH_PREC_FP32_ELEM, _lib.GH_PATH_AUTO)
    ev.sync(stream)
    evs = ev.create(2)
    ev.record(evs[0], stream)
    for _ in range(steps):
        ctx.assemble(_lib.GH_FORM_DIFFUSION, _lib.GH_PREC_FP32_ELEM, _lib.GH_PATH_AUTO)
    ev.record(evs[1], stream)
    ms = ev.elapsed_ms(evs[0], evs[1]) / steps
    ev.destroy(evs)
    nz, b = ctx.values_get()
    ok = bool(np.all(np.isfinite(nz)) and np.all(np.isfinite(b)))
    N = V.nfree
    nverts = model.num_vertices()
    algo = 8 * nnz + 4 * 10 * C + 8 * C + 4 * 4 * C + 24 * nverts + 8 * N
    flops = C * 14 * 2 * (3 * 3 * 10 + 10 * 10 * 3)
    ctx.close()
    return {
        "workload": "P2 Poisson on %d^3 x 6 jittered tets (seed %d), κ lognormal σ=1, "
                    "degree 4 (14 points), Dirichlet boundary; FP64 geometry, FP32 "
                    "contraction, FP64 nzval/b" % (n, seed),
        "cells": C, "free_dofs": N, "nnz": int(nnz),
        "ms_per_assembly": ms, "value": nnz / (ms * 1e-3), "unit": "NNZ/s",
        "roofline": {"bound": "hbm", "achieved": algo / (ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": algo / (ms * 1e-3) / HBM_PEAK,
                     "algorithmic_bytes_per_launch": algo},
        "element_tflops_dense_count": flops / (ms * 1e-3) / 1e12,
        "symbolic_s": symbolic_s, "host_numbering_s": host_s, "finite": ok,
    }


def timed_ms(ev, stream, fn, steps, warmup=1):
    """Mean duration of fn() over steps calls, HIP events on stream around them (fn's
    work is enqueued on that stream)."""
    for _ in range(warmup):
        fn()
    ev.sync(stream)
    evs = ev.create(2)
    ev.record(evs[0], stream)
    for _ in range(steps):
        fn()
    ev.record(evs[1], stream)
    ms = ev.elapsed_ms(evs[0], evs[1]) / steps
    ev.destroy(evs)
    return ms


def plaplacian_leg(n, ev, p=3.0, steps=5, seed=31):
    """Residual and Jacobian of the p-Laplacian (PLaplacianTests.jl's form, p = 3, κ and f
    per cell) at a random FE function on the n^3 Q1 hexahedral mesh, Dirichlet boundary:
    one Newton step's assembly.  Lattice route (element Jacobians and residuals per cell,
    owner-computes columns) and, for comparison, the coloured per-cell scatter; the CPU
    oracle (numpy) on a 48^3 sample."""
    import gridap_hip as G
    from gridap_hip import _lib
    model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), (n, n, n))
    reffe = G.ReferenceFE(G.lagrangian, G.Float64, 1)
    V = G.TestFESpace(model, reffe, dirichlet_tags="boundary")
    C_ = model.num_cells()
    N = V.num_free_dofs()
    rng = np.random.default_rng(seed)
    kappa = np.exp(rng.standard_normal(C_))
    fsrc = rng.uniform(-1, 1, C_)
    u = rng.standard_normal(N)
    q = G.Quadrature(3, 2)
    shp, grad = reffe.tabulate(3, q.coords)
    cd = V.get_cell_dof_ids()
    ptrs = (np.arange(C_ + 1, dtype=np.int64) * cd.shape[1] + 1).astype(np.int32)
    ctx = _lib.Context(0)
    ctx.mesh_cartesian(model.origin, model.sizes, model.partition)
    ctx.space_upload(_lib.GH_REFFE_QK, 1, 1, cd.ravel(), ptrs, N, V.num_dirichlet_dofs())
    ctx.reffe_tabulation(q.weights, shp, grad)
    nnz = ctx.symbolic()
    ctx.coef_upload(kappa, fsrc, None)
    ctx.fefunction_upload(u, np.zeros(V.num_dirichlet_dofs()))
    ctx.plaplacian_params(p)
    stream = ctx.stream()
    # nzval and the residual written; κ, f, the cell dof ids and uh read
    algo = 8 * nnz + 8 * N + 16 * C_ + 32 * C_ + 8 * N
    out = {"workload": "p-Laplacian (p=%g) residual + Jacobian at a random FE function, "
                       "%d^3 HEX Q1, degree 2, κ=exp(N(0,1)), f~U(-1,1) per cell, "
                       "Dirichlet boundary" % (p, n),
           "cells": C_, "free_dofs": N, "nnz": int(nnz), "algorithmic_bytes_per_launch": algo}
    ref = None
    for name, prec in (("fp64", _lib.GH_PREC_FP64), ("fp32", _lib.GH_PREC_FP32_ELEM)):
        run = lambda: ctx.assemble(_lib.GH_FORM_PLAPLACIAN, prec, _lib.GH_PATH_AUTO)
        ms = timed_ms(ev, stream, run, steps)
        kern = ctx.pattern_info()["last_kernel"]
        nz, b = ctx.values_get()
        assert np.all(np.isfinite(nz)) and np.all(np.isfinite(b))
        cell = timed_ms(ev, stream, lambda: ctx.assemble(_lib.GH_FORM_PLAPLACIAN, prec,
                                                         _lib.GH_PATH_GENERAL), steps)
        nzc, bc = ctx.values_get()
        r = {"ms_per_assembly": ms, "value": nnz / (ms * 1e-3), "unit": "NNZ/s",
             "kernel": kern, "cell_kernel_ms": cell,
             "rel_max_vs_cell_kernel": float(np.abs(nz - nzc).max() / np.abs(nzc).max()),
             "roofline": {"bound": "hbm", "achieved": algo / (ms * 1e-3) / 1e9,
                          "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                          "frac": algo / (ms * 1e-3) / HBM_PEAK}}
        if ref is None:
            ref = nz
        else:
            r["rel_frobenius_vs_fp64"] = float(np.linalg.norm(nz - ref) / np.linalg.norm(ref))
        out[name] = r
    ctx.close()
    # CPU oracle on a bounded sample
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ns = 48
    ms_ = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), (ns, ns, ns))
    Vs = G.TestFESpace(ms_, reffe, dirichlet_tags="boundary")
    cds = Vs.get_cell_dof_ids()
    us = rng.standard_normal(Vs.num_free_dofs())
    t = time.time()
    uc = O.cell_values(cds, us, np.zeros(Vs.num_dirichlet_dofs()))
    K, rr = O.plaplacian_cells(3, ms_.cell_node_ids(), ms_.node_coordinates(), q.weights, shp,
                               grad, uc, p, kappa[:ms_.num_cells()], fsrc[:ms_.num_cells()],
                               cartesian_sizes=ms_.sizes)
    _, _, nzs, _ = O.plaplacian_system(cds, Vs.num_free_dofs(), K, rr)
    cpu_s = time.time() - t
    out["cpu_baseline"] = {"value": len(nzs) / cpu_s, "unit": "NNZ/s", "cores": 1,
                           "kind": "port",
                           "sample": "%d^3 cells (%d nonzeros): the oracle's numpy "
                                     "restatement of residual!/jacobian! + assembly, %.2f s"
                                     % (ns, len(nzs), cpu_s)}
    return out


def _boundary_cells(n):
    """Cells of the n^3 lattice with a face on the boundary (x fastest)."""
    i = np.arange(n)
    x, y, z = np.meshgrid(i, i, i, indexing="ij")
    on = (x == 0) | (x == n - 1) | (y == 0) | (y == n - 1) | (z == 0) | (z == n - 1)
    return (x + n * (y + n * z))[on]


def boundary_leg(n, ev, steps=5):
    """Cell terms plus boundary terms in one assembly: ∫κ∇u⋅∇v + ∫_Γ α u v = ∫ f v +
    ∫_Γ g v on the n^3 Q1 hexahedral mesh with a Robin condition (α = 2) and a Neumann
    load g(x) on all six faces (no Dirichlet dofs), the same system a
    BoundaryTriangulation term list assembles (PoissonTests.jl:101-108).  The face kernel
    time is the difference with the cell terms alone on the same pattern."""
    import gridap_hip as G
    model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), (n, n, n))
    labels = model.get_face_labeling()
    Ω = G.Triangulation(model)
    Γ = G.BoundaryTriangulation(model, labels, tags="boundary")
    dΩ, dΓ = G.Measure(Ω, 2), G.Measure(Γ, 2)
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1))
    U = G.TrialFESpace(V)
    kappa = np.exp(np.random.default_rng(5).standard_normal(model.num_cells()))
    k = G.CellField(kappa, Ω)
    g = lambda x: np.sin(3 * x[..., 0]) + x[..., 1] * x[..., 2]
    a = lambda u, v: (G.integrate(k * G.inner(G.gradient(v), G.gradient(u)), dΩ) +
                      G.integrate(2.0 * v * u, dΓ))
    l = lambda v: G.integrate(1.0 * v, dΩ) + G.integrate(v * g, dΓ)
    t = time.time()
    op = G.AffineFEOperator(a, l, U, V)
    first_s = time.time() - t
    assem = op.assem
    ctx = assem.ctx
    stream = ctx.stream()
    nnz = int(assem.colptr[-1] - 1)
    nf = 6 * n * n
    from gridap_hip import _lib
    run = lambda: ctx.assemble(_lib.GH_FORM_DIFFUSION, assem.precision, assem.path)
    ms = timed_ms(ev, stream, run, steps)
    kern = ctx.pattern_info()["last_kernel"]
    ctx.boundary_upload(np.zeros(0, np.int32), np.zeros(0, np.int32))
    ms_cells = timed_ms(ev, stream, run, steps)
    ctx.close()
    # what the face terms move: per face its cell, side, α and 4 g values (48 B), its
    # 8 loads written and read back (128 B), its matrix triangle written (288 B) and its
    # 64 contributions gathered with their 4-B indices (768 B): ~1 KB with the loads;
    # each distinct nonzero the faces touch read and written once (16 B), and each
    # touched b entry (16 B)
    cd = V.get_cell_dof_ids()
    fd = cd[_boundary_cells(n)].astype(np.int64)
    pairs = np.unique((fd[:, :, None] * (1 << 32) + fd[:, None, :]).ravel())
    ndst = len(pairs)
    nown = len(np.unique(fd))
    face_bytes = nf * (48 + 128 + 288 + 768) + ndst * 16 + nown * 16
    dms = max(ms - ms_cells, 1e-6)
    return {"workload": "Q1 Poisson %d^3 HEX with Robin (α=2) + Neumann g(x) terms on all "
                        "%d boundary faces, κ=exp(N(0,1)); lattice cell kernel, then the face kernel and the per-nonzero face sum" % (n, nf),
            "nnz": nnz, "boundary_faces": nf, "face_nonzeros": ndst, "first_assembly_s": first_s,
            "ms_per_assembly": ms, "value": nnz / (ms * 1e-3), "unit": "NNZ/s",
            "kernel": kern, "cells_only_ms": ms_cells, "face_terms_ms": ms - ms_cells,
            "face_roofline": {"bound": "hbm", "achieved": face_bytes / (dms * 1e-3) / 1e9,
                              "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                              "frac": face_bytes / (dms * 1e-3) / HBM_PEAK,
                              "algorithmic_bytes_per_launch": face_bytes}}


def multifield_leg(n, reps=3):
    """Two-field Q1 system (BlockSparseMatrixAssemblersTests.jl's biform2/liform2:
    ∫∇u1⋅∇v1 + u2 v2 - u1 v2, ∫v1 - v2) on the n^3 hexahedral mesh, Dirichlet boundary,
    consecutive layout: assemble_matrix_and_vector end to end (host lowering, the three
    touched pairs on the device, each moved into the consecutive composite on the device,
    one copy of the composite to the host), wall clock; and the composition alone (its
    wait for the moves and the copy)."""
    import gridap_hip as G
    model = G.CartesianDiscreteModel((0, 1, 0, 1, 0, 1), (n, n, n))
    reffe = G.ReferenceFE(G.lagrangian, G.Float64, 1)
    V = G.TestFESpace(model, reffe, dirichlet_tags="boundary")
    U = G.TrialFESpace(V, lambda x: x[..., 0] + x[..., 1])
    dΩ = G.Measure(G.Triangulation(model), 2)
    X, Y = G.MultiFieldFESpace([U, U]), G.MultiFieldFESpace([V, V])
    a = lambda u, v: G.integrate(G.inner(G.gradient(u[0]), G.gradient(v[0])) +
                                 G.inner(u[1], v[1]) - G.inner(u[0], v[1]), dΩ)
    l = lambda v: G.integrate(v[0] * 1.0 - v[1] * 1.0, dΩ)
    assem = G.SparseMatrixAssembler(X, Y)
    t = time.time()
    A, b = G.assemble_matrix_and_vector(a, l, X, Y, assem)
    first_s = time.time() - t
    # time the composition inside the same calls
    compose, spent = assem._compose_device, []

    def timed_compose(*args, **kw):
        t0 = time.time()
        r = compose(*args, **kw)
        spent.append(time.time() - t0)
        return r

    assem._compose_device = timed_compose
    walls, comps = [], []
    try:
        for _ in range(reps):
            del spent[:]
            t = time.time()
            A, b = G.assemble_matrix_and_vector(a, l, X, Y, assem)
            walls.append(time.time() - t)
            comps.append(sum(spent))
    finally:
        del assem._compose_device
    # assemble_matrix_and_vector! into the same A and b: no new host pages to fault in
    inplace = []
    for _ in range(reps):
        t = time.time()
        G.assemble_matrix_and_vector_inplace(a, l, A, b, assem)
        inplace.append(time.time() - t)
    assert np.all(np.isfinite(A.nzval)) and np.all(np.isfinite(b))
    k = int(np.argmin(walls))
    nnz = int(A.colptr[-1] - 1)
    return {"workload": "two-field Q1 system (∇u1⋅∇v1 + u2 v2 - u1 v2) on %d^3 HEX, "
                        "Dirichlet boundary, consecutive layout; assemble_matrix_and_vector "
                        "end to end (host lowering, 3 device pair assemblies, each moved "
                        "into the composite on the device, one copy to the host), wall "
                        "clock" % n,
            "free_dofs": 2 * V.num_free_dofs(), "nnz": nnz, "first_assembly_s": first_s,
            "wall_s_per_assembly": walls[k], "compose_and_copy_s": comps[k],
            "wall_s_per_inplace_assembly": min(inplace),
            "value": nnz / walls[k], "unit": "NNZ/s"}


def stokes_leg(ev, n=256, reps=2, steps=5):
    """Taylor–Hood Stokes (StokesTaylorHoodTests.jl's form; vector Q2 velocity, Q1
    pressure) on (0,2)² with n x n cells, velocity Dirichlet data on the boundary, the
    pressure fixed at one vertex: assemble_matrix_and_vector end to end (wall clock),
    and each block's device numeric phase (HIP events on its context's stream): the
    vector Laplacian, q (∇⋅u) and (∇⋅v) p."""
    import gridap_hip as G
    from gridap_hip import _lib
    model = G.CartesianDiscreteModel((0, 2, 0, 2), (n, n))
    ue = lambda x: np.stack([x[..., 0] ** 2 + 2 * x[..., 1] ** 2, -x[..., 0] ** 2], axis=-1)
    pe = lambda x: x[..., 0] + 3 * x[..., 1]
    V = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.VectorValue[2, G.Float64], 2),
                      dirichlet_tags="boundary")
    Q = G.TestFESpace(model, G.ReferenceFE(G.lagrangian, G.Float64, 1), dirichlet_tags="tag_1")
    U, P = G.TrialFESpace(V, ue), G.TrialFESpace(Q, pe)
    Y, X = G.MultiFieldFESpace([V, Q]), G.MultiFieldFESpace([U, P])
    dΩ = G.Measure(G.Triangulation(model), 2)
    g = lambda x: 2.0 * x[..., 0]
    a = lambda u, v: G.integrate(G.inner(G.gradient(v[0]), G.gradient(u[0])) -
                                 G.divergence(v[0]) * u[1] + v[1] * G.divergence(u[0]), dΩ)
    l = lambda v: G.integrate(G.dot(v[0], G.VectorValue(-5.0, 5.0)) + v[1] * g, dΩ)
    assem = G.SparseMatrixAssembler(X, Y)
    t = time.time()
    A, b = G.assemble_matrix_and_vector(a, l, X, Y, assem)
    first_s = time.time() - t
    walls = []
    for _ in range(reps):
        t = time.time()
        A, b = G.assemble_matrix_and_vector(a, l, X, Y, assem)
        walls.append(time.time() - t)
    assert np.all(np.isfinite(A.nzval)) and np.all(np.isfinite(b))
    blocks = {}
    for name, (i, j), form in (("vector_laplacian", (0, 0), _lib.GH_FORM_DIFFUSION),
                               ("q_div_u", (1, 0), _lib.GH_FORM_DIVERGENCE),
                               ("div_v_p", (0, 1), _lib.GH_FORM_DIVERGENCE)):
        pa = assem._pair(i, j)
        ctx = pa.ctx
        path = _lib.GH_PATH_GENERAL if form == _lib.GH_FORM_DIVERGENCE else pa.path
        # the pair's matrix assembly as it last ran: its coefficients are on the device
        if form == _lib.GH_FORM_DIFFUSION:
            ctx.coef_upload(None, np.zeros(model.num_cells()), None)
            ctx.elasticity_params(0.3, np.zeros(3))
        run = lambda: ctx.assemble(form, _lib.GH_PREC_FP64, path)
        ms = timed_ms(ev, ctx.stream(), run, steps)
        blocks[name] = {"ms": ms, "nnz": int(ctx.nnz), "kernel": ctx.pattern_info()["last_kernel"],
                        "value": ctx.nnz / (ms * 1e-3), "unit": "NNZ/s"}
    return {"workload": "Taylor-Hood Stokes (vector Q2 / Q1) on (0,2)^2, %d x %d cells, "
                        "velocity Dirichlet boundary, pressure fixed at one vertex" % (n, n),
            "free_dofs": int(V.nfree + Q.nfree), "nnz": int(A.colptr[-1] - 1),
            "first_assembly_s": first_s, "wall_s_per_assembly": min(walls),
            "blocks_fp64_device": blocks}


def fsai_leg(k):
    """Learned-FSAI-PCG vs Jacobi-PCG to relative residual 1e-8 on held-out 64^3 Q1
    Poisson instances with κ = exp(GRF), using the committed trained weights."""
    from gridap_hip import fsai as FS
    if not os.path.exists(FS.DEFAULT_WEIGHTS):
        return {"skipped": "no trained weights at %s" % FS.DEFAULT_WEIGHTS}
    inst = FS.PoissonInstances(64)
    net = FS.LearnedFSAI.load(FS.DEFAULT_WEIGHTS)
    net.attach(inst.ctx)
    rows = FS.evaluate(inst, list(FS.HELDOUT_SEEDS)[:k])
    sm = FS.summarize(rows)
    inst.ctx.close()
    N, nnz = int(inst.nfree), int(inst.nnz)
    nnz_l = (nnz - N) // 2 + N  # entries of tril(A), the pattern of G
    # algorithmic bytes per iteration (CSR SpMV + vector updates; the FSAI applications
    # add G and Gᵀ as CSR rows) against the measured time per iteration, launch gaps and
    # convergence checks included; the working set fits the MALL, so this fraction
    # can exceed what HBM alone would allow
    jac_bytes = 12 * nnz + 4 * (N + 1) + 17 * 8 * N
    fsai_bytes = jac_bytes + 2 * (12 * nnz_l + 20 * N)
    jac_us = sm["jacobi_ms"] * 1e3 / sm["jacobi_iters"]
    fsai_us = sm["fsai_ms"] * 1e3 / sm["fsai_iters"]

    def roof(b, us):
        return {"bound": "hbm", "achieved": b / (us * 1e-6) / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": b / (us * 1e-6) / HBM_PEAK,
                "algorithmic_bytes_per_iteration": b, "us_per_iteration": us}
    return {
        "workload": "64^3 HEX Q1 Poisson, κ=exp(GRF: 64 Fourier modes, ℓ=0.1), held-out "
                    "seeds %d..%d, FP64 PCG from x0=0 to ‖r‖/‖b‖<=1e-8" %
                    (FS.HELDOUT_SEEDS[0], FS.HELDOUT_SEEDS[0] + k - 1),
        "free_dofs": int(inst.nfree), "nnz": int(inst.nnz), "instances": k,
        "jacobi": {"iterations": sm["jacobi_iters"], "seconds": sm["jacobi_ms"] * 1e-3},
        "learned_fsai": {"iterations": sm["fsai_iters"], "seconds": sm["fsai_ms"] * 1e-3,
                         "setup_seconds": sm["fsai_setup_ms"] * 1e-3,
                         "seconds_with_setup": (sm["fsai_ms"] + sm["fsai_setup_ms"]) * 1e-3,
                         "fallbacks_to_jacobi": sm["fallbacks"]},
        "speedup_solve": sm["time_ratio_solve"],
        "speedup_with_setup": sm["time_ratio_with_setup"],
        "iteration_ratio": sm["iteration_ratio"],
        "roofline_jacobi_iteration": roof(jac_bytes, jac_us),
        "roofline_fsai_iteration": roof(fsai_bytes, fsai_us),
    }


if __name__ == "__main__":
    main()
