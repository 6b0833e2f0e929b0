#!/usr/bin/env python3
"""Benchmark: P2 anisotropic-diffusion assembly on one MI355X.

Workload (one "step" = one full assembly pass): unit cube, 64^3 sub-cubes x 6
Kuhn tets = 1,572,864 tets, interior nodes jittered by +-0.2h (seed 1), P2
Lagrange elements (2,146,689 DOFs, 60,859,905 nnz before the Dirichlet drop),
per-tet SPD tensor D = R diag(1, 10^a, 10^b) R^T (a, b ~ U[-2, 2], R Haar),
4-point quadrature, unit load, homogeneous Dirichlet on the boundary.  A step
forms every element matrix and load vector and writes each FP64 CSR value and
RHS entry once (FP64 geometry, FP64 contraction, FP64 accumulation): a
geometry launch and a row-block launch.  Inputs are resident in HBM before
timing.

value  = element matrices assembled per second (whole job; 1 GPU).
Extra fields: nnz/s, the FP32-contraction mode, the element-matrix-only
kernels (FP64 / FP32 output), Jacobi-PCG on the 16^3 P1 Poisson system,
the roofline of the assembly pass and the CPU baseline (reference element
core + sequential scatter on the host, whole workload, one thread).

--copy-calibration MIB runs one plain device copy of MIB MiB (torch copy_,
outside the timed region) so that a counter run of the same command can
scale FETCH_SIZE / WRITE_SIZE by what they read for known bytes.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "inmost-fem_amd"))

HBM_PEAK_GBS = 8000.0


def haar_rotations(rng, n):
    G = rng.normal(size=(n, 3, 3))
    Q, R = np.linalg.qr(G)
    s = np.sign(np.diagonal(R, axis1=1, axis2=2))
    Q = Q * s[:, None, :]
    neg = np.linalg.det(Q) < 0
    Q[neg, :, 0] *= -1
    return Q


def aniso_packed(rng, n):
    """(6, n) packed upper-triangle SoA of R diag(1, 10^a, 10^b) R^T."""
    R = haar_rotations(rng, n)
    ab = rng.uniform(-2, 2, (n, 2))
    lam = np.stack([np.ones(n), 10 ** ab[:, 0], 10 ** ab[:, 1]], axis=1)
    D = np.einsum("rij,rj,rkj->rik", R, lam, R)
    idx = [(0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2)]
    return np.stack([D[:, i, j] for i, j in idx]), D


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(coords, tets_reordered, Dfull, dofs, slot, is_dir, nnz, ndof):
    """The reference path on the host cores, whole workload: the reference
    element core (oracle/_ref, compiled from the AniFem++ sources) forms every
    element matrix and load vector, element-level homogeneous Dirichlet
    (dc_on_dof.h:38-45) follows per cell, then the scatter into the same CSR.
    Multi-threaded form: the cells split into contiguous chunks over all host
    CPUs (the reference evaluates cells independently, assembler.inl:381),
    then a colour-parallel scatter (no two cells of a colour share a node).
    Single-thread form: one thread for the cells, then the oracle's
    sequential cell-order scatter (assembler.inl:397-481), whose values are
    returned for the parity check.  Falls back to the oracle's own element
    restatement where the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from concurrent.futures import ThreadPoolExecutor
    import anifem_testlib as T

    ref = T.ref_lib()
    kind = "reference" if ref is not None else "port"
    O = T.oracle()
    n = len(tets_reordered)
    xyz = np.ascontiguousarray(coords[tets_reordered].reshape(n, 12))
    Df = np.ascontiguousarray(Dfull.transpose(0, 2, 1).reshape(n, 9))
    one = np.ones(1)
    dp = ctypes.POINTER(ctypes.c_double)
    dofs = np.ascontiguousarray(dofs, dtype=np.int32)
    is_dir = np.ascontiguousarray(is_dir, dtype=np.uint8)
    if ref is not None:
        ref.ref_fem3dtet.restype = ctypes.c_int

    def cells(lo, hi, A, F):
        if ref is not None:
            # (GRAD P2, GRAD P2) with a per-tet symmetric tensor; (IDEN P0, IDEN P2) with f = 1
            assert ref.ref_fem3dtet(1, 3, 0, 2, ctypes.c_long(hi - lo), xyz[lo:hi].ctypes.data_as(dp),
                                    Df[lo:hi].ctypes.data_as(dp), ctypes.c_long(9), A[lo:hi].ctypes.data_as(dp)) == 0
            assert ref.ref_fem3dtet(6, 2, 0, 2, ctypes.c_long(hi - lo), xyz[lo:hi].ctypes.data_as(dp),
                                    one.ctypes.data_as(dp), ctypes.c_long(0), F[lo:hi].ctypes.data_as(dp)) == 0
        else:
            A[lo:hi] = T.orc_elem(2, 3, 1, 2, 3, 1, 3, 1, 2, xyz[lo:hi], Df[lo:hi])
            F[lo:hi] = T.orc_elem(1, 1, 1, 1, 3, 1, 2, 0, 2, xyz[lo:hi], np.ones((1, 1)))
        O.orc_apply_dir_cells(ctypes.c_long(hi - lo), 10, T.P(dofs[lo:hi]), T.P(is_dir), None, T.P(A[lo:hi]),
                              T.P(F[lo:hi]))

    def zeros(*shape):
        # touched before any timed region: first-touch page faults of a fresh
        # allocation are not part of the path
        a = np.empty(shape)
        a.fill(0.0)
        return a

    # one thread: cells, then the sequential scatter
    A1, F1 = zeros(n, 100), zeros(n, 10)
    t0 = time.perf_counter()
    cells(0, n, A1, F1)
    t_cells_1 = time.perf_counter() - t0
    val_seq, rhs_seq = zeros(nnz), zeros(ndof)
    t0 = time.perf_counter()
    rc = O.orc_scatter(ctypes.c_long(n), 10, T.P(dofs), T.P(slot), T.P(A1), T.P(F1), ctypes.c_double(1e-100),
                       T.P(val_seq), T.P(rhs_seq))
    t_scatter_1 = time.perf_counter() - t0
    assert rc == 0, rc
    # all host CPUs: cells in chunks, colour-parallel scatter (the colouring
    # is setup, like the CSR pattern, and is not timed)
    threads = max(1, os.cpu_count() or 1)
    col = np.zeros(n, np.int32)
    tt = np.ascontiguousarray(tets_reordered, dtype=np.int32)
    n_col = O.orc_greedy_colour(ctypes.c_long(len(coords)), ctypes.c_long(n), T.P(tt), T.P(col))
    col_tets = np.argsort(col, kind="stable").astype(np.int32)
    col_ptr = np.concatenate([[0], np.cumsum(np.bincount(col, minlength=n_col))]).astype(np.int64)
    A, F = zeros(n, 100), zeros(n, 10)
    cuts = np.linspace(0, n, 4 * threads + 1).astype(int)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(lambda k: cells(int(cuts[k]), int(cuts[k + 1]), A, F), range(4 * threads)))
    t_cells_mt = time.perf_counter() - t0
    assert np.array_equal(A, A1) and np.array_equal(F, F1)
    val_col, rhs_col = zeros(nnz), zeros(ndof)
    t0 = time.perf_counter()
    rc = O.orc_scatter_coloured(int(n_col), T.P(col_ptr), T.P(col_tets), 10, T.P(dofs), T.P(slot), T.P(A), T.P(F),
                                ctypes.c_double(1e-100), threads, T.P(val_col), T.P(rhs_col))
    t_scatter_mt = time.perf_counter() - t0
    assert rc == 0, rc
    scale = max(np.abs(val_seq).max(), 1e-300)
    col_err = float(np.abs(val_col - val_seq).max() / scale)
    res = {"value": n / (t_cells_mt + t_scatter_mt), "unit": "element matrices/s", "cores": threads,
           "kind": kind, "host_cpus": os.cpu_count(), "cpu_model": cpu_model(),
           "single_thread_value": n / (t_cells_1 + t_scatter_1),
           "seconds": {"cells_all_cpus": t_cells_mt, "scatter_coloured_all_cpus": t_scatter_mt,
                       "cells_one_thread": t_cells_1, "scatter_sequential": t_scatter_1},
           "colours": int(n_col), "coloured_vs_sequential_max_err_rel_max": col_err,
           "sample": f"all {n} tets: P2 GRAD-GRAD element matrices with per-tet SPD D and unit-load vectors "
                     f"({'reference fem3Dtet compiled from the AniFem++ sources' if ref else 'oracle restatement'}), "
                     f"element-level Dirichlet, scatter into the same CSR: {threads} threads "
                     f"({t_cells_mt:.2f}s cells + {t_scatter_mt:.2f}s coloured scatter, {n_col} colours); "
                     f"one thread {t_cells_1:.2f}s cells + {t_scatter_1:.2f}s sequential scatter"}
    return res, val_seq, rhs_seq


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "assembly_p2_aniso_64cube_fp64_traffic.txt")
TRAFFIC_FILE_C3 = os.path.join(ROOT, "profiles", "batched_uneven_256_traffic.txt")
TRAFFIC_FILE_C4 = os.path.join(ROOT, "profiles", "elasticity_p2vec_48cube_traffic.txt")


def timed_events(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def hbm_roofline(alg_bytes, ms, kernel, per_unit, traffic_file=None):
    """Roofline entry of a section: algorithmic bytes of one pass over its
    measured time against the HBM peak; with the measured HBM bytes per pass
    of the committed counter summary `traffic_file` when it is present."""
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_pass": int(alg_bytes), "kernel": kernel,
           "bytes_model": per_unit}
    if traffic_file:
        traffic, src = read_traffic(traffic_file)
        out["traffic"] = traffic
        out["traffic_source"] = src
    return out


def assembly_bytes(n_tets, nf, coef_bytes, nnz, ndof):
    """Algorithmic bytes of one FP64 assembly pass: per tet 96 B coordinates,
    the coefficient, 4 nf^2 B slot map and 4 nf B DOF ids; 16 B per stored
    nonzero and per DOF (read + write of the FP64 value)."""
    return n_tets * (96 + coef_bytes + 4 * nf * nf + 4 * nf) + 16 * nnz + 16 * ndof


def bench_batched(args, torch):
    """256 jittered cubes, side n log-uniform in [4, 40], P1/P2/P3 uniform,
    per-tet kappa = 10^U[-1,1], f = 1, boundary Dirichlet: one graph launch."""
    import anifem_amd as af
    rng = np.random.default_rng(3)
    sides = np.round(np.exp(rng.uniform(np.log(4), np.log(40), 256))).astype(int)
    fems = rng.integers(af.FEM_P1, af.FEM_P3 + 1, 256)
    meshes = [af.kuhn_cube(int(n), 0.2, 1000 + i) + (int(f),) for i, (n, f) in enumerate(zip(sides, fems))]
    t0 = time.perf_counter()
    B = af.Batch(meshes, autotune=True)
    setup_s = time.perf_counter() - t0
    kappa = torch.from_numpy(10 ** rng.uniform(-1, 1, B.n_tets)).cuda()
    fv = torch.ones(1, dtype=torch.float64, device="cuda")
    val, rhs, st = B.run(kappa, fv)
    assert st.code == 0
    ms = timed_events(lambda: B.run(kappa, fv, val, rhs, status=False), 10, 2, torch)
    tuning = {f"P{f - 1}": {"budget_slots_entries_pairs": [int(x) for x in B.tuning(f)[0]],
                            "candidate_ms": [float(x) for x in B.tuning(f)[1]],
                            "launch_shape": B.launch_tuning(f)} for f in (2, 3, 4)}
    nf_of = {af.FEM_P1: 4, af.FEM_P2: 10, af.FEM_P3: 20}
    per_mesh_tets = [6 * int(m) ** 3 for m in sides]
    alg = sum(assembly_bytes(nt, nf_of[int(fm)], 8, 0, 0) for nt, fm in zip(per_mesh_tets, fems)) \
        + 16 * B.nnz + 16 * B.ndof
    out = {"meshes": 256, "n_tets": B.n_tets, "ndof": B.ndof, "nnz": B.nnz, "row_blocks": B.n_blocks,
           "ms_per_pass": ms, "element_matrices_per_s": B.n_tets / ms * 1e3, "nnz_per_s": B.nnz / ms * 1e3,
           "setup_s": setup_s, "autotune": tuning,
           "roofline": hbm_roofline(alg, ms, "geometry_kernel + row_block_persistent per element order (one graph)",
                                    "per tet 96 B xyz + 8 B kappa + 4 nf^2 B slots + 4 nf B dofs; "
                                    "16 B per nnz and per DOF", TRAFFIC_FILE_C3),
           "note": "one geometry launch for all meshes and one row-block launch per element order, replayed from "
                   "a HIP graph (FP64)"}
    del B, val, rhs
    return out


def bench_elasticity(args, torch):
    """P2^3 linear elasticity on the 48^3 jittered cube (663,552 tets):
    per-tet Lame parameters from E log-uniform [1, 1e3], nu ~ U[0.2, 0.45],
    gravity load, boundary Dirichlet; FP64 and FP32 contraction."""
    import anifem_amd as af
    rng = np.random.default_rng(4)
    coords, tets = af.kuhn_cube(48, 0.2, 1)
    n = len(tets)
    E = 10 ** rng.uniform(0, 3, n)
    nu = rng.uniform(0.2, 0.45, n)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    t0 = time.perf_counter()
    A = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P2, 3), order=2, tensor_type=af.TENSOR_LAME,
                     tensor_storage=af.COEF_PER_TET, rhs_tensor_type=af.TENSOR_GENERAL,
                     rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_BOUNDARY)
    setup_s = time.perf_counter() - t0
    D = torch.from_numpy(np.stack([lam, mu])).cuda()
    g = torch.tensor([0.0, 0.0, -1.0], dtype=torch.float64, device="cuda")
    val = torch.empty(A.nnz, dtype=torch.float64, device="cuda")
    rhs = torch.empty(A.ndof, dtype=torch.float64, device="cuda")
    _, _, st = A.assemble(D=D, F=g, val=val, rhs=rhs)
    assert st.code == 0
    ref, rhs64 = val.clone(), rhs.clone()
    res = {"n_tets": n, "ndof": A.ndof, "nnz": A.nnz, "setup_s": setup_s}
    for prec, name in [(af.PREC_FP64, "fp64"), (af.PREC_FP32, "fp32")]:
        ms = timed_events(lambda: A.assemble(D=D, F=g, precision=prec, val=val, rhs=rhs, status=False), 10, 2,
                          torch)
        res[name] = {"ms_per_pass": ms, "element_matrices_per_s": n / ms * 1e3, "nnz_per_s": A.nnz / ms * 1e3}
        if name == "fp64":
            res["roofline"] = hbm_roofline(assembly_bytes(n, 30, 16, A.nnz, A.ndof), ms,
                                           "geometry_kernel<3,3,3,3,double,false,true> + row_block_persistent_vec<3,3,3,10,double,true> (P2^3, compact Lame rows)",
                                           "per tet 96 B xyz + 16 B (lambda, mu) + 4 nf^2 B slots + 4 nf B dofs; "
                                           "16 B per nnz and per DOF", TRAFFIC_FILE_C4)
    A.assemble(D=D, F=g, precision=af.PREC_FP32, val=val, rhs=rhs, status=False)
    torch.cuda.synchronize()
    res["fp32_max_entry_error_rel_to_max"] = float((val - ref).abs().max() / ref.abs().max())
    # element matrices (30 x 30) of every tet: FP32 contraction against FP64
    xyz = torch.from_numpy(np.ascontiguousarray(coords[tets].reshape(n, 12).T)).to("cuda")
    op = af.Op(af.GRAD, af.FEM_P2, 3)
    E64 = af.fem3dtet_batch(op, op, af.TENSOR_LAME, af.COEF_PER_TET, D, 2, xyz, af.PREC_FP64)
    E32 = af.fem3dtet_batch(op, op, af.TENSOR_LAME, af.COEF_PER_TET, D, 2, xyz, af.PREC_FP32)
    err = (E32.double() - E64).flatten(1).norm(dim=1) / E64.flatten(1).norm(dim=1)
    res["fp32_element_rel_frobenius_error"] = {"max": err.max().item(), "median": err.median().item()}
    af.fem3dtet_batch(op, op, af.TENSOR_LAME, af.COEF_PER_TET, D, 2, xyz, af.PREC_FP16X3, out=E32)
    err = (E32.double() - E64).flatten(1).norm(dim=1) / E64.flatten(1).norm(dim=1)
    res["fp16x3_mfma_element_rel_frobenius_error"] = {"max": err.max().item(), "median": err.median().item()}
    elem = {}
    for prec, name, outp in [(af.PREC_FP64, "fp64", E64), (af.PREC_FP32, "fp32", E32),
                             (af.PREC_FP16X3, "fp16x3_mfma", E32)]:
        ms = timed_events(lambda: af.fem3dtet_batch(op, op, af.TENSOR_LAME, af.COEF_PER_TET, D, 2, xyz, prec,
                                                    out=outp), 5, 1, torch)
        bytes_ = n * (96 + 16 + 900 * outp.element_size())
        elem[name] = {"ms": ms, "element_matrices_per_s": n / ms * 1e3, "GB_per_s": bytes_ / ms / 1e6}
    res["element_only_30x30"] = elem
    del E64, E32, err, xyz
    # the FP16 hi + lo MFMA mode end to end: element matrices and load vectors
    # on the reordered tets, scattered through the row-block lists
    xyzr = torch.from_numpy(np.ascontiguousarray(coords[A.tets()].reshape(n, 12).T)).to("cuda")
    p0, pv = af.Op(af.IDEN, af.FEM_P0), af.Op(af.IDEN, af.FEM_P2, 3)
    Ae = torch.empty((n, 30, 30), dtype=torch.float32, device="cuda")
    Fe = torch.empty((n, 1, 30), dtype=torch.float32, device="cuda")
    val16 = torch.empty_like(val)
    rhs16 = torch.empty_like(rhs)

    def mfma_pass():
        af.fem3dtet_batch(op, op, af.TENSOR_LAME, af.COEF_PER_TET, D, 2, xyzr, af.PREC_FP16X3, out=Ae)
        af.fem3dtet_batch(p0, pv, af.TENSOR_GENERAL, af.COEF_CONST, g, 2, xyzr, af.PREC_FP16X3, out=Fe)
        A.assemble_elements(Ae, Fe.view(n, 30), val=val16, rhs=rhs16, status=False, symmetric=True)

    ms16 = timed_events(mfma_pass, 5, 1, torch)
    res["fp16x3_mfma_elements_then_scatter"] = {
        "ms_per_pass": ms16, "element_matrices_per_s": n / ms16 * 1e3,
        "max_entry_error_rel_to_max": float((val16 - ref).abs().max() / ref.abs().max())}
    del Ae, Fe, xyzr
    # displacements: Jacobi-PCG to rel-res 1e-10 on the FP64-assembled system
    # and on the system (matrix and load) assembled with the FP32 contraction
    rp, ci = A.csr()
    u64 = torch.zeros(A.ndof, dtype=torch.float64, device="cuda")
    u32 = torch.zeros_like(u64)
    t0 = time.perf_counter()
    r64, ok64 = af.pcg_jacobi(A.ndof, rp, ci, ref, rhs64, u64, rtol=1e-10, maxit=100000)
    r32, ok32 = af.pcg_jacobi(A.ndof, rp, ci, val, rhs, u32, rtol=1e-10, maxit=100000)
    u16 = torch.zeros_like(u64)
    r16, ok16 = af.pcg_jacobi(A.ndof, rp, ci, val16, rhs16, u16, rtol=1e-10, maxit=100000)
    torch.cuda.synchronize()
    res["displacement"] = {
        "rtol": 1e-10, "converged": bool(ok64 and ok32 and ok16), "iterations_fp64": r64.iterations,
        "iterations_fp32": r32.iterations, "iterations_fp16x3_mfma": r16.iterations,
        "solve_seconds_all": time.perf_counter() - t0,
        "fp32_rel_l2_error": float((u32 - u64).norm() / u64.norm()),
        "fp32_max_error_rel_to_max": float((u32 - u64).abs().max() / u64.abs().max()),
        "fp16x3_mfma_rel_l2_error": float((u16 - u64).norm() / u64.norm()),
        "fp16x3_mfma_max_error_rel_to_max": float((u16 - u64).abs().max() / u64.abs().max())}
    del A, val, rhs, ref, rhs64, u64, u32, u16, val16, rhs16
    return res


def read_traffic(path=TRAFFIC_FILE):
    """HBM bytes per assembly pass from a committed FETCH_SIZE / WRITE_SIZE
    summary (copy-kernel-calibrated), or None when it is absent."""
    try:
        kv = dict(line.split(": ", 1) for line in open(path).read().splitlines() if ": " in line)
        return int(kv["traffic_bytes_per_pass"]), os.path.relpath(path, ROOT)
    except (OSError, KeyError, ValueError):
        return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=64, help="sub-cubes per axis")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-extras", action="store_true", help="only the headline assembly (no --sections)")
    ap.add_argument("--sections", default="precision,pcg,batched,elasticity",
                    help="comma list of extra sections: precision, pcg, batched, elasticity")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    ap.add_argument("--learned-train", type=int, default=2000, help="training systems (learned preconditioner)")
    ap.add_argument("--learned-eval", type=int, default=200, help="held-out systems (learned preconditioner)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only; runs on several GPUs are not provided "
              "(the owner-computes row blocks shard by owned rows with one ghost layer of tets)", file=sys.stderr)
        sys.exit(2)

    import torch
    import anifem_amd as af

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(1)

    n = args.n
    coords, tets = af.kuhn_cube(n, 0.2, 1)
    n_tets = len(tets)
    t0 = time.perf_counter()
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P2), order=2, tensor_type=af.TENSOR_SYMMETRIC,
                       tensor_storage=af.COEF_PER_TET, rhs_order=2, rhs_tensor_type=af.TENSOR_SCALAR,
                       rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_BOUNDARY)
    setup_s = time.perf_counter() - t0
    Dp_np, Dfull = aniso_packed(rng, n_tets)
    Dp = torch.from_numpy(np.ascontiguousarray(Dp_np)).to(dev)
    f = torch.ones(1, dtype=torch.float64, device=dev)
    val = torch.empty(Asm.nnz, dtype=torch.float64, device=dev)
    rhs = torch.empty(Asm.ndof, dtype=torch.float64, device=dev)
    print(f"# setup: {n_tets} tets, {Asm.ndof} dofs, {Asm.nnz} nnz, {Asm.n_blocks} row blocks, {setup_s:.1f}s",
          file=sys.stderr, flush=True)

    def step(prec):
        Asm.assemble(D=Dp, F=f, precision=prec, val=val, rhs=rhs, status=False, stream=stream)

    def timed(prec, steps, warmup, events=None):
        for _ in range(warmup):
            step(prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        if events:
            events[0].record(stream)
        for _ in range(steps):
            step(prec)
        if events:
            events[1].record(stream)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    if args.copy_calibration:
        # known-size traffic: a DMA-style copy (copyBuffer) and an elementwise
        # kernel that reads and writes the same bytes
        src = torch.ones(args.copy_calibration * 2 ** 17, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    _, _, st = Asm.assemble(D=Dp, F=f, val=val, rhs=rhs, status=True, stream=stream)
    assert st.code == 0 and st.n_degenerate == 0, st.code
    # HIP events on the launch stream bracket the timed passes
    ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
    sec64 = timed(af.PREC_FP64, args.steps, args.warmup, ev)
    ms_pass_events = ev[0].elapsed_time(ev[1]) / args.steps
    print(f"# fp64 assembly {sec64 * 1e3:.3f} ms/step", file=sys.stderr, flush=True)

    # per-phase GPU times (HIP events on the launch stream)
    Asm.enable_timers(True)
    reps = 5
    tm = np.zeros(2)
    for _ in range(reps):
        Asm.assemble(D=Dp, F=f, val=val, rhs=rhs, status=False, stream=stream)
        t = Asm.timers()
        tm += [t.ms_geometry, t.ms_gather]
    tm /= reps
    nlaunch = Asm.timers().n_launches
    Asm.enable_timers(False)
    # algorithmic bytes of the fused assembly unit (FP64): per tet 96 B
    # coordinates + 48 B packed tensor + 4*nf^2 B slot map + 4*nf B DOF ids,
    # plus 16 B per stored nonzero and 16 B per DOF (read + write of the FP64
    # value); priced against the two kernels of the pass
    per_tet = 96 + 48 + 4 * 100 + 4 * 10
    alg_bytes = n_tets * per_tet + 16 * Asm.nnz + 16 * Asm.ndof
    achieved = alg_bytes / (ms_pass_events * 1e-3) / 1e9
    traffic, traffic_src = read_traffic()

    out = {
        "metric": "element matrices/s (P2 anisotropic-diffusion assembly into FP64 CSR + RHS)",
        "value": n_tets / sec64,
        "unit": "element matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec64 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered Kuhn cube, seeded Haar rotations)",
        "config": {"workload": f"P2 anisotropic diffusion assembly, {n}^3 Kuhn cube", "n_tets": n_tets,
                   "ndof": Asm.ndof, "nnz": Asm.nnz, "quadrature_points": 4, "row_blocks": Asm.n_blocks,
                   "dirichlet": "homogeneous, boundary"},
        "nnz_per_s": Asm.nnz / sec64,
        "phase_ms": {"geometry": tm[0], "row_blocks": tm[1], "launches": nlaunch},
        "setup_s": setup_s,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "assembly pass: geometry_kernel<3,3,1,1,double,true,false> + row_block_persistent<6,1,5,double,true,10,false,false>",
                     "ms_per_launch": ms_pass_events,
                     "algorithmic_bytes_per_pass": alg_bytes,
                     "traffic_source": traffic_src},
    }

    sections = set() if args.skip_extras else {x.strip() for x in args.sections.split(",") if x.strip()}
    if "precision" in sections:
        # the coloured scatter (greedy colouring, one launch per colour) beside the row blocks
        Asm.set_scatter(af.SCATTER_COLOURED)
        n_col, _ = Asm.colours()
        sec_col = timed(af.PREC_FP64, max(5, args.steps // 2), 2)
        Asm.set_scatter(af.SCATTER_ROW_BLOCKS)
        out["coloured_scatter_fp64"] = {"ms_per_step": sec_col * 1e3, "n_colours": n_col,
                                        "row_blocks_speedup": sec_col / sec64}
        sec32 = timed(af.PREC_FP32, max(5, args.steps // 2), 2)
        out["fp32_contraction"] = {"ms_per_step": sec32 * 1e3, "element_matrices_per_s": n_tets / sec32,
                                   "speedup_vs_fp64": sec64 / sec32}
        # element matrices only (no scatter): batch API on SoA coordinates
        xyz = torch.from_numpy(np.ascontiguousarray(coords[tets].reshape(n_tets, 12).T)).to(dev)
        op = af.Op(af.GRAD, af.FEM_P2)
        res = {}
        for prec, name in [(af.PREC_FP64, "fp64"), (af.PREC_FP32, "fp32"), (af.PREC_FP16X3, "fp16x3_mfma")]:
            dt = torch.float64 if prec == af.PREC_FP64 else torch.float32
            A = torch.empty((n_tets, 10, 10), dtype=dt, device=dev)
            for _ in range(2):
                af.fem3dtet_batch(op, op, af.TENSOR_SYMMETRIC, af.COEF_PER_TET, Dp, 2, xyz, prec, out=A)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = 10
            e0.record()
            for _ in range(reps):
                af.fem3dtet_batch(op, op, af.TENSOR_SYMMETRIC, af.COEF_PER_TET, Dp, 2, xyz, prec, out=A)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            bytes_ = n_tets * (96 + 48 + 100 * (8 if prec == af.PREC_FP64 else 4))
            res[name] = {"ms": ms, "element_matrices_per_s": n_tets / (ms * 1e-3),
                         "GB_per_s": bytes_ / (ms * 1e-3) / 1e9}
            del A
        res["fp32_speedup_vs_fp64"] = res["fp64"]["ms"] / res["fp32"]["ms"]
        # per-element error of the FP32 contraction against FP64
        A64 = torch.empty((n_tets, 10, 10), dtype=torch.float64, device=dev)
        A32 = torch.empty((n_tets, 10, 10), dtype=torch.float32, device=dev)
        af.fem3dtet_batch(op, op, af.TENSOR_SYMMETRIC, af.COEF_PER_TET, Dp, 2, xyz, af.PREC_FP64, out=A64)
        af.fem3dtet_batch(op, op, af.TENSOR_SYMMETRIC, af.COEF_PER_TET, Dp, 2, xyz, af.PREC_FP32, out=A32)
        err = (A32.double() - A64).flatten(1).norm(dim=1) / A64.flatten(1).norm(dim=1)
        res["fp32_element_rel_frobenius_error"] = {"max": err.max().item(), "median": err.median().item(),
                                                   "p99.9": torch.quantile(err.float()[::7], 0.999).item()}
        af.fem3dtet_batch(op, op, af.TENSOR_SYMMETRIC, af.COEF_PER_TET, Dp, 2, xyz, af.PREC_FP16X3, out=A32)
        err = (A32.double() - A64).flatten(1).norm(dim=1) / A64.flatten(1).norm(dim=1)
        res["fp16x3_mfma_element_rel_frobenius_error"] = {
            "max": err.max().item(), "median": err.median().item(),
            "p99.9": torch.quantile(err.float()[::7], 0.999).item()}
        # FP16X3 (MFMA) elements on the reordered tets, scattered through the row blocks
        xyzr = torch.from_numpy(np.ascontiguousarray(coords[Asm.tets()].reshape(n_tets, 12).T)).to(dev)
        p0, p2 = af.Op(af.IDEN, af.FEM_P0), af.Op(af.IDEN, af.FEM_P2)
        Fe = torch.empty((n_tets, 1, 10), dtype=torch.float32, device=dev)
        val16 = torch.empty_like(val)
        rhs16 = torch.empty_like(rhs)

        def mfma_pass():
            af.fem3dtet_batch(op, op, af.TENSOR_SYMMETRIC, af.COEF_PER_TET, Dp, 2, xyzr, af.PREC_FP16X3, out=A32)
            af.fem3dtet_batch(p0, p2, af.TENSOR_SCALAR, af.COEF_CONST, f, 2, xyzr, af.PREC_FP16X3, out=Fe)
            Asm.assemble_elements(A32, Fe.view(n_tets, 10), val=val16, rhs=rhs16, status=False)

        ms16 = timed_events(mfma_pass, 5, 1, torch)
        Asm.assemble(D=Dp, F=f, val=val, rhs=rhs, status=False, stream=stream)
        torch.cuda.synchronize()
        res["fp16x3_mfma_elements_then_scatter"] = {
            "ms_per_pass": ms16, "element_matrices_per_s": n_tets / ms16 * 1e3,
            "max_entry_error_rel_to_max": float((val16 - val).abs().max() / val.abs().max())}
        del A64, A32, err, Fe, val16, rhs16, xyzr
        out["element_only"] = res
        # the plain FP64 method on the GPU (reference algorithm per tet, FP64
        # atomic scatter) against the mixed-precision row-block assembly
        plain_ms = timed_events(lambda: Asm.assemble_plain(D=Dp, F=f, val=val, rhs=rhs, status=False), 5, 1, torch)
        out["mixed_precision"] = {
            "plain_fp64_kernel_ms": plain_ms, "rowblock_fp64_ms": sec64 * 1e3, "rowblock_fp32_ms": sec32 * 1e3,
            "fp32_speedup_vs_plain_fp64": plain_ms / (sec32 * 1e3),
            "fp32_speedup_vs_rowblock_fp64": sec64 / sec32,
            "fp32_element_rel_frobenius_error_max": res["fp32_element_rel_frobenius_error"]["max"],
            "fp16x3_mfma_element_ms": res["fp16x3_mfma"]["ms"],
            "fp16x3_mfma_element_rel_frobenius_error_max": res["fp16x3_mfma_element_rel_frobenius_error"]["max"],
            "fp16x3_mfma_element_speedup_vs_plain_fp64": plain_ms / res["fp16x3_mfma"]["ms"]}
        del xyz
    if "pcg" in sections:
        # Jacobi-PCG on the 16^3 P1 Poisson system (rel-res 1e-10)
        c1, t1 = af.kuhn_cube(16, 0.2, 1)
        A1 = af.Assembler(c1, t1, af.Op(af.GRAD, af.FEM_P1), dirichlet=af.DIRICHLET_BOUNDARY)
        v1, b1, _ = A1.assemble(D=None, F=f)
        rp, ci = A1.csr()
        x = torch.zeros(A1.ndof, dtype=torch.float64, device=dev)
        af.pcg_jacobi(A1.ndof, rp, ci, v1, b1, x, rtol=1e-10)
        x.zero_()
        torch.cuda.synchronize()
        tp = time.perf_counter()
        r, ok = af.pcg_jacobi(A1.ndof, rp, ci, v1, b1, x, rtol=1e-10)
        torch.cuda.synchronize()
        out["pcg_jacobi_p1_16"] = {"ms": (time.perf_counter() - tp) * 1e3, "iterations": r.iterations,
                                   "true_relres": r.true_relres, "converged": ok}
        # learned preconditioner: P1 GRF diffusion at 32^3, trained on 2,000
        # seeded systems from this assembler, 200 held-out seeds, rel-res 1e-8
        from anifem_amd.learned import GrfSystems, train, compare
        grf = GrfSystems(n=32, ell=0.1, sigma=1.5)
        net, losses, train_s = train(grf, np.arange(args.learned_train), lr=2e-3, epochs=1)
        rows = compare(grf, net, np.arange(1_000_000, 1_000_000 + args.learned_eval), rtol=1e-8)
        jm = np.array([r["jacobi_ms"] for r in rows])
        lm = np.array([r["learned_ms"] for r in rows])
        out["pcg_learned_vs_jacobi_grf32"] = {
            "ndof": grf.n, "nnz": grf.asm.nnz, "train_systems": int(args.learned_train),
            "train_seconds": train_s, "loss_first_100": float(losses[:100].mean()),
            "loss_last_100": float(losses[-100:].mean()), "eval_systems": len(rows), "rtol": 1e-8,
            "jacobi_ms_mean": float(jm.mean()), "learned_ms_mean": float(lm.mean()),
            "speedup_mean": float(jm.mean() / lm.mean()),
            "factor_ms_mean": float(np.mean([r["factor_ms"] for r in rows])),
            "jacobi_iterations_mean": float(np.mean([r["jacobi_it"] for r in rows])),
            "learned_iterations_mean": float(np.mean([r["learned_it"] for r in rows])),
            "learned_faster_fraction": float(np.mean(lm < jm)),
            "fallbacks": int(sum(r["fallback"] for r in rows)),
            "max_true_relres": float(max(max(r["learned_true"], r["jacobi_true"]) for r in rows)),
            "all_converged": bool(all(r["ok"] for r in rows)),
            "note": "wall time of each solve call (learned includes the network's factor), FP64 true residual"}
        # the non-learned factor on the same pattern (FSAI), both factors in
        # the same PCG with M^-1 = G G^T applied in turn
        from anifem_amd.learned import compare_fsai
        frows = compare_fsai(grf, net, np.arange(1_000_000, 1_000_000 + min(20, args.learned_eval)), rtol=1e-8)
        out["pcg_learned_vs_jacobi_grf32"]["vs_fsai_same_pattern"] = {
            "systems": len(frows),
            "fsai_iterations_mean": float(np.mean([r["fsai_it"] for r in frows])),
            "learned_iterations_mean": float(np.mean([r["learned_it"] for r in frows])),
            "fsai_solve_ms_mean": float(np.mean([r["fsai_ms"] for r in frows])),
            "learned_solve_ms_mean": float(np.mean([r["learned_ms"] for r in frows])),
            "fsai_host_setup_ms_mean": float(np.mean([r["fsai_setup_ms"] for r in frows])),
            "all_converged": bool(all(r["ok"] for r in frows)),
            "note": "anifem_pcg_factored with each factor; the learned time includes the network's factor, the "
                    "FSAI factor is built on the host outside the timed solve"}
        del net, grf
    if "batched" in sections:
        out["batched_uneven_256"] = bench_batched(args, torch)
    if "elasticity" in sections:
        out["elasticity_p2vec_48"] = bench_elasticity(args, torch)

    if not args.no_cpu_baseline:
        s = Asm.setup()
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import anifem_testlib as T
        # the assembler's node reorder (ordering.inl:8-26) before the element calls
        tt = np.ascontiguousarray(tets.copy())
        T.oracle().orc_reorder_nodes(ctypes.c_long(len(tt)), T.P(tt), T.P(coords))
        out["cpu_baseline"], val_ref, rhs_ref = cpu_baseline(coords, tt, Dfull, s["dofs"], s["slot"],
                                                             s["dirichlet"], Asm.nnz, Asm.ndof)
        # full-size parity: the timed FP64 pass against the reference path's
        # result (reference element core + sequential scatter), entry error
        # relative to the row's largest entry
        Asm.assemble(D=Dp, F=f, val=val, rhs=rhs, status=False, stream=stream)
        torch.cuda.synchronize()
        v = val.cpu().numpy()
        r = rhs.cpu().numpy()
        rp = s["row_ptr"]
        rowmax = np.maximum.reduceat(np.abs(val_ref), rp[:-1])
        ent = float((np.abs(v - val_ref) / np.repeat(rowmax, np.diff(rp))).max())
        rerr = float(np.abs(r - rhs_ref).max() / np.abs(rhs_ref).max())
        out["parity_full_size"] = {
            "against": "reference fem3Dtet element matrices and load vectors + applyDir + sequential scatter",
            "fp64_max_entry_err_rel_row_max": ent, "fp64_max_rhs_err_rel_max": rerr,
            "bound": 1e-12, "ok": bool(ent <= 1e-12 and rerr <= 1e-12)}
        del val_ref, rhs_ref, v, r
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
