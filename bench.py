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


def cpu_baseline(coords, tets_reordered, Dfull, dofs, slot, is_dir, nnz, ndof):
    """Reference element core (oracle/_ref, compiled from the AniFem++ sources)
    on a bounded sample of the workload's tets, element-level homogeneous
    Dirichlet (dc_on_dof.h:38-45), then the oracle's sequential cell-order
    scatter; single thread on the host.  Falls back to the oracle's own
    element restatement where the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import anifem_testlib as T

    ref = T.ref_lib()
    kind = "reference" if ref is not None else "port"
    n_all = len(tets_reordered)
    n = n_all
    xyz = coords[tets_reordered[:n]].reshape(n, 12)
    Df = np.ascontiguousarray(Dfull[:n].transpose(0, 2, 1).reshape(n, 9))
    t0 = time.perf_counter()
    if ref is not None:
        A = np.zeros(n * 100)
        dp = ctypes.POINTER(ctypes.c_double)
        ref.ref_fem3dtet.restype = ctypes.c_int
        rc = ref.ref_fem3dtet(1, 3, 0, 2, ctypes.c_long(n), xyz.ctypes.data_as(dp), Df.ctypes.data_as(dp),
                              ctypes.c_long(9), A.ctypes.data_as(dp))
        assert rc == 0
        A = A.reshape(n, 100)
    else:
        A = T.orc_elem(2, 3, 1, 2, 3, 1, 3, 1, 2, xyz, Df)
    t1 = time.perf_counter()
    # applyDir with value 0 on every flagged local DOF: zero row and column, unit diagonal
    dmask = is_dir[dofs[:n]].astype(bool)  # (n, 10) local flags
    A3 = A.reshape(n, 10, 10)  # [tet, col ia, row ib] (column-major per tet)
    A3[dmask[:, :, None] | dmask[:, None, :]] = 0.0
    idx = np.nonzero(dmask)
    A3[idx[0], idx[1], idx[1]] = 1.0
    rc, _, _ = T.oracle_scatter(n, 10, np.ascontiguousarray(dofs[:n]), np.ascontiguousarray(slot[:n]), A, None, nnz,
                                ndof)
    assert rc == 0, rc
    t2 = time.perf_counter()
    rate = n / (t2 - t0)
    return {"value": rate, "unit": "element matrices/s", "cores": 1, "kind": kind,
            "sample": f"all {n} of {n_all} tets: P2 GRAD-GRAD element matrices with per-tet SPD D "
                      f"({'reference fem3Dtet compiled from the AniFem++ sources' if ref else 'oracle restatement'}, "
                      f"{t1 - t0:.2f}s) + oracle sequential scatter into the same CSR ({t2 - t1:.2f}s)"}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "assembly_p2_aniso_64cube_fp64_traffic.txt")


def read_traffic():
    """HBM bytes per assembly pass from the committed FETCH_SIZE / WRITE_SIZE
    summary (copy-kernel-calibrated), or None when it is absent."""
    try:
        kv = dict(line.split(": ", 1) for line in open(TRAFFIC_FILE).read().splitlines() if ": " in line)
        return int(kv["traffic_bytes_per_pass"]), os.path.relpath(TRAFFIC_FILE, ROOT)
    except (OSError, KeyError, ValueError):
        return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=64, help="sub-cubes per axis")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-extras", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    ap.add_argument("--learned-train", type=int, default=2000, help="training systems (learned preconditioner)")
    ap.add_argument("--learned-eval", type=int, default=200, help="held-out systems (learned preconditioner)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build targets one GPU; multi-GPU runs are not part of it "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
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
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "assembly pass: geometry_kernel<3,3,1,1,double> + row_block_kernel<3,3,1,1,double,true>",
                     "ms_per_launch": ms_pass_events,
                     "algorithmic_bytes_per_pass": alg_bytes,
                     "traffic_source": traffic_src},
    }

    if not args.skip_extras:
        sec32 = timed(af.PREC_FP32, max(5, args.steps // 2), 2)
        out["fp32_contraction"] = {"ms_per_step": sec32 * 1e3, "element_matrices_per_s": n_tets / sec32,
                                   "speedup_vs_fp64": sec64 / sec32}
        # element matrices only (no scatter): batch API on SoA coordinates
        xyz = torch.from_numpy(np.ascontiguousarray(coords[tets].reshape(n_tets, 12).T)).to(dev)
        op = af.Op(af.GRAD, af.FEM_P2)
        res = {}
        for prec, name in [(af.PREC_FP64, "fp64"), (af.PREC_FP32, "fp32")]:
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
        out["element_only"] = res
        del xyz
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
            "jacobi_iterations_mean": float(np.mean([r["jacobi_it"] for r in rows])),
            "learned_iterations_mean": float(np.mean([r["learned_it"] for r in rows])),
            "learned_faster_fraction": float(np.mean(lm < jm)),
            "fallbacks": int(sum(r["fallback"] for r in rows)),
            "max_true_relres": float(max(max(r["learned_true"], r["jacobi_true"]) for r in rows)),
            "all_converged": bool(all(r["ok"] for r in rows)),
            "note": "wall time of each solve call (learned includes the network's factor), FP64 true residual"}
        del net, grf

    if not args.no_cpu_baseline:
        s = Asm.setup()
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import anifem_testlib as T
        # the assembler's node reorder (ordering.inl:8-26) before the element calls
        tt = np.ascontiguousarray(tets.copy())
        T.oracle().orc_reorder_nodes(ctypes.c_long(len(tt)), T.P(tt), T.P(coords))
        out["cpu_baseline"] = cpu_baseline(coords, tt, Dfull, s["dofs"], s["slot"], s["dirichlet"], Asm.nnz,
                                           Asm.ndof)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
