"""Batched assembly of uneven meshes with mixed element order (anifem_batch_*):
every mesh's CSR values and rhs are bitwise identical to assembling that mesh
alone (anifem_assembly_run), replays of the captured graph reproduce them, and
the autotuner reports a budget per element order."""
import re
from pathlib import Path

import numpy as np
import pytest

import anifem_amd as af

pytestmark = pytest.mark.gpu

SPECS = [(3, af.FEM_P1), (5, af.FEM_P2), (4, af.FEM_P3), (6, af.FEM_P1), (2, af.FEM_P3), (7, af.FEM_P2),
         (9, af.FEM_P1)]


def build(specs, seed=0):
    import torch
    rng = np.random.default_rng(seed)
    meshes, kappas = [], []
    for i, (n, fem) in enumerate(specs):
        c, t = af.kuhn_cube(n, 0.2, 100 + i)
        meshes.append((c, t, fem))
        kappas.append(10 ** rng.uniform(-1, 1, len(t)))
    kappa = torch.from_numpy(np.concatenate(kappas)).cuda()
    return meshes, kappas, kappa


def test_batch_matches_single_mesh_assembly_bitwise():
    import torch
    meshes, kappas, kappa = build(SPECS)
    f = torch.full((1,), 1.5, dtype=torch.float64, device="cuda")
    B = af.Batch(meshes, dirichlet=af.DIRICHLET_BOUNDARY, autotune=True)
    assert B.n_tets == sum(len(t) for _, t, _ in meshes)
    val, rhs, st = B.run(kappa, f)
    assert st.code == 0 and st.n_degenerate == 0
    val_h, rhs_h = val.cpu().numpy(), rhs.cpu().numpy()
    for i, (c, t, fem) in enumerate(meshes):
        order = 4 if fem == af.FEM_P3 else 2
        A = af.Assembler(c, t, af.Op(af.GRAD, fem), order=order, tensor_type=af.TENSOR_SCALAR,
                         tensor_storage=af.COEF_PER_TET, rhs_order=order, rhs_tensor_type=af.TENSOR_SCALAR,
                         rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_BOUNDARY)
        v1, r1, _ = A.assemble(D=torch.from_numpy(kappas[i][None, :].copy()).cuda(), F=f)
        inf = B.info(i)
        assert (inf["ndof"], inf["nnz"]) == (A.ndof, A.nnz)
        sl = slice(inf["nnz_offset"], inf["nnz_offset"] + inf["nnz"])
        sr = slice(inf["dof_offset"], inf["dof_offset"] + inf["ndof"])
        np.testing.assert_array_equal(val_h[sl], v1.cpu().numpy())
        np.testing.assert_array_equal(rhs_h[sr], r1.cpu().numpy())
    # replays of the captured graph, with other device work between them
    for rep in range(3):
        junk = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        v2, r2, st2 = B.run(kappa, f, val=val, rhs=rhs)
        assert st2.code == 0
        np.testing.assert_array_equal(v2.cpu().numpy(), val_h)
        np.testing.assert_array_equal(r2.cpu().numpy(), rhs_h)
        del junk
    # a new coefficient through the same graph (new pointers: captured again):
    # doubling kappa doubles every entry exactly, except the Dirichlet
    # diagonals (incident-tet counts, independent of kappa)
    kappa2 = kappa * 2.0
    v3, r3, _ = B.run(kappa2, f)
    v3 = v3.cpu().numpy()
    same = v3 == val_h
    np.testing.assert_array_equal(v3[~same], 2.0 * val_h[~same])
    assert np.all(val_h[same] == np.round(val_h[same])) and np.all(val_h[same] >= 1)
    np.testing.assert_array_equal(r3.cpu().numpy(), rhs_h)


def test_autotuner_reports_budgets():
    meshes, _, _ = build([(6, af.FEM_P1), (5, af.FEM_P2), (3, af.FEM_P3)])
    B = af.Batch(meshes, autotune=True)
    cands = af.BATCH_BUDGET_CANDIDATES
    hdr = (Path(__file__).resolve().parents[1] / "include" / "anifem_gpu.h").read_text()
    assert len(cands) == int(re.search(r"#define ANIFEM_BATCH_BUDGET_CANDIDATES (\d+)", hdr).group(1))
    for fem in (af.FEM_P1, af.FEM_P2, af.FEM_P3):
        caps, ms = B.tuning(fem)
        assert tuple(int(x) for x in caps) in cands and len(ms) == len(cands)
        assert np.all((ms > 0) | (ms == -1)) and np.any(ms > 0)
        k = cands.index(tuple(int(x) for x in caps))
        assert ms[k] == ms[ms > 0].min()


def test_launch_shapes_timed_at_first_run_and_values_unchanged():
    """The launch-shape tuner times each order's row-block pass at the first
    run (fixed-size or general form, or two lanes per pair for P3; compact or
    16-bit positions; workgroups per CU); the tuned graph gives the untuned
    batch's values bit for bit."""
    import torch
    meshes, _, kappa = build(SPECS, seed=5)
    f = torch.ones(1, dtype=torch.float64, device="cuda")
    B1 = af.Batch(meshes, autotune=True)
    for fem in (af.FEM_P1, af.FEM_P2, af.FEM_P3):
        assert B1.launch_tuning(fem)["ms"] == -1
    v1, r1, st = B1.run(kappa, f)
    assert st.code == 0
    for fem in (af.FEM_P1, af.FEM_P2, af.FEM_P3):
        t = B1.launch_tuning(fem)
        assert t["ms"] > 0 and t["workgroups_per_cu"] >= 1 and t["compact"] in (0, 1)
        if fem == af.FEM_P3:
            caps, _ = B1.tuning(fem)
            assert t["form"] in ((0, 2) if caps[2] <= 128 else (0,))
            assert t["form"] == 0 or t["compact"] == 0
        else:
            assert t["form"] in (0, 1)
    B2 = af.Batch(meshes, autotune=False)
    v2, r2, _ = B2.run(kappa, f)
    assert B2.launch_tuning(af.FEM_P2)["ms"] == -1
    assert torch.equal(v1, v2) and torch.equal(r1, r2)


def test_batch_persistent_and_per_block_bitwise_equal():
    """The persistent row-block pass (element kinds restaged as a workgroup's
    block range crosses meshes of another order) against one workgroup per
    block: identical values and rhs."""
    import torch
    meshes, kappas, kappa = build(SPECS, seed=3)
    f = torch.ones(1, dtype=torch.float64, device="cuda")
    B = af.Batch(meshes, autotune=False)
    v1, r1, st = B.run(kappa, f)
    assert st.code == 0
    v1, r1 = v1.clone(), r1.clone()
    B.set_persistent(False)
    v2, r2, _ = B.run(kappa, f)
    assert torch.equal(v1, v2) and torch.equal(r1, r2)
    B.set_persistent(True)
    v3, r3, _ = B.run(kappa, f)
    assert torch.equal(v1, v3) and torch.equal(r1, r3)
