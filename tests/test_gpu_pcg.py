"""GPU Jacobi-PCG (anifem_pcg_jacobi) against the oracle's CG on the same
assembled system: the FP64 true residual meets the tolerance, the iteration
count matches within 2 and the solutions agree."""
import ctypes

import numpy as np
import pytest

import anifem_amd as af
import anifem_testlib as T

pytestmark = pytest.mark.gpu


def test_pcg_poisson_c1_shape():
    import torch
    coords, tets = af.kuhn_cube(16, 0.2, 1)
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P1), order=2, tensor_type=af.TENSOR_NULL,
                       tensor_storage=af.COEF_CONST, rhs_tensor_type=af.TENSOR_SCALAR,
                       rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_BOUNDARY)
    val, rhs, st = Asm.assemble(D=None, F=torch.ones(1, dtype=torch.float64, device="cuda"))
    assert st.code == 0
    rp, ci = Asm.csr()
    x = torch.zeros(Asm.ndof, dtype=torch.float64, device="cuda")
    res, ok = af.pcg_jacobi(Asm.ndof, rp, ci, val, rhs, x, rtol=1e-10, maxit=5000)
    assert ok and res.true_relres <= 1.5e-10
    s = Asm.setup()
    vh, bh = val.cpu().numpy(), rhs.cpu().numpy()
    xo = np.zeros(Asm.ndof)
    rr = ctypes.c_double()
    it = T.oracle().orc_jacobi_pcg(ctypes.c_long(Asm.ndof), T.P(s["row_ptr"]), T.P(s["col_idx"]), T.P(vh), T.P(bh),
                                   T.P(xo), ctypes.c_double(1e-10), ctypes.c_long(5000), ctypes.byref(rr))
    assert it > 0
    assert abs(res.iterations - it) <= 2, (res.iterations, it)
    xg = x.cpu().numpy()
    assert np.linalg.norm(xg - xo) <= 1e-8 * np.linalg.norm(xo)
    # the discrete solution of -lap u = 1 on the unit cube peaks near 0.0562
    assert 0.05 < xg.max() < 0.06


def test_pcg_large_system_three_launch_iteration():
    """P2 Poisson on the 32^3 cube (274,625 DOFs, 7.7 M stored entries): large
    enough for the iteration whose SpMV gathers one vector per entry.  Against
    the oracle's CG: iterations within 2, solution within 1e-8, and the FP64
    true residual at the tolerance."""
    import torch
    coords, tets = af.kuhn_cube(32, 0.2, 2)
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P2), order=2, rhs_tensor_type=af.TENSOR_SCALAR,
                       rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_BOUNDARY)
    assert Asm.nnz >= 1 << 22
    val, rhs, st = Asm.assemble(F=torch.ones(1, dtype=torch.float64, device="cuda"))
    assert st.code == 0
    rp, ci = Asm.csr()
    x = torch.zeros(Asm.ndof, dtype=torch.float64, device="cuda")
    res, ok = af.pcg_jacobi(Asm.ndof, rp, ci, val, rhs, x, rtol=1e-10, maxit=20000)
    assert ok and res.true_relres <= 1.5e-10
    s = Asm.setup()
    vh, bh = val.cpu().numpy(), rhs.cpu().numpy()
    xo = np.zeros(Asm.ndof)
    rr = ctypes.c_double()
    it = T.oracle().orc_jacobi_pcg(ctypes.c_long(Asm.ndof), T.P(s["row_ptr"]), T.P(s["col_idx"]), T.P(vh), T.P(bh),
                                   T.P(xo), ctypes.c_double(1e-10), ctypes.c_long(20000), ctypes.byref(rr))
    assert it > 0 and abs(res.iterations - it) <= 2, (res.iterations, it)
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-8 * np.linalg.norm(xo)


def test_pcg_zero_rhs_and_maxit():
    import torch
    coords, tets = af.kuhn_cube(4, 0.2, 2)
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P1), dirichlet=af.DIRICHLET_BOUNDARY)
    val, _, _ = Asm.assemble(D=None)
    rp, ci = Asm.csr()
    b = torch.zeros(Asm.ndof, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    res, ok = af.pcg_jacobi(Asm.ndof, rp, ci, val, b, x, rtol=1e-8)
    assert ok and res.iterations == 0
    b = torch.ones_like(b)
    res, ok = af.pcg_jacobi(Asm.ndof, rp, ci, val, b, x, rtol=1e-14, maxit=3)
    assert not ok and res.iterations == 3


def test_pcg_matrix_with_diagonal_m_is_jacobi():
    """M = diag(a_ii)^-1 as an explicit matrix: the same method as Jacobi-PCG."""
    import torch
    coords, tets = af.kuhn_cube(10, 0.2, 3)
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P2), order=2, tensor_type=af.TENSOR_NULL,
                       tensor_storage=af.COEF_CONST, rhs_tensor_type=af.TENSOR_SCALAR,
                       rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_BOUNDARY)
    val, rhs, st = Asm.assemble(D=None, F=torch.ones(1, dtype=torch.float64, device="cuda"))
    assert st.code == 0
    rp, ci = Asm.csr()
    s = Asm.setup()
    row = np.repeat(np.arange(Asm.ndof), np.diff(s["row_ptr"]))
    vh = val.cpu().numpy()
    diag = np.zeros(Asm.ndof)
    isd = row == s["col_idx"]
    diag[row[isd]] = vh[isd]
    mrp = torch.arange(Asm.ndof + 1, dtype=torch.int64, device="cuda")
    mci = torch.arange(Asm.ndof, dtype=torch.int32, device="cuda")
    mv = torch.from_numpy(1.0 / diag).cuda()
    x = torch.zeros(Asm.ndof, dtype=torch.float64, device="cuda")
    rj, okj = af.pcg_jacobi(Asm.ndof, rp, ci, val, rhs, x, rtol=1e-10, maxit=5000)
    xj = x.clone()
    x.zero_()
    rm, okm = af.pcg_matrix(Asm.ndof, rp, ci, val, mrp, mci, mv, rhs, x, rtol=1e-10, maxit=5000)
    assert okj and okm
    assert abs(rm.iterations - rj.iterations) <= 1, (rm.iterations, rj.iterations)
    assert rm.true_relres <= 1.01e-10 and rj.true_relres <= 1.01e-10
    assert torch.max(torch.abs(x - xj)).item() <= 1e-8 * torch.max(torch.abs(xj)).item()
    # maxit bounds and the zero right-hand side
    x.zero_()
    rm, okm = af.pcg_matrix(Asm.ndof, rp, ci, val, mrp, mci, mv, rhs, x, rtol=1e-14, maxit=5)
    assert not okm and rm.iterations == 5
    x.zero_()
    rm, okm = af.pcg_matrix(Asm.ndof, rp, ci, val, mrp, mci, mv, torch.zeros_like(rhs), x, rtol=1e-8)
    assert okm and rm.iterations == 0


def test_pcg_iteration_counts_across_poll_chunks():
    """Convergence inside any iteration of a poll chunk reports that iteration
    (maxit just below / at the converged count)."""
    import torch
    coords, tets = af.kuhn_cube(8, 0.2, 4)
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P1), dirichlet=af.DIRICHLET_BOUNDARY,
                       rhs_tensor_type=af.TENSOR_SCALAR, rhs_tensor_storage=af.COEF_CONST)
    val, rhs, _ = Asm.assemble(D=None, F=torch.ones(1, dtype=torch.float64, device="cuda"))
    rp, ci = Asm.csr()
    x = torch.zeros(Asm.ndof, dtype=torch.float64, device="cuda")
    r0, ok0 = af.pcg_jacobi(Asm.ndof, rp, ci, val, rhs, x, rtol=1e-9)
    assert ok0
    k = r0.iterations
    x.zero_()
    r1, ok1 = af.pcg_jacobi(Asm.ndof, rp, ci, val, rhs, x, rtol=1e-9, maxit=k)
    assert ok1 and r1.iterations == k
    x.zero_()
    r2, ok2 = af.pcg_jacobi(Asm.ndof, rp, ci, val, rhs, x, rtol=1e-9, maxit=k - 1)
    assert not ok2 and r2.iterations == k - 1


def test_pcg_workspace_reuse_and_release():
    """Repeated solves reuse the cached workspace (bitwise-equal results), a
    larger system grows it, and after anifem_pcg_release the next solve
    allocates afresh."""
    import torch
    out = []
    for n in (6, 9, 6):
        coords, tets = af.kuhn_cube(n, 0.2, 5)
        Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P1), dirichlet=af.DIRICHLET_BOUNDARY,
                           rhs_tensor_type=af.TENSOR_SCALAR, rhs_tensor_storage=af.COEF_CONST)
        val, rhs, _ = Asm.assemble(D=None, F=torch.ones(1, dtype=torch.float64, device="cuda"))
        rp, ci = Asm.csr()
        x = torch.zeros(Asm.ndof, dtype=torch.float64, device="cuda")
        r, ok = af.pcg_jacobi(Asm.ndof, rp, ci, val, rhs, x, rtol=1e-9)
        assert ok
        out.append(x.cpu().numpy())
    np.testing.assert_array_equal(out[0], out[2])
    assert af.lib().anifem_pcg_release() == af.OK
    x.zero_()
    r, ok = af.pcg_jacobi(Asm.ndof, rp, ci, val, rhs, x, rtol=1e-9)
    assert ok
    np.testing.assert_array_equal(x.cpu().numpy(), out[2])
