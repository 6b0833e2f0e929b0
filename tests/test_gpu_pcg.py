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
