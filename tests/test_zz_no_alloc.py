"""anifem_assembly_run_noalloc: the assembly pass with its geometry rows in the
library's fixed scratch array instead of a buffer it allocates.  Called
through the C ABI, it gives anifem_assembly_run's values and rhs bit for bit
(same kernels, same order), on two calls in a row, the second on a smaller
mesh that leaves stale rows of the first in the scratch; the device's free
memory does not change across a call on a context that never ran before; a
problem whose geometry rows exceed the scratch is refused."""
import ctypes

import numpy as np
import pytest

import anifem_amd as af

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 128 << 20


def _noalloc(Asm, D, F, val, rhs, precision=af.PREC_FP64, dir_values=None):
    L = af.lib()
    f = L.anifem_assembly_run_noalloc
    f.restype = ctypes.c_int32
    f.argtypes = [ctypes.c_void_p, ctypes.POINTER(af.anifem_tensor), ctypes.POINTER(af.anifem_tensor),
                  ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.POINTER(af.anifem_status), ctypes.c_void_p]
    tD = af.anifem_tensor(Asm.tensor_type, Asm.tensor_storage, af._ptr(D))
    tF = af.anifem_tensor(Asm.rhs_tensor_type, Asm.rhs_tensor_storage, af._ptr(F))
    st = af.anifem_status()
    rc = f(Asm._h, ctypes.byref(tD), ctypes.byref(tF), af._ptr(dir_values), precision, af._ptr(val), af._ptr(rhs),
           ctypes.byref(st), None)
    return rc, st


def _free_bytes():
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _problem(n, seed, general=False):
    """P2 diffusion with a per-tet SPD tensor, stored packed (SYMMETRIC) or in
    full (GENERAL: the geometry rows are then not packed, 80 B per tet)."""
    import torch
    rng = np.random.default_rng(seed)
    coords, tets = af.kuhn_cube(n, 0.2, seed)
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P2), order=2,
                       tensor_type=af.TENSOR_GENERAL if general else af.TENSOR_SYMMETRIC,
                       tensor_storage=af.COEF_PER_TET, rhs_order=2, rhs_tensor_type=af.TENSOR_SCALAR,
                       rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_BOUNDARY)
    m = rng.normal(size=(len(tets), 3, 3))
    Dm = np.einsum("rij,rkj->rik", m, m) + np.eye(3)
    if general:
        Dp = np.stack([Dm[:, i, j] for j in range(3) for i in range(3)])  # column-major
    else:
        Dp = np.stack([Dm[:, i, j] for i, j in [(0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2)]])
    return Asm, torch.from_numpy(np.ascontiguousarray(Dp)).cuda(), torch.full((1,), 1.25, dtype=torch.float64,
                                                                              device="cuda")


def test_noalloc_matches_run_on_two_calls_second_smaller():
    import torch
    for n, seed in [(40, 1), (9, 2)]:
        print(f"noalloc: P2 anisotropic {n}^3", flush=True)
        Asm, D, F = _problem(n, seed)
        val = torch.full((Asm.nnz,), np.nan, dtype=torch.float64, device="cuda")
        rhs = torch.full((Asm.ndof,), np.nan, dtype=torch.float64, device="cuda")
        before = _free_bytes()
        rc, st = _noalloc(Asm, D, F, val, rhs)
        after = _free_bytes()
        assert rc == 0 and st.code == 0
        # the 40^3 pass's geometry rows are 31 MB: an allocation of them would show
        assert before - after < (4 << 20), (before, after)
        v_ref, r_ref, st2 = Asm.assemble(D=D, F=F)
        assert st2.code == 0
        assert torch.equal(val, v_ref) and torch.equal(rhs, r_ref)
        # FP32 contraction and inhomogeneous Dirichlet values through the scratch
        bc = torch.from_numpy(np.where(Asm.setup()["dirichlet"] != 0, 0.5, 0.0)).cuda()
        rc, st = _noalloc(Asm, D, F, val, rhs, precision=af.PREC_FP32, dir_values=bc)
        assert rc == 0 and st.code == 0
        v32, r32, _ = Asm.assemble(D=D, F=F, precision=af.PREC_FP32, dir_values=bc)
        assert torch.equal(val, v32) and torch.equal(rhs, r32)


def test_noalloc_refuses_beyond_the_scratch():
    import torch
    # P2 FP64 geometry rows of a GENERAL tensor take 80 bytes per tet: 66^3
    # Kuhn tets need 138 MB
    n = 66
    assert 6 * n ** 3 * 80 > SCRATCH_BYTES
    Asm, D, F = _problem(n, 3, general=True)
    val = torch.empty(Asm.nnz, dtype=torch.float64, device="cuda")
    rhs = torch.empty(Asm.ndof, dtype=torch.float64, device="cuda")
    rc, _ = _noalloc(Asm, D, F, val, rhs)
    assert rc == af.ERR_DIMS
    # the FP32 rows (48 bytes per tet) fit
    rc, st = _noalloc(Asm, D, F, val, rhs, precision=af.PREC_FP32)
    assert rc == 0 and st.code == 0
