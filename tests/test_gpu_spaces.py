"""The other FemFix spaces of the reference (spaces/*.h) on the GPU: RT0 and
ND0 (vector-valued; Piola-mapped reference fields with per-cell factors per
basis function), CR1, B2 and B4 (polynomial).  Volume integrals are checked
against the fixtures of the compiled reference (test_gpu_element.py, elem_sp_*)
and the oracle for every storage; here: face integrals against closed forms and
the oracle, evaluation, and the cases the entry points refuse.
FP64: <= 1e-12 relative; FP32: <= 5e-6."""
import ctypes

import numpy as np
import pytest

import anifem_amd as af
import anifem_testlib as T

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def face_areas(xyz):
    Pv = xyz.reshape(-1, 4, 3)
    out = []
    for k in range(4):
        a, b, c = Pv[:, k], Pv[:, (k + 1) % 4], Pv[:, (k + 2) % 4]
        out.append(np.linalg.norm(np.cross(b - a, c - a), axis=1) / 2)
    return np.stack(out, 1)


def run_face(opA, opB, tt, storage, D, order, face, xyz, dn, prec=af.PREC_FP64):
    import torch
    out = af.fem3dface_batch(opA, opB, tt, storage, D, order, face, dev(T.soa_xyz(xyz)), dot_normal=dn,
                             precision=prec)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(len(xyz), -1).astype(np.float64)


@pytest.mark.parametrize("prec,tol", [(af.PREC_FP64, 1e-12), (af.PREC_FP32, 5e-6)])
def test_rt0_fluxes_and_cr1_face_means(prec, tol):
    """phi_i . n = delta_ij on face j for RT0 (fem3DfaceN of IDEN RT0 against
    P0: delta_ij |f_j|); CR1 phi_i has mean delta_ij on face j."""
    rng = np.random.default_rng(0)
    n = 500
    xyz = rng.normal(size=(n, 12))
    area = face_areas(xyz)
    p0 = af.Op(af.IDEN, af.FEM_P0)
    for face in range(4):
        ref = np.zeros((n, 4))
        ref[:, face] = area[:, face]
        F = run_face(af.Op(af.IDEN, af.FEM_RT0), p0, af.TENSOR_NULL, af.COEF_CONST, None, 2, face, xyz, True, prec)
        assert np.abs(F - ref).max() <= tol * area.max(), face
        C = run_face(af.Op(af.IDEN, af.FEM_CR1), p0, af.TENSOR_NULL, af.COEF_CONST, None, 2, face, xyz, False, prec)
        assert np.abs(C - ref).max() <= tol * area.max(), face


@pytest.mark.parametrize("fem", [af.FEM_RT0, af.FEM_ND0])
def test_vector_space_face_mass_vs_oracle(fem):
    """Face mass matrices of RT0 / ND0 with a per-tet GENERAL 3 x 3 tensor
    against the oracle's physical-space restatement, all four faces."""
    rng = np.random.default_rng(fem)
    n = 64
    xyz = rng.normal(size=(n, 12))
    D = rng.normal(size=(n, 9))
    op = af.Op(af.IDEN, fem)
    for face in range(4):
        A = run_face(op, op, af.TENSOR_GENERAL, af.COEF_PER_TET, dev(D.T.copy()), 3, face, xyz, False)
        Ao = T.orc_face(af.IDEN, fem, 1, af.IDEN, fem, 1, af.TENSOR_GENERAL, 1, face, 0, xyz, D)
        assert T.rel_frob(A, Ao) <= 1e-12, face


@pytest.mark.parametrize("op,fem", [(af.IDEN, af.FEM_CR1), (af.GRAD, af.FEM_CR1), (af.IDEN, af.FEM_B2),
                                    (af.GRAD, af.FEM_B2), (af.IDEN, af.FEM_B4), (af.GRAD, af.FEM_B4)])
def test_apply_polynomial_spaces_vs_oracle(op, fem):
    """fem3DapplyL of CR1, B2 and B4 against the oracle."""
    rng = np.random.default_rng(op * 7 + fem)
    n, q = 50, 5
    xyz = rng.normal(size=(n, 12))
    xyl = rng.dirichlet(np.ones(4), q)
    o = af.Op(op, fem)
    dofs = rng.normal(size=(n, o.nfa))
    out = af.fem3dapply_batch(o, xyl, dev(T.soa_xyz(xyz)), dev(dofs)).cpu().numpy()
    ref = T.orc_apply(op, fem, 1, xyl, xyz, dofs)
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("fem", [af.FEM_RT0, af.FEM_ND0])
def test_apply_grad_of_vector_spaces_vs_oracle(fem):
    """fem3DapplyL of GRAD RT0 / GRAD ND0 (9 components) against the oracle,
    whose gradients are pinned by finite differences of the fields."""
    rng = np.random.default_rng(fem)
    n, q = 50, 3
    xyz = rng.normal(size=(n, 12))
    xyl = rng.dirichlet(np.ones(4), q)
    o = af.Op(af.GRAD, fem)
    dofs = rng.normal(size=(n, o.nfa))
    out = af.fem3dapply_batch(o, xyl, dev(T.soa_xyz(xyz)), dev(dofs)).cpu().numpy()
    ref = T.orc_apply(af.GRAD, fem, 1, xyl, xyz, dofs)
    assert out.shape == (n, q, 9)
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()


def test_derivatives_of_p0_are_zero():
    """GRAD and DUDX..DUDZ of P0 (poly_0.h:11-59): the constant's derivatives
    vanish, so every element matrix with them is exactly zero, as is its
    evaluation; IDEN P0 is unchanged (its mass is |T|)."""
    rng = np.random.default_rng(4)
    n = 40
    xyz = rng.normal(size=(n, 12))
    X = dev(T.soa_xyz(xyz))
    D = dev(rng.normal(size=(9, n)))
    g0, g1 = af.Op(af.GRAD, af.FEM_P0), af.Op(af.GRAD, af.FEM_P1)
    for prec in (af.PREC_FP64, af.PREC_FP32):
        A = af.fem3dtet_batch(g0, g1, af.TENSOR_GENERAL, af.COEF_PER_TET, D, 2, X, prec).cpu().numpy()
        assert A.shape == (n, 1, 4) and not A.any()
        A = af.fem3dtet_batch(g0, g0, af.TENSOR_NULL, af.COEF_CONST, None, 2, X, prec).cpu().numpy()
        assert A.shape == (n, 1, 1) and not A.any()
    for op in (af.DUDX, af.DUDY, af.DUDZ):
        A = af.fem3dtet_batch(af.Op(op, af.FEM_P0), af.Op(op, af.FEM_P2), af.TENSOR_NULL, af.COEF_CONST, None, 2,
                              X).cpu().numpy()
        assert A.shape == (n, 1, 10) and not A.any()
    xyl = rng.dirichlet(np.ones(4), 3)
    out = af.fem3dapply_batch(g0, xyl, X, dev(rng.normal(size=(n, 1)))).cpu().numpy()
    assert out.shape == (n, 3, 3) and not out.any()
    i0 = af.Op(af.IDEN, af.FEM_P0)
    M = af.fem3dtet_batch(i0, i0, af.TENSOR_NULL, af.COEF_CONST, None, 2, X).cpu().numpy().reshape(n)
    vol = np.abs(np.linalg.det((xyz.reshape(n, 4, 3)[:, 1:] - xyz.reshape(n, 4, 3)[:, :1]))) / 6
    assert np.abs(M - vol).max() <= 1e-14 * vol.max()


def test_refused_cases():
    """GRAD of RT0 / ND0 paired with another operator or in the normal form,
    vector copies of RT0 / ND0 (elements and global assembly) and the MFMA mode
    return ANIFEM_ERR_ARG."""
    import torch
    L = af.lib()
    xyz = dev(T.soa_xyz(np.random.default_rng(1).normal(size=(4, 12))))
    out = torch.empty((4, 36, 36), dtype=torch.float64, device="cuda")
    t = af.anifem_tensor(af.TENSOR_NULL, af.COEF_CONST, None)
    call = lambda a, b, prec: L.anifem_fem3dtet_batch(a.c(), b.c(), ctypes.byref(t), 2, 4, af._ptr(xyz), prec,
                                                      af._ptr(out), None)
    rt, nd = af.Op(af.IDEN, af.FEM_RT0), af.Op(af.IDEN, af.FEM_ND0)
    assert call(rt, rt, af.PREC_FP64) == 0
    assert call(nd, nd, af.PREC_FP32) == 0
    assert call(rt, rt, af.PREC_FP16X3) == af.ERR_ARG
    assert call(af.Op(af.GRAD, af.FEM_RT0), af.Op(af.GRAD, af.FEM_RT0), af.PREC_FP64) == 0
    assert call(af.Op(af.GRAD, af.FEM_RT0), af.Op(af.GRAD, af.FEM_RT0), af.PREC_FP16X3) == af.ERR_ARG
    # 9 components against a 9-component gradient of a vector P1 space: a GENERAL tensor fits the dims, the
    # pairing is not built
    eye9 = dev(np.eye(9).ravel())
    tg = af.anifem_tensor(af.TENSOR_GENERAL, af.COEF_CONST, af._ptr(eye9))
    assert L.anifem_fem3dtet_batch(af.Op(af.GRAD, af.FEM_RT0).c(), af.Op(af.GRAD, af.FEM_P1, 3).c(), ctypes.byref(tg),
                                   2, 4, af._ptr(xyz), af.PREC_FP64, af._ptr(out), None) == af.ERR_ARG
    assert call(af.Op(af.IDEN, af.FEM_ND0, 3), af.Op(af.IDEN, af.FEM_ND0, 3), af.PREC_FP64) == af.ERR_ARG
    assert af.Op(af.GRAD, af.FEM_ND0).nfa == 6 and af.Op(af.GRAD, af.FEM_ND0).dim == 9
    coords, tets = af.kuhn_cube(2, 0.1, 1)
    for fem in (af.FEM_CR1, af.FEM_B2, af.FEM_RT0):
        with pytest.raises(af.AnifemError):
            af.Assembler(coords, tets, af.Op(af.IDEN, fem, 3))
