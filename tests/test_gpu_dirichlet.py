"""Inhomogeneous Dirichlet values through anifem_assembly_run (dir_values):
the reference's element-level applyDir with bc != 0 (dc_on_dof.h:38-45,
applied per cell as ex1.cpp:100-105: F -= A(:, k) bc, F(k) = bc, row and column
k zeroed, diagonal 1), then the sequential cell-order scatter.  The GPU values
and rhs agree with the oracle to 1e-12 (relative to the row maximum / the
largest rhs entry); a manufactured P2 solution is reproduced exactly."""
import numpy as np
import pytest

import anifem_amd as af
import anifem_testlib as T
from test_gpu_assembly import compare, dev

pytestmark = pytest.mark.gpu


def lame_tensor(lam, mu):
    """C_ijkl = lam d_ij d_kl + mu (d_ik d_jl + d_il d_jk) as the 9 x 9
    column-major tensor of each tet (lin_elast.cpp:104-115)."""
    d = np.eye(3)
    C = (lam[:, None, None, None, None] * np.einsum("ij,kl->ijkl", d, d)[None] +
         mu[:, None, None, None, None] * (np.einsum("ik,jl->ijkl", d, d) + np.einsum("il,jk->ijkl", d, d))[None])
    # M[i + 3j, k + 3l] = C_ijkl, stored column-major
    M = C.transpose(0, 2, 1, 4, 3).reshape(len(lam), 9, 9)
    return np.ascontiguousarray(M.transpose(0, 2, 1).reshape(len(lam), 81))


def _case(fem, nc, persistent, precision=af.PREC_FP64, seed=0):
    rng = np.random.default_rng(seed)
    coords, tets = af.kuhn_cube(4 if nc == 1 else 3, 0.2, 3 + seed)
    coords, tets = T.shuffled_mesh(coords, tets, 50 + seed)
    n = len(tets)
    order = 4 if fem == af.FEM_P3 else 2
    if nc == 1:
        kappa = 10 ** rng.uniform(-1, 1, (n, 1))
        Asm = af.Assembler(coords, tets, af.Op(af.GRAD, fem), order=order, tensor_type=af.TENSOR_SCALAR,
                           tensor_storage=af.COEF_PER_TET, rhs_order=order, rhs_tensor_type=af.TENSOR_SCALAR,
                           rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_BOUNDARY)
        D, D_or, tt = dev(kappa.T.copy()), kappa, af.TENSOR_SCALAR
        F, F_or, ftt = dev(np.array([1.5])), np.full((1, 1), 1.5), af.TENSOR_SCALAR
    else:
        E = 10 ** rng.uniform(0, 3, n)
        nu = rng.uniform(0.2, 0.45, n)
        lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
        Asm = af.Assembler(coords, tets, af.Op(af.GRAD, fem, 3), order=2, tensor_type=af.TENSOR_LAME,
                           tensor_storage=af.COEF_PER_TET, rhs_tensor_type=af.TENSOR_GENERAL,
                           rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_BOUNDARY)
        D, D_or, tt = dev(np.stack([lam, mu])), lame_tensor(lam, mu), af.TENSOR_SYMMETRIC
        F, F_or, ftt = dev(np.array([0.3, -0.2, -1.0])), np.array([[0.3, -0.2, -1.0]]), af.TENSOR_GENERAL
    Asm.set_persistent(persistent)
    s = Asm.setup()
    bc = np.where(s["dirichlet"] != 0, rng.normal(size=Asm.ndof), 0.0)
    val, rhs, st = Asm.assemble(D=D, F=F, precision=precision, dir_values=dev(bc))
    assert st.code == 0
    S = T.oracle_assembly(coords, tets, fem, nc, dirichlet_mask=s["dirichlet"])
    xyz = coords[S["tets"]].reshape(n, 12)
    A = T.orc_elem(2, fem, nc, 2, fem, nc, tt, 1, order, xyz, D_or)
    Fe = T.orc_elem(1, 1, 1, 1, fem, nc, ftt, 0, order, xyz, F_or)
    A, Fe = T.apply_dir_elementwise(A, Fe, S["dofs"], s["dirichlet"], bc)
    rc, v_or, r_or = T.oracle_scatter(n, S["dofs"].shape[1], S["dofs"], S["slot"], A, Fe, len(S["col_idx"]),
                                      S["ndof"])
    assert rc == 0
    return Asm, val, rhs, S, v_or, r_or, bc, D, F, s


@pytest.mark.parametrize("fem,nc,persistent", [(af.FEM_P1, 1, True), (af.FEM_P2, 1, True), (af.FEM_P2, 1, False),
                                               (af.FEM_P3, 1, True), (af.FEM_P2, 3, True), (af.FEM_P2, 3, False)])
def test_inhomogeneous_dirichlet_matches_oracle(fem, nc, persistent):
    Asm, val, rhs, S, v_or, r_or, bc, D, F, s = _case(fem, nc, persistent)
    compare(val.cpu().numpy(), rhs.cpu().numpy(), S, v_or, r_or)
    # Dirichlet rows: diagonal = incident-tet count, rhs = count * bc (T3)
    r = rhs.cpu().numpy()
    inc = np.bincount(s["dofs"].ravel(), minlength=Asm.ndof)
    d = np.nonzero(s["dirichlet"])[0]
    np.testing.assert_allclose(r[d], inc[d] * bc[d], rtol=1e-14, atol=0)
    # the matrix does not depend on the values; homogeneous values give the
    # homogeneous pass's rhs bit for bit
    v0, r0, _ = Asm.assemble(D=D, F=F)
    assert np.array_equal(v0.cpu().numpy(), val.cpu().numpy())
    vz, rz, _ = Asm.assemble(D=D, F=F, dir_values=dev(np.zeros(Asm.ndof)))
    assert np.array_equal(rz.cpu().numpy(), r0.cpu().numpy())


def test_inhomogeneous_dirichlet_fp32_contraction():
    Asm, val, rhs, S, v_or, r_or, bc, D, F, s = _case(af.FEM_P2, 1, True, precision=af.PREC_FP32, seed=2)
    compare(val.cpu().numpy(), rhs.cpu().numpy(), S, v_or, r_or, tol=2e-5)


def test_inhomogeneous_dirichlet_without_load():
    """F = NULL: the rhs is the lift alone (the reference's local F starts at 0)."""
    import torch
    rng = np.random.default_rng(7)
    coords, tets = af.kuhn_cube(3, 0.2, 1)
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P2), dirichlet=af.DIRICHLET_BOUNDARY)
    s = Asm.setup()
    bc = np.where(s["dirichlet"] != 0, rng.normal(size=Asm.ndof), 0.0)
    val, rhs, st = Asm.assemble(dir_values=dev(bc))
    assert st.code == 0
    S = T.oracle_assembly(coords, tets, 3, 1, dirichlet_mask=s["dirichlet"])
    xyz = coords[S["tets"]].reshape(len(tets), 12)
    A = T.orc_elem(2, 3, 1, 2, 3, 1, af.TENSOR_NULL, 0, 2, xyz, None)
    A, Fe = T.apply_dir_elementwise(A, np.zeros((len(tets), 10)), S["dofs"], s["dirichlet"], bc)
    rc, v_or, r_or = T.oracle_scatter(len(tets), 10, S["dofs"], S["slot"], A, Fe, len(S["col_idx"]), S["ndof"])
    compare(val.cpu().numpy(), rhs.cpu().numpy(), S, v_or, r_or)
    assert isinstance(rhs, torch.Tensor)


def test_manufactured_p2_solution_is_exact():
    """-div(grad u) = f with u = 1 + x - 2y + x^2 + 3yz - z^2 / 2 (f = -1): P2
    reproduces quadratics and the load is integrated exactly, so with u's
    values on the boundary DOFs the discrete solution equals u at every DOF."""
    import torch
    coords, tets = af.kuhn_cube(5, 0.2, 6)
    u = lambda x: 1 + x[..., 0] - 2 * x[..., 1] + x[..., 0] ** 2 + 3 * x[..., 1] * x[..., 2] - 0.5 * x[..., 2] ** 2
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P2), order=2, rhs_order=2,
                       rhs_tensor_type=af.TENSOR_SCALAR, rhs_tensor_storage=af.COEF_CONST,
                       dirichlet=af.DIRICHLET_BOUNDARY)
    s = Asm.setup()
    pos = T.dof_positions(coords, Asm.tets(), s["dofs"], T.lagrange_nodes(af.FEM_P2))
    uex = np.zeros(Asm.ndof)
    uex[s["dofs"].ravel()] = u(pos.reshape(-1, 3))
    val, rhs, st = Asm.assemble(F=dev(np.array([-1.0])), dir_values=dev(uex))
    assert st.code == 0
    rp, ci = Asm.csr()
    x = torch.zeros(Asm.ndof, dtype=torch.float64, device="cuda")
    res, ok = af.pcg_jacobi(Asm.ndof, rp, ci, val, rhs, x, rtol=1e-13)
    assert ok
    assert np.abs(x.cpu().numpy() - uex).max() < 1e-10 * np.abs(uex).max()
