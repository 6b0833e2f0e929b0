"""Mixed Poisson in RT0 x P0 through a two-variable context
(anifem_assembly_create_vars + anifem_assembly_run_elements): the flux
sigma = -grad u in RT0 (variable 0) and u in P0 (variable 1), with
    (sigma, tau) - (u, div tau) = 0,   (div sigma, v) = (f, v)
and u = 0 on the boundary (natural in this form).  The element matrices are
the caller's, formed from anifem_fem3dtet_batch blocks in the cells' local
basis; the scatter applies the RT0 orientation signs (assembler.inl:154-164).
Checked against the oracle's blocks, signs and sequential scatter, and by a
manufactured solution whose L2 errors in u and sigma halve with the mesh
size (a direct sparse solve on the host: the system is indefinite)."""
import numpy as np
import pytest

import anifem_amd as af
import anifem_testlib as T
from test_gpu_assembly import compare, dev

pytestmark = pytest.mark.gpu

RT0, P0 = af.FEM_RT0, af.FEM_P0


def _blocks_gpu(xs, kappa=None, f_tensor=None, f_storage=af.COEF_PER_TET, rhs_order=2):
    """Element matrices (n, 5, 5) in the column-major layout of
    fem3dtet_batch (E[r, ia, ib] = A(ib, ia)) and load vectors (n, 5)."""
    import torch
    n = xs.shape[1]
    sig, div, u = af.Op(af.IDEN, RT0), af.Op(af.DIV, RT0), af.Op(af.IDEN, P0)
    if kappa is None:
        M = af.fem3dtet_batch(sig, sig, af.TENSOR_NULL, af.COEF_CONST, None, 2, xs)
    else:
        M = af.fem3dtet_batch(sig, sig, af.TENSOR_SCALAR, af.COEF_PER_TET, kappa, 2, xs)
    Bt = af.fem3dtet_batch(u, div, af.TENSOR_NULL, af.COEF_CONST, None, 2, xs)   # (n, 1, 4): A(ib, 0) = (1, div tau_ib)
    Bd = af.fem3dtet_batch(div, u, af.TENSOR_NULL, af.COEF_CONST, None, 2, xs)   # (n, 4, 1): A(0, ia) = (div sigma_ia, 1)
    Fv = af.fem3dtet_batch(u, u, af.TENSOR_SCALAR, f_storage, f_tensor, rhs_order, xs)  # (n, 1, 1)
    E = torch.zeros((n, 5, 5), dtype=torch.float64, device=xs.device)
    E[:, :4, :4] = M
    E[:, 4, :4] = -Bt[:, 0, :]
    E[:, :4, 4] = Bd[:, :, 0]
    F = torch.zeros((n, 5), dtype=torch.float64, device=xs.device)
    F[:, 4] = Fv[:, 0, 0]
    return E.contiguous(), F


def _context(coords, tets, dirichlet=af.DIRICHLET_NONE, mask=None):
    return af.Assembler(coords, tets, variables=[af.Op(af.IDEN, RT0), af.Op(af.IDEN, P0)], dirichlet=dirichlet,
                        dirichlet_dof_mask=mask)


@pytest.mark.parametrize("flux_bc", [False, True])
def test_mixed_assembly_matches_oracle(flux_bc):
    """Per-tet kappa in the flux mass and a per-tet load; with flux_bc the
    boundary faces' RT0 DOFs are Dirichlet (sigma . n = 0, applyDir)."""
    coords, tets = af.kuhn_cube(3, 0.2, 8)
    coords, tets = T.shuffled_mesh(coords, tets, 9)
    n = len(tets)
    mask = None
    if flux_bc:
        R = af.Assembler(coords, tets, af.Op(af.IDEN, RT0), host_only=True, dirichlet=af.DIRICHLET_BOUNDARY)
        mask = np.concatenate([R.setup()["dirichlet"], np.zeros(n, np.uint8)])
    Asm = _context(coords, tets, af.DIRICHLET_DOF_MASK if flux_bc else af.DIRICHLET_NONE, mask)
    assert Asm.nf == 5
    rt = Asm.tets()
    xyz = coords[rt].reshape(n, 12)
    rng = np.random.default_rng(3 + flux_bc)
    kappa = 10 ** rng.uniform(-1, 1, (n, 1))
    f = rng.uniform(-2, 2, (n, 1))
    E, F = _blocks_gpu(dev(T.soa_xyz(xyz)), dev(kappa.T.copy()), dev(f.T.copy()))
    val, rhs, st = Asm.assemble_elements(E, F)
    assert st.code == 0
    # oracle: the same blocks (column-major nfB x nfA per tet), composed
    M = T.orc_elem(af.IDEN, RT0, 1, af.IDEN, RT0, 1, af.TENSOR_SCALAR, 1, 2, xyz, kappa).reshape(n, 4, 4)
    Bt = T.orc_elem(af.IDEN, P0, 1, af.DIV, RT0, 1, af.TENSOR_NULL, 0, 2, xyz, None).reshape(n, 1, 4)
    Bd = T.orc_elem(af.DIV, RT0, 1, af.IDEN, P0, 1, af.TENSOR_NULL, 0, 2, xyz, None).reshape(n, 4, 1)
    Fv = T.orc_elem(af.IDEN, P0, 1, af.IDEN, P0, 1, af.TENSOR_SCALAR, 1, 2, xyz, f).reshape(n)
    C = np.zeros((n, 5, 5))  # C[r, ia, ib] = A(ib, ia)
    C[:, :4, :4] = M
    C[:, 4, :4] = -Bt[:, 0, :]
    C[:, :4, 4] = Bd[:, :, 0]
    Fo = np.zeros((n, 5))
    Fo[:, 4] = Fv
    sg = Asm.signs().astype(np.float64)
    assert np.all(sg[:, 4] == 1) and np.any(sg[:, :4] < 0)
    C = C * sg[:, :, None] * sg[:, None, :]
    Fo = Fo * sg
    s = Asm.setup()
    A2, F2 = T.apply_dir_elementwise(C.reshape(n, 25), Fo, s["dofs"], s["dirichlet"])
    rc, v_or, r_or = T.oracle_scatter(n, 5, s["dofs"], s["slot"], A2, F2, Asm.nnz, Asm.ndof)
    assert rc == 0
    compare(val.cpu().numpy(), rhs.cpu().numpy(), dict(row_ptr=s["row_ptr"]), v_or, r_or)


def _mixed_errors(n):
    """L2 errors of u (P0) and sigma (RT0) for u = sin(pi x) sin(pi y) sin(pi z)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    pi = np.pi
    coords, tets = af.kuhn_cube(n, 0.2, 5)
    Asm = _context(coords, tets)
    rt = Asm.tets()
    nt = len(tets)
    xs = dev(T.soa_xyz(coords[rt].reshape(nt, 12)))
    ue = lambda X: np.sin(pi * X[..., 0]) * np.sin(pi * X[..., 1]) * np.sin(pi * X[..., 2])
    se = lambda X: -pi * np.stack([np.cos(pi * X[..., 0]) * np.sin(pi * X[..., 1]) * np.sin(pi * X[..., 2]),
                                   np.sin(pi * X[..., 0]) * np.cos(pi * X[..., 1]) * np.sin(pi * X[..., 2]),
                                   np.sin(pi * X[..., 0]) * np.sin(pi * X[..., 1]) * np.cos(pi * X[..., 2])], axis=-1)
    X = af.quadrature_points_batch(6, xs)  # (3, q, n)
    Xh = X.cpu().numpy().transpose(2, 1, 0)  # (n, q, 3)
    fq = dev(np.ascontiguousarray((3 * pi ** 2 * ue(Xh)).T))  # (q, n): f = -lap u
    E, F = _blocks_gpu(xs, f_tensor=fq, f_storage=af.COEF_PER_POINT, rhs_order=6)
    val, rhs, st = Asm.assemble_elements(E, F)
    assert st.code == 0
    s = Asm.setup()
    A = sp.csr_matrix((val.cpu().numpy(), s["col_idx"], s["row_ptr"]), shape=(Asm.ndof, Asm.ndof))
    x = spla.spsolve(A.tocsc(), rhs.cpu().numpy())
    assert np.linalg.norm(A @ x - rhs.cpu().numpy()) < 1e-9 * np.linalg.norm(rhs.cpu().numpy())
    nfc = Asm.ndof - nt
    p, w = af.quadrature(6)
    P = coords[rt]
    vol = np.abs(np.linalg.det(np.stack([P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]], axis=1))) / 6
    eu = np.sqrt((vol[:, None] * w[None, :] * (x[nfc:][:, None] - ue(Xh)) ** 2).sum())
    # the flux through a one-variable RT0 context: variable 0's numbering is the space's own
    R = af.Assembler(coords, tets, af.Op(af.IDEN, RT0), order=2)
    assert R.ndof == nfc
    import torch
    sh = R.evaluate(torch.from_numpy(np.ascontiguousarray(x[:nfc])).cuda(), p, op=af.IDEN).cpu().numpy()  # (n, q, 3)
    es = np.sqrt((vol[:, None] * w[None, :] * ((sh - se(Xh)) ** 2).sum(axis=2)).sum())
    return eu, es


def test_mixed_poisson_converges_first_order():
    e4, e8 = _mixed_errors(4), _mixed_errors(8)
    ru, rs = e4[0] / e8[0], e4[1] / e8[1]
    print(f"mixed Poisson L2 errors n=4 {e4}, n=8 {e8}; ratios u {ru:.2f}, sigma {rs:.2f}")
    assert 1.7 < ru < 2.4 and 1.7 < rs < 2.4


def test_vars_context_refuses_the_fused_run():
    coords, tets = af.kuhn_cube(2, 0.1, 1)
    Asm = _context(coords, tets)
    with pytest.raises(af.AnifemError):
        Asm.assemble(D=None)
