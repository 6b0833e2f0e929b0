"""Global assembly of RT0, ND0, CR1, B2, B4 and P0 (anifem_assembly_run on
the element-matrix path): against the oracle's element matrices with the
orientation signs of the global basis (assembler.inl:154-164), element-level
Dirichlet and the sequential scatter, and through manufactured solutions of
an H(curl) problem (curl curl E + E = f, ND0) and an H(div) problem
(-grad div s + s = f, RT0) whose L2 errors halve with the mesh size."""
import numpy as np
import pytest

import anifem_amd as af
import anifem_testlib as T
from test_gpu_assembly import compare, dev

pytestmark = pytest.mark.gpu

RT0, ND0, CR1, B2, B4, P0 = af.FEM_RT0, af.FEM_ND0, af.FEM_CR1, af.FEM_B2, af.FEM_B4, af.FEM_P0


@pytest.mark.parametrize("op,fem", [(af.DIV, RT0), (af.IDEN, RT0), (af.CURL, ND0), (af.IDEN, ND0), (af.GRAD, CR1),
                                    (af.GRAD, B2), (af.IDEN, B4), (af.IDEN, P0), (af.GRAD, RT0), (af.GRAD, ND0)])
def test_space_assembly_matches_oracle(op, fem):
    coords, tets = af.kuhn_cube(3, 0.2, 5)
    coords, tets = T.shuffled_mesh(coords, tets, 6)
    n = len(tets)
    rng = np.random.default_rng(op * 10 + fem)
    kappa = 10 ** rng.uniform(-1, 1, (n, 1))
    vec = fem in (RT0, ND0)
    Asm = af.Assembler(coords, tets, af.Op(op, fem), order=2, tensor_type=af.TENSOR_SCALAR,
                       tensor_storage=af.COEF_PER_TET, rhs_order=2,
                       rhs_tensor_type=af.TENSOR_GENERAL if vec else af.TENSOR_SCALAR,
                       rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_BOUNDARY)
    f = np.array([0.5, -1.0, 2.0]) if vec else np.array([1.5])
    val, rhs, st = Asm.assemble(D=dev(kappa.T.copy()), F=dev(f))
    assert st.code == 0
    s = Asm.setup()
    rt, sg = Asm.tets(), Asm.signs().astype(np.float64)
    nf = Asm.nf
    xyz = coords[rt].reshape(n, 12)
    A = T.orc_elem(op, fem, 1, op, fem, 1, af.TENSOR_SCALAR, 1, 2, xyz, kappa)
    F = T.orc_elem(af.IDEN, P0, 1, af.IDEN, fem, 1, af.TENSOR_GENERAL if vec else af.TENSOR_SCALAR, 0, 2, xyz,
                   f[None, :])
    # global basis = sign x local basis
    A = (A.reshape(n, nf, nf) * sg[:, None, :] * sg[:, :, None]).reshape(n, -1)
    F = F * sg
    A, F = T.apply_dir_elementwise(A, F, s["dofs"], s["dirichlet"])
    rc, v_or, r_or = T.oracle_scatter(n, nf, s["dofs"], s["slot"], A, F, Asm.nnz, Asm.ndof)
    assert rc == 0
    compare(val.cpu().numpy(), rhs.cpu().numpy(), dict(row_ptr=s["row_ptr"]), v_or, r_or)


def _l2_error(Asm, coords, x, exact, order=6):
    """||u - u_h||_L2 with the rule of `order`: u_h from anifem_assembly_eval
    (signs included) at the rule's barycentric points of every cell."""
    p, w = af.quadrature(order)
    uh = Asm.evaluate(x, p, op=af.IDEN).cpu().numpy()  # (n, q, 3)
    X = Asm.quadrature_points(order).cpu().numpy().transpose(2, 1, 0)  # (n, q, 3)
    rt = Asm.tets()
    P = coords[rt]
    vol = np.abs(np.linalg.det(np.stack([P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]], axis=1))) / 6
    e2 = ((uh - exact(X)) ** 2).sum(axis=2)
    return float(np.sqrt((vol[:, None] * w[None, :] * e2).sum()))


def _solve(fem, ops, n, exact, rhs_factor):
    """Assemble the sum of the operator pairs, the load of f = rhs_factor x
    exact (per quadrature point), zero boundary trace, Jacobi-PCG; L2 error."""
    import torch
    coords, tets = af.kuhn_cube(n, 0.2, 3)
    total, rhs_t, A0 = None, None, None
    for i, op in enumerate(ops):
        Asm = af.Assembler(coords, tets, af.Op(op, fem), order=2, rhs_order=6,
                           rhs_tensor_type=af.TENSOR_GENERAL, rhs_tensor_storage=af.COEF_PER_POINT,
                           dirichlet=af.DIRICHLET_BOUNDARY)
        F = None
        if i == 0:
            X = Asm.quadrature_points(6)  # (3, q, n)
            fx = rhs_factor * exact(X.cpu().numpy().transpose(2, 1, 0))  # (n, q, 3)
            F = dev(np.ascontiguousarray(fx.transpose(2, 1, 0)))  # (3, q, n): data[(v*q + k)*n + r]
        val, rhs, st = Asm.assemble(F=F)
        assert st.code == 0
        total = val.clone() if total is None else total + val
        if i == 0:
            rhs_t, A0 = rhs, Asm
    rp, ci = A0.csr()
    x = torch.zeros(A0.ndof, dtype=torch.float64, device="cuda")
    res, ok = af.pcg_jacobi(A0.ndof, rp, ci, total, rhs_t, x, rtol=1e-12, maxit=50000)
    assert ok
    return _l2_error(A0, coords, x, exact)


def test_nd0_maxwell_converges_first_order():
    """E = (sin pi y sin pi z, sin pi z sin pi x, sin pi x sin pi y) has zero
    tangential trace on the cube and curl curl E = 2 pi^2 E."""
    pi = np.pi
    E = lambda X: np.stack([np.sin(pi * X[..., 1]) * np.sin(pi * X[..., 2]),
                            np.sin(pi * X[..., 2]) * np.sin(pi * X[..., 0]),
                            np.sin(pi * X[..., 0]) * np.sin(pi * X[..., 1])], axis=-1)
    e = [_solve(ND0, [af.IDEN, af.CURL], n, E, 2 * pi ** 2 + 1) for n in (4, 8)]
    print(f"ND0 L2 errors {e}, ratio {e[0] / e[1]:.2f}")
    assert 1.7 < e[0] / e[1] < 2.4


def test_rt0_hdiv_converges_first_order():
    """s = grad(cos pi x cos pi y cos pi z) has zero normal trace on the cube
    and -grad div s = 3 pi^2 s."""
    pi = np.pi
    c = np.cos
    S = lambda X: -pi * np.stack([np.sin(pi * X[..., 0]) * c(pi * X[..., 1]) * c(pi * X[..., 2]),
                                  c(pi * X[..., 0]) * np.sin(pi * X[..., 1]) * c(pi * X[..., 2]),
                                  c(pi * X[..., 0]) * c(pi * X[..., 1]) * np.sin(pi * X[..., 2])], axis=-1)
    e = [_solve(RT0, [af.IDEN, af.DIV], n, S, 3 * pi ** 2 + 1) for n in (4, 8)]
    print(f"RT0 L2 errors {e}, ratio {e[0] / e[1]:.2f}")
    assert 1.7 < e[0] / e[1] < 2.4
