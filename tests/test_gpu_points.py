"""Physical quadrature points on the device (internalFem3DtetGeomInit_XYG,
core.inl:250-270): anifem_quadrature_points_batch and
anifem_assembly_quadrature_points.  An x-dependent coefficient evaluated on
the GPU at those points reproduces the compiled reference's element matrices
(its functor evaluated at its own XYG) and the oracle's assembly, with no
host round-trip of the points."""
import numpy as np
import pytest

import anifem_amd as af
import anifem_testlib as T

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def xdep_field_dev(X):
    """m(x) = 1 + x0/2 - x1/4 + x2^2/8 on a (3, q, n) device array: (q, n)."""
    return 1.0 + 0.5 * X[0] - 0.25 * X[1] + 0.125 * X[2] ** 2


@pytest.mark.parametrize("case", [c[0] for c in T.golden_cases() if c[0].startswith("pp_")])
def test_device_points_reproduce_reference_xdep_fixtures(case):
    z = dict(T.golden_cases())[case]
    pair, tt, pp, order, oa, fa, ca, ob, fb, cb = [int(v) for v in z["meta"]]
    assert pp == 2
    xyz, D, Aref = z["xyz"], z["D"], z["A"]
    n = len(xyz)
    opA, opB = af.Op(oa, fa, ca), af.Op(ob, fb, cb)
    xs = dev(T.soa_xyz(xyz))
    X = af.quadrature_points_batch(order, xs)
    q = X.shape[1]
    Xh = X.cpu().numpy()
    np.testing.assert_allclose(Xh.transpose(2, 1, 0), T.product_physical_points(xyz, order), rtol=0, atol=1e-14)
    Dp = T.pack_sym(D, opB.dim) if tt == af.TENSOR_SYMMETRIC else D
    Dv = dev(np.ascontiguousarray(np.asarray(Dp).T))  # (v, n)
    coef = (Dv[:, None, :] * xdep_field_dev(X)[None]).contiguous()  # (v, q, n): data[(v*q + k)*n + r]
    A = af.fem3dtet_batch(opA, opB, tt, af.COEF_PER_POINT, coef, order, xs).cpu().numpy().reshape(n, -1)
    assert T.rel_frob(A, Aref) <= 1e-12, case


@pytest.mark.parametrize("order", [3, 10, 20])
def test_device_points_of_large_rules(order):
    """Rules beyond 24 points (conical products, orders 7..20) read their
    barycentric points from device memory; the 8-point order-3 rule from the
    kernel arguments."""
    rng = np.random.default_rng(order)
    n = 300
    xyz = rng.uniform(-1, 1, (n, 12))
    X = af.quadrature_points_batch(order, dev(T.soa_xyz(xyz))).cpu().numpy()
    assert X.shape[1] == len(af.quadrature(order)[1])
    np.testing.assert_allclose(X.transpose(2, 1, 0), T.product_physical_points(xyz, order), rtol=0, atol=1e-14)


def test_context_points_drive_an_x_dependent_assembly():
    """P2 diffusion with kappa(x) = m(x) evaluated on the device at the
    context's points (reordered cells of a shuffled mesh), against the
    oracle's element matrices with m at its own physical points, applyDir
    and the sequential scatter."""
    coords, tets = af.kuhn_cube(4, 0.2, 12)
    coords, tets = T.shuffled_mesh(coords, tets, 13)
    n = len(tets)
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P2), order=2, tensor_type=af.TENSOR_SCALAR,
                       tensor_storage=af.COEF_PER_POINT, dirichlet=af.DIRICHLET_BOUNDARY)
    X = Asm.quadrature_points()
    rt = Asm.tets()
    np.testing.assert_allclose(X.cpu().numpy().transpose(2, 1, 0),
                               T.product_physical_points(coords[rt].reshape(n, 12), 2), rtol=0, atol=1e-14)
    kappa = xdep_field_dev(X).contiguous()  # (q, n): one scalar per point
    val, _, st = Asm.assemble(D=kappa)
    assert st.code == 0
    s = Asm.setup()
    S = T.oracle_assembly(coords, tets, 3, 1, dirichlet_mask=s["dirichlet"])
    xr = coords[S["tets"]].reshape(n, 12)
    Xo = T.physical_points(xr, 2)  # the oracle's rule, in its own point order
    A = T.orc_elem(2, 3, 1, 2, 3, 1, af.TENSOR_SCALAR, 2, 2, xr, T.xdep_field(Xo).reshape(n, -1))
    A, _ = T.apply_dir_elementwise(A, np.zeros((n, 10)), S["dofs"], s["dirichlet"])
    rc, v_or, _ = T.oracle_scatter(n, 10, S["dofs"], S["slot"], A, None, len(S["col_idx"]), S["ndof"])
    assert rc == 0
    rp = S["row_ptr"]
    rowmax = np.maximum.reduceat(np.abs(v_or), rp[:-1])
    assert np.all(np.abs(val.cpu().numpy() - v_or) <= 1e-12 * np.repeat(rowmax, np.diff(rp)))
