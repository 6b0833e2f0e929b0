"""GPU element matrices (anifem_fem3dtet_batch) against the reference fixtures
and the FP64 oracle.  FP64 mode: <= 1e-12 relative Frobenius per element.
FP32 mode (FP64 geometry, FP32 contraction): <= 5e-6 relative Frobenius per
element.  FP16X3 mode (FP16 hi + lo split on MFMA, FP32 accumulation): <= 1e-6
relative Frobenius per element, the bound stated for the fast mode."""
import numpy as np
import pytest

import anifem_amd as af
import anifem_testlib as T

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 5e-6
FP16X3_TOL = 1e-6


def _torch():
    import torch
    return torch


def dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def product_coef(tt, storage, D, n, dimA, dimB, q=None):
    """Oracle-style D (per tet (n, ddim_full) or per point (n, q, ddim_full))
    -> product SoA layout."""
    if tt == af.TENSOR_NULL:
        return None
    if tt == af.TENSOR_SYMMETRIC:
        if storage == af.COEF_PER_POINT:
            Dp = T.pack_sym(D.reshape(n * q, -1), dimB).reshape(n, q, -1)
        else:
            Dp = T.pack_sym(D, dimB)
    else:
        Dp = D
    if storage == af.COEF_CONST:
        return dev(np.ascontiguousarray(Dp).ravel())
    if storage == af.COEF_PER_TET:
        return dev(np.ascontiguousarray(np.asarray(Dp).reshape(n, -1).T))
    Dp = np.asarray(Dp).reshape(n, q, -1)
    return dev(np.ascontiguousarray(Dp.transpose(2, 1, 0)))  # [v][point][tet]


def run_product(opA, opB, tt, storage, Dprod, order, xyz, prec=af.PREC_FP64):
    out = af.fem3dtet_batch(opA, opB, tt, storage, Dprod, order, dev(T.soa_xyz(xyz)), precision=prec)
    _torch().cuda.synchronize()
    return out.cpu().numpy().reshape(len(xyz), -1).astype(np.float64)


@pytest.mark.parametrize("case", [c[0] for c in T.golden_cases()])
def test_element_vs_reference_fixture(case):
    z = dict(T.golden_cases())[case]
    pair, tt, pp, order, oa, fa, ca, ob, fb, cb = [int(v) for v in z["meta"]]
    xyz, D, Aref = z["xyz"], z["D"], z["A"]
    n = len(xyz)
    opA, opB = af.Op(oa, fa, ca), af.Op(ob, fb, cb)
    dA, dB = opA.dim, opB.dim
    if pp == 2:
        X = T.product_physical_points(xyz, order)  # the product's point order
        q = X.shape[1]
        Dpp = T.xdep_field(X)[:, :, None] * D[:, None, :]
        storage, Dprod = af.COEF_PER_POINT, product_coef(tt, af.COEF_PER_POINT, Dpp, n, dA, dB, q)
    else:
        storage, Dprod = af.COEF_PER_TET, product_coef(tt, af.COEF_PER_TET, D, n, dA, dB)
    precs = [(af.PREC_FP64, FP64_TOL), (af.PREC_FP32, FP32_TOL)]
    if fa in (af.FEM_RT0, af.FEM_ND0) or fb in (af.FEM_RT0, af.FEM_ND0):
        # the MFMA mode takes the spaces whose values are a fixed map of reference values
        with pytest.raises(af.AnifemError):
            run_product(opA, opB, tt, storage, Dprod, order, xyz, af.PREC_FP16X3)
    elif storage == af.COEF_PER_TET:
        precs.append((af.PREC_FP16X3, FP16X3_TOL))
    else:  # the MFMA mode takes tensors constant over each tet only
        with pytest.raises(af.AnifemError):
            run_product(opA, opB, tt, storage, Dprod, order, xyz, af.PREC_FP16X3)
    for prec, tol in precs:
        A = run_product(opA, opB, tt, storage, Dprod, order, xyz, prec)
        err = T.rel_frob(A, Aref)
        print(f"{case} prec={prec} max rel Frobenius error {err:.3e}", flush=True)
        assert err <= tol, (case, prec, err)
    if "lame" in z:
        Dl = dev(np.ascontiguousarray(z["lame"].T))
        A = run_product(opA, opB, af.TENSOR_LAME, af.COEF_PER_TET, Dl, order, xyz)
        assert T.rel_frob(A, Aref) <= FP64_TOL
        A = run_product(opA, opB, af.TENSOR_LAME, af.COEF_PER_TET, Dl, order, xyz, af.PREC_FP16X3)
        assert T.rel_frob(A, Aref) <= FP16X3_TOL


PAIRS = [(2, 2, 1, 2, 2, 1), (2, 3, 1, 2, 3, 1), (2, 4, 1, 2, 4, 1), (2, 3, 3, 2, 3, 3), (1, 1, 1, 1, 3, 1),
         (1, 1, 1, 1, 4, 3), (1, 3, 1, 1, 3, 1), (2, 4, 1, 2, 2, 3), (2, 3, 1, 1, 3, 1), (1, 2, 1, 2, 2, 1),
         (1, 1, 1, 1, 2, 3),
         # RT0 values and divergences, ND0 values and curls, CR1, B2, B4 (spaces/*.h)
         (1, 21, 1, 1, 21, 1), (3, 21, 1, 1, 1, 1), (1, 25, 1, 1, 25, 1), (4, 25, 1, 1, 21, 1),
         (2, 31, 1, 2, 31, 1), (2, 1002, 1, 1, 1002, 1), (1, 1004, 1, 1, 2, 1)]


@pytest.mark.parametrize("pair", PAIRS)
@pytest.mark.parametrize("tt", [af.TENSOR_NULL, af.TENSOR_SCALAR, af.TENSOR_SYMMETRIC, af.TENSOR_GENERAL])
def test_element_vs_oracle_all_storages(pair, tt):
    oa, fa, ca, ob, fb, cb = pair
    opA, opB = af.Op(oa, fa, ca), af.Op(ob, fb, cb)
    dA, dB = opA.dim, opB.dim
    p0b = dA != dB and opA.nfa == 1 and dA == 1
    if tt in (af.TENSOR_NULL, af.TENSOR_SCALAR) and dA != dB and not p0b:
        pytest.skip("identity tensor needs equal dims")
    if tt == af.TENSOR_SYMMETRIC and dA != dB:
        pytest.skip("symmetric tensor needs a square D")
    rng = np.random.default_rng(hash(pair) % 1000 + tt)
    n = 300
    xyz = np.stack([rng.uniform(-1, 1, 12) for _ in range(n)])
    ddim = {af.TENSOR_NULL: 0, af.TENSOR_SCALAR: 1, af.TENSOR_SYMMETRIC: dA * dB, af.TENSOR_GENERAL: dA * dB}[tt]
    for order in [1, 2, 4, 6]:
        q = len(T.orc_quadrature(order)[1])
        for storage in [af.COEF_CONST, af.COEF_PER_TET, af.COEF_PER_POINT]:
            if tt == af.TENSOR_NULL:
                Do = None
                Dprod = None
            else:
                shape = {af.COEF_CONST: (1, ddim), af.COEF_PER_TET: (n, ddim), af.COEF_PER_POINT: (n, q, ddim)}[storage]
                Do = rng.uniform(0.5, 2.0, shape)
                if tt == af.TENSOR_SYMMETRIC:
                    M = Do.reshape(-1, dB, dB)
                    Do = (M + M.transpose(0, 2, 1)).reshape(shape)
                if storage == af.COEF_PER_POINT:  # value of product point j = value of the matching oracle point
                    Dprod = product_coef(tt, storage, Do[:, T.point_permutation(order), :], n, dA, dB, q)
                else:
                    Dprod = product_coef(tt, storage, Do.reshape(shape[0], -1), shape[0], dA, dB, q)
            A_or = T.orc_elem(oa, fa, ca, ob, fb, cb, tt, storage, order, xyz,
                              None if Do is None else Do.reshape(1 if storage == af.COEF_CONST else n, -1))
            A = run_product(opA, opB, tt, storage, Dprod, order, xyz)
            err = T.rel_frob(A, A_or)
            assert err <= FP64_TOL, (pair, tt, order, storage, err)


def test_element_sliver_and_large_batch():
    """Near-slivers (altitude 1e-4) stay within the FP64 bound; a batch larger
    than one launch tile with a ragged tail."""
    rng = np.random.default_rng(1)
    n = 70001
    base = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0.3, 0.3, 1e-4], dtype=float)
    xyz = base[None, :] + rng.uniform(-1e-6, 1e-6, (n, 12)) + rng.uniform(0, 10, (n, 1))
    D = np.array([aniso for aniso in rng.uniform(0.1, 10, (n, 1))])
    opA = af.Op(af.GRAD, af.FEM_P2)
    A = run_product(opA, opA, af.TENSOR_SCALAR, af.COEF_PER_TET, dev(D.T.copy()), 2, xyz)
    sel = rng.choice(n, 500, replace=False)
    sel[-1] = n - 1
    A_or = T.orc_elem(2, 3, 1, 2, 3, 1, af.TENSOR_SCALAR, 1, 2, xyz[sel], D[sel])
    assert T.rel_frob(A[sel], A_or) <= FP64_TOL


def test_fp16x3_mfma_high_contrast_and_slivers():
    """The split-precision MFMA mode against the FP64 mode on the cases that
    stress it: P2^3 elasticity with E log-uniform in [1, 1e3] and nu up to
    0.45 on jittered Kuhn tets, P2 / P3 anisotropic diffusion (eigenvalues
    1e-2 .. 1e2) on near-slivers, and the IDEN / mixed operator pairs."""
    rng = np.random.default_rng(11)
    coords, tets = af.kuhn_cube(6, 0.2, 5)
    xyz = coords[tets].reshape(len(tets), 12)
    n = len(xyz)
    E = 10 ** rng.uniform(0, 3, n)
    nu = rng.uniform(0.2, 0.45, n)
    lame = np.stack([E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))])
    v2 = af.Op(af.GRAD, af.FEM_P2, 3)
    A64 = run_product(v2, v2, af.TENSOR_LAME, af.COEF_PER_TET, dev(lame), 2, xyz)
    A16 = run_product(v2, v2, af.TENSOR_LAME, af.COEF_PER_TET, dev(lame), 2, xyz, af.PREC_FP16X3)
    err = T.rel_frob(A16, A64)
    print(f"P2^3 lame fp16x3 max rel Frobenius error {err:.3e}", flush=True)
    assert err <= FP16X3_TOL
    # near-slivers with anisotropic SPD tensors
    m = 4000
    base = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0.3, 0.3, 1e-3], dtype=float)
    xs = base[None, :] + rng.uniform(-0.05, 0.05, (m, 12)) * np.array([1.0, 1.0, 0.0] * 4)  # altitude 1e-3
    Q = np.linalg.qr(rng.normal(size=(m, 3, 3)))[0]
    lam = 10 ** rng.uniform(-2, 2, (m, 3))
    Dm = np.einsum("rij,rj,rkj->rik", Q, lam, Q)
    Dp = dev(np.ascontiguousarray(T.pack_sym(Dm.reshape(m, 9), 3).T))
    for fem in (af.FEM_P1, af.FEM_P2, af.FEM_P3):
        op = af.Op(af.GRAD, fem)
        order = 4 if fem == af.FEM_P3 else 2
        A64 = run_product(op, op, af.TENSOR_SYMMETRIC, af.COEF_PER_TET, Dp, order, xs)
        A16 = run_product(op, op, af.TENSOR_SYMMETRIC, af.COEF_PER_TET, Dp, order, xs, af.PREC_FP16X3)
        err = T.rel_frob(A16, A64)
        print(f"fem {fem} sliver fp16x3 max rel Frobenius error {err:.3e}", flush=True)
        assert err <= FP16X3_TOL, (fem, err)
    # IDEN pairs: mass matrix and the load vector shape (IDEN P0 against IDEN P2)
    kap = dev(rng.uniform(0.1, 10, (1, n)))
    for opA, opB in [(af.Op(af.IDEN, af.FEM_P2), af.Op(af.IDEN, af.FEM_P2)),
                     (af.Op(af.IDEN, af.FEM_P0), af.Op(af.IDEN, af.FEM_P2)),
                     (af.Op(af.GRAD, af.FEM_P2), af.Op(af.IDEN, af.FEM_P1, 3))]:
        A64 = run_product(opA, opB, af.TENSOR_SCALAR, af.COEF_PER_TET, kap, 2, xyz)
        A16 = run_product(opA, opB, af.TENSOR_SCALAR, af.COEF_PER_TET, kap, 2, xyz, af.PREC_FP16X3)
        err = T.rel_frob(A16, A64)
        assert err <= FP16X3_TOL, (opA, opB, err)
    # a ragged tail: tets beyond the last full 64-tet workgroup are written, nothing else
    torch = _torch()
    op = af.Op(af.GRAD, af.FEM_P2)
    k = 130
    out = torch.full((k + 5, 10, 10), 7.0, dtype=torch.float32, device="cuda")
    af.fem3dtet_batch(op, op, af.TENSOR_NULL, af.COEF_CONST, None, 2, dev(T.soa_xyz(xyz[:k])), af.PREC_FP16X3,
                      out=out[:k])
    torch.cuda.synchronize()
    assert torch.all(out[k:] == 7.0)
    ref = run_product(op, op, af.TENSOR_NULL, af.COEF_CONST, None, 2, xyz[:k])
    assert T.rel_frob(out[:k].cpu().numpy().reshape(k, -1).astype(np.float64), ref) <= FP16X3_TOL


def test_empty_batch_and_errors():
    opA = af.Op(af.GRAD, af.FEM_P1)
    torch = _torch()
    xyz = torch.zeros((12, 0), dtype=torch.float64, device="cuda")
    out = af.fem3dtet_batch(opA, opA, af.TENSOR_NULL, af.COEF_CONST, None, 2, xyz)
    assert out.shape[0] == 0
    with pytest.raises(af.AnifemError) as e:
        af.fem3dtet_batch(opA, opA, af.TENSOR_NULL, af.COEF_CONST, None, 21, dev(np.zeros((12, 4))))
    assert e.value.code == af.ERR_ORDER  # rules stop at order 20 (quadrature_formulas.cpp:1497-1498)
