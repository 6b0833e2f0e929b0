"""GPU whole-mesh assembly (anifem_assembly_run) against the oracle's
sequential cell-order scatter (assembler.inl:397-481) of oracle element
matrices with element-level Dirichlet (dc_on_dof.h:38-45).  FP64 values agree
to 1e-12 relative to the largest entry of the row (the slot sums run in the
same ascending tet order; only the element values differ by rounding);
repeated runs are bitwise identical."""
import numpy as np
import pytest

import anifem_amd as af
import anifem_testlib as T

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def haar(rng):
    Q, R = np.linalg.qr(rng.normal(size=(3, 3)))
    Q = Q @ np.diag(np.sign(np.diag(R)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    return Q


def aniso(rng, n):
    out = np.zeros((n, 9))
    for r in range(n):
        a, b = rng.uniform(-2, 2, 2)
        R = haar(rng)
        out[r] = (R @ np.diag([1, 10 ** a, 10 ** b]) @ R.T).ravel(order="F")
    return out


def oracle_system(coords, tets, fem, nc, tt, D_or, order, f_or, rhs_tt, rhs_order, dirichlet, storage=1):
    S = T.oracle_assembly(coords, tets, fem, nc)
    xyz = coords[S["tets"]].reshape(len(tets), 12)
    A = T.orc_elem(2, fem, nc, 2, fem, nc, tt, storage, order, xyz, D_or)
    F = T.orc_elem(1, 1, 1, 1, fem, nc, rhs_tt, 0, rhs_order, xyz, f_or)
    return S, xyz, A, F


def compare(A_gpu_val, rhs_gpu, S_or, val_or, rhs_or, tol=1e-12):
    rp = S_or["row_ptr"]
    rowmax = np.maximum.reduceat(np.abs(val_or), rp[:-1])
    scale = np.repeat(rowmax, np.diff(rp))
    err = np.max(np.abs(A_gpu_val - val_or) / scale)
    assert err <= tol, err
    if rhs_or is not None:
        rerr = np.max(np.abs(rhs_gpu - rhs_or)) / max(np.abs(rhs_or).max(), 1e-300)
        assert rerr <= tol, rerr
    return err


@pytest.mark.parametrize("fem,dirichlet", [(2, 1), (2, 0), (3, 1), (4, 0), (4, 1)])
def test_scalar_assembly_matches_oracle(fem, dirichlet):
    rng = np.random.default_rng(fem * 10 + dirichlet)
    coords, tets = af.kuhn_cube(4, 0.2, 3)
    n = len(tets)
    kappa = 10 ** rng.uniform(-1, 1, (n, 1))
    order = 4 if fem == 4 else 2
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, fem), order=order, tensor_type=af.TENSOR_SCALAR,
                       tensor_storage=af.COEF_PER_TET, rhs_order=order, rhs_tensor_type=af.TENSOR_SCALAR,
                       rhs_tensor_storage=af.COEF_CONST, dirichlet=dirichlet)
    val, rhs, st = Asm.assemble(D=dev(kappa.T.copy()), F=dev(np.ones(1)))
    assert st.code == 0 and st.n_nonfinite == 0 and st.n_degenerate == 0
    setup = Asm.setup()
    S, xyz, A, F = oracle_system(coords, tets, fem, 1, af.TENSOR_SCALAR, kappa, order, np.ones((1, 1)),
                                 af.TENSOR_SCALAR, order, dirichlet)
    mask = setup["dirichlet"] if dirichlet else None
    S = T.oracle_assembly(coords, tets, fem, 1, dirichlet_mask=mask)
    if dirichlet:
        A, F = T.apply_dir_elementwise(A, F, S["dofs"], setup["dirichlet"])
    rc, v_or, r_or = T.oracle_scatter(n, S["dofs"].shape[1], S["dofs"], S["slot"], A, F, len(S["col_idx"]), S["ndof"])
    assert rc == 0
    compare(val.cpu().numpy(), rhs.cpu().numpy(), S, v_or, r_or)
    # bitwise-identical repeat (no atomics, fixed summation order)
    val2, rhs2, _ = Asm.assemble(D=dev(kappa.T.copy()), F=dev(np.ones(1)))
    assert np.array_equal(val.cpu().numpy(), val2.cpu().numpy())
    assert np.array_equal(rhs.cpu().numpy(), rhs2.cpu().numpy())


def test_p3_node_beyond_the_default_budget_matches_oracle():
    """A P3 node with 200 incident tets: its row holds 4000 entries, beyond the
    default 2048 per block; the context is built with the wider budgets and
    the assembly equals the oracle's (FP64, within 1e-12 of the row maximum)."""
    coords, tets = T.fan_mesh(100)
    n = len(tets)
    kappa = 10 ** np.random.default_rng(2).uniform(-1, 1, (n, 1))
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P3), order=4, tensor_type=af.TENSOR_SCALAR,
                       tensor_storage=af.COEF_PER_TET, rhs_order=4, rhs_tensor_type=af.TENSOR_SCALAR,
                       rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_NONE)
    val, rhs, st = Asm.assemble(D=dev(kappa.T.copy()), F=dev(np.ones(1)))
    assert st.code == 0 and st.n_degenerate == 0
    S, xyz, A, F = oracle_system(coords, tets, 4, 1, af.TENSOR_SCALAR, kappa, 4, np.ones((1, 1)),
                                 af.TENSOR_SCALAR, 4, 0)
    S = T.oracle_assembly(coords, tets, 4, 1)
    rc, v_or, r_or = T.oracle_scatter(n, S["dofs"].shape[1], S["dofs"], S["slot"], A, F, len(S["col_idx"]), S["ndof"])
    assert rc == 0
    compare(val.cpu().numpy(), rhs.cpu().numpy(), S, v_or, r_or)


@pytest.mark.parametrize("tt,dirichlet", [(af.TENSOR_SCALAR, 0), (af.TENSOR_SCALAR, 1), (af.TENSOR_GENERAL, 1)])
def test_p3_two_lanes_per_pair_match_one_lane_and_oracle(tt, dirichlet):
    """P3 row blocks of at most 128 pairs (the default) run with two lanes per
    pair (entries 0-9 and 10-19 of a pair on lanes t and t+128; with a scalar
    coefficient only the triangle of the symmetric reference tensor is staged,
    with a general 3 x 3 one the full tensor): bitwise equal to the one-lane
    pass over blocks of 256 pairs, equal to the oracle within 1e-12 of the row
    maximum, and the same count of non-finite entries."""
    rng = np.random.default_rng(31 + dirichlet + tt)
    coords, tets = af.kuhn_cube(4, 0.2, 13)
    n = len(tets)
    if tt == af.TENSOR_SCALAR:
        kappa = 10 ** rng.uniform(-1, 1, (n, 1))
    else:  # a non-symmetric coefficient: I + 0.3 N, N standard normal
        kappa = (np.eye(3).ravel(order="F")[None, :] + 0.3 * rng.standard_normal((n, 9)))
    kw = dict(order=4, tensor_type=tt, tensor_storage=af.COEF_PER_TET, rhs_order=4,
              rhs_tensor_type=af.TENSOR_SCALAR, rhs_tensor_storage=af.COEF_CONST, dirichlet=dirichlet)
    A1 = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P3), budget=(4096, 3072, 256), **kw)
    A2 = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P3), **kw)
    assert A1.block_layout().max_block_pairs > 128 >= A2.block_layout().max_block_pairs
    v1, r1, s1 = A1.assemble(D=dev(kappa.T.copy()), F=dev(np.ones(1)))
    v2, r2, s2 = A2.assemble(D=dev(kappa.T.copy()), F=dev(np.ones(1)))
    assert s1.code == 0 and s2.code == 0
    assert np.array_equal(v1.cpu().numpy(), v2.cpu().numpy())
    assert np.array_equal(r1.cpu().numpy(), r2.cpu().numpy())
    setup = A2.setup()
    S, xyz, A, F = oracle_system(coords, tets, 4, 1, tt, kappa, 4, np.ones((1, 1)), af.TENSOR_SCALAR, 4, dirichlet)
    S = T.oracle_assembly(coords, tets, 4, 1, dirichlet_mask=setup["dirichlet"] if dirichlet else None)
    if dirichlet:
        A, F = T.apply_dir_elementwise(A, F, S["dofs"], setup["dirichlet"])
    rc, v_or, r_or = T.oracle_scatter(n, S["dofs"].shape[1], S["dofs"], S["slot"], A, F, len(S["col_idx"]), S["ndof"])
    assert rc == 0
    compare(v2.cpu().numpy(), r2.cpu().numpy(), S, v_or, r_or)
    # FP32 contraction (the FP32 copy of the reference tensor, triangle or full)
    w1, q1, _ = A1.assemble(D=dev(kappa.T.copy()), F=dev(np.ones(1)), precision=af.PREC_FP32)
    w2, q2, _ = A2.assemble(D=dev(kappa.T.copy()), F=dev(np.ones(1)), precision=af.PREC_FP32)
    assert np.array_equal(w1.cpu().numpy(), w2.cpu().numpy())
    assert np.array_equal(q1.cpu().numpy(), q2.cpu().numpy())
    compare(w2.cpu().numpy(), q2.cpu().numpy(), S, v_or, r_or, tol=2e-5)
    # a non-finite coefficient on one tet: both passes count the same entries
    kbad = kappa.copy()
    kbad[n // 2, 0] = np.inf
    _, _, b1 = A1.assemble(D=dev(kbad.T.copy()), F=dev(np.ones(1)))
    _, _, b2 = A2.assemble(D=dev(kbad.T.copy()), F=dev(np.ones(1)))
    assert b1.n_nonfinite > 0 and b2.n_nonfinite == b1.n_nonfinite


def test_anisotropic_p2_assembly_fp64_and_fp32():
    rng = np.random.default_rng(5)
    coords, tets = af.kuhn_cube(5, 0.2, 8)
    n = len(tets)
    D = aniso(rng, n)
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P2), order=2, tensor_type=af.TENSOR_SYMMETRIC,
                       tensor_storage=af.COEF_PER_TET, dirichlet=af.DIRICHLET_BOUNDARY)
    Dp = dev(T.pack_sym(D, 3).T.copy())
    val, _, st = Asm.assemble(D=Dp)
    assert st.code == 0
    setup = Asm.setup()
    S = T.oracle_assembly(coords, tets, 3, 1, dirichlet_mask=setup["dirichlet"])
    xyz = coords[S["tets"]].reshape(n, 12)
    A = T.orc_elem(2, 3, 1, 2, 3, 1, af.TENSOR_SYMMETRIC, 1, 2, xyz, D)
    A, _ = T.apply_dir_elementwise(A, np.zeros((n, 10)), S["dofs"], setup["dirichlet"])
    rc, v_or, _ = T.oracle_scatter(n, 10, S["dofs"], S["slot"], A, None, len(S["col_idx"]), S["ndof"])
    compare(val.cpu().numpy(), None, S, v_or, None)
    val32, _, _ = Asm.assemble(D=Dp, precision=af.PREC_FP32)
    e32 = compare(val32.cpu().numpy(), None, S, v_or, None, tol=2e-5)
    print(f"FP32-contraction assembly: max entry error / row max = {e32:.2e}")


def test_c2_workload_at_32cube_matches_oracle():
    """The P2 anisotropic-diffusion workload (per-tet D = R diag(1, 10^a, 10^b)
    R^T, unit load, boundary Dirichlet) on the 32^3 Kuhn cube (196,608 tets):
    FP64 within 1e-12 of the row maximum and FP32-contraction within 2e-5 of
    the oracle's element matrices, applyDir and sequential scatter."""
    rng = np.random.default_rng(32)
    coords, tets = af.kuhn_cube(32, 0.2, 1)
    n = len(tets)
    R = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
    lam = np.stack([np.ones(n), 10 ** rng.uniform(-2, 2, n), 10 ** rng.uniform(-2, 2, n)], axis=1)
    Dm = np.einsum("rij,rj,rkj->rik", R, lam, R)
    D = np.ascontiguousarray(Dm.transpose(0, 2, 1).reshape(n, 9))
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P2), order=2, tensor_type=af.TENSOR_SYMMETRIC,
                       tensor_storage=af.COEF_PER_TET, rhs_order=2, rhs_tensor_type=af.TENSOR_SCALAR,
                       rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_BOUNDARY)
    Dp = dev(T.pack_sym(D, 3).T.copy())
    f = dev(np.ones(1))
    val, rhs, st = Asm.assemble(D=Dp, F=f)
    assert st.code == 0 and st.n_degenerate == 0
    setup = Asm.setup()
    S = T.oracle_assembly(coords, tets, 3, 1, dirichlet_mask=setup["dirichlet"])
    np.testing.assert_array_equal(S["col_idx"], setup["col_idx"])
    xyz = coords[S["tets"]].reshape(n, 12)
    A = T.orc_elem(2, 3, 1, 2, 3, 1, af.TENSOR_SYMMETRIC, 1, 2, xyz, D)
    F = T.orc_elem(1, 1, 1, 1, 3, 1, af.TENSOR_SCALAR, 0, 2, xyz, np.ones((1, 1)))
    A, F = T.apply_dir_elementwise(A, F, S["dofs"], setup["dirichlet"])
    rc, v_or, r_or = T.oracle_scatter(n, 10, S["dofs"], S["slot"], A, F, len(S["col_idx"]), S["ndof"])
    assert rc == 0
    compare(val.cpu().numpy(), rhs.cpu().numpy(), S, v_or, r_or)
    val32, rhs32, _ = Asm.assemble(D=Dp, F=f, precision=af.PREC_FP32)
    compare(val32.cpu().numpy(), rhs32.cpu().numpy(), S, v_or, r_or, tol=2e-5)


def test_elasticity_p2vec_lame_assembly():
    rng = np.random.default_rng(9)
    coords, tets = af.kuhn_cube(3, 0.2, 4)
    n = len(tets)
    E = 10 ** rng.uniform(0, 3, n)
    nu = rng.uniform(0.2, 0.45, n)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P2, 3), order=2, tensor_type=af.TENSOR_LAME,
                       tensor_storage=af.COEF_PER_TET, rhs_tensor_type=af.TENSOR_GENERAL,
                       rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_BOUNDARY)
    g = np.array([0.0, 0.0, -1.0])
    val, rhs, st = Asm.assemble(D=dev(np.stack([lam, mu])), F=dev(g))
    assert st.code == 0
    setup = Asm.setup()
    S = T.oracle_assembly(coords, tets, 3, 3, dirichlet_mask=setup["dirichlet"])
    xyz = coords[S["tets"]].reshape(n, 12)
    C = np.zeros((n, 81))
    for r in range(n):
        M = np.zeros((9, 9))
        for i in range(3):
            for j in range(3):
                for k in range(3):
                    for l in range(3):
                        M[i + 3 * j, k + 3 * l] = lam[r] * (i == j) * (k == l) + mu[r] * (
                            (i == k) * (j == l) + (i == l) * (j == k))
        C[r] = M.ravel(order="F")
    A = T.orc_elem(2, 3, 3, 2, 3, 3, af.TENSOR_SYMMETRIC, 1, 2, xyz, C)
    F = T.orc_elem(1, 1, 1, 1, 3, 3, af.TENSOR_GENERAL, 0, 2, xyz, g[None, :])
    A, F = T.apply_dir_elementwise(A, F, S["dofs"], setup["dirichlet"])
    rc, v_or, r_or = T.oracle_scatter(n, 30, S["dofs"], S["slot"], A, F, len(S["col_idx"]), S["ndof"])
    assert rc == 0
    compare(val.cpu().numpy(), rhs.cpu().numpy(), S, v_or, r_or)


@pytest.mark.parametrize("op,fem,nc,tt", [(af.DIV, af.FEM_P2, 3, af.TENSOR_SCALAR), (af.CURL, af.FEM_P1, 3, af.TENSOR_NULL),
                                          (af.DUDX, af.FEM_P2, 1, af.TENSOR_SCALAR)])
def test_derivative_operator_assembly_vs_oracle(op, fem, nc, tt):
    """grad-div, curl-curl and d/dx-d/dx terms (operators.h:286-400) through
    the fused assembly (persistent passes for scalar and vector spaces)
    against the oracle's element matrices, applyDir and sequential scatter."""
    rng = np.random.default_rng(op)
    coords, tets = af.kuhn_cube(4, 0.2, 6)
    n = len(tets)
    Asm = af.Assembler(coords, tets, af.Op(op, fem, nc), order=2, tensor_type=tt, tensor_storage=af.COEF_PER_TET,
                       dirichlet=af.DIRICHLET_BOUNDARY)
    kap = rng.uniform(0.5, 2.0, n)
    val, _, st = Asm.assemble(D=dev(kap) if tt != af.TENSOR_NULL else None)
    assert st.code == 0
    setup = Asm.setup()
    S = T.oracle_assembly(coords, tets, fem, nc, dirichlet_mask=setup["dirichlet"])
    xyz = coords[S["tets"]].reshape(n, 12)
    A = T.orc_elem(op, fem, nc, op, fem, nc, tt, 1, 2, xyz, kap[:, None] if tt != af.TENSOR_NULL else None)
    nf = T.nfa(fem, nc)
    A, _ = T.apply_dir_elementwise(A, np.zeros((n, nf)), S["dofs"], setup["dirichlet"])
    rc, v_or, _ = T.oracle_scatter(n, nf, S["dofs"], S["slot"], A, None, len(S["col_idx"]), S["ndof"])
    assert rc == 0
    compare(val.cpu().numpy(), None, S, v_or, None)


def test_nonfinite_status():
    coords, tets = af.kuhn_cube(3, 0.2, 1)
    n = len(tets)
    kappa = np.ones((1, n))
    kappa[0, 17] = np.nan
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P1), tensor_type=af.TENSOR_SCALAR,
                       tensor_storage=af.COEF_PER_TET)
    _, _, st = Asm.assemble(D=dev(kappa))
    # Assemble's -1 when a local matrix holds NaN (assembler.inl:419-424)
    assert st.code == af.ERR_NONFINITE and st.n_nonfinite > 0


@pytest.mark.parametrize("fem,nc", [(af.FEM_P1, 1), (af.FEM_P2, 1), (af.FEM_P2, 3)])
def test_status_counts_per_pass(fem, nc):
    """A node with NaN coordinates: every tet around it is degenerate (its
    Jacobian determinant is not finite), and the status of each pass counts
    exactly those tets and the non-finite entries, pass after pass (the counts
    are rewritten, not accumulated); a clean pass in between reports none."""
    coords, tets = af.kuhn_cube(3, 0.2, 4)
    n = len(tets)
    bad = coords.copy()
    node = 21  # an interior node of the 4 x 4 x 4 lattice
    bad[node] = np.nan
    expect = int((tets == node).any(axis=1).sum())
    assert expect > 0
    if nc == 3:  # elasticity with a Lame tensor: the compact-row geometry launch
        kw = dict(order=2, tensor_type=af.TENSOR_LAME, tensor_storage=af.COEF_CONST)
        D = dev(np.array([1.0, 0.5]))
    else:
        kw = dict(tensor_type=af.TENSOR_SCALAR, tensor_storage=af.COEF_PER_TET)
        D = dev(np.ones((1, n)))
    A_bad = af.Assembler(bad, tets, af.Op(af.GRAD, fem, nc), dirichlet=af.DIRICHLET_BOUNDARY, **kw)
    A_ok = af.Assembler(coords, tets, af.Op(af.GRAD, fem, nc), dirichlet=af.DIRICHLET_BOUNDARY, **kw)
    counts = []
    for A in (A_bad, A_ok, A_bad, A_bad):
        _, _, st = A.assemble(D=D)
        counts.append((st.n_degenerate, st.n_nonfinite, st.code))
    assert counts[0][0] == expect and counts[0][1] > 0 and counts[0][2] == af.ERR_NONFINITE
    assert counts[1] == (0, 0, 0)
    assert counts[2] == counts[0] and counts[3] == counts[0]


@pytest.mark.parametrize("fem", [af.FEM_P1, af.FEM_P2, af.FEM_P3])
def test_nonfinite_counts_equal_across_row_block_kernels(fem):
    """NaN, Inf and overflowing (1e306) coefficients on a few tets, with
    boundary Dirichlet rows: the fixed-size persistent pass (which counts
    entry by entry only for pairs whose K is not finite or beyond 1e300), the
    general one and the per-block pass count the same entries."""
    coords, tets = af.kuhn_cube(4, 0.2, 2)
    n = len(tets)
    kappa = np.ones((1, n))
    kappa[0, [3, 50]] = np.nan
    kappa[0, 100] = np.inf
    kappa[0, [7, 200]] = 1e306
    kappa[0, 201] = 1e299
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, fem), tensor_type=af.TENSOR_SCALAR,
                       tensor_storage=af.COEF_PER_TET, rhs_tensor_type=af.TENSOR_SCALAR,
                       rhs_tensor_storage=af.COEF_PER_TET, dirichlet=af.DIRICHLET_BOUNDARY)
    counts = []
    for mode in (1, 2, 0):
        Asm.set_persistent(mode)
        _, _, st = Asm.assemble(D=dev(kappa), F=dev(kappa))
        counts.append(st.n_nonfinite)
        assert st.code == af.ERR_NONFINITE
    assert counts[0] == counts[1] == counts[2] and counts[0] > 0, counts


@pytest.mark.parametrize("fem", [2, 3])
def test_large_assembly_invariants(fem):
    """At 32^3 (196,608 tets, tens of thousands of row blocks): without
    Dirichlet every row of the stiffness matrix sums to zero, the matrix is
    symmetric, the load vector for f = 1 sums to |unit cube| = 1, and a second
    run is bitwise identical."""
    import scipy.sparse as sp
    rng = np.random.default_rng(fem)
    coords, tets = af.kuhn_cube(32, 0.2, 6)
    kappa = 10 ** rng.uniform(-1, 1, (1, len(tets)))
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, fem), order=2, tensor_type=af.TENSOR_SCALAR,
                       tensor_storage=af.COEF_PER_TET, rhs_order=2, rhs_tensor_type=af.TENSOR_SCALAR,
                       rhs_tensor_storage=af.COEF_CONST, dirichlet=0)
    assert Asm.n_blocks > 1000
    v1, r1, st = Asm.assemble(D=dev(kappa), F=dev(np.ones(1)))
    assert st.code == 0
    v2, r2, _ = Asm.assemble(D=dev(kappa), F=dev(np.ones(1)))
    v1, r1 = v1.cpu().numpy(), r1.cpu().numpy()
    assert np.array_equal(v1, v2.cpu().numpy()) and np.array_equal(r1, r2.cpu().numpy())
    s = Asm.setup()
    M = sp.csr_matrix((v1, s["col_idx"], s["row_ptr"]), shape=(Asm.ndof, Asm.ndof))
    rowmax = abs(M).max(axis=1).toarray().ravel()
    assert np.max(np.abs(M @ np.ones(Asm.ndof)) / rowmax) < 1e-12
    D = (M - M.T).tocsr()
    assert np.max(np.abs(D.data) if D.nnz else [0.0]) <= 1e-12 * rowmax.max()
    assert abs(r1.sum() - 1.0) < 1e-12


@pytest.mark.parametrize("fem,dirichlet", [(2, 1), (3, 1), (4, 0)])
def test_plain_baseline_matches_row_blocks(fem, dirichlet):
    """The plain FP64 baseline (reference algorithm per tet + atomic scatter)
    and the row-block assembly agree to rounding (only the summation order of
    the baseline differs)."""
    rng = np.random.default_rng(fem)
    coords, tets = af.kuhn_cube(5, 0.2, 3)
    order = 4 if fem == 4 else 2
    D = aniso(rng, len(tets))
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, fem), order=order, tensor_type=af.TENSOR_SYMMETRIC,
                       tensor_storage=af.COEF_PER_TET, rhs_order=order, rhs_tensor_type=af.TENSOR_SCALAR,
                       rhs_tensor_storage=af.COEF_CONST, dirichlet=dirichlet)
    Dp = dev(T.pack_sym(D, 3).T.copy())
    f = dev(np.array([1.25]))
    v1, r1, st1 = Asm.assemble(D=Dp, F=f)
    v2, r2, st2 = Asm.assemble_plain(D=Dp, F=f)
    assert st1.code == 0 and st2.code == 0
    s = Asm.setup()
    S = {"row_ptr": s["row_ptr"]}
    compare(v2.cpu().numpy(), r2.cpu().numpy(), S, v1.cpu().numpy(), r1.cpu().numpy(), tol=1e-12)


@pytest.mark.parametrize("fem,dirichlet,n", [(2, 1, 4), (3, 1, 16), (4, 0, 6), (3, 0, 24)])
def test_persistent_and_per_block_kernels_bitwise_equal(fem, dirichlet, n):
    """The persistent row-block pass (default for scalar spaces; at n=16/24
    every workgroup walks several blocks) and the one-workgroup-per-block pass
    sum in the same order: bitwise equal values and rhs, FP64 and FP32."""
    rng = np.random.default_rng(n + fem)
    coords, tets = af.kuhn_cube(n, 0.2, 6)
    order = 4 if fem == 4 else 2
    D = aniso(rng, len(tets))
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, fem), order=order, tensor_type=af.TENSOR_SYMMETRIC,
                       tensor_storage=af.COEF_PER_TET, rhs_order=order, rhs_tensor_type=af.TENSOR_SCALAR,
                       rhs_tensor_storage=af.COEF_PER_TET, dirichlet=dirichlet)
    Dp = dev(T.pack_sym(D, 3).T.copy())
    f = dev(rng.uniform(0.5, 2.0, (1, len(tets))))
    for prec in (af.PREC_FP64, af.PREC_FP32):
        Asm.set_persistent(True)  # fixed-size form
        v1, r1, st1 = Asm.assemble(D=Dp, F=f, precision=prec)
        Asm.set_persistent(False)
        v2, r2, st2 = Asm.assemble(D=Dp, F=f, precision=prec)
        Asm.set_persistent(2)  # general form
        v3, r3, st3 = Asm.assemble(D=Dp, F=f, precision=prec)
        assert st1.code == 0 and st2.code == 0 and st3.code == 0
        for v, r in ((v2, r2), (v3, r3)):
            assert np.array_equal(v1.cpu().numpy(), v.cpu().numpy())
            assert np.array_equal(r1.cpu().numpy(), r.cpu().numpy())


@pytest.mark.parametrize("fem,dirichlet,n,rhs_storage", [(af.FEM_P2, 1, 12, af.COEF_PER_TET),
                                                          (af.FEM_P1, 0, 16, af.COEF_CONST),
                                                          (af.FEM_P2, 0, 3, af.COEF_CONST)])
def test_persistent_vector_and_per_block_kernels_bitwise_equal(fem, dirichlet, n, rhs_storage):
    """Vector spaces (FemVec<3, P1|P2>, Lame tensor): the persistent pass with
    component blocks in registers (at n=12/16 every workgroup walks several
    blocks) and the one-workgroup-per-block pass sum in the same order: bitwise
    equal values and rhs, FP64 and FP32."""
    rng = np.random.default_rng(n + fem)
    coords, tets = af.kuhn_cube(n, 0.2, 8)
    nt = len(tets)
    E = 10 ** rng.uniform(0, 3, nt)
    nu = rng.uniform(0.2, 0.45, nt)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, fem, 3), order=2 if fem == af.FEM_P2 else 1,
                       tensor_type=af.TENSOR_LAME, tensor_storage=af.COEF_PER_TET,
                       rhs_tensor_type=af.TENSOR_GENERAL, rhs_tensor_storage=rhs_storage, dirichlet=dirichlet)
    g = rng.normal(size=(3, nt)) if rhs_storage == af.COEF_PER_TET else np.array([0.0, 0.0, -1.0])
    Dp, gp = dev(np.stack([lam, mu])), dev(g)
    for prec in (af.PREC_FP64, af.PREC_FP32):
        Asm.set_persistent(True)
        v1, r1, st1 = Asm.assemble(D=Dp, F=gp, precision=prec)
        Asm.set_persistent(False)
        v2, r2, st2 = Asm.assemble(D=Dp, F=gp, precision=prec)
        assert st1.code == 0 and st2.code == 0
        assert np.array_equal(v1.cpu().numpy(), v2.cpu().numpy())
        assert np.array_equal(r1.cpu().numpy(), r2.cpu().numpy())


@pytest.mark.parametrize("fem", [af.FEM_P1, af.FEM_P2])
def test_lame_compact_rows_match_full_blocks(fem):
    """A Lame tensor constant over each tet takes the compact geometry rows (T,
    lambda |T|, mu |T|, expanded to the 3 x 3 component blocks in the row-block
    pass); per-point storage with the same values at every point takes the full
    K blocks.  Both against the row maximum of the other: within 1e-13, FP64;
    the persistent and one-workgroup-per-block passes of the compact form are
    bitwise equal, with and without Dirichlet values."""
    rng = np.random.default_rng(17 + fem)
    coords, tets = af.kuhn_cube(4, 0.2, 5)
    n = len(tets)
    lam, mu = rng.uniform(1, 10, n), rng.uniform(1, 10, n)
    order = 2
    kw = dict(order=order, rhs_tensor_type=af.TENSOR_GENERAL, rhs_tensor_storage=af.COEF_CONST,
              dirichlet=af.DIRICHLET_BOUNDARY)
    op = af.Op(af.GRAD, fem, 3)
    A1 = af.Assembler(coords, tets, op, tensor_type=af.TENSOR_LAME, tensor_storage=af.COEF_PER_TET, **kw)
    A2 = af.Assembler(coords, tets, op, tensor_type=af.TENSOR_LAME, tensor_storage=af.COEF_PER_POINT, **kw)
    q = len(af.quadrature(order)[1])
    Dp = np.ascontiguousarray(np.repeat(np.stack([lam, mu])[:, None, :], q, axis=1))
    g = dev(np.array([0.3, 0.0, -1.0]))
    v1, r1, s1 = A1.assemble(D=dev(np.stack([lam, mu])), F=g)
    v2, r2, s2 = A2.assemble(D=dev(Dp), F=g)
    assert s1.code == 0 and s2.code == 0
    S = {"row_ptr": A1.setup()["row_ptr"]}
    compare(v1.cpu().numpy(), r1.cpu().numpy(), S, v2.cpu().numpy(), r2.cpu().numpy(), tol=1e-13)
    bc = dev(np.where(A1.setup()["dirichlet"] != 0, rng.normal(size=A1.ndof), 0.0))
    out = []
    for persistent in (True, False):
        A1.set_persistent(persistent)
        out.append(A1.assemble(D=dev(np.stack([lam, mu])), F=g, dir_values=bc))
    assert np.array_equal(out[0][0].cpu().numpy(), out[1][0].cpu().numpy())
    assert np.array_equal(out[0][1].cpu().numpy(), out[1][1].cpu().numpy())


@pytest.mark.parametrize("fem,nc", [(af.FEM_P2, 1), (af.FEM_P1, 3), (af.FEM_P3, 1)])
def test_assemble_elements_matches_oracle_scatter(fem, nc):
    """Caller-supplied element matrices (random, FP64 and FP32 values) through
    the row-block lists (anifem_assembly_run_elements) against the oracle's
    applyDir + sequential cell-order scatter: equal bit for bit."""
    rng = np.random.default_rng(fem * 7 + nc)
    coords, tets = af.kuhn_cube(3, 0.2, 7)
    # random element matrices couple the components: a Lame-tensor context keeps every structural entry
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, fem, nc), tensor_type=af.TENSOR_LAME if nc == 3 else af.TENSOR_NULL,
                       dirichlet=af.DIRICHLET_BOUNDARY)
    n, nf = Asm.n_tets, Asm.nf
    setup = Asm.setup()
    A = rng.normal(size=(n, nf * nf))
    F = rng.normal(size=(n, nf))
    for dt in (np.float64, np.float32):
        Ad, Fd = A.astype(dt), F.astype(dt)
        val, rhs, st = Asm.assemble_elements(dev(Ad), dev(Fd))
        assert st.code == 0
        Ao, Fo = T.apply_dir_elementwise(Ad.astype(np.float64), Fd.astype(np.float64), setup["dofs"],
                                         setup["dirichlet"])
        rc, v_or, r_or = T.oracle_scatter(n, nf, setup["dofs"], setup["slot"], Ao, Fo, Asm.nnz, Asm.ndof)
        assert rc == 0
        np.testing.assert_array_equal(val.cpu().numpy(), v_or)
        np.testing.assert_array_equal(rhs.cpu().numpy(), r_or)
    # no load vectors: rhs zeroed; a non-finite entry is reported
    import torch
    rhs0 = torch.full((Asm.ndof,), 3.0, dtype=torch.float64, device="cuda")
    Ab = A.copy()
    Ab[n // 2, 0] = np.nan
    _, rhs0, st = Asm.assemble_elements(dev(Ab), None, rhs=rhs0)
    assert torch.all(rhs0 == 0)
    assert st.code == af.ERR_NONFINITE and st.n_nonfinite >= 1


@pytest.mark.parametrize("fem,nc", [(af.FEM_P2, 1), (af.FEM_P2, 3)])
def test_assemble_elements_symmetric_reads_columns(fem, nc):
    """ANIFEM_ELEM_SYMMETRIC: exactly symmetric element matrices give the same
    CSR values bit for bit when the rows are read as columns; for a
    non-symmetric input the flag assembles the transposes."""
    import torch
    rng = np.random.default_rng(nc)
    coords, tets = af.kuhn_cube(3, 0.2, 9)
    # random element matrices couple the components: a Lame-tensor context keeps every structural entry
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, fem, nc), tensor_type=af.TENSOR_LAME if nc == 3 else af.TENSOR_NULL,
                       dirichlet=af.DIRICHLET_BOUNDARY)
    n, nf = Asm.n_tets, Asm.nf
    B = rng.normal(size=(n, nf, nf))
    for dt in (torch.float64, torch.float32):
        S = torch.from_numpy(B + B.transpose(0, 2, 1)).to("cuda", dt)
        v1, _, _ = Asm.assemble_elements(S)
        v2, _, _ = Asm.assemble_elements(S, symmetric=True)
        assert torch.equal(v1, v2)
        N = torch.from_numpy(B).to("cuda", dt)
        v3, _, _ = Asm.assemble_elements(N, symmetric=True)
        v4, _, _ = Asm.assemble_elements(N.transpose(1, 2).contiguous())
        assert torch.equal(v3, v4)


@pytest.mark.parametrize("fem,nc,n", [(af.FEM_P1, 1, 16), (af.FEM_P2, 1, 12), (af.FEM_P3, 1, 6), (af.FEM_P1, 3, 10),
                                      (af.FEM_P2, 3, 6), (af.FEM_P2, 1, 2)])
def test_assemble_elements_persistent_and_per_block_bitwise_equal(fem, nc, n):
    """Element-matrix scatter: the persistent pass (4, 10, 12, 20 and 30 DOFs
    per cell; at the larger sizes every workgroup walks several blocks) and the
    one-workgroup-per-block pass give equal values and rhs bit for bit, FP64
    and FP32, rows read as rows or (symmetric) as columns."""
    import torch
    rng = np.random.default_rng(fem * 10 + nc + n)
    coords, tets = af.kuhn_cube(n, 0.2, 11)
    # random element matrices couple the components: a Lame-tensor context keeps every structural entry
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, fem, nc), tensor_type=af.TENSOR_LAME if nc == 3 else af.TENSOR_NULL,
                       dirichlet=af.DIRICHLET_BOUNDARY)
    nt, nf = Asm.n_tets, Asm.nf
    A = rng.normal(size=(nt, nf, nf))
    F = rng.normal(size=(nt, nf))
    for dt in (torch.float64, torch.float32):
        Ad, Fd = torch.from_numpy(A).to("cuda", dt), torch.from_numpy(F).to("cuda", dt)
        for sym in (False, True):
            Asm.set_persistent(True)
            v1, r1, s1 = Asm.assemble_elements(Ad, Fd, symmetric=sym)
            Asm.set_persistent(False)
            v2, r2, s2 = Asm.assemble_elements(Ad, Fd, symmetric=sym)
            assert s1.code == 0 and s2.code == 0
            assert torch.equal(v1, v2) and torch.equal(r1, r2), (dt, sym)
    Asm.set_persistent(True)
    # a non-finite diagonal entry of a DOF that is not Dirichlet (so stored)
    # is counted by the persistent pass too
    S = Asm.setup()
    t, i = np.argwhere(S["dirichlet"][S["dofs"]] == 0)[0]
    A[t, i, i] = np.inf
    _, _, st = Asm.assemble_elements(torch.from_numpy(A).cuda(), torch.from_numpy(F).cuda())
    assert st.code == af.ERR_NONFINITE and st.n_nonfinite >= 1


def test_assemble_elements_from_batch_matches_fused():
    """Element matrices from fem3dtet_batch on the reordered tets, scattered by
    assemble_elements, against the fused assembly: FP64 elements agree to
    rounding, FP16X3 (MFMA) elements within the mode's element bound."""
    import torch
    rng = np.random.default_rng(21)
    coords, tets = af.kuhn_cube(6, 0.2, 9)
    n = len(tets)
    Dp = dev(T.pack_sym(aniso(rng, n), 3).T.copy())
    f = dev(np.ones(1))
    op = af.Op(af.GRAD, af.FEM_P2)
    Asm = af.Assembler(coords, tets, op, order=2, tensor_type=af.TENSOR_SYMMETRIC, tensor_storage=af.COEF_PER_TET,
                       rhs_order=2, rhs_tensor_type=af.TENSOR_SCALAR, rhs_tensor_storage=af.COEF_CONST,
                       dirichlet=af.DIRICHLET_BOUNDARY)
    val, rhs, st = Asm.assemble(D=Dp, F=f)
    assert st.code == 0
    xyz = dev(T.soa_xyz(coords[Asm.tets()].reshape(n, 12)))
    p0, p2 = af.Op(af.IDEN, af.FEM_P0), af.Op(af.IDEN, af.FEM_P2)
    vmax, rmax = val.abs().max().item(), rhs.abs().max().item()
    for prec, tol in [(af.PREC_FP64, 1e-13), (af.PREC_FP16X3, 1e-6)]:
        Ae = af.fem3dtet_batch(op, op, af.TENSOR_SYMMETRIC, af.COEF_PER_TET, Dp, 2, xyz, prec)
        Fe = af.fem3dtet_batch(p0, p2, af.TENSOR_SCALAR, af.COEF_CONST, f, 2, xyz, prec).reshape(n, 10)
        v2, r2, st2 = Asm.assemble_elements(Ae, Fe)
        assert st2.code == 0
        ev = (v2 - val).abs().max().item() / vmax
        er = (r2 - rhs).abs().max().item() / rmax
        print(f"precision {prec}: elements-then-scatter vs fused, max entry error {ev:.2e}, rhs {er:.2e}")
        assert ev <= tol and er <= tol, (prec, ev, er)
    del torch


@pytest.mark.parametrize("fem,nc,tt", [(af.FEM_P2, 1, af.TENSOR_SYMMETRIC), (af.FEM_P1, 3, af.TENSOR_LAME),
                                       (af.FEM_P3, 1, af.TENSOR_SCALAR)])
def test_coloured_scatter_matches_row_blocks(fem, nc, tt):
    """The coloured scatter (greedy colouring, one launch per colour) against
    the row-block pass: the oracle's first-fit colouring, conflict-free
    colours, values equal up to the order of the additions, repeat runs
    bitwise identical."""
    import ctypes
    import torch
    rng = np.random.default_rng(31 + fem)
    coords, tets = af.kuhn_cube(5, 0.2, 12)
    n = len(tets)
    order = 4 if fem == af.FEM_P3 else 2
    if tt == af.TENSOR_SYMMETRIC:
        D = dev(T.pack_sym(aniso(rng, n), 3).T.copy())
    elif tt == af.TENSOR_LAME:
        D = dev(np.stack([rng.uniform(1, 10, n), rng.uniform(1, 10, n)]))
    else:
        D = dev(10 ** rng.uniform(-1, 1, (1, n)))
    rhs_tt = af.TENSOR_GENERAL if nc == 3 else af.TENSOR_SCALAR
    F = dev(np.array([0.0, 0.0, -1.0])) if nc == 3 else dev(np.ones(1))
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, fem, nc), order=order, tensor_type=tt,
                       tensor_storage=af.COEF_PER_TET, rhs_order=order, rhs_tensor_type=rhs_tt,
                       rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_BOUNDARY)
    v_rb, r_rb, st = Asm.assemble(D=D, F=F)
    assert st.code == 0
    v_rb, r_rb = v_rb.cpu().numpy(), r_rb.cpu().numpy()
    Asm.set_scatter(af.SCATTER_COLOURED)
    ncol, col = Asm.colours()
    col_or = np.zeros(n, np.int32)
    tr = np.ascontiguousarray(Asm.tets())
    assert T.oracle().orc_greedy_colour(ctypes.c_long(len(coords)), ctypes.c_long(n), T.P(tr), T.P(col_or)) == ncol
    np.testing.assert_array_equal(col, col_or)
    for c in range(ncol):  # no node in two tets of one colour
        nodes = tr[col == c].ravel()
        assert len(np.unique(nodes)) == len(nodes)
    v_c, r_c, st = Asm.assemble(D=D, F=F)
    assert st.code == 0
    s = Asm.setup()
    S = {"row_ptr": s["row_ptr"]}
    compare(v_c.cpu().numpy(), r_c.cpu().numpy(), S, v_rb, r_rb, tol=1e-13)
    v_c2, r_c2, _ = Asm.assemble(D=D, F=F)
    assert torch.equal(v_c, v_c2) and torch.equal(r_c, r_c2)
    Asm.set_scatter(af.SCATTER_ROW_BLOCKS)
    v_b, _, _ = Asm.assemble(D=D, F=F)
    np.testing.assert_array_equal(v_b.cpu().numpy(), v_rb)


def test_plain_baseline_after_coloured_scatter():
    """The plain baseline and the coloured scatter share the element maps of a
    context: the plain path builds its own tables even when the coloured
    scatter uploaded the maps first (regression: its tables were tied to the
    maps' upload)."""
    rng = np.random.default_rng(41)
    coords, tets = af.kuhn_cube(4, 0.2, 13)
    n = len(tets)
    D = dev(T.pack_sym(aniso(rng, n), 3).T.copy())
    F = dev(np.ones(1))
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P2), order=2, tensor_type=af.TENSOR_SYMMETRIC,
                       tensor_storage=af.COEF_PER_TET, rhs_order=2, rhs_tensor_type=af.TENSOR_SCALAR,
                       rhs_tensor_storage=af.COEF_CONST, dirichlet=af.DIRICHLET_BOUNDARY)
    v_rb, r_rb, _ = Asm.assemble(D=D, F=F)
    Asm.set_scatter(af.SCATTER_COLOURED)
    Asm.assemble(D=D, F=F)
    Asm.set_scatter(af.SCATTER_ROW_BLOCKS)
    v_p, r_p, st = Asm.assemble_plain(D=D, F=F)
    assert st.code == 0
    S = {"row_ptr": Asm.setup()["row_ptr"]}
    compare(v_p.cpu().numpy(), r_p.cpu().numpy(), S, v_rb.cpu().numpy(), r_rb.cpu().numpy(), tol=1e-12)
