"""Learned preconditioner on the GPU (anifem_gnn_*, anifem_pcg_factored,
anifem_pcg_matrix, anifem_pcg_learned, anifem_grf_sample) against the float64 CPU restatement
tests/gnn_oracle.py and against Jacobi-PCG."""
import numpy as np
import pytest

import anifem_amd as af
import gnn_oracle as O

pytestmark = pytest.mark.gpu


def dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.to("cuda") if dtype is None else t.to("cuda", dtype)


def gpu_system(n=4, seed=0, dirichlet=af.DIRICHLET_BOUNDARY):
    coords, tets = af.kuhn_cube(n, 0.2, 1)
    A = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P1), order=2, tensor_type=af.TENSOR_SCALAR,
                     tensor_storage=af.COEF_PER_TET, rhs_order=2, rhs_tensor_type=af.TENSOR_SCALAR,
                     rhs_tensor_storage=af.COEF_CONST, dirichlet=dirichlet)
    kap = np.exp(np.random.default_rng(seed).normal(0, 1.5, (1, len(tets))))
    val, rhs, st = A.assemble(D=dev(kap), F=dev(np.ones(1)))
    assert st.code == 0
    s = A.setup()
    return A, s, val, rhs


def random_params(net, seed, head=0.1):
    rng = np.random.default_rng(seed)
    theta = af._gnn_init(net.n_params, seed).astype(np.float64)
    theta[-(2 * 32 + 2):] = rng.normal(0, head, 2 * 32 + 2)  # non-zero output heads
    return theta.astype(np.float32)


def test_factor_matches_oracle():
    A, s, val, _ = gpu_system()
    net = af.LearnedPreconditioner(s["row_ptr"], s["col_idx"])
    net.set_params(random_params(net, 1))
    import torch
    G = torch.empty(A.nnz, dtype=torch.float64, device="cuda")
    GT = torch.empty_like(G)
    net.factor(val, G, GT)
    g = O.Graph(s["row_ptr"], s["col_idx"], val.cpu().numpy())
    theta_bf = net.params(bf16=True).astype(np.float64)
    G_or, _ = O.forward(theta_bf, g)
    Gg = G.cpu().numpy()
    assert np.max(np.abs(Gg - G_or)) <= 2e-5 * np.max(np.abs(G_or))
    np.testing.assert_array_equal(GT.cpu().numpy(), Gg[g.tslot])
    # lower triangular on the pattern
    assert np.all(Gg[g.row < g.ci] == 0)


def test_loss_and_gradient_match_oracle():
    A, s, val, _ = gpu_system(seed=2)
    net = af.LearnedPreconditioner(s["row_ptr"], s["col_idx"])
    net.set_params(random_params(net, 3))
    rng = np.random.default_rng(5)
    Z = rng.choice([-1.0, 1.0], size=(A.ndof, 4)).astype(np.float32)
    loss = net.loss_grad(val, probes=dev(Z))
    g_gpu = net.grad().astype(np.float64)
    g = O.Graph(s["row_ptr"], s["col_idx"], val.cpu().numpy())
    theta_bf = net.params(bf16=True).astype(np.float64)
    loss_or = O.loss(theta_bf, g, Z.astype(np.float64))
    g_or = O.grad(theta_bf, g, Z.astype(np.float64))
    assert abs(loss - loss_or) <= 1e-4 * abs(loss_or)
    rel = np.linalg.norm(g_gpu - g_or) / np.linalg.norm(g_or)
    assert rel <= 2e-3, rel
    # every parameter block receives its gradient
    o = 0
    for k, sh in O.PARAM_SHAPES:
        c = int(np.prod(sh))
        blk = slice(o, o + c)
        assert np.linalg.norm(g_gpu[blk] - g_or[blk]) <= 5e-3 * np.linalg.norm(g_or[blk]) + 1e-7, k
        o += c
    # deterministic
    net.loss_grad(val, probes=dev(Z))
    np.testing.assert_array_equal(net.grad(), g_gpu.astype(np.float32))


def test_zero_heads_match_jacobi_pcg():
    import torch
    A, s, val, rhs = gpu_system(n=6, seed=4)
    net = af.LearnedPreconditioner(s["row_ptr"], s["col_idx"])
    net.set_params(af._gnn_init(net.n_params, 0))
    G, GT = net.factor(val)
    rp, ci = A.csr()
    x = torch.zeros(A.ndof, dtype=torch.float64, device="cuda")
    rj, okj = af.pcg_jacobi(A.ndof, rp, ci, val, rhs, x, rtol=1e-10)
    xj = x.clone()
    x.zero_()
    rf, okf = af.pcg_factored(A.ndof, rp, ci, val, G, GT, rhs, x, rtol=1e-10)
    assert okj and okf
    assert abs(rf.iterations - rj.iterations) <= 1
    assert rf.true_relres <= 1e-10 * 1.01
    assert torch.max(torch.abs(x - xj)).item() <= 1e-8 * torch.max(torch.abs(xj)).item()


def test_precond_matrix_is_G_GT():
    A, s, val, _ = gpu_system(n=5, seed=8)
    net = af.LearnedPreconditioner(s["row_ptr"], s["col_idx"])
    net.set_params(random_params(net, 9))
    G_addr, _ = net.factor(val)
    m_nnz, mrp, mci, mv = net.precond()
    n = A.ndof
    M_rp = af.copy_to_host(mrp, n + 1, np.int64)
    M_ci = af.copy_to_host(mci, m_nnz, np.int32)
    M_v = af.copy_to_host(mv, m_nnz, np.float64)
    G = af.copy_to_host(G_addr, A.nnz, np.float64)
    g = O.Graph(s["row_ptr"], s["col_idx"], val.cpu().numpy())
    Gd = np.zeros((n, n))
    Gd[g.row, g.ci] = G
    Md = Gd @ Gd.T
    # pattern: exactly the pairs (i, j) that share a column k of G's pattern
    assert np.all(G[g.row >= g.ci] != 0)
    S = np.zeros((n, n), np.int64)
    S[g.row[g.row >= g.ci], g.ci[g.row >= g.ci]] = 1
    P = (S @ S.T) > 0
    Mh = np.zeros((n, n))
    seen = np.zeros((n, n), bool)
    for i in range(n):
        cols = M_ci[M_rp[i]:M_rp[i + 1]]
        assert np.all(np.diff(cols) > 0)
        Mh[i, cols] = M_v[M_rp[i]:M_rp[i + 1]]
        seen[i, cols] = True
    np.testing.assert_array_equal(seen, P)
    assert np.max(np.abs(Mh - Md)) <= 1e-14 * np.max(np.abs(Md))
    np.testing.assert_array_equal(Mh, Mh.T)  # bitwise symmetric


def test_pcg_matrix_matches_factored():
    import torch
    A, s, val, rhs = gpu_system(n=8, seed=10)
    net = af.LearnedPreconditioner(s["row_ptr"], s["col_idx"])
    net.set_params(random_params(net, 11, head=0.05))
    G, GT = net.factor(val)
    m_nnz, mrp, mci, mv = net.precond()
    rp, ci = A.csr()
    x = torch.zeros(A.ndof, dtype=torch.float64, device="cuda")
    rf, okf = af.pcg_factored(A.ndof, rp, ci, val, G, GT, rhs, x, rtol=1e-10)
    xf = x.clone()
    x.zero_()
    rm, okm = af.pcg_matrix(A.ndof, rp, ci, val, mrp, mci, mv, rhs, x, rtol=1e-10)
    assert okf and okm
    assert abs(rm.iterations - rf.iterations) <= 2, (rm.iterations, rf.iterations)
    assert rm.true_relres <= 1.01e-10
    assert torch.max(torch.abs(x - xf)).item() <= 1e-8 * torch.max(torch.abs(xf)).item()
    # repeated solves are bitwise identical
    x2 = torch.zeros_like(x)
    af.pcg_matrix(A.ndof, rp, ci, val, mrp, mci, mv, rhs, x2, rtol=1e-10)
    assert torch.equal(x, x2)


def test_training_reduces_loss_and_iterations():
    from anifem_amd.learned import GrfSystems, train, compare
    sysm = GrfSystems(n=10, ell=0.1, sigma=1.5)
    net, losses, secs = train(sysm, np.arange(300), lr=3e-3, epochs=2)
    assert np.mean(losses[-50:]) < 0.7 * np.mean(losses[:20]), (losses[:20].mean(), losses[-50:].mean())
    rows = compare(sysm, net, np.arange(10_000, 10_005), rtol=1e-8)
    for r in rows:
        assert r["ok"] and r["fallback"] == 0
        assert r["learned_true"] <= 1.01e-8 and r["jacobi_true"] <= 1.01e-8
    assert np.mean([r["learned_it"] for r in rows]) < 0.8 * np.mean([r["jacobi_it"] for r in rows])


def test_learned_factor_against_fsai_on_the_same_pattern():
    """The trained network's factor and the non-learned FSAI factor on the
    same lower pattern, in the same PCG with M^-1 = G G^T: both converge to
    the tolerance by the FP64 true residual, FSAI needs fewer iterations than
    Jacobi, and the comparison is reported."""
    from anifem_amd.learned import GrfSystems, train, compare_fsai, compare
    sysm = GrfSystems(n=10, ell=0.1, sigma=1.5)
    net, _, _ = train(sysm, np.arange(300), lr=3e-3, epochs=2)
    seeds = np.arange(20_000, 20_005)
    rows = compare_fsai(sysm, net, seeds, rtol=1e-8)
    jac = compare(sysm, net, seeds, rtol=1e-8)
    for r, j in zip(rows, jac):
        assert r["ok"]
        assert r["fsai_true"] <= 1.01e-8 and r["learned_true"] <= 1.01e-8
        assert r["fsai_it"] < j["jacobi_it"]
    f_it = np.mean([r["fsai_it"] for r in rows])
    l_it = np.mean([r["learned_it"] for r in rows])
    j_it = np.mean([j["jacobi_it"] for j in jac])
    print(f"iterations: Jacobi {j_it:.1f}, FSAI {f_it:.1f}, learned {l_it:.1f}")


def test_fallback_to_jacobi_when_the_factor_fails():
    import torch
    A, s, val, rhs = gpu_system(n=5, seed=6)
    net = af.LearnedPreconditioner(s["row_ptr"], s["col_idx"])
    theta = af._gnn_init(net.n_params, 0)
    theta[-1] = 1e4  # c_d: exp overflows, the factor is not finite
    net.set_params(theta)
    x = torch.zeros(A.ndof, dtype=torch.float64, device="cuda")
    res, ok = net.solve(val, rhs, x, rtol=1e-9, maxit=2000)
    assert ok and res.fallback == 1
    assert res.true_relres <= 1.01e-9
    assert not np.isfinite(res.learned_true_relres) or res.learned_true_relres > 1e-9


def test_grf_covariance():
    import torch
    d = np.array([0.0, 0.05, 0.1, 0.2, 0.4])
    pts = np.stack([0.3 + d, np.full(5, 0.5), np.full(5, 0.5)], axis=1)
    xyz = dev(pts)
    sig, ell = 1.5, 0.1
    S = np.stack([af.grf_sample(xyz, ell, sig, 512, seed, lognormal=False).cpu().numpy() for seed in range(600)])
    cov = (S[:, :1] * S).mean(axis=0)
    target = sig ** 2 * np.exp(-d ** 2 / (2 * ell ** 2))
    assert np.max(np.abs(cov - target)) <= 0.2 * sig ** 2, (cov, target)
    assert abs(S.mean()) <= 0.2 * sig
    a = af.grf_sample(xyz, ell, sig, 512, 7).cpu().numpy()
    b = af.grf_sample(xyz, ell, sig, 512, 7).cpu().numpy()
    np.testing.assert_array_equal(a, b)
    assert np.all(a > 0)
    del torch
