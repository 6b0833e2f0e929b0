"""The CPU restatement of the learned preconditioner (tests/gnn_oracle.py):
its hand-written reverse-mode gradient against central finite differences,
and the identity network output being exactly Jacobi.  No GPU."""
import numpy as np

import anifem_amd as af
import anifem_testlib as T
import gnn_oracle as O


def small_system(n=3, seed=0):
    coords, tets = af.kuhn_cube(n, 0.2, 1)
    A = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P1), dirichlet=af.DIRICHLET_BOUNDARY, host_only=True)
    s = A.setup()
    rng = np.random.default_rng(seed)
    kap = np.exp(rng.normal(0, 1, len(tets)))
    S = T.oracle_assembly(coords, tets, 2, 1, dirichlet_mask=s["dirichlet"])
    xyz = coords[S["tets"]].reshape(len(tets), 12)
    Ae = T.orc_elem(2, 2, 1, 2, 2, 1, 2, 1, 2, xyz, kap[:, None])
    Ae, _ = T.apply_dir_elementwise(Ae, np.zeros((len(tets), 4)), S["dofs"], s["dirichlet"])
    rc, val, _ = T.oracle_scatter(len(tets), 4, S["dofs"], S["slot"], Ae, None, len(S["col_idx"]), S["ndof"])
    assert rc == 0
    return O.Graph(S["row_ptr"], S["col_idx"], val)


def test_reverse_mode_gradient_matches_finite_differences():
    g = small_system()
    rng = np.random.default_rng(1)
    theta = O.init_params(1) + rng.normal(0, 0.05, O.n_params())
    Z = rng.choice([-1.0, 1.0], size=(g.n, 4))
    gr = O.grad(theta, g, Z)
    eps = 1e-6
    for i in rng.choice(len(theta), 40, replace=False):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        fd = (O.loss(tp, g, Z) - O.loss(tm, g, Z)) / (2 * eps)
        assert abs(fd - gr[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-5 * abs(fd), (i, fd, gr[i])


def test_zero_heads_give_jacobi():
    g = small_system()
    G, _ = O.forward(O.init_params(2), g)
    r = np.random.default_rng(3).normal(size=g.n)
    np.testing.assert_allclose(O.apply_precond(G, g, r), r * g.dinv ** 2, rtol=1e-14)


def test_planted_mistake_in_gradient_is_caught():
    """A gradient with a dropped term (the G z path of the loss) fails the
    finite-difference check."""
    g = small_system()
    rng = np.random.default_rng(4)
    theta = O.init_params(1) + rng.normal(0, 0.05, O.n_params())
    Z = rng.choice([-1.0, 1.0], size=(g.n, 4))
    p = O.unpack(theta)
    G, c = O.forward(theta, g)
    y = g.spmv(G, Z)
    w = g.spmv(g.a, y)
    res = g.spmv_t(G, w) - Z
    gu = 2.0 * res / (Z.shape[1] * g.n)
    dG_wrong = np.sum(w[g.row] * gu[g.ci], axis=1)  # misses gy z
    dd_wrong = np.bincount(g.row, weights=dG_wrong * (g.isd + c["L"]), minlength=g.n) * c["Dh"]
    i = O.PARAM_SHAPES.index(("c_d", (1,)))
    off = sum(int(np.prod(s)) for _, s in O.PARAM_SHAPES[:i])
    eps = 1e-6
    tp, tm = theta.copy(), theta.copy()
    tp[off] += eps
    tm[off] -= eps
    fd = (O.loss(tp, g, Z) - O.loss(tm, g, Z)) / (2 * eps)
    assert abs(dd_wrong.sum() - fd) > 1e-3 * abs(fd)
    assert abs(O.grad(theta, g, Z)[off] - fd) <= 1e-6 * max(1, abs(fd))
    del p
