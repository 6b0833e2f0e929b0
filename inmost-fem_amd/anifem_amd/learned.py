"""Learned preconditioner workflow on the GPU: Gaussian-random-field diffusion
systems produced by this assembler, training of the message-passing network
(anifem_gnn_* kernels) and PCG with M = G G^T against Jacobi-PCG.

System family (one mesh, many coefficients): P1 diffusion on the unit cube
split into n^3 Kuhn sub-cubes (jittered), kappa = exp(g) per tet with g a
Gaussian random field (squared-exponential covariance, correlation length
ell, standard deviation sigma) at the tet centroids, f = 1, homogeneous
Dirichlet on the boundary.
"""
import time

import numpy as np

from . import (Assembler, LearnedPreconditioner, Op, GRAD, FEM_P1, TENSOR_SCALAR, COEF_PER_TET, COEF_CONST,
               DIRICHLET_BOUNDARY, grf_sample, kuhn_cube, pcg_jacobi)


class GrfSystems:
    """Seeded GRF diffusion systems on one mesh; matrices assembled on demand."""

    def __init__(self, n=32, ell=0.1, sigma=1.5, jitter=0.2, mesh_seed=1, n_features=512):
        import torch
        self.dev = torch.device("cuda", torch.cuda.current_device())
        coords, tets = kuhn_cube(n, jitter, mesh_seed)
        self.asm = Assembler(coords, tets, Op(GRAD, FEM_P1), order=2, tensor_type=TENSOR_SCALAR,
                             tensor_storage=COEF_PER_TET, rhs_order=2, rhs_tensor_type=TENSOR_SCALAR,
                             rhs_tensor_storage=COEF_CONST, dirichlet=DIRICHLET_BOUNDARY)
        cen = coords[tets].mean(axis=1)
        self.centroids = torch.from_numpy(np.ascontiguousarray(cen)).to(self.dev)
        self.ell, self.sigma, self.n_features = ell, sigma, n_features
        s = self.asm.setup()
        self.row_ptr, self.col_idx = s["row_ptr"], s["col_idx"]
        self.rp_dev, self.ci_dev = self.asm.csr()
        self.n = self.asm.ndof
        self.f = torch.ones(1, dtype=torch.float64, device=self.dev)
        self.kappa = torch.empty(len(tets), dtype=torch.float64, device=self.dev)

    def system(self, seed, val=None, rhs=None):
        """(val, rhs) device FP64 for coefficient seed `seed`."""
        grf_sample(self.centroids, self.ell, self.sigma, self.n_features, seed, True, out=self.kappa)
        val, rhs, _ = self.asm.assemble(D=self.kappa.view(1, -1), F=self.f, val=val, rhs=rhs, status=False)
        return val, rhs


def train(systems, seeds, lr=1e-3, epochs=1, init_seed=0, log=None):
    """Adam on the probe loss over the systems of `seeds`; returns (net, losses, seconds)."""
    import torch
    from . import _gnn_init
    net = LearnedPreconditioner(systems.row_ptr, systems.col_idx)
    net.set_params(_gnn_init(net.n_params, init_seed))
    val = torch.empty(systems.asm.nnz, dtype=torch.float64, device=systems.dev)
    rhs = torch.empty(systems.n, dtype=torch.float64, device=systems.dev)
    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step = 0
    for ep in range(epochs):
        for sd in seeds:
            systems.system(int(sd), val, rhs)
            loss = net.loss_grad(val, seed=1_000_003 * (step + 1))
            net.adam_step(lr=lr)
            losses.append(loss)
            step += 1
            if log and step % 200 == 0:
                log(f"step {step} loss {np.mean(losses[-200:]):.5f}")
    torch.cuda.synchronize()
    return net, np.array(losses), time.perf_counter() - t0


def compare(systems, net, seeds, rtol=1e-8, maxit=20000):
    """Per held-out seed: Jacobi-PCG and learned PCG times (ms, synchronized
    wall clock of the solve calls), iterations, true residuals, fallbacks, and
    the time of the network's factor plus M = G G^T alone (factor_ms)."""
    import torch
    val = torch.empty(systems.asm.nnz, dtype=torch.float64, device=systems.dev)
    rhs = torch.empty(systems.n, dtype=torch.float64, device=systems.dev)
    x = torch.zeros(systems.n, dtype=torch.float64, device=systems.dev)
    rows = []
    for sd in seeds:
        systems.system(int(sd), val, rhs)
        x.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        rj, okj = pcg_jacobi(systems.n, systems.rp_dev, systems.ci_dev, val, rhs, x, rtol=rtol, maxit=maxit)
        torch.cuda.synchronize()
        tj = (time.perf_counter() - t) * 1e3
        x.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        rl, okl = net.solve(val, rhs, x, rtol=rtol, maxit=maxit)
        torch.cuda.synchronize()
        tl = (time.perf_counter() - t) * 1e3
        torch.cuda.synchronize()
        t = time.perf_counter()
        net.factor(val)
        net.precond()
        torch.cuda.synchronize()
        tf = (time.perf_counter() - t) * 1e3
        rows.append(dict(seed=int(sd), jacobi_ms=tj, jacobi_it=rj.iterations, jacobi_true=rj.true_relres,
                         learned_ms=tl, learned_it=rl.iterations, learned_true=rl.true_relres,
                         fallback=rl.fallback, factor_ms=tf, ok=bool(okj and okl)))
    return rows


def fsai_factor(row_ptr, col_idx, val):
    """Factorized sparse approximate inverse on the lower pattern of a
    symmetric positive definite CSR matrix (host arrays): the non-learned
    factor the network competes with, on the same pattern and with the same
    application cost.  Column j of the lower-triangular G is nonzero on
    S_j = {i >= j : a_ij != 0}; it solves A[S_j, S_j] y = e_j and is scaled to
    g = y / sqrt(y_j), so that G^T A G has a unit diagonal and (A g_j)_i = 0 for
    the other i in S_j (Kolotilina and Yeremin's construction, applied to
    columns, for M^-1 = G G^T).  Returns (G, GT): values of G and of G^T on the
    CSR pattern (zero outside the lower / upper part)."""
    rp = np.asarray(row_ptr, dtype=np.int64)
    ci = np.asarray(col_idx, dtype=np.int64)
    v = np.asarray(val, dtype=np.float64)
    n = len(rp) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    keys = rows * n + ci  # ascending (rows ascend, columns ascend within a row)
    lower = ci >= rows    # entry (j, i) with i >= j: i in S_j (A symmetric)
    m = np.bincount(rows[lower], minlength=n)  # |S_j|
    start = np.zeros(n + 1, dtype=np.int64)
    start[1:] = np.cumsum(m)
    S_flat = ci[lower]  # S_j in ascending order; S_j[0] = j (the diagonal is stored)
    G = np.zeros(len(v))
    for mj in np.unique(m):
        js = np.nonzero(m == mj)[0]
        idx = start[js][:, None] + np.arange(mj)[None, :]
        S = S_flat[idx]                                   # (J, mj)
        assert np.all(S[:, 0] == js)
        k = S[:, :, None] * n + S[:, None, :]              # keys of A[S, S]
        pos = np.searchsorted(keys, k)
        hit = (pos < len(keys)) & (keys[np.minimum(pos, len(keys) - 1)] == k)
        Ab = np.where(hit, v[np.minimum(pos, len(v) - 1)], 0.0)
        e = np.zeros((len(js), mj))
        e[:, 0] = 1.0
        y = np.linalg.solve(Ab, e[:, :, None])[:, :, 0]
        g = y / np.sqrt(y[:, :1])
        # G(i, j) for i in S_j sits at CSR position (i, j)
        G[np.searchsorted(keys, S * n + js[:, None])] = g
    tpos = np.searchsorted(keys, ci * n + rows)  # position of the transposed entry
    return G, G[tpos]


def compare_fsai(systems, net, seeds, rtol=1e-8, maxit=20000):
    """Per held-out seed: the network's factor and the FSAI factor on the same
    pattern, each in the same flexible PCG with M^-1 = G G^T applied in turn
    (anifem_pcg_factored): iterations, solve time (ms, synchronized wall clock,
    the network's factor included; the FSAI factor is built on the host and not
    timed, fsai_setup_ms gives it), true residuals."""
    import torch
    from . import pcg_factored
    val = torch.empty(systems.asm.nnz, dtype=torch.float64, device=systems.dev)
    rhs = torch.empty(systems.n, dtype=torch.float64, device=systems.dev)
    x = torch.zeros(systems.n, dtype=torch.float64, device=systems.dev)
    G = torch.empty_like(val)
    GT = torch.empty_like(val)
    rows = []
    for sd in seeds:
        systems.system(int(sd), val, rhs)
        t = time.perf_counter()
        Gf, GTf = fsai_factor(systems.row_ptr, systems.col_idx, val.cpu().numpy())
        t_setup = (time.perf_counter() - t) * 1e3
        G.copy_(torch.from_numpy(Gf))
        GT.copy_(torch.from_numpy(GTf))
        x.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        rf, okf = pcg_factored(systems.n, systems.rp_dev, systems.ci_dev, val, G, GT, rhs, x, rtol=rtol, maxit=maxit)
        torch.cuda.synchronize()
        tf = (time.perf_counter() - t) * 1e3
        x.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        net.factor(val, G, GT)
        rl, okl = pcg_factored(systems.n, systems.rp_dev, systems.ci_dev, val, G, GT, rhs, x, rtol=rtol, maxit=maxit)
        torch.cuda.synchronize()
        tl = (time.perf_counter() - t) * 1e3
        rows.append(dict(seed=int(sd), fsai_it=rf.iterations, fsai_ms=tf, fsai_true=rf.true_relres,
                         fsai_setup_ms=t_setup, learned_it=rl.iterations, learned_ms=tl,
                         learned_true=rl.true_relres, ok=bool(okf and okl)))
    return rows
