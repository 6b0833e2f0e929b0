"""CPU restatement (float64 numpy) of the learned preconditioner's network,
loss and reverse-mode gradient.  Test infrastructure: the GPU kernels of
inmost-fem_amd/csrc/gnn.hip are checked against it, and it is checked against
central finite differences.  The network has no reference counterpart (the
reference solves with INMOST's solvers, examples/tutorials/ex1.cpp:161-163).

Graph: the CSR pattern of an assembled SPD matrix A (symmetric pattern).
  dinv_r  = 1/sqrt(a_rr)
  at_s    = a_s dinv_r dinv_c for an off-diagonal slot s = (r, c), 0 on the diagonal
  x_r     = [log a_rr - mean_{c in N(r)} log a_cc,  sum_{off s in row r} at_s,
             (row r has no off-diagonal),  1]
Network (H hidden units, 3 message-passing layers):
  h0 = relu(W_in x + b_in)
  m^l = At h^{l-1};  z^l = W_s^l h^{l-1} + W_m^l m^l + b^l;  h^l = relu(z^l) + h^{l-1}
  lower slot s = (r, c), r > c:  L_s = at_s (w_o . (h_r * h_c) + c_o)   (0 elsewhere)
  d_r = w_d . h_r + c_d;  Dh_r = exp(d_r) dinv_r
  G_s = Dh_r (delta_rc + L_s)   (lower triangular on the pattern of A)
Preconditioner: M^-1 = G G^T (identity network output = Jacobi).
Loss over probe vectors z_p:  (1/(P n)) sum_p || G^T A G z_p - z_p ||^2.
"""
import numpy as np

H = 32
F_IN = 4
N_LAYERS = 3
PARAM_SHAPES = [("W_in", (H, F_IN)), ("b_in", (H,))] + \
    [(f"{k}{l}", s) for l in range(N_LAYERS) for k, s in (("W_s", (H, H)), ("W_m", (H, H)), ("b", (H,)))] + \
    [("w_o", (H,)), ("c_o", (1,)), ("w_d", (H,)), ("c_d", (1,))]


def n_params():
    return sum(int(np.prod(s)) for _, s in PARAM_SHAPES)


def unpack(theta):
    out, o = {}, 0
    for k, s in PARAM_SHAPES:
        c = int(np.prod(s))
        out[k] = theta[o:o + c].reshape(s)
        o += c
    return out


def pack(p):
    return np.concatenate([np.asarray(p[k], dtype=np.float64).ravel() for k, _ in PARAM_SHAPES])


def init_params(seed):
    """Small random weights; output heads at zero (the network starts as Jacobi)."""
    rng = np.random.default_rng(seed)
    p = {}
    for k, s in PARAM_SHAPES:
        if k.startswith("W"):
            p[k] = rng.normal(0, 1.0 / np.sqrt(s[1]), s) * (0.5 if k.startswith("W_m") else 1.0)
        else:
            p[k] = np.zeros(s)
    return pack(p)


class Graph:
    def __init__(self, row_ptr, col_idx, a):
        self.rp = np.asarray(row_ptr, np.int64)
        self.ci = np.asarray(col_idx, np.int64)
        self.a = np.asarray(a, np.float64)
        n = len(self.rp) - 1
        self.n = n
        self.row = np.repeat(np.arange(n), np.diff(self.rp))
        diag = np.zeros(n)
        isd = self.row == self.ci
        diag[self.row[isd]] = self.a[isd]
        self.dinv = 1.0 / np.sqrt(diag)
        self.at = np.where(isd, 0.0, self.a * self.dinv[self.row] * self.dinv[self.ci])
        self.lower = self.row > self.ci
        self.isd = isd
        # transposed slot of each slot (the pattern is symmetric)
        key = self.row * n + self.ci
        tkey = self.ci * n + self.row
        order = np.argsort(key)
        self.tslot = order[np.searchsorted(key[order], tkey)]
        logd = np.log(diag)
        off = ~isd
        deg = np.bincount(self.row[off], minlength=n)
        s_nb = np.bincount(self.row[off], weights=logd[self.ci[off]], minlength=n)
        mean_nb = np.where(deg > 0, s_nb / np.maximum(deg, 1), logd)
        self.x = np.stack([logd - mean_nb, np.bincount(self.row, weights=self.at, minlength=n),
                           (deg == 0).astype(float), np.ones(n)], axis=1)

    def spmm_at(self, h):
        """At h (at symmetric)."""
        out = np.zeros((self.n,) + h.shape[1:])
        np.add.at(out, self.row, self.at[:, None] * h[self.ci] if h.ndim == 2 else self.at * h[self.ci])
        return out

    def spmv(self, vals, v):
        out = np.zeros((self.n,) + v.shape[1:])
        np.add.at(out, self.row, vals.reshape((-1,) + (1,) * (v.ndim - 1)) * v[self.ci])
        return out

    def spmv_t(self, vals, v):
        """(M^T v) for M with values `vals` on the pattern."""
        return self.spmv(vals[self.tslot], v)


def relu(x):
    return np.maximum(x, 0.0)


def forward(theta, g):
    p = unpack(theta)
    cache = {}
    z0 = g.x @ p["W_in"].T + p["b_in"]
    h = relu(z0)
    cache["z0"], cache["h"] = z0, [h]
    cache["m"], cache["z"] = [], []
    for l in range(N_LAYERS):
        m = g.spmm_at(h)
        z = h @ p[f"W_s{l}"].T + m @ p[f"W_m{l}"].T + p[f"b{l}"]
        h = relu(z) + h
        cache["m"].append(m)
        cache["z"].append(z)
        cache["h"].append(h)
    hr, hc = h[g.row], h[g.ci]
    t = (hr * hc) @ p["w_o"] + p["c_o"][0]
    L = np.where(g.lower, g.at * t, 0.0)
    d = h @ p["w_d"] + p["c_d"][0]
    Dh = np.exp(d) * g.dinv
    G = Dh[g.row] * (g.isd.astype(float) + L)
    cache.update(t=t, L=L, d=d, Dh=Dh, G=G)
    return G, cache


def apply_precond(G, g, r):
    return g.spmv(G, g.spmv_t(G, r))


def loss(theta, g, Z):
    G, _ = forward(theta, g)
    y = g.spmv(G, Z)
    w = g.spmv(g.a, y)
    u = g.spmv_t(G, w)
    res = u - Z
    return float(np.sum(res * res) / (Z.shape[1] * g.n))


def grad(theta, g, Z):
    """Reverse-mode gradient of loss(theta, g, Z) with respect to theta."""
    p = unpack(theta)
    G, c = forward(theta, g)
    P = Z.shape[1]
    y = g.spmv(G, Z)
    w = g.spmv(g.a, y)
    u = g.spmv_t(G, w)
    res = u - Z
    gu = 2.0 * res / (P * g.n)
    gw = g.spmv(G, gu)
    gy = g.spmv(g.a, gw)
    # d/dG_rc: u_c = sum_r G_rc w_r ; y_r = sum_c G_rc z_c
    dG = np.sum(w[g.row] * gu[g.ci] + gy[g.row] * Z[g.ci], axis=1)
    dDh = np.bincount(g.row, weights=dG * (g.isd.astype(float) + c["L"]), minlength=g.n)
    dL = np.where(g.lower, dG * c["Dh"][g.row], 0.0)
    dd = dDh * c["Dh"]
    h = c["h"][-1]
    gp = {k: np.zeros(s) for k, s in PARAM_SHAPES}
    dt = dL * g.at
    hr, hc = h[g.row], h[g.ci]
    gp["w_o"] = (dt[:, None] * hr * hc).sum(0)
    gp["c_o"] = np.array([dt.sum()])
    dh = np.zeros_like(h)
    np.add.at(dh, g.row, dt[:, None] * p["w_o"] * hc)
    np.add.at(dh, g.ci, dt[:, None] * p["w_o"] * hr)
    gp["w_d"] = dd @ h
    gp["c_d"] = np.array([dd.sum()])
    dh += dd[:, None] * p["w_d"]
    for l in reversed(range(N_LAYERS)):
        z, m, hp = c["z"][l], c["m"][l], c["h"][l]
        dz = dh * (z > 0)
        gp[f"W_s{l}"] = dz.T @ hp
        gp[f"W_m{l}"] = dz.T @ m
        gp[f"b{l}"] = dz.sum(0)
        dm = dz @ p[f"W_m{l}"]
        dh = dh + dz @ p[f"W_s{l}"] + g.spmm_at(dm)
    dz0 = dh * (c["z0"] > 0)
    gp["W_in"] = dz0.T @ g.x
    gp["b_in"] = dz0.sum(0)
    return pack(gp)
