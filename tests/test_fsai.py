"""The FSAI factor (anifem_amd.learned.fsai_factor), the non-learned factor on
the learned preconditioner's pattern: on a system assembled by the oracle,
G is lower triangular on A's pattern, G^T A G has a unit diagonal, each
column g_j makes A g_j vanish on the rest of its pattern, and a planted wrong
column fails that check."""
import numpy as np

import anifem_amd as af
import anifem_testlib as T
from anifem_amd.learned import fsai_factor


def _system(n=4, seed=3):
    coords, tets = af.kuhn_cube(n, 0.2, seed)
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P1), dirichlet=af.DIRICHLET_BOUNDARY, host_only=True)
    s = Asm.setup()
    S = T.oracle_assembly(coords, tets, 2, 1, dirichlet_mask=s["dirichlet"])
    xyz = coords[S["tets"]].reshape(len(tets), 12)
    kap = np.exp(np.random.default_rng(seed).normal(0, 1.5, (len(tets), 1)))
    A = T.orc_elem(2, 2, 1, 2, 2, 1, 2, 1, 2, xyz, kap)
    A, _ = T.apply_dir_elementwise(A, np.zeros((len(tets), 4)), S["dofs"], s["dirichlet"])
    rc, val, _ = T.oracle_scatter(len(tets), 4, S["dofs"], S["slot"], A, None, len(S["col_idx"]), S["ndof"])
    assert rc == 0
    return S["row_ptr"], S["col_idx"], val, S["ndof"]


def _dense(rp, ci, v, n):
    M = np.zeros((n, n))
    for i in range(n):
        M[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    return M


def _violation(Ad, Gd):
    """Largest |(A g_j)_i| over S_j = {i > j : a_ij != 0} (the lower pattern
    of A), and the largest deviation of diag(G^T A G) from 1."""
    AG = Ad @ Gd
    S = np.tril(Ad != 0, -1)
    off = np.where(S, np.abs(AG), 0.0).max()
    return off, np.abs(np.diag(Gd.T @ Ad @ Gd) - 1).max()


def test_fsai_defining_properties():
    rp, ci, val, n = _system()
    G, GT = fsai_factor(rp, ci, val)
    Ad = _dense(rp, ci, val, n)
    Gd = _dense(rp, ci, G, n)
    assert np.all(np.triu(Gd, 1) == 0)
    np.testing.assert_array_equal(_dense(rp, ci, GT, n), Gd.T)
    off, diag = _violation(Ad, Gd)
    assert off < 1e-12 * np.abs(Ad).max() and diag < 1e-12
    # a better preconditioner than Jacobi: the spread of the spectrum of G^T A G
    ev = np.linalg.eigvalsh(Gd.T @ Ad @ Gd)
    dj = 1 / np.sqrt(np.diag(Ad))
    evj = np.linalg.eigvalsh(dj[:, None] * Ad * dj[None, :])
    assert ev.max() / ev.min() < evj.max() / evj.min()
    # planted: one column replaced by the Jacobi column fails the check
    Gw = Gd.copy()
    j = int(np.argmax((Gd != 0).sum(axis=0)))
    Gw[:, j] = 0
    Gw[j, j] = 1 / np.sqrt(Ad[j, j])
    assert _violation(Ad, Gw)[0] > 1e-6 * np.abs(Ad).max()
