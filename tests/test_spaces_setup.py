"""Global numbering of the spaces beyond P1..P3 (RT0, ND0, CR1, B2, B4, P0)
on shuffled meshes, host setup (no device use).

Reference semantics: the DOF maps of spaces/*.h (one DOF per face for RT0 and
CR1, per edge for ND0 and B2, per cell for B4 and P0), NATURAL order, and the
orientation signs of RT0 / ND0 DOFs (assembler.inl:154-164 with
isPositivePermutationOrient, ordering.inl:152-172).  The signs are pinned by
what they are for: with them the global basis is conforming.  Evaluated with
the oracle's basis (pinned by the compiled reference's fixtures) from every
cell that shares a face or an edge, the normal component of an RT0 basis
function across its face and the tangential component of an ND0 basis
function along its edge agree; dropping the signs breaks both."""
import numpy as np
import pytest

import anifem_amd as af
import anifem_testlib as T

RT0, ND0, CR1, B2, B4, P0 = af.FEM_RT0, af.FEM_ND0, af.FEM_CR1, af.FEM_B2, af.FEM_B4, af.FEM_P0
EDGES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
FACES = [(0, 1, 2), (1, 2, 3), (2, 3, 0), (3, 0, 1)]


def _mesh(n=3, seed=4):
    coords, tets = af.kuhn_cube(n, 0.2, seed)
    return T.shuffled_mesh(coords, tets, seed + 7)


def _entities(tets):
    e = np.sort(np.concatenate([tets[:, [a, b]] for a, b in EDGES]), axis=1)
    f = np.sort(np.concatenate([tets[:, list(k)] for k in FACES]), axis=1)
    return len(np.unique(e, axis=0)), len(np.unique(f, axis=0))


@pytest.mark.parametrize("fem,where", [(RT0, "face"), (ND0, "edge"), (CR1, "face"), (B2, "edge"), (B4, "cell"),
                                       (P0, "cell")])
def test_counts_and_numbering(fem, where):
    coords, tets = _mesh()
    A = af.Assembler(coords, tets, af.Op(af.IDEN, fem), host_only=True)
    ne, nfc = _entities(tets)
    assert A.ndof == {"face": nfc, "edge": ne, "cell": len(tets)}[where]
    s = A.setup()
    dofs = s["dofs"]
    # every DOF is used; a cell DOF is the cell's own number (NATURAL order)
    assert np.array_equal(np.unique(dofs), np.arange(A.ndof))
    if where == "cell":
        np.testing.assert_array_equal(dofs[:, 0], np.arange(len(tets)))
    signs = A.signs()
    assert np.all(np.abs(signs) == 1)
    if fem not in (RT0, ND0):
        assert np.all(signs == 1)


def _eval_local(fem, op, coords, rt, X):
    """Local basis function k of cell t at X[t, k] (n, nf, 3): (n, nf, d)."""
    n, nf = X.shape[0], X.shape[1]
    xyz = np.repeat(coords[rt].reshape(n, 12), nf, axis=0)
    dofs = np.tile(np.eye(nf), (n, 1))
    return T.orc_applyx(op, fem, 1, X.reshape(n * nf, 1, 3), xyz, dofs)[:, 0, :].reshape(n, nf, -1)


def _spread(vals, dofs, ndof):
    """Largest disagreement between the values the cells give each global DOF."""
    ref = np.full(ndof, np.nan)
    ref[dofs.ravel()] = vals.ravel()
    return float(np.abs(vals.ravel() - ref[dofs.ravel()]).max())


def _rt0_normal_flux(coords, rt, signs):
    """n_f . (sign x phi_k) at the centroid of face k of every cell, n_f the
    unit normal of the face fixed by its sorted global vertex ids."""
    P = coords[rt]
    X = np.stack([P[:, list(k)].mean(axis=1) for k in FACES], axis=1)
    v = _eval_local(RT0, af.IDEN, coords, rt, X)
    ids = np.sort(np.stack([rt[:, list(k)] for k in FACES], axis=1), axis=2)
    a, b, c = coords[ids[..., 0]], coords[ids[..., 1]], coords[ids[..., 2]]
    nrm = np.cross(b - a, c - a)
    nrm /= np.linalg.norm(nrm, axis=2, keepdims=True)
    return signs * np.einsum("tkc,tkc->tk", nrm, v)


def _nd0_tangent(coords, rt, signs):
    """t_e . (sign x phi_k) at the midpoint of edge k of every cell, t_e the
    unit tangent from the lower to the higher global vertex id."""
    P = coords[rt]
    X = np.stack([(P[:, a] + P[:, b]) / 2 for a, b in EDGES], axis=1)
    v = _eval_local(ND0, af.IDEN, coords, rt, X)
    lo = np.stack([np.minimum(rt[:, a], rt[:, b]) for a, b in EDGES], axis=1)
    hi = np.stack([np.maximum(rt[:, a], rt[:, b]) for a, b in EDGES], axis=1)
    tg = coords[hi] - coords[lo]
    tg /= np.linalg.norm(tg, axis=2, keepdims=True)
    return signs * np.einsum("tkc,tkc->tk", tg, v)


def test_rt0_normal_component_is_continuous():
    coords, tets = _mesh()
    A = af.Assembler(coords, tets, af.Op(af.IDEN, RT0), host_only=True)
    rt, dofs, signs = A.tets(), A.setup()["dofs"], A.signs()
    flux = _rt0_normal_flux(coords, rt, signs)
    assert np.abs(flux).min() > 1e-3
    assert _spread(flux, dofs, A.ndof) < 1e-12 * np.abs(flux).max()
    # planted: no signs (every cell's outward-oriented local function)
    assert _spread(_rt0_normal_flux(coords, rt, np.ones_like(signs)), dofs, A.ndof) > 1e-2


def test_nd0_tangential_component_is_continuous():
    coords, tets = _mesh()
    A = af.Assembler(coords, tets, af.Op(af.IDEN, ND0), host_only=True)
    rt, dofs, signs = A.tets(), A.setup()["dofs"], A.signs()
    tan = _nd0_tangent(coords, rt, signs)
    assert np.abs(tan).min() > 1e-3
    assert _spread(tan, dofs, A.ndof) < 1e-12 * np.abs(tan).max()
    assert _spread(_nd0_tangent(coords, rt, np.ones_like(signs)), dofs, A.ndof) > 1e-2
    # planted: the reversed edge rule (larger id first) flips every global
    # function: still conforming, but against the tangent fixed by the ids
    assert np.all(np.sign(_nd0_tangent(coords, rt, -signs)) == -np.sign(tan))


def test_cr1_and_b2_values_are_continuous():
    coords, tets = _mesh()
    for fem, pts in [(CR1, FACES), (B2, EDGES)]:
        A = af.Assembler(coords, tets, af.Op(af.IDEN, fem), host_only=True)
        rt, dofs = A.tets(), A.setup()["dofs"]
        P = coords[rt]
        X = np.stack([P[:, list(k)].mean(axis=1) for k in pts], axis=1)
        v = _eval_local(fem, af.IDEN, coords, rt, X)[:, :, 0]
        assert np.abs(v).min() > 1e-3
        assert _spread(v, dofs, A.ndof) < 1e-12


def _alone(coords, tets, fem, nc=1, dirichlet=af.DIRICHLET_NONE):
    A = af.Assembler(coords, tets, af.Op(af.IDEN, fem, nc), host_only=True, dirichlet=dirichlet)
    return A, A.setup()


@pytest.mark.parametrize("vars_", [[(RT0, 1), (P0, 1)], [(af.FEM_P2, 3), (af.FEM_P1, 1)],
                                   [(ND0, 1), (af.FEM_P1, 1), (B4, 1)]])
def test_multi_variable_numbering(vars_):
    """Variables in NATURAL order VAR, DIM, ELEM_TYPE, ELEM_ID, DOF_ID
    (global_enumerator.h:396): each variable numbered as alone, shifted past
    the DOFs of the variables before it; a tet's local DOFs concatenate the
    variables' local DOFs; signs follow each variable's space; the boundary
    flags are each variable's own; the pattern couples every pair of DOFs
    that share a tet."""
    coords, tets = _mesh()
    for dirichlet in (af.DIRICHLET_NONE, af.DIRICHLET_BOUNDARY):
        M = af.Assembler(coords, tets, variables=[af.Op(af.IDEN, f, c) for f, c in vars_], host_only=True,
                         dirichlet=dirichlet)
        s = M.setup()
        sg = M.signs()
        base, lf = 0, 0
        for f, c in vars_:
            A, sa = _alone(coords, tets, f, c, dirichlet)
            nf = A.nf
            np.testing.assert_array_equal(s["dofs"][:, lf:lf + nf], sa["dofs"] + base)
            np.testing.assert_array_equal(sg[:, lf:lf + nf], A.signs())
            np.testing.assert_array_equal(s["dirichlet"][base:base + A.ndof], sa["dirichlet"])
            base += A.ndof
            lf += nf
        assert M.ndof == base and M.nf == lf
        # pattern: sorted union of the tets' DOFs per row, Dirichlet rows and
        # columns reduced to the diagonal (assembler.inl:114-136)
        dofs, isd = s["dofs"], s["dirichlet"].astype(bool)
        rows = [set() for _ in range(M.ndof)]
        for t in range(len(dofs)):
            for r in dofs[t]:
                rows[r].add(r)
                if isd[r]:
                    continue
                rows[r].update(int(c) for c in dofs[t] if not isd[c])
        rp, ci = s["row_ptr"], s["col_idx"]
        for r in range(M.ndof):
            assert list(ci[rp[r]:rp[r + 1]]) == sorted(rows[r])


def test_multi_variable_refusals():
    coords, tets = _mesh()
    with pytest.raises(af.AnifemError):  # vector variables are FemVec<n, P1..P3> only
        af.Assembler(coords, tets, variables=[af.Op(af.IDEN, RT0, 3)], host_only=True)
    with pytest.raises(af.AnifemError):  # 64 local DOFs at most
        af.Assembler(coords, tets, variables=[af.Op(af.IDEN, af.FEM_P3, 3), af.Op(af.IDEN, af.FEM_P2, 1)],
                     host_only=True)
