"""The node reorder and the P3 edge-DOF orientation, pinned independently of
the rule the oracle and the product both restate.

Reference semantics:
  * reorderNodesOnTetrahedron (ordering.inl:8-26): a cell whose
    det(P0-P3, P1-P3, P2-P3) is negative has its local nodes 2 and 3 swapped,
    so every reordered cell is positively oriented;
  * createOrderPermutation + index_on_reorderd_elem (ordering.inl:137-151,
    tetdofmap.inl:98-105): the two DOFs on an edge are numbered by the global
    node ids of its ends, the first global DOF of the pair belonging to the
    end with the smaller id.

The checks below use only geometry and the basis: the location of a local DOF
is read off the oracle's basis functions (themselves pinned by the compiled
reference's fem3DapplyL fixtures), so a numbering is conforming exactly when
every cell that sees a global DOF puts it at the same physical point.  Each
check is shown to fail on a planted mistake."""
import numpy as np

import anifem_amd as af
import anifem_testlib as T

P3 = 4
EDGES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def _mesh(n=3, seed=5):
    coords, tets = af.kuhn_cube(n, 0.2, seed)
    return T.shuffled_mesh(coords, tets, seed + 100)


def _det(coords, tets):
    P = coords[tets]
    return np.linalg.det(np.stack([P[:, 0] - P[:, 3], P[:, 1] - P[:, 3], P[:, 2] - P[:, 3]], axis=1))


def _reorder_ok(coords, tets_in, tets_out):
    """Every output cell positively oriented; a cell is either unchanged or
    has exactly its local nodes 2 and 3 swapped, the swap made exactly when
    the input cell was negatively oriented."""
    d_in = _det(coords, tets_in)
    d_out = _det(coords, tets_out)
    if not np.all(d_out > 0):
        return False
    swapped = tets_in[:, [0, 1, 3, 2]]
    same = np.all(tets_out == tets_in, axis=1)
    sw = np.all(tets_out == swapped, axis=1)
    return bool(np.all(same | sw) and np.all(sw == (d_in < 0)))


def _conforming(pos, dofs, ndof):
    """Max spread of the physical positions that the cells give each global DOF."""
    flat_d = dofs.ravel()
    flat_p = pos.reshape(-1, 3)
    ref = np.full((ndof, 3), np.nan)
    ref[flat_d] = flat_p  # any one cell's view
    return float(np.abs(flat_p - ref[flat_d]).max())


def _pair_follows_global_ids(coords, rtets, dofs, pos):
    """For every edge of every cell: the global DOF with the smaller number of
    the pair sits at (2 X_a + X_b)/3, a the end with the smaller global id."""
    worst = 0.0
    for e, (I, J) in enumerate(EDGES):
        d = dofs[:, 4 + 2 * e:6 + 2 * e]
        p = pos[:, 4 + 2 * e:6 + 2 * e]
        first = np.where(d[:, 0] < d[:, 1], 0, 1)
        p_first = p[np.arange(len(d)), first]
        a = np.minimum(rtets[:, I], rtets[:, J])
        b = np.maximum(rtets[:, I], rtets[:, J])
        expect = (2 * coords[a] + coords[b]) / 3
        worst = max(worst, float(np.abs(p_first - expect).max()))
        # the pair is two consecutive global numbers
        assert np.all(np.abs(d[:, 0] - d[:, 1]) == 1)
    return worst


def _oracle_setup(coords, tets):
    S = T.oracle_assembly(coords, tets, P3, 1, with_slots=False)
    return S["tets"], S["dofs"], S["ndof"]


def test_local_dof_locations_are_the_p3_lattice():
    """The basis puts vertex DOFs on the vertices, DOFs 4+2e and 5+2e of edge
    (I, J) at (2I+J)/3 and (I+2J)/3, and face DOF 16+k at the centroid of a
    face, as read off the basis functions."""
    L = T.lagrange_nodes(P3)
    np.testing.assert_allclose(L[:4], np.eye(4), atol=1e-15)
    for e, (I, J) in enumerate(EDGES):
        a = np.zeros(4)
        a[I], a[J] = 2 / 3, 1 / 3
        b = np.zeros(4)
        b[I], b[J] = 1 / 3, 2 / 3
        np.testing.assert_allclose(L[4 + 2 * e], a, atol=1e-15)
        np.testing.assert_allclose(L[5 + 2 * e], b, atol=1e-15)
    for k in range(4):
        assert np.count_nonzero(np.abs(L[16 + k] - 1 / 3) < 1e-15) == 3 and abs(L[16 + k].sum() - 1) < 1e-15


def test_reorder_gives_positive_cells():
    coords, tets = _mesh()
    rt, _, _ = _oracle_setup(coords, tets)
    assert np.any(_det(coords, tets) < 0), "the shuffled mesh holds negatively oriented cells"
    assert _reorder_ok(coords, tets, rt)
    # the product's host setup reorders the same way
    A = af.Assembler(coords, tets, af.Op(af.GRAD, P3), host_only=True)
    np.testing.assert_array_equal(A.tets(), rt)


def test_reorder_check_catches_an_inverted_sign():
    coords, tets = _mesh()
    d = _det(coords, tets)
    wrong = tets.copy()
    flip = d > 0  # planted: swap when the cell is positive
    wrong[flip] = wrong[flip][:, [0, 1, 3, 2]]
    assert not _reorder_ok(coords, tets, wrong)


def test_p3_numbering_is_conforming_and_ordered_by_global_ids():
    coords, tets = _mesh()
    rt, dofs, ndof = _oracle_setup(coords, tets)
    pos = T.dof_positions(coords, rt, dofs, T.lagrange_nodes(P3))
    assert _conforming(pos, dofs, ndof) < 1e-14
    assert _pair_follows_global_ids(coords, rt, dofs, pos) < 1e-14
    # the product's setup (host phase) numbers the DOFs the same way
    A = af.Assembler(coords, tets, af.Op(af.GRAD, P3), host_only=True)
    np.testing.assert_array_equal(A.setup()["dofs"], dofs)


def test_planted_orientation_rules_fail():
    coords, tets = _mesh()
    rt, dofs, ndof = _oracle_setup(coords, tets)
    L = T.lagrange_nodes(P3)
    pos = T.dof_positions(coords, rt, dofs, L)
    # planted 1: ignore the global ids (every cell keeps its local pair order)
    ign = dofs.copy()
    for e in range(6):
        lo = np.minimum(dofs[:, 4 + 2 * e], dofs[:, 5 + 2 * e])
        ign[:, 4 + 2 * e], ign[:, 5 + 2 * e] = lo, lo + 1
    assert _conforming(pos, ign, ndof) > 1e-2
    # planted 2: the reversed rule (larger id first) is conforming, but the
    # pair no longer follows the global ids
    rev = dofs.copy()
    for e in range(6):
        rev[:, 4 + 2 * e], rev[:, 5 + 2 * e] = dofs[:, 5 + 2 * e], dofs[:, 4 + 2 * e]
    assert _conforming(pos, rev, ndof) < 1e-14
    assert _pair_follows_global_ids(coords, rt, rev, pos) > 1e-2


def _cubic(c):
    f = lambda x: (c[0] + c[1] * x[..., 0] + c[2] * x[..., 1] ** 2 + c[3] * x[..., 0] * x[..., 1] * x[..., 2] +
                   c[4] * x[..., 2] ** 3 + c[5] * x[..., 0] ** 2 * x[..., 1])
    g = lambda x: np.stack([c[1] + c[3] * x[..., 1] * x[..., 2] + 2 * c[5] * x[..., 0] * x[..., 1],
                            2 * c[2] * x[..., 1] + c[3] * x[..., 0] * x[..., 2] + c[5] * x[..., 0] ** 2,
                            c[3] * x[..., 0] * x[..., 1] + 3 * c[4] * x[..., 2] ** 2], axis=-1)
    return f, g


def cubic_global_vector(coords, rt, dofs, ndof, f):
    pos = T.dof_positions(coords, rt, dofs, T.lagrange_nodes(P3))
    u = np.full(ndof, np.nan)
    u[dofs.ravel()] = f(pos.reshape(-1, 3))
    assert not np.isnan(u).any()
    return u


def test_oracle_reproduces_cubics_through_the_numbering():
    """A cubic interpolated at the global DOF points and gathered back per cell
    through the numbering is reproduced by the oracle's fem3DapplyL."""
    coords, tets = _mesh()
    rt, dofs, ndof = _oracle_setup(coords, tets)
    rng = np.random.default_rng(3)
    f, g = _cubic(rng.normal(size=6))
    u = cubic_global_vector(coords, rt, dofs, ndof, f)
    xyl = rng.dirichlet(np.ones(4), 5)
    xyz = coords[rt].reshape(len(rt), 12)
    X = np.einsum("qv,tvc->tqc", xyl, coords[rt])
    val = T.orc_apply(1, P3, 1, xyl, xyz, u[dofs])[:, :, 0]
    grad = T.orc_apply(2, P3, 1, xyl, xyz, u[dofs])
    assert np.abs(val - f(X)).max() <= 1e-12 * np.abs(f(X)).max()
    assert np.abs(grad - g(X)).max() <= 1e-11 * np.abs(g(X)).max()


def test_oracle_p3_energy_of_a_linear_field():
    """u^T K u = |grad u|^2 |Omega| for the P3 stiffness assembled by the
    oracle's sequential scatter, u = 2x - y + 3z at the DOF points."""
    coords, tets = _mesh(2, 7)
    S = T.oracle_assembly(coords, tets, P3, 1)
    rt, dofs = S["tets"], S["dofs"]
    xyz = coords[rt].reshape(len(rt), 12)
    A = T.orc_elem(2, P3, 1, 2, P3, 1, 1, 0, 4, xyz, None)  # GRAD.GRAD, D = I
    rc, val, _ = T.oracle_scatter(len(rt), 20, dofs, S["slot"], A, None, len(S["col_idx"]), S["ndof"])
    assert rc == 0
    u = cubic_global_vector(coords, rt, dofs, S["ndof"], lambda x: x @ np.array([2.0, -1.0, 3.0]))
    rp, ci = S["row_ptr"], S["col_idx"]
    Ku = np.zeros(S["ndof"])
    for i in range(S["ndof"]):
        Ku[i] = val[rp[i]:rp[i + 1]] @ u[ci[rp[i]:rp[i + 1]]]
    assert abs(u @ Ku - 14.0) < 1e-10
