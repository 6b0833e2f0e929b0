"""Host-side assembly setup of the product (enumeration, CSR pattern, slot map,
row blocks) against the oracle, bit for bit.  Runs on the CPU through
anifem_assembly_create_host (no device use)."""
import numpy as np
import pytest

import anifem_amd as af
import anifem_testlib as T


def _shuffled_mesh(n, seed):
    coords, tets = af.kuhn_cube(n, 0.2, seed)
    return T.shuffled_mesh(coords, tets, seed)


def _boundary_mask(A, coords):
    s = A.setup()
    return s["dirichlet"]


@pytest.mark.parametrize("fem", [2, 3, 4])
@pytest.mark.parametrize("ncomp", [1, 3])
@pytest.mark.parametrize("dirichlet", [0, 1])
@pytest.mark.parametrize("mesh", ["kuhn", "shuffled"])
def test_setup_matches_oracle(fem, ncomp, dirichlet, mesh):
    if mesh == "kuhn":
        coords, tets = af.kuhn_cube(3, 0.2, 5)
    else:
        coords, tets = _shuffled_mesh(3, 7)
    # a vector space with the Lame tensor couples its components (every
    # structural entry is kept, as the oracle's pattern)
    tt = af.TENSOR_LAME if ncomp == 3 else af.TENSOR_NULL
    A = af.Assembler(coords, tets, af.Op(af.GRAD, fem, ncomp), tensor_type=tt, dirichlet=dirichlet, host_only=True)
    s = A.setup()
    mask = s["dirichlet"] if dirichlet else None
    O = T.oracle_assembly(coords, tets, fem, ncomp, dirichlet_mask=mask)
    assert A.ndof == O["ndof"]
    np.testing.assert_array_equal(s["dofs"], O["dofs"])
    np.testing.assert_array_equal(s["row_ptr"], O["row_ptr"])
    np.testing.assert_array_equal(s["col_idx"], O["col_idx"])
    np.testing.assert_array_equal(s["slot"], O["slot"])


@pytest.mark.parametrize("op,tt", [(af.GRAD, af.TENSOR_SCALAR), (af.IDEN, af.TENSOR_NULL), (af.GRAD, af.TENSOR_LAME)])
@pytest.mark.parametrize("fem", [2, 3])
def test_vector_pattern_follows_the_value_rule(op, tt, fem):
    """The reference inserts an element entry only when |a| > 1e-100
    (assembler.inl:416-417) and always the diagonal (:114-136).  For a vector
    space with a NULL or SCALAR tensor and component-wise operators the
    entries between components are exactly zero in every cell, so the pattern
    must leave them out; the Lame tensor couples them.  The expected pattern is
    built from the oracle's element matrices of every cell by that rule (with
    the Dirichlet rows and columns removed as applyDir zeroes them).  Entries
    that vanish by a coincidence of the geometry (DUDX of a P2 basis function
    on a right-angled cell, say) stay in this build's pattern: it is fixed at
    set-up, before any value exists."""
    coords, tets = af.kuhn_cube(2, 0.2, 3)
    n = len(tets)
    A = af.Assembler(coords, tets, af.Op(op, fem, 3), tensor_type=tt, tensor_storage=af.COEF_PER_TET,
                     dirichlet=af.DIRICHLET_BOUNDARY, host_only=True)
    s = A.setup()
    rt = A.tets()
    xyz = coords[rt].reshape(n, 12)
    rng = np.random.default_rng(fem)
    D = (rng.uniform(0.5, 2.0, (n, 1)) if tt == af.TENSOR_SCALAR else
         np.stack([rng.uniform(1, 2, n), rng.uniform(0.5, 1, n)], axis=1) if tt == af.TENSOR_LAME else None)
    if tt == af.TENSOR_LAME:  # the oracle takes the full 9 x 9 tensor
        Dfull = np.zeros((n, 81))
        for r in range(n):
            lam, mu = D[r]
            C = np.zeros((3, 3, 3, 3))
            for i in range(3):
                for j in range(3):
                    for k in range(3):
                        for l in range(3):
                            C[i, j, k, l] = lam * (i == k) * (j == l) + mu * ((i == j) * (k == l) + (i == l) * (j == k))
            Dfull[r] = C.reshape(9, 9).ravel(order="F")
        Ae = T.orc_elem(op, fem, 3, op, fem, 3, af.TENSOR_GENERAL, 1, 2, xyz, Dfull)
    else:
        Ae = T.orc_elem(op, fem, 3, op, fem, 3, tt, 1 if D is not None else 0, 2, xyz, D)
    nf = A.nf
    dofs, isd = s["dofs"], s["dirichlet"].astype(bool)
    rows = [{r} for r in range(A.ndof)]
    for t in range(n):
        M = Ae[t].reshape(nf, nf, order="F")  # M[ib, ia]
        for ib in range(nf):
            for ia in range(nf):
                r, c = dofs[t, ib], dofs[t, ia]
                if abs(M[ib, ia]) > 1e-100 and not isd[r] and not isd[c]:
                    rows[r].add(int(c))
    rp, ci = s["row_ptr"], s["col_idx"]
    for r in range(A.ndof):
        assert list(ci[rp[r]:rp[r + 1]]) == sorted(rows[r]), r
    coupled = sum(len(x) for x in rows)
    assert A.nnz == coupled


def test_boundary_dirichlet_dofs_are_on_boundary():
    coords, tets = af.kuhn_cube(4, 0.2, 2)
    for fem in [2, 3, 4]:
        A = af.Assembler(coords, tets, af.Op(af.GRAD, fem), dirichlet=af.DIRICHLET_BOUNDARY, host_only=True)
        s = A.setup()
        onb = np.any((coords < 1e-12) | (coords > 1 - 1e-12), axis=1)
        # node DOFs: flagged exactly on the boundary nodes
        np.testing.assert_array_equal(s["dirichlet"][:len(coords)].astype(bool), onb)
        if fem >= 3:
            # an edge DOF is Dirichlet iff the edge midpoint lies on the boundary
            tets_r = T.oracle_assembly(coords, tets, fem, 1, with_slots=False)["tets"]
            for t in range(len(tets)):
                P = coords[tets_r[t]]
                for e, (a, b) in enumerate([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]):
                    mid = 0.5 * (P[a] + P[b])
                    mb = bool(np.any((mid < 1e-12) | (mid > 1 - 1e-12)))
                    d = s["dofs"][t, 4 + (e if fem == 3 else 2 * e)]
                    assert bool(s["dirichlet"][d]) == mb


@pytest.mark.parametrize("n,fem,ndof,nnz", [(16, 2, 4913, 66961), (16, 3, 35937, 977793), (16, 4, 117649, 5488081),
                                            (32, 2, 35937, 513313)])
def test_counts_match_mesh_tables(n, fem, ndof, nnz):
    coords, tets = af.kuhn_cube(n, 0.2, 1)
    A = af.Assembler(coords, tets, af.Op(af.GRAD, fem), host_only=True)
    assert (A.ndof, A.nnz) == (ndof, nnz)


def test_dirichlet_pattern_drop_counts():
    # C1: 46,941 stored entries after the Dirichlet row/column drop (T4)
    coords, tets = af.kuhn_cube(16, 0.2, 1)
    A = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P1), dirichlet=af.DIRICHLET_BOUNDARY, host_only=True)
    assert A.nnz == 46941


def test_connectivity_error():
    coords, tets = af.kuhn_cube(2, 0.2, 1)
    bad = tets.copy()
    bad[3, 1] = bad[3, 0]  # repeated node: broken cell (assembler.inl:309-312 returns -2)
    with pytest.raises(af.AnifemError) as e:
        af.Assembler(coords, bad, af.Op(af.GRAD, af.FEM_P1), host_only=True)
    assert e.value.code == af.ERR_CONNECTIVITY


def test_mesh_generator_properties():
    n = 8
    coords, tets = af.kuhn_cube(n, 0.2, 9)
    assert coords.shape == ((n + 1) ** 3, 3) and tets.shape == (6 * n ** 3, 4)
    grid = np.stack(np.meshgrid(*[np.arange(n + 1)] * 3, indexing="ij"), -1).reshape(-1, 3) / n
    d = coords - grid
    assert np.abs(d).max() <= 0.2 / n + 1e-15
    onb = np.any((grid == 0) | (grid == 1), axis=1)
    assert np.abs(d[onb]).max() == 0.0
    # Kuhn split with jitter 0.2h: every tet keeps positive volume after reorder
    vol = 0.0
    for t in tets:
        P = coords[t]
        v = abs(np.linalg.det(np.array([P[1] - P[0], P[2] - P[0], P[3] - P[0]]))) / 6
        assert v > 0
        vol += v
    assert abs(vol - 1.0) < 1e-12
    c2, _ = af.kuhn_cube(n, 0.2, 9)
    np.testing.assert_array_equal(coords, c2)  # seeded reproducibility


def _emulate_blocks(A, F, s, B, nf):
    """The row-block kernel replayed on the host from the block lists: each
    pair stores its element-matrix row and load-vector entry at the listed
    positions, then each slot (row) adds its entries in list order."""
    rp = s["row_ptr"]
    ndof, nnz = len(rp) - 1, rp[-1]
    val, rhs = np.zeros(nnz), np.zeros(ndof)
    nb = len(B["blk_row0"]) - 1
    for b in range(nb):
        r0, r1 = B["blk_row0"][b], B["blk_row0"][b + 1]
        p0, p1 = B["blk_p0"][b], B["blk_p0"][b + 1]
        nv = B["blk_v0"][b + 1] - B["blk_v0"][b]
        vals = np.full(nv, np.nan)
        vtet = np.full(nv, -1)
        fvals = np.zeros(p1 - p0)
        for p in range(p0, p1):
            t, ibf = B["pair_tet"][p], int(B["pair_ib"][p])
            ib, dirrow = ibf & 63, bool(ibf & 0x80)
            for ia in range(nf):
                ps = int(B["pair_pos"][p, ia])
                if ps == 0xFFFF:
                    continue
                assert np.isnan(vals[ps]), "each list position is written once"
                if dirrow:
                    assert ia == ib, "a kept entry of a Dirichlet row is its diagonal"
                vals[ps] = 1.0 if dirrow else A[t, ib + nf * ia]
                vtet[ps] = t
            fvals[p - p0] = 0.0 if dirrow else F[t, ib]
        assert not np.isnan(vals).any(), "every list position is written"
        s0, s1 = rp[r0], rp[r1]
        for u in range(s0, s1):
            lo = B["sstart"][u]
            hi = B["sstart"][u + 1] if u + 1 < s1 else nv
            assert np.all(np.diff(vtet[lo:hi]) > 0), "a slot's entries in ascending tet order"
            acc = 0.0
            for x in vals[lo:hi]:
                acc += x
            val[u] = acc
        for r in range(r0, r1):
            lo = B["pstart"][r]
            hi = B["pstart"][r + 1] if r + 1 < r1 else p1 - p0
            acc = 0.0
            for x in fvals[lo:hi]:
                acc += x
            rhs[r] = acc
    return val, rhs


@pytest.mark.parametrize("fem,ncomp", [(2, 1), (3, 1), (4, 1), (3, 3)])
def test_row_blocks_reproduce_oracle_scatter(fem, ncomp):
    """The block lists place every kept (tet, entry) contribution of every slot
    once, in ascending tet order: replaying them gives the oracle's scatter bit
    for bit (random element matrices with element-level Dirichlet)."""
    coords, tets = _shuffled_mesh(3, 11) if fem != 3 else af.kuhn_cube(4, 0.2, 2)
    tt = af.TENSOR_LAME if ncomp == 3 else af.TENSOR_NULL  # components coupled: every structural entry kept
    A = af.Assembler(coords, tets, af.Op(af.GRAD, fem, ncomp), tensor_type=tt, dirichlet=af.DIRICHLET_BOUNDARY,
                     host_only=True)
    s = A.setup()
    B = A.blocks()
    nf, E = A.nf, A.nf * A.nf
    rng = np.random.default_rng(fem * 10 + ncomp)
    Ae = rng.standard_normal((len(tets), E))
    Fe = rng.standard_normal((len(tets), nf))
    Ae, Fe = T.apply_dir_elementwise(Ae, Fe, s["dofs"], s["dirichlet"])
    rc, val_o, rhs_o = T.oracle_scatter(len(tets), nf, s["dofs"], s["slot"], Ae, Fe, A.nnz, A.ndof)
    assert rc == 0
    val, rhs = _emulate_blocks(Ae, Fe, s, B, nf)
    np.testing.assert_array_equal(val, val_o)
    np.testing.assert_array_equal(rhs, rhs_o)
    # budgets: rows contiguous and covering, slot / value / pair limits
    r0 = B["blk_row0"]
    assert r0[0] == 0 and r0[-1] == A.ndof and np.all(np.diff(r0) > 0)
    # default budgets by entries per pair: (slots, values, pairs) = (2048,
    # 2048, 256) up to nf = 10, (4096, 2048, 128) for scalar spaces up to 20
    # (4096, 3072, 256) for vector ones, (4096, 8192, 256) above
    if nf <= 10:
        max_slots, max_values, max_pairs = 2048, 2048, 256
    elif nf <= 20:
        max_slots, max_values, max_pairs = (4096, 2048, 128) if ncomp == 1 else (4096, 3072, 256)
    else:
        max_slots, max_values, max_pairs = 4096, 8192, 256
    assert np.all(np.diff(s["row_ptr"][r0]) <= max_slots)
    assert np.all(np.diff(B["blk_v0"]) <= max_values)
    assert np.all(np.diff(B["blk_p0"]) <= max_pairs)
    assert B["blk_v0"][-1] == int((s["slot"] >= 0).sum())
    assert B["blk_p0"][-1] == len(tets) * nf


def test_row_beyond_default_budget_takes_wider_blocks():
    """A P3 node with 200 incident tets (4000 entries in its row, beyond the
    default 2048 per block): the host setup falls back to the wider budgets
    (8192 slots and entries) instead of refusing the mesh; the block lists
    still replay the oracle's scatter bit for bit."""
    coords, tets = T.fan_mesh(100)
    A = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P3), dirichlet=af.DIRICHLET_NONE, host_only=True)
    s = A.setup()
    B = A.blocks()
    nf = A.nf
    vmax = np.diff(B["blk_v0"]).max()
    assert 2048 < vmax <= 8192
    assert np.diff(s["row_ptr"][B["blk_row0"]]).max() <= 8192
    rng = np.random.default_rng(7)
    Ae = rng.standard_normal((len(tets), nf * nf))
    Fe = rng.standard_normal((len(tets), nf))
    rc, val_o, rhs_o = T.oracle_scatter(len(tets), nf, s["dofs"], s["slot"], Ae, Fe, A.nnz, A.ndof)
    assert rc == 0
    val, rhs = _emulate_blocks(Ae, Fe, s, B, nf)
    np.testing.assert_array_equal(val, val_o)
    np.testing.assert_array_equal(rhs, rhs_o)
