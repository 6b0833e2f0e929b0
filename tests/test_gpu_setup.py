"""Device enumeration phase of the assembly setup against the host phase.

A device context numbers edges and faces by first appearance, builds the
NATURAL DOF ids, the Dirichlet flags, the CSR pattern and the slot map on the
GPU (radix sorts); a host-only context builds them on the host
(setup.cpp, itself checked bit for bit against the oracle in
test_setup.py).  The two must agree exactly: every array, every count.  The
meshes are shuffled (tet order and local vertex order) so that first
appearance, the node reorder and the P3 edge-pair orientation are exercised
away from the generator's tidy order."""
import numpy as np
import pytest

import anifem_amd as af

pytestmark = pytest.mark.gpu


def shuffled_cube(n, seed):
    coords, tets = af.kuhn_cube(n, 0.2, seed)
    rng = np.random.default_rng(seed)
    tets = tets[rng.permutation(len(tets))]
    tets = np.take_along_axis(tets, np.argsort(rng.random(tets.shape), axis=1), axis=1)
    return coords, np.ascontiguousarray(tets, dtype=np.int32)


def both(coords, tets, op, **kw):
    dev = af.Assembler(coords, tets, op, **kw)
    host = af.Assembler(coords, tets, op, host_only=True, **kw)
    return dev, host


@pytest.mark.parametrize("fem,nc,dirichlet,n", [
    (af.FEM_P1, 1, af.DIRICHLET_NONE, 5),
    (af.FEM_P1, 1, af.DIRICHLET_BOUNDARY, 6),
    (af.FEM_P2, 1, af.DIRICHLET_BOUNDARY, 5),
    (af.FEM_P2, 1, af.DIRICHLET_NONE, 4),
    (af.FEM_P3, 1, af.DIRICHLET_BOUNDARY, 4),
    (af.FEM_P3, 1, af.DIRICHLET_NONE, 3),
    (af.FEM_P1, 3, af.DIRICHLET_BOUNDARY, 4),
    (af.FEM_P2, 3, af.DIRICHLET_BOUNDARY, 3),
])
def test_device_setup_matches_host(fem, nc, dirichlet, n):
    coords, tets = shuffled_cube(n, 10 * fem + nc + n)
    dev, host = both(coords, tets, af.Op(af.GRAD, fem, nc), dirichlet=dirichlet)
    assert (dev.ndof, dev.nnz, dev.n_blocks) == (host.ndof, host.nnz, host.n_blocks)
    sd, sh = dev.setup(), host.setup()
    for key in sh:
        assert np.array_equal(sd[key], sh[key]), key
    ld, lh = dev.block_layout(), host.block_layout()
    for name, _ in type(lh)._fields_:
        assert getattr(ld, name) == getattr(lh, name), name


def test_device_setup_dof_mask_and_unused_nodes():
    """A caller's DOF mask, and nodes that no tet uses (rows holding only
    their diagonal)."""
    coords, tets = shuffled_cube(4, 3)
    extra = np.random.default_rng(0).random((7, 3))
    coords = np.ascontiguousarray(np.vstack([coords, extra]))
    op = af.Op(af.GRAD, af.FEM_P2)
    probe = af.Assembler(coords, tets, op, host_only=True)
    mask = (np.random.default_rng(1).random(probe.ndof) < 0.2).astype(np.uint8)
    dev, host = both(coords, tets, op, dirichlet=af.DIRICHLET_DOF_MASK, dirichlet_dof_mask=mask)
    sd, sh = dev.setup(), host.setup()
    for key in sh:
        assert np.array_equal(sd[key], sh[key]), key
    assert np.array_equal(sd["dirichlet"], mask)


def test_device_setup_c2_size_counts():
    """At the 64^3 P2 size the device phase reproduces the published counts
    (2,146,689 DOFs; 57,611,761 stored nonzeros after the Dirichlet drop) and
    the host phase's arrays."""
    coords, tets = af.kuhn_cube(64, 0.2, 1)
    dev, host = both(coords, tets, af.Op(af.GRAD, af.FEM_P2), dirichlet=af.DIRICHLET_BOUNDARY)
    assert dev.ndof == 2146689 and dev.nnz == 57611761
    sd, sh = dev.setup(), host.setup()
    for key in sh:
        assert np.array_equal(sd[key], sh[key]), key
