"""Test helpers: the CPU oracle (oracle/liboracle.so, the checker) and the
conversions between the oracle's input layouts and the product's C ABI
layouts.  Only tests import this module."""
import ctypes
import glob
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libanifem_ref.so")
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")

NF = {1: 1, 2: 4, 3: 10, 4: 20, 21: 4, 25: 6, 31: 4, 1002: 6, 1004: 1}

_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True, capture_output=True)
        L = ctypes.CDLL(ORACLE_LIB)
        L.orc_enumerate.restype = ctypes.c_long
        L.orc_csr_pattern.restype = ctypes.c_long
        L.orc_jacobi_pcg.restype = ctypes.c_long
        L.orc_inverse3x3.restype = ctypes.c_double
        _oracle = L
    return _oracle


def ref_lib():
    """The compiled reference element core, or None where it is not built."""
    if not os.path.exists(REF_LIB):
        return None
    return ctypes.CDLL(REF_LIB)


def P(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def nfa(fem, nc=1):
    return NF[fem] * nc


def dim(op, nc=1, fem=None):
    if fem in (21, 25):  # RT0, ND0: IDEN and CURL 3, DIV 1, GRAD 9
        return 1 if op == 3 else (9 if op == 2 else 3)
    return (3 if op == 2 else 1) * nc


def orc_quadrature(order):
    p = np.zeros(4 * 1331)
    w = np.zeros(1331)
    q = oracle().orc_quadrature(order, P(p), P(w))
    return p[:4 * q].reshape(q, 4), w[:q]


def orc_elem(opA, femA, ncA, opB, femB, ncB, ttype, mode, order, xyz, D):
    """Oracle element matrices: returns (n, nfB*nfA) column-major per tet."""
    n = len(xyz)
    A = np.zeros(n * nfa(femA, ncA) * nfa(femB, ncB))
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    D = None if D is None else np.ascontiguousarray(D, dtype=np.float64)
    rc = oracle().orc_fem3dtet(opA, femA, ncA, opB, femB, ncB, ttype, mode, order, ctypes.c_long(n), P(xyz), P(D),
                               P(A))
    if rc != 0:
        raise RuntimeError(f"oracle fem3dtet rc={rc}")
    return A.reshape(n, -1)


def orc_face_quadrature(face):
    p = np.zeros(4 * 36)
    w = np.zeros(36)
    q = oracle().orc_face_quadrature(face, P(p), P(w))
    return p[:4 * q].reshape(q, 4), w[:q]


def orc_face(opA, femA, ncA, opB, femB, ncB, ttype, mode, face, dot_normal, xyz, D):
    """Oracle face integrals (fem3Dface / fem3DfaceN): (n, nfB*nfA)."""
    n = len(xyz)
    A = np.zeros(n * nfa(femA, ncA) * nfa(femB, ncB))
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    D = None if D is None else np.ascontiguousarray(D, dtype=np.float64)
    rc = oracle().orc_fem3dface(opA, femA, ncA, opB, femB, ncB, ttype, mode, face, 1 if dot_normal else 0,
                                ctypes.c_long(n), P(xyz), P(D), P(A))
    if rc != 0:
        raise RuntimeError(f"oracle fem3dface rc={rc}")
    return A.reshape(n, -1)


def points_physical(xyz, p):
    """Physical points (n, q, 3) of barycentric points p (q, 4)."""
    P0 = xyz[:, 0:3]
    X = np.repeat(P0[:, None, :], len(p), axis=1).copy()
    for l in range(3):
        X += p[None, :, l + 1, None] * (xyz[:, None, 3 * (l + 1):3 * (l + 2)] - P0[:, None, :])
    return X


def orc_apply(op, fem, nc, xyl, xyz, dofs):
    """Oracle fem3DapplyL: (n, q, dim)."""
    n, q = len(xyz), len(xyl)
    d = dim(op, nc, fem)
    out = np.zeros(n * q * d)
    xyl = np.ascontiguousarray(xyl, dtype=np.float64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    dofs = np.ascontiguousarray(dofs, dtype=np.float64)
    rc = oracle().orc_fem3dapply(op, fem, nc, q, P(xyl), ctypes.c_long(n), P(xyz), P(dofs), P(out))
    if rc != 0:
        raise RuntimeError(f"oracle fem3dapply rc={rc}")
    return out.reshape(n, q, d)


def orc_applyx(op, fem, nc, X, xyz, dofs):
    """Oracle fem3DapplyX at per-tet physical points X (n, q, 3): (n, q, dim)."""
    n, q = X.shape[0], X.shape[1]
    d = dim(op, nc, fem)
    out = np.zeros(n * q * d)
    X = np.ascontiguousarray(X, dtype=np.float64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    dofs = np.ascontiguousarray(dofs, dtype=np.float64)
    rc = oracle().orc_fem3dapplyx(op, fem, nc, q, ctypes.c_long(n), P(xyz), P(X), P(dofs), P(out))
    if rc != 0:
        raise RuntimeError(f"oracle fem3dapplyx rc={rc}")
    return out.reshape(n, q, d)


def applyx_golden_cases():
    out = []
    for f in sorted(glob.glob(os.path.join(GOLDEN_DIR, "applyx_*.npz"))):
        z = np.load(f)
        out.append((os.path.basename(f)[7:-4], {k: z[k] for k in z.files}))
    return out


def apply_golden_cases():
    out = []
    for f in sorted(glob.glob(os.path.join(GOLDEN_DIR, "apply_*.npz"))):
        z = np.load(f)
        out.append((os.path.basename(f)[6:-4], {k: z[k] for k in z.files}))
    return out


def face_golden_cases():
    out = []
    for f in sorted(glob.glob(os.path.join(GOLDEN_DIR, "face_*.npz"))):
        z = np.load(f)
        out.append((os.path.basename(f)[5:-4], {k: z[k] for k in z.files}))
    return out


def physical_points(xyz, order):
    """Physical quadrature points (n, q, 3) of the rule (core.inl:250-270)."""
    p, _ = orc_quadrature(order)
    P0 = xyz[:, 0:3]
    X = np.repeat(P0[:, None, :], len(p), axis=1).copy()
    for l in range(3):
        X += p[None, :, l + 1, None] * (xyz[:, None, 3 * (l + 1):3 * (l + 2)] - P0[:, None, :])
    return X


def product_physical_points(xyz, order):
    """Physical points (n, q, 3) of the product's rule, in the point order of
    anifem_quadrature (the order its per-point coefficient arrays follow)."""
    import anifem_amd as af
    p, _ = af.quadrature(order)
    P0 = xyz[:, 0:3]
    X = np.repeat(P0[:, None, :], len(p), axis=1).copy()
    for l in range(3):
        X += p[None, :, l + 1, None] * (xyz[:, None, 3 * (l + 1):3 * (l + 2)] - P0[:, None, :])
    return X


def point_permutation(order):
    """perm[j] = index of the oracle's point equal to the product's point j."""
    import anifem_amd as af
    pp, _ = af.quadrature(order)
    po, _ = orc_quadrature(order)
    perm = np.array([int(np.argmin(np.abs(po - pp[j]).max(axis=1))) for j in range(len(pp))])
    assert np.abs(po[perm] - pp).max() < 1e-15 and len(set(perm.tolist())) == len(pp)
    return perm


def xdep_field(X):
    """m(x) = 1 + x0/2 - x1/4 + x2^2/8 (the same field as oracle/ref_harness.cpp)."""
    return 1.0 + 0.5 * X[..., 0] - 0.25 * X[..., 1] + 0.125 * X[..., 2] ** 2


def pack_sym(Dfull, d):
    """(n, d*d) column-major full -> (n, d(d+1)/2) packed upper triangle."""
    n = len(Dfull)
    M = Dfull.reshape(n, d, d).transpose(0, 2, 1)  # M[r, i, j] = D(i, j)
    out = np.zeros((n, d * (d + 1) // 2))
    for j in range(d):
        for i in range(j + 1):
            out[:, i + j * (j + 1) // 2] = M[:, i, j]
    return out


def soa_xyz(xyz):
    """(n, 12) per-tet vertices -> (12, n) SoA of the C ABI."""
    return np.ascontiguousarray(np.asarray(xyz).T)


def rel_frob(A, B):
    """max over tets of ||A-B||_F / ||B||_F."""
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    nb = np.linalg.norm(B, axis=1)
    return float(np.max(np.linalg.norm(A - B, axis=1) / np.maximum(nb, 1e-300)))


def golden_cases():
    out = []
    for f in sorted(glob.glob(os.path.join(GOLDEN_DIR, "elem_*.npz"))):
        z = np.load(f, allow_pickle=False)
        out.append((os.path.basename(f)[5:-4], {k: z[k] for k in z.files}))
    return out


def oracle_assembly(coords, tets, fem, nc, dirichlet_mask=None, with_slots=True):
    """Oracle enumeration / pattern / slot map of a mesh (reorder applied)."""
    O = oracle()
    tt = np.ascontiguousarray(tets, dtype=np.int32).copy()
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    O.orc_reorder_nodes(ctypes.c_long(len(tt)), P(tt), P(coords))
    nf = nfa(fem, nc)
    dofs = np.zeros((len(tt), nf), np.int32)
    ne, nfc = ctypes.c_long(), ctypes.c_long()
    ndof = O.orc_enumerate(fem, nc, ctypes.c_long(len(coords)), ctypes.c_long(len(tt)), P(tt), P(dofs),
                           ctypes.byref(ne), ctypes.byref(nfc))
    mask = None if dirichlet_mask is None else np.ascontiguousarray(dirichlet_mask, dtype=np.uint8)
    rp = np.zeros(ndof + 1, np.int64)
    nnz = O.orc_csr_pattern(ctypes.c_long(ndof), ctypes.c_long(len(tt)), nf, P(dofs), P(mask), P(rp), None)
    ci = np.zeros(nnz, np.int32)
    O.orc_csr_pattern(ctypes.c_long(ndof), ctypes.c_long(len(tt)), nf, P(dofs), P(mask), P(rp), P(ci))
    out = dict(tets=tt, dofs=dofs, ndof=ndof, row_ptr=rp, col_idx=ci, n_edges=ne.value, n_faces=nfc.value)
    if with_slots:
        sl = np.zeros((len(tt), nf * nf), np.int64)
        O.orc_slot_map(ctypes.c_long(len(tt)), nf, P(dofs), P(rp), P(ci), P(sl))
        out["slot"] = sl
    return out


def oracle_scatter(n_tets, nf, dofs, slot, A, F, nnz, ndof, drop_val=1e-100):
    val = np.zeros(nnz)
    rhs = np.zeros(ndof)
    A = np.ascontiguousarray(A, dtype=np.float64)
    F = None if F is None else np.ascontiguousarray(F, dtype=np.float64)
    rc = oracle().orc_scatter(ctypes.c_long(n_tets), nf, P(dofs), P(slot), P(A), P(F), ctypes.c_double(drop_val),
                              P(val), P(rhs))
    return rc, val, rhs


def apply_dir_elementwise(A, F, dofs, is_dir, bc=None):
    """Element-level Dirichlet through the oracle's applyDir on every flagged
    local DOF of every cell (bc: value per global DOF, None = homogeneous)."""
    n, nf = dofs.shape
    A = np.ascontiguousarray(A, dtype=np.float64).copy()
    F = None if F is None else np.ascontiguousarray(F, dtype=np.float64).copy()
    dofs = np.ascontiguousarray(dofs, dtype=np.int32)
    is_dir = np.ascontiguousarray(is_dir, dtype=np.uint8)
    bc = None if bc is None else np.ascontiguousarray(bc, dtype=np.float64)
    oracle().orc_apply_dir_cells(ctypes.c_long(n), nf, P(dofs), P(is_dir), P(bc), P(A), P(F))
    return A, F


def shuffled_mesh(coords, tets, seed):
    """The mesh with relabelled nodes, shuffled cells and a shuffled local
    vertex order in every cell: exercises the node reorder (ordering.inl:8-26)
    and the P3 edge orientation by global node id (ordering.inl:137-151)."""
    rng = np.random.default_rng(seed)
    relabel = rng.permutation(len(coords))
    inv = np.empty_like(relabel)
    inv[relabel] = np.arange(len(coords))
    coords2 = np.ascontiguousarray(coords[inv])
    tets2 = relabel[tets]
    tets2 = tets2[rng.permutation(len(tets2))]
    for r in range(len(tets2)):
        tets2[r] = tets2[r][rng.permutation(4)]
    return coords2, np.ascontiguousarray(tets2, dtype=np.int32)


def lagrange_nodes(fem):
    """Barycentric location (nf, 4) of every local DOF of a Lagrange space,
    read off the oracle's basis (pinned by the compiled reference's
    fem3DapplyL fixtures): DOF i sits at the lattice point where phi_i = 1 and
    every other basis function vanishes."""
    p = {2: 1, 3: 2, 4: 3}[fem]
    lat = np.array([(a, b, c, p - a - b - c) for a in range(p + 1) for b in range(p + 1 - a)
                    for c in range(p + 1 - a - b)], dtype=np.float64) / p
    nf = NF[fem]
    xyz = np.tile(np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], dtype=np.float64), (nf, 1))
    phi = orc_apply(1, fem, 1, lat, xyz, np.eye(nf))[:, :, 0]  # phi[i, point]
    out = np.zeros((nf, 4))
    for i in range(nf):
        j = int(np.argmax(phi[i]))
        assert abs(phi[i, j] - 1) < 1e-12 and np.abs(np.delete(phi[:, j], i)).max() < 1e-12
        out[i] = lat[j]
    return out


def dof_positions(coords, rtets, dofs, nodes):
    """Physical position of every global DOF seen from every cell: (n_tets, nf, 3)."""
    P = coords[rtets]  # (n, 4, 3)
    return np.einsum("iv,tvc->tic", nodes, P)


def fan_mesh(m):
    """2m tets around the origin: a ring of m points in the plane z = 0 joined
    to (0, 0, 1) and (0, 0, -1)."""
    th = 2 * np.pi * np.arange(m) / m
    coords = np.concatenate([np.zeros((1, 3)), np.stack([np.cos(th), np.sin(th), np.zeros(m)], 1),
                             np.array([[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])])
    k = np.arange(m)
    r0, r1 = 1 + k, 1 + (k + 1) % m
    top = np.stack([np.zeros(m, int), r0, r1, np.full(m, m + 1)], 1)
    bot = np.stack([np.zeros(m, int), r1, r0, np.full(m, m + 2)], 1)
    return coords, np.concatenate([top, bot]).astype(np.int32)
