#!/usr/bin/env python3
"""Write the golden element-matrix fixtures (tests/golden/elem_*.npz),
face-integral fixtures (tests/golden/face_*.npz) and solution-evaluation
fixtures (tests/golden/apply_*.npz).

Runs in the development container only: it loads oracle/_ref/libanifem_ref.so,
the reference AniFem++ element core compiled from its own sources by
oracle/Makefile, feeds it seeded tetrahedra and coefficients shaped like the
benchmark configurations, and stores inputs and outputs as FP64 arrays.
The fixtures are data (inputs + reference outputs); no reference source is
stored.  Regenerate with:  make -C oracle && python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libanifem_ref.so")

# harness pair codes (oracle/ref_harness.cpp) -> (opA, femA, ncA, opB, femB, ncB)
PAIRS = {
    0: (2, 2, 1, 2, 2, 1), 1: (2, 3, 1, 2, 3, 1), 2: (2, 4, 1, 2, 4, 1), 3: (2, 2, 3, 2, 2, 3),
    4: (2, 3, 3, 2, 3, 3), 5: (1, 1, 1, 1, 2, 1), 6: (1, 1, 1, 1, 3, 1), 7: (1, 1, 1, 1, 4, 1),
    8: (1, 1, 1, 1, 3, 3), 9: (1, 2, 1, 1, 2, 1), 10: (1, 3, 1, 1, 3, 1), 11: (1, 4, 1, 1, 4, 1),
    12: (2, 4, 1, 2, 2, 3), 13: (1, 1, 1, 1, 2, 3), 14: (2, 3, 1, 1, 3, 1),
    15: (3, 3, 3, 3, 3, 3), 16: (3, 3, 3, 1, 2, 1), 17: (4, 2, 3, 4, 2, 3), 18: (4, 3, 3, 1, 3, 3),
    19: (6, 3, 1, 7, 3, 1), 20: (8, 2, 1, 1, 2, 1),
    21: (1, 21, 1, 1, 21, 1), 22: (3, 21, 1, 3, 21, 1), 23: (3, 21, 1, 1, 1, 1), 24: (1, 25, 1, 1, 25, 1),
    25: (4, 25, 1, 4, 25, 1), 26: (4, 25, 1, 1, 21, 1), 27: (2, 31, 1, 2, 31, 1), 28: (1, 31, 1, 1, 2, 1),
    29: (2, 1002, 1, 2, 1002, 1), 30: (1, 1004, 1, 1, 1004, 1), 31: (2, 1004, 1, 2, 2, 1), 32: (1, 2, 1, 1, 25, 1),
    33: (2, 1004, 1, 2, 1004, 1), 34: (2, 21, 1, 2, 21, 1), 35: (2, 25, 1, 2, 25, 1), 36: (2, 25, 1, 2, 21, 1),
}
# basis functions per component: P0..P3, RT0, ND0, CR1, B2, B4
NF = {1: 1, 2: 4, 3: 10, 4: 20, 21: 4, 25: 6, 31: 4, 1002: 6, 1004: 1}


def kuhn_tets(rng, n_tets, n=16, jitter=0.2):
    """Random Kuhn tets of a jittered n^3 cube (positively oriented after the
    2<->3 swap rule), xyz (n_tets, 12)."""
    h = 1.0 / n
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    out = np.zeros((n_tets, 12))
    for r in range(n_tets):
        c0 = rng.integers(0, n, 3)
        p = perms[rng.integers(0, 6)]
        verts = [c0.copy()]
        c = c0.copy()
        for s in range(3):
            c[p[s]] += 1
            verts.append(c.copy())
        P = np.array(verts, dtype=float) * h
        for v in range(4):
            interior = np.all((verts[v] > 0) & (verts[v] < n))
            if interior:
                P[v] += rng.uniform(-jitter * h, jitter * h, 3)
        m = np.array([P[0] - P[3], P[1] - P[3], P[2] - P[3]])
        if np.linalg.det(m) < 0:
            P[[2, 3]] = P[[3, 2]]
        out[r] = P.ravel()
    return out


def haar_rotation(rng):
    Q, R = np.linalg.qr(rng.normal(size=(3, 3)))
    Q = Q @ np.diag(np.sign(np.diag(R)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


def aniso_tensor(rng):
    a, b = rng.uniform(-2, 2, 2)
    R = haar_rotation(rng)
    return R @ np.diag([1.0, 10.0 ** a, 10.0 ** b]) @ R.T


def lame_tensor(lam, mu):
    """9x9 C((i + 3j), (k + 3l)) of examples/tutorials/lin_elast.cpp:101-117."""
    C = np.zeros((9, 9))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    C[i + 3 * j, k + 3 * l] = lam * (i == j) * (k == l) + mu * ((i == k) * (j == l) + (i == l) * (j == k))
    return C


def cases(rng):
    """(name, pair, tensor_type, per_point, order, xyz, D full (n, dimB*dimA) or (n,1))."""
    out = []
    n = 64
    xyz = kuhn_tets(rng, n)
    out.append(("c1_p1_null", 0, 1, 0, 2, xyz, np.zeros((n, 1))))
    D = np.array([aniso_tensor(rng).ravel(order="F") for _ in range(n)])
    out.append(("c2_p2_sym", 1, 3, 0, 2, xyz, D))
    for name, pair, order in [("c3_p1_scalar", 0, 2), ("c3_p2_scalar", 1, 2), ("c3_p3_scalar", 2, 4)]:
        out.append((name, pair, 2, 0, order, xyz, 10.0 ** rng.uniform(-1, 1, (n, 1))))
    ne = 16
    E = 10.0 ** rng.uniform(0, 3, ne)
    nu = rng.uniform(0.2, 0.45, ne)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.array([lame_tensor(lam[r], mu[r]).ravel(order="F") for r in range(ne)])
    out.append(("c4_p2vec_lame", 4, 3, 0, 2, xyz[:ne], D, np.stack([lam, mu], axis=1)))
    for name, pair, order in [("rhs_p1", 5, 2), ("rhs_p2", 6, 2), ("rhs_p3", 7, 4)]:
        out.append((name, pair, 2, 0, order, xyz, np.ones((n, 1))))
    out.append(("rhs_vec_p2", 8, 4, 0, 2, xyz, rng.normal(size=(n, 3))))
    out.append(("mass_p2", 10, 2, 0, 2, xyz, 10.0 ** rng.uniform(-1, 1, (n, 1))))
    out.append(("mass_p3", 11, 2, 0, 5, xyz, 10.0 ** rng.uniform(-1, 1, (n, 1))))
    D = np.array([aniso_tensor(rng).ravel(order="F") for _ in range(n)])
    out.append(("pp_p2_general_xdep", 1, 4, 2, 4, xyz, D))
    out.append(("pp_p3_scalar_xdep", 2, 2, 2, 6, xyz, 10.0 ** rng.uniform(-1, 1, (n, 1))))
    out.append(("mixed_p3_vecp1_general", 12, 4, 0, 5, xyz[:32], rng.normal(size=(32, 27))))
    out.append(("grad_p2_iden_p2_general", 14, 4, 0, 2, xyz, rng.normal(size=(n, 3))))
    out.append(("vec_p1_general", 3, 4, 0, 2, xyz[:32], rng.normal(size=(32, 81))))
    return out


def grad9_cases(rng):
    """GRAD of RT0 and ND0 (9 components, raviant_thomas_0.h:55-113,
    nedelec_0.h:149-183) with 9 x 9 tensors."""
    n = 32
    xyz = kuhn_tets(rng, n)
    pos = lambda: 10.0 ** rng.uniform(-1, 1, (n, 1))
    return [
        ("g9_rt0_rt0_null", 34, 1, 0, 1, xyz, np.zeros((n, 1))),
        ("g9_rt0_rt0_general", 34, 4, 0, 2, xyz, rng.normal(size=(n, 81))),
        ("g9_nd0_nd0_scalar", 35, 2, 0, 1, xyz, pos()),
        ("g9_nd0_nd0_general_xdep", 35, 4, 2, 3, xyz, rng.normal(size=(n, 81))),
        ("g9_nd0_rt0_general", 36, 4, 0, 2, xyz, rng.normal(size=(n, 81))),
    ]


def high_cases(rng):
    """Orders 3 and 7..20 (the reference's 8-point order-3 rule and its rules
    of orders 7..20, the latter through the external-memory fem3Dtet).  Every
    integrand is a polynomial of degree <= order, so any exact rule of that
    order gives these matrices to rounding."""
    n = 24
    xyz = kuhn_tets(rng, n)
    kappa = 10.0 ** rng.uniform(-1, 1, (n, 1))
    D = np.array([aniso_tensor(rng).ravel(order="F") for _ in range(n)])
    return [
        ("hi_o3_p2_grad_scalar", 1, 2, 0, 3, xyz, kappa),                      # degree 2
        ("hi_o3_rt0_iden_null", 21, 1, 0, 3, xyz, np.zeros((n, 1))),           # degree 2
        ("hi_o7_p3_grad_general", 2, 4, 0, 7, xyz, D),                         # degree 4
        ("hi_o8_p2_grad_general_xdep", 1, 4, 2, 8, xyz, D),                    # degree 2 + 2
        ("hi_o9_nd0_curl_scalar", 25, 2, 0, 9, xyz, kappa),                    # degree 0
        ("hi_o12_p3_mass_scalar_xdep", 11, 2, 2, 12, xyz, kappa),              # degree 6 + 2
        ("hi_o15_p2vec_grad_general", 4, 4, 0, 15, xyz[:12], rng.normal(size=(12, 81))),  # degree 2
        ("hi_o20_p3_mass_scalar", 11, 2, 0, 20, xyz, kappa),                   # degree 6
    ]


def op_cases(rng):
    """DIV / CURL / DUDX..DUDZ operators (operators.h:286-400)."""
    n = 32
    xyz = kuhn_tets(rng, n)
    return [
        ("div_v2_div_v2_scalar", 15, 2, 0, 2, xyz, 10.0 ** rng.uniform(-1, 1, (n, 1))),
        ("div_v2_iden_p1_null", 16, 1, 0, 2, xyz, np.zeros((n, 1))),
        ("curl_v1_curl_v1_scalar", 17, 2, 0, 2, xyz, 10.0 ** rng.uniform(-1, 1, (n, 1))),
        ("curl_v2_iden_v2_general", 18, 4, 0, 3, xyz, rng.normal(size=(n, 9))),
        ("dudx_p2_dudy_p2_scalar", 19, 2, 0, 2, xyz, rng.uniform(0.5, 2, (n, 1))),
        ("dudz_p1_iden_p1_null", 20, 1, 0, 2, xyz, np.zeros((n, 1))),
    ]


def space_cases(rng):
    """RT0, ND0, CR1, B2 and B4 (spaces/*.h)."""
    n = 32
    xyz = kuhn_tets(rng, n)
    pos = lambda: 10.0 ** rng.uniform(-1, 1, (n, 1))
    spd = lambda: np.stack([aniso_tensor(rng).ravel(order="F") for _ in range(n)])
    return [
        ("sp_rt0_rt0_general", 21, 4, 0, 2, xyz, rng.normal(size=(n, 9))),
        ("sp_rt0_rt0_xdep", 21, 4, 2, 2, xyz, rng.normal(size=(n, 9))),
        ("sp_divrt0_divrt0_scalar", 22, 2, 0, 1, xyz, pos()),
        ("sp_divrt0_p0_null", 23, 1, 0, 1, xyz, np.zeros((n, 1))),
        ("sp_nd0_nd0_general", 24, 4, 0, 2, xyz, rng.normal(size=(n, 9))),
        ("sp_curlnd0_curlnd0_scalar", 25, 2, 0, 1, xyz, pos()),
        ("sp_curlnd0_rt0_null", 26, 1, 0, 2, xyz, np.zeros((n, 1))),
        ("sp_cr1_cr1_scalar", 27, 2, 0, 1, xyz, pos()),
        ("sp_cr1_p1_null", 28, 1, 0, 2, xyz, np.zeros((n, 1))),
        ("sp_b2_b2_general", 29, 4, 0, 2, xyz, rng.normal(size=(n, 9))),
        ("sp_b4_b4_null", 30, 1, 0, 6, xyz, np.zeros((n, 1))),
        ("sp_b4_b4_aniso", 33, 4, 0, 6, xyz, spd()),
        ("sp_p1_nd0_general", 32, 4, 0, 2, xyz, rng.normal(size=(n, 3))),
    ]


# face pair codes (ref_fem3dface) -> (opA, femA, ncA, opB, femB, ncB, dot_normal)
FACE_PAIRS = {
    0: (1, 2, 1, 1, 2, 1, 0), 1: (1, 3, 1, 1, 3, 1, 0), 2: (1, 4, 1, 1, 4, 1, 0), 3: (1, 1, 1, 1, 3, 1, 0),
    4: (1, 1, 1, 1, 3, 3, 0), 5: (2, 3, 1, 1, 3, 1, 1), 6: (2, 3, 3, 1, 3, 3, 1), 7: (2, 2, 1, 2, 2, 1, 0),
    8: (2, 4, 1, 1, 4, 1, 1),
}


def face_cases(rng):
    """(name, face pair, tensor_type, per_point, order, xyz, D (n, ddim))."""
    n = 32
    xyz = kuhn_tets(rng, n)
    out = [
        ("mass_p1_null", 0, 1, 0, 2, xyz, np.zeros((n, 1))),
        ("mass_p2_scalar", 1, 2, 0, 4, xyz, 10.0 ** rng.uniform(-1, 1, (n, 1))),
        ("mass_p3_null", 2, 1, 0, 6, xyz, np.zeros((n, 1))),
        ("load_p2", 3, 2, 0, 2, xyz, rng.uniform(0.5, 2, (n, 1))),
        ("traction_vec_p2", 4, 4, 0, 2, xyz, rng.normal(size=(n, 3))),
        ("flux_p2_normal", 5, 2, 0, 3, xyz, 10.0 ** rng.uniform(-1, 1, (n, 1))),
        ("grad_p1_general", 7, 4, 0, 1, xyz, np.array([aniso_tensor(rng).ravel(order="F") for _ in range(n)])),
        ("flux_p3_normal_general", 8, 4, 0, 5, xyz, rng.normal(size=(n, 9))),
        ("pp_mass_p2_scalar_xdep", 1, 2, 2, 6, xyz, 10.0 ** rng.uniform(-1, 1, (n, 1))),
    ]
    ne = 8
    E = 10.0 ** rng.uniform(0, 3, ne)
    nu = rng.uniform(0.2, 0.45, ne)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.array([lame_tensor(lam[r], mu[r]).ravel(order="F") for r in range(ne)])
    out.append(("traction_lame_normal", 6, 4, 0, 3, xyz[:ne], D))
    return out


def write_faces(L, rng):
    dp = ctypes.POINTER(ctypes.c_double)
    L.ref_fem3dface.restype = ctypes.c_int
    for name, pair, tt, pp, order, xyz, D in face_cases(rng):
        oa, fa, ca, ob, fb, cb, dn = FACE_PAIRS[pair]
        nfa, nfb = NF[fa] * ca, NF[fb] * cb
        n = len(xyz)
        xyz = np.ascontiguousarray(xyz)
        D = np.ascontiguousarray(D)
        A = np.zeros((4, n * nfa * nfb))
        for face in range(4):
            Af = np.zeros(n * nfa * nfb)
            rc = L.ref_fem3dface(pair, tt, pp, order, face, ctypes.c_long(n), xyz.ctypes.data_as(dp),
                                 D.ctypes.data_as(dp), ctypes.c_long(D.shape[1]), Af.ctypes.data_as(dp))
            if rc != 0:
                sys.exit(f"face {name}: reference harness returned {rc}")
            A[face] = Af
        meta = np.array([pair, tt, pp, order, oa, fa, ca, ob, fb, cb, dn], dtype=np.int64)
        np.savez_compressed(os.path.join(HERE, f"face_{name}.npz"), meta=meta, xyz=xyz, D=D,
                            A=A.reshape(4, n, nfa * nfb))
        print(f"face {name}: {n} tets x 4 faces, {nfb}x{nfa}")


# fem3DapplyL / fem3DapplyX op codes (ref_fem3dapply, ref_fem3dapplyx) -> (op, fem, ncomp)
APPLY_OPS = {0: (1, 2, 1), 1: (1, 3, 1), 2: (2, 3, 1), 3: (1, 4, 1), 4: (2, 4, 1), 5: (1, 3, 3), 6: (2, 3, 3),
             7: (1, 21, 1), 8: (3, 21, 1), 9: (1, 25, 1), 10: (4, 25, 1), 11: (1, 31, 1), 12: (2, 1002, 1),
             13: (2, 1004, 1)}


def op_dim(op, fem, nc):
    if fem in (21, 25):
        return 1 if op == 3 else 3
    return (3 if op == 2 else 1) * nc


def write_apply(L, rng, codes, at_x=False):
    dp = ctypes.POINTER(ctypes.c_double)
    fn = L.ref_fem3dapplyx if at_x else L.ref_fem3dapply
    fn.restype = ctypes.c_int
    n, q = 24, 7
    xyz = np.ascontiguousarray(kuhn_tets(rng, n))
    lam = rng.dirichlet(np.ones(4), q)  # interior barycentric points
    xyl = np.ascontiguousarray(lam)
    if at_x:
        # per tet its own q physical points (inside and, for the last two, outside the tet)
        lam_t = rng.dirichlet(np.ones(4), (n, q))
        lam_t[:, -2:] = lam_t[:, -2:] * 1.5 - 0.125
        X = np.ascontiguousarray(np.einsum("rqv,rvc->rqc", lam_t, xyz.reshape(n, 4, 3)))
    for code in codes:
        op, fem, nc = APPLY_OPS[code]
        nf = NF[fem] * nc
        dim = op_dim(op, fem, nc)
        dofs = np.ascontiguousarray(rng.normal(size=(n, nf)))
        out = np.zeros(n * dim * q)
        pts = X if at_x else xyl
        rc = fn(code, q, pts.ctypes.data_as(dp), ctypes.c_long(n), xyz.ctypes.data_as(dp), dofs.ctypes.data_as(dp),
                out.ctypes.data_as(dp))
        if rc != 0:
            sys.exit(f"apply {code}: reference harness returned {rc}")
        meta = np.array([code, op, fem, nc], dtype=np.int64)
        name = f"applyx_{code}.npz" if at_x else f"apply_{code}.npz"
        extra = dict(X=X) if at_x else dict(xyl=xyl)
        np.savez_compressed(os.path.join(HERE, name), meta=meta, xyz=xyz, dofs=dofs, out=out.reshape(n, q, dim),
                            **extra)
        print(f"{name}: {n} tets x {q} points, dim {dim}")


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: build it with `make -C oracle`")
    L = ctypes.CDLL(REF)
    L.ref_fem3dtet.restype = ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    # groups to write (default all): elem (with the operator and space cases), faces, apply;
    # "spaces" writes the cases of the RT0 / ND0 / CR1 / B2 / B4 spaces and fem3DapplyX only;
    # "high" the cases of quadrature orders 3 and 7..20; "grad9" those of GRAD of RT0 / ND0
    groups = set(sys.argv[1:]) or {"elem", "faces", "apply"}
    rng = np.random.default_rng(20260101)
    todo = []
    if "elem" in groups:
        todo += cases(rng) + op_cases(np.random.default_rng(20260404))
    if "elem" in groups or "spaces" in groups:
        todo += space_cases(np.random.default_rng(20260505))
    if "elem" in groups or "high" in groups:
        todo += high_cases(np.random.default_rng(20260808))
    if "elem" in groups or "grad9" in groups:
        todo += grad9_cases(np.random.default_rng(20260909))
    for case in todo:
        name, pair, tt, pp, order, xyz, D = case[:7]
        extra = case[7] if len(case) > 7 else None
        oa, fa, ca, ob, fb, cb = PAIRS[pair]
        nfa, nfb = NF[fa] * ca, NF[fb] * cb
        n = len(xyz)
        xyz = np.ascontiguousarray(xyz)
        D = np.ascontiguousarray(D)
        A = np.zeros(n * nfa * nfb)
        rc = L.ref_fem3dtet(pair, tt, pp, order, ctypes.c_long(n), xyz.ctypes.data_as(dp), D.ctypes.data_as(dp),
                            ctypes.c_long(D.shape[1]), A.ctypes.data_as(dp))
        if rc != 0:
            sys.exit(f"{name}: reference harness returned {rc}")
        meta = np.array([pair, tt, pp, order, oa, fa, ca, ob, fb, cb], dtype=np.int64)
        arrays = dict(meta=meta, xyz=xyz, D=D, A=A.reshape(n, nfa * nfb))
        if extra is not None:
            arrays["lame"] = extra
        np.savez_compressed(os.path.join(HERE, f"elem_{name}.npz"), **arrays)
        print(f"{name}: {n} tets, {nfb}x{nfa}")
    if "faces" in groups:
        write_faces(L, np.random.default_rng(20260202))
    if "apply" in groups:
        write_apply(L, np.random.default_rng(20260303), range(7))
    if "apply" in groups or "spaces" in groups:
        write_apply(L, np.random.default_rng(20260606), range(7, 14))
        write_apply(L, np.random.default_rng(20260707), range(14), at_x=True)


if __name__ == "__main__":
    main()
