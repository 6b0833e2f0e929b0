#!/usr/bin/env python3
"""Generate the symmetric triangle quadrature tables of the face integrals.

Face integrals (anifem++/fem/operations/int_face.inl: fem3Dface, fem3DfaceN)
integrate over one face of a tet with ``triangle_quadrature_formulas(order)``
(anifem++/fem/quadrature_formulas.cpp:109-...).  Every integrand on the path is
a polynomial, so any rule exact to the order gives the same integral up to
rounding; the rules here are fully symmetric, interior, positive-weight rules
of the following orbit structures (barycentric (l0, l1, l2)):

    order 0,1 -> 1 point  (centroid)                         degree 1
    order 2   -> 3 points (one (a, a, 1-2a) orbit)           degree 2
    order 3,4 -> 6 points (two (a, a, 1-2a) orbits)          degree 4
    order 5   -> 7 points (centroid + two (a, a, 1-2a))      degree 5
    order 6   -> 12 points (two (a, a, 1-2a) + one (a, b, c)) degree 6

The values are not copied from anywhere: for each structure the script solves
the moment equations  sum_i w_i l(p_i)^alpha = 2! alpha! / (|alpha| + 2)!  for
every barycentric monomial of degree <= d by a seeded multi-start
least-squares search, keeps the admissible solution (points interior, weights
positive; where there are several, the one whose smallest barycentric
coordinate is largest) and refines it to 50 digits with Gauss-Newton in
mpmath.

Usage: python gen_tri_rules.py > tri_rules.inc
"""
import math
import sys

import mpmath as mp
import numpy as np
from scipy.optimize import least_squares

mp.mp.dps = 50

NPAR = {"t1": 0, "t3": 1, "t6": 2}
RULES = [
    ((0, 1), ["t1"], 1),
    ((2,), ["t3"], 2),
    ((3, 4), ["t3", "t3"], 4),
    ((5,), ["t1", "t3", "t3"], 5),
    ((6,), ["t3", "t3", "t6"], 6),
]


def monomials(deg):
    return [(a, b, s - a - b) for s in range(deg + 1) for a in range(s + 1) for b in range(s - a + 1)]


def mono_mean(al, fac=math.factorial):
    a, b, c = al
    return 2 * fac(a) * fac(b) * fac(c) / fac(a + b + c + 2)


def orbit(kind, p, one=1.0):
    if kind == "t1":
        return [[one / 3, one / 3, one / 3]]
    if kind == "t3":
        a = p[0]
        pts = []
        for i in range(3):
            x = [a] * 3
            x[i] = one - 2 * a
            pts.append(x)
        return pts
    if kind == "t6":
        a, b = p
        c = one - a - b
        v = [a, b, c]
        return [[v[i], v[j], v[3 - i - j]] for i in range(3) for j in range(3) if i != j]
    raise ValueError(kind)


def expand(struct, z, one=1.0):
    pts, ws, k = [], [], 0
    wv = z[len(z) - len(struct):]
    for o, kind in enumerate(struct):
        p = z[k:k + NPAR[kind]]
        k += NPAR[kind]
        P = orbit(kind, p, one)
        pts += P
        ws += [wv[o]] * len(P)
    return pts, ws


def residual_np(z, struct, E, rhs):
    P, W = expand(struct, list(z))
    P = np.array(P)
    V = np.prod(P[:, None, :] ** E[None, :, :], axis=2)
    return np.array(W) @ V - rhs


def search(struct, deg, tries=200, seed=11):
    ms = monomials(deg)
    E = np.array(ms, dtype=float)
    rhs = np.array([mono_mean(m) for m in ms])
    rng = np.random.default_rng(seed)
    found = []
    for _ in range(tries):
        z = []
        for kind in struct:
            if kind == "t3":
                z.append(rng.uniform(0, 0.5))
            elif kind == "t6":
                z += [rng.uniform(0, 0.5), rng.uniform(0, 0.5)]
        z += list(rng.uniform(0.02, 0.4, len(struct)))
        r = least_squares(residual_np, np.array(z), args=(struct, E, rhs),
                          xtol=1e-15, ftol=1e-15, gtol=1e-15, max_nfev=400)
        if np.max(np.abs(r.fun)) > 1e-12:
            continue
        P, W = expand(struct, list(r.x))
        if np.min(P) <= 1e-9 or np.min(W) <= 1e-12:
            continue
        key = tuple(np.round(np.sort(np.array(P).ravel()), 7))
        if all(key != f[0] for f in found):
            found.append((key, r.x))
    if not found:
        raise RuntimeError(f"{struct}: no admissible rule")
    # several admissible rules (degree 6): the most interior one
    return max(found, key=lambda f: np.min(expand(struct, list(f[1]))[0]))[1]


def refine(struct, deg, z0):
    ms = monomials(deg)
    rhs = [mp.mpf(2) * mp.factorial(a) * mp.factorial(b) * mp.factorial(c) / mp.factorial(a + b + c + 2)
           for a, b, c in ms]
    z = [mp.mpf(float(v)) for v in z0]
    one = mp.mpf(1)

    def F(zz):
        P, W = expand(struct, zz, one)
        return [sum(W[i] * P[i][0] ** m[0] * P[i][1] ** m[1] * P[i][2] ** m[2] for i in range(len(P))) - rhs[j]
                for j, m in enumerate(ms)]

    for _ in range(60):
        f0 = F(z)
        J = mp.matrix(len(ms), len(z))
        h = mp.mpf(10) ** -30
        for k in range(len(z)):
            zp = list(z)
            zp[k] += h
            fk = F(zp)
            for j in range(len(ms)):
                J[j, k] = (fk[j] - f0[j]) / h
        dz = mp.lu_solve(J.T * J, -(J.T * mp.matrix(f0)))
        z = [z[k] + dz[k] for k in range(len(z))]
        if max(abs(v) for v in F(z)) < mp.mpf(10) ** -45:
            break
    res = max(abs(v) for v in F(z))
    if res > mp.mpf(10) ** -40:
        raise RuntimeError(f"{struct}: refinement did not converge ({res})")
    return z


def main():
    out = sys.stdout
    out.write("// Generated by gen_tri_rules.py -- symmetric triangle quadrature rules.\n")
    out.write("// Points are barycentric (l0, l1, l2) on the face; weights sum to 1 (multiply\n")
    out.write("// by the face area).\n")
    out.write("#pragma once\n\n")
    names = []
    for orders, struct, deg in RULES:
        if struct == ["t1"]:
            P, W = expand(struct, [mp.mpf(1)], mp.mpf(1))
            W = [mp.mpf(1)]
        else:
            z = refine(struct, deg, search(struct, deg))
            P, W = expand(struct, z, mp.mpf(1))
        n = len(P)
        name = f"ANIFEM_TRI_Q{n}"
        names.append((orders, name, n))
        out.write(f"// {n} points, exact for degree {deg}; orbits {'+'.join(struct)}\n")
        out.write(f"static const double {name}_P[{3 * n}] = {{\n")
        for p in P:
            out.write("    " + ", ".join(mp.nstr(v, 20, strip_zeros=False) for v in p) + ",\n")
        out.write("};\n")
        out.write(f"static const double {name}_W[{n}] = {{\n")
        for w in W:
            out.write("    " + mp.nstr(w, 20, strip_zeros=False) + ",\n")
        out.write("};\n\n")
    out.write("#define ANIFEM_TRI_MAX_ORDER 6\n")
    out.write("#define ANIFEM_TRI_MAX_NPTS 12\n")
    npts = [0] * 7
    for orders, name, n in names:
        for o in orders:
            npts[o] = n
    out.write("static const int ANIFEM_TRI_NPTS_BY_ORDER[7] = {" + ", ".join(map(str, npts)) + "};\n")


if __name__ == "__main__":
    main()
