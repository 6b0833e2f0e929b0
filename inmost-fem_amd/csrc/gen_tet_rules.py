#!/usr/bin/env python3
"""Generate the symmetric tetrahedron quadrature tables used by the HIP kernels.

The rules are the fully symmetric, positive-weight, interior-point rules that
AniFem++ selects in ``tetrahedron_quadrature_formulas(order)``
(anifem++/fem/quadrature_formulas.cpp:526-1503, order map :1493-1501):

    order 0,1 -> 1 point (centroid)
    order 2   -> 4 points, one S4 orbit                     (degree 2)
    order 3   -> served here by the 14-point degree-5 rule (see below)
    order 4,5 -> 14 points, two S4 orbits + one S6 orbit    (degree 5)
    order 6   -> 24 points, three S4 orbits + one S12 orbit (degree 6)

The point/weight values are not copied: for each orbit structure this script
solves the moment equations  sum_i w_i lambda(p_i)^alpha = 3! alpha! / (|alpha|+3)!
for every barycentric monomial of degree <= d by a seeded multi-start
least-squares search (numpy/scipy) and then refines the single admissible
solution (all points interior, all weights positive) to 50 digits with
Gauss-Newton in mpmath.  For the 4/14/24-point structures the admissible
solution is unique, so the result equals the reference's rule up to rounding.
The reference's 8-point order-3 rule is one member of a one-parameter family
that the moment equations do not fix; order 3 therefore uses the 14-point
degree-5 rule, which integrates every polynomial of degree <= 3 exactly, so
results agree to rounding for polynomial integrands of that degree.

Orbit point order: an S4 orbit lists the point whose distinct coordinate sits
at position 0,1,2,3 (the same order as the reference's TWIST4 macro); S6 and
S12 orbits use lexicographic position order.

Usage: python gen_tet_rules.py > tet_rules.inc
"""
import itertools
import math
import sys

import mpmath as mp
import numpy as np
from scipy.optimize import least_squares

mp.mp.dps = 50

NPAR = {"s1": 0, "s4": 1, "s6": 1, "s12": 2}
# (orders served, orbit structure, polynomial degree)
RULES = [
    ((0, 1), ["s1"], 1),
    ((2,), ["s4"], 2),
    ((3, 4, 5), ["s4", "s4", "s6"], 5),
    ((6,), ["s4", "s4", "s4", "s12"], 6),
]


def monomials(deg):
    out = []
    for s in range(deg + 1):
        for a in range(s + 1):
            for b in range(s - a + 1):
                for c in range(s - a - b + 1):
                    out.append((a, b, c, s - a - b - c))
    return out


def mono_mean(al, fac=math.factorial):
    a, b, c, d = al
    return 6 * fac(a) * fac(b) * fac(c) * fac(d) / fac(a + b + c + d + 3)


def orbit(kind, p, one=1.0):
    """Points of one orbit; p holds the orbit parameters."""
    if kind == "s1":
        q = one / 4
        return [[q, q, q, q]]
    if kind == "s4":
        a = p[0]
        pts = []
        for i in range(4):
            x = [a] * 4
            x[i] = one - 3 * a
            pts.append(x)
        return pts
    if kind == "s6":
        a = p[0]
        b = one / 2 - a
        pts = []
        for pair in itertools.combinations(range(4), 2):
            x = [b] * 4
            for k in pair:
                x[k] = a
            pts.append(x)
        return pts
    if kind == "s12":
        a, b = p
        c = one - 2 * a - b
        pts = []
        for i in range(4):
            for j in range(4):
                if i != j:
                    x = [a] * 4
                    x[i] = b
                    x[j] = c
                    pts.append(x)
        return pts
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


def search(struct, deg, tries=150, seed=7):
    ms = monomials(deg)
    E = np.array(ms, dtype=float)
    rhs = np.array([mono_mean(m) for m in ms])
    rng = np.random.default_rng(seed)
    found = []
    for _ in range(tries):
        z = []
        for kind in struct:
            if kind == "s4":
                z.append(rng.uniform(0, 1 / 3))
            elif kind == "s6":
                z.append(rng.uniform(0, 0.5))
            elif kind == "s12":
                z += [rng.uniform(0, 0.5), rng.uniform(0, 1)]
        z += list(rng.uniform(0.01, 0.2, len(struct)))
        r = least_squares(residual_np, np.array(z), args=(struct, E, rhs),
                          xtol=1e-15, ftol=1e-15, gtol=1e-15, max_nfev=300)
        if np.max(np.abs(r.fun)) > 1e-12:
            continue
        P, W = expand(struct, list(r.x))
        if np.min(P) <= 0 or np.min(W) <= 0:
            continue
        key = tuple(np.round(np.sort(np.array(P).ravel()), 7))
        if all(key != f[0] for f in found):
            found.append((key, r.x))
    if len(found) != 1:
        raise RuntimeError(f"{struct}: expected a unique admissible rule, found {len(found)}")
    return found[0][1]


def refine(struct, deg, z0):
    ms = monomials(deg)
    rhs = [mp.mpf(6) * mp.factorial(a) * mp.factorial(b) * mp.factorial(c) * mp.factorial(d)
           / mp.factorial(a + b + c + d + 3) for a, b, c, d in ms]
    z = [mp.mpf(float(v)) for v in z0]
    one = mp.mpf(1)

    def F(zz):
        P, W = expand(struct, zz, one)
        return [sum(W[i] * P[i][0] ** m[0] * P[i][1] ** m[1] * P[i][2] ** m[2] * P[i][3] ** m[3]
                    for i in range(len(P))) - rhs[j] for j, m in enumerate(ms)]

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
        fv = mp.matrix(f0)
        dz = mp.lu_solve(J.T * J, -(J.T * fv))
        z = [z[k] + dz[k] for k in range(len(z))]
        if max(abs(v) for v in F(z)) < mp.mpf(10) ** -45:
            break
    res = max(abs(v) for v in F(z))
    if res > mp.mpf(10) ** -40:
        raise RuntimeError(f"{struct}: refinement did not converge ({res})")
    return z


def main():
    out = sys.stdout
    out.write("// Generated by gen_tet_rules.py -- symmetric tetrahedron quadrature rules.\n")
    out.write("// Points are barycentric (l0, l1, l2, l3); weights sum to 1 (mean-value form,\n")
    out.write("// multiply by |T|), as in anifem++/fem/quadrature_formulas.h:148-207.\n")
    out.write("#pragma once\n\n")
    names = []
    for orders, struct, deg in RULES:
        if struct == ["s1"]:
            z = []
            P, W = expand(struct, [mp.mpf(1)], mp.mpf(1))
            W = [mp.mpf(1)]
        else:
            z = refine(struct, deg, search(struct, deg))
            P, W = expand(struct, z, mp.mpf(1))
        n = len(P)
        name = f"ANIFEM_TET_Q{n}"
        names.append((orders, name, n, deg))
        out.write(f"// {n} points, exact for degree {deg}; orbits {'+'.join(struct)}\n")
        out.write(f"static const double {name}_P[{4 * n}] = {{\n")
        for p in P:
            out.write("    " + ", ".join(mp.nstr(v, 20, strip_zeros=False) for v in p) + ",\n")
        out.write("};\n")
        out.write(f"static const double {name}_W[{n}] = {{\n")
        for w in W:
            out.write("    " + mp.nstr(w, 20, strip_zeros=False) + ",\n")
        out.write("};\n\n")
    out.write("#define ANIFEM_TET_MAX_ORDER 6\n")
    out.write("#define ANIFEM_TET_MAX_NPTS 24\n")
    npts = [0] * 7
    for orders, name, n, deg in names:
        for o in orders:
            npts[o] = n
    out.write("static const int ANIFEM_TET_NPTS_BY_ORDER[7] = {" + ", ".join(map(str, npts)) + "};\n")


if __name__ == "__main__":
    main()
