#!/usr/bin/env python3
"""Generate the symmetric tetrahedron quadrature tables used by the HIP kernels.

AniFem++ selects its rule in ``tetrahedron_quadrature_formulas(order)``
(anifem++/fem/quadrature_formulas.cpp:526-1503, order map :1493-1501) for
orders 0..20.  This script tabulates orders 0..6:

    order 0,1 -> 1 point (centroid)
    order 2   -> 4 points, one S4 orbit                     (degree 2)
    order 3   -> 8 points, two S4 orbits, equal weights     (degree 3)
    order 4,5 -> 14 points, two S4 orbits + one S6 orbit    (degree 5)
    order 6   -> 24 points, three S4 orbits + one S12 orbit (degree 6)

Orders 7..20 are built at run time by basis.cpp as conical (collapsed) products
of Gauss-Jacobi rules; see quadrature() there.

The point/weight values are not copied: for the 4/14/24-point structures this
script solves the moment equations
    sum_i w_i lambda(p_i)^alpha = 3! alpha! / (|alpha|+3)!
for every barycentric monomial of degree <= d by a seeded multi-start
least-squares search (numpy/scipy) and then refines the single admissible
solution (all points interior, all weights positive) to 50 digits with
Gauss-Newton in mpmath.  For these structures the admissible solution is
unique, so the result equals the reference's rule up to rounding.

Two S4 orbits of degree 3 leave a one-parameter family (three moment
equations, four unknowns); the reference's 8-point rule is one member of it.
This script takes the member with all eight weights equal to 1/8.  With the
distinct coordinate of an orbit written 1/4 + d (the other three 1/4 - d/3),
the moments of l0^2 and l0^3 read d1^2 + d2^2 = 9/40 and d1^3 + d2^3 = 9/160,
so s = d1 + d2 solves 80 s^3 - 54 s + 9 = 0 and d1, d2 are the roots of
t^2 - s t + (s^2 - 9/40)/2.  Exactly one root s gives real d1 != d2 with every
coordinate in (0, 1).  Its points differ from the reference's member (the
reference's weights are unequal); both integrate every polynomial of degree
<= 3 exactly, so results agree to rounding for such integrands.

Orbit point order: an S4 orbit lists the point whose distinct coordinate sits
at position 0,1,2,3 (the same order as the reference's TWIST4 macro); S6 and
S12 orbits use lexicographic position order.  The 8-point rule lists the orbit
with the smaller distinct coordinate first.

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
    ((3,), ["s4", "s4"], 3),
    ((4, 5), ["s4", "s4", "s6"], 5),
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


def equal_weight_rule8():
    """The 8-point degree-3 rule with all weights 1/8 (closed form above):
    the S4 parameters (repeated coordinates) of the two orbits and the weights."""
    sols = []
    for s in mp.polyroots([80, 0, -54, 9], maxsteps=200, extraprec=200):
        s = mp.re(s)
        p = (s * s - mp.mpf(9) / 40) / 2
        disc = s * s - 4 * p
        if disc <= 0:
            continue
        d = sorted([(s - mp.sqrt(disc)) / 2, (s + mp.sqrt(disc)) / 2])
        coords = [mp.mpf(1) / 4 + x for x in d] + [mp.mpf(1) / 4 - x / 3 for x in d]
        if all(0 < c < 1 for c in coords):
            sols.append(d)
    if len(sols) != 1:
        raise RuntimeError(f"equal-weight 8-point rule: {len(sols)} admissible roots")
    d1, d2 = sols[0]
    return [mp.mpf(1) / 4 - d1 / 3, mp.mpf(1) / 4 - d2 / 3, mp.mpf(1) / 8, mp.mpf(1) / 8]


def check_exact(P, W, deg):
    worst = max(abs(sum(W[i] * P[i][0] ** m[0] * P[i][1] ** m[1] * P[i][2] ** m[2] * P[i][3] ** m[3]
                        for i in range(len(P))) - mp.mpf(6) * mp.factorial(m[0]) * mp.factorial(m[1])
                    * mp.factorial(m[2]) * mp.factorial(m[3]) / mp.factorial(sum(m) + 3)) for m in monomials(deg))
    if worst > mp.mpf(10) ** -40:
        raise RuntimeError(f"rule of degree {deg} is not exact ({worst})")


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
        elif struct == ["s4", "s4"] and deg == 3:
            P, W = expand(struct, equal_weight_rule8(), mp.mpf(1))
        else:
            z = refine(struct, deg, search(struct, deg))
            P, W = expand(struct, z, mp.mpf(1))
        check_exact(P, W, deg)
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
    out.write("// orders above ANIFEM_TET_TABLE_MAX_ORDER are built at run time (basis.cpp)\n")
    out.write("#define ANIFEM_TET_TABLE_MAX_ORDER 6\n")
    out.write("#define ANIFEM_TET_TABLE_MAX_NPTS 24\n")
    npts = [0] * 7
    for orders, name, n, deg in names:
        for o in orders:
            npts[o] = n
    out.write("static const int ANIFEM_TET_NPTS_BY_ORDER[7] = {" + ", ".join(map(str, npts)) + "};\n")


if __name__ == "__main__":
    main()
