/* FP64 CPU oracle for the AniFem++ element-matrix and global-assembly path.
 *
 * TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the bench's
 * CPU-baseline leg use it as the checker.  The product never links it.
 *
 * Element part: a restatement of the reference's per-tet algorithm
 *   geometry   core.inl:218-270, geometry.h:25-45
 *   operators  poly_0.h:62-81, poly_1.h:12-133, poly_2.h:12-102, poly_3.h:12-119,
 *              operators.h:127-155 (FemVec block structure); spaces/
 *              crouzeix_raviart_1.h, bubble_2.h, bubble_4.h,
 *              raviant_thomas_0.h, nedelec_0.h (physical-space formulas)
 *   D*U        diff_tensor.h:62-97 (DFunc wrapper), :280-995 (applyDU)
 *   V^T*DU     core.inl:29-60 (fusive_AT_mul_B), :278-367
 * in the same data layouts (U[k + dim*(n + q*i)], A[ib + nfB*ia]) and the same
 * summation order (j = k + dimB*n).
 *
 * Assembly part: a restatement of assembler.inl:114-184,314-488,
 * ordering.inl:8-172, tetdofmap.inl:98-127 and the NATURAL enumeration of
 * global_enumerator.cpp:562-777.  INMOST (which assigns edge/face ids in the
 * reference) is absent; the element-id rule used here is stated at
 * orc_enumerate.
 */
#include "anifem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Quadrature (quadrature_formulas.cpp:526-1503 selects these orbit shapes). */

typedef long double ld;

static void orc_orbit(int kind, const ld* p, ld pts[][4], int* np) {
    int m = 0;
    if (kind == 1) {
        for (int k = 0; k < 4; ++k) pts[m][k] = 0.25L;
        m = 1;
    } else if (kind == 4) {
        for (int i = 0; i < 4; ++i, ++m) {
            for (int k = 0; k < 4; ++k) pts[m][k] = p[0];
            pts[m][i] = 1.0L - 3.0L * p[0];
        }
    } else if (kind == 6) {
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j, ++m) {
                for (int k = 0; k < 4; ++k) pts[m][k] = 0.5L - p[0];
                pts[m][i] = p[0];
                pts[m][j] = p[0];
            }
    } else if (kind == 12) {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                if (i == j) continue;
                for (int k = 0; k < 4; ++k) pts[m][k] = p[0];
                pts[m][i] = p[1];
                pts[m][j] = 1.0L - 2.0L * p[0] - p[1];
                ++m;
            }
    }
    *np = m;
}

typedef struct {
    int nkind;
    int kinds[4];
    int deg;
    ld seed[9];
} orc_rule_spec;

/* Coarse seeds (4 digits) of the unique admissible solutions; Gauss-Newton
 * polishes them against the moment equations. */
static const orc_rule_spec ORC_RULES[3] = {
    {1, {4}, 2, {0.1382L, 0.25L}},
    {3, {4, 4, 6}, 5, {0.3109L, 0.09274L, 0.4545L, 0.1127L, 0.07349L, 0.04255L}},
    {4, {4, 4, 4, 12}, 6, {0.3223L, 0.2146L, 0.04067L, 0.06366L, 0.6030L, 0.05536L, 0.03992L, 0.01008L, 0.04821L}},
};

static int orc_nparam(int kind) { return kind == 12 ? 2 : (kind == 1 ? 0 : 1); }

static int orc_expand(const orc_rule_spec* s, const ld* z, ld pts[][4], ld* w) {
    int k = 0, m = 0, nw = 0;
    for (int o = 0; o < s->nkind; ++o) nw += 1;
    const ld* wv = z;
    for (int o = 0; o < s->nkind; ++o) k += orc_nparam(s->kinds[o]);
    wv = z + k;
    k = 0;
    for (int o = 0; o < s->nkind; ++o) {
        int np;
        orc_orbit(s->kinds[o], z + k, pts + m, &np);
        for (int i = 0; i < np; ++i) w[m + i] = wv[o];
        m += np;
        k += orc_nparam(s->kinds[o]);
    }
    (void)nw;
    return m;
}

static ld orc_factl(int n) {
    ld r = 1;
    for (int i = 2; i <= n; ++i) r *= i;
    return r;
}

static int orc_monos(int deg, int (*al)[4]) {
    int m = 0;
    for (int s = 0; s <= deg; ++s)
        for (int a = 0; a <= s; ++a)
            for (int b = 0; b <= s - a; ++b)
                for (int c = 0; c <= s - a - b; ++c) {
                    al[m][0] = a, al[m][1] = b, al[m][2] = c, al[m][3] = s - a - b - c;
                    ++m;
                }
    return m;
}

static void orc_resid(const orc_rule_spec* s, const ld* z, int nm, int (*al)[4], ld* r) {
    ld pts[24][4], w[24];
    int np = orc_expand(s, z, pts, w);
    for (int j = 0; j < nm; ++j) {
        ld exact = 6.0L * orc_factl(al[j][0]) * orc_factl(al[j][1]) * orc_factl(al[j][2]) * orc_factl(al[j][3]) /
                   orc_factl(al[j][0] + al[j][1] + al[j][2] + al[j][3] + 3);
        ld acc = 0;
        for (int i = 0; i < np; ++i) {
            ld v = w[i];
            for (int k = 0; k < 4; ++k)
                for (int e = 0; e < al[j][k]; ++e) v *= pts[i][k];
            acc += v;
        }
        r[j] = acc - exact;
    }
}

/* Small dense solve (Gaussian elimination with partial pivoting). */
static void orc_solve_ld(int n, ld* M, ld* b) {
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (fabsl(M[r * n + c]) > fabsl(M[p * n + c])) p = r;
        if (p != c) {
            for (int k = 0; k < n; ++k) {
                ld t = M[c * n + k];
                M[c * n + k] = M[p * n + k];
                M[p * n + k] = t;
            }
            ld t = b[c];
            b[c] = b[p];
            b[p] = t;
        }
        for (int r = c + 1; r < n; ++r) {
            ld f = M[r * n + c] / M[c * n + c];
            for (int k = c; k < n; ++k) M[r * n + k] -= f * M[c * n + k];
            b[r] -= f * b[c];
        }
    }
    for (int c = n - 1; c >= 0; --c) {
        ld s = b[c];
        for (int k = c + 1; k < n; ++k) s -= M[c * n + k] * b[k];
        b[c] = s / M[c * n + c];
    }
}

static int orc_rule_cache_ready[3];
static double orc_rule_p[3][4 * 24], orc_rule_w[3][24];
static int orc_rule_n[3];

static void orc_build_rule(int id) {
    const orc_rule_spec* s = &ORC_RULES[id];
    static int al[300][4];
    int nm = orc_monos(s->deg, al);
    int nz = s->nkind;
    for (int o = 0; o < s->nkind; ++o) nz += orc_nparam(s->kinds[o]);
    ld z[9];
    for (int i = 0; i < nz; ++i) z[i] = s->seed[i];
    ld r[300], r2[300], J[300][9];
    for (int it = 0; it < 50; ++it) {
        orc_resid(s, z, nm, al, r);
        ld h = 1e-9L;
        for (int k = 0; k < nz; ++k) {
            ld zk = z[k];
            z[k] = zk + h;
            orc_resid(s, z, nm, al, r2);
            z[k] = zk - h;
            ld r3[300];
            orc_resid(s, z, nm, al, r3);
            z[k] = zk;
            for (int j = 0; j < nm; ++j) J[j][k] = (r2[j] - r3[j]) / (2 * h);
        }
        ld M[81], g[9];
        for (int a = 0; a < nz; ++a) {
            g[a] = 0;
            for (int j = 0; j < nm; ++j) g[a] -= J[j][a] * r[j];
            for (int b = 0; b < nz; ++b) {
                ld acc = 0;
                for (int j = 0; j < nm; ++j) acc += J[j][a] * J[j][b];
                M[a * nz + b] = acc;
            }
        }
        orc_solve_ld(nz, M, g);
        ld step = 0;
        for (int k = 0; k < nz; ++k) {
            z[k] += g[k];
            if (fabsl(g[k]) > step) step = fabsl(g[k]);
        }
        if (step < 1e-19L) break;
    }
    ld pts[24][4], w[24];
    int np = orc_expand(s, z, pts, w);
    for (int i = 0; i < np; ++i) {
        for (int k = 0; k < 4; ++k) orc_rule_p[id][4 * i + k] = (double)pts[i][k];
        orc_rule_w[id][i] = (double)w[i];
    }
    orc_rule_n[id] = np;
    orc_rule_cache_ready[id] = 1;
}

/* Order 3: the 8-point rule of two S4 orbits with all weights 1/8 (the
 * reference's order-3 rule, quadrature_formulas.cpp:542-549, is another member
 * of the same one-parameter family).  Newton on the moments of l0^2 and l0^3
 * (1/10 and 1/20) in the two distinct coordinates a1 < a2, from coarse seeds. */
static void orc_rule8(double* p, double* w) {
    ld a[2] = {0.0134L, 0.6611L};
    for (int it = 0; it < 60; ++it) {
        ld f[2] = {-0.1L, -0.05L}, J[2][2];
        for (int o = 0; o < 2; ++o) {
            ld b = (1 - a[o]) / 3;
            f[0] += (a[o] * a[o] + 3 * b * b) / 8;
            f[1] += (a[o] * a[o] * a[o] + 3 * b * b * b) / 8;
            J[0][o] = (2 * a[o] - 2 * b) / 8;
            J[1][o] = (3 * a[o] * a[o] - 3 * b * b) / 8;
        }
        ld det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        ld d0 = (f[0] * J[1][1] - f[1] * J[0][1]) / det, d1 = (J[0][0] * f[1] - J[1][0] * f[0]) / det;
        a[0] -= d0, a[1] -= d1;
        if (fabsl(d0) + fabsl(d1) < 1e-20L) break;
    }
    for (int o = 0; o < 2; ++o)
        for (int i = 0; i < 4; ++i) {
            for (int k = 0; k < 4; ++k) p[4 * (4 * o + i) + k] = (double)((1 - a[o]) / 3);
            p[4 * (4 * o + i) + i] = (double)a[o];
            w[4 * o + i] = 0.125;
        }
}

/* Orders 7..20: Stroud's conical product.  x = u, y = (1-u) v,
 * z = (1-u)(1-v) s maps the unit cube onto the tet with volume element
 * (1-u)^2 (1-v); Gauss-Jacobi rules with m = order/2 + 1 points for the
 * weights (1-u)^2, (1-v), 1 on [0,1].  Nodes: zeros of the Jacobi polynomial
 * P_m^(alpha,0)(x), x = 2t - 1 (standard three-term recurrence), by Newton
 * with deflation; weights from the closed form for beta = 0,
 * w = 2^(alpha+1) / ((1 - x^2) P_m'(x)^2) on [-1,1], i.e.
 * 1 / ((1 - x^2) P_m'(x)^2) on [0,1].  Point order: u outer, s inner, each
 * ascending; weights scaled to sum to 1. */
static void orc_jacobi(int m, int alpha, ld x, ld* P, ld* dP) {
    ld al = alpha, p0 = 1, p1 = ((al + 2) * x + al) / 2;
    for (int n = 2; n <= m; ++n) {
        ld c = 2 * n + al;
        ld p2 = ((c - 1) * (c * (c - 2) * x + al * al) * p1 - 2 * (n + al - 1) * (n - 1) * c * p0) /
                (2 * n * (n + al) * (c - 2));
        p0 = p1, p1 = p2;
    }
    *P = m == 0 ? 1 : p1;
    /* (2m+a)(1-x^2) P_m' = m (a - (2m+a) x) P_m + 2 (m+a) m P_{m-1} */
    ld c = 2 * m + al;
    *dP = (m * (al - c * x) * p1 + 2 * (m + al) * m * p0) / (c * (1 - x * x));
}

static void orc_gauss_jacobi01(int m, int alpha, ld* t, ld* w) {
    ld x[16];
    const ld pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < m; ++i) {
        ld z = cosl(pi * (i + 0.75L) / (m + 0.5L)), P, dP;
        for (int it = 0; it < 200; ++it) {
            orc_jacobi(m, alpha, z, &P, &dP);
            ld defl = 0;
            for (int j = 0; j < i; ++j) defl += 1 / (z - x[j]);
            ld dz = P / (dP - P * defl);
            z -= dz;
            if (fabsl(dz) < 1e-19L) break;
        }
        x[i] = z;
    }
    for (int i = 1; i < m; ++i) /* ascending */
        for (int j = i; j > 0 && x[j - 1] > x[j]; --j) {
            ld tmp = x[j];
            x[j] = x[j - 1];
            x[j - 1] = tmp;
        }
    for (int i = 0; i < m; ++i) {
        ld P, dP;
        orc_jacobi(m, alpha, x[i], &P, &dP);
        t[i] = (1 + x[i]) / 2;
        w[i] = 1 / ((1 - x[i] * x[i]) * dP * dP);
    }
}

#define ORC_MAX_ORDER 20
#define ORC_MAX_Q 1331

/* Computed on every call (no shared state: callers may run in threads). */
static int orc_product_rule(int order, double* p, double* w) {
    int m = order / 2 + 1;
    ld tu[16], wu[16], tv[16], wv[16], ts[16], ws[16];
    orc_gauss_jacobi01(m, 2, tu, wu);
    orc_gauss_jacobi01(m, 1, tv, wv);
    orc_gauss_jacobi01(m, 0, ts, ws);
    int n = 0;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j)
            for (int k = 0; k < m; ++k, ++n) {
                ld x = tu[i], y = (1 - tu[i]) * tv[j], z = (1 - tu[i]) * (1 - tv[j]) * ts[k];
                if (p) {
                    p[4 * n] = (double)(1 - x - y - z);
                    p[4 * n + 1] = (double)x;
                    p[4 * n + 2] = (double)y;
                    p[4 * n + 3] = (double)z;
                }
                if (w) w[n] = (double)(6 * wu[i] * wv[j] * ws[k]);
            }
    return n;
}

int orc_quadrature(int order, double* p, double* w) {
    if (order < 0 || order > ORC_MAX_ORDER) return -1;
    if (order <= 1) {
        if (p)
            for (int k = 0; k < 4; ++k) p[k] = 0.25;
        if (w) w[0] = 1.0;
        return 1;
    }
    if (order == 3) {
        double pp[32], ww[8];
        orc_rule8(pp, ww);
        if (p) memcpy(p, pp, sizeof pp);
        if (w) memcpy(w, ww, sizeof ww);
        return 8;
    }
    if (order >= 7) return orc_product_rule(order, p, w);
    int id = order == 2 ? 0 : (order <= 5 ? 1 : 2);
    if (!orc_rule_cache_ready[id]) orc_build_rule(id);
    int q = orc_rule_n[id];
    if (p) memcpy(p, orc_rule_p[id], sizeof(double) * 4 * q);
    if (w) memcpy(w, orc_rule_w[id], sizeof(double) * q);
    return q;
}

/* ------------------------------------------------------------------------- */
/* Geometry: geometry.h:25-45, core.inl:218-270                              */

double orc_inverse3x3(const double* m, double* inv) {
#define M_(I, J) m[(I) + 3 * (J)]
    double da = M_(1, 1) * M_(2, 2) - M_(1, 2) * M_(2, 1);
    double db = M_(1, 2) * M_(2, 0) - M_(1, 0) * M_(2, 2);
    double dc = M_(1, 0) * M_(2, 1) - M_(1, 1) * M_(2, 0);
    double det = M_(0, 0) * da + M_(0, 1) * db + M_(0, 2) * dc;
    inv[0] = da / det;
    inv[1] = db / det;
    inv[2] = dc / det;
    for (int i = 1; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            inv[j + 3 * i] = (M_((i + 1) % 3, (j + 1) % 3) * M_((i + 2) % 3, (j + 2) % 3) -
                              M_((i + 1) % 3, (j + 2) % 3) * M_((i + 2) % 3, (j + 1) % 3)) /
                             det;
#undef M_
    return det;
}

/* ------------------------------------------------------------------------- */
/* Operators on scalar spaces (poly_*.h).  Layout of the scalar-space table:
 * U[k + dim*(n + q*i)], i over the nf basis functions.                       */

static int orc_nf_scalar(int fem) {
    switch (fem) {
        case ORC_P0: return 1;
        case ORC_P1: return 4;
        case ORC_P2: return 10;
        case ORC_P3: return 20;
        case ORC_CR1: return 4;
        case ORC_B2: return 6;
        case ORC_B4: return 1;
        case ORC_RT0: return 4;
        case ORC_ND0: return 6;
    }
    return -1;
}

static int orc_is_vector(int fem) { return fem == ORC_RT0 || fem == ORC_ND0; }

/* operator dimension on (fem, ncomp): raviant_thomas_0.h and nedelec_0.h
 * declare Dim 3 for IDEN and CURL, 1 for DIV */
static int orc_opdim(int op, int fem, int ncomp) {
    if (orc_is_vector(fem)) return op == ORC_DIV ? 1 : (op == ORC_GRAD ? 9 : 3);
    return orc_dim(op, ncomp);
}

int orc_nfa(int fem, int ncomp) {
    int nf = orc_nf_scalar(fem);
    return nf < 0 ? nf : nf * ncomp;
}

int orc_dim(int op, int ncomp) {
    switch (op) {
        case ORC_IDEN: return ncomp;
        case ORC_GRAD: return 3 * ncomp;
        case ORC_DIV: return 1;
        case ORC_CURL: return 3;
        default: return ncomp;
    }
}

/* operator code valid on (fem, ncomp): operators.h:37-45, :321-400 */
static int orc_op_ok(int op, int fem, int ncomp) {
    if (orc_is_vector(fem)) return ncomp == 1 && (op == ORC_IDEN || op == ORC_GRAD || op == ORC_DIV || op == ORC_CURL);
    if (op != ORC_IDEN && op != ORC_GRAD && op != ORC_DIV && op != ORC_CURL && (op < ORC_DUDX || op > ORC_DUDZ))
        return 0;
    if (op != ORC_IDEN && fem == ORC_P0) return 0;
    if ((op == ORC_DIV || op == ORC_CURL) && ncomp != 3) return 0;
    return 1;
}

static const unsigned char ORC_EI[6] = {0, 0, 0, 1, 1, 2}, ORC_EJ[6] = {1, 2, 3, 2, 3, 3};
static const unsigned char ORC_IPF[16] = {0, 1, 2, 3, 1, 2, 3, 0, 0, 2, 3, 1, 0, 1, 3, 2};
static const double ORC_GP1[12] = {-1, -1, -1, 1, 0, 0, 0, 1, 0, 0, 0, 1}; /* col i = grad lambda_i */

/* Basis values (IDEN) at one point (poly_0.h:62-81, poly_1.h:12-40, poly_2.h:12-49, poly_3.h:12-55). */
static void orc_iden(int fem, const double* L, double* phi) {
    switch (fem) {
        case ORC_P0: phi[0] = 1; break;
        case ORC_P1:
            for (int i = 0; i < 4; ++i) phi[i] = L[i];
            break;
        case ORC_P2:
            for (int i = 0; i < 4; ++i) phi[i] = L[i] * (2 * L[i] - 1);
            for (int e = 0; e < 6; ++e) phi[4 + e] = 4 * L[ORC_EI[e]] * L[ORC_EJ[e]];
            break;
        case ORC_P3:
            for (int i = 0; i < 4; ++i) phi[i] = L[i] * (3 * L[i] - 1) * (3 * L[i] - 2) / 2;
            for (int e = 0; e < 6; ++e) {
                double s1 = L[ORC_EI[e]], s2 = L[ORC_EJ[e]];
                phi[4 + 2 * e] = s1 * (3 * s1 - 1) * s2 * 4.5;
                phi[4 + 2 * e + 1] = s1 * (3 * s2 - 1) * s2 * 4.5;
            }
            for (int f = 0; f < 4; ++f)
                phi[16 + f] = 27 * L[ORC_IPF[4 * f]] * L[ORC_IPF[4 * f + 1]] * L[ORC_IPF[4 * f + 2]];
            break;
        case ORC_CR1: /* crouzeix_raviart_1.h:34-37: 1 - 3 l_{IPF3[i]}, IPF3 = {3, 0, 1, 2} */
            for (int i = 0; i < 4; ++i) phi[i] = 1 - 3 * L[(i + 3) % 4];
            break;
        case ORC_B2: /* bubble_2.h:31-34 */
            for (int e = 0; e < 6; ++e) phi[e] = 4 * L[ORC_EI[e]] * L[ORC_EJ[e]];
            break;
        case ORC_B4: /* bubble_4.h:30-33 */
            phi[0] = 256 * L[0] * L[1] * L[2] * L[3];
            break;
    }
}

/* Reference gradients at one point, g[d + 3*i] (poly_1.h:91-133,
 * poly_2.h:52-102, poly_3.h:58-119). */
static void orc_refgrad(int fem, const double* L, double* g) {
    switch (fem) {
        case ORC_P0: g[0] = g[1] = g[2] = 0; break;
        case ORC_P1:
            for (int k = 0; k < 12; ++k) g[k] = ORC_GP1[k];
            break;
        case ORC_P2:
            for (int i = 0; i < 4; ++i)
                for (int d = 0; d < 3; ++d) g[d + 3 * i] = ORC_GP1[d + 3 * i] * (4 * L[i] - 1);
            for (int e = 0; e < 6; ++e) {
                double s1 = L[ORC_EI[e]], s2 = L[ORC_EJ[e]];
                for (int d = 0; d < 3; ++d)
                    g[d + 3 * (4 + e)] = 4 * (ORC_GP1[d + 3 * ORC_EJ[e]] * s1 + ORC_GP1[d + 3 * ORC_EI[e]] * s2);
            }
            break;
        case ORC_P3:
            for (int i = 0; i < 4; ++i) {
                double s1 = L[i];
                for (int d = 0; d < 3; ++d) g[d + 3 * i] = ORC_GP1[d + 3 * i] * ((13.5 * s1 - 9) * s1 + 1);
            }
            for (int e = 0; e < 6; ++e) {
                int I = ORC_EI[e], J = ORC_EJ[e];
                double s1 = L[I], s2 = L[J];
                for (int d = 0; d < 3; ++d) {
                    g[d + 3 * (4 + 2 * e)] =
                        (ORC_GP1[d + 3 * I] * (6 * s1 - 1) * s2 + ORC_GP1[d + 3 * J] * (3 * s1 - 1) * s1) * 4.5;
                    g[d + 3 * (4 + 2 * e + 1)] =
                        (ORC_GP1[d + 3 * J] * (6 * s2 - 1) * s1 + ORC_GP1[d + 3 * I] * (3 * s2 - 1) * s2) * 4.5;
                }
            }
            for (int f = 0; f < 4; ++f) {
                int a = ORC_IPF[4 * f], b = ORC_IPF[4 * f + 1], c = ORC_IPF[4 * f + 2];
                double s1 = L[a], s2 = L[b], s3 = L[c];
                for (int d = 0; d < 3; ++d)
                    g[d + 3 * (16 + f)] = 27 * (ORC_GP1[d + 3 * a] * s2 * s3 + s1 * ORC_GP1[d + 3 * b] * s3 +
                                               s1 * s2 * ORC_GP1[d + 3 * c]);
            }
            break;
        case ORC_CR1: /* crouzeix_raviart_1.h:61-73: -3 grad l_{IPF3[i]} */
            for (int i = 0; i < 4; ++i)
                for (int d = 0; d < 3; ++d) g[d + 3 * i] = -3 * ORC_GP1[d + 3 * ((i + 3) % 4)];
            break;
        case ORC_B2: /* bubble_2.h: 4 (l_I grad l_J + l_J grad l_I) */
            for (int e = 0; e < 6; ++e)
                for (int d = 0; d < 3; ++d)
                    g[d + 3 * e] = 4 * (ORC_GP1[d + 3 * ORC_EI[e]] * L[ORC_EJ[e]] + L[ORC_EI[e]] * ORC_GP1[d + 3 * ORC_EJ[e]]);
            break;
        case ORC_B4: { /* bubble_4.h:70-80: 256 sum_v grad l_v prod_{w != v} l_w */
            double l[4] = {L[1] * L[2] * L[3], L[0] * L[2] * L[3], L[0] * L[1] * L[3], L[0] * L[1] * L[2]};
            for (int d = 0; d < 3; ++d)
                g[d] = 256 * (ORC_GP1[d] * l[0] + ORC_GP1[d + 3] * l[1] + ORC_GP1[d + 6] * l[2] + ORC_GP1[d + 9] * l[3]);
            break;
        }
    }
}

/* (p0 - p2) x (p1 - p2) (geometry.h:55-63) */
static void orc_cross3(const double* p0, const double* p1, const double* p2, double* c) {
    double a[3] = {p0[0] - p2[0], p0[1] - p2[1], p0[2] - p2[2]}, b[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = -a[0] * b[2] + a[2] * b[0];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* Operator tables of the vector-valued spaces for one tet with vertices X[3v + j]
 * and det J = det: U[k + dim*(n + q*i)].
 *   RT0 IDEN raviant_thomas_0.h:35-47, DIV :136-160, CURL :190-196 (zero)
 *   ND0 IDEN nedelec_0.h:35-53, CURL :100-135, DIV :74-78 (zero)            */
static void orc_vector_op(int op, int fem, int q, const double* XYL, const double* PSI, const double* X, double det,
                          double* U) {
    static const unsigned char IPR[24] = {0, 1, 2, 3, 0, 2, 3, 1, 0, 3, 1, 2, 1, 2, 0, 3, 1, 3, 2, 0, 2, 3, 0, 1};
    int dim = orc_opdim(op, fem, 1), nf = orc_nf_scalar(fem);
    memset(U, 0, sizeof(double) * (size_t)dim * q * nf);
    double V[4][3]; /* V[v][j] = d lambda_v / d x_j */
    for (int v = 0; v < 4; ++v)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += ORC_GP1[k + 3 * v] * PSI[k + 3 * j];
            V[v][j] = s;
        }
    if (fem == ORC_RT0) {
        if (op == ORC_CURL) return;
        for (int i = 0; i < 4; ++i) {
            const unsigned char* ip = ORC_IPF + 4 * i;
            const double *x0 = X + 3 * ip[0], *x1 = X + 3 * ip[1], *x2 = X + 3 * ip[2], *x3 = X + 3 * ip[3];
            double c[3];
            orc_cross3(x0, x1, x2, c);
            double sqr = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) / 2 / (fabs(det) / 2);
            if (op == ORC_IDEN) {
                for (int n = 0; n < q; ++n)
                    for (int j = 0; j < 3; ++j)
                        U[j + 3 * (n + q * i)] = sqr * (XYL[ip[0] + 4 * n] * (x0[j] - x3[j]) +
                                                        XYL[ip[1] + 4 * n] * (x1[j] - x3[j]) +
                                                        XYL[ip[2] + 4 * n] * (x2[j] - x3[j]));
            } else if (op == ORC_GRAD) {
                /* grad phi_i, entry k + 3 j: sqr sum_m dlambda_{ip[m]}/dx_k (x_{ip[m]} - x_{ip[3]})_j
                 * (raviant_thomas_0.h:101-104), the same at every point */
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) {
                        double g = sqr * (V[ip[0]][k] * (x0[j] - x3[j]) + V[ip[1]][k] * (x1[j] - x3[j]) +
                                          V[ip[2]][k] * (x2[j] - x3[j]));
                        for (int n = 0; n < q; ++n) U[k + 3 * j + 9 * (n + q * i)] = g;
                    }
            } else {
                double s = 0;
                for (int j = 0; j < 3; ++j)
                    s += V[ip[0]][j] * (x0[j] - x3[j]) + V[ip[1]][j] * (x1[j] - x3[j]) + V[ip[2]][j] * (x2[j] - x3[j]);
                for (int n = 0; n < q; ++n) U[n + q * i] = sqr * s;
            }
        }
        return;
    }
    if (op == ORC_DIV) return;
    for (int i = 0; i < 6; ++i) {
        const unsigned char* ip = IPR + 4 * i;
        const double *a = X + 3 * ip[0], *b = X + 3 * ip[1];
        double len = sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
        double xyn[3], xym[3];
        orc_cross3(X + 3 * ip[2], X + 3 * ip[3], X + 3 * ip[0], xyn);
        orc_cross3(X + 3 * ip[2], X + 3 * ip[3], X + 3 * ip[1], xym);
        double sqr = len / det;
        if (op == ORC_IDEN) {
            for (int n = 0; n < q; ++n)
                for (int k = 0; k < 3; ++k)
                    U[k + 3 * (n + q * i)] = sqr * (xyn[k] * XYL[ip[0] + 4 * n] + xym[k] * XYL[ip[1] + 4 * n]);
        } else {
            double cu[3];
            for (int k = 0; k < 3; ++k) {
                int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
                cu[k] = sqr * ((V[ip[0]][k1] * xyn[k2] + V[ip[1]][k1] * xym[k2]) -
                               (V[ip[0]][k2] * xyn[k1] + V[ip[1]][k2] * xym[k1]));
            }
            if (op == ORC_GRAD) {
                /* grad phi_i: the skew matrix of curl/2, entries 0..8 as
                 * nedelec_0.h:175-179 lays them out */
                double a0 = cu[0] / 2, a1 = cu[1] / 2, a2 = cu[2] / 2;
                const double g[9] = {0, -a2, a1, a2, 0, -a0, -a1, a0, 0};
                for (int n = 0; n < q; ++n)
                    for (int l = 0; l < 9; ++l) U[l + 9 * (n + q * i)] = g[l];
            } else {
                for (int k = 0; k < 3; ++k)
                    for (int n = 0; n < q; ++n) U[k + 3 * (n + q * i)] = cu[k];
            }
        }
    }
}

/* Full (FemVec-expanded) operator table U[k + dim*(n + q*i)] for one tet. */
static void orc_apply_op(int op, int fem, int ncomp, int q, const double* XYL, const double* PSI, const double* X,
                         double det, double* U) {
    if (orc_is_vector(fem)) {
        orc_vector_op(op, fem, q, XYL, PSI, X, det, U);
        return;
    }
    int nf0 = orc_nf_scalar(fem), dim0 = (op == ORC_IDEN ? 1 : 3);
    int dim = orc_dim(op, ncomp), nf = nf0 * ncomp;
    memset(U, 0, sizeof(double) * (size_t)dim * q * nf);
    double tab[3 * 20];
    for (int n = 0; n < q; ++n) {
        const double* L = XYL + 4 * n;
        if (op == ORC_IDEN) {
            orc_iden(fem, L, tab);
        } else {
            double g[3 * 20];
            orc_refgrad(fem, L, g);
            if (fem == ORC_P1) {
                /* poly_1.h:112-126 forms the gradient rows directly from PSI */
                for (int k = 0; k < 3; ++k) {
                    tab[k] = -PSI[0 + 3 * k] - PSI[1 + 3 * k] - PSI[2 + 3 * k];
                    for (int i = 1; i < 4; ++i) tab[k + 3 * i] = PSI[(i - 1) + 3 * k];
                }
            } else {
                for (int i = 0; i < nf0; ++i)
                    for (int k = 0; k < 3; ++k) {
                        double s = 0;
                        for (int j = 0; j < 3; ++j) s += PSI[j + 3 * k] * g[j + 3 * i];
                        tab[k + 3 * i] = s;
                    }
            }
        }
        for (int c = 0; c < ncomp; ++c)
            for (int i = 0; i < nf0; ++i) {
                double* u = U + dim * (n + q * (i + nf0 * c));
                const double* t = tab + dim0 * i;
                switch (op) {
                    case ORC_IDEN: u[c] = t[0]; break;
                    case ORC_GRAD:
                        for (int k = 0; k < 3; ++k) u[k + 3 * c] = t[k];
                        break;
                    case ORC_DIV: u[0] = t[c]; break; /* operators.h:321-349 */
                    case ORC_CURL:                    /* operators.h:356-400: (curl u)_i = eps_idc du_c/dx_d */
                        for (int d = 0; d < 3; ++d)
                            if (d != c) {
                                int ic = 3 - d - c;
                                u[ic] = (d == (ic + 1) % 3) ? t[d] : -t[d];
                            }
                        break;
                    default: u[c] = t[op - ORC_DUDX]; /* operators.h:286-316 */
                }
            }
    }
}

/* ------------------------------------------------------------------------- */
/* fem3Dtet (core.inl:278-367 with applyDU of diff_tensor.h:280-995)          */

int orc_fem3dtet(int opA, int femA, int ncA, int opB, int femB, int ncB, int tensor_type, int coef_mode,
                 int order, long n, const double* xyz, const double* D, double* A) {
    int nfA = orc_nfa(femA, ncA), nfB = orc_nfa(femB, ncB);
    if (nfA <= 0 || nfB <= 0 || !orc_op_ok(opA, femA, ncA) || !orc_op_ok(opB, femB, ncB)) return -1;
    double* XYL = malloc(sizeof(double) * 4 * ORC_MAX_Q);
    double* WG = malloc(sizeof(double) * ORC_MAX_Q);
    int q = orc_quadrature(order, XYL, WG);
    if (q <= 0) {
        free(XYL);
        free(WG);
        return -2;
    }
    int dimA = orc_opdim(opA, femA, ncA), dimB = orc_opdim(opB, femB, ncB);
    int ddim = tensor_type == ORC_TENSOR_NULL ? 0 : (tensor_type == ORC_TENSOR_SCALAR ? 1 : dimA * dimB);
    if ((tensor_type == ORC_TENSOR_NULL || tensor_type == ORC_TENSOR_SCALAR) && dimA != dimB &&
        !(nfA == 1 && dimA == 1)) {
        free(XYL);
        free(WG);
        return -3;
    }
    double* U = malloc(sizeof(double) * (size_t)dimA * q * nfA);
    double* V = malloc(sizeof(double) * (size_t)dimB * q * nfB);
    double* DU = malloc(sizeof(double) * (size_t)dimB * q * nfA);
    double* Dfull = malloc(sizeof(double) * (size_t)(dimA * dimB > 0 ? dimA * dimB : 1));
    for (long r = 0; r < n; ++r) {
        const double* P = xyz + 12 * r;
        double XYP[9], PSI[9];
        for (int l = 0; l < 3; ++l)
            for (int i = 0; i < 3; ++i) XYP[i + 3 * l] = P[3 * (l + 1) + i] - P[i];
        double det = orc_inverse3x3(XYP, PSI);
        double mes = fabs(det) / 6;
        orc_apply_op(opA, femA, ncA, q, XYL, PSI, P, det, U);
        orc_apply_op(opB, femB, ncB, q, XYL, PSI, P, det, V);
        for (int n_ = 0; n_ < q; ++n_) {
            const double* Dn = NULL;
            if (coef_mode == ORC_COEF_CONST) Dn = D;
            else if (coef_mode == ORC_COEF_PER_TET) Dn = D + (size_t)ddim * r;
            else Dn = D + (size_t)ddim * ((size_t)q * r + n_);
            double wv = WG[n_] * mes;
            for (int i = 0; i < nfA; ++i) {
                const double* u = U + dimA * (n_ + q * i);
                double* du = DU + dimB * (n_ + q * i);
                if (tensor_type == ORC_TENSOR_NULL) {
                    /* DU = w|T| U (diff_tensor.h:374-408); for the P0-trial,
                     * vector-test case the first row is replicated (:401-406) */
                    for (int k = 0; k < dimB; ++k) du[k] = wv * (dimA == dimB ? u[k] : u[0]);
                } else if (tensor_type == ORC_TENSOR_SCALAR) {
                    double s = Dn[0];
                    double f = coef_mode == ORC_COEF_CONST
                                   ? WG[n_] * (s * mes)  /* constant tensor: |T| scaled first (:434-440) */
                                   : (s * WG[n_]) * mes; /* tensor evaluated per point (:845-851) */
                    for (int k = 0; k < dimB; ++k) du[k] = f * (dimA == dimB ? u[k] : u[0]);
                } else {
                    /* GENERAL/SYMMETRIC: DU_k = w|T| sum_j D(k,j) U_j with D
                     * column-major dimB x dimA (the DFunc wrapper transposes
                     * GENERAL into DIFF[j + dimA*k], diff_tensor.h:77-83) */
                    for (int k = 0; k < dimB; ++k) {
                        double s = 0;
                        for (int j = 0; j < dimA; ++j) s += Dn[k + dimB * j] * u[j];
                        du[k] = wv * s;
                    }
                }
            }
        }
        /* A = V^T DU, summation over j = k + dimB*n (core.inl:37-47) */
        double* Ar = A + (size_t)nfA * nfB * r;
        int nrow = q * dimB;
        for (int ia = 0; ia < nfA; ++ia)
            for (int ib = 0; ib < nfB; ++ib) {
                double s = 0;
                for (int j = 0; j < nrow; ++j) s += DU[j + nrow * ia] * V[j + nrow * ib];
                Ar[ib + nfB * ia] = s;
            }
    }
    free(U);
    free(V);
    free(DU);
    free(Dfull);
    free(XYL);
    free(WG);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Face integrals: fem3Dface / fem3DfaceN (int_face.inl:12-82)                */

/* Gauss-Legendre nodes and weights on [0, 1] (Newton on P_m, long double). */
static void orc_gauss01(int m, ld* x, ld* w) {
    const ld pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < m; ++i) {
        ld z = cosl(pi * (i + 0.75L) / (m + 0.5L)), dp = 0;
        for (int it = 0; it < 100; ++it) {
            ld p0 = 1, p1 = z;
            for (int k = 2; k <= m; ++k) {
                ld p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = m * (z * p1 - p0) / (z * z - 1);
            ld dz = p1 / dp;
            z -= dz;
            if (fabsl(dz) < 1e-19L) break;
        }
        x[i] = (1 - z) / 2;
        w[i] = 1 / ((1 - z * z) * dp * dp);
    }
}

/* The oracle's own triangle rule: a collapsed (Duffy) product of 6-point
 * Gauss rules, exact for degree <= 10 on the triangle, placed on face `face`
 * (barycentric coordinates face, face+1, face+2 mod 4 as int_face.inl:31-36;
 * the opposite vertex 0).  p[4*36], w[36] summing to 1; returns 36. */
int orc_face_quadrature(int face, double* p, double* w) {
    enum { M = 6 };
    ld x[M], wx[M];
    orc_gauss01(M, x, wx);
    int q = 0;
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j, ++q) {
            ld l1 = x[i] * (1 - x[j]), l2 = x[j], l0 = 1 - l1 - l2;
            if (p) {
                for (int k = 0; k < 4; ++k) p[4 * q + k] = 0;
                p[4 * q + face] = (double)l0;
                p[4 * q + (face + 1) % 4] = (double)l1;
                p[4 * q + (face + 2) % 4] = (double)l2;
            }
            if (w) w[q] = (double)(2 * wx[i] * wx[j] * (1 - x[j]));
        }
    return q;
}

/* Area of the triangle p0 p1 p2 (geometry.h:67-72) and, for face_normal
 * (geometry.h:76-93), the unit normal pointing away from p3. */
static double orc_face_geom(const double* p0, const double* p1, const double* p2, const double* p3, double* nrm) {
    double a[3], b[3], c[3];
    for (int i = 0; i < 3; ++i) a[i] = p1[i] - p0[i], b[i] = p2[i] - p0[i];
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
    double len = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    if (nrm) {
        double h = 0;
        for (int i = 0; i < 3; ++i) h += c[i] * (p3[i] - p0[i]);
        for (int i = 0; i < 3; ++i) nrm[i] = (h > 0 ? -c[i] : c[i]) / len;
    }
    return len / 2;
}

/* fem3Dface (dot_normal = 0) / fem3DfaceN (dot_normal = 1) over face `face`
 * of n tets, with the oracle's own rule (orc_face_quadrature); per-point D
 * (coef_mode 2) is given at those points.  In the normal form D has
 * 3*dimB rows and DU is contracted with the normal as core.h:176-195. */
int orc_fem3dface(int opA, int femA, int ncA, int opB, int femB, int ncB, int tensor_type, int coef_mode, int face,
                  int dot_normal, long n, const double* xyz, const double* D, double* A) {
    int nfA = orc_nfa(femA, ncA), nfB = orc_nfa(femB, ncB);
    if (nfA <= 0 || nfB <= 0 || face < 0 || face > 3) return -1;
    if (!orc_op_ok(opA, femA, ncA) || !orc_op_ok(opB, femB, ncB)) return -1;
    double XYL[4 * 36], WG[36];
    int q = orc_face_quadrature(face, XYL, WG);
    int dimA = orc_opdim(opA, femA, ncA), dimB = orc_opdim(opB, femB, ncB);
    int nD = dot_normal ? 3 : 1, dimBt = dimB * nD;
    int ddim = tensor_type == ORC_TENSOR_NULL ? 0 : (tensor_type == ORC_TENSOR_SCALAR ? 1 : dimA * dimBt);
    int bcast = nfA == 1 && dimA == 1;
    if ((tensor_type == ORC_TENSOR_NULL || tensor_type == ORC_TENSOR_SCALAR) && dimA != dimBt && !bcast) return -3;
    double* U = malloc(sizeof(double) * (size_t)dimA * q * nfA);
    double* V = malloc(sizeof(double) * (size_t)dimB * q * nfB);
    double* DU = malloc(sizeof(double) * (size_t)dimBt * q * nfA);
    for (long r = 0; r < n; ++r) {
        const double* P = xyz + 12 * r;
        double XYP[12], PSI[9], nrm[3];
        for (int i = 0; i < 3; ++i) XYP[i] = 0;
        for (int l = 0; l < 3; ++l)
            for (int i = 0; i < 3; ++i) XYP[3 * (l + 1) + i] = P[3 * (l + 1) + i] - P[i];
        double det = orc_inverse3x3(XYP + 3, PSI);
        const double* x0 = XYP + 3 * face;
        const double* x1 = XYP + 3 * ((face + 1) % 4);
        const double* x2 = XYP + 3 * ((face + 2) % 4);
        const double* x3 = XYP + 3 * ((face + 3) % 4);
        double mes = orc_face_geom(x0, x1, x2, x3, dot_normal ? nrm : NULL);
        orc_apply_op(opA, femA, ncA, q, XYL, PSI, XYP, det, U);
        orc_apply_op(opB, femB, ncB, q, XYL, PSI, XYP, det, V);
        for (int n_ = 0; n_ < q; ++n_) {
            const double* Dn = coef_mode == ORC_COEF_CONST      ? D
                               : coef_mode == ORC_COEF_PER_TET ? D + (size_t)ddim * r
                                                               : D + (size_t)ddim * ((size_t)q * r + n_);
            double wv = WG[n_] * mes;
            for (int i = 0; i < nfA; ++i) {
                const double* u = U + dimA * (n_ + q * i);
                double* du = DU + dimBt * (n_ + q * i);
                for (int k = 0; k < dimBt; ++k) {
                    double s;
                    if (tensor_type == ORC_TENSOR_NULL) s = dimA == dimBt ? u[k] : u[0];
                    else if (tensor_type == ORC_TENSOR_SCALAR) s = Dn[0] * (dimA == dimBt ? u[k] : u[0]);
                    else {
                        s = 0;
                        for (int j = 0; j < dimA; ++j) s += Dn[k + dimBt * j] * u[j];
                    }
                    du[k] = wv * s;
                }
            }
        }
        /* normal contraction: DU_k = sum_d DU_{d + 3k} n_d, packed in place */
        if (dot_normal)
            for (int i = 0; i < nfA; ++i)
                for (int n_ = 0; n_ < q; ++n_) {
                    const double* src = DU + dimBt * (n_ + q * i);
                    double tmp[9];
                    for (int k = 0; k < dimB; ++k)
                        tmp[k] = src[3 * k] * nrm[0] + src[3 * k + 1] * nrm[1] + src[3 * k + 2] * nrm[2];
                    for (int k = 0; k < dimB; ++k) DU[k + dimB * (n_ + q * i)] = tmp[k];
                }
        double* Ar = A + (size_t)nfA * nfB * r;
        int nrow = q * dimB;
        for (int ia = 0; ia < nfA; ++ia)
            for (int ib = 0; ib < nfB; ++ib) {
                double s = 0;
                for (int j = 0; j < nrow; ++j) s += DU[j + nrow * ia] * V[j + nrow * ib];
                Ar[ib + nfB * ia] = s;
            }
    }
    free(U);
    free(V);
    free(DU);
    return 0;
}

/* fem3DapplyL (eval.inl:84-100, core.inl:370-405): Op(u) at q barycentric
 * points xyl of each tet, opU = V dofs with V[k + dim*n, i] the operator table
 * of orc_apply_op; out[r*dim*q + k + dim*n]. */
int orc_fem3dapply(int op, int fem, int ncomp, int q, const double* xyl, long n, const double* xyz,
                   const double* dofs, double* out) {
    int nf = orc_nfa(fem, ncomp), dim = orc_opdim(op, fem, ncomp);
    if (nf <= 0 || q <= 0 || !orc_op_ok(op, fem, ncomp)) return -1;
    double* U = malloc(sizeof(double) * (size_t)dim * q * nf);
    for (long r = 0; r < n; ++r) {
        const double* P = xyz + 12 * r;
        double XYP[9], PSI[9];
        for (int l = 0; l < 3; ++l)
            for (int i = 0; i < 3; ++i) XYP[i + 3 * l] = P[3 * (l + 1) + i] - P[i];
        double det = orc_inverse3x3(XYP, PSI);
        orc_apply_op(op, fem, ncomp, q, xyl, PSI, P, det, U);
        for (int j = 0; j < dim * q; ++j) {
            double s = 0;
            for (int i = 0; i < nf; ++i) s += U[j + dim * q * i] * dofs[(size_t)nf * r + i];
            out[(size_t)dim * q * r + j] = s;
        }
    }
    free(U);
    return 0;
}

/* fem3DapplyX (eval.inl:103-125): q physical points per tet X[(r*q + n)*3 + c],
 * their barycentric coordinates from getBaryCoord (core.inl:120-137), then
 * fem3DapplyL at them; out as orc_fem3dapply. */
int orc_fem3dapplyx(int op, int fem, int ncomp, int q, long n, const double* xyz, const double* X,
                    const double* dofs, double* out) {
    int nf = orc_nfa(fem, ncomp), dim = orc_opdim(op, fem, ncomp);
    if (nf <= 0 || q <= 0 || !orc_op_ok(op, fem, ncomp)) return -1;
    double* XYL = malloc(sizeof(double) * 4 * (size_t)q);
    for (long r = 0; r < n; ++r) {
        const double* P = xyz + 12 * r;
        double XYP[9], PSI[9];
        for (int l = 0; l < 3; ++l)
            for (int i = 0; i < 3; ++i) XYP[i + 3 * l] = P[3 * (l + 1) + i] - P[i];
        orc_inverse3x3(XYP, PSI);
        for (int m = 0; m < q; ++m) {
            const double* x = X + 3 * ((size_t)q * r + m);
            double dr[3] = {x[0] - P[0], x[1] - P[1], x[2] - P[2]};
            for (int i = 0; i < 3; ++i) XYL[4 * m + i + 1] = 0;
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) XYL[4 * m + i + 1] += PSI[i + 3 * j] * dr[j];
            XYL[4 * m] = 1 - XYL[4 * m + 1] - XYL[4 * m + 2] - XYL[4 * m + 3];
        }
        int rc = orc_fem3dapply(op, fem, ncomp, q, XYL, 1, P, dofs + (size_t)nf * r, out + (size_t)dim * q * r);
        if (rc) {
            free(XYL);
            return rc;
        }
    }
    free(XYL);
    return 0;
}

void orc_apply_dir(int nf, double* A, double* F, int k, double bc) {
    for (int i = 0; i < nf; ++i) F[i] -= A[i + nf * k] * bc;
    F[k] = bc;
    for (int i = 0; i < nf; ++i) A[i + nf * k] = 0; /* column k (dc_on_dof.h:28) */
    for (int i = 0; i < nf; ++i) A[k + nf * i] = 0; /* row k (dc_on_dof.h:29) */
    A[k + nf * k] = 1.0;
}

/* Element-level Dirichlet on every cell, as the reference's local function
 * applies it (ex1.cpp:100-105): applyDir on each flagged local DOF k in
 * ascending local order, with value bc[dofs[k]] (0 when bc is NULL). */
void orc_apply_dir_cells(long n_tets, int nf, const int32_t* dofs, const uint8_t* is_dir, const double* bc,
                         double* A, double* F) {
    double Fs[64];
    for (long t = 0; t < n_tets; ++t) {
        double* At = A + (size_t)nf * nf * t;
        double* Ft = F ? F + (size_t)nf * t : Fs;
        if (!F)
            for (int i = 0; i < nf; ++i) Fs[i] = 0.0;
        for (int k = 0; k < nf; ++k) {
            const int32_t d = dofs[(size_t)nf * t + k];
            if (is_dir[d]) orc_apply_dir(nf, At, Ft, k, bc ? bc[d] : 0.0);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Assembly-side restatement                                                  */

void orc_reorder_nodes(long n_tets, int32_t* tets, const double* X) {
    for (long t = 0; t < n_tets; ++t) {
        int32_t* v = tets + 4 * t;
        double m[3][3];
        for (int j = 0; j < 3; ++j) {
            m[0][j] = X[3 * v[0] + j] - X[3 * v[3] + j];
            m[1][j] = X[3 * v[1] + j] - X[3 * v[3] + j];
            m[2][j] = X[3 * v[2] + j] - X[3 * v[3] + j];
        }
        double det = 0;
        det += m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]);
        det += m[0][1] * (m[1][2] * m[2][0] - m[1][0] * m[2][2]);
        det += m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        if (det < 0) {
            int32_t s = v[2];
            v[2] = v[3];
            v[3] = s;
        }
    }
}

typedef struct {
    int32_t k[3];
    long occ;
} orc_key;

static int orc_key_cmp(const void* a, const void* b) {
    const orc_key *x = a, *y = b;
    for (int i = 0; i < 3; ++i)
        if (x->k[i] != y->k[i]) return x->k[i] < y->k[i] ? -1 : 1;
    return (x->occ > y->occ) - (x->occ < y->occ);
}

static int orc_long_cmp(const void* a, const void* b) {
    long x = *(const long*)a, y = *(const long*)b;
    return (x > y) - (x < y);
}

/* Number entities (edges: nv=2, faces: nv=3) by first appearance in cell
 * order; ids[t*ne + e] receives the id.  Returns the entity count. */
static long orc_number_entities(long n_tets, const int32_t* tets, int ne, int nv, const unsigned char* lv,
                                long* ids) {
    long m = n_tets * ne;
    orc_key* keys = malloc(sizeof(orc_key) * (size_t)m);
    for (long t = 0; t < n_tets; ++t)
        for (int e = 0; e < ne; ++e) {
            int32_t a[3] = {0, 0, 0};
            for (int i = 0; i < nv; ++i) a[i] = tets[4 * t + lv[nv * e + i]];
            /* sort ascending */
            for (int i = 0; i < nv; ++i)
                for (int j = i + 1; j < nv; ++j)
                    if (a[j] < a[i]) {
                        int32_t s = a[i];
                        a[i] = a[j];
                        a[j] = s;
                    }
            orc_key* k = &keys[t * ne + e];
            k->k[0] = a[0], k->k[1] = a[1], k->k[2] = a[2];
            k->occ = t * ne + e;
        }
    qsort(keys, (size_t)m, sizeof(orc_key), orc_key_cmp);
    /* first occurrence of every unique key, then rank the uniques by it */
    long* first = malloc(sizeof(long) * (size_t)m);
    long nu = 0;
    for (long i = 0; i < m; ++i)
        if (i == 0 || memcmp(keys[i].k, keys[i - 1].k, sizeof(keys[i].k)) != 0) first[nu++] = keys[i].occ;
    long* sorted_first = malloc(sizeof(long) * (size_t)nu);
    memcpy(sorted_first, first, sizeof(long) * (size_t)nu);
    qsort(sorted_first, (size_t)nu, sizeof(long), orc_long_cmp);
    long u = -1;
    for (long i = 0; i < m; ++i) {
        if (i == 0 || memcmp(keys[i].k, keys[i - 1].k, sizeof(keys[i].k)) != 0) ++u;
        /* id = rank of first[u] among sorted_first */
        long lo = 0, hi = nu - 1, f = first[u];
        while (lo < hi) {
            long mid = (lo + hi) / 2;
            if (sorted_first[mid] < f) lo = mid + 1;
            else hi = mid;
        }
        ids[keys[i].occ] = lo;
    }
    free(keys);
    free(first);
    free(sorted_first);
    return nu;
}

/* Vertex ranks by global node id: rank[k] = number of vertices of the cell
 * whose id is smaller than g[k] (the permutation createOrderPermutation
 * returns, ordering.inl:137-151, for the distinct ids of a valid cell). */
static void orc_vertex_rank(const long* g, unsigned char* rank) {
    for (int k = 0; k < 4; ++k) {
        unsigned char r = 0;
        for (int m = 0; m < 4; ++m) r = (unsigned char)(r + (g[m] < g[k]));
        rank[k] = r;
    }
}

long orc_enumerate(int fem, int ncomp, long n_nodes, long n_tets, const int32_t* tets, int32_t* dofs, long* n_edges,
                   long* n_faces) {
    int nf0 = orc_nf_scalar(fem);
    if (nf0 < 4) return -1; /* P1..P3 */
    static const unsigned char ev[12] = {0, 1, 0, 2, 0, 3, 1, 2, 1, 3, 2, 3};
    static const unsigned char fv[12] = {0, 1, 2, 1, 2, 3, 2, 3, 0, 3, 0, 1};
    long* eid = NULL;
    long* fid = NULL;
    long ne = 0, nfc = 0;
    if (fem >= ORC_P2) {
        eid = malloc(sizeof(long) * (size_t)n_tets * 6);
        ne = orc_number_entities(n_tets, tets, 6, 2, ev, eid);
    }
    if (fem >= ORC_P3) {
        fid = malloc(sizeof(long) * (size_t)n_tets * 4);
        nfc = orc_number_entities(n_tets, tets, 4, 3, fv, fid);
    }
    if (n_edges) *n_edges = ne;
    if (n_faces) *n_faces = nfc;
    long per_edge = fem == ORC_P2 ? 1 : (fem == ORC_P3 ? 2 : 0);
    long per_face = fem == ORC_P3 ? 1 : 0;
    long nscal = n_nodes + per_edge * ne + per_face * nfc;
    int nf = nf0 * ncomp;
    for (long t = 0; t < n_tets; ++t) {
        const int32_t* v = tets + 4 * t;
        long gl[4] = {v[0], v[1], v[2], v[3]};
        unsigned char perm[4] = {0, 1, 2, 3};
        if (fem == ORC_P3) orc_vertex_rank(gl, perm);
        long loc[20];
        for (int i = 0; i < 4; ++i) loc[i] = v[i];
        if (fem == ORC_P2)
            for (int e = 0; e < 6; ++e) loc[4 + e] = n_nodes + eid[6 * t + e];
        if (fem == ORC_P3) {
            for (int e = 0; e < 6; ++e) {
                long base = n_nodes + 2 * eid[6 * t + e];
                /* index_on_reorderd_elem (tetdofmap.inl:99-104): keep the pair
                 * order when perm[I] < perm[J], else swap */
                int keep = perm[ORC_EI[e]] < perm[ORC_EJ[e]];
                loc[4 + 2 * e] = base + (keep ? 0 : 1);
                loc[4 + 2 * e + 1] = base + (keep ? 1 : 0);
            }
            for (int f = 0; f < 4; ++f) loc[16 + f] = n_nodes + 2 * ne + fid[4 * t + f];
        }
        for (int c = 0; c < ncomp; ++c)
            for (int i = 0; i < nf0; ++i) dofs[(size_t)nf * t + nf0 * c + i] = (int32_t)(c * nscal + loc[i]);
    }
    free(eid);
    free(fid);
    return nscal * ncomp;
}

static int orc_int_cmp(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

long orc_csr_pattern(long ndof, long n_tets, int nf, const int32_t* dofs, const uint8_t* is_dir, int64_t* row_ptr,
                     int32_t* col_idx) {
    /* gather candidate columns per row, then sort + unique */
    int64_t* cnt = calloc((size_t)ndof + 1, sizeof(int64_t));
    for (long t = 0; t < n_tets; ++t)
        for (int i = 0; i < nf; ++i) cnt[dofs[(size_t)nf * t + i] + 1] += nf;
    for (long d = 0; d < ndof; ++d) cnt[d + 1] += cnt[d] + 1; /* +1 for the diagonal */
    int32_t* tmp = malloc(sizeof(int32_t) * (size_t)cnt[ndof]);
    int64_t* fill = malloc(sizeof(int64_t) * (size_t)ndof);
    for (long d = 0; d < ndof; ++d) {
        fill[d] = cnt[d];
        tmp[fill[d]++] = (int32_t)d;
    }
    for (long t = 0; t < n_tets; ++t)
        for (int i = 0; i < nf; ++i) {
            int32_t r = dofs[(size_t)nf * t + i];
            for (int j = 0; j < nf; ++j) {
                int32_t c = dofs[(size_t)nf * t + j];
                int keep = !(is_dir && r != c && (is_dir[r] || is_dir[c]));
                tmp[fill[r]++] = keep ? c : r;
            }
        }
    long nnz = 0;
    row_ptr[0] = 0;
    for (long d = 0; d < ndof; ++d) {
        int32_t* b = tmp + cnt[d];
        long m = fill[d] - cnt[d];
        qsort(b, (size_t)m, sizeof(int32_t), orc_int_cmp);
        long u = 0;
        for (long k = 0; k < m; ++k)
            if (k == 0 || b[k] != b[k - 1]) {
                if (col_idx) col_idx[nnz + u] = b[k];
                ++u;
            }
        nnz += u;
        row_ptr[d + 1] = nnz;
    }
    free(cnt);
    free(tmp);
    free(fill);
    return nnz;
}

int orc_slot_map(long n_tets, int nf, const int32_t* dofs, const int64_t* row_ptr, const int32_t* col_idx,
                 int64_t* slot) {
    for (long t = 0; t < n_tets; ++t)
        for (int ia = 0; ia < nf; ++ia)
            for (int ib = 0; ib < nf; ++ib) {
                int32_t r = dofs[(size_t)nf * t + ib], c = dofs[(size_t)nf * t + ia];
                int64_t lo = row_ptr[r], hi = row_ptr[r + 1], pos = -1;
                while (lo < hi) {
                    int64_t mid = (lo + hi) / 2;
                    if (col_idx[mid] < c) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < row_ptr[r + 1] && col_idx[lo] == c) pos = lo;
                slot[(size_t)nf * nf * t + ib + nf * ia] = pos;
            }
    return 0;
}

int orc_greedy_colour(long n_nodes, long n_tets, const int32_t* tets, int32_t* colour) {
    /* node -> incident tets (CSR) */
    int64_t* ptr = calloc((size_t)n_nodes + 1, sizeof(int64_t));
    for (long t = 0; t < 4 * n_tets; ++t) ptr[tets[t] + 1]++;
    for (long i = 0; i < n_nodes; ++i) ptr[i + 1] += ptr[i];
    int64_t* fill = malloc(sizeof(int64_t) * (size_t)n_nodes);
    memcpy(fill, ptr, sizeof(int64_t) * (size_t)n_nodes);
    int32_t* adj = malloc(sizeof(int32_t) * (size_t)(4 * n_tets));
    for (long t = 0; t < n_tets; ++t)
        for (int k = 0; k < 4; ++k) adj[fill[tets[4 * t + k]]++] = (int32_t)t;
    int ncol = 0;
    int64_t used[8] = {0}; /* bitmask of up to 512 colours */
    for (long t = 0; t < n_tets; ++t) colour[t] = -1;
    for (long t = 0; t < n_tets; ++t) {
        memset(used, 0, sizeof(used));
        for (int k = 0; k < 4; ++k) {
            int32_t v = tets[4 * t + k];
            for (int64_t p = ptr[v]; p < ptr[v + 1]; ++p) {
                int32_t c = colour[adj[p]];
                if (c >= 0) used[c >> 6] |= (int64_t)1 << (c & 63);
            }
        }
        int c = 0;
        while (used[c >> 6] & ((int64_t)1 << (c & 63))) ++c;
        colour[t] = c;
        if (c + 1 > ncol) ncol = c + 1;
    }
    free(ptr);
    free(fill);
    free(adj);
    return ncol;
}

int orc_scatter(long n_tets, int nf, const int32_t* dofs, const int64_t* slot, const double* A, const double* F,
                double drop_val, double* val, double* rhs) {
    for (long t = 0; t < n_tets; ++t) {
        for (int i = 0; i < nf; ++i) {
            if (F) {
                double f = F[(size_t)nf * t + i];
                rhs[dofs[(size_t)nf * t + i]] += f;
                if (!isfinite(f)) return -1;
            }
            for (int j = 0; j < nf; ++j) {
                double a = A[(size_t)nf * nf * t + i + nf * j];
                int64_t s = slot[(size_t)nf * nf * t + i + nf * j];
                if (!isfinite(a)) return -1;
                if (fabs(a) > drop_val) {
                    if (s < 0) return -3;
                    val[s] += a;
                }
            }
        }
    }
    return 0;
}

/* The scatter of orc_scatter parallel over colours (the CPU baseline's
 * multi-threaded form): tets of one colour share no node, hence no DOF, so a
 * colour's tets are added by nthreads threads without conflicts, colour after
 * colour.  colour_ptr (n_colours + 1) and colour_tets list each colour's tets
 * in ascending order.  Values equal orc_scatter's up to the order of the
 * additions into each slot. */
int orc_scatter_coloured(int n_colours, const int64_t* colour_ptr, const int32_t* colour_tets, int nf,
                         const int32_t* dofs, const int64_t* slot, const double* A, const double* F, double drop_val,
                         int nthreads, double* val, double* rhs) {
    int bad = 0;
    for (int c = 0; c < n_colours; ++c) {
        const long n_c = (long)(colour_ptr[c + 1] - colour_ptr[c]);
        const int32_t* ct = colour_tets + colour_ptr[c];
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(| : bad)
        for (long k = 0; k < n_c; ++k) {
            const long t = ct[k];
            for (int i = 0; i < nf; ++i) {
                if (F) {
                    const double f = F[(size_t)nf * t + i];
                    rhs[dofs[(size_t)nf * t + i]] += f;
                    bad |= isfinite(f) ? 0 : 1;
                }
                for (int j = 0; j < nf; ++j) {
                    const double a = A[(size_t)nf * nf * t + i + nf * j];
                    const int64_t s = slot[(size_t)nf * nf * t + i + nf * j];
                    bad |= isfinite(a) ? 0 : 1;
                    if (fabs(a) > drop_val) {
                        if (s < 0)
                            bad |= 2;
                        else
                            val[s] += a;
                    }
                }
            }
        }
    }
    return bad & 2 ? -3 : (bad ? -1 : 0);
}

void orc_spmv(long nrow, const int64_t* row_ptr, const int32_t* col_idx, const double* val, const double* x,
              double* y) {
    for (long i = 0; i < nrow; ++i) {
        double s = 0;
        for (int64_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p) s += val[p] * x[col_idx[p]];
        y[i] = s;
    }
}

long orc_jacobi_pcg(long nrow, const int64_t* row_ptr, const int32_t* col_idx, const double* val, const double* b,
                    double* x, double rtol, long maxit, double* final_relres) {
    double *r = malloc(sizeof(double) * (size_t)nrow), *z = malloc(sizeof(double) * (size_t)nrow);
    double *p = malloc(sizeof(double) * (size_t)nrow), *Ap = malloc(sizeof(double) * (size_t)nrow);
    double* dinv = malloc(sizeof(double) * (size_t)nrow);
    for (long i = 0; i < nrow; ++i) {
        double d = 0;
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
            if (col_idx[k] == i) d = val[k];
        dinv[i] = d != 0 ? 1.0 / d : 1.0;
    }
    orc_spmv(nrow, row_ptr, col_idx, val, x, Ap);
    double bn = 0, rz = 0;
    for (long i = 0; i < nrow; ++i) {
        r[i] = b[i] - Ap[i];
        z[i] = dinv[i] * r[i];
        p[i] = z[i];
        bn += b[i] * b[i];
        rz += r[i] * z[i];
    }
    bn = sqrt(bn);
    if (bn == 0) bn = 1;
    long it = 0;
    double rn = 0;
    for (long i = 0; i < nrow; ++i) rn += r[i] * r[i];
    rn = sqrt(rn);
    while (rn > rtol * bn && it < maxit) {
        orc_spmv(nrow, row_ptr, col_idx, val, p, Ap);
        double pAp = 0;
        for (long i = 0; i < nrow; ++i) pAp += p[i] * Ap[i];
        double alpha = rz / pAp;
        double rz_new = 0;
        rn = 0;
        for (long i = 0; i < nrow; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
            z[i] = dinv[i] * r[i];
            rz_new += r[i] * z[i];
            rn += r[i] * r[i];
        }
        rn = sqrt(rn);
        double beta = rz_new / rz;
        rz = rz_new;
        for (long i = 0; i < nrow; ++i) p[i] = z[i] + beta * p[i];
        ++it;
    }
    if (final_relres) *final_relres = rn / bn;
    free(r);
    free(z);
    free(p);
    free(Ap);
    free(dinv);
    return rn <= rtol * bn ? it : -1;
}
