// Host-side basis tables and reference tensors.
//
// The element matrix of fem3Dtet (anifem++/fem/operations/core.inl:278-367)
//   A(ib, ia) = sum_n w_n |T| sum_{k,j} V_k(n, ib) D_n(k, j) U_j(n, ia)
// is evaluated on the GPU in factored form.  For scalar spaces the physical
// operator values are a per-tet linear map of reference-space values,
// U_j(n, i) = sum_a T_A(j, a) Uh_a(n, i) with T = PSI^T for GRAD
// (poly_2.h:94-100: U = PSI^T GRAD_P2) and T = 1 for IDEN, so
//   A(ib, ia) = sum_n sum_{b,a} K_n(b, a) Vh_b(n, ib) Uh_a(n, ia),
//   K_n = w_n |T| T_B^T D_n T_A.
// K (per tet, FP64) is computed on the GPU; the products Vh_b Uh_a depend only
// on the element pair and the quadrature rule and are tabulated here once
// (summed over n with the weights when D is constant over the tet).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

#include "anifem_common.h"
#include "basis_eval.h"
#include "tet_rules.inc"
#include "tri_rules.inc"

namespace anifem {

int nf_scalar(int fem) {
    switch (fem) {
        case ANIFEM_FEM_P0: return 1;
        case ANIFEM_FEM_P1: return 4;
        case ANIFEM_FEM_P2: return 10;
        case ANIFEM_FEM_P3: return 20;
        case ANIFEM_FEM_CR1: return 4;
        case ANIFEM_FEM_B2: return 6;
        case ANIFEM_FEM_B4: return 1;
        case ANIFEM_FEM_RT0: return 4;
        case ANIFEM_FEM_ND0: return 6;
    }
    return -1;
}

bool fem_lagrange(int fem) { return fem >= ANIFEM_FEM_P0 && fem <= ANIFEM_FEM_P3; }
bool fem_vector(int fem) { return fem == ANIFEM_FEM_RT0 || fem == ANIFEM_FEM_ND0; }

int ref_dim(int op, int fem) {
    if (fem_vector(fem)) {
        switch (op) {
            case ANIFEM_IDEN:
            case ANIFEM_CURL: return 3;
            case ANIFEM_DIV: return 1;
            // GRAD: 9 physical components from a reference field of 1 (RT0)
            // or 3 (ND0) components (space_map maps 3 and 4)
            case ANIFEM_GRAD: return fem == ANIFEM_FEM_RT0 ? 1 : 3;
            default: return -1;
        }
    }
    switch (op) {
        case ANIFEM_IDEN: return 1;
        case ANIFEM_GRAD:
        case ANIFEM_DIV:
        case ANIFEM_CURL:
        case ANIFEM_DUDX:
        case ANIFEM_DUDY:
        case ANIFEM_DUDZ: return 3;
        default: return -1;
    }
}

int op_dim(const anifem_op& op) {
    if (fem_vector(op.fem)) return op.op == ANIFEM_GRAD ? 9 : ref_dim(op.op, op.fem);
    switch (op.op) {
        case ANIFEM_IDEN: return op.ncomp;
        case ANIFEM_GRAD: return 3 * op.ncomp;
        case ANIFEM_DIV: return 1;
        case ANIFEM_CURL: return 3;
        default: return op.ncomp;  // one partial derivative per component
    }
}

bool op_valid(const anifem_op& op) {
    if (ref_dim(op.op, op.fem) < 0 || nf_scalar(op.fem) < 0 || op.ncomp < 1 || op.ncomp > 3) return false;
    if (fem_vector(op.fem)) return op.ncomp == 1;
    // derivatives of P0 are zero operators of the usual dims (poly_0.h:11-59)
    if ((op.op == ANIFEM_DIV || op.op == ANIFEM_CURL) && op.ncomp != 3) return false;
    return true;
}

void space_map(int op, int fem, int& map, int& scale) {
    map = 0;
    scale = 0;
    if (fem == ANIFEM_FEM_RT0) {
        // RT0: phi_i = c_i (x - x_{(i+3)%4}) = c_i J (xi - xi_{(i+3)%4}),
        // div phi_i = 3 c_i (raviant_thomas_0.h:35-47, :136-160); curl = 0
        if (op == ANIFEM_IDEN) map = 1, scale = 1;
        else if (op == ANIFEM_DIV) scale = 1;
        // grad phi_i = c_i I (raviant_thomas_0.h:101-104: the sum over the
        // vertices reduces to the identity): map 3, T = vec(I) (9 x 1)
        else if (op == ANIFEM_GRAD) map = 3, scale = 1;
    } else if (fem == ANIFEM_FEM_ND0) {
        // ND0 on edge (a, b): phi = |e| (l_a grad l_b - l_b grad l_a)
        // = |e| PSI^T (l_a gh_b - l_b gh_a); curl phi = 2 |e| grad l_a x grad l_b
        // = (|e| / det J) J (2 gh_a x gh_b) (nedelec_0.h:35-53, :100-135); div = 0
        if (op == ANIFEM_IDEN) scale = 2;
        else if (op == ANIFEM_CURL) map = 1, scale = 3;
        // grad phi is the skew matrix of curl phi / 2 (nedelec_0.h:167-180):
        // (|e| / det J) S J (gh_a x gh_b), map 4, T = S J (9 x 3) with S the
        // vector -> skew-matrix map in the reference's entry order
        else if (op == ANIFEM_GRAD) map = 4, scale = 3;
    }
}

// Conical product rule on the tet for orders above the tabulated ones: with
// x = u, y = (1 - u) v, z = (1 - u)(1 - v) s the volume element is
// (1 - u)^2 (1 - v) du dv ds, so Gauss-Jacobi rules on [0, 1] for the weights
// (1 - u)^2, (1 - v) and 1 with m = order/2 + 1 points each integrate every
// polynomial of total degree <= order exactly (Stroud's collapsed product).
// Nodes are the zeros of the monic orthogonal polynomial of degree m for the
// weight (1 - t)^alpha, found by Newton's method with deflation on the
// three-term recurrence; weights by the Christoffel sum
// w_i = mu0 / sum_{k<m} p_k(t_i)^2 / (b_1 ... b_k).  Points are listed with u
// outermost and s innermost, each in ascending order; the weights are scaled
// to sum to 1 (mean-value form).  The reference serves these orders with
// fully symmetric rules of fewer points (quadrature_formulas.cpp:602-1489);
// both integrate polynomials of degree <= order exactly.
static void gauss_jacobi01(int m, int alpha, std::vector<long double>& t, std::vector<long double>& w) {
    // monic Jacobi recurrence on [-1, 1] for (1 - x)^alpha:
    // p_{k+1} = (x - a_k) p_k - b_k p_{k-1}
    std::vector<long double> ak(m + 1), bk(m + 1);
    const long double al = alpha;
    for (int k = 0; k <= m; ++k) {
        const long double s = 2 * k + al;
        ak[k] = (k == 0) ? -al / (al + 2) : -al * al / (s * (s + 2));
        bk[k] = (k == 0) ? 0 : 4.0L * k * (k + al) * k * (k + al) / (s * s * (s + 1) * (s - 1));
    }
    auto eval = [&](long double x, long double& p, long double& dp) {
        long double p0 = 1, p1 = x - ak[0], d0 = 0, d1 = 1;
        for (int k = 1; k < m; ++k) {
            const long double p2 = (x - ak[k]) * p1 - bk[k] * p0;
            const long double d2 = (x - ak[k]) * d1 + p1 - bk[k] * d0;
            p0 = p1, p1 = p2, d0 = d1, d1 = d2;
        }
        p = p1, dp = d1;
    };
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<long double> x;
    for (int i = 0; i < m; ++i) {
        long double z = std::cos(pi * (i + 0.75L) / (m + 0.5L));
        for (int it = 0; it < 200; ++it) {
            long double p, dp;
            eval(z, p, dp);
            long double defl = 0;
            for (long double r : x) defl += 1.0L / (z - r);
            const long double dz = p / (dp - p * defl);
            z -= dz;
            if (std::fabs(dz) < 1e-19L) break;
        }
        x.push_back(z);
    }
    std::sort(x.begin(), x.end());
    const long double mu0 = std::pow(2.0L, al + 1) / (al + 1);  // int_{-1}^{1} (1 - x)^alpha
    t.resize(m);
    w.resize(m);
    for (int i = 0; i < m; ++i) {
        long double p0 = 1, p1 = x[i] - ak[0], sum = 1, nrm = 1;
        for (int k = 1; k < m; ++k) {
            nrm *= bk[k];
            sum += p1 * p1 / nrm;
            const long double p2 = (x[i] - ak[k]) * p1 - bk[k] * p0;
            p0 = p1, p1 = p2;
        }
        t[i] = (1 + x[i]) / 2;
        w[i] = mu0 / sum / std::pow(2.0L, al + 1);  // dx = 2 dt, (1 - x)^a = 2^a (1 - t)^a
    }
}

struct ProductRule {
    std::vector<double> p, w;
};

static const ProductRule& product_rule(int order) {
    static const std::vector<ProductRule> rules = [] {
        std::vector<ProductRule> r(kTetMaxOrder + 1);
        for (int o = ANIFEM_TET_TABLE_MAX_ORDER + 1; o <= kTetMaxOrder; ++o) {
            const int m = o / 2 + 1;
            std::vector<long double> tu, wu, tv, wv, ts, ws;
            gauss_jacobi01(m, 2, tu, wu);
            gauss_jacobi01(m, 1, tv, wv);
            gauss_jacobi01(m, 0, ts, ws);
            for (int i = 0; i < m; ++i)
                for (int j = 0; j < m; ++j)
                    for (int k = 0; k < m; ++k) {
                        const long double x = tu[i], y = (1 - tu[i]) * tv[j], z = (1 - tu[i]) * (1 - tv[j]) * ts[k];
                        r[o].p.push_back((double)(1 - x - y - z));
                        r[o].p.push_back((double)x);
                        r[o].p.push_back((double)y);
                        r[o].p.push_back((double)z);
                        r[o].w.push_back((double)(6 * wu[i] * wv[j] * ws[k]));
                    }
        }
        return r;
    }();
    return rules[order];
}

int quadrature(int order, const double** p, const double** w) {
    if (order < 0 || order > kTetMaxOrder) return -1;
    if (order > ANIFEM_TET_TABLE_MAX_ORDER) {
        const ProductRule& r = product_rule(order);
        *p = r.p.data();
        *w = r.w.data();
        return (int)r.w.size();
    }
    int q = ANIFEM_TET_NPTS_BY_ORDER[order];
    switch (q) {
        case 1: *p = ANIFEM_TET_Q1_P, *w = ANIFEM_TET_Q1_W; break;
        case 4: *p = ANIFEM_TET_Q4_P, *w = ANIFEM_TET_Q4_W; break;
        case 8: *p = ANIFEM_TET_Q8_P, *w = ANIFEM_TET_Q8_W; break;
        case 14: *p = ANIFEM_TET_Q14_P, *w = ANIFEM_TET_Q14_W; break;
        case 24: *p = ANIFEM_TET_Q24_P, *w = ANIFEM_TET_Q24_W; break;
        default: return -1;
    }
    return q;
}

// Triangle rules above the tabulated orders: the collapsed product x = u,
// y = (1 - u) v with area element (1 - u) du dv, Gauss-Jacobi rules for (1 - u)
// and 1 with m = order/2 + 1 points each (exact for total degree <= order);
// barycentric (1 - x - y, x, y), weights summing to 1.  The reference's
// triangle rules stop at order 20 (quadrature_formulas.cpp:520-521).
static const ProductRule& triangle_product_rule(int order) {
    static const std::vector<ProductRule> rules = [] {
        std::vector<ProductRule> r(kTriMaxOrder + 1);
        for (int o = ANIFEM_TRI_MAX_ORDER + 1; o <= kTriMaxOrder; ++o) {
            const int m = o / 2 + 1;
            std::vector<long double> tu, wu, tv, wv;
            gauss_jacobi01(m, 1, tu, wu);
            gauss_jacobi01(m, 0, tv, wv);
            for (int i = 0; i < m; ++i)
                for (int j = 0; j < m; ++j) {
                    const long double x = tu[i], y = (1 - tu[i]) * tv[j];
                    r[o].p.push_back((double)(1 - x - y));
                    r[o].p.push_back((double)x);
                    r[o].p.push_back((double)y);
                    r[o].w.push_back((double)(2 * wu[i] * wv[j]));
                }
        }
        return r;
    }();
    return rules[order];
}

int triangle_quadrature(int order, const double** p, const double** w) {
    if (order < 0 || order > kTriMaxOrder) return -1;
    if (order > ANIFEM_TRI_MAX_ORDER) {
        const ProductRule& r = triangle_product_rule(order);
        *p = r.p.data();
        *w = r.w.data();
        return (int)r.w.size();
    }
    int q = ANIFEM_TRI_NPTS_BY_ORDER[order];
    switch (q) {
        case 1: *p = ANIFEM_TRI_Q1_P, *w = ANIFEM_TRI_Q1_W; break;
        case 3: *p = ANIFEM_TRI_Q3_P, *w = ANIFEM_TRI_Q3_W; break;
        case 6: *p = ANIFEM_TRI_Q6_P, *w = ANIFEM_TRI_Q6_W; break;
        case 7: *p = ANIFEM_TRI_Q7_P, *w = ANIFEM_TRI_Q7_W; break;
        case 12: *p = ANIFEM_TRI_Q12_P, *w = ANIFEM_TRI_Q12_W; break;
        default: return -1;
    }
    return q;
}

// Points of the rule for `order` as tet barycentric coordinates: the volume
// rule (face < 0), or the triangle rule on face `face` (vertices face,
// face+1, face+2 mod 4; the fourth coordinate 0), placed as in
// int_face.inl:31-36.
int rule_points(int order, int face, std::vector<double>& pts, std::vector<double>& wts) {
    const double *p, *w;
    if (face < 0) {
        const int q = quadrature(order, &p, &w);
        if (q < 0) return q;
        pts.assign(p, p + 4 * q);
        wts.assign(w, w + q);
        return q;
    }
    if (face > 3) return -1;
    const int q = triangle_quadrature(order, &p, &w);
    if (q < 0) return q;
    pts.assign((size_t)4 * q, 0.0);
    for (int n = 0; n < q; ++n) {
        pts[4 * n + face] = p[3 * n + 0];
        pts[4 * n + (face + 1) % 4] = p[3 * n + 1];
        pts[4 * n + (face + 2) % 4] = p[3 * n + 2];
    }
    wts.assign(w, w + q);
    return q;
}

int tensor_ddim(int type, int dimA, int dimB) {
    switch (type) {
        case ANIFEM_TENSOR_NULL: return 0;
        case ANIFEM_TENSOR_SCALAR: return 1;
        case ANIFEM_TENSOR_SYMMETRIC: return dimA == dimB ? dimB * (dimB + 1) / 2 : -1;
        case ANIFEM_TENSOR_GENERAL: return dimA * dimB;
        case ANIFEM_TENSOR_LAME: return (dimA == 9 && dimB == 9) ? 2 : -1;
    }
    return -1;
}

void reference_table(int op, int fem, int q, const double* p, std::vector<double>& tab) {
    int nf = nf_scalar(fem), d = ref_dim(op, fem);
    tab.assign((size_t)q * nf * d, 0.0);
    double buf[60];
    for (int n = 0; n < q; ++n) {
        // the reference recomputes l0 = 1 - l1 - l2 - l3 for stability
        // (quadrature_formulas.h:195-200); the generated table already satisfies it
        const double* L = p + 4 * n;
        if (fem_vector(fem)) vector_field(op, fem, L, buf);
        else if (op == ANIFEM_IDEN) basis_values(fem, L, buf);
        else basis_gradients(fem, L, buf);
        for (int i = 0; i < nf; ++i)
            for (int a = 0; a < d; ++a) tab[((size_t)n * nf + i) * d + a] = buf[i * d + a];
    }
}

void reference_tensor(int opA, int femA, int opB, int femB, int order, bool per_point, std::vector<double>& R,
                      int face) {
    std::vector<double> pts, wts;
    int q = rule_points(order, face, pts, wts);
    const double* p = pts.data();
    const double* w = wts.data();
    std::vector<double> TA, TB;
    reference_table(opA, femA, q, p, TA);
    reference_table(opB, femB, q, p, TB);
    int nfA = nf_scalar(femA), nfB = nf_scalar(femB), dA = ref_dim(opA, femA), dB = ref_dim(opB, femB);
    int KB = dA * dB, E0 = nfA * nfB, qq = per_point ? q : 1;
    R.assign((size_t)E0 * qq * KB, 0.0);
    for (int ia = 0; ia < nfA; ++ia)
        for (int ib = 0; ib < nfB; ++ib) {
            int rr = ib + nfB * ia;
            for (int n = 0; n < q; ++n)
                for (int b = 0; b < dB; ++b)
                    for (int a = 0; a < dA; ++a) {
                        double v = TB[((size_t)n * nfB + ib) * dB + b] * TA[((size_t)n * nfA + ia) * dA + a];
                        if (per_point) R[((size_t)rr * q + n) * KB + b * dA + a] = v;
                        else R[(size_t)rr * KB + b * dA + a] += w[n] * v;
                    }
        }
}

// hi = fp16(x), lo = fp16(x - hi) of n floats: out[0..n) hi, out[n..2n) lo
__global__ void split_f16_kernel(const float* x, int64_t n, _Float16* out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const _Float16 h = (_Float16)x[i];
    out[i] = h;
    out[n + i] = (_Float16)(x[i] - (float)h);
}

thread_local bool g_ref_lookup_only = false;

int get_ref_tensor(int opA, int femA, int opB, int femB, int order, bool per_point, RefTensorDev* out, int face) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, int, int, int, int, int, int>, RefTensorDev> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return ANIFEM_ERR_HIP;
    auto key = std::make_tuple(dev, opA, femA, opB, femB, order, (int)per_point, face);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
        *out = it->second;
        return ANIFEM_OK;
    }
    if (g_ref_lookup_only) return ANIFEM_ERR_ARG;  // building the table would allocate
    std::vector<double> R;
    reference_tensor(opA, femA, opB, femB, order, per_point, R, face);
    std::vector<float> Rf(R.begin(), R.end());
    RefTensorDev t;
    std::vector<double> pts, wts;
    const int nq = rule_points(order, face, pts, wts);
    const double* ww = wts.data();
    t.q = per_point ? nq : 1;
    t.E0 = nf_scalar(femA) * nf_scalar(femB);
    t.KB = ref_dim(opA, femA) * ref_dim(opB, femB);
    void* d64 = nullptr;
    void* d32 = nullptr;
    void* dw = nullptr;
    ANIFEM_HIP_CHECK(hipMalloc(&dw, sizeof(double) * (size_t)nq));
    ANIFEM_HIP_CHECK(hipMemcpy(dw, ww, sizeof(double) * (size_t)nq, hipMemcpyHostToDevice));
    t.w = static_cast<const double*>(dw);
    ANIFEM_HIP_CHECK(hipMalloc(&d64, R.size() * sizeof(double)));
    ANIFEM_HIP_CHECK(hipMalloc(&d32, Rf.size() * sizeof(float)));
    ANIFEM_HIP_CHECK(hipMemcpy(d64, R.data(), R.size() * sizeof(double), hipMemcpyHostToDevice));
    ANIFEM_HIP_CHECK(hipMemcpy(d32, Rf.data(), Rf.size() * sizeof(float), hipMemcpyHostToDevice));
    t.r64 = static_cast<const double*>(d64);
    t.r32 = static_cast<const float*>(d32);
    if (!per_point && t.KB == 9) {
        // columns kk = b*3 + a of K(b, a), summed onto the packed K(a, b)
        std::vector<double> Rs((size_t)t.E0 * 6, 0.0);
        for (int e = 0; e < t.E0; ++e)
            for (int b = 0; b < 3; ++b)
                for (int a = 0; a < 3; ++a) Rs[(size_t)e * 6 + sym_pack(a, b)] += R[(size_t)e * 9 + b * 3 + a];
        std::vector<float> Rsf(Rs.begin(), Rs.end());
        void* ds64 = nullptr;
        void* ds32 = nullptr;
        ANIFEM_HIP_CHECK(hipMalloc(&ds64, Rs.size() * sizeof(double)));
        ANIFEM_HIP_CHECK(hipMalloc(&ds32, Rsf.size() * sizeof(float)));
        ANIFEM_HIP_CHECK(hipMemcpy(ds64, Rs.data(), Rs.size() * sizeof(double), hipMemcpyHostToDevice));
        ANIFEM_HIP_CHECK(hipMemcpy(ds32, Rsf.data(), Rsf.size() * sizeof(float), hipMemcpyHostToDevice));
        t.rs64 = static_cast<const double*>(ds64);
        t.rs32 = static_cast<const float*>(ds32);
        const int nA = nf_scalar(femA), nB = nf_scalar(femB);
        bool sym = nA == nB;
        for (int ia = 0; ia < nA && sym; ++ia)
            for (int ib = 0; ib < nB && sym; ++ib)
                for (int p = 0; p < 6; ++p)
                    sym = sym && Rs[(size_t)(ib + nB * ia) * 6 + p] == Rs[(size_t)(ia + nB * ib) * 6 + p];
        t.rs_sym = sym;
    }
    if (!per_point && t.KB <= 16) {
        // R / max|R| in [-1, 1], rows padded to 16 columns and to a multiple of 16 rows
        double mx = 0;
        for (double v : R) mx = std::max(mx, std::fabs(v));
        if (mx == 0) mx = 1;
        t.e0pad = (t.E0 + 15) / 16 * 16;
        t.r16_scale = (float)mx;
        std::vector<float> xs((size_t)t.e0pad * 16, 0.0f);
        for (int e = 0; e < t.E0; ++e)
            for (int kk = 0; kk < t.KB; ++kk) xs[(size_t)e * 16 + kk] = (float)(R[(size_t)e * t.KB + kk] / mx);
        void* dx = nullptr;
        void* d16 = nullptr;
        const int64_t nx = (int64_t)xs.size();
        ANIFEM_HIP_CHECK(hipMalloc(&dx, nx * sizeof(float)));
        ANIFEM_HIP_CHECK(hipMalloc(&d16, 2 * nx * sizeof(_Float16)));
        ANIFEM_HIP_CHECK(hipMemcpy(dx, xs.data(), nx * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(split_f16_kernel, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, 0,
                           static_cast<const float*>(dx), nx, static_cast<_Float16*>(d16));
        ANIFEM_HIP_CHECK(hipGetLastError());
        ANIFEM_HIP_CHECK(hipDeviceSynchronize());
        ANIFEM_HIP_CHECK(hipFree(dx));
        t.r16 = d16;
    }
    cache[key] = t;
    *out = t;
    return ANIFEM_OK;
}

}  // namespace anifem

extern "C" {

int anifem_op_nfa(anifem_op op) { return anifem::op_valid(op) ? anifem::nf_scalar(op.fem) * op.ncomp : -1; }
int anifem_op_dim(anifem_op op) { return anifem::op_valid(op) ? anifem::op_dim(op) : -1; }

int anifem_quadrature_npoints(int order) {
    const double *p, *w;
    return anifem::quadrature(order, &p, &w);
}

int anifem_quadrature(int order, double* p, double* w) {
    const double *pp, *ww;
    int q = anifem::quadrature(order, &pp, &ww);
    if (q < 0) return ANIFEM_ERR_ORDER;
    for (int i = 0; i < 4 * q; ++i) p[i] = pp[i];
    for (int i = 0; i < q; ++i) w[i] = ww[i];
    return q;
}

const char* anifem_version(void) { return "anifem-mi355x 0.1 (gfx950)"; }

}  // extern "C"
