// Test harness around the reference AniFem++ element core, compiled from the
// reference's own sources (anifem++/fem) by oracle/Makefile into
// oracle/_ref/libanifem_ref.so.  It is test infrastructure only: the golden
// fixture generator (tests/golden/make_golden.py), the reference-vs-oracle tests
// and the bench's CPU baseline load it; nothing in the product path does.
//
// The harness exposes a flat C interface over a fixed set of template
// instantiations of Ani::fem3Dtet (anifem++/fem/operations/int_tet.inl:92-122)
// so that Python can feed it seeded tetrahedra and per-tet tensors.
#include "anifem++/fem/operations/operations.h"
#include "anifem++/fem/operators.h"
#include "anifem++/fem/spaces/spaces.h"
#include "anifem++/fem/quadrature_formulas.h"

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

using namespace Ani;

namespace {

struct TetData {
    const double* D;   // dimB x dimA column-major tensor for this tet (or 1 value for SCALAR)
    long ddim;
};

// D callback returning the tensor stored in user data; TT is the declared type.
// With XDEP the stored tensor is multiplied by the scalar field
// m(x) = 1 + x0/2 - x1/4 + x2^2 / 8, evaluated at the physical quadrature point,
// which exercises the per-point path with a tensor that varies inside the tet.
inline double xdep_field(const std::array<double, 3>& X) { return 1.0 + 0.5 * X[0] - 0.25 * X[1] + 0.125 * X[2] * X[2]; }

template <int TT, bool XDEP>
struct StoredTensor {
    TensorType operator()(const std::array<double, 3>& X, double* D, std::pair<ulong, ulong> dims,
                          void* user_data, int iTet) const {
        (void)iTet;
        const TetData* td = static_cast<const TetData*>(user_data);
        const double m = XDEP ? xdep_field(X) : 1.0;
        if (TT == TENSOR_SCALAR) {
            D[0] = m * td->D[0];
        } else if (TT != TENSOR_NULL) {
            for (std::size_t k = 0; k < dims.first * dims.second; ++k) D[k] = m * td->D[k];
        }
        return static_cast<TensorType>(TT);
    }
};

template <typename OpA, typename OpB, int TT, int SPARSITY, bool XDEP = false>
int run_case(int order, long n, const double* xyz, const double* D, long ddim, double* A) {
    constexpr long nfa = OpA::Nfa::value, nfb = OpB::Nfa::value;
    double buf[nfa * nfb > 1 ? nfa * nfb : 1];
    for (long r = 0; r < n; ++r) {
        const double* p = xyz + 12 * r;
        Tetra<const double> XYZ(p, p + 3, p + 6, p + 9);
        DenseMatrix<double> Am(buf, nfb, nfa, nfa * nfb);
        TetData td{D ? D + ddim * r : nullptr, ddim};
        if (order <= 6) {
            fem3Dtet<OpA, OpB, DfuncTraits<SPARSITY, false>>(XYZ, StoredTensor<TT, XDEP>(), Am, order, &td);
        } else {
            // the stack-memory version stops at order 6 (int_tet.inl:99-100);
            // higher orders run through the external-memory overload (int_tet.h:23-29)
            PlainMemory<double, int> req = fem3Dtet_memory_requirements<OpA, OpB, double, int>(order, 1);
            std::vector<double> dm(req.dSize);
            std::vector<int> im(req.iSize);
            PlainMemory<double, int> mem;
            mem.ddata = dm.data();
            mem.idata = im.data();
            mem.dSize = dm.size();
            mem.iSize = im.size();
            fem3Dtet<OpA, OpB, DfuncTraits<SPARSITY, false>>(XYZ, StoredTensor<TT, XDEP>(), Am, mem, order, &td);
        }
        std::memcpy(A + nfa * nfb * r, buf, sizeof(double) * nfa * nfb);
    }
    return 0;
}

template <typename OpA, typename OpB>
int dispatch_tensor(int tt, int per_point, int order, long n, const double* xyz, const double* D, long ddim,
                    double* A) {
    if (per_point == 2) {
        switch (tt) {
            case TENSOR_SCALAR: return run_case<OpA, OpB, TENSOR_SCALAR, PerPoint, true>(order, n, xyz, D, ddim, A);
            case TENSOR_SYMMETRIC: return run_case<OpA, OpB, TENSOR_SYMMETRIC, PerPoint, true>(order, n, xyz, D, ddim, A);
            case TENSOR_GENERAL: return run_case<OpA, OpB, TENSOR_GENERAL, PerPoint, true>(order, n, xyz, D, ddim, A);
        }
    } else if (per_point) {
        switch (tt) {
            case TENSOR_NULL: return run_case<OpA, OpB, TENSOR_NULL, PerPoint>(order, n, xyz, D, ddim, A);
            case TENSOR_SCALAR: return run_case<OpA, OpB, TENSOR_SCALAR, PerPoint>(order, n, xyz, D, ddim, A);
            case TENSOR_SYMMETRIC: return run_case<OpA, OpB, TENSOR_SYMMETRIC, PerPoint>(order, n, xyz, D, ddim, A);
            case TENSOR_GENERAL: return run_case<OpA, OpB, TENSOR_GENERAL, PerPoint>(order, n, xyz, D, ddim, A);
        }
    } else {
        switch (tt) {
            case TENSOR_NULL: return run_case<OpA, OpB, TENSOR_NULL, TENSOR_NULL>(order, n, xyz, D, ddim, A);
            case TENSOR_SCALAR: return run_case<OpA, OpB, TENSOR_SCALAR, TENSOR_SCALAR>(order, n, xyz, D, ddim, A);
            case TENSOR_SYMMETRIC:
                return run_case<OpA, OpB, TENSOR_SYMMETRIC, TENSOR_SYMMETRIC>(order, n, xyz, D, ddim, A);
            case TENSOR_GENERAL:
                return run_case<OpA, OpB, TENSOR_GENERAL, TENSOR_GENERAL>(order, n, xyz, D, ddim, A);
        }
    }
    return -3;
}

// Face integrals over face `face` of each tet: fem3Dface, or fem3DfaceN when
// DN (anifem++/fem/operations/int_face.inl:12-82).
template <typename OpA, typename OpB, int TT, int SPARSITY, bool DN, bool XDEP = false>
int run_face_case(int order, int face, long n, const double* xyz, const double* D, long ddim, double* A) {
    constexpr long nfa = OpA::Nfa::value, nfb = OpB::Nfa::value;
    double buf[nfa * nfb > 1 ? nfa * nfb : 1];
    for (long r = 0; r < n; ++r) {
        const double* p = xyz + 12 * r;
        Tetra<const double> XYZ(p, p + 3, p + 6, p + 9);
        DenseMatrix<double> Am(buf, nfb, nfa, nfa * nfb);
        TetData td{D ? D + ddim * r : nullptr, ddim};
        if (DN)
            fem3DfaceN<OpA, OpB, DfuncTraits<SPARSITY, false>>(XYZ, face, StoredTensor<TT, XDEP>(), Am, order, &td);
        else
            fem3Dface<OpA, OpB, DfuncTraits<SPARSITY, false>>(XYZ, face, StoredTensor<TT, XDEP>(), Am, order, &td);
        std::memcpy(A + nfa * nfb * r, buf, sizeof(double) * nfa * nfb);
    }
    return 0;
}

template <typename OpA, typename OpB, bool DN>
int dispatch_face(int tt, int per_point, int order, int face, long n, const double* xyz, const double* D, long ddim,
                  double* A) {
    if (per_point == 2) {
        switch (tt) {
            case TENSOR_SCALAR:
                return run_face_case<OpA, OpB, TENSOR_SCALAR, PerPoint, DN, true>(order, face, n, xyz, D, ddim, A);
            case TENSOR_GENERAL:
                return run_face_case<OpA, OpB, TENSOR_GENERAL, PerPoint, DN, true>(order, face, n, xyz, D, ddim, A);
        }
        return -3;
    }
    switch (tt) {
        case TENSOR_NULL: return run_face_case<OpA, OpB, TENSOR_NULL, TENSOR_NULL, DN>(order, face, n, xyz, D, ddim, A);
        case TENSOR_SCALAR:
            return run_face_case<OpA, OpB, TENSOR_SCALAR, TENSOR_SCALAR, DN>(order, face, n, xyz, D, ddim, A);
        case TENSOR_GENERAL:
            return run_face_case<OpA, OpB, TENSOR_GENERAL, TENSOR_GENERAL, DN>(order, face, n, xyz, D, ddim, A);
    }
    return -3;
}

template <typename Op>
int run_apply(int q, const double* xyl, long n, const double* xyz, const double* dofs, double* out) {
    constexpr long nfa = Op::Nfa::value, dim = Op::Dim::value;
    std::vector<double> L(xyl, xyl + 4 * q), dv(nfa), o(dim * q);
    for (long r = 0; r < n; ++r) {
        const double* p = xyz + 12 * r;
        Tetra<const double> XYZ(p, p + 3, p + 6, p + 9);
        std::copy(dofs + nfa * r, dofs + nfa * (r + 1), dv.begin());
        DenseMatrix<double> D(dv.data(), nfa, 1, nfa), O(o.data(), dim * q, 1, dim * q);
        fem3DapplyL<Op>(XYZ, ArrayView<double>(L.data(), 4 * q), D, O);
        std::copy(o.begin(), o.end(), out + dim * q * r);
    }
    return 0;
}

template <typename Op>
int run_applyx(int q, long n, const double* xyz, const double* X, const double* dofs, double* out) {
    constexpr long nfa = Op::Nfa::value, dim = Op::Dim::value;
    std::vector<double> dv(nfa), o(dim * q);
    for (long r = 0; r < n; ++r) {
        const double* p = xyz + 12 * r;
        Tetra<const double> XYZ(p, p + 3, p + 6, p + 9);
        std::copy(dofs + nfa * r, dofs + nfa * (r + 1), dv.begin());
        fem3DapplyX<Op>(XYZ, ArrayView<const double>(X + 3 * q * r, 3 * q), ArrayView<double>(dv.data(), nfa),
                        ArrayView<double>(o.data(), dim * q));
        std::copy(o.begin(), o.end(), out + dim * q * r);
    }
    return 0;
}

template <typename Op>
int run_apply_any(bool at_x, int q, const double* pts, long n, const double* xyz, const double* dofs, double* out) {
    return at_x ? run_applyx<Op>(q, n, xyz, pts, dofs, out) : run_apply<Op>(q, pts, n, xyz, dofs, out);
}

int apply_switch(int code, bool at_x, int q, const double* pts, long n, const double* xyz, const double* dofs,
                 double* out) {
    using P1 = FemFix<FEM_P1>;
    using P2 = FemFix<FEM_P2>;
    using P3 = FemFix<FEM_P3>;
    using V2 = FemVec<3, FEM_P2>;
    using RT0 = FemFix<FEM_RT0>;
    using ND0 = FemFix<FEM_ND0>;
    using CR1 = FemFix<FEM_CR1>;
    using B2 = FemFix<FEM_B2>;
    using B4 = FemFix<FEM_B4>;
    try {
        switch (code) {
            case 0: return run_apply_any<Operator<IDEN, P1>>(at_x, q, pts, n, xyz, dofs, out);
            case 1: return run_apply_any<Operator<IDEN, P2>>(at_x, q, pts, n, xyz, dofs, out);
            case 2: return run_apply_any<Operator<GRAD, P2>>(at_x, q, pts, n, xyz, dofs, out);
            case 3: return run_apply_any<Operator<IDEN, P3>>(at_x, q, pts, n, xyz, dofs, out);
            case 4: return run_apply_any<Operator<GRAD, P3>>(at_x, q, pts, n, xyz, dofs, out);
            case 5: return run_apply_any<Operator<IDEN, V2>>(at_x, q, pts, n, xyz, dofs, out);
            case 6: return run_apply_any<Operator<GRAD, V2>>(at_x, q, pts, n, xyz, dofs, out);
            case 7: return run_apply_any<Operator<IDEN, RT0>>(at_x, q, pts, n, xyz, dofs, out);
            case 8: return run_apply_any<Operator<DIV, RT0>>(at_x, q, pts, n, xyz, dofs, out);
            case 9: return run_apply_any<Operator<IDEN, ND0>>(at_x, q, pts, n, xyz, dofs, out);
            case 10: return run_apply_any<Operator<CURL, ND0>>(at_x, q, pts, n, xyz, dofs, out);
            case 11: return run_apply_any<Operator<IDEN, CR1>>(at_x, q, pts, n, xyz, dofs, out);
            case 12: return run_apply_any<Operator<GRAD, B2>>(at_x, q, pts, n, xyz, dofs, out);
            case 13: return run_apply_any<Operator<GRAD, B4>>(at_x, q, pts, n, xyz, dofs, out);
            default: break;
        }
    } catch (const std::exception&) {
        return -2;
    }
    return -1;
}

}  // namespace

extern "C" {

// Number of points of the reference triangle rule; fills p[3*q], w[q].
int ref_triangle_quadrature(int order, double* p, double* w) {
    try {
        auto f = triangle_quadrature_formulas(order);
        int q = f.GetNumPoints();
        if (p) std::memcpy(p, f.p, sizeof(double) * 3 * q);
        if (w) std::memcpy(w, f.w, sizeof(double) * q);
        return q;
    } catch (...) {
        return -1;
    }
}

// fem3DapplyL (anifem++/fem/operations/eval.inl:84-100) per tet.  Op codes:
// 0 IDEN P1, 1 IDEN P2, 2 GRAD P2, 3 IDEN P3, 4 GRAD P3, 5 IDEN V2, 6 GRAD V2,
// 7 IDEN RT0, 8 DIV RT0, 9 IDEN ND0, 10 CURL ND0, 11 IDEN CR1, 12 GRAD B2, 13 GRAD B4.
// xyl: 4q barycentric points; dofs n x nfa; out n x dim*q.
int ref_fem3dapply(int code, int q, const double* xyl, long n, const double* xyz, const double* dofs, double* out) {
    return apply_switch(code, false, q, xyl, n, xyz, dofs, out);
}

// fem3DapplyX at q physical points per tet, X[(r*q + n)*3 + c]; codes as ref_fem3dapply.
int ref_fem3dapplyx(int code, int q, const double* X, long n, const double* xyz, const double* dofs, double* out) {
    return apply_switch(code, true, q, X, n, xyz, dofs, out);
}

// Face pair codes (OpA, OpB, dot_normal): 0=(IDEN P1, IDEN P1); 1=(IDEN P2, IDEN P2); 2=(IDEN P3, IDEN P3);
// 3=(IDEN P0, IDEN P2); 4=(IDEN P0, IDEN V2); 5=(GRAD P2, IDEN P2, normal); 6=(GRAD V2, IDEN V2, normal);
// 7=(GRAD P1, GRAD P1); 8=(GRAD P3, IDEN P3, normal).  Layouts as ref_fem3dtet.
int ref_fem3dface(int pair, int tensor_type, int per_point, int order, int face, long n, const double* xyz,
                  const double* D, long ddim, double* A) {
    using P0 = FemFix<FEM_P0>;
    using P1 = FemFix<FEM_P1>;
    using P2 = FemFix<FEM_P2>;
    using P3 = FemFix<FEM_P3>;
    using V2 = FemVec<3, FEM_P2>;
    try {
        switch (pair) {
            case 0: return dispatch_face<Operator<IDEN, P1>, Operator<IDEN, P1>, false>(tensor_type, per_point, order, face, n, xyz, D, ddim, A);
            case 1: return dispatch_face<Operator<IDEN, P2>, Operator<IDEN, P2>, false>(tensor_type, per_point, order, face, n, xyz, D, ddim, A);
            case 2: return dispatch_face<Operator<IDEN, P3>, Operator<IDEN, P3>, false>(tensor_type, per_point, order, face, n, xyz, D, ddim, A);
            case 3: return dispatch_face<Operator<IDEN, P0>, Operator<IDEN, P2>, false>(tensor_type, per_point, order, face, n, xyz, D, ddim, A);
            case 4: return dispatch_face<Operator<IDEN, P0>, Operator<IDEN, V2>, false>(tensor_type, per_point, order, face, n, xyz, D, ddim, A);
            case 5: return dispatch_face<Operator<GRAD, P2>, Operator<IDEN, P2>, true>(tensor_type, per_point, order, face, n, xyz, D, ddim, A);
            case 6: return dispatch_face<Operator<GRAD, V2>, Operator<IDEN, V2>, true>(tensor_type, per_point, order, face, n, xyz, D, ddim, A);
            case 7: return dispatch_face<Operator<GRAD, P1>, Operator<GRAD, P1>, false>(tensor_type, per_point, order, face, n, xyz, D, ddim, A);
            case 8: return dispatch_face<Operator<GRAD, P3>, Operator<IDEN, P3>, true>(tensor_type, per_point, order, face, n, xyz, D, ddim, A);
            default: break;
        }
    } catch (const std::exception&) {
        return -2;
    }
    return -1;
}

// Number of points of the reference rule; fills p[4*q], w[q] when non-null.
int ref_quadrature(int order, double* p, double* w) {
    try {
        auto f = tetrahedron_quadrature_formulas(order);
        int q = f.GetNumPoints();
        if (p) std::memcpy(p, f.p, sizeof(double) * 4 * q);
        if (w) std::memcpy(w, f.w, sizeof(double) * q);
        return q;
    } catch (...) {
        return -1;
    }
}

// Pair codes (OpA, OpB): 0=(GRAD P1, GRAD P1); 1=(GRAD P2, GRAD P2); 2=(GRAD P3, GRAD P3); 3=(GRAD V1, GRAD V1); 4=(GRAD V2, GRAD V2); 5=(IDEN P0, IDEN P1); 6=(IDEN P0, IDEN P2); 7=(IDEN P0, IDEN P3); 8=(IDEN P0, IDEN V2); 9=(IDEN P1, IDEN P1); 10=(IDEN P2, IDEN P2); 11=(IDEN P3, IDEN P3); 12=(GRAD P3, GRAD V1); 13=(IDEN P0, IDEN V1); 14=(GRAD P2, IDEN P2);
// 15=(DIV V2, DIV V2); 16=(DIV V2, IDEN P1); 17=(CURL V1, CURL V1); 18=(CURL V2, IDEN V2);
// 19=(DUDX P2, DUDY P2); 20=(DUDZ P1, IDEN P1);
// 21=(IDEN RT0, IDEN RT0); 22=(DIV RT0, DIV RT0); 23=(DIV RT0, IDEN P0); 24=(IDEN ND0, IDEN ND0);
// 25=(CURL ND0, CURL ND0); 26=(CURL ND0, IDEN RT0); 27=(GRAD CR1, GRAD CR1); 28=(IDEN CR1, IDEN P1);
// 29=(GRAD B2, GRAD B2); 30=(IDEN B4, IDEN B4); 31=(GRAD B4, GRAD P1); 32=(IDEN P1, IDEN ND0);
// 33=(GRAD B4, GRAD B4); 34=(GRAD RT0, GRAD RT0); 35=(GRAD ND0, GRAD ND0); 36=(GRAD ND0, GRAD RT0).
//  xyz is n x 12 (P0,P1,P2,P3 per tet), D is n x ddim.
// A receives n element matrices, each nfB x nfA column-major.
int ref_fem3dtet(int pair, int tensor_type, int per_point, int order, long n, const double* xyz, const double* D,
                 long ddim, double* A) {
    using P0 = FemFix<FEM_P0>;
    using P1 = FemFix<FEM_P1>;
    using P2 = FemFix<FEM_P2>;
    using P3 = FemFix<FEM_P3>;
    using V1 = FemVec<3, FEM_P1>;
    using V2 = FemVec<3, FEM_P2>;
    using RT0 = FemFix<FEM_RT0>;
    using ND0 = FemFix<FEM_ND0>;
    using CR1 = FemFix<FEM_CR1>;
    using B2 = FemFix<FEM_B2>;
    using B4 = FemFix<FEM_B4>;
    try {
        switch (pair) {
            case 0: return dispatch_tensor<Operator<GRAD, P1>, Operator<GRAD, P1>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 1: return dispatch_tensor<Operator<GRAD, P2>, Operator<GRAD, P2>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 2: return dispatch_tensor<Operator<GRAD, P3>, Operator<GRAD, P3>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 3: return dispatch_tensor<Operator<GRAD, V1>, Operator<GRAD, V1>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 4: return dispatch_tensor<Operator<GRAD, V2>, Operator<GRAD, V2>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 5: return dispatch_tensor<Operator<IDEN, P0>, Operator<IDEN, P1>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 6: return dispatch_tensor<Operator<IDEN, P0>, Operator<IDEN, P2>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 7: return dispatch_tensor<Operator<IDEN, P0>, Operator<IDEN, P3>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 8: return dispatch_tensor<Operator<IDEN, P0>, Operator<IDEN, V2>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 9: return dispatch_tensor<Operator<IDEN, P1>, Operator<IDEN, P1>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 10: return dispatch_tensor<Operator<IDEN, P2>, Operator<IDEN, P2>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 11: return dispatch_tensor<Operator<IDEN, P3>, Operator<IDEN, P3>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 12: return dispatch_tensor<Operator<GRAD, P3>, Operator<GRAD, V1>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 13: return dispatch_tensor<Operator<IDEN, P0>, Operator<IDEN, V1>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 14: return dispatch_tensor<Operator<GRAD, P2>, Operator<IDEN, P2>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 15: return dispatch_tensor<Operator<DIV, V2>, Operator<DIV, V2>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 16: return dispatch_tensor<Operator<DIV, V2>, Operator<IDEN, P1>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 17: return dispatch_tensor<Operator<CURL, V1>, Operator<CURL, V1>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 18: return dispatch_tensor<Operator<CURL, V2>, Operator<IDEN, V2>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 19: return dispatch_tensor<Operator<DUDX, P2>, Operator<DUDY, P2>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 20: return dispatch_tensor<Operator<DUDZ, P1>, Operator<IDEN, P1>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 21: return dispatch_tensor<Operator<IDEN, RT0>, Operator<IDEN, RT0>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 22: return dispatch_tensor<Operator<DIV, RT0>, Operator<DIV, RT0>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 23: return dispatch_tensor<Operator<DIV, RT0>, Operator<IDEN, P0>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 24: return dispatch_tensor<Operator<IDEN, ND0>, Operator<IDEN, ND0>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 25: return dispatch_tensor<Operator<CURL, ND0>, Operator<CURL, ND0>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 26: return dispatch_tensor<Operator<CURL, ND0>, Operator<IDEN, RT0>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 27: return dispatch_tensor<Operator<GRAD, CR1>, Operator<GRAD, CR1>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 28: return dispatch_tensor<Operator<IDEN, CR1>, Operator<IDEN, P1>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 29: return dispatch_tensor<Operator<GRAD, B2>, Operator<GRAD, B2>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 30: return dispatch_tensor<Operator<IDEN, B4>, Operator<IDEN, B4>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 31: return dispatch_tensor<Operator<GRAD, B4>, Operator<GRAD, P1>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 32: return dispatch_tensor<Operator<IDEN, P1>, Operator<IDEN, ND0>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 33: return dispatch_tensor<Operator<GRAD, B4>, Operator<GRAD, B4>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 34: return dispatch_tensor<Operator<GRAD, RT0>, Operator<GRAD, RT0>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 35: return dispatch_tensor<Operator<GRAD, ND0>, Operator<GRAD, ND0>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            case 36: return dispatch_tensor<Operator<GRAD, ND0>, Operator<GRAD, RT0>>(tensor_type, per_point, order, n, xyz, D, ddim, A);
            default: break;
        }
    } catch (const std::exception&) {
        return -2;
    }
    return -1;
}

}  // extern "C"
