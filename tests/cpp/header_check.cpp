// Host-only check of include/anifem_fem3dtet.hpp: the reference-style tags map
// to the C ABI's operators, the traits reject mismatched views, and the
// template call reaches the library's argument checks (no device use).
#include <cstdio>

#include "anifem_fem3dtet.hpp"

using namespace anifem_gpu;

#define CHECK(c)                                            \
    do {                                                    \
        if (!(c)) {                                         \
            std::printf("FAIL line %d: %s\n", __LINE__, #c); \
            return 1;                                       \
        }                                                   \
    } while (0)

int main() {
    using GP2 = Operator<GRAD, FemFix<FEM_P2>>;
    using IP3 = Operator<IDEN, FemFix<FEM_P3>>;
    using GV2 = Operator<GRAD, FemVec<3, FEM_P2>>;
    constexpr anifem_op g = GP2::c();
    static_assert(g.op == ANIFEM_GRAD && g.fem == ANIFEM_FEM_P2 && g.ncomp == 1, "GRAD P2 tag");
    constexpr anifem_op v = GV2::c();
    static_assert(v.ncomp == 3 && v.fem == ANIFEM_FEM_P2, "FemVec tag");
    // Nfa / Dim (poly_2.h, poly_3.h, operators.h:127-155)
    CHECK(nfa<GP2>() == 10 && dim<GP2>() == 3);
    CHECK(nfa<IP3>() == 20 && dim<IP3>() == 1);
    CHECK(nfa<GV2>() == 30 && dim<GV2>() == 9);
    // DIV / CURL of FemVec<3, P2>, DUDX of P2 (operators.h:286-400)
    CHECK((nfa<Operator<DIV, FemVec<3, FEM_P2>>>() == 30 && dim<Operator<DIV, FemVec<3, FEM_P2>>>() == 1));
    CHECK((dim<Operator<CURL, FemVec<3, FEM_P2>>>() == 3 && dim<Operator<DUDX, FemFix<FEM_P2>>>() == 1));
    CHECK((dim<Operator<DIV, FemFix<FEM_P2>>>() < 0));
    // traits against views
    const TensorView sym_tet{TENSOR_SYMMETRIC, ANIFEM_COEF_PER_TET, nullptr};
    CHECK((check_view<DfuncTraits<TENSOR_SYMMETRIC, false>>(sym_tet) == ANIFEM_OK));
    CHECK((check_view<DfuncTraits<TENSOR_SCALAR, false>>(sym_tet) == ANIFEM_ERR_ARG));
    CHECK((check_view<DfuncTraits<TENSOR_SYMMETRIC, true>>(sym_tet) == ANIFEM_ERR_ARG));
    CHECK((check_view<DfuncTraits<PerTetra, false>>(sym_tet) == ANIFEM_OK));
    // the batch works on device memory, as the reference's dynamic-memory
    // fem3Dtet does, whose rules stop at order 20 (quadrature_formulas.cpp:1497-1498)
    double buf[16] = {0};
    CHECK((fem3Dtet<GP2, GP2, DfuncTraits<TENSOR_SYMMETRIC, false>>(buf, 1, sym_tet, buf, 21) == ANIFEM_ERR_ORDER));
    // identity tensor between GRAD (dim 3) and IDEN (dim 1) (diff_tensor.h:313-315)
    const TensorView null_const{TENSOR_NULL, ANIFEM_COEF_CONST, nullptr};
    CHECK((fem3Dtet<GP2, Operator<IDEN, FemFix<FEM_P2>>, DfuncTraits<TENSOR_NULL, true>>(buf, 1, null_const, buf, 2) ==
           ANIFEM_ERR_DIMS));
    // face integrals: the triangle rules stop at order 20; faces are 0..3
    using IP2 = Operator<IDEN, FemFix<FEM_P2>>;
    CHECK((fem3Dface<IP2, IP2, DfuncTraits<TENSOR_NULL, true>>(buf, 1, 0, null_const, buf, 21) == ANIFEM_ERR_ORDER));
    CHECK((fem3Dface<IP2, IP2, DfuncTraits<TENSOR_NULL, true>>(buf, 1, 4, null_const, buf, 2) == ANIFEM_ERR_ARG));
    // the normal form of an identity needs 3 dimB == dimA (GRAD against IDEN: 3 == 3)
    CHECK((fem3DfaceN<IP2, IP2, DfuncTraits<TENSOR_NULL, true>>(buf, 1, 0, null_const, buf, 2) == ANIFEM_ERR_DIMS));
    // a context that could not be created reports its status from every call
    Assembler<GP2, GP2> bad(0, nullptr, 0, nullptr, 2, sym_tet, 2, null_const, ANIFEM_DIRICHLET_NONE);
    CHECK(bad.status() != ANIFEM_OK);
    CHECK(bad.AssembleElements<double>(nullptr, nullptr, nullptr, nullptr) == bad.status());
    CHECK(bad.AssembleElements<float>(nullptr, nullptr, nullptr, nullptr) == bad.status());
    CHECK(bad.AssembleRHS<double>(nullptr, nullptr) == bad.status());
    // operators and spaces chosen at run time (operators.h:418-533, fem_space.h:404-596)
    const FemSpace p2 = P2Space{}, p0 = P0Space{}, rt0 = RT0Space{}, nd0 = ND0Space{};
    const ApplyOp gp2 = p2.getOP(GRAD);
    CHECK(gp2.Nfa() == 10 && gp2.Dim() == 3 && gp2.Order() == 1 && gp2.valid());
    CHECK((gp2 == ApplyOpFromTemplate<GRAD, FemFix<FEM_P2>>()));
    CHECK((p2.getOP(IDEN).Order() == 2 && p2.dimOP(IDEN) == 1 && p2.dim() == 1 && p2.nfa() == 10));
    const FemSpace v2 = p2 ^ 3;
    CHECK((v2.dim() == 3 && v2.nfa() == 30 && v2.dimOP(GRAD) == 9 && v2.dimOP(DIV) == 1 && v2.dimOP(CURL) == 3));
    CHECK((v2.getOP(GRAD) == ApplyOpFromTemplate<GRAD, FemVec<3, FEM_P2>>()));
    CHECK((rt0.dim() == 3 && rt0.nfa() == 4 && rt0.dimOP(DIV) == 1 && rt0.orderOP(DIV) == 0 && rt0.orderOP(IDEN) == 1));
    CHECK((nd0.nfa() == 6 && nd0.dimOP(CURL) == 3 && nd0.orderOP(CURL) == 0));
    CHECK((FemSpace(Bubble4Space{}).orderOP(GRAD) == 3 && FemSpace(CR1Space{}).orderOP(IDEN) == 1));
    CHECK((FemSpace(P0Space{}).orderOP(GRAD) == 0 && FemSpace(P1Space{}).orderOP(GRAD) == 0));
    CHECK((p2.getOP(DIV).valid() == false && p2.getOP(DIV).Dim() < 0));  // DIV of a scalar space
    // a product space (mixed Poisson: RT0 x P0): its variables, no single operator
    const FemSpace mixed = rt0 * p0;
    CHECK((mixed.isComplex() && mixed.nparts() == 2 && mixed.dim() == 4 && mixed.nfa() == 5));
    CHECK((!mixed.getOP(IDEN).valid() && mixed.part(1) == p0 && mixed.part(0) == rt0));
    // vector copies of RT0 / ND0 and more than 3 components are not built: no operator
    CHECK((!(rt0 ^ 2).getOP(IDEN).valid() && !(p2 ^ 4).getOP(GRAD).valid() && (v2 ^ 2).empty()));
    anifem_vars_desc vd{};
    CHECK((mixed.vars(&vd) == ANIFEM_OK && vd.n_var == 2 && vd.var[0].fem == ANIFEM_FEM_RT0 && vd.var[1].fem == ANIFEM_FEM_P0));
    CHECK(((p2 * p2 * p2 * p2 * p2).vars(&vd) == ANIFEM_ERR_ARG));
    // the runtime calls reach the same argument checks as the template ones
    CHECK((fem3Dtet(buf, 1, sym_tet, gp2, gp2, buf, 21) == ANIFEM_ERR_ORDER));
    CHECK((fem3Dtet(buf, 1, sym_tet, ApplyOp(), gp2, buf, 2) == ANIFEM_ERR_ARG));
    CHECK((fem3Dface(buf, 1, 4, null_const, p2.getOP(IDEN), p2.getOP(IDEN), buf, 2) == ANIFEM_ERR_ARG));
    RuntimeAssembler bad_rt(ApplyOp(), gp2, 0, nullptr, 0, nullptr, 2, sym_tet, 2, null_const, ANIFEM_DIRICHLET_NONE);
    CHECK(bad_rt.status() == ANIFEM_ERR_ARG);
    std::printf("header ok\n");
    return 0;
}
