// Internal declarations shared by the host and device translation units of
// libanifem_gpu.so.  Not part of the C ABI (see include/anifem_gpu.h).
#pragma once

#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/anifem_gpu.h"

namespace anifem {

// Finite-element space data (poly_0.h .. poly_3.h and spaces/*.h of the reference).
int nf_scalar(int fem);          // basis functions per component: 1, 4, 10, 20 for P0..P3, ...; -1 otherwise
bool fem_lagrange(int fem);      // P0..P3: the spaces of the global assembly
bool fem_vector(int fem);        // RT0, ND0: vector-valued bases, one component
// per component: 1 for IDEN of a scalar space, 3 for the gradient-based
// operators; for RT0 / ND0 the reference field's dimension; -1 otherwise
int ref_dim(int op, int fem);
int op_dim(const anifem_op& op);  // the operator's dimension (Dim of operators.h)
bool op_valid(const anifem_op& op);
// How the physical values of OP(phi_i) follow from the reference table:
//   map   0: U = PSI^T Uh (gradients, ND0 values) or U = Uh (dimension 1);
//         1: U = J Uh (RT0 values, ND0 curls: the Piola map without 1/det J)
//   scale 0: none; 1: |face i| / (|det J| / 2) (RT0); 2: |edge i| (ND0 values);
//         3: |edge i| / det J (ND0 curls), a per-tet factor of basis function i
void space_map(int op, int fem, int& map, int& scale);

// Quadrature rule for `order` (0..kTetMaxOrder; generated table tet_rules.inc
// up to order 6, conical Gauss-Jacobi products above); returns q or -1.
constexpr int kTetMaxOrder = 20;
constexpr int kTriMaxOrder = 20;  // triangle rules (face integrals), same construction above order 6
int quadrature(int order, const double** p, const double** w);

// Reference-space operator table of a space at the rule's points:
// tab[(n*nf + i)*d + a] = component a (d = ref_dim) of OP(phi_i) at point n,
// with gradients taken in the reference tetrahedron's coordinates
// (x1, x2, x3) = (l1, l2, l3).
void reference_table(int op, int fem, int q, const double* p, std::vector<double>& tab);

// Reference tensor of the element contraction for scalar bases (OpA0, OpB0).
//   per_point == false:  R[rr*KB + b*dA + a] = sum_n w_n V_b(n, ib) U_a(n, ia)
//   per_point == true :  R[(rr*q + n)*KB + b*dA + a] = V_b(n, ib) U_a(n, ia)
// with rr = ib + nfB0*ia and KB = dA*dB.
// face < 0: volume rule; face 0..3: the triangle rule on that face (face integrals)
void reference_tensor(int opA, int femA, int opB, int femB, int order, bool per_point, std::vector<double>& R,
                      int face = -1);
int triangle_quadrature(int order, const double** p, const double** w);
int rule_points(int order, int face, std::vector<double>& pts, std::vector<double>& wts);

// Device-resident copies of reference tensors, built lazily per (pair, order,
// per_point) and kept for the life of the process.
struct RefTensorDev {
    const double* r64 = nullptr;
    const float* r32 = nullptr;
    int E0 = 0;  // nfA0*nfB0
    int KB = 0;  // dA*dB
    int q = 0;   // points (per-point tensors) or 1
    const double* w = nullptr;  // quadrature weights on the device (q values)
    // FP16 hi + lo split of R / r16_scale for the MFMA contraction (q == 1 and
    // KB <= 16 only): [2][e0pad][16] halves, row e0, column kk (zero padded)
    const void* r16 = nullptr;
    float r16_scale = 1.0f;
    int e0pad = 0;
    // 3 x 3 tensors constant over the tet (KB == 9, q == 1): R with the
    // columns of K(a, b) and K(b, a) summed, for a symmetric K packed as
    // K00 K01 K11 K02 K12 K22 (the contraction sum_kk K R then runs over 6
    // terms); E0 x 6, FP64 and FP32
    const double* rs64 = nullptr;
    const float* rs32 = nullptr;
    // rs64 row (ib + nfB ia) equals row (ia + nfB ib) bit for bit (one space
    // and one operator on both sides; checked when the table is built)
    bool rs_sym = false;
};
// Packed index of K(a, b) = K(b, a) in the symmetric layout above.
__host__ __device__ inline int sym_pack(int a, int b) { return a <= b ? a + b * (b + 1) / 2 : b + a * (a + 1) / 2; }
int get_ref_tensor(int opA, int femA, int opB, int femB, int order, bool per_point, RefTensorDev* out,
                   int face = -1);
// Set by entry points that must not allocate device memory: get_ref_tensor
// then only looks tables up and fails with ANIFEM_ERR_ARG where one would
// have to be built.
extern thread_local bool g_ref_lookup_only;

// Symmetric-tensor packing: D(i,j), i<=j, at i + j*(j+1)/2.
inline int sym_index(int i, int j) { return i <= j ? i + j * (j + 1) / 2 : j + i * (i + 1) / 2; }

// Tensor data size per tet (or per point) for a tensor type and operator dims.
int tensor_ddim(int type, int dimA, int dimB);

#define ANIFEM_HIP_CHECK(expr)                                                      \
    do {                                                                           \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            std::fprintf(stderr, "anifem: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            return ANIFEM_ERR_HIP;                                                  \
        }                                                                          \
    } while (0)

}  // namespace anifem
