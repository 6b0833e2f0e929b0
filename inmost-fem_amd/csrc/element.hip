// Batched element matrices: anifem_fem3dtet_batch (C ABI) and the launch
// helpers shared with the fused assembly.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "anifem_common.h"
#include "element_kernels.h"

namespace anifem {

// Operator-dependent part of ElemArgs; validates the pair like the reference's
// compile-time / runtime checks (int_tet.inl:94-104, diff_tensor.h:313-315).
// Element arguments of a volume integral (face < 0) or of the integral over
// one face, plain (fem3Dface) or of (D OpA(u)) . n (fem3DfaceN, dot_normal):
// there D has 3 dimB rows, and a.dimB is that row count (the tensor's layout).
static int prepare_args(const anifem_op& opA, const anifem_op& opB, const anifem_tensor* D, int order, int face,
                        int dot_normal, ElemArgs& a) {
    if (!op_valid(opA) || !op_valid(opB) || !D) return ANIFEM_ERR_ARG;
    if (face > 3 || (face < 0 && dot_normal)) return ANIFEM_ERR_ARG;
    std::vector<double> pts, wts;
    int q = rule_points(order, face, pts, wts);
    if (q < 0) return ANIFEM_ERR_ORDER;
    a.face = face < 0 ? -1 : face;
    a.dot_normal = dot_normal ? 1 : 0;
    a.nfA0 = nf_scalar(opA.fem);
    a.nfB0 = nf_scalar(opB.fem);
    // DIV / CURL / DUDX..DUDZ run as the gradient with a signed selector in
    // the tensor (element_kernels.h: op_map); the normal form takes them on
    // the trial side only
    if (dot_normal && opB.op != ANIFEM_IDEN && opB.op != ANIFEM_GRAD) return ANIFEM_ERR_ARG;
    // RT0 / ND0: the operator is the space's own (no selector); their values
    // follow from the reference fields through space_map's map and factors
    a.selA = fem_vector(opA.fem) ? ANIFEM_IDEN : opA.op;
    a.selB = fem_vector(opB.fem) ? ANIFEM_IDEN : opB.op;
    space_map(opA.op, opA.fem, a.mapA, a.scA);
    space_map(opB.op, opB.fem, a.mapB, a.scB);
    // GRAD of RT0 / ND0 (9 components) pairs with GRAD of RT0 / ND0 only, and
    // not in the normal form
    if ((a.mapA >= 3) != (a.mapB >= 3) || (a.mapA >= 3 && dot_normal)) return ANIFEM_ERR_ARG;
    a.dimA = op_dim(opA);
    a.dimB = op_dim(opB) * (dot_normal ? 3 : 1);
    a.E = a.nfA0 * opA.ncomp * a.nfB0 * opB.ncomp;
    a.ttype = D->type;
    a.storage = D->storage;
    a.D = D->data;
    const int nfA = a.nfA0 * opA.ncomp;
    a.p0_bcast = (a.dimA != a.dimB && nfA == 1 && a.dimA == 1) ? 1 : 0;
    if ((D->type == ANIFEM_TENSOR_NULL || D->type == ANIFEM_TENSOR_SCALAR) && a.dimA != a.dimB && !a.p0_bcast)
        return ANIFEM_ERR_DIMS;
    if (tensor_ddim(D->type, a.dimA, a.dimB) < 0) return ANIFEM_ERR_ARG;
    if (D->type != ANIFEM_TENSOR_NULL && !D->data) return ANIFEM_ERR_ARG;
    if (D->storage < ANIFEM_COEF_CONST || D->storage > ANIFEM_COEF_PER_POINT) return ANIFEM_ERR_ARG;
    RefTensorDev R;
    int rc = get_ref_tensor(opA.op, opA.fem, opB.op, opB.fem, order, D->storage == ANIFEM_COEF_PER_POINT, &R, a.face);
    if (rc) return rc;
    a.q = R.q;
    a.Rd = R.r64;
    a.Rf = R.r32;
    a.Rds = R.rs64;
    a.Rfs = R.rs32;
    a.rds_sym = R.rs_sym ? 1 : 0;
    a.wq = R.w;
    a.r16 = R.r16;
    a.r16_scale = R.r16_scale;
    a.e0pad = R.e0pad;
    return ANIFEM_OK;
}

int prepare_elem_args(const anifem_op& opA, const anifem_op& opB, const anifem_tensor* D, int order, ElemArgs& a) {
    return prepare_args(opA, opB, D, order, -1, 0, a);
}

// Tets per workgroup: as many as keep the LDS geometry tensors <= 32 KiB
// (several workgroups per CU), at most `cap`.
static int choose_tile(int kstride, int elem_bytes, int cap) {
    int t = cap;
    while (t > 4 && (size_t)t * kstride * elem_bytes > 32768) t /= 2;
    return t;
}

template <int DA0, int DB0, int NCA, int NCB, typename T>
static int launch_one(ElemArgs a, int cap, hipStream_t s) {
    const int kstride = NCA * NCB * a.q * DA0 * DB0;
    a.tile = choose_tile(kstride, (int)sizeof(T), cap);
    size_t lds = (size_t)a.tile * kstride * sizeof(T);
    if (a.scA | a.scB)
        lds = element_scale_offset(a.tile, kstride, (int)sizeof(T)) + (size_t)a.tile * (a.nfA0 + a.nfB0) * sizeof(double);
    if (lds > 64 * 1024) return ANIFEM_ERR_ARG;
    const int64_t blocks = (a.n + a.tile - 1) / a.tile;
    if (blocks == 0) return ANIFEM_OK;
    if (a.mapA | a.mapB | a.scA | a.scB)
        hipLaunchKernelGGL((element_kernel<DA0, DB0, NCA, NCB, T, true>), dim3((unsigned)blocks), dim3(kBlock), lds, s,
                           a);
    else
        hipLaunchKernelGGL((element_kernel<DA0, DB0, NCA, NCB, T, false>), dim3((unsigned)blocks), dim3(kBlock), lds,
                           s, a);
    ANIFEM_HIP_CHECK(hipGetLastError());
    return ANIFEM_OK;
}

template <typename T, int DA0, int DB0>
static int launch_nc(int ncA, int ncB, const ElemArgs& a, int cap, hipStream_t s) {
    if (ncA == 1 && ncB == 1) return launch_one<DA0, DB0, 1, 1, T>(a, cap, s);
    if (ncA == 3 && ncB == 3) return launch_one<DA0, DB0, 3, 3, T>(a, cap, s);
    if (ncA == 1 && ncB == 3) return launch_one<DA0, DB0, 1, 3, T>(a, cap, s);
    if (ncA == 3 && ncB == 1) return launch_one<DA0, DB0, 3, 1, T>(a, cap, s);
    return ANIFEM_ERR_ARG;
}

template <typename T>
int launch_element(const anifem_op& opA, const anifem_op& opB, const ElemArgs& a, int cap, hipStream_t s) {
    const int dA = ref_dim(opA.op, opA.fem), dB = ref_dim(opB.op, opB.fem);
    if (dA == 3 && dB == 3) return launch_nc<T, 3, 3>(opA.ncomp, opB.ncomp, a, cap, s);
    if (dA == 1 && dB == 1) return launch_nc<T, 1, 1>(opA.ncomp, opB.ncomp, a, cap, s);
    if (dA == 3 && dB == 1) return launch_nc<T, 3, 1>(opA.ncomp, opB.ncomp, a, cap, s);
    if (dA == 1 && dB == 3) return launch_nc<T, 1, 3>(opA.ncomp, opB.ncomp, a, cap, s);
    return ANIFEM_ERR_ARG;
}

template <int DA0, int DB0, int NCA, int NCB>
static int launch_mfma_one(const ElemArgs& a, hipStream_t s) {
    if (NCA * NCB == 1) {
        const int64_t blocks = (a.n + kMfmaTets - 1) / kMfmaTets;
        hipLaunchKernelGGL((element_mfma_kernel<DA0, DB0, NCA, NCB>), dim3((unsigned)blocks), dim3(kBlock), 0, s, a);
    } else {
        const int64_t blocks = (a.n + kMfmaVecTets - 1) / kMfmaVecTets;
        hipLaunchKernelGGL((element_mfma_vec_kernel<DA0, DB0, NCA, NCB>), dim3((unsigned)blocks), dim3(kBlock), 0, s,
                           a);
    }
    ANIFEM_HIP_CHECK(hipGetLastError());
    return ANIFEM_OK;
}

template <int DA0, int DB0>
static int launch_mfma_nc(int ncA, int ncB, const ElemArgs& a, hipStream_t s) {
    if (ncA == 1 && ncB == 1) return launch_mfma_one<DA0, DB0, 1, 1>(a, s);
    if (ncA == 3 && ncB == 3) return launch_mfma_one<DA0, DB0, 3, 3>(a, s);
    if (ncA == 1 && ncB == 3) return launch_mfma_one<DA0, DB0, 1, 3>(a, s);
    if (ncA == 3 && ncB == 1) return launch_mfma_one<DA0, DB0, 3, 1>(a, s);
    return ANIFEM_ERR_ARG;
}

// FP16 hi + lo MFMA contraction: tensors constant over each tet only
int launch_element_mfma(const anifem_op& opA, const anifem_op& opB, const ElemArgs& a, hipStream_t s) {
    if (a.q != 1 || !a.r16) return ANIFEM_ERR_ARG;
    if (a.mapA | a.mapB | a.scA | a.scB) return ANIFEM_ERR_ARG;  // RT0 / ND0: FP64 and FP32 modes only
    const int dA = ref_dim(opA.op, opA.fem), dB = ref_dim(opB.op, opB.fem);
    if (dA == 3 && dB == 3) return launch_mfma_nc<3, 3>(opA.ncomp, opB.ncomp, a, s);
    if (dA == 1 && dB == 1) return launch_mfma_nc<1, 1>(opA.ncomp, opB.ncomp, a, s);
    if (dA == 3 && dB == 1) return launch_mfma_nc<3, 1>(opA.ncomp, opB.ncomp, a, s);
    if (dA == 1 && dB == 3) return launch_mfma_nc<1, 3>(opA.ncomp, opB.ncomp, a, s);
    return ANIFEM_ERR_ARG;
}

// Physical quadrature points of every tet (internalFem3DtetGeomInit_XYG,
// core.inl:250-270): X_n = P0 + sum_l lambda_{l+1,n} (P_{l+1} - P0), one lane
// per tet, stored X[(c*q + n)*n + r] (coalesced over tets; the layout of
// per-point coefficient arrays).
struct RulePoints {
    double lam[24][3];    // lambda_1..3 of each point (rules of up to 24 points)
    const double* lamd;   // the same on the device, 3 per point (larger rules), or null
    int q;
};

template <bool MESH>
__global__ __launch_bounds__(kBlock) void physical_points_kernel(ElemArgs a, RulePoints rp, double* __restrict__ X) {
    const int64_t r = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (r >= a.n) return;
    double P[4][3];
    load_tet<MESH>(a, r, P);
    for (int n = 0; n < rp.q; ++n)
        for (int c = 0; c < 3; ++c) {
            double x = P[0][c];
            for (int l = 0; l < 3; ++l) x += (rp.lamd ? rp.lamd[3 * n + l] : rp.lam[n][l]) * (P[l + 1][c] - P[0][c]);
            X[((int64_t)c * rp.q + n) * a.n + r] = x;
        }
}

// lambda_1..3 of the points of rule `order` on the current device, uploaded
// once per device and kept until the process ends.
static const double* rule_lambda_dev(int order, int q, const double* p) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, double*> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({dev, order});
    if (it != cache.end()) return it->second;
    std::vector<double> lam((size_t)3 * q);
    for (int k = 0; k < q; ++k)
        for (int l = 0; l < 3; ++l) lam[3 * k + l] = p[4 * k + l + 1];
    double* d = nullptr;
    if (hipMalloc(&d, lam.size() * sizeof(double)) != hipSuccess) return nullptr;
    if (hipMemcpy(d, lam.data(), lam.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    cache[{dev, order}] = d;
    return d;
}

// Physical points of the volume rule `order` for n tets, from SoA vertices
// (xyz) or from a mesh (coords + SoA tet nodes, tets [pos0, pos0 + n)).
int launch_physical_points(int order, int64_t n, const double* xyz, const double* coords, const int32_t* tetnodes,
                           int64_t nodes_stride, double* X, hipStream_t s) {
    const double *p, *w;
    const int q = quadrature(order, &p, &w);
    if (q < 0) return ANIFEM_ERR_ORDER;
    if (n < 0 || !X || (!xyz && (!coords || !tetnodes))) return ANIFEM_ERR_ARG;
    RulePoints rp{};
    rp.q = q;
    if (q <= 24) {
        for (int k = 0; k < q; ++k)
            for (int l = 0; l < 3; ++l) rp.lam[k][l] = p[4 * k + l + 1];
    } else {
        rp.lamd = rule_lambda_dev(order, q, p);
        if (!rp.lamd) return ANIFEM_ERR_HIP;
    }
    ElemArgs a{};
    a.n = n;
    a.xyz = xyz;
    a.coords = coords;
    a.tetnodes = tetnodes;
    a.nodes_stride = nodes_stride;
    a.pos0 = 0;
    if (n == 0) return ANIFEM_OK;
    const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    if (xyz)
        hipLaunchKernelGGL(physical_points_kernel<false>, dim3(blocks), dim3(kBlock), 0, s, a, rp, X);
    else
        hipLaunchKernelGGL(physical_points_kernel<true>, dim3(blocks), dim3(kBlock), 0, s, a, rp, X);
    ANIFEM_HIP_CHECK(hipGetLastError());
    return ANIFEM_OK;
}

}  // namespace anifem

extern "C" int anifem_quadrature_points_batch(int order, int64_t n_tets, const double* xyz, double* X, void* stream) {
    if (!xyz) return ANIFEM_ERR_ARG;
    return anifem::launch_physical_points(order, n_tets, xyz, nullptr, nullptr, 0, X,
                                          static_cast<hipStream_t>(stream));
}

extern "C" int anifem_fem3dtet_batch(anifem_op opA, anifem_op opB, const anifem_tensor* D, int order, int64_t n_tets,
                                     const double* xyz, int precision, void* A_out, void* stream) {
    using namespace anifem;
    if (n_tets < 0 || (n_tets > 0 && (!xyz || !A_out))) return ANIFEM_ERR_ARG;
    ElemArgs a{};
    int rc = prepare_elem_args(opA, opB, D, order, a);
    if (rc) return rc;
    if (n_tets == 0) return ANIFEM_OK;
    a.n = n_tets;
    a.coef_n = n_tets;
    a.xyz = xyz;
    a.out = A_out;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (precision == ANIFEM_PREC_FP64) return launch_element<double>(opA, opB, a, 256, s);
    if (precision == ANIFEM_PREC_FP32) return launch_element<float>(opA, opB, a, 256, s);
    if (precision == ANIFEM_PREC_FP16X3) return launch_element_mfma(opA, opB, a, s);
    return ANIFEM_ERR_ARG;
}

extern "C" int anifem_fem3dface_batch(anifem_op opA, anifem_op opB, const anifem_tensor* D, int order, int face,
                                      int dot_normal, int64_t n_tets, const double* xyz, int precision, void* A_out,
                                      void* stream) {
    using namespace anifem;
    if (n_tets < 0 || (n_tets > 0 && (!xyz || !A_out)) || face < 0 || face > 3) return ANIFEM_ERR_ARG;
    ElemArgs a{};
    int rc = prepare_args(opA, opB, D, order, face, dot_normal, a);
    if (rc) return rc;
    if (n_tets == 0) return ANIFEM_OK;
    a.n = n_tets;
    a.coef_n = n_tets;
    a.xyz = xyz;
    a.out = A_out;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (precision == ANIFEM_PREC_FP64) return launch_element<double>(opA, opB, a, 256, s);
    if (precision == ANIFEM_PREC_FP32) return launch_element<float>(opA, opB, a, 256, s);
    return ANIFEM_ERR_ARG;  // the MFMA mode computes volume integrals only
}

namespace anifem {
// Operator-dependent part of ApplyArgs: IDEN / GRAD of the scalar spaces,
// IDEN / DIV / CURL of RT0 and ND0.
static int prepare_apply(const anifem_op& op, int q, ApplyArgs& a) {
    if (!op_valid(op) || q < 1 || q > 96) return ANIFEM_ERR_ARG;
    if (!fem_vector(op.fem) && op.op != ANIFEM_IDEN && op.op != ANIFEM_GRAD) return ANIFEM_ERR_ARG;
    int m = 0, sc = 0;
    space_map(op.op, op.fem, m, sc);
    a.d = ref_dim(op.op, op.fem);
    a.map = m >= 3 ? m
                   : (m == 1 ? 2 : ((op.op == ANIFEM_GRAD || (op.fem == ANIFEM_FEM_ND0 && op.op == ANIFEM_IDEN)) ? 1 : 0));
    a.scale = sc;
    a.op = op.op;
    a.fem = op.fem;
    a.nf0 = nf_scalar(op.fem);
    a.ncomp = op.ncomp;
    a.q = q;
    return ANIFEM_OK;
}

static int run_apply(ApplyArgs a, bool mesh, hipStream_t s) {
    const int64_t blocks = (a.n * a.q + kBlock - 1) / kBlock;
    if (mesh) hipLaunchKernelGGL(apply_kernel<true>, dim3((unsigned)blocks), dim3(kBlock), 0, s, a);
    else hipLaunchKernelGGL(apply_kernel<false>, dim3((unsigned)blocks), dim3(kBlock), 0, s, a);
    return hipGetLastError() == hipSuccess ? ANIFEM_OK : ANIFEM_ERR_HIP;
}

// Reference table of op at q barycentric points, on the device, and the
// launch of apply_kernel (the table is freed after the kernel has run).
int launch_apply(const anifem_op& op, const double* xyl, int q, ApplyArgs a, bool mesh, hipStream_t s) {
    if (!xyl) return ANIFEM_ERR_ARG;
    int rc = prepare_apply(op, q, a);
    if (rc) return rc;
    if (a.n == 0) return ANIFEM_OK;
    std::vector<double> tab;
    reference_table(op.op, op.fem, q, xyl, tab);
    double* dtab = nullptr;
    ANIFEM_HIP_CHECK(hipMalloc(&dtab, tab.size() * sizeof(double)));
    hipError_t e = hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        a.tab = dtab;
        a.X = nullptr;
        if (run_apply(a, mesh, s)) e = hipErrorLaunchFailure;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dtab);
    ANIFEM_HIP_CHECK(e);
    return ANIFEM_OK;
}
}  // namespace anifem

extern "C" int anifem_fem3dapplyx_batch(anifem_op op, int q, int64_t n_tets, const double* xyz, const double* X,
                                        const double* dofs, double* opU, void* stream) {
    using namespace anifem;
    if (n_tets < 0 || (n_tets > 0 && (!xyz || !X || !dofs || !opU))) return ANIFEM_ERR_ARG;
    ApplyArgs a{};
    int rc = prepare_apply(op, q, a);
    if (rc) return rc;
    if (n_tets == 0) return ANIFEM_OK;
    a.n = n_tets;
    a.xyz = xyz;
    a.X = X;
    a.tab = nullptr;
    a.dofs = dofs;
    a.out = opU;
    return run_apply(a, false, static_cast<hipStream_t>(stream));
}

extern "C" int anifem_fem3dapply_batch(anifem_op op, const double* xyl, int q, int64_t n_tets, const double* xyz,
                                       const double* dofs, double* opU, void* stream) {
    using namespace anifem;
    if (n_tets < 0 || (n_tets > 0 && (!xyz || !dofs || !opU))) return ANIFEM_ERR_ARG;
    ApplyArgs a{};
    a.n = n_tets;
    a.xyz = xyz;
    a.dofs = dofs;
    a.out = opU;
    return launch_apply(op, xyl, q, a, false, static_cast<hipStream_t>(stream));
}

extern "C" int anifem_face_quadrature(int order, int face, double* p, double* w) {
    std::vector<double> pts, wts;
    const int q = anifem::rule_points(order, face, pts, wts);
    if (q < 0 || face < 0 || face > 3) return q < 0 ? ANIFEM_ERR_ORDER : ANIFEM_ERR_ARG;
    if (p) std::copy(pts.begin(), pts.end(), p);
    if (w) std::copy(wts.begin(), wts.end(), w);
    return q;
}
