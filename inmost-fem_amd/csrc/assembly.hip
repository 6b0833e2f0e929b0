// Whole-mesh assembly on the GPU: anifem_assembly_* (C ABI).
//
// create ~ AssemblerT::PrepareProblem (assembler.inl:194-275): the host setup
// (setup.cpp) is uploaded once.  run ~ AssemblerT::Assemble (assembler.inl:
// 314-488): per colour, one launch of the element kernel in scatter mode adds
// every tet's matrix into the FP64 CSR values through the slot map, then one
// launch adds the load vector.  Tets of one colour share no DOF, launches are
// stream-ordered, so there are no atomics on values and the summation order of
// every entry is fixed (colour order, then nothing concurrent).
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>

#include "anifem_common.h"
#include "element_kernels.h"
#include "setup.h"

namespace anifem {
int prepare_elem_args(const anifem_op& opA, const anifem_op& opB, const anifem_tensor* D, int order, ElemArgs& a);
template <typename T, bool MESH, int MODE>
int launch_element(const anifem_op& opA, const anifem_op& opB, const ElemArgs& a, int cap, hipStream_t s);
}  // namespace anifem

struct anifem_assembly {
    anifem::AssemblySetup S;
    anifem_problem_desc prob{};
    int device = 0;
    double* d_coords = nullptr;     // n_nodes x 3
    int32_t* d_tetnodes = nullptr;  // 4 x n_tets SoA, assembly order
    int32_t* d_orig = nullptr;      // assembly position -> caller tet id
    int32_t* d_slot = nullptr;      // n_tets x E, assembly order
    int32_t* d_dofs = nullptr;      // n_tets x nf, assembly order
    uint64_t* d_dirmask = nullptr;  // n_tets, assembly order (null without Dirichlet)
    int64_t* d_row_ptr = nullptr;
    int32_t* d_col_idx = nullptr;
    int* d_status = nullptr;        // [0] non-finite, [1] det <= 0
    bool host_only = false;         // built by anifem_assembly_create_host
    bool timers_on = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    anifem_timers timers{};
};

namespace {

__global__ void fill_kernel(double* p, int64_t n, double v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void fill_int_kernel(int* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

int fill(double* p, int64_t n, hipStream_t s) {
    if (n <= 0) return ANIFEM_OK;
    int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, 0.0);
    ANIFEM_HIP_CHECK(hipGetLastError());
    return ANIFEM_OK;
}

template <typename T>
int upload(T** dst, const T* src, size_t count, hipStream_t s) {
    ANIFEM_HIP_CHECK(hipMalloc((void**)dst, std::max<size_t>(count, 1) * sizeof(T)));
    if (count) ANIFEM_HIP_CHECK(hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    return ANIFEM_OK;
}

}  // namespace

extern "C" {

int anifem_assembly_create(anifem_assembly** out, const anifem_mesh_desc* mesh, const anifem_problem_desc* prob,
                           void* stream) {
    using namespace anifem;
    if (!out || !mesh || !prob || !mesh->coords || !mesh->tets) return ANIFEM_ERR_ARG;
    if (!op_valid(prob->trial) || !op_valid(prob->test)) return ANIFEM_ERR_ARG;
    std::unique_ptr<anifem_assembly> ctx(new anifem_assembly());
    ctx->prob = *prob;
    ctx->prob.dirichlet_dof_mask = nullptr;
    int rc = build_setup(*mesh, *prob, ctx->S);
    if (rc) return rc;
    AssemblySetup& S = ctx->S;
    hipStream_t s = static_cast<hipStream_t>(stream);
    ANIFEM_HIP_CHECK(hipGetDevice(&ctx->device));
    const int64_t NT = S.n_tets;
    std::vector<int32_t> tn(4 * NT), dofs_asm((size_t)S.nf * NT);
    std::vector<uint64_t> dm;
    bool any_dir = false;
    for (uint8_t f : S.is_dir) any_dir = any_dir || f;
    if (any_dir) dm.assign(NT, 0);
    for (int64_t pos = 0; pos < NT; ++pos) {
        const int32_t t = S.perm[pos];
        for (int v = 0; v < 4; ++v) tn[(size_t)v * NT + pos] = S.tets[4 * (size_t)t + v];
        for (int i = 0; i < S.nf; ++i) {
            const int32_t d = S.dofs[(size_t)S.nf * t + i];
            dofs_asm[(size_t)S.nf * pos + i] = d;
            if (any_dir && S.is_dir[d]) dm[pos] |= 1ull << i;
        }
    }
    if ((rc = upload(&ctx->d_coords, mesh->coords, 3 * (size_t)S.n_nodes, s))) return rc;
    if ((rc = upload(&ctx->d_tetnodes, tn.data(), tn.size(), s))) return rc;
    if ((rc = upload(&ctx->d_orig, S.perm.data(), S.perm.size(), s))) return rc;
    if ((rc = upload(&ctx->d_slot, S.slot.data(), S.slot.size(), s))) return rc;
    if ((rc = upload(&ctx->d_dofs, dofs_asm.data(), dofs_asm.size(), s))) return rc;
    if (any_dir && (rc = upload(&ctx->d_dirmask, dm.data(), dm.size(), s))) return rc;
    if ((rc = upload(&ctx->d_row_ptr, S.row_ptr.data(), S.row_ptr.size(), s))) return rc;
    if ((rc = upload(&ctx->d_col_idx, S.col_idx.data(), S.col_idx.size(), s))) return rc;
    ANIFEM_HIP_CHECK(hipMalloc((void**)&ctx->d_status, 4 * sizeof(int)));
    ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
    // the slot map lives on the device now; keep the host copy for inspection
    *out = ctx.release();
    return ANIFEM_OK;
}

int anifem_assembly_create_host(anifem_assembly** out, const anifem_mesh_desc* mesh, const anifem_problem_desc* prob) {
    using namespace anifem;
    if (!out || !mesh || !prob || !mesh->coords || !mesh->tets) return ANIFEM_ERR_ARG;
    if (!op_valid(prob->trial) || !op_valid(prob->test)) return ANIFEM_ERR_ARG;
    std::unique_ptr<anifem_assembly> ctx(new anifem_assembly());
    ctx->prob = *prob;
    ctx->prob.dirichlet_dof_mask = nullptr;
    ctx->host_only = true;
    int rc = build_setup(*mesh, *prob, ctx->S);
    if (rc) return rc;
    *out = ctx.release();
    return ANIFEM_OK;
}

int anifem_assembly_destroy(anifem_assembly* ctx) {
    if (!ctx) return ANIFEM_OK;
    for (hipEvent_t e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (!ctx->host_only) {
        void* bufs[] = {ctx->d_coords, ctx->d_tetnodes, ctx->d_orig, ctx->d_slot, ctx->d_dofs,
                        ctx->d_dirmask, ctx->d_row_ptr, ctx->d_col_idx, ctx->d_status};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
    }
    delete ctx;
    return ANIFEM_OK;
}

int anifem_assembly_info(const anifem_assembly* ctx, int64_t* ndof, int64_t* nnz, int32_t* n_colours, int32_t* nf) {
    if (!ctx) return ANIFEM_ERR_ARG;
    if (ndof) *ndof = ctx->S.ndof;
    if (nnz) *nnz = ctx->S.nnz;
    if (n_colours) *n_colours = ctx->S.n_colours;
    if (nf) *nf = ctx->S.nf;
    return ANIFEM_OK;
}

int anifem_assembly_csr(const anifem_assembly* ctx, const int64_t** row_ptr, const int32_t** col_idx) {
    if (!ctx) return ANIFEM_ERR_ARG;
    if (row_ptr) *row_ptr = ctx->d_row_ptr;
    if (col_idx) *col_idx = ctx->d_col_idx;
    return ANIFEM_OK;
}

int anifem_assembly_copy_setup(const anifem_assembly* ctx, int64_t* row_ptr, int32_t* col_idx, int32_t* dofs,
                               int64_t* slot, int32_t* colour, uint8_t* dirichlet) {
    if (!ctx) return ANIFEM_ERR_ARG;
    const anifem::AssemblySetup& S = ctx->S;
    if (row_ptr) std::memcpy(row_ptr, S.row_ptr.data(), S.row_ptr.size() * sizeof(int64_t));
    if (col_idx) std::memcpy(col_idx, S.col_idx.data(), S.col_idx.size() * sizeof(int32_t));
    if (dofs) std::memcpy(dofs, S.dofs.data(), S.dofs.size() * sizeof(int32_t));
    if (colour) std::memcpy(colour, S.colour.data(), S.colour.size() * sizeof(int32_t));
    if (dirichlet) std::memcpy(dirichlet, S.is_dir.data(), S.is_dir.size());
    if (slot) {
        // back to the caller's tet order
        for (int64_t pos = 0; pos < S.n_tets; ++pos) {
            const int32_t t = S.perm[pos];
            for (int e = 0; e < S.E; ++e) slot[(size_t)S.E * t + e] = S.slot[(size_t)S.E * pos + e];
        }
    }
    return ANIFEM_OK;
}

int anifem_assembly_enable_timers(anifem_assembly* ctx, int on) {
    if (!ctx || ctx->host_only) return ANIFEM_ERR_ARG;
    if (on && !ctx->ev[0])
        for (auto& e : ctx->ev) ANIFEM_HIP_CHECK(hipEventCreate(&e));
    ctx->timers_on = on != 0;
    return ANIFEM_OK;
}

int anifem_assembly_timers(const anifem_assembly* ctx, anifem_timers* out) {
    if (!ctx || !out) return ANIFEM_ERR_ARG;
    *out = ctx->timers;
    return ANIFEM_OK;
}

int anifem_assembly_coords(anifem_assembly* ctx, double** coords) {
    if (!ctx || ctx->host_only || !coords) return ANIFEM_ERR_ARG;
    *coords = ctx->d_coords;
    return ANIFEM_OK;
}

int anifem_assembly_run(anifem_assembly* ctx, const anifem_tensor* D, const anifem_tensor* F,
                        const double* dir_values, int precision, double* csr_val, double* rhs,
                        anifem_status* status, void* stream) {
    using namespace anifem;
    if (!ctx || ctx->host_only || !D || !csr_val) return ANIFEM_ERR_ARG;
    if (dir_values) return ANIFEM_ERR_ARG;  // inhomogeneous Dirichlet values: not supported yet
    if (precision != ANIFEM_PREC_FP64 && precision != ANIFEM_PREC_FP32) return ANIFEM_ERR_ARG;
    if (D->type != ctx->prob.tensor_type || D->storage != ctx->prob.tensor_storage) return ANIFEM_ERR_ARG;
    const AssemblySetup& S = ctx->S;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;
    if (ctx->timers_on) ANIFEM_HIP_CHECK(hipEventRecord(ctx->ev[0], s));
    if ((rc = fill(csr_val, S.nnz, s))) return rc;
    if (rhs && (rc = fill(rhs, S.ndof, s))) return rc;
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, ctx->d_status, 4);
    ANIFEM_HIP_CHECK(hipGetLastError());

    ElemArgs a{};
    if ((rc = prepare_elem_args(ctx->prob.trial, ctx->prob.test, D, ctx->prob.order, a))) return rc;
    a.coef_n = S.n_tets;
    a.coords = ctx->d_coords;
    a.tetnodes = ctx->d_tetnodes;
    a.nodes_stride = S.n_tets;
    a.orig = ctx->d_orig;
    a.slot = ctx->d_slot;
    a.dirmask = ctx->d_dirmask;
    a.val = csr_val;
    a.status = ctx->d_status;

    ElemArgs af{};
    anifem_op p0{ANIFEM_IDEN, ANIFEM_FEM_P0, 1};
    anifem_op test_iden{ANIFEM_IDEN, ctx->prob.test.fem, ctx->prob.test.ncomp};
    const bool do_rhs = F && rhs;
    if (do_rhs) {
        if ((rc = prepare_elem_args(p0, test_iden, F, ctx->prob.rhs_order, af))) return rc;
        af.coef_n = S.n_tets;
        af.coords = ctx->d_coords;
        af.tetnodes = ctx->d_tetnodes;
        af.nodes_stride = S.n_tets;
        af.orig = ctx->d_orig;
        af.slot = ctx->d_dofs;
        af.dirmask = ctx->d_dirmask;
        af.val = rhs;
        af.status = ctx->d_status;
    }
    if (ctx->timers_on) ANIFEM_HIP_CHECK(hipEventRecord(ctx->ev[1], s));
    for (int c = 0; c < S.n_colours; ++c) {
        a.pos0 = S.colour_ptr[c];
        a.n = S.colour_ptr[c + 1] - S.colour_ptr[c];
        rc = precision == ANIFEM_PREC_FP64
                 ? launch_element<double, true, 1>(ctx->prob.trial, ctx->prob.test, a, 64, s)
                 : launch_element<float, true, 1>(ctx->prob.trial, ctx->prob.test, a, 64, s);
        if (rc) return rc;
    }
    if (ctx->timers_on) ANIFEM_HIP_CHECK(hipEventRecord(ctx->ev[2], s));
    for (int c = 0; do_rhs && c < S.n_colours; ++c) {
        af.pos0 = S.colour_ptr[c];
        af.n = S.colour_ptr[c + 1] - S.colour_ptr[c];
        rc = precision == ANIFEM_PREC_FP64 ? launch_element<double, true, 2>(p0, test_iden, af, 64, s)
                                           : launch_element<float, true, 2>(p0, test_iden, af, 64, s);
        if (rc) return rc;
    }
    if (ctx->timers_on) {
        ANIFEM_HIP_CHECK(hipEventRecord(ctx->ev[3], s));
        ANIFEM_HIP_CHECK(hipEventSynchronize(ctx->ev[3]));
        float t01 = 0, t12 = 0, t23 = 0;
        ANIFEM_HIP_CHECK(hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]));
        ANIFEM_HIP_CHECK(hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]));
        ANIFEM_HIP_CHECK(hipEventElapsedTime(&t23, ctx->ev[2], ctx->ev[3]));
        ctx->timers.ms_fill = t01;
        ctx->timers.ms_matrix = t12;
        ctx->timers.ms_rhs = t23;
        ctx->timers.n_matrix_launches = S.n_colours;
        ctx->timers.n_rhs_launches = do_rhs ? S.n_colours : 0;
    }
    if (status) {
        int h[4] = {0, 0, 0, 0};
        ANIFEM_HIP_CHECK(hipMemcpyAsync(h, ctx->d_status, sizeof(h), hipMemcpyDeviceToHost, s));
        ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
        status->n_nonfinite = h[0];
        status->n_nonpositive = h[1];
        status->code = h[0] ? ANIFEM_ERR_NONFINITE : ANIFEM_OK;
        status->reserved = 0;
        return status->code;
    }
    return ANIFEM_OK;
}

}  // extern "C"
