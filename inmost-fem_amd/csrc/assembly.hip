// Whole-mesh assembly on the GPU: anifem_assembly_* (C ABI).
//
// create ~ AssemblerT::PrepareProblem (assembler.inl:194-275): the host setup
// (setup.cpp: enumeration, pattern, slot map; blocks.cpp: row blocks) is
// uploaded once.
// run ~ AssemblerT::Assemble (assembler.inl:314-488): a geometry launch (one
// lane per tet) and one row-block launch (assembly_kernels.h) that forms every
// CSR value and rhs entry exactly once, summing the element contributions in
// ascending tet order.  No atomics on values, so repeated runs are bitwise
// identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>

#include "anifem_common.h"
#include "assembly_kernels.h"
#include "setup.h"

namespace anifem {
int prepare_elem_args(const anifem_op& opA, const anifem_op& opB, const anifem_tensor* D, int order, ElemArgs& a);
}  // namespace anifem

struct anifem_assembly {
    anifem::AssemblySetup S;
    anifem_problem_desc prob{};
    int device = 0;
    double* d_coords = nullptr;       // n_nodes x 3
    int32_t* d_tetnodes = nullptr;    // 4 x n_tets SoA
    int32_t* d_blk_row0 = nullptr;
    int64_t* d_blk_s0 = nullptr;
    int64_t* d_blk_tet0 = nullptr;
    int32_t* d_blk_tets = nullptr;
    int64_t* d_blk_c0 = nullptr;
    void* d_codes = nullptr;          // uint16 or uint32 (S.code16)
    uint16_t* d_sstart = nullptr;
    int64_t* d_blk_g0 = nullptr;
    uint16_t* d_fcodes = nullptr;
    uint16_t* d_rstart = nullptr;
    int64_t* d_row_ptr = nullptr;
    int32_t* d_col_idx = nullptr;
    int* d_status = nullptr;          // [0] non-finite, [1] degenerate
    void* d_kbuf = nullptr;           // per-tet geometry tensors of the last run
    size_t kbuf_bytes = 0;
    bool host_only = false;           // built by anifem_assembly_create_host
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

template <int DA0, int DB0, int NCA, int NCB, typename T>
int launch_blocks(anifem_assembly* ctx, anifem::BlockArgs a, hipStream_t s) {
    using namespace anifem;
    const AssemblySetup& S = ctx->S;
    const int kstride = NCA * NCB * a.m.q * DA0 * DB0;
    const int kstrideF = a.has_rhs ? NCB * a.f.q : 0;
    // row_gather indexes a block's K values with 16-bit arithmetic
    if ((int64_t)a.max_tets * std::max(kstride, 1) >= 65536 || (int64_t)a.max_tets * std::max(kstrideF, 1) >= 65536)
        return ANIFEM_ERR_DIMS;
    int max_lds = 0;
    ANIFEM_HIP_CHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
    const int lenp = pad16<T>(a.m.q * DA0 * DB0);
    const int lenFp = a.has_rhs ? pad16<T>(a.f.q) : 0;
    auto lds_of = [&](bool with_R, bool with_RF) {
        return block_lds_layout(a.max_tets, NCA * NCB * lenp, NCB * lenFp, (int)sizeof(T),
                                with_R ? a.R_rows * lenp : 0, with_RF ? a.RF_rows * lenFp : 0, a.m.E, a.nf,
                                a.has_rhs, a.max_contrib, a.max_fcontrib, a.max_slots)
            .total;
    };
    // the reference tensors are staged while four workgroups still fit a CU
    const size_t soft = (size_t)max_lds / 4;
    a.stage_R = lds_of(true, false) <= soft;
    a.stage_RF = a.has_rhs && lds_of(a.stage_R, true) <= soft;
    const size_t lds = lds_of(a.stage_R, a.stage_RF);
    if (lds > (size_t)max_lds) return ANIFEM_ERR_DIMS;
    // geometry tensors of every tet (kept on the device between runs)
    const size_t need = (size_t)S.n_tets * (kstride + kstrideF) * sizeof(T);
    if (ctx->kbuf_bytes < need) {
        if (ctx->d_kbuf) ANIFEM_HIP_CHECK(hipFree(ctx->d_kbuf));
        ctx->d_kbuf = nullptr;
        ctx->kbuf_bytes = 0;
        ANIFEM_HIP_CHECK(hipMalloc(&ctx->d_kbuf, need));
        ctx->kbuf_bytes = need;
    }
    T* Kg = static_cast<T*>(ctx->d_kbuf);
    T* KFg = Kg + (size_t)S.n_tets * kstride;
    const int64_t gblocks = (S.n_tets + kBlock - 1) / kBlock;
    hipLaunchKernelGGL((geometry_kernel<DA0, DB0, NCA, NCB, T>), dim3((unsigned)gblocks), dim3(kBlock), 0, s, a, Kg,
                       KFg, S.n_tets);
    ANIFEM_HIP_CHECK(hipGetLastError());
    if (ctx->timers_on) ANIFEM_HIP_CHECK(hipEventRecord(ctx->ev[1], s));
    if (S.n_blocks > 0) {
        if (S.code16)
            hipLaunchKernelGGL((row_block_kernel<DA0, DB0, NCA, NCB, T, uint16_t>), dim3((unsigned)S.n_blocks),
                               dim3(kBlock), lds, s, a, (const T*)Kg, (const T*)KFg);
        else
            hipLaunchKernelGGL((row_block_kernel<DA0, DB0, NCA, NCB, T, uint32_t>), dim3((unsigned)S.n_blocks),
                               dim3(kBlock), lds, s, a, (const T*)Kg, (const T*)KFg);
        ANIFEM_HIP_CHECK(hipGetLastError());
    }
    return ANIFEM_OK;
}

template <typename T>
int dispatch_blocks(anifem_assembly* ctx, const anifem::BlockArgs& a, hipStream_t s) {
    const int dA = anifem::ref_dim(ctx->prob.trial.op), dB = anifem::ref_dim(ctx->prob.test.op);
    const int nc = ctx->prob.trial.ncomp;
    if (dA == 3 && dB == 3 && nc == 1) return launch_blocks<3, 3, 1, 1, T>(ctx, a, s);
    if (dA == 3 && dB == 3 && nc == 3) return launch_blocks<3, 3, 3, 3, T>(ctx, a, s);
    if (dA == 1 && dB == 1 && nc == 1) return launch_blocks<1, 1, 1, 1, T>(ctx, a, s);
    if (dA == 1 && dB == 1 && nc == 3) return launch_blocks<1, 1, 3, 3, T>(ctx, a, s);
    if (dA == 3 && dB == 1 && nc == 1) return launch_blocks<3, 1, 1, 1, T>(ctx, a, s);
    if (dA == 1 && dB == 3 && nc == 1) return launch_blocks<1, 3, 1, 1, T>(ctx, a, s);
    return ANIFEM_ERR_ARG;
}

int build_host(anifem_assembly* ctx, const anifem_mesh_desc* mesh, const anifem_problem_desc* prob) {
    using namespace anifem;
    if (!mesh || !prob || !mesh->coords || !mesh->tets) return ANIFEM_ERR_ARG;
    if (!op_valid(prob->trial) || !op_valid(prob->test)) return ANIFEM_ERR_ARG;
    ctx->prob = *prob;
    ctx->prob.dirichlet_dof_mask = nullptr;
    return build_setup(*mesh, *prob, ctx->S);
}

}  // namespace

extern "C" {

int anifem_assembly_create(anifem_assembly** out, const anifem_mesh_desc* mesh, const anifem_problem_desc* prob,
                           void* stream) {
    using namespace anifem;
    if (!out) return ANIFEM_ERR_ARG;
    std::unique_ptr<anifem_assembly> ctx(new anifem_assembly());
    int rc = build_host(ctx.get(), mesh, prob);
    if (rc) return rc;
    AssemblySetup& S = ctx->S;
    hipStream_t s = static_cast<hipStream_t>(stream);
    ANIFEM_HIP_CHECK(hipGetDevice(&ctx->device));
    const int64_t NT = S.n_tets;
    std::vector<int32_t> tn(4 * NT);
    for (int64_t t = 0; t < NT; ++t)
        for (int v = 0; v < 4; ++v) tn[(size_t)v * NT + t] = S.tets[4 * (size_t)t + v];
    if ((rc = upload(&ctx->d_coords, mesh->coords, 3 * (size_t)S.n_nodes, s))) return rc;
    if ((rc = upload(&ctx->d_tetnodes, tn.data(), tn.size(), s))) return rc;
    if ((rc = upload(&ctx->d_blk_row0, S.blk_row0.data(), S.blk_row0.size(), s))) return rc;
    if ((rc = upload(&ctx->d_blk_s0, S.blk_s0.data(), S.blk_s0.size(), s))) return rc;
    if ((rc = upload(&ctx->d_blk_tet0, S.blk_tet0.data(), S.blk_tet0.size(), s))) return rc;
    if ((rc = upload(&ctx->d_blk_tets, S.blk_tets.data(), S.blk_tets.size(), s))) return rc;
    if ((rc = upload(&ctx->d_blk_c0, S.blk_c0.data(), S.blk_c0.size(), s))) return rc;
    std::vector<uint16_t> codes16;
    if (S.code16) {
        codes16.resize(S.codes.size());
        for (size_t i = 0; i < S.codes.size(); ++i) codes16[i] = (uint16_t)S.codes[i];
        uint16_t* d16 = nullptr;
        rc = upload(&d16, codes16.data(), codes16.size(), s);
        ctx->d_codes = d16;
    } else {
        uint32_t* d32 = nullptr;
        rc = upload(&d32, S.codes.data(), S.codes.size(), s);
        ctx->d_codes = d32;
    }
    if (rc) return rc;
    if ((rc = upload(&ctx->d_sstart, S.sstart.data(), S.sstart.size(), s))) return rc;
    if ((rc = upload(&ctx->d_blk_g0, S.blk_g0.data(), S.blk_g0.size(), s))) return rc;
    if ((rc = upload(&ctx->d_fcodes, S.fcodes.data(), S.fcodes.size(), s))) return rc;
    if ((rc = upload(&ctx->d_rstart, S.rstart.data(), S.rstart.size(), s))) return rc;
    if ((rc = upload(&ctx->d_row_ptr, S.row_ptr.data(), S.row_ptr.size(), s))) return rc;
    if ((rc = upload(&ctx->d_col_idx, S.col_idx.data(), S.col_idx.size(), s))) return rc;
    ANIFEM_HIP_CHECK(hipMalloc((void**)&ctx->d_status, 4 * sizeof(int)));
    ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
    // the device holds the contribution lists; the host keeps the rest for inspection
    std::vector<uint32_t>().swap(S.codes);
    std::vector<uint16_t>().swap(S.sstart);
    std::vector<uint16_t>().swap(S.fcodes);
    std::vector<uint16_t>().swap(S.rstart);
    *out = ctx.release();
    return ANIFEM_OK;
}

int anifem_assembly_create_host(anifem_assembly** out, const anifem_mesh_desc* mesh, const anifem_problem_desc* prob) {
    if (!out) return ANIFEM_ERR_ARG;
    std::unique_ptr<anifem_assembly> ctx(new anifem_assembly());
    ctx->host_only = true;
    int rc = build_host(ctx.get(), mesh, prob);
    if (rc) return rc;
    *out = ctx.release();
    return ANIFEM_OK;
}

int anifem_assembly_destroy(anifem_assembly* ctx) {
    if (!ctx) return ANIFEM_OK;
    for (hipEvent_t e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (!ctx->host_only) {
        void* bufs[] = {ctx->d_coords,  ctx->d_tetnodes, ctx->d_blk_row0, ctx->d_blk_s0, ctx->d_blk_tet0,
                        ctx->d_blk_tets, ctx->d_blk_c0,   ctx->d_codes,   ctx->d_sstart,   ctx->d_blk_g0,
                        ctx->d_fcodes,  ctx->d_rstart,   ctx->d_row_ptr, ctx->d_col_idx,  ctx->d_status,
                        ctx->d_kbuf};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
    }
    delete ctx;
    return ANIFEM_OK;
}

int anifem_assembly_info(const anifem_assembly* ctx, int64_t* ndof, int64_t* nnz, int64_t* n_blocks, int32_t* nf) {
    if (!ctx) return ANIFEM_ERR_ARG;
    if (ndof) *ndof = ctx->S.ndof;
    if (nnz) *nnz = ctx->S.nnz;
    if (n_blocks) *n_blocks = ctx->S.n_blocks;
    if (nf) *nf = ctx->S.nf;
    return ANIFEM_OK;
}

int anifem_assembly_csr(const anifem_assembly* ctx, const int64_t** row_ptr, const int32_t** col_idx) {
    if (!ctx || ctx->host_only) return ANIFEM_ERR_ARG;
    if (row_ptr) *row_ptr = ctx->d_row_ptr;
    if (col_idx) *col_idx = ctx->d_col_idx;
    return ANIFEM_OK;
}

int anifem_assembly_copy_setup(const anifem_assembly* ctx, int64_t* row_ptr, int32_t* col_idx, int32_t* dofs,
                               int64_t* slot, uint8_t* dirichlet) {
    if (!ctx) return ANIFEM_ERR_ARG;
    const anifem::AssemblySetup& S = ctx->S;
    if (row_ptr) std::memcpy(row_ptr, S.row_ptr.data(), S.row_ptr.size() * sizeof(int64_t));
    if (col_idx) std::memcpy(col_idx, S.col_idx.data(), S.col_idx.size() * sizeof(int32_t));
    if (dofs) std::memcpy(dofs, S.dofs.data(), S.dofs.size() * sizeof(int32_t));
    if (dirichlet) std::memcpy(dirichlet, S.is_dir.data(), S.is_dir.size());
    if (slot)
        for (size_t i = 0; i < S.slot.size(); ++i) slot[i] = S.slot[i];
    return ANIFEM_OK;
}

int anifem_assembly_blocks(const anifem_assembly* ctx, anifem_block_layout* out) {
    if (!ctx || !out) return ANIFEM_ERR_ARG;
    const anifem::AssemblySetup& S = ctx->S;
    out->n_blocks = S.n_blocks;
    out->n_block_tets = (int64_t)S.blk_tets.size();
    out->n_codes = S.blk_c0.empty() ? 0 : S.blk_c0.back();
    out->n_fcodes = S.blk_g0.empty() ? 0 : S.blk_g0.back();
    out->code_bytes = S.code16 ? 2 : 4;
    out->block_tet_cap = S.block_tet_cap;
    out->e_bits = S.ebits;
    out->f_bits = S.fbits;
    return ANIFEM_OK;
}

int anifem_assembly_copy_blocks(const anifem_assembly* ctx, int32_t* blk_row0, int64_t* blk_tet0, int32_t* blk_tets,
                                int64_t* blk_c0, uint32_t* codes, uint16_t* sstart, int64_t* blk_g0,
                                uint16_t* fcodes, uint16_t* rstart) {
    if (!ctx) return ANIFEM_ERR_ARG;
    const anifem::AssemblySetup& S = ctx->S;
    // contexts built for the device keep only the block offsets on the host
    if (!ctx->host_only && (codes || sstart || fcodes || rstart)) return ANIFEM_ERR_ARG;
    auto cp = [](auto* dst, const auto& v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(blk_row0, S.blk_row0);
    cp(blk_tet0, S.blk_tet0);
    cp(blk_tets, S.blk_tets);
    cp(blk_c0, S.blk_c0);
    cp(codes, S.codes);
    cp(sstart, S.sstart);
    cp(blk_g0, S.blk_g0);
    cp(fcodes, S.fcodes);
    cp(rstart, S.rstart);
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
    if (dir_values) return ANIFEM_ERR_ARG;  // inhomogeneous Dirichlet values: not supported
    if (precision != ANIFEM_PREC_FP64 && precision != ANIFEM_PREC_FP32) return ANIFEM_ERR_ARG;
    if (D->type != ctx->prob.tensor_type || D->storage != ctx->prob.tensor_storage) return ANIFEM_ERR_ARG;
    const bool do_rhs = F && rhs;
    if (do_rhs && (F->type != ctx->prob.rhs_tensor_type || F->storage != ctx->prob.rhs_tensor_storage))
        return ANIFEM_ERR_ARG;
    const AssemblySetup& S = ctx->S;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;

    BlockArgs a{};
    if ((rc = prepare_elem_args(ctx->prob.trial, ctx->prob.test, D, ctx->prob.order, a.m))) return rc;
    a.m.coef_n = S.n_tets;
    a.m.coords = ctx->d_coords;
    a.m.tetnodes = ctx->d_tetnodes;
    a.m.nodes_stride = S.n_tets;
    a.m.pos0 = 0;
    a.R_rows = a.m.nfA0 * a.m.nfB0;
    if (do_rhs) {
        anifem_op p0{ANIFEM_IDEN, ANIFEM_FEM_P0, 1};
        anifem_op test_iden{ANIFEM_IDEN, ctx->prob.test.fem, ctx->prob.test.ncomp};
        if ((rc = prepare_elem_args(p0, test_iden, F, ctx->prob.rhs_order, a.f))) return rc;
        a.f.coef_n = S.n_tets;
        a.RF_rows = a.f.nfB0;
    }
    a.has_rhs = do_rhs ? 1 : 0;
    a.nf = S.nf;
    a.blk_row0 = ctx->d_blk_row0;
    a.blk_s0 = ctx->d_blk_s0;
    a.blk_tet0 = ctx->d_blk_tet0;
    a.blk_tets = ctx->d_blk_tets;
    a.blk_c0 = ctx->d_blk_c0;
    a.codes = ctx->d_codes;
    a.sstart = ctx->d_sstart;
    a.blk_g0 = ctx->d_blk_g0;
    a.fcodes = ctx->d_fcodes;
    a.rstart = ctx->d_rstart;
    a.ebits = S.ebits;
    a.fbits = S.fbits;
    a.val = csr_val;
    a.rhs = rhs;
    a.status = ctx->d_status;
    a.max_tets = (int)std::max<int64_t>(1, S.max_block_tets);
    a.max_contrib = (int)S.max_block_contrib;
    a.max_fcontrib = do_rhs ? (int)S.max_block_fcontrib : 0;
    a.max_slots = (int)S.max_block_slots;
    if (S.max_block_contrib > (int64_t)kBlockCodesPerLane * kBlock) return ANIFEM_ERR_DIMS;

    if (ctx->timers_on) ANIFEM_HIP_CHECK(hipEventRecord(ctx->ev[0], s));
    // a rhs without a load vector is zero
    if (rhs && !do_rhs && (rc = fill(rhs, S.ndof, s))) return rc;
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, ctx->d_status, 4);
    ANIFEM_HIP_CHECK(hipGetLastError());
    rc = precision == ANIFEM_PREC_FP64 ? dispatch_blocks<double>(ctx, a, s) : dispatch_blocks<float>(ctx, a, s);
    if (rc) return rc;
    if (ctx->timers_on) {
        ANIFEM_HIP_CHECK(hipEventRecord(ctx->ev[2], s));
        ANIFEM_HIP_CHECK(hipEventSynchronize(ctx->ev[2]));
        float t01 = 0, t12 = 0;
        ANIFEM_HIP_CHECK(hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]));
        ANIFEM_HIP_CHECK(hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]));
        ctx->timers.ms_geometry = t01;
        ctx->timers.ms_gather = t12;
        ctx->timers.n_launches = 2 + (rhs && !do_rhs ? 1 : 0) + (S.n_blocks > 0 ? 1 : 0);
        ctx->timers.reserved = 0;
    }
    if (status) {
        int h[4] = {0, 0, 0, 0};
        ANIFEM_HIP_CHECK(hipMemcpyAsync(h, ctx->d_status, sizeof(h), hipMemcpyDeviceToHost, s));
        ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
        status->n_nonfinite = h[0];
        status->n_degenerate = h[1];
        status->code = h[0] ? ANIFEM_ERR_NONFINITE : ANIFEM_OK;
        status->reserved = 0;
        return status->code;
    }
    return ANIFEM_OK;
}

}  // extern "C"
