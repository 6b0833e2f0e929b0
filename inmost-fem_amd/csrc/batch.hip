// Batched assembly of many meshes of uneven size and mixed element order:
// anifem_batch_* (C ABI).
//
// Every mesh is a scalar diffusion problem -div(kappa grad u) = f with P1, P2
// or P3 elements (fem3Dtet<GRAD FemFix<Pk>, GRAD FemFix<Pk>> with a per-tet
// SCALAR kappa, load vector fem3Dtet<IDEN P0, IDEN Pk> with a constant f,
// quadrature order 2 for P1/P2 and 4 for P3; int_tet.h:17-20) and homogeneous
// Dirichlet on its boundary.  The batch is assembled as one block-diagonal
// system: the host setup of each mesh (setup.cpp, blocks.cpp) is concatenated
// with offsets, so a single geometry launch covers all tets.  The row blocks
// are grouped by element order, and each order's blocks run in one persistent
// row-block launch of that order's instantiation (the fixed-size form for P1
// and P2, compact value positions where every mesh of the order has them),
// each block selecting the reference tensor of its order (KindTable).  Each
// mesh's CSR values and rhs land in their own slices of the output arrays and
// are bitwise identical to assembling that mesh alone.  The launch sequence
// is captured once in a HIP graph and replayed.
//
// Autotuning, per element order: the row-block budgets (slots / entries /
// pairs per block) by timing candidate budgets on the GPU on a representative
// mesh of that order; then, at the first run, the launch shape of the order's
// row-block pass on the batch's own blocks: the phase-A form (fixed-size or
// general), the value positions (compact or 16-bit) and the workgroups per CU
// of the persistent grid (the resident waves per SIMD).  Every candidate sums
// in the same order, so the choice changes timings only, never values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <atomic>
#include <memory>
#include <thread>
#include <vector>

#include "anifem_common.h"
#include "assembly_kernels.h"
#include "setup.h"

namespace anifem {
int prepare_elem_args(const anifem_op& opA, const anifem_op& opB, const anifem_tensor* D, int order, ElemArgs& a);
int assembly_create_caps(anifem_assembly** out, const anifem_mesh_desc* mesh, const anifem_problem_desc* prob,
                         const BlockCaps& caps, void* stream, const anifem_vars_desc* vars = nullptr);
}  // namespace anifem

namespace {

constexpr int kKinds = 3;  // P1, P2, P3
// Candidate budgets {slots, entries, pairs}: the last two let three P3
// workgroups (20 entries per pair) share a CU's LDS, with fuller blocks of
// two workgroups per CU as the alternative.
constexpr int kCandidates = ANIFEM_BATCH_BUDGET_CANDIDATES;
const anifem::BlockCaps kCandidateCaps[kCandidates] = {{2048, 1024, 128}, {2048, 2048, 256}, {4096, 4096, 256},
                                                       {1024, 1024, 256}, {4096, 3072, 256}, {4096, 2560, 256},
                                                       {4096, 2560, 128}, {4096, 2048, 128}};

int kind_of(int fem) { return fem - ANIFEM_FEM_P1; }
int order_of(int fem) { return fem == ANIFEM_FEM_P3 ? 4 : 2; }

anifem_problem_desc problem_of(int fem, int dirichlet) {
    anifem_problem_desc p{};
    p.trial = anifem_op{ANIFEM_GRAD, fem, 1};
    p.test = p.trial;
    p.order = order_of(fem);
    p.tensor_type = ANIFEM_TENSOR_SCALAR;
    p.tensor_storage = ANIFEM_COEF_PER_TET;
    p.rhs_order = order_of(fem);
    p.rhs_tensor_type = ANIFEM_TENSOR_SCALAR;
    p.rhs_tensor_storage = ANIFEM_COEF_CONST;
    p.dirichlet = dirichlet;
    return p;
}

__global__ void batch_status_clear(int* p) {
    if (threadIdx.x < 4) p[threadIdx.x] = 0;
}

}  // namespace

struct anifem_batch {
    int n_meshes = 0;
    int dirichlet = 0;
    int device = 0;
    std::vector<int32_t> fem;
    std::vector<int64_t> tet_off, node_off, dof_off, nnz_off, rp_off;
    int64_t n_tets = 0, n_nodes = 0, ndof = 0, nnz = 0, n_blocks = 0;
    int max_values = 0, max_pairs = 0, max_slots = 0;
    // blocks (metadata order) and pairs of each element order: contiguous
    int64_t kind_blk[kKinds + 1] = {0, 0, 0, 0};
    int64_t kind_pair0[kKinds] = {0, 0, 0};
    int kmax_values[kKinds] = {0, 0, 0}, kmax_pairs[kKinds] = {0, 0, 0}, kmax_slots[kKinds] = {0, 0, 0};
    uint32_t* d_pos8[kKinds] = {nullptr, nullptr, nullptr};  // compact positions of the order's pairs, or null
    // launch shape per order: phase-A form (0 general, 1 fixed-size, 2 two
    // lanes per pair), compact positions, workgroups per CU (0: occupancy) and
    // its time per launch (ms; -1 until timed)
    bool tune_launch = false, launch_tuned = false;
    int shape[kKinds][3] = {{1, 1, 0}, {1, 1, 0}, {0, 1, 0}};
    double shape_ms[kKinds] = {-1, -1, -1};
    anifem::BlockCaps caps[kKinds];
    double tune_ms[kKinds][kCandidates] = {};
    bool kind_used[kKinds] = {false, false, false};
    bool persistent = true;  // persistent row-block pass (else one workgroup per block)
    // device
    double* d_coords = nullptr;
    int32_t* d_tetnodes = nullptr;
    anifem::BlockMeta* d_meta = nullptr;
    int32_t* d_pair_tet = nullptr;
    uint8_t* d_pair_ib = nullptr;
    uint16_t* d_pair_pos = nullptr;
    uint16_t* d_sstart = nullptr;
    uint16_t* d_pstart = nullptr;
    int64_t* d_row_ptr = nullptr;  // each mesh's own row_ptr (ndof+1 entries), concatenated
    int32_t* d_col_idx = nullptr;
    int* d_status = nullptr;
    double* d_K = nullptr;
    // captured launch sequence
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* key[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<void*> bufs;
};

namespace {

template <typename T>
int upload_b(anifem_batch* b, T** dst, const std::vector<T>& v, hipStream_t s) {
    ANIFEM_HIP_CHECK(hipMalloc((void**)dst, std::max<size_t>(v.size(), 1) * sizeof(T)));
    b->bufs.push_back(*dst);
    if (!v.empty()) ANIFEM_HIP_CHECK(hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return ANIFEM_OK;
}

// Times the row-block pass of one mesh under each candidate budget.
int autotune_kind(anifem_batch* b, int kind, const anifem_batch_mesh& m, hipStream_t s) {
    const int fem = kind + ANIFEM_FEM_P1;
    anifem_mesh_desc md{m.n_nodes, m.n_tets, m.coords, m.tets};
    const anifem_problem_desc prob = problem_of(fem, b->dirichlet);
    double* kap = nullptr;
    double* f = nullptr;
    ANIFEM_HIP_CHECK(hipMalloc((void**)&kap, m.n_tets * sizeof(double)));
    ANIFEM_HIP_CHECK(hipMalloc((void**)&f, sizeof(double)));
    std::vector<double> ones(m.n_tets, 1.0);
    ANIFEM_HIP_CHECK(hipMemcpy(kap, ones.data(), m.n_tets * sizeof(double), hipMemcpyHostToDevice));
    ANIFEM_HIP_CHECK(hipMemcpy(f, ones.data(), sizeof(double), hipMemcpyHostToDevice));
    double best = 1e300;
    int best_i = 1;
    int rc = ANIFEM_OK;
    for (int c = 0; c < kCandidates && !rc; ++c) {
        anifem_assembly* ctx = nullptr;
        rc = anifem::assembly_create_caps(&ctx, &md, &prob, kCandidateCaps[c], s);
        if (rc == ANIFEM_ERR_DIMS) {  // a row does not fit this budget
            rc = ANIFEM_OK;
            b->tune_ms[kind][c] = -1;
            continue;
        }
        if (rc) break;
        int64_t ndof = 0, nnz = 0;
        anifem_assembly_info(ctx, &ndof, &nnz, nullptr, nullptr);
        double *val = nullptr, *rhs = nullptr;
        if (hipMalloc((void**)&val, nnz * sizeof(double)) != hipSuccess ||
            hipMalloc((void**)&rhs, ndof * sizeof(double)) != hipSuccess) {
            rc = ANIFEM_ERR_NOMEM;
        }
        const anifem_tensor D{ANIFEM_TENSOR_SCALAR, ANIFEM_COEF_PER_TET, kap};
        const anifem_tensor F{ANIFEM_TENSOR_SCALAR, ANIFEM_COEF_CONST, f};
        double ms = 0;
        if (!rc) rc = anifem_assembly_run(ctx, &D, &F, nullptr, ANIFEM_PREC_FP64, val, rhs, nullptr, s);
        if (!rc) rc = anifem_assembly_enable_timers(ctx, 1);
        const int reps = 3;
        for (int r = 0; r < reps && !rc; ++r) {
            rc = anifem_assembly_run(ctx, &D, &F, nullptr, ANIFEM_PREC_FP64, val, rhs, nullptr, s);
            anifem_timers t{};
            anifem_assembly_timers(ctx, &t);
            ms += t.ms_gather / reps;
        }
        if (val) (void)hipFree(val);
        if (rhs) (void)hipFree(rhs);
        anifem_assembly_destroy(ctx);
        b->tune_ms[kind][c] = ms;
        if (!rc && ms < best) {
            best = ms;
            best_i = c;
        }
    }
    (void)hipFree(kap);
    (void)hipFree(f);
    b->caps[kind] = kCandidateCaps[best_i];
    return rc;
}

// The persistent row-block instantiation of an element order and phase-A form
// (0 general, 1 fixed-size for P1 (4 DOFs) and P2 (10), 2 general with two
// lanes per pair for P3 blocks of at most kBlock/2 pairs, with the triangle of
// the reference tensor in LDS when it is symmetric); compact or 16-bit value
// positions (16-bit only in form 2).
const void* kind_kernel(int kind, bool p8, int form, bool tri = false) {
    using namespace anifem;
    if (kind == 2 && form == 2)
        return tri ? reinterpret_cast<const void*>(&row_block_persistent<6, 1, 10, double, false, 0, true, true>)
                   : reinterpret_cast<const void*>(&row_block_persistent<6, 1, 10, double, false, 0, true>);
    const int v = kind * 4 + (form == 1 && kind < 2 ? 2 : 0) + (p8 ? 1 : 0);
    switch (v) {
        case 0: return reinterpret_cast<const void*>(&row_block_persistent<6, 1, 2, double, false>);
        case 1: return reinterpret_cast<const void*>(&row_block_persistent<6, 1, 2, double, true>);
        case 2: return reinterpret_cast<const void*>(&row_block_persistent<6, 1, 2, double, false, 4>);
        case 3: return reinterpret_cast<const void*>(&row_block_persistent<6, 1, 2, double, true, 4>);
        case 4: return reinterpret_cast<const void*>(&row_block_persistent<6, 1, 5, double, false>);
        case 5: return reinterpret_cast<const void*>(&row_block_persistent<6, 1, 5, double, true>);
        case 6: return reinterpret_cast<const void*>(&row_block_persistent<6, 1, 5, double, false, 10>);
        case 7: return reinterpret_cast<const void*>(&row_block_persistent<6, 1, 5, double, true, 10>);
        case 8: return reinterpret_cast<const void*>(&row_block_persistent<6, 1, 10, double, false>);
        default: return reinterpret_cast<const void*>(&row_block_persistent<6, 1, 10, double, true>);
    }
}

// Persistent grid of a row-block launch: wpc workgroups per CU (at most the
// occupancy; 0: the occupancy), whole XCD groups, at most one per block.
int persistent_grid(const void* fn, size_t lds, int wpc, int64_t n_blocks, int device, int64_t* grid, int* occ_out) {
    int occ = 0, n_cu = 0;
    ANIFEM_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, anifem::kBlock, lds));
    ANIFEM_HIP_CHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device));
    occ = std::max(occ, 1);
    const int w = wpc > 0 ? std::min(wpc, occ) : occ;
    int64_t G = std::min<int64_t>((int64_t)w * std::max(n_cu, 1), n_blocks);
    *grid = std::max<int64_t>(8, G / 8 * 8);
    if (occ_out) *occ_out = occ;
    return ANIFEM_OK;
}

int capture(anifem_batch* b, const double* kappa, const double* f, double* val, double* rhs, hipStream_t rs) {
    using namespace anifem;
    BlockArgs a{};
    int rc;
    const anifem_tensor D{ANIFEM_TENSOR_SCALAR, ANIFEM_COEF_PER_TET, kappa};
    const anifem_tensor F{ANIFEM_TENSOR_SCALAR, ANIFEM_COEF_CONST, f};
    int first = -1;
    bool rsym[kKinds] = {false, false, false};  // reference tensor of the order symmetric in (ia, ib)
    for (int k = 0; k < kKinds; ++k) {
        if (!b->kind_used[k]) continue;
        const int fem = k + ANIFEM_FEM_P1;
        ElemArgs em{}, ef{};
        if ((rc = prepare_elem_args(anifem_op{ANIFEM_GRAD, fem, 1}, anifem_op{ANIFEM_GRAD, fem, 1}, &D, order_of(fem),
                                    em)))
            return rc;
        if ((rc = prepare_elem_args(anifem_op{ANIFEM_IDEN, ANIFEM_FEM_P0, 1}, anifem_op{ANIFEM_IDEN, fem, 1}, &F,
                                    order_of(fem), ef)))
            return rc;
        // kappa PSI^T PSI is symmetric: K packed to 6 values, the column-summed R
        if (!em.Rds) return ANIFEM_ERR_ARG;
        a.kinds[k] = KindTable{em.Rds, em.Rfs, ef.Rd, ef.Rf, em.nfA0, em.nfB0, em.nfB0, em.nfA0 * em.nfB0, ef.nfB0, 0};
        rsym[k] = em.rds_sym != 0;
        a.max_R_rows = std::max(a.max_R_rows, a.kinds[k].R_rows);
        a.max_RF_rows = std::max(a.max_RF_rows, a.kinds[k].RF_rows);
        if (first < 0) {
            first = k;
            a.m = em;  // the geometry tensors of all kinds coincide: K = |T| kappa PSI^T PSI, K_F = |T| f
            a.f = ef;
        }
    }
    if (first < 0) return ANIFEM_ERR_ARG;
    a.n_kinds = kKinds;
    a.m.coef_n = b->n_tets;
    a.m.coords = b->d_coords;
    a.m.tetnodes = b->d_tetnodes;
    a.m.nodes_stride = b->n_tets;
    a.m.pos0 = 0;
    a.f.coef_n = b->n_tets;
    a.has_rhs = 1;
    a.rowlen = pad16<double>(6 + 1);
    a.n_blocks = b->n_blocks;
    a.meta = b->d_meta;
    a.pair_tet = b->d_pair_tet;
    a.pair_ib = b->d_pair_ib;
    a.pair_pos = b->d_pair_pos;
    a.sstart = b->d_sstart;
    a.pstart = b->d_pstart;
    a.val = val;
    a.rhs = rhs;
    a.status = b->d_status;
    a.max_values = b->max_values;
    a.max_pairs = b->max_pairs;
    a.max_slots = b->max_slots;
    if (a.max_pairs > kBlock) return ANIFEM_ERR_DIMS;
    int max_lds = 0;
    ANIFEM_HIP_CHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device));
    const size_t lds_staged =
        block_lds_layout(a.max_values, a.max_pairs, a.max_slots, 8, a.max_R_rows * pad16<double>(6),
                         a.max_RF_rows * pad16<double>(1))
            .total;
    const bool staged = lds_staged <= (size_t)max_lds / 2;
    const size_t lds = staged ? lds_staged
                              : block_lds_layout(a.max_values, a.max_pairs, a.max_slots, 8, 0, 0).total;
    if (lds > (size_t)max_lds) return ANIFEM_ERR_DIMS;

    // persistent row-block passes, one per element order over that order's
    // blocks, in the order's launch shape (default: fixed-size form for P1 and
    // P2, compact positions where present, occupancy x CUs workgroups); LDS
    // sized by the order's own block budgets
    struct KindLaunch {
        const void* fn = nullptr;
        BlockArgs a{};
        size_t lds = 0;
        int64_t grid = 0;
        // form 2 of P3 with the triangle of a symmetric reference tensor
        bool tri = false;
        BlockArgs at{};
        size_t lds_t = 0;
    };
    KindLaunch kl[kKinds];
    bool all_persistent = b->persistent;
    for (int k = 0; k < kKinds && all_persistent; ++k) {
        if (b->kind_blk[k + 1] == b->kind_blk[k]) continue;
        KindLaunch& L = kl[k];
        L.a = a;
        L.a.meta = b->d_meta + b->kind_blk[k];
        L.a.n_blocks = b->kind_blk[k + 1] - b->kind_blk[k];
        L.a.max_values = b->kmax_values[k];
        L.a.max_pairs = b->kmax_pairs[k];
        L.a.max_slots = b->kmax_slots[k];
        L.a.max_R_rows = a.kinds[k].R_rows;
        L.a.max_RF_rows = a.kinds[k].RF_rows;
        L.a.pair_pos8 = b->d_pos8[k];
        L.a.pos8_base = b->kind_pair0[k];
        L.lds = block_lds_layout(L.a.max_values, L.a.max_pairs, L.a.max_slots, 8, L.a.max_R_rows * pad16<double>(6),
                                 L.a.max_RF_rows * pad16<double>(1))
                    .total;
        if (L.lds > (size_t)max_lds / 2) {  // the pass stages R: two workgroups per CU at least
            all_persistent = false;
            break;
        }
        const KindTable& KT = a.kinds[k];
        L.tri = k == 2 && rsym[k] && KT.nfA0 == 20 && KT.nfB0 == 20 && KT.R_rows == 400;
        if (L.tri) {
            L.at = L.a;
            L.at.max_R_rows = 20 * 21 / 2;
            L.lds_t = block_lds_layout(L.at.max_values, L.at.max_pairs, L.at.max_slots, 8,
                                       L.at.max_R_rows * pad16<double>(6), L.at.max_RF_rows * pad16<double>(1))
                          .total;
        }
        if (!b->d_pos8[k]) b->shape[k][1] = 0;
        if (k == 2 && !b->launch_tuned) {  // P3: two lanes per pair when the blocks allow it
            b->shape[k][0] = L.a.max_pairs <= kBlock / 2 ? 2 : 0;
            if (b->shape[k][0] == 2) b->shape[k][1] = 0;
        }
    }
    const int64_t gblocks = (b->n_tets + kBlock - 1) / kBlock;
    const size_t glds = geometry_stage_bytes<1, 1, double>(a.rowlen);
    // launch shapes timed once on the batch's own blocks (the geometry rows of
    // this call first; every candidate writes the same values)
    if (all_persistent && b->tune_launch && !b->launch_tuned) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ANIFEM_HIP_CHECK(hipEventCreate(&e0));
        ANIFEM_HIP_CHECK(hipEventCreate(&e1));
        hipLaunchKernelGGL(batch_status_clear, dim3(1), dim3(64), 0, rs, b->d_status);
        hipLaunchKernelGGL((geometry_kernel<3, 3, 1, 1, double, true>), dim3((unsigned)gblocks), dim3(kBlock), glds, rs,
                           a, b->d_K, b->n_tets);
        const double* Kt = b->d_K;
        int rc = ANIFEM_OK;
        for (int k = 0; k < kKinds && !rc; ++k) {
            if (b->kind_blk[k + 1] == b->kind_blk[k]) continue;
            KindLaunch& L = kl[k];
            double best = 1e300;
            // phase-A forms of this order: fixed-size and general (P1, P2);
            // two lanes per pair where the blocks allow it, and general (P3)
            const int fx_first = k < 2 ? 1 : (L.a.max_pairs <= kBlock / 2 ? 2 : 0);
            for (int fx = fx_first; fx >= 0 && !rc; fx = fx == 2 ? 0 : fx - 1)
                for (int p8 = (b->d_pos8[k] && fx != 2 ? 1 : 0); p8 >= 0 && !rc; --p8) {
                    const bool tri = fx == 2 && L.tri;
                    const void* fn = kind_kernel(k, p8 != 0, fx, tri);
                    BlockArgs& A = tri ? L.at : L.a;
                    const size_t lx = tri ? L.lds_t : L.lds;
                    int occ = 1;
                    int64_t G = 0;
                    if ((rc = persistent_grid(fn, lx, 0, A.n_blocks, b->device, &G, &occ))) break;
                    for (int w = occ; w >= std::max(1, occ - 2) && !rc; --w) {
                        if ((rc = persistent_grid(fn, lx, w, A.n_blocks, b->device, &G, nullptr))) break;
                        void* args[] = {&A, &Kt};
                        const int reps = 3;
                        hipError_t le0 = hipLaunchKernel(fn, dim3((unsigned)G), dim3(kBlock), args, lx, rs);
                        le0 = le0 == hipSuccess ? hipEventRecord(e0, rs) : le0;
                        for (int r = 0; r < reps && le0 == hipSuccess; ++r)
                            le0 = hipLaunchKernel(fn, dim3((unsigned)G), dim3(kBlock), args, lx, rs);
                        le0 = le0 == hipSuccess ? hipEventRecord(e1, rs) : le0;
                        le0 = le0 == hipSuccess ? hipEventSynchronize(e1) : le0;
                        float t = 0.f;
                        le0 = le0 == hipSuccess ? hipEventElapsedTime(&t, e0, e1) : le0;
                        if (le0 != hipSuccess) {
                            rc = ANIFEM_ERR_HIP;
                            break;
                        }
                        if (t / reps < best) {
                            best = t / reps;
                            b->shape[k][0] = fx;
                            b->shape[k][1] = p8;
                            b->shape[k][2] = w;
                            b->shape_ms[k] = best;
                        }
                    }
                }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
        b->launch_tuned = true;
    }
    for (int k = 0; k < kKinds && all_persistent; ++k) {
        if (b->kind_blk[k + 1] == b->kind_blk[k]) continue;
        KindLaunch& L = kl[k];
        const bool tri = b->shape[k][0] == 2 && L.tri;
        L.fn = kind_kernel(k, b->shape[k][1] != 0, b->shape[k][0], tri);
        if (tri) {  // the launch takes the triangle's arguments and LDS
            L.a = L.at;
            L.lds = L.lds_t;
        }
        int rc = persistent_grid(L.fn, L.lds, b->shape[k][2], L.a.n_blocks, b->device, &L.grid, nullptr);
        if (rc) return rc;
    }
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    hipStream_t s = b->cap_stream;
    ANIFEM_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(batch_status_clear, dim3(1), dim3(64), 0, s, b->d_status);
    hipLaunchKernelGGL((geometry_kernel<3, 3, 1, 1, double, true>), dim3((unsigned)gblocks), dim3(kBlock), glds, s, a,
                       b->d_K, b->n_tets);
    const unsigned grid = (unsigned)(8 * ((b->n_blocks + 7) / 8));
    const double* Kc = b->d_K;
    hipError_t le = hipSuccess;
    if (all_persistent) {
        for (int k = 0; k < kKinds && le == hipSuccess; ++k) {
            if (!kl[k].fn) continue;
            void* args[] = {&kl[k].a, &Kc};
            le = hipLaunchKernel(kl[k].fn, dim3((unsigned)kl[k].grid), dim3(kBlock), args, kl[k].lds, s);
        }
    } else if (staged) {
        hipLaunchKernelGGL((row_block_kernel<6, 1, 1, 1, double, true>), dim3(grid), dim3(kBlock), lds, s, a, Kc);
    } else {
        hipLaunchKernelGGL((row_block_kernel<6, 1, 1, 1, double, false>), dim3(grid), dim3(kBlock), lds, s, a, Kc);
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (e != hipSuccess || le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return ANIFEM_ERR_HIP;
    }
    b->graph = g;
    ANIFEM_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->key[0] = kappa;
    b->key[1] = f;
    b->key[2] = val;
    b->key[3] = rhs;
    return ANIFEM_OK;
}

}  // namespace

extern "C" {

int anifem_batch_create(anifem_batch** out, int n_meshes, const anifem_batch_mesh* meshes, int dirichlet,
                        int autotune, void* stream) {
    using namespace anifem;
    if (!out || n_meshes <= 0 || !meshes) return ANIFEM_ERR_ARG;
    if (dirichlet != ANIFEM_DIRICHLET_NONE && dirichlet != ANIFEM_DIRICHLET_BOUNDARY) return ANIFEM_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::unique_ptr<anifem_batch> b(new anifem_batch());
    b->n_meshes = n_meshes;
    b->dirichlet = dirichlet;
    b->tune_launch = autotune != 0;
    ANIFEM_HIP_CHECK(hipGetDevice(&b->device));
    for (int i = 0; i < n_meshes; ++i) {
        const anifem_batch_mesh& m = meshes[i];
        if (m.fem < ANIFEM_FEM_P1 || m.fem > ANIFEM_FEM_P3 || m.n_tets <= 0 || m.n_nodes <= 0 || !m.coords || !m.tets)
            return ANIFEM_ERR_ARG;
        b->fem.push_back(m.fem);
        b->kind_used[kind_of(m.fem)] = true;
    }
    // budgets per element order: timed candidates on the largest mesh of that
    // order within 200k tets (else the smallest); untuned, those of a single
    // mesh of that order
    for (int k = 0; k < kKinds; ++k) {
        b->caps[k] = default_caps(k == 0 ? 4 : k == 1 ? 10 : 20, true);
        if (!b->kind_used[k] || !autotune) continue;
        int pick = -1;
        for (int i = 0; i < n_meshes; ++i) {
            if (kind_of(meshes[i].fem) != k) continue;
            const bool fits = meshes[i].n_tets <= 200000;
            if (pick < 0) {
                pick = i;
                continue;
            }
            const bool pfits = meshes[pick].n_tets <= 200000;
            if ((fits && (!pfits || meshes[i].n_tets > meshes[pick].n_tets)) ||
                (!fits && !pfits && meshes[i].n_tets < meshes[pick].n_tets))
                pick = i;
        }
        const int rc = autotune_kind(b.get(), k, meshes[pick], s);
        if (rc) return rc;
    }
    // per-mesh setup, concatenated with offsets
    std::vector<double> coords;
    std::vector<int32_t> tets;  // global node ids, AoS (after the node reorder)
    std::vector<BlockMeta> meta;
    std::vector<int32_t> pair_tet, col_idx;
    std::vector<uint8_t> pair_ib;
    std::vector<uint16_t> pair_pos, sstart, pstart;
    std::vector<int64_t> row_ptr;
    // per-mesh setups on host threads, each with a stream of its own for the
    // device phase
    std::vector<AssemblySetup> Ss(n_meshes);
    std::vector<std::vector<uint32_t>> P8(n_meshes);  // compact positions of each mesh, when they fit
    std::vector<int> p8_words(n_meshes, 0);
    {
        std::vector<int> rcs(n_meshes, ANIFEM_OK);
        std::atomic<int> next{0};
        const int W = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (int w = 0; w < std::min(W, n_meshes); ++w)
            th.emplace_back([&] {
                hipStream_t ws = nullptr;
                const bool ok = hipSetDevice(b->device) == hipSuccess && hipStreamCreate(&ws) == hipSuccess;
                for (;;) {
                    const int j = next.fetch_add(1);
                    if (j >= n_meshes) break;
                    if (!ok) {
                        rcs[j] = ANIFEM_ERR_HIP;
                        continue;
                    }
                    const anifem_batch_mesh& mj = meshes[j];
                    anifem_mesh_desc mdj{mj.n_nodes, mj.n_tets, mj.coords, mj.tets};
                    rcs[j] = build_setup_device(mdj, problem_of(mj.fem, dirichlet), Ss[j], b->caps[kind_of(mj.fem)], ws);
                    if (!rcs[j]) p8_words[j] = compact_positions(Ss[j], P8[j]);
                    // what the concatenation does not read
                    std::vector<int32_t>().swap(Ss[j].slot);
                    std::vector<int32_t>().swap(Ss[j].dofs);
                }
                if (ws) (void)hipStreamDestroy(ws);
            });
        for (auto& x : th) x.join();
        for (int rc : rcs)
            if (rc) return rc;
    }
    // offsets of each mesh in the concatenated arrays
    b->tet_off.assign(1, 0);
    b->node_off.assign(1, 0);
    b->dof_off.assign(1, 0);
    b->nnz_off.assign(1, 0);
    b->rp_off.assign(1, 0);
    for (int i = 0; i < n_meshes; ++i) {
        const AssemblySetup& S = Ss[i];
        b->tet_off.push_back(b->tet_off.back() + S.n_tets);
        b->node_off.push_back(b->node_off.back() + S.n_nodes);
        b->dof_off.push_back(b->dof_off.back() + S.ndof);
        b->nnz_off.push_back(b->nnz_off.back() + S.nnz);
        b->rp_off.push_back(b->rp_off.back() + S.ndof + 1);
        b->max_values = std::max<int>(b->max_values, (int)S.max_block_values);
        b->max_pairs = std::max<int>(b->max_pairs, (int)S.max_block_pairs);
        b->max_slots = std::max<int>(b->max_slots, (int)S.max_block_slots);
        if (b->tet_off.back() >= INT32_MAX || b->dof_off.back() >= INT32_MAX) return ANIFEM_ERR_ARG;
    }
    // pairs, value positions and block metadata in element-order order (the
    // meshes of one order in mesh order), so that each order's blocks and
    // pairs are contiguous; slots, rows and tets stay in mesh order
    std::vector<int64_t> pair_off(n_meshes), pos_off(n_meshes), meta_off(n_meshes), p8_off(n_meshes, 0);
    std::vector<uint32_t> P8k[kKinds];
    {
        int64_t np = 0, nq = 0, nm = 0;
        for (int k = 0; k < kKinds; ++k) {
            b->kind_blk[k] = nm;
            b->kind_pair0[k] = np;
            const int words = (2 + (k == 0 ? 4 : k == 1 ? 10 : 20) + 3) / 4;
            bool p8 = b->kind_used[k];
            int64_t n8 = 0;
            for (int i = 0; i < n_meshes; ++i) {
                if (kind_of(meshes[i].fem) != k) continue;
                const AssemblySetup& S = Ss[i];
                pair_off[i] = np;
                pos_off[i] = nq;
                meta_off[i] = nm;
                p8_off[i] = n8;
                np += (int64_t)S.pair_tet.size();
                nq += (int64_t)S.pair_pos.size();
                nm += (int64_t)S.blk_meta.size();
                n8 += (int64_t)S.pair_tet.size() * words;
                p8 = p8 && p8_words[i] == words;
                b->kmax_values[k] = std::max<int>(b->kmax_values[k], (int)S.max_block_values);
                b->kmax_pairs[k] = std::max<int>(b->kmax_pairs[k], (int)S.max_block_pairs);
                b->kmax_slots[k] = std::max<int>(b->kmax_slots[k], (int)S.max_block_slots);
            }
            if (p8) P8k[k].resize(n8);
        }
        b->kind_blk[kKinds] = nm;
    }
    int64_t pos_total = 0, pair_total = 0;
    for (int i = 0; i < n_meshes; ++i) {
        pair_total += (int64_t)Ss[i].pair_tet.size();
        pos_total += (int64_t)Ss[i].pair_pos.size();
    }
    coords.resize(3 * (size_t)b->node_off.back());
    tets.resize(4 * (size_t)b->tet_off.back());
    meta.resize(b->kind_blk[kKinds]);
    pair_tet.resize(pair_total);
    pair_ib.resize(pair_total);
    pair_pos.resize(pos_total);
    sstart.resize(b->nnz_off.back());
    pstart.resize(b->dof_off.back());
    row_ptr.resize(b->rp_off.back());
    col_idx.resize(b->nnz_off.back());
    // each mesh fills its ranges (ids shifted by the mesh's offsets), then
    // releases its setup
    host_parallel_for(n_meshes, 1, [&](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) {
            AssemblySetup& S = Ss[i];
            const anifem_batch_mesh& m = meshes[i];
            const int k = kind_of(m.fem);
            const int64_t t0 = b->tet_off[i], n0 = b->node_off[i], d0 = b->dof_off[i], z0 = b->nnz_off[i];
            std::copy(m.coords, m.coords + 3 * m.n_nodes, coords.begin() + 3 * n0);
            for (int64_t q = 0; q < 4 * S.n_tets; ++q) tets[4 * t0 + q] = (int32_t)(S.tets[q] + n0);
            for (size_t u = 0; u < S.blk_meta.size(); ++u) {
                BlockMeta M = S.blk_meta[u];
                M.q0 = pos_off[i] + M.p0 * S.nf;
                M.s0 += z0;
                M.p0 += pair_off[i];
                M.r0 += (int32_t)d0;
                M.kind = (uint32_t)k;
                meta[meta_off[i] + u] = M;
            }
            for (size_t u = 0; u < S.pair_tet.size(); ++u) pair_tet[pair_off[i] + u] = (int32_t)(S.pair_tet[u] + t0);
            std::copy(S.pair_ib.begin(), S.pair_ib.end(), pair_ib.begin() + pair_off[i]);
            std::copy(S.pair_pos.begin(), S.pair_pos.end(), pair_pos.begin() + pos_off[i]);
            if (!P8k[k].empty()) std::copy(P8[i].begin(), P8[i].end(), P8k[k].begin() + p8_off[i]);
            std::vector<uint32_t>().swap(P8[i]);
            std::copy(S.sstart.begin(), S.sstart.end(), sstart.begin() + z0);
            std::copy(S.pstart.begin(), S.pstart.end(), pstart.begin() + d0);
            std::copy(S.row_ptr.begin(), S.row_ptr.end(), row_ptr.begin() + b->rp_off[i]);
            std::copy(S.col_idx.begin(), S.col_idx.end(), col_idx.begin() + z0);
            S = AssemblySetup();
        }
    });
    b->n_tets = b->tet_off.back();
    b->n_nodes = b->node_off.back();
    b->ndof = b->dof_off.back();
    b->nnz = b->nnz_off.back();
    b->n_blocks = (int64_t)meta.size();
    std::vector<int32_t> tn(4 * (size_t)b->n_tets);
    for (int64_t t = 0; t < b->n_tets; ++t)
        for (int v = 0; v < 4; ++v) tn[(size_t)v * b->n_tets + t] = tets[4 * (size_t)t + v];
    std::vector<int32_t>().swap(tets);
    anifem_batch* p = b.get();
    int rc = 0;
    if ((rc = upload_b(p, &p->d_coords, coords, s))) return rc;
    if ((rc = upload_b(p, &p->d_tetnodes, tn, s))) return rc;
    if ((rc = upload_b(p, &p->d_meta, meta, s))) return rc;
    if ((rc = upload_b(p, &p->d_pair_tet, pair_tet, s))) return rc;
    if ((rc = upload_b(p, &p->d_pair_ib, pair_ib, s))) return rc;
    if ((rc = upload_b(p, &p->d_pair_pos, pair_pos, s))) return rc;
    for (int k = 0; k < kKinds; ++k)
        if (!P8k[k].empty() && (rc = upload_b(p, &p->d_pos8[k], P8k[k], s))) return rc;
    sstart.push_back(0);  // padding: slot starts are read as aligned 32-bit words
    if ((rc = upload_b(p, &p->d_sstart, sstart, s))) return rc;
    if ((rc = upload_b(p, &p->d_pstart, pstart, s))) return rc;
    if ((rc = upload_b(p, &p->d_row_ptr, row_ptr, s))) return rc;
    if ((rc = upload_b(p, &p->d_col_idx, col_idx, s))) return rc;
    ANIFEM_HIP_CHECK(hipMalloc((void**)&p->d_status, 4 * sizeof(int)));
    p->bufs.push_back(p->d_status);
    ANIFEM_HIP_CHECK(hipMalloc((void**)&p->d_K, (size_t)p->n_tets * pad16<double>(6 + 1) * sizeof(double)));
    p->bufs.push_back(p->d_K);
    ANIFEM_HIP_CHECK(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
    ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
    *out = b.release();
    return ANIFEM_OK;
}

int anifem_batch_destroy(anifem_batch* b) {
    if (!b) return ANIFEM_OK;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    for (void* p : b->bufs) (void)hipFree(p);
    delete b;
    return ANIFEM_OK;
}

int anifem_batch_totals(const anifem_batch* b, int64_t* n_tets, int64_t* ndof, int64_t* nnz, int64_t* n_blocks) {
    if (!b) return ANIFEM_ERR_ARG;
    if (n_tets) *n_tets = b->n_tets;
    if (ndof) *ndof = b->ndof;
    if (nnz) *nnz = b->nnz;
    if (n_blocks) *n_blocks = b->n_blocks;
    return ANIFEM_OK;
}

int anifem_batch_info(const anifem_batch* b, int mesh, int64_t* tet_offset, int64_t* dof_offset, int64_t* nnz_offset,
                      int64_t* ndof, int64_t* nnz) {
    if (!b || mesh < 0 || mesh >= b->n_meshes) return ANIFEM_ERR_ARG;
    if (tet_offset) *tet_offset = b->tet_off[mesh];
    if (dof_offset) *dof_offset = b->dof_off[mesh];
    if (nnz_offset) *nnz_offset = b->nnz_off[mesh];
    if (ndof) *ndof = b->dof_off[mesh + 1] - b->dof_off[mesh];
    if (nnz) *nnz = b->nnz_off[mesh + 1] - b->nnz_off[mesh];
    return ANIFEM_OK;
}

int anifem_batch_csr(const anifem_batch* b, int mesh, const int64_t** row_ptr, const int32_t** col_idx) {
    if (!b || mesh < 0 || mesh >= b->n_meshes) return ANIFEM_ERR_ARG;
    if (row_ptr) *row_ptr = b->d_row_ptr + b->rp_off[mesh];
    if (col_idx) *col_idx = b->d_col_idx + b->nnz_off[mesh];
    return ANIFEM_OK;
}

int anifem_batch_tuning(const anifem_batch* b, int fem, int32_t* caps, double* ms) {
    if (!b || fem < ANIFEM_FEM_P1 || fem > ANIFEM_FEM_P3) return ANIFEM_ERR_ARG;
    const int k = kind_of(fem);
    if (caps) {
        caps[0] = b->caps[k].max_slots;
        caps[1] = b->caps[k].max_values;
        caps[2] = b->caps[k].max_pairs;
    }
    if (ms)
        for (int c = 0; c < kCandidates; ++c) ms[c] = b->tune_ms[k][c];
    return ANIFEM_OK;
}

int anifem_batch_launch_tuning(const anifem_batch* b, int fem, int32_t* shape, double* ms) {
    if (!b || fem < ANIFEM_FEM_P1 || fem > ANIFEM_FEM_P3) return ANIFEM_ERR_ARG;
    const int k = kind_of(fem);
    if (shape)
        for (int i = 0; i < 3; ++i) shape[i] = b->shape[k][i];
    if (ms) *ms = b->shape_ms[k];
    return ANIFEM_OK;
}

int anifem_batch_set_persistent(anifem_batch* b, int on) {
    if (!b) return ANIFEM_ERR_ARG;
    if (b->persistent != (on != 0)) {
        b->persistent = on != 0;
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;  // captured again on the next run
        b->graph = nullptr;
    }
    return ANIFEM_OK;
}

int anifem_batch_run(anifem_batch* b, const double* kappa, const double* f, double* val, double* rhs,
                     anifem_status* status, void* stream) {
    if (!b || !kappa || !f || !val || !rhs) return ANIFEM_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!b->exec || b->key[0] != kappa || b->key[1] != f || b->key[2] != val || b->key[3] != rhs) {
        const int rc = capture(b, kappa, f, val, rhs, s);
        if (rc) return rc;
    }
    ANIFEM_HIP_CHECK(hipGraphLaunch(b->exec, s));
    if (status) {
        int h[4] = {0, 0, 0, 0};
        ANIFEM_HIP_CHECK(hipMemcpyAsync(h, b->d_status, sizeof(h), hipMemcpyDeviceToHost, s));
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
