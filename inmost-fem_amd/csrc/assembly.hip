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
#include <mutex>

#include "anifem_common.h"
#include "assembly_kernels.h"
#include "setup.h"

namespace anifem {
int prepare_elem_args(const anifem_op& opA, const anifem_op& opB, const anifem_tensor* D, int order, ElemArgs& a);
int launch_physical_points(int order, int64_t n, const double* xyz, const double* coords, const int32_t* tetnodes,
                           int64_t nodes_stride, double* X, hipStream_t s);
int launch_apply(const anifem_op& op, const double* xyl, int q, ApplyArgs a, bool mesh, hipStream_t s);
}  // namespace anifem

struct anifem_assembly {
    anifem::AssemblySetup S;
    anifem_problem_desc prob{};
    int device = 0;
    double* d_coords = nullptr;       // n_nodes x 3
    int32_t* d_tetnodes = nullptr;    // 4 x n_tets SoA
    int32_t* d_blk_row0 = nullptr;
    anifem::BlockMeta* d_blk_meta = nullptr;
    int64_t* d_blk_s0 = nullptr;
    int64_t* d_blk_p0 = nullptr;
    int64_t* d_blk_v0 = nullptr;
    int32_t* d_pair_tet = nullptr;
    uint8_t* d_pair_ib = nullptr;
    uint16_t* d_pair_pos = nullptr;
    uint32_t* d_pair_pos8 = nullptr;  // compact positions (compact_positions), when they fit
    int pos8_words = 0;
    uint16_t* d_sstart = nullptr;
    uint16_t* d_pstart = nullptr;
    int64_t* d_row_ptr = nullptr;
    int32_t* d_col_idx = nullptr;
    int* d_status = nullptr;          // [0] non-finite, [1] degenerate
    int* d_deg = nullptr;             // degenerate cells per geometry workgroup (fused pass)
    void* d_kbuf = nullptr;           // per-tet geometry tensors of the last run
    // plain baseline (anifem_assembly_run_plain): slot map, DOFs, flags, tables
    int32_t* d_slot = nullptr;
    int32_t* d_dofs = nullptr;
    uint8_t* d_isdir = nullptr;
    double* d_plain_tab = nullptr;
    size_t kbuf_bytes = 0;
    bool host_only = false;           // built by anifem_assembly_create_host
    bool timers_on = false;
    bool persistent = true;           // persistent row-block pass where it applies
    // coloured scatter (anifem_assembly_set_scatter): tets by colour, in
    // ascending tet order within a colour; colour c is [colour_ptr[c], colour_ptr[c+1])
    bool coloured = false;
    int32_t* d_col_tets = nullptr;
    std::vector<int64_t> colour_ptr;
    // inhomogeneous Dirichlet values: the rows they change (LiftRow records)
    anifem::LiftRow* d_lift = nullptr;
    int64_t n_lift = 0;
    // spaces assembled through element matrices (Piola-mapped RT0 / ND0 and
    // one-DOF spaces): orientation signs, the cells' vertices and the element
    // matrices and load vectors of the last run
    bool elem_path = false;
    int8_t* d_sign = nullptr;
    uint64_t* d_signmask = nullptr;   // per tet: bit j = local DOF j has sign -1 (oriented spaces)
    bool vars_only = false;           // anifem_assembly_create_vars: caller element matrices only
    bool fixed_nf = true;             // fixed-size persistent instantiations where they apply
    double* d_xyz = nullptr;
    void* d_ebuf = nullptr;
    void* d_fbuf = nullptr;
    size_t ebuf_bytes = 0, fbuf_bytes = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    anifem_timers timers{};
    const double* run_bc = nullptr;   // the running call's Dirichlet values (device), or null
    void* run_scratch = nullptr;      // no-allocation call: this device's scratch array, else null
    void* run_kg = nullptr;           // geometry rows of the running call
};

namespace {

// Scratch of anifem_assembly_run_noalloc: the geometry rows of one pass live
// in this array instead of a buffer the context allocates.  Every device the
// library runs on holds its own copy for the life of the process; host code
// uses the device address from scratch_address(), never the array's name.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// This device's address of `scratch`, looked up once per device (with that
// device current) and kept.
int scratch_address(int device, void** out) {
    static std::mutex mu;
    static void* addr[256] = {};
    if (device < 0 || device >= 256) return ANIFEM_ERR_ARG;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            std::fprintf(stderr, "anifem: hipGetSymbolAddress(scratch): %s\n", hipGetErrorString(e));
            return ANIFEM_ERR_HIP;
        }
        addr[device] = p;
    }
    *out = addr[device];
    return ANIFEM_OK;
}

__global__ void fill_kernel(double* p, int64_t n, double v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void fill_int_kernel(int* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
// *out = the sum of the per-workgroup degenerate counts (one workgroup).
__global__ __launch_bounds__(anifem::kBlock) void sum_counts_kernel(const int* __restrict__ c, int64_t n, int* out) {
    constexpr int B = anifem::kBlock;
    __shared__ int part[B];
    int s = 0;
    for (int64_t i = threadIdx.x; i < n; i += B) s += c[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = B / 2; h > 0; h /= 2) {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0];
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

// One colour of the coloured scatter: one lane per (tet of the colour, entry
// e = ib + nf*ia).  The entry is formed from the tet's geometry row as in the
// row-block pass and added to its CSR slot; tets of one colour share no node,
// hence no DOF, so no two lanes of a launch touch one value and the colours
// run one launch after another: no atomics, a fixed (colour) order.  Element-
// level Dirichlet as applyDir (dc_on_dof.h:38-45): a kept entry of a
// Dirichlet row is its diagonal, 1; the row's load-vector entry is 0.
template <int DA0, int DB0, int NCA, int NCB, typename T>
__global__ __launch_bounds__(anifem::kBlock) void colour_scatter_kernel(anifem::BlockArgs a, const T* __restrict__ Kg,
                                                                        const int32_t* __restrict__ ctets, int64_t n_c,
                                                                        const int32_t* __restrict__ slot,
                                                                        const int32_t* __restrict__ dofs,
                                                                        const uint8_t* __restrict__ isdir) {
    using namespace anifem;
    const KindTable& KT = a.kinds[0];
    const int nf = KT.nf, nfA0 = KT.nfA0, nfB0 = KT.nfB0, E = nf * nf;
    const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (i >= n_c * E) return;
    const int64_t t = ctets[i / E];
    const int e = (int)(i % E);
    const int ib = e % nf, ia = e / nf;
    const int ib0 = ib % nfB0, cb = ib / nfB0, ia0 = ia % nfA0, ca = ia / nfA0;
    const int len = a.m.q * DA0 * DB0;
    const int kstride = a.lamec ? anifem::kLameCompact : NCA * NCB * len;
    const T* R = reinterpret_cast<const T*>(sizeof(T) == 8 ? (const void*)KT.Rd : (const void*)KT.Rf);
    const T* Kb = Kg + t * a.rowlen + (cb * NCA + ca) * len;
    const T* Rr = R + (ib0 + nfB0 * ia0) * len;
    T acc = 0;
    if constexpr (NCA == 3 && NCB == 3 && DA0 * DB0 == 9) {
        // compact Lame row: the (cb, ca) block in closed form
        if (a.lamec) acc = anifem::lame_entry(Kg + t * a.rowlen, cb, ca, Rr);
    }
    if (!a.lamec)
        for (int j = 0; j < len; ++j) acc += Kb[j] * Rr[j];
    const int32_t dof = dofs[t * nf + ib];
    const bool dir = isdir[dof] != 0;
    int bad = 0;
    const int32_t sl = slot[t * E + e];
    if (sl >= 0) {
        const double v = dir ? 1.0 : (double)acc;
        bad += isfinite(v) ? 0 : 1;
        a.val[sl] += v;
    }
    if (a.has_rhs && ia == 0) {
        const int lenF = a.f.q;
        const T* RF = reinterpret_cast<const T*>(sizeof(T) == 8 ? (const void*)KT.RFd : (const void*)KT.RFf);
        const T* KF = Kg + t * a.rowlen + kstride + cb * lenF;
        T fa = 0;
        for (int j = 0; j < lenF; ++j) fa += KF[j] * RF[ib0 * lenF + j];
        const double fv = dir ? 0.0 : (double)fa;
        bad += isfinite(fv) ? 0 : 1;
        a.rhs[dof] += fv;
    }
    if (bad) atomicAdd(&a.status[0], bad);
}

template <int DA0, int DB0, int NCA, int NCB, typename T>
int launch_coloured(anifem_assembly* ctx, const anifem::BlockArgs& a, const T* Kg, hipStream_t s) {
    using namespace anifem;
    const AssemblySetup& S = ctx->S;
    int rc;
    if ((rc = fill(a.val, S.nnz, s))) return rc;
    if (a.has_rhs && (rc = fill(a.rhs, S.ndof, s))) return rc;
    const int64_t E = (int64_t)S.nf * S.nf;
    for (size_t c = 0; c + 1 < ctx->colour_ptr.size(); ++c) {
        const int64_t n_c = ctx->colour_ptr[c + 1] - ctx->colour_ptr[c];
        const int64_t blocks = (n_c * E + kBlock - 1) / kBlock;
        if (!ctx->d_slot || !ctx->d_dofs || !ctx->d_isdir || !ctx->d_col_tets) return ANIFEM_ERR_HIP;
        hipLaunchKernelGGL((colour_scatter_kernel<DA0, DB0, NCA, NCB, T>), dim3((unsigned)blocks), dim3(kBlock), 0, s,
                           a, Kg, ctx->d_col_tets + ctx->colour_ptr[c], n_c, ctx->d_slot, ctx->d_dofs, ctx->d_isdir);
    }
    ANIFEM_HIP_CHECK(hipGetLastError());
    return ANIFEM_OK;
}

// Element-to-slot map, element DOFs and Dirichlet flags on the device, for the
// paths that scatter per element (plain baseline, coloured scatter); uploaded
// once, each array on its own.
int upload_maps(anifem_assembly* ctx, hipStream_t s) {
    const anifem::AssemblySetup& S = ctx->S;
    int rc;
    if (!ctx->d_slot && (rc = upload(&ctx->d_slot, S.slot.data(), S.slot.size(), s))) return rc;
    if (!ctx->d_dofs && (rc = upload(&ctx->d_dofs, S.dofs.data(), S.dofs.size(), s))) return rc;
    if (!ctx->d_isdir && (rc = upload(&ctx->d_isdir, S.is_dir.data(), S.is_dir.size(), s))) return rc;
    return ANIFEM_OK;
}

// Persistent row-block pass: occupancy x CUs workgroups, whole XCD groups
// (surplus workgroups exit at once).
template <int DA0, int DB0, int PW, typename T, bool P8, int NFX = 0, bool SPLIT = false, bool TRI = false>
int launch_persistent_p(anifem_assembly* ctx, const anifem::BlockArgs& a, const T* Kg, size_t lds, hipStream_t s) {
    using namespace anifem;
    int occ = 0, n_cu = 0;
    ANIFEM_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &occ, reinterpret_cast<const void*>(&row_block_persistent<DA0, DB0, PW, T, P8, NFX, SPLIT, TRI>), kBlock,
        lds));
    ANIFEM_HIP_CHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    int64_t G = (int64_t)std::max(occ, 1) * std::max(n_cu, 1);
    G = std::min<int64_t>(G, ctx->S.n_blocks);
    G = std::max<int64_t>(8, G / 8 * 8);
    hipLaunchKernelGGL((row_block_persistent<DA0, DB0, PW, T, P8, NFX, SPLIT, TRI>), dim3((unsigned)G), dim3(kBlock),
                       lds, s, a, Kg);
    ANIFEM_HIP_CHECK(hipGetLastError());
    return ANIFEM_OK;
}

// The compact positions when the context has them in this instantiation's
// word count, else the 16-bit positions (identical values either way).
// A single element kind of exactly 2*PW local DOFs (P1: 4, P2: 10) takes the
// fixed-size instantiation (for P3 the general one measured faster: 0.62 against
// 0.65 ms for the row blocks of 384,000 tets).  Blocks of at most kBlock/2
// pairs with longer pairs (P3) take two lanes per pair.
template <int DA0, int DB0, int PW, typename T>
int launch_persistent(anifem_assembly* ctx, const anifem::BlockArgs& a, const T* Kg, size_t lds, hipStream_t s) {
    const bool fixed = PW <= 5 && ctx->fixed_nf && a.n_kinds == 1 && a.kinds[0].nf == 2 * PW &&
                       a.kinds[0].nfA0 == 2 * PW && a.kinds[0].nfB0 == 2 * PW;
    constexpr int NFX = PW <= 5 ? 2 * PW : 0;
    if constexpr (PW > 5 && PW % 2 == 0) {
        if (ctx->S.max_block_pairs <= anifem::kBlock / 2) {
            // a packed symmetric K against a reference tensor symmetric in
            // (ia, ib): its triangle only in LDS
            if constexpr (DA0 == 6 && DB0 == 1) {
                const anifem::KindTable& K0 = a.kinds[0];
                constexpr int NF = 2 * PW;
                if (a.n_kinds == 1 && a.m.rds_sym && K0.Rd == a.m.Rds && K0.nfA0 == NF && K0.nfB0 == NF &&
                    K0.R_rows == NF * NF) {
                    anifem::BlockArgs at = a;
                    at.max_R_rows = NF * (NF + 1) / 2;
                    const size_t lds_t =
                        anifem::block_lds_layout(at.max_values, at.max_pairs, at.max_slots, (int)sizeof(T),
                                                 at.max_R_rows * anifem::pad16<T>(DA0 * DB0),
                                                 at.has_rhs ? at.max_RF_rows * anifem::pad16<T>(1) : 0)
                            .total;
                    return launch_persistent_p<DA0, DB0, PW, T, false, 0, true, true>(ctx, at, Kg, lds_t, s);
                }
            }
            return launch_persistent_p<DA0, DB0, PW, T, false, 0, true>(ctx, a, Kg, lds, s);
        }
    }
    if (ctx->d_pair_pos8 && ctx->pos8_words == (2 + 2 * PW + 3) / 4)
        return fixed ? launch_persistent_p<DA0, DB0, PW, T, true, NFX>(ctx, a, Kg, lds, s)
                     : launch_persistent_p<DA0, DB0, PW, T, true>(ctx, a, Kg, lds, s);
    return fixed ? launch_persistent_p<DA0, DB0, PW, T, false, NFX>(ctx, a, Kg, lds, s)
                 : launch_persistent_p<DA0, DB0, PW, T, false>(ctx, a, Kg, lds, s);
}

// Persistent row-block pass for vector spaces (NC components of NF0 scalar
// DOFs on both sides): occupancy x CUs workgroups, whole XCD groups.
template <int DA0, int DB0, int NC, int NF0, typename T, bool LC = false>
int launch_persistent_vec(anifem_assembly* ctx, const anifem::BlockArgs& a, const T* Kg, size_t lds, hipStream_t s) {
    using namespace anifem;
    int occ = 0, n_cu = 0;
    ANIFEM_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &occ, reinterpret_cast<const void*>(&row_block_persistent_vec<DA0, DB0, NC, NF0, T, LC>), kBlock, lds));
    ANIFEM_HIP_CHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    int64_t G = (int64_t)std::max(occ, 1) * std::max(n_cu, 1);
    G = std::min<int64_t>(G, ctx->S.n_blocks);
    G = std::max<int64_t>(8, G / 8 * 8);
    hipLaunchKernelGGL((row_block_persistent_vec<DA0, DB0, NC, NF0, T, LC>), dim3((unsigned)G), dim3(kBlock), lds, s,
                       a, Kg);
    ANIFEM_HIP_CHECK(hipGetLastError());
    return ANIFEM_OK;
}

// Compact Lame rows (lame_kblock) for GRAD x GRAD of 3-component spaces with
// the Lame tensor constant over each tet, volume integrals.
template <int NCA, int NCB, int KDA, int KDB>
bool lame_compact(const anifem::BlockArgs& a) {
    if constexpr (NCA == 3 && NCB == 3 && KDA == 3 && KDB == 3)
        return a.m.ttype == ANIFEM_TENSOR_LAME && a.m.q == 1 && a.m.face < 0 && a.m.selA == ANIFEM_GRAD &&
               a.m.selB == ANIFEM_GRAD && a.m.mapA == 0 && a.m.mapB == 0 && a.m.scA == 0 && a.m.scB == 0 &&
               (a.m.storage == ANIFEM_COEF_CONST || a.m.storage == ANIFEM_COEF_PER_TET);
    return false;
}

// SYM: K symmetric (the same 3-dimensional operator on both sides of a scalar
// space, a NULL / SCALAR / SYMMETRIC tensor constant over each tet): the
// geometry rows hold K packed to 6 values and the row-block passes run as
// for 6 x 1 blocks against the column-summed R of the kind tables.
template <int DA0, int DB0, int NCA, int NCB, typename T, bool SYM>
int launch_blocks(anifem_assembly* ctx, anifem::BlockArgs a, hipStream_t s) {
    using namespace anifem;
    constexpr int KDA = SYM ? 6 : DA0, KDB = SYM ? 1 : DB0;  // contraction block of the row-block passes
    const AssemblySetup& S = ctx->S;
    const int len = a.m.q * KDA * KDB, lenF = a.has_rhs ? a.f.q : 0;
    a.lamec = lame_compact<NCA, NCB, KDA, KDB>(a) ? 1 : 0;
    const int kstride = a.lamec ? kLameCompact : NCA * NCB * len;
    a.rowlen = pad16<T>(kstride + (a.has_rhs ? NCB * lenF : 0));
    int max_lds = 0;
    ANIFEM_HIP_CHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
    auto lds_of = [&](bool with_R) {
        return block_lds_layout(a.max_values, a.max_pairs, a.max_slots, (int)sizeof(T),
                                with_R ? a.max_R_rows * pad16<T>(len) : 0,
                                with_R && a.has_rhs ? a.max_RF_rows * pad16<T>(lenF) : 0)
            .total;
    };
    // the reference tensors are staged in LDS while two workgroups fit a CU
    // (and the staging's 16-bit indexing holds)
    const bool staged = lds_of(true) <= (size_t)max_lds / 2 && a.max_R_rows * len < 65536 &&
                        a.max_RF_rows * lenF < 65536;
    if (a.max_pairs > kBlock) return ANIFEM_ERR_DIMS;
    const size_t lds = lds_of(staged);
    if (lds > (size_t)max_lds) return ANIFEM_ERR_DIMS;
    // geometry rows of every tet: in the no-allocation call's scratch array,
    // else in a buffer the context keeps between runs
    const size_t need = (size_t)S.n_tets * a.rowlen * sizeof(T);
    T* Kg;
    if (ctx->run_scratch) {
        if (need > kScratchBytes) return ANIFEM_ERR_DIMS;
        Kg = static_cast<T*>(ctx->run_scratch);
    } else {
        if (ctx->kbuf_bytes < need) {
            if (ctx->d_kbuf) ANIFEM_HIP_CHECK(hipFree(ctx->d_kbuf));
            ctx->d_kbuf = nullptr;
            ctx->kbuf_bytes = 0;
            ANIFEM_HIP_CHECK(hipMalloc(&ctx->d_kbuf, need));
            ctx->kbuf_bytes = need;
        }
        Kg = static_cast<T*>(ctx->d_kbuf);
    }
    ctx->run_kg = Kg;
    const int64_t gblocks = (S.n_tets + kBlock - 1) / kBlock;
    const size_t glds = geometry_stage_bytes<NCA, NCB, T>(a.rowlen);
    if constexpr (NCA == 3 && NCB == 3 && DA0 == 3 && DB0 == 3) {
        const size_t glds_lc = geometry_stage_bytes<3, 3, T, true>(a.rowlen);
        if (a.lamec)
            hipLaunchKernelGGL((geometry_kernel<3, 3, 3, 3, T, false, true>), dim3((unsigned)gblocks), dim3(kBlock),
                               glds_lc, s, a, Kg, S.n_tets);
        else
            hipLaunchKernelGGL((geometry_kernel<3, 3, 3, 3, T>), dim3((unsigned)gblocks), dim3(kBlock), glds, s, a, Kg,
                               S.n_tets);
    } else {
        hipLaunchKernelGGL((geometry_kernel<DA0, DB0, NCA, NCB, T, SYM>), dim3((unsigned)gblocks), dim3(kBlock), glds,
                           s, a, Kg, S.n_tets);
    }
    ANIFEM_HIP_CHECK(hipGetLastError());
    if (ctx->timers_on) ANIFEM_HIP_CHECK(hipEventRecord(ctx->ev[1], s));
    if (ctx->coloured) return launch_coloured<KDA, KDB, NCA, NCB, T>(ctx, a, Kg, s);
    if (S.n_blocks == 0) return ANIFEM_OK;
    // persistent pass for scalar spaces with one carried point
    constexpr bool scalar_space = NCA == 1 && NCB == 1;
    if (scalar_space && ctx->persistent && staged && len == KDA * KDB && (!a.has_rhs || lenF == 1) &&
        a.kinds[0].nf <= 20) {
        // position words per pair: the shortest instantiation that holds nf
        const int nf = a.kinds[0].nf;
        if (nf <= 4) return launch_persistent<KDA, KDB, 2, T>(ctx, a, Kg, lds, s);
        if (nf <= 10) return launch_persistent<KDA, KDB, 5, T>(ctx, a, Kg, lds, s);
        return launch_persistent<KDA, KDB, 10, T>(ctx, a, Kg, lds, s);
    }
    // persistent pass for 3-component vector spaces of P1 or P2 components:
    // the reference tensors staged in LDS whenever one workgroup fits a CU
    // (vector blocks hold one or two workgroups per CU either way)
    if constexpr (NCA == 3 && NCB == 3) {
        const KindTable& K0 = a.kinds[0];
        const size_t lds_v = lds_of(true);
        if (ctx->persistent && lds_v <= (size_t)max_lds && a.max_R_rows * len < 65536 && len == KDA * KDB &&
            (!a.has_rhs || lenF == 1) && K0.nfA0 == K0.nfB0) {
            if constexpr (DA0 == 3 && DB0 == 3) {
                if (a.lamec && K0.nfA0 == 4)
                    return launch_persistent_vec<3, 3, 3, 4, T, true>(ctx, a, (const T*)Kg, lds_v, s);
                if (a.lamec && K0.nfA0 == 10)
                    return launch_persistent_vec<3, 3, 3, 10, T, true>(ctx, a, (const T*)Kg, lds_v, s);
            }
            if (K0.nfA0 == 4) return launch_persistent_vec<DA0, DB0, 3, 4, T>(ctx, a, (const T*)Kg, lds_v, s);
            if (K0.nfA0 == 10) return launch_persistent_vec<DA0, DB0, 3, 10, T>(ctx, a, (const T*)Kg, lds_v, s);
        }
    }
    // 8 equal XCD ranges of the processing order (padding workgroups exit)
    const unsigned grid = (unsigned)(8 * ((S.n_blocks + 7) / 8));
    if (staged)
        hipLaunchKernelGGL((row_block_kernel<KDA, KDB, NCA, NCB, T, true>), dim3(grid), dim3(kBlock), lds, s, a,
                           (const T*)Kg);
    else
        hipLaunchKernelGGL((row_block_kernel<KDA, KDB, NCA, NCB, T, false>), dim3(grid), dim3(kBlock), lds, s, a,
                           (const T*)Kg);
    ANIFEM_HIP_CHECK(hipGetLastError());
    return ANIFEM_OK;
}

// The row-block pass, then the rows an inhomogeneous Dirichlet condition
// changes (when the call has Dirichlet values and a rhs).
template <int DA0, int DB0, int NCA, int NCB, typename T, bool SYM = false>
int launch_blocks_lift(anifem_assembly* ctx, anifem::BlockArgs a, hipStream_t s) {
    using namespace anifem;
    constexpr int KDA = SYM ? 6 : DA0, KDB = SYM ? 1 : DB0;
    int rc = launch_blocks<DA0, DB0, NCA, NCB, T, SYM>(ctx, a, s);
    if (rc || !ctx->run_bc || !a.rhs || ctx->n_lift == 0) return rc;
    const int len = a.m.q * KDA * KDB, lenF = a.has_rhs ? a.f.q : 0;
    // as launch_blocks sets them
    a.lamec = lame_compact<NCA, NCB, KDA, KDB>(a) ? 1 : 0;
    a.rowlen = pad16<T>((a.lamec ? kLameCompact : NCA * NCB * len) + (a.has_rhs ? NCB * lenF : 0));
    const int64_t blocks = (ctx->n_lift + kBlock - 1) / kBlock;
    hipLaunchKernelGGL((dirichlet_lift_kernel<KDA, KDB, NCA, NCB, T>), dim3((unsigned)blocks), dim3(kBlock), 0, s, a,
                       static_cast<const T*>(ctx->run_kg), ctx->d_lift, ctx->n_lift, ctx->d_dofs, ctx->d_isdir,
                       ctx->run_bc);
    ANIFEM_HIP_CHECK(hipGetLastError());
    return ANIFEM_OK;
}

template <typename T>
int dispatch_blocks(anifem_assembly* ctx, const anifem::BlockArgs& a, bool sym, hipStream_t s) {
    const int dA = anifem::ref_dim(ctx->prob.trial.op, ctx->prob.trial.fem),
              dB = anifem::ref_dim(ctx->prob.test.op, ctx->prob.test.fem);
    const int nc = ctx->prob.trial.ncomp;
    if (dA == 3 && dB == 3 && nc == 1 && sym) return launch_blocks_lift<3, 3, 1, 1, T, true>(ctx, a, s);
    if (dA == 3 && dB == 3 && nc == 1) return launch_blocks_lift<3, 3, 1, 1, T>(ctx, a, s);
    if (dA == 3 && dB == 3 && nc == 3) return launch_blocks_lift<3, 3, 3, 3, T>(ctx, a, s);
    if (dA == 1 && dB == 1 && nc == 1) return launch_blocks_lift<1, 1, 1, 1, T>(ctx, a, s);
    if (dA == 1 && dB == 1 && nc == 3) return launch_blocks_lift<1, 1, 3, 3, T>(ctx, a, s);
    if (dA == 3 && dB == 1 && nc == 1) return launch_blocks_lift<3, 1, 1, 1, T>(ctx, a, s);
    if (dA == 1 && dB == 3 && nc == 1) return launch_blocks_lift<1, 3, 1, 1, T>(ctx, a, s);
    return ANIFEM_ERR_ARG;
}

// Plain FP64 method, one lane per tet, following the reference's algorithm
// step by step (core.inl:218-367): J, J^-1, |T|; at each quadrature point the
// physical gradients U = PSI^T Ghat (poly_*.h), DU = w |T| D U
// (diff_tensor.h applyDU), A += V^T DU (fusive_AT_mul_B, core.inl:29-60);
// the load vector with the basis values; element-level Dirichlet
// (dc_on_dof.h:38-45); then FP64 atomic adds into the CSR through the
// element-to-slot map (assembler.inl:397-481).  A measurement baseline: the
// order of the atomic adds is not fixed.
template <int NF>
__global__ __launch_bounds__(128) void plain_kernel(anifem::ElemArgs m, anifem::ElemArgs f, int has_rhs,
                                                    const double* gtab, const double* wq, int q,
                                                    const double* vtab, const double* wf, int qf,
                                                    const int32_t* slot, const int32_t* dofs, const uint8_t* isdir,
                                                    double* val, double* rhs, int64_t n, int* status) {
    using namespace anifem;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n) return;
    double P[4][3], PSI[9], det;
    load_tet<true>(m, t, P);
    tet_geometry(P, PSI, det);
    const double mes = fabs(det) / 6.0;
    double A[NF * NF];
    for (int e = 0; e < NF * NF; ++e) A[e] = 0.0;
    for (int qn = 0; qn < q; ++qn) {
        double U[NF][3], DU[NF][3], D[3][3];
        for (int i = 0; i < NF; ++i)
            for (int k = 0; k < 3; ++k) {
                double acc = 0;
                for (int j = 0; j < 3; ++j) acc += PSI[j + 3 * k] * gtab[(qn * NF + i) * 3 + j];
                U[i][k] = acc;
            }
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) D[k][j] = tensor_at(m, t, qn, k, j);
        const double s = wq[qn] * mes;
        for (int i = 0; i < NF; ++i)
            for (int k = 0; k < 3; ++k) DU[i][k] = s * (D[k][0] * U[i][0] + D[k][1] * U[i][1] + D[k][2] * U[i][2]);
        for (int ia = 0; ia < NF; ++ia)
            for (int ib = 0; ib < NF; ++ib)
                A[ib + NF * ia] += U[ib][0] * DU[ia][0] + U[ib][1] * DU[ia][1] + U[ib][2] * DU[ia][2];
    }
    double F[NF];
    for (int i = 0; i < NF; ++i) F[i] = 0.0;
    if (has_rhs)
        for (int qn = 0; qn < qf; ++qn) {
            const double s = wf[qn] * mes * tensor_at(f, t, qn, 0, 0);
            for (int i = 0; i < NF; ++i) F[i] += s * vtab[qn * NF + i];
        }
    // applyDir with value 0 on the flagged local DOFs
    const int32_t* td = dofs + t * NF;
    for (int k = 0; k < NF; ++k) {
        if (!isdir[td[k]]) continue;
        for (int i = 0; i < NF; ++i) {
            A[k + NF * i] = 0.0;
            A[i + NF * k] = 0.0;
        }
        A[k + NF * k] = 1.0;
        F[k] = 0.0;
    }
    int bad = 0;
    const int32_t* sl = slot + t * NF * NF;
    for (int e = 0; e < NF * NF; ++e) {
        bad += isfinite(A[e]) ? 0 : 1;
        if (sl[e] >= 0 && fabs(A[e]) > 1e-100) atomicAdd(val + sl[e], A[e]);
    }
    if (has_rhs)
        for (int i = 0; i < NF; ++i) {
            bad += isfinite(F[i]) ? 0 : 1;
            atomicAdd(rhs + td[i], F[i]);
        }
    if (bad) atomicAdd(status, bad);
}

// First-fit greedy colouring of the tets in ascending order: two tets
// conflict when they share a node (then they may share a DOF; tets with no
// common node share none).  Returns the number of colours.
int greedy_colour(int64_t n_nodes, int64_t n_tets, const int32_t* tets, std::vector<int32_t>& colour) {
    std::vector<int64_t> ptr((size_t)n_nodes + 1, 0);
    for (int64_t k = 0; k < 4 * n_tets; ++k) ptr[(size_t)tets[k] + 1]++;
    for (int64_t i = 0; i < n_nodes; ++i) ptr[i + 1] += ptr[i];
    std::vector<int64_t> fillp(ptr.begin(), ptr.end() - 1);
    std::vector<int32_t> adj((size_t)(4 * n_tets));
    for (int64_t t = 0; t < n_tets; ++t)
        for (int k = 0; k < 4; ++k) adj[(size_t)fillp[(size_t)tets[4 * t + k]]++] = (int32_t)t;
    colour.assign((size_t)n_tets, -1);
    std::vector<uint64_t> used;
    int ncol = 0;
    for (int64_t t = 0; t < n_tets; ++t) {
        used.assign((size_t)ncol / 64 + 1, 0);
        for (int k = 0; k < 4; ++k) {
            const int32_t v = tets[4 * t + k];
            for (int64_t q = ptr[v]; q < ptr[v + 1]; ++q) {
                const int32_t c = colour[(size_t)adj[(size_t)q]];
                if (c >= 0) used[(size_t)c >> 6] |= uint64_t(1) << (c & 63);
            }
        }
        int c = 0;
        while ((size_t)(c >> 6) < used.size() && (used[(size_t)c >> 6] >> (c & 63)) & 1) ++c;
        colour[(size_t)t] = c;
        ncol = std::max(ncol, c + 1);
    }
    return ncol;
}

// The setup of a context: its enumeration phase on the device for a device
// context, on the host for a host-only one (identical results).
int build_host(anifem_assembly* ctx, const anifem_mesh_desc* mesh, const anifem_problem_desc* prob,
               const anifem::BlockCaps* caps = nullptr, bool on_device = false, void* stream = nullptr,
               const anifem_vars_desc* vars = nullptr) {
    using namespace anifem;
    if (!mesh || !prob || !mesh->coords || !mesh->tets) return ANIFEM_ERR_ARG;
    if (vars) {
        // a multi-variable problem: element matrices only, numbering on the host
        if (vars->n_var < 1 || vars->n_var > kMaxVars) return ANIFEM_ERR_ARG;
        int nf = 0;
        for (int v = 0; v < vars->n_var; ++v) {
            if (!space_layout(vars->var[v].fem).valid || vars->var[v].ncomp < 1) return ANIFEM_ERR_ARG;
            nf += nf_scalar(vars->var[v].fem) * vars->var[v].ncomp;
        }
        ctx->prob = *prob;
        ctx->prob.dirichlet_dof_mask = nullptr;
        ctx->prob.trial = ctx->prob.test = anifem_op{ANIFEM_IDEN, vars->var[0].fem, vars->var[0].ncomp};
        ctx->vars_only = true;
        const BlockCaps c = caps ? *caps : default_caps(nf, false);
        const int rc = build_setup_vars(*mesh, vars->n_var, vars->var, *prob, ctx->S, c);
        if (rc != ANIFEM_ERR_DIMS || caps) return rc;
        ctx->S = AssemblySetup();  // the default budgets, then wider ones when one row exceeds them
        return build_setup_vars(*mesh, vars->n_var, vars->var, *prob, ctx->S, wide_caps(nf));
    }
    if (!op_valid(prob->trial) || !op_valid(prob->test)) return ANIFEM_ERR_ARG;
    // the global assembly numbers the DOFs of the spaces of space_layout
    if (!space_layout(prob->trial.fem).valid || !space_layout(prob->test.fem).valid) return ANIFEM_ERR_ARG;
    ctx->prob = *prob;
    ctx->prob.dirichlet_dof_mask = nullptr;
    const int nf = nf_scalar(prob->test.fem) * prob->test.ncomp;
    const BlockCaps c = caps ? *caps : default_caps(nf, prob->test.ncomp == 1);
    const int rc =
        on_device ? build_setup_device(*mesh, *prob, ctx->S, c, stream) : build_setup(*mesh, *prob, ctx->S, c);
    if (rc != ANIFEM_ERR_DIMS || caps) return rc;
    ctx->S = AssemblySetup();  // the default budgets, then wider ones when one row exceeds them
    const BlockCaps w = wide_caps(nf);
    return on_device ? build_setup_device(*mesh, *prob, ctx->S, w, stream) : build_setup(*mesh, *prob, ctx->S, w);
}

}  // namespace

namespace anifem {
int assembly_create_caps(anifem_assembly** out, const anifem_mesh_desc* mesh, const anifem_problem_desc* prob,
                         const BlockCaps& caps, void* stream, const anifem_vars_desc* vars = nullptr);
}

namespace {
// The default budgets, then wider ones when a single row exceeds them.
int create_default(anifem_assembly** out, const anifem_mesh_desc* mesh, const anifem_problem_desc* prob, int nf,
                   void* stream, const anifem_vars_desc* vars) {
    using namespace anifem;
    const bool scalar_space = !vars && prob->test.ncomp == 1;
    const int rc = assembly_create_caps(out, mesh, prob, default_caps(nf, scalar_space), stream, vars);
    if (rc != ANIFEM_ERR_DIMS) return rc;
    return assembly_create_caps(out, mesh, prob, wide_caps(nf), stream, vars);
}
}  // namespace

extern "C" int anifem_assembly_create_vars(anifem_assembly** out, const anifem_mesh_desc* mesh,
                                           const anifem_vars_desc* vars, const anifem_problem_desc* prob,
                                           void* stream) {
    using namespace anifem;
    if (!vars || vars->n_var < 1 || vars->n_var > kMaxVars) return ANIFEM_ERR_ARG;
    int nf = 0;
    for (int v = 0; v < vars->n_var; ++v) {
        if (!space_layout(vars->var[v].fem).valid || vars->var[v].ncomp < 1) return ANIFEM_ERR_ARG;
        nf += nf_scalar(vars->var[v].fem) * vars->var[v].ncomp;
    }
    return create_default(out, mesh, prob, nf, stream, vars);
}

extern "C" int anifem_assembly_create(anifem_assembly** out, const anifem_mesh_desc* mesh,
                                      const anifem_problem_desc* prob, void* stream) {
    if (!prob || !anifem::op_valid(prob->test) || !anifem::space_layout(prob->test.fem).valid) return ANIFEM_ERR_ARG;
    const int nf = anifem::nf_scalar(prob->test.fem) * prob->test.ncomp;
    return create_default(out, mesh, prob, nf, stream, nullptr);
}

extern "C" int anifem_assembly_create_budget(anifem_assembly** out, const anifem_mesh_desc* mesh,
                                             const anifem_problem_desc* prob, int max_slots, int max_values,
                                             int max_pairs, void* stream) {
    anifem::BlockCaps c;
    c.max_slots = max_slots;
    c.max_values = max_values;
    c.max_pairs = max_pairs;
    return anifem::assembly_create_caps(out, mesh, prob, c, stream);
}

// Assembly context with given row-block budgets (the batch autotuner's trials).
int anifem::assembly_create_caps(anifem_assembly** out, const anifem_mesh_desc* mesh,
                                 const anifem_problem_desc* prob, const BlockCaps& caps, void* stream,
                                 const anifem_vars_desc* vars) {
    using namespace anifem;
    if (!out) return ANIFEM_ERR_ARG;
    std::unique_ptr<anifem_assembly> ctx(new anifem_assembly());
    int rc = build_host(ctx.get(), mesh, prob, &caps, true, stream, vars);
    if (rc) return rc;
    AssemblySetup& S = ctx->S;
    hipStream_t s = static_cast<hipStream_t>(stream);
    ANIFEM_HIP_CHECK(hipGetDevice(&ctx->device));
    // set-up for anifem_assembly_run_noalloc on this device: the scratch
    // array's address, and the reference tensors of the problem's pairs
    // (built here, so that the run only looks them up; a pair the run would
    // refuse is left for the run to report)
    {
        void* sp = nullptr;
        if ((rc = scratch_address(ctx->device, &sp))) return rc;
        RefTensorDev R;
        std::vector<double> pts, wts;
        const bool pp = prob->tensor_storage == ANIFEM_COEF_PER_POINT;
        if (!vars && rule_points(prob->order, -1, pts, wts) > 0)
            (void)get_ref_tensor(prob->trial.op, prob->trial.fem, prob->test.op, prob->test.fem, prob->order, pp, &R);
        const bool ppf = prob->rhs_tensor_storage == ANIFEM_COEF_PER_POINT;
        if (!vars && rule_points(prob->rhs_order, -1, pts, wts) > 0)
            (void)get_ref_tensor(ANIFEM_IDEN, ANIFEM_FEM_P0, ANIFEM_IDEN, prob->test.fem, prob->rhs_order, ppf, &R);
    }
    const int64_t NT = S.n_tets;
    std::vector<int32_t> tn(4 * NT);
    for (int64_t t = 0; t < NT; ++t)
        for (int v = 0; v < 4; ++v) tn[(size_t)v * NT + t] = S.tets[4 * (size_t)t + v];
    if ((rc = upload(&ctx->d_coords, mesh->coords, 3 * (size_t)S.n_nodes, s))) return rc;
    if ((rc = upload(&ctx->d_tetnodes, tn.data(), tn.size(), s))) return rc;
    if ((rc = upload(&ctx->d_blk_row0, S.blk_row0.data(), S.blk_row0.size(), s))) return rc;
    if ((rc = upload(&ctx->d_blk_meta, S.blk_meta.data(), S.blk_meta.size(), s))) return rc;
    if ((rc = upload(&ctx->d_blk_s0, S.blk_s0.data(), S.blk_s0.size(), s))) return rc;
    if ((rc = upload(&ctx->d_blk_p0, S.blk_p0.data(), S.blk_p0.size(), s))) return rc;
    if ((rc = upload(&ctx->d_blk_v0, S.blk_v0.data(), S.blk_v0.size(), s))) return rc;
    if ((rc = upload(&ctx->d_pair_tet, S.pair_tet.data(), S.pair_tet.size(), s))) return rc;
    if ((rc = upload(&ctx->d_pair_ib, S.pair_ib.data(), S.pair_ib.size(), s))) return rc;
    if ((rc = upload(&ctx->d_pair_pos, S.pair_pos.data(), S.pair_pos.size(), s))) return rc;
    {
        std::vector<uint32_t> p8;
        ctx->pos8_words = compact_positions(S, p8);
        if (ctx->pos8_words && (rc = upload(&ctx->d_pair_pos8, p8.data(), p8.size(), s))) return rc;
        ANIFEM_HIP_CHECK(hipStreamSynchronize(s));  // p8 is a host temporary
    }
    // one entry of padding: the persistent pass reads slot starts as aligned
    // 32-bit words, the last of which may reach one entry past the array
    S.sstart.push_back(0);
    if ((rc = upload(&ctx->d_sstart, S.sstart.data(), S.sstart.size(), s))) return rc;
    if ((rc = upload(&ctx->d_pstart, S.pstart.data(), S.pstart.size(), s))) return rc;
    if ((rc = upload(&ctx->d_row_ptr, S.row_ptr.data(), S.row_ptr.size(), s))) return rc;
    if ((rc = upload(&ctx->d_col_idx, S.col_idx.data(), S.col_idx.size(), s))) return rc;
    ANIFEM_HIP_CHECK(hipMalloc((void**)&ctx->d_status, 4 * sizeof(int)));
    if (S.n_tets > 0)
        ANIFEM_HIP_CHECK(hipMalloc((void**)&ctx->d_deg, (size_t)((S.n_tets + kBlock - 1) / kBlock) * sizeof(int)));
    // spaces whose elements need the Piola map and per-DOF factors (RT0,
    // ND0) or have one DOF (P0, B4) are assembled from element matrices
    ctx->elem_path = vars || fem_vector(prob->trial.fem) || (S.nf & 1);
    if (!S.sign.empty()) {
        if ((rc = upload(&ctx->d_sign, S.sign.data(), S.sign.size(), s))) return rc;
        std::vector<uint64_t> mask((size_t)S.n_tets, 0);
        for (int64_t t = 0; t < S.n_tets; ++t)
            for (int j = 0; j < S.nf; ++j)
                if (S.sign[(size_t)t * S.nf + j] < 0) mask[(size_t)t] |= uint64_t(1) << j;
        if ((rc = upload(&ctx->d_signmask, mask.data(), mask.size(), s))) return rc;
        ANIFEM_HIP_CHECK(hipStreamSynchronize(s));  // mask is a host temporary
    }
    // rows an inhomogeneous Dirichlet condition changes: Dirichlet rows and
    // rows with a dropped (Dirichlet-column) entry in one of their pairs
    {
        std::vector<LiftRow> lift;
        const int nf = S.nf;
        for (int64_t b = 0; b < S.n_blocks; ++b) {
            const int32_t r0 = S.blk_row0[b], r1 = S.blk_row0[b + 1];
            for (int32_t r = r0; r < r1; ++r) {
                const int64_t pb = S.blk_p0[b] + S.pstart[r];
                const int64_t pe = r + 1 < r1 ? S.blk_p0[b] + S.pstart[r + 1] : S.blk_p0[b + 1];
                bool lifted = S.is_dir[r] != 0;
                for (int64_t p = pb; p < pe && !lifted; ++p) {
                    const int32_t* td = &S.dofs[(size_t)S.pair_tet[p] * nf];
                    for (int ia = 0; ia < nf && !lifted; ++ia)
                        lifted = S.pair_pos[(size_t)p * nf + ia] == kPosDropped && S.is_dir[td[ia]];
                }
                if (lifted) lift.push_back(LiftRow{pb, r, (int32_t)(pe - pb)});
            }
        }
        ctx->n_lift = (int64_t)lift.size();
        if ((rc = upload(&ctx->d_lift, lift.data(), lift.size(), s))) return rc;
        if (ctx->n_lift && (rc = upload(&ctx->d_dofs, S.dofs.data(), S.dofs.size(), s))) return rc;
        if (ctx->n_lift && (rc = upload(&ctx->d_isdir, S.is_dir.data(), S.is_dir.size(), s))) return rc;
    }
    ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
    // the device holds the pair lists; the host keeps the rest for inspection
    std::vector<int32_t>().swap(S.pair_tet);
    std::vector<uint8_t>().swap(S.pair_ib);
    std::vector<uint16_t>().swap(S.pair_pos);
    std::vector<uint16_t>().swap(S.sstart);
    std::vector<uint16_t>().swap(S.pstart);
    *out = ctx.release();
    return ANIFEM_OK;
}

extern "C" {

int anifem_assembly_create_host(anifem_assembly** out, const anifem_mesh_desc* mesh, const anifem_problem_desc* prob) {
    if (!out) return ANIFEM_ERR_ARG;
    std::unique_ptr<anifem_assembly> ctx(new anifem_assembly());
    ctx->host_only = true;
    int rc = build_host(ctx.get(), mesh, prob);
    if (rc) return rc;
    *out = ctx.release();
    return ANIFEM_OK;
}

int anifem_assembly_create_vars_host(anifem_assembly** out, const anifem_mesh_desc* mesh,
                                     const anifem_vars_desc* vars, const anifem_problem_desc* prob) {
    if (!out || !vars) return ANIFEM_ERR_ARG;
    std::unique_ptr<anifem_assembly> ctx(new anifem_assembly());
    ctx->host_only = true;
    int rc = build_host(ctx.get(), mesh, prob, nullptr, false, nullptr, vars);
    if (rc) return rc;
    *out = ctx.release();
    return ANIFEM_OK;
}

int anifem_assembly_destroy(anifem_assembly* ctx) {
    if (!ctx) return ANIFEM_OK;
    for (hipEvent_t e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (!ctx->host_only) {
        void* bufs[] = {ctx->d_coords,   ctx->d_tetnodes, ctx->d_blk_row0, ctx->d_blk_meta, ctx->d_blk_s0, ctx->d_blk_p0,
                        ctx->d_blk_v0,   ctx->d_pair_tet, ctx->d_pair_ib,  ctx->d_pair_pos, ctx->d_sstart,
                        ctx->d_pstart,   ctx->d_row_ptr,  ctx->d_col_idx,  ctx->d_status,   ctx->d_kbuf,
                        ctx->d_slot,     ctx->d_dofs,     ctx->d_isdir,    ctx->d_plain_tab, ctx->d_col_tets,
                        ctx->d_lift,     ctx->d_pair_pos8, ctx->d_sign,   ctx->d_xyz,      ctx->d_ebuf,
                        ctx->d_fbuf,     ctx->d_signmask, ctx->d_deg};
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

int anifem_assembly_copy_signs(const anifem_assembly* ctx, int8_t* sign) {
    if (!ctx || !sign) return ANIFEM_ERR_ARG;
    const anifem::AssemblySetup& S = ctx->S;
    if (S.sign.empty()) std::memset(sign, 1, (size_t)S.n_tets * S.nf);
    else std::memcpy(sign, S.sign.data(), S.sign.size());
    return ANIFEM_OK;
}

int anifem_assembly_blocks(const anifem_assembly* ctx, anifem_block_layout* out) {
    if (!ctx || !out) return ANIFEM_ERR_ARG;
    const anifem::AssemblySetup& S = ctx->S;
    out->n_blocks = S.n_blocks;
    out->n_pairs = S.blk_p0.empty() ? 0 : S.blk_p0.back();
    out->n_values = S.blk_v0.empty() ? 0 : S.blk_v0.back();
    out->max_block_pairs = S.max_block_pairs;
    out->max_block_values = S.max_block_values;
    out->max_block_slots = S.max_block_slots;
    return ANIFEM_OK;
}

int anifem_assembly_copy_blocks(const anifem_assembly* ctx, int32_t* blk_row0, int64_t* blk_p0, int64_t* blk_v0,
                                int32_t* pair_tet, uint8_t* pair_ib, uint16_t* pair_pos, uint16_t* sstart,
                                uint16_t* pstart) {
    if (!ctx) return ANIFEM_ERR_ARG;
    const anifem::AssemblySetup& S = ctx->S;
    // contexts built for the device keep only the block offsets on the host
    if (!ctx->host_only && (pair_tet || pair_ib || pair_pos || sstart || pstart)) return ANIFEM_ERR_ARG;
    auto cp = [](auto* dst, const auto& v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(blk_row0, S.blk_row0);
    cp(blk_p0, S.blk_p0);
    cp(blk_v0, S.blk_v0);
    cp(pair_tet, S.pair_tet);
    cp(pair_ib, S.pair_ib);
    cp(pair_pos, S.pair_pos);
    cp(sstart, S.sstart);
    cp(pstart, S.pstart);
    return ANIFEM_OK;
}

int anifem_assembly_set_scatter(anifem_assembly* ctx, int method) {
    using namespace anifem;
    if (!ctx || ctx->host_only || (ctx->vars_only && method != ANIFEM_SCATTER_ROW_BLOCKS)) return ANIFEM_ERR_ARG;
    if (method == ANIFEM_SCATTER_ROW_BLOCKS) {
        ctx->coloured = false;
        return ANIFEM_OK;
    }
    if (method != ANIFEM_SCATTER_COLOURED) return ANIFEM_ERR_ARG;
    if (!ctx->d_col_tets) {
        const AssemblySetup& S = ctx->S;
        std::vector<int32_t> colour;
        const int nc = greedy_colour(S.n_nodes, S.n_tets, S.tets.data(), colour);
        if (nc < 0) return ANIFEM_ERR_NOMEM;
        ctx->colour_ptr.assign((size_t)nc + 1, 0);
        for (int64_t t = 0; t < S.n_tets; ++t) ctx->colour_ptr[(size_t)colour[t] + 1]++;
        for (int c = 0; c < nc; ++c) ctx->colour_ptr[c + 1] += ctx->colour_ptr[c];
        std::vector<int64_t> fillp(ctx->colour_ptr.begin(), ctx->colour_ptr.end() - 1);
        std::vector<int32_t> ct((size_t)S.n_tets);
        for (int64_t t = 0; t < S.n_tets; ++t) ct[(size_t)fillp[(size_t)colour[t]]++] = (int32_t)t;
        hipStream_t s = nullptr;
        int rc;
        if ((rc = upload_maps(ctx, s))) return rc;
        if ((rc = upload(&ctx->d_col_tets, ct.data(), ct.size(), s))) return rc;
        ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
    }
    ctx->coloured = true;
    return ANIFEM_OK;
}

int anifem_assembly_colours(const anifem_assembly* ctx, int32_t* n_colours, int32_t* colour) {
    if (!ctx || !n_colours) return ANIFEM_ERR_ARG;
    if (ctx->colour_ptr.empty()) return ANIFEM_ERR_ARG;  // anifem_assembly_set_scatter(COLOURED) first
    *n_colours = (int32_t)ctx->colour_ptr.size() - 1;
    if (colour) {
        std::vector<int32_t> ct((size_t)ctx->S.n_tets);
        ANIFEM_HIP_CHECK(hipMemcpy(ct.data(), ctx->d_col_tets, ct.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (size_t c = 0; c + 1 < ctx->colour_ptr.size(); ++c)
            for (int64_t k = ctx->colour_ptr[c]; k < ctx->colour_ptr[c + 1]; ++k) colour[ct[(size_t)k]] = (int32_t)c;
    }
    return ANIFEM_OK;
}

int anifem_assembly_set_persistent(anifem_assembly* ctx, int on) {
    if (!ctx || on < 0 || on > 2) return ANIFEM_ERR_ARG;
    ctx->persistent = on != 0;
    ctx->fixed_nf = on == 1;
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

int anifem_assembly_quadrature_points(const anifem_assembly* ctx, int order, double* X, void* stream) {
    if (!ctx || ctx->host_only) return ANIFEM_ERR_ARG;
    return anifem::launch_physical_points(order, ctx->S.n_tets, nullptr, ctx->d_coords, ctx->d_tetnodes,
                                          ctx->S.n_tets, X, static_cast<hipStream_t>(stream));
}

int anifem_assembly_coords(anifem_assembly* ctx, double** coords) {
    if (!ctx || ctx->host_only || !coords) return ANIFEM_ERR_ARG;
    *coords = ctx->d_coords;
    return ANIFEM_OK;
}

}  // extern "C"

namespace {

// The cells' vertices as the 12 x n SoA of the element batch.
__global__ void gather_xyz_kernel(const double* __restrict__ coords, const int32_t* __restrict__ tetnodes, int64_t n,
                                  double* __restrict__ xyz) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int v = 0; v < 4; ++v) {
        const int64_t node = tetnodes[v * n + r];
        for (int c = 0; c < 3; ++c) xyz[(int64_t)(3 * v + c) * n + r] = coords[3 * node + c];
    }
}

template <typename T>
int grow(void** buf, size_t* have, size_t need) {
    if (*have >= need) return ANIFEM_OK;
    if (*buf) ANIFEM_HIP_CHECK(hipFree(*buf));
    *buf = nullptr;
    *have = 0;
    ANIFEM_HIP_CHECK(hipMalloc(buf, need));
    *have = need;
    return ANIFEM_OK;
}

// Assembly through element matrices, for the spaces whose elements the
// fused pass does not form (RT0 / ND0: Piola map, per-DOF factors and
// orientation signs; one-DOF spaces): the element batch (anifem_fem3dtet_batch,
// the reference's runtime fem3Dtet over any operator pair, int_tet.inl:2-27)
// on the reordered cells, then the row-block scatter of element matrices
// (anifem_assembly_run_elements, which applies the signs of oriented DOFs)
// with element-level Dirichlet.
int run_via_elements(anifem_assembly* ctx, const anifem_tensor* D, const anifem_tensor* F, int precision,
                     double* csr_val, double* rhs, anifem_status* status, hipStream_t s) {
    using namespace anifem;
    const AssemblySetup& S = ctx->S;
    const int64_t n = S.n_tets;
    const int nf = S.nf;
    const size_t esz = precision == ANIFEM_PREC_FP64 ? 8 : 4;
    int rc;
    size_t xyz_bytes = ctx->d_xyz ? (size_t)12 * n * sizeof(double) : 0;
    void* xp = ctx->d_xyz;
    if ((rc = grow<double>(&xp, &xyz_bytes, (size_t)12 * n * sizeof(double)))) return rc;
    ctx->d_xyz = static_cast<double*>(xp);
    if ((rc = grow<double>(&ctx->d_ebuf, &ctx->ebuf_bytes, (size_t)n * nf * nf * esz))) return rc;
    const bool do_rhs = F && rhs;
    if (do_rhs && (rc = grow<double>(&ctx->d_fbuf, &ctx->fbuf_bytes, (size_t)n * nf * esz))) return rc;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(gather_xyz_kernel, dim3(blocks), dim3(256), 0, s, ctx->d_coords, ctx->d_tetnodes, n,
                       ctx->d_xyz);
    ANIFEM_HIP_CHECK(hipGetLastError());
    if ((rc = anifem_fem3dtet_batch(ctx->prob.trial, ctx->prob.test, D, ctx->prob.order, n, ctx->d_xyz, precision,
                                    ctx->d_ebuf, s)))
        return rc;
    if (do_rhs) {
        const anifem_op p0{ANIFEM_IDEN, ANIFEM_FEM_P0, 1};
        const anifem_op tb{ANIFEM_IDEN, ctx->prob.test.fem, ctx->prob.test.ncomp};
        if ((rc = anifem_fem3dtet_batch(p0, tb, F, ctx->prob.rhs_order, n, ctx->d_xyz, precision, ctx->d_fbuf, s)))
            return rc;
    }
    return anifem_assembly_run_elements(ctx, ctx->d_ebuf, do_rhs ? ctx->d_fbuf : nullptr, precision, csr_val, rhs,
                                        status, s);
}

// anifem_assembly_run, and its no-allocation form (scratch: this device's
// scratch array, which then holds the geometry rows).
int assembly_run(anifem_assembly* ctx, const anifem_tensor* D, const anifem_tensor* F, const double* dir_values,
                 int precision, double* csr_val, double* rhs, anifem_status* status, void* stream, void* scratch) {
    using namespace anifem;
    if (!ctx || ctx->host_only || ctx->vars_only || !D || !csr_val) return ANIFEM_ERR_ARG;
    // Dirichlet values are applied by the row-block scatter (the coloured
    // scatter resolves homogeneous conditions only)
    if (dir_values && ctx->coloured) return ANIFEM_ERR_ARG;
    if (precision != ANIFEM_PREC_FP64 && precision != ANIFEM_PREC_FP32) return ANIFEM_ERR_ARG;
    if (D->type != ctx->prob.tensor_type || D->storage != ctx->prob.tensor_storage) return ANIFEM_ERR_ARG;
    const bool do_rhs = F && rhs;
    if (do_rhs && (F->type != ctx->prob.rhs_tensor_type || F->storage != ctx->prob.rhs_tensor_storage))
        return ANIFEM_ERR_ARG;
    if (ctx->elem_path) {
        // element matrices in buffers of the context: no Dirichlet values, no
        // no-allocation form, no coloured scatter
        if (dir_values || scratch || ctx->coloured) return ANIFEM_ERR_ARG;
        return run_via_elements(ctx, D, F, precision, csr_val, rhs, status, static_cast<hipStream_t>(stream));
    }
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
    if (do_rhs) {
        anifem_op p0{ANIFEM_IDEN, ANIFEM_FEM_P0, 1};
        anifem_op test_iden{ANIFEM_IDEN, ctx->prob.test.fem, ctx->prob.test.ncomp};
        if ((rc = prepare_elem_args(p0, test_iden, F, ctx->prob.rhs_order, a.f))) return rc;
        a.f.coef_n = S.n_tets;
    }
    a.has_rhs = do_rhs ? 1 : 0;
    a.n_kinds = 1;
    // K symmetric: the same 3-dimensional operator of one scalar space on both
    // sides (same selector, no Piola map) and a symmetric tensor constant over
    // each tet
    const anifem_op& tr = ctx->prob.trial;
    const anifem_op& te = ctx->prob.test;
    const bool sym = tr.op == te.op && tr.fem == te.fem && tr.ncomp == 1 && te.ncomp == 1 && a.m.Rds &&
                     a.m.q == 1 && ref_dim(tr.op, tr.fem) == 3 && !(a.m.mapA | a.m.mapB | a.m.scA | a.m.scB) &&
                     (D->type == ANIFEM_TENSOR_NULL || D->type == ANIFEM_TENSOR_SCALAR ||
                      D->type == ANIFEM_TENSOR_SYMMETRIC);
    a.kinds[0] = KindTable{sym ? a.m.Rds : a.m.Rd, sym ? a.m.Rfs : a.m.Rf, a.f.Rd, a.f.Rf, a.m.nfA0, a.m.nfB0, S.nf,
                           a.m.nfA0 * a.m.nfB0, do_rhs ? a.f.nfB0 : 0, 0};
    a.max_R_rows = a.kinds[0].R_rows;
    a.max_RF_rows = a.kinds[0].RF_rows;
    a.n_blocks = S.n_blocks;
    a.meta = ctx->d_blk_meta;
    a.blk_row0 = ctx->d_blk_row0;
    a.blk_s0 = ctx->d_blk_s0;
    a.blk_p0 = ctx->d_blk_p0;
    a.blk_v0 = ctx->d_blk_v0;
    a.pair_tet = ctx->d_pair_tet;
    a.pair_ib = ctx->d_pair_ib;
    a.pair_pos = ctx->d_pair_pos;
    a.pair_pos8 = ctx->d_pair_pos8;
    a.sstart = ctx->d_sstart;
    a.pstart = ctx->d_pstart;
    a.signmask = ctx->d_signmask;
    a.val = csr_val;
    a.rhs = rhs;
    a.status = ctx->d_status;
    a.max_values = (int)S.max_block_values;
    a.max_pairs = (int)S.max_block_pairs;
    a.max_slots = (int)S.max_block_slots;

    if (ctx->timers_on) ANIFEM_HIP_CHECK(hipEventRecord(ctx->ev[0], s));
    // a rhs without a load vector is zero
    if (rhs && !do_rhs && (rc = fill(rhs, S.ndof, s))) return rc;
    // the geometry launch clears the non-finite count and stores a degenerate
    // count per workgroup (summed below when the status is asked for)
    const bool per_wg = ctx->d_deg != nullptr;
    if (per_wg) {
        a.deg = ctx->d_deg;
    } else {
        hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, ctx->d_status, 4);
        ANIFEM_HIP_CHECK(hipGetLastError());
    }
    ctx->run_bc = dir_values;
    ctx->run_scratch = scratch;
    rc = precision == ANIFEM_PREC_FP64 ? dispatch_blocks<double>(ctx, a, sym, s) : dispatch_blocks<float>(ctx, a, sym, s);
    ctx->run_bc = nullptr;
    ctx->run_scratch = nullptr;
    ctx->run_kg = nullptr;
    if (rc) return rc;
    if (ctx->timers_on) {
        ANIFEM_HIP_CHECK(hipEventRecord(ctx->ev[2], s));
        ANIFEM_HIP_CHECK(hipEventSynchronize(ctx->ev[2]));
        float t01 = 0, t12 = 0;
        ANIFEM_HIP_CHECK(hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]));
        ANIFEM_HIP_CHECK(hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]));
        ctx->timers.ms_geometry = t01;
        ctx->timers.ms_gather = t12;
        ctx->timers.n_launches = (per_wg ? 1 : 2) + (rhs && !do_rhs ? 1 : 0) + (S.n_blocks > 0 ? 1 : 0);
        ctx->timers.reserved = 0;
    }
    if (status) {
        if (per_wg) {
            hipLaunchKernelGGL(sum_counts_kernel, dim3(1), dim3(kBlock), 0, s, ctx->d_deg,
                               (S.n_tets + kBlock - 1) / kBlock, ctx->d_status + 1);
            ANIFEM_HIP_CHECK(hipGetLastError());
        }
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

}  // namespace

extern "C" {

int anifem_assembly_run(anifem_assembly* ctx, const anifem_tensor* D, const anifem_tensor* F,
                        const double* dir_values, int precision, double* csr_val, double* rhs,
                        anifem_status* status, void* stream) {
    return assembly_run(ctx, D, F, dir_values, precision, csr_val, rhs, status, stream, nullptr);
}

int anifem_assembly_run_noalloc(anifem_assembly* ctx, const anifem_tensor* D, const anifem_tensor* F,
                                const double* dir_values, int precision, double* csr_val, double* rhs,
                                anifem_status* status, void* stream) {
    using namespace anifem;
    if (!ctx || ctx->host_only) return ANIFEM_ERR_ARG;
    // the coloured scatter builds device tables on first use
    if (ctx->coloured) return ANIFEM_ERR_ARG;
    void* sp = nullptr;
    int rc = scratch_address(ctx->device, &sp);  // looked up at create: no HIP call here
    if (rc) return rc;
    g_ref_lookup_only = true;
    rc = assembly_run(ctx, D, F, dir_values, precision, csr_val, rhs, status, stream, sp);
    g_ref_lookup_only = false;
    return rc;
}

int anifem_assembly_run_rhs(anifem_assembly* ctx, const void* F_elem, int elem_precision, double* rhs, int accumulate,
                            anifem_status* status, void* stream) {
    using namespace anifem;
    if (!ctx || ctx->host_only || !F_elem || !rhs) return ANIFEM_ERR_ARG;
    if (elem_precision != ANIFEM_PREC_FP64 && elem_precision != ANIFEM_PREC_FP32) return ANIFEM_ERR_ARG;
    const AssemblySetup& S = ctx->S;
    hipStream_t s = static_cast<hipStream_t>(stream);
    BlockArgs a{};
    a.has_rhs = 1;
    a.n_kinds = 1;
    a.kinds[0] = KindTable{nullptr, nullptr, nullptr, nullptr, 0, 0, S.nf, 0, 0, 0};
    a.n_blocks = S.n_blocks;
    a.meta = ctx->d_blk_meta;
    a.pair_tet = ctx->d_pair_tet;
    a.pair_ib = ctx->d_pair_ib;
    a.pstart = ctx->d_pstart;
    a.signmask = ctx->d_signmask;
    a.rhs = rhs;
    a.status = ctx->d_status;
    if (S.max_block_pairs > kBlock) return ANIFEM_ERR_DIMS;
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, ctx->d_status, 4);
    ANIFEM_HIP_CHECK(hipGetLastError());
    if (S.n_blocks > 0) {
        const unsigned grid = (unsigned)(8 * ((S.n_blocks + 7) / 8));
        if (elem_precision == ANIFEM_PREC_FP64)
            hipLaunchKernelGGL(row_block_rhs_kernel<double>, dim3(grid), dim3(kBlock), 0, s, a,
                               static_cast<const double*>(F_elem), accumulate ? 1 : 0);
        else
            hipLaunchKernelGGL(row_block_rhs_kernel<float>, dim3(grid), dim3(kBlock), 0, s, a,
                               static_cast<const float*>(F_elem), accumulate ? 1 : 0);
        ANIFEM_HIP_CHECK(hipGetLastError());
    }
    if (status) {
        int h[4] = {0, 0, 0, 0};
        ANIFEM_HIP_CHECK(hipMemcpyAsync(h, ctx->d_status, sizeof(h), hipMemcpyDeviceToHost, s));
        ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
        status->n_nonfinite = h[0];
        status->n_degenerate = 0;
        status->code = h[0] ? ANIFEM_ERR_NONFINITE : ANIFEM_OK;
        status->reserved = 0;
        return status->code;
    }
    return ANIFEM_OK;
}

int anifem_assembly_eval(anifem_assembly* ctx, int op, const double* xyl, int q, const double* u, double* opU,
                         void* stream) {
    using namespace anifem;
    if (!ctx || ctx->host_only || ctx->vars_only || !u || !opU) return ANIFEM_ERR_ARG;
    // IDEN / GRAD, and DIV / CURL of the vector-valued spaces
    if (op != ANIFEM_IDEN && op != ANIFEM_GRAD && !(fem_vector(ctx->prob.trial.fem) && op != ANIFEM_GRAD))
        return ANIFEM_ERR_ARG;
    const AssemblySetup& S = ctx->S;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;
    if (!ctx->d_dofs && (rc = upload(&ctx->d_dofs, S.dofs.data(), S.dofs.size(), s))) return rc;
    ApplyArgs a{};
    a.n = S.n_tets;
    a.coords = ctx->d_coords;
    a.tetnodes = ctx->d_tetnodes;
    a.dofmap = ctx->d_dofs;
    a.sign = ctx->d_sign;
    a.u = u;
    a.out = opU;
    return launch_apply(anifem_op{op, ctx->prob.trial.fem, ctx->prob.trial.ncomp}, xyl, q, a, true, s);
}

int anifem_assembly_tets(const anifem_assembly* ctx, int32_t* tets) {
    if (!ctx || !tets) return ANIFEM_ERR_ARG;
    std::memcpy(tets, ctx->S.tets.data(), ctx->S.tets.size() * sizeof(int32_t));
    return ANIFEM_OK;
}

}  // extern "C"

namespace {

// Persistent element scatter for the common element sizes (occupancy x CUs
// workgroups, whole XCD groups), else one workgroup per block.
template <typename TE, int NF, bool SYM>
int launch_elements_persistent(anifem_assembly* ctx, const anifem::BlockArgs& a, const TE* A, const TE* F, size_t lds,
                               hipStream_t s) {
    using namespace anifem;
    int occ = 0, n_cu = 0;
    ANIFEM_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &occ, reinterpret_cast<const void*>(&row_block_elements_persistent<TE, NF, SYM>), kBlock, lds));
    ANIFEM_HIP_CHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    int64_t G = (int64_t)std::max(occ, 1) * std::max(n_cu, 1);
    G = std::min<int64_t>(G, ctx->S.n_blocks);
    G = std::max<int64_t>(8, G / 8 * 8);
    hipLaunchKernelGGL((row_block_elements_persistent<TE, NF, SYM>), dim3((unsigned)G), dim3(kBlock), lds, s, a, A, F);
    ANIFEM_HIP_CHECK(hipGetLastError());
    return ANIFEM_OK;
}

template <typename TE, bool SYM>
int launch_elements(anifem_assembly* ctx, const anifem::BlockArgs& a, const TE* A, const TE* F, size_t lds,
                    unsigned grid, hipStream_t s) {
    using namespace anifem;
    if (ctx->persistent) {
        switch (ctx->S.nf) {
            case 4: return launch_elements_persistent<TE, 4, SYM>(ctx, a, A, F, lds, s);
            case 10: return launch_elements_persistent<TE, 10, SYM>(ctx, a, A, F, lds, s);
            case 12: return launch_elements_persistent<TE, 12, SYM>(ctx, a, A, F, lds, s);
            case 20: return launch_elements_persistent<TE, 20, SYM>(ctx, a, A, F, lds, s);
            case 30: return launch_elements_persistent<TE, 30, SYM>(ctx, a, A, F, lds, s);
            default: break;
        }
    }
    hipLaunchKernelGGL((row_block_elements_kernel<TE, SYM>), dim3(grid), dim3(kBlock), lds, s, a, A, F);
    ANIFEM_HIP_CHECK(hipGetLastError());
    return ANIFEM_OK;
}

}  // namespace

extern "C" {

int anifem_assembly_run_elements(anifem_assembly* ctx, const void* A_elem, const void* F_elem, int elem_precision,
                                 double* csr_val, double* rhs, anifem_status* status, void* stream) {
    using namespace anifem;
    if (!ctx || ctx->host_only || !A_elem || !csr_val) return ANIFEM_ERR_ARG;
    const bool sym = (elem_precision & ANIFEM_ELEM_SYMMETRIC) != 0;
    elem_precision &= ~ANIFEM_ELEM_SYMMETRIC;
    if (elem_precision != ANIFEM_PREC_FP64 && elem_precision != ANIFEM_PREC_FP32) return ANIFEM_ERR_ARG;
    const AssemblySetup& S = ctx->S;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;
    BlockArgs a{};
    a.has_rhs = (F_elem && rhs) ? 1 : 0;
    a.n_kinds = 1;
    a.kinds[0] = KindTable{nullptr, nullptr, nullptr, nullptr, 0, 0, S.nf, 0, 0, 0};
    a.n_blocks = S.n_blocks;
    a.meta = ctx->d_blk_meta;
    a.blk_row0 = ctx->d_blk_row0;
    a.blk_s0 = ctx->d_blk_s0;
    a.blk_p0 = ctx->d_blk_p0;
    a.blk_v0 = ctx->d_blk_v0;
    a.pair_tet = ctx->d_pair_tet;
    a.pair_ib = ctx->d_pair_ib;
    a.pair_pos = ctx->d_pair_pos;
    a.pair_pos8 = ctx->d_pair_pos8;
    a.sstart = ctx->d_sstart;
    a.pstart = ctx->d_pstart;
    a.signmask = ctx->d_signmask;
    a.val = csr_val;
    a.rhs = rhs;
    a.status = ctx->d_status;
    a.max_values = (int)S.max_block_values;
    a.max_pairs = (int)S.max_block_pairs;
    a.max_slots = (int)S.max_block_slots;
    // a rhs without load vectors is zero
    if (rhs && !a.has_rhs && (rc = fill(rhs, S.ndof, s))) return rc;
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, ctx->d_status, 4);
    ANIFEM_HIP_CHECK(hipGetLastError());
    if (S.n_blocks > 0) {
        const size_t lds = block_lds_layout(a.max_values, a.max_pairs, a.max_slots, 8, 0, 0).total;
        const unsigned grid = (unsigned)(8 * ((S.n_blocks + 7) / 8));
        const double* A64 = static_cast<const double*>(A_elem);
        const double* F64 = static_cast<const double*>(F_elem);
        const float* A32 = static_cast<const float*>(A_elem);
        const float* F32 = static_cast<const float*>(F_elem);
        if (elem_precision == ANIFEM_PREC_FP64 && sym) rc = launch_elements<double, true>(ctx, a, A64, F64, lds, grid, s);
        else if (elem_precision == ANIFEM_PREC_FP64) rc = launch_elements<double, false>(ctx, a, A64, F64, lds, grid, s);
        else if (sym) rc = launch_elements<float, true>(ctx, a, A32, F32, lds, grid, s);
        else rc = launch_elements<float, false>(ctx, a, A32, F32, lds, grid, s);
        if (rc) return rc;
    }
    if (status) {
        int h[4] = {0, 0, 0, 0};
        ANIFEM_HIP_CHECK(hipMemcpyAsync(h, ctx->d_status, sizeof(h), hipMemcpyDeviceToHost, s));
        ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
        status->n_nonfinite = h[0];
        status->n_degenerate = 0;
        status->code = h[0] ? ANIFEM_ERR_NONFINITE : ANIFEM_OK;
        status->reserved = 0;
        return status->code;
    }
    return ANIFEM_OK;
}

int anifem_assembly_run_plain(anifem_assembly* ctx, const anifem_tensor* D, const anifem_tensor* F, double* csr_val,
                              double* rhs, anifem_status* status, void* stream) {
    using namespace anifem;
    if (!ctx || ctx->host_only || ctx->vars_only || !D || !csr_val) return ANIFEM_ERR_ARG;
    const anifem_problem_desc& P = ctx->prob;
    // scalar GRAD x GRAD spaces only (the C2 / C3 problems)
    if (P.trial.op != ANIFEM_GRAD || P.test.op != ANIFEM_GRAD || P.trial.ncomp != 1) return ANIFEM_ERR_ARG;
    if (D->type != P.tensor_type || D->storage != P.tensor_storage || D->type == ANIFEM_TENSOR_LAME)
        return ANIFEM_ERR_ARG;
    const bool do_rhs = F && rhs;
    if (do_rhs && (F->type != P.rhs_tensor_type || F->storage != P.rhs_tensor_storage)) return ANIFEM_ERR_ARG;
    AssemblySetup& S = ctx->S;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;
    const int nf = S.nf;
    const double *pq, *wq, *pf, *wf;
    const int q = quadrature(P.order, &pq, &wq);
    const int qf = quadrature(P.rhs_order, &pf, &wf);
    if (q < 0 || qf < 0) return ANIFEM_ERR_ORDER;
    if ((rc = upload_maps(ctx, s))) return rc;
    // the reference tables of this path are built on its first call, apart
    // from the element maps (which the coloured scatter may have uploaded)
    if (!ctx->d_plain_tab) {
        std::vector<double> g, v, tab;
        reference_table(ANIFEM_GRAD, S.fem, q, pq, g);
        reference_table(ANIFEM_IDEN, S.fem, qf, pf, v);
        tab = g;
        tab.insert(tab.end(), wq, wq + q);
        tab.insert(tab.end(), v.begin(), v.end());
        tab.insert(tab.end(), wf, wf + qf);
        if ((rc = upload(&ctx->d_plain_tab, tab.data(), tab.size(), s))) return rc;
    }
    if (!ctx->d_plain_tab || !ctx->d_slot || !ctx->d_dofs || !ctx->d_isdir) return ANIFEM_ERR_HIP;
    const double* gtab = ctx->d_plain_tab;
    const double* wqd = gtab + (size_t)q * nf * 3;
    const double* vtab = wqd + q;
    const double* wfd = vtab + (size_t)qf * nf;
    ElemArgs m{}, fa{};
    if ((rc = prepare_elem_args(P.trial, P.test, D, P.order, m))) return rc;
    m.coef_n = S.n_tets;
    m.coords = ctx->d_coords;
    m.tetnodes = ctx->d_tetnodes;
    m.nodes_stride = S.n_tets;
    m.pos0 = 0;
    // the plain method evaluates D at every quadrature point
    m.q = D->storage == ANIFEM_COEF_PER_POINT ? q : 1;
    if (do_rhs) {
        if ((rc = prepare_elem_args(anifem_op{ANIFEM_IDEN, ANIFEM_FEM_P0, 1}, anifem_op{ANIFEM_IDEN, S.fem, 1}, F,
                                    P.rhs_order, fa)))
            return rc;
        fa.coef_n = S.n_tets;
        fa.q = F->storage == ANIFEM_COEF_PER_POINT ? qf : 1;
    }
    if (ctx->timers_on) ANIFEM_HIP_CHECK(hipEventRecord(ctx->ev[0], s));
    if ((rc = fill(csr_val, S.nnz, s))) return rc;
    if (rhs && (rc = fill(rhs, S.ndof, s))) return rc;
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, ctx->d_status, 4);
    if (ctx->timers_on) ANIFEM_HIP_CHECK(hipEventRecord(ctx->ev[1], s));
    const unsigned blocks = (unsigned)((S.n_tets + 127) / 128);
    // tensor_at reads per-point tensors at index qn; constant / per-tet ignore it
    const int qpts = q;
    switch (nf) {
        case 4:
            hipLaunchKernelGGL(plain_kernel<4>, dim3(blocks), dim3(128), 0, s, m, fa, do_rhs ? 1 : 0, gtab, wqd,
                               qpts, vtab, wfd, qf, ctx->d_slot, ctx->d_dofs, ctx->d_isdir, csr_val, rhs, S.n_tets,
                               ctx->d_status);
            break;
        case 10:
            hipLaunchKernelGGL(plain_kernel<10>, dim3(blocks), dim3(128), 0, s, m, fa, do_rhs ? 1 : 0, gtab, wqd,
                               qpts, vtab, wfd, qf, ctx->d_slot, ctx->d_dofs, ctx->d_isdir, csr_val, rhs, S.n_tets,
                               ctx->d_status);
            break;
        case 20:
            hipLaunchKernelGGL(plain_kernel<20>, dim3(blocks), dim3(128), 0, s, m, fa, do_rhs ? 1 : 0, gtab, wqd,
                               qpts, vtab, wfd, qf, ctx->d_slot, ctx->d_dofs, ctx->d_isdir, csr_val, rhs, S.n_tets,
                               ctx->d_status);
            break;
        default:
            return ANIFEM_ERR_ARG;
    }
    ANIFEM_HIP_CHECK(hipGetLastError());
    if (ctx->timers_on) {
        ANIFEM_HIP_CHECK(hipEventRecord(ctx->ev[2], s));
        ANIFEM_HIP_CHECK(hipEventSynchronize(ctx->ev[2]));
        float t01 = 0, t12 = 0;
        ANIFEM_HIP_CHECK(hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]));
        ANIFEM_HIP_CHECK(hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]));
        ctx->timers.ms_geometry = t01;
        ctx->timers.ms_gather = t12;
        ctx->timers.n_launches = rhs ? 4 : 3;
        ctx->timers.reserved = 0;
    }
    if (status) {
        int h[4] = {0, 0, 0, 0};
        ANIFEM_HIP_CHECK(hipMemcpyAsync(h, ctx->d_status, sizeof(h), hipMemcpyDeviceToHost, s));
        ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
        status->n_nonfinite = h[0];
        status->n_degenerate = 0;
        status->code = h[0] ? ANIFEM_ERR_NONFINITE : ANIFEM_OK;
        status->reserved = 0;
        return status->code;
    }
    return ANIFEM_OK;
}

}  // extern "C"
