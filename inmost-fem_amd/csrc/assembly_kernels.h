// Whole-mesh assembly kernels: a geometry pass over the tets, then one
// owner-computes pass over row blocks of the CSR matrix.
//
//   geometry_kernel   one lane per tet: FP64 Jacobian, inverse and |T|
//                     (core.inl:218-248) and the geometry tensors K of the
//                     matrix and K_F of the load vector, stored per tet.
//   row_block_kernel  one workgroup per block of consecutive CSR rows: the
//                     block's tets' K rows and the reference tensors are
//                     staged in LDS; the lanes evaluate the block's listed
//                     element entries sum_kk K[kk] R[rr][kk] (with
//                     element-level Dirichlet, dc_on_dof.h:38-45) in a flat,
//                     evenly shared loop into an LDS value list; then one lane
//                     per slot adds its entries in ascending tet order and
//                     stores the slot value once; likewise one lane per row for
//                     the right-hand side.
// Every CSR value and rhs entry is written exactly once, by plain coalesced
// stores: no zero-fill, no read-modify-write, no atomics on values, and the
// sums are formed in a fixed order.
#pragma once

#include "element_kernels.h"

namespace anifem {

// codes per lane of a row-block workgroup: a block lists at most
// kBlockCodesPerLane * kBlock contributions
constexpr int kBlockCodesPerLane = 8;

struct BlockArgs {
    ElemArgs m;   // matrix pair (geometry source, tensor, reference tensor)
    ElemArgs f;   // load-vector pair IDEN P0 x IDEN test (when has_rhs)
    int has_rhs;
    int nf;       // load-vector entries per tet
    const int32_t* blk_row0;
    const int64_t* blk_s0;
    const int64_t* blk_tet0;
    const int32_t* blk_tets;
    const int64_t* blk_c0;
    const void* codes;          // uint16 or uint32: local_tet << ebits | entry (E: Dirichlet diagonal)
    const uint16_t* sstart;     // per slot: first code, relative to its block
    const int64_t* blk_g0;
    const uint16_t* fcodes;     // local_tet << fbits | i
    const uint16_t* rstart;     // per row: first load-vector code, relative to its block
    double* val;
    double* rhs;
    int* status;
    int max_tets, max_contrib, max_fcontrib, max_slots;
    int ebits, fbits;
    int stage_R, stage_RF;      // reference tensors copied to LDS
    int R_rows, RF_rows;        // rows of the reference tensors (len / lenF values each)
};

// Entry e = ib + nfB*ia decoded for the contraction, packed in 26 bits:
// reference-tensor row rr (10 bits), component block cb*NCA + ca (4 bits),
// ib (6 bits), ia (6 bits).
__device__ __forceinline__ uint32_t decode_entry(int e, int nfA0, int nfB0, int NCA, int NCB) {
    const int nfB = nfB0 * NCB;
    const int ib = e % nfB, ia = e / nfB;
    const int ib0 = ib % nfB0, cb = ib / nfB0;
    const int ia0 = ia % nfA0, ca = ia / nfA0;
    const int rr = ib0 + nfB0 * ia0;
    return (uint32_t)rr | ((uint32_t)(cb * NCA + ca) << 10) | ((uint32_t)ib << 14) | ((uint32_t)ia << 20);
}

// floor(x / d) for 0 <= x < 2^16 and 1 <= d < 2^16: float estimate, one correction
__device__ __forceinline__ int div_small(int x, int d, float rd) {
    int q = (int)((float)x * rd);
    q += (q + 1) * d <= x ? 1 : 0;
    q -= q * d > x ? 1 : 0;
    return q;
}

// floor(x / d) for quotients below 2^16 (32-bit codes): the float estimate
// is within one unit of the quotient
__device__ __forceinline__ uint32_t div_code(uint32_t x, uint32_t d, float rd) {
    uint32_t q = (uint32_t)((float)x * rd);
    while (q * d > x) --q;
    while ((q + 1) * d <= x) ++q;
    return q;
}

// sum_j k[j] r[j] over one component block: KB terms per carried point
template <int KB, typename T>
__device__ __forceinline__ T dot_block(const T* k, const T* r, int len) {
    T acc = 0;
    if (len == KB) {
#pragma unroll
        for (int j = 0; j < KB; ++j) acc += k[j] * r[j];
    } else {
        for (int j = 0; j < len; ++j) acc += k[j] * r[j];
    }
    return acc;
}

// Global -> LDS copy by the whole workgroup with 8 loads in flight per lane.
template <typename V>
__device__ __forceinline__ void block_copy(V* __restrict__ dst, const V* __restrict__ src, int n) {
    constexpr int U = 8;
    for (int base = threadIdx.x; base < n; base += U * kBlock) {
        V v[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int i = base + k * kBlock;
            v[k] = i < n ? src[i] : V(0);
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int i = base + k * kBlock;
            if (i < n) dst[i] = v[k];
        }
    }
}

// Gather of the geometry tensors of a block's tets into LDS, each row of len
// values (one component block) padded to lenp:
// dst[(t*nblk + blk)*lenp + j] = src[tets[t]*nblk*len + blk*len + j] for
// t < nt (nt*nblk*len < 2^16), 8 loads in flight per lane.
template <typename V>
__device__ __forceinline__ void k_gather(V* __restrict__ dst, const V* __restrict__ src,
                                         const int32_t* __restrict__ tets, int nt, int nblk, int len, int lenp) {
    constexpr int U = 8;
    const int stride = nblk * len;
    const int n = nt * stride;
    const float rs = 1.0f / (float)stride, rl = 1.0f / (float)len;
    for (int base = threadIdx.x; base < n; base += U * kBlock) {
        V v[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int i = base + k * kBlock;
            if (i < n) {
                const int t = div_small(i, stride, rs);
                v[k] = src[(int64_t)tets[t] * stride + (i - t * stride)];
            } else {
                v[k] = V(0);
            }
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int i = base + k * kBlock;
            if (i < n) {
                const int row = div_small(i, len, rl);
                dst[row * lenp + (i - row * len)] = v[k];
            }
        }
    }
}

// Rows of len values copied to LDS rows of lenp values.
template <typename V>
__device__ __forceinline__ void pad_copy(V* __restrict__ dst, const V* __restrict__ src, int rows, int len, int lenp) {
    const int n = rows * len;
    const float rl = 1.0f / (float)len;
    for (int i = threadIdx.x; i < n; i += kBlock) {
        const int row = div_small(i, len, rl);
        dst[row * lenp + (i - row * len)] = src[i];
    }
}

// Row length padded to a 16-byte multiple: LDS rows read as 16-byte vectors
// stay aligned.
template <typename T>
__host__ __device__ inline int pad16(int len) {
    constexpr int v = 16 / (int)sizeof(T);
    return (len + v - 1) / v * v;
}

// Geometry pass over all tets (one lane per tet): FP64 Jacobian, inverse and
// |T| (core.inl:218-248), then the geometry tensors of the matrix and of the
// load vector, stored per tet to Kg / KFg (kstride + kstrideF values per tet,
// in the T precision of the contraction).
template <int DA0, int DB0, int NCA, int NCB, typename T>
__global__ __launch_bounds__(kBlock) void geometry_kernel(BlockArgs a, T* __restrict__ Kg, T* __restrict__ KFg,
                                                         int64_t n_tets) {
    const int64_t r = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    int degenerate = 0;
    if (r < n_tets) {
        const int kstride = NCA * NCB * a.m.q * DA0 * DB0;
        const int kstrideF = a.has_rhs ? NCB * a.f.q : 0;
        double P[4][3], PSI[9], det;
        load_tet<true>(a.m, r, P);
        tet_geometry(P, PSI, det);
        degenerate = !(fabs(det) > 0.0 && isfinite(det));
        const double mes = fabs(det) / 6.0;
        geometry_tensor<DA0, DB0, NCA, NCB, T>(a.m, r, PSI, mes, Kg + r * kstride);
        if (a.has_rhs) geometry_tensor<1, 1, 1, NCB, T>(a.f, r, PSI, mes, KFg + r * kstrideF);
    }
    // one counter update per wavefront
    const unsigned long long ballot = __ballot(degenerate);
    if (ballot && (threadIdx.x & 63) == 0) atomicAdd(&a.status[1], (int)__popcll(ballot));
}

// LDS layout of one row-block workgroup, in bytes from the start (all parts
// 16-byte aligned): entry values | K | K_F | R | R_F | tet ids | entry decode
// table | load-vector decode table | load-vector codes | slot starts.
struct BlockLds {
    size_t vals, K, KF, R, RF, tets, etab, ftab, fl, sst, total;
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// kpad / kpadF: padded K / K_F values per tet; R_vals / RF_vals: padded
// values of the staged reference tensors (0 when not staged).
__host__ __device__ inline BlockLds block_lds_layout(int max_tets, int kpad, int kpadF, int elem, int R_vals,
                                                     int RF_vals, int E, int nf, int has_rhs, int max_contrib,
                                                     int max_fcontrib, int max_slots) {
    BlockLds L;
    size_t o = 0;
    L.vals = o;
    o = align16(o + 8 * (size_t)max_contrib);
    L.K = o;
    o = align16(o + (size_t)elem * max_tets * kpad);
    L.KF = o;
    o = align16(o + (size_t)elem * max_tets * kpadF);
    L.R = o;
    o = align16(o + (size_t)elem * R_vals);
    L.RF = o;
    o = align16(o + (size_t)elem * RF_vals);
    L.tets = o;
    o = align16(o + 4 * (size_t)max_tets);
    L.etab = o;
    o = align16(o + 4 * (size_t)E);
    L.ftab = o;
    o = align16(o + 4 * (size_t)(has_rhs ? nf : 0));
    L.fl = o;
    o = align16(o + 2 * (size_t)max_fcontrib);
    L.sst = o;
    o = align16(o + 2 * (size_t)max_slots);
    L.total = o;
    return L;
}

template <int DA0, int DB0, int NCA, int NCB, typename T, typename C>
__global__ __launch_bounds__(kBlock) void row_block_kernel(BlockArgs a, const T* __restrict__ Kg,
                                                          const T* __restrict__ KFg) {
    constexpr int KB = DA0 * DB0;
    constexpr int NBLK = NCA * NCB;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int E = a.m.E;
    const int len = a.m.q * KB;
    const int lenp = pad16<T>(len);
    const int lenF = a.has_rhs ? a.f.q : 0;
    const int lenFp = a.has_rhs ? pad16<T>(lenF) : 0;
    const int nf = a.nf;
    const BlockLds L = block_lds_layout(a.max_tets, NBLK * lenp, NCB * lenFp, (int)sizeof(T),
                                        a.stage_R ? a.R_rows * lenp : 0, a.stage_RF ? a.RF_rows * lenFp : 0, E, nf,
                                        a.has_rhs, a.max_contrib, a.max_fcontrib, a.max_slots);
    double* vals = reinterpret_cast<double*>(smem_raw + L.vals);
    T* K = reinterpret_cast<T*>(smem_raw + L.K);
    T* KF = reinterpret_cast<T*>(smem_raw + L.KF);
    T* Rs = reinterpret_cast<T*>(smem_raw + L.R);
    T* RFs = reinterpret_cast<T*>(smem_raw + L.RF);
    int32_t* tl = reinterpret_cast<int32_t*>(smem_raw + L.tets);
    uint32_t* etab = reinterpret_cast<uint32_t*>(smem_raw + L.etab);
    uint32_t* ftab = reinterpret_cast<uint32_t*>(smem_raw + L.ftab);
    uint16_t* fl = reinterpret_cast<uint16_t*>(smem_raw + L.fl);
    uint16_t* sst = reinterpret_cast<uint16_t*>(smem_raw + L.sst);

    const int b = blockIdx.x;
    const int32_t r0 = a.blk_row0[b];
    const int nr = a.blk_row0[b + 1] - r0;
    const int64_t s0 = a.blk_s0[b];
    const int ns = (int)(a.blk_s0[b + 1] - s0);
    const int64_t t0 = a.blk_tet0[b];
    const int nt = (int)(a.blk_tet0[b + 1] - t0);
    const int64_t c0 = a.blk_c0[b];
    const int nc = (int)(a.blk_c0[b + 1] - c0);
    const int64_t g0 = a.has_rhs ? a.blk_g0[b] : 0;
    const int ng = a.has_rhs ? (int)(a.blk_g0[b + 1] - g0) : 0;
    const C* codes = static_cast<const C*>(a.codes) + c0;

    // this lane's codes (lane + k*kBlock), requested first so that their loads
    // overlap the staging below
    constexpr int CPL = kBlockCodesPerLane;
    uint32_t code[CPL];
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
        const int c = threadIdx.x + k * kBlock;
        code[k] = c < nc ? (uint32_t)codes[c] : 0u;
    }

    block_copy(tl, a.blk_tets + t0, nt);
    block_copy(sst, a.sstart + s0, ns);
    if (a.has_rhs) block_copy(fl, a.fcodes + g0, ng);
    const T* Rg;
    const T* RFg;
    if constexpr (sizeof(T) == 8) {
        Rg = reinterpret_cast<const T*>(a.m.Rd);
        RFg = reinterpret_cast<const T*>(a.f.Rd);
    } else {
        Rg = reinterpret_cast<const T*>(a.m.Rf);
        RFg = reinterpret_cast<const T*>(a.f.Rf);
    }
    if (a.stage_R) pad_copy(Rs, Rg, a.R_rows, len, lenp);
    if (a.stage_RF) pad_copy(RFs, RFg, a.RF_rows, lenF, lenFp);
    for (int e = threadIdx.x; e < E; e += kBlock) etab[e] = decode_entry(e, a.m.nfA0, a.m.nfB0, NCA, NCB);
    if (a.has_rhs)
        for (int i = threadIdx.x; i < nf; i += kBlock) ftab[i] = decode_entry(i, 1, a.f.nfB0, 1, NCB);
    const T* R = a.stage_R ? Rs : Rg;
    const int rstride = a.stage_R ? lenp : len;
    const T* RF = a.stage_RF ? RFs : RFg;
    const int rfstride = a.stage_RF ? lenFp : lenF;
    __syncthreads();  // tet ids in LDS

    k_gather(K, Kg, tl, nt, NBLK, len, lenp);
    if (a.has_rhs) k_gather(KF, KFg, tl, nt, NCB, lenF, lenFp);
    __syncthreads();

    // phase A: every listed entry
    int nonfinite = 0;
    const uint32_t emask = (1u << a.ebits) - 1u;
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
        const int c = threadIdx.x + k * kBlock;
        if (c >= nc) break;
        const int t = (int)(code[k] >> a.ebits);
        const int e = (int)(code[k] & emask);
        double v = 1.0;  // Dirichlet diagonal: applyDir sets it to 1
        if (e != E) {
            const uint32_t ent = etab[e];
            const int rr = ent & 1023, blk = (ent >> 10) & 15;
            v = (double)dot_block<KB, T>(K + (t * NBLK + blk) * lenp, R + rr * rstride, len);
            if (!isfinite(v)) ++nonfinite;
        }
        vals[c] = v;
    }
    __syncthreads();

    // phase B: one lane per slot, entries added in list (ascending tet) order
    for (int u = threadIdx.x; u < ns; u += kBlock) {
        const int cb = sst[u];
        const int ce = u + 1 < ns ? sst[u + 1] : nc;
        double sum = 0.0;
        int c = cb;
        // loads issued four at a time, added in order
        for (; c + 4 <= ce; c += 4) {
            const double v0 = vals[c], v1 = vals[c + 1], v2 = vals[c + 2], v3 = vals[c + 3];
            sum += v0;
            sum += v1;
            sum += v2;
            sum += v3;
        }
        for (; c < ce; ++c) sum += vals[c];
        a.val[s0 + u] = sum;
    }
    if (a.has_rhs) {
        const uint32_t fmask = (1u << a.fbits) - 1u;
        for (int u = threadIdx.x; u < nr; u += kBlock) {
            const int cb = a.rstart[r0 + u];
            const int ce = u + 1 < nr ? a.rstart[r0 + u + 1] : ng;
            double sum = 0.0;
            for (int c = cb; c < ce; ++c) {
                const uint32_t code = fl[c];
                const int t = (int)(code >> a.fbits);
                const uint32_t ent = ftab[code & fmask];
                const int rr = ent & 1023, blk = (ent >> 10) & 15;
                const double v =
                    (double)dot_block<1, T>(KF + (t * NCB + blk) * lenFp, RF + rr * rfstride, lenF);
                if (!isfinite(v)) ++nonfinite;
                sum += v;
            }
            a.rhs[r0 + u] = sum;
        }
    }
    if (nonfinite) atomicAdd(&a.status[0], nonfinite);
}

}  // namespace anifem
