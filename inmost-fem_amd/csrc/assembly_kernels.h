// Whole-mesh assembly kernels: a geometry pass over the tets, then one
// owner-computes pass over row blocks of the CSR matrix.
//
//   geometry_kernel   one lane per tet: FP64 Jacobian, inverse and |T|
//                     (core.inl:218-248) and the geometry tensors K of the
//                     matrix and K_F of the load vector, stored per tet.
//   row_block_kernel  one workgroup per block of consecutive CSR rows: the
//                     block's tets' K rows, the per-slot contribution codes
//                     and the reference tensors are staged in LDS; one lane
//                     per slot forms each listed element entry
//                     sum_kk K[kk] R[rr][kk], applies element-level Dirichlet
//                     (dc_on_dof.h:38-45), adds them in ascending tet order
//                     and stores the slot value once; likewise one lane per
//                     row for the right-hand side.
// Every CSR value and rhs entry is written exactly once, by plain coalesced
// stores: no zero-fill, no read-modify-write, no atomics on values, and the
// sums are formed in a fixed order.
#pragma once

#include "element_kernels.h"

namespace anifem {

struct BlockArgs {
    ElemArgs m;   // matrix pair (geometry source, tensor, reference tensor)
    ElemArgs f;   // load-vector pair IDEN P0 x IDEN test (when has_rhs)
    int has_rhs;
    int nf;       // load-vector entries per tet
    const int32_t* blk_row0;
    const int64_t* row_ptr;
    const int64_t* blk_tet0;
    const int32_t* blk_tets;
    const int64_t* blk_c0;
    const void* codes;          // uint16 or uint32: local_tet*E + entry
    const uint16_t* sstart;     // per slot: first code, relative to its block
    const int64_t* blk_g0;
    const uint16_t* fcodes;     // local_tet*nf + i
    const uint16_t* rstart;     // per row: first load-vector code, relative to its block
    const uint64_t* dirmask;    // per tet (or null)
    double* val;
    double* rhs;
    int* status;
    int max_tets, max_contrib, max_fcontrib;
    int stage_R, stage_RF;      // reference tensors copied to LDS
    int R_len, RF_len;          // elements of the reference tensors
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

// Gather of per-tet rows of `stride` values: dst[t*stride + j] =
// src[tets[t]*stride + j] for t < nt (nt*stride < 2^16), 8 loads in flight.
template <typename V>
__device__ __forceinline__ void row_gather(V* __restrict__ dst, const V* __restrict__ src,
                                           const int32_t* __restrict__ tets, int nt, int stride) {
    constexpr int U = 8;
    const int n = nt * stride;
    const float rs = 1.0f / (float)stride;
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
            if (i < n) dst[i] = v[k];
        }
    }
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
// 8-byte aligned): dirichlet masks | K | K_F | R | R_F | tet ids | entry
// decode tables | codes | load-vector codes.
struct BlockLds {
    size_t dm, K, KF, R, RF, tets, etab, ftab, cl, fl, total;
};

__host__ __device__ inline size_t align8(size_t x) { return (x + 7) & ~(size_t)7; }

__host__ __device__ inline BlockLds block_lds_layout(int max_tets, int kstride, int kstrideF, int elem, int R_len,
                                                     int RF_len, int E, int nf, int has_rhs, int max_contrib,
                                                     int max_fcontrib, int code_bytes) {
    BlockLds L;
    size_t o = 0;
    L.dm = o;
    o = align8(o + 8 * (size_t)max_tets);
    L.K = o;
    o = align8(o + (size_t)elem * max_tets * kstride);
    L.KF = o;
    o = align8(o + (size_t)elem * max_tets * kstrideF);
    L.R = o;
    o = align8(o + (size_t)elem * R_len);
    L.RF = o;
    o = align8(o + (size_t)elem * RF_len);
    L.tets = o;
    o = align8(o + 4 * (size_t)max_tets);
    L.etab = o;
    o = align8(o + 4 * (size_t)E);
    L.ftab = o;
    o = align8(o + 4 * (size_t)(has_rhs ? nf : 0));
    L.cl = o;
    o = align8(o + (size_t)code_bytes * max_contrib);
    L.fl = o;
    o = align8(o + 2 * (size_t)max_fcontrib);
    L.total = o;
    return L;
}

template <int DA0, int DB0, int NCA, int NCB, typename T, typename C>
__global__ __launch_bounds__(kBlock) void row_block_kernel(BlockArgs a, const T* __restrict__ Kg,
                                                          const T* __restrict__ KFg) {
    constexpr int KB = DA0 * DB0;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int E = a.m.E;
    const int len = a.m.q * KB;
    const int kstride = NCA * NCB * len;
    const int lenF = a.has_rhs ? a.f.q : 0;
    const int kstrideF = NCB * lenF;
    const int nf = a.nf;
    const BlockLds L = block_lds_layout(a.max_tets, kstride, kstrideF, (int)sizeof(T), a.stage_R ? a.R_len : 0,
                                        a.stage_RF ? a.RF_len : 0, E, nf, a.has_rhs, a.max_contrib,
                                        a.max_fcontrib, (int)sizeof(C));
    uint64_t* dm = reinterpret_cast<uint64_t*>(smem_raw + L.dm);
    T* K = reinterpret_cast<T*>(smem_raw + L.K);
    T* KF = reinterpret_cast<T*>(smem_raw + L.KF);
    T* Rs = reinterpret_cast<T*>(smem_raw + L.R);
    T* RFs = reinterpret_cast<T*>(smem_raw + L.RF);
    int32_t* tl = reinterpret_cast<int32_t*>(smem_raw + L.tets);
    uint32_t* etab = reinterpret_cast<uint32_t*>(smem_raw + L.etab);
    uint32_t* ftab = reinterpret_cast<uint32_t*>(smem_raw + L.ftab);
    C* cl = reinterpret_cast<C*>(smem_raw + L.cl);
    uint16_t* fl = reinterpret_cast<uint16_t*>(smem_raw + L.fl);

    const int b = blockIdx.x;
    const int32_t r0 = a.blk_row0[b];
    const int nr = a.blk_row0[b + 1] - r0;
    const int64_t s0 = a.row_ptr[r0];
    const int ns = (int)(a.row_ptr[r0 + nr] - s0);
    const int64_t t0 = a.blk_tet0[b];
    const int nt = (int)(a.blk_tet0[b + 1] - t0);
    const int64_t c0 = a.blk_c0[b];
    const int nc = (int)(a.blk_c0[b + 1] - c0);
    const int64_t g0 = a.has_rhs ? a.blk_g0[b] : 0;
    const int ng = a.has_rhs ? (int)(a.blk_g0[b + 1] - g0) : 0;

    block_copy(tl, a.blk_tets + t0, nt);
    block_copy(cl, static_cast<const C*>(a.codes) + c0, nc);
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
    if (a.stage_R) block_copy(Rs, Rg, a.R_len);
    if (a.stage_RF) block_copy(RFs, RFg, a.RF_len);
    for (int e = threadIdx.x; e < E; e += kBlock) etab[e] = decode_entry(e, a.m.nfA0, a.m.nfB0, NCA, NCB);
    if (a.has_rhs)
        for (int i = threadIdx.x; i < nf; i += kBlock) ftab[i] = decode_entry(i, 1, a.f.nfB0, 1, NCB);
    const T* R = a.stage_R ? Rs : Rg;
    const T* RF = a.stage_RF ? RFs : RFg;
    __syncthreads();  // tet ids in LDS

    row_gather(K, Kg, tl, nt, kstride);
    if (a.has_rhs) row_gather(KF, KFg, tl, nt, kstrideF);
    for (int t = threadIdx.x; t < nt; t += kBlock) dm[t] = a.dirmask ? a.dirmask[tl[t]] : 0ull;
    __syncthreads();

    int nonfinite = 0;
    const float rE = 1.0f / (float)E;
    // UN slots per lane per sweep: their list bounds are loaded before any
    // contribution is evaluated, so the loads overlap
    constexpr int UN = 4;
    for (int base = threadIdx.x; base < ns; base += UN * kBlock) {
        int cb[UN], ce[UN];
#pragma unroll
        for (int k = 0; k < UN; ++k) {
            const int u = base + k * kBlock;
            cb[k] = u < ns ? a.sstart[s0 + u] : 0;
            ce[k] = u + 1 < ns ? a.sstart[s0 + u + 1] : nc;
        }
#pragma unroll
        for (int k = 0; k < UN; ++k) {
            const int u = base + k * kBlock;
            if (u >= ns) continue;
            double sum = 0.0;
            for (int c = cb[k]; c < ce[k]; ++c) {
                const uint32_t code = cl[c];
                int t;
                if constexpr (sizeof(C) == 2) t = div_small((int)code, E, rE);
                else t = (int)div_code(code, (uint32_t)E, rE);
                const uint32_t ent = etab[code - (uint32_t)t * E];
                const int rr = ent & 1023, blk = (ent >> 10) & 15;
                double v = (double)dot_block<KB, T>(K + (size_t)t * kstride + blk * len, R + rr * len, len);
                const uint64_t m = dm[t];
                if (m) {
                    // applyDir with value 0: zero row and column, unit diagonal
                    const int ib = (ent >> 14) & 63, ia = (ent >> 20) & 63;
                    if (((m >> ib) | (m >> ia)) & 1) v = (ib == ia) ? 1.0 : 0.0;
                }
                if (!isfinite(v)) ++nonfinite;
                sum += v;
            }
            a.val[s0 + u] = sum;
        }
    }
    if (a.has_rhs) {
        const float rnf = 1.0f / (float)nf;
        for (int u = threadIdx.x; u < nr; u += kBlock) {
            const int cb = a.rstart[r0 + u];
            const int ce = u + 1 < nr ? a.rstart[r0 + u + 1] : ng;
            double sum = 0.0;
            for (int c = cb; c < ce; ++c) {
                const int code = fl[c];
                const int t = div_small(code, nf, rnf);
                const int i = code - t * nf;
                const uint32_t ent = ftab[i];
                const int rr = ent & 1023, blk = (ent >> 10) & 15;
                double v = (double)dot_block<1, T>(KF + (size_t)t * kstrideF + blk * lenF, RF + rr * lenF, lenF);
                if ((dm[t] >> i) & 1) v = 0.0;
                if (!isfinite(v)) ++nonfinite;
                sum += v;
            }
            a.rhs[r0 + u] = sum;
        }
    }
    if (nonfinite) atomicAdd(&a.status[0], nonfinite);
}

}  // namespace anifem
