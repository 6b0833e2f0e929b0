// Whole-mesh assembly kernels: a geometry pass over the tets, then one
// owner-computes pass over row blocks of the CSR matrix.
//
//   geometry_kernel   one lane per tet: FP64 Jacobian, inverse and |T|
//                     (core.inl:218-248) and the geometry tensors K of the
//                     matrix and K_F of the load vector, stored as one row of
//                     rowlen values per tet.
//   row_block_kernel  one workgroup per block of consecutive CSR rows.
//                     Phase A, one lane per (row, incident tet) pair: the
//                     tet's K row is read once into registers and the pair's
//                     row of the element matrix, A(ib, ia) = sum_kk K[kk]
//                     R[rr(ib, ia)][kk], is formed against the reference
//                     tensor staged in LDS (lanes of one local row index read
//                     the same R rows), each entry stored to its place in the
//                     block's slot-ordered LDS value list; likewise the pair's
//                     load-vector entry.  Phase B, one lane per slot (row):
//                     the slot's (row's) entries are added in ascending tet
//                     order and the CSR value (rhs entry) is stored once.
// Every CSR value and rhs entry is written exactly once, by plain coalesced
// stores: no zero-fill, no read-modify-write, no atomics on values, and the
// sums are formed in a fixed order.  Element-level homogeneous Dirichlet
// (dc_on_dof.h:38-45) is resolved in the block lists (a kept entry on a
// Dirichlet row or column is a diagonal entry of value 1; a Dirichlet row's
// load-vector entries are 0).
#pragma once

#include <type_traits>

#include "element_kernels.h"
#include "setup.h"

namespace anifem {

// Element kinds of one launch: the batched assembly mixes element orders, one
// reference tensor per kind (a single-mesh assembly has one kind).
constexpr int kMaxKinds = 4;
struct KindTable {
    const double* Rd;   // reference tensor R (R_rows x len), FP64 / FP32 copies
    const float* Rf;
    const double* RFd;  // load-vector reference tensor (RF_rows x lenF)
    const float* RFf;
    int nfA0, nfB0, nf, R_rows, RF_rows, pad;
};


struct BlockArgs {
    ElemArgs m;   // matrix pair (geometry source, tensor, reference tensor)
    ElemArgs f;   // load-vector pair IDEN P0 x IDEN test (when has_rhs)
    int has_rhs;
    int rowlen;   // values per tet of the geometry rows
    int n_kinds;
    KindTable kinds[kMaxKinds];
    int64_t n_blocks;
    const BlockMeta* meta;      // per block, in processing order
    const int32_t* blk_row0;
    const int64_t* blk_s0;
    const int64_t* blk_p0;
    const int64_t* blk_v0;
    const int32_t* pair_tet;
    const uint8_t* pair_ib;     // local row index | kPairDirichlet
    const uint16_t* pair_pos;   // nf per pair: value-list position or kPosDropped
    const uint32_t* pair_pos8;  // compact positions (compact_positions), or null
    int64_t pos8_base;          // pair whose compact positions start pair_pos8 (0; batches: first pair of a kind)
    const uint16_t* sstart;     // per slot: first value, relative to its block
    const uint16_t* pstart;     // per row: first pair, relative to its block
    const uint64_t* signmask;   // element-matrix passes: per tet, bit j = local DOF j has orientation sign -1, or null
    int lamec;                  // geometry rows hold the compact Lame form (lame_kblock), not the K blocks
    double* val;
    double* rhs;
    int* status;
    // geometry_kernel: one degenerate-cell count per workgroup (and status[0]
    // cleared by the first workgroup), or null: counts added to status[1]
    int* deg;
    int max_values, max_pairs, max_slots;  // over all blocks (at most the workgroup size of pairs)
    int max_R_rows, max_RF_rows;           // over all kinds
};

// Row length padded to a 16-byte multiple: rows read as 16-byte vectors stay
// aligned.
template <typename T>
__host__ __device__ inline int pad16(int len) {
    constexpr int v = 16 / (int)sizeof(T);
    return (len + v - 1) / v * v;
}

// floor(x / d) for 0 <= x < 2^16 and 1 <= d < 2^16: float estimate, one correction
__device__ __forceinline__ int div_small(int x, int d, float rd) {
    int q = (int)((float)x * rd);
    q += (q + 1) * d <= x ? 1 : 0;
    q -= q * d > x ? 1 : 0;
    return q;
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

// Rows of len values copied to LDS rows of lenp values (rows*len < 2^16).
template <typename V>
__device__ __forceinline__ void pad_copy(V* __restrict__ dst, const V* __restrict__ src, int rows, int len, int lenp) {
    const int n = rows * len;
    const float rl = 1.0f / (float)len;
    for (int i = threadIdx.x; i < n; i += kBlock) {
        const int row = div_small(i, len, rl);
        dst[row * lenp + (i - row * len)] = src[i];
    }
}

// Row (lo, hi) of a reference tensor symmetric in its two DOF indices, lo <= hi
// < NF, in the triangle layout: the rows of column lo (hi = lo .. NF-1) start
// at lo NF - lo (lo - 1) / 2.
template <int NF>
__host__ __device__ constexpr int tri_row(int lo, int hi) {
    return lo * NF - ((lo * (lo - 1)) >> 1) + (hi - lo);
}

// The rows ib + NF ia with ib >= ia of an NF^2 x len table (the triangle of a
// table symmetric in (ia, ib)) copied to LDS rows of lenp values.
template <int NF, int LEN, typename V>
__device__ __forceinline__ void tri_copy(V* __restrict__ dst, const V* __restrict__ src, int lenp) {
    for (int i = threadIdx.x; i < NF * NF * LEN; i += kBlock) {
        const int r = i / LEN, j = i - (i / LEN) * LEN;
        const int ia = r / NF, ib = r - (r / NF) * NF;
        if (ib >= ia) dst[tri_row<NF>(ia, ib) * lenp + j] = src[i];
    }
}

// Compact rows of the Lame tensor (lin_elast.cpp:101-117) on FemVec<3, .>
// GRAD x GRAD: with C((i,j),(k,l)) = lam d_ij d_kl + mu (d_ik d_jl + d_il d_jk)
// and the gradient map T(j, a) = Tm[a + 3 j] (PSI), the sum over gradient
// entries of T^T C T collapses to
//   K_(cb,ca)(b, a) = lam T(cb,b) T(ca,a) + mu (d_{cb,ca} (T^T T)(b,a) + T(ca,b) T(cb,a)),
// lam and mu scaled by the cell measure.  A compact row holds Tm, lam, mu
// (kLameCompact values) instead of the 81 values of the nine blocks; every
// pass expands the blocks it needs with this function, with its roundings
// fixed (no contraction left to the compiler), so all passes agree bitwise.
constexpr int kLameCompact = 11;

// u = row cb of T, v = row ca, same = (cb == ca); callers read u and v from
// the row (or from c with fixed indices): a runtime index into a register
// array would move the array to scratch memory.
template <typename T>
__device__ __forceinline__ void lame_kblock(const T* c, const T* u, const T* v, bool same, T* Kb) {
#pragma clang fp contract(off)
    const T lam = c[9], mu = c[10];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T e = lam * (u[b] * v[a]);
            e = fma(mu, v[b] * u[a], e);
            const T g = fma(c[b + 6], c[a + 6], fma(c[b + 3], c[a + 3], c[b] * c[a]));
            const T eg = fma(mu, g, e);
            Kb[b * 3 + a] = same ? eg : e;
        }
}

// The (cb, ca) block of a compact row in global memory, contracted with one
// reference row r of 9 values (the contraction order of the row-block passes).
template <typename T>
__device__ __forceinline__ T lame_entry(const T* Kt, int cb, int ca, const T* r) {
    T c[kLameCompact], u[3], v[3], kb[9];
#pragma unroll
    for (int j = 0; j < kLameCompact; ++j) c[j] = Kt[j];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        u[k] = Kt[3 * cb + k];
        v[k] = Kt[3 * ca + k];
    }
    lame_kblock(c, u, v, cb == ca, kb);
    T acc = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) acc += kb[j] * r[j];
    return acc;
}

// Geometry pass over all tets (one lane per tet): FP64 Jacobian, inverse and
// |T| (core.inl:218-248), then the geometry tensors of the matrix (NCA*NCB
// blocks of q*DA0*DB0 values) and of the load vector (NCB blocks of qF
// values), stored to Kg[r*rowlen ...] in the T precision of the contraction.
// SYM (3 x 3 tensors of a symmetric K, one component, one carried point): the
// matrix part of the row is K packed to its 6 distinct values (sym_pack),
// which the row-block passes contract with the column-summed R.
// Dynamic LDS of a geometry launch: short rows (scalar spaces or compact Lame
// rows, at most 16 values) are staged, one row per lane; the launch passes
// exactly this size.
template <int NCA, int NCB, typename T, bool LC = false>
__host__ __device__ inline size_t geometry_stage_bytes(int rowlen) {
    return ((NCA * NCB == 1 || LC) && rowlen <= 16) ? (size_t)kBlock * rowlen * sizeof(T) : 0;
}

// LC: compact Lame rows (lame_kblock; a tensor constant over each tet).
template <int DA0, int DB0, int NCA, int NCB, typename T, bool SYM = false, bool LC = false>
__global__ __launch_bounds__(kBlock) void geometry_kernel(BlockArgs a, T* __restrict__ Kg, int64_t n_tets) {
    static_assert(!SYM || (DA0 == 3 && DB0 == 3 && NCA == 1 && NCB == 1), "packed K: 3 x 3 scalar blocks");
    static_assert(!LC || (DA0 == 3 && DB0 == 3 && NCA == 3 && NCB == 3), "compact Lame rows: FemVec<3> gradients");
    // short rows are assembled in LDS (sized to the row, so that more
    // workgroups fit a CU) and stored by the workgroup as one contiguous
    // range: coalesced stores instead of one strided store per value and lane
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* stage = reinterpret_cast<T*>(smem_raw);
    const bool staged = geometry_stage_bytes<NCA, NCB, T, LC>(a.rowlen) > 0;
    const int64_t r0 = blockIdx.x * (int64_t)kBlock;
    const int64_t r = r0 + threadIdx.x;
    int degenerate = 0;
    if (r < n_tets) {
        const int kstride = LC ? kLameCompact : (SYM ? 6 : NCA * NCB * a.m.q * DA0 * DB0);
        double P[4][3], PSI[9], det;
        load_tet<true>(a.m, r, P);
        tet_geometry(P, PSI, det);
        degenerate = !(fabs(det) > 0.0 && isfinite(det));
        const double mes = fabs(det) / 6.0;
        T* row = staged ? stage + threadIdx.x * a.rowlen : Kg + r * a.rowlen;
        if constexpr (LC) {
            const bool cst = a.m.storage == ANIFEM_COEF_CONST;
            const double lam = cst ? a.m.D[0] : a.m.D[r], mu = cst ? a.m.D[1] : a.m.D[a.m.coef_n + r];
#pragma unroll
            for (int i = 0; i < 9; ++i) row[i] = (T)PSI[i];
            row[9] = (T)(mes * lam);
            row[10] = (T)(mes * mu);
        } else if constexpr (SYM) {
            double K9[9];  // K(b, a) at b*3 + a
            geometry_block<3, 3, double>(a.m, r, PSI, mes, 0, 0, 0, K9);
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int c = 0; c <= b; ++c) row[sym_pack(c, b)] = (T)K9[b * 3 + c];
        } else {
            geometry_tensor<DA0, DB0, NCA, NCB, T>(a.m, r, PSI, mes, row);
        }
        if (a.has_rhs) geometry_tensor<1, 1, 1, NCB, T>(a.f, r, PSI, mes, row + kstride);
    }
    if (staged) {  // uniform over the workgroup
        __syncthreads();
        const int64_t nt = n_tets - r0 < kBlock ? n_tets - r0 : kBlock;
        const int nv = (int)nt * a.rowlen;
        T* dst = Kg + r0 * a.rowlen;
        for (int i = threadIdx.x; i < nv; i += kBlock) dst[i] = stage[i];
    }
    const unsigned long long ballot = __ballot(degenerate);
    if (a.deg) {  // uniform: every workgroup stores its count, so a pass clears nothing beforehand
        __shared__ int wave_count[kBlock / 64];
        if ((threadIdx.x & 63) == 0) wave_count[threadIdx.x / 64] = (int)__popcll(ballot);
        __syncthreads();
        if (threadIdx.x == 0) {
            int c = 0;
            for (int w = 0; w < kBlock / 64; ++w) c += wave_count[w];
            a.deg[blockIdx.x] = c;
            // the non-finite count of this pass is added by the passes after this one
            if (blockIdx.x == 0) a.status[0] = 0;
        }
    } else if (ballot && (threadIdx.x & 63) == 0) {  // one counter update per wavefront
        atomicAdd(&a.status[1], (int)__popcll(ballot));
    }
}


// LDS layout of one row-block workgroup, in bytes from the start (all parts
// 16-byte aligned): entry values | load-vector values | R | R_F | slot starts.
struct BlockLds {
    size_t vals, fvals, R, RF, sst, pst, total;
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

__host__ __device__ inline BlockLds block_lds_layout(int max_values, int max_pairs, int max_slots, int elem,
                                                     int R_vals, int RF_vals) {
    BlockLds L;
    size_t o = 0;
    L.vals = o;
    o = align16(o + 8 * ((size_t)max_values + 1));  // + one junk slot for dropped entries
    L.fvals = o;
    o = align16(o + 8 * (size_t)max_pairs);
    L.R = o;
    o = align16(o + (size_t)elem * R_vals);
    L.RF = o;
    o = align16(o + (size_t)elem * RF_vals);
    L.sst = o;
    o = align16(o + 2 * (size_t)max_slots + 4);  // + one word: slot starts may be staged from an odd entry
    L.pst = o;
    o = align16(o + 4 * (size_t)max_pairs);
    L.total = o;
    return L;
}

// Phase B of a row block: one lane per slot adds the slot's entries in list
// (ascending tet) order and stores the CSR value once; one lane per row does
// the same for its pairs' load-vector entries.
__device__ __forceinline__ void row_block_finish(const BlockArgs& a, int32_t r0, int nr, int64_t s0, int ns, int np,
                                                 int nv, const double* vals, const double* fvals,
                                                 const uint16_t* sst) {
    for (int u = threadIdx.x; u < ns; u += kBlock) {
        const int cb = sst[u];
        const int ce = u + 1 < ns ? sst[u + 1] : nv;
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
        __builtin_nontemporal_store(sum, &a.val[s0 + u]);  // written once, never re-read by the pass
    }
    // one lane per row: its pairs' load-vector entries, ascending tet order
    if (a.has_rhs) {
        for (int u = threadIdx.x; u < nr; u += kBlock) {
            const int pb = a.pstart[r0 + u];
            const int pe = u + 1 < nr ? a.pstart[r0 + u + 1] : np;
            double sum = 0.0;
            for (int p = pb; p < pe; ++p) sum += fvals[p];
            a.rhs[r0 + u] = sum;
        }
    }
}

// RS: the reference tensors are staged in LDS (else read from global memory
// when they do not fit); fixed per instantiation so that every load has a
// known address space.
template <int DA0, int DB0, int NCA, int NCB, typename T, bool RS>
__global__ __launch_bounds__(kBlock) void row_block_kernel(BlockArgs a, const T* __restrict__ Kg) {
    constexpr int KB = DA0 * DB0;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int len = a.m.q * KB, lenp = pad16<T>(len);
    const int lenF = a.has_rhs ? a.f.q : 0, lenFp = a.has_rhs ? pad16<T>(lenF) : 0;
    const int kstride = a.lamec ? kLameCompact : NCA * NCB * len;
    const BlockLds L = block_lds_layout(a.max_values, a.max_pairs, a.max_slots, (int)sizeof(T),
                                        RS ? a.max_R_rows * lenp : 0, RS && a.has_rhs ? a.max_RF_rows * lenFp : 0);
    double* vals = reinterpret_cast<double*>(smem_raw + L.vals);
    double* fvals = reinterpret_cast<double*>(smem_raw + L.fvals);
    uint16_t* sst = reinterpret_cast<uint16_t*>(smem_raw + L.sst);

    // workgroup w runs on XCD w % 8: each XCD takes a contiguous range of the
    // processing order, so blocks sharing tets meet in one L2
    const int64_t per_xcd = (a.n_blocks + 7) / 8;
    const int64_t k = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (k >= a.n_blocks) return;  // uniform over the workgroup
    const BlockMeta M = a.meta[k];
    const int32_t r0 = M.r0;
    const int nr = M.nr;
    const int64_t s0 = M.s0;
    const int ns = M.ns;
    const int64_t p0 = M.p0;
    const int np = M.np;
    const int nv = M.nv;
    const KindTable& KT = a.kinds[M.kind];  // uniform over the workgroup
    const int nf = KT.nf, nfA0 = KT.nfA0, nfB0 = KT.nfB0;

    // the pair of this lane and its tet's geometry row, requested before the
    // staging so that their loads overlap it
    const bool has_pair = (int)threadIdx.x < np;
    const int64_t gp = p0 + (has_pair ? threadIdx.x : 0);
    const int64_t qp = M.q0 + (int64_t)(has_pair ? threadIdx.x : 0) * nf;  // this pair's value positions
    const int32_t tet = has_pair ? a.pair_tet[gp] : 0;
    const int ibf = has_pair ? a.pair_ib[gp] : 0;
    const int ib = ibf & 63;
    const int cb = ib / nfB0;
    const T* Kt = Kg + (int64_t)tet * a.rowlen;
    const bool qc = len == KB;  // one carried point: K blocks held in registers
    T kreg[NCA][KB];
    if constexpr (NCA == 3 && NCB == 3 && KB == 9) {
        if (a.lamec) {  // compact Lame row: the pair's three blocks in closed form
            T c[kLameCompact], u[3];
#pragma unroll
            for (int j = 0; j < kLameCompact; ++j) c[j] = has_pair ? Kt[j] : T(0);
#pragma unroll
            for (int k = 0; k < 3; ++k) u[k] = has_pair ? Kt[3 * cb + k] : T(0);
#pragma unroll
            for (int ca = 0; ca < NCA; ++ca) lame_kblock(c, u, c + 3 * ca, cb == ca, kreg[ca]);
        }
    }
    if (qc && !a.lamec) {
#pragma unroll
        for (int ca = 0; ca < NCA; ++ca)
#pragma unroll
            for (int j = 0; j < KB; ++j) kreg[ca][j] = has_pair ? Kt[(cb * NCA + ca) * KB + j] : T(0);
    }
    // value positions as 16-bit pairs (nf is even), held in registers for
    // scalar spaces (nf <= 2*PW)
    constexpr int PW = 10;
    const bool posr = NCA == 1 && nf <= 2 * PW;
    uint32_t posw[PW];
    {
        const uint32_t* pw = reinterpret_cast<const uint32_t*>(a.pair_pos + qp);
#pragma unroll
        for (int w = 0; w < PW; ++w) posw[w] = (has_pair && posr && 2 * w < nf) ? pw[w] : 0u;
    }

    const T* Rg;
    const T* RFg;
    if constexpr (sizeof(T) == 8) {
        Rg = reinterpret_cast<const T*>(KT.Rd);
        RFg = reinterpret_cast<const T*>(KT.RFd);
    } else {
        Rg = reinterpret_cast<const T*>(KT.Rf);
        RFg = reinterpret_cast<const T*>(KT.RFf);
    }
    const T* R;
    const T* RF;
    int rstride, rfstride;
    if constexpr (RS) {
        T* Rs = reinterpret_cast<T*>(smem_raw + L.R);
        T* RFs = reinterpret_cast<T*>(smem_raw + L.RF);
        pad_copy(Rs, Rg, KT.R_rows, len, lenp);
        if (a.has_rhs) pad_copy(RFs, RFg, KT.RF_rows, lenF, lenFp);
        R = Rs;
        RF = RFs;
        rstride = lenp;
        rfstride = lenFp;
    } else {
        R = Rg;
        RF = RFg;
        rstride = len;
        rfstride = lenF;
    }
    block_copy(sst, a.sstart + s0, ns);
    __syncthreads();

    // phase A: the pair's element-matrix row and load-vector entry
    int nonfinite = 0;
    if (has_pair) {
        const bool dir = (ibf & kPairDirichlet) != 0;
        const int ib0 = ib % nfB0;
        const uint16_t* pos = a.pair_pos + qp;
        auto entry = [&](int ps, const T* Kb, const T* kr, int ia0) {
            const T* Rr = R + (ib0 + nfB0 * ia0) * rstride;
            T acc = 0;
            if (qc) {
#pragma unroll
                for (int j = 0; j < KB; ++j) acc += kr[j] * Rr[j];
            } else {
                for (int j = 0; j < len; ++j) acc += Kb[j] * Rr[j];
            }
            // a kept entry of a Dirichlet row is its diagonal: applyDir sets it to 1
            const double v = dir ? 1.0 : (double)acc;
            nonfinite += isfinite(v) ? 0 : 1;
            vals[ps] = v;
        };
        if (posr) {
            // scalar space: entries unrolled, positions from registers
#pragma unroll
            for (int ia0 = 0; ia0 < 2 * PW; ++ia0) {
                const uint32_t w = posw[ia0 >> 1];
                const int ps = (int)((ia0 & 1) ? (w >> 16) : (w & 0xFFFFu));
                if (ia0 < nfA0 && ps != kPosDropped) entry(ps, Kt + cb * len, kreg[0], ia0);
            }
        } else {
#pragma unroll
            for (int ca = 0; ca < NCA; ++ca) {
                for (int ia0 = 0; ia0 < nfA0; ++ia0) {
                    const int ps = pos[ca * nfA0 + ia0];
                    if (ps != kPosDropped) entry(ps, Kt + (cb * NCA + ca) * len, kreg[ca], ia0);
                }
            }
        }
        if (a.has_rhs) {
            const T* KFb = Kt + kstride + cb * lenF;
            const T* RFr = RF + ib0 * rfstride;
            T acc = 0;
            for (int j = 0; j < lenF; ++j) acc += KFb[j] * RFr[j];
            // applyDir zeroes the load-vector entries of a Dirichlet row
            const double fv = dir ? 0.0 : (double)acc;
            nonfinite += isfinite(fv) ? 0 : 1;
            fvals[threadIdx.x] = fv;
        }
    }
    __syncthreads();

    row_block_finish(a, r0, nr, s0, ns, np, nv, vals, fvals, sst);
    if (nonfinite) atomicAdd(&a.status[0], nonfinite);
}

// Rows whose load-vector entries an inhomogeneous Dirichlet condition changes:
// a Dirichlet row, or a row with a Dirichlet column in one of its cells.  Its
// pairs (the row's incident tets, ascending) are the global pair range
// [p0, p0 + np).
struct LiftRow {
    int64_t p0;
    int32_t row;
    int32_t np;
};

// Inhomogeneous Dirichlet values (dc_on_dof.h:38-45, applied per cell as in
// ex1.cpp:100-105): after the row-block pass, one lane per lifted row forms
// the row's rhs again with each cell's load-vector entry corrected as applyDir
// leaves it, then sums the cells' entries in ascending tet order (the order of
// the pass and of the reference's element loop):
//   * Dirichlet row k: F(k) = bc_k in every cell, so the row gets the sum of
//     bc_k over its incident tets;
//   * other row i: F(i) = F0(i) - A(i, k1) bc_k1 - A(i, k2) bc_k2 - ... over the
//     cell's Dirichlet DOFs in local order; applyDir on an earlier DOF zeroes
//     only its own row and column, so every A(i, k) is the original element
//     entry.  Those entries are exactly the ones the pattern dropped
//     (positions kPosDropped), formed here from the tet's geometry row and the
//     reference tensor as phase A forms the kept ones.
// Rows that no Dirichlet DOF touches keep the pass's values.
template <int DA0, int DB0, int NCA, int NCB, typename T>
__global__ __launch_bounds__(kBlock) void dirichlet_lift_kernel(BlockArgs a, const T* __restrict__ Kg,
                                                                const LiftRow* __restrict__ lift, int64_t n_lift,
                                                                const int32_t* __restrict__ dofs,
                                                                const uint8_t* __restrict__ is_dir,
                                                                const double* __restrict__ bc) {
    constexpr int KB = DA0 * DB0;
    const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (i >= n_lift) return;
    const KindTable& KT = a.kinds[0];
    const int nf = KT.nf, nfA0 = KT.nfA0, nfB0 = KT.nfB0;
    const int len = a.m.q * KB, lenF = a.has_rhs ? a.f.q : 0;
    const int kstride = a.lamec ? kLameCompact : NCA * NCB * len;
    const T* R = reinterpret_cast<const T*>(sizeof(T) == 8 ? (const void*)KT.Rd : (const void*)KT.Rf);
    const T* RF = reinterpret_cast<const T*>(sizeof(T) == 8 ? (const void*)KT.RFd : (const void*)KT.RFf);
    const LiftRow L = lift[i];
    const double bc_row = bc[L.row];
    double sum = 0.0;
    int nonfinite = 0;
    for (int64_t gp = L.p0; gp < L.p0 + L.np; ++gp) {
        const int ibf = a.pair_ib[gp];
        double fv;
        if (ibf & kPairDirichlet) {
            fv = bc_row;
        } else {
            const int64_t tet = a.pair_tet[gp];
            const int ib = ibf & 63, ib0 = ib % nfB0, cb = ib / nfB0;
            const T* Kt = Kg + tet * a.rowlen;
            T f0 = 0;
            for (int j = 0; j < lenF; ++j) f0 += Kt[kstride + cb * lenF + j] * RF[ib0 * lenF + j];
            fv = (double)f0;
            const uint16_t* pos = a.pair_pos + gp * nf;
            for (int ia = 0; ia < nf; ++ia) {
                // dropped columns of Dirichlet DOFs (a decoupled pattern also
                // drops zero entries between components)
                if (pos[ia] != kPosDropped || !is_dir[dofs[tet * nf + ia]]) continue;
                const int ca = ia / nfA0, ia0 = ia % nfA0;
                const T* Kb = Kt + (cb * NCA + ca) * len;
                const T* Rr = R + (ib0 + nfB0 * ia0) * len;
                T acc = 0;
                if constexpr (NCA == 3 && NCB == 3 && KB == 9) {
                    if (a.lamec) acc = lame_entry(Kt, cb, ca, Rr);
                }
                if (!a.lamec)
                    for (int j = 0; j < len; ++j) acc += Kb[j] * Rr[j];
                fv -= (double)acc * bc[dofs[tet * nf + ia]];
            }
        }
        nonfinite += isfinite(fv) ? 0 : 1;
        sum += fv;
    }
    a.rhs[L.row] = sum;
    if (nonfinite) atomicAdd(&a.status[0], nonfinite);
}

// Row-block pass over caller-supplied element matrices (the local matrices a
// MatFuncWrap returns, func_wrap.h:96-187): Ae holds n_tets nf x nf matrices,
// column-major (entry (ib, ia) at Ae[t*nf*nf + ib + nf*ia]), Fe n_tets load
// vectors of nf (or null), in each tet's local DOF order after the node
// reorder.  Phase A reads the pair's row of its tet's matrix instead of forming
// it; Dirichlet rows and phase B are those of row_block_kernel.
// SYM: the caller's matrices are symmetric; the pair's row is read as the
// matching column, which is contiguous in the column-major layout.
template <typename TE, bool SYM>
__global__ __launch_bounds__(kBlock) void row_block_elements_kernel(BlockArgs a, const TE* __restrict__ Ae,
                                                                    const TE* __restrict__ Fe) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const BlockLds L = block_lds_layout(a.max_values, a.max_pairs, a.max_slots, 8, 0, 0);
    double* vals = reinterpret_cast<double*>(smem_raw + L.vals);
    double* fvals = reinterpret_cast<double*>(smem_raw + L.fvals);
    uint16_t* sst = reinterpret_cast<uint16_t*>(smem_raw + L.sst);
    const int64_t per_xcd = (a.n_blocks + 7) / 8;
    const int64_t k = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (k >= a.n_blocks) return;  // uniform over the workgroup
    const BlockMeta M = a.meta[k];
    const int nf = a.kinds[M.kind].nf;
    const bool has_pair = (int)threadIdx.x < M.np;
    block_copy(sst, a.sstart + M.s0, M.ns);
    int nonfinite = 0;
    if (has_pair) {
        const int64_t gp = M.p0 + threadIdx.x;
        const int64_t qp = M.q0 + (int64_t)threadIdx.x * nf;
        const int64_t tet = a.pair_tet[gp];
        const int ibf = a.pair_ib[gp];
        const int ib = ibf & 63;
        const bool dir = (ibf & kPairDirichlet) != 0;
        // orientation signs (assembler.inl:154-164): entry (ib, ia) of the
        // global basis is s_ib s_ia times the local one
        const uint64_t sm = a.signmask ? a.signmask[tet] : 0;
        const uint64_t flip = ((sm >> ib) & 1) ? ~sm : sm;  // bit ia: s_ib s_ia = -1
        const TE* At = Ae + tet * nf * nf + (SYM ? (int64_t)nf * ib : ib);
        const int64_t st = SYM ? 1 : nf;
        // eight entries at a time: their positions and values are requested
        // together (every value is inside the tet's matrix, dropped or not),
        // then the kept ones are stored
        constexpr int U = 8;
        for (int ia0 = 0; ia0 < nf; ia0 += U) {
            int ps[U];
            TE v[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const int ia = ia0 + j < nf ? ia0 + j : nf - 1;
                ps[j] = ia0 + j < nf ? (int)a.pair_pos[qp + ia] : kPosDropped;
                v[j] = At[st * ia];
            }
#pragma unroll
            for (int j = 0; j < U; ++j) {
                if (ps[j] == kPosDropped) continue;
                const double x = dir ? 1.0 : ((flip >> ((ia0 + j) & 63)) & 1 ? -(double)v[j] : (double)v[j]);
                nonfinite += isfinite(x) ? 0 : 1;
                vals[ps[j]] = x;
            }
        }
        if (a.has_rhs) {
            const double f0 = (double)Fe[tet * nf + ib];
            const double fv = dir ? 0.0 : ((sm >> ib) & 1 ? -f0 : f0);
            nonfinite += isfinite(fv) ? 0 : 1;
            fvals[threadIdx.x] = fv;
        }
    }
    __syncthreads();
    row_block_finish(a, M.r0, M.nr, M.s0, M.ns, M.np, M.nv, vals, fvals, sst);
    if (nonfinite) atomicAdd(&a.status[0], nonfinite);
}

// Right-hand side only (AssembleRHS, assembler.inl:497-561): one workgroup per
// block; each pair's load-vector entry (0 on a Dirichlet row, applyDir) goes
// to LDS, then one lane per row adds its pairs' entries in ascending tet order
// and stores (accumulate = 0) or adds the sum to rhs (rhs <- rhs + assembled).
template <typename TE>
__global__ __launch_bounds__(kBlock) void row_block_rhs_kernel(BlockArgs a, const TE* __restrict__ Fe, int accumulate) {
    __shared__ double fv[kBlock];
    const int64_t per_xcd = (a.n_blocks + 7) / 8;
    const int64_t k = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (k >= a.n_blocks) return;  // uniform over the workgroup
    const BlockMeta M = a.meta[k];
    const int nf = a.kinds[M.kind].nf;
    int nonfinite = 0;
    if ((int)threadIdx.x < M.np) {
        const int64_t gp = M.p0 + threadIdx.x;
        const int64_t tet = a.pair_tet[gp];
        const int ibf = a.pair_ib[gp];
        const uint64_t sm = a.signmask ? a.signmask[tet] : 0;
        const double f0 = (double)Fe[tet * nf + (ibf & 63)];
        const double v = (ibf & kPairDirichlet) ? 0.0 : ((sm >> (ibf & 63)) & 1 ? -f0 : f0);
        nonfinite += isfinite(v) ? 0 : 1;
        fv[threadIdx.x] = v;
    }
    __syncthreads();
    for (int u = threadIdx.x; u < M.nr; u += kBlock) {
        const int pb = a.pstart[M.r0 + u];
        const int pe = u + 1 < M.nr ? a.pstart[M.r0 + u + 1] : M.np;
        double sum = 0.0;
        for (int p = pb; p < pe; ++p) sum += fv[p];
        a.rhs[M.r0 + u] = accumulate ? a.rhs[M.r0 + u] + sum : sum;
    }
    if (nonfinite) atomicAdd(&a.status[0], nonfinite);
}

}  // namespace anifem

namespace anifem {

// Block sequence of a persistent workgroup: workgroup w runs on XCD w % 8, and
// the blocks of the processing order are dealt to the XCDs in contiguous
// ranges (the grid is a multiple of 8 and at most n_blocks).
struct BlockWalk {
    int64_t next, end, step;
    __device__ void init(int64_t n_blocks) {
        const int64_t G = gridDim.x, w = blockIdx.x;
        const int64_t x = w % 8, l = w / 8;
        next = n_blocks * x / 8 + l;
        end = n_blocks * (x + 1) / 8;
        step = G / 8;
    }
    __device__ int64_t take() {
        const int64_t b = next < end ? next : -1;
        next += step;
        return b;
    }
};

// One lane's share of a block's records: its pair's tet, geometry row (K block
// and load-vector factor) and value positions as 16-bit pairs (PW words), and
// SW 32-bit words of the block's slot starts and one row's first pair, which
// are moved to LDS when the block's turn comes.
template <int KB, int PW, int SW, typename T>
struct PairRegs {
    int32_t tet;
    int ibf;
    uint32_t posw[PW];
    T k[KB];
    T kf;
    uint32_t sw[SW > 0 ? SW : 1];
    int pst;
};

// Slot starts of a block as aligned 32-bit words: the words covering
// sstart[s0 .. s0+ns) (the last may reach one entry past the array, which the
// upload pads).
__device__ __forceinline__ int slot_words(int64_t s0, int ns) { return (int)(((s0 & 1) + ns + 1) >> 1); }

// Persistent row-block pass for scalar spaces with one carried point (the
// matrix tensor constant over each tet) and elements of at most 2*PW entries
// per row: each workgroup stages the reference tensors of its blocks' element
// kind (restaged when the kind changes) and walks its blocks; while block b is
// evaluated, the next block's pair records, value positions, slot starts and
// row starts are in flight, and its geometry rows are requested once the
// records have arrived, so the gather latency hides behind phases A and B.
// Same arithmetic and summation order as row_block_kernel: the results are
// bitwise identical.
// P8: the value positions come from the compact list (a 16-bit base and one
// byte per entry, P8 words per pair) instead of 16-bit positions.
// NFX > 0: every block is of one element kind with NFX local DOFs (a scalar
// space, ib < NFX): the reference-tensor rows of a pair sit at compile-time
// offsets, and phase A runs without per-entry branches: a dropped entry is
// stored to a junk slot after the value list, a Dirichlet pair contracts
// K = 0 onto the value 1, and non-finite entries are counted entry by entry
// only for a pair whose K is not finite or beyond 1e300 (below that no entry
// can overflow).  Same values and sums as NFX = 0.
// SPLIT (general form, 16-bit positions, blocks of at most kBlock/2 pairs):
// two lanes per pair, lane t and lane t + kBlock/2 forming entries 0..PW-1 and
// PW..2*PW-1 of pair t, so that every lane works in phase A when a block's
// pairs are long (P3: 20 entries).  Same values and sums.
// TRI (with SPLIT, K packed to 6 values, one kind of 2*PW DOFs on both sides
// whose reference tensor is symmetric in (ia, ib) bit for bit): only the
// triangle ib >= ia of the reference tensor is staged (P3: 210 of 400 rows,
// 10 KB less LDS per workgroup, a.max_R_rows set to the triangle's rows);
// entry (ib, ia) reads row (min, max).  Same values and sums.
template <int DA0, int DB0, int PW, typename T, bool P8 = false, int NFX = 0, bool SPLIT = false, bool TRI = false>
// Five waves per SIMD (the LDS of five 30 KB workgroups fills a CU): at most
// 96 VGPRs, measured faster than four unconstrained waves and than spilling
// to six or eight.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void row_block_persistent(BlockArgs a, const T* __restrict__ Kg) {
    static_assert(!SPLIT || (!P8 && NFX == 0 && PW % 2 == 0), "split pairs: general form, 16-bit positions");
    static_assert(!TRI || (SPLIT && DA0 == 6 && DB0 == 1), "triangle of a symmetric packed reference tensor");
    constexpr int KB = DA0 * DB0;
    constexpr int NF = 2 * PW;  // DOFs per cell of the TRI kind
    // position words per lane: a pair's, or half of them when two lanes share it
    constexpr int NPW = P8 ? (2 + 2 * PW + 3) / 4 : (SPLIT ? PW / 2 : PW);
    constexpr int NE = SPLIT ? PW : 2 * PW;  // entries per lane
    constexpr int kHalf = kBlock / 2;
    // the lane's pair and its first entry
    const int pl = SPLIT ? ((int)threadIdx.x & (kHalf - 1)) : (int)threadIdx.x;
    const int ia_lane = SPLIT ? ((int)threadIdx.x / kHalf) * NE : 0;
    const bool rhs_lane = !SPLIT || (int)threadIdx.x < kHalf;
    // slot starts carried in registers for blocks of at most 2*SW*kBlock - 2
    // slots (the registers freed by the shorter position lists of P1 and P2);
    // else copied to LDS when the block's turn comes
    constexpr int SW = PW <= 5 ? 5 : 0;
    const bool sw_regs = SW > 0 && a.max_slots <= 2 * SW * kBlock - 2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int lenp = pad16<T>(KB);
    const int lenFp = a.has_rhs ? pad16<T>(1) : 0;
    const BlockLds L = block_lds_layout(a.max_values, a.max_pairs, a.max_slots, (int)sizeof(T), a.max_R_rows * lenp,
                                        a.has_rhs ? a.max_RF_rows * lenFp : 0);
    double* vals = reinterpret_cast<double*>(smem_raw + L.vals);
    double* fvals = reinterpret_cast<double*>(smem_raw + L.fvals);
    uint16_t* sst = reinterpret_cast<uint16_t*>(smem_raw + L.sst);
    int* pstl = reinterpret_cast<int*>(smem_raw + L.pst);
    T* R = reinterpret_cast<T*>(smem_raw + L.R);
    T* RF = reinterpret_cast<T*>(smem_raw + L.RF);
    // reference tensors of one element kind in LDS (a workgroup's contiguous
    // range of the processing order rarely crosses kinds: they are restaged
    // only when it does)
    auto stage = [&](int kind) {
        const KindTable& KT = a.kinds[kind];
        const T* Rsrc = reinterpret_cast<const T*>(sizeof(T) == 8 ? (const void*)KT.Rd : (const void*)KT.Rf);
        const T* RFsrc = reinterpret_cast<const T*>(sizeof(T) == 8 ? (const void*)KT.RFd : (const void*)KT.RFf);
        if constexpr (TRI)
            tri_copy<NF, KB>(R, Rsrc, lenp);
        else
            pad_copy(R, Rsrc, KT.R_rows, KB, lenp);
        if (a.has_rhs) pad_copy(RF, RFsrc, KT.RF_rows, 1, lenFp);
    };
    const int kstride = KB;  // one component block, one point

    BlockWalk W;
    W.init(a.n_blocks);
    int64_t b = W.take();
    if (b < 0) return;  // uniform over the workgroup

    using Regs = PairRegs<KB, NPW, SW, T>;
    auto load_records = [&](const BlockMeta& Mx, Regs& P) {
        const int nf = a.kinds[Mx.kind].nf;
        const bool hp = pl < Mx.np;
        const int64_t gp = Mx.p0 + (hp ? pl : 0);
        P.tet = hp ? a.pair_tet[gp] : 0;
        P.ibf = hp ? a.pair_ib[gp] : 0;
        if constexpr (P8) {
            const uint32_t* pw = a.pair_pos8 + (gp - a.pos8_base) * NPW;
#pragma unroll
            for (int w = 0; w < NPW; ++w) P.posw[w] = hp ? pw[w] : 0u;
        } else {
            const int w0 = ia_lane / 2;
            const uint32_t* pw =
                reinterpret_cast<const uint32_t*>(a.pair_pos + Mx.q0 + (int64_t)(hp ? pl : 0) * nf) + w0;
#pragma unroll
            for (int w = 0; w < NPW; ++w) P.posw[w] = (hp && 2 * (w0 + w) < nf) ? pw[w] : 0u;
        }
    };
    auto load_k = [&](const BlockMeta& Mx, Regs& P) {
        const bool hp = pl < Mx.np;
        const T* Kt = Kg + (int64_t)P.tet * a.rowlen;
#pragma unroll
        for (int j = 0; j < KB; ++j) P.k[j] = hp ? Kt[j] : T(0);
        P.kf = (hp && a.has_rhs) ? Kt[kstride] : T(0);
        // slot and row starts, needed only at the block's phase B
        if constexpr (SW > 0) {
            if (sw_regs) {
                const uint32_t* sg = reinterpret_cast<const uint32_t*>(a.sstart) + (Mx.s0 >> 1);
                const int nw = slot_words(Mx.s0, (int)Mx.ns);
#pragma unroll
                for (int w = 0; w < SW; ++w) {
                    const int i = w * kBlock + (int)threadIdx.x;
                    P.sw[w] = i < nw ? sg[i] : 0u;
                }
            }
        }
        P.pst = (a.has_rhs && (int)threadIdx.x < Mx.nr) ? (int)a.pstart[Mx.r0 + threadIdx.x] : 0;
    };

    BlockMeta M = a.meta[b];
    int staged_kind = (int)M.kind;
    stage(staged_kind);
    Regs cur, nxt, nn;
    load_records(M, cur);
    load_k(M, cur);
    int64_t bn = W.take();
    BlockMeta Mn{};
    if (bn >= 0) Mn = a.meta[bn];
    // fixed-size form: records two blocks ahead (a block's geometry rows are
    // then requested when the block before it starts, not after its phase A)
    int64_t bn2 = -1;
    BlockMeta Mn2{};
    if constexpr (NFX > 0) {
        load_records(Mn, nxt);
        bn2 = W.take();
        if (bn2 >= 0) Mn2 = a.meta[bn2];
    }
    __syncthreads();  // reference tensors in LDS

    int nonfinite = 0;
    for (;;) {
        if ((int)M.kind != staged_kind) {  // uniform: the previous block ended with a barrier
            staged_kind = (int)M.kind;
            stage(staged_kind);
            __syncthreads();
        }
        const int nfA0 = a.kinds[staged_kind].nfA0, nfB0 = a.kinds[staged_kind].nfB0;
        // slot and row starts of b to LDS (read in phase B), records of the
        // next block
        const uint16_t* sstb = sst;
        bool sst_copied = false;
        if constexpr (SW > 0) {
            if (sw_regs) {
                uint32_t* s32 = reinterpret_cast<uint32_t*>(sst);
                const int nw = slot_words(M.s0, (int)M.ns);
#pragma unroll
                for (int w = 0; w < SW; ++w) {
                    const int i = w * kBlock + (int)threadIdx.x;
                    if (i < nw) s32[i] = cur.sw[w];
                }
                sstb = sst + (M.s0 & 1);
                sst_copied = true;
            }
        }
        if (!sst_copied) block_copy(sst, a.sstart + M.s0, (int)M.ns);
        if (a.has_rhs && (int)threadIdx.x < M.nr) pstl[threadIdx.x] = cur.pst;
        int64_t bnn = -1;
        BlockMeta Mnn{};
        if constexpr (NFX > 0) {
            load_k(Mn, nxt);        // records of b+1 arrived during the previous block
            load_records(Mn2, nn);  // past the last block: no pairs, nothing is read
            bnn = W.take();         // b+3
            if (bnn >= 0) Mnn = a.meta[bnn];
        } else {
            load_records(Mn, nxt);  // past the last block: Mn has no pairs, nothing is read
            bnn = W.take();
            if (bnn >= 0) Mnn = a.meta[bnn];
        }

        // phase A of b
        if (NFX > 0 && (int)threadIdx.x < M.np) {  // uniform over NFX
            static_assert(NFX <= 2 * PW, "NFX entries per pair");
            const int ib = cur.ibf & 63;
            const bool dir = (cur.ibf & kPairDirichlet) != 0;
            const T* Rb = R + ib * lenp;
            const int base8 = (int)(cur.posw[0] & 0xFFFFu);
            const int junk = a.max_values;
            // finite K below 1e300 (R is finite and O(1)): no entry overflows
            const T big = sizeof(T) == 8 ? T(1e300) : T(1e36);
            bool fast = true;
#pragma unroll
            for (int j = 0; j < KB; ++j) fast = fast && fabs(cur.k[j]) <= big;
            fast = fast && (!a.has_rhs || fabs(cur.kf) <= big);
            // applyDir: a Dirichlet pair's kept entry (its diagonal) is 1
            const T bias = dir ? T(1) : T(0);
#pragma unroll
            for (int j = 0; j < KB; ++j) cur.k[j] = dir ? T(0) : cur.k[j];
            // value-list position of entry ia0 (the junk slot for a dropped one)
            auto pos = [&](int ia0) {
                if constexpr (P8) {
                    const uint32_t o = (cur.posw[(2 + ia0) >> 2] >> (((2 + ia0) & 3) * 8)) & 0xFFu;
                    return o == 0xFFu ? junk : base8 + (int)o;
                } else {
                    const uint32_t w = cur.posw[ia0 >> 1];
                    const int p16 = (int)((ia0 & 1) ? (w >> 16) : (w & 0xFFFFu));
                    return p16 == (int)kPosDropped ? junk : p16;
                }
            };
#pragma unroll
            for (int ia0 = 0; ia0 < NFX; ++ia0) {
                const T* Rr = Rb + NFX * ia0 * lenp;
                T acc = bias;
#pragma unroll
                for (int j = 0; j < KB; ++j) acc += cur.k[j] * Rr[j];
                vals[pos(ia0)] = (double)acc;
            }
            if (!fast && !dir) {  // count the kept entries that are not finite
#pragma unroll
                for (int ia0 = 0; ia0 < NFX; ++ia0) {
                    const int ps = pos(ia0);
                    if (ps != junk) nonfinite += isfinite(vals[ps]) ? 0 : 1;
                }
            }
            if (a.has_rhs) {
                const T acc = cur.kf * RF[ib * lenFp];
                const double fv = dir ? 0.0 : (double)acc;
                if (!fast) nonfinite += isfinite(fv) ? 0 : 1;
                fvals[threadIdx.x] = fv;
            }
        } else if (NFX == 0 && pl < M.np) {
            const int ib = cur.ibf & 63;
            const bool dir = (cur.ibf & kPairDirichlet) != 0;
            const T* Rb = R + (ib % nfB0) * lenp;
            const int base8 = (int)(cur.posw[0] & 0xFFFFu);
            const int tri_ib = TRI ? tri_row<NF>(ib, 0) : 0;  // row (ib, ia) = tri_ib + ia for ia >= ib
            // entries IA .. IA+NE-1 of the pair (IA = 0, or NE on the second lane of a split pair)
            auto entries = [&](auto ia_c) {
                constexpr int IA = decltype(ia_c)::value;
#pragma unroll
                for (int i = 0; i < NE; ++i) {
                    const int ia0 = IA + i;
                    int ps;
                    if constexpr (P8) {
                        // byte 2 + ia0 of the pair's words: offset from the base, 0xFF dropped
                        const uint32_t o = (cur.posw[(2 + ia0) >> 2] >> (((2 + ia0) & 3) * 8)) & 0xFFu;
                        ps = o == 0xFFu ? (int)kPosDropped : base8 + (int)o;
                    } else {
                        const uint32_t w = cur.posw[i >> 1];
                        ps = (int)((i & 1) ? (w >> 16) : (w & 0xFFFFu));
                    }
                    if (ia0 < nfA0 && ps != kPosDropped) {
                        const T* Rr;
                        if constexpr (TRI)
                            Rr = R + (ia0 <= ib ? tri_row<NF>(ia0, 0) + ib : tri_ib + ia0) * lenp;
                        else
                            Rr = Rb + nfB0 * ia0 * lenp;
                        T acc = 0;
#pragma unroll
                        for (int j = 0; j < KB; ++j) acc += cur.k[j] * Rr[j];
                        const double v = dir ? 1.0 : (double)acc;
                        nonfinite += isfinite(v) ? 0 : 1;
                        vals[ps] = v;
                    }
                }
            };
            if constexpr (SPLIT) {
                if ((int)threadIdx.x < kHalf)  // uniform over each wave
                    entries(std::integral_constant<int, 0>{});
                else
                    entries(std::integral_constant<int, NE>{});
            } else {
                entries(std::integral_constant<int, 0>{});
            }
            if (a.has_rhs && rhs_lane) {
                const T acc = cur.kf * RF[(ib % nfB0) * lenFp];
                const double fv = dir ? 0.0 : (double)acc;
                nonfinite += isfinite(fv) ? 0 : 1;
                fvals[pl] = fv;
            }
        }
        // geometry rows of the next block (its records have arrived by now)
        if constexpr (NFX == 0) {
            __builtin_amdgcn_sched_barrier(0);
            load_k(Mn, nxt);
        }
        __syncthreads();

        // phase B of b
        const int ns = M.ns, nv = M.nv, nr = M.nr, np = M.np;
        for (int u = threadIdx.x; u < ns; u += kBlock) {
            const int cb = sstb[u];
            const int ce = u + 1 < ns ? sstb[u + 1] : nv;
            double sum = 0.0;
            int c = cb;
            for (; c + 4 <= ce; c += 4) {
                const double v0 = vals[c], v1 = vals[c + 1], v2 = vals[c + 2], v3 = vals[c + 3];
                sum += v0;
                sum += v1;
                sum += v2;
                sum += v3;
            }
            for (; c < ce; ++c) sum += vals[c];
            __builtin_nontemporal_store(sum, &a.val[M.s0 + u]);  // written once, never re-read by the pass
        }
        if (a.has_rhs) {
            for (int u = threadIdx.x; u < nr; u += kBlock) {
                const int pb = pstl[u];
                const int pe = u + 1 < nr ? pstl[u + 1] : np;
                double sum = 0.0;
                for (int p = pb; p < pe; ++p) sum += fvals[p];
                a.rhs[M.r0 + u] = sum;
            }
        }
        __syncthreads();
        if (bn < 0) break;
        cur = nxt;
        M = Mn;
        b = bn;
        if constexpr (NFX > 0) {
            nxt.tet = nn.tet;
            nxt.ibf = nn.ibf;
#pragma unroll
            for (int w = 0; w < NPW; ++w) nxt.posw[w] = nn.posw[w];
            Mn = Mn2;
            bn = bn2;
            Mn2 = Mnn;
            bn2 = bnn;
        } else {
            Mn = Mnn;
            bn = bnn;
        }
    }
    if (nonfinite) atomicAdd(&a.status[0], nonfinite);
}

// Persistent row-block pass for vector spaces FemVec<NC, FemFix<..>> x
// FemVec<NC, ..> (NF0 scalar DOFs per component on both sides, one carried
// point): the scheme of row_block_persistent with a pair's NC component
// blocks (cb, 0..NC-1) of its tet's geometry row in registers.  Vector
// elements fill a CU's LDS with one or two workgroups, so the next block's
// records, geometry rows, slot starts and row starts are all carried in
// registers (no occupancy cap on the register count), which is what hides
// the gather latency.  Same arithmetic and summation order as
// row_block_kernel: the results are bitwise identical.
// LC: compact Lame rows; a pair carries 11 values and expands its three
// blocks in phase A (lame_kblock).
template <int DA0, int DB0, int NC, int NF0, typename T, bool LC = false>
__global__ __launch_bounds__(kBlock) void row_block_persistent_vec(BlockArgs a, const T* __restrict__ Kg) {
    static_assert(!LC || (NC == 3 && DA0 == 3 && DB0 == 3), "compact Lame rows: FemVec<3> gradients");
    constexpr int KB = DA0 * DB0;
    constexpr int NF = NC * NF0;       // entries per pair
    constexpr int PW = NF / 2;         // position words per pair
    static_assert(NF % 2 == 0, "positions are read as 32-bit words");
    constexpr int SW = 9;              // slot-start words per lane: blocks of at most 4606 slots
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int lenp = pad16<T>(KB);
    const int lenFp = a.has_rhs ? pad16<T>(1) : 0;
    const BlockLds L = block_lds_layout(a.max_values, a.max_pairs, a.max_slots, (int)sizeof(T), a.max_R_rows * lenp,
                                        a.has_rhs ? a.max_RF_rows * lenFp : 0);
    double* vals = reinterpret_cast<double*>(smem_raw + L.vals);
    double* fvals = reinterpret_cast<double*>(smem_raw + L.fvals);
    uint16_t* sst = reinterpret_cast<uint16_t*>(smem_raw + L.sst);
    int* pstl = reinterpret_cast<int*>(smem_raw + L.pst);
    T* R = reinterpret_cast<T*>(smem_raw + L.R);
    T* RF = reinterpret_cast<T*>(smem_raw + L.RF);
    const bool sw_regs = a.max_slots <= 2 * SW * kBlock - 2;
    const int kstride = LC ? kLameCompact : NC * NC * KB;  // NC x NC component blocks, one point

    BlockWalk W;
    W.init(a.n_blocks);
    int64_t b = W.take();
    if (b < 0) return;  // uniform over the workgroup

    {  // reference tensors of the (single) element kind
        const KindTable& KT = a.kinds[0];
        if constexpr (sizeof(T) == 8) {
            pad_copy(R, reinterpret_cast<const T*>(KT.Rd), KT.R_rows, KB, lenp);
            if (a.has_rhs) pad_copy(RF, reinterpret_cast<const T*>(KT.RFd), KT.RF_rows, 1, lenFp);
        } else {
            pad_copy(R, reinterpret_cast<const T*>(KT.Rf), KT.R_rows, KB, lenp);
            if (a.has_rhs) pad_copy(RF, reinterpret_cast<const T*>(KT.RFf), KT.RF_rows, 1, lenFp);
        }
    }

    struct Regs {
        int32_t tet;
        int ibf;
        uint32_t posw[PW];
        T k[LC ? 1 : NC][LC ? 1 : KB];     // the pair's component blocks
        T c[LC ? kLameCompact : 1];        // or its compact Lame row
        T u[LC ? 3 : 1];                   // and row cb of its T
        T kf;
        uint32_t sw[SW];
        int pst;
    };
    auto load_records = [&](const BlockMeta& Mx, Regs& P) {
        const bool hp = (int)threadIdx.x < Mx.np;
        const int64_t gp = Mx.p0 + (hp ? threadIdx.x : 0);
        P.tet = hp ? a.pair_tet[gp] : 0;
        P.ibf = hp ? a.pair_ib[gp] : 0;
        // NF even: a pair's positions start at an even entry (q0 = p0 * NF)
        const uint32_t* pw =
            reinterpret_cast<const uint32_t*>(a.pair_pos + Mx.q0 + (int64_t)(hp ? threadIdx.x : 0) * NF);
#pragma unroll
        for (int w = 0; w < PW; ++w) P.posw[w] = hp ? pw[w] : 0u;
    };
    auto load_k = [&](const BlockMeta& Mx, Regs& P) {
        const bool hp = (int)threadIdx.x < Mx.np;
        const int cb = (P.ibf & 63) / NF0;
        const T* Kt = Kg + (int64_t)P.tet * a.rowlen;
        if constexpr (LC) {
#pragma unroll
            for (int j = 0; j < kLameCompact; ++j) P.c[j] = hp ? Kt[j] : T(0);
#pragma unroll
            for (int k = 0; k < 3; ++k) P.u[k] = hp ? Kt[3 * cb + k] : T(0);
        } else {
#pragma unroll
            for (int ca = 0; ca < NC; ++ca)
#pragma unroll
                for (int j = 0; j < KB; ++j) P.k[ca][j] = hp ? Kt[(cb * NC + ca) * KB + j] : T(0);
        }
        P.kf = (hp && a.has_rhs) ? Kt[kstride + cb] : T(0);
        if (sw_regs) {
            const uint32_t* sg = reinterpret_cast<const uint32_t*>(a.sstart) + (Mx.s0 >> 1);
            const int nw = slot_words(Mx.s0, (int)Mx.ns);
#pragma unroll
            for (int w = 0; w < SW; ++w) {
                const int i = w * kBlock + (int)threadIdx.x;
                P.sw[w] = i < nw ? sg[i] : 0u;
            }
        }
        P.pst = (a.has_rhs && (int)threadIdx.x < Mx.nr) ? (int)a.pstart[Mx.r0 + threadIdx.x] : 0;
    };

    BlockMeta M = a.meta[b];
    Regs cur, nxt;
    load_records(M, cur);
    load_k(M, cur);
    int64_t bn = W.take();
    BlockMeta Mn{};
    if (bn >= 0) Mn = a.meta[bn];
    __syncthreads();  // reference tensors in LDS

    int nonfinite = 0;
    for (;;) {
        // slot and row starts of b to LDS (read in phase B), records of the
        // next block
        const uint16_t* sstb = sst;
        if (sw_regs) {
            uint32_t* s32 = reinterpret_cast<uint32_t*>(sst);
            const int nw = slot_words(M.s0, (int)M.ns);
#pragma unroll
            for (int w = 0; w < SW; ++w) {
                const int i = w * kBlock + (int)threadIdx.x;
                if (i < nw) s32[i] = cur.sw[w];
            }
            sstb = sst + (M.s0 & 1);
        } else {
            block_copy(sst, a.sstart + M.s0, (int)M.ns);
        }
        if (a.has_rhs && (int)threadIdx.x < M.nr) pstl[threadIdx.x] = cur.pst;
        load_records(Mn, nxt);  // past the last block: Mn has no pairs, nothing is read
        const int64_t bnn = W.take();
        BlockMeta Mnn{};
        if (bnn >= 0) Mnn = a.meta[bnn];

        // phase A of b: entries (ib, ca*NF0 + ia0) of the pair's row
        if ((int)threadIdx.x < M.np) {
            const int ib = cur.ibf & 63;
            const int ib0 = ib % NF0;
            const bool dir = (cur.ibf & kPairDirichlet) != 0;
            const T* Rb = R + ib0 * lenp;
            T kx[LC ? NC : 1][LC ? KB : 1];  // the expanded blocks of a compact row
            if constexpr (LC) {
#pragma unroll
                for (int ca = 0; ca < NC; ++ca) lame_kblock(cur.c, cur.u, cur.c + 3 * ca, ib / NF0 == ca, kx[ca]);
            }
            // one reference row per ia0, shared by the NC component entries
#pragma unroll
            for (int ia0 = 0; ia0 < NF0; ++ia0) {
                T r[KB];
#pragma unroll
                for (int j = 0; j < KB; ++j) r[j] = Rb[NF0 * ia0 * lenp + j];
#pragma unroll
                for (int ca = 0; ca < NC; ++ca) {
                    const int e = ca * NF0 + ia0;
                    const uint32_t w = cur.posw[e >> 1];
                    const int ps = (int)((e & 1) ? (w >> 16) : (w & 0xFFFFu));
                    if (ps != kPosDropped) {
                        T acc = 0;
#pragma unroll
                        for (int j = 0; j < KB; ++j) acc += (LC ? kx[LC ? ca : 0][LC ? j : 0] : cur.k[LC ? 0 : ca][LC ? 0 : j]) * r[j];
                        const double v = dir ? 1.0 : (double)acc;
                        nonfinite += isfinite(v) ? 0 : 1;
                        vals[ps] = v;
                    }
                }
            }
            if (a.has_rhs) {
                const T acc = cur.kf * RF[ib0 * lenFp];
                const double fv = dir ? 0.0 : (double)acc;
                nonfinite += isfinite(fv) ? 0 : 1;
                fvals[threadIdx.x] = fv;
            }
        }
        // geometry rows, slot and row starts of the next block
        __builtin_amdgcn_sched_barrier(0);
        load_k(Mn, nxt);
        __syncthreads();

        // phase B of b
        const int ns = M.ns, nv = M.nv, nr = M.nr, np = M.np;
        for (int u = threadIdx.x; u < ns; u += kBlock) {
            const int cb = sstb[u];
            const int ce = u + 1 < ns ? sstb[u + 1] : nv;
            double sum = 0.0;
            int c = cb;
            for (; c + 4 <= ce; c += 4) {
                const double v0 = vals[c], v1 = vals[c + 1], v2 = vals[c + 2], v3 = vals[c + 3];
                sum += v0;
                sum += v1;
                sum += v2;
                sum += v3;
            }
            for (; c < ce; ++c) sum += vals[c];
            __builtin_nontemporal_store(sum, &a.val[M.s0 + u]);  // written once, never re-read by the pass
        }
        if (a.has_rhs) {
            for (int u = threadIdx.x; u < nr; u += kBlock) {
                const int pb = pstl[u];
                const int pe = u + 1 < nr ? pstl[u + 1] : np;
                double sum = 0.0;
                for (int p = pb; p < pe; ++p) sum += fvals[p];
                a.rhs[M.r0 + u] = sum;
            }
        }
        __syncthreads();
        if (bn < 0) break;
        cur = nxt;
        M = Mn;
        Mn = Mnn;
        b = bn;
        bn = bnn;
    }
    if (nonfinite) atomicAdd(&a.status[0], nonfinite);
}

// Persistent pass over caller-supplied element matrices (NF entries per
// pair): the walk of row_block_persistent_vec with the next block's pair
// records in flight during phase A and its matrix rows, load-vector entries,
// slot and row starts requested after it.  Phase A only places values, so
// the results are bitwise those of row_block_elements_kernel.
template <typename TE, int NF, bool SYM>
__global__ __launch_bounds__(kBlock) void row_block_elements_persistent(BlockArgs a, const TE* __restrict__ Ae,
                                                                        const TE* __restrict__ Fe) {
    constexpr int PW = NF / 2;
    constexpr int SW = 9;
    static_assert(NF % 2 == 0, "positions are read as 32-bit words");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const BlockLds L = block_lds_layout(a.max_values, a.max_pairs, a.max_slots, 8, 0, 0);
    double* vals = reinterpret_cast<double*>(smem_raw + L.vals);
    double* fvals = reinterpret_cast<double*>(smem_raw + L.fvals);
    uint16_t* sst = reinterpret_cast<uint16_t*>(smem_raw + L.sst);
    int* pstl = reinterpret_cast<int*>(smem_raw + L.pst);
    const bool sw_regs = a.max_slots <= 2 * SW * kBlock - 2;

    BlockWalk W;
    W.init(a.n_blocks);
    int64_t b = W.take();
    if (b < 0) return;  // uniform over the workgroup

    struct Regs {
        int32_t tet;
        int ibf;
        uint32_t posw[PW];
        TE v[NF];
        TE fv;
        uint64_t flip;  // bit ia: s_ib s_ia = -1 (orientation signs)
        uint32_t sw[SW];
        int pst;
    };
    auto load_records = [&](const BlockMeta& Mx, Regs& P) {
        const bool hp = (int)threadIdx.x < Mx.np;
        const int64_t gp = Mx.p0 + (hp ? threadIdx.x : 0);
        P.tet = hp ? a.pair_tet[gp] : 0;
        P.ibf = hp ? a.pair_ib[gp] : 0;
        const uint32_t* pw =
            reinterpret_cast<const uint32_t*>(a.pair_pos + Mx.q0 + (int64_t)(hp ? threadIdx.x : 0) * NF);
#pragma unroll
        for (int w = 0; w < PW; ++w) P.posw[w] = hp ? pw[w] : 0u;
    };
    auto load_rows = [&](const BlockMeta& Mx, Regs& P) {
        const bool hp = (int)threadIdx.x < Mx.np;
        const int ib = P.ibf & 63;
        const TE* At = Ae + (int64_t)P.tet * NF * NF + (SYM ? (int64_t)NF * ib : ib);
#pragma unroll
        for (int e = 0; e < NF; ++e) P.v[e] = hp ? At[(SYM ? 1 : NF) * e] : TE(0);
        P.fv = (hp && a.has_rhs) ? Fe[(int64_t)P.tet * NF + ib] : TE(0);
        // orientation signs (assembler.inl:154-164): s_ib on the load entry
        const uint64_t sm = (hp && a.signmask) ? a.signmask[P.tet] : 0;
        const bool nb = (sm >> ib) & 1;
        P.flip = nb ? ~sm : sm;
        if (nb) P.fv = -P.fv;
        if (sw_regs) {
            const uint32_t* sg = reinterpret_cast<const uint32_t*>(a.sstart) + (Mx.s0 >> 1);
            const int nw = slot_words(Mx.s0, (int)Mx.ns);
#pragma unroll
            for (int w = 0; w < SW; ++w) {
                const int i = w * kBlock + (int)threadIdx.x;
                P.sw[w] = i < nw ? sg[i] : 0u;
            }
        }
        P.pst = (a.has_rhs && (int)threadIdx.x < Mx.nr) ? (int)a.pstart[Mx.r0 + threadIdx.x] : 0;
    };

    BlockMeta M = a.meta[b];
    Regs cur, nxt;
    load_records(M, cur);
    load_rows(M, cur);
    int64_t bn = W.take();
    BlockMeta Mn{};
    if (bn >= 0) Mn = a.meta[bn];

    int nonfinite = 0;
    for (;;) {
        const uint16_t* sstb = sst;
        if (sw_regs) {
            uint32_t* s32 = reinterpret_cast<uint32_t*>(sst);
            const int nw = slot_words(M.s0, (int)M.ns);
#pragma unroll
            for (int w = 0; w < SW; ++w) {
                const int i = w * kBlock + (int)threadIdx.x;
                if (i < nw) s32[i] = cur.sw[w];
            }
            sstb = sst + (M.s0 & 1);
        } else {
            block_copy(sst, a.sstart + M.s0, (int)M.ns);
        }
        if (a.has_rhs && (int)threadIdx.x < M.nr) pstl[threadIdx.x] = cur.pst;
        load_records(Mn, nxt);  // past the last block: Mn has no pairs, nothing is read
        const int64_t bnn = W.take();
        BlockMeta Mnn{};
        if (bnn >= 0) Mnn = a.meta[bnn];

        // phase A of b: the pair's row placed at its listed positions
        if ((int)threadIdx.x < M.np) {
            const bool dir = (cur.ibf & kPairDirichlet) != 0;
#pragma unroll
            for (int e = 0; e < NF; ++e) {
                const uint32_t w = cur.posw[e >> 1];
                const int ps = (int)((e & 1) ? (w >> 16) : (w & 0xFFFFu));
                if (ps != kPosDropped) {
                    const double v = dir ? 1.0 : ((cur.flip >> e) & 1 ? -(double)cur.v[e] : (double)cur.v[e]);
                    nonfinite += isfinite(v) ? 0 : 1;
                    vals[ps] = v;
                }
            }
            if (a.has_rhs) {
                const double fv = dir ? 0.0 : (double)cur.fv;
                nonfinite += isfinite(fv) ? 0 : 1;
                fvals[threadIdx.x] = fv;
            }
        }
        load_rows(Mn, nxt);
        __syncthreads();

        // phase B of b
        const int ns = M.ns, nv = M.nv, nr = M.nr, np = M.np;
        for (int u = threadIdx.x; u < ns; u += kBlock) {
            const int cb = sstb[u];
            const int ce = u + 1 < ns ? sstb[u + 1] : nv;
            double sum = 0.0;
            int c = cb;
            for (; c + 4 <= ce; c += 4) {
                const double v0 = vals[c], v1 = vals[c + 1], v2 = vals[c + 2], v3 = vals[c + 3];
                sum += v0;
                sum += v1;
                sum += v2;
                sum += v3;
            }
            for (; c < ce; ++c) sum += vals[c];
            __builtin_nontemporal_store(sum, &a.val[M.s0 + u]);  // written once, never re-read by the pass
        }
        if (a.has_rhs) {
            for (int u = threadIdx.x; u < nr; u += kBlock) {
                const int pb = pstl[u];
                const int pe = u + 1 < nr ? pstl[u + 1] : np;
                double sum = 0.0;
                for (int p = pb; p < pe; ++p) sum += fvals[p];
                a.rhs[M.r0 + u] = sum;
            }
        }
        __syncthreads();
        if (bn < 0) break;
        cur = nxt;
        M = Mn;
        Mn = Mnn;
        b = bn;
        bn = bnn;
    }
    if (nonfinite) atomicAdd(&a.status[0], nonfinite);
}

}  // namespace anifem
