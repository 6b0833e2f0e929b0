// Device code of the element evaluator, shared by the batch entry point
// (element.hip) and the fused colour-by-colour assembly (assembly.hip).
//
// Work decomposition (one workgroup of 256 lanes = 4 wavefronts per tile of
// TILE tets):
//   phase 1  one lane per tet: FP64 Jacobian J = [P1-P0, P2-P0, P3-P0], its
//            inverse PSI and |T| = |det J|/6 (core.inl:218-248, geometry.h:25-45),
//            then the geometry tensor K = |T| T_B^T D T_A per component block
//            (and per quadrature point when D varies inside the tet), stored
//            to LDS.
//   phase 2  lanes own element entries e = ib + nfB*ia; a group of G lanes
//            covers all entries of one tet, NG = 256/G groups walk the tile's
//            tets.  Each lane keeps its entry's reference-tensor row in
//            registers and reads K from LDS (a broadcast within a group for
//            scalar spaces).  Output rows of one tet are contiguous, so a
//            group's stores (or slot-map loads) are coalesced.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/anifem_gpu.h"

namespace anifem {

constexpr int kBlock = 256;

struct ElemArgs {
    int64_t n;         // tets in this launch
    int64_t coef_n;    // SoA stride of per-tet / per-point coefficient arrays
    int nfA0, nfB0;    // scalar basis sizes
    int E;             // nfA*nfB entries per tet
    int q;             // points carried in K (1 when D is constant over the tet)
    int ttype, storage;
    int dimA, dimB;    // full operator dims (ncomp * ref dim)
    int p0_bcast;      // IDEN P0 trial against a vector test: D column replicated
    int tile;          // tets per workgroup
    // geometry source: either SoA coordinates (batch mode) or mesh gather
    const double* xyz;        // 12 x n
    const double* coords;     // n_nodes x 3
    const int32_t* tetnodes;  // 4 x nodes_stride (SoA), assembly order
    int64_t nodes_stride;     // SoA stride of tetnodes (all tets of the mesh)
    int64_t pos0;             // assembly position of this launch's first tet
    const int32_t* orig;      // assembly position -> caller's tet index (coefficients), or null
    const double* D;
    const double* Rd;
    const float* Rf;
    const double* wq;  // quadrature weights (per-point K)
    void* out;         // batch mode: element matrices
};

__device__ __forceinline__ void tet_geometry(const double P[4][3], double PSI[9], double& det) {
    double m[9];
    for (int l = 0; l < 3; ++l)
        for (int i = 0; i < 3; ++i) m[i + 3 * l] = P[l + 1][i] - P[0][i];
    // cofactor inverse, column-major (geometry.h:25-45)
    double da = m[4] * m[8] - m[7] * m[5];
    double db = m[7] * m[2] - m[1] * m[8];
    double dc = m[1] * m[5] - m[4] * m[2];
    det = m[0] * da + m[3] * db + m[6] * dc;
    double rd = 1.0 / det;
    PSI[0] = da * rd;
    PSI[1] = db * rd;
    PSI[2] = dc * rd;
    PSI[3] = (m[5] * m[6] - m[3] * m[8]) * rd;
    PSI[4] = (m[8] * m[0] - m[6] * m[2]) * rd;
    PSI[5] = (m[2] * m[3] - m[0] * m[5]) * rd;
    PSI[6] = (m[3] * m[7] - m[6] * m[4]) * rd;
    PSI[7] = (m[6] * m[1] - m[0] * m[7]) * rd;
    PSI[8] = (m[0] * m[4] - m[3] * m[1]) * rd;
}

// Value D(k, j) of the dimB x dimA tensor at (tet, point), k row, j column.
__device__ __forceinline__ double tensor_at(const ElemArgs& a, int64_t cr, int n, int k, int j) {
    const int64_t N = a.coef_n;
    auto ld = [&](int v) -> double {
        if (a.storage == ANIFEM_COEF_CONST) return a.D[v];
        if (a.storage == ANIFEM_COEF_PER_TET) return a.D[(int64_t)v * N + cr];
        return a.D[((int64_t)v * a.q + n) * N + cr];
    };
    switch (a.ttype) {
        case ANIFEM_TENSOR_NULL: return (a.p0_bcast || k == j) ? 1.0 : 0.0;
        case ANIFEM_TENSOR_SCALAR: return (a.p0_bcast || k == j) ? ld(0) : 0.0;
        case ANIFEM_TENSOR_SYMMETRIC: return ld(k <= j ? k + j * (j + 1) / 2 : j + k * (k + 1) / 2);
        case ANIFEM_TENSOR_GENERAL: return ld(k + a.dimB * j);
        case ANIFEM_TENSOR_LAME: {
            // row (i, jj) = (k%3, k/3), column (kk, l) = (j%3, j/3)
            int i = k % 3, jj = k / 3, kk = j % 3, l = j / 3;
            double lam = ld(0), mu = ld(1);
            return lam * (i == jj && kk == l) + mu * ((i == kk && jj == l) + (i == l && jj == kk));
        }
    }
    return 0.0;
}

// K[(cb*NCA + ca)*q*KB + n*KB + b*DA0 + a] for one tet.
template <int DA0, int DB0, int NCA, int NCB, typename KT>
__device__ __forceinline__ void geometry_tensor(const ElemArgs& a, int64_t cr, const double PSI[9], double mes,
                                                KT* K) {
    constexpr int KB = DA0 * DB0;
    for (int n = 0; n < a.q; ++n) {
        const double s = (a.q > 1) ? mes * a.wq[n] : mes;
        for (int cb = 0; cb < NCB; ++cb)
            for (int ca = 0; ca < NCA; ++ca) {
                double Dblk[DB0][DA0];
                for (int k0 = 0; k0 < DB0; ++k0)
                    for (int j0 = 0; j0 < DA0; ++j0) Dblk[k0][j0] = tensor_at(a, cr, n, k0 + DB0 * cb, j0 + DA0 * ca);
                // K(b, a) = s * sum_{k0, j0} TB(k0, b) D(k0, j0) TA(j0, a); T = PSI^T for GRAD
                double M[DB0][DA0];  // (D * TA)
                for (int k0 = 0; k0 < DB0; ++k0)
                    for (int a0 = 0; a0 < DA0; ++a0) {
                        double acc = 0;
                        if (DA0 == 3) {
                            for (int j0 = 0; j0 < 3; ++j0) acc += Dblk[k0][j0] * PSI[a0 + 3 * j0];
                        } else {
                            acc = Dblk[k0][0];
                        }
                        M[k0][a0] = acc;
                    }
                KT* Kb = K + ((cb * NCA + ca) * a.q + n) * KB;
                for (int b0 = 0; b0 < DB0; ++b0)
                    for (int a0 = 0; a0 < DA0; ++a0) {
                        double acc = 0;
                        if (DB0 == 3) {
                            for (int k0 = 0; k0 < 3; ++k0) acc += PSI[b0 + 3 * k0] * M[k0][a0];
                        } else {
                            acc = M[0][a0];
                        }
                        Kb[b0 * DA0 + a0] = (KT)(s * acc);
                    }
            }
    }
}

template <bool MESH>
__device__ __forceinline__ void load_tet(const ElemArgs& a, int64_t r, double P[4][3]) {
    if (MESH) {
        for (int v = 0; v < 4; ++v) {
            int32_t node = a.tetnodes[(int64_t)v * a.nodes_stride + a.pos0 + r];
            for (int c = 0; c < 3; ++c) P[v][c] = a.coords[3 * (int64_t)node + c];
        }
    } else {
        for (int v = 0; v < 4; ++v)
            for (int c = 0; c < 3; ++c) P[v][c] = a.xyz[(int64_t)(3 * v + c) * a.n + r];
    }
}

// Element matrices of a tile of tets, stored in T precision to a.out
// (n_tets matrices of nfB x nfA, column-major, one after the other).
template <int DA0, int DB0, int NCA, int NCB, typename T>
__global__ __launch_bounds__(kBlock) void element_kernel(ElemArgs a) {
    constexpr int KB = DA0 * DB0;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* Ks = reinterpret_cast<T*>(smem_raw);
    const int kstride = NCA * NCB * a.q * KB;
    const int64_t tile0 = (int64_t)blockIdx.x * a.tile;
    const int ntile = (int)((a.n - tile0) < a.tile ? (a.n - tile0) : a.tile);

    for (int t = threadIdx.x; t < ntile; t += kBlock) {
        const int64_t r = tile0 + t;
        double P[4][3], PSI[9], det;
        load_tet<false>(a, r, P);
        tet_geometry(P, PSI, det);
        const double mes = fabs(det) / 6.0;
        geometry_tensor<DA0, DB0, NCA, NCB, T>(a, r, PSI, mes, Ks + (int64_t)t * kstride);
    }
    __syncthreads();

    const int E = a.E;
    const int G = E >= kBlock ? kBlock : ((E + 15) / 16) * 16;
    const int NG = kBlock / G;
    const int es = threadIdx.x % G, tg = threadIdx.x / G;
    if (tg >= NG) return;
    const int nfB = a.nfB0 * NCB;
    const T* R = nullptr;
    if constexpr (sizeof(T) == 8) R = reinterpret_cast<const T*>(a.Rd);
    else R = reinterpret_cast<const T*>(a.Rf);

    for (int e = es; e < E; e += G) {
        const int ib = e % nfB, ia = e / nfB;
        const int ib0 = ib % a.nfB0, cb = ib / a.nfB0;
        const int ia0 = ia % a.nfA0, ca = ia / a.nfA0;
        const int rr = ib0 + a.nfB0 * ia0;
        const int koff = (cb * NCA + ca) * a.q * KB;
        T Rreg[KB];
        if (a.q == 1) {
            for (int kk = 0; kk < KB; ++kk) Rreg[kk] = R[rr * KB + kk];
        }
        for (int t = tg; t < ntile; t += NG) {
            const T* Kt = Ks + (int64_t)t * kstride + koff;
            T acc = 0;
            if (a.q == 1) {
                for (int kk = 0; kk < KB; ++kk) acc += Kt[kk] * Rreg[kk];
            } else {
                const T* Rn = R + (int64_t)rr * a.q * KB;
                for (int j = 0; j < a.q * KB; ++j) acc += Kt[j] * Rn[j];
            }
            static_cast<T*>(a.out)[(tile0 + t) * E + e] = acc;
        }
    }
}

}  // namespace anifem
