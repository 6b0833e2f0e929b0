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
#include "basis_eval.h"

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
    int selA, selB;    // operator codes: DIV / CURL / DUDX..DUDZ select from the gradient (op_map)
    int mapA, mapB;    // 0: T = PSI^T (or 1 for dimension 1); 1: T = J (Piola map of RT0 values, ND0 curls)
    int scA, scB;      // per-DOF factor of the basis (dof_scale): 0 none, 1 RT0, 2 ND0 values, 3 ND0 curls
    int face;          // -1: volume integral; 0..3: integral over that face (int_face.inl)
    int dot_normal;    // face integral of (D OpA(u)) . n against OpB (fem3DfaceN): D has 3 dimB rows
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
    const double* Rds;  // R for a symmetric K in packed form (RefTensorDev::rs64), or null
    const float* Rfs;
    const double* wq;  // quadrature weights (per-point K)
    void* out;         // batch mode: element matrices
    const void* r16;   // FP16 hi + lo split of R / r16_scale ([2][e0pad][16])
    float r16_scale;
    int e0pad;
    int rds_sym;       // Rds is symmetric in (ia, ib) bit for bit (RefTensorDev::rs_sym)
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

// JT[a + 3j] = J(j, a) = P[a+1][j] - P[0][j]: the map of contravariant
// (Piola) reference fields, x = P0 + J xi.
__device__ __forceinline__ void tet_jacobian_t(const double P[4][3], double JT[9]) {
    for (int a = 0; a < 3; ++a)
        for (int j = 0; j < 3; ++j) JT[a + 3 * j] = P[a + 1][j] - P[0][j];
}

// Per-DOF factor of basis function i of the vector-valued spaces:
// 1: RT0, |face i| / (|det J| / 2), face i through vertices i, i+1, i+2
//    (raviant_thomas_0.h:35-45); 2: ND0 values, |edge i| (nedelec_0.h:37-46);
// 3: ND0 curls, |edge i| / det J (nedelec_0.h:111-122); edges 01,02,03,12,13,23.
__device__ __forceinline__ double dof_scale(int kind, const double P[4][3], double det, int i) {
    if (kind == 1) {
        const double* p0 = P[i];
        const double* p1 = P[(i + 1) % 4];
        const double* p2 = P[(i + 2) % 4];
        const double u[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const double v[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        const double c[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        return sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) / fabs(det);
    }
    const int ea = i < 3 ? 0 : (i < 5 ? 1 : 2);
    const int eb = i < 3 ? i + 1 : (i == 3 ? 2 : 3);
    const double d0 = P[eb][0] - P[ea][0], d1 = P[eb][1] - P[ea][1], d2 = P[eb][2] - P[ea][2];
    const double len = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    return kind == 3 ? len / det : len;
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

// Component of a gradient-based operator that gradient entry g = d + 3c
// (d/dx_d of component c) feeds, with its sign, or -1 (operators.h:321-400):
// DIV: div u = sum_c du_c/dx_c; CURL: (curl u)_i = eps_{idc} du_c/dx_d;
// DUDX..DUDZ: du_c/dx_{0,1,2}; GRAD and IDEN: g itself.
__device__ __forceinline__ int op_map(int op, int g, int& sign) {
    sign = 1;
    const int d = g % 3, c = g / 3;
    switch (op) {
        case ANIFEM_DIV: return d == c ? 0 : -1;
        case ANIFEM_CURL: {
            if (d == c) return -1;
            const int i = 3 - d - c;
            sign = (d == (i + 1) % 3) ? 1 : -1;
            return i;
        }
        case ANIFEM_DUDX:
        case ANIFEM_DUDY:
        case ANIFEM_DUDZ: return d == op - ANIFEM_DUDX ? c : -1;
        default: return g;
    }
}

// Tensor entry between gradient entries (or values) k of the test side and j
// of the trial side: the operators' selectors applied to D.
__device__ __forceinline__ double tensor_sel(const ElemArgs& a, int64_t cr, int n, int k, int j) {
    if (a.selA <= ANIFEM_GRAD && a.selB <= ANIFEM_GRAD) return tensor_at(a, cr, n, k, j);
    int sA, sB;
    const int mA = op_map(a.selA, j, sA), mB = op_map(a.selB, k, sB);
    if (mA < 0 || mB < 0) return 0.0;
    const double v = tensor_at(a, cr, n, mB, mA);
    return sA * sB > 0 ? v : -v;
}

// Area of face f of a tet (vertices f, f+1, f+2 mod 4) and, when nrm is given,
// its unit normal pointing away from the fourth vertex (geometry.h:67-93).
__device__ __forceinline__ double face_measure(const double P[4][3], int f, double* nrm) {
    const double* p0 = P[f];
    const double* p1 = P[(f + 1) % 4];
    const double* p2 = P[(f + 2) % 4];
    const double u[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double v[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const double c[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    const double len = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    if (nrm) {
        const double* p3 = P[(f + 3) % 4];
        const double h = c[0] * (p3[0] - p0[0]) + c[1] * (p3[1] - p0[1]) + c[2] * (p3[2] - p0[2]);
        const double sgn = h > 0 ? -1.0 : 1.0;
        for (int i = 0; i < 3; ++i) nrm[i] = sgn * c[i] / len;
    }
    return len / 2;
}

// K block (cb, ca) at point n of one tet, scaled by s:
// Kb[b*DA0 + a] = s * sum_{k0, j0} TB(k0, b) D(k0 + DB0 cb, j0 + DA0 ca) TA(j0, a),
// T = PSI^T for GRAD, 1 for IDEN, J for the Piola-mapped fields (map 1, with
// JT given).  With a face normal nrm, D(k, j) is the normal contraction
// sum_d D(d + 3k, j) nrm[d] of the 3 dimB x dimA tensor (core.h:176-195).
// 9 x DA0 map of a GRAD of RT0 (map 3: T = vec(I), DA0 = 1) or ND0 (map 4:
// T = S J, DA0 = 3, S the vector -> skew-matrix map in the entry order of
// nedelec_0.h:175-179: entries 1, 2, 3, 5, 6, 7 take -a2, a1, a2, -a0, -a1, a0).
__device__ __forceinline__ double grad9_map(int map, const double* JT, int l, int c) {
    if (map == 3) return (l == 0 || l == 4 || l == 8) ? 1.0 : 0.0;
    // row l of S: (component m of the axial vector, sign)
    const int m = l == 1 || l == 3 ? 2 : (l == 2 || l == 6 ? 1 : (l == 5 || l == 7 ? 0 : -1));
    if (m < 0) return 0.0;
    const double sg = (l == 2 || l == 3 || l == 7) ? 1.0 : -1.0;
    return sg * JT[c + 3 * m];  // J(m, c)
}

template <int DA0, int DB0, typename KT>
__device__ __forceinline__ void geometry_block(const ElemArgs& a, int64_t cr, const double PSI[9], double s, int n,
                                               int cb, int ca, KT* Kb, const double* nrm = nullptr,
                                               const double* JT = nullptr) {
    if (JT && a.mapA >= 3) {
        // GRAD of RT0 / ND0 on both sides: 9 physical components,
        // Kb = s TB^T D TA with the 9 x 9 tensor and the 9 x D0 maps
        double M[9][DA0];
        for (int k = 0; k < 9; ++k)
            for (int a0 = 0; a0 < DA0; ++a0) M[k][a0] = 0.0;
        for (int j = 0; j < 9; ++j) {
            double tA[DA0];
            bool any = false;
            for (int a0 = 0; a0 < DA0; ++a0) {
                tA[a0] = grad9_map(a.mapA, JT, j, a0);
                any = any || tA[a0] != 0.0;
            }
            if (!any) continue;
            for (int k = 0; k < 9; ++k) {
                const double d = tensor_sel(a, cr, n, k, j);
                for (int a0 = 0; a0 < DA0; ++a0) M[k][a0] += d * tA[a0];
            }
        }
        for (int b0 = 0; b0 < DB0; ++b0)
            for (int a0 = 0; a0 < DA0; ++a0) {
                double acc = 0;
                for (int k = 0; k < 9; ++k) acc += grad9_map(a.mapB, JT, k, b0) * M[k][a0];
                Kb[b0 * DA0 + a0] = (KT)(s * acc);
            }
        return;
    }
    const double* TAm = (JT && a.mapA) ? JT : PSI;
    const double* TBm = (JT && a.mapB) ? JT : PSI;
    double Dblk[DB0][DA0];
    for (int k0 = 0; k0 < DB0; ++k0)
        for (int j0 = 0; j0 < DA0; ++j0) {
            if (nrm) {
                const int k = k0 + DB0 * cb;
                double v = 0;
                for (int d = 0; d < 3; ++d) v += tensor_sel(a, cr, n, d + 3 * k, j0 + DA0 * ca) * nrm[d];
                Dblk[k0][j0] = v;
            } else {
                Dblk[k0][j0] = tensor_sel(a, cr, n, k0 + DB0 * cb, j0 + DA0 * ca);
            }
        }
    double M[DB0][DA0];  // (D * TA)
    for (int k0 = 0; k0 < DB0; ++k0)
        for (int a0 = 0; a0 < DA0; ++a0) {
            double acc = 0;
            if (DA0 == 3) {
                for (int j0 = 0; j0 < 3; ++j0) acc += Dblk[k0][j0] * TAm[a0 + 3 * j0];
            } else {
                acc = Dblk[k0][0];
            }
            M[k0][a0] = acc;
        }
    for (int b0 = 0; b0 < DB0; ++b0)
        for (int a0 = 0; a0 < DA0; ++a0) {
            double acc = 0;
            if (DB0 == 3) {
                for (int k0 = 0; k0 < 3; ++k0) acc += TBm[b0 + 3 * k0] * M[k0][a0];
            } else {
                acc = M[0][a0];
            }
            Kb[b0 * DA0 + a0] = (KT)(s * acc);
        }
}

// K[(cb*NCA + ca)*q*KB + n*KB + b*DA0 + a] for one tet.
template <int DA0, int DB0, int NCA, int NCB, typename KT>
__device__ __forceinline__ void geometry_tensor(const ElemArgs& a, int64_t cr, const double PSI[9], double mes,
                                                KT* K, const double* nrm = nullptr, const double* JT = nullptr) {
    constexpr int KB = DA0 * DB0;
    for (int n = 0; n < a.q; ++n) {
        const double s = (a.q > 1) ? mes * a.wq[n] : mes;
        for (int cb = 0; cb < NCB; ++cb)
            for (int ca = 0; ca < NCA; ++ca)
                geometry_block<DA0, DB0, KT>(a, cr, PSI, s, n, cb, ca, K + ((cb * NCA + ca) * a.q + n) * KB, nrm,
                                             JT);
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

// LDS byte offset of the per-DOF factors that follow the K tensors of a tile.
__host__ __device__ inline size_t element_scale_offset(int tile, int kstride, int elem_bytes) {
    return ((size_t)tile * kstride * elem_bytes + 15) / 16 * 16;
}

// Element matrices of a tile of tets, stored in T precision to a.out
// (n_tets matrices of nfB x nfA, column-major, one after the other).
// PIOLA (RT0, ND0 on either side): the J map of geometry_block, and a per-tet
// factor per basis function applied to the contracted entry,
// A(ib, ia) = sB_ib sA_ia sum_kk K R; a separate instantiation keeps the
// registers of the other spaces' kernel unchanged.
template <int DA0, int DB0, int NCA, int NCB, typename T, bool PIOLA = false>
__global__ __launch_bounds__(kBlock) void element_kernel(ElemArgs a) {
    constexpr int KB = DA0 * DB0;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* Ks = reinterpret_cast<T*>(smem_raw);
    const int kstride = NCA * NCB * a.q * KB;
    const int64_t tile0 = (int64_t)blockIdx.x * a.tile;
    const int ntile = (int)((a.n - tile0) < a.tile ? (a.n - tile0) : a.tile);
    const bool scaled = PIOLA && (a.scA | a.scB) != 0;
    const int ns = a.nfA0 + a.nfB0;
    double* Ss = reinterpret_cast<double*>(smem_raw + element_scale_offset(a.tile, kstride, (int)sizeof(T)));

    for (int t = threadIdx.x; t < ntile; t += kBlock) {
        const int64_t r = tile0 + t;
        double P[4][3], PSI[9], det, nrm[3];
        load_tet<false>(a, r, P);
        tet_geometry(P, PSI, det);
        // volume |det|/6, or the face's area (int_face.inl:40-43, 79-84)
        const double mes = a.face < 0 ? fabs(det) / 6.0 : face_measure(P, a.face, a.dot_normal ? nrm : nullptr);
        if constexpr (PIOLA) {
            double JT[9];
            tet_jacobian_t(P, JT);
            geometry_tensor<DA0, DB0, NCA, NCB, T>(a, r, PSI, mes, Ks + (int64_t)t * kstride,
                                                   a.face >= 0 && a.dot_normal ? nrm : nullptr, JT);
        } else {
            geometry_tensor<DA0, DB0, NCA, NCB, T>(a, r, PSI, mes, Ks + (int64_t)t * kstride,
                                                   a.face >= 0 && a.dot_normal ? nrm : nullptr);
        }
        if (scaled) {
            double* St = Ss + (int64_t)t * ns;
            for (int i = 0; i < a.nfA0; ++i) St[i] = a.scA ? dof_scale(a.scA, P, det, i) : 1.0;
            for (int i = 0; i < a.nfB0; ++i) St[a.nfA0 + i] = a.scB ? dof_scale(a.scB, P, det, i) : 1.0;
        }
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
            if constexpr (PIOLA) {
                if (scaled) {
                    const double* St = Ss + (int64_t)t * ns;
                    acc = (T)((double)acc * (St[ia0] * St[a.nfA0 + ib0]));
                }
            }
            static_cast<T*>(a.out)[(tile0 + t) * E + e] = acc;
        }
    }
}

// ---- solution evaluation: fem3DapplyL (eval.inl:84-100, core.inl:370-405) ----
struct ApplyArgs {
    int d;                 // reference components per basis function and component (1 or 3)
    int map;               // 0: as tabulated; 1: PSI^T (gradients, ND0 values); 2: J (RT0 values, ND0 curls);
                           // 3 / 4: GRAD of RT0 / ND0 (9 components, grad9_map)
    int scale;             // per-DOF factor (dof_scale kind), 0 none
    int op, fem;           // operator and space: the basis is evaluated on the device in X mode
    int nf0, ncomp, q;     // basis size per component, components, points per tet
    const double* tab;     // L mode: reference table at the common points, tab[(n*nf0 + i)*d + a]
    const double* X;       // X mode: physical points X[(r*q + n)*3 + c] (tab unused)
    int64_t n;             // tets
    const double* xyz;     // batch: 12 x n SoA coordinates
    const double* coords;  // mesh: node coordinates (n_nodes x 3) ...
    const int32_t* tetnodes;  // ... and 4 x n SoA node lists
    const double* dofs;    // batch: per-tet DOF values, dofs[r*nf + i]
    const int32_t* dofmap; // mesh: per-tet global DOF ids, dofmap[r*nf + i] ...
    const int8_t* sign;    // ... their orientation signs (oriented spaces), or null ...
    const double* u;       // ... into the global vector u
    double* out;           // opU[r*dim*q + k + dim*n]
};

// One lane per (tet, point): Op(u)(x_n) = sum_i s_i T U_i(x_n) dof_i with the
// reference values U_i tabulated (fem3DapplyL, barycentric points common to
// all tets) or evaluated here at the barycentric coordinates of the tet's own
// physical point (fem3DapplyX: l_{k+1} = sum_j PSI(k, j) (x_j - P0_j),
// l_0 = 1 - l_1 - l_2 - l_3, core.inl:120-137); T = 1, PSI^T or J and s_i the
// per-DOF factor of RT0 / ND0 (component k + 3c = du_c/dx_k for GRAD of a
// vector space).
template <bool MESH>
__global__ __launch_bounds__(kBlock) void apply_kernel(ApplyArgs a) {
    const int64_t idx = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (idx >= a.n * a.q) return;
    const int64_t r = idx / a.q;
    const int pt = (int)(idx % a.q);
    const int nf = a.nf0 * a.ncomp, d = a.d, dim = a.map >= 3 ? 9 : d * a.ncomp;
    double P[4][3], PSI[9], JT[9], det = 1.0;
    if (a.map != 0 || a.scale != 0 || a.X) {
        for (int v = 0; v < 4; ++v)
            for (int c = 0; c < 3; ++c)
                P[v][c] = MESH ? a.coords[3 * (int64_t)a.tetnodes[(int64_t)v * a.n + r] + c]
                               : a.xyz[(int64_t)(3 * v + c) * a.n + r];
        tet_geometry(P, PSI, det);
        tet_jacobian_t(P, JT);
    }
    double loc[60];
    const double* T = a.tab ? a.tab + (int64_t)pt * a.nf0 * d : loc;
    if (a.X) {
        const double* x = a.X + (r * a.q + pt) * 3;
        const double dr[3] = {x[0] - P[0][0], x[1] - P[0][1], x[2] - P[0][2]};
        double L[4];
        for (int i = 0; i < 3; ++i) L[i + 1] = PSI[i] * dr[0] + PSI[i + 3] * dr[1] + PSI[i + 6] * dr[2];
        L[0] = 1.0 - L[1] - L[2] - L[3];
        if (a.fem == ANIFEM_FEM_RT0 || a.fem == ANIFEM_FEM_ND0) vector_field(a.op, a.fem, L, loc);
        else if (a.op == ANIFEM_GRAD) basis_gradients(a.fem, L, loc);
        else basis_values(a.fem, L, loc);
    }
    double* o = a.out + r * dim * a.q + (int64_t)dim * pt;
    for (int c = 0; c < a.ncomp; ++c) {
        double g[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < a.nf0; ++i) {
            const int64_t e = r * nf + c * a.nf0 + i;
            double dv = MESH ? a.u[a.dofmap[e]] : a.dofs[e];
            if (MESH && a.sign) dv *= (double)a.sign[e];
            if (a.scale) dv *= dof_scale(a.scale, P, det, i);
            for (int j = 0; j < d; ++j) g[j] += T[i * d + j] * dv;
        }
        if (a.map >= 3) {  // GRAD of RT0 (vec(I) g) / ND0 (S J g): 9 components
            for (int l = 0; l < 9; ++l) {
                double v = 0.0;
                for (int j = 0; j < d; ++j) v += grad9_map(a.map, JT, l, j) * g[j];
                o[l] = v;
            }
        } else if (a.map == 1) {
            for (int k = 0; k < 3; ++k) o[k + 3 * c] = PSI[0 + 3 * k] * g[0] + PSI[1 + 3 * k] * g[1] + PSI[2 + 3 * k] * g[2];
        } else if (a.map == 2) {
            for (int k = 0; k < 3; ++k) o[k + 3 * c] = JT[0 + 3 * k] * g[0] + JT[1 + 3 * k] * g[1] + JT[2 + 3 * k] * g[2];
        } else {
            for (int j = 0; j < d; ++j) o[j + d * c] = g[j];
        }
    }
}

// ---- split-precision MFMA contraction (ANIFEM_PREC_FP16X3) -----------------
// For a tensor constant over each tet the element matrices of a block of tets
// are one product per component block (cb, ca):
//   C[e0][t] = sum_kk R[e0][kk] K_t[kk],  e0 = ib0 + nfB0 ia0 < E0, kk < KB <= 16,
// run on v_mfma_f32_16x16x16f16 (16 entries x 16 tets x 16 kk per instruction).
// R / max|R| and K_t / max|K_t| lie in [-1, 1]; each is split as hi = fp16(x),
// lo = fp16(x - hi) (22 significant bits), and C = Rl Kh + Rh Kl + Rh Kh in
// FP32 (the Rl Kl term, below 2^-22 of |R||K|, is dropped), then scaled back.
// Operand layouts of the 16x16x16 instruction, lane l (j = l % 16, g = l / 16):
//   A (16 x 16): row j, columns 4g..4g+3;  B (16 x 16): rows 4g..4g+3, column j;
//   D (16 x 16): rows 4g..4g+3, column j.
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kMfmaTets = 256;  // tets per workgroup: 64 per wavefront, in 4 groups of 16

template <int DA0, int DB0, int NCA, int NCB>
__global__ __launch_bounds__(kBlock) void element_mfma_kernel(ElemArgs a) {
    constexpr int KB = DA0 * DB0, NB = NCA * NCB;
    static_assert(KB <= 16, "the 16-wide kk dimension holds K");
    // one component block at a time: K_t / max|K_t| of that block, hi and lo
    __shared__ __align__(16) _Float16 Kh[kMfmaTets][16], Kl[kMfmaTets][16];
    __shared__ float ks[kMfmaTets];
    const int64_t tile0 = (int64_t)blockIdx.x * kMfmaTets;
    const int ntile = (int)((a.n - tile0) < kMfmaTets ? (a.n - tile0) : kMfmaTets);
    // FP64 geometry of this lane's tet, kept in registers across the blocks
    const int t = threadIdx.x;
    const int64_t r = tile0 + t;
    double PSI[9], mes = 0;
    if (t < ntile) {
        double P[4][3], det;
        load_tet<false>(a, r, P);
        tet_geometry(P, PSI, det);
        mes = fabs(det) / 6.0;
    }
    const int w = threadIdx.x / 64, l = threadIdx.x % 64, j = l % 16, g = l / 16;
    const int E = a.E, nfB0 = a.nfB0, nfA0 = a.nfA0, nfB = nfB0 * NCB;
    const int E0 = nfA0 * nfB0;
    // a lane's 4 entries are 16 contiguous, aligned bytes
    const bool vec4 = NB == 1 && (E0 % 4) == 0 && (reinterpret_cast<uintptr_t>(a.out) & 15) == 0;
    const _Float16* Rh = static_cast<const _Float16*>(a.r16);
    const _Float16* Rl = Rh + (int64_t)a.e0pad * 16;
    for (int b = 0; b < NB; ++b) {
        const int cb = b / NCA, ca = b % NCA;
        {
            double Kb[KB];
            double mx = 0;
            if (t < ntile) {
                geometry_block<DA0, DB0, double>(a, r, PSI, mes, 0, cb, ca, Kb);
                for (int i = 0; i < KB; ++i) mx = fmax(mx, fabs(Kb[i]));
            } else {
                for (int i = 0; i < KB; ++i) Kb[i] = 0.0;
            }
            const double sc = mx > 0 ? mx : 1.0;
            ks[t] = (float)sc;
            for (int kk = 0; kk < 16; ++kk) {
                const float x = kk < KB ? (float)(Kb[kk < KB ? kk : 0] / sc) : 0.0f;
                const _Float16 h = (_Float16)x;
                Kh[t][kk] = h;
                Kl[t][kk] = (_Float16)(x - (float)h);
            }
        }
        __syncthreads();
        // wavefront w: tets 64w + 16q + j, q = 0..3
        half4_t bh[4], bl[4];
        float sc[4];
        bool live[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int tl = 64 * w + 16 * q + j;
            bh[q] = *reinterpret_cast<const half4_t*>(&Kh[tl][4 * g]);
            bl[q] = *reinterpret_cast<const half4_t*>(&Kl[tl][4 * g]);
            sc[q] = ks[tl] * a.r16_scale;
            live[q] = tl < ntile;
        }
        for (int et = 0; et < a.e0pad / 16; ++et) {
            const int row = et * 16 + j;
            const half4_t ah = *reinterpret_cast<const half4_t*>(Rh + row * 16 + 4 * g);
            const half4_t al = *reinterpret_cast<const half4_t*>(Rl + row * 16 + 4 * g);
            const int e00 = et * 16 + 4 * g;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float4_t c = {0.0f, 0.0f, 0.0f, 0.0f};
                c = __builtin_amdgcn_mfma_f32_16x16x16f16(al, bh[q], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bl[q], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bh[q], c, 0, 0, 0);
                if (!live[q]) continue;
                float* out = static_cast<float*>(a.out) + (tile0 + 64 * w + 16 * q + j) * E;
                if (vec4) {
                    if (e00 < E0) {
                        const float4_t o = {c[0] * sc[q], c[1] * sc[q], c[2] * sc[q], c[3] * sc[q]};
                        *reinterpret_cast<float4_t*>(out + e00) = o;
                    }
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int e0 = e00 + v;
                        if (e0 < E0) {
                            const int ib0 = e0 % nfB0, ia0 = e0 / nfB0;
                            out[(ib0 + nfB0 * cb) + nfB * (ia0 + nfA0 * ca)] = c[v] * sc[q];
                        }
                    }
                }
            }
        }
        __syncthreads();  // Kh / Kl / ks are rewritten for the next block
    }
}

// Vector spaces (NCA * NCB > 1): 64 tets per workgroup, all component blocks
// of K_t (each scaled by its own max|K|) split into LDS, four lanes per tet.
// The MFMA tiles then walk the output by columns: tile h of column
// ia = ia0 + nfA0 ca covers rows ib = 16h .. 16h + 15, whose component rows
// cb = ib / nfB0 are one k-step each (the A rows of other components are zero),
// so every column of every tet is written once, contiguously.
constexpr int kMfmaVecTets = 64;

template <int DA0, int DB0, int NCA, int NCB>
__global__ __launch_bounds__(kBlock) void element_mfma_vec_kernel(ElemArgs a) {
    constexpr int KB = DA0 * DB0, NB = NCA * NCB;
    static_assert(KB <= 16, "the 16-wide kk dimension holds K");
    __shared__ __align__(16) _Float16 Kh[NB][kMfmaVecTets][16], Kl[NB][kMfmaVecTets][16];
    __shared__ float ks[NB][kMfmaVecTets];
    const int64_t tile0 = (int64_t)blockIdx.x * kMfmaVecTets;
    const int ntile = (int)((a.n - tile0) < kMfmaVecTets ? (a.n - tile0) : kMfmaVecTets);
    {
        const int t = threadIdx.x / 4, part = threadIdx.x % 4;
        const int64_t r = tile0 + t;
        double PSI[9], mes = 0;
        if (t < ntile) {
            double P[4][3], det;
            load_tet<false>(a, r, P);
            tet_geometry(P, PSI, det);
            mes = fabs(det) / 6.0;
        }
        for (int b = part; b < NB; b += 4) {
            double Kb[KB];
            double mx = 0;
            if (t < ntile) {
                geometry_block<DA0, DB0, double>(a, r, PSI, mes, 0, b / NCA, b % NCA, Kb);
                for (int i = 0; i < KB; ++i) mx = fmax(mx, fabs(Kb[i]));
            } else {
                for (int i = 0; i < KB; ++i) Kb[i] = 0.0;
            }
            const double sc = mx > 0 ? mx : 1.0;
            ks[b][t] = (float)sc;
            for (int kk = 0; kk < 16; ++kk) {
                const float x = kk < KB ? (float)(Kb[kk < KB ? kk : 0] / sc) : 0.0f;
                const _Float16 h = (_Float16)x;
                Kh[b][t][kk] = h;
                Kl[b][t][kk] = (_Float16)(x - (float)h);
            }
        }
    }
    __syncthreads();
    // D = K^T R^T: rows are tets, columns entries ib, so that for each of
    // its four values a lane group of 16 stores 16 consecutive entries of one
    // tet's column (64-byte runs)
    const int w = threadIdx.x / 64, l = threadIdx.x % 64, j = l % 16, g = l / 16;
    const int tl = 16 * w + j;  // the tet whose K row this lane supplies (A operand row)
    const int nfB0 = a.nfB0, nfA0 = a.nfA0, nfB = nfB0 * NCB, nfA = nfA0 * NCA;
    const _Float16* Rh = static_cast<const _Float16*>(a.r16);
    const _Float16* Rl = Rh + (int64_t)a.e0pad * 16;
    const int ntiles = (nfB + 15) / 16;
    const half4_t zero4 = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
    for (int ia = 0; ia < nfA; ++ia) {
        const int ca = ia / nfA0, ia0 = ia % nfA0;
        for (int h = 0; h < ntiles; ++h) {
            const int ibB = 16 * h + j;  // B operand column (entry) of this lane
            const int cb_lo = (16 * h) / nfB0;
            const int cb_hi = min(NCB - 1, (16 * h + 15) / nfB0);
            float4_t c = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int cb = cb_lo; cb <= cb_hi; ++cb) {
                const int b = cb * NCA + ca;
                half4_t rh = zero4, rl = zero4;
                if (ibB < nfB && ibB / nfB0 == cb) {
                    const int e0 = ibB % nfB0 + nfB0 * ia0;
                    rh = *reinterpret_cast<const half4_t*>(Rh + e0 * 16 + 4 * g);
                    rl = *reinterpret_cast<const half4_t*>(Rl + e0 * 16 + 4 * g);
                }
                const half4_t kh = *reinterpret_cast<const half4_t*>(&Kh[b][tl][4 * g]);
                const half4_t kl = *reinterpret_cast<const half4_t*>(&Kl[b][tl][4 * g]);
                c = __builtin_amdgcn_mfma_f32_16x16x16f16(kh, rl, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x16f16(kl, rh, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x16f16(kh, rh, c, 0, 0, 0);
            }
            if (ibB >= nfB) continue;
            const int kb = (ibB / nfB0) * NCA + ca;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int tv = 16 * w + 4 * g + v;  // D row of this lane's value v
                if (tv < ntile)
                    static_cast<float*>(a.out)[(tile0 + tv) * a.E + ibB + nfB * ia] = c[v] * ks[kb][tv] * a.r16_scale;
            }
        }
    }
}

}  // namespace anifem
