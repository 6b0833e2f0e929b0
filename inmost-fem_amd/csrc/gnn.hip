// Learned preconditioner: a message-passing network over the pattern of an
// assembled SPD matrix A that outputs a lower-triangular factor G on that
// pattern, M^-1 = G G^T (applied by two sparse products, see pcg.hip).
//
// No reference counterpart: the reference solves with INMOST's external
// solvers (examples/tutorials/ex1.cpp:161-163).  The network, loss and
// reverse-mode gradient follow the float64 restatement in tests/gnn_oracle.py
// (same symbols):
//   dinv_r = 1/sqrt(a_rr); at_s = a_s dinv_r dinv_c (off-diagonal slots)
//   x_r    = [log a_rr - mean_nb log a_cc, sum_off at, (no off-diagonal), 1]
//   h0 = relu(W_in x + b_in); m = At h; z = W_s h + W_m m + b; h <- relu(z) + h  (3 layers)
//   L_s = at_s (w_o.(h_r*h_c) + c_o) on lower slots; d_r = w_d.h_r + c_d
//   G_s = exp(d_r) dinv_r (delta_rc + L_s)
//   loss = (1/(P n)) sum_p ||G^T A G z_p - z_p||^2 over P Rademacher probes
// The forward and backward passes read BF16 weights (16-bit words, rounded
// to nearest even from the FP32 master weights after each Adam step) and
// compute in FP32; FP32 master weights and Adam moments.
// Every reduction is a fixed two-stage tree: results are bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "anifem_common.h"

namespace anifem {
namespace gnn {

constexpr int H = 32;        // hidden units
constexpr int FIN = 4;       // input features
constexpr int NL = 3;        // message-passing layers
constexpr int P = 4;         // probe vectors per loss evaluation
constexpr int NT = 256;      // threads per workgroup
constexpr int NODES = NT / H;  // nodes per workgroup in node x feature kernels
constexpr int CHUNK = 128;     // nodes per partial sum of a weight gradient

// flat parameter layout (tests/gnn_oracle.py PARAM_SHAPES)
constexpr int OFF_WIN = 0, OFF_BIN = OFF_WIN + H * FIN;
constexpr int OFF_L0 = OFF_BIN + H;
constexpr int LAYER = 2 * H * H + H;  // W_s, W_m, b
constexpr int OFF_WO = OFF_L0 + NL * LAYER, OFF_CO = OFF_WO + H, OFF_WD = OFF_CO + 1, OFF_CD = OFF_WD + H;
constexpr int NPARAM = OFF_CD + 1;
__host__ __device__ constexpr int off_ws(int l) { return OFF_L0 + l * LAYER; }
__host__ __device__ constexpr int off_wm(int l) { return OFF_L0 + l * LAYER + H * H; }
__host__ __device__ constexpr int off_b(int l) { return OFF_L0 + l * LAYER + 2 * H * H; }

// BF16 by round to nearest even: the 16 high bits of the rounded FP32 value
__device__ __forceinline__ uint16_t bf16_bits(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
// a stored BF16 weight, widened to FP32 exactly
__device__ __forceinline__ float wt(const uint16_t* pw, int i) { return __uint_as_float((uint32_t)pw[i] << 16); }

// ---- features -------------------------------------------------------------
__global__ void diag_kernel(int64_t n, const int64_t* rp, const int32_t* ci, const double* a, float* dinv,
                            float* logd) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    double d = 0;
    for (int64_t s = rp[r]; s < rp[r + 1]; ++s)
        if (ci[s] == r) d = a[s];
    dinv[r] = (float)(1.0 / sqrt(d));
    logd[r] = (float)log(d);
}

__global__ void feat_kernel(int64_t n, const int64_t* rp, const int32_t* ci, const double* a, const float* dinv,
                            const float* logd, float* at, float* a32, float* x) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    float sat = 0, snb = 0;
    int deg = 0;
    for (int64_t s = rp[r]; s < rp[r + 1]; ++s) {
        const int32_t c = ci[s];
        a32[s] = (float)a[s];
        float v = 0;
        if (c != r) {
            v = (float)a[s] * dinv[r] * dinv[c];
            snb += logd[c];
            ++deg;
        }
        at[s] = v;
        sat += v;
    }
    x[r * FIN + 0] = deg ? logd[r] - snb / (float)deg : 0.0f;
    x[r * FIN + 1] = sat;
    x[r * FIN + 2] = deg ? 0.0f : 1.0f;
    x[r * FIN + 3] = 1.0f;
}

// ---- forward --------------------------------------------------------------
__global__ void input_kernel(int64_t n, const float* x, const uint16_t* pw, float* z0, float* h0) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n * H) return;
    const int64_t r = i / H;
    const int k = (int)(i % H);
    float z = wt(pw, OFF_BIN + k);
#pragma unroll
    for (int j = 0; j < FIN; ++j) z += wt(pw, OFF_WIN + k * FIN + j) * x[r * FIN + j];
    z0[i] = z;
    h0[i] = fmaxf(z, 0.0f);
}

// sum_{s in [s0, s1)} at_s in[ci_s][k] added to acc in ascending s, with four
// gathers in flight at a time
__device__ __forceinline__ float gather_rows(float acc, int64_t s0, int64_t s1, const int32_t* ci, const float* at,
                                             const float* in, int k) {
    int64_t s = s0;
    for (; s + 4 <= s1; s += 4) {
        const int32_t c0 = ci[s], c1 = ci[s + 1], c2 = ci[s + 2], c3 = ci[s + 3];
        const float a0 = at[s], a1 = at[s + 1], a2 = at[s + 2], a3 = at[s + 3];
        const float v0 = in[(int64_t)c0 * H + k], v1 = in[(int64_t)c1 * H + k];
        const float v2 = in[(int64_t)c2 * H + k], v3 = in[(int64_t)c3 * H + k];
        acc += a0 * v0;
        acc += a1 * v1;
        acc += a2 * v2;
        acc += a3 * v3;
    }
    for (; s < s1; ++s) acc += at[s] * in[(int64_t)ci[s] * H + k];
    return acc;
}

// out[r][k] = sum_{s in row r} at_s in[ci_s][k] (+ add[r][k])
__global__ void spmm_kernel(int64_t n, const int64_t* rp, const int32_t* ci, const float* at, const float* in,
                            const float* add, float* out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n * H) return;
    const int64_t r = i / H;
    const int k = (int)(i % H);
    out[i] = gather_rows(add ? add[i] : 0.0f, rp[r], rp[r + 1], ci, at, in, k);
}

// One message-passing layer for LN nodes per workgroup:
//   m = At h;  z = W_s h + W_m m + b;  hn = relu(z) + h
// Lane (g, k) = (threadIdx / H, threadIdx % H) owns feature k of nodes
// g, g + 8, ...  The weights sit in LDS transposed (Wt[j][k] = W[k][j]), so a
// wave's 32 lanes of one node read 32 consecutive words; h and m rows are
// broadcasts.
constexpr int LN = 32;
__global__ __launch_bounds__(NT) void mp_layer_kernel(int64_t n, const int64_t* rp, const int32_t* ci,
                                                      const float* at, const float* h, const uint16_t* pw, int l,
                                                      float* m, float* z, float* hn) {
    __shared__ float Wst[H][H + 1], Wmt[H][H + 1], hs[LN][H], ms[LN][H];
    for (int i = threadIdx.x; i < H * H; i += NT) {
        const int k = i / H, j = i % H;
        Wst[j][k] = wt(pw, off_ws(l) + i);
        Wmt[j][k] = wt(pw, off_wm(l) + i);
    }
    constexpr int G = NT / H, PER = LN / G;
    const int k = threadIdx.x % H, g = threadIdx.x / H;
    const int64_t base = (int64_t)blockIdx.x * LN;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
        const int ln = g + G * t;
        const int64_t r = base + ln;
        float hv = 0, mv = 0;
        if (r < n) {
            hv = h[r * H + k];
            mv = gather_rows(0.0f, rp[r], rp[r + 1], ci, at, h, k);
            m[r * H + k] = mv;
        }
        hs[ln][k] = hv;
        ms[ln][k] = mv;
    }
    __syncthreads();
    const float b = wt(pw, off_b(l) + k);
    float acc[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) acc[t] = b;
#pragma unroll 4
    for (int j = 0; j < H; ++j) {
        const float ws = Wst[j][k], wm = Wmt[j][k];
#pragma unroll
        for (int t = 0; t < PER; ++t) acc[t] += ws * hs[g + G * t][j] + wm * ms[g + G * t][j];
    }
#pragma unroll
    for (int t = 0; t < PER; ++t) {
        const int ln = g + G * t;
        const int64_t r = base + ln;
        if (r < n) {
            z[r * H + k] = acc[t];
            hn[r * H + k] = fmaxf(acc[t], 0.0f) + hs[ln][k];
        }
    }
}


__global__ void head_edge_kernel(int64_t nnz, const int32_t* row, const int32_t* ci, const float* at, const float* h,
                                 const uint16_t* pw, float* Lv) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    const int32_t r = row[s], c = ci[s];
    float v = 0;
    if (r > c) {
        float t = wt(pw, OFF_CO);
        for (int k = 0; k < H; ++k) t += wt(pw, OFF_WO + k) * h[(int64_t)r * H + k] * h[(int64_t)c * H + k];
        v = at[s] * t;
    }
    Lv[s] = v;
}

__global__ void head_node_kernel(int64_t n, const int64_t* rp, const int32_t* ci, const float* h, const uint16_t* pw,
                                 const float* dinv, const float* Lv, float* Dh, float* G32, double* G64) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    float d = wt(pw, OFF_CD);
    for (int k = 0; k < H; ++k) d += wt(pw, OFF_WD + k) * h[r * H + k];
    const float dh = expf(d) * dinv[r];
    Dh[r] = dh;
    for (int64_t s = rp[r]; s < rp[r + 1]; ++s) {
        const float g = dh * ((ci[s] == r ? 1.0f : 0.0f) + Lv[s]);
        G32[s] = g;
        if (G64) G64[s] = (double)g;
    }
}

template <typename T>
__global__ void transpose_vals_kernel(int64_t nnz, const int32_t* tslot, const T* v, T* vt) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s < nnz) vt[s] = v[tslot[s]];
}

// M(e) = sum over the entry's terms of G_ik G_jk, in ascending k (the two
// mirror entries (i,j), (j,i) list the same products in the same order, so M
// is bitwise symmetric)
__global__ void form_m_kernel(int64_t m_nnz, const int64_t* tp, const int32_t* ta, const int32_t* tb, const double* G,
                              double* mv) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= m_nnz) return;
    double s = 0;
    for (int64_t t = tp[e]; t < tp[e + 1]; ++t) s = fma(G[ta[t]], G[tb[t]], s);
    mv[e] = s;
}

// ---- loss and its gradient --------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void probes_kernel(int64_t n, uint64_t seed, float* Z) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n * P) return;
    Z[i] = (mix64(seed * 0x100000001B3ull + (uint64_t)i) >> 63) ? 1.0f : -1.0f;
}

// Y[r][p] = sum_{s in row r} v_s X[ci_s][p]
__global__ void spmv_multi_kernel(int64_t n, const int64_t* rp, const int32_t* ci, const float* v, const float* X,
                                  float* Y) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    float acc[P] = {0, 0, 0, 0};
    for (int64_t s = rp[r]; s < rp[r + 1]; ++s) {
        const float vs = v[s];
        const float4 xv = *reinterpret_cast<const float4*>(X + (int64_t)ci[s] * P);
        acc[0] += vs * xv.x;
        acc[1] += vs * xv.y;
        acc[2] += vs * xv.z;
        acc[3] += vs * xv.w;
    }
    *reinterpret_cast<float4*>(Y + r * P) = make_float4(acc[0], acc[1], acc[2], acc[3]);
}

__device__ __forceinline__ float block_sum_f(float v) {
    __shared__ float sh[NT / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    float r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NT / 64; ++i) r += sh[i];
    return r;
}

// res = U - Z; GU = 2 res / (P n); per-workgroup partial sums of res^2
__global__ __launch_bounds__(NT) void loss_kernel(int64_t n, const float* U, const float* Z, float* GU,
                                                  float* partial) {
    float acc = 0;
    const float sc = 2.0f / (float)(P * n);
    for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < n * P; i += (int64_t)gridDim.x * NT) {
        const float res = U[i] - Z[i];
        GU[i] = sc * res;
        acc += res * res;
    }
    const float s = block_sum_f(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// dG_s = sum_p W[r][p] GU[c][p] + GY[r][p] Z[c][p]
__global__ void dG_kernel(int64_t nnz, const int32_t* row, const int32_t* ci, const float* W, const float* GU,
                          const float* GY, const float* Z, float* dG) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    const int64_t r = row[s], c = ci[s];
    float acc = 0;
#pragma unroll
    for (int p = 0; p < P; ++p) acc += W[r * P + p] * GU[c * P + p] + GY[r * P + p] * Z[c * P + p];
    dG[s] = acc;
}

// Head backward, one lane per (node r, feature k):
//   dDh_r = sum_{s in row r} dG_s (delta + L_s); dd_r = dDh_r Dh_r
//   dt_s  = dG_s Dh_r at_s on lower slots s = (r, c)
//   dh[r] = dd_r w_d + sum_{off-diagonal s in row r} dt_(lower of s) w_o * h[c_s]
// and the per-node terms of the head's weight gradients.
__global__ void head_back_kernel(int64_t n, const int64_t* rp, const int32_t* ci, const int32_t* tslot,
                                 const float* dG, const float* Lv, const float* Dh, const float* at, const float* h,
                                 const uint16_t* pw, float* dh, float* gwo, float* gwd, float* gsc) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n * H) return;
    const int64_t r = i / H;
    const int k = (int)(i % H);
    float dDh = 0;
    for (int64_t s = rp[r]; s < rp[r + 1]; ++s) dDh += dG[s] * ((ci[s] == r ? 1.0f : 0.0f) + Lv[s]);
    const float dd = dDh * Dh[r];
    const float hrk = h[r * H + k];
    float acc = dd * wt(pw, OFF_WD + k);
    float wo_part = 0, co_part = 0;
    for (int64_t s = rp[r]; s < rp[r + 1]; ++s) {
        const int32_t c = ci[s];
        if (c == r) continue;
        float dt;
        if (c < r) {
            dt = dG[s] * Dh[r] * at[s];
            const float hck = h[(int64_t)c * H + k];
            wo_part += dt * hrk * hck;
            co_part += dt;
        } else {
            const int32_t q = tslot[s];  // lower slot (c, r) in row c
            dt = dG[q] * Dh[c] * at[q];
        }
        acc += dt * wt(pw, OFF_WO + k) * h[(int64_t)c * H + k];
    }
    dh[i] = acc;
    gwo[i] = wo_part;
    gwd[i] = dd * hrk;
    if (k == 0) {
        gsc[2 * r] = co_part;
        gsc[2 * r + 1] = dd;
    }
}

// Layer backward: dz = dh * [z > 0]; dm = W_m^T dz; tmp = dh + W_s^T dz
__global__ __launch_bounds__(NT) void layer_back_kernel(int64_t n, const float* dh, const float* z, const uint16_t* pw,
                                                        int l, float* dz, float* dm, float* tmp) {
    __shared__ float Ws[H * H], Wm[H * H], dzs[NODES][H];
    for (int i = threadIdx.x; i < H * H; i += NT) {
        Ws[i] = wt(pw, off_ws(l) + i);
        Wm[i] = wt(pw, off_wm(l) + i);
    }
    const int ln = threadIdx.x / H, k = threadIdx.x % H;
    const int64_t r = (int64_t)blockIdx.x * NODES + ln;
    float g = 0, dhk = 0;
    if (r < n) {
        dhk = dh[r * H + k];
        g = z[r * H + k] > 0.0f ? dhk : 0.0f;
        dz[r * H + k] = g;
    }
    dzs[ln][k] = g;
    __syncthreads();
    if (r >= n) return;
    float a = 0, b = dhk;
#pragma unroll 8
    for (int j = 0; j < H; ++j) {
        a += Wm[j * H + k] * dzs[ln][j];
        b += Ws[j * H + k] * dzs[ln][j];
    }
    dm[r * H + k] = a;
    tmp[r * H + k] = b;
}

__global__ void relu_mask_kernel(int64_t n, const float* dh, const float* z, float* dz) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n * H) dz[i] = z[i] > 0.0f ? dh[i] : 0.0f;
}

// Partial sums over CHUNK nodes of A[r][k] * B[r][j] (k < H, j < KB) and,
// with sums, of A[r][k]: partial[blk][k*KB + j], partial[blk][H*KB + k].
__global__ __launch_bounds__(NT) void outer_partial_kernel(int64_t n, const float* A, const float* B, int KB,
                                                           int sums, float* partial) {
    // the chunk's rows of A (CHUNK x H) and B (CHUNK x KB, KB <= H) in LDS;
    // lane e = k * KB + j then reads A[r][k] as a broadcast and B[r][j] from
    // consecutive words
    __shared__ float As[CHUNK][H], Bs[CHUNK][H];
    const int64_t r0 = (int64_t)blockIdx.x * CHUNK;
    const int rows = (int)std::min<int64_t>(CHUNK, n - r0);
    for (int i = threadIdx.x; i < CHUNK * H; i += NT) {
        const int r = i / H, c = i % H;
        As[r][c] = r < rows ? A[(r0 + r) * H + c] : 0.0f;
        if (c < KB) Bs[r][c] = r < rows ? B[(r0 + r) * KB + c] : 0.0f;
    }
    __syncthreads();
    const int len = H * KB + (sums ? H : 0);
    float* out = partial + (int64_t)blockIdx.x * len;
    for (int e = threadIdx.x; e < len; e += NT) {
        float acc = 0;
        if (e < H * KB) {
            const int k = e / KB, j = e % KB;
            for (int r = 0; r < rows; ++r) acc += As[r][k] * Bs[r][j];
        } else {
            const int k = e - H * KB;
            for (int r = 0; r < rows; ++r) acc += As[r][k];
        }
        out[e] = acc;
    }
}

// Column sums over CHUNK rows of X (n x C): partial[blk][c]
__global__ __launch_bounds__(NT) void colsum_partial_kernel(int64_t n, const float* X, int C, float* partial) {
    const int64_t r0 = (int64_t)blockIdx.x * CHUNK;
    const int64_t r1 = std::min<int64_t>(n, r0 + CHUNK);
    for (int c = threadIdx.x; c < C; c += NT) {
        float acc = 0;
        for (int64_t r = r0; r < r1; ++r) acc += X[r * C + c];
        partial[(int64_t)blockIdx.x * C + c] = acc;
    }
}

// grad[base + e] = sum_b partial[b * stride + e] for e < count: 32 entries per
// workgroup, eight lanes per entry each adding the blocks b = sub (mod 8) in
// order, then the eight subsums in order (a fixed summation order)
constexpr int kRedSub = NT / 32;
__global__ __launch_bounds__(NT) void reduce_partials_kernel(int nblk, int stride, int count, const float* partial,
                                                             float* grad, int base) {
    __shared__ float sh[kRedSub][32];
    const int el = threadIdx.x % 32, sub = threadIdx.x / 32;
    const int e = blockIdx.x * 32 + el;
    float acc = 0;
    if (e < count)
        for (int b = sub; b < nblk; b += kRedSub) acc += partial[(int64_t)b * stride + e];
    sh[sub][el] = acc;
    __syncthreads();
    if (sub == 0 && e < count) {
        float t = 0;
        for (int k = 0; k < kRedSub; ++k) t += sh[k][el];
        grad[base + e] = t;
    }
}

__global__ void loss_total_kernel(int nblk, const float* partial, float scale, float* loss) {
    if (threadIdx.x != 0) return;
    float acc = 0;
    for (int b = 0; b < nblk; ++b) acc += partial[b];
    *loss = acc * scale;
}

__global__ void adam_kernel(int np, float* theta, const float* g, float* m, float* v, float lr, float b1, float b2,
                            float eps, float bc1, float bc2, uint16_t* pw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1 - b1) * gi;
    const float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float t = theta[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    theta[i] = t;
    pw[i] = bf16_bits(t);
}

__global__ void round_params_kernel(int np, const float* theta, uint16_t* pw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) pw[i] = bf16_bits(theta[i]);
}

// Gaussian random field by random Fourier features (squared-exponential
// covariance sigma^2 exp(-|x-y|^2 / (2 ell^2))): g(x) = sigma sqrt(2/M)
// sum_m cos(w_m . x + phi_m), w_m ~ N(0, I/ell^2), phi_m ~ U[0, 2 pi).
__global__ void grf_kernel(int64_t n, const double* xyz, int M, const double* wphi, double amp, int lognormal,
                           double* out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    double acc = 0;
    for (int m = 0; m < M; ++m) {
        const double* f = wphi + 4 * m;
        acc += cos(f[0] * x + f[1] * y + f[2] * z + f[3]);
    }
    const double g = amp * acc;
    out[i] = lognormal ? exp(g) : g;
}

}  // namespace gnn
}  // namespace anifem

using namespace anifem::gnn;

struct anifem_gnn {
    int64_t n = 0, nnz = 0;
    int64_t *rp = nullptr;
    int32_t *ci = nullptr, *row = nullptr, *tslot = nullptr;
    // per slot
    float *at = nullptr, *a32 = nullptr, *Lv = nullptr, *G32 = nullptr, *GT32 = nullptr, *dG = nullptr;
    double *G64 = nullptr, *GT64 = nullptr;
    // per node
    float *dinv = nullptr, *logd = nullptr, *x = nullptr, *Dh = nullptr, *gsc = nullptr;
    float* z0 = nullptr;
    float* h[NL + 1] = {};
    float* m[NL] = {};
    float* z[NL] = {};
    float *dh = nullptr, *dz = nullptr, *dm = nullptr, *tmp = nullptr, *gwo = nullptr, *gwd = nullptr;
    // probes
    float *Z = nullptr, *Y = nullptr, *W = nullptr, *U = nullptr, *GU = nullptr, *GW = nullptr, *GY = nullptr;
    // parameters
    float *theta = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
    uint16_t* pw = nullptr;  // BF16 weights of the forward and backward passes
    float *partial = nullptr, *loss = nullptr;
    // M = G G^T as an explicit matrix: its CSR pattern, and per entry the slot
    // pairs (G_ik, G_jk) of its terms in ascending k
    int64_t m_nnz = 0;
    int64_t *m_rp = nullptr, *m_tp = nullptr;
    int32_t *m_ci = nullptr, *m_ta = nullptr, *m_tb = nullptr;
    double* m_val = nullptr;
    int64_t nchunk = 0, adam_t = 0;
    std::vector<void*> bufs;
};

namespace {

template <typename T>
int dalloc(anifem_gnn* g, T** p, size_t count) {
    ANIFEM_HIP_CHECK(hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
    g->bufs.push_back(*p);
    return ANIFEM_OK;
}

unsigned nb(int64_t work, int threads = NT) { return (unsigned)std::max<int64_t>(1, (work + threads - 1) / threads); }

// entries [0, count) of the chunk partials (row stride `stride`) summed into grad[base ..]
void reduce_partials(hipStream_t s, int64_t nblk, int stride, int count, const float* partial, float* grad,
                     int base) {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(nb(count, 32)), dim3(NT), 0, s, (int)nblk, stride, count, partial,
                       grad, base);
}

// forward through the network: G32 / G64 (if non-null) and, for the loss,
// GT32
int forward(anifem_gnn* g, const double* a, double* G64, bool gt32, hipStream_t s) {
    const int64_t n = g->n, nnz = g->nnz;
    hipLaunchKernelGGL(diag_kernel, dim3(nb(n)), dim3(NT), 0, s, n, g->rp, g->ci, a, g->dinv, g->logd);
    hipLaunchKernelGGL(feat_kernel, dim3(nb(n)), dim3(NT), 0, s, n, g->rp, g->ci, a, g->dinv, g->logd, g->at, g->a32,
                       g->x);
    hipLaunchKernelGGL(input_kernel, dim3(nb(n * H)), dim3(NT), 0, s, n, g->x, g->pw, g->z0, g->h[0]);
    for (int l = 0; l < NL; ++l)
        hipLaunchKernelGGL(mp_layer_kernel, dim3(nb(n, LN)), dim3(NT), 0, s, n, g->rp, g->ci, g->at, g->h[l], g->pw,
                           l, g->m[l], g->z[l], g->h[l + 1]);
    hipLaunchKernelGGL(head_edge_kernel, dim3(nb(nnz)), dim3(NT), 0, s, nnz, g->row, g->ci, g->at, g->h[NL], g->pw,
                       g->Lv);
    hipLaunchKernelGGL(head_node_kernel, dim3(nb(n)), dim3(NT), 0, s, n, g->rp, g->ci, g->h[NL], g->pw, g->dinv, g->Lv,
                       g->Dh, g->G32, G64);
    if (gt32)
        hipLaunchKernelGGL(transpose_vals_kernel<float>, dim3(nb(nnz)), dim3(NT), 0, s, nnz, g->tslot, g->G32,
                           g->GT32);
    ANIFEM_HIP_CHECK(hipGetLastError());
    return ANIFEM_OK;
}

}  // namespace

extern "C" {

int anifem_gnn_n_params(void) { return NPARAM; }

int anifem_grf_sample(int64_t n, const double* xyz, double ell, double sigma, int n_features, uint64_t seed,
                      int lognormal, double* out, void* stream) {
    if (n <= 0 || !xyz || !out || !(ell > 0) || !(sigma >= 0) || n_features <= 0) return ANIFEM_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // features drawn on the host from a counter-based generator (splitmix64)
    std::vector<double> wphi(4 * (size_t)n_features);
    uint64_t ctr = seed * 0x2545F4914F6CDD1Dull;
    auto uni = [&]() {
        uint64_t z = (ctr += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);  // (0, 1)
    };
    const double two_pi = 6.283185307179586;
    for (int m = 0; m < n_features; ++m) {
        for (int c = 0; c < 3; ++c) {
            // Box-Muller
            const double u1 = uni(), u2 = uni();
            wphi[4 * m + c] = std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2) / ell;
        }
        wphi[4 * m + 3] = two_pi * uni();
    }
    double* d = nullptr;
    ANIFEM_HIP_CHECK(hipMalloc((void**)&d, wphi.size() * sizeof(double)));
    int rc = ANIFEM_OK;
    if (hipMemcpyAsync(d, wphi.data(), wphi.size() * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess)
        rc = ANIFEM_ERR_HIP;
    if (!rc) {
        hipLaunchKernelGGL(grf_kernel, dim3(nb(n)), dim3(NT), 0, s, n, xyz, n_features, (const double*)d,
                           sigma * std::sqrt(2.0 / n_features), lognormal, out);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) rc = ANIFEM_ERR_HIP;
    }
    (void)hipFree(d);
    return rc;
}

int anifem_gnn_create(anifem_gnn** out, int64_t n, const int64_t* row_ptr, const int32_t* col_idx, void* stream) {
    if (!out || n <= 0 || !row_ptr || !col_idx) return ANIFEM_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::unique_ptr<anifem_gnn> g(new anifem_gnn());
    g->n = n;
    g->nnz = row_ptr[n];
    const int64_t nnz = g->nnz;
    if (nnz <= 0 || nnz >= INT32_MAX) return ANIFEM_ERR_ARG;
    // slot rows and transposed slots (the pattern must be symmetric)
    std::vector<int32_t> row(nnz), tslot(nnz);
    for (int64_t r = 0; r < n; ++r)
        for (int64_t q = row_ptr[r]; q < row_ptr[r + 1]; ++q) row[q] = (int32_t)r;
    for (int64_t q = 0; q < nnz; ++q) {
        const int32_t c = col_idx[q];
        if (c < 0 || c >= n) return ANIFEM_ERR_ARG;
        const int32_t* b = col_idx + row_ptr[c];
        const int32_t* e = col_idx + row_ptr[c + 1];
        const int32_t* it = std::lower_bound(b, e, row[q]);
        if (it == e || *it != row[q]) return ANIFEM_ERR_ARG;
        tslot[q] = (int32_t)(it - col_idx);
    }
    // Symbolic M = G G^T.  G has the lower triangle of the pattern: its column
    // k holds slots (i, k), i >= k (the transposes of row k's slots (k, i)).
    // Row i of M gathers column k's entries for every (i, k) in G.
    std::vector<int64_t> cptr(n + 1, 0);
    for (int64_t k = 0; k < n; ++k) {
        int64_t c = 0;
        for (int64_t q = row_ptr[k]; q < row_ptr[k + 1]; ++q) c += col_idx[q] >= k;
        cptr[k + 1] = cptr[k] + c;
    }
    std::vector<int32_t> cslot(cptr[n]), crow(cptr[n]);
    for (int64_t k = 0; k < n; ++k) {
        int64_t o = cptr[k];
        for (int64_t q = row_ptr[k]; q < row_ptr[k + 1]; ++q)
            if (col_idx[q] >= k) {
                cslot[o] = tslot[q];
                crow[o++] = col_idx[q];
            }
    }
    int64_t n_terms = 0;
    for (int64_t k = 0; k < n; ++k) n_terms += (cptr[k + 1] - cptr[k]) * (cptr[k + 1] - cptr[k]);
    constexpr int64_t kMaxTerms = int64_t(1) << 30;  // beyond this M is not formed
    std::vector<int64_t> m_rp, m_tp;
    std::vector<int32_t> m_ci, m_ta, m_tb;
    if (n_terms <= kMaxTerms) {
        m_rp.assign(n + 1, 0);
        m_ta.reserve(n_terms);
        m_tb.reserve(n_terms);
        m_tp.push_back(0);
        struct Term {
            int32_t j, ta, tb;
        };
        std::vector<Term> terms;
        for (int64_t i = 0; i < n; ++i) {
            terms.clear();
            for (int64_t q = row_ptr[i]; q < row_ptr[i + 1]; ++q) {
                const int32_t k = col_idx[q];
                if (k > i) break;  // columns ascend: (i, k) in G for k <= i, in ascending k
                for (int64_t o = cptr[k]; o < cptr[k + 1]; ++o) terms.push_back({crow[o], (int32_t)q, cslot[o]});
            }
            std::stable_sort(terms.begin(), terms.end(), [](const Term& a, const Term& b) { return a.j < b.j; });
            for (size_t u = 0; u < terms.size(); ++u) {
                if (u == 0 || terms[u].j != terms[u - 1].j) {
                    if (u) m_tp.push_back((int64_t)m_ta.size());
                    m_ci.push_back(terms[u].j);
                }
                m_ta.push_back(terms[u].ta);
                m_tb.push_back(terms[u].tb);
            }
            if (!terms.empty()) m_tp.push_back((int64_t)m_ta.size());
            m_rp[i + 1] = (int64_t)m_ci.size();
        }
        g->m_nnz = (int64_t)m_ci.size();
        if (g->m_nnz >= INT32_MAX) {
            g->m_nnz = 0;
            m_ci.clear();
        }
    }
    anifem_gnn* p = g.get();
    int rc = 0;
#define ALLOC(ptr, cnt) \
    if ((rc = dalloc(p, &p->ptr, (size_t)(cnt)))) return rc
    ALLOC(rp, n + 1);
    ALLOC(ci, nnz);
    ALLOC(row, nnz);
    ALLOC(tslot, nnz);
    ALLOC(at, nnz);
    ALLOC(a32, nnz);
    ALLOC(Lv, nnz);
    ALLOC(G32, nnz);
    ALLOC(GT32, nnz);
    ALLOC(dG, nnz);
    ALLOC(G64, nnz);
    ALLOC(GT64, nnz);
    ALLOC(dinv, n);
    ALLOC(logd, n);
    ALLOC(x, n * FIN);
    ALLOC(Dh, n);
    ALLOC(gsc, 2 * n);
    ALLOC(z0, n * H);
    for (int l = 0; l <= NL; ++l) ALLOC(h[l], n * H);
    for (int l = 0; l < NL; ++l) {
        ALLOC(m[l], n * H);
        ALLOC(z[l], n * H);
    }
    ALLOC(dh, n * H);
    ALLOC(dz, n * H);
    ALLOC(dm, n * H);
    ALLOC(tmp, n * H);
    ALLOC(gwo, n * H);
    ALLOC(gwd, n * H);
    ALLOC(Z, n * P);
    ALLOC(Y, n * P);
    ALLOC(W, n * P);
    ALLOC(U, n * P);
    ALLOC(GU, n * P);
    ALLOC(GW, n * P);
    ALLOC(GY, n * P);
    ALLOC(theta, NPARAM);
    ALLOC(pw, NPARAM);
    ALLOC(grad, NPARAM);
    ALLOC(am, NPARAM);
    ALLOC(av, NPARAM);
    g->nchunk = (n + CHUNK - 1) / CHUNK;
    ALLOC(partial, g->nchunk * (H * H + H) + 1024);
    ALLOC(loss, 1);
    if (g->m_nnz > 0) {
        ALLOC(m_rp, n + 1);
        ALLOC(m_tp, g->m_nnz + 1);
        ALLOC(m_ci, g->m_nnz);
        ALLOC(m_ta, m_ta.size());
        ALLOC(m_tb, m_tb.size());
        ALLOC(m_val, g->m_nnz);
    }
#undef ALLOC
    ANIFEM_HIP_CHECK(hipMemcpyAsync(g->rp, row_ptr, (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    ANIFEM_HIP_CHECK(hipMemcpyAsync(g->ci, col_idx, nnz * sizeof(int32_t), hipMemcpyHostToDevice, s));
    ANIFEM_HIP_CHECK(hipMemcpyAsync(g->row, row.data(), nnz * sizeof(int32_t), hipMemcpyHostToDevice, s));
    ANIFEM_HIP_CHECK(hipMemcpyAsync(g->tslot, tslot.data(), nnz * sizeof(int32_t), hipMemcpyHostToDevice, s));
    if (g->m_nnz > 0) {
        ANIFEM_HIP_CHECK(hipMemcpyAsync(g->m_rp, m_rp.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
        ANIFEM_HIP_CHECK(hipMemcpyAsync(g->m_tp, m_tp.data(), m_tp.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
        ANIFEM_HIP_CHECK(hipMemcpyAsync(g->m_ci, m_ci.data(), m_ci.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
        ANIFEM_HIP_CHECK(hipMemcpyAsync(g->m_ta, m_ta.data(), m_ta.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
        ANIFEM_HIP_CHECK(hipMemcpyAsync(g->m_tb, m_tb.data(), m_tb.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    }
    ANIFEM_HIP_CHECK(hipMemsetAsync(g->theta, 0, NPARAM * sizeof(float), s));
    ANIFEM_HIP_CHECK(hipMemsetAsync(g->pw, 0, NPARAM * sizeof(uint16_t), s));
    ANIFEM_HIP_CHECK(hipMemsetAsync(g->grad, 0, NPARAM * sizeof(float), s));
    ANIFEM_HIP_CHECK(hipMemsetAsync(g->am, 0, NPARAM * sizeof(float), s));
    ANIFEM_HIP_CHECK(hipMemsetAsync(g->av, 0, NPARAM * sizeof(float), s));
    ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
    *out = g.release();
    return ANIFEM_OK;
}

int anifem_gnn_destroy(anifem_gnn* g) {
    if (!g) return ANIFEM_OK;
    for (void* b : g->bufs) (void)hipFree(b);
    delete g;
    return ANIFEM_OK;
}

int anifem_gnn_set_params(anifem_gnn* g, const float* theta) {
    if (!g || !theta) return ANIFEM_ERR_ARG;
    ANIFEM_HIP_CHECK(hipMemcpy(g->theta, theta, NPARAM * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(round_params_kernel, dim3(nb(NPARAM)), dim3(NT), 0, 0, NPARAM, g->theta, g->pw);
    ANIFEM_HIP_CHECK(hipMemset(g->am, 0, NPARAM * sizeof(float)));
    ANIFEM_HIP_CHECK(hipMemset(g->av, 0, NPARAM * sizeof(float)));
    g->adam_t = 0;
    ANIFEM_HIP_CHECK(hipDeviceSynchronize());
    return ANIFEM_OK;
}

int anifem_gnn_get_params(const anifem_gnn* g, float* theta, float* theta_bf16) {
    if (!g) return ANIFEM_ERR_ARG;
    if (theta) ANIFEM_HIP_CHECK(hipMemcpy(theta, g->theta, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    if (theta_bf16) {  // the BF16 weights, widened to FP32
        std::vector<uint16_t> b(NPARAM);
        ANIFEM_HIP_CHECK(hipMemcpy(b.data(), g->pw, NPARAM * sizeof(uint16_t), hipMemcpyDeviceToHost));
        for (int i = 0; i < NPARAM; ++i) {
            const uint32_t u = (uint32_t)b[i] << 16;
            std::memcpy(&theta_bf16[i], &u, sizeof(float));
        }
    }
    return ANIFEM_OK;
}

int anifem_gnn_get_grad(const anifem_gnn* g, float* grad) {
    if (!g || !grad) return ANIFEM_ERR_ARG;
    ANIFEM_HIP_CHECK(hipMemcpy(grad, g->grad, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return ANIFEM_OK;
}

int anifem_gnn_factor(anifem_gnn* g, const double* a, double* G, double* GT, void* stream) {
    if (!g || !a) return ANIFEM_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = forward(g, a, G ? G : g->G64, false, s);
    if (rc) return rc;
    hipLaunchKernelGGL(transpose_vals_kernel<double>, dim3(nb(g->nnz)), dim3(NT), 0, s, g->nnz, g->tslot,
                       (const double*)(G ? G : g->G64), GT ? GT : g->GT64);
    ANIFEM_HIP_CHECK(hipGetLastError());
    return ANIFEM_OK;
}

int anifem_gnn_views(const anifem_gnn* g, int64_t* n, const int64_t** row_ptr, const int32_t** col_idx,
                     const double** G, const double** GT) {
    if (!g) return ANIFEM_ERR_ARG;
    if (n) *n = g->n;
    if (row_ptr) *row_ptr = g->rp;
    if (col_idx) *col_idx = g->ci;
    if (G) *G = g->G64;
    if (GT) *GT = g->GT64;
    return ANIFEM_OK;
}

int anifem_gnn_precond(anifem_gnn* g, int64_t* m_nnz, const int64_t** m_row_ptr, const int32_t** m_col_idx,
                       const double** m_val, void* stream) {
    if (!g) return ANIFEM_ERR_ARG;
    if (g->m_nnz <= 0) return ANIFEM_ERR_NOMEM;  // M was not formed (more than 2^30 terms)
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(form_m_kernel, dim3(nb(g->m_nnz)), dim3(NT), 0, s, g->m_nnz, g->m_tp, g->m_ta, g->m_tb,
                       (const double*)g->G64, g->m_val);
    ANIFEM_HIP_CHECK(hipGetLastError());
    if (m_nnz) *m_nnz = g->m_nnz;
    if (m_row_ptr) *m_row_ptr = g->m_rp;
    if (m_col_idx) *m_col_idx = g->m_ci;
    if (m_val) *m_val = g->m_val;
    return ANIFEM_OK;
}

int anifem_gnn_loss_grad(anifem_gnn* g, const double* a, const float* probes, uint64_t seed, float* loss,
                         void* stream) {
    if (!g || !a) return ANIFEM_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t n = g->n, nnz = g->nnz;
    int rc = forward(g, a, nullptr, true, s);
    if (rc) return rc;
    if (probes)
        ANIFEM_HIP_CHECK(hipMemcpyAsync(g->Z, probes, n * P * sizeof(float), hipMemcpyDeviceToDevice, s));
    else
        hipLaunchKernelGGL(probes_kernel, dim3(nb(n * P)), dim3(NT), 0, s, n, seed, g->Z);
    // loss: Y = G Z, W = A Y, U = G^T W
    hipLaunchKernelGGL(spmv_multi_kernel, dim3(nb(n)), dim3(NT), 0, s, n, g->rp, g->ci, g->G32, g->Z, g->Y);
    hipLaunchKernelGGL(spmv_multi_kernel, dim3(nb(n)), dim3(NT), 0, s, n, g->rp, g->ci, g->a32, g->Y, g->W);
    hipLaunchKernelGGL(spmv_multi_kernel, dim3(nb(n)), dim3(NT), 0, s, n, g->rp, g->ci, g->GT32, g->W, g->U);
    const unsigned lb = (unsigned)std::min<int64_t>(1024, nb(n * P));
    hipLaunchKernelGGL(loss_kernel, dim3(lb), dim3(NT), 0, s, n, g->U, g->Z, g->GU, g->partial);
    hipLaunchKernelGGL(loss_total_kernel, dim3(1), dim3(64), 0, s, (int)lb, g->partial, 1.0f / (float)(P * n),
                       g->loss);
    // backward through the loss: GW = G GU, GY = A GW, dG
    hipLaunchKernelGGL(spmv_multi_kernel, dim3(nb(n)), dim3(NT), 0, s, n, g->rp, g->ci, g->G32, g->GU, g->GW);
    hipLaunchKernelGGL(spmv_multi_kernel, dim3(nb(n)), dim3(NT), 0, s, n, g->rp, g->ci, g->a32, g->GW, g->GY);
    hipLaunchKernelGGL(dG_kernel, dim3(nb(nnz)), dim3(NT), 0, s, nnz, g->row, g->ci, g->W, g->GU, g->GY, g->Z,
                       g->dG);
    // heads
    hipLaunchKernelGGL(head_back_kernel, dim3(nb(n * H)), dim3(NT), 0, s, n, g->rp, g->ci, g->tslot, g->dG, g->Lv,
                       g->Dh, g->at, g->h[NL], g->pw, g->dh, g->gwo, g->gwd, g->gsc);
    hipLaunchKernelGGL(colsum_partial_kernel, dim3((unsigned)g->nchunk), dim3(NT), 0, s, n, g->gwo, H, g->partial);
    reduce_partials(s, g->nchunk, H, H, g->partial, g->grad, OFF_WO);
    hipLaunchKernelGGL(colsum_partial_kernel, dim3((unsigned)g->nchunk), dim3(NT), 0, s, n, g->gwd, H, g->partial);
    reduce_partials(s, g->nchunk, H, H, g->partial, g->grad, OFF_WD);
    hipLaunchKernelGGL(colsum_partial_kernel, dim3((unsigned)g->nchunk), dim3(NT), 0, s, n, g->gsc, 2, g->partial);
    reduce_partials(s, g->nchunk, 2, 1, g->partial, g->grad, OFF_CO);
    reduce_partials(s, g->nchunk, 2, 1, g->partial + 1, g->grad, OFF_CD);
    // message-passing layers, last to first
    for (int l = NL - 1; l >= 0; --l) {
        hipLaunchKernelGGL(layer_back_kernel, dim3(nb(n, NODES)), dim3(NT), 0, s, n, g->dh, g->z[l], g->pw, l, g->dz,
                           g->dm, g->tmp);
        // dW_s = dz^T h, dW_m = dz^T m, db = sum dz
        hipLaunchKernelGGL(outer_partial_kernel, dim3((unsigned)g->nchunk), dim3(NT), 0, s, n, g->dz, g->h[l], H, 1,
                           g->partial);
        reduce_partials(s, g->nchunk, H * H + H, H * H, g->partial, g->grad, off_ws(l));
        reduce_partials(s, g->nchunk, H * H + H, H, g->partial + H * H, g->grad, off_b(l));
        hipLaunchKernelGGL(outer_partial_kernel, dim3((unsigned)g->nchunk), dim3(NT), 0, s, n, g->dz, g->m[l], H, 0,
                           g->partial);
        reduce_partials(s, g->nchunk, H * H, H * H, g->partial, g->grad, off_wm(l));
        // dh_prev = tmp + At dm
        hipLaunchKernelGGL(spmm_kernel, dim3(nb(n * H)), dim3(NT), 0, s, n, g->rp, g->ci, g->at, g->dm, g->tmp,
                           g->dh);
    }
    // input layer
    hipLaunchKernelGGL(relu_mask_kernel, dim3(nb(n * H)), dim3(NT), 0, s, n, g->dh, g->z0, g->dz);
    hipLaunchKernelGGL(outer_partial_kernel, dim3((unsigned)g->nchunk), dim3(NT), 0, s, n, g->dz, g->x, FIN, 1,
                       g->partial);
    reduce_partials(s, g->nchunk, H * FIN + H, H * FIN + H, g->partial, g->grad, OFF_WIN);
    ANIFEM_HIP_CHECK(hipGetLastError());
    if (loss) {
        ANIFEM_HIP_CHECK(hipMemcpyAsync(loss, g->loss, sizeof(float), hipMemcpyDeviceToHost, s));
        ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
    }
    return ANIFEM_OK;
}

int anifem_gnn_adam_step(anifem_gnn* g, float lr, float beta1, float beta2, float eps, void* stream) {
    if (!g || !(lr > 0)) return ANIFEM_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    g->adam_t += 1;
    const float bc1 = 1.0f - (float)std::pow((double)beta1, (double)g->adam_t);
    const float bc2 = 1.0f - (float)std::pow((double)beta2, (double)g->adam_t);
    hipLaunchKernelGGL(adam_kernel, dim3(nb(NPARAM)), dim3(NT), 0, s, NPARAM, g->theta, g->grad, g->am, g->av, lr,
                       beta1, beta2, eps, bc1, bc2, g->pw);
    ANIFEM_HIP_CHECK(hipGetLastError());
    return ANIFEM_OK;
}

}  // extern "C"
