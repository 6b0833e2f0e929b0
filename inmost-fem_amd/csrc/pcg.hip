// FP64 Jacobi-preconditioned conjugate gradients on a device CSR matrix.
//
// The reference solves with INMOST's external Solver (examples/tutorials/
// ex1.cpp:161-163); this solver is the GPU counterpart used on the assembled
// systems.  All scalars stay on the device: every kernel reads a `done` flag
// and becomes a no-op once the residual test has passed, so the host only
// polls every few iterations yet the iteration count and x are exactly those
// of the iteration that converged.  Reductions use a fixed two-stage tree, so
// results are bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "anifem_common.h"

namespace anifem {

struct PcgState {
    double rz, rz_new, pAp, rr, bb, tol2, rzo;
    int64_t it;
    int done;
};

constexpr int kRedBlocks = 1024;  // partial sums per reduction
constexpr int kThreads = 256;

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[kThreads / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kThreads / 64; ++i) r += sh[i];
    __syncthreads();
    return r;  // valid in thread 0
}

// y = A x with L lanes per row; optionally partial dot(x, y) per block.
template <int L>
__global__ __launch_bounds__(kThreads) void spmv_kernel(int64_t n, const int64_t* __restrict__ rp,
                                                         const int32_t* __restrict__ ci,
                                                         const double* __restrict__ v, const double* __restrict__ x,
                                                         double* __restrict__ y, double* partial,
                                                         const PcgState* st) {
    if (st && st->done) return;
    const int64_t rows_per_block = kThreads / L;
    double dot = 0;
    for (int64_t base = (int64_t)blockIdx.x * rows_per_block; base < n; base += (int64_t)gridDim.x * rows_per_block) {
        const int64_t row = base + threadIdx.x / L;
        const int lane = threadIdx.x % L;
        double s = 0;
        if (row < n) {
            for (int64_t k = rp[row] + lane; k < rp[row + 1]; k += L) s += v[k] * x[ci[k]];
        }
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_down(s, off, L);
        if (row < n && lane == 0) {
            y[row] = s;
            if (partial) dot += s * x[row];
        }
    }
    if (partial) {
        double b = block_sum(dot);
        if (threadIdx.x == 0) partial[blockIdx.x] = b;
    }
}

__global__ void jacobi_inverse_kernel(int64_t n, const int64_t* rp, const int32_t* ci, const double* v, double* dinv) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double d = 0;
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) d = v[k];
        dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
    }
}

// r = b - Ax, z = Dinv r, p = z; partials: r.z, r.r, b.b
__global__ __launch_bounds__(kThreads) void init_kernel(int64_t n, const double* b, const double* Ax,
                                                         const double* dinv, double* r, double* z, double* p,
                                                         double* partial) {
    double rz = 0, rr = 0, bb = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double ri = b[i] - Ax[i];
        double zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        rz += ri * zi;
        rr += ri * ri;
        bb += b[i] * b[i];
    }
    double s0 = block_sum(rz), s1 = block_sum(rr), s2 = block_sum(bb);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s0;
        partial[kRedBlocks + blockIdx.x] = s1;
        partial[2 * kRedBlocks + blockIdx.x] = s2;
    }
}

// Second reduction stage; `which`: 0 init (rz, rr, bb), 1 pAp, 2 (rz_new, rr),
// 3 init of the factored method (rr, bb), 4 rr after an update, 5 (rz_new, rzo).
__global__ __launch_bounds__(kThreads) void reduce_kernel(const double* partial, int nparts, int which,
                                                           PcgState* st, double rtol) {
    if (which != 0 && which != 3 && st->done) return;
    double a = 0, b = 0, c = 0;
    for (int i = threadIdx.x; i < nparts; i += kThreads) {
        a += partial[i];
        if (which != 1 && which != 4) b += partial[kRedBlocks + i];
        if (which == 0) c += partial[2 * kRedBlocks + i];
    }
    a = block_sum(a);
    b = block_sum(b);
    c = block_sum(c);
    if (threadIdx.x == 0) {
        if (which == 0) {
            st->rz = a;
            st->rr = b;
            st->bb = c > 0 ? c : 1.0;
            st->tol2 = rtol * rtol * st->bb;
            st->it = 0;
            st->done = (b <= st->tol2) ? 1 : 0;
        } else if (which == 1) {
            st->pAp = a;
        } else if (which == 3) {
            st->rr = a;
            st->bb = b > 0 ? b : 1.0;
            st->tol2 = rtol * rtol * st->bb;
            st->it = 0;
            st->done = (a <= st->tol2) ? 1 : 0;
        } else if (which == 4) {
            st->rr = a;
            st->it += 1;
            if (a <= st->tol2) st->done = 1;
        } else if (which == 5) {
            st->rz_new = a;
            st->rzo = b;
        } else {
            st->rz_new = a;
            st->rr = b;
            st->it += 1;
            if (b <= st->tol2) st->done = 1;
        }
    }
}

// x += alpha p, r -= alpha Ap, z = Dinv r; partials r.z, r.r
__global__ __launch_bounds__(kThreads) void update_kernel(int64_t n, double* x, double* r, double* z,
                                                           const double* p, const double* Ap, const double* dinv,
                                                           double* partial, const PcgState* st) {
    if (st->done) return;
    const double alpha = st->rz / st->pAp;
    double rz = 0, rr = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        double ri = r[i] - alpha * Ap[i];
        double zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        rz += ri * zi;
        rr += ri * ri;
    }
    double s0 = block_sum(rz), s1 = block_sum(rr);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s0;
        partial[kRedBlocks + blockIdx.x] = s1;
    }
}

// p = z + beta p with beta = rz_new / rz; then rz = rz_new
__global__ void direction_kernel(int64_t n, const double* z, double* p, PcgState* st) {
    if (st->done) return;
    const double beta = st->rz_new / st->rz;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z[i] + beta * p[i];
}
__global__ void commit_kernel(PcgState* st) {
    if (!st->done) st->rz = st->rz_new;
}

// ---- factored preconditioner M^-1 = G G^T (flexible CG) -------------------
// r = b - Ax; partials r.r, b.b
__global__ __launch_bounds__(kThreads) void init_f_kernel(int64_t n, const double* b, const double* Ax, double* r,
                                                           double* partial) {
    double rr = 0, bb = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double ri = b[i] - Ax[i];
        r[i] = ri;
        rr += ri * ri;
        bb += b[i] * b[i];
    }
    const double s0 = block_sum(rr), s1 = block_sum(bb);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s0;
        partial[kRedBlocks + blockIdx.x] = s1;
    }
}

// x += alpha p, r -= alpha Ap; partial r.r
__global__ __launch_bounds__(kThreads) void update_f_kernel(int64_t n, double* x, double* r, const double* p,
                                                             const double* Ap, double* partial, const PcgState* st) {
    if (st->done) return;
    const double alpha = st->rz / st->pAp;
    double rr = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * Ap[i];
        r[i] = ri;
        rr += ri * ri;
    }
    const double s = block_sum(rr);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// z = V t (one lane group per row) with partials r.z and r.zo (zo may be null)
template <int L>
__global__ __launch_bounds__(kThreads) void prec_spmv_kernel(int64_t n, const int64_t* __restrict__ rp,
                                                              const int32_t* __restrict__ ci,
                                                              const double* __restrict__ v, const double* t,
                                                              double* z, const double* r, const double* zo,
                                                              double* partial, const PcgState* st) {
    if (st->done) return;
    const int64_t rows_per_block = kThreads / L;
    double d0 = 0, d1 = 0;
    for (int64_t base = (int64_t)blockIdx.x * rows_per_block; base < n; base += (int64_t)gridDim.x * rows_per_block) {
        const int64_t row = base + threadIdx.x / L;
        const int lane = threadIdx.x % L;
        double s = 0;
        if (row < n)
            for (int64_t k = rp[row] + lane; k < rp[row + 1]; k += L) s += v[k] * t[ci[k]];
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_down(s, off, L);
        if (row < n && lane == 0) {
            z[row] = s;
            d0 += r[row] * s;
            if (zo) d1 += r[row] * zo[row];
        }
    }
    const double s0 = block_sum(d0), s1 = block_sum(d1);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s0;
        partial[kRedBlocks + blockIdx.x] = s1;
    }
}

// p = z + beta p, beta = (rz_new - rzo) / rz (Polak-Ribiere; 0 on the first step)
__global__ void direction_f_kernel(int64_t n, const double* z, double* p, PcgState* st, int first) {
    if (st->done) return;
    const double beta = first ? 0.0 : (st->rz_new - st->rzo) / st->rz;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z[i] + beta * p[i];
}

__global__ __launch_bounds__(kThreads) void residual_kernel(int64_t n, const double* b, const double* Ax,
                                                             double* partial) {
    double rr = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double ri = b[i] - Ax[i];
        rr += ri * ri;
    }
    double s = block_sum(rr);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Returns the number of workgroups launched (= partial sums written).
template <int L>
static int launch_spmv(int64_t n, const int64_t* rp, const int32_t* ci, const double* v, const double* x, double* y,
                       double* partial, const PcgState* st, hipStream_t s) {
    const int64_t rows_per_block = kThreads / L;
    int64_t blocks = (n + rows_per_block - 1) / rows_per_block;
    if (partial) blocks = std::min<int64_t>(blocks, kRedBlocks);
    blocks = std::max<int64_t>(blocks, 1);
    hipLaunchKernelGGL(spmv_kernel<L>, dim3((unsigned)blocks), dim3(kThreads), 0, s, n, rp, ci, v, x, y, partial, st);
    return (int)blocks;
}

// Lanes per row from the average row length.
int spmv_dispatch(int64_t n, int64_t nnz, const int64_t* rp, const int32_t* ci, const double* v, const double* x,
                  double* y, double* partial, const PcgState* st, hipStream_t s) {
    const double avg = n > 0 ? (double)nnz / n : 0;
    if (avg <= 12) return launch_spmv<4>(n, rp, ci, v, x, y, partial, st, s);
    if (avg <= 24) return launch_spmv<8>(n, rp, ci, v, x, y, partial, st, s);
    if (avg <= 48) return launch_spmv<16>(n, rp, ci, v, x, y, partial, st, s);
    return launch_spmv<32>(n, rp, ci, v, x, y, partial, st, s);
}

template <int L>
static int launch_prec(int64_t n, const int64_t* rp, const int32_t* ci, const double* v, const double* t, double* z,
                       const double* r, const double* zo, double* partial, const PcgState* st, hipStream_t s) {
    const int64_t rows_per_block = kThreads / L;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n + rows_per_block - 1) / rows_per_block, kRedBlocks));
    hipLaunchKernelGGL(prec_spmv_kernel<L>, dim3((unsigned)blocks), dim3(kThreads), 0, s, n, rp, ci, v, t, z, r, zo,
                       partial, st);
    return (int)blocks;
}

int prec_dispatch(int64_t n, int64_t nnz, const int64_t* rp, const int32_t* ci, const double* v, const double* t,
                  double* z, const double* r, const double* zo, double* partial, const PcgState* st, hipStream_t s) {
    const double avg = n > 0 ? (double)nnz / n : 0;
    if (avg <= 12) return launch_prec<4>(n, rp, ci, v, t, z, r, zo, partial, st, s);
    if (avg <= 24) return launch_prec<8>(n, rp, ci, v, t, z, r, zo, partial, st, s);
    if (avg <= 48) return launch_prec<16>(n, rp, ci, v, t, z, r, zo, partial, st, s);
    return launch_prec<32>(n, rp, ci, v, t, z, r, zo, partial, st, s);
}

}  // namespace anifem

extern "C" int anifem_pcg_jacobi(int64_t n, const int64_t* row_ptr, const int32_t* col_idx, const double* val,
                                 const double* b, double* x, double rtol, int64_t maxit, anifem_pcg_result* res,
                                 void* stream) {
    using namespace anifem;
    if (n <= 0 || !row_ptr || !col_idx || !val || !b || !x || rtol <= 0 || maxit < 0) return ANIFEM_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t nnz = 0;
    ANIFEM_HIP_CHECK(hipMemcpyAsync(&nnz, row_ptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
    double* work = nullptr;
    const size_t nv = (size_t)n;
    ANIFEM_HIP_CHECK(hipMalloc((void**)&work, sizeof(double) * (5 * nv + 3 * kRedBlocks) + sizeof(PcgState)));
    double *r = work, *z = r + nv, *p = z + nv, *Ap = p + nv, *dinv = Ap + nv, *part = dinv + nv;
    PcgState* st = reinterpret_cast<PcgState*>(part + 3 * kRedBlocks);
    const unsigned vb = (unsigned)std::min<int64_t>((n + kThreads - 1) / kThreads, kRedBlocks);
    int rc = ANIFEM_OK;
    hipLaunchKernelGGL(jacobi_inverse_kernel, dim3(vb), dim3(kThreads), 0, s, n, row_ptr, col_idx, val, dinv);
    spmv_dispatch(n, nnz, row_ptr, col_idx, val, x, Ap, nullptr, nullptr, s);
    hipLaunchKernelGGL(init_kernel, dim3(vb), dim3(kThreads), 0, s, n, b, Ap, dinv, r, z, p, part);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kThreads), 0, s, part, (int)vb, 0, st, rtol);
    PcgState hs{};
    const int64_t poll = 16;
    for (int64_t it = 0; it < maxit; ++it) {
        const int spmv_blocks = spmv_dispatch(n, nnz, row_ptr, col_idx, val, p, Ap, part, st, s);
        hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kThreads), 0, s, part, spmv_blocks, 1, st, rtol);
        hipLaunchKernelGGL(update_kernel, dim3(vb), dim3(kThreads), 0, s, n, x, r, z, p, Ap, dinv, part, st);
        hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kThreads), 0, s, part, (int)vb, 2, st, rtol);
        hipLaunchKernelGGL(direction_kernel, dim3(vb), dim3(kThreads), 0, s, n, z, p, st);
        hipLaunchKernelGGL(commit_kernel, dim3(1), dim3(1), 0, s, st);
        if ((it + 1) % poll == 0 || it + 1 == maxit) {
            if (hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                rc = ANIFEM_ERR_HIP;
                break;
            }
            if (hs.done) break;
        }
    }
    if (rc == ANIFEM_OK) {
        // true residual from x
        spmv_dispatch(n, nnz, row_ptr, col_idx, val, x, Ap, nullptr, nullptr, s);
        hipLaunchKernelGGL(residual_kernel, dim3(vb), dim3(kThreads), 0, s, n, b, Ap, part);
        double* hp = new double[vb];
        if (hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(hp, part, sizeof(double) * vb, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = ANIFEM_ERR_HIP;
        } else {
            double tr = 0;
            for (unsigned i = 0; i < vb; ++i) tr += hp[i];
            if (res) {
                res->iterations = hs.it;
                res->relres = std::sqrt(hs.rr / hs.bb);
                res->true_relres = std::sqrt(tr / hs.bb);
            }
            if (!hs.done) rc = ANIFEM_ERR_NOT_CONVERGED;
        }
        delete[] hp;
    }
    (void)hipFree(work);
    return rc;
}

// Flexible PCG with the factored preconditioner M^-1 = G G^T (G, GT: values of
// G and of G^T on the pattern of A).  Same stopping rule and reporting as
// anifem_pcg_jacobi.
extern "C" int anifem_pcg_factored(int64_t n, const int64_t* row_ptr, const int32_t* col_idx, const double* val,
                                   const double* G, const double* GT, const double* b, double* x, double rtol,
                                   int64_t maxit, anifem_pcg_result* res, void* stream) {
    using namespace anifem;
    if (n <= 0 || !row_ptr || !col_idx || !val || !G || !GT || !b || !x || rtol <= 0 || maxit < 0)
        return ANIFEM_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t nnz = 0;
    ANIFEM_HIP_CHECK(hipMemcpyAsync(&nnz, row_ptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
    double* work = nullptr;
    const size_t nv = (size_t)n;
    ANIFEM_HIP_CHECK(hipMalloc((void**)&work, sizeof(double) * (6 * nv + 3 * kRedBlocks) + sizeof(PcgState)));
    double *r = work, *t = r + nv, *za = t + nv, *zb = za + nv, *p = zb + nv, *Ap = p + nv, *part = Ap + nv;
    PcgState* st = reinterpret_cast<PcgState*>(part + 3 * kRedBlocks);
    const unsigned vb = (unsigned)std::min<int64_t>((n + kThreads - 1) / kThreads, kRedBlocks);
    int rc = ANIFEM_OK;
    // r0 = b - A x0, z0 = G G^T r0, p0 = z0
    spmv_dispatch(n, nnz, row_ptr, col_idx, val, x, Ap, nullptr, nullptr, s);
    hipLaunchKernelGGL(init_f_kernel, dim3(vb), dim3(kThreads), 0, s, n, b, Ap, r, part);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kThreads), 0, s, part, (int)vb, 3, st, rtol);
    spmv_dispatch(n, nnz, row_ptr, col_idx, GT, r, t, nullptr, st, s);
    int pb = prec_dispatch(n, nnz, row_ptr, col_idx, G, t, za, r, nullptr, part, st, s);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kThreads), 0, s, part, pb, 5, st, rtol);
    hipLaunchKernelGGL(direction_f_kernel, dim3(vb), dim3(kThreads), 0, s, n, za, p, st, 1);
    hipLaunchKernelGGL(commit_kernel, dim3(1), dim3(1), 0, s, st);
    double* zold = za;
    double* znew = zb;
    PcgState hs{};
    const int64_t poll = 16;
    for (int64_t it = 0; it < maxit; ++it) {
        const int spmv_blocks = spmv_dispatch(n, nnz, row_ptr, col_idx, val, p, Ap, part, st, s);
        hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kThreads), 0, s, part, spmv_blocks, 1, st, rtol);
        hipLaunchKernelGGL(update_f_kernel, dim3(vb), dim3(kThreads), 0, s, n, x, r, p, Ap, part, st);
        hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kThreads), 0, s, part, (int)vb, 4, st, rtol);
        spmv_dispatch(n, nnz, row_ptr, col_idx, GT, r, t, nullptr, st, s);
        pb = prec_dispatch(n, nnz, row_ptr, col_idx, G, t, znew, r, zold, part, st, s);
        hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kThreads), 0, s, part, pb, 5, st, rtol);
        hipLaunchKernelGGL(direction_f_kernel, dim3(vb), dim3(kThreads), 0, s, n, znew, p, st, 0);
        hipLaunchKernelGGL(commit_kernel, dim3(1), dim3(1), 0, s, st);
        std::swap(zold, znew);
        if ((it + 1) % poll == 0 || it + 1 == maxit) {
            if (hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                rc = ANIFEM_ERR_HIP;
                break;
            }
            if (hs.done) break;
        }
    }
    if (rc == ANIFEM_OK) {
        spmv_dispatch(n, nnz, row_ptr, col_idx, val, x, Ap, nullptr, nullptr, s);
        hipLaunchKernelGGL(residual_kernel, dim3(vb), dim3(kThreads), 0, s, n, b, Ap, part);
        std::vector<double> hp(vb);
        if (hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(hp.data(), part, sizeof(double) * vb, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = ANIFEM_ERR_HIP;
        } else {
            double tr = 0;
            for (unsigned i = 0; i < vb; ++i) tr += hp[i];
            if (res) {
                res->iterations = hs.it;
                res->relres = std::sqrt(hs.rr / hs.bb);
                res->true_relres = std::sqrt(tr / hs.bb);
            }
            if (!hs.done) rc = ANIFEM_ERR_NOT_CONVERGED;
        }
    }
    (void)hipFree(work);
    return rc;
}

// Learned-preconditioned solve: the network's factor G for this matrix, then
// flexible PCG with M^-1 = G G^T; the FP64 true residual judges the result
// (||b - A x|| <= rtol ||b||, up to 1% for the recursive-residual drift) and
// Jacobi-PCG from the same initial guess takes over when it fails.
extern "C" int anifem_pcg_learned(anifem_gnn* g, const double* val, const double* b, double* x, double rtol,
                                  int64_t maxit, anifem_pcg_learned_result* res, void* stream) {
    using namespace anifem;
    if (!g || !val || !b || !x) return ANIFEM_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t n = 0;
    const int64_t* rp = nullptr;
    const int32_t* ci = nullptr;
    const double *G = nullptr, *GT = nullptr;
    int rc = anifem_gnn_views(g, &n, &rp, &ci, &G, &GT);
    if (rc) return rc;
    double* x0 = nullptr;
    ANIFEM_HIP_CHECK(hipMalloc((void**)&x0, n * sizeof(double)));
    anifem_pcg_learned_result out{};
    if (hipMemcpyAsync(x0, x, n * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess) rc = ANIFEM_ERR_HIP;
    if (!rc) rc = anifem_gnn_factor(g, val, nullptr, nullptr, stream);
    anifem_pcg_result r1{};
    int rc1 = rc ? rc : anifem_pcg_factored(n, rp, ci, val, G, GT, b, x, rtol, maxit, &r1, stream);
    out.learned_iterations = r1.iterations;
    out.learned_true_relres = r1.true_relres;
    const bool ok = (rc1 == ANIFEM_OK || rc1 == ANIFEM_ERR_NOT_CONVERGED) && r1.true_relres <= 1.01 * rtol &&
                    std::isfinite(r1.true_relres);
    if (rc == ANIFEM_OK && (rc1 == ANIFEM_ERR_HIP || rc1 == ANIFEM_ERR_NOMEM)) rc = rc1;
    if (!rc) {
        if (ok) {
            out.iterations = r1.iterations;
            out.relres = r1.relres;
            out.true_relres = r1.true_relres;
            out.fallback = 0;
        } else {
            anifem_pcg_result r2{};
            if (hipMemcpyAsync(x, x0, n * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess) rc = ANIFEM_ERR_HIP;
            const int rc2 = rc ? rc : anifem_pcg_jacobi(n, rp, ci, val, b, x, rtol, maxit, &r2, stream);
            out.iterations = r2.iterations;
            out.relres = r2.relres;
            out.true_relres = r2.true_relres;
            out.fallback = 1;
            if (!rc) rc = rc2;
        }
    }
    (void)hipFree(x0);
    if (res) *res = out;
    return rc;
}
