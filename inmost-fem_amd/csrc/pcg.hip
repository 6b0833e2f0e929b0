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
#include <vector>

#include "anifem_common.h"

namespace anifem {

struct PcgState {
    double rr, bb, tol2;
    double rzb[2];  // r.z of iteration k at rzb[k & 1]
    int64_t it;
    int done;
};

constexpr int kRedBlocks = 2048;  // partial sums per reduction
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

// Second reduction stage of the set-up; `which`: 0 Jacobi init (rz, rr, bb),
// 3 factored init (rr, bb), 5 factored init r.z.
__global__ __launch_bounds__(kThreads) void reduce_kernel(const double* partial, int nparts, int which,
                                                           PcgState* st, double rtol) {
    if (which == 5 && st->done) return;
    double a = 0, b = 0, c = 0;
    for (int i = threadIdx.x; i < nparts; i += kThreads) {
        a += partial[i];
        if (which != 5) b += partial[kRedBlocks + i];
        if (which == 0) c += partial[2 * kRedBlocks + i];
    }
    a = block_sum(a);
    b = block_sum(b);
    c = block_sum(c);
    if (threadIdx.x == 0) {
        if (which == 0) {
            st->rzb[0] = a;
            st->rr = b;
            st->bb = c > 0 ? c : 1.0;
            st->tol2 = rtol * rtol * st->bb;
            st->it = 0;
            st->done = (b <= st->tol2) ? 1 : 0;
        } else if (which == 3) {
            st->rr = a;
            st->bb = b > 0 ? b : 1.0;
            st->tol2 = rtol * rtol * st->bb;
            st->it = 0;
            st->done = (a <= st->tol2) ? 1 : 0;
        } else {
            st->rzb[0] = a;
        }
    }
}

// ---- fused iteration: each kernel sums the partial sums of the previous one
// itself (every workgroup in the same fixed order), and the iteration's r.z
// alternates between rzb[k & 1] and rzb[(k + 1) & 1] -------------------------
// this thread's share of a partial-sum array: its loads are issued at once and
// can overlap other loads; block_bcast then finishes the sum (fixed order)
__device__ __forceinline__ double thread_partial(const double* part, int nparts) {
    double a = 0;
#pragma unroll
    for (int j = 0; j < kRedBlocks / kThreads; ++j) {
        const int i = threadIdx.x + j * kThreads;
        if (i < nparts) a += part[i];
    }
    return a;
}

__device__ double block_bcast(double a) {
    a = block_sum(a);
    __shared__ double bc;
    if (threadIdx.x == 0) bc = a;
    __syncthreads();
    const double r = bc;
    __syncthreads();
    return r;
}

__device__ double block_total(const double* part, int nparts) { return block_bcast(thread_partial(part, nparts)); }

// NV block sums at once, every thread receiving the totals: one barrier pair
// for all of them.  Wave trees then waves 0..3 in order, as block_sum.
template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV]) {
    __shared__ double sh[NV][kThreads / 64];
#pragma unroll
    for (int i = 0; i < NV; ++i)
        for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_down(v[i], off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) sh[i][w] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double r = 0;
        for (int j = 0; j < kThreads / 64; ++j) r += sh[i][j];
        v[i] = r;
    }
    __syncthreads();
}

// ---- two launches per iteration ----------------------------------------
// Iteration k runs dir_spmv(k) then update(k):
//   dir_spmv(k): test r_k.r_k (done, it = k); beta = rz_k / rz_(k-1), or
//                (rz_k - r_k.z_(k-1)) / rz_(k-1) (flexible, Polak-Ribiere);
//                p_k = z_k + beta p_(k-1) (p_0 = z_0); Ap_k = A p_k; partials p_k.Ap_k
//   update(k):   alpha = rz_k / p_k.Ap_k; x += alpha p_k; r_(k+1) = r_k - alpha Ap_k;
//                z_(k+1) = M r_(k+1); partials r.z, r.r (and r.z_old if flexible)
// A row of A p_k needs p_k at its neighbours, which other workgroups write in the
// same launch, so every reader forms p_k = fma(beta, p_(k-1), z_k) itself from
// the previous buffer (p alternates between two buffers); likewise r_(k+1) for
// z = M r_(k+1).  The explicit fma makes every lane's copy bitwise equal.
// `hdone` (host-mapped, may be null) receives the done flag for the host's poll.
// Entries per lane whose operands are loaded before the scalars of the step
// are known (rows up to kPre * L entries; longer rows finish in a second loop).
constexpr int kPre = 8;

template <int L>
__global__ __launch_bounds__(kThreads) void dir_spmv_kernel(int64_t n, const int64_t* __restrict__ rp,
                                                             const int32_t* __restrict__ ci,
                                                             const double* __restrict__ v, const double* z,
                                                             const double* p_prev, double* p, double* Ap,
                                                             double* partA, const double* partU, int nU,
                                                             int flexible, PcgState* st, int64_t k,
                                                             int check_only, int* hdone) {
    const bool first = k == 0;
    const int64_t rows_per_block = kThreads / L;
    const int lane = threadIdx.x % L;
    // loads that do not depend on beta: the first row pass's entries and gathers
    const int64_t row0 = (int64_t)blockIdx.x * rows_per_block + threadIdx.x / L;
    // a row's first kPre entries per lane: all loads issued before any is used
    double ev[kPre], ez[kPre], ep[kPre];
    auto load_row = [&](int64_t row, int64_t& qa, int64_t& qb) {
        qa = qb = 0;
        if (row < n) {
            qa = rp[row];
            qb = rp[row + 1];
        }
#pragma unroll
        for (int j = 0; j < kPre; ++j) {
            const int64_t q = qa + lane + (int64_t)j * L;
            ev[j] = ez[j] = ep[j] = 0;
            if (q < qb) {
                const int32_t c = ci[q];
                ev[j] = v[q];
                ez[j] = z[c];
                if (!first) ep[j] = p_prev[c];
            }
        }
    };
    int64_t q0, q1;
    load_row(row0, q0, q1);
    double zr = 0, pr = 0;
    if (row0 < n && lane == 0) {
        zr = z[row0];
        if (!first) pr = p_prev[row0];
    }
    double tot[3] = {0.0, 0.0, 0.0};  // r.r, r.z, r.z_old
    if (!first) {
        tot[0] = thread_partial(partU + kRedBlocks, nU);
        tot[1] = thread_partial(partU, nU);
        if (flexible) tot[2] = thread_partial(partU + 2 * kRedBlocks, nU);
    }
    if (st->done) return;
    if (!first) block_sums(tot);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const double rr = first ? st->rr : tot[0];
    if (rr <= st->tol2 || check_only) {
        if (lead) {
            st->rr = rr;
            st->it = k;
            if (rr <= st->tol2) {
                st->done = 1;
                if (hdone) *hdone = 1;
            }
        }
        return;
    }
    double beta = 0;
    if (!first) {
        // Polak-Ribiere (flexible) for a preconditioner that may not be exactly linear
        const double rz = tot[1];
        beta = (rz - tot[2]) / st->rzb[(k - 1) & 1];
        if (lead) st->rzb[k & 1] = rz;
    }
    if (lead) {
        st->rr = rr;
        st->it = k;
    }
    double dot = 0;
    for (int64_t base = (int64_t)blockIdx.x * rows_per_block; base < n; base += (int64_t)gridDim.x * rows_per_block) {
        const int64_t row = base + threadIdx.x / L;
        const bool pre = row == row0;  // the first pass, loaded above
        if (!pre) load_row(row, q0, q1);
        double s = 0;
#pragma unroll
        for (int j = 0; j < kPre; ++j) s += ev[j] * (first ? ez[j] : fma(beta, ep[j], ez[j]));
        if (row < n) {
            const int64_t qs = q0 + lane + (int64_t)kPre * L;
            const int64_t qe = q1;
            for (int64_t q = qs; q < qe; q += L) {
                const int32_t c = ci[q];
                s += v[q] * (first ? z[c] : fma(beta, p_prev[c], z[c]));
            }
        }
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_down(s, off, L);
        if (row < n && lane == 0) {
            const double zz = pre ? zr : z[row];
            const double pk = first ? zz : fma(beta, pre ? pr : p_prev[row], zz);
            p[row] = pk;
            Ap[row] = s;
            dot += s * pk;
        }
    }
    double out[1] = {dot};
    block_sums(out);
    if (threadIdx.x == 0) partA[blockIdx.x] = out[0];
}

// ---- three launches per iteration (large systems) ------------------------
// The direction update as a row-local launch of its own, so that the SpMV
// gathers one vector per entry instead of forming p at every neighbour:
//   direction(k): the convergence test and beta of dir_spmv; p_k = z_k + beta p_(k-1)
//   spmv_pdot(k): Ap_k = A p_k; partials p_k.Ap_k
//   update(k):    as in the two-launch iteration.
// p_k is the same fma of the same operands as in dir_spmv, so both forms give
// bitwise identical iterates; the split adds one launch and 3 vectors of row
// traffic, and halves the gathers of the SpMV.
__global__ __launch_bounds__(kThreads) void direction_kernel(int64_t n, const double* z, const double* p_prev,
                                                              double* p, const double* partU, int nU, int flexible,
                                                              PcgState* st, int64_t k, int check_only, int* hdone) {
    const bool first = k == 0;
    const int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double z0 = 0, p0 = 0;
    if (i0 < n) {
        z0 = z[i0];
        if (!first) p0 = p_prev[i0];
    }
    double tot[3] = {0.0, 0.0, 0.0};  // r.r, r.z, r.z_old
    if (!first) {
        tot[0] = thread_partial(partU + kRedBlocks, nU);
        tot[1] = thread_partial(partU, nU);
        if (flexible) tot[2] = thread_partial(partU + 2 * kRedBlocks, nU);
    }
    if (st->done) return;
    if (!first) block_sums(tot);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const double rr = first ? st->rr : tot[0];
    if (rr <= st->tol2 || check_only) {
        if (lead) {
            st->rr = rr;
            st->it = k;
            if (rr <= st->tol2) {
                st->done = 1;
                if (hdone) *hdone = 1;
            }
        }
        return;
    }
    double beta = 0;
    if (!first) {
        const double rz = tot[1];
        beta = (rz - tot[2]) / st->rzb[(k - 1) & 1];
        if (lead) st->rzb[k & 1] = rz;
    }
    if (lead) {
        st->rr = rr;
        st->it = k;
    }
    for (int64_t i = i0; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const bool pre = i == i0;
        const double zi = pre ? z0 : z[i];
        p[i] = first ? zi : fma(beta, pre ? p0 : p_prev[i], zi);
    }
}

// Entries per lane of a row that the three-launch SpMV loads ahead.
constexpr int kPreS = 8;

// Ap = A p with L lanes per row, one gathered operand per entry; partials
// p.Ap.  Two rows are in flight per lane group: the next row's entries and
// column indices are requested before the current row is summed, and its
// operands are gathered right after (the first row's loads are issued before
// the done test).  The sum of a row is formed in the same order as in
// dir_spmv_kernel.  (Giving each XCD a contiguous eighth of the rows cut the
// HBM fetch from 4.4 to 3.5 GB per launch at the elasticity size but made the
// launch slower, 0.96 against 0.84 ms per iteration.)
template <int L, int KP>
__global__ __launch_bounds__(kThreads) void spmv_pdot_kernel(int64_t n, const int64_t* __restrict__ rp,
                                                              const int32_t* __restrict__ ci,
                                                              const double* __restrict__ v,
                                                              const double* __restrict__ p, double* __restrict__ Ap,
                                                              double* partA, const PcgState* st) {
    const int64_t rows_per_block = kThreads / L;
    const int64_t stride = (int64_t)gridDim.x * rows_per_block;
    const int lane = threadIdx.x % L;
    double ev[KP], ep[KP], nv[KP];
    int32_t nc[KP];
    int64_t q0 = 0, q1 = 0, n0 = 0, n1 = 0;
    auto load_entries = [&](int64_t row) {
        n0 = n1 = 0;
        if (row < n) {
            n0 = rp[row];
            n1 = rp[row + 1];
        }
#pragma unroll
        for (int j = 0; j < KP; ++j) {
            const int64_t q = n0 + lane + (int64_t)j * L;
            nv[j] = 0;
            nc[j] = -1;
            if (q < n1) {
                nv[j] = v[q];
                nc[j] = ci[q];
            }
        }
    };
    auto gather = [&]() {
#pragma unroll
        for (int j = 0; j < KP; ++j) {
            ev[j] = nv[j];
            ep[j] = nc[j] >= 0 ? p[nc[j]] : 0.0;
        }
        q0 = n0;
        q1 = n1;
    };
    int64_t row = (int64_t)blockIdx.x * rows_per_block + threadIdx.x / L;
    load_entries(row);
    if (st->done) return;
    gather();
    double dot = 0;
    for (int64_t base = (int64_t)blockIdx.x * rows_per_block; base < n; base += stride) {
        load_entries(row + stride);
        double s = 0;
#pragma unroll
        for (int j = 0; j < KP; ++j) s += ev[j] * ep[j];
        if (row < n)
            for (int64_t q = q0 + lane + (int64_t)KP * L; q < q1; q += L) s += v[q] * p[ci[q]];
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_down(s, off, L);
        if (row < n && lane == 0) {
            Ap[row] = s;
            dot += s * p[row];
        }
        gather();
        row += stride;
    }
    double out[1] = {dot};
    block_sums(out);
    if (threadIdx.x == 0) partA[blockIdx.x] = out[0];
}

// Jacobi update: z = Dinv r (row-local, r updated in place)
__global__ __launch_bounds__(kThreads) void update_jacobi_kernel(int64_t n, double* x, double* r, double* z,
                                                                  const double* p, const double* Ap,
                                                                  const double* dinv, const double* partA, int nA,
                                                                  double* partU, const PcgState* st, int64_t k) {
    const int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double x0 = 0, p0 = 0, a0 = 0, r0 = 0, d0 = 0;
    if (i0 < n) {
        x0 = x[i0];
        p0 = p[i0];
        a0 = Ap[i0];
        r0 = r[i0];
        d0 = dinv[i0];
    }
    double pap[1] = {thread_partial(partA, nA)};
    if (st->done) return;
    block_sums(pap);
    const double alpha = st->rzb[k & 1] / pap[0];
    double rz = 0, rr = 0;
    for (int64_t i = i0; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const bool pre = i == i0;
        x[i] = fma(alpha, pre ? p0 : p[i], pre ? x0 : x[i]);
        const double ri = fma(-alpha, pre ? a0 : Ap[i], pre ? r0 : r[i]);
        const double zi = (pre ? d0 : dinv[i]) * ri;
        r[i] = ri;
        z[i] = zi;
        rz += ri * zi;
        rr += ri * ri;
    }
    double out[2] = {rz, rr};
    block_sums(out);
    if (threadIdx.x == 0) {
        partU[blockIdx.x] = out[0];
        partU[kRedBlocks + blockIdx.x] = out[1];
    }
}

// explicit preconditioner update: z = M r_(k+1), r_(k+1) formed by every reader
template <int L>
__global__ __launch_bounds__(kThreads) void update_matrix_kernel(int64_t n, const int64_t* __restrict__ mrp,
                                                                  const int32_t* __restrict__ mci,
                                                                  const double* __restrict__ mv, double* x,
                                                                  const double* r_prev, double* r, double* z,
                                                                  const double* p, const double* Ap,
                                                                  const double* partA, int nA, double* partU,
                                                                  const PcgState* st, int64_t k) {
    const int64_t rows_per_block = kThreads / L;
    const int lane = threadIdx.x % L;
    const int64_t row0 = (int64_t)blockIdx.x * rows_per_block + threadIdx.x / L;
    int64_t q0 = 0, q1 = 0;
    if (row0 < n) {
        q0 = mrp[row0];
        q1 = mrp[row0 + 1];
    }
    double ev[kPre], ea[kPre], er[kPre];
#pragma unroll
    for (int j = 0; j < kPre; ++j) {
        const int64_t q = q0 + lane + (int64_t)j * L;
        ev[j] = ea[j] = er[j] = 0;
        if (q < q1) {
            const int32_t c = mci[q];
            ev[j] = mv[q];
            ea[j] = Ap[c];
            er[j] = r_prev[c];
        }
    }
    double xr = 0, pr = 0, ar = 0, rr0 = 0, zo0 = 0;
    if (row0 < n && lane == 0) {
        xr = x[row0];
        pr = p[row0];
        ar = Ap[row0];
        rr0 = r_prev[row0];
        zo0 = z[row0];
    }
    double pap[1] = {thread_partial(partA, nA)};
    if (st->done) return;
    block_sums(pap);
    const double alpha = st->rzb[k & 1] / pap[0];
    double rz = 0, rr = 0, rzo = 0;
    for (int64_t base = (int64_t)blockIdx.x * rows_per_block; base < n; base += (int64_t)gridDim.x * rows_per_block) {
        const int64_t row = base + threadIdx.x / L;
        const bool pre = row == row0;
        double s = 0;
        if (pre) {
#pragma unroll
            for (int j = 0; j < kPre; ++j) s += ev[j] * fma(-alpha, ea[j], er[j]);
        }
        if (row < n) {
            const int64_t qs = pre ? q0 + lane + (int64_t)kPre * L : mrp[row] + lane;
            const int64_t qe = mrp[row + 1];
            for (int64_t q = qs; q < qe; q += L) {
                const int32_t c = mci[q];
                s += mv[q] * fma(-alpha, Ap[c], r_prev[c]);
            }
        }
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_down(s, off, L);
        if (row < n && lane == 0) {
            const double ri = fma(-alpha, pre ? ar : Ap[row], pre ? rr0 : r_prev[row]);
            x[row] = fma(alpha, pre ? pr : p[row], pre ? xr : x[row]);
            rzo += ri * (pre ? zo0 : z[row]);
            r[row] = ri;
            z[row] = s;
            rz += ri * s;
            rr += ri * ri;
        }
    }
    double out[3] = {rz, rr, rzo};
    block_sums(out);
    if (threadIdx.x == 0) {
        partU[blockIdx.x] = out[0];
        partU[kRedBlocks + blockIdx.x] = out[1];
        partU[2 * kRedBlocks + blockIdx.x] = out[2];
    }
}

// factored: x += alpha p, r -= alpha Ap; partial r.r
__global__ __launch_bounds__(kThreads) void update_fk_kernel(int64_t n, double* x, double* r, const double* p,
                                                              const double* Ap, const double* partA, int nA,
                                                              double* partB, const PcgState* st, int64_t k) {
    if (st->done) return;
    const double pAp = block_total(partA, nA);
    const double alpha = st->rzb[k & 1] / pAp;
    double rr = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * Ap[i];
        r[i] = ri;
        rr += ri * ri;
    }
    const double s = block_sum(rr);
    if (threadIdx.x == 0) partB[blockIdx.x] = s;
}

// factored: convergence test, then t = G^T r (L lanes per row)
template <int L>
__global__ __launch_bounds__(kThreads) void gt_spmv_kernel(int64_t n, const int64_t* __restrict__ rp,
                                                            const int32_t* __restrict__ ci,
                                                            const double* __restrict__ v, const double* r, double* t,
                                                            const double* partB, int nB, PcgState* st, int64_t k,
                                                            int* hdone) {
    if (st->done) return;
    const double rr = block_total(partB, nB);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (rr <= st->tol2) {
        if (lead) {
            st->rr = rr;
            st->it = k + 1;
            st->done = 1;
            if (hdone) *hdone = 1;
        }
        return;
    }
    if (lead) st->rr = rr;
    const int64_t rows_per_block = kThreads / L;
    for (int64_t base = (int64_t)blockIdx.x * rows_per_block; base < n; base += (int64_t)gridDim.x * rows_per_block) {
        const int64_t row = base + threadIdx.x / L;
        const int lane = threadIdx.x % L;
        double s = 0;
        if (row < n)
            for (int64_t q = rp[row] + lane; q < rp[row + 1]; q += L) s += v[q] * r[ci[q]];
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_down(s, off, L);
        if (row < n && lane == 0) t[row] = s;
    }
}

// factored: p = z + beta p, beta = (rz_new - rzo) / rz_k (Polak-Ribiere)
__global__ __launch_bounds__(kThreads) void direction_fk_kernel(int64_t n, const double* z, double* p,
                                                                 const double* partD, int nD, PcgState* st,
                                                                 int64_t k) {
    if (st->done) return;
    const double rz_new = block_total(partD, nD);
    const double rzo = block_total(partD + kRedBlocks, nD);
    const double beta = (rz_new - rzo) / st->rzb[k & 1];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z[i] + beta * p[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->rzb[(k + 1) & 1] = rz_new;
        st->it = k + 1;
    }
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

// first direction p = z
__global__ void copy_direction_kernel(int64_t n, const double* z, double* p, const PcgState* st) {
    if (st->done) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z[i];
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

template <int L>
static void launch_gt(int64_t n, const int64_t* rp, const int32_t* ci, const double* v, const double* r, double* t,
                      const double* partB, int nB, PcgState* st, int64_t k, int* hdone, hipStream_t s) {
    const int64_t rows_per_block = kThreads / L;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n + rows_per_block - 1) / rows_per_block, kRedBlocks));
    hipLaunchKernelGGL(gt_spmv_kernel<L>, dim3((unsigned)blocks), dim3(kThreads), 0, s, n, rp, ci, v, r, t, partB, nB,
                       st, k, hdone);
}

void gt_dispatch(int64_t n, int64_t nnz, const int64_t* rp, const int32_t* ci, const double* v, const double* r,
                 double* t, const double* partB, int nB, PcgState* st, int64_t k, int* hdone, hipStream_t s) {
    const double avg = n > 0 ? (double)nnz / n : 0;
    if (avg <= 12) return launch_gt<4>(n, rp, ci, v, r, t, partB, nB, st, k, hdone, s);
    if (avg <= 24) return launch_gt<8>(n, rp, ci, v, r, t, partB, nB, st, k, hdone, s);
    if (avg <= 48) return launch_gt<16>(n, rp, ci, v, r, t, partB, nB, st, k, hdone, s);
    return launch_gt<32>(n, rp, ci, v, r, t, partB, nB, st, k, hdone, s);
}

// Lanes per row for the two-launch iteration: about six entries per lane, so
// that a row's operands fit the kPre prefetch registers and a workgroup
// covers many rows (fewer partial sums for the next kernel to read).
int lanes_per_row(int64_t n, int64_t nnz) {
    const double avg = n > 0 ? (double)nnz / n : 0;
    return avg <= 12 ? 2 : avg <= 24 ? 4 : avg <= 48 ? 8 : avg <= 96 ? 16 : 32;
}

int red_grid(int64_t n, int L) {
    const int64_t rows_per_block = kThreads / L;
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + rows_per_block - 1) / rows_per_block, kRedBlocks));
}

// A grid-stride launch of no more workgroups than are resident at once: a
// second, partial round would leave most CUs idle while it runs.
int resident_grid(const void* kernel, int grid) {
    int dev = 0, ncu = 0, occ = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, kThreads, 0) == hipSuccess && occ > 0 && ncu > 0)
        grid = (int)std::min<int64_t>(grid, (int64_t)occ * ncu);
    return std::max(grid, 1);
}

const void* dir_kernel_of(int L) {
    if (L == 2) return reinterpret_cast<const void*>(&dir_spmv_kernel<2>);
    if (L == 4) return reinterpret_cast<const void*>(&dir_spmv_kernel<4>);
    if (L == 8) return reinterpret_cast<const void*>(&dir_spmv_kernel<8>);
    if (L == 16) return reinterpret_cast<const void*>(&dir_spmv_kernel<16>);
    return reinterpret_cast<const void*>(&dir_spmv_kernel<32>);
}

const void* update_matrix_kernel_of(int L) {
    if (L == 2) return reinterpret_cast<const void*>(&update_matrix_kernel<2>);
    if (L == 4) return reinterpret_cast<const void*>(&update_matrix_kernel<4>);
    if (L == 8) return reinterpret_cast<const void*>(&update_matrix_kernel<8>);
    if (L == 16) return reinterpret_cast<const void*>(&update_matrix_kernel<16>);
    return reinterpret_cast<const void*>(&update_matrix_kernel<32>);
}

void launch_dir(int L, int grid, int64_t n, const int64_t* rp, const int32_t* ci, const double* v, const double* z,
                const double* p_prev, double* p, double* Ap, double* partA, const double* partU, int nU, int flexible,
                PcgState* st, int64_t k, int check_only, int* hdone, hipStream_t s) {
#define ANIFEM_DIR(LL)                                                                                               \
    hipLaunchKernelGGL(dir_spmv_kernel<LL>, dim3((unsigned)grid), dim3(kThreads), 0, s, n, rp, ci, v, z, p_prev, p, \
                       Ap, partA, partU, nU, flexible, st, k, check_only, hdone)
    if (L == 2) ANIFEM_DIR(2);
    else if (L == 4) ANIFEM_DIR(4);
    else if (L == 8) ANIFEM_DIR(8);
    else if (L == 16) ANIFEM_DIR(16);
    else ANIFEM_DIR(32);
#undef ANIFEM_DIR
}

const void* spmv_pdot_kernel_of(int L) {
    if (L == 2) return reinterpret_cast<const void*>(&spmv_pdot_kernel<2, kPreS>);
    if (L == 4) return reinterpret_cast<const void*>(&spmv_pdot_kernel<4, kPreS>);
    if (L == 8) return reinterpret_cast<const void*>(&spmv_pdot_kernel<8, kPreS>);
    if (L == 16) return reinterpret_cast<const void*>(&spmv_pdot_kernel<16, kPreS>);
    return reinterpret_cast<const void*>(&spmv_pdot_kernel<32, kPreS>);
}

void launch_spmv_pdot(int L, int grid, int64_t n, const int64_t* rp, const int32_t* ci, const double* v,
                      const double* p, double* Ap, double* partA, const PcgState* st, hipStream_t s) {
#define ANIFEM_SPD(LL) \
    hipLaunchKernelGGL((spmv_pdot_kernel<LL, kPreS>), dim3((unsigned)grid), dim3(kThreads), 0, s, n, rp, ci, v, p, Ap, partA, st)
    if (L == 2) ANIFEM_SPD(2);
    else if (L == 4) ANIFEM_SPD(4);
    else if (L == 8) ANIFEM_SPD(8);
    else if (L == 16) ANIFEM_SPD(16);
    else ANIFEM_SPD(32);
#undef ANIFEM_SPD
}

// Systems whose SpMV outweighs a launch take the three-launch iteration.
// (Streaming the matrix with nontemporal loads was measured slower at the
// elasticity size, 1.16 against 0.97 ms per iteration, and is not used.)
constexpr int64_t kSplitNnz = int64_t(1) << 22;

void launch_update_matrix(int L, int grid, int64_t n, const int64_t* mrp, const int32_t* mci, const double* mv,
                          double* x, const double* r_prev, double* r, double* z, const double* p, const double* Ap,
                          const double* partA, int nA, double* partU, const PcgState* st, int64_t k, hipStream_t s) {
#define ANIFEM_UPD(LL)                                                                                             \
    hipLaunchKernelGGL(update_matrix_kernel<LL>, dim3((unsigned)grid), dim3(kThreads), 0, s, n, mrp, mci, mv, x, \
                       r_prev, r, z, p, Ap, partA, nA, partU, st, k)
    if (L == 2) ANIFEM_UPD(2);
    else if (L == 4) ANIFEM_UPD(4);
    else if (L == 8) ANIFEM_UPD(8);
    else if (L == 16) ANIFEM_UPD(16);
    else ANIFEM_UPD(32);
#undef ANIFEM_UPD
}

// The done flag as the host sees it: one pinned, host-mapped word per host
// thread (allocated on the thread's first solve and held until the process
// ends).  Null if pinned memory is unavailable; the poll then copies the state.
int* host_done_flag(int** dev) {
    thread_local int* h = nullptr;
    thread_local int* d = nullptr;
    if (!h) {
        int* p = nullptr;
        if (hipHostMalloc((void**)&p, 64, hipHostMallocMapped | hipHostMallocCoherent | hipHostMallocPortable) !=
            hipSuccess)
            return *dev = nullptr;
        if (hipHostGetDevicePointer((void**)&d, p, 0) != hipSuccess) {
            (void)hipHostFree(p);
            return *dev = nullptr;
        }
        h = p;
    }
    *dev = d;
    return h;
}

// Queue the iterations in chunks; after queueing chunk c, wait for chunk c-1
// and stop once a kernel has raised the done flag.  The device always has a
// chunk queued, and at most one chunk of no-op launches follows convergence.
// enqueue(k, check_only, hdone) queues iteration k; after maxit iterations it
// is called once more with check_only = 1.
template <class F>
int run_iterations(int64_t maxit, const PcgState* st, F&& enqueue, hipStream_t s) {
    int* dflag = nullptr;
    volatile int* hflag = host_done_flag(&dflag);
    if (hflag) *hflag = 0;
    const int64_t chunk = 8;
    hipEvent_t ev[2] = {nullptr, nullptr};
    for (auto& e : ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return ANIFEM_ERR_HIP;
    int rc = ANIFEM_OK;
    int64_t k = 0;
    for (int64_t c = 0; k < maxit; ++c) {
        const int64_t e = std::min(maxit, k + chunk);
        for (; k < e; ++k) enqueue(k, 0, dflag);
        if (hipEventRecord(ev[c & 1], s) != hipSuccess) {
            rc = ANIFEM_ERR_HIP;
            break;
        }
        if (c == 0) continue;
        if (hipEventSynchronize(ev[(c - 1) & 1]) != hipSuccess) {
            rc = ANIFEM_ERR_HIP;
            break;
        }
        int done = 0;
        if (hflag) {
            done = *hflag;
        } else {
            PcgState hs{};
            if (hipMemcpy(&hs, st, sizeof(hs), hipMemcpyDeviceToHost) != hipSuccess) {
                rc = ANIFEM_ERR_HIP;
                break;
            }
            done = hs.done;
        }
        if (done) break;
    }
    if (rc == ANIFEM_OK && k == maxit) enqueue(maxit, 1, dflag);
    for (auto e : ev) (void)hipEventDestroy(e);
    return rc;
}

// Device workspaces of the solvers, per host thread, device and slot: grown
// on demand and kept for the next call, so a solve allocates nothing once its
// sizes have been seen (anifem_pcg_release frees them).  Every solver returns
// after synchronizing its stream, so a workspace is idle between calls.
struct WorkEntry {
    int dev, slot;
    void* p;
    size_t bytes;
};
constexpr int kWorkEntries = 16;
thread_local WorkEntry tl_work[kWorkEntries];
thread_local int tl_work_used = 0;

int workspace(int slot, size_t bytes, void** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return ANIFEM_ERR_HIP;
    WorkEntry* e = nullptr;
    for (int i = 0; i < tl_work_used; ++i)
        if (tl_work[i].dev == dev && tl_work[i].slot == slot) e = &tl_work[i];
    if (!e) {
        if (tl_work_used == kWorkEntries) return ANIFEM_ERR_NOMEM;
        e = &tl_work[tl_work_used++];
        *e = WorkEntry{dev, slot, nullptr, 0};
    }
    if (e->bytes < bytes) {
        if (e->p) (void)hipFree(e->p);
        e->p = nullptr;
        e->bytes = 0;
        if (hipMalloc(&e->p, bytes) != hipSuccess) return ANIFEM_ERR_NOMEM;
        e->bytes = bytes;
    }
    *out = e->p;
    return ANIFEM_OK;
}

// Final report: the true residual ||b - A x|| from x, and the state's counters.
int finish(int64_t n, int64_t nnz, const int64_t* rp, const int32_t* ci, const double* v, const double* b,
           const double* x, double* Ax, double* part, const PcgState* st, anifem_pcg_result* res, hipStream_t s) {
    const unsigned vb = (unsigned)std::min<int64_t>((n + kThreads - 1) / kThreads, kRedBlocks);
    spmv_dispatch(n, nnz, rp, ci, v, x, Ax, nullptr, nullptr, s);
    hipLaunchKernelGGL(residual_kernel, dim3(vb), dim3(kThreads), 0, s, n, b, Ax, part);
    std::vector<double> hp(vb);
    PcgState hs{};
    if (hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(hp.data(), part, sizeof(double) * vb, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return ANIFEM_ERR_HIP;
    double tr = 0;
    for (unsigned i = 0; i < vb; ++i) tr += hp[i];
    if (res) {
        res->iterations = hs.it;
        res->relres = std::sqrt(hs.rr / hs.bb);
        res->true_relres = std::sqrt(tr / hs.bb);
    }
    return hs.done ? ANIFEM_OK : ANIFEM_ERR_NOT_CONVERGED;
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
    {
        const int wrc = workspace(0, sizeof(double) * (6 * nv + 3 * kRedBlocks) + sizeof(PcgState), (void**)&work);
        if (wrc) return wrc;
    }
    double *r = work, *z = r + nv, *pb[2] = {z + nv, z + 2 * nv}, *Ap = z + 3 * nv, *dinv = Ap + nv, *part = dinv + nv;
    PcgState* st = reinterpret_cast<PcgState*>(part + 3 * kRedBlocks);
    double *partA = part, *partU = part + kRedBlocks;  // p.Ap; (r.z, r.r)
    const unsigned vb = (unsigned)std::min<int64_t>((n + kThreads - 1) / kThreads, kRedBlocks);
    // r0 = b - A x0, z0 = Dinv r0
    hipLaunchKernelGGL(jacobi_inverse_kernel, dim3(vb), dim3(kThreads), 0, s, n, row_ptr, col_idx, val, dinv);
    spmv_dispatch(n, nnz, row_ptr, col_idx, val, x, Ap, nullptr, nullptr, s);
    hipLaunchKernelGGL(init_kernel, dim3(vb), dim3(kThreads), 0, s, n, b, Ap, dinv, r, z, pb[1], part);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kThreads), 0, s, part, (int)vb, 0, st, rtol);
    const int L = lanes_per_row(n, nnz);
    const bool split = nnz >= kSplitNnz;
    const int nD = resident_grid(split ? spmv_pdot_kernel_of(L) : dir_kernel_of(L), red_grid(n, L));
    int rc = run_iterations(
        maxit, st,
        [&](int64_t k, int check_only, int* hdone) {
            if (split) {
                hipLaunchKernelGGL(direction_kernel, dim3(vb), dim3(kThreads), 0, s, n, z, pb[(k + 1) & 1], pb[k & 1],
                                   partU, (int)vb, 0, st, k, check_only, hdone);
                if (!check_only) launch_spmv_pdot(L, nD, n, row_ptr, col_idx, val, pb[k & 1], Ap, partA, st, s);
            } else {
                launch_dir(L, nD, n, row_ptr, col_idx, val, z, pb[(k + 1) & 1], pb[k & 1], Ap, partA, partU, (int)vb,
                           0, st, k, check_only, hdone, s);
            }
            if (!check_only)
                hipLaunchKernelGGL(update_jacobi_kernel, dim3(vb), dim3(kThreads), 0, s, n, x, r, z, pb[k & 1], Ap,
                                   dinv, partA, nD, partU, st, k);
        },
        s);
    if (rc == ANIFEM_OK) rc = finish(n, nnz, row_ptr, col_idx, val, b, x, Ap, part, st, res, s);
    if (rc != ANIFEM_OK) (void)hipStreamSynchronize(s);  // the workspace is idle on return
    return rc;
}

// Flexible PCG (Polak-Ribiere beta) with an explicit SPD preconditioner
// matrix M (z = M r), for example M = G G^T of the learned factor.  Same
// stopping rule and reporting as anifem_pcg_jacobi; two launches per
// iteration.
extern "C" int anifem_pcg_matrix(int64_t n, const int64_t* row_ptr, const int32_t* col_idx, const double* val,
                                 const int64_t* m_row_ptr, const int32_t* m_col_idx, const double* m_val,
                                 const double* b, double* x, double rtol, int64_t maxit, anifem_pcg_result* res,
                                 void* stream) {
    using namespace anifem;
    if (n <= 0 || !row_ptr || !col_idx || !val || !m_row_ptr || !m_col_idx || !m_val || !b || !x || rtol <= 0 ||
        maxit < 0)
        return ANIFEM_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t nnz = 0, mnnz = 0;
    ANIFEM_HIP_CHECK(hipMemcpyAsync(&nnz, row_ptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    ANIFEM_HIP_CHECK(hipMemcpyAsync(&mnnz, m_row_ptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
    double* work = nullptr;
    const size_t nv = (size_t)n;
    {
        const int wrc = workspace(0, sizeof(double) * (6 * nv + 4 * kRedBlocks) + sizeof(PcgState), (void**)&work);
        if (wrc) return wrc;
    }
    double *rb[2] = {work, work + nv}, *z = work + 2 * nv, *pb[2] = {z + nv, z + 2 * nv}, *Ap = z + 3 * nv,
           *part = Ap + nv;
    PcgState* st = reinterpret_cast<PcgState*>(part + 4 * kRedBlocks);
    double *partA = part, *partU = part + kRedBlocks;  // p.Ap; (r.z, r.r, r.z_old)
    const unsigned vb = (unsigned)std::min<int64_t>((n + kThreads - 1) / kThreads, kRedBlocks);
    // r0 = b - A x0, z0 = M r0
    spmv_dispatch(n, nnz, row_ptr, col_idx, val, x, Ap, nullptr, nullptr, s);
    hipLaunchKernelGGL(init_f_kernel, dim3(vb), dim3(kThreads), 0, s, n, b, Ap, rb[0], part);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kThreads), 0, s, part, (int)vb, 3, st, rtol);
    const int pz = spmv_dispatch(n, mnnz, m_row_ptr, m_col_idx, m_val, rb[0], z, part, st, s);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kThreads), 0, s, part, pz, 5, st, rtol);
    const int L = lanes_per_row(n, nnz), nD = resident_grid(dir_kernel_of(L), red_grid(n, L));
    const int LM = lanes_per_row(n, mnnz), nU = resident_grid(update_matrix_kernel_of(LM), red_grid(n, LM));
    int rc = run_iterations(
        maxit, st,
        [&](int64_t k, int check_only, int* hdone) {
            launch_dir(L, nD, n, row_ptr, col_idx, val, z, pb[(k + 1) & 1], pb[k & 1], Ap, partA, partU, nU, 1, st,
                       k, check_only, hdone, s);
            if (!check_only)
                launch_update_matrix(LM, nU, n, m_row_ptr, m_col_idx, m_val, x, rb[k & 1], rb[(k + 1) & 1], z,
                                     pb[k & 1], Ap, partA, nD, partU, st, k, s);
        },
        s);
    if (rc == ANIFEM_OK) rc = finish(n, nnz, row_ptr, col_idx, val, b, x, Ap, part, st, res, s);
    if (rc != ANIFEM_OK) (void)hipStreamSynchronize(s);  // the workspace is idle on return
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
    {
        const int wrc = workspace(0, sizeof(double) * (6 * nv + 5 * kRedBlocks) + sizeof(PcgState), (void**)&work);
        if (wrc) return wrc;
    }
    double *r = work, *t = r + nv, *za = t + nv, *zb = za + nv, *p = zb + nv, *Ap = p + nv, *part = Ap + nv;
    PcgState* st = reinterpret_cast<PcgState*>(part + 5 * kRedBlocks);
    double* partB = part + 2 * kRedBlocks;  // update partials (r.r)
    double* partD = part + 3 * kRedBlocks;  // preconditioner partials (r.z, r.zo) at partD, partD + kRedBlocks
    const unsigned vb = (unsigned)std::min<int64_t>((n + kThreads - 1) / kThreads, kRedBlocks);
    int rc = ANIFEM_OK;
    // r0 = b - A x0, z0 = G G^T r0, p0 = z0
    spmv_dispatch(n, nnz, row_ptr, col_idx, val, x, Ap, nullptr, nullptr, s);
    hipLaunchKernelGGL(init_f_kernel, dim3(vb), dim3(kThreads), 0, s, n, b, Ap, r, part);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kThreads), 0, s, part, (int)vb, 3, st, rtol);
    spmv_dispatch(n, nnz, row_ptr, col_idx, GT, r, t, nullptr, st, s);
    int pb = prec_dispatch(n, nnz, row_ptr, col_idx, G, t, za, r, nullptr, part, st, s);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kThreads), 0, s, part, pb, 5, st, rtol);
    hipLaunchKernelGGL(copy_direction_kernel, dim3(vb), dim3(kThreads), 0, s, n, za, p, st);
    double* zz[2] = {za, zb};
    rc = run_iterations(
        maxit, st,
        [&](int64_t k, int check_only, int* hdone) {
            if (check_only) return;
            const int spmv_blocks = spmv_dispatch(n, nnz, row_ptr, col_idx, val, p, Ap, part, st, s);
            hipLaunchKernelGGL(update_fk_kernel, dim3(vb), dim3(kThreads), 0, s, n, x, r, p, Ap, part, spmv_blocks,
                               partB, st, k);
            gt_dispatch(n, nnz, row_ptr, col_idx, GT, r, t, partB, (int)vb, st, k, hdone, s);
            const int nd = prec_dispatch(n, nnz, row_ptr, col_idx, G, t, zz[(k + 1) & 1], r, zz[k & 1], partD, st, s);
            hipLaunchKernelGGL(direction_fk_kernel, dim3(vb), dim3(kThreads), 0, s, n, zz[(k + 1) & 1], p, partD, nd,
                               st, k);
        },
        s);
    if (rc == ANIFEM_OK) rc = finish(n, nnz, row_ptr, col_idx, val, b, x, Ap, part, st, res, s);
    if (rc != ANIFEM_OK) (void)hipStreamSynchronize(s);  // the workspace is idle on return
    return rc;
}

// Learned-preconditioned solve: the network's factor G for this matrix, then
// flexible PCG with M^-1 = G G^T; the FP64 true residual decides convergence
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
    rc = workspace(1, n * sizeof(double), (void**)&x0);
    if (rc) return rc;
    anifem_pcg_learned_result out{};
    if (hipMemcpyAsync(x0, x, n * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess) rc = ANIFEM_ERR_HIP;
    if (!rc) rc = anifem_gnn_factor(g, val, nullptr, nullptr, stream);
    // M = G G^T formed once, then two launches per iteration; without M (too
    // many terms) the factored form applies G^T and G per iteration
    int64_t mnnz = 0;
    const int64_t* mrp = nullptr;
    const int32_t* mci = nullptr;
    const double* mv = nullptr;
    const bool have_m = !rc && anifem_gnn_precond(g, &mnnz, &mrp, &mci, &mv, stream) == ANIFEM_OK;
    anifem_pcg_result r1{};
    int rc1 = rc ? rc
                 : have_m ? anifem_pcg_matrix(n, rp, ci, val, mrp, mci, mv, b, x, rtol, maxit, &r1, stream)
                          : anifem_pcg_factored(n, rp, ci, val, G, GT, b, x, rtol, maxit, &r1, stream);
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
    if (rc != ANIFEM_OK) (void)hipStreamSynchronize(s);
    if (res) *res = out;
    return rc;
}

extern "C" int anifem_pcg_release(void) {
    using namespace anifem;
    for (int i = 0; i < tl_work_used; ++i)
        if (tl_work[i].p) (void)hipFree(tl_work[i].p);
    tl_work_used = 0;
    return ANIFEM_OK;
}
