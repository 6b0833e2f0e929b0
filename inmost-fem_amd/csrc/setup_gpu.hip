// Device-side enumeration phase of the assembly setup: the same arrays as
// setup_enumerate_host (setup.cpp), bit for bit, built with radix sorts
// instead of host sorts and per-row merges.
//
//   * edges (and faces) by first appearance (the reference takes the ids from
//     INMOST; setup.cpp documents the rule): the (tet, local entity)
//     appearances are sorted by entity key with a stable radix sort, so each
//     run starts at the entity's first appearance; flagging those appearances
//     and an exclusive scan over the appearance order ranks the entities by
//     first appearance;
//   * NATURAL DOF ids per tet (global_enumerator.cpp:562-777), P3 edge pairs
//     by global node order (ordering.inl:137-151, tetdofmap.inl:98-105);
//   * boundary Dirichlet flags: the closure of the faces that belong to one
//     tet;
//   * CSR pattern and slot map (assembler.inl:114-136, 416): every element
//     entry (row, col) that the drop rule keeps, plus each row's diagonal,
//     sorted by row * ndof + col; run heads are the CSR slots in order, and
//     each kept entry's run is its slot.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "anifem_common.h"
#include "setup.h"

namespace anifem {

int nf_scalar(int fem);

namespace {

constexpr int kT = 256;

__constant__ int kEdgeD[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
__constant__ int kFaceD[4][3] = {{0, 1, 2}, {1, 2, 3}, {2, 3, 0}, {3, 0, 1}};

// Device buffers of one setup, freed together.
struct Pool {
    std::vector<void*> p;
    ~Pool() {
        for (void* x : p) (void)hipFree(x);
    }
    template <typename T>
    hipError_t get(T** out, size_t n) {
        void* x = nullptr;
        hipError_t e = hipMalloc(&x, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) p.push_back(x);
        *out = static_cast<T*>(x);
        return e;
    }
};

unsigned grid_of(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

__global__ void edge_keys_kernel(const int32_t* __restrict__ tets, int64_t m, uint64_t NN, uint64_t* keys,
                                 int32_t* vals) {
    const int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x;
    if (i >= m) return;
    const int64_t t = i / 6;
    const int k = (int)(i % 6);
    uint64_t a = (uint64_t)tets[4 * t + kEdgeD[k][0]], c = (uint64_t)tets[4 * t + kEdgeD[k][1]];
    if (a > c) {
        const uint64_t x = a;
        a = c;
        c = x;
    }
    keys[i] = a * NN + c;
    vals[i] = (int32_t)i;
}

// sorted node triple packed in 21-bit fields (nodes < 2^21)
__global__ void face_keys_kernel(const int32_t* __restrict__ tets, int64_t m, uint64_t* keys, int32_t* vals) {
    const int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x;
    if (i >= m) return;
    const int64_t t = i / 4;
    const int k = (int)(i % 4);
    uint64_t v0 = (uint64_t)tets[4 * t + kFaceD[k][0]], v1 = (uint64_t)tets[4 * t + kFaceD[k][1]],
             v2 = (uint64_t)tets[4 * t + kFaceD[k][2]];
    uint64_t x;
    if (v0 > v1) x = v0, v0 = v1, v1 = x;
    if (v1 > v2) x = v1, v1 = v2, v2 = x;
    if (v0 > v1) x = v0, v0 = v1, v1 = x;
    keys[i] = (v0 << 42) | (v1 << 21) | v2;
    vals[i] = (int32_t)i;
}

// run heads of sorted keys; keys equal to `dropped` start no run
__global__ void heads_kernel(const uint64_t* __restrict__ keys, int64_t m, uint64_t dropped, int32_t* head) {
    const int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x;
    if (i >= m) return;
    const uint64_t k = keys[i];
    head[i] = (k != dropped && (i == 0 || keys[i - 1] != k)) ? 1 : 0;
}

// at each run head: the run's first appearance (flagged in appearance order)
// and its start in the sorted order
__global__ void first_kernel(const int32_t* __restrict__ head, const int32_t* __restrict__ run,
                             const int32_t* __restrict__ vals, int64_t m, int32_t* flag, int32_t* first,
                             int32_t* start) {
    const int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x;
    if (i >= m || !head[i]) return;
    const int32_t u = run[i] - 1;
    flag[vals[i]] = 1;
    first[u] = vals[i];
    start[u] = (int32_t)i;
}

// entity id of every appearance: the rank of its run's first appearance
__global__ void ids_kernel(const int32_t* __restrict__ run, const int32_t* __restrict__ vals,
                           const int32_t* __restrict__ first, const int32_t* __restrict__ rankpos, int64_t m,
                           int32_t* ids) {
    const int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x;
    if (i >= m) return;
    ids[vals[i]] = rankpos[first[run[i] - 1]];
}

// appearances per entity, by entity id
__global__ void mult_kernel(const int32_t* __restrict__ start, const int32_t* __restrict__ first,
                            const int32_t* __restrict__ rankpos, int64_t nu, int64_t m, int32_t* mult) {
    const int64_t u = blockIdx.x * (int64_t)kT + threadIdx.x;
    if (u >= nu) return;
    const int32_t end = u + 1 < nu ? start[u + 1] : (int32_t)m;
    mult[rankpos[first[u]]] = end - start[u];
}

// P3 edge DOF pair order: the pair of local edge (I, J) keeps its local order
// when gid(I) < gid(J), which is what the reference's vertex ranking by global
// node id gives (ordering.inl:137-151, tetdofmap.inl:98-105).
__global__ void dofs_kernel(const int32_t* __restrict__ tets, int64_t NT, int64_t NN, const int32_t* __restrict__ eid,
                            const int32_t* __restrict__ fid, int64_t n_edges, int fem, int nf0, int ncomp,
                            int64_t nscal, int32_t* dofs) {
    const int64_t t = blockIdx.x * (int64_t)kT + threadIdx.x;
    if (t >= NT) return;
    const int32_t* v = tets + 4 * t;
    int64_t loc[20];
    for (int i = 0; i < 4; ++i) loc[i] = v[i];
    if (fem == ANIFEM_FEM_P2)
        for (int k = 0; k < 6; ++k) loc[4 + k] = NN + eid[6 * t + k];
    if (fem == ANIFEM_FEM_P3) {
        for (int k = 0; k < 6; ++k) {
            const int64_t base = NN + 2 * (int64_t)eid[6 * t + k];
            const bool keep = v[kEdgeD[k][0]] < v[kEdgeD[k][1]];
            loc[4 + 2 * k] = base + (keep ? 0 : 1);
            loc[5 + 2 * k] = base + (keep ? 1 : 0);
        }
        for (int k = 0; k < 4; ++k) loc[16 + k] = NN + 2 * n_edges + fid[4 * t + k];
    }
    const int nf = nf0 * ncomp;
    for (int c = 0; c < ncomp; ++c)
        for (int i = 0; i < nf0; ++i) dofs[nf * t + nf0 * c + i] = (int32_t)(c * nscal + loc[i]);
}

// closure of the boundary faces (faces of one tet): every writer stores 1
__global__ void boundary_kernel(const int32_t* __restrict__ tets, int64_t m, const int32_t* __restrict__ eid,
                                const int32_t* __restrict__ fid, const int32_t* __restrict__ fmult, uint8_t* snode,
                                uint8_t* sedge, uint8_t* sface) {
    const int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x;
    if (i >= m) return;
    const int64_t t = i / 4;
    const int k = (int)(i % 4);
    const int32_t f = fid[i];
    if (fmult[f] != 1) return;
    sface[f] = 1;
    for (int a = 0; a < 3; ++a) snode[tets[4 * t + kFaceD[k][a]]] = 1;
    for (int q = 0; q < 6; ++q) {
        bool in = false, in2 = false;
        for (int a = 0; a < 3; ++a) {
            in = in || kFaceD[k][a] == kEdgeD[q][0];
            in2 = in2 || kFaceD[k][a] == kEdgeD[q][1];
        }
        if (in && in2) sedge[eid[6 * t + q]] = 1;
    }
}

__global__ void dirichlet_expand_kernel(int64_t ndof, int64_t nscal, int64_t NN, int per_edge, int64_t n_edges,
                                        const uint8_t* __restrict__ snode, const uint8_t* __restrict__ sedge,
                                        const uint8_t* __restrict__ sface, uint8_t* is_dir) {
    const int64_t d = blockIdx.x * (int64_t)kT + threadIdx.x;
    if (d >= ndof) return;
    const int64_t i = d % nscal;
    uint8_t v;
    if (i < NN) v = snode[i];
    else if (i < NN + per_edge * n_edges) v = sedge[(i - NN) / per_edge];
    else v = sface[i - NN - per_edge * n_edges];
    is_dir[d] = v;
}

__global__ void mask_kernel(int64_t ndof, const uint8_t* __restrict__ mask, uint8_t* is_dir) {
    const int64_t d = blockIdx.x * (int64_t)kT + threadIdx.x;
    if (d < ndof) is_dir[d] = mask[d] ? 1 : 0;
}

// element entries (ib + nf*ia of tet t at i = t*E + ib + nf*ia) and one
// diagonal per row; dropped entries get the key ndof^2 (sorted last)
// (nscal_dec > 0: decoupled components, an entry between components
// r / nscal_dec != c / nscal_dec is dropped as well)
__global__ void entry_keys_kernel(const int32_t* __restrict__ dofs, const uint8_t* __restrict__ is_dir, int64_t m_el,
                                  int nf, int64_t ndof, int64_t nscal_dec, uint64_t* keys, int32_t* vals) {
    const int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x;
    const int64_t E = (int64_t)nf * nf;
    if (i < m_el) {
        const int64_t t = i / E;
        const int e = (int)(i % E);
        const int ib = e % nf, ia = e / nf;
        const int64_t r = dofs[nf * t + ib], c = dofs[nf * t + ia];
        const bool keep = r == c || (!(is_dir[r] || is_dir[c]) && (nscal_dec == 0 || r / nscal_dec == c / nscal_dec));
        keys[i] = keep ? (uint64_t)r * (uint64_t)ndof + (uint64_t)c : (uint64_t)ndof * (uint64_t)ndof;
        vals[i] = (int32_t)i;
    } else if (i < m_el + ndof) {
        const int64_t d = i - m_el;
        keys[i] = (uint64_t)d * (uint64_t)ndof + (uint64_t)d;
        vals[i] = -1;
    }
}

// CSR from the sorted entries: slot u = run - 1; the first slot of each row
// starts the row (every row holds its diagonal)
__global__ void csr_kernel(const uint64_t* __restrict__ keys, const int32_t* __restrict__ vals,
                           const int32_t* __restrict__ head, const int32_t* __restrict__ run, int64_t m, int64_t ndof,
                           int32_t* col_idx, int64_t* row_ptr, int32_t* slot) {
    const int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x;
    if (i >= m) return;
    const uint64_t k = keys[i];
    if (k == (uint64_t)ndof * (uint64_t)ndof) return;  // dropped (slot stays -1)
    const int32_t u = run[i] - 1;
    if (vals[i] >= 0) slot[vals[i]] = u;
    if (!head[i]) return;
    const int64_t r = (int64_t)(k / (uint64_t)ndof);
    col_idx[u] = (int32_t)(k % (uint64_t)ndof);
    if (i == 0 || keys[i - 1] / (uint64_t)ndof != (uint64_t)r) row_ptr[r] = u;
}

__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
    const int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x;
    if (i < n) p[i] = v;
}

int bits_for(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b) != 0) ++b;
    return b > 0 ? b : 1;
}

// Sort (key, appearance) pairs, then number the distinct keys by first
// appearance: ids[appearance] (m entries); mult by id when asked.
int number_entities(Pool& P, uint64_t* keys, int32_t* vals, int64_t m, int bits, int32_t* ids, int32_t** mult,
                    int64_t* n_out, hipStream_t s) {
    uint64_t* k2;
    int32_t *v2, *head, *run, *flag, *rankpos, *first, *start;
    ANIFEM_HIP_CHECK(P.get(&k2, m));
    ANIFEM_HIP_CHECK(P.get(&v2, m));
    size_t tb = 0;
    ANIFEM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, k2, vals, v2, (int)m, 0, bits, s));
    void* tmp;
    ANIFEM_HIP_CHECK(P.get(reinterpret_cast<char**>(&tmp), tb));
    ANIFEM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, k2, vals, v2, (int)m, 0, bits, s));
    ANIFEM_HIP_CHECK(P.get(&head, m));
    ANIFEM_HIP_CHECK(P.get(&run, m));
    hipLaunchKernelGGL(heads_kernel, dim3(grid_of(m)), dim3(kT), 0, s, k2, m, ~0ull, head);
    size_t sb = 0;
    ANIFEM_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, sb, head, run, (int)m, s));
    void* stmp;
    ANIFEM_HIP_CHECK(P.get(reinterpret_cast<char**>(&stmp), sb));
    ANIFEM_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(stmp, sb, head, run, (int)m, s));
    int32_t nu = 0;
    ANIFEM_HIP_CHECK(hipMemcpyAsync(&nu, run + (m - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
    ANIFEM_HIP_CHECK(P.get(&flag, m));
    ANIFEM_HIP_CHECK(P.get(&rankpos, m));
    ANIFEM_HIP_CHECK(P.get(&first, nu));
    ANIFEM_HIP_CHECK(P.get(&start, nu));
    hipLaunchKernelGGL(fill_i32_kernel, dim3(grid_of(m)), dim3(kT), 0, s, flag, m, 0);
    hipLaunchKernelGGL(first_kernel, dim3(grid_of(m)), dim3(kT), 0, s, head, run, v2, m, flag, first, start);
    size_t eb = 0;
    ANIFEM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, eb, flag, rankpos, (int)m, s));
    void* etmp;
    ANIFEM_HIP_CHECK(P.get(reinterpret_cast<char**>(&etmp), eb));
    ANIFEM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(etmp, eb, flag, rankpos, (int)m, s));
    hipLaunchKernelGGL(ids_kernel, dim3(grid_of(m)), dim3(kT), 0, s, run, v2, first, rankpos, m, ids);
    if (mult) {
        ANIFEM_HIP_CHECK(P.get(mult, nu));
        hipLaunchKernelGGL(mult_kernel, dim3(grid_of(nu)), dim3(kT), 0, s, start, first, rankpos, (int64_t)nu, m,
                           *mult);
    }
    ANIFEM_HIP_CHECK(hipGetLastError());
    *n_out = nu;
    return ANIFEM_OK;
}

}  // namespace

// Returns ANIFEM_OK, an error, or 1 when the mesh exceeds the device path's
// index ranges (the caller then runs setup_enumerate_host).
int setup_enumerate_device(const anifem_problem_desc& prob, AssemblySetup& S, void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int fem = S.fem, nf0 = nf_scalar(fem), nf = S.nf;
    const int64_t NT = S.n_tets, NN = S.n_nodes, E = S.E;
    if (S.n_var != 1 || fem < ANIFEM_FEM_P1 || fem > ANIFEM_FEM_P3) return 1;  // the other spaces: host phase
    const bool need_edges = fem >= ANIFEM_FEM_P2 || prob.dirichlet == ANIFEM_DIRICHLET_BOUNDARY;
    const bool need_faces = fem >= ANIFEM_FEM_P3 || prob.dirichlet == ANIFEM_DIRICHLET_BOUNDARY;
    if (need_faces && NN >= (int64_t(1) << 21)) return 1;
    if (NT * E + 4 * NT >= INT32_MAX) return 1;
    if (prob.dirichlet != ANIFEM_DIRICHLET_NONE && prob.dirichlet != ANIFEM_DIRICHLET_BOUNDARY &&
        prob.dirichlet != ANIFEM_DIRICHLET_DOF_MASK)
        return ANIFEM_ERR_ARG;
    if (prob.dirichlet == ANIFEM_DIRICHLET_DOF_MASK && !prob.dirichlet_dof_mask) return ANIFEM_ERR_ARG;
    Pool P;
    int rc;
    int32_t* tets;
    ANIFEM_HIP_CHECK(P.get(&tets, 4 * NT));
    ANIFEM_HIP_CHECK(hipMemcpyAsync(tets, S.tets.data(), 4 * NT * sizeof(int32_t), hipMemcpyHostToDevice, s));

    int32_t *eid = nullptr, *fid = nullptr, *fmult = nullptr;
    int64_t n_edges = 0, n_faces = 0;
    if (need_edges) {
        const int64_t m = 6 * NT;
        uint64_t* keys;
        int32_t* vals;
        ANIFEM_HIP_CHECK(P.get(&keys, m));
        ANIFEM_HIP_CHECK(P.get(&vals, m));
        ANIFEM_HIP_CHECK(P.get(&eid, m));
        hipLaunchKernelGGL(edge_keys_kernel, dim3(grid_of(m)), dim3(kT), 0, s, tets, m, (uint64_t)NN, keys, vals);
        if ((rc = number_entities(P, keys, vals, m, bits_for((uint64_t)NN * (uint64_t)NN), eid, nullptr, &n_edges, s)))
            return rc;
    }
    if (need_faces) {
        const int64_t m = 4 * NT;
        uint64_t* keys;
        int32_t* vals;
        ANIFEM_HIP_CHECK(P.get(&keys, m));
        ANIFEM_HIP_CHECK(P.get(&vals, m));
        ANIFEM_HIP_CHECK(P.get(&fid, m));
        hipLaunchKernelGGL(face_keys_kernel, dim3(grid_of(m)), dim3(kT), 0, s, tets, m, keys, vals);
        if ((rc = number_entities(P, keys, vals, m, 63, fid, &fmult, &n_faces, s))) return rc;
    }
    S.n_edges = n_edges;
    S.n_faces = n_faces;

    // NATURAL enumeration
    const int per_edge = fem == ANIFEM_FEM_P2 ? 1 : (fem == ANIFEM_FEM_P3 ? 2 : 0);
    const int per_face = fem == ANIFEM_FEM_P3 ? 1 : 0;
    const int64_t nscal = NN + per_edge * n_edges + per_face * n_faces;
    S.ndof = nscal * S.ncomp;
    if (S.ndof >= INT32_MAX) return ANIFEM_ERR_ARG;
    const int64_t ndof = S.ndof;
    if (NT * E + ndof >= INT32_MAX) return 1;  // entries are sorted with 32-bit counts
    int32_t* dofs;
    ANIFEM_HIP_CHECK(P.get(&dofs, nf * NT));
    hipLaunchKernelGGL(dofs_kernel, dim3(grid_of(NT)), dim3(kT), 0, s, tets, NT, NN, eid, fid, n_edges, fem, nf0,
                       S.ncomp, nscal, dofs);

    // Dirichlet flags
    uint8_t* is_dir;
    ANIFEM_HIP_CHECK(P.get(&is_dir, ndof));
    ANIFEM_HIP_CHECK(hipMemsetAsync(is_dir, 0, ndof, s));
    if (prob.dirichlet == ANIFEM_DIRICHLET_DOF_MASK) {
        uint8_t* mask;
        ANIFEM_HIP_CHECK(P.get(&mask, ndof));
        ANIFEM_HIP_CHECK(hipMemcpyAsync(mask, prob.dirichlet_dof_mask, ndof, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(mask_kernel, dim3(grid_of(ndof)), dim3(kT), 0, s, ndof, mask, is_dir);
    } else if (prob.dirichlet == ANIFEM_DIRICHLET_BOUNDARY) {
        uint8_t *snode, *sedge, *sface;
        ANIFEM_HIP_CHECK(P.get(&snode, NN));
        ANIFEM_HIP_CHECK(P.get(&sedge, n_edges));
        ANIFEM_HIP_CHECK(P.get(&sface, n_faces));
        ANIFEM_HIP_CHECK(hipMemsetAsync(snode, 0, NN, s));
        ANIFEM_HIP_CHECK(hipMemsetAsync(sedge, 0, std::max<int64_t>(n_edges, 1), s));
        ANIFEM_HIP_CHECK(hipMemsetAsync(sface, 0, std::max<int64_t>(n_faces, 1), s));
        hipLaunchKernelGGL(boundary_kernel, dim3(grid_of(4 * NT)), dim3(kT), 0, s, tets, 4 * NT, eid, fid, fmult,
                           snode, sedge, sface);
        hipLaunchKernelGGL(dirichlet_expand_kernel, dim3(grid_of(ndof)), dim3(kT), 0, s, ndof, nscal, NN, per_edge,
                           n_edges, snode, sedge, sface, is_dir);
    }
    ANIFEM_HIP_CHECK(hipGetLastError());

    // CSR pattern and slot map
    const int64_t m_el = NT * E, m = m_el + ndof;
    uint64_t *keys, *k2;
    int32_t *vals, *v2, *head, *run, *col_idx, *slot;
    int64_t* row_ptr;
    ANIFEM_HIP_CHECK(P.get(&keys, m));
    ANIFEM_HIP_CHECK(P.get(&vals, m));
    const int64_t nscal_dec = S.decoupled ? ndof / S.ncomp : 0;
    hipLaunchKernelGGL(entry_keys_kernel, dim3(grid_of(m)), dim3(kT), 0, s, dofs, is_dir, m_el, nf, ndof, nscal_dec,
                       keys, vals);
    ANIFEM_HIP_CHECK(P.get(&k2, m));
    ANIFEM_HIP_CHECK(P.get(&v2, m));
    const uint64_t dropped = (uint64_t)ndof * (uint64_t)ndof;
    size_t tb = 0;
    const int bits = bits_for(dropped);
    ANIFEM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, k2, vals, v2, (int)m, 0, bits, s));
    void* tmp;
    ANIFEM_HIP_CHECK(P.get(reinterpret_cast<char**>(&tmp), tb));
    ANIFEM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, k2, vals, v2, (int)m, 0, bits, s));
    ANIFEM_HIP_CHECK(P.get(&head, m));
    ANIFEM_HIP_CHECK(P.get(&run, m));
    hipLaunchKernelGGL(heads_kernel, dim3(grid_of(m)), dim3(kT), 0, s, k2, m, dropped, head);
    size_t sb = 0;
    ANIFEM_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, sb, head, run, (int)m, s));
    void* stmp;
    ANIFEM_HIP_CHECK(P.get(reinterpret_cast<char**>(&stmp), sb));
    ANIFEM_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(stmp, sb, head, run, (int)m, s));
    int32_t nnz = 0;
    ANIFEM_HIP_CHECK(hipMemcpyAsync(&nnz, run + (m - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
    ANIFEM_HIP_CHECK(P.get(&col_idx, nnz));
    ANIFEM_HIP_CHECK(P.get(&row_ptr, ndof + 1));
    ANIFEM_HIP_CHECK(P.get(&slot, m_el));
    hipLaunchKernelGGL(fill_i32_kernel, dim3(grid_of(m_el)), dim3(kT), 0, s, slot, m_el, -1);
    hipLaunchKernelGGL(csr_kernel, dim3(grid_of(m)), dim3(kT), 0, s, k2, v2, head, run, m, ndof, col_idx, row_ptr,
                       slot);
    ANIFEM_HIP_CHECK(hipGetLastError());
    const int64_t nnz64 = nnz;
    ANIFEM_HIP_CHECK(hipMemcpyAsync(row_ptr + ndof, &nnz64, sizeof(int64_t), hipMemcpyHostToDevice, s));

    S.nnz = nnz;
    S.dofs.resize((size_t)nf * NT);
    S.is_dir.resize(ndof);
    S.row_ptr.resize(ndof + 1);
    S.col_idx.resize(nnz);
    S.slot.resize((size_t)m_el);
    ANIFEM_HIP_CHECK(hipMemcpyAsync(S.dofs.data(), dofs, S.dofs.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    ANIFEM_HIP_CHECK(hipMemcpyAsync(S.is_dir.data(), is_dir, ndof, hipMemcpyDeviceToHost, s));
    ANIFEM_HIP_CHECK(hipMemcpyAsync(S.row_ptr.data(), row_ptr, (ndof + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    ANIFEM_HIP_CHECK(hipMemcpyAsync(S.col_idx.data(), col_idx, (size_t)nnz * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    ANIFEM_HIP_CHECK(hipMemcpyAsync(S.slot.data(), slot, (size_t)m_el * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    ANIFEM_HIP_CHECK(hipStreamSynchronize(s));
    return ANIFEM_OK;
}

}  // namespace anifem

namespace anifem {

int build_setup_device(const anifem_mesh_desc& mesh, const anifem_problem_desc& prob, AssemblySetup& S,
                       const BlockCaps& caps, void* stream) {
    int rc = setup_prepare(mesh, prob, S);
    if (rc) return rc;
    rc = setup_enumerate_device(prob, S, stream);
    if (rc == 1) rc = setup_enumerate_host(prob, S);  // beyond the device path's index ranges
    if (rc) return rc;
    std::vector<int64_t> aptr;
    std::vector<int32_t> adj;
    dof_adjacency(S, aptr, adj);
    return build_blocks(S, aptr, adj, caps);
}

}  // namespace anifem
