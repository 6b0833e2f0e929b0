// Host-side assembly setup (see setup.h).
//
// Follows the reference's assembly-side semantics:
//   * local node reorder: swap nodes 2 and 3 when det(P0-P3, P1-P3, P2-P3) < 0
//     (inmost_interface/ordering.inl:8-26; on by default, assembler.h:298);
//   * NATURAL global enumeration: (component, geometric type {node, edge,
//     face}, element id, dof on element) (global_enumerator.h:392-404,
//     global_enumerator.cpp:562-777);
//   * P3 edge DOF pairs ordered by global node id: the pair of local edge
//     (I, J) keeps its local order when gid(I) < gid(J), the reference's
//     vertex-rank rule (ordering.inl:137-151, tetdofmap.inl:98-105, applied
//     in fill_assemble_templates, assembler.inl:140-184);
//   * CSR pattern: every DOF pair sharing a tet plus the diagonal
//     (assembler.inl:114-136), minus the off-diagonals of Dirichlet rows and
//     columns, whose element entries applyDir makes exactly zero and the drop
//     rule |a| <= drop_val never inserts (assembler.inl:416, dc_on_dof.h:38-45).
// Edge and face ids are assigned by first appearance in the caller's cell
// order over the local edges 01,02,03,12,13,23 and local faces
// {k, k+1, k+2 mod 4}; the reference takes them from INMOST (absent here).
#include "setup.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <functional>
#include <thread>

namespace anifem {

int nf_scalar(int fem);
int quadrature(int order, const double** p, const double** w);

void host_parallel_for(int64_t n, int64_t grain, const std::function<void(int64_t, int64_t)>& fn) {
    unsigned hw = std::thread::hardware_concurrency();
    int64_t nt = std::max<int64_t>(1, std::min<int64_t>(hw ? hw : 1, 32));
    if (n <= grain || nt == 1) {
        fn(0, n);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> th;
    for (int64_t t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (;;) {
                int64_t b = next.fetch_add(grain);
                if (b >= n) break;
                fn(b, std::min(n, b + grain));
            }
        });
    for (auto& x : th) x.join();
}

namespace {

const int kEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
const int kFace[4][3] = {{0, 1, 2}, {1, 2, 3}, {2, 3, 0}, {3, 0, 1}};

// Ids by first appearance.  keys[t*ne + e] identifies the entity; returns the
// number of distinct entities and fills ids (same indexing as keys).
template <typename Key>
int64_t number_by_first_appearance(const std::vector<Key>& keys, std::vector<int64_t>& ids,
                                   std::vector<int32_t>* mult = nullptr) {
    const int64_t m = (int64_t)keys.size();
    std::vector<int64_t> order(m);
    for (int64_t i = 0; i < m; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return keys[a] < keys[b] || (keys[a] == keys[b] && a < b);
    });
    // first occurrence of each distinct key, in key order
    std::vector<int64_t> first;
    std::vector<int64_t> group(m);
    for (int64_t i = 0; i < m; ++i) {
        if (i == 0 || keys[order[i]] != keys[order[i - 1]]) first.push_back(order[i]);
        group[order[i]] = (int64_t)first.size() - 1;
    }
    const int64_t nu = (int64_t)first.size();
    std::vector<int64_t> rank(nu), by_first(nu);
    for (int64_t u = 0; u < nu; ++u) by_first[u] = u;
    std::sort(by_first.begin(), by_first.end(), [&](int64_t a, int64_t b) { return first[a] < first[b]; });
    for (int64_t r = 0; r < nu; ++r) rank[by_first[r]] = r;
    ids.resize(m);
    for (int64_t i = 0; i < m; ++i) ids[i] = rank[group[i]];
    if (mult) {
        mult->assign(nu, 0);
        for (int64_t i = 0; i < m; ++i) (*mult)[ids[i]]++;
    }
    return nu;
}

// P3 edge DOF pair order.  The reference ranks the four vertices of a cell by
// global node id (createOrderPermutation, ordering.inl:137-151) and keeps the
// pair of local edge (I, J) in local order when rank(I) < rank(J)
// (index_on_reorderd_elem, tetdofmap.inl:98-105).  For distinct ids the ranks
// compare as the ids do, so the pair keeps its order exactly when the edge's
// first local vertex carries the smaller global node id.
inline bool edge_pair_kept(const int32_t* v, int k) { return v[kEdge[k][0]] < v[kEdge[k][1]]; }

}  // namespace

SpaceLayout space_layout(int fem) {
    SpaceLayout L;
    L.valid = true;
    switch (fem) {
        case ANIFEM_FEM_P0: L.cell = 1; break;                           // poly_0.h:86-87
        case ANIFEM_FEM_P1: L.node = 1; break;                           // poly_1.h:135-137
        case ANIFEM_FEM_P2: L.node = 1, L.edge = 1; break;               // poly_2.h:104-107
        case ANIFEM_FEM_P3: L.node = 1, L.edge = 2, L.face = 1; break;   // poly_3.h:122-125
        case ANIFEM_FEM_RT0: L.face = 1, L.face_oriented = true; break;  // raviant_thomas_0.h:202-203
        case ANIFEM_FEM_ND0: L.edge = 1, L.edge_oriented = true; break;  // nedelec_0.h:186-187
        case ANIFEM_FEM_CR1: L.face = 1; break;                          // crouzeix_raviart_1.h:90-91
        case ANIFEM_FEM_B2: L.edge = 1; break;                           // bubble_2.h:97-98
        case ANIFEM_FEM_B4: L.cell = 1; break;                           // bubble_4.h:90-91
        default: L.valid = false;
    }
    return L;
}

namespace {
// Orientation sign of a cell's local edge k or face k for its global basis
// function (global = sign x local, assembler.inl:154-164).
//  * Edge (I, J): positive when gid(I) < gid(J), the reference's rule
//    (isPositivePermutationOrient, ordering.inl:152-158): the global ND0
//    function then has unit tangential component from the lower to the
//    higher global id in every cell.
//  * Face k = (k, k+1, k+2): the local RT0 function points out of the cell.
//    In a reordered cell (det(P0-P3, P1-P3, P2-P3) >= 0) the vertex order
//    k, k+1, k+2 turns counter-clockwise seen from outside for even k and
//    clockwise for odd k, so the sign (-1)^k x (parity of the face's global
//    ids in that order) gives every cell the same normal, the right-hand
//    normal of the face's vertices in ascending global id.  (The reference's
//    face test, ordering.inl:159-164, compares the third face vertex with the
//    opposite vertex; read as written it does not give a normal-continuous
//    global space, and this build uses the sign that does.)
inline int8_t edge_sign(const int32_t* v, int k) { return v[kEdge[k][0]] < v[kEdge[k][1]] ? 1 : -1; }
inline int8_t face_sign(const int32_t* v, int k) {
    const int32_t a = v[k], b = v[(k + 1) % 4], c = v[(k + 2) % 4];
    const int inversions = (a > b) + (a > c) + (b > c);
    return ((k + inversions) & 1) ? -1 : 1;
}
}  // namespace

namespace {
// The variables' spaces and sizes; every space is one of space_layout's,
// vector variables FemVec<ncomp, P1..P3> only, at most 64 local DOFs in all.
int set_vars(int n_var, const anifem_op* vars, AssemblySetup& S) {
    if (n_var < 1 || n_var > kMaxVars) return ANIFEM_ERR_ARG;
    S.n_var = n_var;
    S.nf = 0;
    for (int v = 0; v < n_var; ++v) {
        const int fem = vars[v].fem, nc = vars[v].ncomp;
        if (!space_layout(fem).valid || nc < 1) return ANIFEM_ERR_ARG;
        if (nc != 1 && (fem < ANIFEM_FEM_P1 || fem > ANIFEM_FEM_P3)) return ANIFEM_ERR_ARG;
        S.var_fem[v] = fem;
        S.var_ncomp[v] = nc;
        S.nf += nf_scalar(fem) * nc;
        if (S.nf > 64) return ANIFEM_ERR_ARG;  // Dirichlet masks are 64-bit
    }
    S.fem = vars[0].fem;
    S.ncomp = vars[0].ncomp;
    S.E = S.nf * S.nf;
    return ANIFEM_OK;
}

int reorder_nodes(const anifem_mesh_desc& mesh, AssemblySetup& S);
}  // namespace

int setup_prepare(const anifem_mesh_desc& mesh, const anifem_problem_desc& prob, AssemblySetup& S) {
    if (prob.test.fem != prob.trial.fem || prob.test.ncomp != prob.trial.ncomp) return ANIFEM_ERR_ARG;
    int rc = set_vars(1, &prob.trial, S);
    auto componentwise = [](int op) { return op == ANIFEM_IDEN || op == ANIFEM_GRAD || (op >= ANIFEM_DUDX && op <= ANIFEM_DUDZ); };
    S.decoupled = prob.trial.ncomp > 1 && componentwise(prob.trial.op) && componentwise(prob.test.op) &&
                  (prob.tensor_type == ANIFEM_TENSOR_NULL || prob.tensor_type == ANIFEM_TENSOR_SCALAR);
    return rc ? rc : reorder_nodes(mesh, S);
}

int setup_prepare_vars(const anifem_mesh_desc& mesh, int n_var, const anifem_op* vars, AssemblySetup& S) {
    if (!vars) return ANIFEM_ERR_ARG;
    int rc = set_vars(n_var, vars, S);
    return rc ? rc : reorder_nodes(mesh, S);
}

namespace {
int reorder_nodes(const anifem_mesh_desc& mesh, AssemblySetup& S) {
    S.n_nodes = mesh.n_nodes;
    S.n_tets = mesh.n_tets;
    const int64_t NT = S.n_tets, NN = S.n_nodes;
    if (NT <= 0 || NN <= 0 || NT > INT32_MAX / 64) return ANIFEM_ERR_ARG;

    // node reorder (ordering.inl:8-26) and connectivity check
    S.tets.assign(mesh.tets, mesh.tets + 4 * NT);
    std::atomic<int64_t> bad{0}, nonpos{0};
    const double* X = mesh.coords;
    host_parallel_for(NT, 1 << 14, [&](int64_t b, int64_t e) {
        int64_t lbad = 0, lnp = 0;
        for (int64_t t = b; t < e; ++t) {
            int32_t* v = &S.tets[4 * t];
            bool ok = true;
            for (int k = 0; k < 4; ++k) ok = ok && v[k] >= 0 && v[k] < NN;
            for (int k = 0; k < 4 && ok; ++k)
                for (int l = k + 1; l < 4; ++l) ok = ok && v[k] != v[l];
            if (!ok) {
                ++lbad;
                continue;
            }
            double m[3][3];
            for (int j = 0; j < 3; ++j) {
                m[0][j] = X[3 * v[0] + j] - X[3 * v[3] + j];
                m[1][j] = X[3 * v[1] + j] - X[3 * v[3] + j];
                m[2][j] = X[3 * v[2] + j] - X[3 * v[3] + j];
            }
            double det = 0;
            det += m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]);
            det += m[0][1] * (m[1][2] * m[2][0] - m[1][0] * m[2][2]);
            det += m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
            if (det < 0) std::swap(v[2], v[3]);
            if (det == 0 || !std::isfinite(det)) ++lnp;
        }
        bad += lbad;
        nonpos += lnp;
    });
    if (bad.load()) return ANIFEM_ERR_CONNECTIVITY;
    S.n_degenerate = nonpos.load();
    return ANIFEM_OK;
}
}  // namespace

int setup_enumerate_host(const anifem_problem_desc& prob, AssemblySetup& S) {
    const int64_t NT = S.n_tets, NN = S.n_nodes;
    SpaceLayout L[kMaxVars];
    bool any_edge = false, any_face = false;
    for (int v = 0; v < S.n_var; ++v) {
        L[v] = space_layout(S.var_fem[v]);
        any_edge = any_edge || L[v].edge > 0;
        any_face = any_face || L[v].face > 0;
    }

    // edges and faces by first appearance
    std::vector<int64_t> eid, fid;
    std::vector<int32_t> fmult;
    const bool need_edges = any_edge || prob.dirichlet == ANIFEM_DIRICHLET_BOUNDARY;
    const bool need_faces = any_face || prob.dirichlet == ANIFEM_DIRICHLET_BOUNDARY;
    if (need_edges) {
        std::vector<uint64_t> keys(6 * NT);
        host_parallel_for(NT, 1 << 14, [&](int64_t b, int64_t e) {
            for (int64_t t = b; t < e; ++t)
                for (int k = 0; k < 6; ++k) {
                    uint64_t a = (uint64_t)S.tets[4 * t + kEdge[k][0]], c = (uint64_t)S.tets[4 * t + kEdge[k][1]];
                    if (a > c) std::swap(a, c);
                    keys[6 * t + k] = a * (uint64_t)NN + c;
                }
        });
        S.n_edges = number_by_first_appearance(keys, eid);
    }
    if (need_faces) {
        std::vector<unsigned __int128> keys(4 * NT);
        host_parallel_for(NT, 1 << 14, [&](int64_t b, int64_t e) {
            for (int64_t t = b; t < e; ++t)
                for (int k = 0; k < 4; ++k) {
                    uint64_t v[3];
                    for (int i = 0; i < 3; ++i) v[i] = (uint64_t)S.tets[4 * t + kFace[k][i]];
                    std::sort(v, v + 3);
                    keys[4 * t + k] = ((unsigned __int128)v[0] << 64) | ((unsigned __int128)v[1] << 32) | v[2];
                }
        });
        S.n_faces = number_by_first_appearance(keys, fid, &fmult);
    }

    // NATURAL enumeration (global_enumerator.h:392-404): per variable, per
    // component, nodes, edges, faces, cells; variable v starts at var_base[v]
    struct VarNum {
        int64_t edge_base, face_base, cell_base, nscal, base;
        int nf0, lf0;  // scalar local DOFs, first local DOF of the variable
    };
    VarNum V[kMaxVars];
    int64_t total = 0;
    int lf = 0;
    for (int v = 0; v < S.n_var; ++v) {
        VarNum& n = V[v];
        n.edge_base = NN * L[v].node;
        n.face_base = n.edge_base + (int64_t)L[v].edge * S.n_edges;
        n.cell_base = n.face_base + (int64_t)L[v].face * S.n_faces;
        n.nscal = n.cell_base + (int64_t)L[v].cell * NT;
        n.base = total;
        n.nf0 = nf_scalar(S.var_fem[v]);
        n.lf0 = lf;
        total += n.nscal * S.var_ncomp[v];
        lf += n.nf0 * S.var_ncomp[v];
    }
    S.ndof = total;
    if (S.ndof >= INT32_MAX) return ANIFEM_ERR_ARG;
    S.dofs.resize((size_t)S.nf * NT);
    bool oriented = false;
    for (int v = 0; v < S.n_var; ++v) oriented = oriented || L[v].edge_oriented || L[v].face_oriented;
    if (oriented) S.sign.assign((size_t)S.nf * NT, 1);
    else std::vector<int8_t>().swap(S.sign);
    host_parallel_for(NT, 1 << 14, [&](int64_t b, int64_t e) {
        for (int64_t t = b; t < e; ++t) {
            const int32_t* vt = &S.tets[4 * t];
            for (int var = 0; var < S.n_var; ++var) {
                const SpaceLayout& Lv = L[var];
                const VarNum& n = V[var];
                const int64_t per_edge = Lv.edge, per_face = Lv.face;
                int64_t loc[20];
                int8_t sg[20];
                int i = 0;
                for (int k = 0; k < 4 && Lv.node; ++k, ++i) loc[i] = vt[k], sg[i] = 1;
                for (int k = 0; k < 6 && per_edge; ++k) {
                    const int64_t base = n.edge_base + per_edge * eid[(size_t)6 * t + k];
                    if (per_edge == 1) {
                        loc[i] = base;
                        sg[i++] = Lv.edge_oriented ? edge_sign(vt, k) : 1;
                    } else if (per_edge == 2) {
                        const bool keep = edge_pair_kept(vt, k);
                        loc[i] = base + (keep ? 0 : 1), sg[i++] = 1;
                        loc[i] = base + (keep ? 1 : 0), sg[i++] = 1;
                    }
                }
                for (int k = 0; k < 4 && per_face; ++k, ++i) {
                    loc[i] = n.face_base + fid[(size_t)4 * t + k];
                    sg[i] = Lv.face_oriented ? face_sign(vt, k) : 1;
                }
                if (Lv.cell) loc[i] = n.cell_base + t, sg[i++] = 1;
                for (int c = 0; c < S.var_ncomp[var]; ++c)
                    for (int j = 0; j < n.nf0; ++j) {
                        const size_t at = (size_t)S.nf * t + n.lf0 + n.nf0 * c + j;
                        S.dofs[at] = (int32_t)(n.base + c * n.nscal + loc[j]);
                        if (oriented) S.sign[at] = sg[j];
                    }
            }
        }
    });

    // Dirichlet flags
    S.is_dir.assign(S.ndof, 0);
    if (prob.dirichlet == ANIFEM_DIRICHLET_DOF_MASK) {
        if (!prob.dirichlet_dof_mask) return ANIFEM_ERR_ARG;
        for (int64_t d = 0; d < S.ndof; ++d) S.is_dir[d] = prob.dirichlet_dof_mask[d] ? 1 : 0;
    } else if (prob.dirichlet == ANIFEM_DIRICHLET_BOUNDARY) {
        // boundary faces belong to exactly one tet; their closure is Dirichlet
        std::vector<uint8_t> snode(NN, 0), sedge(S.n_edges, 0), sface(S.n_faces, 0);
        for (int64_t t = 0; t < NT; ++t)
            for (int k = 0; k < 4; ++k) {
                if (fmult[fid[4 * t + k]] != 1) continue;
                sface[fid[4 * t + k]] = 1;
                for (int i = 0; i < 3; ++i) snode[S.tets[4 * t + kFace[k][i]]] = 1;
                for (int q = 0; q < 6; ++q) {
                    int a = kEdge[q][0], c = kEdge[q][1];
                    bool in = false, in2 = false;
                    for (int i = 0; i < 3; ++i) {
                        in = in || kFace[k][i] == a;
                        in2 = in2 || kFace[k][i] == c;
                    }
                    if (in && in2) sedge[eid[6 * t + q]] = 1;
                }
            }
        for (int var = 0; var < S.n_var; ++var) {
            const VarNum& n = V[var];
            const int64_t per_edge = L[var].edge;
            for (int c = 0; c < S.var_ncomp[var]; ++c) {
                uint8_t* d = S.is_dir.data() + n.base + c * n.nscal;
                for (int64_t i = 0; i < NN && L[var].node; ++i) d[i] = snode[i];
                for (int64_t i = 0; i < S.n_edges; ++i)
                    for (int s = 0; s < per_edge; ++s) d[n.edge_base + per_edge * i + s] = sedge[i];
                for (int64_t i = 0; i < S.n_faces && L[var].face; ++i) d[n.face_base + i] = sface[i];
            }
        }
    } else if (prob.dirichlet != ANIFEM_DIRICHLET_NONE) {
        return ANIFEM_ERR_ARG;
    }

    std::vector<int64_t> aptr;
    std::vector<int32_t> adj;
    dof_adjacency(S, aptr, adj);

    // CSR pattern: per row, sort + unique of the DOFs of its tets (of its own
    // component when the components are decoupled)
    S.row_ptr.assign(S.ndof + 1, 0);
    const int nf = S.nf;
    const int64_t nscal_dec = S.decoupled ? S.ndof / S.ncomp : 0;
    auto row_cols = [&](int64_t d, std::vector<int32_t>& buf) {
        buf.clear();
        buf.push_back((int32_t)d);
        const bool rd = S.is_dir[d];
        for (int64_t p = aptr[d]; p < aptr[d + 1]; ++p) {
            const int32_t* td = &S.dofs[(size_t)nf * adj[p]];
            for (int j = 0; j < nf; ++j)
                if (!(rd || S.is_dir[td[j]]) && (nscal_dec == 0 || td[j] / nscal_dec == d / nscal_dec))
                    buf.push_back(td[j]);
        }
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    };
    host_parallel_for(S.ndof, 4096, [&](int64_t b, int64_t e) {
        std::vector<int32_t> buf;
        for (int64_t d = b; d < e; ++d) {
            row_cols(d, buf);
            S.row_ptr[d + 1] = (int64_t)buf.size();
        }
    });
    for (int64_t d = 0; d < S.ndof; ++d) S.row_ptr[d + 1] += S.row_ptr[d];
    S.nnz = S.row_ptr[S.ndof];
    if (S.nnz >= INT32_MAX) return ANIFEM_ERR_ARG;  // slots are int32
    S.col_idx.resize(S.nnz);
    host_parallel_for(S.ndof, 4096, [&](int64_t b, int64_t e) {
        std::vector<int32_t> buf;
        for (int64_t d = b; d < e; ++d) {
            row_cols(d, buf);
            std::copy(buf.begin(), buf.end(), S.col_idx.begin() + S.row_ptr[d]);
        }
    });

    // slot map in the caller's tet order: slot[t*E + ib + nf*ia]
    S.slot.resize((size_t)S.E * NT);
    host_parallel_for(NT, 1 << 12, [&](int64_t b, int64_t e) {
        for (int64_t t = b; t < e; ++t) {
            const int32_t* td = &S.dofs[(size_t)nf * t];
            int32_t* sl = &S.slot[(size_t)S.E * t];
            for (int ia = 0; ia < nf; ++ia)
                for (int ib = 0; ib < nf; ++ib) {
                    const int32_t r = td[ib], c = td[ia];
                    const int32_t* b0 = S.col_idx.data() + S.row_ptr[r];
                    const int32_t* e0 = S.col_idx.data() + S.row_ptr[r + 1];
                    const int32_t* it = std::lower_bound(b0, e0, c);
                    sl[ib + nf * ia] = (it != e0 && *it == c) ? (int32_t)(it - S.col_idx.data()) : -1;
                }
        }
    });
    return ANIFEM_OK;
}

void dof_adjacency(const AssemblySetup& S, std::vector<int64_t>& aptr, std::vector<int32_t>& adj) {
    aptr.assign(S.ndof + 1, 0);
    for (int64_t i = 0; i < (int64_t)S.dofs.size(); ++i) aptr[S.dofs[i] + 1]++;
    for (int64_t d = 0; d < S.ndof; ++d) aptr[d + 1] += aptr[d];
    adj.resize(aptr[S.ndof]);
    std::vector<int64_t> fill(aptr.begin(), aptr.end() - 1);
    for (int64_t t = 0; t < S.n_tets; ++t)
        for (int i = 0; i < S.nf; ++i) adj[fill[S.dofs[(size_t)S.nf * t + i]]++] = (int32_t)t;
}

int build_setup_vars(const anifem_mesh_desc& mesh, int n_var, const anifem_op* vars, const anifem_problem_desc& prob,
                     AssemblySetup& S, const BlockCaps& caps) {
    int rc = setup_prepare_vars(mesh, n_var, vars, S);
    if (!rc) rc = setup_enumerate_host(prob, S);
    if (rc) return rc;
    std::vector<int64_t> aptr;
    std::vector<int32_t> adj;
    dof_adjacency(S, aptr, adj);
    return build_blocks(S, aptr, adj, caps);
}

int build_setup(const anifem_mesh_desc& mesh, const anifem_problem_desc& prob, AssemblySetup& S,
                const BlockCaps& caps) {
    int rc = setup_prepare(mesh, prob, S);
    if (!rc) rc = setup_enumerate_host(prob, S);
    if (rc) return rc;
    std::vector<int64_t> aptr;
    std::vector<int32_t> adj;
    dof_adjacency(S, aptr, adj);
    return build_blocks(S, aptr, adj, caps);
}

}  // namespace anifem
