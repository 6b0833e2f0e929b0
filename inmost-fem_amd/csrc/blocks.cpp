// Row blocks of the owner-computes assembly (see setup.h).
//
// The CSR rows are cut into blocks of consecutive rows.  A block lists the
// distinct tets incident to its rows (ascending tet id) and, for every CSR
// slot of its rows, the (tet, entry) pairs that add into it, in ascending tet
// order: the order in which the reference's element loop adds the element
// matrices into the global matrix (assembler.inl:381-430).  One workgroup per
// block then forms every slot value of its rows exactly once.  Element-level
// homogeneous Dirichlet (dc_on_dof.h:38-45) is resolved here: the only listed
// entries it touches are diagonal ones, whose value becomes exactly 1, and
// it zeroes every load-vector entry of a Dirichlet row.
#include <algorithm>
#include <atomic>

#include "setup.h"

namespace anifem {

int build_blocks(AssemblySetup& S, const std::vector<int64_t>& aptr, const std::vector<int32_t>& adj) {
    const int64_t ND = S.ndof;
    const int nf = S.nf, E = S.E;

    // local index of row r in each incident tet, and the row's contributions
    // (a Dirichlet row has no load-vector contributions: applyDir makes them 0)
    std::vector<uint8_t> lidx(adj.size());
    std::vector<int32_t> row_nc(ND, 0);
    host_parallel_for(ND, 4096, [&](int64_t b, int64_t e) {
        for (int64_t r = b; r < e; ++r) {
            int32_t nc = 0;
            for (int64_t p = aptr[r]; p < aptr[r + 1]; ++p) {
                const int64_t t = adj[p];
                const int32_t* td = &S.dofs[(size_t)nf * t];
                int ib = 0;
                while (td[ib] != r) ++ib;
                lidx[p] = (uint8_t)ib;
                const int32_t* sl = &S.slot[(size_t)E * t + ib];
                for (int ia = 0; ia < nf; ++ia) nc += sl[(size_t)nf * ia] >= 0;
            }
            row_nc[r] = nc;
        }
    });
    int64_t max_valence = 0;
    for (int64_t r = 0; r < ND; ++r) max_valence = std::max<int64_t>(max_valence, aptr[r + 1] - aptr[r]);

    // tets per block: the LDS share of the geometry tensors sets the default,
    // raised to the largest DOF valence (a row's tets must fit one block);
    // codes are 16-bit when tet*E + entry fits
    int64_t cap = std::max<int64_t>(1, std::min<int64_t>(kBlockMaxTets, kBlockKBytes / (8 * S.k_per_tet)));
    cap = std::max(cap, max_valence);
    if (cap > kBlockHardMaxTets || cap * 8 * S.k_per_tet > kBlockMaxKBytes) return ANIFEM_ERR_DIMS;
    // codes: local_tet << ebits | entry, entry == E marking a Dirichlet
    // diagonal entry; load-vector codes local_tet << fbits | i
    S.ebits = 1;
    while ((1 << S.ebits) <= E) ++S.ebits;
    S.fbits = 0;
    while ((1 << S.fbits) < nf) ++S.fbits;
    S.code16 = (cap << S.ebits) <= 65536;
    if (!S.code16 && (65536 >> S.ebits) >= std::max<int64_t>(64, max_valence)) {
        cap = 65536 >> S.ebits;
        S.code16 = true;
    }
    if ((cap << S.fbits) > 65536) return ANIFEM_ERR_DIMS;
    S.block_tet_cap = (int)cap;

    // greedy cut: extend the current block while every budget holds
    S.blk_row0.assign(1, 0);
    S.blk_tet0.assign(1, 0);
    S.blk_c0.assign(1, 0);
    S.blk_g0.assign(1, 0);
    {
        std::vector<int32_t> stamp(S.n_tets, -1);
        int32_t cur = 0;
        int64_t slots = 0, nc = 0, ng = 0, nt = 0;
        auto close = [&](int64_t r) {
            S.blk_row0.push_back((int32_t)r);
            S.blk_tet0.push_back(S.blk_tet0.back() + nt);
            S.blk_c0.push_back(S.blk_c0.back() + nc);
            S.blk_g0.push_back(S.blk_g0.back() + ng);
            ++cur;
            slots = nc = ng = nt = 0;
        };
        for (int64_t r = 0; r < ND; ++r) {
            const int64_t rs = S.row_ptr[r + 1] - S.row_ptr[r], rc = row_nc[r];
            const int64_t rg = S.is_dir[r] ? 0 : aptr[r + 1] - aptr[r];
            if (rs > kBlockMaxSlots || rc > kBlockMaxContrib || rg > kBlockMaxContrib) return ANIFEM_ERR_DIMS;
            int64_t fresh = 0;
            for (int64_t p = aptr[r]; p < aptr[r + 1]; ++p) fresh += stamp[adj[p]] != cur;
            if (slots && (slots + rs > kBlockMaxSlots || nc + rc > kBlockMaxContrib || ng + rg > kBlockMaxContrib ||
                          nt + fresh > S.block_tet_cap)) {
                close(r);
                fresh = aptr[r + 1] - aptr[r];
            }
            for (int64_t p = aptr[r]; p < aptr[r + 1]; ++p) stamp[adj[p]] = cur;
            slots += rs;
            nc += rc;
            ng += rg;
            nt += fresh;
        }
        if (slots || ND == 0) close(ND);
    }
    S.n_blocks = (int64_t)S.blk_row0.size() - 1;
    S.blk_s0.resize(S.n_blocks + 1);
    for (int64_t b = 0; b <= S.n_blocks; ++b) S.blk_s0[b] = S.row_ptr[S.blk_row0[b]];

    // fill the lists, one block at a time
    S.blk_tets.resize(S.blk_tet0.back());
    S.codes.resize(S.blk_c0.back());
    S.sstart.resize(S.nnz);
    S.fcodes.resize(S.blk_g0.back());
    S.rstart.resize(ND);
    std::atomic<int64_t> max_tets{0}, max_c{0}, max_g{0};
    std::atomic<bool> mismatch{false};
    host_parallel_for(S.n_blocks, 16, [&](int64_t b0, int64_t b1) {
        std::vector<int32_t> pos, buf;
        int64_t lt = 0, lc = 0, lg = 0;
        for (int64_t b = b0; b < b1; ++b) {
            const int32_t r0 = S.blk_row0[b], r1 = S.blk_row0[b + 1];
            const int64_t ntb = S.blk_tet0[b + 1] - S.blk_tet0[b];
            buf.assign(adj.begin() + aptr[r0], adj.begin() + aptr[r1]);
            std::sort(buf.begin(), buf.end());
            buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
            const int64_t k = (int64_t)buf.size();
            if (k != ntb) {
                mismatch = true;
                continue;
            }
            int32_t* tl = S.blk_tets.data() + S.blk_tet0[b];
            std::copy(buf.begin(), buf.end(), tl);
            uint32_t* cd = S.codes.data() + S.blk_c0[b];
            uint16_t* fc = S.fcodes.data() + S.blk_g0[b];
            int64_t c = 0, g = 0;
            for (int32_t r = r0; r < r1; ++r) {
                const int64_t base = S.row_ptr[r], rs = S.row_ptr[r + 1] - base;
                pos.assign(rs + 1, 0);
                for (int64_t p = aptr[r]; p < aptr[r + 1]; ++p) {
                    const int32_t* sl = &S.slot[(size_t)E * adj[p] + lidx[p]];
                    for (int ia = 0; ia < nf; ++ia)
                        if (sl[(size_t)nf * ia] >= 0) pos[sl[(size_t)nf * ia] - base + 1]++;
                }
                pos[0] = (int32_t)c;
                for (int64_t u = 0; u < rs; ++u) pos[u + 1] += pos[u];
                const int64_t row_end = pos[rs];
                for (int64_t u = 0; u < rs; ++u) S.sstart[base + u] = (uint16_t)pos[u];
                S.rstart[r] = (uint16_t)g;
                const bool rdir = S.is_dir[r] != 0;
                for (int64_t p = aptr[r]; p < aptr[r + 1]; ++p) {
                    const int32_t t = adj[p];
                    const uint32_t tloc = (uint32_t)(std::lower_bound(tl, tl + k, t) - tl);
                    const int ib = lidx[p];
                    const int32_t* sl = &S.slot[(size_t)E * t + ib];
                    const int32_t* td = &S.dofs[(size_t)nf * t];
                    for (int ia = 0; ia < nf; ++ia) {
                        const int32_t s = sl[(size_t)nf * ia];
                        if (s < 0) continue;
                        // a listed entry on a Dirichlet row or column is the
                        // diagonal one (the pattern drops the others), which
                        // applyDir sets to 1 (dc_on_dof.h:38-45)
                        const bool dir = rdir || S.is_dir[td[ia]];
                        cd[pos[s - base]++] = (tloc << S.ebits) | (uint32_t)(dir ? E : ib + nf * ia);
                    }
                    if (!rdir) fc[g++] = (uint16_t)((tloc << S.fbits) | (uint32_t)ib);
                }
                c = row_end;
            }
            lt = std::max(lt, ntb);
            lc = std::max(lc, c);
            lg = std::max(lg, g);
        }
        int64_t v;
        while ((v = max_tets.load()) < lt && !max_tets.compare_exchange_weak(v, lt)) {}
        while ((v = max_c.load()) < lc && !max_c.compare_exchange_weak(v, lc)) {}
        while ((v = max_g.load()) < lg && !max_g.compare_exchange_weak(v, lg)) {}
    });
    if (mismatch.load()) return ANIFEM_ERR_ARG;
    S.max_block_tets = max_tets.load();
    S.max_block_contrib = max_c.load();
    S.max_block_fcontrib = max_g.load();
    for (int64_t b = 0; b < S.n_blocks; ++b)
        S.max_block_slots = std::max(S.max_block_slots, S.row_ptr[S.blk_row0[b + 1]] - S.row_ptr[S.blk_row0[b]]);
    return ANIFEM_OK;
}

}  // namespace anifem
