// Row blocks of the owner-computes assembly (see setup.h).
//
// The CSR rows are cut into blocks of consecutive rows.  Within a block:
//   * a pair is one (row r, incident tet t): it forms row ib of t's element
//     matrix (ib = the local index of r in t) and entry ib of t's load vector;
//     pairs are ordered by row, then ascending tet id;
//   * the block's element entries are laid out in a value list grouped by CSR
//     slot, each slot's entries in ascending tet id: the order in which the
//     reference's element loop adds the element matrices into the global
//     matrix (assembler.inl:381-430); pair_pos gives, for every entry of a
//     pair, its place in that list, or kPosDropped for an entry the pattern
//     drops.
// One workgroup per block then forms every slot value and rhs entry of its
// rows exactly once.  Element-level homogeneous Dirichlet (dc_on_dof.h:38-45)
// is resolved here: the pattern keeps only the diagonal entry of a Dirichlet
// row or column, whose value applyDir sets to 1, and a Dirichlet row's load
// vector entries are 0; such a pair carries kPairDirichlet.
#include <algorithm>
#include <atomic>

#include "setup.h"

namespace anifem {

int build_blocks(AssemblySetup& S, const std::vector<int64_t>& aptr, const std::vector<int32_t>& adj,
                 const BlockCaps& caps) {
    if (caps.max_slots < 1 || caps.max_slots > 65535 || caps.max_values < 1 || caps.max_values > 65535 ||
        caps.max_pairs < 1 || caps.max_pairs > 256)
        return ANIFEM_ERR_ARG;
    const int64_t ND = S.ndof;
    const int nf = S.nf, E = S.E;

    // local index of row r in each incident tet, and the row's entries
    std::vector<uint8_t> lidx(adj.size());
    std::vector<int32_t> row_nv(ND, 0);
    host_parallel_for(ND, 4096, [&](int64_t b, int64_t e) {
        for (int64_t r = b; r < e; ++r) {
            int32_t nv = 0;
            for (int64_t p = aptr[r]; p < aptr[r + 1]; ++p) {
                const int64_t t = adj[p];
                const int32_t* td = &S.dofs[(size_t)nf * t];
                int ib = 0;
                while (td[ib] != r) ++ib;
                lidx[p] = (uint8_t)ib;
                const int32_t* sl = &S.slot[(size_t)E * t + ib];
                for (int ia = 0; ia < nf; ++ia) nv += sl[(size_t)nf * ia] >= 0;
            }
            row_nv[r] = nv;
        }
    });

    // greedy cut: extend the current block while every budget holds
    S.blk_row0.assign(1, 0);
    S.blk_p0.assign(1, 0);
    S.blk_v0.assign(1, 0);
    {
        int64_t slots = 0, nv = 0, np = 0;
        auto close = [&](int64_t r) {
            S.blk_row0.push_back((int32_t)r);
            S.blk_p0.push_back(S.blk_p0.back() + np);
            S.blk_v0.push_back(S.blk_v0.back() + nv);
            slots = nv = np = 0;
        };
        for (int64_t r = 0; r < ND; ++r) {
            const int64_t rs = S.row_ptr[r + 1] - S.row_ptr[r], rv = row_nv[r], rp = aptr[r + 1] - aptr[r];
            if (rs > caps.max_slots || rv > caps.max_values || rp > caps.max_pairs) return ANIFEM_ERR_DIMS;
            if (slots && (slots + rs > caps.max_slots || nv + rv > caps.max_values || np + rp > caps.max_pairs))
                close(r);
            slots += rs;
            nv += rv;
            np += rp;
        }
        if (slots || ND == 0) close(ND);
    }
    S.n_blocks = (int64_t)S.blk_row0.size() - 1;
    S.blk_s0.resize(S.n_blocks + 1);
    for (int64_t b = 0; b <= S.n_blocks; ++b) S.blk_s0[b] = S.row_ptr[S.blk_row0[b]];

    // fill the pair lists and value positions, one block at a time
    const int64_t NP = S.blk_p0.back();
    S.pair_tet.resize(NP);
    S.pair_ib.resize(NP);
    S.pair_pos.resize((size_t)NP * nf);
    S.sstart.resize(S.nnz);
    S.pstart.resize(ND);
    host_parallel_for(S.n_blocks, 16, [&](int64_t b0, int64_t b1) {
        std::vector<int32_t> cur;
        for (int64_t b = b0; b < b1; ++b) {
            const int32_t r0 = S.blk_row0[b], r1 = S.blk_row0[b + 1];
            int64_t p = S.blk_p0[b];
            int64_t v = 0;  // value offset within the block
            for (int32_t r = r0; r < r1; ++r) {
                const int64_t base = S.row_ptr[r], rs = S.row_ptr[r + 1] - base;
                // the row's slots: entries counted, then laid out by ascending tet
                cur.assign(rs + 1, 0);
                for (int64_t q = aptr[r]; q < aptr[r + 1]; ++q) {
                    const int32_t* sl = &S.slot[(size_t)E * adj[q] + lidx[q]];
                    for (int ia = 0; ia < nf; ++ia)
                        if (sl[(size_t)nf * ia] >= 0) cur[sl[(size_t)nf * ia] - base + 1]++;
                }
                cur[0] = (int32_t)v;
                for (int64_t u = 0; u < rs; ++u) cur[u + 1] += cur[u];
                const int64_t row_end = cur[rs];
                for (int64_t u = 0; u < rs; ++u) S.sstart[base + u] = (uint16_t)cur[u];
                S.pstart[r] = (uint16_t)(p - S.blk_p0[b]);
                const bool rdir = S.is_dir[r] != 0;
                for (int64_t q = aptr[r]; q < aptr[r + 1]; ++q, ++p) {
                    const int32_t t = adj[q];
                    const int ib = lidx[q];
                    S.pair_tet[p] = t;
                    S.pair_ib[p] = (uint8_t)(ib | (rdir ? kPairDirichlet : 0));
                    const int32_t* sl = &S.slot[(size_t)E * t + ib];
                    uint16_t* pos = &S.pair_pos[(size_t)nf * p];
                    for (int ia = 0; ia < nf; ++ia) {
                        const int32_t s = sl[(size_t)nf * ia];
                        pos[ia] = s >= 0 ? (uint16_t)cur[s - base]++ : kPosDropped;
                    }
                }
                v = row_end;
            }
        }
    });
    // processing order: by the largest id of the blocks' tets, so that blocks
    // sharing tets (node rows and the edge / face rows around them, far apart
    // in row order) run close in time and find the tets' geometry rows in L2.
    // A node row's tets lie in a few separate id ranges (the cells around the
    // node); the block runs with the last of them, while the earlier ones are
    // still cached.  (A replay of the pass through a 4 MB LRU cache per XCD
    // fetches each K row about twice with this key, four times with the mean
    // id.)
    {
        std::vector<int64_t> key(S.n_blocks);
        for (int64_t b = 0; b < S.n_blocks; ++b) {
            int64_t mx = -1;
            for (int64_t p = S.blk_p0[b]; p < S.blk_p0[b + 1]; ++p) mx = std::max<int64_t>(mx, S.pair_tet[p]);
            key[b] = mx;
        }
        S.blk_order.resize(S.n_blocks);
        for (int64_t b = 0; b < S.n_blocks; ++b) S.blk_order[b] = (int32_t)b;
        std::stable_sort(S.blk_order.begin(), S.blk_order.end(),
                         [&](int32_t x, int32_t y) { return key[x] < key[y]; });
        S.blk_meta.resize(S.n_blocks);
        for (int64_t k = 0; k < S.n_blocks; ++k) {
            const int64_t b = S.blk_order[k];
            BlockMeta& M = S.blk_meta[k];
            M.s0 = S.blk_s0[b];
            M.p0 = S.blk_p0[b];
            M.q0 = S.blk_p0[b] * nf;
            M.r0 = S.blk_row0[b];
            M.nr = (uint16_t)(S.blk_row0[b + 1] - S.blk_row0[b]);
            M.ns = (uint16_t)(S.blk_s0[b + 1] - S.blk_s0[b]);
            M.np = (uint16_t)(S.blk_p0[b + 1] - S.blk_p0[b]);
            M.nv = (uint16_t)(S.blk_v0[b + 1] - S.blk_v0[b]);
            M.kind = 0;
            M.pad = 0;
        }
    }
    for (int64_t b = 0; b < S.n_blocks; ++b) {
        S.max_block_values = std::max(S.max_block_values, S.blk_v0[b + 1] - S.blk_v0[b]);
        S.max_block_pairs = std::max(S.max_block_pairs, S.blk_p0[b + 1] - S.blk_p0[b]);
        S.max_block_slots = std::max(S.max_block_slots, S.blk_s0[b + 1] - S.blk_s0[b]);
    }
    return ANIFEM_OK;
}

int compact_positions(const AssemblySetup& S, std::vector<uint32_t>& out) {
    const int nf = S.nf;
    const int words = (2 + nf + 3) / 4;
    const int64_t NP = (int64_t)S.pair_tet.size();
    out.assign((size_t)NP * words, 0u);
    std::atomic<bool> ok{true};
    host_parallel_for(NP, 1 << 14, [&](int64_t b, int64_t e) {
        for (int64_t p = b; p < e && ok.load(std::memory_order_relaxed); ++p) {
            const uint16_t* pos = &S.pair_pos[(size_t)p * nf];
            int lo = 0xFFFF, hi = -1;
            for (int ia = 0; ia < nf; ++ia)
                if (pos[ia] != kPosDropped) {
                    lo = std::min<int>(lo, pos[ia]);
                    hi = std::max<int>(hi, pos[ia]);
                }
            if (hi < 0 || hi - lo >= 0xFF) {
                ok = false;
                return;
            }
            uint8_t bytes[4 * 16] = {};
            bytes[0] = (uint8_t)(lo & 0xFF);
            bytes[1] = (uint8_t)(lo >> 8);
            for (int ia = 0; ia < nf; ++ia) bytes[2 + ia] = pos[ia] == kPosDropped ? 0xFF : (uint8_t)(pos[ia] - lo);
            for (int w = 0; w < words; ++w)
                out[(size_t)p * words + w] = (uint32_t)bytes[4 * w] | (uint32_t)bytes[4 * w + 1] << 8 |
                                             (uint32_t)bytes[4 * w + 2] << 16 | (uint32_t)bytes[4 * w + 3] << 24;
        }
    });
    if (!ok.load()) {
        std::vector<uint32_t>().swap(out);
        return 0;
    }
    return words;
}

}  // namespace anifem
