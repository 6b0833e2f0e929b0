// Synthetic jittered Kuhn-cube meshes (host): anifem_mesh_kuhn_cube.
//
// The unit cube is split into n^3 sub-cubes and every sub-cube into the 6
// Kuhn (Freudenthal) tetrahedra that share its main diagonal: for each
// permutation s of the axes, [v000, v000+e_s0, v000+e_s0+e_s1, v111].  Node
// (i, j, k) sits at (i, j, k) h with id (i*(n+1)+j)*(n+1)+k, the V_ID
// numbering of the reference's cube generator (utils/mesh_utils.cpp:124).
// Interior nodes move by i.i.d. U[-jitter h, jitter h] per coordinate drawn
// from Philox-4x32-10 (Salmon et al., SC'11) keyed by the seed with the node id
// as counter; with jitter < 1/4 no Kuhn tet can invert.
#include <cmath>
#include <cstdint>

#include "../../include/anifem_gpu.h"

namespace {

inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    uint64_t p = (uint64_t)a * b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

// Philox-4x32 with 10 rounds.
void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(M0, c[0], hi0, lo0);
        mulhilo(M1, c[2], hi1, lo1);
        uint32_t n0 = hi1 ^ c[1] ^ k0, n1 = lo1, n2 = hi0 ^ c[3] ^ k1, n3 = lo0;
        c[0] = n0, c[1] = n1, c[2] = n2, c[3] = n3;
        k0 += W0;
        k1 += W1;
    }
}

}  // namespace

extern "C" int anifem_mesh_kuhn_cube(int64_t n, double jitter, uint64_t seed, double* coords, int32_t* tets) {
    if (n < 1 || n > 1200 || !(jitter >= 0.0 && jitter < 0.25)) return ANIFEM_ERR_ARG;
    const int64_t m = n + 1;
    if (m * m * m >= INT32_MAX) return ANIFEM_ERR_ARG;
    const double h = 1.0 / (double)n;
    if (coords) {
        for (int64_t i = 0; i < m; ++i)
            for (int64_t j = 0; j < m; ++j)
                for (int64_t k = 0; k < m; ++k) {
                    const int64_t id = (i * m + j) * m + k;
                    double x[3] = {i * h, j * h, k * h};
                    const bool interior = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
                    if (interior && jitter > 0) {
                        uint32_t c[4] = {(uint32_t)id, (uint32_t)(id >> 32), 0u, 0u};
                        philox4x32(c, (uint32_t)seed, (uint32_t)(seed >> 32));
                        for (int d = 0; d < 3; ++d) {
                            const double u = ((double)c[d] + 0.5) * (1.0 / 4294967296.0);
                            x[d] += (2.0 * u - 1.0) * jitter * h;
                        }
                    }
                    for (int d = 0; d < 3; ++d) coords[3 * id + d] = x[d];
                }
    }
    if (tets) {
        static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        int64_t t = 0;
        for (int64_t i = 0; i < n; ++i)
            for (int64_t j = 0; j < n; ++j)
                for (int64_t k = 0; k < n; ++k)
                    for (int p = 0; p < 6; ++p) {
                        int64_t c[3] = {i, j, k};
                        int32_t* v = tets + 4 * t;
                        v[0] = (int32_t)((c[0] * m + c[1]) * m + c[2]);
                        for (int s = 0; s < 3; ++s) {
                            c[perms[p][s]] += 1;
                            v[s + 1] = (int32_t)((c[0] * m + c[1]) * m + c[2]);
                        }
                        ++t;
                    }
    }
    return ANIFEM_OK;
}
