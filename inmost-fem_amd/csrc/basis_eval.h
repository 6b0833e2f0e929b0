// Basis functions of the finite-element spaces in reference coordinates,
// shared by the host tables (basis.cpp) and the device evaluation at points
// known only on the device (fem3DapplyX).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/anifem_gpu.h"

namespace anifem {

// Edges of the local tetrahedron in the order 01,02,03,12,13,23 and the P3
// face node triples {0,1,2},{1,2,3},{2,3,0},{3,0,1}
// (docs: local numbering; poly_2.h:75, poly_3.h:83).
__host__ __device__ inline int edge_a(int e) { return e < 3 ? 0 : (e < 5 ? 1 : 2); }
__host__ __device__ inline int edge_b(int e) { return e < 3 ? e + 1 : (e == 3 ? 2 : 3); }
__host__ __device__ inline int face_vertex(int f, int k) { return (f + k) % 4; }

// d lambda_i / d x_a in reference coordinates (lambda_0 = 1 - x1 - x2 - x3).
__host__ __device__ inline double dlam(int i, int a) { return i == 0 ? -1.0 : (i - 1 == a ? 1.0 : 0.0); }

__host__ __device__ inline void basis_values(int fem, const double* L, double* phi) {
    switch (fem) {
        case ANIFEM_FEM_P0: phi[0] = 1.0; break;
        case ANIFEM_FEM_P1:
            for (int i = 0; i < 4; ++i) phi[i] = L[i];
            break;
        case ANIFEM_FEM_P2:
            for (int i = 0; i < 4; ++i) phi[i] = L[i] * (2.0 * L[i] - 1.0);
            for (int e = 0; e < 6; ++e) phi[4 + e] = 4.0 * L[edge_a(e)] * L[edge_b(e)];
            break;
        case ANIFEM_FEM_P3:
            for (int i = 0; i < 4; ++i) phi[i] = 0.5 * L[i] * (3.0 * L[i] - 1.0) * (3.0 * L[i] - 2.0);
            for (int e = 0; e < 6; ++e) {
                // two DOFs per edge (I, J): the one near I first (poly_3.h:39-44)
                double a = L[edge_a(e)], b = L[edge_b(e)];
                phi[4 + 2 * e] = 4.5 * a * (3.0 * a - 1.0) * b;
                phi[5 + 2 * e] = 4.5 * a * (3.0 * b - 1.0) * b;
            }
            for (int f = 0; f < 4; ++f) phi[16 + f] = 27.0 * L[face_vertex(f, 0)] * L[face_vertex(f, 1)] * L[face_vertex(f, 2)];
            break;
        case ANIFEM_FEM_CR1:
            // face i (vertices i, i+1, i+2): 1 - 3 l_{(i+3)%4}
            for (int i = 0; i < 4; ++i) phi[i] = 1.0 - 3.0 * L[(i + 3) % 4];
            break;
        case ANIFEM_FEM_B2:
            for (int e = 0; e < 6; ++e) phi[e] = 4.0 * L[edge_a(e)] * L[edge_b(e)];
            break;
        case ANIFEM_FEM_B4: phi[0] = 256.0 * L[0] * L[1] * L[2] * L[3]; break;
    }
}

// g[i*3 + a] = d phi_i / d x_a (reference coordinates)
__host__ __device__ inline void basis_gradients(int fem, const double* L, double* g) {
    switch (fem) {
        case ANIFEM_FEM_P0:  // the constant: a zero gradient (poly_0.h:11-31)
            g[0] = g[1] = g[2] = 0.0;
            break;
        case ANIFEM_FEM_P1:
            for (int i = 0; i < 4; ++i)
                for (int a = 0; a < 3; ++a) g[i * 3 + a] = dlam(i, a);
            break;
        case ANIFEM_FEM_P2:
            for (int i = 0; i < 4; ++i)
                for (int a = 0; a < 3; ++a) g[i * 3 + a] = (4.0 * L[i] - 1.0) * dlam(i, a);
            for (int e = 0; e < 6; ++e) {
                int I = edge_a(e), J = edge_b(e);
                for (int a = 0; a < 3; ++a) g[(4 + e) * 3 + a] = 4.0 * (L[I] * dlam(J, a) + L[J] * dlam(I, a));
            }
            break;
        case ANIFEM_FEM_P3:
            for (int i = 0; i < 4; ++i) {
                double s = L[i];
                for (int a = 0; a < 3; ++a) g[i * 3 + a] = ((13.5 * s - 9.0) * s + 1.0) * dlam(i, a);
            }
            for (int e = 0; e < 6; ++e) {
                int I = edge_a(e), J = edge_b(e);
                double s = L[I], t = L[J];
                for (int a = 0; a < 3; ++a) {
                    g[(4 + 2 * e) * 3 + a] = 4.5 * ((6.0 * s - 1.0) * t * dlam(I, a) + (3.0 * s - 1.0) * s * dlam(J, a));
                    g[(5 + 2 * e) * 3 + a] = 4.5 * ((6.0 * t - 1.0) * s * dlam(J, a) + (3.0 * t - 1.0) * t * dlam(I, a));
                }
            }
            for (int f = 0; f < 4; ++f) {
                int A = face_vertex(f, 0), B = face_vertex(f, 1), C = face_vertex(f, 2);
                for (int a = 0; a < 3; ++a)
                    g[(16 + f) * 3 + a] =
                        27.0 * (dlam(A, a) * L[B] * L[C] + L[A] * dlam(B, a) * L[C] + L[A] * L[B] * dlam(C, a));
            }
            break;
        case ANIFEM_FEM_CR1:
            for (int i = 0; i < 4; ++i)
                for (int a = 0; a < 3; ++a) g[i * 3 + a] = -3.0 * dlam((i + 3) % 4, a);
            break;
        case ANIFEM_FEM_B2:
            for (int e = 0; e < 6; ++e) {
                int I = edge_a(e), J = edge_b(e);
                for (int a = 0; a < 3; ++a) g[e * 3 + a] = 4.0 * (L[I] * dlam(J, a) + L[J] * dlam(I, a));
            }
            break;
        case ANIFEM_FEM_B4:
            for (int a = 0; a < 3; ++a) {
                double s = 0;
                for (int v = 0; v < 4; ++v) {
                    double prod = dlam(v, a);
                    for (int w = 0; w < 4; ++w)
                        if (w != v) prod *= L[w];
                    s += prod;
                }
                g[a] = 256.0 * s;
            }
            break;
        default:
            for (int a = 0; a < 3; ++a) g[a] = 0.0;
    }
}

// Reference fields of the vector-valued spaces (the part of OP(phi_i) that
// space_map's map and scale turn into physical values), d components each:
//   RT0 IDEN: xi - xi_{(i+3)%4} (xi_a = l_{a+1}); DIV: 3; CURL: 0; GRAD: 1
//   ND0 IDEN on edge (a, b): l_a gh_b - l_b gh_a; CURL: 2 gh_a x gh_b; DIV: 0;
//   GRAD: gh_a x gh_b (the axial vector of the skew gradient)
__host__ __device__ inline void vector_field(int op, int fem, const double* L, double* u) {
    if (fem == ANIFEM_FEM_RT0) {
        for (int i = 0; i < 4; ++i) {
            const int v = (i + 3) % 4;
            if (op == ANIFEM_IDEN)
                for (int a = 0; a < 3; ++a) u[i * 3 + a] = L[a + 1] - (v == a + 1 ? 1.0 : 0.0);
            else if (op == ANIFEM_DIV) u[i] = 3.0;
            else if (op == ANIFEM_GRAD) u[i] = 1.0;
            else
                for (int a = 0; a < 3; ++a) u[i * 3 + a] = 0.0;
        }
    } else {
        for (int e = 0; e < 6; ++e) {
            const int A = edge_a(e), B = edge_b(e);
            if (op == ANIFEM_IDEN) {
                for (int c = 0; c < 3; ++c) u[e * 3 + c] = L[A] * dlam(B, c) - L[B] * dlam(A, c);
            } else if (op == ANIFEM_CURL || op == ANIFEM_GRAD) {
                const double f = op == ANIFEM_CURL ? 2.0 : 1.0;
                for (int c = 0; c < 3; ++c) {
                    const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                    u[e * 3 + c] = f * (dlam(A, c1) * dlam(B, c2) - dlam(A, c2) * dlam(B, c1));
                }
            } else {
                u[e] = 0.0;
            }
        }
    }
}


}  // namespace anifem
