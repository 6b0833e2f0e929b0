/* FP64 CPU oracle for the AniFem++ element-matrix and global-assembly path.
 *
 * TEST INFRASTRUCTURE ONLY.  This restatement is the checker the tests, the
 * smoke check and the bench's CPU-baseline leg compare the HIP path against.
 * Nothing in the product (the HIP library, the C++ host mirror, the Python
 * loader) includes, links or loads it.
 *
 * Every function cites the reference file:line it follows.  Parity status of
 * each part (what pins it) is stated next to its declaration.
 */
#ifndef ANIFEM_ORACLE_H
#define ANIFEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Operator / element codes use the reference enums
 * (anifem++/fem/operators.h:25-45): IDEN=1, GRAD=2; FEM_P0=1 .. FEM_P3=4. */
enum { ORC_IDEN = 1, ORC_GRAD = 2, ORC_DIV = 3, ORC_CURL = 4, ORC_DUDX = 6, ORC_DUDY = 7, ORC_DUDZ = 8 };
enum { ORC_P0 = 1, ORC_P1 = 2, ORC_P2 = 3, ORC_P3 = 4 };
/* operators.h:26-35 */
enum { ORC_RT0 = 21, ORC_ND0 = 25, ORC_CR1 = 31, ORC_B2 = 1002, ORC_B4 = 1004 };
/* Tensor types (anifem++/fem/diff_tensor.h:17-22). */
enum { ORC_TENSOR_NULL = 1, ORC_TENSOR_SCALAR = 2, ORC_TENSOR_SYMMETRIC = 3, ORC_TENSOR_GENERAL = 4 };
/* Coefficient storage: one tensor for all tets, one per tet, one per (tet, point). */
enum { ORC_COEF_CONST = 0, ORC_COEF_PER_TET = 1, ORC_COEF_PER_POINT = 2 };

/* Symmetric tetrahedron rule of the given order (0..6).  Returns the number of
 * points; fills p[4q] (barycentric) and w[q] when non-null.  Built by solving
 * the moment equations with Gauss-Newton in long double from coarse seeds.
 * Pinned by: polynomial exactness (tests/fem/quadrature_formulas_test.cpp:7-45
 * restated) and the reference rules of oracle/_ref where that library exists. */
int orc_quadrature(int order, double* p, double* w);
int orc_face_quadrature(int face, double* p, double* w);

/* inverse3x3 (anifem++/fem/geometry.h:25-45): column-major inverse, returns det. */
double orc_inverse3x3(const double* m, double* inv);

/* Element matrices of fem3Dtet<Operator<opA,femA^ncA>, Operator<opB,femB^ncB>>
 * over n tetrahedra (anifem++/fem/operations/int_tet.inl:92-122 ->
 * core.inl:218-367).  xyz is n x 12 (P0,P1,P2,P3 per tet).  D layout by
 * coef_mode: CONST -> ddim values; PER_TET -> n x ddim; PER_POINT -> n x q x ddim,
 * where ddim = 0 (NULL), 1 (SCALAR) or dimB*dimA (column-major dimB x dimA,
 * SYMMETRIC and GENERAL).  A receives n matrices of nfB x nfA, column-major
 * (rows = test functions of B).  Returns 0, or a negative code:
 * -1 bad operator, -2 bad order, -3 incompatible dimensions for NULL/SCALAR
 * (the reference throws at diff_tensor.h:313-315).
 * Pinned by: reference fixtures (tests/golden) and the reference's own
 * known-answer tables (tests/fem/operations/int_tet_test.cpp:191-500). */
int orc_fem3dtet(int opA, int femA, int ncA, int opB, int femB, int ncB, int tensor_type, int coef_mode,
                 int order, long n, const double* xyz, const double* D, double* A);

/* Face integrals over face `face` (vertices face, face+1, face+2 mod 4) of n
 * tetrahedra: fem3Dface (dot_normal 0) and fem3DfaceN (dot_normal 1: D has
 * 3*dimB rows, DU contracted with the outward unit normal) of
 * anifem++/fem/operations/int_face.inl:12-82, core.h:176-195.  The oracle's own
 * triangle rule (orc_face_quadrature: 36-point collapsed Gauss, degree 10)
 * replaces the reference's order-selected rule; every integrand on the path is
 * a polynomial of lower degree, so the integrals agree to rounding.  PER_POINT
 * D is given at the orc_face_quadrature points.  Layouts and codes as
 * orc_fem3dtet.
 * Pinned by: reference fixtures (tests/golden/face_*.npz from the compiled
 * reference's fem3Dface / fem3DfaceN) and closed forms (face mass
 * area/12 (1 + delta_ij) for P1, int_f phi_i = area/3). */
int orc_fem3dface(int opA, int femA, int ncA, int opB, int femB, int ncB, int tensor_type, int coef_mode, int face,
                  int dot_normal, long n, const double* xyz, const double* D, double* A);

/* fem3DapplyL (anifem++/fem/operations/eval.inl:84-100, core.inl:370-405):
 * Op(u) at q barycentric points xyl (4q) of n tets with per-tet DOF values
 * dofs (n x nfa); out (n x dim*q) receives out[r*dim*q + k + dim*n].
 * Pinned by: fixtures of the compiled reference's fem3DapplyL
 * (tests/golden/apply_*.npz) and the reproduction of polynomials of the
 * space's degree from their nodal interpolants. */
int orc_fem3dapply(int op, int fem, int ncomp, int q, const double* xyl, long n, const double* xyz,
                   const double* dofs, double* out);
/* fem3DapplyX: q physical points per tet, X[(r*q + n)*3 + c]. */
int orc_fem3dapplyx(int op, int fem, int ncomp, int q, long n, const double* xyz, const double* X,
                    const double* dofs, double* out);

/* Number of basis functions / operator dimension for (op, fem, ncomp). */
int orc_nfa(int fem, int ncomp);
int orc_dim(int op, int ncomp);

/* applyDir (anifem++/fem/operations/dc_on_dof.h:38-45) on a column-major
 * nf x nf matrix A and vector F. */
void orc_apply_dir(int nf, double* A, double* F, int k, double bc);

/* applyDir on every flagged local DOF of every cell, in ascending local order
 * (ex1.cpp:100-105): A n_tets column-major nf x nf matrices, F n_tets vectors
 * of nf (may be NULL), bc the Dirichlet value of each global DOF (NULL: 0). */
void orc_apply_dir_cells(long n_tets, int nf, const int32_t* dofs, const uint8_t* is_dir, const double* bc,
                         double* A, double* F);

/* ---- global assembly (assembler.inl, ordering.inl, global_enumerator.cpp) ---- */

/* Local node reorder of the assembler (ordering.inl:8-26), applied in place to
 * tets (n x 4 node ids): swap local nodes 2 and 3 when det(P0-P3,P1-P3,P2-P3)<0. */
void orc_reorder_nodes(long n_tets, int32_t* tets, const double* coords);

/* Global DOF enumeration in NATURAL order (global_enumerator.h:392-404):
 * (component, geometric type {node, edge, face}, element id, dof on element).
 * Edges and faces get ids by first appearance in cell order over the local
 * edge order 01,02,03,12,13,23 and local faces {k,k+1,k+2 mod 4}.  P3 edge DOF
 * pairs follow the global-node-id orientation (ordering.inl:137-151,
 * tetdofmap.inl:98-105).  dofs receives n_tets x (ncomp*nf) global ids in the
 * reference's local order (component-major for FemVec).  Returns the number of
 * DOFs, or <0 on error.  n_edges/n_faces (nullable) receive the entity counts. */
long orc_enumerate(int fem, int ncomp, long n_nodes, long n_tets, const int32_t* tets, int32_t* dofs,
                   long* n_edges, long* n_faces);

/* CSR pattern: structural pattern of the DOF graph (every pair of DOFs sharing
 * a tet, diagonal always, assembler.inl:114-136), minus the off-diagonal
 * entries of Dirichlet rows and columns (their element entries are exactly zero
 * after applyDir and the drop rule |a| <= 1e-100 never inserts them,
 * assembler.inl:416).  is_dir may be NULL.  Two-call protocol: with col_idx ==
 * NULL only row_ptr (ndof+1) is filled and nnz is returned. */
long orc_csr_pattern(long ndof, long n_tets, int nf, const int32_t* dofs, const uint8_t* is_dir, int64_t* row_ptr,
                     int32_t* col_idx);

/* Element-to-slot map: slot[t*nf*nf + ib + nf*ia] = CSR position of
 * (dofs[t][ib], dofs[t][ia]) or -1 when the entry is dropped. */
int orc_slot_map(long n_tets, int nf, const int32_t* dofs, const int64_t* row_ptr, const int32_t* col_idx,
                 int64_t* slot);

/* Greedy colouring of tets (two tets conflict when they share a node).
 * Returns the number of colours; colour receives one colour per tet. */
int orc_greedy_colour(long n_nodes, long n_tets, const int32_t* tets, int32_t* colour);

/* Sequential cell-order scatter (assembler.inl:397-481): val[slot] += A entry
 * when |a| > drop_val, rhs[dof] += F.  Returns 0, -1 when a non-finite value
 * is met (assembler.inl:419-424,475-479), or -3 when a non-dropped entry has no
 * slot. */
int orc_scatter(long n_tets, int nf, const int32_t* dofs, const int64_t* slot, const double* A, const double* F,
                double drop_val, double* val, double* rhs);

/* orc_scatter parallel over colours (no two tets of a colour share a node):
 * colour c holds colour_tets[colour_ptr[c] .. colour_ptr[c+1]).  Same return
 * codes; values equal orc_scatter's up to the order of the additions. */
int orc_scatter_coloured(int n_colours, const int64_t* colour_ptr, const int32_t* colour_tets, int nf,
                         const int32_t* dofs, const int64_t* slot, const double* A, const double* F, double drop_val,
                         int nthreads, double* val, double* rhs);

/* y = A x for CSR A. */
void orc_spmv(long nrow, const int64_t* row_ptr, const int32_t* col_idx, const double* val, const double* x,
              double* y);

/* Jacobi-preconditioned CG to ||b - A x|| <= rtol ||b||.  x holds the initial
 * guess.  Returns the iteration count, or -1 if maxit was reached. */
long orc_jacobi_pcg(long nrow, const int64_t* row_ptr, const int32_t* col_idx, const double* val, const double* b,
                    double* x, double rtol, long maxit, double* final_relres);

#ifdef __cplusplus
}
#endif
#endif
