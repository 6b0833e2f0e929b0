/* anifem_gpu.h -- C ABI of the MI355X element-matrix / global-assembly library.
 *
 * Plain C: pointers, sizes and enums only.  Device buffers are HIP device
 * pointers owned by the caller; `stream` is a hipStream_t passed as void*
 * (NULL = the default stream).  All calls are stream-ordered; calls on
 * different streams with separate buffers are independent.
 *
 * Reference interfaces each entry point replaces are cited per declaration
 * (paths relative to the AniFem++ source tree).
 */
#ifndef ANIFEM_GPU_H
#define ANIFEM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define ANIFEM_OK 0
#define ANIFEM_ERR_NONFINITE (-1)   /* Assemble's -1: NaN/Inf in a local matrix or rhs (assembler.inl:419-424) */
#define ANIFEM_ERR_CONNECTIVITY (-2) /* Assemble's -2: broken cell connectivity (assembler.inl:309-312) */
#define ANIFEM_ERR_ARG (-3)          /* invalid argument / unsupported combination */
#define ANIFEM_ERR_ORDER (-4)        /* quadrature order outside 0..20 (quadrature_formulas.cpp:1497-1498 throws) */
#define ANIFEM_ERR_DIMS (-5)         /* NULL/SCALAR tensor with incompatible operator dims (diff_tensor.h:313-315) */
#define ANIFEM_ERR_HIP (-6)          /* HIP runtime failure */
#define ANIFEM_ERR_NOMEM (-7)        /* allocation failure */
#define ANIFEM_ERR_NOT_CONVERGED (-8) /* PCG reached maxit */

/* ---- operators: Operator<OP, FemFix<FEM>> or Operator<OP, FemVec<ncomp, FEM>>
 * Numbering follows the reference enums (anifem++/fem/operators.h:25-45). */
/* Operators (the reference's OperatorType, operators.h:37-45).  DIV and CURL
 * act on 3-component vector spaces, DUDX/DUDY/DUDZ on any space (one partial
 * derivative per component); all three are formed from the gradient. */
enum {
    ANIFEM_IDEN = 1,
    ANIFEM_GRAD = 2,
    ANIFEM_DIV = 3,
    ANIFEM_CURL = 4,
    ANIFEM_DUDX = 6,
    ANIFEM_DUDY = 7,
    ANIFEM_DUDZ = 8
};
enum { ANIFEM_FEM_P0 = 1, ANIFEM_FEM_P1 = 2, ANIFEM_FEM_P2 = 3, ANIFEM_FEM_P3 = 4 };
/* The other FemFix spaces of the reference (operators.h:26-35; the spaces/ headers),
 * same codes.  They are accepted by the element-level entry points
 * (anifem_fem3dtet_batch in FP64 / FP32, anifem_fem3dface_batch,
 * anifem_fem3dapply_batch, anifem_fem3dapplyx_batch), not by the global
 * assembly:
 *   CR1 Crouzeix-Raviart, 1 - 3 lambda_{(i+3)%4} per face (crouzeix_raviart_1.h)
 *   B2  edge bubbles 4 lambda_i lambda_j, edges 01,02,03,12,13,23 (bubble_2.h)
 *   B4  cell bubble 256 lambda_0 lambda_1 lambda_2 lambda_3 (bubble_4.h)
 *   RT0 Raviart-Thomas, one vector field per face: IDEN, DIV (CURL = 0)
 *       (raviant_thomas_0.h)
 *   ND0 Nedelec, one vector field per edge: IDEN, CURL (DIV = 0)
 *       (nedelec_0.h)
 * RT0 and ND0 take ncomp = 1; their GRAD is not provided. */
enum {
    ANIFEM_FEM_RT0 = 21,
    ANIFEM_FEM_ND0 = 25,
    ANIFEM_FEM_CR1 = 31,
    ANIFEM_FEM_B2 = 1002,
    ANIFEM_FEM_B4 = 1004
};

typedef struct anifem_op {
    int32_t op;    /* ANIFEM_IDEN, ANIFEM_GRAD, ANIFEM_DIV, ANIFEM_CURL, ANIFEM_DUDX..DUDZ */
    int32_t fem;   /* ANIFEM_FEM_P0 .. ANIFEM_FEM_P3, or one of the spaces above */
    int32_t ncomp; /* 1 for FemFix, >1 for FemVec<ncomp, fem> (operators.h:63-67,127-155) */
} anifem_op;

/* ---- tensor coefficient D (dimB x dimA), the Dfnc of fem3Dtet ----------
 * Types follow TensorType (diff_tensor.h:17-22); LAME is an isotropic
 * elasticity tensor C((i,j),(k,l)) = lambda d_ij d_kl + mu (d_ik d_jl + d_il d_jk)
 * over 9x9 gradient components (the 9x9 TENSOR_SYMMETRIC of
 * examples/tutorials/lin_elast.cpp:101-117) given by its 2 parameters.
 *
 * Storage (device, FP64, structure-of-arrays over tets):
 *   CONST     : ddim values shared by all tets  (DfuncTraits<..., true>)
 *   PER_TET   : value v of tet r at data[v*n_tets + r]   (per-tet tensor, PerTetra)
 *   PER_POINT : value v of tet r at quadrature point n at data[(v*q + n)*n_tets + r]
 *               (a host functor evaluated at the physical quadrature points:
 *                the DfuncTraitsFusive layout of diff_tensor.h:31-32,52-53)
 * ddim: NULL 0; SCALAR 1; SYMMETRIC dimB*(dimB+1)/2 packed upper triangle,
 * column-major (D(i,j), i<=j, at i + j*(j+1)/2); GENERAL dimB*dimA
 * column-major; LAME 2 (lambda, mu). */
enum {
    ANIFEM_TENSOR_NULL = 1,
    ANIFEM_TENSOR_SCALAR = 2,
    ANIFEM_TENSOR_SYMMETRIC = 3,
    ANIFEM_TENSOR_GENERAL = 4,
    ANIFEM_TENSOR_LAME = 101
};
enum { ANIFEM_COEF_CONST = 0, ANIFEM_COEF_PER_TET = 1, ANIFEM_COEF_PER_POINT = 2 };

typedef struct anifem_tensor {
    int32_t type;       /* ANIFEM_TENSOR_* */
    int32_t storage;    /* ANIFEM_COEF_* */
    const double* data; /* device pointer (may be NULL for TENSOR_NULL) */
} anifem_tensor;

/* Arithmetic of the element contraction.  Geometry (Jacobian, inverse,
 * determinant) is always FP64; global CSR / RHS accumulation is always FP64. */
enum {
    ANIFEM_PREC_FP64 = 0,  /* FP64 contraction, FP64 element output */
    ANIFEM_PREC_FP32 = 1,  /* FP32 contraction, FP32 element output */
    /* anifem_fem3dtet_batch only, tensors constant over each tet: R and the
     * per-tet geometry tensor, each scaled to [-1, 1], split into FP16 hi + lo
     * and contracted on MFMA (three products, FP32 accumulation), FP32 element
     * output.  No reference counterpart (the reference is FP64 throughout). */
    ANIFEM_PREC_FP16X3 = 2
};

/* ---- element level ------------------------------------------------------
 * Batched fem3Dtet<OpA, OpB, DfuncTraits>(XYZ, Dfnc, A, order): the element
 * matrices  A_r(ib, ia) = int_{T_r} (D OpA(phi_ia)) . OpB(phi_ib) dx  of n_tets
 * tetrahedra (anifem++/fem/operations/int_tet.h:17-20, int_tet.inl:92-122,
 * core.inl:278-367).
 *   xyz   : device, 12 x n_tets SoA: coordinate c of vertex v of tet r at
 *           xyz[(3*v + c)*n_tets + r]  (the Tetras XY0..XY3 columns)
 *   A_out : device, n_tets matrices of nfB x nfA, column-major, one after the
 *           other (the reference's fused layout A[ib + nfB*(ia + nfA*r)]);
 *           double for ANIFEM_PREC_FP64, float for ANIFEM_PREC_FP32 and
 *           ANIFEM_PREC_FP16X3.
 * Errors mirror the reference's throws: ANIFEM_ERR_ORDER (order outside 0..20),
 * ANIFEM_ERR_DIMS (NULL/SCALAR with dimA != dimB), ANIFEM_ERR_ARG. */
int anifem_fem3dtet_batch(anifem_op opA, anifem_op opB, const anifem_tensor* D, int order, int64_t n_tets,
                          const double* xyz, int precision, void* A_out, void* stream);

/* Batched face integrals over face `face` (0..3: the face through vertices
 * face, face+1, face+2 mod 4) of each tet:
 *   dot_normal = 0: fem3Dface<OpA, OpB>(XYZ, face, Dfnc, A, order)
 *     A_r(ib, ia) = int_{f_r} (D OpA(phi_ia)) . OpB(phi_ib) ds
 *     (anifem++/fem/operations/int_face.h:15-20, int_face.inl:12-44);
 *   dot_normal = 1: fem3DfaceN<OpA, OpB>(XYZ, face, Dfnc, A, order)
 *     A_r(ib, ia) = int_{f_r} ((D OpA(phi_ia)) . n) . OpB(phi_ib) ds with n the
 *     unit normal pointing out of the tet; D has 3*dimB rows, row d + 3k
 *     pairing normal component d with output component k
 *     (int_face.h:32-37, int_face.inl:46-82, core.h:176-195).
 * The triangle rule of `order` (0..20: symmetric rules up to 6, collapsed
 * Gauss-Jacobi products above) on the face; per-point coefficients are
 * given at the points anifem_face_quadrature returns.  Layouts, errors and
 * precisions as anifem_fem3dtet_batch, FP64 and FP32 only. */
int anifem_fem3dface_batch(anifem_op opA, anifem_op opB, const anifem_tensor* D, int order, int face, int dot_normal,
                           int64_t n_tets, const double* xyz, int precision, void* A_out, void* stream);
/* Batched fem3DapplyL<Op>(XYZ, XYL, dofs, opU) (anifem++/fem/operations/
 * eval.h:13-18, eval.inl:84-100, core.inl:370-405): Op(u) at q barycentric
 * points xyl (host, 4q, the same for every tet) of each tet, u given by its
 * per-tet DOF values dofs (device, n_tets x nfa: dofs[r*nfa + i]).  opU
 * (device) receives n_tets blocks of dim*q values: opU[r*dim*q + k + dim*n]
 * (for a vector space, component k + 3c of GRAD is du_c/dx_k).  Operators:
 * IDEN and GRAD of P0..P3, CR1, B2, B4; IDEN, DIV and CURL of RT0 and ND0. */
int anifem_fem3dapply_batch(anifem_op op, const double* xyl, int q, int64_t n_tets, const double* xyz,
                            const double* dofs, double* opU, void* stream);
/* Batched fem3DapplyX<Op>(XYZ, X, dofs, opU) (eval.h:26-31, eval.inl:103-125):
 * Op(u) at q physical points per tet, X (device, n_tets x q x 3:
 * X[(r*q + n)*3 + c]); their barycentric coordinates are formed on the device
 * from the inverse Jacobian as getBaryCoord does (core.inl:120-137).  Other
 * arguments, layouts and operators as anifem_fem3dapply_batch. */
int anifem_fem3dapplyx_batch(anifem_op op, int q, int64_t n_tets, const double* xyz, const double* X,
                             const double* dofs, double* opU, void* stream);
/* Triangle rule of `order` placed on face `face`, as tet barycentric points
 * p[4q] (coordinate of the opposite vertex 0) and weights w[q] summing to 1;
 * returns q (or <0). */
int anifem_face_quadrature(int order, int face, double* p, double* w);

/* Number of basis functions (Nfa) and operator dimension (Dim) of an operator
 * (the Nfa/Dim typedefs of poly_*.h and operators.h:127-155).  <0 if invalid. */
int anifem_op_nfa(anifem_op op);
int anifem_op_dim(anifem_op op);
/* Number of quadrature points of the rule used for `order` (<0 if invalid). */
int anifem_quadrature_npoints(int order);
/* Quadrature rule (barycentric p[4q], weights w[q] summing to 1), host arrays. */
int anifem_quadrature(int order, double* p, double* w);

/* Physical quadrature points X_n = P0 + sum_l lambda_{l+1,n} (P_{l+1} - P0)
 * of the volume rule `order` (the anifem_quadrature points) in every tet, on
 * the device: the reference's internalFem3DtetGeomInit_XYG (core.inl:250-270),
 * where it evaluates the coefficient functor (diff_tensor.h:147-153).
 *   xyz : device, 12 x n_tets SoA as anifem_fem3dtet_batch
 *   X   : device, 3 x q x n_tets: coordinate c of point n of tet r at
 *         X[(c*q + n)*n_tets + r], the layout of PER_POINT coefficients, so
 *         that a coefficient evaluated at X component-wise is a PER_POINT
 *         tensor value without any reordering. */
int anifem_quadrature_points_batch(int order, int64_t n_tets, const double* xyz, double* X, void* stream);

/* ---- whole-mesh assembly --------------------------------------------------
 * The reference assembles cell by cell through a virtual plugin
 * (MatFuncWrap::operator(), inmost_interface/func_wrap.h:96-187) called from
 * AssemblerT::Assemble (assembler.inl:314-488).  A per-cell virtual call
 * cannot feed a GPU, so the drop-in sits at the whole-mesh call:
 * anifem_assembly_create ~ PrepareProblem (assembler.inl:194-275) and
 * anifem_assembly_run ~ Assemble(matrix, rhs) with the element evaluator
 * being fem3Dtet<OpA, OpB> on every tet plus the load vector
 * fem3Dtet<IDEN P0, IDEN test> (examples/tutorials/ex1.cpp:92-106). */
typedef struct anifem_assembly anifem_assembly;

typedef struct anifem_mesh_desc {
    int64_t n_nodes;
    int64_t n_tets;
    const double* coords; /* host, n_nodes x 3 */
    const int32_t* tets;  /* host, n_tets x 4 node ids (cell order) */
} anifem_mesh_desc;

enum { ANIFEM_DIRICHLET_NONE = 0, ANIFEM_DIRICHLET_BOUNDARY = 1, ANIFEM_DIRICHLET_DOF_MASK = 2 };

typedef struct anifem_problem_desc {
    anifem_op trial;        /* OpA: the unknown's operator, e.g. GRAD FemFix<P2> */
    anifem_op test;         /* OpB: must live on the same space as trial */
    int32_t order;          /* quadrature order (0..20) */
    int32_t tensor_type;    /* ANIFEM_TENSOR_* of D */
    int32_t tensor_storage; /* ANIFEM_COEF_* of D */
    int32_t rhs_order;      /* quadrature order of the load vector */
    int32_t rhs_tensor_type;    /* ANIFEM_TENSOR_SCALAR/NULL (or GENERAL ncomp x 1) */
    int32_t rhs_tensor_storage; /* ANIFEM_COEF_* of f */
    int32_t dirichlet;          /* ANIFEM_DIRICHLET_* (homogeneous values by default) */
    const uint8_t* dirichlet_dof_mask; /* host, ndof flags, for ANIFEM_DIRICHLET_DOF_MASK */
} anifem_problem_desc;

typedef struct anifem_status {
    int32_t code;            /* ANIFEM_OK or ANIFEM_ERR_NONFINITE */
    int32_t reserved;
    int64_t n_nonfinite;     /* local matrix / rhs entries that were NaN/Inf */
    int64_t n_degenerate;    /* tets with a zero or non-finite Jacobian determinant (MES = |det|/6, core.inl:241) */
} anifem_status;

/* Builds on the host, then uploads: node reorder (ordering.inl:8-26), NATURAL
 * DOF enumeration with P3 edge orientation (global_enumerator.cpp:562-777,
 * ordering.inl:137-151, tetdofmap.inl:98-127), CSR pattern with the diagonal
 * and Dirichlet drop rules (assembler.inl:114-136,416) -- for a vector space
 * with component-wise operators (IDEN, GRAD, DUDX..DUDZ) and a NULL or SCALAR
 * tensor also without the entries between components, which are zero in every
 * cell and which the reference's |a| > 1e-100 rule therefore never inserts --
 * the element-to-CSR-slot
 * map and the row blocks of the owner-computes assembly.  `stream` orders the
 * uploads. */
int anifem_assembly_create(anifem_assembly** ctx, const anifem_mesh_desc* mesh, const anifem_problem_desc* prob,
                           void* stream);
/* anifem_assembly_create with explicit row-block budgets (CSR slots, element
 * entries and (row, tet) pairs per block; at most 65535, 65535 and 256).  The
 * default (anifem_assembly_create) chooses them by the entries per pair:
 * 2048 / 2048 / 256 up to 10; up to 20, 4096 / 2048 / 128 for a scalar space
 * (its pass then runs two lanes per pair) and 4096 / 3072 / 256 for a vector
 * one; 4096 / 8192 / 256 above; and 8192 / 8192 (16384 above 20) / 256 when
 * one row alone exceeds those.  Returns ANIFEM_ERR_DIMS when one row alone exceeds the budgets.  No
 * reference counterpart. */
int anifem_assembly_create_budget(anifem_assembly** ctx, const anifem_mesh_desc* mesh,
                                  const anifem_problem_desc* prob, int max_slots, int max_values, int max_pairs,
                                  void* stream);
/* Problems over several variables (the reference's assembler over a list of
 * FemSpace variables, e.g. mixed Poisson: a flux in RT0 and a pressure in
 * P0).  NATURAL numbering VAR, DIM, ELEM_TYPE, ELEM_ID, DOF_ID
 * (global_enumerator.h:396): every DOF of variable v precedes those of
 * variable v+1, and each variable is numbered as it would be alone (vector
 * variables by component).  A tet's local DOFs list the variables in order,
 * each in its own local order.  Such a context assembles element matrices the
 * caller forms over that local order (anifem_assembly_run_elements,
 * anifem_assembly_run_rhs: the AssemblerT::Assemble path with a local
 * function); anifem_assembly_run, anifem_assembly_run_plain and
 * anifem_assembly_eval return ANIFEM_ERR_ARG.  Of prob only dirichlet and
 * dirichlet_dof_mask are read (BOUNDARY flags the boundary-closure DOFs of
 * every variable).  At most 4 variables and 64 local DOFs. */
#define ANIFEM_MAX_VARS 4
typedef struct anifem_vars_desc {
    int32_t n_var;
    anifem_op var[ANIFEM_MAX_VARS]; /* fem and ncomp of each variable (op is not read) */
} anifem_vars_desc;
int anifem_assembly_create_vars(anifem_assembly** ctx, const anifem_mesh_desc* mesh, const anifem_vars_desc* vars,
                                const anifem_problem_desc* prob, void* stream);
/* Host-only variant (inspection of the numbering and pattern). */
int anifem_assembly_create_vars_host(anifem_assembly** ctx, const anifem_mesh_desc* mesh,
                                     const anifem_vars_desc* vars, const anifem_problem_desc* prob);
int anifem_assembly_destroy(anifem_assembly* ctx);
/* Host-only variant: builds the same setup without touching the device
 * (inspection of the enumeration, pattern, slot map and row blocks through
 * the info / copy_setup / copy_blocks calls; run and coords return
 * ANIFEM_ERR_ARG). */
int anifem_assembly_create_host(anifem_assembly** ctx, const anifem_mesh_desc* mesh, const anifem_problem_desc* prob);

/* Sizes: number of DOFs, stored nonzeros, row blocks, local DOFs per tet. */
int anifem_assembly_info(const anifem_assembly* ctx, int64_t* ndof, int64_t* nnz, int64_t* n_blocks,
                         int32_t* nf);
/* Device views of the CSR pattern (row_ptr: ndof+1 int64, col_idx: nnz int32). */
int anifem_assembly_csr(const anifem_assembly* ctx, const int64_t** row_ptr, const int32_t** col_idx);
/* Host copies for inspection: pattern, per-tet global DOFs (n_tets x nf, the
 * reference's local order), slot map (n_tets x nf*nf, entry ib + nf*ia, -1 =
 * dropped), Dirichlet DOF flags.  Any pointer may be NULL. */
int anifem_assembly_copy_setup(const anifem_assembly* ctx, int64_t* row_ptr, int32_t* col_idx, int32_t* dofs,
                               int64_t* slot, uint8_t* dirichlet);
/* Orientation signs of the cells' local DOFs (host, n_tets x nf, +1 / -1):
 * the global basis function of an oriented DOF (RT0 faces, ND0 edges) is
 * sign x the cell's local one (assembler.inl:154-164, ordering.inl:152-172);
 * every sign is +1 for the other spaces. */
int anifem_assembly_copy_signs(const anifem_assembly* ctx, int8_t* sign);
/* Row blocks of the owner-computes assembly.  Block b owns the CSR rows
 * [blk_row0[b], blk_row0[b+1]) and forms every slot value and rhs entry of
 * them exactly once.  Its pairs [blk_p0[b], blk_p0[b+1]) are the (row, incident
 * tet) pairs of its rows, by row then ascending tet: pair p forms row
 * pair_ib[p] & 63 of tet pair_tet[p]'s element matrix and that entry of its
 * load vector; bit 0x80 of pair_ib marks a Dirichlet row (applyDir: the kept
 * diagonal entry is 1, the load-vector entry 0).  The block's element entries
 * form a value list of blk_v0[b+1] - blk_v0[b] entries grouped by CSR slot,
 * each slot's entries in ascending tet order (the reference's element-loop
 * order); pair_pos[p*nf + ia] is the list position of entry (ib, ia), 0xFFFF
 * for an entry the pattern drops.  sstart[s] is the first list position of
 * slot s and pstart[r] the first pair of row r, both relative to the block.
 * No reference counterpart: this replaces the row locks of the threaded
 * reference assembler (assembler.inl:339-340,406,474).  Sizes first (layout),
 * then the host copies (host-only contexts keep the pair lists; any pointer
 * may be NULL). */
typedef struct anifem_block_layout {
    int64_t n_blocks;
    int64_t n_pairs;
    int64_t n_values;         /* element entries kept in the pattern, all blocks */
    int64_t max_block_pairs;
    int64_t max_block_values;
    int64_t max_block_slots;
} anifem_block_layout;
int anifem_assembly_blocks(const anifem_assembly* ctx, anifem_block_layout* out);
int anifem_assembly_copy_blocks(const anifem_assembly* ctx, int32_t* blk_row0, int64_t* blk_p0, int64_t* blk_v0,
                                int32_t* pair_tet, uint8_t* pair_ib, uint16_t* pair_pos, uint16_t* sstart,
                                uint16_t* pstart);
/* The physical quadrature points of the context's tets, in the caller's tet
 * order and the local node order after the orientation reorder
 * (ordering.inl:8-26): the points at which anifem_assembly_run expects
 * PER_POINT coefficients, as the reference evaluates them at XYG of the
 * reordered cell.  Layout as anifem_quadrature_points_batch. */
int anifem_assembly_quadrature_points(const anifem_assembly* ctx, int order, double* X, void* stream);
/* Device node coordinates (n_nodes x 3) used by run; the caller may update
 * them in place (moving mesh) between runs. */
int anifem_assembly_coords(anifem_assembly* ctx, double** coords);

/* Assemble: csr_val (nnz, device FP64) and rhs (ndof, device FP64) are
 * overwritten with  sum_T P^T A_T P  and  sum_T P^T F_T  (assembler.inl:397-481)
 * after element-level Dirichlet (dc_on_dof.h:38-45) on the flagged DOFs.
 * One workgroup per row block forms each slot value (and rhs entry) as the
 * sum of its element contributions in ascending tet order, the order of the
 * reference's element loop; there are no atomics on values and repeated runs
 * are bitwise identical.
 * D and F use the problem's tensor kinds; F may be NULL (the load vectors
 * are then zero; rhs may be NULL too).  dir_values (device FP64, ndof, may be
 * NULL = homogeneous) gives the value of every Dirichlet DOF (the others are
 * ignored); each cell then applies applyDir with those values
 * (dc_on_dof.h:38-45, per cell as ex1.cpp:100-105): its load vector loses
 * A(:, k) bc_k for each Dirichlet local DOF k and takes bc_k on row k, so a
 * Dirichlet row's rhs is (incident tets) x bc_k.  The matrix does not depend
 * on the values.  Values need the row-block scatter (ANIFEM_ERR_ARG with the
 * coloured one).  If status is non-NULL the call synchronizes the stream and
 * fills it; otherwise it returns after enqueueing. */
int anifem_assembly_run(anifem_assembly* ctx, const anifem_tensor* D, const anifem_tensor* F,
                        const double* dir_values, int precision, double* csr_val, double* rhs,
                        anifem_status* status, void* stream);

/* anifem_assembly_run without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The geometry rows of the pass, which anifem_assembly_run keeps in a buffer
 * the context allocates on its first run, live in a fixed scratch array of
 * the library instead: 128 MiB (134,217,728 bytes) per device, held by every
 * device the library runs on until the process ends.  A problem needs
 * n_tets x rowlen x (8 in FP64, 4 in FP32) bytes, rowlen being the tensor
 * values per tet padded to 16 bytes (GRAD-GRAD of a scalar space with a load:
 * 64 / 32 bytes per tet in FP64 / FP32 for a symmetric tensor, whose K is
 * kept packed, so up to 2,097,152 / 4,194,304 tets; 80 / 48 bytes for a
 * general one);
 * a larger one returns ANIFEM_ERR_DIMS before anything is launched.  The
 * reference tensors of the problem are built at anifem_assembly_create; the
 * coloured scatter is refused (ANIFEM_ERR_ARG).  Arguments, results and
 * status as anifem_assembly_run (the status read-back synchronizes the
 * stream, as there).  All calls on one device share its scratch array:
 * calls that overlap in time on one device (other streams, other threads)
 * race and must be serialized by the caller. */
int anifem_assembly_run_noalloc(anifem_assembly* ctx, const anifem_tensor* D, const anifem_tensor* F,
                                const double* dir_values, int precision, double* csr_val, double* rhs,
                                anifem_status* status, void* stream);

/* Row-block kernel choice (default 1): a persistent pass (each workgroup
 * walks many blocks and prefetches the next one's data) where it applies
 * (scalar spaces, tensor constant over each tet), in its fixed-size form for
 * P1 / P2 / P3; 2 takes the persistent pass in its general form (any element
 * size); 0 forces one workgroup per block.  All sum in the same order: the
 * results are bitwise identical.  No reference counterpart. */
int anifem_assembly_set_persistent(anifem_assembly* ctx, int on);

/* Scatter method of anifem_assembly_run (no reference counterpart: the
 * reference scatters with one thread, assembler.inl:381-481).
 *   ANIFEM_SCATTER_ROW_BLOCKS (default): owner-computes row blocks, each value
 *     written once, sums in the reference's ascending tet order.
 *   ANIFEM_SCATTER_COLOURED: first-fit greedy colouring of the tets (no two
 *     tets of a colour share a node), one launch per colour adding each entry
 *     to its slot; deterministic (colour order), no atomics; values equal to
 *     the row blocks' up to the order of the additions.
 * The colouring is built on the first switch to COLOURED. */
enum { ANIFEM_SCATTER_ROW_BLOCKS = 0, ANIFEM_SCATTER_COLOURED = 1 };
int anifem_assembly_set_scatter(anifem_assembly* ctx, int method);
/* the colouring (after set_scatter(COLOURED)): count and, if colour is not
 * NULL, the colour of each tet (host, n_tets) */
int anifem_assembly_colours(const anifem_assembly* ctx, int32_t* n_colours, int32_t* colour);

/* Assembly of caller-supplied element matrices, the drop-in for
 * AssemblerT::Assemble with an arbitrary local function (the MatFuncWrap
 * plugin, func_wrap.h:96-187, whose result is the dense column-major local
 * matrix plus the local load vector, func_wrap.inl:138-141), evaluated for all
 * cells beforehand.  A_elem: device, n_tets matrices of nf x nf, column-major
 * (entry (ib, ia) at A_elem[t*nf*nf + ib + nf*ia]); F_elem: device, n_tets
 * vectors of nf, or NULL (rhs is then zeroed); both in the caller's tet order
 * and in the local DOF order of the tet's nodes as anifem_assembly_tets
 * returns them (after the orientation reorder, ordering.inl:8-26, as the
 * reference hands the reordered nodes to the local function), in the cell's
 * local basis: on oriented DOFs (RT0 faces, ND0 edges) the scatter multiplies
 * entry (ib, ia) by s_ib s_ia and load entry ib by s_ib, the signs of
 * anifem_assembly_copy_signs (assembler.inl:154-164).  The values go into the
 * context's pattern: for a vector space whose problem has component-wise
 * operators and a NULL or SCALAR tensor, the pattern leaves out the entries
 * between components (zero for such a problem; the reference's value rule,
 * assembler.inl:416, never inserts them), so matrices that couple components
 * need a context whose tensor couples them.  elem_precision:
 * ANIFEM_PREC_FP64 (double) or ANIFEM_PREC_FP32 (float) element values; the
 * CSR and rhs accumulate in FP64, with element-level homogeneous Dirichlet and
 * the same deterministic order as anifem_assembly_run. */
int anifem_assembly_run_elements(anifem_assembly* ctx, const void* A_elem, const void* F_elem, int elem_precision,
                                 double* csr_val, double* rhs, anifem_status* status, void* stream);
/* Flag for elem_precision: the element matrices are symmetric (the caller's
 * promise); each row is read as the matching column, contiguous in the
 * column-major layout, i.e. the scatter uses A(ia, ib) for A(ib, ia). */
enum { ANIFEM_ELEM_SYMMETRIC = 0x100 };
/* AssembleRHS (anifem++/inmost_interface/assembler.inl:497-561): the global
 * right-hand side from per-cell load vectors F_elem (n_tets x nf, FP64 or FP32,
 * local DOF order after the node reorder), with element-level homogeneous
 * Dirichlet (a Dirichlet row receives 0).  accumulate = 0 stores the sum,
 * accumulate = 1 adds it: rhs <- rhs + assembled, as the reference.  Each row
 * sums its cells' entries in ascending cell order (deterministic; no atomics).
 * Boundary load terms (Neumann) enter as the face integrals of
 * anifem_fem3dface_batch added into the load vector of the face's cell. */
int anifem_assembly_run_rhs(anifem_assembly* ctx, const void* F_elem, int elem_precision, double* rhs, int accumulate,
                            anifem_status* status, void* stream);
/* Evaluation of a global FE vector u (device, ndof; e.g. the PCG solution) in
 * every cell: op = ANIFEM_IDEN or ANIFEM_GRAD of the context's trial space at
 * q barycentric points xyl (host, 4q) of each cell, layout as
 * anifem_fem3dapply_batch, caller's cell order.  The cell's DOF values are
 * gathered through the context's DOF map (NATURAL numbering, P3 edge
 * orientation), the step after the solve of examples/tutorials/ex1.cpp:167-168. */
int anifem_assembly_eval(anifem_assembly* ctx, int op, const double* xyl, int q, const double* u, double* opU,
                         void* stream);
/* Host copy of the tets' node lists after the orientation reorder, caller's
 * tet order (n_tets x 4). */
int anifem_assembly_tets(const anifem_assembly* ctx, int32_t* tets);

/* The plain FP64 method as a measurement baseline: one lane per tet follows
 * the reference's algorithm step by step (U = PSI^T Ghat and DU = w|T| D U at
 * each quadrature point, A = V^T DU; core.inl:218-367), element-level
 * Dirichlet, then FP64 atomic adds through the element-to-slot map
 * (assembler.inl:397-481).  Same outputs as anifem_assembly_run up to the
 * (unfixed) order of the atomic adds.  Scalar GRAD x GRAD problems only. */
int anifem_assembly_run_plain(anifem_assembly* ctx, const anifem_tensor* D, const anifem_tensor* F, double* csr_val,
                              double* rhs, anifem_status* status, void* stream);

/* Per-phase GPU timers of the last run, the counterpart of the assembler
 * timers of the reference (assembler.h:234-248, GetTimeEvalLocFunc and
 * friends).  When enabled, run records HIP events on its stream around each
 * phase and synchronizes at its end.  Times in milliseconds; launch counts. */
typedef struct anifem_timers {
    double ms_geometry; /* per-tet Jacobians and geometry tensors (matrix and rhs) */
    double ms_gather;   /* row blocks: element entries formed and summed into csr_val / rhs */
    int32_t n_launches; /* kernel launches of the run */
    int32_t reserved;
} anifem_timers;
int anifem_assembly_enable_timers(anifem_assembly* ctx, int on);
int anifem_assembly_timers(const anifem_assembly* ctx, anifem_timers* out);

/* ---- batched assembly of many meshes --------------------------------------
 * No reference counterpart (the reference assembles one mesh per Assembler).
 * Each mesh is scalar diffusion -div(kappa grad u) = f: the element matrices
 * fem3Dtet<GRAD FemFix<Pk>, GRAD FemFix<Pk>, DfuncTraits<TENSOR_SCALAR, false>>
 * with quadrature order 2 (P1, P2) or 4 (P3), load vectors of a constant f,
 * homogeneous Dirichlet on the boundary (or none).  All meshes are assembled
 * by one geometry launch and one row-block launch, captured in a HIP graph:
 * mesh i's CSR values occupy val[nnz_offset_i ..] and its rhs
 * rhs[dof_offset_i ..], bitwise identical to assembling it alone with
 * anifem_assembly_run.  With autotune, the row-block budgets are chosen per
 * element order by timing candidates on the GPU (anifem_batch_tuning). */
typedef struct anifem_batch anifem_batch;
typedef struct anifem_batch_mesh {
    int64_t n_nodes;
    int64_t n_tets;
    const double* coords; /* host, n_nodes x 3 */
    const int32_t* tets;  /* host, n_tets x 4 */
    int32_t fem;          /* ANIFEM_FEM_P1 .. ANIFEM_FEM_P3 */
    int32_t reserved;
} anifem_batch_mesh;
int anifem_batch_create(anifem_batch** ctx, int n_meshes, const anifem_batch_mesh* meshes, int dirichlet,
                        int autotune, void* stream);
int anifem_batch_destroy(anifem_batch* ctx);
int anifem_batch_totals(const anifem_batch* ctx, int64_t* n_tets, int64_t* ndof, int64_t* nnz, int64_t* n_blocks);
int anifem_batch_info(const anifem_batch* ctx, int mesh, int64_t* tet_offset, int64_t* dof_offset,
                      int64_t* nnz_offset, int64_t* ndof, int64_t* nnz);
/* device views of mesh i's own CSR pattern (row_ptr starts at 0) */
int anifem_batch_csr(const anifem_batch* ctx, int mesh, const int64_t** row_ptr, const int32_t** col_idx);
/* chosen budgets {slots, entries, pairs} per block and the times of the
 * ANIFEM_BATCH_BUDGET_CANDIDATES candidates (ms[] of that length; -1: not
 * applicable), in the order {2048, 1024, 128}, {2048, 2048, 256},
 * {4096, 4096, 256}, {1024, 1024, 256}, {4096, 3072, 256}, {4096, 2560, 256},
 * {4096, 2560, 128}, {4096, 2048, 128} */
#define ANIFEM_BATCH_BUDGET_CANDIDATES 8
int anifem_batch_tuning(const anifem_batch* ctx, int fem, int32_t* caps, double* ms);
/* Launch shape of element order fem's row-block pass: shape = {phase-A form
 * (0 general, 1 fixed-size for P1 / P2, 2 two lanes per pair for P3 blocks of
 * at most 128 pairs), compact value positions, workgroups per CU (0: the
 * occupancy)}; with autotune, timed at the first anifem_batch_run on the
 * batch's own blocks over the order's forms, both position lists where the
 * order has compact ones (form 2 takes 16-bit ones), and the occupancy and the
 * two counts below it; ms: its time per launch (-1 until timed).  No
 * counterpart in the reference. */
int anifem_batch_launch_tuning(const anifem_batch* ctx, int fem, int32_t* shape, double* ms);
/* Row-block kernel choice of the batch's graph (default 1: the persistent
 * pass, element kinds restaged per workgroup range; 0: one workgroup per
 * block).  Both sum in the same order: the results are bitwise identical. */
int anifem_batch_set_persistent(anifem_batch* b, int on);
/* kappa: device, all tets in mesh order; f: device, one FP64 value; val (total
 * nnz) and rhs (total ndof): device FP64, overwritten */
int anifem_batch_run(anifem_batch* ctx, const double* kappa, const double* f, double* val, double* rhs,
                     anifem_status* status, void* stream);

/* ---- solver ---------------------------------------------------------------
 * Jacobi-preconditioned CG on a device CSR matrix in FP64 (the reference
 * solves with INMOST's Solver, examples/tutorials/ex1.cpp:161-163; there is
 * no CG in the reference).  Stops when ||b - A x||_2 <= rtol ||b||_2 measured
 * on the recursively updated residual; true_relres receives the FP64
 * residual recomputed from x at exit.  x holds the initial guess. */
typedef struct anifem_pcg_result {
    int64_t iterations;
    double relres;      /* recursive residual at exit */
    double true_relres; /* ||b - A x|| / ||b|| recomputed at exit */
} anifem_pcg_result;

int anifem_pcg_jacobi(int64_t n, const int64_t* row_ptr, const int32_t* col_idx, const double* val,
                      const double* b, double* x, double rtol, int64_t maxit, anifem_pcg_result* res,
                      void* stream);

/* The solvers keep their device workspaces per host thread and device
 * (grown on demand, reused by later calls, so a solve allocates nothing once
 * its sizes have been seen); this frees the calling thread's workspaces. */
int anifem_pcg_release(void);

/* Flexible PCG (Polak-Ribiere beta) with a factored preconditioner
 * M^-1 = G G^T: G and GT are device FP64 values of G and of G^T on the CSR
 * pattern of A (which must be symmetric).  Stopping rule and reporting as for
 * anifem_pcg_jacobi.  No reference counterpart. */
int anifem_pcg_factored(int64_t n, const int64_t* row_ptr, const int32_t* col_idx, const double* val,
                        const double* G, const double* GT, const double* b, double* x, double rtol, int64_t maxit,
                        anifem_pcg_result* res, void* stream);

/* Flexible PCG (Polak-Ribiere beta) with an explicit SPD preconditioner
 * matrix M (z = M r; device CSR
 * m_row_ptr / m_col_idx / m_val, any pattern), for example M = G G^T from
 * anifem_gnn_precond.  Stopping rule and reporting as for anifem_pcg_jacobi.
 * No reference counterpart. */
int anifem_pcg_matrix(int64_t n, const int64_t* row_ptr, const int32_t* col_idx, const double* val,
                      const int64_t* m_row_ptr, const int32_t* m_col_idx, const double* m_val, const double* b,
                      double* x, double rtol, int64_t maxit, anifem_pcg_result* res, void* stream);

/* ---- learned preconditioner ----------------------------------------------
 * No reference counterpart.  A 3-layer message-passing network (32 hidden
 * units; BF16 weights, stored as 16-bit words and widened in the kernels, with
 * FP32 arithmetic and accumulation on VALU; FP32 master weights and Adam
 * moments) reads an assembled SPD matrix A and
 * writes a lower-triangular factor G on A's pattern; M^-1 = G G^T.  With the
 * output heads at zero the network is exactly Jacobi (G = diag(a_ii)^-1/2).
 * Training minimises (1/(P n)) sum_p ||G^T A G z_p - z_p||^2 over P = 4
 * Rademacher probes, with the network's own forward / backward kernels and
 * Adam on the GPU.  Parameters: a flat FP32 vector of anifem_gnn_n_params()
 * values in the order W_in(32x4), b_in, then per layer W_s(32x32), W_m(32x32),
 * b, then w_o, c_o, w_d, c_d (row-major, out x in). */
typedef struct anifem_gnn anifem_gnn;
int anifem_gnn_n_params(void);
/* row_ptr / col_idx: host CSR pattern of the matrices the context will see */
int anifem_gnn_create(anifem_gnn** ctx, int64_t n, const int64_t* row_ptr, const int32_t* col_idx, void* stream);
int anifem_gnn_destroy(anifem_gnn* ctx);
int anifem_gnn_set_params(anifem_gnn* ctx, const float* theta);                   /* host; resets Adam */
int anifem_gnn_get_params(const anifem_gnn* ctx, float* theta, float* theta_bf16); /* host, either may be NULL */
int anifem_gnn_get_grad(const anifem_gnn* ctx, float* grad);                       /* host, last loss_grad */
/* G (and G^T) of the matrix with device values a; G/GT NULL: kept in the context */
int anifem_gnn_factor(anifem_gnn* ctx, const double* a, double* G, double* GT, void* stream);
/* device views of the context's pattern and last factor */
int anifem_gnn_views(const anifem_gnn* ctx, int64_t* n, const int64_t** row_ptr, const int32_t** col_idx,
                     const double** G, const double** GT);
/* M = G G^T of the last factor kept in the context (anifem_gnn_factor with
 * G = NULL), as a device CSR matrix owned by the context.  Its pattern and
 * term lists are built by anifem_gnn_create; ANIFEM_ERR_NOMEM when M has more
 * than 2^30 terms and was not built. */
int anifem_gnn_precond(anifem_gnn* ctx, int64_t* m_nnz, const int64_t** m_row_ptr, const int32_t** m_col_idx,
                       const double** m_val, void* stream);
/* forward + loss + gradient for matrix values a; probes: device n x 4 FP32
 * (row-major) or NULL for Rademacher probes from seed; loss (host) may be NULL */
int anifem_gnn_loss_grad(anifem_gnn* ctx, const double* a, const float* probes, uint64_t seed, float* loss,
                         void* stream);
int anifem_gnn_adam_step(anifem_gnn* ctx, float lr, float beta1, float beta2, float eps, void* stream);

typedef struct anifem_pcg_learned_result {
    int64_t iterations;          /* of the accepted solve */
    double relres;
    double true_relres;          /* FP64 ||b - A x|| / ||b|| of the returned x */
    int64_t learned_iterations;  /* of the learned-preconditioned attempt */
    double learned_true_relres;
    int32_t fallback;            /* 1: Jacobi-PCG was used */
    int32_t reserved;
} anifem_pcg_learned_result;
/* The network's factor G for A (values val on the context's pattern), then
 * flexible PCG with M = G G^T formed explicitly (anifem_pcg_matrix; with G
 * and G^T applied in turn, anifem_pcg_factored, when M was not built); when its FP64 true residual exceeds rtol
 * (1% slack for the recursive-residual drift) or it fails, Jacobi-PCG from the
 * same initial guess. */
int anifem_pcg_learned(anifem_gnn* ctx, const double* val, const double* b, double* x, double rtol, int64_t maxit,
                       anifem_pcg_learned_result* res, void* stream);

/* Gaussian random field at n points (device xyz, n x 3): random Fourier
 * features for the covariance sigma^2 exp(-|x - y|^2 / (2 ell^2)),
 * g(x) = sigma sqrt(2/M) sum_m cos(w_m . x + phi_m), w_m ~ N(0, I/ell^2),
 * phi_m ~ U[0, 2 pi) drawn from seed; out = g, or exp(g) with lognormal. */
int anifem_grf_sample(int64_t n, const double* xyz, double ell, double sigma, int n_features, uint64_t seed,
                      int lognormal, double* out, void* stream);

/* ---- synthetic meshes (host) ----------------------------------------------
 * Unit cube split into n^3 sub-cubes, each into 6 Kuhn tets; node
 * (i,j,k) has id (i*(n+1)+j)*(n+1)+k (the V_ID of utils/mesh_utils.cpp:124).
 * Interior nodes are moved by i.i.d. U[-jitter*h, jitter*h] per coordinate from
 * a Philox-4x32-10 stream keyed by seed.  coords: (n+1)^3 x 3, tets: 6 n^3 x 4. */
int anifem_mesh_kuhn_cube(int64_t n, double jitter, uint64_t seed, double* coords, int32_t* tets);

/* Library version string. */
const char* anifem_version(void);

#ifdef __cplusplus
}
#endif
#endif /* ANIFEM_GPU_H */
