"""Python access to libanifem_gpu.so (the MI355X element-matrix / assembly
library) through its C ABI (include/anifem_gpu.h).

PyTorch is used only for device memory and streams: tensors passed in must be
contiguous, on the current HIP device, and of the dtype documented per call.
The library is loaded from this package's directory; importing the GPU entry
points when it has not been built raises immediately -- there is no CPU
fallback.

Names mirror the reference interface: Operator/FemFix/FemVec, TensorType,
fem3Dtet, Assembler (anifem++/fem/operators.h, diff_tensor.h,
operations/int_tet.h, inmost_interface/assembler.h).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libanifem_gpu.so")

# ---- enums (include/anifem_gpu.h) -----------------------------------------
OK, ERR_NONFINITE, ERR_CONNECTIVITY, ERR_ARG, ERR_ORDER, ERR_DIMS, ERR_HIP, ERR_NOMEM, ERR_NOT_CONVERGED = (
    0, -1, -2, -3, -4, -5, -6, -7, -8)
IDEN, GRAD, DIV, CURL, DUDX, DUDY, DUDZ = 1, 2, 3, 4, 6, 7, 8
ELEM_SYMMETRIC = 0x100
FEM_P0, FEM_P1, FEM_P2, FEM_P3 = 1, 2, 3, 4
# element-level spaces (fem3dtet_batch, fem3dface_batch; not the global assembly)
FEM_RT0, FEM_ND0, FEM_CR1, FEM_B2, FEM_B4 = 21, 25, 31, 1002, 1004
TENSOR_NULL, TENSOR_SCALAR, TENSOR_SYMMETRIC, TENSOR_GENERAL, TENSOR_LAME = 1, 2, 3, 4, 101
COEF_CONST, COEF_PER_TET, COEF_PER_POINT = 0, 1, 2
PREC_FP64, PREC_FP32, PREC_FP16X3 = 0, 1, 2
SCATTER_ROW_BLOCKS, SCATTER_COLOURED = 0, 1
DIRICHLET_NONE, DIRICHLET_BOUNDARY, DIRICHLET_DOF_MASK = 0, 1, 2

_ERRNAMES = {ERR_NONFINITE: "non-finite local entry", ERR_CONNECTIVITY: "broken connectivity",
             ERR_ARG: "invalid argument", ERR_ORDER: "unsupported quadrature order",
             ERR_DIMS: "identity/scalar tensor needs equal operator dimensions", ERR_HIP: "HIP runtime error",
             ERR_NOMEM: "out of memory", ERR_NOT_CONVERGED: "solver did not converge"}


class AnifemError(RuntimeError):
    def __init__(self, code, what=""):
        super().__init__(f"{what}: {_ERRNAMES.get(code, 'error')} (code {code})")
        self.code = code


class anifem_op(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int32), ("fem", ctypes.c_int32), ("ncomp", ctypes.c_int32)]


class anifem_tensor(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("storage", ctypes.c_int32), ("data", ctypes.c_void_p)]


class anifem_mesh_desc(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int64), ("n_tets", ctypes.c_int64),
                ("coords", ctypes.c_void_p), ("tets", ctypes.c_void_p)]


class anifem_problem_desc(ctypes.Structure):
    _fields_ = [("trial", anifem_op), ("test", anifem_op), ("order", ctypes.c_int32),
                ("tensor_type", ctypes.c_int32), ("tensor_storage", ctypes.c_int32),
                ("rhs_order", ctypes.c_int32), ("rhs_tensor_type", ctypes.c_int32),
                ("rhs_tensor_storage", ctypes.c_int32), ("dirichlet", ctypes.c_int32),
                ("dirichlet_dof_mask", ctypes.c_void_p)]


class anifem_vars_desc(ctypes.Structure):
    _fields_ = [("n_var", ctypes.c_int32), ("var", anifem_op * 4)]


class anifem_status(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("n_nonfinite", ctypes.c_int64), ("n_degenerate", ctypes.c_int64)]


class anifem_timers(ctypes.Structure):
    _fields_ = [("ms_geometry", ctypes.c_double), ("ms_gather", ctypes.c_double),
                ("n_launches", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class anifem_block_layout(ctypes.Structure):
    _fields_ = [("n_blocks", ctypes.c_int64), ("n_pairs", ctypes.c_int64), ("n_values", ctypes.c_int64),
                ("max_block_pairs", ctypes.c_int64), ("max_block_values", ctypes.c_int64),
                ("max_block_slots", ctypes.c_int64)]


class anifem_pcg_result(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("relres", ctypes.c_double), ("true_relres", ctypes.c_double)]


class anifem_batch_mesh(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int64), ("n_tets", ctypes.c_int64), ("coords", ctypes.c_void_p),
                ("tets", ctypes.c_void_p), ("fem", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class anifem_pcg_learned_result(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("relres", ctypes.c_double), ("true_relres", ctypes.c_double),
                ("learned_iterations", ctypes.c_int64), ("learned_true_relres", ctypes.c_double),
                ("fallback", ctypes.c_int32), ("reserved", ctypes.c_int32)]


# Every symbol include/anifem_gpu.h declares, with its ctypes signature.
_VP, _I32, _I64, _D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
SIGNATURES = {
    "anifem_fem3dtet_batch": (_I32, [anifem_op, anifem_op, ctypes.POINTER(anifem_tensor), _I32, _I64, _VP, _I32, _VP,
                                     _VP]),
    "anifem_op_nfa": (_I32, [anifem_op]),
    "anifem_op_dim": (_I32, [anifem_op]),
    "anifem_quadrature_npoints": (_I32, [_I32]),
    "anifem_quadrature": (_I32, [_I32, _VP, _VP]),
    "anifem_quadrature_points_batch": (_I32, [_I32, _I64, _VP, _VP, _VP]),
    "anifem_assembly_quadrature_points": (_I32, [_VP, _I32, _VP, _VP]),
    "anifem_fem3dface_batch": (_I32, [anifem_op, anifem_op, ctypes.POINTER(anifem_tensor), _I32, _I32, _I32, _I64,
                                      _VP, _I32, _VP, _VP]),
    "anifem_face_quadrature": (_I32, [_I32, _I32, _VP, _VP]),
    "anifem_fem3dapply_batch": (_I32, [anifem_op, _VP, _I32, _I64, _VP, _VP, _VP, _VP]),
    "anifem_fem3dapplyx_batch": (_I32, [anifem_op, _I32, _I64, _VP, _VP, _VP, _VP, _VP]),
    "anifem_assembly_eval": (_I32, [_VP, _I32, _VP, _I32, _VP, _VP, _VP]),
    "anifem_assembly_create_budget": (_I32, [ctypes.POINTER(_VP), ctypes.POINTER(anifem_mesh_desc),
                                             ctypes.POINTER(anifem_problem_desc), _I32, _I32, _I32, _VP]),
    "anifem_assembly_create": (_I32, [ctypes.POINTER(_VP), ctypes.POINTER(anifem_mesh_desc),
                                      ctypes.POINTER(anifem_problem_desc), _VP]),
    "anifem_assembly_create_host": (_I32, [ctypes.POINTER(_VP), ctypes.POINTER(anifem_mesh_desc),
                                           ctypes.POINTER(anifem_problem_desc)]),
    "anifem_assembly_create_vars": (_I32, [ctypes.POINTER(_VP), ctypes.POINTER(anifem_mesh_desc),
                                           ctypes.POINTER(anifem_vars_desc), ctypes.POINTER(anifem_problem_desc),
                                           _VP]),
    "anifem_assembly_create_vars_host": (_I32, [ctypes.POINTER(_VP), ctypes.POINTER(anifem_mesh_desc),
                                                ctypes.POINTER(anifem_vars_desc),
                                                ctypes.POINTER(anifem_problem_desc)]),
    "anifem_assembly_destroy": (_I32, [_VP]),
    "anifem_assembly_info": (_I32, [_VP, _VP, _VP, _VP, _VP]),
    "anifem_assembly_csr": (_I32, [_VP, _VP, _VP]),
    "anifem_assembly_copy_setup": (_I32, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "anifem_assembly_copy_signs": (_I32, [_VP, _VP]),
    "anifem_assembly_blocks": (_I32, [_VP, ctypes.POINTER(anifem_block_layout)]),
    "anifem_assembly_copy_blocks": (_I32, [_VP] + [_VP] * 8),
    "anifem_assembly_coords": (_I32, [_VP, _VP]),
    "anifem_assembly_run": (_I32, [_VP, ctypes.POINTER(anifem_tensor), ctypes.POINTER(anifem_tensor), _VP, _I32, _VP,
                                   _VP, ctypes.POINTER(anifem_status), _VP]),
    "anifem_assembly_run_noalloc": (_I32, [_VP, ctypes.POINTER(anifem_tensor), ctypes.POINTER(anifem_tensor), _VP,
                                           _I32, _VP, _VP, ctypes.POINTER(anifem_status), _VP]),
    "anifem_assembly_run_elements": (_I32, [_VP, _VP, _VP, _I32, _VP, _VP, ctypes.POINTER(anifem_status), _VP]),
    "anifem_assembly_tets": (_I32, [_VP, _VP]),
    "anifem_assembly_run_rhs": (_I32, [_VP, _VP, _I32, _VP, _I32, ctypes.POINTER(anifem_status), _VP]),
    "anifem_assembly_set_scatter": (_I32, [_VP, _I32]),
    "anifem_assembly_colours": (_I32, [_VP, _VP, _VP]),
    "anifem_assembly_run_plain": (_I32, [_VP, ctypes.POINTER(anifem_tensor), ctypes.POINTER(anifem_tensor), _VP, _VP,
                                         ctypes.POINTER(anifem_status), _VP]),
    "anifem_assembly_set_persistent": (_I32, [_VP, _I32]),
    "anifem_assembly_enable_timers": (_I32, [_VP, _I32]),
    "anifem_assembly_timers": (_I32, [_VP, ctypes.POINTER(anifem_timers)]),
    "anifem_pcg_jacobi": (_I32, [_I64, _VP, _VP, _VP, _VP, _VP, _D, _I64, ctypes.POINTER(anifem_pcg_result), _VP]),
    "anifem_batch_create": (_I32, [ctypes.POINTER(_VP), _I32, ctypes.POINTER(anifem_batch_mesh), _I32, _I32, _VP]),
    "anifem_batch_destroy": (_I32, [_VP]),
    "anifem_batch_totals": (_I32, [_VP, _VP, _VP, _VP, _VP]),
    "anifem_batch_info": (_I32, [_VP, _I32, _VP, _VP, _VP, _VP, _VP]),
    "anifem_batch_csr": (_I32, [_VP, _I32, _VP, _VP]),
    "anifem_batch_tuning": (_I32, [_VP, _I32, _VP, _VP]),
    "anifem_batch_launch_tuning": (_I32, [_VP, _I32, _VP, _VP]),
    "anifem_batch_set_persistent": (_I32, [_VP, _I32]),
    "anifem_batch_run": (_I32, [_VP, _VP, _VP, _VP, _VP, ctypes.POINTER(anifem_status), _VP]),
    "anifem_pcg_factored": (_I32, [_I64, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _D, _I64,
                                   ctypes.POINTER(anifem_pcg_result), _VP]),
    "anifem_pcg_release": (_I32, []),
    "anifem_pcg_matrix": (_I32, [_I64, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _D, _I64,
                                 ctypes.POINTER(anifem_pcg_result), _VP]),
    "anifem_gnn_n_params": (_I32, []),
    "anifem_gnn_create": (_I32, [ctypes.POINTER(_VP), _I64, _VP, _VP, _VP]),
    "anifem_gnn_destroy": (_I32, [_VP]),
    "anifem_gnn_set_params": (_I32, [_VP, _VP]),
    "anifem_gnn_get_params": (_I32, [_VP, _VP, _VP]),
    "anifem_gnn_get_grad": (_I32, [_VP, _VP]),
    "anifem_gnn_factor": (_I32, [_VP, _VP, _VP, _VP, _VP]),
    "anifem_gnn_views": (_I32, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "anifem_gnn_precond": (_I32, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "anifem_gnn_loss_grad": (_I32, [_VP, _VP, _VP, ctypes.c_uint64, _VP, _VP]),
    "anifem_gnn_adam_step": (_I32, [_VP, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _VP]),
    "anifem_pcg_learned": (_I32, [_VP, _VP, _VP, _VP, _D, _I64, ctypes.POINTER(anifem_pcg_learned_result), _VP]),
    "anifem_grf_sample": (_I32, [_I64, _VP, _D, _D, _I32, ctypes.c_uint64, _I32, _VP, _VP]),
    "anifem_mesh_kuhn_cube": (_I32, [_I64, _D, ctypes.c_uint64, _VP, _VP]),
    "anifem_version": (ctypes.c_char_p, []),
}

_lib = None


def lib():
    """The loaded shared library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run `make -C inmost-fem_amd` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc != OK:
        raise AnifemError(rc, what)


def _ptr(a):
    """Address of a numpy array or torch tensor (None -> NULL; an int is an address)."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def _stream_handle(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


# ---- operators (Operator<OP, FemFix<FEM>> / FemVec<n, FEM>) ------------------
@dataclass(frozen=True)
class Op:
    op: int
    fem: int
    ncomp: int = 1

    def c(self):
        return anifem_op(self.op, self.fem, self.ncomp)

    @property
    def nfa(self):
        return lib().anifem_op_nfa(self.c())

    @property
    def dim(self):
        return lib().anifem_op_dim(self.c())


def Operator(op, fem, ncomp=1):
    return Op(op, fem, ncomp)


def tensor_ddim(ttype, dimA, dimB):
    return {TENSOR_NULL: 0, TENSOR_SCALAR: 1, TENSOR_SYMMETRIC: dimB * (dimB + 1) // 2,
            TENSOR_GENERAL: dimA * dimB, TENSOR_LAME: 2}[ttype]


def quadrature(order):
    q = lib().anifem_quadrature_npoints(order)
    if q < 0:
        raise AnifemError(ERR_ORDER, "quadrature")
    p = np.zeros(4 * q)
    w = np.zeros(q)
    rc = lib().anifem_quadrature(order, _ptr(p), _ptr(w))
    if rc < 0:
        raise AnifemError(rc, "quadrature")
    return p.reshape(q, 4), w


# ---- element level -------------------------------------------------------
def fem3dtet_batch(opA: Op, opB: Op, ttype, storage, D, order, xyz, precision=PREC_FP64, out=None, stream=None):
    """Element matrices of n tets (anifem_fem3dtet_batch).

    xyz: device float64 (12, n) SoA; D: device float64 coefficient array in the
    documented layout (or None for TENSOR_NULL); returns out of shape
    (n, nfA, nfB): tet r's nfB x nfA matrix stored column-major, so out[r].T
    is that matrix.
    """
    import torch
    n = xyz.shape[1]
    nfa, nfb = opA.nfa, opB.nfa
    dt = torch.float64 if precision == PREC_FP64 else torch.float32
    if out is None:
        out = torch.empty((n, nfa, nfb), dtype=dt, device=xyz.device)
    t = anifem_tensor(ttype, storage, _ptr(D))
    rc = lib().anifem_fem3dtet_batch(opA.c(), opB.c(), ctypes.byref(t), order, n, _ptr(xyz), precision, _ptr(out),
                                     _stream_handle(stream))
    _check(rc, "anifem_fem3dtet_batch")
    return out


def quadrature_points_batch(order, xyz, out=None, stream=None):
    """Physical points of the volume rule in every tet (device): xyz the
    (12, n) SoA vertices; returns (3, q, n), the PER_POINT coefficient layout."""
    import torch
    n = xyz.shape[1]
    q = len(quadrature(order)[1])
    if out is None:
        out = torch.empty((3, q, n), dtype=torch.float64, device=xyz.device)
    _check(lib().anifem_quadrature_points_batch(order, n, _ptr(xyz), _ptr(out), _stream_handle(stream)),
           "anifem_quadrature_points_batch")
    return out


def face_quadrature(order, face):
    """Triangle rule of `order` on face `face` of the tet: (p (q, 4) barycentric,
    w (q,) summing to 1) -- the points of per-point face coefficients."""
    q = lib().anifem_face_quadrature(order, face, None, None)
    if q < 0:
        raise AnifemError(q, "face_quadrature")
    p = np.zeros(4 * q)
    w = np.zeros(q)
    q = lib().anifem_face_quadrature(order, face, _ptr(p), _ptr(w))
    return p[:4 * q].reshape(q, 4), w[:q]


def fem3dface_batch(opA: Op, opB: Op, ttype, storage, D, order, face, xyz, dot_normal=False, precision=PREC_FP64,
                    out=None, stream=None):
    """Face integrals over face `face` of n tets (anifem_fem3dface_batch):
    fem3Dface, or fem3DfaceN with dot_normal (D then has 3*dimB rows).
    Layouts as fem3dtet_batch."""
    import torch
    n = xyz.shape[1]
    nfa, nfb = opA.nfa, opB.nfa
    dt = torch.float64 if precision == PREC_FP64 else torch.float32
    if out is None:
        out = torch.empty((n, nfa, nfb), dtype=dt, device=xyz.device)
    t = anifem_tensor(ttype, storage, _ptr(D))
    rc = lib().anifem_fem3dface_batch(opA.c(), opB.c(), ctypes.byref(t), order, face, 1 if dot_normal else 0, n,
                                      _ptr(xyz), precision, _ptr(out), _stream_handle(stream))
    _check(rc, "anifem_fem3dface_batch")
    return out


def fem3dapply_batch(op: Op, xyl, xyz, dofs, out=None, stream=None):
    """Op(u) at the barycentric points xyl (q, 4) of n tets with per-tet DOF
    values dofs (device (n, nfa)) (anifem_fem3dapply_batch): (n, q, dim)."""
    import torch
    xyl = np.ascontiguousarray(xyl, dtype=np.float64)
    q, n = len(xyl), xyz.shape[1]
    if op.dim < 0:
        raise AnifemError(ERR_ARG, "anifem_fem3dapply_batch: invalid operator")
    if out is None:
        out = torch.empty((n, q, op.dim), dtype=torch.float64, device=xyz.device)
    _check(lib().anifem_fem3dapply_batch(op.c(), _ptr(xyl), q, n, _ptr(xyz), _ptr(dofs), _ptr(out),
                                         _stream_handle(stream)), "anifem_fem3dapply_batch")
    return out


def fem3dapplyx_batch(op: Op, X, xyz, dofs, out=None, stream=None):
    """Op(u) at q physical points per tet, X device (n, q, 3), with per-tet DOF
    values dofs (device (n, nfa)) (anifem_fem3dapplyx_batch): (n, q, dim)."""
    import torch
    n, q = X.shape[0], X.shape[1]
    if op.dim < 0:
        raise AnifemError(ERR_ARG, "anifem_fem3dapplyx_batch: invalid operator")
    if out is None:
        out = torch.empty((n, q, op.dim), dtype=torch.float64, device=X.device)
    _check(lib().anifem_fem3dapplyx_batch(op.c(), q, n, _ptr(xyz), _ptr(X), _ptr(dofs), _ptr(out),
                                          _stream_handle(stream)), "anifem_fem3dapplyx_batch")
    return out


def write_vtu(path, coords, tets, point_data=None, cell_data=None):
    """Unstructured-grid VTK file (XML .vtu, ASCII) of a tetrahedral mesh with
    per-node and per-cell fields (arrays of n or (n, k) values), the output
    step after the solve (the reference saves through its mesh library,
    examples/tutorials/ex1.cpp:167-168)."""
    coords = np.asarray(coords, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int64)

    def arr(name, a):
        a = np.asarray(a, dtype=np.float64)
        k = 1 if a.ndim == 1 else a.shape[1]
        body = " ".join(repr(float(v)) for v in a.ravel())
        return (f'<DataArray type="Float64" Name="{name}" NumberOfComponents="{k}" format="ascii">'
                f"{body}</DataArray>\n")

    lines = ['<?xml version="1.0"?>\n',
             '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n<UnstructuredGrid>\n',
             f'<Piece NumberOfPoints="{len(coords)}" NumberOfCells="{len(tets)}">\n']
    lines.append("<PointData>\n" + "".join(arr(k, v) for k, v in (point_data or {}).items()) + "</PointData>\n")
    lines.append("<CellData>\n" + "".join(arr(k, v) for k, v in (cell_data or {}).items()) + "</CellData>\n")
    lines.append("<Points>\n" + arr("Points", coords) + "</Points>\n")
    lines.append("<Cells>\n")
    lines.append('<DataArray type="Int64" Name="connectivity" format="ascii">' +
                 " ".join(map(str, tets.ravel())) + "</DataArray>\n")
    lines.append('<DataArray type="Int64" Name="offsets" format="ascii">' +
                 " ".join(map(str, 4 * np.arange(1, len(tets) + 1))) + "</DataArray>\n")
    lines.append('<DataArray type="UInt8" Name="types" format="ascii">' + " ".join(["10"] * len(tets)) +
                 "</DataArray>\n")  # VTK_TETRA
    lines.append("</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")
    with open(path, "w") as fh:
        fh.write("".join(lines))


# ---- meshes ----------------------------------------------------------------
def kuhn_cube(n, jitter=0.2, seed=0):
    """Jittered Kuhn cube: (coords (N,3) float64, tets (6n^3, 4) int32)."""
    m = n + 1
    coords = np.zeros((m ** 3, 3))
    tets = np.zeros((6 * n ** 3, 4), dtype=np.int32)
    _check(lib().anifem_mesh_kuhn_cube(n, jitter, seed, _ptr(coords), _ptr(tets)), "anifem_mesh_kuhn_cube")
    return coords, tets


# ---- global assembly ---------------------------------------------------------
class Assembler:
    """Whole-mesh GPU assembler (anifem_assembly_*), the counterpart of
    AssemblerT::PrepareProblem + Assemble (inmost_interface/assembler.h:232-446).
    With host_only=True only the setup is built (no device use)."""

    def __init__(self, coords, tets, trial: Op = None, test: Op = None, order=2, tensor_type=TENSOR_NULL,
                 tensor_storage=COEF_CONST, rhs_order=None, rhs_tensor_type=TENSOR_SCALAR,
                 rhs_tensor_storage=COEF_CONST, dirichlet=DIRICHLET_NONE, dirichlet_dof_mask=None, host_only=False,
                 budget=None, stream=None, variables=None):
        """variables: a list of Op (fem and ncomp are read) for a
        multi-variable problem (anifem_assembly_create_vars): caller element
        matrices over the variables' concatenated local DOFs only."""
        self._coords = np.ascontiguousarray(coords, dtype=np.float64)
        self._tets = np.ascontiguousarray(tets, dtype=np.int32)
        if variables is not None:
            trial = trial or Op(IDEN, variables[0].fem, variables[0].ncomp)
        test = test or trial
        self.trial, self.test = trial, test
        self.tensor_type, self.tensor_storage = tensor_type, tensor_storage
        self.order = order
        self.rhs_tensor_type, self.rhs_tensor_storage = rhs_tensor_type, rhs_tensor_storage
        self._mask = None if dirichlet_dof_mask is None else np.ascontiguousarray(dirichlet_dof_mask, dtype=np.uint8)
        mesh = anifem_mesh_desc(len(self._coords), len(self._tets), _ptr(self._coords), _ptr(self._tets))
        prob = anifem_problem_desc(trial.c(), test.c(), order, tensor_type, tensor_storage,
                                   order if rhs_order is None else rhs_order, rhs_tensor_type, rhs_tensor_storage,
                                   dirichlet, _ptr(self._mask))
        h = ctypes.c_void_p()
        L = lib()
        if variables is not None:
            if len(variables) > 4:
                raise ValueError("at most 4 variables")
            vd = anifem_vars_desc()
            vd.n_var = len(variables)
            for k, v in enumerate(variables):
                vd.var[k] = v.c()
            self.variables = list(variables)
            if host_only:
                rc = L.anifem_assembly_create_vars_host(ctypes.byref(h), ctypes.byref(mesh), ctypes.byref(vd),
                                                        ctypes.byref(prob))
            else:
                rc = L.anifem_assembly_create_vars(ctypes.byref(h), ctypes.byref(mesh), ctypes.byref(vd),
                                                   ctypes.byref(prob), _stream_handle(stream))
        elif host_only:
            rc = L.anifem_assembly_create_host(ctypes.byref(h), ctypes.byref(mesh), ctypes.byref(prob))
        elif budget is not None:  # (max_slots, max_values, max_pairs) per row block
            rc = L.anifem_assembly_create_budget(ctypes.byref(h), ctypes.byref(mesh), ctypes.byref(prob),
                                                 int(budget[0]), int(budget[1]), int(budget[2]),
                                                 _stream_handle(stream))
        else:
            rc = L.anifem_assembly_create(ctypes.byref(h), ctypes.byref(mesh), ctypes.byref(prob),
                                          _stream_handle(stream))
        _check(rc, "anifem_assembly_create")
        self._h = h
        self.host_only = host_only
        ndof, nnz, nblk, nf = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        L.anifem_assembly_info(h, ctypes.byref(ndof), ctypes.byref(nnz), ctypes.byref(nblk), ctypes.byref(nf))
        self.ndof, self.nnz, self.n_blocks, self.nf = ndof.value, nnz.value, nblk.value, nf.value
        self.n_tets = len(self._tets)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.anifem_assembly_destroy(h)
            self._h = None

    def setup(self):
        """Host copies: dict of row_ptr, col_idx, dofs, slot, dirichlet."""
        rp = np.zeros(self.ndof + 1, np.int64)
        ci = np.zeros(self.nnz, np.int32)
        dofs = np.zeros((self.n_tets, self.nf), np.int32)
        slot = np.zeros((self.n_tets, self.nf * self.nf), np.int64)
        dr = np.zeros(self.ndof, np.uint8)
        _check(lib().anifem_assembly_copy_setup(self._h, _ptr(rp), _ptr(ci), _ptr(dofs), _ptr(slot), _ptr(dr)),
               "copy_setup")
        return dict(row_ptr=rp, col_idx=ci, dofs=dofs, slot=slot, dirichlet=dr)

    def signs(self):
        """(n_tets, nf) int8 orientation signs of the local DOFs (+1 / -1)."""
        out = np.zeros((self.n_tets, self.nf), np.int8)
        _check(lib().anifem_assembly_copy_signs(self._h, _ptr(out)), "anifem_assembly_copy_signs")
        return out

    def block_layout(self):
        lay = anifem_block_layout()
        _check(lib().anifem_assembly_blocks(self._h, ctypes.byref(lay)), "blocks")
        return lay

    def blocks(self):
        """Host copies of the row blocks (host-only contexts): dict of blk_row0,
        blk_p0, blk_v0, pair_tet, pair_ib, pair_pos (n_pairs x nf), sstart, pstart."""
        lay = self.block_layout()
        nb, npair = lay.n_blocks, lay.n_pairs
        out = dict(blk_row0=np.zeros(nb + 1, np.int32), blk_p0=np.zeros(nb + 1, np.int64),
                   blk_v0=np.zeros(nb + 1, np.int64), pair_tet=np.zeros(npair, np.int32),
                   pair_ib=np.zeros(npair, np.uint8), pair_pos=np.zeros((npair, self.nf), np.uint16),
                   sstart=np.zeros(self.nnz, np.uint16), pstart=np.zeros(self.ndof, np.uint16))
        k = ["blk_row0", "blk_p0", "blk_v0", "pair_tet", "pair_ib", "pair_pos", "sstart", "pstart"]
        _check(lib().anifem_assembly_copy_blocks(self._h, *[_ptr(out[x]) for x in k]), "copy_blocks")
        out["layout"] = lay
        return out

    def csr(self):
        """Device (row_ptr, col_idx) addresses."""
        rp, ci = ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib().anifem_assembly_csr(self._h, ctypes.byref(rp), ctypes.byref(ci)), "csr")
        return rp.value, ci.value

    def assemble(self, D=None, F=None, precision=PREC_FP64, val=None, rhs=None, status=True, stream=None,
                 dir_values=None):
        """Assemble into device FP64 val (nnz) and rhs (ndof); returns (val,
        rhs, status).  dir_values: device FP64 (ndof) Dirichlet values of the
        flagged DOFs (None: homogeneous), applied per cell as applyDir does."""
        import torch
        dev = torch.device("cuda", torch.cuda.current_device())
        if val is None:
            val = torch.empty(self.nnz, dtype=torch.float64, device=dev)
        if rhs is None and (F is not None or dir_values is not None):
            rhs = torch.empty(self.ndof, dtype=torch.float64, device=dev)
        tD = anifem_tensor(self.tensor_type, self.tensor_storage, _ptr(D))
        tF = anifem_tensor(self.rhs_tensor_type, self.rhs_tensor_storage, _ptr(F)) if F is not None else None
        st = anifem_status()
        rc = lib().anifem_assembly_run(self._h, ctypes.byref(tD), ctypes.byref(tF) if tF is not None else None,
                                       _ptr(dir_values), precision, _ptr(val), _ptr(rhs),
                                       ctypes.byref(st) if status else None, _stream_handle(stream))
        if rc not in (OK, ERR_NONFINITE):
            raise AnifemError(rc, "anifem_assembly_run")
        return val, rhs, (st if status else None)

    def assemble_plain(self, D=None, F=None, val=None, rhs=None, status=True, stream=None):
        """The plain FP64 baseline (anifem_assembly_run_plain)."""
        import torch
        dev = torch.device("cuda", torch.cuda.current_device())
        if val is None:
            val = torch.empty(self.nnz, dtype=torch.float64, device=dev)
        if rhs is None and F is not None:
            rhs = torch.empty(self.ndof, dtype=torch.float64, device=dev)
        tD = anifem_tensor(self.tensor_type, self.tensor_storage, _ptr(D))
        tF = anifem_tensor(self.rhs_tensor_type, self.rhs_tensor_storage, _ptr(F)) if F is not None else None
        st = anifem_status()
        rc = lib().anifem_assembly_run_plain(self._h, ctypes.byref(tD), ctypes.byref(tF) if tF is not None else None,
                                             _ptr(val), _ptr(rhs), ctypes.byref(st) if status else None,
                                             _stream_handle(stream))
        if rc not in (OK, ERR_NONFINITE):
            raise AnifemError(rc, "anifem_assembly_run_plain")
        return val, rhs, (st if status else None)

    def tets(self):
        """(n_tets, 4) int32 node lists after the orientation reorder: the
        local node order element matrices for assemble_elements refer to."""
        out = np.zeros((self.n_tets, 4), np.int32)
        _check(lib().anifem_assembly_tets(self._h, _ptr(out)), "anifem_assembly_tets")
        return out

    def assemble_elements(self, A_elem, F_elem=None, val=None, rhs=None, status=True, stream=None, symmetric=False):
        """Scatter caller-supplied element matrices (n_tets, nf, nf) in the
        column-major layout of fem3dtet_batch (float64 or float32) and load
        vectors (n_tets, nf) into the FP64 CSR (anifem_assembly_run_elements)."""
        import torch
        dev = torch.device("cuda", torch.cuda.current_device())
        if val is None:
            val = torch.empty(self.nnz, dtype=torch.float64, device=dev)
        if rhs is None and F_elem is not None:
            rhs = torch.empty(self.ndof, dtype=torch.float64, device=dev)
        prec = PREC_FP64 if A_elem.dtype == torch.float64 else PREC_FP32
        if A_elem.dtype not in (torch.float64, torch.float32) or (F_elem is not None and F_elem.dtype != A_elem.dtype):
            raise AnifemError(ERR_ARG, "element arrays must be float64 or float32, both of one type")
        st = anifem_status()
        if symmetric:
            prec |= ELEM_SYMMETRIC
        rc = lib().anifem_assembly_run_elements(self._h, _ptr(A_elem), _ptr(F_elem), prec, _ptr(val), _ptr(rhs),
                                                ctypes.byref(st) if status else None, _stream_handle(stream))
        if rc not in (OK, ERR_NONFINITE):
            raise AnifemError(rc, "anifem_assembly_run_elements")
        return val, rhs, (st if status else None)

    def assemble_rhs(self, F_elem, rhs=None, accumulate=False, status=True, stream=None):
        """AssembleRHS: the rhs from per-cell load vectors (n_tets, nf), float64
        or float32 (anifem_assembly_run_rhs); accumulate adds to rhs."""
        import torch
        if rhs is None:
            rhs = torch.zeros(self.ndof, dtype=torch.float64, device=F_elem.device)
        if F_elem.dtype not in (torch.float64, torch.float32):
            raise AnifemError(ERR_ARG, "load vectors must be float64 or float32")
        prec = PREC_FP64 if F_elem.dtype == torch.float64 else PREC_FP32
        st = anifem_status()
        rc = lib().anifem_assembly_run_rhs(self._h, _ptr(F_elem), prec, _ptr(rhs), 1 if accumulate else 0,
                                           ctypes.byref(st) if status else None, _stream_handle(stream))
        if rc not in (OK, ERR_NONFINITE):
            raise AnifemError(rc, "anifem_assembly_run_rhs")
        return rhs, (st if status else None)

    def node_values(self, u):
        """Per-node values (n_nodes, ncomp) of a global vector u (host or
        device): the vertex DOFs, which come first in each component's block of
        the NATURAL numbering (global_enumerator.cpp:562-777)."""
        u = np.asarray(u.cpu() if hasattr(u, "cpu") else u)
        nc = self.trial.ncomp
        nscal = self.ndof // nc
        nn = len(self._coords)
        return np.stack([u[c * nscal:c * nscal + nn] for c in range(nc)], axis=1)

    def evaluate(self, u, xyl, op=GRAD, out=None, stream=None):
        """op (IDEN or GRAD; IDEN, DIV or CURL of RT0 / ND0) of the trial
        space's global vector u (device, ndof) at the barycentric points xyl
        (q, 4) of every cell
        (anifem_assembly_eval): (n_tets, q, dim)."""
        import torch
        xyl = np.ascontiguousarray(xyl, dtype=np.float64)
        q = len(xyl)
        if self.trial.fem in (FEM_RT0, FEM_ND0):
            dim = 1 if op == DIV else (9 if op == GRAD else 3)  # vector-valued spaces: IDEN, CURL 3, DIV 1, GRAD 9
        else:
            dim = (3 if op == GRAD else 1) * self.trial.ncomp
        if out is None:
            out = torch.empty((self.n_tets, q, dim), dtype=torch.float64, device=u.device)
        _check(lib().anifem_assembly_eval(self._h, op, _ptr(xyl), q, _ptr(u), _ptr(out), _stream_handle(stream)),
               "anifem_assembly_eval")
        return out

    def quadrature_points(self, order=None, out=None, stream=None):
        """Physical points (3, q, n_tets) of the volume rule in every tet of the
        context (device): where PER_POINT coefficients are expected."""
        import torch
        order = self.order if order is None else order
        q = len(quadrature(order)[1])
        if out is None:
            out = torch.empty((3, q, self.n_tets), dtype=torch.float64,
                              device=torch.device("cuda", torch.cuda.current_device()))
        _check(lib().anifem_assembly_quadrature_points(self._h, order, _ptr(out), _stream_handle(stream)),
               "anifem_assembly_quadrature_points")
        return out

    def set_persistent(self, on=True):
        """True / 1: persistent pass (fixed-size form for P1..P3); 2: its general
        form; False / 0: one workgroup per block."""
        mode = on if isinstance(on, int) and not isinstance(on, bool) else (1 if on else 0)
        _check(lib().anifem_assembly_set_persistent(self._h, mode), "set_persistent")

    def set_scatter(self, method):
        """SCATTER_ROW_BLOCKS (default) or SCATTER_COLOURED."""
        _check(lib().anifem_assembly_set_scatter(self._h, method), "anifem_assembly_set_scatter")

    def colours(self):
        """(n_colours, colour per tet) of the coloured scatter."""
        n = ctypes.c_int32()
        col = np.zeros(self.n_tets, np.int32)
        _check(lib().anifem_assembly_colours(self._h, ctypes.byref(n), _ptr(col)), "anifem_assembly_colours")
        return n.value, col

    def enable_timers(self, on=True):
        _check(lib().anifem_assembly_enable_timers(self._h, 1 if on else 0), "enable_timers")

    def timers(self):
        t = anifem_timers()
        _check(lib().anifem_assembly_timers(self._h, ctypes.byref(t)), "timers")
        return t

    def device_csr_tensors(self):
        """(row_ptr, col_idx) as torch tensors copied from the device views."""
        import torch
        s = self.setup()
        dev = torch.device("cuda", torch.cuda.current_device())
        return torch.from_numpy(s["row_ptr"]).to(dev), torch.from_numpy(s["col_idx"]).to(dev)


def pcg_jacobi(n, row_ptr, col_idx, val, b, x, rtol=1e-8, maxit=100000, stream=None):
    """FP64 Jacobi-PCG on device CSR (addresses or tensors); returns result struct."""
    res = anifem_pcg_result()
    rp = row_ptr if isinstance(row_ptr, int) else _ptr(row_ptr)
    ci = col_idx if isinstance(col_idx, int) else _ptr(col_idx)
    rc = lib().anifem_pcg_jacobi(n, rp, ci, _ptr(val), _ptr(b), _ptr(x), rtol, maxit, ctypes.byref(res),
                                 _stream_handle(stream))
    if rc not in (OK, ERR_NOT_CONVERGED):
        raise AnifemError(rc, "anifem_pcg_jacobi")
    return res, rc == OK


def pcg_factored(n, row_ptr, col_idx, val, G, GT, b, x, rtol=1e-8, maxit=100000, stream=None):
    """Flexible PCG with M^-1 = G G^T on device CSR; returns (result, converged)."""
    res = anifem_pcg_result()
    rp = row_ptr if isinstance(row_ptr, int) else _ptr(row_ptr)
    ci = col_idx if isinstance(col_idx, int) else _ptr(col_idx)
    rc = lib().anifem_pcg_factored(n, rp, ci, _ptr(val), _ptr(G), _ptr(GT), _ptr(b), _ptr(x), rtol, maxit,
                                   ctypes.byref(res), _stream_handle(stream))
    if rc not in (OK, ERR_NOT_CONVERGED):
        raise AnifemError(rc, "anifem_pcg_factored")
    return res, rc == OK


def copy_to_host(addr, count, dtype):
    """numpy copy of `count` elements of `dtype` at a device address (for the
    device views the library returns: G, M = G G^T)."""
    hip = ctypes.CDLL("libamdhip64.so.7")  # the HIP runtime libanifem_gpu.so links
    out = np.empty(count, dtype)
    if hip.hipDeviceSynchronize() != 0 or hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(addr),
                                                         ctypes.c_size_t(out.nbytes), 2) != 0:
        raise AnifemError(ERR_HIP, "hipMemcpy")
    return out


def pcg_matrix(n, row_ptr, col_idx, val, m_row_ptr, m_col_idx, m_val, b, x, rtol=1e-8, maxit=100000,
               stream=None):
    """PCG with an explicit SPD preconditioner matrix M (z = M r) on device
    CSR; returns (result, converged)."""
    res = anifem_pcg_result()
    a = [p if isinstance(p, int) else _ptr(p) for p in (row_ptr, col_idx, val, m_row_ptr, m_col_idx, m_val)]
    rc = lib().anifem_pcg_matrix(n, *a, _ptr(b), _ptr(x), rtol, maxit, ctypes.byref(res), _stream_handle(stream))
    if rc not in (OK, ERR_NOT_CONVERGED):
        raise AnifemError(rc, "anifem_pcg_matrix")
    return res, rc == OK


def grf_sample(xyz, ell, sigma, n_features=512, seed=0, lognormal=True, out=None, stream=None):
    """Gaussian random field (or its exponential) at device points xyz (n x 3, float64)."""
    import torch
    n = xyz.shape[0]
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device=xyz.device)
    _check(lib().anifem_grf_sample(n, _ptr(xyz), ell, sigma, n_features, seed, 1 if lognormal else 0, _ptr(out),
                                   _stream_handle(stream)), "anifem_grf_sample")
    return out


class LearnedPreconditioner:
    """The message-passing network of anifem_gnn_* over one CSR pattern (host
    row_ptr / col_idx); matrices are device FP64 value arrays on that pattern."""

    def __init__(self, row_ptr, col_idx, stream=None):
        self._rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
        self._ci = np.ascontiguousarray(col_idx, dtype=np.int32)
        self.n = len(self._rp) - 1
        self.nnz = int(self._rp[-1])
        h = ctypes.c_void_p()
        _check(lib().anifem_gnn_create(ctypes.byref(h), self.n, _ptr(self._rp), _ptr(self._ci),
                                       _stream_handle(stream)), "anifem_gnn_create")
        self._h = h
        self.n_params = lib().anifem_gnn_n_params()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.anifem_gnn_destroy(h)
            self._h = None

    def set_params(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        assert t.size == self.n_params
        _check(lib().anifem_gnn_set_params(self._h, _ptr(t)), "anifem_gnn_set_params")

    def params(self, bf16=False):
        t = np.zeros(self.n_params, np.float32)
        if bf16:
            _check(lib().anifem_gnn_get_params(self._h, None, _ptr(t)), "anifem_gnn_get_params")
        else:
            _check(lib().anifem_gnn_get_params(self._h, _ptr(t), None), "anifem_gnn_get_params")
        return t

    def grad(self):
        g = np.zeros(self.n_params, np.float32)
        _check(lib().anifem_gnn_get_grad(self._h, _ptr(g)), "anifem_gnn_get_grad")
        return g

    def loss_grad(self, a, probes=None, seed=0, stream=None):
        loss = ctypes.c_float()
        _check(lib().anifem_gnn_loss_grad(self._h, _ptr(a), _ptr(probes), seed, ctypes.byref(loss),
                                          _stream_handle(stream)), "anifem_gnn_loss_grad")
        return loss.value

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        _check(lib().anifem_gnn_adam_step(self._h, lr, beta1, beta2, eps, _stream_handle(stream)),
               "anifem_gnn_adam_step")

    def factor(self, a, G=None, GT=None, stream=None):
        """Factor for matrix values a: into the device tensors G / GT when given
        (returned), else (G, GT) device addresses of the context's copies."""
        _check(lib().anifem_gnn_factor(self._h, _ptr(a), _ptr(G), _ptr(GT), _stream_handle(stream)),
               "anifem_gnn_factor")
        if G is not None and GT is not None:
            return G, GT
        G, GT = ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib().anifem_gnn_views(self._h, None, None, None, ctypes.byref(G), ctypes.byref(GT)),
               "anifem_gnn_views")
        return G.value, GT.value

    def precond(self, stream=None):
        """M = G G^T of the last factor kept in the context (factor(a) with no
        G / GT): (m_nnz, row_ptr, col_idx, val) device addresses."""
        m = ctypes.c_int64()
        rp, ci, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib().anifem_gnn_precond(self._h, ctypes.byref(m), ctypes.byref(rp), ctypes.byref(ci),
                                        ctypes.byref(v), _stream_handle(stream)), "anifem_gnn_precond")
        return m.value, rp.value, ci.value, v.value

    def solve(self, a, b, x, rtol=1e-8, maxit=100000, stream=None):
        """Learned-preconditioned PCG (M = G G^T) with Jacobi-PCG fallback."""
        res = anifem_pcg_learned_result()
        rc = lib().anifem_pcg_learned(self._h, _ptr(a), _ptr(b), _ptr(x), rtol, maxit, ctypes.byref(res),
                                      _stream_handle(stream))
        if rc not in (OK, ERR_NOT_CONVERGED):
            raise AnifemError(rc, "anifem_pcg_learned")
        return res, rc == OK


# the batch autotuner's candidate row-block budgets {slots, entries, pairs}
# (ANIFEM_BATCH_BUDGET_CANDIDATES, in the order of anifem_batch_tuning's times)
BATCH_BUDGET_CANDIDATES = [(2048, 1024, 128), (2048, 2048, 256), (4096, 4096, 256), (1024, 1024, 256),
                           (4096, 3072, 256), (4096, 2560, 256), (4096, 2560, 128), (4096, 2048, 128)]


class Batch:
    """Batched scalar-diffusion assembly of many meshes (anifem_batch_*).
    meshes: list of (coords (n,3) float64, tets (m,4) int32, fem)."""

    def __init__(self, meshes, dirichlet=DIRICHLET_BOUNDARY, autotune=True, stream=None):
        self._keep = []
        arr = (anifem_batch_mesh * len(meshes))()
        for i, (c, t, fem) in enumerate(meshes):
            c = np.ascontiguousarray(c, dtype=np.float64)
            t = np.ascontiguousarray(t, dtype=np.int32)
            self._keep += [c, t]
            arr[i] = anifem_batch_mesh(len(c), len(t), _ptr(c), _ptr(t), fem, 0)
        h = ctypes.c_void_p()
        _check(lib().anifem_batch_create(ctypes.byref(h), len(meshes), arr, dirichlet, 1 if autotune else 0,
                                         _stream_handle(stream)), "anifem_batch_create")
        self._h = h
        self._keep = []
        self.n_meshes = len(meshes)
        nt, nd, nz, nb = (ctypes.c_int64() for _ in range(4))
        lib().anifem_batch_totals(h, ctypes.byref(nt), ctypes.byref(nd), ctypes.byref(nz), ctypes.byref(nb))
        self.n_tets, self.ndof, self.nnz, self.n_blocks = nt.value, nd.value, nz.value, nb.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.anifem_batch_destroy(h)
            self._h = None

    def info(self, i):
        v = [ctypes.c_int64() for _ in range(5)]
        _check(lib().anifem_batch_info(self._h, i, *[ctypes.byref(x) for x in v]), "anifem_batch_info")
        return dict(zip(["tet_offset", "dof_offset", "nnz_offset", "ndof", "nnz"], [x.value for x in v]))

    def csr(self, i):
        """Host copies (row_ptr, col_idx) of mesh i's own CSR pattern."""
        inf = self.info(i)
        rp, ci = ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib().anifem_batch_csr(self._h, i, ctypes.byref(rp), ctypes.byref(ci)), "anifem_batch_csr")
        return copy_to_host(rp.value, inf["ndof"] + 1, np.int64), copy_to_host(ci.value, inf["nnz"], np.int32)

    def csr_all(self):
        """Host copies (row_ptr, col_idx) of every mesh's pattern, concatenated:
        mesh i's row_ptr occupies ndof_i + 1 entries starting at 0, and its
        columns are its own DOF numbers."""
        rp, ci = ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib().anifem_batch_csr(self._h, 0, ctypes.byref(rp), ctypes.byref(ci)), "anifem_batch_csr")
        return (copy_to_host(rp.value, self.ndof + self.n_meshes, np.int64),
                copy_to_host(ci.value, self.nnz, np.int32))

    def tuning(self, fem):
        caps = np.zeros(3, np.int32)
        ms = np.zeros(len(BATCH_BUDGET_CANDIDATES))
        _check(lib().anifem_batch_tuning(self._h, fem, _ptr(caps), _ptr(ms)), "anifem_batch_tuning")
        return caps, ms

    def launch_tuning(self, fem):
        """Launch shape of order fem's row-block pass: phase-A form (0
        general, 1 fixed-size, 2 two lanes per pair), compact positions,
        workgroups per CU (0: occupancy), ms per launch (-1 until timed at the
        first run)."""
        shape = np.zeros(3, np.int32)
        ms = ctypes.c_double()
        _check(lib().anifem_batch_launch_tuning(self._h, fem, _ptr(shape), ctypes.byref(ms)),
               "anifem_batch_launch_tuning")
        return {"form": int(shape[0]), "compact": int(shape[1]), "workgroups_per_cu": int(shape[2]), "ms": ms.value}

    def set_persistent(self, on=True):
        _check(lib().anifem_batch_set_persistent(self._h, 1 if on else 0), "anifem_batch_set_persistent")

    def run(self, kappa, f, val=None, rhs=None, status=True, stream=None):
        import torch
        dev = kappa.device
        if val is None:
            val = torch.empty(self.nnz, dtype=torch.float64, device=dev)
        if rhs is None:
            rhs = torch.empty(self.ndof, dtype=torch.float64, device=dev)
        st = anifem_status()
        rc = lib().anifem_batch_run(self._h, _ptr(kappa), _ptr(f), _ptr(val), _ptr(rhs),
                                    ctypes.byref(st) if status else None, _stream_handle(stream))
        if rc not in (OK, ERR_NONFINITE):
            raise AnifemError(rc, "anifem_batch_run")
        return val, rhs, (st if status else None)


def _gnn_init(n_params, seed):
    """Initial network parameters: Gaussian weights (scaled by fan-in), zero
    biases and zero output heads, so that the initial factor is Jacobi."""
    H, F = 32, 4
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, 1 / np.sqrt(F), (H, F)).ravel(), np.zeros(H)]
    for _ in range(3):
        parts += [rng.normal(0, 1 / np.sqrt(H), (H, H)).ravel(), 0.5 * rng.normal(0, 1 / np.sqrt(H), (H, H)).ravel(),
                  np.zeros(H)]
    parts += [np.zeros(H), np.zeros(1), np.zeros(H), np.zeros(1)]
    theta = np.concatenate(parts).astype(np.float32)
    assert theta.size == n_params
    return theta


def version():
    return lib().anifem_version().decode()
