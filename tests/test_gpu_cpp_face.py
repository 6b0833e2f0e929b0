"""The C++ face (include/anifem_fem3dtet.hpp) with data on the GPU: a host
program compiled against the header calls fem3Dtet<GRAD P2, GRAD P2,
DfuncTraits<TENSOR_SYMMETRIC, false>> on a batch of tets and
Assembler<GRAD P2, GRAD P2>::Assemble on a C2-shaped 5^3 mesh (per-tet
anisotropic tensor, unit load, boundary Dirichlet, then Dirichlet values),
the call shapes of int_tet.h:17-20 and assembler.h:383-386.  Its outputs
match the oracle within 1e-12."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import anifem_amd as af
import anifem_testlib as T

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path / "assemble_gpu"
    libdir = os.path.join(ROOT, "inmost-fem_amd")
    subprocess.run([gxx, "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
                    "-I", "/opt/rocm/include", os.path.join(ROOT, "tests", "cpp", "assemble_gpu.cpp"),
                    "-L", libdir, "-lanifem_gpu", "-L", "/opt/rocm/lib", "-lamdhip64",
                    "-Wl,-rpath," + libdir, "-Wl,-rpath,/opt/rocm/lib", "-o", str(exe)], check=True)
    return exe


def test_cpp_face_element_batch_and_assembler(tmp_path):
    exe = _build(tmp_path)
    rng = np.random.default_rng(55)
    coords, tets = af.kuhn_cube(5, 0.2, 8)
    n = len(tets)
    R = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
    lam = np.stack([np.ones(n), 10 ** rng.uniform(-2, 2, n), 10 ** rng.uniform(-2, 2, n)], axis=1)
    Dm = np.einsum("rij,rj,rkj->rik", R, lam, R)
    D = np.ascontiguousarray(Dm.transpose(0, 2, 1).reshape(n, 9))
    dpack = np.ascontiguousarray(T.pack_sym(D, 3).T)
    xyz = coords[tets].reshape(n, 12)
    d = str(tmp_path)
    np.array([len(coords), n], np.int64).tofile(os.path.join(d, "meta.bin"))
    np.ascontiguousarray(coords).tofile(os.path.join(d, "coords.bin"))
    np.ascontiguousarray(tets, dtype=np.int32).tofile(os.path.join(d, "tets.bin"))
    dpack.tofile(os.path.join(d, "dpacked.bin"))
    np.ascontiguousarray(xyz.T).tofile(os.path.join(d, "xyz.bin"))
    out = subprocess.run([str(exe), d], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "cpp face ok" in out.stdout

    # element level
    A = np.fromfile(os.path.join(d, "elem.bin")).reshape(n, 100)
    assert T.rel_frob(A, T.orc_elem(2, 3, 1, 2, 3, 1, 3, 1, 2, xyz, D)) <= 1e-12

    # global level against applyDir + the sequential scatter
    ndof, nnz = np.fromfile(os.path.join(d, "sizes.bin"), np.int64)
    val = np.fromfile(os.path.join(d, "val.bin"))
    rhs = np.fromfile(os.path.join(d, "rhs.bin"))
    rp = np.fromfile(os.path.join(d, "row_ptr.bin"), np.int64)
    ci = np.fromfile(os.path.join(d, "col_idx.bin"), np.int32)
    pos0 = T.oracle_assembly(coords, tets, 3, 1, with_slots=False)
    dpos = T.dof_positions(coords, pos0["tets"], pos0["dofs"], T.lagrange_nodes(3))
    mask = np.zeros(pos0["ndof"], np.uint8)
    mask[pos0["dofs"][np.any((dpos < 1e-12) | (dpos > 1 - 1e-12), axis=-1)]] = 1
    S = T.oracle_assembly(coords, tets, 3, 1, dirichlet_mask=mask)
    assert (S["ndof"], len(S["col_idx"])) == (ndof, nnz)
    np.testing.assert_array_equal(rp, S["row_ptr"])
    np.testing.assert_array_equal(ci, S["col_idx"])
    np.testing.assert_array_equal(np.fromfile(os.path.join(d, "tets_reordered.bin"), np.int32).reshape(n, 4),
                                  S["tets"])
    xr = coords[S["tets"]].reshape(n, 12)
    Ae = T.orc_elem(2, 3, 1, 2, 3, 1, 3, 1, 2, xr, D)
    Fe = T.orc_elem(1, 1, 1, 1, 3, 1, 2, 0, 2, xr, np.ones((1, 1)))
    for name, bc in [("rhs.bin", None), ("rhs_bc.bin", 1.0 + np.arange(ndof) / ndof)]:
        A1, F1 = T.apply_dir_elementwise(Ae, Fe, S["dofs"], mask, bc)
        rc, v_or, r_or = T.oracle_scatter(n, 10, S["dofs"], S["slot"], A1, F1, nnz, ndof)
        assert rc == 0
        rowmax = np.maximum.reduceat(np.abs(v_or), rp[:-1])
        assert np.all(np.abs(val - v_or) <= 1e-12 * np.repeat(rowmax, np.diff(rp)))
        r = np.fromfile(os.path.join(d, name))
        assert np.abs(r - r_or).max() <= 1e-12 * np.abs(r_or).max(), name
    # the product space RT0 x P0 through RuntimeAssembler(FemSpace): the
    # several-variable context of the Python binding (host-only numbering)
    mnd, mnz = np.fromfile(os.path.join(d, "mixed_sizes.bin"), np.int64)
    H = af.Assembler(coords, tets, None, variables=[af.Op(af.IDEN, af.FEM_RT0), af.Op(af.IDEN, af.FEM_P0)],
                     host_only=True)
    assert (mnd, mnz) == (H.ndof, H.nnz)
    faces = {tuple(sorted(t[[k, (k + 1) % 4, (k + 2) % 4]])) for t in tets for k in range(4)}
    assert mnd == len(faces) + n
