"""The C-ABI library loads and exports every symbol include/anifem_gpu.h
declares; argument validation needs no device."""
import ctypes
import os
import re

import numpy as np
import pytest

import anifem_amd as af
import anifem_testlib as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "anifem_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(anifem_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol():
    L = af.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
        assert n in af.SIGNATURES, f"{n} has no ctypes signature"


def test_version_and_op_sizes():
    assert "gfx950" in af.version()
    assert af.Op(af.GRAD, af.FEM_P2).nfa == 10 and af.Op(af.GRAD, af.FEM_P2).dim == 3
    assert af.Op(af.GRAD, af.FEM_P2, 3).nfa == 30 and af.Op(af.GRAD, af.FEM_P2, 3).dim == 9
    assert af.Op(af.IDEN, af.FEM_P3).nfa == 20 and af.Op(af.IDEN, af.FEM_P0).nfa == 1
    # derivatives of P0: zero operators with the dims of poly_0.h:11-59
    assert af.Op(af.GRAD, af.FEM_P0).nfa == 1 and af.Op(af.GRAD, af.FEM_P0).dim == 3
    assert af.Op(af.DUDY, af.FEM_P0).dim == 1 and af.Op(af.DIV, af.FEM_P0, 3).dim == 1
    # DIV / CURL need 3 components (operators.h:286-400)
    assert af.lib().anifem_op_nfa(af.anifem_op(af.DIV, af.FEM_P1, 1)) < 0


def test_quadrature_rules():
    counts = [(0, 1), (1, 1), (2, 4), (3, 8), (4, 14), (5, 14), (6, 24)]
    counts += [(o, (o // 2 + 1) ** 3) for o in range(7, 21)]  # conical products
    for order, q in counts:
        p, w = af.quadrature(order)
        assert len(w) == q and abs(w.sum() - 1) < 1e-15 and np.all(p > 0) and np.all(w > 0)
        assert np.allclose(p.sum(axis=1), 1, atol=1e-15)
        if order == 3 or order >= 7:  # same construction, same point order as the oracle's
            po, wo = T.orc_quadrature(order)
            assert np.abs(po - p).max() < 1e-15 and np.abs(wo - w).max() < 1e-17
    for bad in (-1, 21):  # quadrature_formulas.cpp:1497-1498 throws outside 0..20
        with pytest.raises(af.AnifemError):
            af.quadrature(bad)


def test_argument_errors_without_device():
    L = af.lib()
    t = af.anifem_tensor(af.TENSOR_NULL, af.COEF_CONST, None)
    g1, g2 = af.Op(af.GRAD, af.FEM_P1).c(), af.Op(af.IDEN, af.FEM_P1).c()
    buf = np.zeros(16)
    # order > 20 (quadrature_formulas.cpp:1497-1498 throws)
    assert L.anifem_fem3dtet_batch(g1, g1, ctypes.byref(t), 21, 1, buf.ctypes.data, 0, buf.ctypes.data, None) \
        == af.ERR_ORDER
    # identity tensor between GRAD (dim 3) and IDEN (dim 1) (diff_tensor.h:313-315)
    assert L.anifem_fem3dtet_batch(g1, g2, ctypes.byref(t), 2, 1, buf.ctypes.data, 0, buf.ctypes.data, None) \
        == af.ERR_DIMS
    bad = af.anifem_op(af.DIV, af.FEM_P1, 1)
    assert L.anifem_fem3dtet_batch(bad, g1, ctypes.byref(t), 2, 1, buf.ctypes.data, 0, buf.ctypes.data, None) \
        == af.ERR_ARG
    assert L.anifem_mesh_kuhn_cube(0, 0.2, 0, None, None) == af.ERR_ARG
    assert L.anifem_mesh_kuhn_cube(4, 0.3, 0, None, None) == af.ERR_ARG


def test_cpp_header_mirrors_reference_tags(tmp_path):
    """include/anifem_fem3dtet.hpp compiles with the host compiler and maps the
    reference's Operator / FemFix / FemVec / DfuncTraits tags onto the C ABI."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "inmost-fem_amd")
    exe = tmp_path / "header_check"
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "cpp", "header_check.cpp"), "-L", libdir, "-lanifem_gpu",
                    "-Wl,-rpath," + libdir, "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "header ok" in out.stdout


def test_write_vtu_round_trip(tmp_path):
    """The VTU writer: an XML unstructured grid whose points, connectivity,
    offsets, cell types and fields read back as written."""
    import xml.etree.ElementTree as ET
    coords, tets = af.kuhn_cube(2, 0.2, 1)
    u = np.sin(coords[:, 0]) + coords[:, 1] ** 2
    g = np.random.default_rng(0).normal(size=(len(tets), 3))
    path = tmp_path / "mesh.vtu"
    af.write_vtu(str(path), coords, tets, point_data={"u": u}, cell_data={"g": g})
    root = ET.parse(str(path)).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfPoints")) == len(coords) and int(piece.get("NumberOfCells")) == len(tets)
    arrays = {a.get("Name"): a for a in root.iter("DataArray")}
    rd = lambda name: np.array(arrays[name].text.split(), dtype=np.float64)
    assert np.array_equal(rd("Points").reshape(-1, 3), coords)
    assert np.array_equal(rd("connectivity").reshape(-1, 4), tets)
    assert np.array_equal(rd("offsets"), 4 * np.arange(1, len(tets) + 1))
    assert np.all(rd("types") == 10)
    assert np.array_equal(rd("u"), u) and np.array_equal(rd("g").reshape(-1, 3), g)
    assert arrays["g"].get("NumberOfComponents") == "3"
