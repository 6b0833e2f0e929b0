"""P3 through the device setup and the GPU kernels on a shuffled mesh (relabelled
nodes, shuffled cells and local vertex orders): the numbering built on the GPU
puts every global DOF at one physical point, so a cubic interpolated through
it is reproduced in every cell by anifem_assembly_eval, and the GPU-assembled
P3 stiffness gives a linear field its exact energy.  The DOF locations come
from the basis alone (see test_p3_orientation.py)."""
import numpy as np
import pytest

import anifem_amd as af
import anifem_testlib as T
from test_p3_orientation import _cubic, _mesh, cubic_global_vector

pytestmark = pytest.mark.gpu

P3 = 4


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def test_gpu_numbering_reproduces_cubics():
    coords, tets = _mesh(4, 9)
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, P3), order=4)
    rt = Asm.tets()
    dofs = Asm.setup()["dofs"]
    rng = np.random.default_rng(5)
    f, g = _cubic(rng.normal(size=6))
    u = cubic_global_vector(coords, rt, dofs, Asm.ndof, f)
    xyl = rng.dirichlet(np.ones(4), 6)
    X = np.einsum("qv,tvc->tqc", xyl, coords[rt])
    val = Asm.evaluate(dev(u), xyl, op=af.IDEN).cpu().numpy()[:, :, 0]
    grad = Asm.evaluate(dev(u), xyl, op=af.GRAD).cpu().numpy()
    assert np.abs(val - f(X)).max() <= 1e-12 * np.abs(f(X)).max()
    assert np.abs(grad - g(X)).max() <= 1e-11 * np.abs(g(X)).max()


def test_gpu_p3_energy_of_a_linear_field():
    coords, tets = _mesh(3, 4)
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, P3), order=4)
    val, _, st = Asm.assemble()
    assert st.code == 0
    s = Asm.setup()
    u = cubic_global_vector(coords, Asm.tets(), s["dofs"], Asm.ndof, lambda x: x @ np.array([2.0, -1.0, 3.0]))
    v = val.cpu().numpy()
    rp, ci = s["row_ptr"], s["col_idx"]
    rows = np.repeat(np.arange(Asm.ndof), np.diff(rp))
    Ku = np.bincount(rows, weights=v * u[ci], minlength=Asm.ndof)
    assert abs(u @ Ku - 14.0) < 1e-10
    # stiffness annihilates constants
    assert np.abs(np.bincount(rows, weights=v, minlength=Asm.ndof)).max() < 1e-11 * np.abs(v).max()
