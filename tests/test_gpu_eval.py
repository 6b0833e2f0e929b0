"""GPU solution evaluation: anifem_fem3dapply_batch (fem3DapplyL of
anifem++/fem/operations/eval.inl) against the fixtures of the reference
element core compiled here, and anifem_assembly_eval (a global vector
evaluated in every cell through the DOF map) against polynomials its space
reproduces.  FP64: <= 1e-12 relative to the largest value."""
import numpy as np
import pytest

import anifem_amd as af
import anifem_testlib as T

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.mark.parametrize("case", [c[0] for c in T.apply_golden_cases()])
def test_apply_batch_vs_reference_fixture(case):
    z = dict(T.apply_golden_cases())[case]
    code, op, fem, nc = [int(v) for v in z["meta"]]
    out = af.fem3dapply_batch(af.Op(op, fem, nc), z["xyl"], dev(T.soa_xyz(z["xyz"])), dev(z["dofs"]))
    out = out.cpu().numpy()
    ref = z["out"]
    assert out.shape == ref.shape
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max(), case


@pytest.mark.parametrize("case", [c[0] for c in T.applyx_golden_cases()])
def test_applyx_batch_vs_reference_fixture(case):
    """anifem_fem3dapplyx_batch (fem3DapplyX) at the fixtures' per-tet physical
    points against the compiled reference."""
    z = dict(T.applyx_golden_cases())[case]
    code, op, fem, nc = [int(v) for v in z["meta"]]
    out = af.fem3dapplyx_batch(af.Op(op, fem, nc), dev(z["X"]), dev(T.soa_xyz(z["xyz"])), dev(z["dofs"]))
    out = out.cpu().numpy()
    ref = z["out"]
    assert out.shape == ref.shape
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max(), case


def test_applyx_matches_applyl_at_common_points():
    """At physical points that are the same barycentric points in every tet,
    the X form equals the L form (up to the rounding of the barycentric
    coordinates) for P3 gradients and ND0 curls."""
    rng = np.random.default_rng(9)
    n, q = 200, 6
    xyz = rng.normal(size=(n, 12))
    xyl = rng.dirichlet(np.ones(4), q)
    X = T.points_physical(xyz, xyl)
    for op in (af.Op(af.GRAD, af.FEM_P3), af.Op(af.CURL, af.FEM_ND0), af.Op(af.IDEN, af.FEM_RT0)):
        dofs = dev(rng.normal(size=(n, op.nfa)))
        a = af.fem3dapply_batch(op, xyl, dev(T.soa_xyz(xyz)), dofs).cpu().numpy()
        b = af.fem3dapplyx_batch(op, dev(X), dev(T.soa_xyz(xyz)), dofs).cpu().numpy()
        assert np.abs(a - b).max() <= 1e-10 * np.abs(a).max()


def quadratic(c):
    f = lambda x: (c[0] + c[1] * x[..., 0] + c[2] * x[..., 1] + c[3] * x[..., 2] + c[4] * x[..., 0] * x[..., 1] +
                   c[5] * x[..., 2] ** 2 + c[6] * x[..., 1] * x[..., 2])
    g = lambda x: np.stack([c[1] + c[4] * x[..., 1], c[2] + c[4] * x[..., 0] + c[6] * x[..., 2],
                            c[3] + 2 * c[5] * x[..., 2] + c[6] * x[..., 1]], axis=-1)
    return f, g


@pytest.mark.parametrize("nc", [1, 3])
def test_assembly_eval_reproduces_quadratics(nc):
    """A P2 (or P2^3) global vector holding quadratics at the DOF nodes
    (vertices and edge midpoints, through the context's NATURAL numbering) is
    evaluated in every cell: values and gradients equal the quadratics'."""
    rng = np.random.default_rng(nc)
    coords, tets = af.kuhn_cube(3, 0.2, 4)
    tets = tets[rng.permutation(len(tets))]
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P2, nc), dirichlet=af.DIRICHLET_BOUNDARY)
    rt = Asm.tets()
    dofs = Asm.setup()["dofs"]
    E = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    fs = [quadratic(rng.normal(size=7)) for _ in range(nc)]
    u = np.full(Asm.ndof, np.nan)
    for t in range(len(rt)):
        pts = [coords[rt[t, i]] for i in range(4)] + [(coords[rt[t, i]] + coords[rt[t, j]]) / 2 for i, j in E]
        for c in range(nc):
            for i, x in enumerate(pts):
                u[dofs[t, 10 * c + i]] = fs[c][0](x)
    assert not np.isnan(u).any()
    xyl = rng.dirichlet(np.ones(4), 5)
    X = np.einsum("qv,tvc->tqc", xyl, coords[rt])
    val = Asm.evaluate(dev(u), xyl, op=af.IDEN).cpu().numpy()
    grad = Asm.evaluate(dev(u), xyl, op=af.GRAD).cpu().numpy()
    for c in range(nc):
        f, g = fs[c]
        fv, gv = f(X), g(X)
        assert np.abs(val[:, :, c] - fv).max() <= 1e-12 * np.abs(fv).max()
        # component k + 3c of GRAD is du_c / dx_k
        assert np.abs(grad[:, :, 3 * c:3 * c + 3] - gv).max() <= 1e-11 * np.abs(gv).max()


def test_solve_evaluate_and_write_vtu(tmp_path):
    """The steps after the assembly (ex1.cpp:160-168): Jacobi-PCG solve of a
    P2 Poisson problem, per-node values of the solution, its gradient at the
    cells' centroids, and the VTU file of both."""
    import torch
    import xml.etree.ElementTree as ET
    coords, tets = af.kuhn_cube(4, 0.2, 3)
    Asm = af.Assembler(coords, tets, af.Op(af.GRAD, af.FEM_P2), order=2, rhs_order=2,
                       rhs_tensor_type=af.TENSOR_SCALAR, rhs_tensor_storage=af.COEF_CONST,
                       dirichlet=af.DIRICHLET_BOUNDARY)
    val, rhs, st = Asm.assemble(F=dev(np.array([1.0])))
    assert st.code == 0
    x = torch.zeros(Asm.ndof, dtype=torch.float64, device="cuda")
    rp, ci = Asm.csr()
    res, ok = af.pcg_jacobi(Asm.ndof, rp, ci, val, rhs, x, rtol=1e-10)
    assert ok and res.true_relres < 1e-9
    un = Asm.node_values(x)
    assert un.shape == (len(coords), 1) and un.max() > 0
    g = Asm.evaluate(x, np.full((1, 4), 0.25), op=af.GRAD).cpu().numpy()[:, 0, :]
    path = tmp_path / "u.vtu"
    af.write_vtu(str(path), coords, Asm.tets(), point_data={"u": un[:, 0]}, cell_data={"grad_u": g})
    arrays = {a.get("Name"): a for a in ET.parse(str(path)).getroot().iter("DataArray")}
    assert np.allclose(np.array(arrays["u"].text.split(), dtype=float), un[:, 0], rtol=0, atol=0)
