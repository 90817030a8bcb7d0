"""GPU parity on meshes read from gmsh files (sdg_mesh_create_gmsh): an unstructured triangulation
with far-field and slip-wall groups or with $Periodic links, Euler with and without artificial
viscosity, residual and 5 SSP-RK3 steps against the oracle at the FP64 bar."""
import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L
from test_gmsh import triangle_file

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order,av,periodic", [(2, False, False), (3, False, False), (3, True, False), (3, False, True)],
                         ids=["p2", "p3", "p3-av", "p3-periodic"])
def test_triangle_file_mesh_parity(tmp_path, order, av, periodic):
    print(f"gmsh triangle parity P{order} av={av} periodic={periodic}", flush=True)
    path = str(tmp_path / "t.msh")
    triangle_file(path, n=6, bc=(1, 1, 2, 2), periodic=periodic)
    mesh = L.Mesh.gmsh(path, order, {1: L.BC_RIEMANN_FARFIELD, 2: L.BC_ADIABATIC_SLIP_WALL})
    prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_HLLC, cp=3.5, cv=2.5,
                          shock_capturing=L.SHOCK_ARTIFICIAL_VISCOSITY if av else L.SHOCK_NONE,
                          empirical_tolerance=0.5)
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, 2)
    jump = np.where(xq[..., 0] < 0.47, 1.0, 0.5)  # a mild jump: the projection stays positive
    prim = np.stack([jump, 0 * jump, 0 * jump, 0 * jump + 1.0], axis=-1) if av else np.stack(
        [1.0 + 0.1 * np.sin(2 * np.pi * xq[..., 0]), 0.3 + 0 * xq[..., 0], 0.05 * np.sin(np.pi * xq[..., 1]),
         1.0 + 0.05 * np.cos(2 * np.pi * xq[..., 1])], axis=-1)
    bprim = np.tile(np.array([1.0, 0.3, 0.0, 1.0]), (mesh.Fb, mesh.nfq, 1))
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    if mesh.Fb:
        s.set_boundary_primitive(bprim)
    o, x = O.Oracle(prob, mesh), O.Oracle(prob, mesh, extended=True)
    for k in (o, x):
        k.set_state(s.get_state())
        if mesh.Fb:
            k.set_boundary_primitive(bprim)
    u0 = o.get_state()
    P.fp64_parity(s.residual(), o.residual(), x.residual(), u0, 2, tol=1e-12)
    if av:
        assert o.artificial_viscosity().max() > 0.0
        assert np.abs(s.artificial_viscosity() - o.artificial_viscosity()).max() <= 1e-13 * o.artificial_viscosity().max()
    dt = o.compute_dt(0.05)
    o.step(dt, 5)
    x.step(dt, 5)
    _, nan, taken = s.step(dt, 5)
    assert nan == 0 and taken == 5
    uo = o.get_state()
    P.fp64_parity(s.get_state(), uo, x.get_state(), uo, 2, tol=1e-12)


def test_cpp_driver_reads_a_mesh_file(tmp_path):
    """setMesh(path) in the C++ drop-in: a triangulation with far-field (group 1) and slip-wall
    (group 2) sides keeps a uniform stream along x uniform."""
    import os
    import subprocess

    path = str(tmp_path / "t.msh")
    triangle_file(path, n=6, bc=(1, 1, 2, 2))
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "examples", "bin",
                       "mesh_file_2d_ceuler")
    r = subprocess.run([exe, path, "3", "40"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "mesh file: 40 steps" in r.stdout


def test_cpp_driver_generates_a_mesh(tmp_path):
    """setMesh(path, generate) in the C++ drop-in with a path that does not exist: the callback's
    in-memory triangulation (jittered interior nodes, far field at x = 0 and 2, slip walls at
    y = 0 and 1) goes through sdg_mesh_create_from_arrays and keeps a uniform stream uniform."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "examples", "bin",
                       "mesh_file_2d_ceuler")
    path = str(tmp_path / "absent.msh")
    r = subprocess.run([exe, path, "2", "40"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "mesh file: 40 steps" in r.stdout
    assert not os.path.exists(path)


def test_cpp_driver_reads_a_curved_mesh_file(tmp_path):
    """setMesh(path) in the C++ drop-in on a file of 9-node quadrangles (an annular sector, every
    geometry node on its circle, boundary lines of 3 nodes, all in the far-field group): a uniform
    stream stays uniform on the curved cells."""
    import os
    import subprocess

    import msh_writer as W
    from test_curved import _annulus_arrays

    g, na, nr = 2, 5, 2
    gid, xyz, cells = _annulus_arrays(g, n=(na, nr))
    radial = np.array([[gid[i, j * g], gid[i, (j + 1) * g], gid[i, j * g + 1]] for j in range(nr) for i in (0, na * g)])
    arcs = np.array([[gid[i * g, j], gid[(i + 1) * g, j], gid[i * g + 1, j]] for i in range(na) for j in (0, nr * g)])
    path = str(tmp_path / "annulus.msh")
    W.write_msh(path, 2, np.arange(1, len(xyz) + 1), xyz, "quad9", cells + 1,
                [(1, "line3", radial + 1), (1, "line3", arcs + 1)])
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "examples", "bin",
                       "mesh_file_2d_ceuler")
    r = subprocess.run([exe, path, "3", "40"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "mesh file: 40 steps" in r.stdout
