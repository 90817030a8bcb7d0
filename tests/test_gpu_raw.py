"""Raw state files from the GPU engine: the payload layout (per element the coefficients and, for
Navier-Stokes, the lifted-gradient coefficients; per boundary face its parent's blocks; nodal
artificial viscosity), restart from the last step (bit-identical continuation), and the start of
an order-P run from an order-(P-1) file (SpecificFile, InitialCondition.cpp:41-83)."""
import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L

pytestmark = pytest.mark.gpu

GRAD_TOL = 5e-11


def _payload(path):
    return np.frombuffer(L.raw_file_read(path), dtype=np.float64)


@pytest.mark.parametrize("order", [2, 3])
def test_euler_file_layout_and_restart(tmp_path, order):
    """Order 2 runs the vector kernels, order 3 the matrix-core kernels; both restart exactly."""
    mesh, prob, prim = P.euler_vortex(n=6, order=order)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    dt = s.compute_dt(0.1)
    s.step(dt, 3)
    path = str(tmp_path / "euler_3.zst")
    s.write_raw(path, node_count=49)
    u = s.get_state()
    data = _payload(path)
    assert data.size == u.size + 49
    assert np.array_equal(data[: u.size], u.reshape(-1))
    assert np.all(data[u.size:] == 0.0)
    # LastStep restart continues bit-identically
    r = L.Solver(prob, mesh)
    r.read_raw(path, L.IC_LAST_STEP)
    assert np.array_equal(r.get_state(), u)
    s.step(dt, 4)
    r.step(dt, 4)
    assert np.array_equal(r.get_state(), s.get_state())


@pytest.mark.parametrize("visc", [L.VISC_BR1, L.VISC_BR2])
def test_navier_stokes_file_holds_gradients(tmp_path, visc):
    mesh, prob, prim = P.ns_tgv(n=2, order=3, visc=visc)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    path = str(tmp_path / "ns.zst")
    s.write_raw(path)
    u, g = s.get_state(), s.gradient()
    data = _payload(path)
    nu, ng = mesh.nb * mesh.nv, mesh.nb * mesh.nv * mesh.dim
    assert data.size == mesh.E * (nu + ng)
    blocks = data.reshape(mesh.E, nu + ng)
    assert np.array_equal(blocks[:, :nu].reshape(u.shape), u)
    assert np.array_equal(blocks[:, nu:].reshape(g.shape), g)
    r = L.Solver(prob, mesh)
    r.read_raw(path, L.IC_LAST_STEP)
    assert np.array_equal(r.get_state(), u)


def test_specific_file_prolongs_from_order_p_minus_one(tmp_path):
    lo_mesh, prob, prim = P.euler_vortex(n=5, order=2)
    s = L.Solver(prob, lo_mesh)
    s.set_initial_primitive(prim)
    path = str(tmp_path / "p2.zst")
    s.write_raw(path)
    hi_mesh, prob_hi, _ = P.euler_vortex(n=5, order=3)
    r = L.Solver(prob_hi, hi_mesh)
    r.read_raw(path, L.IC_SPECIFIC_FILE)
    lo, hi = s.get_state(), r.get_state()
    assert np.array_equal(hi[:, : lo_mesh.nb, :], lo)
    assert np.all(hi[:, lo_mesh.nb:, :] == 0.0)


@pytest.mark.parametrize("order", [3, 4])
def test_specific_file_on_lines(tmp_path, order):
    """A line run of order P starts from an order P-1 file (basis sizes P and P+1)."""
    print(f"specific file on lines P{order}", flush=True)
    prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_ROE, cp=3.5, cv=2.5)
    meshes = [L.Mesh(L.ELEM_LINE, p, (8,), (0.0,), (1.0,)) for p in (order - 1, order)]
    x = meshes[0].field("quad_coord").reshape(meshes[0].E, meshes[0].nq, 1)[..., 0]
    prim = np.stack([1.0 + 0.2 * np.sin(2 * np.pi * x), 0.3 + 0.0 * x, 1.0 + 0.0 * x], axis=-1)
    s = L.Solver(prob, meshes[0])
    s.set_initial_primitive(prim)
    path = str(tmp_path / "line.zst")
    s.write_raw(path)
    r = L.Solver(prob, meshes[1])
    r.read_raw(path, L.IC_SPECIFIC_FILE)
    lo, hi = s.get_state(), r.get_state()
    assert np.array_equal(hi[:, : meshes[0].nb, :], lo)
    assert np.all(hi[:, meshes[0].nb:, :] == 0.0)


def test_boundary_faces_carry_their_parents_blocks(tmp_path):
    mesh, prob, prim, bprim = P.euler_box(n=4, order=2)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    s.set_boundary_primitive(bprim)
    path = str(tmp_path / "box.zst")
    s.write_raw(path)
    u = s.get_state()
    data = _payload(path)
    nu = mesh.nb * mesh.nv
    par = mesh.field("boundary_face_parent").reshape(-1, 2)[:, 0]
    assert data.size == (mesh.E + par.size) * nu
    blocks = data.reshape(-1, nu)
    assert np.array_equal(blocks[: mesh.E].reshape(u.shape), u)
    assert np.array_equal(blocks[mesh.E:], u.reshape(mesh.E, nu)[par])


@pytest.mark.parametrize("element,order", [(L.ELEM_QUADRANGLE, 3), (L.ELEM_QUADRANGLE, 2), (L.ELEM_HEXAHEDRON, 2),
                                           (L.ELEM_HEXAHEDRON, 3)])
def test_br2_boundary_face_blocks(tmp_path, element, order):
    """Navier-Stokes BR2 with boundary faces (far field, no-slip walls): each boundary face's block
    is its parent's state and the parent's volume gradient plus that face's BR2 lift
    (RawBinary.cpp:93-110), at the gradient bar against the oracle's per-face coefficients; the
    element blocks are the state and the lifted total gradient."""
    print(f"BR2 boundary blocks {element} P{order}", flush=True)
    dim = 2 if element == L.ELEM_QUADRANGLE else 3
    wall, far = L.BC_ISOTHERMAL_NONSLIP_WALL, L.BC_RIEMANN_FARFIELD
    bcs = ((far, far), (wall, wall)) + ((None,) if dim == 3 else ())
    n = 3
    mesh = L.Mesh(element, order, (n,) * dim, (0.0,) * dim, (1.0,) * dim, boundary=bcs)
    cp, cv = P.tgv_thermo(0.1)
    prob = L.problem_desc(equation=L.EQ_NS, convective_flux=L.FLUX_ROE, viscous_flux=L.VISC_BR2,
                          transport=L.TRANSPORT_CONSTANT, cp=cp, cv=cv, dynamic_viscosity=1e-3)
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, dim)
    prim = np.zeros((mesh.E, mesh.nq, dim + 2))
    prim[..., 0] = 1.0 + 0.1 * np.sin(2 * np.pi * xq[..., 0]) * np.cos(2 * np.pi * xq[..., 1])
    prim[..., 1] = 0.3 + 0.05 * np.cos(2 * np.pi * xq[..., 1])
    prim[..., 2] = 0.1 * np.sin(2 * np.pi * xq[..., 0])
    prim[..., -1] = 1.0 + 0.05 * np.cos(2 * np.pi * (xq[..., 0] + xq[..., 1]))
    bprim = np.zeros((mesh.Fb, mesh.nfq, dim + 2))
    bprim[..., 0], bprim[..., 1], bprim[..., -1] = 1.0, 0.3, 1.0
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    s.set_boundary_primitive(bprim)
    path = str(tmp_path / "br2.zst")
    s.write_raw(path)
    u = s.get_state()
    o = O.Oracle(prob, mesh)
    o.set_state(u)
    o.set_boundary_primitive(bprim)
    o.residual()
    fg = o.face_gradient_coefficients()
    data = _payload(path)
    nu, ng = mesh.nb * mesh.nv, mesh.nb * mesh.nv * dim
    bpar = mesh.field("boundary_face_parent").reshape(-1, 2)
    assert data.size == (mesh.E + mesh.Fb) * (nu + ng)
    blocks = data.reshape(-1, nu + ng)
    assert np.array_equal(blocks[: mesh.E, :nu].reshape(u.shape), u)
    assert np.array_equal(blocks[mesh.E:, :nu], u.reshape(mesh.E, nu)[bpar[:, 0]])
    got = blocks[mesh.E:, nu:].reshape(mesh.Fb, mesh.nb, mesh.nv * dim)
    ref = fg[bpar[:, 0], bpar[:, 1]]
    h = float(mesh.field("min_edge").min())
    err = P.parity_error(got.reshape(-1, mesh.nv * dim), ref.reshape(-1, mesh.nv * dim), u.reshape(-1, mesh.nv), dim,
                         h=h)
    print(f"  boundary gradient blocks vs oracle {err:.2e}", flush=True)
    assert err <= GRAD_TOL
    # the blocks differ from the parents' total gradients (the other faces' lifts are not in them)
    assert np.abs(got - s.gradient()[bpar[:, 0]]).max() > 1e-6 * np.abs(got).max()


def test_cpp_driver_restart_example(tmp_path):
    """The C++ drop-in driver writes raw files every io_interval steps; a LastStep restart from the
    middle file ends bit-identical, and an order-4 SpecificFile start reads the order-3 file."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "examples", "bin",
                       "shear_layer_2d_cns_restart")
    r = subprocess.run([exe, str(tmp_path), "8", "10"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("max |difference| = 0.000e+00") == 2
    assert sorted(os.listdir(tmp_path / "raw")) == ["shear_0.zst", "shear_10.zst", "shear_5.zst"]
    # the monitor log in the reference's format (CommandLine.cpp:70-127): header, time 0, one line
    # per step; the restart without setDeltaTime took its time step from the time of step 1
    lines = (tmp_path / "error.txt").read_text().splitlines()
    assert lines[0] == "|    Time     |     rho     |    rho*u    |    rho*v    |    rho*E    |"
    assert lines[1] == "|" + "|".join([" 0.00000e+00 "] * 5) + "|"
    assert len(lines) == 12
    for k, line in enumerate(lines[2:], start=1):
        cells = line.split("|")[1:-1]
        assert len(cells) == 5 and all(len(c) == 13 for c in cells)
        assert float(cells[0]) == float(f"{k * 2.0e-5:.5e}")
    assert "logged time step 2.000000e-05" in r.stdout
