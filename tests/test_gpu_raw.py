"""Raw state files from the GPU engine: the payload layout (per element the coefficients and, for
Navier-Stokes, the lifted-gradient coefficients; per boundary face its parent's blocks; nodal
artificial viscosity), restart from the last step (bit-identical continuation), and the start of
an order-P run from an order-(P-1) file (SpecificFile, InitialCondition.cpp:41-83)."""
import numpy as np
import pytest

import problems as P
from subrosadg_amd import _lib as L

pytestmark = pytest.mark.gpu


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


def test_br2_with_boundary_faces_is_refused(tmp_path):
    mesh, prob, prim, bprim = P.euler_box(n=3, order=3)
    cp, cv = P.tgv_thermo(0.1)
    ns = L.problem_desc(equation=L.EQ_NS, convective_flux=L.FLUX_ROE, viscous_flux=L.VISC_BR2,
                        transport=L.TRANSPORT_CONSTANT, cp=cp, cv=cv, dynamic_viscosity=1e-3)
    s = L.Solver(ns, mesh)
    s.set_initial_primitive(prim)
    s.set_boundary_primitive(bprim)
    with pytest.raises(L.SDGError):
        s.write_raw(str(tmp_path / "br2.zst"))


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
    assert "max |difference| = 0.000e+00" in r.stdout
    assert sorted(os.listdir(tmp_path / "raw")) == ["shear_0.zst", "shear_10.zst", "shear_5.zst"]
