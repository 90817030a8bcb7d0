"""The C++ drop-in driver examples without a test elsewhere (taylor_green_3d_cns, lid_driven_cavity_2d_cns):
each run writes its final raw state, which equals the same problem driven through the C-ABI from
Python (the same library calls: projection, CFL time step, steps) at the FP64 bar; the two sides
form their constants (p0 = 1/(gamma Ma^2), c_v, c_p) in different orders, so their inputs differ in
the last bit. The Python path is held to the oracle by the parity tests."""
import os
import subprocess

import numpy as np
import pytest

import problems as P
from subrosadg_amd import _lib as L

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "examples", "bin")


def _run(name, args, out):
    r = subprocess.run([os.path.join(BIN, name), *map(str, args), str(out)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-2000:], r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _raw_state(path, shape, dim):
    """The state blocks of a Navier-Stokes raw file: per element NV x N_b coefficients followed by
    its (NV D) x N_b gradient coefficients (RawBinary.cpp:76-90)."""
    payload = np.frombuffer(L.raw_file_read(str(path)), dtype=np.float64)
    E, nb, nv = shape
    nu, ng = nb * nv, nb * nv * dim
    return payload[:E * (nu + ng)].reshape(E, nu + ng)[:, :nu].reshape(shape)


def test_taylor_green_example(tmp_path):
    n, steps = 4, 6
    out = _run("taylor_green_3d_cns", [n, steps], tmp_path)
    assert out.count("step ") == steps
    mesh, prob, prim = P.ns_tgv(n=n, order=3)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    dt = s.compute_dt(0.1)
    s.step(dt, steps)
    ref = s.get_state()
    got = _raw_state(tmp_path / "raw" / f"tgv_{steps}.zst", ref.shape, 3)
    err = P.parity_error(got, ref, ref, 3)
    print(f"  driver vs python {err:.3e}", flush=True)
    assert err <= 1e-11


def test_lid_driven_cavity_example(tmp_path):
    n, steps = 6, 20
    _run("lid_driven_cavity_2d_cns", [n, steps], tmp_path)
    wall = L.BC_ISOTHERMAL_NONSLIP_WALL
    mesh = L.Mesh(L.ELEM_QUADRANGLE, 3, (n, n), (0.0, 0.0), (1.0, 1.0), boundary=((wall, wall), (wall, wall)))
    cp, cv = P.tgv_thermo(0.1)
    prob = L.problem_desc(equation=L.EQ_NS, convective_flux=L.FLUX_ROE, viscous_flux=L.VISC_BR2,
                          transport=L.TRANSPORT_CONSTANT, cp=cp, cv=cv, dynamic_viscosity=1.0 / 100.0)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(np.tile([1.0, 0.0, 0.0, 1.0], (mesh.E, mesh.nq, 1)))
    bn = mesh.field("boundary_face_normal").reshape(mesh.Fb, mesh.nfq, 2)
    b = np.tile([1.0, 0.0, 0.0, 1.0], (mesh.Fb, mesh.nfq, 1))
    b[..., 1] = np.where(bn[..., 1] > 0.5, 1.0, 0.0)  # the lid at y = 1
    s.set_boundary_primitive(b)
    dt = s.compute_dt(0.1)
    s.step(dt, steps)
    ref = s.get_state()
    got = _raw_state(tmp_path / "raw" / f"cavity_{steps}.zst", ref.shape, 2)
    err = P.parity_error(got, ref, ref, 2)
    print(f"  driver vs python {err:.3e}", flush=True)
    assert err <= 1e-11
