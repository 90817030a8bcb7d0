"""1-D Euler and Navier-Stokes on line meshes against the oracle at the FP64 bar (periodic and
bounded), and the reference's 1-D shock tube (sod_1d_ceuler: HLLC, artificial viscosity) against
the exact Riemann solution."""
import math

import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L
from test_gpu_av import _sod_exact

pytestmark = pytest.mark.gpu

CASES = [(L.EQ_EULER, 2, None), (L.EQ_EULER, 4, [(L.BC_RIEMANN_FARFIELD, L.BC_ADIABATIC_SLIP_WALL)]),
         (L.EQ_NS, 3, None), (L.EQ_NS, 3, [(L.BC_RIEMANN_FARFIELD, L.BC_ADIABATIC_NONSLIP_WALL)])]


@pytest.mark.parametrize("eq,order,bc", CASES, ids=["euler-p2", "euler-p4-bounded", "ns-p3", "ns-p3-bounded"])
def test_line_parity(eq, order, bc):
    print(f"line parity eq={eq} P{order} bc={bc is not None}", flush=True)
    m = L.Mesh(L.ELEM_LINE, order, (12,), (0.0,), (1.0,), boundary=bc)
    if eq == L.EQ_EULER:
        prob = L.problem_desc(equation=eq, convective_flux=L.FLUX_ROE, cp=3.5, cv=2.5)
    else:
        prob = L.problem_desc(equation=eq, convective_flux=L.FLUX_HLLC, viscous_flux=L.VISC_BR2,
                              transport=L.TRANSPORT_CONSTANT, cp=3.5, cv=2.5, dynamic_viscosity=2e-3)
    x = m.field("quad_coord").reshape(m.E, m.nq, 1)[..., 0]
    prim = np.stack([1.0 + 0.2 * np.sin(2 * np.pi * x), 0.3 + 0.1 * np.cos(2 * np.pi * x),
                     1.0 + 0.05 * np.sin(4 * np.pi * x)], axis=-1)
    s = L.Solver(prob, m)
    s.set_initial_primitive(prim)
    o, xo = O.Oracle(prob, m), O.Oracle(prob, m, extended=True)
    bprim = np.tile(np.array([1.0, 0.3, 1.0]), (m.Fb, m.nfq, 1))
    if m.Fb:
        s.set_boundary_primitive(bprim)
    for k in (o, xo):
        k.set_state(s.get_state())
        if m.Fb:
            k.set_boundary_primitive(bprim)
    u0 = o.get_state()
    P.fp64_parity(s.residual(), o.residual(), xo.residual(), u0, 1, tol=1e-12)
    dt = o.compute_dt(0.1)
    assert abs(s.compute_dt(0.1) - dt) <= 1e-13 * dt
    o.step(dt, 5)
    xo.step(dt, 5)
    _, nan, taken = s.step(dt, 5)
    assert nan == 0 and taken == 5
    uo = o.get_state()
    P.fp64_parity(s.get_state(), uo, xo.get_state(), uo, 1, tol=1e-12)


def test_sod_1d_with_artificial_viscosity():
    """Sod's tube on [0, 1] (200 lines P2, far-field ends) to t = 0.2 with
    setArtificialViscosity(0.5): node values match the oracle after 5 steps, and at the end the
    density's L1 error against the exact solution is that of a shock-capturing scheme."""
    order, n = 2, 200
    prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_HLLC, cp=3.5, cv=2.5,
                          shock_capturing=L.SHOCK_ARTIFICIAL_VISCOSITY, empirical_tolerance=0.5)
    m = L.Mesh(L.ELEM_LINE, order, (n,), (0.0,), (1.0,), boundary=[(L.BC_RIEMANN_FARFIELD,) * 2])
    xq = m.field("quad_coord").reshape(m.E, m.nq, 1)
    s = L.Solver(prob, m)
    s.set_initial_primitive(P.sod_primitive(xq, 0.5))
    bprim = P.sod_primitive(P.boundary_points(m), 0.5)
    s.set_boundary_primitive(bprim)
    o = O.Oracle(prob, m)
    o.set_state(s.get_state())
    o.set_boundary_primitive(bprim)
    dt0 = o.compute_dt(0.1)
    o.step(dt0, 5)
    s.step(dt0, 5)
    assert np.abs(s.artificial_viscosity() - o.artificial_viscosity()).max() <= 1e-12 * o.artificial_viscosity().max()
    P.fp64_parity(s.get_state(), o.get_state(), o.get_state(), o.get_state(), 1, tol=1e-11)
    t_end = 0.2
    steps = math.ceil((t_end - 5 * dt0) / s.compute_dt(0.1))
    dt = (t_end - 5 * dt0) / steps
    _, nan, taken = s.step(dt, steps)
    assert nan == 0 and taken == steps
    V = m.field("basis_value").reshape(m.nq, m.nb)
    jw = m.field("jacobian_det_w").reshape(m.E, m.nq)
    rho = np.einsum("qb,eb->eq", V, s.get_state()[:, :, 0])
    l1 = float((np.abs(rho - _sod_exact(xq[..., 0], t_end, 0.5)) * jw).sum() / jw.sum())
    print(f"  1-D Sod L1(rho) = {l1:.4f}", flush=True)
    assert rho.min() > 0.1 and l1 < 0.01


def test_cpp_sod_1d_example():
    """DimensionEnum::D1 with artificial viscosity through the C++ drop-in driver."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "examples", "bin",
                       "sod_1d_ceuler")
    r = subprocess.run([exe, "200"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "sod 1-D: 200 cells" in r.stdout
