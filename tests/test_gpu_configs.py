"""The benchmark configurations at their full sizes (32^2 quad P2 Euler vortex; 64^2 quad P3 NS
shear layer; 32^3 hex P3 NS Taylor-Green; 64^3 hex P4 NS Taylor-Green): element-by-element parity
with the oracle where it finishes in seconds, and size-independent properties where it does not --
discrete conservation of mass, momentum and energy on the periodic box (to rounding of the sums),
free-stream preservation, and the mixed mode's deviation from the FP64 engine."""
import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L

pytestmark = pytest.mark.gpu


def _integrals(mesh, U):
    """Domain integrals of the conserved variables, int U = sum_e sum_q |J| w_q (V C_e)_q."""
    V = mesh.field("basis_value").reshape(mesh.nq, mesh.nb)
    jw = mesh.field("jacobian_det_w").reshape(mesh.E, mesh.nq)
    return np.einsum("eb,ebv->v", jw @ V, U)


def test_c1_euler_vortex_32x32_p2_steps_match_the_oracle():
    mesh, prob, prim = P.euler_vortex(n=32, order=2)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    o = O.Oracle(prob, mesh)
    o.set_state(s.get_state())
    x = O.Oracle(prob, mesh, extended=True)
    x.set_state(s.get_state())
    dt = o.compute_dt(0.1)
    o.step(dt, 20)
    x.step(dt, 20)
    _, nan, taken = s.step(dt, 20)
    assert nan == 0 and taken == 20
    uo = o.get_state()
    P.fp64_parity(s.get_state(), uo, x.get_state(), uo, 2, tol=1e-12)


def test_c4_shear_layer_64x64_p3_steps_match_the_oracle():
    mesh, prob, prim = P.ns_shear(n=64, order=3, flux=L.FLUX_ROE, visc=L.VISC_BR2, perturb=True)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    o = O.Oracle(prob, mesh)
    o.set_state(s.get_state())
    x = O.Oracle(prob, mesh, extended=True)
    x.set_state(s.get_state())
    dt = o.compute_dt(0.1)
    o.step(dt, 5)
    x.step(dt, 5)
    s.step(dt, 5)
    uo = o.get_state()
    P.fp64_parity(s.get_state(), uo, x.get_state(), uo, 2, tol=1e-12)


@pytest.mark.parametrize("n,order,steps", [(32, 3, 20), (64, 4, 3)], ids=["c2-32cube-p3", "c5-64cube-p4"])
def test_taylor_green_full_size_conservation(n, order, steps):
    """Periodic box: the domain integrals of rho, rho u and rho E do not change (every face flux
    enters its two elements with opposite signs), to 1e-11 of the integral's scale (the rounding
    of sums over 2e6-3e7 point values)."""
    print(f"TGV {n}^3 P{order}", flush=True)
    mesh, prob, prim = P.ns_tgv(n=n, order=order)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    i0 = _integrals(mesh, s.get_state())
    dt = s.compute_dt(0.1)
    _, nan, taken = s.step(dt, steps)
    assert nan == 0 and taken == steps
    i1 = _integrals(mesh, s.get_state())
    scale = np.array([abs(i0[0]), abs(i0[0]), abs(i0[0]), abs(i0[0]), abs(i0[4])])
    assert np.all(np.abs(i1 - i0) <= 1e-11 * scale), (i1 - i0) / scale


@pytest.mark.parametrize("n,order", [(32, 3), (64, 4)], ids=["c2-32cube-p3", "c5-64cube-p4"])
def test_taylor_green_full_size_free_stream(n, order):
    """A uniform moving state is a steady solution: its residual vanishes to rounding."""
    mesh, prob, _ = P.ns_tgv(n=n, order=order)
    prim = np.zeros((mesh.E, mesh.nq, 5))
    prim[..., 0] = 1.0
    prim[..., 1] = 0.3
    prim[..., 2] = -0.2
    prim[..., 3] = 0.1
    prim[..., 4] = 1.0
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    r = s.residual()
    u = s.get_state()
    assert np.abs(r).max() <= 1e-10 * np.abs(u).max()


def test_c2_mixed_mode_full_size():
    """32^3 hex P3 Taylor-Green: 20 steps of the mixed mode within 1e-6 (relative L2 over the domain
    per variable group) of the FP64 engine."""
    mesh, prob, prim = P.ns_tgv(n=32, order=3)
    pm = L.ProblemDesc.from_buffer_copy(prob)
    pm.precision = L.PREC_MIXED
    a = L.Solver(prob, mesh)
    a.set_initial_primitive(prim)
    b = L.Solver(pm, mesh)
    b.set_state(a.get_state())
    dt = a.compute_dt(0.1)
    a.step(dt, 20)
    b.step(dt, 20)
    dev = P.function_l2(mesh, b.get_state(), a.get_state())
    print(f"  mixed L2 deviation {dev:.3e}", flush=True)
    assert dev <= 1e-6
