"""Boundary conditions of the oracle (BoundaryCondition.cpp:79-547, SpatialDiscrete.cpp:750-968):
pointwise definitions of every boundary type, and the invariants they imply for the whole
residual (free-stream preservation, zero wall mass and energy flux)."""
import math

import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L

G = 1.4


def _prob():
    return L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_ROE, cp=3.5, cv=2.5)


def _conserved(prim):
    rho, u, T = prim[0], np.asarray(prim[1:-1]), prim[-1]
    return np.concatenate([[rho], rho * u, [rho * (2.5 * T + 0.5 * u @ u)]])


def _sound(rho, p):
    return math.sqrt(G * p / rho)


def _wl(prim):
    rho, u, T = prim[0], np.asarray(prim[1:-1]), prim[-1]
    e = 2.5 * T
    return rho, u, e, (G - 1.0) * rho * e


@pytest.mark.parametrize("inflow", [True, False])
def test_farfield_subsonic_characteristics(inflow):
    """Subsonic farfield: the outgoing invariant u_n + 2c/(g-1) comes from inside, the incoming
    u_n - 2c/(g-1) from the boundary data; entropy p/rho^g and tangential velocity from the
    upwind side (outside for inflow, inside for outflow)."""
    rng = np.random.default_rng(3)
    n = np.array([0.6, 0.8])
    sgn = -1.0 if inflow else 1.0
    inside = np.array([1.1, sgn * 0.3 * 0.6 + 0.1 * 0.8, sgn * 0.3 * 0.8 - 0.1 * 0.6, 0.9])
    outside = np.array([0.95, sgn * 0.25 * 0.6 - 0.2 * 0.8, sgn * 0.25 * 0.8 + 0.2 * 0.6, 1.05])
    wb = O.point_boundary_state(_prob(), 2, L.BC_RIEMANN_FARFIELD, n, _conserved(inside), outside)
    rb, ub, eb, pb = wb[0], wb[1:3], wb[3], wb[4]
    cb = _sound(rb, pb)
    ri, ui, ei, pi = _wl(inside)
    ro, uo, eo, po = _wl(outside)
    ci, co = _sound(ri, pi), _sound(ro, po)
    assert abs((ub @ n + 2 * cb / (G - 1)) - (ui @ n + 2 * ci / (G - 1))) < 1e-13
    assert abs((ub @ n - 2 * cb / (G - 1)) - (uo @ n - 2 * co / (G - 1))) < 1e-13
    src_r, src_u, src_p = (ro, uo, po) if inflow else (ri, ui, pi)
    assert abs(pb / rb ** G - src_p / src_r ** G) < 1e-13 * (src_p / src_r ** G)
    t = np.array([-n[1], n[0]])
    assert abs(ub @ t - src_u @ t) < 1e-14
    assert abs(pb - (G - 1) * rb * eb) < 1e-13 * pb
    del rng


@pytest.mark.parametrize("inflow", [True, False])
def test_farfield_supersonic_takes_one_side(inflow):
    n = np.array([1.0, 0.0])
    inside = np.array([1.0, -3.0 if inflow else 3.0, 0.1, 1.0])
    outside = np.array([1.2, -2.5 if inflow else 2.5, -0.1, 0.8])
    wb = O.point_boundary_state(_prob(), 2, L.BC_RIEMANN_FARFIELD, n, _conserved(inside), outside)
    src = outside if inflow else inside
    r, u, e, p = _wl(src)
    np.testing.assert_allclose(wb, [r, *u, e, p], rtol=1e-14)


def test_velocity_inflow_and_pressure_outflow():
    n = np.array([0.0, 1.0])
    inside = np.array([1.1, 0.1, -0.2, 0.9])  # subsonic, entering through the boundary (u.n < 0)
    data = np.array([0.9, 0.2, -0.3, 1.1])
    wb = O.point_boundary_state(_prob(), 2, L.BC_VELOCITY_INFLOW, n, _conserved(inside), data)
    r, u, e, p = _wl(data)
    np.testing.assert_allclose(wb[:4], [r, *u, e], rtol=1e-14)
    assert abs(wb[4] - _wl(inside)[3]) < 1e-14  # pressure from inside
    inside_out = np.array([1.1, 0.1, 0.2, 0.9])
    wb = O.point_boundary_state(_prob(), 2, L.BC_PRESSURE_OUTFLOW, n, _conserved(inside_out), data)
    r, u, e, p = _wl(inside_out)
    np.testing.assert_allclose(wb[:4], [r, *u, e], rtol=1e-14)
    assert abs(wb[4] - _wl(data)[3]) < 1e-14  # pressure from the boundary data


def test_walls():
    n = np.array([0.6, -0.8])
    inside = np.array([1.1, 0.4, 0.3, 0.9])
    data = np.array([2.0, 0.05, -0.02, 1.3])  # wall velocity and temperature
    ri, ui, ei, pi = _wl(inside)
    wb = O.point_boundary_state(_prob(), 2, L.BC_ADIABATIC_SLIP_WALL, n, _conserved(inside), data)
    assert abs(wb[1:3] @ n) < 1e-15
    t = np.array([-n[1], n[0]])
    assert abs(wb[1:3] @ t - ui @ t) < 1e-15
    np.testing.assert_allclose([wb[0], wb[3], wb[4]], [ri, ei, pi], rtol=1e-14)
    wb = O.point_boundary_state(_prob(), 2, L.BC_ADIABATIC_NONSLIP_WALL, n, _conserved(inside), data)
    np.testing.assert_allclose(wb, [ri, data[1], data[2], ei, pi], rtol=1e-14)
    wb = O.point_boundary_state(_prob(), 2, L.BC_ISOTHERMAL_NONSLIP_WALL, n, _conserved(inside), data)
    ew = 2.5 * data[3]
    np.testing.assert_allclose(wb, [ri, data[1], data[2], ew, (G - 1) * ri * ew], rtol=1e-14)


FREE_STREAM = [
    ("farfield-subsonic", ((L.BC_RIEMANN_FARFIELD,) * 2, (L.BC_RIEMANN_FARFIELD,) * 2), (1.0, 0.3, 0.2, 1.0)),
    ("farfield-supersonic", ((L.BC_RIEMANN_FARFIELD,) * 2, (L.BC_RIEMANN_FARFIELD,) * 2), (1.0, 2.5, 0.4, 1.0)),
    ("inflow-outflow", ((L.BC_VELOCITY_INFLOW, L.BC_PRESSURE_OUTFLOW), None), (1.0, 0.3, 0.0, 1.0)),
    ("slip-walls", (None, (L.BC_ADIABATIC_SLIP_WALL,) * 2), (1.0, 0.5, 0.0, 1.0)),
]


@pytest.mark.parametrize("name,bc,state", FREE_STREAM, ids=[c[0] for c in FREE_STREAM])
@pytest.mark.parametrize("flux", [L.FLUX_ROE, L.FLUX_HLLC, L.FLUX_LAX_FRIEDRICHS])
def test_free_stream_with_boundaries(name, bc, state, flux):
    """A uniform state whose boundary data equal it is a steady solution for every boundary type
    consistent with it: the residual vanishes to rounding."""
    mesh, prob, prim, bprim = P.euler_box(n=4, order=2, flux=flux, boundary=bc, state=state)
    o = O.Oracle(prob, mesh)
    o.set_initial_primitive(prim)
    o.set_boundary_primitive(bprim)
    r = o.residual()
    scale = 1.0 + np.abs(o.get_state()).max() * max(1.0, abs(state[1]))
    assert np.abs(r).max() <= 1e-13 * scale, np.abs(r).max()


def test_slip_walls_conserve_mass_and_energy():
    """With slip walls (and periodic sides) no mass or energy crosses the boundary: the residual
    tested with the constant function sums to zero over the mesh."""
    mesh = L.Mesh(L.ELEM_QUADRANGLE, 3, (6, 6), (-5.0, -5.0), (5.0, 5.0),
                  boundary=(None, (L.BC_ADIABATIC_SLIP_WALL,) * 2))
    prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_ROE, cp=3.5, cv=2.5)
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, 2)
    prim = P.isentropic_vortex_primitive(xq - np.array([0.0, 1.5]))  # off-centre: unequal wall pressures
    o = O.Oracle(prob, mesh)
    o.set_initial_primitive(prim)
    o.set_boundary_primitive(np.tile([1.0, 0.0, 0.0, 1.0], (mesh.Fb, mesh.nfq, 1)))
    r = o.residual()
    one = P.project_conserved(mesh, np.ones((mesh.E, mesh.nq, 1)))[..., 0]
    tot = np.einsum("eb,ebv->v", one, r)
    scale = np.abs(r).max()
    assert abs(tot[0]) <= 1e-12 * scale and abs(tot[3]) <= 1e-12 * scale, tot
    # the wall pressure force makes the momentum normal to the walls not conserved
    assert abs(tot[2]) > 1e-6 * scale


def test_ns_fluid_at_rest_between_isothermal_walls():
    """Navier-Stokes gas at rest between isothermal no-slip walls at the gas temperature: steady."""
    mesh = L.Mesh(L.ELEM_QUADRANGLE, 3, (4, 4), (0.0, 0.0), (1.0, 1.0),
                  boundary=(None, (L.BC_ISOTHERMAL_NONSLIP_WALL,) * 2))
    cp, cv = P.tgv_thermo(0.1)
    prob = L.problem_desc(equation=L.EQ_NS, convective_flux=L.FLUX_ROE, viscous_flux=L.VISC_BR2,
                          transport=L.TRANSPORT_CONSTANT, cp=cp, cv=cv, dynamic_viscosity=1e-3)
    prim = np.tile([1.0, 0.0, 0.0, 1.0], (mesh.E, mesh.nq, 1))
    o = O.Oracle(prob, mesh)
    o.set_initial_primitive(prim)
    o.set_boundary_primitive(np.tile([1.0, 0.0, 0.0, 1.0], (mesh.Fb, mesh.nfq, 1)))
    r = o.residual()
    assert np.abs(r).max() <= 1e-12 * np.abs(o.get_state()).max(), np.abs(r).max()
