"""Closed-form pins of oracle functions that the whole-scheme invariants cannot see.

* BR2 / BR1 face lift: for a piecewise-constant state with a jump [u] = u_ext - u_int across one
  face of an element (outward normal n, width h normal to the face), the lift of a face value g
  onto the degree-P polynomials of the element has, at that face, the value g (P+1)^2 / h and, at
  the opposite face, g (-1)^P (P+1) / h (the reproducing kernel of the Legendre polynomials at the
  end points). With the reference's jump term g = [u] n / 2 (SpatialDiscrete.cpp:899-906) and
  face lift factor 1 (VariableConvertor.cpp:674-688), the BR2 face gradient at the jump face is
  [u] n (P+1)^2 / h and at the opposite face [u] n (-1)^P (P+1) / (2h); BR1's is twice that at
  the opposite face. Lines P1-P5 and the tensor analogue on quadrangles.
* CFL time step: on a uniform state dt = CFL h / ((|u| + c)(P+1)^2), h the minimum edge
  (TimeIntegration.cpp:104-131), ideal gas and weakly-compressible EOS.
* Residual monitor: mean over the points of |R V^T| averaged over the elements
  (TimeIntegration.cpp:279-324): an element-constant residual c_e gives mean_e |c_e|; a residual
  a + b xi (|b| < a, symmetric rule) gives a; the monitor of a forward-Euler step is that of the
  residual of its start state.
* Exact flux of the weakly-compressible model (ConvectiveFlux.cpp:352-414): the star state satisfies
  the two rarefaction invariants u* + c0 ln rho* = u_L + c0 ln rho_L and
  u* - c0 ln rho* = u_R - c0 ln rho_R, the upwind side is chosen by the sign of u*, and the
  flux is the physical flux of that state.
* HLLC (ConvectiveFlux.cpp:136-238): for a symmetric pair (equal density and pressure, opposite
  normal velocities) the contact speed is 0, so the mass and energy fluxes vanish and the normal
  momentum flux equals the star pressure whatever it is. The star pressure itself is the
  reference's estimate max(0, (p_L+p_R)/2 - (u_R-u_L) rho_avg c_avg) (ConvectiveFlux.cpp:162-165),
  which differs from Toro's PVRS by its factor 1 on the acoustic term; that value is the
  reference's choice, checked here at symmetric pairs where it is the only term left, and the
  vacuum case (strong rarefaction) gives p* = 0 and a zero flux.
* Sutherland transport (PhysicalModel.cpp:102-122): mu(T) = mu T^(3/2) (1+S)/(T+S) with
  S = 110.4/273.15, kappa scaled by the same ratio, through the viscous flux of a pure shear and a
  pure conduction at T != 1.
"""
import math

import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L

GAMMA = 1.4


def _piecewise_constant_state(m, region_of_element, prims):
    """Conserved coefficients of a state constant on each element (exact projection)."""
    prim = np.zeros((m.E, m.nq, m.nv))
    for e in range(m.E):
        prim[e, :, :] = prims[region_of_element[e]]
    return prim


def _ns_problem(visc, cv=2.5, cp=3.5, mu=0.01):
    return L.problem_desc(equation=L.EQ_NS, convective_flux=L.FLUX_ROE, viscous_flux=visc,
                          transport=L.TRANSPORT_CONSTANT, cp=cp, cv=cv, dynamic_viscosity=mu)


def _conserved(prim, dim, cv=2.5):
    rho, vel, T = prim[0], np.asarray(prim[1:dim + 1]), prim[dim + 1]
    return np.concatenate([[rho], rho * vel, [rho * (cv * T + 0.5 * vel @ vel)]])


@pytest.mark.parametrize("visc", [L.VISC_BR2, L.VISC_BR1])
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_face_lift_closed_form_on_lines(order, visc):
    n, h = 8, 1.0 / 8
    m = L.Mesh(L.ELEM_LINE, order, (n,), (0.0,), (1.0,))
    prims = [np.array([1.0, 0.3, 1.0]), np.array([1.3, -0.2, 0.7])]
    xc = m.field("quad_coord").reshape(m.E, m.nq)[:, :].mean(1)
    region = (xc > 0.5).astype(int)  # jumps at x = 0.5 and (periodically) x = 0
    o = O.Oracle(_ns_problem(visc), m)
    o.set_initial_primitive(_piecewise_constant_state(m, region, prims))
    o.residual()
    fg = o.face_gradient()  # E x 2 x 1 x NV
    u = [_conserved(p, 1) for p in prims]
    factor = 1.0 if visc == L.VISC_BR2 else 2.0  # BR1: the whole lifted gradient at the far face
    scale = np.abs(fg).max()
    # local face 0 is the element's x-low end (outward normal -1), face 1 its x-high end (+1)
    for e in range(m.E):
        expect = np.zeros((2, m.nv))
        for s, (nbr, nrm) in enumerate((((e - 1) % n, -1.0), ((e + 1) % n, 1.0))):
            jump = u[region[nbr]] - u[region[e]]
            if not np.any(jump):
                continue
            expect[s] += jump * nrm * (order + 1) ** 2 / h                        # at the jump face
            expect[1 - s] += factor * jump * nrm * (-1) ** order * (order + 1) / (2 * h)  # at the far face
        np.testing.assert_allclose(fg[e, :, 0, :], expect, rtol=1e-12, atol=1e-13 * scale)


@pytest.mark.parametrize("visc", [L.VISC_BR2, L.VISC_BR1])
@pytest.mark.parametrize("order", [2, 3])
def test_face_lift_closed_form_on_quadrangles(order, visc):
    """Tensor analogue: a state constant in y with jumps across vertical lines; the lift is a
    function of x only, so at the x-faces it takes the 1-D values with h = h_x, at every face point."""
    nx, ny = 8, 3
    hx = 1.0 / nx
    m = L.Mesh(L.ELEM_QUADRANGLE, order, (nx, ny), (0.0, 0.0), (1.0, 0.6))
    prims = [np.array([1.0, 0.3, -0.1, 1.0]), np.array([1.2, -0.2, 0.05, 0.8])]
    xq = m.field("quad_coord").reshape(m.E, m.nq, 2)
    ix = np.floor(xq[..., 0].mean(1) / hx).astype(int)
    region = (ix >= nx // 2).astype(int)
    o = O.Oracle(_ns_problem(visc), m)
    o.set_initial_primitive(_piecewise_constant_state(m, region, prims))
    o.residual()
    fg = o.face_gradient()  # E x 4 x N_fq x (NV*2), row var*2 + dim
    u = [_conserved(p, 2) for p in prims]
    factor = 1.0 if visc == L.VISC_BR2 else 2.0
    region_at = {i: (1 if i >= nx // 2 else 0) for i in range(nx)}
    scale = np.abs(fg).max()
    # local faces of the quadrangle: 1 = x-high (normal +x), 3 = x-low (normal -x)
    for e in range(m.E):
        i = ix[e]
        expect = {1: np.zeros(m.nv), 3: np.zeros(m.nv)}
        for s, nbr, nrm in ((3, (i - 1) % nx, -1.0), (1, (i + 1) % nx, 1.0)):
            jump = u[region_at[nbr]] - u[region[e]]
            if not np.any(jump):
                continue
            other = 1 if s == 3 else 3
            expect[s] += jump * nrm * (order + 1) ** 2 / hx
            expect[other] += factor * jump * nrm * (-1) ** order * (order + 1) / (2 * hx)
        for s in (1, 3):
            gx = fg[e, s, :, 0::2]  # d/dx of every conserved variable at every face point
            gy = fg[e, s, :, 1::2]
            np.testing.assert_allclose(gx, np.broadcast_to(expect[s], gx.shape), rtol=1e-12, atol=1e-13 * scale)
            np.testing.assert_allclose(gy, 0.0, atol=1e-13 * scale)


@pytest.mark.parametrize("kind", ["line", "quad", "tri", "hex"])
@pytest.mark.parametrize("order", [1, 2, 4])
@pytest.mark.parametrize("eos", [L.EOS_IDEAL_GAS, L.EOS_WEAK_COMPRESSIBLE])
def test_cfl_time_step_closed_form(kind, order, eos):
    if kind == "line":
        m, h = L.Mesh(L.ELEM_LINE, order, (7,), (0.0,), (1.4,)), 0.2
    elif kind == "quad":
        m, h = L.Mesh(L.ELEM_QUADRANGLE, order, (4, 5), (0.0, 0.0), (1.0, 1.0)), 0.2
    elif kind == "tri":
        m, h = L.Mesh.split(order, (4, 5), (0.0, 0.0), (1.0, 1.0), 1.0, 3), 0.2
    else:
        m, h = L.Mesh(L.ELEM_HEXAHEDRON, order, (2, 3, 4), (0.0,) * 3, (1.0, 1.2, 1.0)), 0.25
    assert np.allclose(m.field("min_edge"), h, rtol=1e-14)
    c0 = 7.0
    prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_ROE, cp=3.5, cv=2.5, equation_of_state=eos,
                          reference_sound_speed=c0, reference_density=1.1)
    dim = m.dim
    rho, T = 1.15, 0.9
    vel = np.array([0.3, -0.4, 0.2])[:dim]
    prim = np.zeros((m.E, m.nq, m.nv))
    prim[..., 0], prim[..., 1:dim + 1], prim[..., dim + 1] = rho, vel, T
    o = O.Oracle(prob, m)
    o.set_initial_primitive(prim)
    c = math.sqrt(GAMMA * (GAMMA - 1.0) * 2.5 * T) if eos == L.EOS_IDEAL_GAS else c0  # p = (g-1) rho cv T
    cfl = 0.37
    expect = cfl * h / ((np.linalg.norm(vel) + c) * (order + 1) ** 2)
    assert abs(o.compute_dt(cfl) - expect) <= 1e-12 * expect  # the projected uniform state carries rounding


@pytest.mark.parametrize("kind,order", [("line", 3), ("quad", 2), ("tri", 3), ("hex", 2)])
def test_monitor_closed_forms(kind, order):
    if kind == "line":
        m = L.Mesh(L.ELEM_LINE, order, (6,), (0.0,), (1.0,))
    elif kind == "quad":
        m = L.Mesh(L.ELEM_QUADRANGLE, order, (3, 4), (0.0, 0.0), (1.0, 1.0))
    elif kind == "tri":
        m = L.Mesh.split(order, (3, 3), (0.0, 0.0), (1.0, 1.0), 1.0, 2)
    else:
        m = L.Mesh(L.ELEM_HEXAHEDRON, order, (2, 2, 2), (0.0,) * 3, (1.0,) * 3)
    prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_ROE, cp=3.5, cv=2.5)
    o = O.Oracle(prob, m)
    V = m.field("basis_value").reshape(m.nq, m.nb)
    ones = np.linalg.lstsq(V, np.ones(m.nq), rcond=None)[0]  # coefficients of the constant 1
    assert np.allclose(V @ ones, 1.0, atol=1e-13)
    rng = np.random.default_rng(4)
    c = rng.normal(size=(m.E, m.nv))
    R = ones[None, :, None] * c[:, None, :]
    np.testing.assert_allclose(o.monitor(R), np.abs(c).mean(0), rtol=1e-13)
    # a + b xi_0 with |b| < a on a rule symmetric in xi_0: the mean of |.| over the points is a
    xi0 = m.field("quad_coord").reshape(m.E, m.nq, m.dim)[0, :, 0]
    xi0 = (xi0 - xi0.mean())
    if kind != "tri":  # the tensor rules are symmetric about the element centre
        lin = np.linalg.lstsq(V, xi0 / np.abs(xi0).max(), rcond=None)[0]
        a = np.abs(c) + 1.0
        b = 0.9 * np.sign(c)
        R = ones[None, :, None] * a[:, None, :] + lin[None, :, None] * b[:, None, :]
        np.testing.assert_allclose(o.monitor(R), a.mean(0), rtol=1e-13)
    # a forward-Euler step reports the monitor of its start state's residual
    mesh, prob_v, prim = P.euler_vortex(n=4, order=2, rk=L.RK_FORWARD_EULER)
    ov = O.Oracle(prob_v, mesh)
    ov.set_initial_primitive(prim)
    r0 = ov.residual()
    dt = ov.compute_dt(0.1)
    rel, _, _ = ov.step(dt, 1)
    np.testing.assert_allclose(rel, ov.monitor(r0), rtol=1e-15)


def _physical_flux_wc(rho, vel, rho_e, n, c0, rho0):
    """F(U).n of the weakly-compressible model, p = c0^2 (rho - rho0) + 0.01 rho0 c0^2
    (PhysicalModel.cpp:58-75)."""
    p = c0 * c0 * (rho - rho0) + 0.01 * rho0 * c0 * c0
    un = vel @ n
    rho_E = rho_e + 0.5 * rho * vel @ vel
    return np.concatenate([[rho * un], rho * un * vel + p * n, [(rho_E + p) * un]])


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_exact_flux_rarefaction_invariants(dim):
    c0, rho0, cv = 10.0, 1.0, 2.5
    prob = L.problem_desc(equation_of_state=L.EOS_WEAK_COMPRESSIBLE, reference_sound_speed=c0,
                          reference_density=rho0, cp=3.5, cv=cv)
    rng = np.random.default_rng(17 + dim)
    upwind_left = upwind_right = 0
    for _ in range(48):
        n = rng.normal(size=dim)
        n /= np.linalg.norm(n)
        rho_l, rho_r, rho_s = rng.uniform(0.9, 1.1, 3)
        u_s = rng.uniform(-0.5, 0.5)
        # left and right normal velocities on the two invariants through the star state
        u_l = u_s + c0 * (math.log(rho_s) - math.log(rho_l))
        u_r = u_s - c0 * (math.log(rho_s) - math.log(rho_r))
        vt = rng.normal(size=(2, dim)) * 0.3
        vt -= np.outer(vt @ n, n)  # tangential parts
        vel_l, vel_r = vt[0] + u_l * n, vt[1] + u_r * n
        T = rng.uniform(0.5, 2.0, 2)
        us = [np.concatenate([[r], r * v, [r * (cv * t + 0.5 * v @ v)]])
              for r, v, t in ((rho_l, vel_l, T[0]), (rho_r, vel_r, T[1]))]
        h = O.point_flux(prob, dim, L.FLUX_EXACT, n, us[0], us[1])
        # upwind side by the sign of u*: its tangential velocity and its internal energy per volume
        if u_s >= 0.0:
            vel_s, rho_e = vt[0] + u_s * n, rho_l * cv * T[0]
            upwind_left += 1
        else:
            vel_s, rho_e = vt[1] + u_s * n, rho_r * cv * T[1]
            upwind_right += 1
        expect = _physical_flux_wc(rho_s, vel_s, rho_e, n, c0, rho0)
        np.testing.assert_allclose(h, expect, rtol=1e-11, atol=1e-11 * np.abs(expect).max())
    assert upwind_left > 5 and upwind_right > 5


@pytest.mark.parametrize("dim", [2, 3])
def test_hllc_symmetric_pairs(dim):
    prob = L.problem_desc(cp=3.5, cv=2.5)
    rng = np.random.default_rng(23 + dim)
    for _ in range(24):
        n = rng.normal(size=dim)
        n /= np.linalg.norm(n)
        rho, p = rng.uniform(0.5, 2.0, 2)
        c = math.sqrt(GAMMA * p / rho)
        t = rng.normal(size=dim)
        t -= (t @ n) * n
        for sign in (1.0, -1.0):  # +1: two shocks (colliding), -1: two rarefactions
            u = sign * rng.uniform(0.05, 0.3) * c
            vl, vr = u * n + 0.2 * t, -u * n + 0.2 * t
            ul = np.concatenate([[rho], rho * vl, [p / (GAMMA - 1) + 0.5 * rho * vl @ vl]])
            ur = np.concatenate([[rho], rho * vr, [p / (GAMMA - 1) + 0.5 * rho * vr @ vr]])
            hflux = O.point_flux(prob, dim, L.FLUX_HLLC, n, ul, ur)
            p_star = max(0.0, p + 2.0 * u * rho * c)  # (p_L+p_R)/2 - (u_R-u_L) rho c with u_R-u_L = -2u
            scale = p + rho * c * c
            assert abs(hflux[0]) <= 1e-13 * scale and abs(hflux[-1]) <= 1e-12 * scale * c
            np.testing.assert_allclose(hflux[1:dim + 1], p_star * n, rtol=1e-12, atol=1e-13 * scale)
    # vacuum: a strong symmetric rarefaction clamps p* at 0 and the whole flux vanishes
    n = np.eye(dim)[0]
    rho, p = 1.0, 1.0
    c = math.sqrt(GAMMA * p / rho)
    u = 0.9 * c
    vl, vr = -u * n, u * n
    ul = np.concatenate([[rho], rho * vl, [p / (GAMMA - 1) + 0.5 * rho * u * u]])
    ur = np.concatenate([[rho], rho * vr, [p / (GAMMA - 1) + 0.5 * rho * u * u]])
    np.testing.assert_allclose(O.point_flux(prob, dim, L.FLUX_HLLC, n, ul, ur), 0.0, atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_sutherland_viscous_flux(dim):
    mu, cp, cv = 0.01, 3.5, 2.5
    S = 110.4 / 273.15
    prob = L.problem_desc(equation=L.EQ_NS, viscous_flux=L.VISC_BR2, transport=L.TRANSPORT_SUTHERLAND,
                          cp=cp, cv=cv, dynamic_viscosity=mu)
    nv = dim + 2
    for T in (0.6, 1.0, 1.7):
        ratio = T ** 1.5 * (1.0 + S) / (T + S)
        rho = 1.3
        vel = np.arange(1, dim + 1) * 0.1
        u = np.concatenate([[rho], rho * vel, [rho * (cv * T + 0.5 * vel @ vel)]])
        n = np.zeros(dim)
        n[1] = 1.0
        s = 0.7
        g = np.zeros(nv * dim)
        g[1 * dim + 1] = rho * s
        g[(dim + 1) * dim + 1] = rho * vel[0] * s
        f = O.point_viscous_flux(prob, dim, n, u, g)
        expect = np.zeros(nv)
        expect[1] = mu * ratio * s
        expect[dim + 1] = mu * ratio * s * vel[0]
        np.testing.assert_allclose(f, expect, rtol=1e-12, atol=1e-15)
        u0 = np.concatenate([[rho], np.zeros(dim), [rho * cv * T]])
        g = np.zeros(nv * dim)
        g[(dim + 1) * dim + 1] = rho * cv * 0.4
        f = O.point_viscous_flux(prob, dim, n, u0, g)
        expect = np.zeros(nv)
        expect[dim + 1] = cp * mu / 0.71 * ratio * 0.4
        np.testing.assert_allclose(f, expect, rtol=1e-12, atol=1e-15)
