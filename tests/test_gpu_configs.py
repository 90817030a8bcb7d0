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


def test_c5_mixed_mode_full_size():
    """64^3 hex P4 Taylor-Green: 10 steps of the mixed mode within 1e-6 (relative L2 over the domain
    per variable group) of the FP64 engine."""
    print("C5 mixed vs FP64, 64^3 P4", flush=True)
    mesh, prob, prim = P.ns_tgv(n=64, order=4)
    pm = L.ProblemDesc.from_buffer_copy(prob)
    pm.precision = L.PREC_MIXED
    a = L.Solver(prob, mesh)
    a.set_initial_primitive(prim)
    u0 = a.get_state()
    dt = a.compute_dt(0.1)
    a.step(dt, 10)
    ua = a.get_state()
    a.close()
    b = L.Solver(pm, mesh)
    b.set_state(u0)
    b.step(dt, 10)
    ub = b.get_state()
    b.close()
    dev = _chunked_l2(mesh, ub, ua)
    print(f"  mixed L2 deviation {dev:.3e}", flush=True)
    assert dev <= 1e-6


def _chunked_l2(mesh, a, b, chunk=32768):
    """P.function_l2 in element chunks (the point values of 64^3 P4 are 1.3 GB per array)."""
    V = mesh.field("basis_value").reshape(mesh.nq, mesh.nb)
    jw = mesh.field("jacobian_det_w").reshape(mesh.E, mesh.nq)
    worst = 0.0
    for g in ([0], [1, 2, 3], [4]):
        num = den = 0.0
        for c0 in range(0, mesh.E, chunk):
            c1 = min(mesh.E, c0 + chunk)
            qa = np.einsum("qb,ebv->eqv", V, (a[c0:c1] - b[c0:c1])[..., g])
            qb = np.einsum("qb,ebv->eqv", V, b[c0:c1][..., g])
            num += float((jw[c0:c1, :, None] * qa ** 2).sum())
            den += float((jw[c0:c1, :, None] * qb ** 2).sum())
        worst = max(worst, np.sqrt(num / den))
    return worst


def _tgv_primitive(mesh, amplitude):
    cp, cv = P.tgv_thermo(0.1)
    p0 = cp - cv
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, 3)
    x, y, z = xq[..., 0], xq[..., 1], xq[..., 2]
    p = p0 + amplitude ** 2 * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0) / 16.0
    prim = np.zeros((mesh.E, mesh.nq, 5))
    prim[..., 0] = p / p0
    prim[..., 1] = amplitude * np.sin(x) * np.cos(y) * np.cos(z)
    prim[..., 2] = -amplitude * np.cos(x) * np.sin(y) * np.cos(z)
    prim[..., 4] = 1.0
    return prim


def test_c5_adjoint_amplitude_gradient_full_size():
    """64^3 hex P4 Taylor-Green (HLLC + BR2, FP64): the gradient of the time-integrated kinetic-energy
    dissipation K(0) - K(T) over 4 SSP-RK3 steps with respect to the initial velocity amplitude A
    (u scaled by A, the pressure perturbation by A^2), from the checkpointed adjoint with 2 snapshots
    (binomial recomputation), equals the central difference of the objective in A (step 1e-3) to
    1e-5 relative."""
    print("C5 adjoint, 64^3 P4, 4 steps", flush=True)
    steps = 4
    mesh, prob, _ = P.ns_tgv(n=64, order=4)
    s = L.Solver(prob, mesh)

    def u_of(amp):
        s.set_initial_primitive(_tgv_primitive(mesh, amp))
        return s.get_state()

    u0 = u_of(1.0)
    dt = s.compute_dt(0.1)
    K0, dK0 = P.kinetic_energy(mesh, u0)
    s.adjoint_forward(dt, steps, 2)
    KT, dKT = P.kinetic_energy(mesh, s.get_state())
    lam0 = s.adjoint_backward(-dKT)
    h = 1e-4
    du0 = (u_of(1.0 + h) - u_of(1.0 - h)) / (2 * h)
    g_adj = float(((dK0 + lam0) * du0).sum())
    eps = 1e-3
    js = []
    for sign in (1.0, -1.0):
        ua = u_of(1.0 + sign * eps)
        ka = P.kinetic_energy(mesh, ua, gradient=False)[0]
        s.step(dt, steps)
        js.append(ka - P.kinetic_energy(mesh, s.get_state(), gradient=False)[0])
    g_fd = (js[0] - js[1]) / (2 * eps)
    rel = abs(g_adj - g_fd) / abs(g_fd)
    print(f"  dissipation {K0 - KT:.6e}: adjoint {g_adj:.10e} fd {g_fd:.10e} rel {rel:.2e}", flush=True)
    assert rel <= 1e-5
