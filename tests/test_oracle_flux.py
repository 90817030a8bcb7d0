"""Pointwise physics of the CPU oracle, pinned by properties fixed by the mathematics.

* every numerical flux is consistent, H(u, u; n) = F(u).n, conservative,
  H(uL, uR; n) = -H(uR, uL; -n), and rotation invariant;
* Roe equals 1/2 (F_L + F_R) - 1/2 |A(u_Roe)| (U_R - U_L), with |A| built here from a numpy
  eigen-decomposition of the complex-step flux Jacobian at the Roe state (independent of the
  reference's wave-by-wave formulas, ConvectiveFlux.cpp:240-350), and the Roe state satisfies
  A (U_R - U_L) = F_R - F_L;
* HLLC upwinds supersonic states and resolves a stationary contact exactly;
* Lax-Friedrichs matches its closed form;
* the viscous flux vanishes for uniform states and equals mu du/dy and kappa dT/dx for a pure
  shear and a pure conduction gradient.
"""
import numpy as np
import pytest

import oracle as O
from subrosadg_amd import _lib as L

GAMMA = 1.4
FLUXES = [L.FLUX_CENTRAL, L.FLUX_LAX_FRIEDRICHS, L.FLUX_HLLC, L.FLUX_ROE]
PROB = L.problem_desc(cp=3.5, cv=2.5)


def conserved(rho, vel, p):
    vel = np.asarray(vel, dtype=complex if np.iscomplexobj(vel) or np.iscomplexobj(rho) else float)
    e = p / ((GAMMA - 1) * rho)
    return np.concatenate([[rho], rho * vel, [rho * (e + 0.5 * np.dot(vel, vel))]])


def physical_flux(u, n):
    """F(U).n for the ideal gas, written for complex-step differentiation."""
    dim = len(n)
    rho = u[0]
    vel = u[1:dim + 1] / rho
    p = (GAMMA - 1) * (u[dim + 1] - 0.5 * rho * np.dot(vel, vel))
    un = np.dot(vel, n)
    return np.concatenate([[rho * un], rho * un * vel + p * np.asarray(n), [(u[dim + 1] + p) * un]])


def random_state(rng, dim, mach=0.4):
    rho = rng.uniform(0.5, 2.0)
    p = rng.uniform(0.5, 2.0)
    c = np.sqrt(GAMMA * p / rho)
    d = rng.normal(size=dim)
    vel = mach * c * rng.uniform(0, 1) * d / np.linalg.norm(d)
    return conserved(rho, vel, p)


def random_normal(rng, dim):
    n = rng.normal(size=dim)
    return n / np.linalg.norm(n)


def rotation(rng, dim):
    q, r = np.linalg.qr(rng.normal(size=(dim, dim)))
    q = q @ np.diag(np.sign(np.diag(r)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def rotate_state(u, Q):
    dim = Q.shape[0]
    v = u.copy()
    v[1:dim + 1] = Q @ u[1:dim + 1]
    return v


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("flux", FLUXES)
def test_consistency_conservation_rotation(dim, flux):
    rng = np.random.default_rng(10 * dim + flux)
    for _ in range(32):
        ul, ur, n = random_state(rng, dim), random_state(rng, dim), random_normal(rng, dim)
        h = O.point_flux(PROB, dim, flux, n, ul, ul)
        np.testing.assert_allclose(h, physical_flux(ul, n), rtol=1e-13, atol=1e-13)
        h = O.point_flux(PROB, dim, flux, n, ul, ur)
        hm = O.point_flux(PROB, dim, flux, -n, ur, ul)
        np.testing.assert_allclose(h, -hm, rtol=1e-13, atol=1e-13)
        Q = rotation(rng, dim)
        hq = O.point_flux(PROB, dim, flux, Q @ n, rotate_state(ul, Q), rotate_state(ur, Q))
        np.testing.assert_allclose(hq, rotate_state(h, Q), rtol=1e-12, atol=1e-12)


def complex_step_jacobian(u, n):
    m = len(u)
    A = np.zeros((m, m))
    for j in range(m):
        du = np.zeros(m, dtype=complex)
        du[j] = 1e-30j
        A[:, j] = physical_flux(u.astype(complex) + du, n).imag / 1e-30
    return A


def roe_state(ul, ur, dim):
    def prim(u):
        rho = u[0]
        vel = u[1:dim + 1] / rho
        p = (GAMMA - 1) * (u[dim + 1] - 0.5 * rho * vel @ vel)
        return rho, vel, (u[dim + 1] + p) / rho
    rl, vl, hl = prim(ul)
    rr, vr, hr = prim(ur)
    sl, sr = np.sqrt(rl), np.sqrt(rr)
    vel = (sl * vl + sr * vr) / (sl + sr)
    H = (sl * hl + sr * hr) / (sl + sr)
    rho = np.sqrt(rl * rr)
    E = (H + (GAMMA - 1) * 0.5 * vel @ vel) / GAMMA  # total energy with enthalpy H
    return np.concatenate([[rho], rho * vel, [rho * E]]), vel, H


@pytest.mark.parametrize("dim", [2, 3])
def test_roe_against_eigendecomposition(dim):
    rng = np.random.default_rng(7 + dim)
    checked = 0
    for _ in range(200):
        ul, ur, n = random_state(rng, dim), random_state(rng, dim), random_normal(rng, dim)
        ut, vel, H = roe_state(ul, ur, dim)
        c = np.sqrt((GAMMA - 1) * (H - 0.5 * vel @ vel))
        un = vel @ n
        if min(abs(un - c), abs(un + c)) <= c / 20 * 1.01:
            continue  # Harten's entropy fix active (ConvectiveFlux.cpp:300-312)
        A = complex_step_jacobian(ut, n)
        # Roe property: A(U_roe) (U_R - U_L) = F_R - F_L
        np.testing.assert_allclose(A @ (ur - ul), physical_flux(ur, n) - physical_flux(ul, n), rtol=1e-10,
                                   atol=1e-10)
        lam, R = np.linalg.eig(A)
        absA = (R @ np.diag(np.abs(lam)) @ np.linalg.inv(R)).real
        expect = 0.5 * (physical_flux(ul, n) + physical_flux(ur, n)) - 0.5 * absA @ (ur - ul)
        got = O.point_flux(PROB, dim, L.FLUX_ROE, n, ul, ur)
        np.testing.assert_allclose(got, expect, rtol=1e-10, atol=1e-10)
        checked += 1
    assert checked > 100


@pytest.mark.parametrize("dim", [2, 3])
def test_lax_friedrichs_closed_form(dim):
    rng = np.random.default_rng(3)
    for _ in range(32):
        ul, ur, n = random_state(rng, dim), random_state(rng, dim), random_normal(rng, dim)

        def speed(u):
            rho = u[0]
            vel = u[1:dim + 1] / rho
            p = (GAMMA - 1) * (u[dim + 1] - 0.5 * rho * vel @ vel)
            return abs(vel @ n) + np.sqrt(GAMMA * p / rho)
        lam = max(speed(ul), speed(ur))
        expect = 0.5 * (physical_flux(ul, n) + physical_flux(ur, n)) - 0.5 * lam * (ur - ul)
        np.testing.assert_allclose(O.point_flux(PROB, dim, L.FLUX_LAX_FRIEDRICHS, n, ul, ur), expect, rtol=1e-13,
                                   atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_hllc_upwinding_and_contact(dim):
    rng = np.random.default_rng(5)
    for _ in range(16):
        n = random_normal(rng, dim)
        rl, rr, pl, pr = rng.uniform(0.5, 2, 4)
        cl, cr = np.sqrt(GAMMA * pl / rl), np.sqrt(GAMMA * pr / rr)
        # both states supersonic along +n: flux is the left physical flux
        ul, ur = conserved(rl, 3 * cl * n, pl), conserved(rr, 3.2 * cr * n, pr)
        np.testing.assert_allclose(O.point_flux(PROB, dim, L.FLUX_HLLC, n, ul, ur), physical_flux(ul, n),
                                   rtol=1e-13, atol=1e-13)
        # ... and along -n: the right physical flux
        ul, ur = conserved(rl, -3.2 * cl * n, pl), conserved(rr, -3 * cr * n, pr)
        np.testing.assert_allclose(O.point_flux(PROB, dim, L.FLUX_HLLC, n, ul, ur), physical_flux(ur, n),
                                   rtol=1e-13, atol=1e-13)
        # stationary contact: equal pressure, no motion, density jump -> (0, p n, 0) exactly
        ul, ur = conserved(rl, np.zeros(dim), pl), conserved(rr, np.zeros(dim), pl)
        expect = np.concatenate([[0.0], pl * n, [0.0]])
        np.testing.assert_allclose(O.point_flux(PROB, dim, L.FLUX_HLLC, n, ul, ur), expect, atol=1e-13)
        # a contact moving with tangential velocity only: mass flux vanishes
        t = np.linalg.svd(n.reshape(1, -1))[2][-1]
        ul, ur = conserved(rl, 0.3 * t, pl), conserved(rr, -0.2 * t, pl)
        h = O.point_flux(PROB, dim, L.FLUX_HLLC, n, ul, ur)
        assert abs(h[0]) < 1e-13


def test_exact_weak_compressible_consistency():
    """Exact flux of the weakly-compressible model (ConvectiveFlux.cpp:352-414): only consistency
    and conservation are pinned here (parity unpinned beyond these properties)."""
    prob = L.problem_desc(equation_of_state=L.EOS_WEAK_COMPRESSIBLE, reference_sound_speed=10.0,
                          reference_density=1.0, cp=3.5, cv=2.5)
    rng = np.random.default_rng(11)
    for dim in (2, 3):
        for _ in range(16):
            n = random_normal(rng, dim)
            rho = rng.uniform(0.9, 1.1, 2)
            vel = rng.normal(size=(2, dim)) * 0.3
            T = rng.uniform(0.5, 2, 2)
            us = [np.concatenate([[rho[i]], rho[i] * vel[i], [rho[i] * (2.5 * T[i] + 0.5 * vel[i] @ vel[i])]])
                  for i in range(2)]
            h_same = O.point_flux(prob, dim, L.FLUX_EXACT, n, us[0], us[0])
            p = 100.0 * (rho[0] - 1.0) + 0.01 * 100.0
            un = vel[0] @ n
            expect = np.concatenate([[rho[0] * un], rho[0] * un * vel[0] + p * n,
                                     [(us[0][-1] + p) * un]])
            np.testing.assert_allclose(h_same, expect, rtol=1e-12, atol=1e-12)
            h = O.point_flux(prob, dim, L.FLUX_EXACT, n, us[0], us[1])
            hm = O.point_flux(prob, dim, L.FLUX_EXACT, -n, us[1], us[0])
            np.testing.assert_allclose(h, -hm, rtol=1e-12, atol=1e-12)


def test_viscous_flux_closed_forms():
    mu, cp, cv = 0.01, 3.5, 2.5
    prob = L.problem_desc(equation=L.EQ_NS, viscous_flux=L.VISC_BR2, transport=L.TRANSPORT_CONSTANT,
                          cp=cp, cv=cv, dynamic_viscosity=mu)
    kappa = cp * mu / 0.71
    for dim in (2, 3):
        nv = dim + 2
        rho, T = 1.3, 0.8
        vel = np.arange(1, dim + 1) * 0.1
        u = np.concatenate([[rho], rho * vel, [rho * (cv * T + 0.5 * vel @ vel)]])
        n = np.zeros(dim)
        n[1] = 1.0
        # uniform state: zero gradient -> zero viscous flux
        np.testing.assert_allclose(O.point_viscous_flux(prob, dim, n, u, np.zeros(nv * dim)), 0.0, atol=1e-15)
        # pure shear du/dy = s at uniform density and temperature
        s = 0.7
        g = np.zeros(nv * dim)
        g[1 * dim + 1] = rho * s                     # d(rho u)/dy
        g[(dim + 1) * dim + 1] = rho * vel[0] * s    # d(rho E)/dy = rho u du/dy
        f = O.point_viscous_flux(prob, dim, n, u, g)
        expect = np.zeros(nv)
        expect[1] = mu * s                            # tau_xy
        expect[dim + 1] = mu * s * vel[0]             # tau_yx u
        np.testing.assert_allclose(f, expect, rtol=1e-12, atol=1e-14)
        # pure conduction dT/dy = t at rest
        u0 = np.concatenate([[rho], np.zeros(dim), [rho * cv * T]])
        t = 0.4
        g = np.zeros(nv * dim)
        g[(dim + 1) * dim + 1] = rho * cv * t
        f = O.point_viscous_flux(prob, dim, n, u0, g)
        expect = np.zeros(nv)
        expect[dim + 1] = kappa * t
        np.testing.assert_allclose(f, expect, rtol=1e-12, atol=1e-14)
