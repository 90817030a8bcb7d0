"""The oracle's tangent-linear build (forward-mode numbers through the same passes) on the CPU:
its values are the FP64 oracle's, its derivative part is linear in the seed, and it equals central
differences of the FP64 oracle's residual and of its RK steps. It is the CPU side of the adjoint's
dot-product test (<J v, w> = <v, J^T w>, tests/test_gpu_adjoint.py)."""
import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L

CASES = [
    ("vortex-quad-p2-roe", lambda: P.euler_vortex(n=4, order=2)),
    ("vortex-quad-p3-hllc", lambda: P.euler_vortex(n=4, order=3, flux=L.FLUX_HLLC)),
    ("shear-quad-p3-roe-br2", lambda: P.ns_shear(n=4, order=3, perturb=True)),
    ("shear-quad-p3-hllc-br1", lambda: P.ns_shear(n=4, order=3, flux=L.FLUX_HLLC, visc=L.VISC_BR1, perturb=True)),
    ("tgv-hex-p2-lf-br2", lambda: P.ns_tgv(n=2, order=2, flux=L.FLUX_LAX_FRIEDRICHS)),
]


def _setup(make, seed=3):
    mesh, prob, prim = make()
    o = O.Oracle(prob, mesh)
    o.set_initial_primitive(prim)
    u0 = o.get_state()
    rng = np.random.default_rng(seed)
    # off the kinks of |u.n| and of the wave-speed bounds (see test_gpu_adjoint.py)
    scale = np.ones_like(u0)
    for g in ([0], list(range(1, mesh.dim + 1)), [mesh.dim + 1]):
        scale[..., g] = np.abs(u0[..., g]).max()
    u0 = u0 + 1e-3 * scale * rng.standard_normal(u0.shape)
    v = scale * rng.standard_normal(u0.shape)
    return mesh, prob, o, u0, v, rng, scale


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_tangent_residual(name, make):
    mesh, prob, o, u0, v, rng, scale = _setup(make)
    t = O.Oracle(prob, mesh, tangent=True)
    o.set_state(u0)
    r_fp64 = o.residual()
    r, jv = t.residual_tangent(u0, v)
    # the value part is the FP64 oracle's residual (same arithmetic; contraction of a*b+c into an
    # FMA may differ between the two builds)
    assert np.abs(r - r_fp64).max() <= 1e-13 * np.abs(r_fp64).max()
    # central differences of the FP64 oracle, O(eps^2) truncation and O(1e-16 / eps) rounding
    eps = 1e-6
    o.set_state(u0 + eps * v)
    rp = o.residual()
    o.set_state(u0 - eps * v)
    rm = o.residual()
    fd = (rp - rm) / (2 * eps)
    err = np.abs(jv - fd).max() / np.abs(fd).max()
    print(f"{name}: |Jv - FD| / |FD| = {err:.2e}")
    assert err < 1e-6
    # linear in the seed, to rounding
    v2 = scale * rng.standard_normal(u0.shape)
    _, jv2 = t.residual_tangent(u0, v2)
    _, jv12 = t.residual_tangent(u0, 2.0 * v - 3.0 * v2)
    assert np.abs(jv12 - (2.0 * jv - 3.0 * jv2)).max() <= 1e-12 * np.abs(jv12).max()


@pytest.mark.parametrize("name,make", CASES[2:4], ids=[c[0] for c in CASES[2:4]])
def test_tangent_steps(name, make):
    """d U_n / d U_0 v through 3 SSP-RK3 steps of fixed size against central differences."""
    mesh, prob, o, u0, v, rng, scale = _setup(make)
    o.set_state(u0)
    dt = o.compute_dt(0.1)
    t = O.Oracle(prob, mesh, tangent=True)
    un, dun = t.step_tangent(u0, v, dt, 3)
    o.set_state(u0)
    o.step(dt, 3)
    ref = o.get_state()
    assert np.abs(un - ref).max() <= 1e-13 * np.abs(ref).max()
    eps = 1e-6
    outs = []
    for sgn in (1.0, -1.0):
        o.set_state(u0 + sgn * eps * v)
        o.step(dt, 3)
        outs.append(o.get_state())
    fd = (outs[0] - outs[1]) / (2 * eps)
    err = np.abs(dun - fd).max() / np.abs(fd).max()
    print(f"{name}: |dU - FD| / |FD| = {err:.2e}")
    assert err < 1e-6


@pytest.mark.parametrize("flux", [L.FLUX_ROE, L.FLUX_HLLC, L.FLUX_LAX_FRIEDRICHS, L.FLUX_CENTRAL])
def test_tangent_euler_homogeneity(flux):
    """The ideal-gas Euler fluxes are homogeneous of degree one in the conserved state (scaling
    rho, rho u, rho E by s leaves u and c unchanged and scales p by s), and so is the residual:
    R(s U) = s R(U). Euler's theorem then fixes the tangent along the state itself: J(U) U = R(U)
    exactly, for every flux, with no finite-difference error in the check."""
    mesh, prob, prim = P.euler_vortex(n=4, order=3, flux=flux)
    t = O.Oracle(prob, mesh, tangent=True)
    o = O.Oracle(prob, mesh)
    o.set_initial_primitive(prim)
    u0 = o.get_state()
    r, ju = t.residual_tangent(u0, u0)
    assert np.abs(ju - r).max() <= 1e-12 * np.abs(r).max()
