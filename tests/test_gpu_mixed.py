"""Mixed-precision mode (FP32 matrix-core contractions, FP64 state, geometry, pointwise physics and
RK accumulation) against the FP64 oracle.

Bars (written per test): the mixed residual agrees with the FP64 residual to FP32 contraction
accuracy, and the state after SSP-RK3 steps stays within 1e-6 relative L2 deviation of the FP64
oracle per variable group (density, momentum, energy), both in the L2 norm of the solution over
the domain and in the Euclidean norm of the modal coefficients.
"""
import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L

pytestmark = pytest.mark.gpu

# FP32 operands carry 2^-24 ~ 6e-8 relative rounding; a residual row sums ~N_q D + N_aq products,
# so its deviation from FP64 is O(1e-6) of the contributing terms (in units of the state / h).
RESIDUAL_TOL = 2e-5
# headline bar: relative L2 deviation of the stepped state from the FP64 oracle
STATE_L2_TOL = 1e-6


def _mixed(prob):
    q = L.ProblemDesc.from_buffer_copy(prob)
    q.precision = L.PREC_MIXED
    return q


def group_l2(a, b, dim):
    """max over variable groups (density, momentum, energy) of ||a - b||_2 / ||b||_2."""
    nv = dim + 2
    a2, b2 = a.reshape(-1, nv), b.reshape(-1, nv)
    out = []
    for g in ([0], list(range(1, dim + 1)), [dim + 1]):
        out.append(np.linalg.norm(a2[:, g] - b2[:, g]) / np.linalg.norm(b2[:, g]))
    return max(out)


CASES = [
    ("tgv-hex-p3-hllc-br2", lambda: P.ns_tgv(n=3, order=3)),
    ("tgv-hex-p2-hllc-br1", lambda: P.ns_tgv(n=3, order=2, visc=L.VISC_BR1)),
    ("shear-quad-p3-roe-br2", lambda: P.ns_shear(n=4, order=3, perturb=True)),
    ("vortex-quad-p3-roe", lambda: P.euler_vortex(n=6, order=3)),
    ("vortex-quad-p4-lf", lambda: P.euler_vortex(n=6, order=4, flux=L.FLUX_LAX_FRIEDRICHS)),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_mixed_residual_matches_fp64(name, make):
    print(f"mixed residual {name}", flush=True)
    mesh, prob, prim = make()
    s = L.Solver(_mixed(prob), mesh)
    s.set_initial_primitive(prim)
    u0 = s.get_state()
    o = O.Oracle(prob, mesh)
    o.set_state(u0)
    # Each group is measured against the largest group's scale (floor 1): a nearly vanishing
    # group (the mass residual of a nearly incompressible flow) carries the absolute FP32 noise of
    # the fluxes that cancel in it, not a relative error of its own size.
    groups = P.parity_error(s.residual(), o.residual(), u0, mesh.dim, floor=1.0, per_group=True)
    print(f"  parity error (density, momentum, energy) {groups}", flush=True)
    assert max(groups) <= RESIDUAL_TOL, groups


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_mixed_steps_l2_deviation(name, make):
    print(f"mixed steps {name}", flush=True)
    mesh, prob, prim = make()
    s = L.Solver(_mixed(prob), mesh)
    s.set_initial_primitive(prim)
    o = O.Oracle(prob, mesh)
    o.set_state(s.get_state())
    dt = o.compute_dt(0.1)
    o.step(dt, 20)
    _, nan, taken = s.step(dt, 20)
    assert nan == 0 and taken == 20
    us, uo = s.get_state(), o.get_state()
    dev = P.function_l2(mesh, us, uo)
    dev_coeff = group_l2(us, uo, mesh.dim)
    print(f"  relative L2 deviation {dev:.3e} (modal coefficients {dev_coeff:.3e})", flush=True)
    assert dev <= STATE_L2_TOL, dev
    # the hierarchic coefficients themselves (a stricter norm: the high modes' basis functions
    # are small, so their coefficients weigh more than their share of the solution)
    assert dev_coeff <= STATE_L2_TOL, dev_coeff


def test_mixed_fluid_at_rest_has_no_fp32_pressure_noise():
    """A uniform gas at rest at Ma-0.1 pressure (p ~ 71): the deviation form and the reference
    pressure taken out of the momentum fluxes leave a residual at FP64 rounding, far below the
    FP32 rounding of p / h (~1e-5) that a plain FP32 contraction of the pressure would give."""
    mesh, prob, prim = P.ns_tgv(n=3, order=3)
    prim[..., 1:4] = 0.0
    prim[..., 0] = 1.0
    s = L.Solver(_mixed(prob), mesh)
    s.set_initial_primitive(prim)
    r = s.residual()
    assert np.abs(r).max() <= 1e-9, np.abs(r).max()


def test_mixed_rejects_vector_only_orders():
    mesh, prob, _ = P.euler_vortex(n=4, order=1)
    with pytest.raises(L.SDGError):
        L.Solver(_mixed(prob), mesh)
