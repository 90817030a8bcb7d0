"""GPU parity of the physics variants off the benchmark path: the weakly-compressible equation of
state with the Exact flux (ConvectiveFlux.cpp:352-414, PhysicalModel.cpp:58-75) and Sutherland
transport (PhysicalModel.cpp:102-122), on every kernel family (vector, matrix-core, sum-factorised
hexahedra, triangles, lines): residual, lifted gradient and 5 SSP-RK3 steps at the FP64 bar
against the oracle (the oracle's functions are pinned in test_oracle_pins.py)."""
import math

import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L

pytestmark = pytest.mark.gpu

TOL, GRAD_TOL = 1e-12, 5e-11
C0, RHO0 = 10.0, 1.0


def _mesh(kind, order):
    if kind == "quad":
        return L.Mesh(L.ELEM_QUADRANGLE, order, (5, 4), (0.0, 0.0), (1.0, 1.0))
    if kind == "tri":
        return L.Mesh.split(order, (4, 4), (0.0, 0.0), (1.0, 1.0), 1.0, 7)
    if kind == "hex":
        return L.Mesh(L.ELEM_HEXAHEDRON, order, (3, 3, 3), (0.0,) * 3, (1.0,) * 3)
    return L.Mesh(L.ELEM_LINE, order, (10,), (0.0,), (1.0,))


def _prim(m, weak):
    xq = m.field("quad_coord").reshape(m.E, m.nq, m.dim)
    x = xq[..., 0]
    y = xq[..., 1] if m.dim > 1 else 0.3 * x
    z = xq[..., 2] if m.dim > 2 else 0.0
    s = np.sin(2 * math.pi * x) * np.cos(2 * math.pi * y) * np.cos(2 * math.pi * z)
    prim = np.zeros((m.E, m.nq, m.nv))
    prim[..., 0] = RHO0 * (1.0 + (0.01 if weak else 0.1) * s)
    vel = [0.3 + 0.1 * s, -0.2 + 0.1 * np.cos(2 * math.pi * x), 0.1 * np.sin(2 * math.pi * y)]
    for d in range(m.dim):
        prim[..., 1 + d] = vel[d]
    prim[..., m.dim + 1] = 1.0 + 0.2 * np.cos(2 * math.pi * (x + y))
    return prim


CASES = [
    # name, element, order, equation, flux, viscous scheme, transport, weakly compressible
    ("exact-wc-euler-quad-p2", "quad", 2, L.EQ_EULER, L.FLUX_EXACT, L.VISC_NONE, L.TRANSPORT_NONE, True),
    ("exact-wc-euler-quad-p3", "quad", 3, L.EQ_EULER, L.FLUX_EXACT, L.VISC_NONE, L.TRANSPORT_NONE, True),
    ("exact-wc-euler-tri-p2", "tri", 2, L.EQ_EULER, L.FLUX_EXACT, L.VISC_NONE, L.TRANSPORT_NONE, True),
    ("exact-wc-euler-line-p3", "line", 3, L.EQ_EULER, L.FLUX_EXACT, L.VISC_NONE, L.TRANSPORT_NONE, True),
    ("exact-wc-ns-quad-p3-br1", "quad", 3, L.EQ_NS, L.FLUX_EXACT, L.VISC_BR1, L.TRANSPORT_CONSTANT, True),
    ("exact-wc-ns-hex-p2-br2", "hex", 2, L.EQ_NS, L.FLUX_EXACT, L.VISC_BR2, L.TRANSPORT_CONSTANT, True),
    ("exact-wc-ns-hex-p3-sutherland", "hex", 3, L.EQ_NS, L.FLUX_EXACT, L.VISC_BR2, L.TRANSPORT_SUTHERLAND, True),
    ("sutherland-quad-p2-br1", "quad", 2, L.EQ_NS, L.FLUX_ROE, L.VISC_BR1, L.TRANSPORT_SUTHERLAND, False),
    ("sutherland-quad-p3-br2", "quad", 3, L.EQ_NS, L.FLUX_ROE, L.VISC_BR2, L.TRANSPORT_SUTHERLAND, False),
    ("sutherland-tri-p3-br2", "tri", 3, L.EQ_NS, L.FLUX_HLLC, L.VISC_BR2, L.TRANSPORT_SUTHERLAND, False),
    ("sutherland-hex-p3-br2", "hex", 3, L.EQ_NS, L.FLUX_HLLC, L.VISC_BR2, L.TRANSPORT_SUTHERLAND, False),
    ("sutherland-hex-p4-br1", "hex", 4, L.EQ_NS, L.FLUX_LAX_FRIEDRICHS, L.VISC_BR1, L.TRANSPORT_SUTHERLAND, False),
    ("sutherland-line-p4-br2", "line", 4, L.EQ_NS, L.FLUX_ROE, L.VISC_BR2, L.TRANSPORT_SUTHERLAND, False),
]


@pytest.mark.parametrize("name,kind,order,eq,flux,visc,transport,weak", CASES, ids=[c[0] for c in CASES])
def test_physics_variant_parity(name, kind, order, eq, flux, visc, transport, weak):
    print(f"physics variant {name}", flush=True)
    m = _mesh(kind, order)
    kw = dict(equation=eq, convective_flux=flux, cp=3.5, cv=2.5)
    if weak:
        kw.update(equation_of_state=L.EOS_WEAK_COMPRESSIBLE, reference_sound_speed=C0, reference_density=RHO0)
    if eq == L.EQ_NS:
        kw.update(viscous_flux=visc, transport=transport, dynamic_viscosity=5e-3)
    prob = L.problem_desc(**kw)
    s = L.Solver(prob, m)
    s.set_initial_primitive(_prim(m, weak))
    u0 = s.get_state()
    o, x = O.Oracle(prob, m), O.Oracle(prob, m, extended=True)
    o.set_state(u0)
    x.set_state(u0)
    P.fp64_parity(s.residual(), o.residual(), x.residual(), u0, m.dim, tol=TOL)
    if eq == L.EQ_NS:
        h = float(m.field("min_edge").min())
        P.fp64_parity(s.gradient(), o.gradient(), x.gradient(), u0, m.dim, tol=GRAD_TOL, h=h)
    dt = o.compute_dt(0.1)
    assert abs(s.compute_dt(0.1) - dt) <= 1e-13 * dt
    o.step(dt, 5)
    x.step(dt, 5)
    _, nan, taken = s.step(dt, 5)
    assert nan == 0 and taken == 5
    uo = o.get_state()
    P.fp64_parity(s.get_state(), uo, x.get_state(), uo, m.dim, tol=TOL)
