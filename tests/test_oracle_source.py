"""Boussinesq source term of the oracle (SourceTerm.cpp:29-58, SpatialDiscrete.cpp:254-262,
1026-1029): rho beta (T - T_ref) g, g = 1, on the last momentum component, added to the residual
as S V. Pinned by an independent numpy evaluation of S V from the state at the quadrature points,
by its absence at T = T_ref, and by the closed-form forward-Euler step of a uniform fluid at rest
(every flux term vanishes, so d(rho w)/dt = rho beta (T - T_ref) at every point)."""
import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L


def _box(element, order, n, eq, beta, tref, precision=L.PREC_FP64):
    dim = 2 if element == L.ELEM_QUADRANGLE else 3
    mesh = L.Mesh(element, order, (n,) * dim, (0.0,) * dim, (1.0,) * dim)
    kw = dict(equation=eq, convective_flux=L.FLUX_ROE, cp=3.5, cv=2.5, source_term=L.SOURCE_BOUSSINESQ,
              thermal_expansion_coefficient=beta, reference_temperature=tref, precision=precision)
    if eq == L.EQ_NS:
        kw.update(viscous_flux=L.VISC_BR2, transport=L.TRANSPORT_CONSTANT, dynamic_viscosity=1e-3)
    return mesh, L.problem_desc(**kw)


def _wavy_primitive(mesh):
    dim = mesh.dim
    x = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, dim)
    prim = np.zeros((mesh.E, mesh.nq, dim + 2))
    prim[..., 0] = 1.0 + 0.1 * np.sin(2 * np.pi * x[..., 0])
    for d in range(dim):
        prim[..., 1 + d] = 0.2 * np.cos(2 * np.pi * x[..., (d + 1) % dim])
    prim[..., -1] = 1.0 + 0.05 * np.cos(2 * np.pi * x[..., dim - 1])
    return prim


@pytest.mark.parametrize("element,order,eq", [(L.ELEM_QUADRANGLE, 2, L.EQ_EULER), (L.ELEM_QUADRANGLE, 3, L.EQ_NS),
                                              (L.ELEM_HEXAHEDRON, 2, L.EQ_NS)])
def test_source_equals_numpy_projection(element, order, eq):
    mesh, prob = _box(element, order, 3, eq, beta=0.7, tref=0.98)
    off = L.ProblemDesc.from_buffer_copy(prob)
    off.source_term = L.SOURCE_NONE
    o = O.Oracle(prob, mesh)
    o.set_initial_primitive(_wavy_primitive(mesh))
    u = o.get_state()
    o0 = O.Oracle(off, mesh)
    o0.set_state(u)
    diff = o.residual() - o0.residual()
    # numpy: state at the points, T = (E/rho - |m|^2 / (2 rho^2)) / cv, S = rho beta (T - T_ref) |J| w
    V = mesh.field("basis_value").reshape(mesh.nq, mesh.nb)
    jw = mesh.field("jacobian_det_w").reshape(mesh.E, mesh.nq)
    uq = np.einsum("qb,ebv->eqv", V, u)
    rho, m, E = uq[..., 0], uq[..., 1:1 + mesh.dim], uq[..., -1]
    T = (E / rho - 0.5 * (m * m).sum(-1) / rho ** 2) / 2.5
    S = rho * 0.7 * (T - 0.98) * jw
    ref = np.einsum("eq,qb->eb", S, V)
    comp = mesh.dim  # the last momentum component
    assert np.abs(diff[..., comp] - ref).max() <= 1e-13 * np.abs(ref).max()
    others = [v for v in range(mesh.nv) if v != comp]
    assert np.abs(diff[..., others]).max() <= 1e-13 * np.abs(ref).max()


def test_no_source_at_the_reference_temperature():
    mesh, prob = _box(L.ELEM_QUADRANGLE, 2, 3, L.EQ_EULER, beta=0.5, tref=1.0)
    prim = np.zeros((mesh.E, mesh.nq, 4))
    prim[..., 0] = 1.3
    prim[..., 1] = 0.2
    prim[..., 3] = 1.0
    o = O.Oracle(prob, mesh)
    o.set_initial_primitive(prim)
    assert np.abs(o.residual()).max() <= 1e-13


@pytest.mark.parametrize("element", [L.ELEM_QUADRANGLE, L.ELEM_HEXAHEDRON])
def test_fluid_at_rest_accelerates_at_rho_beta_delta_t(element):
    """Forward Euler from a uniform fluid at rest: w-momentum becomes dt rho beta (T - T_ref)."""
    rho, T, beta, tref, dt = 1.2, 1.1, 0.4, 1.0, 1e-3
    mesh, prob = _box(element, 2, 2, L.EQ_NS, beta=beta, tref=tref)
    prob.time_integration = L.RK_FORWARD_EULER
    prim = np.zeros((mesh.E, mesh.nq, mesh.nv))
    prim[..., 0] = rho
    prim[..., -1] = T
    o = O.Oracle(prob, mesh)
    o.set_initial_primitive(prim)
    o.step(dt, 1)
    V = mesh.field("basis_value").reshape(mesh.nq, mesh.nb)
    uq = np.einsum("qb,ebv->eqv", V, o.get_state())
    expect = dt * rho * beta * (T - tref)
    assert np.abs(uq[..., mesh.dim] - expect).max() <= 1e-10 * expect  # rounding of O(1) terms: 1e-16
    assert np.abs(uq[..., 1:mesh.dim]).max() <= 1e-14
