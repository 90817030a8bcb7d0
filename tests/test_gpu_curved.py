"""GPU parity on curved (high-order geometry) meshes and on tetrahedra: the vector kernels with
per-point Jacobians and per-element mass inverses (Geometry.cpp:44-100) against the oracle, which
consumes the same per-point tables. Residual, lifted gradient (Navier-Stokes) and 5 SSP-RK3 steps at
the FP64 bar."""
import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L

pytestmark = pytest.mark.gpu


def _smooth(mesh, D):
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, D)
    prim = np.zeros((mesh.E, mesh.nq, D + 2))
    prim[..., 0] = 1.0 + 0.1 * np.sin(2 * np.pi * xq[..., 0]) * np.cos(np.pi * xq[..., 1])
    prim[..., 1] = 0.3 + 0.05 * np.cos(2 * np.pi * xq[..., 1])
    prim[..., 2] = 0.05 * np.sin(np.pi * xq[..., 0])
    prim[..., D + 1] = 1.0 + 0.05 * np.cos(2 * np.pi * (xq[..., 0] - xq[..., 1]))
    return prim


def _check(mesh, prob, bprim=None, steps=5):
    D = mesh.dim
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(_smooth(mesh, D))
    if mesh.Fb:
        s.set_boundary_primitive(bprim)
    o, x = O.Oracle(prob, mesh), O.Oracle(prob, mesh, extended=True)
    for k in (o, x):
        k.set_state(s.get_state())
        if mesh.Fb:
            k.set_boundary_primitive(bprim)
    u0 = o.get_state()
    P.fp64_parity(s.residual(), o.residual(), x.residual(), u0, D, tol=1e-12)
    if prob.equation == L.EQ_NS:
        h = float(np.min(mesh.field("min_edge")))
        P.fp64_parity(s.gradient(), o.gradient(), x.gradient(), u0, D, tol=5e-11, h=h)
    dt = o.compute_dt(0.05)
    assert abs(s.compute_dt(0.05) - dt) <= 1e-13 * dt
    o.step(dt, steps)
    x.step(dt, steps)
    _, nan, taken = s.step(dt, steps)
    assert nan == 0 and taken == steps
    uo = o.get_state()
    P.fp64_parity(s.get_state(), uo, x.get_state(), uo, D, tol=1e-12)
    s.close()


def _euler(flux=L.FLUX_HLLC):
    return L.problem_desc(equation=L.EQ_EULER, convective_flux=flux, cp=3.5, cv=2.5)


def _ns(visc=L.VISC_BR2, flux=L.FLUX_ROE):
    return L.problem_desc(equation=L.EQ_NS, convective_flux=flux, viscous_flux=visc, cp=3.5, cv=2.5,
                          transport=L.TRANSPORT_CONSTANT, dynamic_viscosity=0.01)


@pytest.mark.parametrize("g,order,ns", [(2, 3, False), (3, 3, True), (2, 2, True), (2, 4, False)],
                         ids=["g2-p3-euler", "g3-p3-ns", "g2-p2-ns", "g2-p4-euler"])
def test_curved_annulus_parity(g, order, ns):
    """Quadrangles of geometry order g on an annular sector, far field on the radial sides and a
    slip wall on the arcs."""
    print(f"curved annulus g{g} P{order} ns={ns}", flush=True)
    mesh = P.curved_quad_annulus(n=(6, 3), order=order, geometry_order=g)
    assert mesh.info.affine == 0
    bprim = np.tile(np.array([1.0, 0.3, 0.0, 1.0]), (mesh.Fb, mesh.nfq, 1))
    _check(mesh, _ns() if ns else _euler(), bprim)


@pytest.mark.parametrize("element,g,order,visc", [(L.ELEM_TRIANGLE, 2, 2, L.VISC_BR1), (L.ELEM_TRIANGLE, 3, 3, L.VISC_BR2),
                                                  (L.ELEM_QUADRANGLE, 1, 3, L.VISC_BR2),
                                                  (L.ELEM_HEXAHEDRON, 2, 2, L.VISC_BR2),
                                                  (L.ELEM_HEXAHEDRON, 1, 2, L.VISC_BR1),
                                                  (L.ELEM_TETRAHEDRON, 2, 2, L.VISC_BR2)],
                         ids=["tri6-p2-br1", "tri10-p3-br2", "bilinear-quad-p3-br2", "hex27-p2-br2",
                              "trilinear-hex-p2-br1", "tet10-p2-br2"])
def test_curved_periodic_parity(element, g, order, visc):
    """Warped periodic meshes (curved, or non-parallelogram quadrangles / non-parallelepiped
    hexahedra), Navier-Stokes."""
    print(f"curved periodic element {element} g{g} P{order} visc {visc}", flush=True)
    mesh, _, _ = P.lattice_mesh(element, 2 if element in (L.ELEM_HEXAHEDRON, L.ELEM_TETRAHEDRON) else 3, g,
                                order=order, warp=P.bump(0.06), periodic=True)
    assert mesh.info.affine == 0
    _check(mesh, _ns(visc))


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_tetrahedra_parity(order):
    """Jittered periodic Kuhn tetrahedra, Euler (HLLC), P1-P4."""
    print(f"tetrahedra P{order} euler", flush=True)
    mesh = P.tet_box(n=2, order=order, jitter=0.2)
    _check(mesh, _euler())


@pytest.mark.parametrize("order,visc", [(2, L.VISC_BR2), (3, L.VISC_BR1)])
def test_tetrahedra_navier_stokes_with_boundaries(order, visc):
    """Jittered tetrahedra of a box with far-field sides, Navier-Stokes."""
    print(f"tetrahedra P{order} ns visc {visc} bounded", flush=True)
    mesh = P.tet_box(n=2, order=order, periodic=False, jitter=0.2)
    bprim = np.tile(np.array([1.0, 0.3, 0.0, 0.0, 1.0]), (mesh.Fb, mesh.nfq, 1))
    _check(mesh, _ns(visc), bprim)



def test_tetrahedra_artificial_viscosity_parity():
    """Euler with artificial viscosity on jittered tetrahedra (P2): node viscosities and the residual."""
    print("tetrahedra P2 artificial viscosity", flush=True)
    mesh = P.tet_box(n=2, order=2, jitter=0.2)
    prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_HLLC, cp=3.5, cv=2.5,
                          shock_capturing=L.SHOCK_ARTIFICIAL_VISCOSITY, empirical_tolerance=4.0)
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, 3)
    # a steep but smooth front: a jump inside a P2 tetrahedron projects to negative face densities
    rho = 0.75 + 0.25 * np.tanh((0.47 - xq[..., 0]) / 0.03)
    prim = np.stack([rho, 0 * rho, 0 * rho, 0 * rho, 0 * rho + 1.0], axis=-1)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    o, x = O.Oracle(prob, mesh), O.Oracle(prob, mesh, extended=True)
    for k in (o, x):
        k.set_state(s.get_state())
    u0 = o.get_state()
    P.fp64_parity(s.residual(), o.residual(), x.residual(), u0, 3, tol=1e-12)
    assert o.artificial_viscosity().max() > 0.0
    assert np.abs(s.artificial_viscosity() - o.artificial_viscosity()).max() <= 1e-13 * o.artificial_viscosity().max()
    s.close()


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_pyramids_parity(order):
    """Six pyramids per box around jittered apexes, periodic, Euler (Roe), P1-P4."""
    print(f"pyramids P{order} euler", flush=True)
    mesh = P.pyramid_box(n=2, order=order, jitter=0.15)
    _check(mesh, _euler(L.FLUX_ROE))


def test_pyramids_navier_stokes_with_boundaries():
    """Pyramids of a box with far-field sides, Navier-Stokes BR2, P2."""
    print("pyramids P2 ns bounded", flush=True)
    mesh = P.pyramid_box(n=2, order=2, periodic=False, jitter=0.15)
    bprim = np.tile(np.array([1.0, 0.3, 0.0, 0.0, 1.0]), (mesh.Fb, mesh.nfq, 1))
    _check(mesh, _ns(L.VISC_BR2), bprim)
