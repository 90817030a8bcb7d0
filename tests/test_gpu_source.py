"""Boussinesq source term on the GPU (SourceTerm.cpp:29-58): residual and 5-step parity with the
oracle on every kernel family that carries it -- vector kernels (Euler and Navier-Stokes, small
bases), matrix-core kernels (Euler and Navier-Stokes quadrangles, dense hexahedra) and the
sum-factorised hexahedron kernels (the source formed at their nodes) --, the mixed mode, and the
refusal of the adjoint."""
import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L
from test_oracle_source import _box, _wavy_primitive

pytestmark = pytest.mark.gpu

TOL = 1e-12
GRAD_TOL = 5e-11

CASES = [("euler-quad-p2-valu", L.ELEM_QUADRANGLE, 2, L.EQ_EULER, False),
         ("euler-quad-p3-mfma", L.ELEM_QUADRANGLE, 3, L.EQ_EULER, False),
         ("ns-quad-p2-valu", L.ELEM_QUADRANGLE, 2, L.EQ_NS, False),
         ("ns-quad-p3-mfma", L.ELEM_QUADRANGLE, 3, L.EQ_NS, False),
         ("ns-hex-p2-sumfact", L.ELEM_HEXAHEDRON, 2, L.EQ_NS, False),
         ("ns-hex-p3-sumfact", L.ELEM_HEXAHEDRON, 3, L.EQ_NS, False),
         ("ns-hex-p3-dense", L.ELEM_HEXAHEDRON, 3, L.EQ_NS, True)]


@pytest.mark.parametrize("name,element,order,eq,dense", CASES, ids=[c[0] for c in CASES])
def test_source_residual_and_step_parity(name, element, order, eq, dense):
    print(f"source parity {name}", flush=True)
    mesh, prob = _box(element, order, 3 if element == L.ELEM_HEXAHEDRON else 4, eq, beta=0.7, tref=0.98)
    s = L.Solver(P.with_kernels(prob, L.KERNELS_DENSE) if dense else prob, mesh)
    s.set_initial_primitive(_wavy_primitive(mesh))
    u0 = s.get_state()
    o = O.Oracle(prob, mesh)
    o.set_state(u0)
    x = O.Oracle(prob, mesh, extended=True)
    x.set_state(u0)
    ro, rx = o.residual(), x.residual()
    if eq == L.EQ_NS:
        h = float(mesh.field("min_edge").min())
        P.fp64_parity(s.gradient(), o.gradient(), x.gradient(), u0, mesh.dim, tol=GRAD_TOL, h=h)
    P.fp64_parity(s.residual(), ro, rx, u0, mesh.dim, tol=TOL)
    # the source moves the last momentum residual by a visible amount
    off = L.ProblemDesc.from_buffer_copy(prob)
    off.source_term = L.SOURCE_NONE
    o0 = O.Oracle(off, mesh)
    o0.set_state(u0)
    assert np.abs(ro - o0.residual())[..., mesh.dim].max() > 1e-3 * np.abs(ro[..., mesh.dim]).max()
    dt = o.compute_dt(0.1)
    o.step(dt, 5)
    x.step(dt, 5)
    _, nan, taken = s.step(dt, 5)
    assert nan == 0 and taken == 5
    uo = o.get_state()
    P.fp64_parity(s.get_state(), uo, x.get_state(), uo, mesh.dim, tol=TOL)


def test_source_with_isothermal_walls_on_hexahedra():
    """Thermal-cavity layout: isothermal no-slip walls in z, periodic in x and y."""
    mesh = L.Mesh(L.ELEM_HEXAHEDRON, 3, (3, 3, 3), (0.0,) * 3, (1.0,) * 3,
                  boundary=(None, None, (L.BC_ISOTHERMAL_NONSLIP_WALL, L.BC_ISOTHERMAL_NONSLIP_WALL)))
    prob = L.problem_desc(equation=L.EQ_NS, convective_flux=L.FLUX_ROE, viscous_flux=L.VISC_BR2,
                          transport=L.TRANSPORT_CONSTANT, cp=3.5, cv=2.5, dynamic_viscosity=2e-3,
                          source_term=L.SOURCE_BOUSSINESQ, thermal_expansion_coefficient=0.5, reference_temperature=1.0)
    prim = _wavy_primitive(mesh)
    bprim = np.zeros((mesh.Fb, mesh.nfq, 5))
    bprim[..., 0] = 1.0
    bprim[..., -1] = np.where(P.boundary_points(mesh)[..., 2] < 0.5, 1.05, 0.95)  # hot floor, cold lid
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    s.set_boundary_primitive(bprim)
    u0 = s.get_state()
    o = O.Oracle(prob, mesh)
    o.set_state(u0)
    o.set_boundary_primitive(bprim)
    x = O.Oracle(prob, mesh, extended=True)
    x.set_state(u0)
    x.set_boundary_primitive(bprim)
    P.fp64_parity(s.residual(), o.residual(), x.residual(), u0, 3, tol=TOL)
    dt = o.compute_dt(0.1)
    o.step(dt, 5)
    x.step(dt, 5)
    s.step(dt, 5)
    uo = o.get_state()
    P.fp64_parity(s.get_state(), uo, x.get_state(), uo, 3, tol=TOL)


@pytest.mark.parametrize("element", [L.ELEM_QUADRANGLE, L.ELEM_HEXAHEDRON], ids=["quad-p3", "hex-p3"])
def test_source_in_the_mixed_mode(element):
    """20 mixed-precision steps with the source against the FP64 oracle (quadrangles: the dense
    matrix-core kernels with the source against the orthonormal test functions; hexahedra: the
    sum-factorised kernels with the nodal source)."""
    mesh, prob = _box(element, 3, 3 if element == L.ELEM_HEXAHEDRON else 4, L.EQ_NS, beta=0.7, tref=0.98,
                      precision=L.PREC_MIXED)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(_wavy_primitive(mesh))
    p64 = L.ProblemDesc.from_buffer_copy(prob)
    p64.precision = L.PREC_FP64
    o = O.Oracle(p64, mesh)
    o.set_state(s.get_state())
    dt = o.compute_dt(0.1)
    o.step(dt, 20)
    s.step(dt, 20)
    nv = mesh.nv
    a, b = s.get_state().reshape(-1, nv), o.get_state().reshape(-1, nv)
    dev = max(np.linalg.norm(a[:, g] - b[:, g]) / np.linalg.norm(b[:, g])
              for g in ([0], list(range(1, nv - 1)), [nv - 1]))
    print(f"  mixed L2 deviation {dev:.3e}", flush=True)
    assert dev <= 1e-6, dev


def test_paths_without_a_source_refuse_it():
    mesh, prob = _box(L.ELEM_QUADRANGLE, 3, 4, L.EQ_NS, beta=0.7, tref=0.98)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(_wavy_primitive(mesh))
    with pytest.raises(L.SDGError):
        s.vjp_residual(np.ones_like(s.get_state()))


def test_cpp_driver_thermal_cavity_example():
    """The C++ drop-in driver with SourceTermEnum::Boussinesq and setSourceTerm on a heated cavity
    (sum-factorised hexahedron kernels with walls and the nodal source): the run completes with a
    finite state whose z momentum, zero initially, has grown."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "examples", "bin",
                       "thermal_cavity_3d_cns")
    r = subprocess.run([exe, "4", "40"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    wmax = float(r.stdout.strip().splitlines()[-1].split()[-1])
    assert 0.0 < wmax < 1.0
