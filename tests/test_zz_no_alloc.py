"""sdg_compute_dt_noalloc, the allocation-free variant of sdg_compute_dt (its per-element minima
live in the library's fixed device array), called through the C-ABI: the same time step bit for
bit as sdg_compute_dt on two calls in a row, the second on a smaller mesh (the array keeps the
first call's larger contents), on both element families, and the refusal of meshes beyond the
array's 2^21 elements."""
import ctypes

import numpy as np
import pytest

import problems as P
from subrosadg_amd import _lib as L

pytestmark = pytest.mark.gpu


def _both(s, cfl):
    ref, got = ctypes.c_double(), ctypes.c_double()
    assert s.lib.sdg_compute_dt(s.handle, cfl, ctypes.byref(ref)) == 0
    rc = s.lib.sdg_compute_dt_noalloc(s.handle, cfl, ctypes.byref(got))
    assert rc == 0, s.lib.sdg_last_error(s.handle)
    return ref.value, got.value


def test_noalloc_dt_equals_compute_dt_large_then_small():
    print("noalloc dt: tgv hex P3 16^3 then vortex quad P2 6^2", flush=True)
    cases = [P.ns_tgv(n=16, order=3), P.euler_vortex(n=6, order=2)]
    for mesh, prob, prim in cases:
        s = L.Solver(prob, mesh)
        s.set_initial_primitive(prim)
        for cfl in (0.1, 0.37):
            ref, got = _both(s, cfl)
            print(f"  E={mesh.E} cfl={cfl}: compute_dt {ref!r} noalloc {got!r}", flush=True)
            assert np.isfinite(ref) and ref > 0
            assert got == ref
        s.close()


def test_noalloc_dt_after_steps_and_in_sequence():
    """Interleaved with steps of the same context and with another context on the device."""
    print("noalloc dt: shear quad P3 8^2 steps, then split grid P2", flush=True)
    mesh, prob, prim = P.ns_shear(n=8, order=3, perturb=True)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    for _ in range(3):
        ref, got = _both(s, 0.1)
        assert got == ref
        s.step(ref, 2)
    m2, p2, pr2 = P.euler_vortex(n=5, order=1)
    s2 = L.Solver(p2, m2)
    s2.set_initial_primitive(pr2)
    ref2, got2 = _both(s2, 0.2)
    assert got2 == ref2
    ref, got = _both(s, 0.1)
    assert got == ref
    s2.close()
    s.close()


def test_noalloc_dt_refuses_meshes_beyond_the_array():
    print("noalloc dt: refusal at 1449^2 quad P1", flush=True)
    mesh = L.Mesh(L.ELEM_QUADRANGLE, 1, (1449, 1449), (0.0, 0.0), (1.0, 1.0))
    assert mesh.E > (1 << 21)
    prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_ROE, cp=3.5, cv=2.5)
    s = L.Solver(prob, mesh)
    dt = ctypes.c_double(-1.0)
    rc = s.lib.sdg_compute_dt_noalloc(s.handle, 0.1, ctypes.byref(dt))
    assert rc == L.E_INVALID
    assert b"2^21" in s.lib.sdg_last_error(s.handle)
    s.close()
