"""Triangles on the GPU (single-problem engine against the oracle) and the batched solver: batch
against the single-problem engine on single-type grids and against the mixed-mesh oracle on mixed
triangle/quadrangle grids, invariants (free stream, conservation), mixed precision against FP64,
the autotuner, and the full 48-problem uneven workload."""
import math

import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _vortex(xq):
    return P.isentropic_vortex_primitive(xq)


def _euler(flux=L.FLUX_LAX_FRIEDRICHS, precision=L.PREC_FP64):
    return L.problem_desc(equation=L.EQ_EULER, convective_flux=flux, cp=3.5, cv=2.5, precision=precision)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("flux", [L.FLUX_ROE, L.FLUX_LAX_FRIEDRICHS])
def test_triangle_euler_parity(order, flux):
    print(f"triangles P{order} flux {flux}", flush=True)
    m = L.Mesh.split(order, (5, 5), (-5.0, -5.0), (5.0, 5.0), 1.0, 3)
    prob = _euler(flux)
    xq = m.field("quad_coord").reshape(m.E, m.nq, 2)
    s = L.Solver(prob, m)
    s.set_initial_primitive(_vortex(xq))
    o = O.Oracle(prob, m)
    o.set_state(s.get_state())
    u0 = o.get_state()
    assert P.parity_error(s.residual(), o.residual(), u0, 2) <= TOL
    dt = o.compute_dt(0.1)
    o.step(dt, 10)
    s.step(dt, 10)
    uo = o.get_state()
    assert P.parity_error(s.get_state(), uo, uo, 2) <= TOL


def test_triangle_ns_parity():
    m = L.Mesh.split(3, (4, 4), (0.0, 0.0), (1.0, 1.0), 1.0, 5)
    cp, cv = P.tgv_thermo(0.1)
    prob = L.problem_desc(equation=L.EQ_NS, convective_flux=L.FLUX_ROE, viscous_flux=L.VISC_BR2,
                          transport=L.TRANSPORT_CONSTANT, cp=cp, cv=cv, dynamic_viscosity=1e-3)
    xq = m.field("quad_coord").reshape(m.E, m.nq, 2)
    s = L.Solver(prob, m)
    s.set_initial_primitive(P.shear_layer_primitive(xq, perturb=True))
    o = O.Oracle(prob, m)
    o.set_state(s.get_state())
    u0 = o.get_state()
    ro = o.residual()  # also fills the oracle's lifted gradient
    h = float(m.field("min_edge").min())
    assert P.parity_error(s.gradient(), o.gradient(), u0, 2, h=h) <= 5e-11
    assert P.parity_error(s.residual(), ro, u0, 2) <= TOL
    dt = o.compute_dt(0.1)
    o.step(dt, 5)
    s.step(dt, 5)
    uo = o.get_state()
    assert P.parity_error(s.get_state(), uo, uo, 2) <= TOL


def _batch_problems(specs, precision=L.PREC_FP64, flux=L.FLUX_LAX_FRIEDRICHS):
    meshes = [L.Mesh.split(order, (n, n), (-5.0, -5.0), (5.0, 5.0), p, seed) for (order, n, p, seed) in specs]
    b = L.Batch(_euler(flux, precision), meshes)
    for i, m in enumerate(meshes):
        xq = m.field("quad_coord").reshape(-1, 2)
        b.set_initial_primitive(i, _vortex(xq))
    return meshes, b


def test_batch_matches_single_problem_engine():
    """On single-type grids every problem of the batch follows the single-problem engine."""
    specs = [(3, 6, 0.0, 1), (2, 5, 1.0, 2), (1, 7, 0.0, 3), (4, 4, 1.0, 4), (4, 5, 0.0, 5), (1, 6, 1.0, 6)]
    meshes, b = _batch_problems(specs)
    dt = b.compute_dt(0.1)
    singles = []
    for i, m in enumerate(meshes):
        s = L.Solver(_euler(), m)
        s.set_state(b.get_state(i).reshape(m.E, m.nb, 4))
        assert abs(s.compute_dt(0.1) - dt[i]) <= 1e-13 * dt[i]
        singles.append(s)
    b.step(dt, 10)
    for i, (m, s) in enumerate(zip(meshes, singles)):
        s.step(dt[i], 10)
        us = s.get_state()
        err = P.parity_error(b.get_state(i).reshape(us.shape), us, us, 2)
        print(f"  problem {i}: {err:.2e}", flush=True)
        assert err <= TOL, (i, err)


def _integrals(meshes, b):
    """Domain integrals of the conserved variables per problem (quadrature of the modal state)."""
    out = []
    for i, m in enumerate(meshes):
        c = b.get_state(i)
        et = m.field("element_type")
        jw = m.field("jacobian_det_w")
        refs = {t: L.Mesh.split(m.order, (2, 2), (0, 0), (1, 1), 1.0 if t == L.ELEM_TRIANGLE else 0.0, 0)
                for t in (L.ELEM_TRIANGLE, L.ELEM_QUADRANGLE)}
        V = {t: r.field("basis_value").reshape(r.nq, r.nb) for t, r in refs.items()}
        tot = np.zeros(4)
        co, qo = 0, 0
        for t in et:
            Vt = V[int(t)]
            nq, nb = Vt.shape
            C = c[co:co + nb * 4].reshape(nb, 4)
            tot += (jw[qo:qo + nq, None] * (Vt @ C)).sum(0)
            co += nb * 4
            qo += nq
        out.append(tot)
    return np.array(out)


def test_batch_mixed_grids_conserve_and_keep_free_stream():
    specs = [(1, 8, 0.5, 11), (2, 9, 0.5, 12), (3, 7, 0.5, 13), (4, 6, 0.5, 14)]
    meshes, b = _batch_problems(specs, flux=L.FLUX_ROE)
    assert all(m.mixed for m in meshes)
    before = _integrals(meshes, b)
    dt = b.compute_dt(0.1)
    b.step(dt, 20)
    after = _integrals(meshes, b)
    assert np.all(np.isfinite(after))
    rel = np.abs(after - before) / np.abs(before).max(axis=1, keepdims=True)
    assert rel.max() <= 1e-12, rel
    # free stream
    for i, m in enumerate(meshes):
        nq = m.field("jacobian_det_w").size
        b.set_initial_primitive(i, np.tile([1.0, 0.3, 0.2, 1.0], (nq, 1)))
    u0 = [b.get_state(i) for i in range(len(meshes))]
    b.step(dt, 5)
    for i in range(len(meshes)):
        assert np.abs(b.get_state(i) - u0[i]).max() <= 1e-11 * np.abs(u0[i]).max()


def test_batch_mixed_precision_deviation():
    specs = [(1, 8, 0.5, 21), (2, 8, 0.5, 22), (3, 6, 0.5, 23), (4, 6, 0.5, 24)]
    _, b64 = _batch_problems(specs)
    _, b32 = _batch_problems(specs, precision=L.PREC_MIXED)
    dt = b64.compute_dt(0.1)
    b64.step(dt, 20)
    b32.step(dt, 20)
    worst = 0.0
    for i in range(len(specs)):
        a, r = b32.get_state(i).reshape(-1, 4), b64.get_state(i).reshape(-1, 4)
        for g in ([0], [1, 2], [3]):
            worst = max(worst, np.linalg.norm(a[:, g] - r[:, g]) / np.linalg.norm(r[:, g]))
    print(f"  mixed-precision batch L2 deviation {worst:.3e}", flush=True)
    assert worst <= 1e-6, worst


def test_batch_tuner_keeps_state():
    specs = [(1, 12, 0.5, 31), (4, 8, 0.5, 32), (2, 10, 0.0, 33)]
    meshes, b = _batch_problems(specs)
    u0 = [b.get_state(i) for i in range(len(meshes))]
    ms = b.tune()
    chunk, epb, grid = b.config()
    print(f"  tuned stage {ms:.4f} ms, tiles per item {chunk.tolist()}, elements per tile {epb.tolist()}, grid {grid}",
          flush=True)
    assert ms > 0 and grid > 0
    for i in range(len(meshes)):
        assert np.array_equal(b.get_state(i), u0[i])
    dt = b.compute_dt(0.1)
    b.step(dt, 3)
    assert all(np.all(np.isfinite(b.get_state(i))) for i in range(len(meshes)))


@pytest.mark.parametrize("flux", [L.FLUX_LAX_FRIEDRICHS, L.FLUX_ROE, L.FLUX_HLLC])
def test_batch_mixed_grids_match_the_oracle(flux):
    """Mixed triangle/quadrangle grids of orders 1-4 in one batch: each problem's CFL time step and
    its state after 10 SSP-RK3 steps equal the mixed-mesh oracle's (the reference's triangle_ and
    quadrangle_ element solvers in one Solver) at the FP64 bar."""
    specs = [(1, 6, 0.5, 41), (2, 5, 0.5, 42), (3, 5, 0.4, 43), (4, 4, 0.6, 44)]
    print(f"batch vs mixed-mesh oracle, flux {flux}", flush=True)
    meshes, b = _batch_problems(specs, flux=flux)
    assert all(m.mixed for m in meshes)
    dt = b.compute_dt(0.1)
    oracles = []
    for i, m in enumerate(meshes):
        o = O.MixedOracle(_euler(flux), m)
        o.set_state(b.get_state(i))
        odt = o.compute_dt(0.1)
        assert abs(odt - dt[i]) <= 1e-13 * odt, (i, odt, dt[i])
        oracles.append(o)
    b.step(dt, 10)
    for i, (m, o) in enumerate(zip(meshes, oracles)):
        o.step(dt[i], 10)
        uo = o.get_state().reshape(-1, 4)
        err = P.parity_error(b.get_state(i).reshape(-1, 4), uo, uo, 2)
        print(f"  problem {i} (P{m.order}, {m.E} elements): {err:.2e}", flush=True)
        assert err <= TOL, (i, err)


def test_c3_workload_48_uneven_problems():
    """The full batched workload: 48 isentropic-vortex problems (n ~ U{8..96}, P ~ U{1..4},
    Bernoulli(0.5) triangle split, seed 20261202), Lax-Friedrichs, SSP-RK3 at each problem's CFL-0.1
    step, 100 steps, tuned launch. FP64: mass, momentum and energy conserved to 1e-12 of their
    magnitude in every problem, and a uniform moving stream stays uniform to rounding. Mixed
    precision: relative L2(domain) deviation from the FP64 run at most 1e-6 in every problem and
    variable group, and a uniform gas at rest stays at rest (its deviations from the reference
    state and its fluxes vanish; a moving stream keeps its FP32 contraction rounding, ~1e-7 of
    the mass flux per step)."""
    specs = P.c3_specs()
    print(f"C3: {len(specs)} problems", flush=True)
    meshes = [L.Mesh.split(order, (n, n), (-5.0, -5.0), (5.0, 5.0), 0.5, sd) for n, order, sd in specs]
    prims = [_vortex(m.field("quad_coord").reshape(-1, 2)) for m in meshes]
    states, dt = {}, None
    for prec in (L.PREC_FP64, L.PREC_MIXED):
        b = L.Batch(_euler(L.FLUX_LAX_FRIEDRICHS, prec), meshes)
        for i, pr in enumerate(prims):
            b.set_initial_primitive(i, pr)
        u0 = [b.get_state(i) for i in range(len(meshes))]
        if dt is None:
            dt = b.compute_dt(0.1)
        b.tune()
        b.step(dt, 100)
        states[prec] = [b.get_state(i) for i in range(len(meshes))]
        if prec == L.PREC_FP64:
            worst = 0.0
            for m, a, c in zip(meshes, u0, states[prec]):
                i0, i1 = P.split_integrals(m, a), P.split_integrals(m, c)
                worst = max(worst, float(np.abs(i1 - i0).max() / np.abs(i0).max()))
            print(f"  FP64 conservation {worst:.2e}", flush=True)
            assert worst <= 1e-12, worst
        uniform = [1.0, 0.3, 0.2, 1.0] if prec == L.PREC_FP64 else [1.0, 0.0, 0.0, 1.0]
        for i, m in enumerate(meshes):
            nq = m.field("jacobian_det_w").size
            b.set_initial_primitive(i, np.tile(uniform, (nq, 1)))
        f0 = [b.get_state(i) for i in range(len(meshes))]
        b.step(dt, 5)
        fs = max(float(np.abs(b.get_state(i) - f0[i]).max() / np.abs(f0[i]).max()) for i in range(len(meshes)))
        print(f"  precision {prec}: uniform state {uniform} after 5 steps {fs:.2e}", flush=True)
        assert fs <= 1e-11, fs
        b.close()
    devs = [P.split_function_l2(m, a, c) for m, a, c in zip(meshes, states[L.PREC_MIXED], states[L.PREC_FP64])]
    k = int(np.argmax(devs))
    print(f"  mixed vs FP64 L2 deviation: max {devs[k]:.3e} (problem {k}: n={specs[k][0]}, P{specs[k][1]}), "
          f"median {float(np.median(devs)):.3e}", flush=True)
    assert max(devs) <= 1e-6, devs
