"""Discrete adjoint of the GPU engine: residual VJP (dot-product tests against central finite
differences of the GPU and of the CPU oracle), the checkpointed adjoint of SSP-RK3 stepping
(gradient of the final kinetic energy with respect to the initial v-perturbation field against
central finite differences), schedule independence of the binomial checkpointing, and the
torch.autograd wrapper."""
import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L

pytestmark = pytest.mark.gpu

CASES = [
    ("shear-quad-p3-roe-br2", lambda: P.ns_shear(n=4, order=3, perturb=True)),
    ("shear-quad-p3-hllc-br1", lambda: P.ns_shear(n=4, order=3, flux=L.FLUX_HLLC, visc=L.VISC_BR1, perturb=True)),
    ("shear-quad-p4-lf-br2", lambda: P.ns_shear(n=4, order=4, flux=L.FLUX_LAX_FRIEDRICHS, perturb=True)),
    ("tgv-hex-p2-hllc-br2", lambda: P.ns_tgv(n=3, order=2)),
    ("tgv-hex-p3-roe-br2", lambda: P.ns_tgv(n=2, order=3, flux=L.FLUX_ROE)),
    ("tgv-hex-p3-hllc-br1", lambda: P.ns_tgv(n=2, order=3, visc=L.VISC_BR1)),
    ("tgv-hex-p4-lf-br1", lambda: P.ns_tgv(n=2, order=4, flux=L.FLUX_LAX_FRIEDRICHS, visc=L.VISC_BR1)),
    ("tgv-hex-p4-hllc-br2", lambda: P.ns_tgv(n=2, order=4)),
    ("tgv-hex-p3-roe-br2-dense", lambda: _dense(P.ns_tgv(n=2, order=3, flux=L.FLUX_ROE))),
]


def _dense(case):
    """The same case on the dense matrix-core kernels (hexahedra default to the sum-factorised
    stage kernels and their adjoint)."""
    mesh, prob, prim = case
    prob.kernels = L.KERNELS_DENSE
    return mesh, prob, prim


def _direction(rng, u0, dim):
    """A random perturbation scaled to each variable group's magnitude."""
    v = rng.standard_normal(u0.shape)
    for g in ([0], list(range(1, dim + 1)), [dim + 1]):
        v[..., g] *= np.abs(u0[..., g]).max()
    return v


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_residual_vjp_dot_product(name, make):
    """<lam, R'(U) v> by central differences of the residual (GPU engine and CPU oracle) equals
    <R'(U)^T lam, v> from the adjoint kernels."""
    print(f"vjp dot product {name}", flush=True)
    mesh, prob, prim = make()
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    rng = np.random.default_rng(7)
    # Break the symmetries of the analytic states (u.n = 0 on whole faces of the Taylor-Green
    # vortex, equal left/right traces) so that no quadrature point sits on a kink of |u.n| or of
    # a min/max wave-speed bound, where central differences and the derivative disagree.
    u0 = s.get_state()
    u0 = u0 + 1e-3 * _direction(rng, u0, mesh.dim)
    s.set_state(u0)
    lam = rng.standard_normal(u0.shape)
    v = _direction(rng, u0, mesh.dim)
    w = s.vjp_residual(lam)
    adj = float((w * v).sum())
    eps = 1e-7
    s.set_state(u0 + eps * v)
    rp = s.residual()
    s.set_state(u0 - eps * v)
    rm = s.residual()
    fd_gpu = float((lam * (rp - rm)).sum() / (2 * eps))
    o = O.Oracle(prob, mesh)
    o.set_state(u0 + eps * v)
    rp = o.residual()
    o.set_state(u0 - eps * v)
    rm = o.residual()
    fd_cpu = float((lam * (rp - rm)).sum() / (2 * eps))
    print(f"  adjoint {adj:.12e}  fd(gpu) {fd_gpu:.12e}  fd(oracle) {fd_cpu:.12e}", flush=True)
    assert abs(adj - fd_gpu) <= 1e-7 * abs(adj), (adj, fd_gpu)
    assert abs(adj - fd_cpu) <= 1e-7 * abs(adj), (adj, fd_cpu)


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_residual_vjp_tangent_dot_product(name, make):
    """<J v, w> with J v from the oracle's tangent-linear build (forward-mode numbers through the
    oracle's passes, no difference quotient) equals <v, J^T w> from the GPU adjoint kernels to
    rounding: 1e-12 of the sum of the absolute products."""
    print(f"vjp tangent dot product {name}", flush=True)
    mesh, prob, prim = make()
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    rng = np.random.default_rng(11)
    u0 = s.get_state()
    u0 = u0 + 1e-3 * _direction(rng, u0, mesh.dim)
    s.set_state(u0)
    w = rng.standard_normal(u0.shape)
    v = _direction(rng, u0, mesh.dim)
    jtw = s.vjp_residual(w)
    _, jv = O.Oracle(prob, mesh, tangent=True).residual_tangent(u0, v)
    lhs, rhs = float((jv * w).sum()), float((v * jtw).sum())
    scale = float(np.abs(jv * w).sum() + np.abs(v * jtw).sum()) / 2
    print(f"  <Jv,w> {lhs:.15e}  <v,J^T w> {rhs:.15e}  rel {abs(lhs - rhs) / abs(lhs):.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, scale)


@pytest.mark.parametrize("snapshots", [4, 12])
def test_adjoint_steps_tangent_dot_product(snapshots):
    """The same through 12 SSP-RK3 steps of fixed size (shear layer 8^2 quads P3, Roe + BR2):
    <dU_n v, w> from the oracle's tangent steps equals <v, lambda_0> from the checkpointed GPU
    adjoint sweep with lambda_n = w, at 1e-12 of the sum of the absolute products."""
    steps = 12
    mesh, prob, prim = P.ns_shear(n=8, order=3, perturb=True)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    dt = s.compute_dt(0.1)
    rng = np.random.default_rng(12)
    u0 = s.get_state()
    u0 = u0 + 1e-3 * _direction(rng, u0, mesh.dim)
    w = rng.standard_normal(u0.shape)
    v = _direction(rng, u0, mesh.dim)
    s.set_state(u0)
    s.adjoint_forward(dt, steps, snapshots)
    lam0 = s.adjoint_backward(w)
    _, dun = O.Oracle(prob, mesh, tangent=True).step_tangent(u0, v, dt, steps)
    lhs, rhs = float((dun * w).sum()), float((v * lam0).sum())
    scale = float(np.abs(dun * w).sum() + np.abs(v * lam0).sum()) / 2
    print(f"  <dU v,w> {lhs:.15e}  <v,lam0> {rhs:.15e}  rel {abs(lhs - rhs) / abs(lhs):.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, scale)


@pytest.mark.parametrize("snapshots", [2, 4])
def test_adjoint_steps_tangent_dot_product_hexahedra(snapshots):
    """The same for the sum-factorised hexahedron kernels (Taylor-Green vortex 2^3 hex P3, HLLC +
    BR2, 4 SSP-RK3 steps): binomial recomputation (2 snapshots) and the recorded run that keeps
    every stage state (4 snapshots)."""
    steps = 4
    mesh, prob, prim = P.ns_tgv(n=2, order=3)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    dt = s.compute_dt(0.1)
    rng = np.random.default_rng(13)
    u0 = s.get_state()
    u0 = u0 + 1e-3 * _direction(rng, u0, mesh.dim)
    w = rng.standard_normal(u0.shape)
    v = _direction(rng, u0, mesh.dim)
    s.set_state(u0)
    s.adjoint_forward(dt, steps, snapshots)
    lam0 = s.adjoint_backward(w)
    _, dun = O.Oracle(prob, mesh, tangent=True).step_tangent(u0, v, dt, steps)
    lhs, rhs = float((dun * w).sum()), float((v * lam0).sum())
    scale = float(np.abs(dun * w).sum() + np.abs(v * lam0).sum()) / 2
    print(f"  <dU v,w> {lhs:.15e}  <v,lam0> {rhs:.15e}  rel {abs(lhs - rhs) / abs(lhs):.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, scale)


def test_hexahedron_adjoint_schedules_agree_bitwise():
    """Sum-factorised kernels: the recorded run (every stage state and its traces kept) and
    binomial recomputation from 1 or 3 snapshots give bit-identical adjoints."""
    steps = 6
    mesh, prob, prim = P.ns_tgv(n=2, order=3)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    dt = s.compute_dt(0.1)
    u0 = s.get_state()
    s.adjoint_forward(dt, steps, steps)
    _, dJ = P.kinetic_energy(mesh, s.get_state())
    ref = s.adjoint_backward(dJ)
    for snapshots in (1, 3):
        s.set_state(u0)
        s.adjoint_forward(dt, steps, snapshots)
        assert np.array_equal(s.adjoint_backward(dJ), ref), snapshots


def _shear_rollout_setup(n, order, steps, cfl=0.1):
    mesh, prob, prim = P.ns_shear(n=n, order=order, flux=L.FLUX_ROE, visc=L.VISC_BR2)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    dt = s.compute_dt(cfl)
    return mesh, prob, prim, s, dt


def _final_ke(s, mesh, prim, dt, steps):
    s.set_initial_primitive(prim)
    _, nan, taken = s.step(dt, steps)
    assert nan == 0 and taken == steps
    return P.kinetic_energy(mesh, s.get_state())[0]


def _dU0_dv(mesh, prim, dv, cv):
    """Derivative of the projected initial coefficients with respect to the v field: m_y = rho v,
    rho E = rho (c_v T + |u|^2 / 2)."""
    du = np.zeros(prim.shape)
    du[..., 2] = prim[..., 0] * dv
    du[..., 3] = prim[..., 0] * prim[..., 2] * dv
    return P.project_conserved(mesh, du)


def test_adjoint_gradient_matches_finite_differences():
    """Double shear layer (64^2 quads, P3, Roe + BR2, Re 1e4, Ma 0.1), 400 SSP-RK3 steps: the
    gradient of the final domain-integrated kinetic energy with respect to the initial
    v-perturbation field, from the checkpointed adjoint, agrees with central finite differences
    along 16 seeded random smooth directions (seed 20261202, each scaled to the L2 norm of the
    initial v perturbation, central step 1e-6 of that norm) to 1e-5 relative."""
    steps = 400
    mesh, prob, prim, s, dt = _shear_rollout_setup(64, 3, steps)
    u0 = s.get_state()
    s.adjoint_forward(dt, steps, 32)
    uN = s.get_state()
    J, dJ = P.kinetic_energy(mesh, uN)
    lam0 = s.adjoint_backward(dJ)
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, 2)
    jw = mesh.field("jacobian_det_w").reshape(mesh.E, mesh.nq)
    rng = np.random.default_rng(20261202)
    worst = 0.0
    for k in range(16):
        # direction with the L2 norm of the initial v perturbation; step 1e-6 * ||v||
        dv = P.smooth_direction(xq, jw, rng, prim[..., 2])
        g_adj = float((lam0 * _dU0_dv(mesh, prim, dv, prob.cv)).sum())
        eps = 1e-6
        pp, pm = prim.copy(), prim.copy()
        pp[..., 2] += eps * dv
        pm[..., 2] -= eps * dv
        g_fd = (_final_ke(s, mesh, pp, dt, steps) - _final_ke(s, mesh, pm, dt, steps)) / (2 * eps)
        rel = abs(g_adj - g_fd) / max(abs(g_fd), 1e-300)
        worst = max(worst, rel)
        print(f"  direction {k}: adjoint {g_adj:.10e} fd {g_fd:.10e} rel {rel:.2e}", flush=True)
    assert worst <= 1e-5, worst
    assert np.isfinite(J)
    del u0


@pytest.mark.parametrize("snapshots", [1, 3, 39])
def test_checkpoint_schedule_does_not_change_the_adjoint(snapshots):
    """Recomputation from binomial checkpoints is deterministic: every snapshot budget gives
    bit-identical adjoints, including the recorded run that keeps every stage state and its traces
    (traces recomputed from a stored state match the stage kernels' own bitwise)."""
    steps = 40
    mesh, prob, prim, s, dt = _shear_rollout_setup(8, 3, steps)
    u0 = s.get_state()
    s.adjoint_forward(dt, steps, steps)
    _, dJ = P.kinetic_energy(mesh, s.get_state())
    stored = s.adjoint_backward(dJ)
    assert np.array_equal(s.get_state(), u0)  # the context is left at the recorded initial state
    s.adjoint_forward(dt, steps, 2)
    ref = s.adjoint_backward(dJ)
    s.set_state(u0)
    s.adjoint_forward(dt, steps, snapshots)
    got = s.adjoint_backward(dJ)
    assert np.array_equal(got, ref)
    assert np.array_equal(stored, ref)
    assert np.array_equal(s.get_state(), u0)


def test_torch_autograd_wrapper():
    torch = pytest.importorskip("torch")
    from subrosadg_amd.autograd import rollout

    steps = 20
    mesh, prob, prim, s, dt = _shear_rollout_setup(8, 3, steps)
    u0 = torch.tensor(s.get_state(), requires_grad=True)
    V = torch.tensor(mesh.field("basis_value").reshape(mesh.nq, mesh.nb))
    jw = torch.tensor(mesh.field("jacobian_det_w").reshape(mesh.E, mesh.nq))
    uN = rollout(s, u0, dt, steps, snapshots=4)
    uq = torch.einsum("qb,ebv->eqv", V, uN)
    ke = (jw * 0.5 * (uq[..., 1] ** 2 + uq[..., 2] ** 2) / uq[..., 0]).sum()
    ke.backward()
    _, dJ = P.kinetic_energy(mesh, uN.detach().numpy())
    s.set_state(u0.detach().numpy())
    s.adjoint_forward(dt, steps, steps)
    ref = s.adjoint_backward(dJ)
    assert np.abs(u0.grad.numpy() - ref).max() <= 1e-12 * np.abs(ref).max()


def test_device_resident_backward_matches_host():
    """sdg_adjoint_backward_device (adjoints in device memory) gives the host entry point's
    adjoint bit for bit, leaves the recorded initial state and refuses host pointers."""
    steps = 6
    mesh, prob, prim = P.ns_tgv(n=2, order=3)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    dt = s.compute_dt(0.1)
    u0 = s.get_state()
    s.adjoint_forward(dt, steps, 2)
    _, dJ = P.kinetic_energy(mesh, s.get_state())
    ref = s.adjoint_backward(dJ)
    s.set_state(u0)
    s.adjoint_forward(dt, steps, 2)
    got = s.adjoint_backward_device(L.DeviceArray.from_numpy(dJ)).numpy()
    assert np.array_equal(got, ref)
    assert np.array_equal(s.get_state(), u0)
    host = np.zeros(s.state_shape())
    assert s.lib.sdg_adjoint_backward_device(s.handle, host.ctypes.data, host.ctypes.data) == L.E_INVALID


def test_large_mesh_stored_sweep_matches_recomputation():
    """On meshes whose state exceeds 2^22 doubles the stored reverse sweep launches its kernels on
    the recording pools directly (no graph over per-step copies): Taylor-Green 24^3 hex P3 (4.4M
    doubles), 3 steps, identical to binomial recomputation from one snapshot bit for bit."""
    steps = 3
    mesh, prob, prim = P.ns_tgv(n=24, order=3)
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    assert mesh.E * mesh.nb * 5 >= 1 << 22
    dt = s.compute_dt(0.1)
    u0 = s.get_state()
    s.adjoint_forward(dt, steps, steps)
    _, dJ = P.kinetic_energy(mesh, s.get_state())
    stored = s.adjoint_backward(dJ)
    assert np.array_equal(s.get_state(), u0)
    s.adjoint_forward(dt, steps, 1)
    assert np.array_equal(s.adjoint_backward(dJ), stored)


def _mixed(prob):
    q = L.ProblemDesc.from_buffer_copy(prob)
    q.precision = L.PREC_MIXED
    return q


def test_mixed_context_vjp_equals_fp64_context():
    """A mixed-precision context differentiates the FP64 residual on FP64 traces of its FP64 state:
    at the same state its residual VJP is the FP64 context's bit for bit (hexahedra P3, BR2 and
    BR1)."""
    for visc in (L.VISC_BR2, L.VISC_BR1):
        mesh, prob, prim = P.ns_tgv(n=2, order=3, visc=visc)
        s64 = L.Solver(prob, mesh)
        s64.set_initial_primitive(prim)
        u0 = s64.get_state()
        u0 = u0 + 1e-3 * _direction(np.random.default_rng(5), u0, mesh.dim)
        s64.set_state(u0)
        smx = L.Solver(_mixed(prob), mesh)
        smx.set_state(u0)
        lam = np.random.default_rng(6).standard_normal(u0.shape)
        assert np.array_equal(smx.vjp_residual(lam), s64.vjp_residual(lam))


def test_mixed_precision_adjoint_sweep():
    """Checkpointed adjoint of a mixed-precision run (Taylor-Green 3^3 hex P3, 6 SSP-RK3 steps):
    schedule-independent bit for bit (every stage kept, 2 binomial snapshots), and within 1e-5 of the
    FP64 run's adjoint (the two trajectories differ by the mixed mode's 1e-8 rounding)."""
    steps = 6
    mesh, prob, prim = P.ns_tgv(n=3, order=3)
    out = {}
    for label, pr in (("fp64", prob), ("mixed", _mixed(prob))):
        s = L.Solver(pr, mesh)
        s.set_initial_primitive(prim)
        dt = s.compute_dt(0.1) if label == "fp64" else dt
        u0 = s.get_state()
        s.adjoint_forward(dt, steps, steps)
        _, dJ = P.kinetic_energy(mesh, s.get_state())
        lam = s.adjoint_backward(dJ)
        assert np.array_equal(s.get_state(), u0)
        s.adjoint_forward(dt, steps, 2)
        assert np.array_equal(s.adjoint_backward(dJ), lam)
        out[label] = lam
    rel = np.linalg.norm(out["mixed"] - out["fp64"]) / np.linalg.norm(out["fp64"])
    print(f"  mixed vs FP64 adjoint: {rel:.2e}", flush=True)
    assert rel <= 1e-5, rel
