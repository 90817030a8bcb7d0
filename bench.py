"""Benchmark: DG residual + SSP-RK3 stepping of the 3-D compressible Navier-Stokes Taylor-Green
vortex on one MI355X, in the FP64 mode and in the mixed-precision mode.

Workload (default): periodic [0, 2 pi]^3, uniform 32^3 hexahedra, P3 (64 modes per element,
5 variables), Re = 1600, Ma = 0.1, Pr = 0.71, HLLC + BR2, SSP-RK3, 500 timed steps. Synthetic
initial state u = sin x cos y cos z, v = -cos x sin y cos z, w = 0,
p = p0 + (cos 2x + cos 2y)(cos 2z + 2)/16, rho = p / (R T0).

A step is one SSP-RK3 time step (3 stages) of the whole mesh with the state resident in HBM.
Metric: DOF.stage/s = E * N_b * stages * steps / wall time of the timed steps. `value` is the
mixed-precision rate (FP32 matrix-core contractions, FP64 state, geometry, pointwise physics and RK
accumulation); the line also carries the FP64 rate, their ratio and the relative L2 deviation of
the mixed state from the FP64 state after the same steps (per variable group, the largest).

Also printed: a roofline entry for the dominant kernel of each mode (HIP events on the solver's
stream) and the CPU baseline (the oracle on a bounded sample of the same workload on the host
cores).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from subrosadg_amd import _lib as L  # noqa: E402

GAMMA = 1.4
PEAK_FP64_TFLOPS = 78.6        # MI355X dense FP64 (vector and matrix)
PEAK_FP32_MFMA_TFLOPS = 157.3  # MI355X dense FP32 matrix
PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the dominant kernel, from rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled
# by a copy kernel of known size (fetch x2.000, write x1.000); kernel summary in
# profiles/ns_tgv32_p3_fp64_mixed_mfma_kernels.md. Keyed by (kernel, precision mode, cells per
# direction, order).
# Sum-factorised kernels (the default on hexahedra): profiles/ns_tgv32_p3_fp64_mixed_tensor_kernels.md.
PROFILED_TRAFFIC = {("nsGradMfma", "fp64", 32, 3): (190207.1 * 2.0 + 614400.0 * 1.0) * 1024,
                    ("nsGradMfma", "mixed", 32, 3): (144194.1 * 2.0 + 307200.0 * 1.0) * 1024,
                    ("nsGradTensor", "fp64", 32, 3): (194245.8 * 2.0 + 204800.0 * 1.0) * 1024,
                    ("nsGradTensor", "mixed", 32, 3): (145675.7 * 2.0 + 102400.0 * 1.0) * 1024,
                    ("nsFaceFluxKernel", "fp64", 32, 3): (151386.9 * 2.0 + 122880.0 * 1.0) * 1024,
                    ("nsFaceFluxKernel", "mixed", 32, 3): (114860.2 * 2.0 + 61509.7 * 1.0) * 1024,
                    ("nsResidualTensor", "fp64", 32, 3): (184933.9 * 2.0 + 232533.4 * 1.0) * 1024,
                    ("nsResidualTensor", "mixed", 32, 3): (133740.9 * 2.0 + 171093.4 * 1.0) * 1024}
# the same for the batched stage kernel on the 48-problem workload (average over its launches;
# profiles/batch48_uneven_tri_quad_euler_kernels.md)
PROFILED_BATCH_TRAFFIC = {"fp64": (346788.6 * 2.0 + 231499.4 * 1.0) * 1024,
                          "mixed": (312995.3 * 2.0 + 136595.1 * 1.0) * 1024}


def tgv_setup(n: int, order: int, mach=0.1, reynolds=1600.0):
    mesh = L.Mesh(L.ELEM_HEXAHEDRON, order, (n, n, n), (0.0, 0.0, 0.0), (2 * math.pi,) * 3)
    p0 = 1.0 / (GAMMA * mach ** 2)
    R = p0
    cv = R / (GAMMA - 1.0)
    cp = GAMMA * cv
    prob = L.problem_desc(equation=L.EQ_NS, convective_flux=L.FLUX_HLLC, viscous_flux=L.VISC_BR2,
                          time_integration=L.RK_SSPRK3, transport=L.TRANSPORT_CONSTANT,
                          cp=cp, cv=cv, dynamic_viscosity=1.0 / reynolds)
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, 3)
    x, y, z = xq[..., 0], xq[..., 1], xq[..., 2]
    p = p0 + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0) / 16.0
    prim = np.zeros((mesh.E, mesh.nq, 5))
    prim[..., 0] = p / (R * 1.0)
    prim[..., 1] = np.sin(x) * np.cos(y) * np.cos(z)
    prim[..., 2] = -np.cos(x) * np.sin(y) * np.cos(z)
    prim[..., 4] = 1.0
    return mesh, prob, prim


def algorithmic_counts(mesh, ns=True):
    """Dense reference-formulation flop counts per element-stage, split by the kernel that performs
    each contraction, and the FP64 state bytes B_alg.

    F_alg = 2 NV N_b (N_q + 2 N_aq + D N_q + N_b)
          + [NS] 2 NV D N_b (D N_q + N_q + 3 N_aq + N_b (1 + N_f))
    gradient kernel: volume interpolation and every gradient-lifting term;
    residual kernel: weak divergence, face lift, mass inverse and next-stage traces."""
    D, nv, nb, nq = mesh.dim, mesh.nv, mesh.nb, mesh.nq
    naq, nf = mesh.nf * mesh.nfq, mesh.nf
    f_grad = 2 * nv * nb * nq + (2 * nv * D * nb * (D * nq + nq + 3 * naq + nb * (1 + nf)) if ns else 0)
    f_res = 2 * nv * nb * (D * nq + 2 * naq + nb)
    b_alg = 8 * (3 * nv * nb + (2 * nv * D * nb if ns else 0))
    return f_grad, f_res, b_alg


def tensor_path(mesh) -> bool:
    """True where the engine runs the sum-factorised Navier-Stokes kernels (hexahedra P2-P4)."""
    return mesh.elem == L.ELEM_HEXAHEDRON and 2 <= mesh.order <= 4 and os.environ.get("SDG_DENSE") != "1"


def tensor_bytes(mesh, tc: int):
    """Algorithmic HBM bytes per element and launch of the three sum-factorised Navier-Stokes stage
    kernels on a periodic mesh (tc: bytes of the contraction type, 8 FP64 / 4 mixed; the state is
    FP64 in both):
      gradient: state U + own and neighbour face traces + the normal viscous flux Vn at the
                element's face points and the weak divergence Rvol of the volume flux at its
                nodes (both written in the contraction type) + geometry (|J|, J^-1, two slot maps,
                and half of each face point's normal and weight, which the neighbour shares);
      face flux (per element, its half of the face points): both sides' traces and normal viscous
                fluxes, normal, weight, point map, and the two signed flux slots written;
      residual: Rvol + flux slots + U + U_n read, U + next traces written, U_n written at the last
                of the 3 stages (1/3 per launch), |J|."""
    nv, nq, naq, D = mesh.nv, mesh.nq, mesh.nf * mesh.nfq, mesh.dim
    geo = 8 + 8 * D * D + naq * 8 + naq * (8 * D + 8) / 2
    grad = 8 * nv * nq + 2 * nv * naq * tc + nv * naq * tc + nv * nq * tc + geo
    face = naq / 2 * (4 * nv * tc + 8 * D + 8 + 4 + 2 * nv * tc)
    res = nv * nq * tc + nv * naq * tc + 3 * 8 * nv * nq + nv * naq * tc + 8 * nv * nq / 3 + 8
    return grad, face, res


def cpu_baseline(order: int, target_s: float = 15.0):
    """The CPU oracle (C++/OpenMP restatement of the reference) on a bounded sample of the same
    workload: an 8^3 mesh of the same element, order, physics and scheme."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # checker library, used here only as the timed CPU baseline

    cores = os.cpu_count() or 1
    threads = max(1, min(cores - 1, 16)) if cores > 1 else 1  # reference rule: physical cores - 1
    O.set_threads(threads)
    n = 8
    mesh, prob, prim = tgv_setup(n, order)
    o = O.Oracle(prob, mesh)
    o.set_initial_primitive(prim)
    dt = o.compute_dt(0.1)
    t0 = time.perf_counter()
    o.step(dt, 1)
    one = time.perf_counter() - t0
    steps = max(1, min(400, int(target_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    o.step(dt, steps)
    t = time.perf_counter() - t0
    value = mesh.E * mesh.nb * 3 * steps / t
    return {"value": value, "unit": "DOF.stage/s", "cores": O.get_threads(), "kind": "port",
            "sample": f"oracle, TGV {n}^3 hex P{order} HLLC+BR2 SSP-RK3 FP64, {steps} steps, {t:.1f} s"}


def copy_probe(mib: int, reps: int = 5):
    """Calibration workload for the HBM byte counters: an elementwise copy of `mib` MiB (read mib,
    write mib per launch), run under `rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` beside the solver
    so that the counters' scale can be read off a kernel of known traffic."""
    import torch

    n = mib * (1 << 20) // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "reps": reps, "bytes_read": n * 8, "bytes_written": n * 8}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=32, help="elements per direction")
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--cfl", type=float, default=0.1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward timing leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven-mesh leg")
    ap.add_argument("--no-shock", action="store_true", help="skip the artificial-viscosity leg")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run the HBM-counter calibration copy of this many MiB")
    ap.add_argument("--profile-leg", choices=["adjoint", "batch", "shock"], default=None,
                    help="only run one secondary leg (for kernel profiling) and print its record")
    args = ap.parse_args()
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
        return
    if args.profile_leg:
        leg = {"adjoint": adjoint_leg, "batch": batch_leg, "shock": shock_leg}[args.profile_leg](args)
        print(json.dumps(leg), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: runs on one GPU only; sharding the element partition over several GPUs "
              "(face-trace halo exchange per stage) is not implemented", file=sys.stderr)
        sys.exit(2)

    t_setup = time.perf_counter()
    mesh, prob, prim = tgv_setup(args.n, args.order)
    prob_mixed = L.ProblemDesc.from_buffer_copy(prob)
    prob_mixed.precision = L.PREC_MIXED
    runs = {}
    for mode, pd in (("fp64", prob), ("mixed", prob_mixed)):
        runs[mode] = timed_run(mode, pd, mesh, prim, args)
    print(f"[bench] setup+runs {time.perf_counter() - t_setup:.1f}s", file=sys.stderr, flush=True)
    # deviation of the mixed-precision state from the FP64 state after the same warmup + timed steps
    # from the same initial coefficients (the FP64 engine matches the oracle to 1e-12)
    a_m, a_d = runs["mixed"]["state"], runs["fp64"]["state"]
    dev = function_l2(mesh, a_m, a_d)
    dev_coeff = group_l2(a_m, a_d, mesh.dim)
    dev_all = float(np.linalg.norm(a_m - a_d) / np.linalg.norm(a_d))
    dof = mesh.E * mesh.nb
    f_grad, f_res, _ = algorithmic_counts(mesh)
    roof = {mode: roofline(r["kms"], f_grad, f_res, mesh, args, mode) for mode, r in runs.items()}
    head = runs["mixed"]

    adjoint = None if args.no_adjoint else adjoint_leg(args)
    batch = None if args.no_batch else batch_leg(args)
    shock = None if args.no_shock else shock_leg(args)
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.order)
    line = {
        "metric": "DOF.stage/s mixed vs FP64 at fixed L2 error (3-D NS Taylor-Green P3)",
        "value": head["value"],
        "unit": "DOF.stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 contractions and work buffers (sum-factorised line passes), f64 state/geometry/"
                 "physics/RK",
        "data": "synthetic",
        "config": {"workload": f"3-D compressible NS Taylor-Green vortex, periodic [0,2pi]^3, {args.n}^3 hex, "
                               f"P{args.order}, Re 1600, Ma 0.1, HLLC + BR2, SSP-RK3, CFL {args.cfl}",
                   "elements": mesh.E, "basis": mesh.nb, "dof_per_field": dof, "precision": "mixed",
                   "parallelism": "single GPU"},
        "fp64": {"value": runs["fp64"]["value"], "ms_per_step": runs["fp64"]["ms_per_step"], "dtype": "f64"},
        "mixed_over_fp64": head["value"] / runs["fp64"]["value"],
        "mixed_l2_rel_dev_vs_fp64": dev,
        "mixed_l2_rel_dev_definition": "max over density, momentum, energy of ||u_mixed - u_fp64||_L2(domain) / "
                                       "||u_fp64||_L2(domain), after the warmup and timed steps",
        "mixed_coefficient_l2_rel_dev_vs_fp64": dev_coeff,
        "mixed_coefficient_l2_rel_dev_vs_fp64_whole_state": dev_all,
        "l2_dev_after_steps": args.warmup + args.steps,
        "roofline": roof["mixed"],
        "roofline_fp64": roof["fp64"],
        "adjoint": adjoint,
        "batch": batch,
        "shock_capturing": shock,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def shock_leg(args, n=256, order=3, steps=100, cfl=0.1):
    """Artificial-viscosity cost on the explosion case (2-D Euler, [0,2]^2 with adiabatic slip
    walls, a circle of radius 0.4 at rho = 1, p = 1 in rho = 0.125, p = 0.1 spread over half a
    cell, 256^2 quadrangles P3, HLLC, SSP-RK3, setArtificialViscosity(4.0)): the step with the
    per-step sensor and the gradient passes (vector kernels) against the plain Euler step on the
    same mesh and state."""
    slip = L.BC_ADIABATIC_SLIP_WALL
    mesh = L.Mesh(L.ELEM_QUADRANGLE, order, (n, n), (0.0, 0.0), (2.0, 2.0), boundary=((slip, slip), (slip, slip)))
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, 2)
    r = np.sqrt((xq[..., 0] - 1.0) ** 2 + (xq[..., 1] - 1.0) ** 2)
    w = 0.5 * (1.0 - np.tanh((r - 0.4) / (1.0 / n)))
    gas = 2.5 - 25.0 / 14.0
    prim = np.zeros((mesh.E, mesh.nq, 4))
    prim[..., 0] = 0.125 + 0.875 * w
    prim[..., 3] = (0.1 + 0.9 * w) / (prim[..., 0] * gas)
    out = {"workload": f"2-D Euler explosion, [0,2]^2 slip walls, {n}^2 quads, P{order}, HLLC, SSP-RK3, CFL {cfl}, "
                       f"artificial viscosity (tolerance 4.0, factor 1)", "elements": mesh.E, "steps": steps}
    for label, sc in (("artificial_viscosity", L.SHOCK_ARTIFICIAL_VISCOSITY), ("plain", L.SHOCK_NONE)):
        prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_HLLC, cp=2.5, cv=25.0 / 14.0,
                              shock_capturing=sc, empirical_tolerance=4.0, artificial_viscosity_factor=1.0)
        s = L.Solver(prob, mesh)
        s.set_initial_primitive(prim)
        s.set_boundary_primitive(np.zeros((mesh.Fb, mesh.nfq, 4)))
        dt = s.compute_dt(cfl)
        s.step(dt, 5)
        kms = s.time_kernels(dt, 5)
        s.synchronize()
        t0 = time.perf_counter()
        _, nan, taken = s.step(dt, steps)
        s.synchronize()
        el = (time.perf_counter() - t0) / steps
        rec = {"ms_per_step": el * 1e3, "dof_stage_per_s": mesh.E * mesh.nb * 3 / el, "finite": bool(nan == 0),
               "kernel_ms": [float(x) for x in kms]}
        if sc:
            av = s.artificial_viscosity()
            rec["nodes_with_viscosity"] = int((av > 0).sum())
            rec["nodes"] = int(av.size)
        out[label] = rec
    out["cost_over_plain"] = out["artificial_viscosity"]["ms_per_step"] / out["plain"]["ms_per_step"]
    print(f"[bench] shock leg: {out}", file=sys.stderr, flush=True)
    return out


def adjoint_leg(args, n=64, order=3, steps=400, cfl=0.1):
    """Adjoint/forward time ratio on the double shear layer (periodic [0,1]^2, 64^2 quads, P3,
    Roe + BR2, Re 1e4, Ma 0.1, 400 SSP-RK3 steps, FP64): forward = the plain 400-step run;
    adjoint = the checkpointed forward (one snapshot per step) plus the reverse sweep that turns
    dJ/dU(t_N) of the final kinetic energy J into dJ/dU(t_0). One gradient direction is checked
    against central differences of J."""
    gamma, mach = 1.4, 0.1
    mesh = L.Mesh(L.ELEM_QUADRANGLE, order, (n, n), (0.0, 0.0), (1.0, 1.0))
    p0 = 1.0 / (gamma * mach ** 2)
    cv = p0 / (gamma - 1.0)
    prob = L.problem_desc(equation=L.EQ_NS, convective_flux=L.FLUX_ROE, viscous_flux=L.VISC_BR2,
                          time_integration=L.RK_SSPRK3, transport=L.TRANSPORT_CONSTANT, cp=gamma * cv, cv=cv,
                          dynamic_viscosity=1.0e-4)
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, 2)
    x, y = xq[..., 0], xq[..., 1]
    prim = np.zeros((mesh.E, mesh.nq, 4))
    prim[..., 0] = 1.0
    prim[..., 1] = np.where(y <= 0.5, np.tanh(80.0 * (y - 0.25)), np.tanh(80.0 * (0.75 - y)))
    prim[..., 2] = 0.05 * np.sin(2.0 * math.pi * x)
    prim[..., 3] = 1.0
    V = mesh.field("basis_value").reshape(mesh.nq, mesh.nb)
    jw = mesh.field("jacobian_det_w").reshape(mesh.E, mesh.nq)
    lsq = mesh.field("basis_lsq_inverse").reshape(mesh.nb, mesh.nb)

    def ke(U):
        uq = np.einsum("qb,ebv->eqv", V, U)
        m2 = uq[..., 1] ** 2 + uq[..., 2] ** 2
        dj = np.zeros_like(uq)
        dj[..., 0] = -0.5 * m2 / uq[..., 0] ** 2
        dj[..., 1:3] = uq[..., 1:3] / uq[..., 0:1]
        return float((jw * 0.5 * m2 / uq[..., 0]).sum()), np.einsum("eq,qb,eqv->ebv", jw, V, dj)

    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    dt = s.compute_dt(cfl)
    s.step(dt, 5)  # warm-up (graph capture, clocks)
    s.set_initial_primitive(prim)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(dt, steps)
    s.synchronize()
    t_fwd = time.perf_counter() - t0
    s.set_initial_primitive(prim)
    t0 = time.perf_counter()
    s.adjoint_forward(dt, steps, steps)
    t_rec = time.perf_counter() - t0
    J, dJ = ke(s.get_state())
    t0 = time.perf_counter()
    lam0 = s.adjoint_backward(dJ)
    t_bwd = time.perf_counter() - t0
    # one finite-difference check along a seeded random smooth v perturbation with the L2 norm of
    # the initial v field (Fourier modes 0 <= kx, ky <= 3), central step 1e-6 of that norm
    rng = np.random.default_rng(20261202)
    dv = np.zeros_like(x)
    for kx in range(4):
        for ky in range(4):
            if kx or ky:
                a, b = rng.standard_normal(2)
                dv += a * np.cos(2 * math.pi * (kx * x + ky * y)) + b * np.sin(2 * math.pi * (kx * x + ky * y))
    dv *= np.sqrt((jw * prim[..., 2] ** 2).sum() / (jw * dv ** 2).sum())
    du = np.zeros(prim.shape)
    du[..., 2] = prim[..., 0] * dv
    du[..., 3] = prim[..., 0] * prim[..., 2] * dv
    g_adj = float((lam0 * np.einsum("eqv,qc,cb->ebv", du, V, lsq)).sum())
    eps = 1e-6
    js = []
    for sign in (1.0, -1.0):
        pp = prim.copy()
        pp[..., 2] += sign * eps * dv
        s.set_initial_primitive(pp)
        s.step(dt, steps)
        js.append(ke(s.get_state())[0])
    g_fd = (js[0] - js[1]) / (2 * eps)
    s.close()
    out = {"workload": f"2-D NS double shear layer, {n}^2 quads, P{order}, Roe + BR2, Re 1e4, Ma 0.1, "
                       f"{steps} SSP-RK3 steps, FP64, d(final kinetic energy)/d(initial state)",
           "forward_s": t_fwd, "checkpointed_forward_s": t_rec, "reverse_s": t_bwd,
           "adjoint_over_forward": (t_rec + t_bwd) / t_fwd, "reverse_over_forward": t_bwd / t_fwd,
           "snapshots": steps, "fd_check_rel_err": abs(g_adj - g_fd) / abs(g_fd)}
    print(f"[bench] adjoint leg: {out}", file=sys.stderr, flush=True)
    return out


def vortex_primitive(xq, eps=5.0):
    """Shu isentropic vortex on the uniform stream rho = u = v = p = 1 (R = 1): primitive (rho, u,
    v, T) at points xq (..., 2)."""
    x, y = xq[..., 0], xq[..., 1]
    r2 = x * x + y * y
    f = np.exp(0.5 * (1.0 - r2))
    T = 1.0 - (GAMMA - 1.0) * eps * eps / (8.0 * GAMMA * math.pi ** 2) * f * f
    prim = np.empty(xq.shape[:-1] + (4,))
    prim[..., 0] = T ** (1.0 / (GAMMA - 1.0))
    prim[..., 1] = 1.0 - eps / (2 * math.pi) * f * y
    prim[..., 2] = 1.0 + eps / (2 * math.pi) * f * x
    prim[..., 3] = T
    return prim


def batch_leg(args, count=48, seed=20261202, steps=100, cfl=0.1):
    """Batched uneven 2-D Euler: `count` isentropic-vortex problems on periodic [-5,5]^2, per
    problem n ~ U{8..96} cells per direction, P ~ U{1..4}, each cell a quadrangle or two triangles
    by Bernoulli(0.5) (all draws from numpy's default generator with the given seed), Lax-Friedrichs,
    SSP-RK3 with a per-problem CFL-0.1 time step, one persistent launch per stage, whole steps
    captured in a hipGraph, work-item and grid sizes from the autotuner. Timed in mixed precision
    (the headline mode of this workload) and in FP64; the rate is DOF.stage/s over all problems."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        n = int(rng.integers(8, 97))
        order = int(rng.integers(1, 5))
        specs.append((n, order, int(rng.integers(0, 2 ** 62))))
    meshes = [L.Mesh.split(order, (n, n), (-5.0, -5.0), (5.0, 5.0), 0.5, sd) for n, order, sd in specs]
    elements = sum(m.E for m in meshes)
    out = {"workload": f"{count} 2-D Euler isentropic vortices, periodic [-5,5]^2, n ~ U{{8..96}}, P ~ U{{1..4}}, "
                       f"Bernoulli(0.5) triangle split, Lax-Friedrichs, SSP-RK3 CFL {cfl}, {steps} steps, seed {seed}",
           "problems": count, "elements": elements}
    states = {}
    for mode, prec in (("fp64", L.PREC_FP64), ("mixed", L.PREC_MIXED)):
        prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_LAX_FRIEDRICHS,
                              time_integration=L.RK_SSPRK3, cp=3.5, cv=2.5, precision=prec)
        b = L.Batch(prob, meshes)
        prims = [vortex_primitive(m.field("quad_coord").reshape(-1, 2)) for m in meshes]
        for i, pr in enumerate(prims):
            b.set_initial_primitive(i, pr)
        dt = b.compute_dt(cfl)
        stage_ms = b.tune()
        b.step(dt, 3)  # graph capture and warm-up
        for i, pr in enumerate(prims):
            b.set_initial_primitive(i, pr)
        b.step(dt, steps)
        ms = b.last_step_ms()
        dof = sum(sz // 4 for sz in b.sizes)
        chunk, _, grid = b.config()
        out[mode] = {"value": dof * 3 * steps / (ms * 1e-3), "ms_per_step": ms / steps,
                     "tuned_stage_ms": stage_ms, "grid": grid, "tiles_per_item": chunk.tolist()}
        out["dof_per_field"] = dof
        # algorithmic HBM bytes of a stage: read U^i and U^n, write U^{i+1} (FP64 state)
        out[mode]["state_gb_per_s"] = 96.0 * dof * 3 * steps / (ms * 1e-3) / 1e9
        if count == 48 and seed == 20261202:  # profiled HBM bytes per stage launch of this workload
            tb = PROFILED_BATCH_TRAFFIC[mode]
            out[mode]["profiled_traffic_mb"] = tb / 1e6
            out[mode]["profiled_hbm_gb_per_s"] = tb * 3 * steps / (ms * 1e-3) / 1e9
            out[mode]["profiled_hbm_frac"] = out[mode]["profiled_hbm_gb_per_s"] / PEAK_HBM_GBS
        states[mode] = [b.get_state(i) for i in range(count)]
        b.close()
    out["mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    out["mixed_l2_rel_dev_vs_fp64"] = max(group_l2(a.reshape(-1, 4), c.reshape(-1, 4), 2)
                                          for a, c in zip(states["mixed"], states["fp64"]))
    print(f"[bench] batch leg: {out}", file=sys.stderr, flush=True)
    return out


def function_l2(mesh, a, b):
    """max over variable groups of the relative L2(domain) norm of the difference of two solutions
    given as modal coefficients (reference layout E x N_b x NV): quadrature with |J| w at the
    volume points, exact for the squared degree-P polynomials."""
    nv = mesh.nv
    V = mesh.field("basis_value").reshape(mesh.nq, mesh.nb)
    jw = mesh.field("jacobian_det_w").reshape(mesh.E, mesh.nq)
    qa = np.einsum("qb,ebv->eqv", V, (a - b).reshape(mesh.E, mesh.nb, nv))
    qb = np.einsum("qb,ebv->eqv", V, b.reshape(mesh.E, mesh.nb, nv))
    out = []
    for g in ([0], list(range(1, mesh.dim + 1)), [mesh.dim + 1]):
        num = float((jw[..., None] * qa[..., g] ** 2).sum())
        den = float((jw[..., None] * qb[..., g] ** 2).sum())
        out.append(np.sqrt(num / den))
    return max(out)


def group_l2(a, b, dim):
    """max over variable groups (density, momentum vector, energy) of ||a - b||_2 / ||b||_2."""
    nv = dim + 2
    a2, b2 = a.reshape(-1, nv), b.reshape(-1, nv)
    return max(float(np.linalg.norm(a2[:, g] - b2[:, g]) / np.linalg.norm(b2[:, g]))
               for g in ([0], list(range(1, dim + 1)), [dim + 1]))


def timed_run(mode, pd, mesh, prim, args):
    solver = L.Solver(pd, mesh)
    solver.set_initial_primitive(prim)
    dt = solver.compute_dt(args.cfl)
    rel, nan, _ = solver.step(dt, args.warmup)
    solver.synchronize()
    t0 = time.perf_counter()
    rel, nan, taken = solver.step(dt, args.steps)
    solver.synchronize()
    wall = time.perf_counter() - t0
    if nan or taken != args.steps or not np.all(np.isfinite(rel)):
        raise RuntimeError(f"bench: {mode} run did not complete cleanly (nan={nan}, taken={taken})")
    value = mesh.E * mesh.nb * 3 * args.steps / wall
    print(f"[bench] {mode}: dt={dt:.4e} {args.steps} steps in {wall:.3f}s ({1000 * wall / args.steps:.3f} ms/step, "
          f"{value:.4e} DOF.stage/s), monitor {rel.tolist()}", file=sys.stderr, flush=True)
    state = solver.get_state()
    # per-kernel averages with HIP events on the solver stream (advances the state further)
    kms = solver.time_kernels(dt, max(2, min(args.steps, 10)))
    solver.close()
    return {"value": value, "ms_per_step": 1000.0 * wall / args.steps, "state": state, "kms": kms}


def roofline(kms, f_grad, f_res, mesh, args, mode):
    """Roofline entry of the dominant kernel (kms: HIP-event averages of the gradient, face-flux,
    residual and monitor kernels). Sum-factorised path: HBM-bound, algorithmic bytes per launch /
    duration. Dense matrix-core path: algorithmic flops per launch / duration."""
    if tensor_path(mesh):
        per = dict(zip(("nsGradTensor", "nsFaceFluxKernel", "nsResidualTensor"),
                       tensor_bytes(mesh, 8 if mode == "fp64" else 4)))
        names = list(per)
        k = int(np.argmax(kms[:3]))
        name, ms = names[k], kms[k]
        achieved = per[name] * mesh.E / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": PROFILED_TRAFFIC.get((name, mode, args.n, args.order)),
                "kernel": name, "algorithmic_bytes_per_element": per[name], "kernel_ms": ms,
                "kernel_ms_all": kms.tolist(), "kernel_ms_names": names + ["monitor"]}
    peak = PEAK_FP64_TFLOPS if mode == "fp64" else PEAK_FP32_MFMA_TFLOPS
    if kms[0] >= kms[1] + kms[2]:
        name, ms, flops = "nsGradMfma", kms[0], f_grad
    else:
        name, ms, flops = "nsFaceFluxKernel+nsResidualMfma", kms[1] + kms[2], f_res
    achieved = flops * mesh.E / (ms * 1e-3) / 1e12
    return {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": PROFILED_TRAFFIC.get((name, mode, args.n, args.order)), "kernel": name,
            "kernel_ms": ms, "kernel_ms_all": kms.tolist()}


if __name__ == "__main__":
    main()
