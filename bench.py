"""Benchmark: DG residual + SSP-RK3 stepping on one MI355X.

Headline workload (`value`): the 3-D compressible Navier-Stokes Taylor-Green vortex, periodic
[0, 2 pi]^3, uniform 32^3 hexahedra, P3 (64 modes per element, 5 variables), Re = 1600, Ma = 0.1,
Pr = 0.71, HLLC + BR2, SSP-RK3, CFL 0.1, with synthetic initial state u = sin x cos y cos z,
v = -cos x sin y cos z, w = 0, p = p0 + (cos 2x + cos 2y)(cos 2z + 2)/16, rho = p / (R T0). A step is
one SSP-RK3 time step (3 stages) of the whole mesh with the state resident in HBM.
Metric: DOF.stage/s = E * N_b * stages * steps / wall time of the timed steps, the median of 5 runs
of K steps each (after W warm-up steps). `value` is the mixed-precision rate (FP32 contractions and
work buffers, FP64 state, geometry, pointwise physics and RK accumulation); the line also carries the
FP64 rate, their ratio and the relative L2 deviation of the mixed state from the FP64 state after the
same steps.

Further legs (same JSON line): the 2-D Euler isentropic vortex (32^2 quadrangles P2, Roe, 200 steps)
with its final state against the oracle; the 64^3 hexahedron P4 Taylor-Green vortex in FP64 and
mixed precision; the 48-problem batched uneven 2-D Euler workload; the discrete adjoint of the 2-D
double shear layer (adjoint / forward time at full storage and at a binomial snapshot budget); the
adjoint of the 64^3 P4 Taylor-Green vortex's kinetic-energy decay with respect to the velocity
amplitude; artificial-viscosity shock capturing. A roofline entry for the dominant kernel (HIP events
on the solver's stream, bytes and flops as defined below) and the CPU baseline (the oracle on the
same 32^3 mesh on the host's cores).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from subrosadg_amd import _lib as L  # noqa: E402

GAMMA = 1.4
PEAK_FP64_TFLOPS = 78.6        # MI355X dense FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3       # MI355X dense FP32 (vector and matrix)
PEAK_HBM_GBS = 8000.0
RUNS = 5                       # timed runs per configuration; the median is reported

# HBM bytes per launch of the stage kernels on the 32^3 P3 workload, from rocprofv3 FETCH_SIZE /
# WRITE_SIZE passes scaled by a copy kernel of known size (fetch x2.000, write x1.000):
# profiles/ns_tgv32_p3_fp64_mixed_tensor_kernels.md. Keyed by (kernel, precision mode, cells per direction, order).
PROFILED_TRAFFIC = {("nsGradTensor", "fp64", 32, 3): (194707.8 * 2.0 + 204800.0 * 1.0) * 1024,
                    ("nsGradTensor", "mixed", 32, 3): (146875.7 * 2.0 + 102400.0 * 1.0) * 1024,
                    ("nsFaceFluxKernel", "fp64", 32, 3): (151378.2 * 2.0 + 122880.0 * 1.0) * 1024,
                    ("nsFaceFluxKernel", "mixed", 32, 3): (114859.4 * 2.0 + 61514.2 * 1.0) * 1024,
                    ("nsResidualTensor", "fp64", 32, 3): (184931.6 * 2.0 + 232533.4 * 1.0) * 1024,
                    ("nsResidualTensor", "mixed", 32, 3): (133738.0 * 2.0 + 171093.4 * 1.0) * 1024}
# the batched stage on the 48-problem workload: one launch per element variant (triangle and quadrangle,
# P1-P4), FETCH_SIZE (x2.0) + WRITE_SIZE (x1.0) KiB per launch, summed over the 8 variants of a stage
# (profiles/batch48_uneven_tri_quad_euler_kernels.md)
PROFILED_BATCH_TRAFFIC = {"fp64": (568801.4) * 1024,
                          "mixed": (494566.9) * 1024}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---- workloads -----------------------------------------------------------------------------------
def tgv_setup(n: int, order: int, mach=0.1, reynolds=1600.0):
    mesh = L.Mesh(L.ELEM_HEXAHEDRON, order, (n, n, n), (0.0, 0.0, 0.0), (2 * math.pi,) * 3)
    p0 = 1.0 / (GAMMA * mach ** 2)
    R = p0
    cv = R / (GAMMA - 1.0)
    cp = GAMMA * cv
    prob = L.problem_desc(equation=L.EQ_NS, convective_flux=L.FLUX_HLLC, viscous_flux=L.VISC_BR2,
                          time_integration=L.RK_SSPRK3, transport=L.TRANSPORT_CONSTANT,
                          cp=cp, cv=cv, dynamic_viscosity=1.0 / reynolds)
    return mesh, prob, tgv_primitive(mesh, p0, R)


def tgv_primitive(mesh, p0, R, amplitude=1.0):
    """Taylor-Green primitive state (rho, u, v, w, T) at the volume points; the velocity scaled by
    `amplitude` and the pressure perturbation by its square (the incompressible pressure of the
    scaled field), rho = p / (R T0), T0 = 1."""
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, 3)
    x, y, z = xq[..., 0], xq[..., 1], xq[..., 2]
    p = p0 + amplitude ** 2 * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0) / 16.0
    prim = np.zeros((mesh.E, mesh.nq, 5))
    prim[..., 0] = p / (R * 1.0)
    prim[..., 1] = amplitude * np.sin(x) * np.cos(y) * np.cos(z)
    prim[..., 2] = -amplitude * np.cos(x) * np.sin(y) * np.cos(z)
    prim[..., 4] = 1.0
    return prim


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


# ---- algorithmic counts ------------------------------------------------------------------------------
def algorithmic_counts(mesh, ns=True, mixed=False):
    """Per element and stage, the dense reference formulation (each distinct contraction counted
    once, pointwise flux flops excluded):
      F_alg = 2 NV N_b (N_q + 2 N_aq + D N_q + N_b) + [NS] 2 NV D N_b (D N_q + N_q + 3 N_aq + N_b (1 + N_f))
      B_alg = 8 (3 NV N_b) + [NS] 8 (2 NV D N_b) in FP64; the gradient term at 4 bytes in mixed mode
    (read U^i and U^n, write U^{i+1}; write and read the gradient coefficients)."""
    D, nv, nb, nq = mesh.dim, mesh.nv, mesh.nb, mesh.nq
    naq, nf = mesh.nf * mesh.nfq, mesh.nf
    f_alg = 2 * nv * nb * (nq + 2 * naq + D * nq + nb)
    if ns:
        f_alg += 2 * nv * D * nb * (D * nq + nq + 3 * naq + nb * (1 + nf))
    b_alg = 8 * 3 * nv * nb + ((4 if mixed else 8) * 2 * nv * D * nb if ns else 0)
    return f_alg, b_alg


def tensor_flops(order: int) -> int:
    """Contraction flops per element and stage that the sum-factorised hexahedron kernels execute
    (N = P+1 nodes per direction, N^3 nodes, 6 N^2 face points; pointwise physics excluded, as in
    F_alg): modal -> nodal state, 3 line passes of N-term sums over NV components (3 NV N^4 MACs);
    reference derivatives along 3 axes (3 NV N^4); the chain rule J^-T (9 NV N^3); the six faces'
    lift corrections (6 NV 2 D N^3); traces of the 15 gradient components (15 6 N^2 N); the weak
    divergence Dn^T along 3 axes after the contravariant map (3 NV N^4 + 9 NV N^3); the face lift
    of the fluxes (6 NV N^3); back to modal coefficients, 3 B^-1 line passes (3 NV N^4); the next
    stage's traces (NV 6 N^2 N). Two flops per MAC."""
    N = order + 1
    nv = 5
    macs = 12 * nv * N ** 4 + (9 + 12 + 9 + 6) * nv * N ** 3 + 15 * 6 * N ** 3 + nv * 6 * N ** 3
    return 2 * macs


def tensor_path(mesh) -> bool:
    """True where the engine runs the sum-factorised Navier-Stokes kernels (hexahedra P2-P4)."""
    return mesh.elem == L.ELEM_HEXAHEDRON and 2 <= mesh.order <= 4


def roofline(kms, mesh, mode, n, kernels="auto"):
    """Roofline entry of the dominant kernel. kms: HIP-event averages of the gradient, face-flux and
    residual kernels (and the monitor) on the solver's stream. `achieved`: the algorithmic bytes
    B_alg per element-stage times the elements one launch processes, over the launch's duration
    (B_alg counts the whole stage's state traffic, so this is the rate the stage would need if the
    kernel were all of it). Also reported: the flop side of the same count (F_alg, the dense
    formulation) and of the flops the sum-factorised kernels execute, and the whole stage (the
    three kernels) against B_alg."""
    mixed = mode == "mixed"
    f_alg, b_alg = algorithmic_counts(mesh, mixed=mixed)
    tensor = tensor_path(mesh) and kernels == "auto"
    if kernels == "vector":  # no separate face-flux kernel: its interval is empty
        names = ["nsGradKernel", "(none)", "nsFluxKernel"]
    elif tensor:
        names = ["nsGradTensor", "nsFaceFluxKernel", "nsResidualTensor"]
    else:
        names = ["nsGradMfma", "nsFaceFluxKernel", "nsResidualMfma"]
    k = int(np.argmax(kms[:3]))
    name, ms = names[k], float(kms[k])
    stage_ms = float(sum(kms[:3]))
    peak_f = PEAK_FP32_TFLOPS if mixed else PEAK_FP64_TFLOPS
    achieved = b_alg * mesh.E / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
           "traffic": PROFILED_TRAFFIC.get((name, mode, n, mesh.order)), "kernel": name, "kernel_ms": ms,
           "algorithmic_bytes_per_element_stage": b_alg, "kernel_ms_all": [float(x) for x in kms],
           "kernel_ms_names": names + ["monitor"], "stage_ms": stage_ms,
           "stage_hbm_frac": b_alg * mesh.E / (stage_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
           "flop_alg_per_element_stage": f_alg,
           "flop_alg_frac_of_peak_kernel": f_alg * mesh.E / (ms * 1e-3) / 1e12 / peak_f,
           "flop_alg_frac_of_peak_stage": f_alg * mesh.E / (stage_ms * 1e-3) / 1e12 / peak_f}
    if tensor:
        f_sf = tensor_flops(mesh.order)
        out["flop_executed_per_element_stage"] = f_sf
        out["flop_executed_frac_of_peak_stage"] = f_sf * mesh.E / (stage_ms * 1e-3) / 1e12 / peak_f
    return out


# ---- measurement helpers -----------------------------------------------------------------------------
def timed_steps(solver, dt, steps, runs=RUNS):
    """`runs` timed runs of `steps` steps each (the state advances through all of them); wall time
    per run between stream synchronisations. Returns the per-run ms/step."""
    out = []
    for _ in range(runs):
        solver.synchronize()
        t0 = time.perf_counter()
        rel, nan, taken = solver.step(dt, steps)
        solver.synchronize()
        wall = time.perf_counter() - t0
        if nan or taken != steps or not np.all(np.isfinite(rel)):
            raise RuntimeError(f"bench: run did not complete cleanly (nan={nan}, taken={taken})")
        out.append(1000.0 * wall / steps)
    return out


def function_l2(mesh, a, b):
    """max over variable groups of the relative L2(domain) norm of the difference of two solutions
    given as modal coefficients (reference layout E x N_b x NV): quadrature with |J| w at the
    volume points, exact for the squared degree-P polynomials."""
    nv = mesh.nv
    V = mesh.field("basis_value").reshape(mesh.nq, mesh.nb)
    jw = mesh.field("jacobian_det_w").reshape(mesh.E, mesh.nq)
    out = []
    for g in ([0], list(range(1, mesh.dim + 1)), [mesh.dim + 1]):
        num = den = 0.0
        for c0 in range(0, mesh.E, 32768):  # chunks: the 64^3 P4 point values do not fit at once
            c1 = min(mesh.E, c0 + 32768)
            da = (a - b).reshape(mesh.E, mesh.nb, nv)[c0:c1][..., g]
            bb = b.reshape(mesh.E, mesh.nb, nv)[c0:c1][..., g]
            qa = np.einsum("qb,ebv->eqv", V, da)
            qb = np.einsum("qb,ebv->eqv", V, bb)
            num += float((jw[c0:c1, :, None] * qa ** 2).sum())
            den += float((jw[c0:c1, :, None] * qb ** 2).sum())
        out.append(math.sqrt(num / den))
    return max(out)


def group_l2(a, b, dim):
    """max over variable groups (density, momentum vector, energy) of ||a - b||_2 / ||b||_2."""
    nv = dim + 2
    a2, b2 = a.reshape(-1, nv), b.reshape(-1, nv)
    return max(float(np.linalg.norm(a2[:, g] - b2[:, g]) / np.linalg.norm(b2[:, g]))
               for g in ([0], list(range(1, dim + 1)), [dim + 1]))


def tgv_modes(mesh, prob, prim, args, label, runs=RUNS, kernel_steps=10, modes=("fp64", "mixed")):
    """FP64 and mixed runs of one Taylor-Green configuration from the same initial coefficients:
    per mode the median rate of `runs` runs of args.steps steps after args.warmup steps, the
    per-kernel HIP-event averages, and the mixed state's deviation from the FP64 state."""
    out, states, coeff0 = {}, {}, None
    for mode in modes:
        pd = L.ProblemDesc.from_buffer_copy(prob)
        pd.precision = L.PREC_MIXED if mode == "mixed" else L.PREC_FP64
        s = L.Solver(pd, mesh)
        if coeff0 is None:
            s.set_initial_primitive(prim)
            coeff0 = s.get_state()
        else:
            s.set_state(coeff0)
        dt = s.compute_dt(args.cfl) if mode == "fp64" else out["fp64"]["dt"]
        s.step(dt, args.warmup)
        ms = timed_steps(s, dt, args.steps, runs)
        states[mode] = s.get_state()
        kms = s.time_kernels(dt, kernel_steps)
        s.close()
        med = statistics.median(ms)
        out[mode] = {"value": mesh.E * mesh.nb * 3 / (med * 1e-3), "ms_per_step": med, "ms_per_step_runs": ms,
                     "dt": dt, "kms": kms}
        log(f"{label} {mode}: {med:.3f} ms/step (runs {[round(x, 3) for x in ms]}), "
            f"{out[mode]['value']:.4e} DOF.stage/s, kernels {[round(float(x), 4) for x in kms]} ms")
    if "mixed" not in states:
        return out
    out["mixed_l2_rel_dev_vs_fp64"] = function_l2(mesh, states["mixed"], states["fp64"])
    out["mixed_coefficient_l2_rel_dev_vs_fp64"] = group_l2(states["mixed"], states["fp64"], mesh.dim)
    out["steps_compared"] = args.warmup + runs * args.steps
    return out


# ---- legs ------------------------------------------------------------------------------------------
def c1_leg(args, n=32, order=2, steps=200):
    """2-D Euler isentropic vortex: periodic [-5, 5]^2, 32^2 quadrangles P2 (3x3 Gauss, the reference's
    order-2P rule), Roe, SSP-RK3 at CFL 0.1, 200 steps, FP64; the final state against the oracle."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    mesh = L.Mesh(L.ELEM_QUADRANGLE, order, (n, n), (-5.0, -5.0), (5.0, 5.0))
    prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_ROE, time_integration=L.RK_SSPRK3,
                          cp=3.5, cv=2.5)
    prim = vortex_primitive(mesh.field("quad_coord").reshape(mesh.E, mesh.nq, 2))
    s = L.Solver(prob, mesh)
    s.set_initial_primitive(prim)
    u0 = s.get_state()
    dt = s.compute_dt(0.1)
    s.step(dt, 5)
    ms = []
    for _ in range(RUNS):
        s.set_state(u0)
        ms += timed_steps(s, dt, steps, runs=1)
    final = s.get_state()
    s.close()
    med = statistics.median(ms)
    rec = {"workload": f"2-D Euler isentropic vortex, periodic [-5,5]^2, {n}^2 quads, P{order}, Roe, SSP-RK3 CFL 0.1, "
                       f"{steps} steps, FP64", "value": mesh.E * mesh.nb * 3 / (med * 1e-3), "unit": "DOF.stage/s",
           "ms_per_step": med, "ms_per_step_runs": ms, "steps": steps}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # checker only: the final state of the same run on the CPU
        o = O.Oracle(prob, mesh)
        o.set_state(u0)
        o.step(dt, steps)
        uo = o.get_state()
        rec["final_state_max_rel_err_vs_oracle"] = float(np.abs(final - uo).max() / np.abs(uo).max())
    log(f"C1: {rec}")
    return rec


def c5_leg(args, n=64, order=4):
    """3-D NS Taylor-Green vortex on 64^3 hexahedra P4 (32.8 M DOF per field): FP64 and mixed."""
    mesh, prob, prim = tgv_setup(n, order)
    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.warmup = args.c5_steps, 2
    rec = tgv_modes(mesh, prob, prim, sub, "C5", runs=3, kernel_steps=2)
    res = {"workload": f"3-D NS Taylor-Green vortex, periodic [0,2pi]^3, {n}^3 hex, P{order}, Re 1600, Ma 0.1, "
                       f"HLLC + BR2, SSP-RK3 CFL {args.cfl}", "elements": mesh.E, "dof_per_field": mesh.E * mesh.nb}
    for mode in ("fp64", "mixed"):
        r = rec[mode]
        res[mode] = {"value": r["value"], "ms_per_step": r["ms_per_step"], "ms_per_step_runs": r["ms_per_step_runs"],
                     "roofline": roofline(r["kms"], mesh, mode, n)}
    res["mixed_over_fp64"] = rec["mixed"]["value"] / rec["fp64"]["value"]
    res["mixed_l2_rel_dev_vs_fp64"] = rec["mixed_l2_rel_dev_vs_fp64"]
    res["steps_compared"] = rec["steps_compared"]
    log(f"C5: {res}")
    return res


def batch_leg(args, count=48, seed=20261202, steps=100, cfl=0.1):
    """Batched uneven 2-D Euler: `count` isentropic-vortex problems on periodic [-5,5]^2, per
    problem n ~ U{8..96} cells per direction, P ~ U{1..4}, each cell a quadrangle or two triangles
    by Bernoulli(0.5) (all draws from numpy's default generator with the given seed), Lax-Friedrichs,
    SSP-RK3 with a per-problem CFL-0.1 time step, one persistent launch per stage, whole steps
    captured in a hipGraph, work-item and grid sizes from the autotuner. Timed in mixed precision
    (the headline mode of this workload) and in FP64 (median of 5 runs of `steps` steps from the
    initial state); the rate is DOF.stage/s over all problems."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import problems as P  # the workload's draws and the split-grid L2 norm

    specs = P.c3_specs(count, seed)
    meshes = [L.Mesh.split(order, (n, n), (-5.0, -5.0), (5.0, 5.0), 0.5, sd) for n, order, sd in specs]
    elements = sum(m.E for m in meshes)
    out = {"workload": f"{count} 2-D Euler isentropic vortices, periodic [-5,5]^2, n ~ U{{8..96}}, P ~ U{{1..4}}, "
                       f"Bernoulli(0.5) triangle split, Lax-Friedrichs, SSP-RK3 CFL {cfl}, {steps} steps, seed {seed}",
           "problems": count, "elements": elements}
    states = {}
    prims = [vortex_primitive(m.field("quad_coord").reshape(-1, 2)) for m in meshes]
    for mode, prec in (("fp64", L.PREC_FP64), ("mixed", L.PREC_MIXED)):
        prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_LAX_FRIEDRICHS,
                              time_integration=L.RK_SSPRK3, cp=3.5, cv=2.5, precision=prec)
        b = L.Batch(prob, meshes)
        for i, pr in enumerate(prims):
            b.set_initial_primitive(i, pr)
        u0 = [b.get_state(i) for i in range(count)]
        dt = b.compute_dt(cfl)
        stage_ms = b.tune()
        b.step(dt, 3)  # graph capture and warm-up
        ms = []
        for _ in range(RUNS):
            for i in range(count):
                b.set_state(i, u0[i])
            b.step(dt, steps)
            ms.append(b.last_step_ms())
        dof = sum(sz // 4 for sz in b.sizes)
        chunk, _, grid = b.config()
        med = statistics.median(ms)
        out[mode] = {"value": dof * 3 * steps / (med * 1e-3), "ms_per_step": med / steps,
                     "ms_per_step_runs": [x / steps for x in ms], "tuned_stage_ms": stage_ms, "grid": grid,
                     "tiles_per_item": chunk.tolist()}
        out["dof_per_field"] = dof
        # algorithmic HBM bytes of a stage: read U^i and U^n, write U^{i+1} (FP64 state)
        out[mode]["state_gb_per_s"] = 96.0 * dof * 3 * steps / (med * 1e-3) / 1e9
        out[mode]["state_hbm_frac"] = out[mode]["state_gb_per_s"] / PEAK_HBM_GBS
        if count == 48 and seed == 20261202:  # profiled HBM bytes per stage launch of this workload
            tb = PROFILED_BATCH_TRAFFIC[mode]
            out[mode]["profiled_traffic_mb"] = tb / 1e6
        states[mode] = [b.get_state(i) for i in range(count)]
        b.close()
    out["mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    devs = [P.split_function_l2(m, a, c) for m, a, c in zip(meshes, states["mixed"], states["fp64"])]
    out["mixed_l2_rel_dev_vs_fp64"] = max(devs)
    out["mixed_l2_rel_dev_definition"] = ("max over problems and over density, momentum, energy of "
                                          "||u_mixed - u_fp64||_L2(domain) / ||u_fp64||_L2(domain) after the steps")
    out["mixed_coefficient_l2_rel_dev_vs_fp64"] = max(group_l2(a.reshape(-1, 4), c.reshape(-1, 4), 2)
                                                      for a, c in zip(states["mixed"], states["fp64"]))
    log(f"batch leg: {out}")
    return out


def adjoint_leg(args, n=64, order=3, steps=400, cfl=0.1, budget=32):
    """Adjoint/forward time ratio on the double shear layer (periodic [0,1]^2, 64^2 quads, P3,
    Roe + BR2, Re 1e4, Ma 0.1, 400 SSP-RK3 steps, FP64): forward = the plain 400-step run;
    adjoint = the checkpointed forward plus the reverse sweep that turns dJ/dU(t_N) of the final
    kinetic energy J into dJ/dU(t_0), with every step stored (no recomputation) and with a binomial
    budget of 32 snapshots (recomputation from the checkpoints). One gradient direction is checked
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
    u0 = s.get_state()
    dt = s.compute_dt(cfl)
    s.step(dt, 5)  # warm-up (graph capture, clocks)

    def sync_time(fn):
        s.synchronize()
        t0 = time.perf_counter()
        r = fn()
        s.synchronize()
        return time.perf_counter() - t0, r

    t_fwd = statistics.median([sync_time(lambda: (s.set_state(u0), s.step(dt, steps)))[0] for _ in range(3)])
    rec = {}
    lam0 = None
    for label, snaps in (("full_storage", steps), ("binomial", budget)):
        s.set_state(u0)
        t_rec, _ = sync_time(lambda: s.adjoint_forward(dt, steps, snaps))
        J, dJ = ke(s.get_state())
        t_bwd, lam = sync_time(lambda: s.adjoint_backward(dJ))
        if lam0 is None:
            lam0 = lam
        rec[label] = {"snapshots": snaps, "checkpointed_forward_s": t_rec, "reverse_s": t_bwd,
                      "adjoint_over_forward": (t_rec + t_bwd) / t_fwd, "identical_to_full_storage":
                      bool(np.array_equal(lam, lam0))}
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
           "forward_s": t_fwd, "forward_dof_stage_per_s": mesh.E * mesh.nb * 3 * steps / t_fwd,
           **rec, "adjoint_over_forward": rec["binomial"]["adjoint_over_forward"],
           "fd_check_rel_err": abs(g_adj - g_fd) / abs(g_fd)}
    log(f"adjoint leg: {out}")
    return out


def mixed_desc(prob):
    q = L.ProblemDesc.from_buffer_copy(prob)
    q.precision = L.PREC_MIXED
    return q


def sync_time_of(solver, fn):
    solver.synchronize()
    t0 = time.perf_counter()
    r = fn()
    solver.synchronize()
    return time.perf_counter() - t0, r


def c5_adjoint_leg(args, n=64, order=4, steps=None):
    """Gradient of the time-integrated kinetic-energy dissipation int_0^T eps_K dt = K(0) - K(T)
    of the 64^3 P4 Taylor-Green vortex (HLLC + BR2, FP64, fixed dt) with respect to the initial
    velocity amplitude A (u = A sin x cos y cos z, ..., p = p0 + A^2 (...)/16): the checkpointed
    adjoint of `steps` SSP-RK3 steps (sum-factorised adjoint kernels), with every stage state kept
    and with a binomial budget of steps/4 snapshots, against the central finite difference of the
    objective in A. The reverse sweep is timed with dJ/dU(T) resident in HBM
    (sdg_adjoint_backward_device); the host entry point's time, which adds the layout conversion
    and the PCIe copies of two 1.3 GB arrays, is reported beside it."""
    steps = steps or args.c5_adjoint_steps
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import problems as P

    mesh, prob, _ = tgv_setup(n, order)
    p0 = 1.0 / (GAMMA * 0.01)
    s = L.Solver(prob, mesh)
    A = 1.0

    def u_of(amp):
        s.set_initial_primitive(tgv_primitive(mesh, p0, p0, amp))
        return s.get_state()

    def sync_time(fn):
        s.synchronize()
        t0 = time.perf_counter()
        r = fn()
        s.synchronize()
        return time.perf_counter() - t0, r

    u0 = u_of(A)
    dt = s.compute_dt(args.cfl)
    K0, dK0 = P.kinetic_energy(mesh, u0)
    s.step(dt, 1)  # warm-up (graph capture)
    fw = []
    for _ in range(3):
        s.set_state(u0)
        fw.append(sync_time(lambda: s.step(dt, steps))[0])
    t_fwd = statistics.median(fw)
    rec = {}
    lam0 = None
    for label, snaps in (("full_storage", steps), ("binomial", max(2, steps // 4))):
        s.set_state(u0)
        t_rec, _ = sync_time(lambda: s.adjoint_forward(dt, steps, snaps))
        KT, dKT = P.kinetic_energy(mesh, s.get_state())
        lam_dev = L.DeviceArray.from_numpy(-dKT)
        out_dev = L.DeviceArray(lam_dev.shape)
        t_bwd, _ = sync_time(lambda: s.adjoint_backward_device(lam_dev, out_dev))
        lam = out_dev.numpy()
        if lam0 is None:
            lam0 = lam
        lam_last = lam
        rec[label] = {"snapshots": snaps, "checkpointed_forward_s": t_rec, "reverse_s": t_bwd,
                      "adjoint_over_forward": (t_rec + t_bwd) / t_fwd,
                      "identical_to_full_storage": bool(np.array_equal(lam, lam0))}
        del lam_dev, out_dev
    # the host entry point on the binomial schedule (PCIe-inclusive; not the figure of record)
    s.set_state(u0)
    s.adjoint_forward(dt, steps, rec["binomial"]["snapshots"])
    KT, dKT = P.kinetic_energy(mesh, s.get_state())
    mdKT = -dKT
    t_host, lam_h = sync_time(lambda: s.adjoint_backward(mdKT))
    rec["binomial"]["reverse_host_api_s"] = t_host
    rec["binomial"]["host_api_identical"] = bool(np.array_equal(lam_h, lam_last))
    del lam_h, lam_last, mdKT
    # dJ/dA = <dK(0)/dU0 + lambda_0, dU0/dA>, dU0/dA by central differences of the projection
    h = 1e-4
    du0 = (u_of(A + h) - u_of(A - h)) / (2 * h)
    g_adj = float(((dK0 + lam0) * du0).sum())
    eps = 1e-3
    js = []
    for sign in (1.0, -1.0):
        ua = u_of(A + sign * eps)
        ka = P.kinetic_energy(mesh, ua, gradient=False)[0]
        s.step(dt, steps)
        js.append(ka - P.kinetic_energy(mesh, s.get_state(), gradient=False)[0])
    g_fd = (js[0] - js[1]) / (2 * eps)
    s.close()
    # the same in a mixed-precision context: FP32 contractions and fluxes in the recorded and
    # recomputed forward runs, the FP64 adjoint kernels on FP64 traces of the stored FP64 states
    sm = L.Solver(mixed_desc(prob), mesh)
    sm.set_initial_primitive(tgv_primitive(mesh, p0, p0, A))
    sm.step(dt, 1)
    fwm = []
    for _ in range(3):
        sm.set_state(u0)
        fwm.append(sync_time_of(sm, lambda: sm.step(dt, steps))[0])
    t_fwd_m = statistics.median(fwm)
    sm.set_state(u0)
    t_rec_m, _ = sync_time_of(sm, lambda: sm.adjoint_forward(dt, steps, steps))
    _, dKTm = P.kinetic_energy(mesh, sm.get_state())
    lam_dev = L.DeviceArray.from_numpy(-dKTm)
    out_dev = L.DeviceArray(lam_dev.shape)
    t_bwd_m, _ = sync_time_of(sm, lambda: sm.adjoint_backward_device(lam_dev, out_dev))
    g_adj_m = float(((dK0 + out_dev.numpy()) * du0).sum())
    del lam_dev, out_dev
    sm.close()
    mixed = {"forward_s": t_fwd_m, "forward_dof_stage_per_s": mesh.E * mesh.nb * 3 * steps / t_fwd_m,
             "full_storage": {"snapshots": steps, "checkpointed_forward_s": t_rec_m, "reverse_s": t_bwd_m,
                              "adjoint_over_forward": (t_rec_m + t_bwd_m) / t_fwd_m,
                              "adjoint_over_fp64_forward": (t_rec_m + t_bwd_m) / t_fwd},
             "gradient": g_adj_m, "rel_diff_vs_fp64_gradient": abs(g_adj_m - g_adj) / abs(g_adj),
             "fd_check_rel_err": abs(g_adj_m - g_fd) / abs(g_fd)}
    out = {"workload": f"3-D NS Taylor-Green vortex {n}^3 hex P{order}, HLLC + BR2, FP64, {steps} SSP-RK3 steps, "
                       f"d(K(0) - K(T))/d(velocity amplitude)", "steps": steps, "forward_s": t_fwd,
           "forward_dof_stage_per_s": mesh.E * mesh.nb * 3 * steps / t_fwd, **rec,
           "adjoint_over_forward": rec["binomial"]["adjoint_over_forward"], "dissipation": K0 - KT,
           "gradient": g_adj, "gradient_fd": g_fd, "fd_check_rel_err": abs(g_adj - g_fd) / abs(g_fd),
           "mixed": mixed}
    log(f"C5 adjoint leg: {out}")
    return out


def shock_leg(args, n=256, order=3, steps=100, cfl=0.1):
    """Artificial-viscosity cost on the explosion case (2-D Euler, [0,2]^2 with adiabatic slip
    walls, a circle of radius 0.4 at rho = 1, p = 1 in rho = 0.125, p = 0.1 spread over half a
    cell, 256^2 quadrangles P3, HLLC, SSP-RK3, setArtificialViscosity(4.0)): the step with the
    per-step sensor and the gradient passes (matrix-core kernels, FP64) against the plain Euler step
    on the same mesh and state."""
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
    V = mesh.field("basis_value").reshape(mesh.nq, mesh.nb)
    jwq = mesh.field("jacobian_det_w").reshape(mesh.E, mesh.nq)
    finals = {}
    for label, sc, prec in (("artificial_viscosity", L.SHOCK_ARTIFICIAL_VISCOSITY, L.PREC_FP64),
                            ("artificial_viscosity_mixed", L.SHOCK_ARTIFICIAL_VISCOSITY, L.PREC_MIXED),
                            ("plain", L.SHOCK_NONE, L.PREC_FP64)):
        prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_HLLC, cp=2.5, cv=25.0 / 14.0,
                              shock_capturing=sc, empirical_tolerance=4.0, artificial_viscosity_factor=1.0,
                              precision=prec)
        s = L.Solver(prob, mesh)
        s.set_initial_primitive(prim)
        s.set_boundary_primitive(np.zeros((mesh.Fb, mesh.nfq, 4)))
        u0 = s.get_state()
        dt = s.compute_dt(cfl)
        s.step(dt, 5)
        kms = s.time_kernels(dt, 5)
        ms = []
        for _ in range(3):  # each run from the initial state (the blast's first `steps` steps)
            s.set_state(u0)
            ms += timed_steps(s, dt, steps, runs=1)
        med = statistics.median(ms)
        rec = {"ms_per_step": med, "dof_stage_per_s": mesh.E * mesh.nb * 3 / (med * 1e-3),
               "kernel_ms": [float(x) for x in kms]}
        finals[label] = np.einsum("qb,ebv->eqv", V, s.get_state())
        if sc:
            av = s.artificial_viscosity()
            rec["nodes_with_viscosity"] = int((av > 0).sum())
            rec["nodes"] = int(av.size)
        s.close()
        out[label] = rec
    out["cost_over_plain"] = out["artificial_viscosity"]["ms_per_step"] / out["plain"]["ms_per_step"]
    # the mixed run against the FP64 run after the same steps: relative L2 over the domain per group
    a, b = finals["artificial_viscosity_mixed"], finals["artificial_viscosity"]
    out["artificial_viscosity_mixed"]["mixed_over_fp64"] = (out["artificial_viscosity"]["ms_per_step"]
                                                            / out["artificial_viscosity_mixed"]["ms_per_step"])
    out["artificial_viscosity_mixed"]["l2_rel_dev_vs_fp64"] = max(
        float(np.sqrt((jwq[..., None] * (a[..., g] - b[..., g]) ** 2).sum() / (jwq[..., None] * b[..., g] ** 2).sum()))
        for g in ([0], [1, 2], [3]))
    log(f"shock leg: {out}")
    return out


def host_cpu():
    """The host's CPU as lscpu reports it (model, sockets, cores, threads per core) and nproc."""
    info = {"nproc": os.cpu_count()}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(n: int, order: int, target_s: float = 12.0):
    """The CPU oracle (C++/OpenMP restatement of the reference, the reference's per-element dense
    pass structure) on the same workload: the 32^3 hexahedron P3 Taylor-Green mesh, FP64, timed over
    a bounded number of steps with T = physical cores - 1 threads (the reference's rule) and with
    all cores. Physical cores from lscpu (cores per socket x sockets)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # checker library, used here only as the timed CPU baseline

    host = host_cpu()
    try:
        physical = int(host["Core(s) per socket"]) * int(host["Socket(s)"])
    except (KeyError, ValueError):
        physical = os.cpu_count() or 1
    mesh, prob, prim = tgv_setup(n, order)
    o = O.Oracle(prob, mesh)
    o.set_initial_primitive(prim)
    dt = o.compute_dt(0.1)
    runs = {}
    for label, threads in (("cores_minus_1", max(1, physical - 1)), ("all_cores", physical)):
        O.set_threads(threads)
        t0 = time.perf_counter()
        o.step(dt, 1)
        one = time.perf_counter() - t0
        steps = max(1, min(50, int(target_s / max(one, 1e-6))))
        t0 = time.perf_counter()
        o.step(dt, steps)
        t = time.perf_counter() - t0
        runs[label] = {"value": mesh.E * mesh.nb * 3 * steps / t, "threads": O.get_threads(), "steps": steps,
                       "seconds": t}
        log(f"cpu baseline {label}: {runs[label]}")
    main = runs["cores_minus_1"]
    return {"value": main["value"], "unit": "DOF.stage/s", "cores": main["threads"], "kind": "port",
            "sample": f"oracle, TGV {n}^3 hex P{order} HLLC+BR2 SSP-RK3 FP64, {main['steps']} steps, "
                      f"{main['seconds']:.1f} s, T = physical cores - 1",
            "all_cores": runs["all_cores"], "host": host}


def copy_probe(mib: int, reps: int = 5):
    """Calibration workload for the HBM byte counters: an elementwise copy of `mib` MiB (read mib,
    write mib per launch), run under `rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` beside the solver
    so that the counters' scale can be read off a kernel of known traffic."""
    import torch

    nel = mib * (1 << 20) // 8
    src = torch.rand(nel, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "reps": reps, "bytes_read": nel * 8, "bytes_written": nel * 8}))


LEGS = {"c1": c1_leg, "c5": c5_leg, "batch": batch_leg, "adjoint": adjoint_leg, "c5_adjoint": c5_adjoint_leg,
        "shock": shock_leg}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed steps per run (5 runs, median)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=32, help="elements per direction")
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--cfl", type=float, default=0.1)
    ap.add_argument("--c5-steps", type=int, default=10, help="timed steps per run of the 64^3 P4 leg")
    ap.add_argument("--c5-adjoint-steps", type=int, default=8, help="steps of the 64^3 P4 adjoint leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip", default="", help="comma-separated legs to skip: " + ",".join(LEGS))
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run the HBM-counter calibration copy of this many MiB")
    ap.add_argument("--profile-leg", choices=["headline"] + list(LEGS), default=None,
                    help="only run one leg (for kernel profiling) and print its record")
    ap.add_argument("--kernels", choices=["auto", "dense", "vector"], default="auto",
                    help="headline kernels: auto (sum-factorised on hexahedra P2-P4), the dense "
                         "matrix-core kernels or the vector kernels (FP64 only)")
    args = ap.parse_args()
    if args.kernels != "auto" and args.profile_leg != "headline":
        ap.error("--kernels other than auto only with --profile-leg headline (the bench line is the default kernels)")
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
        return
    if args.gpus != 1:
        print("bench.py: runs on one GPU only; sharding the element partition over several GPUs "
              "(face-trace halo exchange per stage) is not implemented", file=sys.stderr)
        sys.exit(2)
    if args.profile_leg and args.profile_leg != "headline":
        print(json.dumps(LEGS[args.profile_leg](args)), flush=True)
        return

    t_setup = time.perf_counter()
    mesh, prob, prim = tgv_setup(args.n, args.order)
    prob.kernels = {"auto": L.KERNELS_AUTO, "dense": L.KERNELS_DENSE, "vector": L.KERNELS_VECTOR}[args.kernels]
    rec = tgv_modes(mesh, prob, prim, args, "C2", modes=("fp64",) if args.kernels == "vector" else ("fp64", "mixed"))
    log(f"C2 runs {time.perf_counter() - t_setup:.1f}s")
    roof = {mode: roofline(rec[mode]["kms"], mesh, mode, args.n, args.kernels) for mode in rec if mode in ("fp64", "mixed")}
    if args.profile_leg == "headline":
        log(f"headline kernels {args.kernels}: roofline {roof}")
        return
    head = rec["mixed"]
    skip = set(filter(None, args.skip.split(",")))
    legs = {}
    for name, fn in LEGS.items():
        if name in skip:
            continue
        t0 = time.perf_counter()
        try:
            legs[name] = fn(args)
        except L.SDGError as exc:  # a leg the engine refuses is reported, not fatal
            legs[name] = {"error": str(exc)}
        log(f"leg {name}: {time.perf_counter() - t0:.1f}s")
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.n, args.order)
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
                   "elements": mesh.E, "basis": mesh.nb, "dof_per_field": mesh.E * mesh.nb, "precision": "mixed",
                   "parallelism": "single GPU", "runs": RUNS, "statistic": "median"},
        "ms_per_step_runs": head["ms_per_step_runs"],
        "fp64": {"value": rec["fp64"]["value"], "ms_per_step": rec["fp64"]["ms_per_step"],
                 "ms_per_step_runs": rec["fp64"]["ms_per_step_runs"], "dtype": "f64"},
        "mixed_over_fp64": head["value"] / rec["fp64"]["value"],
        "mixed_l2_rel_dev_vs_fp64": rec["mixed_l2_rel_dev_vs_fp64"],
        "mixed_l2_rel_dev_definition": "max over density, momentum, energy of ||u_mixed - u_fp64||_L2(domain) / "
                                       "||u_fp64||_L2(domain), after the warmup and timed steps",
        "mixed_coefficient_l2_rel_dev_vs_fp64": rec["mixed_coefficient_l2_rel_dev_vs_fp64"],
        "l2_dev_after_steps": rec["steps_compared"],
        "roofline": roof["mixed"],
        "roofline_fp64": roof["fp64"],
        **legs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
