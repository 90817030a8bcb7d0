"""Shared problem set-ups for the tests (synthetic meshes and initial conditions)."""
from __future__ import annotations

import math

import numpy as np

from subrosadg_amd import _lib as L

GAMMA = 1.4


def isentropic_vortex_primitive(xq: np.ndarray, eps: float = 5.0, u0=1.0, v0=1.0, rinf=1.0):
    """Shu's isentropic vortex (strength eps, centre 0) on top of the free stream
    rho = 1, u = v = 1, p = 1 with R = 1: T = 1 - (gamma-1) eps^2/(8 gamma pi^2) e^(1-r^2)."""
    x, y = xq[..., 0], xq[..., 1]
    r2 = x * x + y * y
    f = np.exp((1.0 - r2) / 2.0)
    T = 1.0 - (GAMMA - 1.0) * eps ** 2 / (8.0 * GAMMA * math.pi ** 2) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (GAMMA - 1.0))
    prim = np.zeros(xq.shape[:-1] + (4,))
    prim[..., 0] = rho
    prim[..., 1] = u0 - eps / (2.0 * math.pi) * f * y
    prim[..., 2] = v0 + eps / (2.0 * math.pi) * f * x
    prim[..., 3] = T
    return prim


def euler_vortex(n=8, order=2, flux=L.FLUX_ROE, rk=L.RK_SSPRK3):
    """The 2-D isentropic-vortex case on periodic [-5,5]^2 (R = 1: cv = 2.5, cp = 3.5)."""
    mesh = L.Mesh(L.ELEM_QUADRANGLE, order, (n, n), (-5.0, -5.0), (5.0, 5.0))
    prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=flux, time_integration=rk, cp=3.5, cv=2.5)
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, 2)
    return mesh, prob, isentropic_vortex_primitive(xq)


def tgv_primitive(xq: np.ndarray, mach=0.1, T0=1.0):
    """Taylor-Green vortex: u = sin x cos y cos z, v = -cos x sin y cos z, w = 0,
    p = p0 + (cos 2x + cos 2y)(cos 2z + 2)/16 with p0 = 1/(gamma Ma^2), rho = p/(R T0), R = p0."""
    x, y, z = xq[..., 0], xq[..., 1], xq[..., 2]
    p0 = 1.0 / (GAMMA * mach ** 2)
    R = p0
    p = p0 + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0) / 16.0
    prim = np.zeros(xq.shape[:-1] + (5,))
    prim[..., 0] = p / (R * T0)
    prim[..., 1] = np.sin(x) * np.cos(y) * np.cos(z)
    prim[..., 2] = -np.cos(x) * np.sin(y) * np.cos(z)
    prim[..., 3] = 0.0
    prim[..., 4] = T0
    return prim


def tgv_thermo(mach=0.1):
    p0 = 1.0 / (GAMMA * mach ** 2)
    R = p0
    cv = R / (GAMMA - 1.0)
    return GAMMA * cv, cv


def ns_tgv(n=4, order=3, flux=L.FLUX_HLLC, visc=L.VISC_BR2, reynolds=1600.0, mach=0.1):
    mesh = L.Mesh(L.ELEM_HEXAHEDRON, order, (n, n, n), (0.0, 0.0, 0.0), (2 * math.pi,) * 3)
    cp, cv = tgv_thermo(mach)
    prob = L.problem_desc(equation=L.EQ_NS, convective_flux=flux, viscous_flux=visc, transport=L.TRANSPORT_CONSTANT,
                          cp=cp, cv=cv, dynamic_viscosity=1.0 / reynolds)
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, 3)
    return mesh, prob, tgv_primitive(xq, mach)


def shear_layer_primitive(xq: np.ndarray, mach=0.1, perturb=False):
    """Double shear layer on [0,1]^2: rho = 1, u = tanh(80(y-1/4)) (y <= 1/2) else tanh(80(3/4-y)),
    v = 0.05 sin(2 pi x), p = p0 = 1/(gamma Ma^2), R = p0 (T = 1). Its mass and energy residuals
    vanish analytically; `perturb` adds smooth density and temperature waves so that every
    variable has a genuine residual (used where relative residual errors are compared)."""
    x, y = xq[..., 0], xq[..., 1]
    prim = np.zeros(xq.shape[:-1] + (4,))
    prim[..., 0] = 1.0
    prim[..., 1] = np.where(y <= 0.5, np.tanh(80.0 * (y - 0.25)), np.tanh(80.0 * (0.75 - y)))
    prim[..., 2] = 0.05 * np.sin(2.0 * math.pi * x)
    prim[..., 3] = 1.0
    if perturb:
        prim[..., 0] += 0.1 * np.sin(2.0 * math.pi * x) * np.cos(2.0 * math.pi * y)
        prim[..., 3] += 0.1 * np.cos(2.0 * math.pi * (x + y))
    return prim


def ns_shear(n=8, order=3, flux=L.FLUX_ROE, visc=L.VISC_BR2, reynolds=1.0e4, mach=0.1, perturb=False):
    mesh = L.Mesh(L.ELEM_QUADRANGLE, order, (n, n), (0.0, 0.0), (1.0, 1.0))
    cp, cv = tgv_thermo(mach)
    prob = L.problem_desc(equation=L.EQ_NS, convective_flux=flux, viscous_flux=visc, transport=L.TRANSPORT_CONSTANT,
                          cp=cp, cv=cv, dynamic_viscosity=1.0 / reynolds)
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, 2)
    return mesh, prob, shear_layer_primitive(xq, mach, perturb)


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max |a-b| / max |b| (relative to the reference's scale)."""
    scale = np.abs(b).max()
    return float(np.abs(a - b).max() / (scale if scale > 0 else 1.0))


def rel_err_per_var(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per-variable max |a-b| / max |b| for arrays whose last axis is the variable index."""
    a2 = a.reshape(-1, a.shape[-1])
    b2 = b.reshape(-1, b.shape[-1])
    scale = np.abs(b2).max(axis=0)
    scale = np.where(scale > 0, scale, 1.0)
    return np.abs(a2 - b2).max(axis=0) / scale


def parity_error(a: np.ndarray, b: np.ndarray, state: np.ndarray, dim: int, floor: float = 1e-3,
                 h: float | None = None, per_group: bool = False):
    """Parity metric used for the 1e-12 FP64 bar.

    Variables are grouped physically -- density, the momentum vector, total energy -- and the
    error of a group is max |a - b| over the group divided by the group's scale. The scale is the
    group's magnitude in b, measured in units of the state's magnitude of the same group (so a
    residual or gradient of rho*E is compared with rho*E, not with rho). A group whose
    nondimensional magnitude is below `floor` times the largest one (a field that is zero
    analytically, e.g. the mass residual of a uniform-density flow) is measured against that
    floor instead of its own rounding noise. a, b: (..., NV) or, for gradients, (..., NV*D)
    with variable index = column // D; state: (..., NV).

    For gradients pass the element size h: a gradient cannot be more accurate than its state
    over h, so each group's scale is at least (state magnitude) / h. This floor matters for
    nearly constant fields (e.g. grad rho with rho = 1 + 1e-3 perturbation), whose gradients come
    out of an O(|U| h^(D-1)) face/volume cancellation.
    """
    nv = dim + 2
    groups = [[0], list(range(1, dim + 1)), [dim + 1]]
    ncol = a.shape[-1]
    per = ncol // nv  # 1 for states/residuals, D for gradients
    a2 = a.reshape(-1, ncol)
    b2 = b.reshape(-1, ncol)
    s2 = state.reshape(-1, nv)
    S, nd, diff = [], [], []
    for g in groups:
        cols = [v * per + k for v in g for k in range(per)]
        sg = np.abs(s2[:, g]).max()
        sg = sg if sg > 0 else 1.0
        S.append(sg)
        nd.append(np.abs(b2[:, cols]).max() / sg)
        diff.append(np.abs(a2[:, cols] - b2[:, cols]).max())
    fl = floor * max(nd)
    if h is not None:
        fl = max(fl, 1.0 / h)
    errs = [diff[i] / (S[i] * max(nd[i], fl)) if max(nd[i], fl) > 0 else diff[i] for i in range(3)]
    if per_group:
        return [float(x) for x in errs]
    return float(max(errs))


def fp64_parity(gpu, oracle_val, extended_val, state, dim, tol, h=None):
    """gpu matches the FP64 oracle to tol, or (the oracle's own rounding floor being above tol)
    lies no farther from the extended-precision evaluation than max(tol, 2 x the oracle's own
    distance from it). Returns (error vs oracle, error vs extended, oracle's floor)."""
    e = parity_error(gpu, oracle_val, state, dim, h=h)
    ex = parity_error(gpu, extended_val, state, dim, h=h)
    floor = parity_error(oracle_val, extended_val, state, dim, h=h)
    print(f"  vs oracle {e:.2e}  vs extended {ex:.2e}  oracle floor {floor:.2e}", flush=True)
    assert e <= tol or ex <= max(tol, 2.0 * floor), (e, ex, floor)
    return e, ex, floor


def kinetic_energy(mesh, U: np.ndarray, gradient: bool = True):
    """Domain-integrated kinetic energy J = sum_e sum_q |J|w 1/2 |m|^2 / rho of modal coefficients
    U (E, N_b, NV) and its gradient dJ/dU (same shape; None with gradient=False)."""
    D = mesh.dim
    V = mesh.field("basis_value").reshape(mesh.nq, mesh.nb)
    jw = mesh.field("jacobian_det_w").reshape(mesh.E, mesh.nq)
    uq = np.matmul(V[None], U.reshape(mesh.E, mesh.nb, mesh.nv))  # (E, N_q, NV), batched products
    rho, m = uq[..., 0], uq[..., 1:1 + D]
    m2 = (m * m).sum(-1)
    J = float((jw * 0.5 * m2 / rho).sum())
    if not gradient:
        return J, None
    dj = np.zeros_like(uq)
    dj[..., 0] = -0.5 * m2 / rho ** 2
    dj[..., 1:1 + D] = m / rho[..., None]
    dj *= jw[..., None]
    dJ = np.matmul(V.T[None], dj)
    return J, dJ


def project_conserved(mesh, uq: np.ndarray) -> np.ndarray:
    """Unweighted least-squares projection of point values (E, N_q, NV) onto modal coefficients
    (InitialCondition.cpp:99-101): C = U_q V (V^T V)^-1."""
    V = mesh.field("basis_value").reshape(mesh.nq, mesh.nb)
    lsq = mesh.field("basis_lsq_inverse").reshape(mesh.nb, mesh.nb)
    return np.einsum("eqv,qc,cb->ebv", uq, V, lsq)


def euler_box(n=6, order=2, flux=L.FLUX_ROE, boundary=((L.BC_RIEMANN_FARFIELD,) * 2, (L.BC_RIEMANN_FARFIELD,) * 2),
              state=(1.0, 0.3, 0.2, 1.0)):
    """2-D Euler on [0,1]^2 with boundary types per direction (None = periodic) and a uniform
    primitive state (rho, u, v, T) as initial condition and as boundary data (R = 1)."""
    mesh = L.Mesh(L.ELEM_QUADRANGLE, order, (n, n), (0.0, 0.0), (1.0, 1.0), boundary=boundary)
    prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=flux, cp=3.5, cv=2.5)
    prim = np.tile(np.asarray(state, dtype=np.float64), (mesh.E, mesh.nq, 1))
    bprim = np.tile(np.asarray(state, dtype=np.float64), (mesh.Fb, mesh.nfq, 1))
    return mesh, prob, prim, bprim


def boundary_points(mesh):
    return mesh.field("boundary_face_coord").reshape(mesh.Fb, mesh.nfq, mesh.dim)


def smooth_direction(xq: np.ndarray, jw: np.ndarray, rng, like: np.ndarray, kmax: int = 3) -> np.ndarray:
    """Random smooth periodic field on [0,1]^2 (Gaussian weights on the Fourier modes
    sin/cos(2 pi (kx x + ky y)), 0 <= kx, ky <= kmax), scaled to the quadrature L2 norm of `like`.
    Every such direction has an O(1) component along the low modes, so directional derivatives
    stay well above finite-difference noise."""
    x, y = xq[..., 0], xq[..., 1]
    d = np.zeros_like(x)
    for kx in range(kmax + 1):
        for ky in range(kmax + 1):
            if kx == 0 and ky == 0:
                continue
            ph = 2.0 * np.pi * (kx * x + ky * y)
            a, b = rng.standard_normal(2)
            d += a * np.cos(ph) + b * np.sin(ph)
    return d * np.sqrt((jw * like ** 2).sum() / (jw * d ** 2).sum())


def function_l2(mesh, a: np.ndarray, b: np.ndarray) -> float:
    """max over variable groups (density, momentum, energy) of ||u_a - u_b||_L2 / ||u_b||_L2 over the
    domain for modal coefficients in the reference layout (E x N_b x NV); the volume quadrature
    with |J| w is exact for the squared degree-P polynomials."""
    nv, E = mesh.dim + 2, mesh.E
    V = mesh.field("basis_value").reshape(mesh.nq, mesh.nb)
    jw = mesh.field("jacobian_det_w").reshape(E, mesh.nq)
    qa = np.einsum("qb,ebv->eqv", V, (a - b).reshape(E, mesh.nb, nv))
    qb = np.einsum("qb,ebv->eqv", V, b.reshape(E, mesh.nb, nv))
    out = []
    for g in ([0], list(range(1, mesh.dim + 1)), [mesh.dim + 1]):
        out.append(np.sqrt((jw[..., None] * qa[..., g] ** 2).sum() / (jw[..., None] * qb[..., g] ** 2).sum()))
    return float(max(out))


def sod_primitive(xq: np.ndarray, x0=0.47) -> np.ndarray:
    """Sod's states (rho, u.., T) with R = 1: rho = 1, p = 1 left of x = x0 and rho = 0.125,
    p = 0.1 right of it (T = p / rho), at rest."""
    left = xq[..., 0] < x0
    prim = np.zeros(xq.shape[:-1] + (xq.shape[-1] + 2,))
    prim[..., 0] = np.where(left, 1.0, 0.125)
    prim[..., -1] = np.where(left, 1.0, 0.8)
    return prim


def euler_av(element=L.ELEM_QUADRANGLE, n=6, order=3, flux=L.FLUX_HLLC, tol=0.5, factor=1.0, boundary=None,
             split=False):
    """2-D (or 3-D for hexahedra) Euler with artificial viscosity (setArtificialViscosity(tol,
    factor)) on the unit box, Sod's jump at x = 0.47 inside an element (and the periodic image at x = 0)."""
    prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=flux, cp=3.5, cv=2.5,
                          shock_capturing=L.SHOCK_ARTIFICIAL_VISCOSITY, empirical_tolerance=tol,
                          artificial_viscosity_factor=factor)
    if split:
        mesh = L.Mesh.split(order, (n, n), (0.0, 0.0), (1.0, 1.0), 1.0, 7)
    elif element == L.ELEM_HEXAHEDRON:
        mesh = L.Mesh(element, order, (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), boundary=boundary)
    else:
        mesh = L.Mesh(element, order, (n, n), (0.0, 0.0), (1.0, 1.0), boundary=boundary)
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, mesh.dim)
    return mesh, prob, sod_primitive(xq)


def with_kernels(prob, kernels):
    """A copy of a problem descriptor that selects a kernel family (L.KERNELS_*)."""
    pd = L.ProblemDesc.from_buffer_copy(prob)
    pd.kernels = kernels
    return pd


def split_point_values(mesh, c):
    """Point values of a split-grid state (per element its N_b x 4 block in element order, each
    element's own N_b, as the batched solver and the mixed-mesh oracle store it) at the volume
    quadrature points: a list over the element types present of (|J| w, values), shapes (n_t, N_q)
    and (n_t, N_q, 4)."""
    c = np.asarray(c).reshape(-1)
    et = mesh.field("element_type")
    jw = mesh.field("jacobian_det_w")
    out = []
    refs = {t: L.Mesh.split(mesh.order, (2, 2), (0, 0), (1, 1), 1.0 if t == L.ELEM_TRIANGLE else 0.0, 0)
            for t in (L.ELEM_TRIANGLE, L.ELEM_QUADRANGLE)}
    nbs = np.array([refs[int(t)].nb for t in (L.ELEM_TRIANGLE, L.ELEM_QUADRANGLE)])
    nqs = np.array([refs[int(t)].nq for t in (L.ELEM_TRIANGLE, L.ELEM_QUADRANGLE)])
    is_quad = (et == L.ELEM_QUADRANGLE).astype(np.int64)
    coff = np.concatenate([[0], np.cumsum(nbs[is_quad] * 4)])[:-1]
    qoff = np.concatenate([[0], np.cumsum(nqs[is_quad])])[:-1]
    for k, t in enumerate((L.ELEM_TRIANGLE, L.ELEM_QUADRANGLE)):
        sel = np.nonzero(is_quad == k)[0]
        if sel.size == 0:
            continue
        r = refs[t]
        V = r.field("basis_value").reshape(r.nq, r.nb)
        C = c[coff[sel][:, None] + np.arange(r.nb * 4)[None, :]].reshape(-1, r.nb, 4)
        J = jw[qoff[sel][:, None] + np.arange(r.nq)[None, :]]
        out.append((J, np.einsum("qb,ebv->eqv", V, C)))
    return out


def split_function_l2(mesh, a, b):
    """max over density, momentum, energy of ||a - b||_L2(domain) / ||b||_L2(domain) for split-grid
    states."""
    pa, pb = split_point_values(mesh, a), split_point_values(mesh, b)
    worst = 0.0
    for g in ([0], [1, 2], [3]):
        num = sum(float((J[..., None] * (ua[..., g] - ub[..., g]) ** 2).sum()) for (J, ua), (_, ub) in zip(pa, pb))
        den = sum(float((J[..., None] * ub[..., g] ** 2).sum()) for (J, ub) in pb)
        worst = max(worst, np.sqrt(num / den))
    return worst


def split_integrals(mesh, c):
    """Domain integrals of the four conserved variables of a split-grid state."""
    return sum((J[..., None] * u).sum(axis=(0, 1)) for J, u in split_point_values(mesh, c))


def c3_specs(count=48, seed=20261202):
    """The batched uneven 2-D Euler workload: per problem n ~ U{8..96} cells per direction, P ~
    U{1..4} and a split seed, all from numpy's default generator with the given seed (the draws
    bench.py makes)."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        n = int(rng.integers(8, 97))
        order = int(rng.integers(1, 5))
        specs.append((n, order, int(rng.integers(0, 2 ** 62))))
    return specs


# ---- tetrahedra and curved elements ------------------------------------------------------------
KUHN = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def tet_box(n=2, order=2, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), periodic=True, jitter=0.0, seed=5, bc=None):
    """Tetrahedra of the Kuhn (Freudenthal) split of an n^3 grid of boxes: per box the 6 tetrahedra
    (v, v + e_a, v + e_a + e_b, v + 1) along the permutations (a, b, c) of the axes, conforming across
    boxes. jitter moves the interior nodes by up to that fraction of a box. periodic: the x, y, z
    sides are linked by node pairs (one link per direction); otherwise the boundary faces get the
    types bc[side] (sides x-, x+, y-, y+, z-, z+; physical tag = side + 1)."""
    n1 = n + 1
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    h = (hi - lo) / n
    idx = np.arange(n1 ** 3).reshape(n1, n1, n1)  # [i, j, k]
    ijk = np.stack(np.meshgrid(np.arange(n1), np.arange(n1), np.arange(n1), indexing="ij"), -1).reshape(-1, 3)
    xyz = lo + ijk * h
    if jitter:
        rng = np.random.default_rng(seed)
        inner = np.all((ijk > 0) & (ijk < n), axis=1)
        xyz[inner] += jitter * h * rng.uniform(-1, 1, (inner.sum(), 3))
    tets = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for perm in KUHN:
                    p = np.array([i, j, k])
                    vs = [idx[tuple(p)]]
                    for a in perm:
                        p = p.copy()
                        p[a] += 1
                        vs.append(idx[tuple(p)])
                    tets.append(vs)
    tets = np.array(tets, dtype=np.int64)
    kw = {}
    if periodic:
        links = []
        for d in range(3):
            sl = np.nonzero(ijk[:, d] == n)[0]
            m = ijk[sl].copy()
            m[:, d] = 0
            links.append([(int(s), int(idx[tuple(mm)])) for s, mm in zip(sl, m)])
        kw["periodic"] = links
    else:
        faces = {}
        for t in tets:
            for f in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
                key = tuple(sorted(t[list(f)]))
                faces[key] = faces.get(key, 0) + 1
        bf, bt, bp = [], [], []
        for key, cnt in faces.items():
            if cnt != 1:
                continue
            c = ijk[list(key)]
            for side in range(6):
                d, v = side // 2, 0 if side % 2 == 0 else n
                if np.all(c[:, d] == v):
                    bf.append(key)
                    bt.append((bc or [L.BC_RIEMANN_FARFIELD] * 6)[side])
                    bp.append(side + 1)
                    break
        kw.update(boundary_faces=np.array(bf, dtype=np.int64), boundary_type=bt, boundary_physical=bp)
    return L.Mesh.from_arrays(L.ELEM_TETRAHEDRON, order, xyz, tets, **kw)


def curved_quad_annulus(n=(6, 3), order=3, geometry_order=2, r=(1.0, 2.0), angle=(0.0, np.pi / 2),
                        bc=(L.BC_RIEMANN_FARFIELD, L.BC_ADIABATIC_SLIP_WALL)):
    """Quadrangles of geometry order g on a sector of an annulus: n[0] cells along the angle, n[1]
    along the radius; every geometry node on its exact polar position, so the cells are curved.
    Boundary sides: the radial lines get bc[0] (tag 1), the arcs bc[1] (tag 2). Node order per cell
    as gmsh's: vertices, edge interiors along (0,1), (1,2), (2,3), (3,0), then the interior (as a
    cell of order g - 2)."""
    na, nr = n
    g = geometry_order
    # global lattice of (na g + 1) x (nr g + 1) points
    ta = np.linspace(angle[0], angle[1], na * g + 1)
    tr = np.linspace(r[0], r[1], nr * g + 1)
    A, R = np.meshgrid(ta, tr, indexing="ij")
    xyz = np.stack([R * np.cos(A), R * np.sin(A), 0 * A], -1).reshape(-1, 3)
    gid = np.arange((na * g + 1) * (nr * g + 1)).reshape(na * g + 1, nr * g + 1)
    lat = quad_lattice(g)
    cells = []
    for i in range(na):
        for j in range(nr):  # first lattice index radial, second angular: counter-clockwise cells
            cells.append([gid[i * g + b, j * g + a] for a, b in lat])
    bf, bt, bp = [], [], []
    for j in range(nr):  # radial sides
        for i in (0, na * g):
            bf.append((gid[i, j * g], gid[i, (j + 1) * g]))
            bt.append(bc[0])
            bp.append(1)
    for i in range(na):  # arcs
        for j in (0, nr * g):
            bf.append((gid[i * g, j], gid[(i + 1) * g, j]))
            bt.append(bc[1])
            bp.append(2)
    return L.Mesh.from_arrays(L.ELEM_QUADRANGLE, order, xyz, np.array(cells, dtype=np.int64), bf, bt, bp,
                              geometry_order=g)


def quad_lattice(g):
    """gmsh's node order of a quadrangle of order g as lattice indices (a, b), 0..g."""
    if g == 0:
        return [(0, 0)]
    out = [(0, 0), (g, 0), (g, g), (0, g)]
    out += [(k, 0) for k in range(1, g)] + [(g, k) for k in range(1, g)]
    out += [(g - k, g) for k in range(1, g)] + [(0, g - k) for k in range(1, g)]
    if g >= 2:
        out += [(a + 1, b + 1) for a, b in quad_lattice(g - 2)]
    return out


def tri_lattice(g):
    """gmsh's node order of a triangle of order g as lattice indices (a, b), a + b <= g."""
    if g < 0:
        return []
    if g == 0:
        return [(0, 0)]
    out = [(0, 0), (g, 0), (0, g)]
    out += [(k, 0) for k in range(1, g)] + [(g - k, k) for k in range(1, g)] + [(0, g - k) for k in range(1, g)]
    out += [(a + 1, b + 1) for a, b in tri_lattice(g - 3)]
    return out


def lattice_mesh(element, n, g, order=2, warp=None, lo=0.0, hi=1.0, periodic=False, bc=L.BC_RIEMANN_FARFIELD):
    """Elements of geometry order g on a structured grid of the unit square / cube: quadrangles
    (element QUADRANGLE), pairs of triangles per cell (TRIANGLE), hexahedra (HEXAHEDRON) or Kuhn
    tetrahedra per cell (TETRAHEDRON, g <= 2); geometry nodes in gmsh order on a lattice of
    n g + 1 points per direction, moved by warp(xyz) -> xyz when given (curved cells). Bounded:
    every boundary face gets bc with physical tag side + 1; periodic: one link per direction."""
    dim = 3 if element in (L.ELEM_HEXAHEDRON, L.ELEM_TETRAHEDRON) else 2
    m1 = n * g + 1
    grid = np.stack(np.meshgrid(*([np.arange(m1)] * dim), indexing="ij"), -1).reshape(-1, dim)
    gid = np.arange(m1 ** dim).reshape((m1,) * dim)
    xyz = np.zeros((len(grid), 3))
    xyz[:, :dim] = lo + (hi - lo) * grid / (n * g)
    if warp is not None:
        xyz = warp(xyz)
    cells = []
    hexlat = None
    if element == L.ELEM_HEXAHEDRON:
        if g == 1:
            hexlat = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
        else:
            v = [(0, 0, 0), (2, 0, 0), (2, 2, 0), (0, 2, 0), (0, 0, 2), (2, 0, 2), (2, 2, 2), (0, 2, 2)]
            edges = [(0, 1), (0, 3), (0, 4), (1, 2), (1, 5), (2, 3), (2, 6), (3, 7), (4, 5), (4, 7), (5, 6), (6, 7)]
            faces = [(0, 3, 2, 1), (0, 1, 5, 4), (0, 4, 7, 3), (1, 2, 6, 5), (2, 3, 7, 6), (4, 5, 6, 7)]
            hexlat = v + [tuple((np.add(v[a], v[b]) // 2)) for a, b in edges]
            hexlat += [tuple(np.sum([v[k] for k in f], 0) // 4) for f in faces] + [(1, 1, 1)]
    for c in np.ndindex(*([n] * dim)):
        base = np.array(c) * g
        if element == L.ELEM_QUADRANGLE:
            cells.append([gid[tuple(base + (a, b))] for a, b in quad_lattice(g)])
        elif element == L.ELEM_TRIANGLE:
            # (A, B, C) and (A, C, D) of the cell corners, nodes of each on the cell's lattice
            for tri in (((0, 0), (g, 0), (g, g)), ((0, 0), (g, g), (0, g))):
                p0, p1, p2 = (np.array(t) for t in tri)
                cells.append([gid[tuple(base + p0 + ((p1 - p0) * a + (p2 - p0) * b) // g)] for a, b in tri_lattice(g)])
        elif element == L.ELEM_HEXAHEDRON:
            cells.append([gid[tuple(base + np.array(h) * g // (2 if g == 2 else 1))] for h in hexlat])
        else:
            for perm in KUHN:
                path = [np.zeros(3, int)]
                for a in perm:
                    q = path[-1].copy()
                    q[a] += g
                    path.append(q)
                vs = [gid[tuple(base + p)] for p in path]
                if g == 2:
                    for a, b in ((0, 1), (1, 2), (2, 0), (3, 0), (3, 2), (3, 1)):
                        vs.append(gid[tuple(base + (path[a] + path[b]) // 2)])
                cells.append(vs)
    cells = np.array(cells, dtype=np.int64)
    nv = {L.ELEM_QUADRANGLE: 4, L.ELEM_TRIANGLE: 3, L.ELEM_HEXAHEDRON: 8, L.ELEM_TETRAHEDRON: 4}[element]
    nfv = {L.ELEM_QUADRANGLE: 2, L.ELEM_TRIANGLE: 2, L.ELEM_HEXAHEDRON: 4, L.ELEM_TETRAHEDRON: 3}[element]
    kw = {"geometry_order": g}
    if periodic:
        links = []
        for d in range(dim):
            sl = np.nonzero(grid[:, d] == n * g)[0]
            mm = grid[sl].copy()
            mm[:, d] = 0
            links.append([(int(s), int(gid[tuple(x)])) for s, x in zip(sl, mm)])
        kw["periodic"] = links
    else:
        fvl = {L.ELEM_QUADRANGLE: [(0, 1), (1, 2), (2, 3), (3, 0)], L.ELEM_TRIANGLE: [(0, 1), (1, 2), (2, 0)],
               L.ELEM_HEXAHEDRON: [(0, 3, 2, 1), (0, 1, 5, 4), (0, 4, 7, 3), (1, 2, 6, 5), (2, 3, 7, 6), (4, 5, 6, 7)],
               L.ELEM_TETRAHEDRON: [(0, 2, 1), (0, 1, 3), (0, 3, 2), (3, 1, 2)]}[element]
        count = {}
        for c in cells:
            for f in fvl:
                key = tuple(sorted(c[list(f)]))
                count[key] = count.get(key, 0) + 1
        bf, bp = [], []
        for key, cnt in count.items():
            if cnt != 1:
                continue
            gc = grid[list(key)]
            for side in range(2 * dim):
                d, v = side // 2, 0 if side % 2 == 0 else n * g
                if np.all(gc[:, d] == v):
                    bf.append(key)
                    bp.append(side + 1)
                    break
        kw.update(boundary_faces=np.array(bf, dtype=np.int64).reshape(-1, nfv), boundary_type=[bc] * len(bf),
                  boundary_physical=bp)
    del nv
    return L.Mesh.from_arrays(element, order, xyz, cells, **kw), cells, xyz


def bump(amplitude=0.06):
    """A smooth warp of the unit square / cube that keeps its sides in place."""
    def warp(xyz):
        x = xyz.copy()
        s = np.sin(np.pi * xyz[:, 0]) * np.sin(np.pi * xyz[:, 1])
        x[:, 0] += amplitude * s * np.cos(0.7 * xyz[:, 1])
        x[:, 1] += amplitude * s * np.sin(1.3 * xyz[:, 0] + 0.2)
        if np.any(xyz[:, 2] != 0):
            t = s * np.sin(np.pi * xyz[:, 2])
            x[:, 0] = xyz[:, 0] + amplitude * t
            x[:, 1] = xyz[:, 1] + amplitude * t * np.cos(xyz[:, 0])
            x[:, 2] = xyz[:, 2] + amplitude * t * np.sin(2.0 * xyz[:, 1])
        return x
    return warp


def pyramid_box(n=2, order=2, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), periodic=True, jitter=0.0, seed=3, bc=None):
    """Pyramids: every box of an n^3 grid split into the six pyramids on its faces with a common apex
    near the box centre (moved by up to jitter of a box; the bases stay parallelograms, so the
    elements stay affine). Quadrangle bases pair across boxes, triangles inside a box. periodic:
    one link per direction; otherwise the box sides get bc[side] (tag side + 1)."""
    n1 = n + 1
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    h = (hi - lo) / n
    ijk = np.stack(np.meshgrid(np.arange(n1), np.arange(n1), np.arange(n1), indexing="ij"), -1).reshape(-1, 3)
    idx = np.arange(n1 ** 3).reshape(n1, n1, n1)
    xyz = [lo + p * h for p in ijk]
    rng = np.random.default_rng(seed)
    # hexahedron faces (gmsh order, outward); the pyramid base is the face reversed (inward normal,
    # i.e. counter-clockwise seen from the apex)
    corner = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    faces = [(0, 3, 2, 1), (0, 1, 5, 4), (0, 4, 7, 3), (1, 2, 6, 5), (2, 3, 7, 6), (4, 5, 6, 7)]
    pyrs, sides = [], []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                c = len(xyz)
                xyz.append(lo + (np.array([i, j, k]) + 0.5 + jitter * rng.uniform(-1, 1, 3)) * h)
                verts = [idx[i + a, j + b, k + cc] for a, b, cc in corner]
                for f in faces:
                    base = [verts[f[0]], verts[f[3]], verts[f[2]], verts[f[1]]]
                    pyrs.append(base + [c])
    xyz = np.array(xyz)
    pyrs = np.array(pyrs, dtype=np.int64)
    kw = {}
    if periodic:
        links = []
        for d in range(3):
            sl = np.nonzero(ijk[:, d] == n)[0]
            m = ijk[sl].copy()
            m[:, d] = 0
            links.append([(int(s), int(idx[tuple(mm)])) for s, mm in zip(sl, m)])
        kw["periodic"] = links
    else:
        bf, bt, bp = [], [], []
        for p in pyrs:
            g = ijk[p[:4]]
            for side in range(6):
                d, v = side // 2, 0 if side % 2 == 0 else n
                if np.all(g[:, d] == v):
                    bf.append(p[:4])
                    bt.append((bc or [L.BC_RIEMANN_FARFIELD] * 6)[side])
                    bp.append(side + 1)
        kw.update(boundary_faces=np.array(bf, dtype=np.int64), boundary_type=bt, boundary_physical=bp)
    return L.Mesh.from_arrays(L.ELEM_PYRAMID, order, xyz, pyrs, **kw)
