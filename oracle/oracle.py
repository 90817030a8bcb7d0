"""ctypes wrapper of the CPU ORACLE (oracle/liboracle.so) -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, as the
checker (or the timed CPU baseline), never as part of the product path. The oracle consumes the
same operand tables (basis, geometry, faces) that the product library exports for a mesh, so a
parity test compares the two implementations of the hot path on identical inputs.

`Oracle(..., extended=True)` runs the same source built with x87 extended precision
(liboracle_ld.so, 64-bit mantissa) on the same FP64 tables and states. Its distance from the
FP64 oracle is the FP64 rounding floor of a case: a Navier-Stokes residual that is the small
difference of large volume and face terms (e.g. the mass residual of a divergence-free
Taylor-Green state) cannot be matched between two FP64 evaluations more closely than that floor.

`Oracle(..., tangent=True)` runs the same source built with a forward-mode number type
(liboracle_tl.so): each value carries a directional derivative, so `residual_tangent(U, v)` gives
R(U) and the tangent-linear J v = (dR/dU) v, and `step_tangent` the same through whole RK steps.
It is the CPU side of the adjoint's dot-product test <J v, w> = <v, J^T w>.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_int64, c_longdouble, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
LIB_LD_PATH = os.path.join(HERE, "liboracle_ld.so")
LIB_TL_PATH = os.path.join(HERE, "liboracle_tl.so")
SRC_PATH = os.path.join(HERE, "oracle.cpp")


def build(force: bool = False) -> str:
    """Compile oracle.cpp into liboracle.so (FP64), liboracle_ld.so (x87 extended) and
    liboracle_tl.so (FP64 tangent-linear) with g++."""
    for path, extra in ((LIB_PATH, []), (LIB_LD_PATH, ["-DORACLE_REAL=long double"]),
                        (LIB_TL_PATH, ["-DORACLE_TANGENT"])):
        if force or not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(SRC_PATH):
            cmd = ["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", *extra,
                   "-o", path, SRC_PATH]
            subprocess.run(cmd, check=True)
    return LIB_PATH


_libs = {}


class _Tangent(ctypes.Structure):
    """A scalar of the tangent-linear build passed by value: (value, derivative)."""
    _fields_ = [("v", c_double), ("d", c_double)]


def load(extended: bool = False, tangent: bool = False):
    key = "tangent" if tangent else extended
    lib = _libs.get(key)
    if lib is None:
        path = LIB_TL_PATH if tangent else (LIB_LD_PATH if extended else LIB_PATH)
        if not os.path.exists(path):
            build()
        real = _Tangent if tangent else (c_longdouble if extended else c_double)
        lib = ctypes.CDLL(path)
        lib.oracle_create.restype = c_void_p
        lib.oracle_create.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int64, c_int64] + \
            [c_void_p] * 13
        lib.oracle_destroy.argtypes = [c_void_p]
        for name in ("oracle_set_state", "oracle_get_state", "oracle_set_initial_primitive", "oracle_residual",
                     "oracle_get_gradient", "oracle_get_volume_gradient", "oracle_get_artificial_viscosity"):
            getattr(lib, name).argtypes = [c_void_p, c_void_p]
            getattr(lib, name).restype = c_int
        lib.oracle_get_face_gradient.argtypes = [c_void_p, c_void_p]
        lib.oracle_get_face_gradient.restype = c_int
        lib.oracle_get_face_gradient_coefficients.argtypes = [c_void_p, c_void_p]
        lib.oracle_get_face_gradient_coefficients.restype = c_int
        lib.oracle_monitor.argtypes = [c_void_p, c_void_p, c_void_p]
        lib.oracle_monitor.restype = c_int
        lib.oracle_compute_dt.argtypes = [c_void_p, real, POINTER(real)]
        lib.oracle_step.argtypes = [c_void_p, real, c_int, c_void_p, POINTER(c_int), POINTER(c_int)]
        lib.oracle_point_flux.argtypes = [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]
        lib.oracle_point_viscous_flux.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]
        lib.oracle_point_boundary_state.argtypes = [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]
        lib.oracle_set_boundary.argtypes = [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]
        lib.oracle_set_boundary_primitive.argtypes = [c_void_p, c_void_p]
        lib.oracle_set_artificial_viscosity.argtypes = [c_void_p, c_int, c_int64, c_void_p, c_void_p, c_void_p,
                                                        c_void_p, real, real]
        lib.oracle_set_threads.argtypes = [c_int]
        lib.oracle_get_threads.restype = c_int
        _libs[key] = lib
    return lib


def _ptr(a: np.ndarray):
    return a.ctypes.data


class Oracle:
    """CPU restatement of Solver<SC> on one mesh (extended=True: the x87 extended build; arrays
    still enter and leave as float64, converted exactly on the way in)."""

    _TABLES = ("basis_value", "basis_gradient", "basis_face_value", "basis_lsq_inverse", "jacobian_det_w",
               "jacobian_tiw", "mass_inverse", "min_edge", "face_normal", "face_jacobian_w",
               "boundary_face_normal", "boundary_face_jacobian_w")

    def __init__(self, problem, mesh, extended: bool = False, tangent: bool = False):
        self.extended = extended
        self.tangent = tangent
        self.real = np.complex128 if tangent else (np.longdouble if extended else np.float64)
        self.lib = load(extended, tangent)
        self.mesh = mesh
        self.dim, self.nv, self.nb, self.nq = mesh.dim, mesh.nv, mesh.nb, mesh.nq
        self.E = mesh.E
        self._keep = {k: mesh.field(k) for k in (
            "basis_value", "basis_gradient", "basis_face_value", "basis_lsq_inverse", "jacobian_det_w",
            "jacobian_tiw", "mass_inverse", "min_edge", "face_parent", "face_local", "face_sequence",
            "face_normal", "face_jacobian_w")}
        k = self._keep
        for key in self._TABLES:
            if key in k:
                k[key] = np.ascontiguousarray(k[key], dtype=self.real)
        self.problem = problem
        self.h = self.lib.oracle_create(
            ctypes.byref(problem), mesh.dim, mesh.nb, mesh.nq, mesh.nf, mesh.nfq, mesh.order, mesh.E, mesh.F,
            _ptr(k["basis_value"]), _ptr(k["basis_gradient"]), _ptr(k["basis_face_value"]),
            _ptr(k["basis_lsq_inverse"]), _ptr(k["jacobian_det_w"]), _ptr(k["jacobian_tiw"]),
            _ptr(k["mass_inverse"]), _ptr(k["min_edge"]), _ptr(k["face_parent"]), _ptr(k["face_local"]),
            _ptr(k["face_sequence"]), _ptr(k["face_normal"]), _ptr(k["face_jacobian_w"]))
        if not self.h:
            raise RuntimeError("oracle_create failed")
        self.Fb = getattr(mesh, "Fb", 0)
        if self.Fb:
            for key in ("boundary_face_parent", "boundary_face_type", "boundary_face_normal",
                        "boundary_face_jacobian_w"):
                k[key] = mesh.field(key)
                if key in self._TABLES:
                    k[key] = np.ascontiguousarray(k[key], dtype=self.real)
            self.lib.oracle_set_boundary(self.h, self.Fb, _ptr(k["boundary_face_parent"]),
                                         _ptr(k["boundary_face_type"]), _ptr(k["boundary_face_normal"]),
                                         _ptr(k["boundary_face_jacobian_w"]))
        self.av = getattr(problem, "shock_capturing", 0) == 1
        if self.av:
            for key in ("node_tag", "inner_radius", "nodal_value", "nodal_face_value"):
                k[key] = mesh.field(key)
                if k[key].dtype == np.float64:
                    k[key] = np.ascontiguousarray(k[key], dtype=self.real)
            self.nodes = mesh.nodes
            self.lib.oracle_set_artificial_viscosity(
                self.h, mesh.nvert, mesh.nodes, _ptr(k["node_tag"]), _ptr(k["inner_radius"]),
                _ptr(k["nodal_value"]), _ptr(k["nodal_face_value"]), problem.empirical_tolerance,
                problem.artificial_viscosity_factor)

    def set_boundary_primitive(self, prim: np.ndarray):
        prim = np.ascontiguousarray(prim, dtype=self.real)
        assert prim.size == self.Fb * self.mesh.nfq * self.nv
        self.lib.oracle_set_boundary_primitive(self.h, _ptr(prim))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.oracle_destroy(self.h)
            self.h = None

    def state_shape(self):
        return (self.E, self.nb, self.nv)

    def set_state(self, c: np.ndarray):
        c = np.ascontiguousarray(c, dtype=self.real)
        self.lib.oracle_set_state(self.h, _ptr(c))

    def _out(self, shape) -> np.ndarray:
        return np.empty(shape, dtype=self.real)

    def get_state(self) -> np.ndarray:
        out = self._out(self.state_shape())
        self.lib.oracle_get_state(self.h, _ptr(out))
        return out.astype(np.float64)

    def set_initial_primitive(self, prim: np.ndarray):
        prim = np.ascontiguousarray(prim, dtype=self.real)
        self.lib.oracle_set_initial_primitive(self.h, _ptr(prim))

    def residual(self) -> np.ndarray:
        out = self._out(self.state_shape())
        self.lib.oracle_residual(self.h, _ptr(out))
        return out.astype(np.float64)

    def gradient(self) -> np.ndarray:
        out = self._out((self.E, self.nb, self.nv * self.dim))
        rc = self.lib.oracle_get_gradient(self.h, _ptr(out))
        if rc != 0:
            raise RuntimeError("gradient is only defined for Navier-Stokes")
        return out.astype(np.float64)

    def face_gradient(self) -> np.ndarray:
        """Gradient traces each element uses at its own face points in the viscous face flux of the
        last residual (BR2: volume gradient plus that face's lift), (E, N_f, N_fq, NV D)."""
        out = self._out((self.E, self.mesh.nf, self.mesh.nfq, self.nv * self.dim))
        if self.lib.oracle_get_face_gradient(self.h, _ptr(out)) != 0:
            raise RuntimeError("face gradients are only defined for Navier-Stokes")
        return out.astype(np.float64)

    def face_gradient_coefficients(self) -> np.ndarray:
        """Per-face gradient coefficients of the last residual, (E, N_f, N_b, NV D): BR2 the volume
        gradient plus that face's lift, BR1 the total gradient."""
        out = self._out((self.E, self.mesh.nf, self.nb, self.nv * self.dim))
        if self.lib.oracle_get_face_gradient_coefficients(self.h, _ptr(out)) != 0:
            raise RuntimeError("face gradients are only defined for Navier-Stokes")
        return out.astype(np.float64)

    def monitor(self, residual: np.ndarray) -> np.ndarray:
        """Residual monitor of a given residual (E, N_b, NV); replaces the oracle's residual."""
        r = np.ascontiguousarray(residual, dtype=self.real)
        assert r.shape == self.state_shape()
        out = self._out(self.nv)
        self.lib.oracle_monitor(self.h, _ptr(r), _ptr(out))
        return out.astype(np.float64)

    def volume_gradient(self) -> np.ndarray:
        """Volume-gradient coefficients of the last residual or stage, (E, N_b, NV D)."""
        out = self._out((self.E, self.nb, self.nv * self.dim))
        self.lib.oracle_get_volume_gradient(self.h, _ptr(out))
        return out.astype(np.float64)

    def artificial_viscosity(self) -> np.ndarray:
        """Node values of the last artificial-viscosity evaluation."""
        out = self._out(self.nodes)
        self.lib.oracle_get_artificial_viscosity(self.h, _ptr(out))
        return out.astype(np.float64)

    def compute_dt(self, cfl: float) -> float:
        dt = (c_longdouble if self.extended else c_double)()
        self.lib.oracle_compute_dt(self.h, cfl, ctypes.byref(dt))
        return float(dt.value)

    def residual_tangent(self, u: np.ndarray, v: np.ndarray):
        """(R(u), J v) with J = dR/du at u (tangent build only)."""
        assert self.tangent
        self.set_state(np.asarray(u, dtype=np.float64) + 1j * np.asarray(v, dtype=np.float64))
        out = self._out(self.state_shape())
        self.lib.oracle_residual(self.h, _ptr(out))
        return out.real.copy(), out.imag.copy()

    def step_tangent(self, u: np.ndarray, v: np.ndarray, dt: float, nsteps: int):
        """(U_n, dU_n/du v) after nsteps steps of the fixed size dt from u (tangent build only)."""
        assert self.tangent
        self.set_state(np.asarray(u, dtype=np.float64) + 1j * np.asarray(v, dtype=np.float64))
        rel = self._out(self.nv)
        nan, taken = c_int(), c_int()
        self.lib.oracle_step(self.h, _Tangent(dt, 0.0), nsteps, _ptr(rel), ctypes.byref(nan), ctypes.byref(taken))
        out = self._out(self.state_shape())
        self.lib.oracle_get_state(self.h, _ptr(out))
        return out.real.copy(), out.imag.copy()

    def step(self, dt: float, nsteps: int = 1):
        rel = self._out(self.nv)
        nan = c_int()
        taken = c_int()
        self.lib.oracle_step(self.h, dt, nsteps, _ptr(rel), ctypes.byref(nan), ctypes.byref(taken))
        return rel.astype(np.float64), nan.value, taken.value


def set_threads(n: int):
    load().oracle_set_threads(n)


def get_threads() -> int:
    return load().oracle_get_threads()


def point_flux(problem, dim: int, flux: int, n, ul, ur) -> np.ndarray:
    lib = load()
    n = np.ascontiguousarray(n, dtype=np.float64)
    ul = np.ascontiguousarray(ul, dtype=np.float64)
    ur = np.ascontiguousarray(ur, dtype=np.float64)
    h = np.empty(dim + 2)
    lib.oracle_point_flux(ctypes.byref(problem), dim, flux, _ptr(n), _ptr(ul), _ptr(ur), _ptr(h))
    return h


def point_viscous_flux(problem, dim: int, n, u, g) -> np.ndarray:
    lib = load()
    n = np.ascontiguousarray(n, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    f = np.empty(dim + 2)
    lib.oracle_point_viscous_flux(ctypes.byref(problem), dim, _ptr(n), _ptr(u), _ptr(g), _ptr(f))
    return f


def point_boundary_state(problem, dim: int, bc: int, n, ul, prim) -> np.ndarray:
    """Boundary state (rho, u.., e, p) of boundary type bc from the conserved interior state ul and
    the primitive boundary data prim."""
    lib = load()
    n = np.ascontiguousarray(n, dtype=np.float64)
    ul = np.ascontiguousarray(ul, dtype=np.float64)
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    wb = np.empty(dim + 3)
    lib.oracle_point_boundary_state(ctypes.byref(problem), dim, bc, _ptr(n), _ptr(ul), _ptr(prim), _ptr(wb))
    return wb


class MixedOracle:
    """Euler oracle on a split grid of triangles and quadrangles (the reference's triangle_ and
    quadrangle_ element solvers in one Solver). State layout: per element its N_b x NV block, in
    element order (each element's own N_b), as the batched solver returns a problem's state."""

    def __init__(self, problem, mesh):
        from subrosadg_amd import _lib as L

        self.lib = load()
        lib = self.lib
        if not hasattr(lib, "_mixed_ready"):
            lib.oracle_mixed_create.restype = c_void_p
            lib.oracle_mixed_create.argtypes = [c_void_p, c_int, c_int64, c_int64, c_void_p, c_void_p, c_void_p,
                                                c_void_p, c_int] + [c_void_p] * 13
            lib.oracle_mixed_destroy.argtypes = [c_void_p]
            for name in ("oracle_mixed_set_state", "oracle_mixed_get_state", "oracle_mixed_set_initial_primitive",
                         "oracle_mixed_residual"):
                getattr(lib, name).argtypes = [c_void_p, c_void_p]
            lib.oracle_mixed_compute_dt.argtypes = [c_void_p, c_double, POINTER(c_double)]
            lib.oracle_mixed_step.argtypes = [c_void_p, c_double, c_int, c_void_p]
            lib._mixed_ready = True
        self.mesh = mesh
        order = mesh.order
        refs = [L.Mesh.split(order, (2, 2), (0.0, 0.0), (1.0, 1.0), frac, 0) for frac in (1.0, 0.0)]
        et = mesh.field("element_type")
        self.etype = np.ascontiguousarray((et == L.ELEM_QUADRANGLE).astype(np.int8))
        self.nb = np.array([r.nb for r in refs], dtype=np.int32)
        nq = np.array([r.nq for r in refs], dtype=np.int32)
        nf = np.array([r.nf for r in refs], dtype=np.int32)
        nfq = refs[0].nfq
        keep = {}
        tabs = {}
        for key in ("basis_value", "basis_gradient", "basis_face_value", "basis_lsq_inverse"):
            arrs = [np.ascontiguousarray(r.field(key)) for r in refs]
            keep[key] = arrs
            tabs[key] = (c_void_p * 2)(*[a.ctypes.data for a in arrs])
        for key in ("jacobian_det_w", "jacobian_tiw", "mass_inverse", "min_edge", "face_parent", "face_local",
                    "face_sequence", "face_normal", "face_jacobian_w"):
            keep[key] = mesh.field(key)
        self._keep = (keep, tabs, nq, nf, refs)
        self.E = mesh.E
        self.F = keep["face_parent"].size // 2
        self.size = int(self.nb[self.etype].sum()) * 4
        k = keep
        self.h = lib.oracle_mixed_create(
            ctypes.byref(problem), order, self.E, self.F, _ptr(self.etype), _ptr(self.nb), _ptr(nq), _ptr(nf), nfq,
            ctypes.cast(tabs["basis_value"], c_void_p), ctypes.cast(tabs["basis_gradient"], c_void_p),
            ctypes.cast(tabs["basis_face_value"], c_void_p), ctypes.cast(tabs["basis_lsq_inverse"], c_void_p),
            _ptr(k["jacobian_det_w"]), _ptr(k["jacobian_tiw"]), _ptr(k["mass_inverse"]), _ptr(k["min_edge"]),
            _ptr(k["face_parent"]), _ptr(k["face_local"]), _ptr(k["face_sequence"]), _ptr(k["face_normal"]),
            _ptr(k["face_jacobian_w"]))
        if not self.h:
            raise RuntimeError("oracle_mixed_create failed (Euler only)")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.oracle_mixed_destroy(self.h)
            self.h = None

    def set_state(self, c):
        c = np.ascontiguousarray(c, dtype=np.float64)
        assert c.size == self.size
        self.lib.oracle_mixed_set_state(self.h, _ptr(c))

    def get_state(self):
        out = np.empty(self.size)
        self.lib.oracle_mixed_get_state(self.h, _ptr(out))
        return out

    def set_initial_primitive(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        self.lib.oracle_mixed_set_initial_primitive(self.h, _ptr(prim))

    def residual(self):
        out = np.empty(self.size)
        self.lib.oracle_mixed_residual(self.h, _ptr(out))
        return out

    def compute_dt(self, cfl):
        dt = c_double()
        self.lib.oracle_mixed_compute_dt(self.h, cfl, ctypes.byref(dt))
        return dt.value

    def step(self, dt, nsteps=1):
        rel = np.empty(4)
        self.lib.oracle_mixed_step(self.h, dt, nsteps, _ptr(rel))
        return rel
