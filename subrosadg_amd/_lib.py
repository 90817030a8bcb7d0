"""ctypes binding of libsdg_hip.so (the C-ABI declared in include/sdg.h).

The library is built in-tree (see subrosadg_amd/build.py). Loading fails loudly when it is
missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_char_p, c_double, c_int, c_int32, c_int64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsdg_hip.so")

ABI_VERSION = 5

# status codes
OK, E_INVALID, E_UNSUPPORTED, E_DEVICE, E_NOMEM, E_STATE = 0, -1, -2, -3, -4, -5

# enums (include/sdg.h)
ELEM_LINE, ELEM_TRIANGLE, ELEM_QUADRANGLE, ELEM_TETRAHEDRON, ELEM_PYRAMID, ELEM_HEXAHEDRON = 1, 2, 3, 4, 5, 6
ELEM_MIXED_TRIANGLE_QUADRANGLE = 100
EQ_EULER, EQ_NS = 0, 1
FLUX_CENTRAL, FLUX_LAX_FRIEDRICHS, FLUX_HLLC, FLUX_ROE, FLUX_EXACT = 0, 1, 2, 3, 4
VISC_NONE, VISC_BR1, VISC_BR2 = 0, 1, 2
RK_FORWARD_EULER, RK_HEUN2, RK_SSPRK3 = 0, 1, 2
EOS_IDEAL_GAS, EOS_WEAK_COMPRESSIBLE = 0, 1
TRANSPORT_NONE, TRANSPORT_CONSTANT, TRANSPORT_SUTHERLAND = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
SOURCE_NONE, SOURCE_BOUSSINESQ = 0, 1
SHOCK_NONE, SHOCK_ARTIFICIAL_VISCOSITY = 0, 1
VIEW_VARIABLES = ("Density", "Velocity", "Temperature", "Pressure", "SoundSpeed", "MachNumber", "Entropy", "Vorticity",
                  "HeatFlux", "ArtificialViscosity", "VelocityX", "VelocityY", "VelocityZ", "MachNumberX",
                  "MachNumberY", "MachNumberZ", "VorticityX", "VorticityY", "VorticityZ", "HeatFluxX", "HeatFluxY",
                  "HeatFluxZ")  # enum sdg_view_variable, in order
KERNELS_AUTO, KERNELS_DENSE, KERNELS_VECTOR = 0, 1, 2
LAUNCH_GRAPH, LAUNCH_EAGER = 0, 1
IC_FUNCTION, IC_SPECIFIC_FILE, IC_LAST_STEP = 0, 1, 2
(BC_RIEMANN_FARFIELD, BC_VELOCITY_INFLOW, BC_PRESSURE_OUTFLOW, BC_ISOTHERMAL_NONSLIP_WALL, BC_ADIABATIC_SLIP_WALL,
 BC_ADIABATIC_NONSLIP_WALL, BC_PERIODIC) = range(7)

FIELDS = {
    "basis_value": (0, np.float64),
    "basis_gradient": (1, np.float64),
    "basis_face_value": (2, np.float64),
    "basis_lsq_inverse": (3, np.float64),
    "quad_weight": (4, np.float64),
    "face_quad_weight": (5, np.float64),
    "quad_coord": (6, np.float64),
    "jacobian_det_w": (7, np.float64),
    "jacobian_tiw": (8, np.float64),
    "mass_inverse": (9, np.float64),
    "min_edge": (10, np.float64),
    "face_normal": (11, np.float64),
    "face_jacobian_w": (12, np.float64),
    "face_coord": (13, np.float64),
    "face_parent": (14, np.int32),
    "face_local": (15, np.int32),
    "face_sequence": (16, np.int32),
    "face_rotation": (17, np.int32),
    "quad_point": (18, np.float64),
    "face_quad_point": (19, np.float64),
    "boundary_face_coord": (20, np.float64),
    "boundary_face_normal": (21, np.float64),
    "boundary_face_jacobian_w": (22, np.float64),
    "boundary_face_parent": (23, np.int32),
    "boundary_face_type": (24, np.int32),
    "element_type": (25, np.int32),
    "inner_radius": (26, np.float64),
    "node_tag": (27, np.int32),
    "nodal_value": (28, np.float64),
    "nodal_face_value": (29, np.float64),
    "boundary_face_physical": (30, np.int32),
}


class ProblemDesc(Structure):
    _fields_ = [
        ("abi_version", c_int),
        ("equation", c_int),
        ("convective_flux", c_int),
        ("viscous_flux", c_int),
        ("time_integration", c_int),
        ("equation_of_state", c_int),
        ("transport", c_int),
        ("precision", c_int),
        ("cp", c_double),
        ("cv", c_double),
        ("dynamic_viscosity", c_double),
        ("reference_sound_speed", c_double),
        ("reference_density", c_double),
        ("source_term", c_int),
        ("pad_", c_int),
        ("thermal_expansion_coefficient", c_double),
        ("reference_temperature", c_double),
        ("shock_capturing", c_int),
        ("pad2_", c_int),
        ("empirical_tolerance", c_double),
        ("artificial_viscosity_factor", c_double),
        ("kernels", c_int),
        ("launch", c_int),
    ]


class StructuredDesc(Structure):
    _fields_ = [
        ("element", c_int),
        ("polynomial_order", c_int),
        ("n", c_int * 3),
        ("lo", c_double * 3),
        ("hi", c_double * 3),
        ("bounded", c_int * 3),
        ("boundary", c_int * 6),
    ]


class SplitDesc(Structure):
    _fields_ = [
        ("polynomial_order", c_int),
        ("n", c_int * 2),
        ("lo", c_double * 2),
        ("hi", c_double * 2),
        ("split_probability", c_double),
        ("seed", ctypes.c_uint64),
    ]


class ArrayMeshDesc(Structure):
    _fields_ = [
        ("element", c_int),
        ("polynomial_order", c_int),
        ("geometry_order", c_int),
        ("node_count", c_int64),
        ("nodes", c_void_p),
        ("element_count", c_int64),
        ("elements", c_void_p),
        ("boundary_face_count", c_int64),
        ("boundary_faces", c_void_p),
        ("boundary_type", c_void_p),
        ("boundary_physical", c_void_p),
        ("periodic_link_count", c_int),
        ("periodic_link_size", c_void_p),
        ("periodic_pairs", c_void_p),
    ]


class GmshDesc(Structure):
    _fields_ = [
        ("path", c_char_p),
        ("polynomial_order", c_int),
        ("boundary_count", c_int),
        ("physical_index", POINTER(c_int)),
        ("boundary", POINTER(c_int)),
    ]


class MeshInfo(Structure):
    _fields_ = [
        ("dimension", c_int),
        ("element", c_int),
        ("polynomial_order", c_int),
        ("elements", c_int64),
        ("faces", c_int64),
        ("basis", c_int),
        ("quadrature", c_int),
        ("faces_per_element", c_int),
        ("face_quadrature", c_int),
        ("affine", c_int),
        ("boundary_faces", c_int64),
        ("nodes", c_int64),
        ("vertices", c_int),
        ("pad_", c_int),
    ]


# every symbol declared in include/sdg.h, with its ctypes signature
SIGNATURES = {
    "sdg_abi_version": (c_int, []),
    "sdg_last_error": (c_char_p, [c_void_p]),
    "sdg_mesh_create_structured": (c_int, [POINTER(StructuredDesc), POINTER(c_void_p)]),
    "sdg_mesh_destroy": (None, [c_void_p]),
    "sdg_mesh_get_info": (c_int, [c_void_p, POINTER(MeshInfo)]),
    "sdg_mesh_field_size": (c_int64, [c_void_p, c_int]),
    "sdg_mesh_export": (c_int, [c_void_p, c_int, c_void_p, c_int64]),
    "sdg_create": (c_int, [POINTER(ProblemDesc), c_void_p, POINTER(c_void_p)]),
    "sdg_destroy": (None, [c_void_p]),
    "sdg_set_initial_primitive": (c_int, [c_void_p, c_void_p]),
    "sdg_set_state": (c_int, [c_void_p, c_void_p]),
    "sdg_get_state": (c_int, [c_void_p, c_void_p]),
    "sdg_compute_dt": (c_int, [c_void_p, c_double, POINTER(c_double)]),
    "sdg_compute_dt_noalloc": (c_int, [c_void_p, c_double, POINTER(c_double)]),
    "sdg_residual": (c_int, [c_void_p, c_void_p]),
    "sdg_set_boundary_primitive": (c_int, [c_void_p, c_void_p]),
    "sdg_get_gradient": (c_int, [c_void_p, c_void_p]),
    "sdg_get_artificial_viscosity": (c_int, [c_void_p, c_void_p]),
    "sdg_step": (c_int, [c_void_p, c_double, c_int, c_void_p, POINTER(c_int), POINTER(c_int)]),
    "sdg_last_step_ms": (c_int, [c_void_p, POINTER(c_double)]),
    "sdg_synchronize": (c_int, [c_void_p]),
    "sdg_time_kernels": (c_int, [c_void_p, c_double, c_int, c_void_p, c_int]),
    "sdg_vjp_residual": (c_int, [c_void_p, c_void_p, c_void_p]),
    "sdg_mesh_create_split": (c_int, [POINTER(SplitDesc), POINTER(c_void_p)]),
    "sdg_mesh_create_gmsh": (c_int, [POINTER(GmshDesc), POINTER(c_void_p)]),
    "sdg_mesh_create_from_arrays": (c_int, [POINTER(ArrayMeshDesc), POINTER(c_void_p)]),
    "sdg_batch_create": (c_int, [POINTER(ProblemDesc), c_void_p, c_int, POINTER(c_void_p)]),
    "sdg_batch_destroy": (None, [c_void_p]),
    "sdg_batch_last_error": (c_char_p, [c_void_p]),
    "sdg_batch_set_initial_primitive": (c_int, [c_void_p, c_int, c_void_p]),
    "sdg_batch_set_state": (c_int, [c_void_p, c_int, c_void_p]),
    "sdg_batch_get_state": (c_int, [c_void_p, c_int, c_void_p]),
    "sdg_batch_compute_dt": (c_int, [c_void_p, c_double, c_void_p]),
    "sdg_batch_step": (c_int, [c_void_p, c_void_p, c_int]),
    "sdg_batch_last_step_ms": (c_int, [c_void_p, POINTER(c_double)]),
    "sdg_batch_tune": (c_int, [c_void_p, POINTER(c_double)]),
    "sdg_batch_get_config": (c_int, [c_void_p, c_void_p, c_void_p, POINTER(c_int)]),
    "sdg_adjoint_forward": (c_int, [c_void_p, c_double, c_int, c_int]),
    "sdg_adjoint_backward": (c_int, [c_void_p, c_void_p, c_void_p]),
    "sdg_adjoint_backward_device": (c_int, [c_void_p, c_void_p, c_void_p]),
    "sdg_raw_binary_write": (c_int, [c_void_p, c_char_p, c_int64]),
    "sdg_raw_binary_read": (c_int, [c_void_p, c_char_p, c_int]),
    "sdg_raw_file_write": (c_int, [c_char_p, c_void_p, c_int64]),
    "sdg_write_vtu": (c_int, [c_void_p, POINTER(ProblemDesc), c_void_p, c_void_p, c_int, c_double, c_char_p]),
    "sdg_write_view": (c_int, [c_void_p, POINTER(ProblemDesc), c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                               c_double, c_char_p]),
    "sdg_raw_file_read": (c_int, [c_char_p, c_void_p, c_int64, POINTER(c_int64)]),
}

_lib = None


class SDGError(RuntimeError):
    pass


def load(path: str | None = None):
    """Load libsdg_hip.so; raise if it was not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise SDGError(
            f"{p} is missing: build the HIP extension first (python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.sdg_abi_version() != ABI_VERSION:
        raise SDGError("ABI version mismatch between subrosadg_amd and libsdg_hip.so")
    if path is None:
        _lib = lib
    return lib


def check(rc: int, ctx=None):
    if rc != OK:
        lib = load()
        msg = lib.sdg_last_error(ctx)
        raise SDGError(f"sdg error {rc}: {msg.decode() if msg else ''}")


def problem_desc(**kw) -> ProblemDesc:
    p = ProblemDesc()
    p.abi_version = ABI_VERSION
    p.equation = kw.get("equation", EQ_EULER)
    p.convective_flux = kw.get("convective_flux", FLUX_ROE)
    p.viscous_flux = kw.get("viscous_flux", VISC_NONE)
    p.time_integration = kw.get("time_integration", RK_SSPRK3)
    p.equation_of_state = kw.get("equation_of_state", EOS_IDEAL_GAS)
    p.transport = kw.get("transport", TRANSPORT_NONE)
    p.precision = kw.get("precision", PREC_FP64)
    p.cp = kw.get("cp", 3.5)
    p.cv = kw.get("cv", 2.5)
    p.dynamic_viscosity = kw.get("dynamic_viscosity", 0.0)
    p.reference_sound_speed = kw.get("reference_sound_speed", 0.0)
    p.reference_density = kw.get("reference_density", 0.0)
    p.source_term = kw.get("source_term", SOURCE_NONE)
    p.thermal_expansion_coefficient = kw.get("thermal_expansion_coefficient", 0.0)
    p.reference_temperature = kw.get("reference_temperature", 0.0)
    p.shock_capturing = kw.get("shock_capturing", SHOCK_NONE)
    p.empirical_tolerance = kw.get("empirical_tolerance", 0.0)
    p.artificial_viscosity_factor = kw.get("artificial_viscosity_factor", 1.0)
    p.kernels = kw.get("kernels", KERNELS_AUTO)
    p.launch = kw.get("launch", LAUNCH_GRAPH)
    return p


class Mesh:
    """Structured mesh on a box plus its operand tables (host only). Periodic in every direction
    unless `boundary` gives, per direction, the (low, high) boundary types (BC_* constants) or None
    for a periodic direction."""

    def __init__(self, element: int, order: int, n, lo, hi, boundary=None, _handle=None):
        self.lib = load()
        if _handle is not None:
            self._attach(_handle)
            return
        d = StructuredDesc()
        d.element = element
        d.polynomial_order = order
        for i in range(3):
            d.n[i] = int(n[i]) if i < len(n) else 1
            d.lo[i] = float(lo[i]) if i < len(lo) else 0.0
            d.hi[i] = float(hi[i]) if i < len(hi) else 1.0
        if boundary is not None:
            for i, sides in enumerate(boundary):
                if sides is not None:
                    d.bounded[i] = 1
                    d.boundary[2 * i] = int(sides[0])
                    d.boundary[2 * i + 1] = int(sides[1])
        h = c_void_p()
        check(self.lib.sdg_mesh_create_structured(ctypes.byref(d), ctypes.byref(h)))
        self._attach(h)

    @classmethod
    def split(cls, order: int, n, lo, hi, split_probability: float, seed: int):
        """Periodic split grid: each cell a quadrangle or (with the given probability, seeded) two
        triangles."""
        lib = load()
        d = SplitDesc()
        d.polynomial_order = order
        for i in range(2):
            d.n[i] = int(n[i])
            d.lo[i] = float(lo[i])
            d.hi[i] = float(hi[i])
        d.split_probability = float(split_probability)
        d.seed = int(seed)
        h = c_void_p()
        check(lib.sdg_mesh_create_split(ctypes.byref(d), ctypes.byref(h)))
        return cls(0, order, (), (), (), _handle=h)

    @classmethod
    def gmsh(cls, path: str, order: int, boundary: dict | None = None):
        """Mesh from a gmsh MSH 4.1 ASCII file; `boundary` maps physical tags to BC_* types."""
        lib = load()
        d = GmshDesc()
        d.path = path.encode()
        d.polynomial_order = order
        items = sorted((boundary or {}).items())
        ph = (c_int * max(1, len(items)))(*[int(k) for k, _ in items])
        bc = (c_int * max(1, len(items)))(*[int(v) for _, v in items])
        d.boundary_count = len(items)
        d.physical_index = ph
        d.boundary = bc
        h = c_void_p()
        check(lib.sdg_mesh_create_gmsh(ctypes.byref(d), ctypes.byref(h)))
        return cls(0, order, (), (), (), _handle=h)

    @classmethod
    def from_arrays(cls, element: int, order: int, nodes, elements, boundary_faces=None, boundary_type=None,
                    boundary_physical=None, periodic=None, geometry_order=1):
        """Mesh from host arrays (sdg_mesh_create_from_arrays): nodes (N, 3) or (N, 2), elements (E,
        geometry nodes) of 0-based nodes in gmsh order (the vertices for geometry_order 1); boundary
        faces (F_b, face vertices) with BC_* types and optional physical tags; periodic: a list of
        links, each a list of (slave, master) node pairs."""
        lib = load()
        xyz = np.zeros((len(nodes), 3))
        nodes = np.asarray(nodes, dtype=np.float64)
        xyz[:, : nodes.shape[1]] = nodes
        conn = np.ascontiguousarray(elements, dtype=np.int64)
        keep = [xyz, conn]
        d = ArrayMeshDesc()
        d.element, d.polynomial_order, d.geometry_order = element, order, geometry_order
        d.node_count, d.nodes = xyz.shape[0], xyz.ctypes.data
        d.element_count, d.elements = conn.shape[0], conn.ctypes.data
        if boundary_faces is not None and len(boundary_faces):
            bf = np.ascontiguousarray(boundary_faces, dtype=np.int64)
            bt = np.ascontiguousarray(boundary_type, dtype=np.int32)
            keep += [bf, bt]
            d.boundary_face_count, d.boundary_faces, d.boundary_type = bf.shape[0], bf.ctypes.data, bt.ctypes.data
            if boundary_physical is not None:
                bp = np.ascontiguousarray(boundary_physical, dtype=np.int32)
                keep.append(bp)
                d.boundary_physical = bp.ctypes.data
        if periodic:
            sizes = np.array([len(link) for link in periodic], dtype=np.int64)
            pairs = np.ascontiguousarray(np.concatenate([np.asarray(link, dtype=np.int64).reshape(-1, 2)
                                                         for link in periodic]), dtype=np.int64)
            keep += [sizes, pairs]
            d.periodic_link_count, d.periodic_link_size, d.periodic_pairs = len(periodic), sizes.ctypes.data, \
                pairs.ctypes.data
        h = c_void_p()
        check(lib.sdg_mesh_create_from_arrays(ctypes.byref(d), ctypes.byref(h)))
        del keep
        return cls(0, order, (), (), (), _handle=h)

    def _attach(self, h):
        self.handle = h
        info = MeshInfo()
        check(self.lib.sdg_mesh_get_info(h, ctypes.byref(info)))
        self.info = info
        self.dim = info.dimension
        self.order = info.polynomial_order
        self.E = info.elements
        self.F = info.faces
        self.nb = info.basis
        self.nq = info.quadrature
        self.nf = info.faces_per_element
        self.nfq = info.face_quadrature
        self.Fb = info.boundary_faces
        self.nodes = info.nodes
        self.nvert = info.vertices
        self.nv = self.dim + 2
        self.elem = info.element
        self.mixed = info.element == ELEM_MIXED_TRIANGLE_QUADRANGLE

    def field(self, name: str) -> np.ndarray:
        fid, dt = FIELDS[name]
        n = self.lib.sdg_mesh_field_size(self.handle, fid)
        out = np.empty(n, dtype=dt)
        check(self.lib.sdg_mesh_export(self.handle, fid, out.ctypes.data, n))
        return out

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.sdg_mesh_destroy(h)
            self.handle = None


def _hip_runtime():
    """The HIP runtime the library links (already loaded with it, found by its soname)."""
    load()
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [POINTER(c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [c_void_p]
    lib.hipMemcpy.argtypes = [c_void_p, c_void_p, ctypes.c_size_t, c_int]
    for f in (lib.hipMalloc, lib.hipFree, lib.hipMemcpy):
        f.restype = c_int
    return lib


class DeviceArray:
    """A float64 array in device memory (hipMalloc on the current device), for the entry points
    that take device pointers. Freed with the object."""

    def __init__(self, shape):
        self.shape = tuple(int(x) for x in shape)
        self.nbytes = 8 * int(np.prod(self.shape))
        self._hip = _hip_runtime()
        self._ptr = c_void_p()
        if self._hip.hipMalloc(ctypes.byref(self._ptr), max(self.nbytes, 8)) != 0:
            raise MemoryError(f"hipMalloc of {self.nbytes} bytes failed")

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.shape)
        if d._hip.hipMemcpy(d._ptr, a.ctypes.data, d.nbytes, 1) != 0:  # host to device
            raise RuntimeError("hipMemcpy to the device failed")
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape)
        if self._hip.hipMemcpy(out.ctypes.data, self._ptr, self.nbytes, 2) != 0:  # device to host
            raise RuntimeError("hipMemcpy from the device failed")
        return out

    def data_ptr(self) -> int:
        return self._ptr.value

    def __del__(self):
        if getattr(self, "_ptr", None) is not None and self._ptr.value:
            self._hip.hipFree(self._ptr)
            self._ptr = None


def _device_ptr(x) -> int:
    return x.data_ptr()


class Solver:
    """A device solver context (sdg_ctx) on one mesh: the drop-in for Solver<SC>."""

    def __init__(self, problem: ProblemDesc, mesh: Mesh):
        self.lib = load()
        self.mesh = mesh
        self.problem = problem
        h = c_void_p()
        rc = self.lib.sdg_create(ctypes.byref(problem), mesh.handle, ctypes.byref(h))
        check(rc)
        self.handle = h
        self.E, self.nb, self.nv, self.dim = mesh.E, mesh.nb, mesh.nv, mesh.dim

    def _check(self, rc):
        if rc != OK:
            msg = self.lib.sdg_last_error(self.handle)
            raise SDGError(f"sdg error {rc}: {msg.decode() if msg else ''}")

    def state_shape(self):
        return (self.E, self.nb, self.nv)

    def set_state(self, c: np.ndarray):
        c = np.ascontiguousarray(c, dtype=np.float64)
        assert c.size == self.E * self.nb * self.nv
        self._check(self.lib.sdg_set_state(self.handle, c.ctypes.data))

    def get_state(self) -> np.ndarray:
        out = np.empty(self.state_shape())
        self._check(self.lib.sdg_get_state(self.handle, out.ctypes.data))
        return out

    def set_initial_primitive(self, prim: np.ndarray):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        assert prim.size == self.E * self.mesh.nq * self.nv
        self._check(self.lib.sdg_set_initial_primitive(self.handle, prim.ctypes.data))

    def residual(self) -> np.ndarray:
        out = np.empty(self.state_shape())
        self._check(self.lib.sdg_residual(self.handle, out.ctypes.data))
        return out

    def gradient(self) -> np.ndarray:
        out = np.empty((self.E, self.nb, self.nv * self.dim))
        self._check(self.lib.sdg_get_gradient(self.handle, out.ctypes.data))
        return out

    def write_vtu(self, path: str, step: int = 0, time: float = 0.0):
        """VTU file of the current state (with the node artificial viscosity when shock capturing
        is on); see sdg_write_vtu."""
        nodes = self.artificial_viscosity() if getattr(self.problem, "shock_capturing", 0) else None
        write_vtu(self.mesh, self.problem, self.get_state(), path, node_values=nodes, step=step, time=time)

    def artificial_viscosity(self) -> np.ndarray:
        """Node values of the last artificial-viscosity evaluation (mesh nodes long)."""
        out = np.empty(self.mesh.nodes)
        self._check(self.lib.sdg_get_artificial_viscosity(self.handle, out.ctypes.data))
        return out

    def set_boundary_primitive(self, prim: np.ndarray):
        """Primitive boundary data (rho, u.., T) at every boundary face point (F_b, N_fq, NV)."""
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        assert prim.size == self.mesh.Fb * self.mesh.nfq * self.nv
        self._check(self.lib.sdg_set_boundary_primitive(self.handle, prim.ctypes.data if prim.size else None))

    def vjp_residual(self, lam: np.ndarray) -> np.ndarray:
        """(dR/dU)^T lam at the current state (reference layout)."""
        lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(self.state_shape())
        out = np.empty(self.state_shape())
        self._check(self.lib.sdg_vjp_residual(self.handle, lam.ctypes.data, out.ctypes.data))
        return out

    def adjoint_forward(self, dt: float, nsteps: int, snapshots: int):
        """Advance nsteps steps recording binomial checkpoints (at most `snapshots` states)."""
        self._check(self.lib.sdg_adjoint_forward(self.handle, dt, nsteps, snapshots))

    def adjoint_backward(self, lam_final: np.ndarray) -> np.ndarray:
        """dJ/dU(t_0) from dJ/dU(t_N) for the last adjoint_forward run."""
        lam = np.ascontiguousarray(lam_final, dtype=np.float64).reshape(self.state_shape())
        out = np.empty(self.state_shape())
        self._check(self.lib.sdg_adjoint_backward(self.handle, lam.ctypes.data, out.ctypes.data))
        return out

    def adjoint_backward_device(self, lam_final, lam_initial=None):
        """adjoint_backward with both adjoints in device memory of the solver's device (float64,
        reference layout): DeviceArray buffers or CUDA tensors. Nothing crosses to the host.
        Returns lam_initial (a DeviceArray, or an empty_like tensor, when not given)."""
        assert tuple(lam_final.shape) == self.state_shape()
        if lam_initial is None:
            lam_initial = DeviceArray(lam_final.shape) if isinstance(lam_final, DeviceArray) else \
                lam_final.new_empty(lam_final.shape)
        assert tuple(lam_initial.shape) == self.state_shape()
        self._check(self.lib.sdg_adjoint_backward_device(self.handle, _device_ptr(lam_final), _device_ptr(lam_initial)))
        return lam_initial

    def write_raw(self, path: str, node_count: int = 0):
        """Raw state file in the reference's format (header + zstd frame; RawBinary.cpp:42-191)."""
        self._check(self.lib.sdg_raw_binary_write(self.handle, os.fsencode(path), node_count))

    def read_raw(self, path: str, initial_condition: int = IC_LAST_STEP):
        """Start from a raw file: IC_LAST_STEP (same discretisation) or IC_SPECIFIC_FILE (order P-1)."""
        self._check(self.lib.sdg_raw_binary_read(self.handle, os.fsencode(path), initial_condition))

    def compute_dt(self, cfl: float) -> float:
        dt = c_double()
        self._check(self.lib.sdg_compute_dt(self.handle, cfl, ctypes.byref(dt)))
        return dt.value

    def step(self, dt: float, nsteps: int = 1):
        rel = np.empty(self.nv)
        nan = c_int()
        taken = c_int()
        self._check(self.lib.sdg_step(self.handle, dt, nsteps, rel.ctypes.data, ctypes.byref(nan), ctypes.byref(taken)))
        return rel, nan.value, taken.value

    def last_step_ms(self) -> float:
        ms = c_double()
        self._check(self.lib.sdg_last_step_ms(self.handle, ctypes.byref(ms)))
        return ms.value

    def time_kernels(self, dt: float, nsteps: int) -> np.ndarray:
        out = np.zeros(4)
        self._check(self.lib.sdg_time_kernels(self.handle, dt, nsteps, out.ctypes.data, 4))
        return out

    def synchronize(self):
        self._check(self.lib.sdg_synchronize(self.handle))

    def close(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.sdg_destroy(h)
            self.handle = None

    def __del__(self):
        self.close()


class Batch:
    """Batched solver (sdg_batch): independent 2-D Euler problems on split grids, one persistent
    launch per RK stage for all of them."""

    def __init__(self, problem: ProblemDesc, meshes):
        self.lib = load()
        self.meshes = list(meshes)
        self.problem = problem
        arr = (c_void_p * len(self.meshes))(*[m.handle.value for m in self.meshes])
        h = c_void_p()
        rc = self.lib.sdg_batch_create(ctypes.byref(problem), ctypes.cast(arr, c_void_p), len(self.meshes),
                                       ctypes.byref(h))
        check(rc)
        self.handle = h
        self.sizes = []
        for m in self.meshes:
            et = m.field("element_type")
            nb = {ELEM_TRIANGLE: (m.order + 1) * (m.order + 2) // 2, ELEM_QUADRANGLE: (m.order + 1) ** 2}
            self.sizes.append(int(sum(nb[int(t)] for t in et)) * 4)

    def _check(self, rc):
        if rc != OK:
            msg = self.lib.sdg_batch_last_error(self.handle)
            raise SDGError(f"sdg error {rc}: {msg.decode() if msg else ''}")

    def set_initial_primitive(self, i: int, prim: np.ndarray):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        self._check(self.lib.sdg_batch_set_initial_primitive(self.handle, i, prim.ctypes.data))

    def set_state(self, i: int, c: np.ndarray):
        c = np.ascontiguousarray(c, dtype=np.float64)
        assert c.size == self.sizes[i]
        self._check(self.lib.sdg_batch_set_state(self.handle, i, c.ctypes.data))

    def get_state(self, i: int) -> np.ndarray:
        out = np.empty(self.sizes[i])
        self._check(self.lib.sdg_batch_get_state(self.handle, i, out.ctypes.data))
        return out

    def compute_dt(self, cfl: float) -> np.ndarray:
        out = np.empty(len(self.meshes))
        self._check(self.lib.sdg_batch_compute_dt(self.handle, cfl, out.ctypes.data))
        return out

    def step(self, dt, nsteps: int):
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        self._check(self.lib.sdg_batch_step(self.handle, dt.ctypes.data, nsteps))

    def last_step_ms(self) -> float:
        ms = c_double()
        self._check(self.lib.sdg_batch_last_step_ms(self.handle, ctypes.byref(ms)))
        return ms.value

    def tune(self) -> float:
        ms = c_double()
        self._check(self.lib.sdg_batch_tune(self.handle, ctypes.byref(ms)))
        return ms.value

    def config(self):
        chunk = np.zeros(8, dtype=np.int32)
        epb = np.zeros(8, dtype=np.int32)
        grid = c_int()
        self._check(self.lib.sdg_batch_get_config(self.handle, chunk.ctypes.data, epb.ctypes.data, ctypes.byref(grid)))
        return chunk, epb, grid.value

    def close(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.sdg_batch_destroy(h)
            self.handle = None

    def __del__(self):
        self.close()


def raw_file_write(path: str, payload: bytes):
    """Compress `payload` into a raw file (no device needed)."""
    lib = load()
    buf = ctypes.create_string_buffer(payload, len(payload))
    check(lib.sdg_raw_file_write(os.fsencode(path), ctypes.cast(buf, c_void_p), len(payload)))


def write_vtu(mesh: "Mesh", problem: ProblemDesc, coefficients: np.ndarray, path: str, node_values=None,
              step: int = 0, time: float = 0.0):
    """VTK XML output of a state (host only): see sdg_write_vtu in include/sdg.h."""
    lib = load()
    c = np.ascontiguousarray(coefficients, dtype=np.float64)
    assert c.size == mesh.E * mesh.nb * mesh.nv
    nv = None if node_values is None else np.ascontiguousarray(node_values, dtype=np.float64)
    check(lib.sdg_write_vtu(mesh.handle, ctypes.byref(problem), c.ctypes.data,
                            None if nv is None else nv.ctypes.data, int(step), float(time), path.encode()))


def write_view(mesh: "Mesh", problem: ProblemDesc, coefficients: np.ndarray, path: str, variables,
               gradient=None, node_values=None, step: int = 0, time: float = 0.0):
    """VTK XML output of the named view variables (host only): see sdg_write_view in include/sdg.h."""
    lib = load()
    c = np.ascontiguousarray(coefficients, dtype=np.float64)
    assert c.size == mesh.E * mesh.nb * mesh.nv
    g = None if gradient is None else np.ascontiguousarray(gradient, dtype=np.float64)
    nv = None if node_values is None else np.ascontiguousarray(node_values, dtype=np.float64)
    ids = np.array([VIEW_VARIABLES.index(v) if isinstance(v, str) else int(v) for v in variables], dtype=np.int32)
    check(lib.sdg_write_view(mesh.handle, ctypes.byref(problem), c.ctypes.data, None if g is None else g.ctypes.data,
                             None if nv is None else nv.ctypes.data, ids.ctypes.data, int(ids.size), int(step),
                             float(time), path.encode()))


def raw_file_read(path: str) -> bytes:
    """The decompressed payload of a raw file (no device needed)."""
    lib = load()
    n = c_int64()
    check(lib.sdg_raw_file_read(os.fsencode(path), None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(max(n.value, 1))
    check(lib.sdg_raw_file_read(os.fsencode(path), ctypes.cast(buf, c_void_p), n.value, ctypes.byref(n)))
    return buf.raw[: n.value]
