"""VTU output (sdg_write_vtu, host only; the reference's View::stepView, Paraview.cpp:280-313):
the XML parses, the sample points are the affine images of the equispaced reference points, the
point data equal the closed forms of a state whose conserved fields are polynomials of degree P
(exactly represented, so density, velocity, pressure, temperature and Mach number are known at
every point), the artificial viscosity is the P1 interpolant of the node values, and the linear
sub-cells tile each element with positive orientation (their areas / volumes sum to the domain's)."""
import ctypes
import math
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import problems as P
from subrosadg_amd import _lib as L

GAMMA, CV = 1.4, 2.5


def _fields(x):
    """Conserved fields of degree <= 2 in every coordinate (rho > 0, p > 0 on the unit box)."""
    X = x[..., 0]
    Y = x[..., 1] if x.shape[-1] >= 2 else 0.5 * X
    Z = x[..., 2] if x.shape[-1] == 3 else 0.0
    rho = 1.0 + 0.2 * X + 0.1 * Y * Y + 0.05 * Z
    mom = [0.3 + 0.1 * X * Y, -0.2 + 0.1 * Y, 0.1 * X + 0.05 * Z][: x.shape[-1]]
    rhoE = 3.0 + 0.3 * X * X + 0.2 * Y + 0.1 * Z
    return [rho] + mom + [rhoE]


def _mesh(kind, order):
    if kind == "line":
        return L.Mesh(L.ELEM_LINE, order, (4,), (0.0,), (1.0,))
    if kind == "tri":
        return L.Mesh.split(order, (3, 3), (0.0, 0.0), (1.0, 1.0), 1.0, 2)
    if kind == "hex":
        return L.Mesh(L.ELEM_HEXAHEDRON, order, (2, 2, 2), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    return L.Mesh(L.ELEM_QUADRANGLE, order, (3, 2), (0.0, 0.0), (1.0, 1.0))


def _read(path):
    root = ET.parse(path).getroot()
    arrays = {}
    for da in root.iter("DataArray"):
        name = da.get("Name")
        vals = np.array(da.text.split(), dtype=np.float64)
        comps = int(da.get("NumberOfComponents", "1"))
        arrays[name] = vals.reshape(-1, comps) if comps > 1 else vals
    piece = root.find("UnstructuredGrid/Piece")
    return arrays, int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells"))


@pytest.mark.parametrize("kind,order", [("line", 2), ("line", 4), ("quad", 2), ("quad", 3), ("tri", 2), ("tri", 3),
                                        ("hex", 2)])
def test_vtu_points_fields_and_cells(tmp_path, kind, order):
    m = _mesh(kind, order)
    prob = L.problem_desc(equation=L.EQ_EULER, cp=GAMMA * CV, cv=CV)
    xq = m.field("quad_coord").reshape(m.E, m.nq, m.dim)
    C = P.project_conserved(m, np.stack(_fields(xq), axis=-1))  # exact: degree <= P per coordinate
    node = np.linspace(0.0, 1.0, m.nodes)
    path = str(tmp_path / "out.vtu")
    L.write_vtu(m, prob, C, path, node_values=node, step=7, time=0.25)
    a, npts, ncells = _read(path)
    per = (order + 1) * (order + 2) // 2 if kind == "tri" else (order + 1) ** m.dim
    sub = order * order if kind == "tri" else order ** m.dim
    assert npts == m.E * per and ncells == m.E * sub
    assert a["TMSTEP"] == 7 and a["TimeValue"] == 0.25
    x = a["Points"][:, : m.dim]
    f = _fields(x)
    rho, mom, rhoE = f[0], np.stack(f[1:-1], axis=-1), f[-1]
    vel = mom / rho[:, None]
    p = (GAMMA - 1.0) * (rhoE - 0.5 * (mom * vel).sum(-1))
    assert np.allclose(a["Density"], rho, rtol=1e-12, atol=1e-13)
    assert np.allclose(a["Velocity"][:, : m.dim], vel, rtol=1e-12, atol=1e-12)
    assert np.all(a["Velocity"][:, m.dim:] == 0.0)
    assert np.allclose(a["Pressure"], p, rtol=1e-11)
    assert np.allclose(a["Temperature"], p / ((GAMMA - 1.0) * rho * CV), rtol=1e-11)
    assert np.allclose(a["MachNumber"], np.sqrt((vel ** 2).sum(-1)) / np.sqrt(GAMMA * p / rho), rtol=1e-11)
    # artificial viscosity: node values at the element corners, P1 interpolant elsewhere
    tag = m.field("node_tag").reshape(m.E, m.nvert)
    pts = x.reshape(m.E, per, m.dim)
    av = a["ArtificialViscosity"].reshape(m.E, per)
    corners = [0, order, per - 1] if kind == "tri" else None
    if kind == "line":
        corners = [0, order]
    if kind == "quad":
        corners = [0, order, per - 1, per - 1 - order]
    if kind == "hex":
        n1 = order + 1
        corners = [0, order, n1 * n1 - 1, n1 * order, n1 * n1 * order, n1 * n1 * order + order, per - 1,
                   per - 1 - order]
    assert np.allclose(av[:, corners], node[tag], rtol=1e-14, atol=1e-15)
    # the sub-cells tile the elements with positive orientation
    conn = a["connectivity"].astype(np.int64).reshape(ncells, -1)
    cx = a["Points"][conn][..., : m.dim]
    if m.dim == 1:
        length = cx[:, 1, 0] - cx[:, 0, 0]
        assert conn.shape[1] == 2 and np.all(length > 0.0) and math.isclose(length.sum(), 1.0, rel_tol=1e-12)
        assert np.allclose(length, 1.0 / (m.E * order), rtol=1e-12)  # equispaced points of order P
    elif m.dim == 2:
        xs, ys = cx[..., 0], cx[..., 1]
        area = 0.5 * (xs * np.roll(ys, -1, axis=1) - np.roll(xs, -1, axis=1) * ys).sum(1)
        assert np.all(area > 0.0) and math.isclose(area.sum(), 1.0, rel_tol=1e-12)
    else:
        e1, e3, e4 = cx[:, 1] - cx[:, 0], cx[:, 3] - cx[:, 0], cx[:, 4] - cx[:, 0]
        vol = np.einsum("ci,ci->c", np.cross(e1, e3), e4)
        assert np.all(vol > 0.0) and math.isclose(vol.sum(), 1.0, rel_tol=1e-12)
        assert np.allclose(cx[:, 6], cx[:, 0] + e1 + e3 + e4)  # opposite corner of a box cell
    vtk = {"line": 3, "tri": 5, "quad": 9, "hex": 12}[kind]
    assert set(np.unique(a["types"]).astype(int)) == {vtk}


def test_vtu_without_viscosity_and_errors(tmp_path):
    m = _mesh("quad", 2)
    prob = L.problem_desc(equation=L.EQ_EULER, cp=GAMMA * CV, cv=CV)
    xq = m.field("quad_coord").reshape(m.E, m.nq, m.dim)
    C = P.project_conserved(m, np.stack(_fields(xq), axis=-1))
    path = str(tmp_path / "plain.vtu")
    L.write_vtu(m, prob, C, path)
    a, _, _ = _read(path)
    assert "ArtificialViscosity" not in a
    with pytest.raises(L.SDGError):
        L.write_vtu(m, prob, C, str(tmp_path / "missing" / "x.vtu"))
    mixed = L.Mesh.split(2, (3, 3), (0.0, 0.0), (1.0, 1.0), 0.5, 4)
    buf = np.zeros(mixed.E * 9 * 4)
    rc = L.load().sdg_write_vtu(mixed.handle, ctypes.byref(prob), buf.ctypes.data, None, 0, 0.0, path.encode())
    assert rc == L.E_UNSUPPORTED
