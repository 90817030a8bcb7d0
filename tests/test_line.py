"""1-D line meshes (DimensionEnum::D1 of the reference; end-point faces with normals -1 / +1,
Geometry.cpp:102-112): the operand tables in closed form, and the oracle's free-stream
preservation and discrete conservation on periodic and bounded lines."""
import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L


def line(n, order, boundary=None, lo=0.0, hi=1.0):
    return L.Mesh(L.ELEM_LINE, order, (n,), (lo,), (hi,), boundary=boundary)


@pytest.mark.parametrize("order", [1, 3, 5])
def test_line_tables(order):
    m = line(8, order, lo=-1.0, hi=3.0)
    h = 0.5
    assert (m.dim, m.nb, m.nq, m.nf, m.nfq, m.E, m.F, m.Fb) == (1, order + 1, order + 1, 2, 1, 8, 8, 0)
    assert np.isclose(m.field("jacobian_det_w").sum(), 4.0, rtol=1e-14)
    assert np.allclose(m.field("face_jacobian_w"), 1.0)
    par, loc = m.field("face_parent").reshape(-1, 2), m.field("face_local").reshape(-1, 2)
    # outward from the left (first-visiting) parent: -1 at its first end point, +1 at its second
    assert np.array_equal(m.field("face_normal"), np.where(loc[:, 0] == 0, -1.0, 1.0))
    assert np.all(loc[:, 1] == 1 - loc[:, 0])
    assert np.all(par[:, 1] == np.where(loc[:, 0] == 1, par[:, 0] + 1, par[:, 0] - 1) % 8)
    assert np.allclose(m.field("inner_radius"), h / 2) and np.allclose(m.field("min_edge"), h)
    assert m.nodes == 9
    # the basis is hierarchic: the first P functions are the order P-1 basis (values at points)
    V = m.field("basis_value").reshape(m.nq, m.nb)
    assert np.linalg.matrix_rank(V) == m.nb
    b = line(4, order, boundary=[(L.BC_RIEMANN_FARFIELD, L.BC_ADIABATIC_SLIP_WALL)])
    assert (b.F, b.Fb) == (3, 2)
    assert np.allclose(b.field("boundary_face_normal"), [-1.0, 1.0])
    assert list(b.field("boundary_face_type")) == [L.BC_RIEMANN_FARFIELD, L.BC_ADIABATIC_SLIP_WALL]


def _euler(**kw):
    return L.problem_desc(equation=L.EQ_EULER, convective_flux=kw.pop("flux", L.FLUX_HLLC), cp=3.5, cv=2.5, **kw)


@pytest.mark.parametrize("order", [2, 4])
def test_line_free_stream_and_conservation(order):
    m = line(10, order)
    o = O.Oracle(_euler(), m)
    u = np.array([1.0, 0.4, 1.0])
    o.set_initial_primitive(np.tile(u, (m.E, m.nq, 1)))
    assert np.abs(o.residual()).max() < 1e-13
    xq = m.field("quad_coord").reshape(m.E, m.nq, 1)
    prim = np.stack([1.0 + 0.2 * np.sin(2 * np.pi * xq[..., 0]), 0.4 + 0 * xq[..., 0],
                     1.0 + 0.1 * np.cos(2 * np.pi * xq[..., 0])], axis=-1)
    o.set_initial_primitive(prim)
    r = o.residual()
    one = P.project_conserved(m, np.ones((m.E, m.nq, 3)))[:, :, 0]
    assert np.abs(np.einsum("eb,ebv->v", one, r)).max() < 1e-13 * np.abs(r).max() * m.E
