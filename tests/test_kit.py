"""Operand kit (quadrature, Lobatto basis, structured periodic meshes, affine geometry).

Pinned against numpy (Gauss-Legendre, Legendre polynomials), closed forms for uniform affine
meshes, and the reference's own right-parent point-sequence tables
(getAdjacencyElementQuadratureSequence, SimulationControl.cpp:381-523) as golden vectors.
"""
import itertools
import math

import numpy as np
import numpy.polynomial.legendre as npleg
import pytest

from subrosadg_amd import _lib as L


def lobatto_np(k, x):
    """Solin's Lobatto shape functions from numpy's Legendre polynomials."""
    if k == 0:
        return (1 - x) / 2
    if k == 1:
        return (1 + x) / 2
    ck = np.zeros(k + 1)
    ck[k] = 1
    ck2 = np.zeros(k + 1)
    ck2[k - 2] = 1
    return (npleg.legval(x, ck) - npleg.legval(x, ck2)) / math.sqrt(2 * (2 * k - 1))


def lobatto_np_deriv(k, x):
    if k == 0:
        return -0.5 * np.ones_like(x)
    if k == 1:
        return 0.5 * np.ones_like(x)
    ck = np.zeros(k + 1)
    ck[k - 1] = 1
    return math.sqrt((2 * k - 1) / 2) * npleg.legval(x, ck)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_quad_rules_match_numpy_gauss_legendre(order):
    m = L.Mesh(L.ELEM_QUADRANGLE, order, (2, 2), (0, 0), (1, 1))
    x, w = npleg.leggauss(order + 1)
    pts = m.field("quad_point").reshape(-1, 2)
    wts = m.field("quad_weight")
    expect = np.array([(x[a], x[b]) for b in range(order + 1) for a in range(order + 1)])
    np.testing.assert_allclose(pts, expect, atol=1e-15)
    np.testing.assert_allclose(wts, np.outer(w, w).T.reshape(-1), rtol=1e-14)
    np.testing.assert_allclose(m.field("face_quad_point"), x, atol=1e-15)
    np.testing.assert_allclose(m.field("face_quad_weight"), w, rtol=1e-14)
    # exactness: volume rule order 2P, face rule order 2P+1 (SimulationControl.cpp:275-283)
    for a, b in itertools.product(range(2 * order + 2), repeat=2):
        exact = (2 / (a + 1) if a % 2 == 0 else 0) * (2 / (b + 1) if b % 2 == 0 else 0)
        got = np.sum(wts * pts[:, 0] ** a * pts[:, 1] ** b)
        assert abs(got - exact) < 1e-13


@pytest.mark.parametrize("order", [1, 2, 3])
def test_hex_rule_counts(order):
    m = L.Mesh(L.ELEM_HEXAHEDRON, order, (2, 2, 2), (0, 0, 0), (1, 1, 1))
    # counts of SimulationControl.cpp:268-273 at orders 2P (hex) and 2P+1 (quad faces)
    assert m.nq == (order + 1) ** 3 and m.nfq == (order + 1) ** 2 and m.nb == (order + 1) ** 3


@pytest.mark.parametrize("elem,order", [(L.ELEM_QUADRANGLE, p) for p in (1, 2, 3, 4)] +
                         [(L.ELEM_HEXAHEDRON, p) for p in (1, 2, 3)])
def test_basis_is_tensor_lobatto(elem, order):
    """Every basis column is a product of Lobatto functions, each product appears once, and the
    gradient table is the derivative of the same product."""
    dim = 2 if elem == L.ELEM_QUADRANGLE else 3
    m = L.Mesh(elem, order, (2,) * dim, (0,) * dim, (1,) * dim)
    pts = m.field("quad_point").reshape(-1, dim)
    V = m.field("basis_value").reshape(m.nq, m.nb)
    dV = m.field("basis_gradient").reshape(m.nq, dim, m.nb)
    cands = {}
    for idx in itertools.product(range(order + 1), repeat=dim):
        val = np.prod([lobatto_np(idx[d], pts[:, d]) for d in range(dim)], axis=0)
        cands[idx] = val
    used = set()
    for b in range(m.nb):
        match = [k for k, v in cands.items() if np.allclose(V[:, b], v, atol=1e-13)]
        assert len(match) == 1
        idx = match[0]
        used.add(idx)
        for d in range(dim):
            g = np.prod([lobatto_np_deriv(idx[e], pts[:, e]) if e == d else lobatto_np(idx[e], pts[:, e])
                         for e in range(dim)], axis=0)
            np.testing.assert_allclose(dV[:, d, b], g, atol=1e-12)
    assert len(used) == m.nb
    # hierarchy: the first N_b(P-1) functions use only 1-D indices <= P-1
    if order > 1:
        first = order ** dim
        for b in range(first):
            idx = [k for k, v in cands.items() if np.allclose(V[:, b], v, atol=1e-13)][0]
            assert max(idx) <= order - 1


def test_lsq_inverse_and_projection_reproduce_polynomials():
    order = 3
    m = L.Mesh(L.ELEM_QUADRANGLE, order, (3, 3), (-1, -1), (2, 2))
    V = m.field("basis_value").reshape(m.nq, m.nb)
    lsq = m.field("basis_lsq_inverse").reshape(m.nb, m.nb)
    np.testing.assert_allclose(lsq @ (V.T @ V), np.eye(m.nb), atol=1e-11)
    xq = m.field("quad_coord").reshape(m.E, m.nq, 2)
    f = lambda x, y: 1 + x ** 3 * y ** 2 - 2 * x * y ** 3 + y  # degree <= 3 per variable
    Vf = m.field("basis_face_value").reshape(m.nf * m.nfq, m.nb)
    fx = m.field("face_coord").reshape(m.F, m.nfq, 2)
    par = m.field("face_parent").reshape(-1, 2)
    loc = m.field("face_local").reshape(-1, 2)
    for e in range(m.E):
        c = f(xq[e, :, 0], xq[e, :, 1]) @ V @ lsq
        np.testing.assert_allclose(V @ c, f(xq[e, :, 0], xq[e, :, 1]), rtol=1e-12, atol=1e-12)
    # traces at the face points of the left parent equal f at the physical face points
    for fc in range(m.F):
        e, s = par[fc, 0], loc[fc, 0]
        c = f(xq[e, :, 0], xq[e, :, 1]) @ V @ lsq
        tr = Vf[s * m.nfq:(s + 1) * m.nfq] @ c
        np.testing.assert_allclose(tr, f(fx[fc, :, 0], fx[fc, :, 1]), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("elem", [L.ELEM_QUADRANGLE, L.ELEM_HEXAHEDRON])
def test_uniform_geometry_closed_forms(elem):
    dim = 2 if elem == L.ELEM_QUADRANGLE else 3
    n, lo, hi, order = 4, -1.0, 3.0, 2
    h = (hi - lo) / n
    m = L.Mesh(elem, order, (n,) * dim, (lo,) * dim, (hi,) * dim)
    w = m.field("quad_weight")
    detjw = m.field("jacobian_det_w").reshape(m.E, m.nq)
    np.testing.assert_allclose(detjw, np.tile((h / 2) ** dim * w, (m.E, 1)), rtol=1e-14)
    tiw = m.field("jacobian_tiw").reshape(m.E, m.nq, dim, dim)
    for q in range(m.nq):
        np.testing.assert_allclose(tiw[:, q], np.broadcast_to((2 / h) * (h / 2) ** dim * w[q] * np.eye(dim),
                                                              (m.E, dim, dim)), rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(m.field("min_edge"), h, rtol=1e-14)
    V = m.field("basis_value").reshape(m.nq, m.nb)
    minv = m.field("mass_inverse").reshape(m.E, m.nb, m.nb)
    np.testing.assert_allclose(minv[0], np.linalg.inv(V.T @ np.diag(detjw[0]) @ V), rtol=1e-10, atol=1e-10)
    fw = m.field("face_quad_weight")
    np.testing.assert_allclose(m.field("face_jacobian_w").reshape(m.F, m.nfq),
                               np.tile((h / 2) ** (dim - 1) * fw, (m.F, 1)), rtol=1e-14)
    # normals: unit, axis-aligned, outward from the left parent
    nrm = m.field("face_normal").reshape(m.F, m.nfq, dim)
    np.testing.assert_allclose(np.linalg.norm(nrm, axis=2), 1.0, rtol=1e-14)
    xq = m.field("quad_coord").reshape(m.E, m.nq, dim)
    fx = m.field("face_coord").reshape(m.F, m.nfq, dim)
    par = m.field("face_parent").reshape(-1, 2)
    centre = xq.mean(axis=1)
    for f in range(m.F):
        assert np.all(np.einsum("jd,jd->j", nrm[f], fx[f] - centre[par[f, 0]]) > 0)
        assert np.allclose(np.abs(nrm[f]).max(axis=1), 1.0)


@pytest.mark.parametrize("elem", [L.ELEM_QUADRANGLE, L.ELEM_HEXAHEDRON])
def test_periodic_face_pairing(elem):
    dim = 2 if elem == L.ELEM_QUADRANGLE else 3
    m = L.Mesh(elem, 2, (3, 4, 2)[:dim], (0,) * dim, (1.5, 2.0, 1.0)[:dim])
    period = np.array((1.5, 2.0, 1.0)[:dim])
    par = m.field("face_parent").reshape(-1, 2)
    loc = m.field("face_local").reshape(-1, 2)
    seq = m.field("face_sequence").reshape(-1, m.nfq)
    assert m.F * 2 == m.E * m.nf  # every local face is paired exactly once
    slots = {(int(e), int(s)) for p, l in zip(par, loc) for e, s in zip(p, l)}
    assert len(slots) == m.E * m.nf
    assert np.all(par[:, 0] <= par[:, 1])  # left = first seen in element order
    # matched points coincide modulo the period, seen from either parent
    Vf = m.field("basis_face_value").reshape(m.nf * m.nfq, m.nb)
    V = m.field("basis_value").reshape(m.nq, m.nb)
    lsq = m.field("basis_lsq_inverse").reshape(m.nb, m.nb)
    xq = m.field("quad_coord").reshape(m.E, m.nq, dim)
    for f in range(m.F):
        pts = []
        for side in range(2):
            e, s = par[f, side], loc[f, side]
            c = xq[e] .T @ V @ lsq  # physical coordinates are degree-1: projection is exact
            pts.append(Vf[s * m.nfq:(s + 1) * m.nfq] @ c.T)
        d = pts[0] - pts[1][seq[f]]
        d -= period * np.round(d / period)
        assert np.abs(d).max() < 1e-12
        assert sorted(seq[f]) == list(range(m.nfq))


# Right-parent point sequences of the reference for quadrangle faces at rotations 0 and 1
# (SimulationControl.cpp:454-505); a structured hexahedral mesh produces these two rotations.
REFERENCE_QUAD_FACE_SEQUENCE = {
    1: {0: [0, 2, 1, 3], 1: [2, 3, 0, 1]},
    2: {0: [0, 3, 6, 1, 4, 7, 2, 5, 8], 1: [6, 7, 8, 3, 4, 5, 0, 1, 2]},
    3: {0: [0, 4, 8, 12, 1, 5, 9, 13, 2, 6, 10, 14, 3, 7, 11, 15],
        1: [12, 13, 14, 15, 8, 9, 10, 11, 4, 5, 6, 7, 0, 1, 2, 3]},
    4: {0: [0, 5, 10, 15, 20, 1, 6, 11, 16, 21, 2, 7, 12, 17, 22, 3, 8, 13, 18, 23, 4, 9, 14, 19, 24],
        1: [20, 21, 22, 23, 24, 15, 16, 17, 18, 19, 10, 11, 12, 13, 14, 5, 6, 7, 8, 9, 0, 1, 2, 3, 4]},
}


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_face_sequence_matches_reference_tables(order):
    m = L.Mesh(L.ELEM_HEXAHEDRON, order, (3, 3, 3), (0, 0, 0), (1, 1, 1))
    seq = m.field("face_sequence").reshape(-1, m.nfq)
    rot = m.field("face_rotation")
    assert set(rot.tolist()) == {0, 1}
    for s, r in zip(seq, rot):
        assert s.tolist() == REFERENCE_QUAD_FACE_SEQUENCE[order][int(r)]
    # line faces: always reversed (SimulationControl.cpp:386-402)
    m2 = L.Mesh(L.ELEM_QUADRANGLE, order, (3, 3), (0, 0), (1, 1))
    for s in m2.field("face_sequence").reshape(-1, m2.nfq):
        assert s.tolist() == list(range(order, -1, -1))


@pytest.mark.parametrize("args", [
    (L.ELEM_QUADRANGLE, 2, (1, 4), (0, 0), (1, 1)),   # fewer than 2 cells
    (L.ELEM_QUADRANGLE, 0, (4, 4), (0, 0), (1, 1)),   # order 0
    (L.ELEM_QUADRANGLE, 6, (4, 4), (0, 0), (1, 1)),   # order 6
    (L.ELEM_QUADRANGLE, 2, (4, 4), (0, 0), (0, 1)),   # empty box
    (L.ELEM_TRIANGLE, 2, (4, 4), (0, 0), (1, 1)),     # structured triangles not provided
])
def test_bad_mesh_descriptors_raise(args):
    with pytest.raises(L.SDGError):
        L.Mesh(*args)
