"""Pyramids: the collapsed Gauss-Jacobi volume rule (exactness), the modal basis of gmsh's pyramid
space (it contains the complete polynomials of degree P; (P+1)(P+2)(2P+3)/6 functions), the faces
(triangle faces carry the tetrahedra's face rule padded with zero-weight slots to the (P+1)^2 of the
quadrangle base; the right-parent points of both shapes against the reference's tables), the
geometry of jittered pyramid meshes (volumes by the divergence theorem, closed surfaces), the
artificial-viscosity sensor's prefix, and the oracle's free stream and conservation."""
import math

import numpy as np
import pytest

import oracle as O
import problems as P
from subrosadg_amd import _lib as L
from test_kit import REFERENCE_QUAD_FACE_SEQUENCE
from test_kit_tet import ROTATION


def _exact(a, b, c):
    """int over the reference pyramid (base [-1,1]^2 at z = 0, apex (0,0,1)) of x^a y^b z^c."""
    if a % 2 or b % 2:
        return 0.0
    return (2.0 / (a + 1)) * (2.0 / (b + 1)) * math.factorial(c) * math.factorial(a + b + 2) / math.factorial(a + b + c + 3)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_pyramid_rule_and_basis(order):
    """(P+1)^3 points exact for x^a y^b z^c, a + b + c <= 2P + 1; the basis has (P+1)(P+2)(2P+3)/6
    functions, full rank, reproduces every polynomial of degree <= P with its exact gradient, and
    its mass matrix is well conditioned."""
    m = P.pyramid_box(n=1, order=order, periodic=False)
    assert m.nq == (order + 1) ** 3 and m.nfq == (order + 1) ** 2
    assert m.nb == (order + 1) * (order + 2) * (2 * order + 3) // 6
    w = m.field("quad_weight")
    x = m.field("quad_point").reshape(-1, 3)
    assert np.all(x[:, 2] > 0) and np.all(x[:, 2] < 1) and np.all(np.abs(x[:, :2]) < 1 - x[:, 2:3])
    for a in range(2 * order + 2):
        for b in range(2 * order + 2 - a):
            for c in range(2 * order + 2 - a - b):
                got = (w * x[:, 0] ** a * x[:, 1] ** b * x[:, 2] ** c).sum()
                assert abs(got - _exact(a, b, c)) <= 1e-14, (a, b, c)
    V = m.field("basis_value").reshape(m.nq, m.nb)
    dV = m.field("basis_gradient").reshape(m.nq, 3, m.nb)
    lsq = m.field("basis_lsq_inverse").reshape(m.nb, m.nb)
    assert np.linalg.matrix_rank(V) == m.nb
    assert np.linalg.cond(V.T @ (w[:, None] * V)) < 1e3
    rng = np.random.default_rng(order)
    terms = [(a, b, c) for a in range(order + 1) for b in range(order + 1 - a) for c in range(order + 1 - a - b)]
    coef = rng.standard_normal(len(terms))
    f = sum(k * x[:, 0] ** a * x[:, 1] ** b * x[:, 2] ** c for k, (a, b, c) in zip(coef, terms))
    C = f @ V @ lsq
    np.testing.assert_allclose(V @ C, f, atol=1e-12)
    for d in range(3):
        g = np.zeros(len(x))
        for k, e in zip(coef, terms):
            if e[d]:
                t = k * e[d]
                for dd in range(3):
                    t = t * x[:, dd] ** (e[dd] - (1 if dd == d else 0))
                g += t
        np.testing.assert_allclose(dV[:, d, :] @ C, g, atol=1e-11)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_pyramid_faces(order):
    """Triangle faces: the tetrahedra's face rule in the first slots, the rest zero weight; the
    right-parent points of interior faces equal the reference's rotation tables (triangles between
    the pyramids of a box, quadrangle bases between boxes), padding slots map to themselves."""
    m = P.pyramid_box(n=2, order=order, jitter=0.15)
    ntri = [4, 7, 13, 19][order - 1]
    loc = m.field("face_local").reshape(m.F, 2)
    rot = m.field("face_rotation")
    seq = m.field("face_sequence").reshape(m.F, m.nfq)
    fjw = m.field("face_jacobian_w").reshape(m.F, m.nfq)
    seen = set()
    for f in range(m.F):
        if loc[f, 0] < 4:  # triangle
            assert loc[f, 1] < 4
            assert list(seq[f, :ntri]) == ROTATION[order][rot[f]], (f, rot[f])
            assert list(seq[f, ntri:]) == list(range(ntri, m.nfq))
            assert np.all(fjw[f, ntri:] == 0.0) and np.all(fjw[f, :ntri] != 0.0)
            seen.add(("tri", int(rot[f])))
        else:
            assert loc[f, 1] == 4
            table = REFERENCE_QUAD_FACE_SEQUENCE[order].get(int(rot[f]))
            if table is not None:
                assert list(seq[f]) == table
            seen.add(("quad", int(rot[f])))
    assert any(k[0] == "tri" for k in seen) and ("quad", 0) in seen


@pytest.mark.parametrize("order", [1, 3])
def test_pyramid_geometry(order):
    """Volumes by the divergence theorem (|e| = (1/3) surface integral of x . n) and closed surfaces
    on a jittered bounded mesh; volumes sum to the box."""
    m = P.pyramid_box(n=2, order=order, periodic=False, jitter=0.2, hi=(2.0, 1.0, 1.5))
    vol = m.field("jacobian_det_w").reshape(m.E, m.nq).sum(1)
    assert abs(vol.sum() - 3.0) < 1e-13
    xn = np.zeros(m.E)
    ns = np.zeros((m.E, 3))
    fc = m.field("face_coord").reshape(m.F, m.nfq, 3)
    fn = m.field("face_normal").reshape(m.F, m.nfq, 3) * m.field("face_jacobian_w").reshape(m.F, m.nfq, 1)
    par = m.field("face_parent").reshape(m.F, 2)
    np.add.at(xn, par[:, 0], (fc * fn).sum((1, 2)))
    np.add.at(xn, par[:, 1], -(fc * fn).sum((1, 2)))
    np.add.at(ns, par[:, 0], fn.sum(1))
    np.add.at(ns, par[:, 1], -fn.sum(1))
    bc = m.field("boundary_face_coord").reshape(m.Fb, m.nfq, 3)
    bn = m.field("boundary_face_normal").reshape(m.Fb, m.nfq, 3) * m.field("boundary_face_jacobian_w").reshape(m.Fb, m.nfq, 1)
    bp = m.field("boundary_face_parent").reshape(m.Fb, 2)[:, 0]
    np.add.at(xn, bp, (bc * bn).sum((1, 2)))
    np.add.at(ns, bp, bn.sum(1))
    np.testing.assert_allclose(vol, xn / 3.0, rtol=1e-12)
    assert np.abs(ns).max() < 1e-14


@pytest.mark.parametrize("visc", [None, L.VISC_BR2])
def test_oracle_free_stream_and_conservation_on_pyramids(visc):
    m = P.pyramid_box(n=2, order=2, jitter=0.15)
    if visc is None:
        prob = L.problem_desc(equation=L.EQ_EULER, convective_flux=L.FLUX_ROE, cp=3.5, cv=2.5)
    else:
        prob = L.problem_desc(equation=L.EQ_NS, convective_flux=L.FLUX_HLLC, viscous_flux=visc, cp=3.5, cv=2.5,
                              transport=L.TRANSPORT_CONSTANT, dynamic_viscosity=0.05)
    o = O.Oracle(prob, m)
    u = np.array([1.0, 0.3, -0.2, 0.1, 1.0])
    o.set_initial_primitive(np.tile(u, (m.E, m.nq, 1)))
    assert np.abs(o.residual()).max() < 1e-12
    xq = m.field("quad_coord").reshape(m.E, m.nq, 3)
    prim = np.tile(u, (m.E, m.nq, 1))
    prim[..., 0] += 0.2 * np.sin(2 * np.pi * xq[..., 0]) * np.cos(2 * np.pi * xq[..., 2])
    prim[..., 4] += 0.1 * np.sin(2 * np.pi * (xq[..., 0] + xq[..., 1]))
    o.set_initial_primitive(prim)
    r = o.residual()
    one = P.project_conserved(m, np.ones((m.E, m.nq, 5)))[:, :, 0]
    assert np.abs(np.einsum("eb,ebv->v", one, r)).max() < 1e-12 * np.abs(r).max() * m.E


@pytest.mark.parametrize("order", [2, 3])
def test_sensor_prefix_on_pyramids(order):
    """The modes above order P-1 start at P(P+1)(2P+1)/6: a density of only the first of them gives
    the full viscosity f r / P on the element's nodes."""
    import test_oracle_av as A
    m = P.pyramid_box(n=1, order=order, periodic=False)
    k, f = 0.3, 2.0
    low = order * (order + 1) * (2 * order + 1) // 6
    tag = m.field("node_tag").reshape(m.E, 5)
    for mode in (low, m.nb - 1):
        C = A._conserved(m, A._uniform_state(m))
        C[2, :, 0] = 0.0
        C[2, mode, 0] = 0.7
        o = A._oracle_with_state(m, A._av_problem(k, f), C)
        want = np.zeros(m.nodes)
        want[tag[2]] = f * m.field("inner_radius")[2] / order
        assert np.allclose(o.artificial_viscosity(), want, rtol=1e-13, atol=0.0), mode
