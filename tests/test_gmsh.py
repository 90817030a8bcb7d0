"""gmsh MSH 4.1 reader (sdg_mesh_create_gmsh; the reference reads meshes through gmsh,
ReadControl.cpp:286-335): a file holding a structured grid in the structured builder's order gives
bit-identical operand tables (quadrangles and hexahedra); shuffled node tags, shuffled and mirrored
elements and cyclically rotated vertex lists give the same element integrals of the oracle residual
(matched by centroid); $Periodic node pairs give the structured builder's periodic tables; an
unstructured triangulation preserves the free stream; unsupported input is refused. GPU parity on a file mesh is in test_gpu_gmsh.py."""
import numpy as np
import pytest

import msh_writer as W
import oracle as O
import problems as P
from subrosadg_amd import _lib as L

FAR, SLIP = L.BC_RIEMANN_FARFIELD, L.BC_ADIABATIC_SLIP_WALL
FIELDS = ["jacobian_det_w", "jacobian_tiw", "mass_inverse", "min_edge", "face_normal", "face_jacobian_w",
          "face_coord", "face_parent", "face_local", "face_sequence", "face_rotation", "boundary_face_coord",
          "boundary_face_normal", "boundary_face_jacobian_w", "boundary_face_parent", "boundary_face_type",
          "inner_radius", "node_tag", "quad_coord"]


def _same_tables(a, b):
    assert (a.E, a.F, a.Fb, a.nodes) == (b.E, b.F, b.Fb, b.nodes)
    for f in FIELDS:
        assert np.array_equal(a.field(f), b.field(f)), f


def test_quadrangle_file_equals_the_structured_builder(tmp_path):
    nx, ny = 5, 3
    tags, xyz, elems, sides = W.grid_quads(nx, ny, (0.0, 0.0), (2.0, 1.0))
    path = str(tmp_path / "q.msh")
    W.write_msh(path, 2, tags, xyz, "quad", elems, [(1, "line", sides[0]), (1, "line", sides[1]),
                                                     (2, "line", sides[2]), (2, "line", sides[3])])
    g = L.Mesh.gmsh(path, 3, {1: FAR, 2: SLIP})
    s = L.Mesh(L.ELEM_QUADRANGLE, 3, (nx, ny), (0.0, 0.0), (2.0, 1.0), boundary=((FAR, FAR), (SLIP, SLIP)))
    _same_tables(g, s)


def test_hexahedron_file_equals_the_structured_builder(tmp_path):
    n = 2
    tags, xyz, elems, sides = W.grid_hexes(n)
    path = str(tmp_path / "h.msh")
    kinds = [FAR, FAR, SLIP, SLIP, L.BC_ADIABATIC_NONSLIP_WALL, L.BC_ADIABATIC_NONSLIP_WALL]
    W.write_msh(path, 3, tags, xyz, "hex", elems, [(10 + k, "bquad", sides[k]) for k in range(6)])
    g = L.Mesh.gmsh(path, 2, {10 + k: kinds[k] for k in range(6)})
    s = L.Mesh(L.ELEM_HEXAHEDRON, 2, (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0),
               boundary=((FAR, FAR), (SLIP, SLIP), (L.BC_ADIABATIC_NONSLIP_WALL,) * 2))
    _same_tables(g, s)


def test_periodic_file_equals_the_structured_builder(tmp_path):
    """$Periodic links (x = 1 onto x = 0, y = 1 onto y = 0, the corner a slave in both) give the
    structured builder's periodic tables bit for bit."""
    nx, ny = 4, 3
    tags, xyz, elems, _ = W.grid_quads(nx, ny)
    tag = lambda i, j: 1 + i + (nx + 1) * j  # noqa: E731
    pairs = [[(tag(nx, j), tag(0, j)) for j in range(ny + 1)], [(tag(i, ny), tag(i, 0)) for i in range(nx + 1)]]
    path = str(tmp_path / "p.msh")
    W.write_msh(path, 2, tags, xyz, "quad", elems, [], periodic=pairs)
    g = L.Mesh.gmsh(path, 3, {})
    s = L.Mesh(L.ELEM_QUADRANGLE, 3, (nx, ny), (0.0, 0.0), (1.0, 1.0))
    _same_tables(g, s)


def test_periodic_hexahedra_with_walls(tmp_path):
    n = 2
    tags, xyz, elems, sides = W.grid_hexes(n)
    tag = lambda i, j, k: 1 + i + (n + 1) * (j + (n + 1) * k)  # noqa: E731
    pairs = [(tag(i, j, n), tag(i, j, 0)) for j in range(n + 1) for i in range(n + 1)]  # z periodic
    path = str(tmp_path / "ph.msh")
    W.write_msh(path, 3, tags, xyz, "hex", elems, [(10 + k, "bquad", sides[k]) for k in range(4)], periodic=pairs)
    g = L.Mesh.gmsh(path, 2, {10: FAR, 11: FAR, 12: SLIP, 13: SLIP})
    s = L.Mesh(L.ELEM_HEXAHEDRON, 2, (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0),
               boundary=((FAR, FAR), (SLIP, SLIP), None))
    _same_tables(g, s)


def _euler(flux=L.FLUX_HLLC):
    return L.problem_desc(equation=L.EQ_EULER, convective_flux=flux, cp=3.5, cv=2.5)


def _state(xq):
    prim = np.zeros(xq.shape[:-1] + (4,))
    prim[..., 0] = 1.0 + 0.1 * np.sin(2 * np.pi * xq[..., 0]) * np.cos(np.pi * xq[..., 1])
    prim[..., 1] = 0.3 + 0.05 * np.cos(np.pi * xq[..., 1])
    prim[..., 2] = 0.1 * np.sin(np.pi * xq[..., 0])
    prim[..., 3] = 1.0 + 0.05 * np.cos(2 * np.pi * xq[..., 0])
    return prim


def _element_integrals(mesh, prob, bstate=(1.0, 0.3, 0.0, 1.0)):
    """Per element the residual against the constant test function (frame independent), with the
    element centroids."""
    xq = mesh.field("quad_coord").reshape(mesh.E, mesh.nq, 2)
    o = O.Oracle(prob, mesh)
    o.set_initial_primitive(_state(xq))
    o.set_boundary_primitive(np.tile(np.asarray(bstate), (mesh.Fb, mesh.nfq, 1)))
    r = o.residual()
    one = P.project_conserved(mesh, np.ones((mesh.E, mesh.nq, 4)))[:, :, 0]
    jw = mesh.field("jacobian_det_w").reshape(mesh.E, mesh.nq)
    cen = (xq * jw[..., None]).sum(1) / jw.sum(1)[:, None]
    return np.einsum("eb,ebv->ev", one, r), cen


def test_shuffled_and_mirrored_quadrangles_give_the_same_element_residuals(tmp_path):
    nx, ny = 6, 4
    tags, xyz, elems, sides = W.grid_quads(nx, ny)
    rng = np.random.default_rng(5)
    new_tag = rng.permutation(len(tags)) * 3 + 7  # sparse, shuffled node tags
    remap = dict(zip(tags, new_tag))
    elems = np.vectorize(remap.get)(elems)
    sides = [np.vectorize(remap.get)(s) for s in sides]
    order = rng.permutation(len(elems))
    elems = elems[order]
    for e in range(len(elems)):
        elems[e] = np.roll(elems[e], rng.integers(4))
        if rng.random() < 0.5:
            elems[e] = elems[e][::-1]  # clockwise: mirrored by the reader
    sides = [s[:, ::-1] if k % 2 else s for k, s in enumerate(sides)]
    path = str(tmp_path / "s.msh")
    W.write_msh(path, 2, new_tag, xyz, "quad", elems, [(1, "line", sides[0]), (1, "line", sides[1]),
                                                        (2, "line", sides[2]), (2, "line", sides[3])])
    g = L.Mesh.gmsh(path, 3, {1: FAR, 2: SLIP})
    s = L.Mesh(L.ELEM_QUADRANGLE, 3, (nx, ny), (0.0, 0.0), (1.0, 1.0), boundary=((FAR, FAR), (SLIP, SLIP)))
    assert (g.E, g.F, g.Fb, g.nodes) == (s.E, s.F, s.Fb, s.nodes)
    assert np.array_equal(np.sort(g.field("boundary_face_type")), np.sort(s.field("boundary_face_type")))
    rg, cg = _element_integrals(g, _euler())
    rs, cs = _element_integrals(s, _euler())
    key = lambda c: np.lexsort((np.round(c[:, 0], 9), np.round(c[:, 1], 9)))  # noqa: E731
    ig, is_ = key(cg), key(cs)
    assert np.allclose(cg[ig], cs[is_], atol=1e-13)
    assert np.allclose(rg[ig], rs[is_], rtol=1e-11, atol=1e-12 * np.abs(rs).max())


def triangle_file(path, n=6, seed=3, bc=(1, 1, 1, 1), periodic=False):
    """Unstructured-like triangulation of the unit square: an (n+1)^2 node grid with jittered
    interior nodes, each cell cut along a random diagonal; sides in physical groups bc, or, with
    periodic=True, $Periodic links x = 1 -> x = 0 and y = 1 -> y = 0 and no boundary groups."""
    tags, xyz, quads, sides = W.grid_quads(n, n)
    rng = np.random.default_rng(seed)
    inner = (xyz[:, 0] > 0) & (xyz[:, 0] < 1) & (xyz[:, 1] > 0) & (xyz[:, 1] < 1)
    xyz = xyz.copy()
    xyz[inner, :2] += rng.uniform(-0.2, 0.2, (inner.sum(), 2)) / n
    tris = []
    for q in quads:
        a, b, c, d = q
        tris += [[a, b, c], [a, c, d]] if rng.random() < 0.5 else [[a, b, d], [b, c, d]]
    if periodic:
        tag = lambda i, j: 1 + i + (n + 1) * j  # noqa: E731
        links = [[(tag(n, j), tag(0, j)) for j in range(n + 1)], [(tag(i, n), tag(i, 0)) for i in range(n + 1)]]
        W.write_msh(path, 2, tags, xyz, "tri", np.array(tris), [], periodic=links)
        return
    W.write_msh(path, 2, tags, xyz, "tri", np.array(tris), [(bc[k], "line", sides[k]) for k in range(4)])


@pytest.mark.parametrize("order", [2, 3])
def test_unstructured_triangles_free_stream_and_conservation(tmp_path, order):
    path = str(tmp_path / "t.msh")
    triangle_file(path)
    m = L.Mesh.gmsh(path, order, {1: FAR})
    assert m.elem == L.ELEM_TRIANGLE and m.E == 72 and m.Fb == 24
    assert abs(m.field("jacobian_det_w").sum() - 1.0) < 1e-13
    u = (1.0, 0.3, 0.2, 1.0)
    o = O.Oracle(_euler(), m)
    o.set_initial_primitive(np.tile(np.asarray(u), (m.E, m.nq, 1)))
    o.set_boundary_primitive(np.tile(np.asarray(u), (m.Fb, m.nfq, 1)))
    assert np.abs(o.residual()).max() < 1e-12  # free stream through the far field
    # node tags: the (n+1)^2 grid nodes, shared by the triangles meeting there
    assert m.nodes == 49 and np.bincount(m.field("node_tag")).min() >= 1


def test_periodic_triangulation_free_stream_and_conservation(tmp_path):
    path = str(tmp_path / "pt.msh")
    triangle_file(path, periodic=True)
    m = L.Mesh.gmsh(path, 3, {})
    assert m.Fb == 0 and m.F == 3 * m.E // 2
    u = (1.0, 0.3, 0.2, 1.0)
    o = O.Oracle(_euler(), m)
    o.set_initial_primitive(np.tile(np.asarray(u), (m.E, m.nq, 1)))
    assert np.abs(o.residual()).max() < 1e-12
    xq = m.field("quad_coord").reshape(m.E, m.nq, 2)
    o.set_initial_primitive(_state(xq))
    r = o.residual()
    one = P.project_conserved(m, np.ones((m.E, m.nq, 4)))[:, :, 0]
    assert np.abs(np.einsum("eb,ebv->v", one, r)).max() < 1e-12 * np.abs(r).max() * m.E


def test_refusals(tmp_path):
    tags, xyz, elems, sides = W.grid_quads(3, 3)
    groups = [(1, "line", s) for s in sides]
    p = str(tmp_path / "per.msh")
    W.write_msh(p, 2, tags, xyz, "quad", elems, groups, periodic=[(1, 999)])
    with pytest.raises(L.SDGError, match="unknown node"):
        L.Mesh.gmsh(p, 2, {1: FAR})
    p = str(tmp_path / "serendipity.msh")
    W.write_msh(p, 2, tags, xyz, "quad", elems, groups, elem_type_code=16)  # 8-node serendipity quadrangle
    with pytest.raises(L.SDGError, match="element type"):
        L.Mesh.gmsh(p, 2, {1: FAR})
    p = str(tmp_path / "nonaffine.msh")
    bent = xyz.copy()
    bent[5, 0] += 0.05  # an interior node off the grid: its four cells are no longer parallelograms
    W.write_msh(p, 2, tags, bent, "quad", elems, groups)
    assert L.Mesh.gmsh(p, 2, {1: FAR}).info.affine == 0  # read with their bilinear maps
    p = str(tmp_path / "ok.msh")
    W.write_msh(p, 2, tags, xyz, "quad", elems, groups)
    with pytest.raises(L.SDGError, match="boundary condition"):
        L.Mesh.gmsh(p, 2, {})
    with pytest.raises(L.SDGError, match="cannot open"):
        L.Mesh.gmsh(str(tmp_path / "none.msh"), 2, {1: FAR})


def test_boundary_elements_that_are_not_faces_are_refused(tmp_path):
    """A hexahedron file whose wall group holds triangles (3 nodes against the 4 of a hexahedron
    face) is refused with an error naming the group, instead of reading past the node list."""
    tags, xyz, elems, sides = W.grid_hexes(2)
    groups = [(10 + k, "bquad", sides[k]) for k in range(5)]
    q = sides[5]
    tris = np.concatenate([q[:, [0, 1, 2]], q[:, [0, 2, 3]]])
    p = str(tmp_path / "tri_wall.msh")
    W.write_msh(p, 3, tags, xyz, "hex", elems, groups + [(15, "btri", tris)])
    with pytest.raises(L.SDGError, match="physical group 15"):
        L.Mesh.gmsh(p, 2, {10 + k: FAR for k in range(6)})


def test_array_meshes_equal_the_structured_builder():
    """sdg_mesh_create_from_arrays (a mesh handed over as host arrays, e.g. read out of a gmsh model)
    gives the structured builder's tables bit for bit: quadrangles with boundaries, periodic
    quadrangles, hexahedra with boundaries."""
    nx, ny = 5, 3
    tags, xyz, elems, sides = W.grid_quads(nx, ny, (0.0, 0.0), (2.0, 1.0))
    bf = np.concatenate(sides) - 1
    bt = [FAR] * (2 * ny) + [SLIP] * (2 * nx)
    a = L.Mesh.from_arrays(L.ELEM_QUADRANGLE, 3, xyz, elems - 1, bf, bt, [1] * (2 * ny) + [2] * (2 * nx))
    s = L.Mesh(L.ELEM_QUADRANGLE, 3, (nx, ny), (0.0, 0.0), (2.0, 1.0), boundary=((FAR, FAR), (SLIP, SLIP)))
    _same_tables(a, s)
    # the boundary faces come in element order; each keeps the physical tag it was given with
    assert np.array_equal(a.field("boundary_face_physical"), np.where(a.field("boundary_face_type") == FAR, 1, 2))
    # periodic in both directions through node pairs (the corner a slave in both links)
    nx, ny = 4, 3
    tags, xyz, elems, _ = W.grid_quads(nx, ny)
    node = lambda i, j: i + (nx + 1) * j  # noqa: E731
    pairs = [[(node(nx, j), node(0, j)) for j in range(ny + 1)], [(node(i, ny), node(i, 0)) for i in range(nx + 1)]]
    a = L.Mesh.from_arrays(L.ELEM_QUADRANGLE, 2, xyz[:, :2], elems - 1, periodic=pairs)
    _same_tables(a, L.Mesh(L.ELEM_QUADRANGLE, 2, (nx, ny), (0.0, 0.0), (1.0, 1.0)))
    n = 2
    tags, xyz, elems, sides = W.grid_hexes(n)
    kinds = [FAR, FAR, SLIP, SLIP, L.BC_ADIABATIC_NONSLIP_WALL, L.BC_ADIABATIC_NONSLIP_WALL]
    bt = np.concatenate([[kinds[k]] * len(sides[k]) for k in range(6)])
    a = L.Mesh.from_arrays(L.ELEM_HEXAHEDRON, 2, xyz, elems - 1, np.concatenate(sides) - 1, bt)
    s = L.Mesh(L.ELEM_HEXAHEDRON, 2, (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0),
               boundary=((FAR, FAR), (SLIP, SLIP), (L.BC_ADIABATIC_NONSLIP_WALL,) * 2))
    _same_tables(a, s)


def test_array_mesh_refusals():
    tags, xyz, elems, sides = W.grid_quads(3, 2)
    with pytest.raises(L.SDGError):  # vertex index out of range
        L.Mesh.from_arrays(L.ELEM_QUADRANGLE, 2, xyz, elems)
    with pytest.raises(L.SDGError):  # unpaired faces without boundary data
        L.Mesh.from_arrays(L.ELEM_QUADRANGLE, 2, xyz, elems - 1)
    with pytest.raises(L.SDGError):
        L.Mesh.from_arrays(L.ELEM_LINE, 2, xyz, elems - 1)
