"""Writer of gmsh MSH 4.1 ASCII files for the mesh-reader tests (test data, not a mesh generator):
one domain entity holding the elements, one entity per boundary group with its physical tag."""
from __future__ import annotations

import numpy as np

GMSH_TYPE = {"tri": 2, "quad": 3, "hex": 5, "line": 1, "bquad": 3, "btri": 2, "tet": 4, "tri6": 9, "quad9": 10,
             "line3": 8, "tet10": 11, "hex27": 12, "tri10": 21, "quad16": 36}


def write_msh(path, dim: int, node_tags: np.ndarray, xyz: np.ndarray, elem_type: str, elems: np.ndarray,
              boundary: list[tuple[int, str, np.ndarray]], periodic=None, elem_type_code: int | None = None):
    """node_tags (N,), xyz (N, 3); elems (E, nvert) of node tags; boundary: (physical tag, element
    kind, (n, nfvert) node tags) per group; periodic: the (slave tag, master tag) node pairs of one
    $Periodic link, or a list of such lists (one link each; the affine part is written as the
    identity, which the reader skips).
    elem_type_code overrides the gmsh type of the domain elements (to write unsupported types)."""
    lines = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat"]
    nb = len(boundary)
    counts = [0, 0, 0, 0]
    counts[dim - 1] = nb
    counts[dim] = 1
    lines += ["$Entities", " ".join(str(c) for c in counts)]
    lo, hi = xyz.min(0), xyz.max(0)
    box = " ".join(repr(float(v)) for v in (*lo, *hi))
    for k, (phys, _, _) in enumerate(boundary):
        lines.append(f"{k + 1} {box} 1 {phys} 0")
    lines.append(f"1 {box} 0 0")
    lines.append("$EndEntities")
    lines += ["$Nodes", f"1 {len(node_tags)} {int(node_tags.min())} {int(node_tags.max())}", f"{dim} 1 0 {len(node_tags)}"]
    lines += [str(int(t)) for t in node_tags]
    lines += [" ".join(repr(float(v)) for v in p) for p in xyz]
    lines.append("$EndNodes")
    total = len(elems) + sum(len(b[2]) for b in boundary)
    lines += ["$Elements", f"{1 + nb} {total} 1 {total}"]
    tag = 1
    code = elem_type_code if elem_type_code is not None else GMSH_TYPE[elem_type]
    lines.append(f"{dim} 1 {code} {len(elems)}")
    for e in elems:
        lines.append(f"{tag} " + " ".join(str(int(v)) for v in e))
        tag += 1
    for k, (_, kind, conn) in enumerate(boundary):
        lines.append(f"{dim - 1} {k + 1} {GMSH_TYPE[kind]} {len(conn)}")
        for c in conn:
            lines.append(f"{tag} " + " ".join(str(int(v)) for v in c))
            tag += 1
    lines.append("$EndElements")
    if periodic is not None:
        links = periodic if periodic and isinstance(periodic[0], list) else [periodic]
        lines += ["$Periodic", str(len(links))]
        for k, link in enumerate(links):
            lines += [f"{dim - 1} {2 * k + 1} {2 * k + 2}",
                      "16 " + " ".join("1" if i % 5 == 0 else "0" for i in range(16)), str(len(link))]
            lines += [f"{int(a)} {int(b)}" for a, b in link]
        lines.append("$EndPeriodic")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def grid_quads(nx: int, ny: int, lo=(0.0, 0.0), hi=(1.0, 1.0)):
    """Structured quadrangle grid in the structured builder's element and vertex order: node (i, j)
    has tag 1 + i + (nx + 1) j; returns tags, xyz, elements and the four sides' line lists
    (x-low, x-high, y-low, y-high)."""
    h = [(hi[0] - lo[0]) / nx, (hi[1] - lo[1]) / ny]
    tag = lambda i, j: 1 + i + (nx + 1) * j  # noqa: E731
    tags = np.array([tag(i, j) for j in range(ny + 1) for i in range(nx + 1)])
    xyz = np.array([[lo[0] + i * h[0], lo[1] + j * h[1], 0.0] for j in range(ny + 1) for i in range(nx + 1)])
    elems = np.array([[tag(i, j), tag(i + 1, j), tag(i + 1, j + 1), tag(i, j + 1)] for j in range(ny) for i in range(nx)])
    sides = [np.array([[tag(0, j), tag(0, j + 1)] for j in range(ny)]),
             np.array([[tag(nx, j), tag(nx, j + 1)] for j in range(ny)]),
             np.array([[tag(i, 0), tag(i + 1, 0)] for i in range(nx)]),
             np.array([[tag(i, ny), tag(i + 1, ny)] for i in range(nx)])]
    return tags, xyz, elems, sides


def grid_hexes(n: int):
    """Structured hexahedron grid of the unit cube (n^3 cells) in the structured builder's order;
    sides: x-low, x-high, y-low, y-high, z-low, z-high quadrangle lists."""
    h = 1.0 / n
    tag = lambda i, j, k: 1 + i + (n + 1) * (j + (n + 1) * k)  # noqa: E731
    tags = np.array([tag(i, j, k) for k in range(n + 1) for j in range(n + 1) for i in range(n + 1)])
    xyz = np.array([[i * h, j * h, k * h] for k in range(n + 1) for j in range(n + 1) for i in range(n + 1)])
    off = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    elems = np.array([[tag(i + a, j + b, k + c) for a, b, c in off]
                      for k in range(n) for j in range(n) for i in range(n)])
    sides = []
    for d in range(3):
        for v in (0, n):
            quads = []
            for q in range(n):
                for r in range(n):
                    pts = []
                    for a, b in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        ijk = [0, 0, 0]
                        ijk[d] = v
                        ijk[(d + 1) % 3] = q + a
                        ijk[(d + 2) % 3] = r + b
                        pts.append(tag(*ijk))
                    quads.append(pts)
            sides.append(np.array(quads))
    return tags, xyz, elems, sides
