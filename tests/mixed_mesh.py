"""Test helper: a box of hexahedra and prisms (one mesh, two volume element
types) and an MSH 4.1 ASCII writer for it.

The box is a structured hex8 mesh whose cells with i >= split are each cut
into two 6-node wedges along the diagonal plane through hex nodes 0, 2, 6, 4
(gmsh node order), so every face shared by a hexahedron and a wedge is the
same quadrilateral.  Material groups: "steel" (cells with k < nz/2) and
"alu" (the rest), each holding elements of both types.  Face groups: "left"
(x = 0, hex quads), "right" (x = lx, wedge quads) and "front" (z = lz: hex
quads and wedge triangles, two face types in one group)."""
import numpy as np

from lowlevelfem import mesh as M

WEDGES = ([0, 1, 2, 4, 5, 6], [0, 2, 3, 4, 6, 7])


def hex_prism_box(cells=(4, 3, 4), lx=1.0, ly=0.8, lz=1.2, split=None):
    nx, ny, nz = cells
    split = nx // 2 if split is None else split
    m = M.structured_box_mesh(lx=lx, ly=ly, lz=lz, cells=cells)
    c = np.arange(m.E)
    i, k = c % nx, c // (nx * ny)
    is_hex = i < split
    steel_cell = k < nz // 2
    hconn = np.ascontiguousarray(m.conn[is_hex])
    pc = m.conn[~is_hex]
    pconn = np.ascontiguousarray(np.stack([pc[:, WEDGES[0]], pc[:, WEDGES[1]]], axis=1).reshape(-1, 6))
    hvol = {"steel": np.flatnonzero(steel_cell[is_hex]), "alu": np.flatnonzero(~steel_cell[is_hex])}
    pst = np.repeat(steel_cell[~is_hex], 2)
    pvol = {"steel": np.flatnonzero(pst), "alu": np.flatnonzero(~pst)}
    top = k == nz - 1
    fq = m.conn[top & is_hex][:, [4, 5, 6, 7]]
    ft = m.conn[top & ~is_hex]
    ftri = np.stack([ft[:, [4, 5, 6]], ft[:, [4, 6, 7]]], axis=1).reshape(-1, 3)
    faces = {"left": m.faces["left"], "right": m.faces["right"],
             "front": [M.FaceSet(3, np.ascontiguousarray(fq)), M.FaceSet(2, np.ascontiguousarray(ftri))]}
    return M.Mesh(m.coords, 5, hconn, faces=faces, name="hexprism", volumes=hvol,
                  blocks=[M.Block(6, pconn, pvol)])


def _gmsh_conn(etype, conn):
    """Element-local node order -> gmsh's (high-order hexahedra/quadrangles)."""
    from lowlevelfem.msh import GMSH_PERMUTED, gmsh_to_local
    if etype in GMSH_PERMUTED:
        return conn[:, np.argsort(gmsh_to_local(etype))]
    return conn


def write_msh(mesh, path, binary=False):
    """MSH 4.1, ASCII or binary (little-endian, data size 8): one volume
    entity per (volume group, element type), one point, curve or surface
    entity per (boundary group, element type) by the element type's
    dimension; node tags 1..N; element nodes in gmsh's order."""
    from lowlevelfem import refelem as R
    groups = list(dict.fromkeys(g for b in mesh.element_blocks() for g in b.volumes))
    fnames = list(mesh.faces)
    ptag = {("v", g): i + 1 for i, g in enumerate(groups)}
    ptag.update({("s", f): len(groups) + i + 1 for i, f in enumerate(fnames)})
    lo, hi = mesh.coords.min(axis=0), mesh.coords.max(axis=0)
    # per dimension 0..3: entities (tag, physical tag, element type, connectivity)
    ents_by_dim = [[], [], [], []]
    fdim = {}
    for f in fnames:
        for fs in mesh.face_sets(f):
            d = R.element_properties(fs.etype)[1]
            fdim[f] = d
            ents_by_dim[d].append((len(ents_by_dim[d]) + 1, ptag[("s", f)], fs.etype, _gmsh_conn(fs.etype, fs.conn)))
    for g in groups:
        for b in mesh.element_blocks():
            if g in b.volumes:
                ents_by_dim[3].append((len(ents_by_dim[3]) + 1, ptag[("v", g)], b.etype,
                                       _gmsh_conn(b.etype, b.conn[b.volumes[g]])))
    ents = [(d, t, et, c) for d in range(4) for t, _, et, c in ents_by_dim[d]]
    ne = sum(c.shape[0] for *_, c in ents)
    N = mesh.N
    names = [f'3 {ptag[("v", g)]} "{g}"' for g in groups] + [f'{fdim[f]} {ptag[("s", f)]} "{f}"' for f in fnames]
    if not binary:
        bbox = " ".join(f"{v:.17g}" for v in list(lo) + list(hi))
        out = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$PhysicalNames", str(len(ptag))] + names
        out += ["$EndPhysicalNames", "$Entities", " ".join(str(len(e)) for e in ents_by_dim)]
        for d in range(4):
            for t, p, _, c in ents_by_dim[d]:
                if d == 0:
                    x = mesh.coords[c[0, 0]]
                    out.append(f"{t} " + " ".join(f"{v:.17g}" for v in x) + f" 1 {p}")
                else:
                    out.append(f"{t} {bbox} 1 {p} 0")
        out += ["$EndEntities", "$Nodes", f"1 {N} 1 {N}", f"3 1 0 {N}"]
        out += [str(i + 1) for i in range(N)] + [" ".join(f"{v:.17g}" for v in x) for x in mesh.coords]
        out += ["$EndNodes", "$Elements", f"{len(ents)} {ne} 1 {ne}"]
        tag = 1
        for dim, t, et, c in ents:
            out.append(f"{dim} {t} {et} {c.shape[0]}")
            for row in c:
                out.append(f"{tag} " + " ".join(str(v + 1) for v in row))
                tag += 1
        out.append("$EndElements")
        with open(path, "w") as fh:
            fh.write("\n".join(out) + "\n")
        return
    import struct
    q = lambda *v: struct.pack(f"<{len(v)}Q", *v)
    i4 = lambda *v: struct.pack(f"<{len(v)}i", *v)
    d8 = lambda *v: struct.pack(f"<{len(v)}d", *v)
    parts = [b"$MeshFormat\n4.1 1 8\n", i4(1), b"\n$EndMeshFormat\n$PhysicalNames\n",
             ("\n".join([str(len(ptag))] + names) + "\n").encode(), b"$EndPhysicalNames\n$Entities\n",
             q(*(len(e) for e in ents_by_dim))]
    for d in range(4):
        for t, p, _, c in ents_by_dim[d]:
            if d == 0:
                parts += [i4(t), d8(*mesh.coords[c[0, 0]]), q(1), i4(p)]
            else:
                parts += [i4(t), d8(*lo, *hi), q(1), i4(p), q(0)]
    parts += [b"\n$EndEntities\n$Nodes\n", q(1, N, 1, N), i4(3, 1, 0), q(N),
              np.arange(1, N + 1, dtype="<u8").tobytes(), np.ascontiguousarray(mesh.coords, dtype="<f8").tobytes(),
              b"\n$EndNodes\n$Elements\n", q(len(ents), ne, 1, ne)]
    tag = 1
    for dim, t, et, c in ents:
        rows = np.empty((c.shape[0], c.shape[1] + 1), dtype="<u8")
        rows[:, 0] = np.arange(tag, tag + c.shape[0])
        rows[:, 1:] = c + 1
        tag += c.shape[0]
        parts += [i4(dim, t, et), q(c.shape[0]), rows.tobytes()]
    parts.append(b"\n$EndElements\n")
    with open(path, "wb") as fh:
        fh.write(b"".join(parts))


LINE_TYPE_BY_ORDER = {1: 1, 2: 8, 3: 26, 4: 27}


def with_edge_and_corner(mesh):
    """The box `mesh` with three more boundary groups: "edge" (the line
    elements of the order of the mesh along the box edge y = y_min,
    z = z_min, vertices first, then the interior nodes from the first vertex
    to the second), "corner" (a point element at the node with the largest
    x + y + z) and "corners2" (point elements at two nodes, the second one
    listed twice)."""
    from lowlevelfem import refelem as R
    order = R.element_properties(mesh.etype)[2]
    x = mesh.coords
    lo = x.min(axis=0)
    on = np.flatnonzero((np.abs(x[:, 1] - lo[1]) < 1e-12) & (np.abs(x[:, 2] - lo[2]) < 1e-12))
    on = on[np.argsort(x[on, 0])]
    ne = (on.size - 1) // order
    assert ne * order + 1 == on.size, "edge nodes do not split into elements"
    rows = [[on[order * e], on[order * (e + 1)]] + [on[order * e + k] for k in range(1, order)] for e in range(ne)]
    corner = int(np.argmax(x.sum(axis=1)))
    other = int(np.argmin(x.sum(axis=1)))
    faces = dict(mesh.faces)
    faces["edge"] = M.FaceSet(LINE_TYPE_BY_ORDER[order], np.array(rows, dtype=np.int64))
    faces["corner"] = M.FaceSet(15, np.array([[corner]], dtype=np.int64))
    faces["corners2"] = M.FaceSet(15, np.array([[corner], [other], [other]], dtype=np.int64))
    return M.Mesh(mesh.coords, mesh.etype, mesh.conn, faces=faces, name=mesh.name,
                  volumes=mesh.volumes, blocks=mesh.blocks)
