"""Reader of gmsh MSH 4.1 ASCII files (host side).

The reference loads meshes through the gmsh API (``gmsh.open`` and
``gmsh.model.mesh.getNodes`` / ``getElements`` / ``getElementsByType``, called
at src/multifield.jl:1612-1648 and src/linear.jl:195-243) and derives global
DoFs from node tags, ``dof = 3(tag-1) + c`` (src/linear.jl:309-312), which is
why it requires the tags to be exactly 1..N (src/multifield.jl:1602-1611).
This module reads the published MSH 4.1 ASCII layout directly -- sections
``$PhysicalNames``, ``$Entities``, ``$Nodes``, ``$Elements`` -- and returns the
same ``Mesh`` the structured generators build: node ``tag - 1`` is row
``tag - 1`` of the coordinates (gmsh's own numbering, so the CSC pattern is
the one the reference builds on this file), the volume elements in file order
with gmsh's node order -- one block per volume element type, types in order
of first appearance -- the element indices of every 3D physical group per
block (the material groups), and the elements of every 2D physical group as
a ``FaceSet`` (a list of them when the group has several face types) under
the group's name.  Files ending in ``.gz`` are read through
gzip.

Supported element types are those whose node order here is gmsh's (hex8,
hex27, tet4, tet10, prism6 volumes; tri3, tri6, quad4, quad9 faces).  Binary
MSH files and MSH 2.x are rejected.
"""
from __future__ import annotations

import gzip
import os

import numpy as np

from . import refelem as re
from .mesh import Block, FaceSet, Mesh

# element types whose node order in refelem is gmsh's documented order
GMSH_ORDERED = {2, 3, 4, 5, 6, 9, 10, 11, 12}


def _open(path):
    if str(path).endswith(".gz"):
        return gzip.open(path, "rt")
    return open(path, "r")


def _sections(lines):
    """(name, body lines) for every $Name ... $EndName section."""
    i, n = 0, len(lines)
    while i < n:
        s = lines[i].strip()
        if s.startswith("$") and not s.startswith("$End"):
            name = s[1:]
            j = i + 1
            while j < n and lines[j].strip() != "$End" + name:
                j += 1
            if j == n:
                raise ValueError(f"read_msh: section ${name} is not closed")
            yield name, lines[i + 1:j]
            i = j + 1
        else:
            i += 1


def read_msh(path) -> Mesh:
    """Read a gmsh MSH 4.1 ASCII file into a ``Mesh``."""
    with _open(path) as fh:
        lines = fh.read().splitlines()
    sec = {}
    for name, body in _sections(lines):
        sec.setdefault(name, body)
    if "MeshFormat" not in sec:
        raise ValueError("read_msh: no $MeshFormat section")
    version, ftype = sec["MeshFormat"][0].split()[:2]
    if not version.startswith("4.1") or ftype != "0":
        raise NotImplementedError(f"read_msh: MSH {version} file type {ftype}; only 4.1 ASCII is read")

    # physical groups: (dim, physical tag) -> name
    names = {}
    for ln in sec.get("PhysicalNames", [])[1:]:
        dim, tag, rest = ln.split(maxsplit=2)
        names[(int(dim), int(tag))] = rest.strip().strip('"')

    # entities: (dim, entity tag) -> physical tags
    ent_phys = {}
    if "Entities" in sec:
        body = sec["Entities"]
        counts = [int(x) for x in body[0].split()[:4]]
        k = 1
        for dim, cnt in enumerate(counts):
            for _ in range(cnt):
                tok = body[k].split()
                k += 1
                tag = int(tok[0])
                p = 4 if dim == 0 else 7          # points: tag x y z; others: tag + bounding box
                nph = int(tok[p])
                ent_phys[(dim, tag)] = [int(t) for t in tok[p + 1:p + 1 + nph]]

    # nodes
    if "Nodes" not in sec:
        raise ValueError("read_msh: no $Nodes section")
    body = sec["Nodes"]
    nblocks, nnodes, tmin, tmax = (int(x) for x in body[0].split()[:4])
    tags = np.empty(nnodes, np.int64)
    xyz = np.empty((nnodes, 3))
    k, c = 1, 0
    for _ in range(nblocks):
        edim, etag, param, nb = (int(x) for x in body[k].split()[:4])
        k += 1
        tags[c:c + nb] = [int(body[k + i]) for i in range(nb)]
        k += nb
        for i in range(nb):
            xyz[c + i] = [float(v) for v in body[k + i].split()[:3]]
        k += nb
        c += nb
    if c != nnodes:
        raise ValueError(f"read_msh: {c} nodes read, header says {nnodes}")
    order = np.argsort(tags)
    if not np.array_equal(tags[order], np.arange(1, nnodes + 1)):
        raise ValueError("read_msh: node tags must be 1..N (global DoFs are derived from node tags)")
    coords = np.ascontiguousarray(xyz[order])

    # elements, grouped by (entity dim, entity tag)
    if "Elements" not in sec:
        raise ValueError("read_msh: no $Elements section")
    body = sec["Elements"]
    nblocks = int(body[0].split()[0])
    blocks = []
    k = 1
    for _ in range(nblocks):
        edim, etag, etype, nb = (int(x) for x in body[k].split()[:4])
        k += 1
        rows = np.array([[int(v) for v in body[k + i].split()] for i in range(nb)], np.int64).reshape(nb, -1)
        k += nb
        blocks.append((edim, etag, etype, rows[:, 1:] - 1))

    vol = [b for b in blocks if b[0] == 3]
    if not vol:
        raise ValueError("read_msh: no volume elements")
    # one element block per volume element type, types in order of first
    # appearance, elements of a type in file order
    vtypes = list(dict.fromkeys(b[2] for b in vol))
    for vt in vtypes:
        if vt not in GMSH_ORDERED or re.element_properties(vt)[1] != 3:
            raise NotImplementedError(f"read_msh: volume element type {vt} is not supported")
    eblocks = []
    for vt in vtypes:
        vb = [b for b in vol if b[2] == vt]
        conn = np.ascontiguousarray(np.concatenate([b[3] for b in vb]))
        if conn.min() < 0 or conn.max() >= nnodes:
            raise ValueError("read_msh: element node tag out of range")
        # volume physical groups: element indices (file order within the type) of their entities
        first = np.cumsum([0] + [b[3].shape[0] for b in vb])
        volumes = {}
        for (dim, ptag), name in names.items():
            if dim != 3:
                continue
            idx = [np.arange(first[i], first[i + 1]) for i, b in enumerate(vb)
                   if ptag in ent_phys.get((3, b[1]), [])]
            if idx:
                volumes[name] = np.concatenate(idx)
        eblocks.append(Block(vt, conn, volumes))

    faces = {}
    for (dim, ptag), name in names.items():
        if dim != 2:
            continue
        fb = [b for b in blocks if b[0] == 2 and ptag in ent_phys.get((2, b[1]), [])]
        if not fb:
            continue
        sets = []
        for ft in dict.fromkeys(b[2] for b in fb):
            if ft not in GMSH_ORDERED:
                raise NotImplementedError(f"read_msh: face group '{name}' has element type {ft}")
            sets.append(FaceSet(etype=ft, conn=np.ascontiguousarray(np.concatenate([b[3] for b in fb if b[2] == ft]))))
        faces[name] = sets[0] if len(sets) == 1 else sets
    b0 = eblocks[0]
    return Mesh(coords=coords, etype=b0.etype, conn=b0.conn, faces=faces,
                name=os.path.basename(str(path)).split(".")[0], volumes=b0.volumes, blocks=eblocks[1:])
