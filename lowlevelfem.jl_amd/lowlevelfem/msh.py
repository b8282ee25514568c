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

Both MSH 4.1 encodings are read: ASCII, and binary with 8-byte data
(``4.1 1 8``; little- or big-endian, told by the int written after the
header).  MSH 2.x is rejected.

Supported volume types: hex8, hex27, tet4, tet10, prism6, hex64, hex125; face
types tri3, tri6, quad4, quad9, quad16, quad25.  For orders <= 2 the node
order here is gmsh's documented one.  For hex64/hex125 (and quad16/quad25)
gmsh orders the nodes as vertices, edge nodes, then the face and volume
interiors each as a lower-order element of the same family, recursively
(gmsh's point generators for complete Lagrange hexahedra and quadrangles);
the element-local order here lists those interiors lexicographically, so the
reader permutes them.  gmsh itself is not available here and no file of the
reference holds such elements, so this permutation is parity unpinned: it is
checked for consistency (every gmsh position is a lattice point, round trips
through the writer, and the recursive rule reproduces gmsh's documented
hex27/quad9 orders).
"""
from __future__ import annotations

import gzip
import os

import numpy as np

from . import refelem as re
from .mesh import Block, FaceSet, Mesh

import struct
from functools import lru_cache

# element types whose node order in refelem is gmsh's documented order
# (points, lines of order 1-4, and the surface and volume types below)
GMSH_ORDERED = {1, 2, 3, 4, 5, 6, 8, 9, 10, 11, 12, 15, 26, 27}
# element types read through a node permutation (gmsh's recursive order)
GMSH_PERMUTED = {36, 37, 92, 93}
# nodes per element of gmsh's element type codes (binary element rows carry
# no length), also for types this reader passes over (points, lines, ...)
GMSH_NPE = {1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 6: 6, 7: 5, 8: 3, 9: 6, 10: 9, 11: 10, 12: 27, 13: 18, 14: 14,
            15: 1, 16: 8, 17: 20, 18: 15, 19: 13, 21: 10, 23: 15, 26: 4, 27: 5, 29: 20, 30: 35, 36: 16,
            37: 25, 92: 64, 93: 125}


@lru_cache(maxsize=None)
def gmsh_quad_lattice(p: int) -> tuple:
    """Lattice points (i, j) in [0, p]^2 of the order-p quadrangle in gmsh's
    order: vertices, edge nodes (edges 0-1, 1-2, 2-3, 3-0, from their first
    vertex), then the interior as an order p-2 quadrangle, recursively."""
    if p == 0:
        return ((0, 0),)
    V = [np.array(v, dtype=np.int64) * p for v in re.QUAD_VERTS]
    pts = [tuple(int(x) for x in v) for v in V]
    for a, b in re.QUAD_EDGES:
        for m in range(1, p):
            pts.append(tuple(int(x) for x in V[a] + (V[b] - V[a]) * m // p))
    if p >= 2:
        pts += [(i + 1, j + 1) for i, j in gmsh_quad_lattice(p - 2)]
    return tuple(pts)


@lru_cache(maxsize=None)
def gmsh_hex_lattice(p: int) -> tuple:
    """Lattice points (i, j, k) in [0, p]^3 of the order-p hexahedron in gmsh's
    order: vertices, edge nodes along gmsh's twelve edges, the interior nodes
    of each of gmsh's six faces as an order p-2 quadrangle on the face (first
    axis from the face's first to second vertex, second axis from its first
    to fourth), then the volume interior as an order p-2 hexahedron,
    recursively."""
    if p == 0:
        return ((0, 0, 0),)
    V = [np.array(v, dtype=np.int64) * p for v in re.HEX_VERTS]
    pts = [tuple(int(x) for x in v) for v in V]
    for a, b in re.HEX_EDGES:
        for m in range(1, p):
            pts.append(tuple(int(x) for x in V[a] + (V[b] - V[a]) * m // p))
    if p >= 2:
        for f in re.HEX_FACES:
            v0, v1, v3 = V[f[0]], V[f[1]], V[f[3]]
            for s, t in gmsh_quad_lattice(p - 2):
                pts.append(tuple(int(x) for x in v0 + (v1 - v0) * (s + 1) // p + (v3 - v0) * (t + 1) // p))
        pts += [(i + 1, j + 1, k + 1) for i, j, k in gmsh_hex_lattice(p - 2)]
    return tuple(pts)


@lru_cache(maxsize=None)
def gmsh_to_local(etype: int) -> np.ndarray:
    """perm with local_conn = gmsh_conn[:, perm] (identity for the types whose
    local order is gmsh's)."""
    name, dim, order, npe, family = re.element_properties(etype)
    if etype not in GMSH_PERMUTED:
        return np.arange(npe)
    if family == "hex":
        g, loc = gmsh_hex_lattice(order), re.hex_node_lattice(order)
    else:
        g, loc = gmsh_quad_lattice(order), re.quad_node_lattice(order)
    pos = {tuple(int(x) for x in pt): j for j, pt in enumerate(g)}
    perm = np.array([pos[tuple(int(x) for x in pt)] for pt in loc], np.int64)
    assert len(pos) == npe and np.array_equal(np.sort(perm), np.arange(npe))
    return perm


def _raw(path) -> bytes:
    if str(path).endswith(".gz"):
        with gzip.open(path, "rb") as fh:
            return fh.read()
    with open(path, "rb") as fh:
        return fh.read()


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


def _parse_ascii(text):
    """(names, entity physical tags, node tags, coordinates, element blocks)
    of an ASCII MSH 4.1 file."""
    sec = {}
    for name, body in _sections(text.splitlines()):
        sec.setdefault(name, body)
    names = {}
    for ln in sec.get("PhysicalNames", [])[1:]:
        dim, tag, rest = ln.split(maxsplit=2)
        names[(int(dim), int(tag))] = rest.strip().strip('"')
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
    return names, ent_phys, tags, xyz, blocks


class _Bin:
    """Cursor over a binary MSH 4.1 file: ints are 4 bytes, size_t and double
    8 (data size 8), in the byte order the header's int told."""

    def __init__(self, data: bytes, pos: int, endian: str):
        self.d, self.p, self.e = data, pos, endian

    def take(self, fmt: str, n: int = 1):
        f = f"{self.e}{n}{fmt}"
        size = struct.calcsize(f)
        if self.p + size > len(self.d):
            raise ValueError("read_msh: binary section ends early")
        v = struct.unpack_from(f, self.d, self.p)
        self.p += size
        return v

    def array(self, dtype: str, n: int) -> np.ndarray:
        dt = np.dtype(dtype).newbyteorder(self.e)
        size = dt.itemsize * n
        if self.p + size > len(self.d):
            raise ValueError("read_msh: binary section ends early")
        a = np.frombuffer(self.d, dt, count=n, offset=self.p).astype(dtype)
        self.p += size
        return a

    def end(self, name: str):
        tag = b"$End" + name.encode()
        j = self.d.find(tag, self.p)
        if j < 0 or self.d[self.p:j].strip():
            raise ValueError(f"read_msh: binary section ${name} is not closed where it should be")
        self.p = j + len(tag)


def _parse_binary(data: bytes, pos: int):
    """The same as _parse_ascii for a binary MSH 4.1 file whose $MeshFormat
    header line ends at pos (the endianness int follows it)."""
    one = data[pos:pos + 4]
    endian = "<" if one == (1).to_bytes(4, "little") else ">" if one == (1).to_bytes(4, "big") else None
    if endian is None:
        raise ValueError("read_msh: binary MSH without the endianness int after the header")
    cur = _Bin(data, pos + 4, endian)
    cur.end("MeshFormat")
    names, ent_phys, tags, xyz, blocks = {}, {}, None, None, []
    while True:
        j = data.find(b"$", cur.p)
        if j < 0:
            break
        nl = data.find(b"\n", j)
        name = data[j + 1:nl if nl >= 0 else len(data)].decode("ascii", "replace").strip()
        cur.p = nl + 1 if nl >= 0 else len(data)
        if name == "PhysicalNames":        # ASCII in both encodings
            k = data.find(b"$EndPhysicalNames", cur.p)
            lines = data[cur.p:k].decode().splitlines()
            for ln in lines[1:]:
                if ln.strip():
                    dim, tag, rest = ln.split(maxsplit=2)
                    names[(int(dim), int(tag))] = rest.strip().strip('"')
            cur.p = k + len(b"$EndPhysicalNames")
        elif name == "Entities":
            counts = cur.take("Q", 4)
            for dim, cnt in enumerate(counts):
                for _ in range(cnt):
                    tag = cur.take("i")[0]
                    cur.take("d", 3 if dim == 0 else 6)
                    nph = cur.take("Q")[0]
                    ent_phys[(dim, tag)] = list(cur.take("i", nph)) if nph else []
                    if dim > 0:
                        nbd = cur.take("Q")[0]
                        if nbd:
                            cur.take("i", nbd)
            cur.end("Entities")
        elif name == "Nodes":
            nblocks, nnodes, tmin, tmax = cur.take("Q", 4)
            tags = np.empty(nnodes, np.int64)
            xyz = np.empty((nnodes, 3))
            c = 0
            for _ in range(nblocks):
                edim, etag, param = cur.take("i", 3)
                nb = cur.take("Q")[0]
                tags[c:c + nb] = cur.array("u8", nb).astype(np.int64)
                width = 3 + (edim if param else 0)
                xyz[c:c + nb] = cur.array("f8", nb * width).reshape(nb, width)[:, :3]
                c += nb
            if c != nnodes:
                raise ValueError(f"read_msh: {c} nodes read, header says {nnodes}")
            cur.end("Nodes")
        elif name == "Elements":
            nblocks = cur.take("Q", 4)[0]
            for _ in range(nblocks):
                edim, etag, etype = cur.take("i", 3)
                nb = cur.take("Q")[0]
                npe = GMSH_NPE.get(etype)
                if npe is None:
                    raise NotImplementedError(f"read_msh: element type {etype} in a binary file")
                rows = cur.array("u8", nb * (npe + 1)).astype(np.int64).reshape(nb, npe + 1)
                blocks.append((edim, etag, etype, rows[:, 1:] - 1))
            cur.end("Elements")
        else:                              # sections this reader does not use
            k = data.find(b"$End" + name.encode(), cur.p)
            if k < 0:
                raise ValueError(f"read_msh: section ${name} is not closed")
            cur.p = k + len(name) + 4
    if tags is None:
        raise ValueError("read_msh: no $Nodes section")
    return names, ent_phys, tags, xyz, blocks


def read_msh(path) -> Mesh:
    """Read a gmsh MSH 4.1 file (ASCII or binary) into a ``Mesh``."""
    data = _raw(path)
    i = data.find(b"$MeshFormat")
    if i < 0:
        raise ValueError("read_msh: no $MeshFormat section")
    hdr_start = data.find(b"\n", i) + 1
    hdr_end = data.find(b"\n", hdr_start)
    header = data[hdr_start:hdr_end].decode("ascii", "replace").split()
    if len(header) < 3:
        raise ValueError("read_msh: bad $MeshFormat header")
    version, ftype, dsize = header[:3]
    if not version.startswith("4.1") or ftype not in ("0", "1"):
        raise NotImplementedError(f"read_msh: MSH {version} file type {ftype}; only 4.1 is read")
    if ftype == "1":
        if dsize != "8":
            raise NotImplementedError(f"read_msh: binary MSH with data size {dsize}; only 8 is read")
        names, ent_phys, tags, xyz, blocks = _parse_binary(data, hdr_end + 1)
    else:
        names, ent_phys, tags, xyz, blocks = _parse_ascii(data.decode("utf-8", "replace"))
    nnodes = tags.size
    order = np.argsort(tags)
    if not np.array_equal(tags[order], np.arange(1, nnodes + 1)):
        raise ValueError("read_msh: node tags must be 1..N (global DoFs are derived from node tags)")
    coords = np.ascontiguousarray(xyz[order])
    if not blocks:
        raise ValueError("read_msh: no $Elements section")
    # high-order types: gmsh's node order -> the element-local one
    blocks = [(d, t, et, np.ascontiguousarray(c[:, gmsh_to_local(et)]) if et in GMSH_PERMUTED else c)
              for d, t, et, c in blocks]

    vol = [b for b in blocks if b[0] == 3]
    if not vol:
        raise ValueError("read_msh: no volume elements")
    # one element block per volume element type, types in order of first
    # appearance, elements of a type in file order
    vtypes = list(dict.fromkeys(b[2] for b in vol))
    for vt in vtypes:
        if vt not in GMSH_ORDERED | GMSH_PERMUTED or re.element_properties(vt)[1] != 3:
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

    # boundary groups: surfaces, curves and points (loads and supports)
    faces = {}
    for (dim, ptag), name in names.items():
        if dim not in (0, 1, 2):
            continue
        fb = [b for b in blocks if b[0] == dim and ptag in ent_phys.get((dim, b[1]), [])]
        if not fb:
            continue
        sets = []
        for ft in dict.fromkeys(b[2] for b in fb):
            if ft not in GMSH_ORDERED | GMSH_PERMUTED:
                raise NotImplementedError(f"read_msh: boundary group '{name}' has element type {ft}")
            sets.append(FaceSet(etype=ft, conn=np.ascontiguousarray(np.concatenate([b[3] for b in fb if b[2] == ft]))))
        faces[name] = sets[0] if len(sets) == 1 else sets
    b0 = eblocks[0]
    return Mesh(coords=coords, etype=b0.etype, conn=b0.conn, faces=faces,
                name=os.path.basename(str(path)).split(".")[0], volumes=b0.volumes, blocks=eblocks[1:])
