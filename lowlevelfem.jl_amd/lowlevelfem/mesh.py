"""Structured mesh generators (host side).

They define the meshes the reference builds with gmsh's
``structured_box_mesh`` (src/general.jl:8910-8996): an axis-aligned box with
physical groups ``left``/``right`` (x = x0 / x0+lx), ``bottom``/``top``
(y = y0 / y0+ly), ``rear``/``front`` (z = z0 / z0+lz) and ``body``; the cell
size is ``d = min(dx, dy)`` on every axis (src/general.jl:8948-8963).

Node numbering is lexicographic on the (order*n+1)-point lattice (x fastest);
gmsh's transfinite numbering is not reproduced (it only permutes DoFs).  The
same arrays are handed to the oracle and to the GPU library, so pattern and
value parity are checked on identical inputs.

Element families: ``hex`` (order 1..4, gmsh hex8/hex27 ordering for orders
1-2), ``tet`` (Kuhn/Freudenthal 6-tet split of each cell, P1 or P2), and
``prism`` (two 6-node wedges per cell).
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field

import numpy as np

from . import refelem as re

FACE_AXES = {"left": (0, 0), "right": (0, 1), "bottom": (1, 0),
             "top": (1, 1), "rear": (2, 0), "front": (2, 1)}


@dataclass
class FaceSet:
    etype: int                 # gmsh boundary element type (surface, line or point)
    conn: np.ndarray           # (F, nf) 0-based node ids


@dataclass
class Block:
    """Volume elements of one gmsh element type."""
    etype: int
    conn: np.ndarray           # (E, npe) int64, 0-based node ids
    volumes: dict = field(default_factory=dict)  # volume physical group -> element indices in this block


@dataclass
class Mesh:
    coords: np.ndarray         # (N, 3) float64
    etype: int                 # gmsh volume element type (of the first block)
    conn: np.ndarray           # (E, npe) int64, 0-based node ids
    faces: dict = field(default_factory=dict)   # boundary group (surface, curve or point) name -> FaceSet,
                                                # or a list of FaceSets of several types
    name: str = "structured_box"
    volumes: dict = field(default_factory=dict)  # volume physical group -> element indices
    blocks: list = field(default_factory=list)   # further volume element types: Block per type

    def element_blocks(self) -> list:
        """Every volume element block: (etype, conn, volumes) first, then the
        further element types in file order."""
        return [Block(self.etype, self.conn, self.volumes)] + list(self.blocks)

    def face_sets(self, name: str) -> list:
        """The element sets of boundary physical group `name` (a surface,
        curve or point group), one per element type."""
        if name not in self.faces:
            raise KeyError(f"Physical group '{name}' not found.")
        fs = self.faces[name]
        return list(fs) if isinstance(fs, (list, tuple)) else [fs]

    @property
    def N(self) -> int:
        return self.coords.shape[0]

    @property
    def E(self) -> int:
        return self.conn.shape[0]

    @property
    def npe(self) -> int:
        return self.conn.shape[1]

    def nodes_on(self, group: str) -> np.ndarray:
        """Sorted unique node ids (0-based) of a physical group; the
        reference's nodesOnPhysicalGroup (src/general.jl:5249-5275)."""
        blocks = self.element_blocks()
        if any(group in b.volumes for b in blocks):
            return np.unique(np.concatenate([b.conn[b.volumes[group]].ravel() for b in blocks
                                             if group in b.volumes]))
        if group == "body":
            return np.arange(self.N, dtype=np.int64)
        return np.unique(np.concatenate([fs.conn.ravel() for fs in self.face_sets(group)]))


def _cells(lx, ly, lz, n, dx, dy, dz):
    d = min(dx, dy)
    return (int(math.ceil(lx / d - 1e-12)), int(math.ceil(ly / d - 1e-12)),
            int(math.ceil(lz / d - 1e-12)))


def structured_box_mesh(x0=0.0, y0=0.0, z0=0.0, lx=1.0, ly=1.0, lz=1.0, n=10,
                        dx=None, dy=None, dz=None, order=1, elem="hex",
                        cells=None, perturb=0.0, seed=0) -> Mesh:
    """Structured box mesh.  ``cells=(nx,ny,nz)`` overrides the reference's
    ``d = min(dx,dy)`` rule.  ``perturb`` moves every node not on the box
    boundary by a seeded uniform offset in [-perturb, perturb] * cell size
    per coordinate."""
    dx = lx / n if dx is None else dx
    dy = ly / n if dy is None else dy
    dz = lz / n if dz is None else dz
    nx, ny, nz = cells if cells is not None else _cells(lx, ly, lz, n, dx, dy, dz)
    if elem == "hex":
        m = _hex_mesh(nx, ny, nz, order)
    elif elem == "tet":
        m = _kuhn_tet_mesh(nx, ny, nz, order)
    elif elem == "prism":
        if order != 1:
            raise ValueError("prism meshes are order 1")
        m = _prism_mesh(nx, ny, nz)
    else:
        raise ValueError(elem)
    lat, conn, etype, faces, p = m
    h = np.array([lx / nx, ly / ny, lz / nz])
    coords = np.array([x0, y0, z0])[None, :] + lat * (h / p)[None, :]
    if perturb:
        L = np.array([nx, ny, nz]) * p
        interior = np.all((lat > 0) & (lat < L[None, :]), axis=1)
        rng = np.random.default_rng(seed)
        off = rng.uniform(-perturb, perturb, size=coords.shape) * h[None, :]
        coords = coords + off * interior[:, None]
    return Mesh(np.ascontiguousarray(coords), etype, conn, faces)


def _lattice_ids(shape):
    Ix, Iy, Iz = shape

    def nid(i, j, k):
        return i + Ix * (j + Iy * k)
    return nid


def _all_lattice(shape):
    Ix, Iy, Iz = shape
    k, j, i = np.meshgrid(np.arange(Iz), np.arange(Iy), np.arange(Ix), indexing="ij")
    return np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)


def _hex_mesh(nx, ny, nz, p):
    shape = (p * nx + 1, p * ny + 1, p * nz + 1)
    nid = _lattice_ids(shape)
    lat = _all_lattice(shape)
    loc = re.hex_node_lattice(p)                       # (npe, 3)
    ck, cj, ci = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ci, cj, ck = ci.ravel() * p, cj.ravel() * p, ck.ravel() * p
    conn = nid(ci[:, None] + loc[None, :, 0], cj[:, None] + loc[None, :, 1],
               ck[:, None] + loc[None, :, 2]).astype(np.int64)
    etype = re.HEX_TYPE_BY_ORDER[p]
    qloc = re.quad_node_lattice(p)                     # (nf, 2)
    ftype = re.QUAD_TYPE_BY_ORDER[p]
    faces = {}
    ncell = (nx, ny, nz)
    for name, (ax, side) in FACE_AXES.items():
        a1, a2 = [a for a in range(3) if a != ax]
        g2, g1 = np.meshgrid(np.arange(ncell[a2]), np.arange(ncell[a1]), indexing="ij")
        g1, g2 = g1.ravel() * p, g2.ravel() * p
        idx = [None, None, None]
        idx[ax] = np.full((g1.size, qloc.shape[0]), side * p * ncell[ax])
        idx[a1] = g1[:, None] + qloc[None, :, 0]
        idx[a2] = g2[:, None] + qloc[None, :, 1]
        faces[name] = FaceSet(ftype, nid(*idx).astype(np.int64))
    return lat, conn, etype, faces, p


def _kuhn_tet_mesh(nx, ny, nz, p):
    """Kuhn split: each cell is cut into the 6 tetrahedra that share the
    cell diagonal (0,0,0)-(1,1,1); the split is consistent across cells, so
    every face diagonal matches between neighbours.  P2 nodes live on the
    doubled lattice, which they fill completely."""
    if p not in (1, 2):
        raise ValueError("tet meshes are order 1 or 2")
    s = 2 if p == 2 else 1
    shape = (s * nx + 1, s * ny + 1, s * nz + 1)
    nid = _lattice_ids(shape)
    lat = _all_lattice(shape)
    tets = []
    for perm in itertools.permutations(range(3)):
        v = [np.zeros(3, int)]
        for ax in perm:
            w = v[-1].copy()
            w[ax] = 1
            v.append(w)
        V = np.array(v)
        det = np.linalg.det((V[1:] - V[0]).astype(float))
        if det < 0:
            V = V[[0, 2, 1, 3]]
        tets.append(V)
    tets = np.array(tets)                               # (6, 4, 3) in {0,1}
    if p == 2:
        loc = np.concatenate([2 * tets, np.stack([tets[:, a] + tets[:, b]
                                                  for a, b in re.TET_EDGES], axis=1)], axis=1)
    else:
        loc = tets
    ck, cj, ci = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = np.stack([ci.ravel(), cj.ravel(), ck.ravel()], axis=1) * s   # (C, 3)
    g = base[:, None, None, :] + loc[None, :, :, :]                      # (C, 6, n, 3)
    conn = nid(g[..., 0], g[..., 1], g[..., 2]).reshape(-1, loc.shape[1]).astype(np.int64)
    etype = 11 if p == 2 else 4
    ftype = 9 if p == 2 else 2
    # boundary faces: tet faces whose three vertices lie on a box plane
    tet_faces = [(0, 2, 1), (0, 1, 3), (0, 3, 2), (1, 2, 3)]
    faces = {}
    vl = g[:, :, :4, :].reshape(-1, 4, 3)               # vertex lattice coords
    ncell = (nx, ny, nz)
    for name, (ax, side) in FACE_AXES.items():
        target = side * s * ncell[ax]
        rows = []
        for f in tet_faces:
            on = np.all(vl[:, list(f), ax] == target, axis=1)
            if not np.any(on):
                continue
            tv = vl[on][:, list(f), :]                  # (F, 3, 3)
            if p == 2:
                mids = np.stack([(tv[:, a] + tv[:, b]) // 2 for a, b in re.TRI_EDGES], axis=1)
                tv = np.concatenate([tv, mids], axis=1)
            rows.append(nid(tv[..., 0], tv[..., 1], tv[..., 2]))
        faces[name] = FaceSet(ftype, np.concatenate(rows).astype(np.int64))
    return lat, conn, etype, faces, s


def _prism_mesh(nx, ny, nz):
    shape = (nx + 1, ny + 1, nz + 1)
    nid = _lattice_ids(shape)
    lat = _all_lattice(shape)
    tri = np.array([[(0, 0), (1, 0), (1, 1)], [(0, 0), (1, 1), (0, 1)]])   # ccw
    loc = []
    for t in tri:
        loc.append([(a, b, 0) for a, b in t] + [(a, b, 1) for a, b in t])
    loc = np.array(loc)                                 # (2, 6, 3)
    ck, cj, ci = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = np.stack([ci.ravel(), cj.ravel(), ck.ravel()], axis=1)
    g = base[:, None, None, :] + loc[None]
    conn = nid(g[..., 0], g[..., 1], g[..., 2]).reshape(-1, 6).astype(np.int64)
    faces = {}
    ncell = (nx, ny, nz)
    for name, (ax, side) in FACE_AXES.items():
        a1, a2 = [a for a in range(3) if a != ax]
        if ax == 2:     # triangle faces
            idx_rows = []
            for t in tri:
                g2, g1 = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
                i = g1.ravel()[:, None] + t[None, :, 0]
                j = g2.ravel()[:, None] + t[None, :, 1]
                k = np.full_like(i, side * nz)
                idx_rows.append(nid(i, j, k))
            faces[name] = FaceSet(2, np.concatenate(idx_rows).astype(np.int64))
        else:
            q = re.quad_node_lattice(1)
            g2, g1 = np.meshgrid(np.arange(ncell[a2]), np.arange(ncell[a1]), indexing="ij")
            idx = [None, None, None]
            idx[ax] = np.full((g1.size, 4), side * ncell[ax])
            idx[a1] = g1.ravel()[:, None] + q[None, :, 0]
            idx[a2] = g2.ravel()[:, None] + q[None, :, 1]
            faces[name] = FaceSet(3, nid(*idx).astype(np.int64))
    return lat, conn, 6, faces, 1


def structural_nnz(mesh: Mesh, pdim: int = 3) -> int:
    """Structural non-zeros of the node-adjacency pattern x pdim^2 (host,
    for sizing and cross-checks)."""
    n = mesh.npe
    rows = np.repeat(mesh.conn, n, axis=1).ravel()
    cols = np.tile(mesh.conn, (1, n)).ravel()
    keys = np.unique(cols.astype(np.int64) * mesh.N + rows)
    return int(keys.size) * pdim * pdim
