"""Reference-element tables: quadrature rules and Lagrange bases.

The reference obtains these from gmsh (``getIntegrationPoints``,
``getBasisFunctions(..., "Lagrange" / "GradLagrange")``, called at
src/linear.jl:235-239 and src/multifield.jl:2932-2940).  gmsh is a third-party
C++ library pinned at gmsh_jll 4.13.1+0 (docs/Manifest.toml:851-855) and is not
part of the reference source tree, so its published conventions are restated
here:

* element type codes are gmsh's (5 = hex8, 11 = tet10, ...);
* reference domains: hex/quad [-1,1]^d, tet/tri the unit simplex,
  prism = unit triangle x [-1,1];
* low-order node orderings (hex8, hex27, tet4, tet10, prism6, quad4, quad9,
  tri3, tri6) are gmsh's documented orderings;
* "GaussN" selects a rule exact for polynomial degree N; on tensor cells this
  is ceil((N+1)/2) Gauss-Legendre points per direction (so "Gauss3" on hex8 is
  2x2x2 and "Gauss9" on hex125 is 5x5x5).

Table layout matches gmsh's flattened output as the reference reshapes it
(src/linear.jl:239, ``reshape(dfun, :, numIntPoints)``): for GP q and node a,
``dN[q, a, :]`` holds (d/dxi, d/deta, d/dzeta); ``N[q, a]`` the value.

Simplex rules of degree >= 3 are collapsed (Duffy) Gauss-Jacobi products with
positive weights; gmsh uses other point sets there.  Every rule used here is
exact for the integrands of the affine-element operators it is applied to,
so results agree with any other exact rule up to rounding.

Node orderings of hex64/hex125 and quad16/quad25 follow the same
vertex/edge/face/interior hierarchy as gmsh's, but face and volume interior
nodes are listed lexicographically; meshes and tables built here are
consistent with each other, which is all the assembly kernels require.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np
from scipy.special import roots_jacobi

# gmsh element type codes -> (name, dim, order, nodes per element, family)
ELEMENT_TYPES = {
    1: ("line2", 1, 1, 2, "line"),
    2: ("tri3", 2, 1, 3, "tri"),
    3: ("quad4", 2, 1, 4, "quad"),
    4: ("tet4", 3, 1, 4, "tet"),
    5: ("hex8", 3, 1, 8, "hex"),
    6: ("prism6", 3, 1, 6, "prism"),
    8: ("line3", 1, 2, 3, "line"),
    9: ("tri6", 2, 2, 6, "tri"),
    10: ("quad9", 2, 2, 9, "quad"),
    11: ("tet10", 3, 2, 10, "tet"),
    12: ("hex27", 3, 2, 27, "hex"),
    36: ("quad16", 2, 3, 16, "quad"),
    37: ("quad25", 2, 4, 25, "quad"),
    15: ("point", 0, 0, 1, "point"),
    26: ("line4", 1, 3, 4, "line"),
    27: ("line5", 1, 4, 5, "line"),
    92: ("hex64", 3, 3, 64, "hex"),
    93: ("hex125", 3, 4, 125, "hex"),
}
TYPE_BY_NAME = {v[0]: k for k, v in ELEMENT_TYPES.items()}
HEX_TYPE_BY_ORDER = {1: 5, 2: 12, 3: 92, 4: 93}
QUAD_TYPE_BY_ORDER = {1: 3, 2: 10, 3: 36, 4: 37}

# gmsh hex vertex / edge / face lists (vertex indices)
HEX_VERTS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
             (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
HEX_EDGES = [(0, 1), (0, 3), (0, 4), (1, 2), (1, 5), (2, 3),
             (2, 6), (3, 7), (4, 5), (4, 7), (5, 6), (6, 7)]
HEX_FACES = [(0, 3, 2, 1), (0, 1, 5, 4), (0, 4, 7, 3),
             (1, 2, 6, 5), (2, 3, 7, 6), (4, 5, 6, 7)]
QUAD_VERTS = [(0, 0), (1, 0), (1, 1), (0, 1)]
QUAD_EDGES = [(0, 1), (1, 2), (2, 3), (3, 0)]
TET_EDGES = [(0, 1), (1, 2), (2, 0), (3, 0), (3, 2), (3, 1)]
TRI_EDGES = [(0, 1), (1, 2), (2, 0)]


def element_properties(etype: int):
    """(name, dim, order, npe, family) of a gmsh element type code."""
    if etype not in ELEMENT_TYPES:
        raise ValueError(f"unsupported element type {etype}")
    return ELEMENT_TYPES[etype]


# ----------------------------------------------------------------------------
# node lattices
# ----------------------------------------------------------------------------
@lru_cache(maxsize=None)
def hex_node_lattice(p: int) -> np.ndarray:
    """Integer lattice coordinates (i,j,k) in [0,p]^3 of the order-p hex nodes,
    in element-local node order (vertices, edges, faces, interior)."""
    pts = [tuple(p * c for c in v) for v in HEX_VERTS]
    V = [np.array(v) * p for v in HEX_VERTS]
    for a, b in HEX_EDGES:
        for m in range(1, p):
            pts.append(tuple(V[a] + (V[b] - V[a]) * m // p))
    for f in HEX_FACES:
        v0, v1, _, v3 = (V[i] for i in f)
        for t in range(1, p):
            for s in range(1, p):
                pts.append(tuple(v0 + (v1 - v0) * s // p + (v3 - v0) * t // p))
    for k in range(1, p):
        for j in range(1, p):
            for i in range(1, p):
                pts.append((i, j, k))
    arr = np.array(pts, dtype=np.int64)
    assert len({tuple(r) for r in arr}) == (p + 1) ** 3
    return arr


@lru_cache(maxsize=None)
def quad_node_lattice(p: int) -> np.ndarray:
    V = [np.array(v) * p for v in QUAD_VERTS]
    pts = [tuple(v) for v in V]
    for a, b in QUAD_EDGES:
        for m in range(1, p):
            pts.append(tuple(V[a] + (V[b] - V[a]) * m // p))
    for j in range(1, p):
        for i in range(1, p):
            pts.append((i, j))
    arr = np.array(pts, dtype=np.int64)
    assert len({tuple(r) for r in arr}) == (p + 1) ** 2
    return arr


def _lagrange_1d(p: int, x: np.ndarray):
    """Values and derivatives of the p+1 equispaced Lagrange polynomials on
    [-1,1] at points x.  Returns (L, dL) of shape (len(x), p+1)."""
    nodes = np.linspace(-1.0, 1.0, p + 1)
    x = np.asarray(x, dtype=np.float64)
    L = np.ones((x.size, p + 1))
    dL = np.zeros((x.size, p + 1))
    for i in range(p + 1):
        others = [m for m in range(p + 1) if m != i]
        denom = np.prod([nodes[i] - nodes[m] for m in others])
        num = np.ones(x.size)
        for m in others:
            num = num * (x - nodes[m])
        L[:, i] = num / denom
        d = np.zeros(x.size)
        for skip in others:
            t = np.ones(x.size)
            for m in others:
                if m != skip:
                    t = t * (x - nodes[m])
            d += t
        dL[:, i] = d / denom
    return L, dL


# ----------------------------------------------------------------------------
# quadrature
# ----------------------------------------------------------------------------
def _gl(n: int):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def _gj01(n: int, alpha: int):
    """n-point Gauss-Jacobi rule on [0,1] for weight (1-t)^alpha."""
    x, w = roots_jacobi(n, alpha, 0)
    return (x + 1.0) / 2.0, w / 2.0 ** (alpha + 1)


def _npts(degree: int) -> int:
    return max(1, int(math.ceil((degree + 1) / 2)))


def quadrature(family: str, degree: int):
    """Quadrature rule exact to `degree` on the reference cell of `family`.
    Returns (points (q,3), weights (q,))."""
    if degree < 0:
        degree = 0
    if family == "hex":
        x, w = _gl(_npts(degree))
        Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
        WZ, WY, WX = np.meshgrid(w, w, w, indexing="ij")
        pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        return pts, (WX * WY * WZ).ravel()
    if family == "quad":
        x, w = _gl(_npts(degree))
        Y, X = np.meshgrid(x, x, indexing="ij")
        WY, WX = np.meshgrid(w, w, indexing="ij")
        pts = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
        return pts, (WX * WY).ravel()
    if family == "line":
        x, w = _gl(_npts(degree))
        return np.stack([x, 0 * x, 0 * x], axis=1), w
    if family == "point":
        return np.zeros((1, 3)), np.ones(1)
    if family == "tet":
        if degree <= 1:
            return np.array([[0.25, 0.25, 0.25]]), np.array([1.0 / 6.0])
        if degree == 2:
            a = (5.0 + 3.0 * math.sqrt(5.0)) / 20.0
            b = (5.0 - math.sqrt(5.0)) / 20.0
            pts = np.array([[b, b, b], [a, b, b], [b, a, b], [b, b, a]])
            return pts, np.full(4, 1.0 / 24.0)
        n = _npts(degree)
        t1, w1 = _gj01(n, 2)
        t2, w2 = _gj01(n, 1)
        t3, w3 = _gj01(n, 0)
        pts, wts = [], []
        for i in range(n):
            for j in range(n):
                for k in range(n):
                    z = t1[i]
                    y = t2[j] * (1.0 - z)
                    x = t3[k] * (1.0 - z) * (1.0 - t2[j])
                    pts.append((x, y, z))
                    wts.append(w1[i] * w2[j] * w3[k])
        return np.array(pts), np.array(wts)
    if family == "tri":
        if degree <= 1:
            return np.array([[1 / 3, 1 / 3, 0.0]]), np.array([0.5])
        if degree == 2:
            pts = np.array([[1 / 6, 1 / 6, 0.0], [2 / 3, 1 / 6, 0.0], [1 / 6, 2 / 3, 0.0]])
            return pts, np.full(3, 1.0 / 6.0)
        n = _npts(degree)
        t1, w1 = _gj01(n, 1)
        t2, w2 = _gj01(n, 0)
        pts, wts = [], []
        for i in range(n):
            for k in range(n):
                y = t1[i]
                x = t2[k] * (1.0 - y)
                pts.append((x, y, 0.0))
                wts.append(w1[i] * w2[k])
        return np.array(pts), np.array(wts)
    if family == "prism":
        tp, tw = quadrature("tri", degree)
        x, w = _gl(_npts(degree))
        pts, wts = [], []
        for k in range(x.size):
            for i in range(tw.size):
                pts.append((tp[i, 0], tp[i, 1], x[k]))
                wts.append(tw[i] * w[k])
        return np.array(pts), np.array(wts)
    raise ValueError(f"no quadrature for family {family}")


def gauss_degree(order: int, gauss=None) -> int:
    """Rule degree the reference selects: "Gauss"*(2*order+1) for the default
    ``gauss=:full`` (src/linear.jl:235, src/multifield.jl:2927-2930),
    2*order-1 for ``:reduced``, and 2*order+1+g for an integer offset g."""
    if gauss is None or gauss == "full":
        return 2 * order + 1
    if gauss == "reduced":
        return 2 * order - 1
    return 2 * order + 1 + int(gauss)


# ----------------------------------------------------------------------------
# bases
# ----------------------------------------------------------------------------
def basis(etype: int, pts: np.ndarray):
    """Lagrange values N (q, n) and reference gradients dN (q, n, 3)."""
    name, dim, order, npe, family = element_properties(etype)
    pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
    q = pts.shape[0]
    N = np.zeros((q, npe))
    dN = np.zeros((q, npe, 3))
    if family == "hex":
        lat = hex_node_lattice(order)
        Lx, dLx = _lagrange_1d(order, pts[:, 0])
        Ly, dLy = _lagrange_1d(order, pts[:, 1])
        Lz, dLz = _lagrange_1d(order, pts[:, 2])
        i, j, k = lat[:, 0], lat[:, 1], lat[:, 2]
        N[:] = Lx[:, i] * Ly[:, j] * Lz[:, k]
        dN[:, :, 0] = dLx[:, i] * Ly[:, j] * Lz[:, k]
        dN[:, :, 1] = Lx[:, i] * dLy[:, j] * Lz[:, k]
        dN[:, :, 2] = Lx[:, i] * Ly[:, j] * dLz[:, k]
        return N, dN
    if family == "quad":
        lat = quad_node_lattice(order)
        Lx, dLx = _lagrange_1d(order, pts[:, 0])
        Ly, dLy = _lagrange_1d(order, pts[:, 1])
        i, j = lat[:, 0], lat[:, 1]
        N[:] = Lx[:, i] * Ly[:, j]
        dN[:, :, 0] = dLx[:, i] * Ly[:, j]
        dN[:, :, 1] = Lx[:, i] * dLy[:, j]
        return N, dN
    if family in ("tet", "tri"):
        nv = 4 if family == "tet" else 3
        edges = TET_EDGES if family == "tet" else TRI_EDGES
        x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
        if family == "tet":
            lam = np.stack([1 - x - y - z, x, y, z], axis=1)
            dlam = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
        else:
            lam = np.stack([1 - x - y, x, y], axis=1)
            dlam = np.array([[-1, -1, 0], [1, 0, 0], [0, 1, 0]], float)
        if order == 1:
            N[:] = lam
            dN[:] = dlam[None, :, :]
            return N, dN
        if order == 2:
            for a in range(nv):
                N[:, a] = lam[:, a] * (2 * lam[:, a] - 1)
                dN[:, a, :] = (4 * lam[:, a] - 1)[:, None] * dlam[a][None, :]
            for m, (a, b) in enumerate(edges):
                N[:, nv + m] = 4 * lam[:, a] * lam[:, b]
                dN[:, nv + m, :] = 4 * (lam[:, a][:, None] * dlam[b][None, :]
                                        + lam[:, b][:, None] * dlam[a][None, :])
            return N, dN
    if family == "prism":
        u, v, w = pts[:, 0], pts[:, 1], pts[:, 2]
        lam = [1 - u - v, u, v]
        dlam = [(-1.0, -1.0), (1.0, 0.0), (0.0, 1.0)]
        for lvl, s in enumerate((-1.0, 1.0)):
            fz = (1 + s * w) / 2
            for a in range(3):
                n = 3 * lvl + a
                N[:, n] = lam[a] * fz
                dN[:, n, 0] = dlam[a][0] * fz
                dN[:, n, 1] = dlam[a][1] * fz
                dN[:, n, 2] = lam[a] * s / 2
        return N, dN
    if family == "line":
        L, dL = _lagrange_1d(order, pts[:, 0])
        perm = [0, order] + list(range(1, order))
        N[:] = L[:, perm]
        dN[:, :, 0] = dL[:, perm]
        return N, dN
    if family == "point":
        N[:] = 1.0
        return N, dN
    raise ValueError(f"no basis for element type {etype}")


def node_coords(etype: int) -> np.ndarray:
    """Reference coordinates of the element's nodes (n, 3)."""
    name, dim, order, npe, family = element_properties(etype)
    if family == "hex":
        return -1.0 + 2.0 * hex_node_lattice(order) / order
    if family == "quad":
        lat = -1.0 + 2.0 * quad_node_lattice(order) / order
        return np.concatenate([lat, np.zeros((lat.shape[0], 1))], axis=1)
    if family == "tet":
        V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
        if order == 1:
            return V
        return np.concatenate([V, np.array([(V[a] + V[b]) / 2 for a, b in TET_EDGES])])
    if family == "tri":
        V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
        if order == 1:
            return V
        return np.concatenate([V, np.array([(V[a] + V[b]) / 2 for a, b in TRI_EDGES])])
    if family == "prism":
        return np.array([[0, 0, -1], [1, 0, -1], [0, 1, -1],
                         [0, 0, 1], [1, 0, 1], [0, 1, 1]], float)
    if family == "line":
        x = np.concatenate([[-1.0, 1.0], np.linspace(-1.0, 1.0, order + 1)[1:-1]])
        return np.stack([x, 0 * x, 0 * x], axis=1)
    if family == "point":
        return np.zeros((1, 3))
    raise ValueError(etype)


class RuleTables:
    """Quadrature + basis tables of one element type, in gmsh layout."""

    def __init__(self, etype: int, degree: int):
        name, dim, order, npe, family = element_properties(etype)
        self.etype, self.degree, self.npe, self.dim = etype, degree, npe, dim
        self.points, self.weights = quadrature(family, degree)
        self.q = self.weights.size
        N, dN = basis(etype, self.points)
        self.N = np.ascontiguousarray(N)            # (q, n)
        self.dN = np.ascontiguousarray(dN)          # (q, n, 3)


@lru_cache(maxsize=None)
def rule_tables(etype: int, degree: int) -> RuleTables:
    return RuleTables(etype, degree)
