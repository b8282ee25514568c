"""Reference-element tables (the gmsh substitute): quadrature exactness,
partition of unity, nodal interpolation, and the reference's own quadrature
KATs (test/test_integration.jl:21-79)."""
import itertools

import numpy as np
import pytest

from lowlevelfem import refelem as R
from lowlevelfem import mesh as M

VOLUME = {"hex": 8.0, "quad": 4.0, "tet": 1 / 6, "tri": 0.5, "prism": 1.0, "line": 2.0}


def _exact_monomial(family, a, b, c):
    from math import factorial
    if family == "hex":
        f = lambda k: 0.0 if k % 2 else 2.0 / (k + 1)
        return f(a) * f(b) * f(c)
    if family == "quad":
        f = lambda k: 0.0 if k % 2 else 2.0 / (k + 1)
        return f(a) * f(b) * (1.0 if c == 0 else 0.0)
    if family == "tet":
        return factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)
    if family == "tri":
        return factorial(a) * factorial(b) / factorial(a + b + 2) * (1.0 if c == 0 else 0.0)
    if family == "prism":
        f = lambda k: 0.0 if k % 2 else 2.0 / (k + 1)
        return factorial(a) * factorial(b) / factorial(a + b + 2) * f(c)
    if family == "line":
        return (0.0 if a % 2 else 2.0 / (a + 1)) * (1.0 if b == c == 0 else 0.0)
    raise ValueError(family)


@pytest.mark.parametrize("family,deg", [(f, d) for f in ("hex", "quad", "tet", "tri", "prism", "line")
                                        for d in (1, 2, 3, 5, 9)])
def test_quadrature_exact(family, deg):
    pts, w = R.quadrature(family, deg)
    assert np.all(w > 0)
    assert abs(w.sum() - VOLUME[family]) < 1e-14
    dims = 3 if family in ("hex", "tet", "prism") else 2 if family in ("quad", "tri") else 1
    for a, b, c in itertools.product(range(deg + 1), repeat=3):
        if (dims <= 2 and c) or (dims == 1 and b):
            continue
        if family in ("tet", "tri") and a + b + c > deg:
            continue
        if family == "prism" and (a + b > deg or c > deg):
            continue
        if family in ("hex", "quad", "line") and max(a, b, c) > deg:
            continue
        got = np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c)
        assert abs(got - _exact_monomial(family, a, b, c)) < 1e-13, (a, b, c)


@pytest.mark.parametrize("etype", sorted(R.ELEMENT_TYPES))
def test_basis_partition_and_interpolation(etype):
    name, dim, order, npe, family = R.element_properties(etype)
    pts, _ = R.quadrature(family, 2 * order + 1)
    N, dN = R.basis(etype, pts)
    assert np.abs(N.sum(1) - 1).max() < 1e-13
    assert np.abs(dN.sum(1)).max() < 1e-12
    X = R.node_coords(etype)
    Nn, _ = R.basis(etype, X)
    assert np.abs(Nn - np.eye(npe)).max() < 1e-12
    # gradients reproduce linear functions: sum_a x_a dN_a = I
    G = np.einsum("qai,aj->qji", dN, X)
    eye = np.eye(3)[:dim, :dim]
    if dim:                                          # a point has no gradients
        assert np.abs(G[:, :dim, :dim] - eye[None]).max() < 1e-12


def test_gmsh_low_order_orderings():
    # hex8 / tet10 / prism6 vertex & edge conventions
    assert np.allclose(R.node_coords(5)[:8], [[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                                             [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])
    X = R.node_coords(11)
    assert np.allclose(X[8], (X[3] + X[2]) / 2) and np.allclose(X[9], (X[3] + X[1]) / 2)
    assert np.allclose(R.node_coords(12)[26], 0)
    assert R.rule_tables(5, 3).q == 8 and R.rule_tables(93, 9).q == 125


def _integrate(mesh, f, deg):
    tab = R.rule_tables(mesh.etype, deg)
    X = mesh.coords[mesh.conn]                       # (E, n, 3)
    J = np.einsum("ena,qnb->eqab", X, tab.dN)
    det = np.abs(np.linalg.det(J))
    x = np.einsum("ena,qn->eqa", X, tab.N)
    return np.sum(f(x[..., 0], x[..., 1], x[..., 2]) * det * tab.weights[None, :])


def _integrate_face(mesh, name, f, deg):
    fs = mesh.faces[name]
    tab = R.rule_tables(fs.etype, deg)
    X = mesh.coords[fs.conn]
    t1 = np.einsum("ena,qn->eqa", X, tab.dN[:, :, 0])
    t2 = np.einsum("ena,qn->eqa", X, tab.dN[:, :, 1])
    Ja = np.linalg.norm(np.cross(t1, t2), axis=-1)
    x = np.einsum("ena,qn->eqa", X, tab.N)
    return np.sum(f(x[..., 0], x[..., 1], x[..., 2]) * Ja * tab.weights[None, :])


@pytest.mark.parametrize("elem,order", [("hex", 1), ("hex", 2), ("tet", 1), ("tet", 2), ("prism", 1)])
def test_integration_kats(elem, order):
    """test/test_integration.jl: ∫(1+x+2y+3z) = 4 over the unit cube and 4.5 on
    x = 1; ∫(x²+y²+z²) = 1 with a degree-3 rule."""
    m = M.structured_box_mesh(n=2, order=order, elem=elem)
    f = lambda x, y, z: 1.0 + x + 2.0 * y + 3.0 * z
    assert abs(_integrate(m, f, 2) - 4.0) < 1e-12
    assert abs(_integrate_face(m, "right", f, 2) - 4.5) < 1e-12
    assert abs(_integrate(m, lambda x, y, z: x * x + y * y + z * z, 3) - 1.0) < 1e-12
