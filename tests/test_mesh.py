"""Mesh generators: node/element counts, orientation, structural nnz."""
import numpy as np
import pytest

from lowlevelfem import mesh as M
from lowlevelfem import refelem as R


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_hex8_structural_nnz_closed_form(n):
    # 9 (3n+1)^3 reproduces the reference's 15,944,049 at n = 40
    # (examples/dev/DSL_vs_legacy.ipynb:736)
    m = M.structured_box_mesh(n=n)
    assert M.structural_nnz(m) == 9 * (3 * n + 1) ** 3
    assert 9 * (3 * 40 + 1) ** 3 == 15_944_049


def test_reference_cell_rule_d_min():
    # structured_box_mesh uses d = min(dx, dy) on every axis (src/general.jl:8948)
    m = M.structured_box_mesh(lx=1, ly=1, lz=4, n=4)
    assert m.E == 4 * 4 * 16


@pytest.mark.parametrize("elem,order,npe", [("hex", 1, 8), ("hex", 2, 27), ("hex", 3, 64),
                                            ("hex", 4, 125), ("tet", 1, 4), ("tet", 2, 10),
                                            ("prism", 1, 6)])
def test_positive_jacobians_and_volume(elem, order, npe):
    m = M.structured_box_mesh(n=2, order=order, elem=elem, lx=1.0, ly=2.0, lz=3.0)
    assert m.npe == npe
    tab = R.rule_tables(m.etype, 2 * order + 1)
    X = m.coords[m.conn]
    J = np.einsum("ena,qnb->eqab", X, tab.dN)
    det = np.linalg.det(J)
    assert det.min() > 0
    assert abs(np.sum(det * tab.weights) - 6.0) < 1e-12


def test_kuhn_tet10_counts():
    m = M.structured_box_mesh(lz=4, cells=(4, 4, 16), elem="tet", order=2)
    assert m.N == 9 * 9 * 33 and m.E == 6 * 4 * 4 * 16
    # every face triangle set covers its face exactly once
    for name, area in (("left", 4.0), ("right", 4.0), ("rear", 1.0), ("top", 4.0)):
        fs = m.faces[name]
        v = m.coords[fs.conn[:, :3]]
        a = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
        assert abs(a.sum() - area) < 1e-12


def test_perturbed_interior_only():
    m = M.structured_box_mesh(n=3, order=2, perturb=0.05, seed=7)
    b = m.nodes_on("left")
    assert np.all(m.coords[b, 0] == 0.0)
    r = M.structured_box_mesh(n=3, order=2)
    assert np.abs(m.coords - r.coords).max() > 0
