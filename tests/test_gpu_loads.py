"""Body force, volume Gauss points and pressure on the GPU against the oracle
(oracle.body_force / volume_points / pressure, pinned in
test_oracle_loads.py), and the mirror's loadVector / ∫(Pu ⋅ g) on top of
them, including the reference's resultant check of test/test_dsl.jl:50-66
(∫(Pu ⋅ [1, 1, 1]) over the unit cube and over its face "right" both sum to
[1, 1, 1]); line loads on curve groups and point loads on point groups
(oracle.traction with Ja = |dx/dxi| or 1, pinned in test_oracle_loads.py).  Loads are FP64 sums of a few terms per node: the bar is 1e-13
relative to the largest entry; repeated calls are bit-identical."""
import sys

import numpy as np
import pytest

import lowlevelfem as L
from lowlevelfem import _lib
from lowlevelfem import mesh as M
from lowlevelfem import refelem as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def close(a, b, tol=1e-13):
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)


BODY = [("hex8-perturbed", dict(n=4, perturb=0.2, seed=2), 3), ("tet10", dict(n=3, elem="tet", order=2), 5),
        ("prism6", dict(n=3, elem="prism"), 3), ("hex27", dict(n=2, order=2), 5), ("tet4", dict(n=3, elem="tet"), 3)]


@pytest.mark.parametrize("label,kw,deg", BODY, ids=[b[0] for b in BODY])
def test_body_force_matches_oracle(gpu_ctx, label, kw, deg):
    say(f"[body] {label}")
    m = M.structured_box_mesh(**kw)
    tab = R.rule_tables(m.etype, deg)
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    g = np.array([0.4, -1.3, 2.2])
    f = dev.body(tab, g=g)
    assert close(f, O.body_force(m, tab, g))
    assert np.array_equal(f, dev.body(tab, g=g))
    # Gauss points and a function-valued force on a subset of the elements
    rng = np.random.default_rng(5)
    sel = np.flatnonzero(rng.random(m.E) < 0.5)
    P = dev.volume_points(tab, sel)
    assert close(P, O.volume_points(m, tab, sel), 1e-15)
    G = np.stack([np.sin(P[..., 0]), P[..., 1] * P[..., 2], np.full(P.shape[:2], -0.5)], axis=-1)
    fs = dev.body(tab, g_gp=G, elems=sel)
    assert close(fs, O.body_force(m, tab, g_gp=G, elems=sel))
    dev.close()


def test_body_force_rejects_bad_sets(gpu_ctx):
    m = M.structured_box_mesh(n=2)
    tab = R.rule_tables(m.etype, 3)
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    with pytest.raises(_lib.LLFEMError, match="out of range"):
        dev.body(tab, g=[1, 0, 0], elems=[0, m.E])
    with pytest.raises(_lib.LLFEMError, match="listed twice"):
        dev.body(tab, g=[1, 0, 0], elems=[1, 1])
    dev.close()


def _face_tables(fs):
    order = R.element_properties(fs.etype)[2]
    return R.rule_tables(fs.etype, 2 * order + 1), R.basis(fs.etype, R.node_coords(fs.etype))[1]


def _cylinder(cells=(2, 10, 3), order=2):
    m = M.structured_box_mesh(cells=cells, order=order)
    x, y, z = m.coords.T
    r, th = 1.0 + 0.4 * x, 0.5 * np.pi * y
    return M.Mesh(np.stack([r * np.cos(th), r * np.sin(th), 2.0 * z], axis=1), m.etype, m.conn, m.faces)


PRESS = [("hex8-flat", lambda: M.structured_box_mesh(lx=2.0, cells=(3, 2, 2)), "right"),
         ("hex27-flat", lambda: M.structured_box_mesh(n=2, order=2), "top"),
         ("tet10-flat", lambda: M.structured_box_mesh(n=2, elem="tet", order=2), "front"),
         ("hex27-cylinder", _cylinder, "right"),
         ("hex8-cylinder", lambda: _cylinder((2, 12, 3), 1), "right")]


@pytest.mark.parametrize("label,make,face", PRESS, ids=[p[0] for p in PRESS])
def test_pressure_matches_oracle(gpu_ctx, label, make, face):
    say(f"[pressure] {label}")
    m = make()
    fs = m.faces[face]
    tab, dNn = _face_tables(fs)
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    f = dev.pressure(fs.etype, fs.conn, tab, dNn, p=2.5)
    fo, _ = O.pressure(m, fs, tab, dNn, 2.5)
    assert close(f, fo)
    assert np.array_equal(f, dev.pressure(fs.etype, fs.conn, tab, dNn, p=2.5))
    pn = 1.0 + m.coords[:, 0] ** 2 + 0.5 * m.coords[:, 2]
    fn = dev.pressure(fs.etype, fs.conn, tab, dNn, p_nodal=pn)
    fno, _ = O.pressure(m, fs, tab, dNn, p_nodal=pn)
    assert close(fn, fno)
    dev.close()


def test_dsl_linear_form_resultants(gpu_ctx):
    """test/test_dsl.jl:50-66: ∫(Pu ⋅ [1,1,1]) over the unit cube and over
    its face "right" each sum to [1, 1, 1]."""
    L.structured_box_mesh(n=2, order=1)
    Pu = L.Problem([L.Material("body")], type="VectorField", dim=3, field="u")
    fv = L.integral(Pu, [1.0, 1.0, 1.0])
    assert np.allclose(fv.a[:, 0].reshape(-1, 3).sum(axis=0), [1.0, 1.0, 1.0], rtol=1e-12, atol=1e-12)
    ft = L.integral(Pu, [1.0, 1.0, 1.0], Γ="right")
    assert np.allclose(ft.a[:, 0].reshape(-1, 3).sum(axis=0), [1.0, 1.0, 1.0], rtol=1e-12, atol=1e-12)
    # the same vectors through loadVector (volume group "body" and face "right")
    fb = L.loadVector(Pu, [L.load("body", fx=1.0, fy=1.0, fz=1.0)])
    assert np.array_equal(fb.a, fv.a)
    fr = L.loadVector(Pu, [L.load("right", fx=1.0, fy=1.0, fz=1.0)])
    assert close(fr.a[:, 0], ft.a[:, 0], 1e-15)


def test_loadvector_body_pressure_and_function(gpu_ctx):
    m = L.structured_box_mesh(lx=2.0, ly=1.0, lz=1.0, n=3, order=2)
    P = L.Problem([L.Material("body", E=2e5, nu=0.3)])
    tab = R.rule_tables(m.etype, 5)
    g = lambda x, y, z: 1.0 + x * y                             # noqa: E731
    f = L.loadVector(P, [L.load("body", fz=g), L.load("top", p=3.0)])
    fs = m.faces["top"]
    ftab, dNn = _face_tables(fs)
    Gp = O.volume_points(m, tab)
    G = np.zeros(Gp.shape)
    G[..., 2] = 1.0 + Gp[..., 0] * Gp[..., 1]
    want = O.body_force(m, tab, g_gp=G) + O.pressure(m, fs, ftab, dNn, 3.0)[0]
    assert close(f.a[:, 0], want)
    with pytest.raises(ValueError, match="pressure can be given on a surface"):
        L.loadVector(P, [L.load("body", p=1.0)])


# ---- line and point loads (src/linear.jl:3385-3388) ------------------------
EDGE = [("hex8", dict(cells=(5, 2, 2), lx=2.0)), ("tet10", dict(n=2, elem="tet", order=2)),
        ("hex64", dict(n=2, order=3)), ("hex125", dict(n=2, order=4))]


@pytest.mark.parametrize("label,kw", EDGE, ids=[e[0] for e in EDGE])
def test_line_and_point_loads_match_oracle(gpu_ctx, label, kw):
    """Line loads on the box edge (line2 ... line5 elements, their interior
    nodes moved along the edge so |dx/dxi| varies) and point loads: constant
    and function-valued, against the oracle; the resultant of a uniform
    line load is q L."""
    from mixed_mesh import with_edge_and_corner
    say(f"[edge] {label}")
    m = with_edge_and_corner(M.structured_box_mesh(**kw))
    fs = m.faces["edge"]
    order = R.element_properties(fs.etype)[2]
    if order > 1:            # interior edge nodes slide along x: non-affine x(xi)
        x = m.coords.copy()
        inner = fs.conn[:, 2:].ravel()
        x[inner, 0] += 0.08 * (x[fs.conn[:, 1], 0] - x[fs.conn[:, 0], 0]).repeat(order - 1) / order
        m = M.Mesh(x, m.etype, m.conn, m.faces)
    tab = R.rule_tables(fs.etype, 2 * order + 1)
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    t = np.array([1.5, -0.5, 2.0])
    f = dev.traction(fs.etype, fs.conn, tab, t)
    assert close(f, O.traction(m, fs, tab, t))
    Lx = m.coords[:, 0].max() - m.coords[:, 0].min()
    assert np.allclose(f.reshape(-1, 3).sum(axis=0), t * Lx, rtol=1e-13, atol=0)
    P = dev.face_points(fs.etype, fs.conn, tab)
    assert close(P, O.face_points(m, fs, tab), 1e-15)
    fns = [lambda x, y, z: x * x, None, lambda x, y, z: 1.0 - x]
    ref, T = O.traction_fn(m, fs, tab, fns)
    assert close(dev.traction_gp(fs.etype, fs.conn, tab, T), ref)
    for name in ("corner", "corners2"):
        ps = m.faces[name]
        ptab = R.rule_tables(15, 1)
        fp = dev.traction(15, ps.conn, ptab, t)
        assert np.array_equal(fp, O.traction(m, ps, ptab, t))
    dev.close()


@pytest.mark.parametrize("binary", [False, True], ids=["ascii", "binary"])
def test_loadvector_on_curve_and_point_groups_from_msh(gpu_ctx, tmp_path, binary):
    """The mirror's loadVector on a curve group (line load, one component a
    function of position) and on point groups, read back from an MSH file
    with physical curves and points; a point group also serves as a
    support; pressure on a curve is refused as in the reference."""
    from mixed_mesh import with_edge_and_corner, write_msh
    say(f"[edge-msh] binary={binary}")
    m0 = with_edge_and_corner(M.structured_box_mesh(n=2, elem="tet", order=2))
    m0.volumes = {"body": np.arange(m0.E)}
    path = str(tmp_path / ("edges.msh" + ("b" if binary else "")))
    write_msh(m0, path, binary=binary)
    m = L.read_msh(path)
    assert m.faces["edge"].etype == 8 and m.faces["corner"].etype == 15
    assert np.array_equal(m.faces["corners2"].conn, m0.faces["corners2"].conn)
    P = L.Problem([L.Material("body", E=2e5, nu=0.3)])
    f = L.loadVector(P, [L.load("edge", fx=2.0, fy=lambda x, y, z: 1.0 + x),
                         L.load("corners2", fz=-3.0)])
    fs, ps = m.faces["edge"], m.faces["corners2"]
    tab = R.rule_tables(8, 5)
    want, _ = O.traction_fn(m, fs, tab, [2.0, lambda x, y, z: 1.0 + x, None])
    want = want + O.traction(m, ps, R.rule_tables(15, 1), [0.0, 0.0, -3.0])
    assert close(f.a[:, 0], want)
    assert np.isclose(f.a[:, 0].reshape(-1, 3)[:, 0].sum(), 2.0, rtol=1e-13)
    assert np.array_equal(L.nodesOnPhysicalGroup(P, "corner"), m0.faces["corner"].conn[0])
    with pytest.raises(ValueError, match="pressure can be given on a surface"):
        L.loadVector(P, [L.load("edge", p=1.0)])
