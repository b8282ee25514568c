"""Pins of the oracle's body-force and pressure loads (oracle.body_force,
oracle.pressure) by what the mathematics fixes, on the CPU:

- consistent nodal loads of single affine elements under a uniform body
  force: hex8 V/8 per node, hex27 V (1/6, 2/3, 1/6)^3, tet10 -V/20 on the
  corners and V/5 on the edge nodes (the textbook values);
- the resultant of a uniform body force on the unit cube is g (the
  reference's test/test_dsl.jl:50-66 checks [1, 1, 1]), and for g(x) = x the
  resultant is the first moment ∫ x dV;
- element subsets add up to the whole;
- pressure on a flat face is the traction -p n (the oracle's traction
  routine, pinned separately), with resultant -p n A; p given per node as a
  constant equals the constant p;
- pressure on a quarter-cylinder surface has resultant -p R L (1, 1, 0) up
  to the discretisation error, sign fixed by the faces' orientation;
- on a closed box surface the corner and edge nodes average the normals of
  their faces, and the resultant vanishes;
- line loads (oracle.traction on line elements, Ja = |dx/dxi|): the
  Newton-Cotes nodal weights on straight equispaced elements of order 1-4,
  resultant q L and first moment q L^2 / 2 when the interior nodes are
  moved; point loads (Ja = 1) are the force at the node."""
import numpy as np

from lowlevelfem import mesh as M
from lowlevelfem import refelem as R
from oracle import oracle as O


def _affine(mesh, A, b):
    return M.Mesh(mesh.coords @ A.T + b, mesh.etype, mesh.conn, mesh.faces)


A0 = np.array([[1.3, 0.2, -0.1], [0.1, 0.9, 0.3], [-0.2, 0.15, 1.1]])


def test_hex8_single_element_v_over_8():
    m = _affine(M.structured_box_mesh(n=1), A0, np.array([0.3, -0.2, 0.5]))
    V = abs(np.linalg.det(A0))
    g = np.array([1.5, -2.0, 0.7])
    f = O.body_force(m, R.rule_tables(m.etype, 3), g).reshape(-1, 3)
    assert np.allclose(f, np.tile(g * V / 8, (8, 1)), rtol=1e-13, atol=0)


def test_hex27_single_element_tensor_weights():
    m = _affine(M.structured_box_mesh(n=1, order=2), A0, np.zeros(3))
    V = abs(np.linalg.det(A0))
    f = O.body_force(m, R.rule_tables(m.etype, 5), [0.0, 0.0, 1.0]).reshape(-1, 3)[:, 2]
    w1 = {0: 1 / 6, 1: 2 / 3, 2: 1 / 6}
    lat = np.rint(m.coords @ np.linalg.inv(A0).T * 2).astype(int)   # lattice index 0..2 per axis
    want = np.array([w1[i] * w1[j] * w1[k] * V for i, j, k in lat])
    assert np.allclose(f, want, rtol=1e-13, atol=1e-15)


def test_tet10_corner_and_edge_loads():
    m = M.structured_box_mesh(n=1, elem="tet", order=2)
    sub = M.Mesh(m.coords, m.etype, m.conn[:1])
    X = m.coords[m.conn[0, :4]]
    V = abs(np.linalg.det(X[1:] - X[0])) / 6
    f = O.body_force(sub, R.rule_tables(m.etype, 5), [2.0, 0.0, 0.0]).reshape(-1, 3)[:, 0]
    assert np.allclose(f[m.conn[0, :4]], -2.0 * V / 20, rtol=1e-12, atol=0)
    assert np.allclose(f[m.conn[0, 4:]], 2.0 * V / 5, rtol=1e-12, atol=0)


def test_unit_cube_resultant_and_first_moment():
    for kw, deg in ((dict(n=3), 3), (dict(n=2, elem="tet", order=2), 5), (dict(n=2, elem="prism"), 3),
                    (dict(n=3, perturb=0.2, seed=1), 3)):
        m = M.structured_box_mesh(**kw)
        tab = R.rule_tables(m.etype, deg)
        f = O.body_force(m, tab, [1.0, 1.0, 1.0]).reshape(-1, 3)
        assert np.allclose(f.sum(axis=0), [1.0, 1.0, 1.0], rtol=1e-12, atol=1e-12)
        # g(x) = x: resultant = ∫ x dV = (1/2, 1/2, 1/2) on the unit cube
        P = O.volume_points(m, tab)
        f2 = O.body_force(m, tab, g_gp=P).reshape(-1, 3)
        assert np.allclose(f2.sum(axis=0), [0.5, 0.5, 0.5], rtol=1e-12, atol=1e-12)


def test_element_subsets_add_up():
    m = M.structured_box_mesh(n=3, elem="tet", order=1)
    tab = R.rule_tables(m.etype, 3)
    rng = np.random.default_rng(0)
    sel = rng.random(m.E) < 0.4
    a = O.body_force(m, tab, [0.3, 1.0, -2.0], elems=np.flatnonzero(sel))
    b = O.body_force(m, tab, [0.3, 1.0, -2.0], elems=np.flatnonzero(~sel))
    full = O.body_force(m, tab, [0.3, 1.0, -2.0])
    assert np.allclose(a + b, full, rtol=1e-13, atol=1e-16)


def _face_tables(ftype, order):
    return R.rule_tables(ftype, 2 * order + 1), R.basis(ftype, R.node_coords(ftype))[1]


def test_flat_face_pressure_is_minus_p_n_traction():
    for kw in (dict(lx=2.0, ly=1.0, lz=1.5, cells=(3, 2, 2)), dict(n=2, order=2),
               dict(n=2, elem="tet", order=2)):
        m = M.structured_box_mesh(**kw)
        fs = m.faces["right"]
        order = R.element_properties(fs.etype)[2]
        tab, dNn = _face_tables(fs.etype, order)
        n = O.face_node_normals(m, fs, dNn)
        nhat = n[0, 0]
        assert np.allclose(n, nhat, atol=1e-15)
        p = 3.5
        f, tn = O.pressure(m, fs, tab, dNn, p)
        ft = O.traction(m, fs, tab, -p * nhat)
        assert np.allclose(f, ft, rtol=1e-13, atol=1e-15)
        area = (m.coords[:, 1].max() - m.coords[:, 1].min()) * (m.coords[:, 2].max() - m.coords[:, 2].min())
        assert np.allclose(f.reshape(-1, 3).sum(axis=0), -p * nhat * area, rtol=1e-13, atol=1e-14)
        fn, _ = O.pressure(m, fs, tab, dNn, 0.0, p_nodal=np.full(m.N, p))
        assert np.array_equal(f, fn)


def test_quarter_cylinder_resultant():
    R1, R2, L, p = 1.0, 1.4, 2.0, 2.5
    m = M.structured_box_mesh(cells=(2, 12, 4), order=2)
    x, y, z = m.coords.T
    r, th = R1 + (R2 - R1) * x, 0.5 * np.pi * y
    m = M.Mesh(np.stack([r * np.cos(th), r * np.sin(th), L * z], axis=1), m.etype, m.conn, m.faces)
    fs = m.faces["right"]                         # r = R2
    tab, dNn = _face_tables(fs.etype, 2)
    f, _ = O.pressure(m, fs, tab, dNn, p)
    n = O.face_node_normals(m, fs, dNn)
    X = m.coords[fs.conn[:, 0]]
    s = np.sign(np.einsum("fi,fi->f", n[:, 0, :2], X[:, :2]))
    assert np.all(s == s[0])                      # one orientation over the face set
    want = -p * s[0] * R2 * L * np.array([1.0, 1.0, 0.0])
    got = f.reshape(-1, 3).sum(axis=0)
    assert np.allclose(got, want, rtol=1e-4, atol=1e-9)


def test_closed_box_nodal_normals_average():
    m = M.structured_box_mesh(n=1)
    faces = [m.faces[k] for k in ("left", "right", "bottom", "top", "rear", "front")]
    conn = np.concatenate([fs.conn for fs in faces])
    # orient every face outward, as a closed surface's faces are
    X = m.coords[conn]
    cen = X.mean(axis=1) - 0.5
    tab, dNn = _face_tables(3, 1)
    allf = M.FaceSet(3, conn)
    n = O.face_node_normals(m, allf, dNn)[:, 0]
    flip = np.einsum("fi,fi->f", n, cen) < 0
    conn[flip] = conn[flip][:, ::-1]
    allf = M.FaceSet(3, conn)
    f, tn = O.pressure(m, allf, tab, dNn, 1.0)
    # every corner of the unit cube touches three faces: t = -(outward corner direction)/3
    corner = np.sign(m.coords - 0.5)
    assert np.allclose(tn, -corner / 3, rtol=0, atol=1e-15)
    assert np.allclose(f.reshape(-1, 3).sum(axis=0), 0.0, atol=1e-14)


# ---- line and point loads (src/linear.jl:3385-3388) ------------------------
def _segment_mesh(order, shift, L=3.0, d=(1.0, 2.0, 2.0)):
    """One straight segment of length L along d in 3D as one line element of
    the given order; the interior nodes sit at the equispaced positions plus
    `shift` (fractions of L), so the parametrisation is not affine when
    shift != 0.  Returns (mesh, faceset, arc-length coordinate per node)."""
    u = np.asarray(d) / np.linalg.norm(d)
    s = np.concatenate([[0.0, L], np.linspace(0.0, L, order + 1)[1:-1] + shift * L])
    coords = np.array([0.5, -0.2, 1.0]) + s[:, None] * u[None, :]
    et = {1: 1, 2: 8, 3: 26, 4: 27}[order]
    fs = M.FaceSet(et, np.arange(order + 1, dtype=np.int64)[None, :])
    return M.Mesh(coords, 4, np.zeros((0, 4), np.int64)), fs, s


def test_line_load_textbook_nodal_values():
    """Uniform line load q on a straight element with equispaced nodes: the
    consistent nodal loads are q L times the Newton-Cotes weights (trapezoid,
    Simpson, Simpson 3/8, Boole)."""
    want = {1: [1 / 2, 1 / 2], 2: [1 / 6, 1 / 6, 2 / 3], 3: [1 / 8, 1 / 8, 3 / 8, 3 / 8],
            4: [7 / 90, 7 / 90, 32 / 90, 12 / 90, 32 / 90]}
    t = np.array([0.7, -1.1, 2.5])
    for order, wts in want.items():
        m, fs, s = _segment_mesh(order, 0.0)
        f = O.traction(m, fs, R.rule_tables(fs.etype, 2 * order + 1), t).reshape(-1, 3)
        assert np.allclose(f, 3.0 * np.outer(wts, t), rtol=1e-13, atol=1e-15), order


def test_line_load_resultant_and_moment_on_curved_parametrisation():
    """With the interior nodes moved along the segment, x(xi) is not affine
    and |dx/dxi| varies; still sum_a f_a = q L and sum_a f_a s_a = q L^2 / 2
    (the integrands are polynomials the rule integrates exactly), which a
    missing or wrong Ja would break."""
    for order, shift in ((2, 0.12), (3, 0.05), (1, 0.0)):
        m, fs, s = _segment_mesh(order, shift)
        f = O.traction(m, fs, R.rule_tables(fs.etype, 2 * order + 1), [2.0, 0.0, 0.0]).reshape(-1, 3)[:, 0]
        assert abs(f.sum() - 2.0 * 3.0) < 1e-13
        if order <= 2:
            assert abs(f @ s - 2.0 * 3.0 ** 2 / 2) < 1e-12


def test_line_load_function_of_position():
    """q = s (arc length) on one line3 with its mid-node moved: sum f = L^2 / 2
    and the function-valued path equals the constant one for a constant."""
    m, fs, s = _segment_mesh(2, 0.1)
    u = np.array([1.0, 2.0, 2.0]) / 3.0
    x0 = np.array([0.5, -0.2, 1.0])
    tab = R.rule_tables(fs.etype, 5)
    f, T = O.traction_fn(m, fs, tab, [lambda x, y, z: float(np.dot([x, y, z] - x0, u)), None, None])
    assert abs(f.reshape(-1, 3)[:, 0].sum() - 3.0 ** 2 / 2) < 1e-12
    fc, _ = O.traction_fn(m, fs, tab, [0.5, None, -1.0])
    assert np.allclose(fc, O.traction(m, fs, tab, [0.5, 0.0, -1.0]), rtol=0, atol=1e-15)


def test_point_load_is_the_force_at_its_node():
    """A point element carries the force itself (N = 1, Ja = 1, w = 1); a
    node listed twice gets it twice."""
    m = M.structured_box_mesh(n=2)
    fs = M.FaceSet(15, np.array([[3], [7], [7]], dtype=np.int64))
    tab = R.rule_tables(15, 1)
    assert tab.q == 1 and tab.weights[0] == 1.0 and tab.N[0, 0] == 1.0
    t = np.array([1.5, -2.0, 0.25])
    f = O.traction(m, fs, tab, t).reshape(-1, 3)
    want = np.zeros_like(f)
    want[3] = t
    want[7] = 2 * t
    assert np.array_equal(f, want)
