"""Pins the CPU oracle (oracle/oracle.c) to what the reference fixes:
its recorded golden values, closed-form hex8 element matrices, the DSL ≡
legacy identity (test/test_dsl.jl:31), resultant identities
(test/test_dsl.jl:79-85), the solve residual definition (test_dsl.jl:106),
and brute-force pattern construction."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import hooke_symgrad_D
from lowlevelfem import mesh as M
from lowlevelfem import refelem as R
from oracle import oracle as O

E, NU = 2e5, 0.3
MU = E / (2 * (1 + NU))
LAM = 2 * MU * NU / (1 - 2 * NU)


def hex8_closed_form(h):
    s = (np.array(R.HEX_VERTS) * 2 - 1).astype(float)     # node signs, gmsh order
    I = np.zeros((8, 8, 3, 3))
    for a in range(8):
        for b in range(8):
            for i in range(3):
                for j in range(3):
                    if i == j:
                        I[a, b, i, i] = h / 16 * s[a, i] * s[b, i] * np.prod(
                            [1 + s[a, k] * s[b, k] / 3 for k in range(3) if k != i])
                    else:
                        k = 3 - i - j
                        I[a, b, i, j] = h / 16 * s[a, i] * s[b, j] * (1 + s[a, k] * s[b, k] / 3)
    K = np.zeros((24, 24))
    for a in range(8):
        for b in range(8):
            for i in range(3):
                for j in range(3):
                    K[3 * a + i, 3 * b + j] = (LAM * I[a, b, i, j] + MU * I[a, b, j, i]
                                               + (MU * np.trace(I[a, b]) if i == j else 0.0))
    return K


def test_hex8_element_closed_form():
    h = 0.25
    m = M.structured_box_mesh(lx=h, ly=h, lz=h, n=1)
    tab = R.rule_tables(5, 3)
    d = (3 * m.conn[0][:, None] + np.arange(3)).ravel()      # element-local DoF order
    K = O.stiffness_legacy(m, tab, dropzeros=False).toarray()[np.ix_(d, d)]
    Kc = hex8_closed_form(h)
    assert np.abs(K - Kc).max() < 1e-13 * np.abs(Kc).max()
    nz = O.stiffness_csc(m, tab, hooke_symgrad_D())
    cp, rv = O.csc_pattern(m)
    Kd = sp.csc_matrix((nz, rv, cp), shape=(24, 24)).toarray()[np.ix_(d, d)]
    assert np.abs(Kd - Kc).max() < 1e-13 * np.abs(Kc).max()


def test_golden_sum_abs_K_30():
    """sum(abs, K.A) = 4.0073076923076982e9 for structured_box_mesh(n=30)
    (examples/dev/compare_LLFEM_Ferrite_dealII_Gridap/results.txt:7)."""
    m = M.structured_box_mesh(n=30)
    tab = R.rule_tables(5, 3)
    cp, rv = O.csc_pattern(m)
    nz = O.stiffness_csc(m, tab, hooke_symgrad_D(), cp, rv, nthreads=4)
    assert abs(np.abs(nz).sum() / 4.0073076923076982e9 - 1) < 1e-13


def test_golden_corner_entries_40():
    """K[1,1], K[2,1], K[3,1] of the 40³ corner node
    (examples/dev/DSL_vs_legacy.ipynb:189); the gmsh corner node sits at
    z = 1, which here is node (0,0,40)."""
    m = M.structured_box_mesh(n=40)
    tab = R.rule_tables(5, 3)
    cp, rv = O.csc_pattern(m)
    nz = O.stiffness_csc(m, tab, hooke_symgrad_D(), cp, rv, nthreads=4)
    j = 41 * 41 * 40
    col = slice(cp[3 * j], cp[3 * j + 1])
    vals = dict(zip(rv[col], nz[col]))
    ref = [1175.2136752136657, 400.64102564102234, -400.64102564102285]
    for c in range(3):
        assert abs(vals[3 * j + c] / ref[c] - 1) < 1e-12
    assert cp[-1] == 15_944_049                  # structural nnz (DSL_vs_legacy.ipynb:736)


def test_golden_corner_entries_hex27_20():
    """Order-2 hexahedra: K[1,1], K[2,1], K[3,1] = 877.4928774928592,
    320.51282051281675, -320.51282051281373 at 20³ with Gauss5
    (examples/dev/multi-thread.ipynb:75), at the same gmsh corner (0,0,1)."""
    m = M.structured_box_mesh(n=20, order=2)
    tab = R.rule_tables(12, 5)
    cp, rv = O.csc_pattern(m)
    nz = O.stiffness_csc(m, tab, hooke_symgrad_D(), cp, rv, nthreads=4)
    j = 41 * 41 * 40
    col = slice(cp[3 * j], cp[3 * j + 1])
    vals = dict(zip(rv[col], nz[col]))
    ref = [877.4928774928592, 320.51282051281675, -320.51282051281373]
    for c in range(3):
        assert abs(vals[3 * j + c] / ref[c] - 1) < 1e-12


def test_golden_sum_u_30():
    """sum(u) = 0.07057241535937334: left clamped, traction (1,0,0) on
    right, 30³ hex8 (results.txt:7; Ferrite gives ...37218)."""
    m = M.structured_box_mesh(n=30)
    tab = R.rule_tables(5, 3)
    cp, rv = O.csc_pattern(m)
    nz = O.stiffness_csc(m, tab, hooke_symgrad_D(), cp, rv, nthreads=4)
    K = sp.csc_matrix((nz, rv, cp), shape=(3 * m.N, 3 * m.N))
    fs = m.faces["right"]
    f = O.traction(m, fs, R.rule_tables(fs.etype, 3), [1.0, 0.0, 0.0])
    left = m.nodes_on("left")
    fixed = np.concatenate([3 * left + c for c in range(3)])
    u, it, rel = O.pcg(K, f, fixed, reltol=1e-13)
    assert it > 0 and rel < 1e-12
    assert abs(u.sum() / 0.07057241535937334 - 1) < 1e-11
    free = np.setdiff1d(np.arange(3 * m.N), fixed)
    r = K @ u - f
    assert np.linalg.norm(r[free]) / np.linalg.norm(f[free]) < 1e-10
    assert np.all(u[fixed] == 0)


@pytest.mark.parametrize("elem,order", [("hex", 1), ("hex", 2), ("tet", 2), ("prism", 1)])
def test_dsl_equals_legacy(elem, order):
    """test/test_dsl.jl:28-31: ‖Kdsl − Klegacy‖/‖Klegacy‖ < 1e-12."""
    m = M.structured_box_mesh(n=2, elem=elem, order=order, perturb=0.1, seed=1)
    tab = R.rule_tables(m.etype, 2 * order + 1)
    Kl = O.stiffness_legacy(m, tab)
    cp, rv = O.csc_pattern(m)
    nz = O.stiffness_csc(m, tab, hooke_symgrad_D(), cp, rv)
    Kd = sp.csc_matrix((nz, rv, cp), shape=Kl.shape)
    assert spla.norm(Kd - Kl) / spla.norm(Kl) < 1e-12
    assert spla.norm(Kl - Kl.T) / spla.norm(Kl) < 1e-12   # test/test_assembly.jl:13


def test_threaded_equals_serial():
    """test/test_dsl.jl:36-45."""
    m = M.structured_box_mesh(n=4, elem="tet", order=2)
    tab = R.rule_tables(11, 2)
    cp, rv = O.csc_pattern(m)
    a = O.stiffness_csc(m, tab, hooke_symgrad_D(), cp, rv, nthreads=1)
    b = O.stiffness_csc(m, tab, hooke_symgrad_D(), cp, rv, nthreads=3, chunk=7)
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-12


@pytest.mark.parametrize("elem,order", [("hex", 1), ("tet", 2), ("prism", 1)])
def test_pattern_brute_force(elem, order):
    m = M.structured_box_mesh(n=2, elem=elem, order=order)
    cp, rv = O.csc_pattern(m)
    n = 3 * m.N
    dense = np.zeros((n, n), bool)
    for el in m.conn:
        d = (3 * el[:, None] + np.arange(3)).ravel()
        dense[np.ix_(d, d)] = True
    B = sp.csc_matrix(dense)
    B.sort_indices()
    assert np.array_equal(B.indptr, cp) and np.array_equal(B.indices, rv)


def test_sparse_restatement_matches_scipy():
    rng = np.random.default_rng(0)
    I = rng.integers(0, 20, 500); J = rng.integers(0, 20, 500); V = rng.normal(size=500)
    V[::7] = 0.0
    A = O.sparse(20, I, J, V, dropzeros=False)
    B = sp.coo_matrix((V, (I, J)), shape=(20, 20)).tocsc()
    assert np.allclose(A.toarray(), B.toarray(), atol=1e-14)
    Z = O.sparse(3, [0, 0, 1], [0, 0, 1], [1.0, -1.0, 2.0], dropzeros=True)
    assert Z.nnz == 1 and Z[1, 1] == 2.0


@pytest.mark.parametrize("elem,order", [("hex", 1), ("hex", 2), ("tet", 2)])
def test_traction_resultant(elem, order):
    """test/test_dsl.jl:79-85: traction (1,1,1) on the unit face sums to (1,1,1)."""
    m = M.structured_box_mesh(n=3, elem=elem, order=order)
    fs = m.faces["right"]
    f = O.traction(m, fs, R.rule_tables(fs.etype, 2 * order + 1), [1.0, 1.0, 1.0])
    assert np.allclose(f.reshape(-1, 3).sum(0), 1.0, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("elem,order", [("hex", 1), ("hex", 2), ("tet", 2)])
def test_traction_function_closed_forms(elem, order):
    """Function-valued traction on the right face x = 1 of the unit cube:
    fx = y gives the resultant int y dA = 1/2 and, because the face basis
    reproduces linear functions (sum_a N_a y_a = y), the nodal first moment
    sum_a f_x,a y_a = int y^2 dA = 1/3; fz = 2 + z gives 5/2 and
    sum_a f_z,a z_a = int (2 + z) z dA = 4/3.  A constant function equals
    the constant-traction restatement (oracle.c) to rounding."""
    m = M.structured_box_mesh(n=3, elem=elem, order=order)
    fs = m.faces["right"]
    tab = R.rule_tables(fs.etype, 2 * order + 1)
    f, T = O.traction_fn(m, fs, tab, [lambda x, y, z: y, None, lambda x, y, z: 2.0 + z])
    fx, fz = f[0::3], f[2::3]
    y, z = m.coords[:, 1], m.coords[:, 2]
    assert abs(fx.sum() - 0.5) < 1e-13 and abs(fz.sum() - 2.5) < 1e-13
    assert abs(fx @ y - 1 / 3) < 1e-13 and abs(fz @ z - 4 / 3) < 1e-13
    assert np.allclose(O.face_points(m, fs, tab)[..., 0], 1.0)
    fc, _ = O.traction_fn(m, fs, tab, [0.5, lambda x, y, z: -1.0, 2.0])
    assert np.allclose(fc, O.traction(m, fs, tab, [0.5, -1.0, 2.0]), rtol=0, atol=1e-15)


def test_traction_nodal_values_hex8():
    m = M.structured_box_mesh(n=4)
    fs = m.faces["right"]
    f = O.traction(m, fs, R.rule_tables(3, 3), [0.0, -1.0, 0.0])
    fy = f[1::3][m.nodes_on("right")]
    h2 = 1 / 16
    assert np.allclose(sorted(set(np.round(-fy / h2, 12))), [0.25, 0.5, 1.0])


def test_mass_totals_and_closed_form():
    m = M.structured_box_mesh(n=1, lx=0.5, ly=0.5, lz=0.5)
    d = (3 * m.conn[0][:, None] + np.arange(3)).ravel()
    M1 = O.mass_legacy(m, R.rule_tables(5, 3), rho=2.0).toarray()[np.ix_(d, d)]
    s = (np.array(R.HEX_VERTS) * 2 - 1).astype(float)
    h = 0.5
    for a in range(8):
        for b in range(8):
            ref = 2.0 * h ** 3 / 64 * np.prod(1 + s[a] * s[b] / 3)
            assert abs(M1[3 * a, 3 * b] - ref) < 1e-15
            assert M1[3 * a, 3 * b + 1] == 0.0
    m = M.structured_box_mesh(n=2, elem="tet", order=2, lz=2.0)
    Mc = O.mass_legacy(m, R.rule_tables(11, 5), rho=3.0)
    assert abs(Mc.sum() - 3 * 3.0 * 2.0) < 1e-12
    assert Mc.nnz * 3 == 9 * (M.structural_nnz(m) // 9)    # node adjacency x I3
    Ml = O.mass_legacy(m, R.rule_tables(11, 5), rho=3.0, lumped=True)
    assert abs(Ml.sum() - 18.0) < 1e-12


def test_oracle_pcg_matches_direct():
    m = M.structured_box_mesh(n=4, elem="tet", order=2)
    tab = R.rule_tables(11, 2)
    cp, rv = O.csc_pattern(m)
    K = sp.csc_matrix((O.stiffness_csc(m, tab, hooke_symgrad_D(), cp, rv), rv, cp))
    fs = m.faces["right"]
    f = O.traction(m, fs, R.rule_tables(fs.etype, 5), [0.0, -1.0, 0.0])
    left = m.nodes_on("left")
    fixed = np.concatenate([3 * left + c for c in range(3)])
    free = np.setdiff1d(np.arange(3 * m.N), fixed)
    ud = np.zeros(3 * m.N)
    ud[free] = spla.spsolve(K[free][:, free].tocsc(), f[free])
    u, it, rel = O.pcg(K, f, fixed, reltol=1e-12)
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-9
