"""GPU parity: libllfem (HIP, gfx950) against the CPU oracle on the same mesh
arrays.  Bars: colptr/rowval bit-exact against build_csc_pattern; FP64
values ‖ΔK‖_F/‖K‖_F ≤ 1e-12 (test/test_dsl.jl:31) and per entry
|Δ| ≤ 1e-12·(|K_ij| + max|K_ii|); FP32-integration error reported and bounded."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import hooke_symgrad_D
from lowlevelfem import _lib
from lowlevelfem import mesh as M
from lowlevelfem import refelem as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(ctx, m):
    return _lib.DeviceMesh(ctx, m.coords, m.etype, m.conn)


def _check_values(gpu, ref, tol=1e-12):
    ref = np.asarray(ref)
    d = np.abs(gpu - ref)
    rel = np.linalg.norm(gpu - ref) / np.linalg.norm(ref)
    assert rel <= tol, rel
    assert np.all(d <= tol * (np.abs(ref) + np.abs(ref).max())), d.max()
    return rel


CASES = [
    # (label, mesh kwargs, rule degree)
    ("hex8-20^3", dict(n=20), 3),
    ("hex8-perturbed", dict(n=5, perturb=0.1, seed=3), 3),
    ("tet10-kuhn-4pt", dict(lz=4, cells=(4, 4, 16), elem="tet", order=2), 2),
    ("tet10-kuhn-gauss5", dict(n=3, elem="tet", order=2), 5),
    ("tet4", dict(n=4, elem="tet", order=1), 3),
    ("prism6", dict(n=4, elem="prism"), 3),
    ("hex27", dict(n=3, order=2), 5),
    ("hex27-perturbed", dict(n=2, order=2, perturb=0.05, seed=5), 5),
    ("hex64", dict(n=2, order=3), 7),
]


@pytest.mark.parametrize("label,kw,deg", CASES, ids=[c[0] for c in CASES])
def test_stiffness_parity(gpu_ctx, label, kw, deg):
    say(f"[stiffness] {label}")
    m = M.structured_box_mesh(**kw)
    tab = R.rule_tables(m.etype, deg)
    cp, rv = O.csc_pattern(m)
    ref = O.stiffness_csc(m, tab, hooke_symgrad_D(), cp, rv, nthreads=4)
    dev = _dev(gpu_ctx, m)
    gcp, grv = dev.pattern()
    assert np.array_equal(gcp, cp) and np.array_equal(grv, rv)
    val, st = dev.assemble(_lib.OP_STIFFNESS, tab, E=2e5, nu=0.3)
    rel = _check_values(val, ref)
    say(f"  rel {rel:.2e} kernel {st.kernel_ms:.3f} ms")
    # deterministic: a second run is bitwise identical
    val2, _ = dev.assemble(_lib.OP_STIFFNESS, tab, E=2e5, nu=0.3)
    assert np.array_equal(val, val2)


@pytest.mark.parametrize("label,kw", [("tet10", dict(lz=2, cells=(3, 3, 8), elem="tet", order=2)),
                                      ("tet10-classes", dict(lz=2, cells=(12, 12, 24), elem="tet", order=2)),
                                      ("hex8", dict(n=6)), ("hex8-classes", dict(n=24)),
                                      ("tet4", dict(n=5, elem="tet", order=1))])
def test_fixed_store_count_path_is_bitwise_the_loop_path(gpu_ctx, monkeypatch, label, kw):
    """Affine Hooke stiffness panels leave either in a compile-time number of
    stores per round -- in the step-unrolled round kernels for node classes
    with 4, 6, 8 or 24 incidences, otherwise in the generic kernel -- or
    through the run-time store loop; all sum the same contributions in the
    same order (the kernels may contract multiply-adds differently, so the
    bar is 1e-14 of the largest entry, as for batched against alone)."""
    say(f"[stiffness] fixed vs run-time panel stores {label}")
    m = M.structured_box_mesh(**kw)
    tab = R.rule_tables(m.etype, 2 if m.npe == 10 else 3)
    dev = _dev(gpu_ctx, m)
    fixed, _ = dev.assemble(_lib.OP_STIFFNESS, tab, E=2e5, nu=0.3)
    monkeypatch.setenv("LLFEM_PANEL_RUNTIME_STORES", "1")
    loop, _ = dev.assemble(_lib.OP_STIFFNESS, tab, E=2e5, nu=0.3)
    assert np.abs(fixed - loop).max() <= 1e-14 * np.abs(loop).max()


def test_large_values_copied_out_in_chunks(gpu_ctx):
    """Values larger than two 64 MB pinned staging buffers reach the caller's
    array chunk by chunk (4 chunks here); every entry against the oracle."""
    say("[stiffness] chunked copy-out")
    m = M.structured_box_mesh(lz=2, cells=(20, 20, 32), elem="tet", order=2)
    tab = R.rule_tables(11, 2)
    cp, rv = O.csc_pattern(m)
    assert cp[-1] * 8 > 3 * (64 << 20)
    ref = O.stiffness_csc(m, tab, hooke_symgrad_D(), cp, rv, nthreads=4)
    dev = _dev(gpu_ctx, m)
    val, _ = dev.assemble(_lib.OP_STIFFNESS, tab, E=2e5, nu=0.3)
    _check_values(val, ref)
    assert np.array_equal(dev.values(_lib.OP_STIFFNESS), val)


def test_stiffness_matches_legacy_path(gpu_ctx):
    """GPU vs stiffnessMatrixSolid + sparse + dropzeros! (src/linear.jl:190-330):
    normwise 1e-12, as the reference's own DSL≡legacy test."""
    m = M.structured_box_mesh(n=6, elem="tet", order=2)
    tab = R.rule_tables(11, 5)
    Kl = O.stiffness_legacy(m, tab)
    dev = _dev(gpu_ctx, m)
    val, _ = dev.assemble(_lib.OP_STIFFNESS, tab)
    cp, rv = dev.pattern()
    K = sp.csc_matrix((val, rv, cp), shape=Kl.shape)
    assert sp.linalg.norm(K - Kl) / sp.linalg.norm(Kl) < 1e-12


def test_general_D(gpu_ctx):
    rng = np.random.default_rng(11)
    A = rng.normal(size=(6, 6))
    D = A @ A.T + 6 * np.eye(6)
    for kw, deg in ((dict(n=4, elem="tet", order=2), 2), (dict(n=3, perturb=0.1, seed=2), 3)):
        m = M.structured_box_mesh(**kw)
        tab = R.rule_tables(m.etype, deg)
        ref = O.stiffness_csc(m, tab, D)
        val, _ = _dev(gpu_ctx, m).assemble(_lib.OP_STIFFNESS, tab, D=D)
        _check_values(val, ref)


def test_fp32_integration_error(gpu_ctx):
    m = M.structured_box_mesh(lz=2, cells=(4, 4, 8), elem="tet", order=2)
    tab = R.rule_tables(11, 2)
    ref = O.stiffness_csc(m, tab, hooke_symgrad_D())
    dev = _dev(gpu_ctx, m)
    v32, _ = dev.assemble(_lib.OP_STIFFNESS, tab, precision=_lib.FP32_INT_FP64_SUM)
    rel = np.linalg.norm(v32 - ref) / np.linalg.norm(ref)
    say(f"  FP32-integrate/FP64-sum rel error {rel:.3e}")
    assert 0 < rel < 1e-6


def test_fp64_fp32_alternation_on_one_mesh(gpu_ctx):
    """FP64 assemblies write only the FP64 part of the geometry records; an
    FP32-integration assembly on the same mesh writes its FP32 copy again, so
    FP64, FP32, FP64 in a row give the same bits as each alone."""
    m = M.structured_box_mesh(lz=2, cells=(3, 3, 6), elem="tet", order=2)
    tab = R.rule_tables(11, 2)
    dev = _dev(gpu_ctx, m)
    v1, _ = dev.assemble(_lib.OP_STIFFNESS, tab)
    v32, _ = dev.assemble(_lib.OP_STIFFNESS, tab, precision=_lib.FP32_INT_FP64_SUM)
    v2, _ = dev.assemble(_lib.OP_STIFFNESS, tab)
    w32, _ = _dev(gpu_ctx, m).assemble(_lib.OP_STIFFNESS, tab, precision=_lib.FP32_INT_FP64_SUM)
    assert np.array_equal(v1, v2)
    assert np.array_equal(v32, w32)
    assert 0 < np.linalg.norm(v32 - v1) / np.linalg.norm(v1) < 1e-6


@pytest.mark.parametrize("elem,order,deg", [("hex", 1, 3), ("tet", 2, 5), ("hex", 2, 5),
                                            ("prism", 1, 3)])
def test_mass_parity(gpu_ctx, elem, order, deg):
    m = M.structured_box_mesh(n=3, elem=elem, order=order, perturb=0.05, seed=9)
    tab = R.rule_tables(m.etype, deg)
    ref = O.mass_legacy(m, tab, rho=7.85e-9)
    dev = _dev(gpu_ctx, m)
    val, _ = dev.assemble(_lib.OP_MASS, tab, rho=7.85e-9)
    cp, rv = dev.pattern(_lib.OP_MASS)
    G = sp.csc_matrix((val, rv, cp), shape=ref.shape)
    assert sp.linalg.norm(G - ref) / sp.linalg.norm(ref) < 1e-12
    lump, _ = dev.assemble(_lib.OP_MASS, tab, rho=7.85e-9, lumped=True)
    refl = O.mass_legacy(m, tab, rho=7.85e-9, lumped=True).diagonal()
    assert np.abs(lump - refl).max() <= 1e-12 * np.abs(refl).max()


@pytest.mark.parametrize("elem,order", [("hex", 1), ("tet", 2), ("hex", 2), ("prism", 1)])
def test_traction_parity(gpu_ctx, elem, order):
    m = M.structured_box_mesh(n=4, elem=elem, order=order)
    dev = _dev(gpu_ctx, m)
    for name, t in (("right", [0.0, -1.0, 0.0]), ("top", [0.3, 0.2, -0.7])):
        fs = m.faces[name]
        name_, dim, o, nf, fam = R.element_properties(fs.etype)
        tab = R.rule_tables(fs.etype, 2 * o + 1)
        ref = O.traction(m, fs, tab, t)
        got = dev.traction(fs.etype, fs.conn, tab, t)
        assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("elem,order", [("hex", 1), ("tet", 2), ("hex", 2), ("prism", 1)])
def test_traction_function_parity(gpu_ctx, elem, order):
    """Function-valued traction (fx = f(x, y, z) etc.): the face Gauss-point
    coordinates from llfem_face_points and the load from
    llfem_load_traction_gp against the oracle's per-point restatement, on a
    perturbed mesh so the face Jacobians vary."""
    say(f"[traction fn] {elem}{order}")
    m = M.structured_box_mesh(n=4, elem=elem, order=order, perturb=0.08, seed=11)
    dev = _dev(gpu_ctx, m)
    fns = [lambda x, y, z: np.sin(3 * y) + z, None, lambda x, y, z: x * y - 2 * z * z]
    for name in ("right", "top", "front"):
        fs = m.faces[name]
        name_, dim, o, nf, fam = R.element_properties(fs.etype)
        tab = R.rule_tables(fs.etype, 2 * o + 1)
        P = dev.face_points(fs.etype, fs.conn, tab)
        assert np.abs(P - O.face_points(m, fs, tab)).max() <= 1e-14
        ref, T = O.traction_fn(m, fs, tab, fns)
        got = dev.traction_gp(fs.etype, fs.conn, tab, T)
        assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


def test_solve_golden_sum_u_30(gpu_ctx):
    """sum(u) = 0.07057241535937334 (results.txt:7) on the GPU: assemble,
    traction, Dirichlet-eliminated PCG, both FP64 and mixed precision."""
    m = M.structured_box_mesh(n=30)
    dev = _dev(gpu_ctx, m)
    dev.assemble(_lib.OP_STIFFNESS, R.rule_tables(5, 3), copy_out=False)
    fs = m.faces["right"]
    f = dev.traction(fs.etype, fs.conn, R.rule_tables(3, 3), [1.0, 0.0, 0.0])
    left = m.nodes_on("left")
    fixed = np.concatenate([3 * left + c for c in range(3)])
    u, st = dev.solve(f, fixed, reltol=1e-12)
    say(f"  fp64 pcg: {st.iterations} it, relres {st.relres:.2e}, sum(u) {u.sum()!r}")
    assert st.relres < 1e-10
    assert abs(u.sum() / 0.07057241535937334 - 1) < 1e-10
    assert np.all(u[fixed] == 0)
    um, sm = dev.solve(f, fixed, method=_lib.PCG_MIXED_REFINE, reltol=1e-10)
    say(f"  mixed: {sm.iterations} it, {sm.refinements} refinements, relres {sm.relres:.2e}")
    assert sm.relres < 1e-10
    assert abs(um.sum() / 0.07057241535937334 - 1) < 1e-8
    ur, sr = dev.solve(f, fixed, method=_lib.PCG_MIXED_RELIABLE, reltol=1e-10)
    say(f"  mixed reliable-update: {sr.iterations} it, {sr.refinements} updates, relres {sr.relres:.2e}")
    assert sr.relres < 1e-10
    assert abs(ur.sum() / 0.07057241535937334 - 1) < 1e-8
    assert np.all(ur[fixed] == 0)
    uo, so = dev.solve(f, fixed, method=_lib.PCG_F32OP_RELIABLE, reltol=1e-10)
    say(f"  FP32 operator / FP64 vectors: {so.iterations} it, {so.refinements} updates, relres {so.relres:.2e}")
    assert so.relres < 1e-10
    assert abs(uo.sum() / 0.07057241535937334 - 1) < 1e-8


def test_solve_reliable_inhomogeneous_dirichlet(gpu_ctx):
    """Mixed-precision reliable-update PCG on the mesh path with prescribed
    non-zero displacements, against the oracle's FP64 PCG."""
    m = M.structured_box_mesh(n=4, elem="tet", order=2)
    tab = R.rule_tables(11, 2)
    cp, rv = O.csc_pattern(m)
    K = sp.csc_matrix((O.stiffness_csc(m, tab, hooke_symgrad_D(), cp, rv), rv, cp))
    dev = _dev(gpu_ctx, m)
    dev.assemble(_lib.OP_STIFFNESS, tab, copy_out=False)
    f = np.zeros(3 * m.N)
    left, right = m.nodes_on("left"), m.nodes_on("right")
    fixed = np.concatenate([3 * left + c for c in range(3)] + [3 * right])
    ufix = np.concatenate([np.zeros(3 * left.size), np.full(right.size, 1e-3)])
    u, st = dev.solve(f, fixed, ufix, method=_lib.PCG_MIXED_RELIABLE, reltol=1e-10)
    u3, _ = dev.solve(f, fixed, ufix, method=_lib.PCG_F32OP_RELIABLE, reltol=1e-10)
    uo, it, rel = O.pcg(K, f, fixed, ufix, reltol=1e-12)
    assert np.linalg.norm(u3 - uo) / np.linalg.norm(uo) < 1e-8
    say(f"  reliable: {st.iterations} it, {st.refinements} updates; oracle FP64 {it} it")
    assert np.all(u[3 * left] == 0) and np.abs(u[3 * right] - 1e-3).max() == 0
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8
    free = np.setdiff1d(np.arange(3 * m.N), fixed)
    fk = K[:, fixed] @ u[fixed]
    r = K @ u - f
    assert np.linalg.norm(r[free]) / np.linalg.norm(fk[free]) < 1e-10


def test_solve_csc_and_dirichlet_values(gpu_ctx):
    m = M.structured_box_mesh(n=4, elem="tet", order=2)
    tab = R.rule_tables(11, 2)
    cp, rv = O.csc_pattern(m)
    K = sp.csc_matrix((O.stiffness_csc(m, tab, hooke_symgrad_D(), cp, rv), rv, cp))
    f = np.zeros(3 * m.N)
    left, right = m.nodes_on("left"), m.nodes_on("right")
    fixed = np.concatenate([3 * left + c for c in range(3)] + [3 * right])
    ufix = np.concatenate([np.zeros(3 * left.size), np.full(right.size, 1e-3)])
    u, st = _lib.solve_csc(gpu_ctx, K.indptr, K.indices, K.data, f, fixed, ufix, reltol=1e-12)
    uo, it, rel = O.pcg(K, f, fixed, ufix, reltol=1e-12)
    assert np.all(u[3 * left] == 0)
    assert np.abs(u[3 * right] - 1e-3).max() == 0
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-9
    free = np.setdiff1d(np.arange(3 * m.N), fixed)
    r = K @ u - f
    # inhomogeneous Dirichlet: relative to the kinematic load
    fk = K[:, fixed] @ u[fixed]
    assert np.linalg.norm(r[free]) / np.linalg.norm(fk[free]) < 1e-10


def test_singular_geometry_error(gpu_ctx):
    m = M.structured_box_mesh(n=2)
    c = m.coords.copy()
    c[m.conn[0]] = c[m.conn[0][0]]          # collapse one element to a point
    dev = _lib.DeviceMesh(gpu_ctx, c, m.etype, m.conn)
    with pytest.raises(_lib.LLFEMError, match="singular geometry"):
        dev.assemble(_lib.OP_STIFFNESS, R.rule_tables(5, 3))


def test_bad_connectivity_rejected_on_host(gpu_ctx):
    m = M.structured_box_mesh(n=2)
    conn = m.conn.copy()
    conn[0, 0] = m.N + 5
    with pytest.raises(_lib.LLFEMError, match="outside the trial-space node range"):
        _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, conn)
