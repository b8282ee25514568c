"""The reference's own DSL test (test/test_dsl.jl:1-116) read through the
Python mirror of its API, on the GPU: the DSL stiffness reproduces the legacy
stiffnessMatrix, repeated assembly agrees, the traction resultant is the face
area times the load, and the eliminated solve meets the relative residual on
the free DoFs with the constrained DoFs exactly zero -- for the Jacobi path
and the AMG preconditioner.  Mass shapes: massMatrix (consistent and
lumped) and ∫(Id ⋅ ρ ⋅ Id) with and without a structural csc_matrix, checked
against the closed-form total mass and against each other."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import lowlevelfem as lf
from lowlevelfem import _lib
from lowlevelfem import refelem as R

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _D(mat):
    lam, mu = mat.lam, mat.mu
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] = lam + 2 * mu
    D[np.arange(3, 6), np.arange(3, 6)] = mu
    return D


def test_dsl_reproduces_legacy_and_solves():
    say("[api] test_dsl.jl analogue")
    lf.structured_box_mesh(n=2, order=1)                      # test/test_dsl.jl:9
    mat = lf.Material("body")
    P = lf.Problem([mat])
    Klegacy = lf.stiffnessMatrix(P)
    Kdsl = lf.integral(lf.SymGrad(P), _D(mat), lf.SymGrad(P))
    assert Kdsl.A.shape == Klegacy.A.shape                     # :30
    assert spla.norm(Kdsl.A - Klegacy.A) / spla.norm(Klegacy.A) < 1e-12   # :31
    Kagain = lf.integral(lf.SymGrad(P), _D(mat), lf.SymGrad(P))
    assert np.array_equal(Kagain.A.data, Kdsl.A.data)         # :36-45 (assembly is repeatable)

    f = lf.loadVector(P, [lf.load("right", fx=1.0, fy=1.0, fz=1.0)])
    resultant = f.a[:, 0].reshape(-1, 3).sum(axis=0)          # :68-85, right face area 1
    assert np.allclose(resultant, [1.0, 1.0, 1.0], rtol=1e-12, atol=1e-12)

    support = [lf.displacementConstraint("left", ux=0.0, uy=0.0, uz=0.0)]
    fixed = lf.constrainedDoFs(P, support)
    free = lf.freeDoFs(P, support)
    for pre in ("jacobi", "amg"):
        u = lf.solveDisplacement(Kagain, f, support=support, iterative=True, reltol=1e-10,
                                 preconditioner=pre)
        r = Kagain.A @ u.a[:, 0] - f.a[:, 0]
        rel = np.linalg.norm(r[free]) / np.linalg.norm(f.a[free, 0])
        say(f"  {pre}: relative residual {rel:.2e}")
        assert rel < 1e-10                                     # :103-105
        assert np.all(u.a[fixed, 0] == 0.0)                    # :106


def test_amg_needs_the_device_stiffness():
    lf.structured_box_mesh(n=2, order=1)
    P = lf.Problem([lf.Material("body")])
    K1 = lf.stiffnessMatrix(P)
    lf.stiffnessMatrix(P)                                      # the device now holds a newer K
    f = lf.loadVector(P, [lf.load("right", fy=1.0)])
    with pytest.raises(ValueError, match="not the last one assembled"):
        lf.solveDisplacement(K1, f, support=[lf.displacementConstraint("left", ux=0, uy=0, uz=0)],
                             iterative=True, reltol=1e-10, preconditioner="amg")


@pytest.mark.parametrize("elem,order", [("hex", 1), ("tet", 2), ("hex", 2)])
def test_mass_shapes(elem, order):
    say(f"[api] mass shapes {elem}{order}")
    rho = 7.85e-9
    lf.structured_box_mesh(lx=1.0, ly=2.0, lz=0.5, n=4, elem=elem, order=order)
    P = lf.Problem([lf.Material("body", rho=rho)])
    M = lf.massMatrix(P, lumped=False)
    # closed form: every component's mass is rho * volume
    assert abs(M.A.sum() / (3 * rho * 1.0) - 1) < 1e-12
    Mid = lf.integral(lf.Id(P), rho, lf.Id(P))
    assert spla.norm(Mid.A - M.A) / spla.norm(M.A) < 1e-14
    Kpat = lf.build_csc_pattern(P)
    Mk = lf.integral(lf.Id(P), rho, lf.Id(P), csc_matrix=Kpat)
    assert Mk.A.nnz == Kpat.nnz                                # structural pattern kept
    assert spla.norm(lf.dropzeros(Mk.A) - M.A) / spla.norm(M.A) < 1e-14
    A = Mk.A.tocoo()
    assert np.all(A.data[(A.row % 3) != (A.col % 3)] == 0.0)  # off-diagonal block entries stored zeros
    Ml = lf.massMatrix(P, lumped=True)
    assert np.allclose(Ml.A.diagonal(), np.asarray(M.A.sum(axis=1)).ravel(), rtol=1e-12, atol=0.0)


def test_function_valued_load():
    """load(name; fx = f(x, y, z)) through the mirror (src/linear.jl:3288-3299):
    fx = y on the right face of the unit cube has resultant 1/2 and nodal
    first moment 1/3; a function returning a constant equals the constant
    load; scalar-only functions are evaluated point by point."""
    say("[api] function-valued load")
    m = lf.structured_box_mesh(n=3, order=2)
    P = lf.Problem([lf.Material("body")])
    f = lf.loadVector(P, [lf.load("right", fx=lambda x, y, z: y)]).a[:, 0]
    assert abs(f[0::3].sum() - 0.5) < 1e-13 and abs(f[0::3] @ m.coords[:, 1] - 1 / 3) < 1e-13
    assert np.all(f[1::3] == 0) and np.all(f[2::3] == 0)
    import math
    fs = lf.loadVector(P, [lf.load("top", fz=lambda x, y, z: math.cos(x))]).a[:, 0]
    fv = lf.loadVector(P, [lf.load("top", fz=lambda x, y, z: np.cos(x))]).a[:, 0]
    assert np.abs(fs - fv).max() <= 1e-15
    c1 = lf.loadVector(P, [lf.load("left", fy=lambda x, y, z: -2.0)]).a[:, 0]
    c2 = lf.loadVector(P, [lf.load("left", fy=-2.0)]).a[:, 0]
    assert np.abs(c1 - c2).max() <= 1e-14


def test_mirror_errors():
    lf.structured_box_mesh(n=2, order=1)
    P = lf.Problem([lf.Material("body")])
    with pytest.raises(KeyError, match="Physical group 'nowhere' not found"):
        lf.loadVector(P, [lf.load("nowhere", fy=1.0)])
    with pytest.raises(TypeError, match="loads must be in a vector"):
        lf.loadVector(P, lf.load("right", fy=1.0))
    with pytest.raises(NotImplementedError):
        lf.integral(lf.Id(P), np.eye(3), lf.Id(P))


def test_reference_benchmark_mesh_file():
    """The reference's own benchmark (lowlevelfem_elasticity_benchmark.jl:26-48)
    on the mesh file gmsh wrote for it, through the mirror: K on gmsh's
    numbering and its recorded checksum sum(abs, K) = 4.0073076923076982e9,
    the pattern bit-exact against build_csc_pattern of the oracle, and the
    solve with the left face clamped and traction (1, 0, 0) on the right,
    whose recorded sum(u) is 0.07057241535937334 (results.txt:7)."""
    import os

    from oracle import oracle as O

    say("[api] reference benchmark mesh file")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "elasticity_benchmark_n30.msh.gz")
    m = lf.read_msh(path)
    mat = lf.Material("body")
    P = lf.Problem([mat], type="VectorField", dim=3, field="u")   # benchmark.jl:30, renumbering aside
    K = lf.integral(lf.SymGrad(P), _D(mat), lf.SymGrad(P), csc_matrix=lf.build_csc_pattern(P), threads="auto")
    cp, rv = O.csc_pattern(m)
    assert np.array_equal(K.A.indptr, cp) and np.array_equal(K.A.indices, rv)
    say(f"  sum(abs, K) {np.abs(K.A.data).sum()!r}")
    assert abs(np.abs(K.A.data).sum() / 4.0073076923076982e9 - 1) < 1e-12
    f = lf.loadVector(P, [lf.load("right", fx=1.0)])
    support = [lf.displacementConstraint("left", ux=0.0, uy=0.0, uz=0.0)]
    u = lf.solveField(lf.Symmetric(K), f, support=support)   # lowlevelfem_elasticity_benchmark.jl:47
    say(f"  sum(u) {u.a[:, 0].sum()!r}")
    assert abs(u.a[:, 0].sum() / 0.07057241535937334 - 1) < 1e-10
    assert np.all(u.a[lf.constrainedDoFs(P, support), 0] == 0.0)
    u2 = lf.solveDisplacement(K, f, support=support)
    assert np.array_equal(u2.a, u.a)


def test_several_material_groups():
    """Problem([Material("steel"; ...), Material("alu"; ...)]) on two volume
    groups (src/linear.jl:190-330 loops over problem.material): stiffness and
    consistent mass against the oracle assembled group by group and summed in
    group order; one material on every element equals the plain assembly;
    an unknown group name fails as the reference does."""
    import copy

    from oracle import oracle as O

    say("[api] several material groups")
    m = lf.structured_box_mesh(n=4, elem="tet", order=2)
    xc = m.coords[m.conn].mean(axis=1)[:, 0]
    m.volumes = {"steel": np.nonzero(xc < 0.5)[0], "alu": np.nonzero(xc >= 0.5)[0]}
    steel = lf.Material("steel", E=2.1e5, nu=0.3, rho=7.85e-9)
    alu = lf.Material("alu", E=7.0e4, nu=0.34, rho=2.7e-9)
    P = lf.Problem([steel, alu])
    K = lf.stiffnessMatrix(P, dropzeros_=False)
    M = lf.massMatrix(P, lumped=False)
    cp, rv = O.csc_pattern(m)
    tab = R.rule_tables(m.etype, 5)
    Kref = np.zeros(cp[-1])
    Mref = None
    for mat in (steel, alu):
        sub = copy.copy(m)
        sub.conn = m.conn[m.volumes[mat.phName]]
        Kref = Kref + O.stiffness_csc(sub, tab, _D(mat), cp, rv)
        Mg = O.mass_legacy(sub, tab, rho=mat.rho)
        Mref = Mg if Mref is None else Mref + Mg
    assert np.array_equal(K.A.indptr, cp) and np.array_equal(K.A.indices, rv)
    assert np.abs(K.A.data - Kref).max() <= 1e-12 * np.abs(Kref).max()
    assert spla.norm(M.A - Mref) / spla.norm(Mref) < 1e-12
    # the groups were added on the device: the mesh holds the whole K, and the
    # AMG solve (which solves that device copy) matches a direct solve of it
    dev = P.device()
    assert np.array_equal(dev.values(_lib.OP_STIFFNESS), K.A.data)
    sup = [lf.displacementConstraint("left", ux=0, uy=0, uz=0)]
    f = lf.loadVector(P, [lf.load("right", fy=-1.0)])
    u = lf.solveDisplacement(K, f, support=sup, iterative=True, reltol=1e-10, preconditioner="amg")
    fixed = lf.constrainedDoFs(P, sup)
    free = np.setdiff1d(np.arange(P.ndofs()), fixed)
    Kff = K.A[free][:, free].tocsc()
    uref = spla.spsolve(Kff, f.a[free, 0])
    assert np.linalg.norm(u.a[free, 0] - uref) <= 1e-8 * np.linalg.norm(uref)
    # one group alone leaves an incomplete operator that llfem_solve refuses;
    # accumulating the other group completes it
    sc = np.zeros(m.E)
    sc[m.volumes["steel"]] = 1.0
    dev.assemble(_lib.OP_STIFFNESS, tab, E=steel.E, nu=steel.nu, element_scale=sc, copy_out=False)
    with pytest.raises(_lib.LLFEMError, match="no stiffness assembled"):
        dev.values(_lib.OP_STIFFNESS)
    sc2 = np.zeros(m.E)
    sc2[m.volumes["alu"]] = 1.0
    Kd, _ = dev.assemble(_lib.OP_STIFFNESS, tab, E=alu.E, nu=alu.nu, element_scale=sc2, accumulate=True)
    assert np.abs(Kd - Kref).max() <= 1e-12 * np.abs(Kref).max()
    assert np.array_equal(dev.values(_lib.OP_STIFFNESS), Kd)
    # one material named for a group that holds every element: the plain path
    m.volumes["all"] = np.arange(m.E)
    P1 = lf.Problem([lf.Material("all", E=2.1e5, nu=0.3)])
    K1 = lf.stiffnessMatrix(P1, dropzeros_=False)
    P2 = lf.Problem([lf.Material("body", E=2.1e5, nu=0.3)])
    del m.volumes["all"]
    with pytest.raises(KeyError, match="Physical group 'body' not found"):
        lf.stiffnessMatrix(P2)
    m.volumes = {}
    K2 = lf.stiffnessMatrix(lf.Problem([lf.Material("body", E=2.1e5, nu=0.3)]), dropzeros_=False)
    assert np.array_equal(K1.A.data, K2.A.data)
    with pytest.raises(_lib.LLFEMError, match="element_scale must be finite and >= 0"):
        K1.model.device().assemble(_lib.OP_STIFFNESS, tab, element_scale=-np.ones(m.E))


def test_solve_csc_rejects_bad_indices():
    """llfem_solve_csc checks colptr and rowval before they reach the device
    SpMV: a row index out of range or a decreasing colptr is LLFEM_EINVAL."""
    import ctypes as C
    say("[api] solve_csc index checks")
    ctx = _lib.default_context()
    lib = ctx.lib
    n = 3
    A = sp.csc_matrix(np.array([[4.0, 1, 0], [1, 4, 1], [0, 1, 4]]))
    f = np.ones(n)
    u = np.zeros(n)

    def call(cp, rv):
        cp = np.ascontiguousarray(cp, dtype=np.int64)
        rv = np.ascontiguousarray(rv, dtype=np.int64)
        return lib.llfem_solve_csc(ctx.h, n, _lib.ptr(cp), _lib.ptr(rv), _lib.ptr(A.data), 0, _lib.ptr(f),
                                   None, 0, None, _lib.PCG_JACOBI_FP64, C.c_double(1e-10), 100, _lib.ptr(u),
                                   None)
    assert call(A.indptr, A.indices) == _lib.OK
    assert np.allclose(A @ u, f)
    bad = A.indices.copy()
    bad[2] = 7
    assert call(A.indptr, bad) == _lib.EINVAL
    assert "out of range" in lib.llfem_last_error().decode()
    badp = A.indptr.copy()
    badp[1], badp[2] = badp[2], badp[1]
    assert call(badp, A.indices) == _lib.EINVAL
    assert "non-decreasing" in lib.llfem_last_error().decode()
