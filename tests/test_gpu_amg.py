"""AMG-preconditioned mixed-precision PCG (LLFEM_PCG_AMG_RELIABLE) against the
oracle's FP64 PCG on the same system: the true FP64 relative residual meets
1e-10, the displacements agree to 1e-8, constrained DoFs hold their values
exactly, the hierarchy built for a support is reused for other materials,
repeated solves are bit-identical, and the V-cycle cuts the iteration count
well below Jacobi's."""
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


def _system(kw, deg):
    m = M.structured_box_mesh(**kw)
    tab = R.rule_tables(m.etype, deg)
    face = m.faces["right"]
    f = O.traction(m, face, R.rule_tables(face.etype, deg), [0.0, 1.0, 0.5])
    left = m.nodes_on("left")
    fixed = (3 * left[:, None] + np.arange(3)[None, :]).ravel()
    return m, tab, f, fixed


def _oracle_K(m, tab, E, nu):
    cp, rv = O.csc_pattern(m)
    return sp.csc_matrix((O.stiffness_csc(m, tab, hooke_symgrad_D(E, nu), cp, rv, nthreads=4), rv, cp))


CASES = [
    ("hex8-beam", dict(lx=1.0, ly=0.5, lz=3.0, cells=(6, 3, 18)), 3),
    ("tet10-beam", dict(lx=1.0, ly=1.0, lz=3.0, cells=(4, 4, 12), elem="tet", order=2), 2),
    ("hex8-perturbed", dict(n=6, perturb=0.1, seed=3), 3),
    ("tet4", dict(n=5, elem="tet", order=1), 3),
    ("prism6", dict(n=5, elem="prism"), 3),
    ("hex27", dict(n=3, order=2), 5),
    ("hex8-tiny", dict(n=2), 3),
]


@pytest.mark.parametrize("label,kw,deg", CASES, ids=[c[0] for c in CASES])
def test_amg_solve_matches_oracle(gpu_ctx, label, kw, deg):
    say(f"[amg] {label}")
    m, tab, f, fixed = _system(kw, deg)
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    dev.assemble(_lib.OP_STIFFNESS, tab, copy_out=False)
    u, st = dev.solve(f, fixed, method=_lib.PCG_AMG_RELIABLE, reltol=1e-10)
    uj, sj = dev.solve(f, fixed, method=_lib.PCG_MIXED_RELIABLE, reltol=1e-10)
    K = _oracle_K(m, tab, 2e5, 0.3)
    uo, it, _ = O.pcg(K, f, fixed, reltol=1e-12)
    say(f"  AMG {st.iterations} it ({st.refinements} updates, set-up {st.setup_ms:.1f} ms), "
        f"Jacobi {sj.iterations} it, oracle FP64 {it} it; relres {st.relres:.2e}")
    assert st.relres < 1e-10
    assert np.all(u[fixed] == 0)
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8
    free = np.setdiff1d(np.arange(3 * m.N), fixed)
    r = K @ u - f
    assert np.linalg.norm(r[free]) / np.linalg.norm(f[free]) < 1e-10
    if m.N > 1000:
        assert st.iterations < sj.iterations / 2
    dev.close()


def test_amg_hierarchy_reused_and_deterministic(gpu_ctx):
    kw, deg = dict(lx=1.0, ly=1.0, lz=4.0, cells=(4, 4, 16), elem="tet", order=2), 2
    m, tab, f, fixed = _system(kw, deg)
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    dev.assemble(_lib.OP_STIFFNESS, tab, E=2e5, nu=0.3, copy_out=False)
    u1, s1 = dev.solve(f, fixed, method=_lib.PCG_AMG_RELIABLE, reltol=1e-10)
    u1b, s1b = dev.solve(f, fixed, method=_lib.PCG_AMG_RELIABLE, reltol=1e-10)
    assert np.array_equal(u1, u1b)
    assert s1b.iterations == s1.iterations
    # another material on the same mesh and support: the kept hierarchy serves
    dev.assemble(_lib.OP_STIFFNESS, tab, E=1.3e5, nu=0.44, copy_out=False)
    u2, s2 = dev.solve(f, fixed, method=_lib.PCG_AMG_RELIABLE, reltol=1e-10)
    say(f"  set-up first {s1.setup_ms:.1f} ms, cached {s1b.setup_ms:.1f} ms; nu=0.44: {s2.iterations} it")
    K2 = _oracle_K(m, tab, 1.3e5, 0.44)
    uo, _, _ = O.pcg(K2, f, fixed, reltol=1e-12)
    assert s2.relres < 1e-10
    assert np.linalg.norm(u2 - uo) / np.linalg.norm(uo) < 1e-8
    dev.close()


def test_amg_inhomogeneous_dirichlet(gpu_ctx):
    m = M.structured_box_mesh(n=5, elem="tet", order=2)
    tab = R.rule_tables(11, 2)
    K = _oracle_K(m, tab, 2e5, 0.3)
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    dev.assemble(_lib.OP_STIFFNESS, tab, copy_out=False)
    f = np.zeros(3 * m.N)
    left, right = m.nodes_on("left"), m.nodes_on("right")
    fixed = np.concatenate([3 * left + c for c in range(3)] + [3 * right])
    ufix = np.concatenate([np.zeros(3 * left.size), np.full(right.size, 1e-3)])
    u, st = dev.solve(f, fixed, ufix, method=_lib.PCG_AMG_RELIABLE, reltol=1e-10)
    uo, it, _ = O.pcg(K, f, fixed, ufix, reltol=1e-12)
    say(f"  AMG {st.iterations} it; oracle FP64 {it} it")
    assert np.all(u[3 * left] == 0) and np.abs(u[3 * right] - 1e-3).max() == 0
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8
    free = np.setdiff1d(np.arange(3 * m.N), fixed)
    fk = K[:, fixed] @ u[fixed]
    assert np.linalg.norm((K @ u - f)[free]) / np.linalg.norm(fk[free]) < 1e-10
    dev.close()


def test_rollback_and_refresh_are_counted(gpu_ctx):
    """The reliable-update engine keeps the previous FP64 solution when an FP32
    correction is not finite (a rollback) and recomputes the AMG
    preconditioned residual when the rescaled one is no longer consistent (a
    restart).  A rollback is forced from the host at the second update: the
    solve still converges to the same displacements, and the counters say
    what happened; an unforced solve reports no rollback."""
    import os
    say("[amg] forced rollback")
    m, tab, f, fixed = _system(dict(lx=1.0, ly=1.0, lz=3.0, cells=(4, 4, 12), elem="tet", order=2), 2)
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    dev.assemble(_lib.OP_STIFFNESS, tab, copy_out=False)
    u0, st0 = dev.solve(f, fixed, method=_lib.PCG_AMG_RELIABLE, reltol=1e-10)
    assert st0.rollbacks == 0 and st0.relres <= 1e-10
    os.environ["LLFEM_ENGINE_FORCE_ROLLBACK"] = "2"
    try:
        u1, st1 = dev.solve(f, fixed, method=_lib.PCG_AMG_RELIABLE, reltol=1e-10)
    finally:
        del os.environ["LLFEM_ENGINE_FORCE_ROLLBACK"]
    say(f"  rollbacks {st1.rollbacks}, restarts {st1.restarts} (unforced {st0.restarts}), "
        f"updates {st1.refinements}, relres {st1.relres:.2e}")
    assert st1.rollbacks == 1
    assert st1.relres <= 1e-10
    assert np.linalg.norm(u1 - u0) <= 1e-8 * np.linalg.norm(u0)
    assert st0.restarts >= 0 and st1.restarts >= 0
