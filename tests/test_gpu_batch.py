"""Batched assembly (one persistent launch over many problems) and the
launch-shape autotuner, against the oracle and against per-problem assembly."""
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


def _problems(seed, count):
    """Uneven problems of mixed element types with random Hooke materials."""
    rng = np.random.default_rng(seed)
    kinds = [("hex", 1, 3), ("tet", 2, 2), ("prism", 1, 3)]
    out = []
    for i in range(count):
        elem, order, deg = kinds[i % 3]
        n = int(rng.integers(2, 6))
        cells = (n, int(rng.integers(2, 6)), int(rng.integers(2, 6)))
        m = M.structured_box_mesh(lx=cells[0], ly=cells[1], lz=cells[2], cells=cells, elem=elem,
                                  order=order, perturb=0.1 if i % 4 == 3 else 0.0, seed=i)
        E = float(rng.uniform(1e5, 3e5))
        nu = float(rng.uniform(0.2, 0.45))
        out.append((m, R.rule_tables(m.etype, deg), R.rule_tables(m.etype, 2 * deg), E, nu))
    return out


def _dev(ctx, m):
    return _lib.DeviceMesh(ctx, m.coords, m.etype, m.conn)


def test_batch_matches_oracle_and_single(gpu_ctx):
    probs = _problems(1, 12)
    devs = [_dev(gpu_ctx, p[0]) for p in probs]
    items = []
    for d, (m, tk, tm, E, nu) in zip(devs, probs):
        items.append((d, _lib.OP_STIFFNESS, tk, dict(E=E, nu=nu)))
        items.append((d, _lib.OP_MASS, tk, dict(rho=7.85e-9)))
    say(f"[batch] {len(items)} items")
    st = _lib.assemble_batch(items)
    say(f"  batch kernel {st.kernel_ms:.3f} ms (panel {st.main_kernel_ms:.3f}) elements {st.elements}")
    assert st.elements == 2 * sum(p[0].E for p in probs)
    first = []
    for d, (m, tk, tm, E, nu) in zip(devs, probs):
        K = d.values(_lib.OP_STIFFNESS)
        Mv = d.values(_lib.OP_MASS)
        first.append((K, Mv))
        cp, rv = O.csc_pattern(m)
        ref = O.stiffness_csc(m, tk, hooke_symgrad_D(E, nu), cp, rv, nthreads=4)
        assert np.linalg.norm(K - ref) / np.linalg.norm(ref) <= 1e-12
        assert np.all(np.abs(K - ref) <= 1e-12 * (np.abs(ref) + np.abs(ref).max()))
        mref = O.mass_legacy(m, tk, rho=7.85e-9)
        mcp, mrv = d.pattern(_lib.OP_MASS)
        G = sp.csc_matrix((Mv, mrv, mcp), shape=mref.shape)
        assert sp.linalg.norm(G - mref) / sp.linalg.norm(mref) < 1e-12
        # the same problem assembled alone
        Ks, _ = d.assemble(_lib.OP_STIFFNESS, tk, E=E, nu=nu)
        Ms, _ = d.assemble(_lib.OP_MASS, tk, rho=7.85e-9)
        assert np.abs(Ks - K).max() <= 1e-14 * np.abs(Ks).max()
        assert np.abs(Ms - Mv).max() <= 1e-14 * np.abs(Ms).max()
    # deterministic: a second batch reproduces every value bit for bit
    _lib.assemble_batch(items)
    for d, (K, Mv) in zip(devs, first):
        assert np.array_equal(d.values(_lib.OP_STIFFNESS), K)
        assert np.array_equal(d.values(_lib.OP_MASS), Mv)


def test_batch_fallback_items_and_lumped(gpu_ctx):
    """Items the persistent kernel does not cover (general D, FP32 variant, a
    second rule on the same mesh) are assembled inside the same call."""
    rng = np.random.default_rng(3)
    A = rng.normal(size=(6, 6))
    Dg = A @ A.T + 6 * np.eye(6)
    m1 = M.structured_box_mesh(n=3, elem="tet", order=2)
    m2 = M.structured_box_mesh(n=4)
    t1, t2 = R.rule_tables(m1.etype, 2), R.rule_tables(m2.etype, 3)
    t1m = R.rule_tables(m1.etype, 4)
    d1, d2 = _dev(gpu_ctx, m1), _dev(gpu_ctx, m2)
    items = [(d1, _lib.OP_STIFFNESS, t1, dict(D=Dg)),
             (d1, _lib.OP_MASS, t1m, dict(rho=2.0, lumped=True)),
             (d2, _lib.OP_STIFFNESS, t2, dict(precision=_lib.FP32_INT_FP64_SUM)),
             (d2, _lib.OP_MASS, t2, dict(rho=3.0, lumped=True))]
    _lib.assemble_batch(items)
    ref1 = O.stiffness_csc(m1, t1, Dg)
    K1 = d1.values(_lib.OP_STIFFNESS)
    assert np.linalg.norm(K1 - ref1) / np.linalg.norm(ref1) <= 1e-12
    L1 = d1.values(_lib.OP_MASS, lumped=True)
    r1 = O.mass_legacy(m1, t1m, rho=2.0, lumped=True).diagonal()
    assert np.abs(L1 - r1).max() <= 1e-12 * np.abs(r1).max()
    ref2 = O.stiffness_csc(m2, t2, hooke_symgrad_D(2e5, 0.3))
    K2 = d2.values(_lib.OP_STIFFNESS)
    assert 0 < np.linalg.norm(K2 - ref2) / np.linalg.norm(ref2) < 1e-6
    L2 = d2.values(_lib.OP_MASS, lumped=True)
    r2 = O.mass_legacy(m2, t2, rho=3.0, lumped=True).diagonal()
    assert np.abs(L2 - r2).max() <= 1e-12 * np.abs(r2).max()


def test_autotune_keeps_results(gpu_ctx):
    m = M.structured_box_mesh(n=12, elem="tet", order=2)
    tab = R.rule_tables(m.etype, 2)
    d = _dev(gpu_ctx, m)
    before, _ = d.assemble(_lib.OP_STIFFNESS, tab)
    res = d.autotune(_lib.OP_STIFFNESS, tab)
    say(f"  autotune tet10: mult {res.mult} wpb {res.waves_per_block} streams {res.streams} "
        f"rounds {res.round_kernels} best {res.best_ms:.4f} ms default {res.default_ms:.4f} ms "
        f"over {res.candidates} configurations")
    assert res.candidates == 36 and res.best_ms <= res.default_ms
    assert 1 <= res.streams <= 3 and res.round_kernels in (0, 1)
    after, _ = d.assemble(_lib.OP_STIFFNESS, tab)
    assert np.abs(after - before).max() <= 1e-14 * np.abs(before).max()
    # a mesh created afterwards on the same context uses the tuned shape
    m2 = M.structured_box_mesh(n=5, elem="tet", order=2)
    cp, rv = O.csc_pattern(m2)
    ref = O.stiffness_csc(m2, tab, hooke_symgrad_D(2e5, 0.3), cp, rv)
    v2, _ = _dev(gpu_ctx, m2).assemble(_lib.OP_STIFFNESS, tab)
    assert np.linalg.norm(v2 - ref) / np.linalg.norm(ref) <= 1e-12


def test_mesh_values_errors(gpu_ctx):
    m = M.structured_box_mesh(n=2)
    d = _dev(gpu_ctx, m)
    with pytest.raises(_lib.LLFEMError, match="no stiffness assembled"):
        d.values(_lib.OP_STIFFNESS)
