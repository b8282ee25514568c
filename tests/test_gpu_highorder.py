"""GPU parity of the high-order (order >= 3 hexahedra) Gram/MFMA path against
the CPU oracle: pattern bit-exact, FP64 values within 1e-12 (normwise and per
entry, as the panel path), BF16x3 error reported and bounded."""
import sys
from types import SimpleNamespace

import numpy as np
import pytest

from conftest import hooke_symgrad_D
from lowlevelfem import _lib
from lowlevelfem import mesh as M
from lowlevelfem import refelem as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check(gpu, ref, tol):
    rel = np.linalg.norm(gpu - ref) / np.linalg.norm(ref)
    assert rel <= tol, rel
    assert np.all(np.abs(gpu - ref) <= tol * (np.abs(ref) + np.abs(ref).max()))
    return rel


HO_CASES = [
    ("hex125-2^3", dict(n=2, order=4), 9),
    ("hex125-perturbed", dict(n=2, order=4, perturb=0.02, seed=7), 9),
    ("hex125-3x2x1", dict(lx=3, ly=2, lz=1, cells=(3, 2, 1), order=4), 9),
    ("hex125-gauss7", dict(n=2, order=4), 7),
]


@pytest.mark.parametrize("label,kw,deg", HO_CASES, ids=[c[0] for c in HO_CASES])
def test_highorder_fp64_parity(gpu_ctx, label, kw, deg):
    say(f"[highorder fp64] {label}")
    m = M.structured_box_mesh(**kw)
    tab = R.rule_tables(m.etype, deg)
    cp, rv = O.csc_pattern(m)
    ref = O.stiffness_csc(m, tab, hooke_symgrad_D(), cp, rv, nthreads=4)
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    gcp, grv = dev.pattern()
    assert np.array_equal(gcp, cp) and np.array_equal(grv, rv)
    val, st = dev.assemble(_lib.OP_STIFFNESS, tab, E=2e5, nu=0.3)
    rel = _check(val, ref, 1e-12)
    say(f"  rel {rel:.2e} gram {st.main_kernel_ms:.3f} ms total {st.kernel_ms:.3f} ms")
    assert st.alg_flops > 0
    val2, _ = dev.assemble(_lib.OP_STIFFNESS, tab, E=2e5, nu=0.3)
    assert np.array_equal(val, val2)


def test_highorder_general_D(gpu_ctx):
    rng = np.random.default_rng(5)
    A = rng.normal(size=(6, 6))
    D = A @ A.T + 6 * np.eye(6)
    m = M.structured_box_mesh(n=2, order=4, perturb=0.02, seed=1)
    tab = R.rule_tables(m.etype, 9)
    ref = O.stiffness_csc(m, tab, D)
    val, _ = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn).assemble(_lib.OP_STIFFNESS, tab, D=D)
    _check(val, ref, 1e-12)


def test_hex64_gram_path_matches_panel(gpu_ctx, monkeypatch):
    """hex64 through the Gram path (forced) and the panel path agree, and both
    match the oracle."""
    m = M.structured_box_mesh(n=2, order=3, perturb=0.05, seed=2)
    tab = R.rule_tables(m.etype, 7)
    ref = O.stiffness_csc(m, tab, hooke_symgrad_D())
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    panel, _ = dev.assemble(_lib.OP_STIFFNESS, tab)
    monkeypatch.setenv("LLFEM_HIGHORDER", "1")
    gram, _ = dev.assemble(_lib.OP_STIFFNESS, tab)
    _check(panel, ref, 1e-12)
    _check(gram, ref, 1e-12)


@pytest.mark.parametrize("order,deg", [(4, 9), (3, 7)])
def test_bf16x3_error(gpu_ctx, order, deg):
    """BF16x3 (hi*hi + hi*lo + lo*hi, FP32 accumulation) against FP64: error
    reported; bounded by 1e-4 normwise."""
    m = M.structured_box_mesh(n=2, order=order, perturb=0.02, seed=4)
    tab = R.rule_tables(m.etype, deg)
    ref = O.stiffness_csc(m, tab, hooke_symgrad_D())
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    v, st = dev.assemble(_lib.OP_STIFFNESS, tab, precision=_lib.BF16X3)
    rel = np.linalg.norm(v - ref) / np.linalg.norm(ref)
    say(f"  BF16x3 order {order}: rel error {rel:.3e} (gram {st.main_kernel_ms:.3f} ms)")
    assert 0 < rel < 1e-4


def test_highorder_rejects_negative_weights(gpu_ctx):
    m = M.structured_box_mesh(n=1, order=4)
    tab = R.rule_tables(m.etype, 9)
    w = np.array(tab.weights, dtype=np.float64)
    w[0] = -w[0]
    bad = SimpleNamespace(weights=w, N=tab.N, dN=tab.dN, q=tab.q)
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    with pytest.raises(_lib.LLFEMError, match="positive quadrature weights"):
        dev.assemble(_lib.OP_STIFFNESS, bad)
