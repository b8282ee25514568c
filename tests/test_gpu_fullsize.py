"""Value checks at the sizes the benchmarks time, on the GPU.

- tet10 beam 40x40x160 cells (1.54 M elements, 536 M stored entries, 4.0 GiB
  of values): K t = 0 for the three rigid translations and the three
  infinitesimal rotations, to 1e-10 ||K||_F ||t|| (and 1e-11 of the
  rounding scale || |K| |t| ||); symmetry of sampled columns; and sampled
  columns against the oracle, normwise 1e-12 (test/test_dsl.jl:31), where
  the oracle assembles only the elements incident to the sampled nodes (a
  column of K is the sum over the elements at its node).  The samples cover
  every incidence class (the most-connected nodes included), both ends of
  the node range and random nodes in between.
- hex125 12^3 (the matrix-core Gram path): the same properties; BF16x3
  against FP64.
- the 256-problem batch: every problem's batch values against the same
  problem assembled alone.
- one served request of about 2 M DoF: the true residual, recomputed on the
  host from the assembled K, is at most 1e-10.
Each case prints its progress; host memory peaks near 15 GB (the beam)."""
import sys
import time

import numpy as np
import pytest
import scipy.sparse as sp

from lowlevelfem import _lib
from lowlevelfem import mesh as M
from lowlevelfem import refelem as R
from lowlevelfem import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rigid_modes(X):
    n = X.shape[0]
    T = np.zeros((3 * n, 6))
    for c in range(3):
        T[c::3, c] = 1.0
    x, y, z = (X - X.mean(axis=0)).T
    T[1::3, 3], T[2::3, 3] = -z, y          # e_x × r
    T[0::3, 4], T[2::3, 4] = z, -x          # e_y × r
    T[0::3, 5], T[1::3, 5] = -y, x          # e_z × r
    return T


def sample_nodes(m, rng, nrand):
    inc = np.bincount(m.conn.ravel(), minlength=m.N)
    picks = [rng.choice(m.N, nrand, replace=False), [0, m.N - 1]]
    for c in np.unique(inc):                 # every incidence class
        picks.append(rng.choice(np.flatnonzero(inc == c), min(4, int((inc == c).sum())), replace=False))
    return np.unique(np.concatenate([np.asarray(p, dtype=np.int64) for p in picks]))


def check_properties(m, K, tab, nrand, rng, tag, E=2e5, nu=0.3):
    t0 = time.perf_counter()
    T = rigid_modes(m.coords)
    KT = K @ T
    Kabs = sp.csc_matrix((np.abs(K.data), K.indices, K.indptr), shape=K.shape)
    scale = Kabs @ np.abs(T)
    del Kabs
    knorm = np.sqrt(np.dot(K.data, K.data))
    for c in range(6):
        r = np.linalg.norm(KT[:, c])
        say(f"  {tag} rigid mode {c}: |K t| {r:.3e}, / (|K|_F |t|) {r / (knorm * np.linalg.norm(T[:, c])):.2e}, "
            f"/ ||K||t|| {r / np.linalg.norm(scale[:, c]):.2e}")
        assert r <= 1e-10 * knorm * np.linalg.norm(T[:, c])
        assert r <= 1e-11 * np.linalg.norm(scale[:, c])
    say(f"  {tag} rigid modes checked ({time.perf_counter() - t0:.1f} s)")
    nodes = sample_nodes(m, rng, nrand)
    dofs = (3 * nodes[:, None] + np.arange(3)[None, :]).ravel()
    # symmetry: K[i, j] == K[j, i] for every stored entry of the sampled columns
    worst = 0.0
    for j in dofs:
        lo, hi = K.indptr[j], K.indptr[j + 1]
        rows, vals = K.indices[lo:hi], K.data[lo:hi]
        colmax = np.abs(vals).max()
        for i, v in zip(rows, vals):
            a, b = K.indptr[i], K.indptr[i + 1]
            k = a + np.searchsorted(K.indices[a:b], j)
            assert k < b and K.indices[k] == j
            worst = max(worst, abs(K.data[k] - v) / colmax)
    say(f"  {tag} symmetry over {dofs.size} columns: max |K_ij - K_ji| / max|K_:j| = {worst:.2e}")
    assert worst <= 1e-13
    # sampled columns against the oracle on the elements at the sampled nodes
    t0 = time.perf_counter()
    at = np.isin(m.conn, nodes).any(axis=1)
    sub = M.Mesh(m.coords, m.etype, np.ascontiguousarray(m.conn[at]))
    Ko = O.stiffness_legacy(sub, tab, E, nu, dropzeros=False, nthreads=8)
    A, B = K[:, dofs], Ko.tocsc()[:, dofs]
    err = sp.linalg.norm(A - B) / sp.linalg.norm(B)
    say(f"  {tag} {nodes.size} sampled nodes ({at.sum()} elements): normwise {err:.2e} "
        f"({time.perf_counter() - t0:.1f} s)")
    assert err <= 1e-12


def test_tet10_beam_full_size(gpu_ctx):
    t0 = time.perf_counter()
    m = M.structured_box_mesh(lz=4.0, cells=(40, 40, 160), elem="tet", order=2)
    tab = R.rule_tables(11, 2)
    say(f"[fullsize] tet10 beam: {m.E} elements, {m.N} nodes ({time.perf_counter() - t0:.1f} s)")
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    val, st = dev.assemble(_lib.OP_STIFFNESS, tab)
    colptr, rowval = dev.pattern(_lib.OP_STIFFNESS)
    say(f"  nnz {val.size} ({8 * val.size / 2**30:.2f} GiB of values), kernel {st.kernel_ms:.2f} ms")
    assert val.size == 535_898_889
    K = sp.csc_matrix((val, rowval, colptr), shape=(3 * m.N, 3 * m.N))
    check_properties(m, K, tab, 160, np.random.default_rng(11), "tet10")
    dev.close()


def test_hex125_full_size(gpu_ctx):
    m = M.structured_box_mesh(n=12, order=4, perturb=0.02, seed=12)
    tab = R.rule_tables(m.etype, 9)
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    val, st = dev.assemble(_lib.OP_STIFFNESS, tab)
    colptr, rowval = dev.pattern(_lib.OP_STIFFNESS)
    say(f"[fullsize] hex125 12^3: {m.E} elements, nnz {val.size}, kernel {st.kernel_ms:.2f} ms")
    K = sp.csc_matrix((val, rowval, colptr), shape=(3 * m.N, 3 * m.N))
    check_properties(m, K, tab, 24, np.random.default_rng(12), "hex125")
    vb, _ = dev.assemble(_lib.OP_STIFFNESS, tab, precision=_lib.BF16X3)
    err = np.linalg.norm(vb - val) / np.linalg.norm(val)
    say(f"  BF16x3 vs FP64 normwise {err:.2e}")
    assert err < 1e-4
    dev.close()


def test_batch256_matches_single_assembly(gpu_ctx):
    probs = W.batched_problems(seed=2024, count=256)
    meshes = [p.mesh() for p in probs]
    rules = {k: W.rules(k) for k in W.KINDS}
    devs = [_lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn) for m in meshes]
    items = []
    for p, d in zip(probs, devs):
        tk, tm = rules[p.kind]
        items.append((d, _lib.OP_STIFFNESS, tk, dict(E=p.E, nu=p.nu)))
        items.append((d, _lib.OP_MASS, tm, dict(rho=p.rho)))
    say(f"[fullsize] batch: {len(probs)} problems, {sum(m.E for m in meshes)} elements")
    st = _lib.assemble_batch(items)
    say(f"  batch kernel {st.kernel_ms:.2f} ms")
    worst = 0.0
    for k, (p, d) in enumerate(zip(probs, devs)):
        tk, tm = rules[p.kind]
        Kb, Mb = d.values(_lib.OP_STIFFNESS), d.values(_lib.OP_MASS)
        Ks, _ = d.assemble(_lib.OP_STIFFNESS, tk, E=p.E, nu=p.nu)
        Ms, _ = d.assemble(_lib.OP_MASS, tm, rho=p.rho)
        eK = np.abs(Ks - Kb).max() / np.abs(Ks).max()
        eM = np.abs(Ms - Mb).max() / np.abs(Ms).max()
        worst = max(worst, eK, eM)
        assert eK <= 1e-14 and eM <= 1e-14, (k, p.kind, p.cells, eK, eM)
        if k % 64 == 63:
            say(f"  {k + 1} problems compared, worst {worst:.2e}")
    for d in devs:
        d.close()


def test_served_2m_dof_request_true_residual(gpu_ctx):
    cells = (86, 86, 86)                     # hex8: 87^3 nodes, 1.98 M DoF
    m = M.structured_box_mesh(cells=cells)
    tab = R.rule_tables(m.etype, 3)
    face = m.faces["right"]
    ftab = R.rule_tables(face.etype, 3)
    left = m.nodes_on("left")
    fixed = (3 * left[:, None] + np.arange(3)[None, :]).ravel()
    E, nu, t = 1.7e5, 0.31, [0.0, 0.6, 0.0]
    say(f"[fullsize] served request: {3 * m.N} DoF")
    srv = _lib.Server()
    try:
        u, rs = srv.wait(srv.submit(m, tab, ftab, face, t, fixed, E=E, nu=nu, method=_lib.PCG_AMG_RELIABLE,
                                    reltol=1e-10))
    finally:
        srv.close()
    say(f"  {rs.iterations} it, {rs.updates} updates, rollbacks {rs.rollbacks}, relres {rs.relres:.3e}, "
        f"latency {rs.latency_ms:.0f} ms")
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    val, _ = dev.assemble(_lib.OP_STIFFNESS, tab, E=E, nu=nu)
    colptr, rowval = dev.pattern(_lib.OP_STIFFNESS)
    f = dev.traction(face.etype, face.conn, ftab, t)
    dev.close()
    K = sp.csc_matrix((val, rowval, colptr), shape=(3 * m.N, 3 * m.N))
    free = np.ones(3 * m.N, bool)
    free[fixed] = False
    r = (K @ u - f)[free]
    relres = np.linalg.norm(r) / np.linalg.norm(f[free])
    say(f"  host true residual {relres:.3e}")
    assert np.all(u[fixed] == 0.0)
    assert relres <= 1e-10
    assert rs.rollbacks == 0
