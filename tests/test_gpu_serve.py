"""Serving loop: requests alone and in a mix give the same displacements,
results meet the relative residual and match the oracle's FP64 solve, the
HBM budget admits requests one at a time when only one fits, and a bad
request fails without stopping the server."""
import sys
import threading
import time

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


def _case(kind, cells):
    elem, order = {"hex8": ("hex", 1), "tet10": ("tet", 2)}[kind]
    m = M.structured_box_mesh(lx=1.0, ly=cells[1] / cells[0], lz=cells[2] / cells[0], cells=cells,
                              elem=elem, order=order)
    deg = 3 if kind == "hex8" else 2
    tab = R.rule_tables(m.etype, deg)
    face = m.faces["right"]
    ftab = R.rule_tables(face.etype, 3 if kind == "hex8" else 5)
    left = m.nodes_on("left")
    fixed = (3 * left[:, None] + np.arange(3)[None, :]).ravel()
    return m, tab, face, ftab, fixed


CASES = [("hex8", (6, 4, 4)), ("tet10", (3, 3, 6)), ("hex8", (10, 3, 3))]
MATS = [(2e5, 0.3, -1.0), (1.3e5, 0.42, 0.4), (2.7e5, 0.21, 0.9), (1.1e5, 0.35, -0.25)]


def _submit(srv, case, mat, method=_lib.PCG_MIXED_RELIABLE):
    m, tab, face, ftab, fixed = case
    E, nu, t = mat
    return srv.submit(m, tab, ftab, face, [0.0, t, 0.0], fixed, E=E, nu=nu, method=method, reltol=1e-10)


@pytest.mark.parametrize("method", [_lib.PCG_MIXED_RELIABLE, _lib.PCG_AMG_RELIABLE], ids=["jacobi", "amg"])
def test_alone_equals_mix_and_oracle(method):
    say(f"[serve alone/mix] method {method}")
    cases = [_case(k, c) for k, c in CASES]
    jobs = [(ci, mi) for ci in range(len(cases)) for mi in range(len(MATS))][:9]
    # alone: one server, one request at a time
    srv = _lib.Server()
    alone = {}
    for j in jobs:
        u, st = srv.wait(_submit(srv, cases[j[0]], MATS[j[1]], method))
        assert st.relres < 1e-10 and st.status == 0
        alone[j] = u
    srv.close()
    # mix: every request in flight at once on a fresh server, in reverse order
    srv = _lib.Server()
    ids = {j: _submit(srv, cases[j[0]], MATS[j[1]], method) for j in reversed(jobs)}
    hits = 0
    for j, rid in ids.items():
        u, st = srv.wait(rid)
        hits += st.mesh_cache_hit
        assert st.relres < 1e-10
        assert np.array_equal(u, alone[j]), j          # bit-identical alone and in the mix
    srv.close()
    say(f"  {len(jobs)} requests, mesh cache hits in the mix: {hits}")
    assert hits >= len(jobs) - len(cases)
    # against the oracle's FP64 PCG on the same system
    for ci, mi in jobs[:4]:
        m, tab, face, ftab, fixed = cases[ci]
        E, nu, t = MATS[mi]
        cp, rv = O.csc_pattern(m)
        K = sp.csc_matrix((O.stiffness_csc(m, tab, hooke_symgrad_D(E, nu), cp, rv), rv, cp))
        f = O.traction(m, face, ftab, [0.0, t, 0.0])
        uo, it, rel = O.pcg(K, f, fixed, reltol=1e-12)
        u = alone[(ci, mi)]
        assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8
        assert np.all(u[fixed] == 0)


def test_budget_admits_one_at_a_time():
    case = _case("tet10", (3, 3, 6))
    ref = _lib.Server()
    u0, _ = ref.wait(_submit(ref, case, MATS[0]))
    ref.close()
    # measure what one request needs, then give the server room for about one
    probe = _lib.Context()
    m = case[0]
    dm = _lib.DeviceMesh(probe, m.coords, m.etype, m.conn)
    mesh_bytes = probe.bytes_in_use()
    nnz = dm.pattern(with_arrays=False)
    dm.close()
    probe.close()
    # per request: K (8 B/nz), FP32 K (4 B/nz), vectors, geometry records, slack
    per_req = 12 * nnz + 400 * m.N + 256 * m.E + (4 << 20)
    srv = _lib.Server(hbm_budget=mesh_bytes + int(1.5 * per_req))
    ids = [_submit(srv, case, MATS[i % len(MATS)]) for i in range(4)]
    stats = []
    for i, rid in enumerate(ids):
        u, st = srv.wait(rid)
        stats.append(st)
        if i % len(MATS) == 0:
            assert np.array_equal(u, u0)
    srv.close()
    q = [s.queue_ms for s in stats]
    say(f"  queue ms under the budget: {[round(x, 2) for x in q]}")
    assert all(s.status == 0 for s in stats)


def test_bad_request_fails_alone():
    case = _case("hex8", (4, 3, 3))
    m, tab, face, ftab, fixed = case
    srv = _lib.Server()
    bad = srv.submit(m, tab, ftab, face, [0.0, 1.0, 0.0], np.array([3 * m.N + 7]), method=_lib.PCG_MIXED_RELIABLE)
    good = _submit(srv, case, MATS[1])
    with pytest.raises(_lib.LLFEMError, match="constrained DoF out of range"):
        srv.wait(bad)
    u, st = srv.wait(good)
    assert st.relres < 1e-10
    srv.close()


def test_arrivals_while_running():
    """Requests submitted from another thread while earlier ones iterate are
    admitted into the running batch and complete correctly (Jacobi and AMG
    requests mixed in one batch)."""
    cases = [_case(k, c) for k, c in CASES]
    srv = _lib.Server()
    ids, lock = [], threading.Lock()

    def producer():
        for i in range(6):
            method = _lib.PCG_AMG_RELIABLE if i % 2 else _lib.PCG_MIXED_RELIABLE
            rid = _submit(srv, cases[i % 3], MATS[i % 4], method)
            with lock:
                ids.append(rid)
            time.sleep(0.002)

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    for rid in ids:
        u, st = srv.wait(rid)
        assert st.status == 0 and st.relres < 1e-10
    srv.close()
