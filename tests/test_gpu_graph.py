"""HIP-graph replay of the solver's iteration chunks against the same
launches issued one by one (LLFEM_ENGINE_GRAPH=0): bit-identical solutions,
iteration and update counts, for the reliable-update engine with block
Jacobi and with the AMG V-cycle, through llfem_solve and through the
server.  A chunk graph is replayed many times per solve, with the eager
reliable updates (host-driven) between the replays."""
import os
import sys

import numpy as np
import pytest

from lowlevelfem import _lib
from lowlevelfem import mesh as M
from lowlevelfem import refelem as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(kw, deg):
    m = M.structured_box_mesh(**kw)
    tab = R.rule_tables(m.etype, deg)
    face = m.faces["right"]
    ftab = R.rule_tables(face.etype, deg)
    left = m.nodes_on("left")
    fixed = (3 * left[:, None] + np.arange(3)[None, :]).ravel()
    return m, tab, face, ftab, fixed


def _with_graph(flag, fn):
    old = os.environ.get("LLFEM_ENGINE_GRAPH")
    os.environ["LLFEM_ENGINE_GRAPH"] = flag
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["LLFEM_ENGINE_GRAPH"]
        else:
            os.environ["LLFEM_ENGINE_GRAPH"] = old


@pytest.mark.parametrize("method", [_lib.PCG_MIXED_RELIABLE, _lib.PCG_AMG_RELIABLE], ids=["jacobi", "amg"])
def test_solve_graph_replay_is_bit_identical(gpu_ctx, method):
    m, tab, face, ftab, fixed = _case(dict(lx=1.0, ly=1.0, lz=3.0, cells=(5, 5, 15), elem="tet", order=2), 2)
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    dev.assemble(_lib.OP_STIFFNESS, tab, copy_out=False)
    f = O.traction(m, face, ftab, [0.0, 1.0, 0.3])
    runs = {}
    for flag in ("0", "1", "1", "0"):
        u, st = _with_graph(flag, lambda: dev.solve(f, fixed, method=method, reltol=1e-10))
        runs.setdefault(flag, []).append((u, st.iterations, st.refinements))
        say(f"[graph] method {method} graph={flag}: {st.iterations} it, {st.refinements} updates, "
            f"relres {st.relres:.2e}")
    ref = runs["0"][0]
    for u, it, up in runs["0"] + runs["1"]:
        assert np.array_equal(u, ref[0]) and it == ref[1] and up == ref[2]
    assert ref[1] > 16          # several chunks: the graph was replayed
    dev.close()


def test_server_graph_replay_is_bit_identical(gpu_ctx):
    cases = [_case(dict(n=4), 3), _case(dict(lx=1.0, ly=1.0, lz=2.0, cells=(3, 3, 6), elem="tet", order=2), 2)]
    mats = [(2e5, 0.3, -1.0), (1.3e5, 0.42, 0.4), (2.7e5, 0.21, 0.9)]

    def serve():
        srv = _lib.Server()
        ids = []
        for k in range(6):
            m, tab, face, ftab, fixed = cases[k % 2]
            E, nu, t = mats[k % 3]
            ids.append(srv.submit(m, tab, ftab, face, [0.0, t, 0.0], fixed, E=E, nu=nu,
                                  method=_lib.PCG_AMG_RELIABLE, reltol=1e-10))
        out = [srv.wait(i) for i in ids]
        srv.close()
        return [(u, st.iterations) for u, st in out]

    a = _with_graph("0", serve)
    b = _with_graph("1", serve)
    for (ua, ia), (ub, ib) in zip(a, b):
        assert ia == ib and np.array_equal(ua, ub)
    say(f"[graph] server: {len(a)} requests identical, iterations {[i for _, i in a]}")
