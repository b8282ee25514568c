"""HIP path against the committed golden fixtures (tests/golden/*.npz):
pattern bit-exact, stiffness / consistent mass within 1e-12, traction within
1e-14, and the solves (plain FP64 and mixed precision) within 1e-8 of the
direct-solve displacement at relative residual 1e-10."""
import importlib.util
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from lowlevelfem import _lib
from lowlevelfem import mesh as M
from lowlevelfem import refelem as R

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLD, "make_golden.py"))
MK = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(MK)


@pytest.mark.parametrize("name", sorted(MK.CASES))
def test_gpu_matches_golden(gpu_ctx, name):
    print(f"[golden] {name}", flush=True)
    sys.stdout.flush()
    g = dict(np.load(os.path.join(GOLD, name + ".npz")))
    spec = MK.CASES[name]
    m = M.structured_box_mesh(**spec["mesh"])
    dev = _lib.DeviceMesh(gpu_ctx, m.coords, m.etype, m.conn)
    cp, rv = dev.pattern()
    assert np.array_equal(cp, g["colptr"]) and np.array_equal(rv, g["rowval"])
    tab = R.rule_tables(m.etype, spec["deg"])
    K, _ = dev.assemble(_lib.OP_STIFFNESS, tab, E=MK.E, nu=MK.NU)
    assert np.linalg.norm(K - g["K"]) <= 1e-12 * np.linalg.norm(g["K"])
    Mv, _ = dev.assemble(_lib.OP_MASS, R.rule_tables(m.etype, spec.get("mdeg", spec["deg"])), rho=MK.RHO)
    mcp, mrv = dev.pattern(_lib.OP_MASS)
    n = 3 * m.N
    Mg = sp.csc_matrix((Mv, mrv, mcp), shape=(n, n))
    Mr = sp.csc_matrix((g["M"], g["M_rowval"], g["M_colptr"]), shape=(n, n))
    assert sp.linalg.norm(Mg - Mr) <= 1e-12 * sp.linalg.norm(Mr)
    face = m.faces["right"]
    f = dev.traction(face.etype, face.conn, R.rule_tables(face.etype, spec["fdeg"]), [0.0, -1.0, 0.0])
    assert np.abs(f - g["f"]).max() <= 1e-14 * np.abs(g["f"]).max()
    dev.assemble(_lib.OP_STIFFNESS, tab, E=MK.E, nu=MK.NU, copy_out=False)
    for method in (_lib.PCG_JACOBI_FP64, _lib.PCG_MIXED_RELIABLE):
        u, st = dev.solve(f, g["fixed"], method=method, reltol=1e-10)
        assert st.relres < 1e-10
        assert np.linalg.norm(u - g["u"]) <= 1e-8 * np.linalg.norm(g["u"]), method
