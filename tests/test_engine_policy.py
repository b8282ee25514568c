"""CPU test of the PCG engine's reliable-update decisions (rollback of a
non-finite FP32 correction, stop conditions), through the header the
engine's update step calls (lowlevelfem.jl_amd/csrc/engine_policy.h).

tests/engine_policy_check.cpp runs the engine's process on the host (FP32 CG
on the FP32-rounded K with unit-scaled residual, FP64 reliable updates) on a
clamped tet10 cantilever; overflows are planted in the FP32 correction before
chosen updates.  Checked:
- without plants: converged to the true relative residual, no rollback;
- one planted overflow: the trial is rolled back (one rollback counted), the
  recursion restarts from the kept solution and converges to the same
  tolerance; the solution agrees with the unplanted one to that tolerance;
- overflows at two non-consecutive updates: both rolled back, converged;
- three consecutive overflows: the system stops as non-finite, and the kept
  solution is the last adopted one (finite);
- the stop rules on their own (zero right-hand side, maxiter, a non-finite
  residual) and the AMG keep-or-recompute rule for the rescaled z."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import hooke_symgrad_D
from lowlevelfem import mesh as M
from lowlevelfem import refelem as R
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lowlevelfem.jl_amd", "csrc")
RUNNING, CONVERGED, MAXITER, NONFINITE = 0, 1, -1, -2
TOL = 1e-10


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = tmp_path_factory.mktemp("policy") / "engine_policy_check"
    r = subprocess.run(["g++", "-O2", "-std=c++17", f"-I{CSRC}", os.path.join(ROOT, "tests", "engine_policy_check.cpp"),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


@pytest.fixture(scope="module")
def system(tmp_path_factory):
    m = M.structured_box_mesh(lx=1.0, ly=1.0, lz=3.0, cells=(3, 3, 9), elem="tet", order=2)
    tab = R.rule_tables(m.etype, 2)
    ftab = R.rule_tables(m.faces["right"].etype, 2)
    cp, rv = O.csc_pattern(m)
    K = sp.csc_matrix((O.stiffness_csc(m, tab, hooke_symgrad_D(2e5, 0.3), cp, rv), rv, cp))
    f = O.traction(m, m.faces["right"], ftab, [0.0, 1.0, 0.3])
    left = m.nodes_on("left")
    free = np.ones(3 * m.N, bool)
    free[(3 * left[:, None] + np.arange(3)[None, :]).ravel()] = False
    Kff = K[free][:, free].tocsc()
    Kff.sort_indices()
    d = tmp_path_factory.mktemp("sys")
    Kff.indptr.astype(np.int64).tofile(d / "colptr.bin")
    Kff.indices.astype(np.int64).tofile(d / "rowval.bin")
    Kff.data.astype(np.float64).tofile(d / "nzval.bin")
    f[free].astype(np.float64).tofile(d / "b.bin")
    return d, Kff, f[free]


def run(driver, system, plants=(), maxiter=20000):
    d, K, b = system
    out = subprocess.run([str(driver), str(d), repr(TOL), str(maxiter)] + [str(k) for k in plants],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    w = out.stdout.split()
    res = {w[i]: float(w[i + 1]) for i in range(0, len(w), 2)}
    x = np.fromfile(d / "x.bin", dtype=np.float64)
    true = np.linalg.norm(b - K @ x) / np.linalg.norm(b)
    return res, x, true


def test_no_rollback_converges(driver, system):
    res, x, true = run(driver, system)
    assert res["status"] == CONVERGED and res["rollbacks"] == 0
    assert true <= TOL and abs(true - res["relres"]) <= 1e-3 * TOL
    assert res["updates"] >= 3


def test_one_rollback_restarts_from_kept_solution(driver, system):
    ref, xr, _ = run(driver, system)
    res, x, true = run(driver, system, plants=[2])
    assert res["status"] == CONVERGED and res["rollbacks"] == 1
    assert true <= TOL
    assert np.all(np.isfinite(x))
    # both meet the tolerance, so they agree to the conditioning-scaled tolerance
    assert np.linalg.norm(x - xr) <= 1e-4 * np.linalg.norm(xr)
    # the rolled-back update costs the iterations since the last update
    assert res["iters"] > ref["iters"]


def test_non_consecutive_rollbacks_converge(driver, system):
    res, x, true = run(driver, system, plants=[2, 4])
    assert res["status"] == CONVERGED and res["rollbacks"] == 2 and true <= TOL


def test_three_consecutive_rollbacks_stop_as_non_finite(driver, system):
    _, x1, _ = run(driver, system, plants=[])
    res, x, true = run(driver, system, plants=[2, 3, 4])
    assert res["status"] == NONFINITE and res["rollbacks"] == 3 and res["updates"] == 4
    # the kept solution is the one adopted at update 1: finite, residual below b's
    assert np.all(np.isfinite(x)) and 0.0 < true < 1.0


class _Book(ctypes.Structure):
    _fields_ = [("bad", ctypes.c_int), ("rollbacks", ctypes.c_int64), ("updates", ctypes.c_int64)]


def test_stop_rules(driver, tmp_path):
    # the decision functions on edge cases, compiled from the header into a small shared library
    src = tmp_path / "p.cpp"
    src.write_text('#include "engine_policy.h"\n'
                   'extern "C" int st(double rn, double bn, double tol, long long it, long long mx, '
                   'llfem::UpdateBook* b) { return llfem::update_status(rn, bn, tol, it, mx, *b); }\n'
                   'extern "C" bool ad(double r2, llfem::UpdateBook* b) { return llfem::adopt_trial(r2, b); }\n'
                   'extern "C" bool keep(double rz, double d2, double k) { return llfem::amg_keep_rescaled(rz, d2, k); }\n')
    so = tmp_path / "p.so"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", f"-I{CSRC}", str(src), "-o", str(so)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    L = ctypes.CDLL(str(so))
    L.st.argtypes = [ctypes.c_double] * 3 + [ctypes.c_longlong] * 2 + [ctypes.POINTER(_Book)]
    L.ad.argtypes = [ctypes.c_double, ctypes.POINTER(_Book)]
    L.ad.restype = ctypes.c_bool
    L.keep.argtypes = [ctypes.c_double] * 3
    L.keep.restype = ctypes.c_bool
    b = _Book()
    assert L.st(0.0, 0.0, 1e-8, 0, 10, ctypes.byref(b)) == CONVERGED          # zero right-hand side
    assert L.st(1e-9, 1.0, 1e-8, 5, 10, ctypes.byref(b)) == CONVERGED
    assert L.st(1e-3, 1.0, 1e-8, 10, 10, ctypes.byref(b)) == MAXITER
    assert L.st(float("nan"), 1.0, 1e-8, 5, 10, ctypes.byref(b)) == NONFINITE
    assert L.st(1e-3, 1.0, 1e-8, 5, 10, ctypes.byref(b)) == RUNNING
    assert not L.ad(float("inf"), ctypes.byref(b)) and b.bad == 1 and b.rollbacks == 1
    assert not L.ad(float("nan"), ctypes.byref(b)) and b.bad == 2
    assert L.st(1e-3, 1.0, 1e-8, 5, 10, ctypes.byref(b)) == RUNNING
    assert L.ad(4.0, ctypes.byref(b)) and b.bad == 0 and b.rollbacks == 2
    for _ in range(3):
        L.ad(float("inf"), ctypes.byref(b))
    assert L.st(1e-3, 1.0, 1e-8, 5, 10, ctypes.byref(b)) == NONFINITE
    assert L.keep(0.9, 0.01, 0.25) and not L.keep(0.0, 0.01, 0.25) and not L.keep(-1.0, 0.0, 0.25)
    assert not L.keep(0.9, 0.25 ** 2 * 1.01, 0.25) and not L.keep(float("nan"), 0.0, 0.25)
