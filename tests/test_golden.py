"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle): the oracle still reproduces
them, and they satisfy what the mechanics fixes independently of any code --
symmetry, rigid-body null space, mass and load resultants, equilibrium."""
import importlib.util
import os

import numpy as np
import pytest
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _maker():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLD, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


MK = _maker()
NAMES = sorted(MK.CASES)


def _load(name):
    return dict(np.load(os.path.join(GOLD, name + ".npz")))


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    g = _load(name)
    o = MK.compute(MK.CASES[name])
    for k in ("colptr", "rowval", "M_colptr", "M_rowval", "fixed"):
        assert np.array_equal(o[k], g[k]), k
    for k in ("K", "M", "f", "u"):
        assert np.linalg.norm(o[k] - g[k]) <= 1e-13 * np.linalg.norm(g[k]), k


@pytest.mark.parametrize("name", NAMES)
def test_fixture_mechanics(name):
    g = _load(name)
    n = len(g["u"])
    K = sp.csc_matrix((g["K"], g["rowval"], g["colptr"]), shape=(n, n))
    Mm = sp.csc_matrix((g["M"], g["M_rowval"], g["M_colptr"]), shape=(n, n))
    scale = abs(K).max()
    # symmetric
    assert abs(K - K.T).max() <= 1e-12 * scale
    # rigid translations are in the null space of the unconstrained K
    for c in range(3):
        t = np.zeros(n)
        t[c::3] = 1.0
        assert np.abs(K @ t).max() <= 1e-10 * scale
    # consistent mass: each component carries rho * volume
    spec = MK.CASES[name]["mesh"]
    vol = spec.get("lx", 1.0) * spec.get("ly", 1.0) * spec.get("lz", 1.0)
    for c in range(3):
        t = np.zeros(n)
        t[c::3] = 1.0
        assert abs(t @ (Mm @ t) - MK.RHO * vol) <= 1e-12 * MK.RHO * vol
    # traction fy = -1 on the right face (x = lx): resultant = -area
    area = spec.get("ly", 1.0) * spec.get("lz", 1.0)
    assert abs(g["f"][1::3].sum() + area) <= 1e-12 * area
    assert np.abs(g["f"][0::3]).max() == 0 and np.abs(g["f"][2::3]).max() == 0
    # the displacement solves the eliminated system and is zero where clamped
    free = np.setdiff1d(np.arange(n), g["fixed"])
    r = K[free][:, free] @ g["u"][free] - g["f"][free]
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(g["f"][free])
    assert np.all(g["u"][g["fixed"]] == 0)
    # reaction at the clamp balances the load
    reac = K @ g["u"] - g["f"]
    assert abs(reac[1::3].sum() - area) <= 1e-8 * area
