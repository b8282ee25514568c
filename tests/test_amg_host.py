"""CPU check of the AMG hierarchy build (the host part of the aggregation-AMG
preconditioner; no GPU needed).  tests/amg_host_check.hip is compiled with
hipcc against csrc/amg.hip and run on the node pattern, strong-coupling flags,
coordinates and support of small meshes; it checks the invariants listed in
its header (aggregate membership, orthonormal tentative prolongators spanning
the rigid-body modes, Galerkin contribution lists covering every fine block
once, symmetric coarse patterns).  The strong flags are computed here from the
oracle's stiffness with the same rule the device uses (3x3 block Frobenius
norm >= 0.25 x the row's largest off-diagonal block)."""
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
HIPCC = "/opt/rocm/bin/hipcc"

pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = tmp_path_factory.mktemp("amgh") / "amg_host_check"
    cmd = [HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", f"-I{ROOT}/include", f"-I{ROOT}/lowlevelfem.jl_amd/csrc",
           f"{ROOT}/tests/amg_host_check.hip", f"{ROOT}/lowlevelfem.jl_amd/csrc/ctx.hip", "-o", str(exe),
           "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _inputs(m, tab, fixed, out):
    """Node pattern, strong flags, SoA coordinates and DoF mask as raw files."""
    cp, rv = O.csc_pattern(m)
    K = sp.csc_matrix((O.stiffness_csc(m, tab, hooke_symgrad_D(1.0, 0.3), cp, rv), rv, cp)).tocoo()
    N = m.N
    key = (K.col // 3).astype(np.int64) * N + (K.row // 3)
    order = np.argsort(key, kind="stable")
    uk, first = np.unique(key[order], return_index=True)
    norm2 = np.add.reduceat((K.data ** 2)[order], first)
    col = (uk // N).astype(np.int64)
    row = (uk % N).astype(np.int32)
    adj_ptr = np.zeros(N + 1, np.int64)
    np.add.at(adj_ptr, col + 1, 1)
    adj_ptr = np.cumsum(adj_ptr)
    offd = row != col
    mx = np.zeros(N)
    np.maximum.at(mx, col[offd], norm2[offd])
    strong = (offd & (norm2 > 0) & (norm2 >= 0.25 ** 2 * mx[col])).astype(np.uint8)
    selfpos = np.nonzero(~offd)[0]
    dslot = np.zeros(N, np.int32)
    dslot[col[selfpos]] = (selfpos - adj_ptr[col[selfpos]]).astype(np.int32)
    mask = np.zeros(3 * N, np.uint8)
    mask[np.asarray(fixed, dtype=np.int64)] = 1
    xyz = np.ascontiguousarray(m.coords.T).ravel()
    for name, a in [("adj_ptr", adj_ptr), ("adj", row), ("dslot", dslot), ("strong", strong), ("xyz", xyz),
                    ("mask", mask)]:
        a.tofile(out / f"{name}.bin")


CASES = [
    ("hex8-left-clamped", dict(lx=1.0, ly=0.5, lz=3.0, cells=(6, 3, 18)), 3, "left"),
    ("tet10-left-clamped", dict(lx=1.0, ly=1.0, lz=2.0, cells=(3, 3, 6), elem="tet", order=2), 2, "left"),
    ("prism6-free", dict(n=5, elem="prism"), 3, None),
    ("hex8-perturbed-bottom", dict(n=6, perturb=0.1, seed=3), 3, "bottom"),
    ("hex27-left", dict(n=3, order=2), 5, "left"),
]


@pytest.mark.parametrize("label,kw,deg,face", CASES, ids=[c[0] for c in CASES])
def test_amg_hierarchy_invariants(harness, tmp_path, label, kw, deg, face):
    m = M.structured_box_mesh(**kw)
    tab = R.rule_tables(m.etype, deg)
    fixed = [] if face is None else (3 * m.nodes_on(face)[:, None] + np.arange(3)[None, :]).ravel()
    _inputs(m, tab, fixed, tmp_path)
    r = subprocess.run([str(harness), str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(label, r.stdout.strip())
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
    sizes = [int(x) for x in r.stdout.split()[2:]]
    assert sizes[0] == m.N and all(a > b for a, b in zip(sizes, sizes[1:]))
