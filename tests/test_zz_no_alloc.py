"""llfem_solve_csc_noalloc, the allocation-free FP64 Jacobi-PCG on device CSC
arrays, against llfem_solve_csc (LLFEM_PCG_JACOBI_FP64) on the same system:
bit-identical displacements and iteration counts on two calls in a row, the
second on a smaller system (the shared workspace keeps the first call's
contents), no device memory allocated by the library during the calls, and
the size limit and index checks reported as errors."""
import ctypes as C
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from lowlevelfem import _lib
from lowlevelfem import mesh as M
from lowlevelfem import refelem as R
from oracle import oracle as O
from conftest import hooke_symgrad_D

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _system(n):
    m = M.structured_box_mesh(lx=1.0, ly=1.0, lz=2.0, cells=(n, n, 2 * n))
    tab = R.rule_tables(m.etype, 3)
    cp, rv = O.csc_pattern(m)
    nz = O.stiffness_csc(m, tab, hooke_symgrad_D(2e5, 0.3), cp, rv, nthreads=4)
    face = m.faces["right"]
    f = O.traction(m, face, R.rule_tables(face.etype, 3), [0.0, 1.0, 0.5])
    left = m.nodes_on("left")
    fixed = (3 * left[:, None] + np.arange(3)[None, :]).ravel()
    return cp, rv, nz, f, fixed


class _Dev:
    """A caller-owned device buffer, allocated with the HIP runtime libllfem.so
    links (the caller's allocation, made outside the solve)."""
    hip = None

    def __init__(self, a, dtype=None):
        if _Dev.hip is None:
            _Dev.hip = C.CDLL("libamdhip64.so.7")
            _Dev.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            _Dev.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            _Dev.hip.hipFree.argtypes = [C.c_void_p]
        self.a = np.ascontiguousarray(a, dtype=dtype)
        self.p = C.c_void_p()
        assert _Dev.hip.hipMalloc(C.byref(self.p), max(self.a.nbytes, 8)) == 0
        assert _Dev.hip.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, 1) == 0   # host to device

    @property
    def ptr(self):
        return self.p.value

    def host(self):
        out = np.empty_like(self.a)
        assert _Dev.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0          # device to host
        return out

    def __del__(self):
        if self.p.value:
            _Dev.hip.hipFree(self.p)


def test_noalloc_matches_allocating_solve(gpu_ctx):
    for n in (6, 3):   # the second call is smaller than the first
        cp, rv, nz, f, fixed = _system(n)
        ndof = cp.size - 1
        say(f"[noalloc] box {n}x{n}x{2 * n}: ndof {ndof}, nnz {nz.size}")
        ufix = np.zeros(ndof)
        ufix[fixed[::3]] = 1e-4   # inhomogeneous x on the left face
        u_ref, st_ref = _lib.solve_csc(gpu_ctx, cp, rv, nz, f, fixed, ufix[fixed], reltol=1e-10)
        mask = np.zeros(ndof, np.uint8)
        mask[fixed] = 1
        d_cp, d_rv = _Dev(cp, np.int64), _Dev(rv, np.int32)
        d_nz, d_f = _Dev(nz, np.float64), _Dev(f, np.float64)
        d_mask, d_ufix = _Dev(mask, np.uint8), _Dev(ufix, np.float64)
        d_u = _Dev(np.full(ndof, np.nan))
        before = gpu_ctx.bytes_in_use()
        st = _lib.solve_csc_noalloc(gpu_ctx, ndof, d_cp.ptr, d_rv.ptr, d_nz.ptr, d_f.ptr, d_mask.ptr,
                                    d_ufix.ptr, d_u.ptr, reltol=1e-10)
        assert gpu_ctx.bytes_in_use() == before
        u = d_u.host()
        say(f"  {st.iterations} it (allocating solve {st_ref.iterations}), relres {st.relres:.2e}")
        assert st.iterations == st_ref.iterations
        assert np.array_equal(u, u_ref)
        assert st.relres == st_ref.relres
        assert np.all(u[fixed] == ufix[fixed])
        # converged against the eliminated right-hand side f - K u_fixed
        K = sp.csc_matrix((nz, rv, cp), shape=(ndof, ndof))
        free = mask == 0
        b = (f - K @ ufix)[free]
        assert np.linalg.norm((K @ u - f)[free]) <= 1e-9 * np.linalg.norm(b)


def test_noalloc_rejects_bad_input(gpu_ctx):
    cp, rv, nz, f, fixed = _system(2)
    ndof = cp.size - 1
    mask = np.zeros(ndof, np.uint8)
    mask[fixed] = 1
    d_cp, d_nz, d_f = _Dev(cp, np.int64), _Dev(nz, np.float64), _Dev(f, np.float64)
    d_mask = _Dev(mask, np.uint8)
    d_u = _Dev(np.zeros(ndof))
    bad = rv.astype(np.int32).copy()
    bad[5] = ndof   # one row index out of range
    d_bad = _Dev(bad, np.int32)
    say("[noalloc] row index out of range")
    with pytest.raises(_lib.LLFEMError, match="row index out of range"):
        _lib.solve_csc_noalloc(gpu_ctx, ndof, d_cp.ptr, d_bad.ptr, d_nz.ptr, d_f.ptr, d_mask.ptr, None,
                               d_u.ptr)
    # above the workspace limit: rejected on the host before any device access
    say("[noalloc] size above the limit")
    with pytest.raises(_lib.LLFEMError, match="bytes of workspace; the maximum is"):
        _lib.solve_csc_noalloc(gpu_ctx, 3_000_000, d_cp.ptr, d_bad.ptr, d_nz.ptr, d_f.ptr, d_mask.ptr, None,
                               d_u.ptr)
