"""ctypes binding of libllfem.so (the C ABI declared in include/llfem.h).

The product path has no fallback: if the HIP library is missing or no gfx950
device is present, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LLFEM_LIB", os.path.join(_HERE, "libllfem.so"))

OK, EINVAL, EHIP, ENOMEM, EGEOM, EPATTERN, ENOCONV, EUNSUP = 0, -1, -2, -3, -4, -5, -6, -7
OP_STIFFNESS, OP_MASS = 0, 1
FP64, FP32_INT_FP64_SUM, BF16X3 = 0, 1, 2
PCG_JACOBI_FP64, PCG_MIXED_REFINE, PCG_MIXED_RELIABLE, PCG_F32OP_RELIABLE, PCG_AMG_RELIABLE = 0, 1, 2, 3, 4

EXPORTED = [
    "llfem_ctx_create", "llfem_ctx_destroy", "llfem_ctx_synchronize", "llfem_ctx_bytes_in_use",
    "llfem_last_error", "llfem_mesh_create", "llfem_mesh_destroy", "llfem_mesh_pattern",
    "llfem_assemble", "llfem_load_traction", "llfem_face_points", "llfem_load_traction_gp",
    "llfem_load_body", "llfem_volume_points", "llfem_load_pressure",
    "llfem_solve", "llfem_solve_csc", "llfem_solve_csc_noalloc",
    "llfem_assemble_batch", "llfem_mesh_values", "llfem_autotune",
    "llfem_server_create", "llfem_server_submit", "llfem_server_wait", "llfem_server_destroy",
]


class OpDesc(C.Structure):
    _fields_ = [("op", C.c_int32), ("q", C.c_int32), ("weights", C.c_void_p),
                ("N_ref", C.c_void_p), ("dN_ref", C.c_void_p), ("D", C.c_void_p),
                ("E", C.c_double), ("nu", C.c_double), ("rho", C.c_double),
                ("lumped", C.c_int32), ("precision", C.c_int32), ("element_scale", C.c_void_p),
                ("accumulate", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("main_kernel_ms", C.c_double),
                ("total_ms", C.c_double), ("nnz", C.c_int64), ("elements", C.c_int64),
                ("alg_bytes", C.c_double), ("iterations", C.c_int64),
                ("refinements", C.c_int64), ("relres", C.c_double),
                ("alg_flops", C.c_double), ("setup_ms", C.c_double),
                ("rollbacks", C.c_int64), ("restarts", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneResult(C.Structure):
    _fields_ = [("elem_type", C.c_int32), ("mult", C.c_int32), ("waves_per_block", C.c_int32),
                ("candidates", C.c_int32), ("best_ms", C.c_double), ("default_ms", C.c_double),
                ("streams", C.c_int32), ("round_kernels", C.c_int32)]


class RequestSpec(C.Structure):   # llfem_request
    _fields_ = [("xyz", C.c_void_p), ("N", C.c_int64), ("elem_type", C.c_int32), ("npe", C.c_int32),
                ("conn", C.c_void_p), ("E", C.c_int64), ("stiffness", OpDesc),
                ("face_type", C.c_int32), ("nf", C.c_int32), ("face_conn", C.c_void_p), ("F", C.c_int64),
                ("face_q", C.c_int32), ("face_weights", C.c_void_p), ("face_N_ref", C.c_void_p),
                ("face_dN_ref", C.c_void_p), ("traction", C.c_double * 3), ("fixed", C.c_void_p),
                ("nfixed", C.c_int64), ("method", C.c_int32), ("reltol", C.c_double),
                ("maxiter", C.c_int64), ("mesh_key", C.c_uint64), ("u_out", C.c_void_p)]


class RequestStats(C.Structure):  # llfem_request_stats
    _fields_ = [("latency_ms", C.c_double), ("queue_ms", C.c_double), ("service_ms", C.c_double),
                ("iterations", C.c_int64), ("updates", C.c_int64), ("relres", C.c_double),
                ("mesh_cache_hit", C.c_int32), ("status", C.c_int32),
                ("rollbacks", C.c_int64), ("restarts", C.c_int64)]


class LLFEMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libllfem.so and declare its prototypes (no device access)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"libllfem.so not built ({path}); run __graft_entry__.build()")
    lib = C.CDLL(path)
    P, I32, I64, D, SZ = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_size_t
    lib.llfem_ctx_create.argtypes = [C.c_int, SZ, C.POINTER(P)]
    lib.llfem_ctx_destroy.argtypes = [P]
    lib.llfem_ctx_synchronize.argtypes = [P]
    lib.llfem_ctx_bytes_in_use.argtypes = [P]
    lib.llfem_ctx_bytes_in_use.restype = SZ
    lib.llfem_last_error.restype = C.c_char_p
    lib.llfem_mesh_create.argtypes = [P, P, I64, I32, I32, P, I64, I32, I32, C.POINTER(P)]
    lib.llfem_mesh_destroy.argtypes = [P]
    lib.llfem_mesh_pattern.argtypes = [P, I32, I32, C.POINTER(I64), P, P]
    lib.llfem_assemble.argtypes = [P, C.POINTER(OpDesc), P, C.POINTER(Stats)]
    lib.llfem_load_traction.argtypes = [P, I32, I32, P, I64, I32, I32, P, P, P, P, P]
    lib.llfem_face_points.argtypes = [P, I32, I32, P, I64, I32, I32, P, P]
    lib.llfem_load_traction_gp.argtypes = [P, I32, I32, P, I64, I32, I32, P, P, P, P, P]
    lib.llfem_load_body.argtypes = [P, I32, P, P, P, P, I64, P, P, P]
    lib.llfem_volume_points.argtypes = [P, I32, P, P, I64, P]
    lib.llfem_load_pressure.argtypes = [P, I32, I32, P, I64, I32, I32, P, P, P, P, D, P, P]
    lib.llfem_solve.argtypes = [P, P, P, I64, I32, P, I32, D, I64, P, C.POINTER(Stats)]
    lib.llfem_solve_csc.argtypes = [P, I64, P, P, P, I32, P, P, I64, P, I32, D, I64, P,
                                    C.POINTER(Stats)]
    lib.llfem_solve_csc_noalloc.argtypes = [P, I64, P, P, P, P, P, P, D, I64, P, C.POINTER(Stats)]
    lib.llfem_assemble_batch.argtypes = [I32, P, P, C.POINTER(Stats)]
    lib.llfem_mesh_values.argtypes = [P, I32, I32, P]
    lib.llfem_autotune.argtypes = [P, C.POINTER(OpDesc), C.POINTER(TuneResult)]
    lib.llfem_server_create.argtypes = [C.c_int, SZ, I32, C.POINTER(P)]
    lib.llfem_server_submit.argtypes = [P, C.POINTER(RequestSpec), C.POINTER(I64)]
    lib.llfem_server_wait.argtypes = [P, I64, C.POINTER(RequestStats)]
    lib.llfem_server_destroy.argtypes = [P]
    for name in EXPORTED:
        fn = getattr(lib, name)
        if name not in ("llfem_last_error", "llfem_ctx_bytes_in_use"):
            fn.restype = C.c_int
    _lib = lib
    return lib


def check(code: int):
    if code != OK:
        msg = _lib.llfem_last_error().decode(errors="replace")
        raise LLFEMError(code, msg)


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Context:
    """One HIP stream + memory accounting on one device."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        lib = load_library()
        h = C.c_void_p()
        check(lib.llfem_ctx_create(device, hbm_budget, C.byref(h)))
        self.h = h
        self.lib = lib

    def synchronize(self):
        check(self.lib.llfem_ctx_synchronize(self.h))

    def bytes_in_use(self) -> int:
        return int(self.lib.llfem_ctx_bytes_in_use(self.h))

    def close(self):
        if self.h:
            self.lib.llfem_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(int(os.environ.get("LLFEM_DEVICE", "0")))
    return _default_ctx


def make_desc(op, tables, E=2e5, nu=0.3, rho=7.85e-9, D=None, lumped=False, precision=FP64,
              element_scale=None):
    """OpDesc for llfem_assemble; returns (desc, arrays that must outlive it)."""
    w = np.ascontiguousarray(tables.weights, dtype=np.float64)
    Nr = np.ascontiguousarray(tables.N, dtype=np.float64)
    dN = np.ascontiguousarray(tables.dN, dtype=np.float64)
    Dm = None if D is None else np.ascontiguousarray(D, dtype=np.float64).reshape(6, 6)
    sc = None if element_scale is None else np.ascontiguousarray(element_scale, dtype=np.float64)
    desc = OpDesc(op, tables.q, ptr(w).value, ptr(Nr).value, ptr(dN).value,
                  None if Dm is None else ptr(Dm).value, E, nu, rho, int(lumped), precision,
                  None if sc is None else ptr(sc).value)
    return desc, (w, Nr, dN, Dm, sc)


class Batch:
    """The argument arrays of llfem_assemble_batch for a list of
    (DeviceMesh, op, tables, kwargs) items, built once; run() makes the call
    (one persistent launch; values stay on the device, DeviceMesh.values)."""

    def __init__(self, items):
        self.ctx = items[0][0].ctx if items else None
        self.n = len(items)
        self.meshes = [it[0] for it in items]                  # keep the handles alive
        self.handles = (C.c_void_p * self.n)(*[it[0].h.value for it in items])
        self.descs = (OpDesc * self.n)()
        self.keep = []
        for i, it in enumerate(items):
            d, k = make_desc(it[1], it[2], **(it[3] if len(it) > 3 else {}))
            self.descs[i] = d
            self.keep.append(k)

    def run(self):
        st = Stats()
        if self.n:
            check(self.ctx.lib.llfem_assemble_batch(self.n, self.handles, self.descs, C.byref(st)))
        return st


def assemble_batch(items):
    """One persistent assembly launch over many (DeviceMesh, op, tables, kwargs)
    items (llfem_assemble_batch); values stay on the device (DeviceMesh.values)."""
    return Batch(items).run()


class DeviceMesh:
    """Device-resident mesh + structural pattern (llfem_mesh)."""

    def __init__(self, ctx: Context, coords, etype, conn):
        self.ctx = ctx
        lib = ctx.lib
        xyz = np.ascontiguousarray(coords, dtype=np.float64)
        cn = np.ascontiguousarray(conn, dtype=np.int64)
        h = C.c_void_p()
        check(lib.llfem_mesh_create(ctx.h, ptr(xyz), xyz.shape[0], int(etype), cn.shape[1], ptr(cn),
                                    cn.shape[0], 3, 0, C.byref(h)))
        self.h = h
        self.N = xyz.shape[0]
        self.E = cn.shape[0]

    def pattern(self, op=OP_STIFFNESS, with_arrays=True):
        lib = self.ctx.lib
        nnz = C.c_int64()
        check(lib.llfem_mesh_pattern(self.h, op, 0, C.byref(nnz), None, None))
        if not with_arrays:
            return int(nnz.value)
        colptr = np.empty(3 * self.N + 1, dtype=np.int64)
        rowval = np.empty(int(nnz.value), dtype=np.int64)
        check(lib.llfem_mesh_pattern(self.h, op, 0, C.byref(nnz), ptr(colptr), ptr(rowval)))
        return colptr, rowval

    def assemble(self, op, tables, E=2e5, nu=0.3, rho=7.85e-9, D=None, lumped=False,
                 precision=FP64, copy_out=True, element_scale=None, out=None, accumulate=False):
        """Values to a new host array (copy_out), to `out` (a float64 array of
        the right size), or left on the device (copy_out=False).  With
        accumulate=True the values are added to the mesh's resident operator
        and the sum is what is returned."""
        if element_scale is not None and len(element_scale) != self.E:
            raise ValueError(f"element_scale has {len(element_scale)} entries, the mesh {self.E} elements")
        desc, _keep = make_desc(op, tables, E, nu, rho, D, lumped, precision, element_scale)
        desc.accumulate = int(bool(accumulate))
        st = Stats()
        n = 3 * self.N if (op == OP_MASS and lumped) else self.pattern(op, with_arrays=False)
        if out is not None:
            if out.dtype != np.float64 or out.size != n or not out.flags.c_contiguous:
                raise ValueError(f"assemble: out must be a contiguous float64 array of {n} values")
        elif copy_out:
            out = np.empty(n, dtype=np.float64)
        check(self.ctx.lib.llfem_assemble(self.h, C.byref(desc), ptr(out), C.byref(st)))
        return out, st

    def values(self, op=OP_STIFFNESS, lumped=False):
        """Device-resident values of the last assembly of `op` (llfem_mesh_values)."""
        n = 3 * self.N if (op == OP_MASS and lumped) else self.pattern(op, with_arrays=False)
        out = np.empty(n, dtype=np.float64)
        check(self.ctx.lib.llfem_mesh_values(self.h, op, int(lumped), ptr(out)))
        return out

    def autotune(self, op, tables, **kw):
        """Time launch shapes on this mesh and keep the best for its element type."""
        desc, _keep = make_desc(op, tables, **kw)
        res = TuneResult()
        check(self.ctx.lib.llfem_autotune(self.h, C.byref(desc), C.byref(res)))
        return res

    def traction(self, face_type, face_conn, tables, t):
        fc = np.ascontiguousarray(face_conn, dtype=np.int64)
        f = np.empty(3 * self.N, dtype=np.float64)
        tt = np.ascontiguousarray(t, dtype=np.float64)
        w = np.ascontiguousarray(tables.weights)
        Nr = np.ascontiguousarray(tables.N)
        dN = np.ascontiguousarray(tables.dN)
        check(self.ctx.lib.llfem_load_traction(self.h, int(face_type), fc.shape[1], ptr(fc),
                                               fc.shape[0], 0, tables.q, ptr(w), ptr(Nr), ptr(dN),
                                               ptr(tt), ptr(f)))
        return f

    def face_points(self, face_type, face_conn, tables):
        """Physical coordinates of the face Gauss points, (F, q, 3)."""
        fc = np.ascontiguousarray(face_conn, dtype=np.int64)
        Nr = np.ascontiguousarray(tables.N)
        out = np.empty((fc.shape[0], tables.q, 3), dtype=np.float64)
        check(self.ctx.lib.llfem_face_points(self.h, int(face_type), fc.shape[1], ptr(fc), fc.shape[0], 0,
                                             tables.q, ptr(Nr), ptr(out)))
        return out

    def traction_gp(self, face_type, face_conn, tables, t_gp):
        """Load vector of a traction given per face Gauss point, t_gp (F, q, 3)."""
        fc = np.ascontiguousarray(face_conn, dtype=np.int64)
        tg = np.ascontiguousarray(t_gp, dtype=np.float64)
        if tg.shape != (fc.shape[0], tables.q, 3):
            raise ValueError(f"traction_gp: t_gp has shape {tg.shape}, expected {(fc.shape[0], tables.q, 3)}")
        f = np.empty(3 * self.N, dtype=np.float64)
        w = np.ascontiguousarray(tables.weights)
        Nr = np.ascontiguousarray(tables.N)
        dN = np.ascontiguousarray(tables.dN)
        check(self.ctx.lib.llfem_load_traction_gp(self.h, int(face_type), fc.shape[1], ptr(fc), fc.shape[0], 0,
                                                  tables.q, ptr(w), ptr(Nr), ptr(dN), ptr(tg), ptr(f)))
        return f

    def _elems(self, elems):
        return (None, 0) if elems is None else (np.ascontiguousarray(elems, dtype=np.int64),
                                                len(elems))

    def body(self, tables, g=None, g_gp=None, elems=None):
        """Body-force load vector over the elements `elems` (0-based, None =
        all): constant g [3], or g_gp (ne, q, 3) per Gauss point."""
        el, ne = self._elems(elems)
        w, Nr, dN = (np.ascontiguousarray(a, dtype=np.float64) for a in (tables.weights, tables.N, tables.dN))
        gg = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
        gp = None
        if g_gp is not None:
            gp = np.ascontiguousarray(g_gp, dtype=np.float64)
            nset = self.E if el is None else ne
            if gp.shape != (nset, tables.q, 3):
                raise ValueError(f"body: g_gp has shape {gp.shape}, expected {(nset, tables.q, 3)}")
        f = np.empty(3 * self.N, dtype=np.float64)
        check(self.ctx.lib.llfem_load_body(self.h, tables.q, ptr(w), ptr(Nr), ptr(dN), ptr(el), ne, ptr(gg),
                                           ptr(gp), ptr(f)))
        return f

    def volume_points(self, tables, elems=None):
        """Physical coordinates of the volume Gauss points, (ne, q, 3)."""
        el, ne = self._elems(elems)
        Nr = np.ascontiguousarray(tables.N, dtype=np.float64)
        out = np.empty((self.E if el is None else ne, tables.q, 3), dtype=np.float64)
        check(self.ctx.lib.llfem_volume_points(self.h, tables.q, ptr(Nr), ptr(el), ne, ptr(out)))
        return out

    def pressure(self, face_type, face_conn, tables, dN_nodes, p=0.0, p_nodal=None):
        """Load vector of a pressure on a face set: constant p, or p_nodal [N]."""
        fc = np.ascontiguousarray(face_conn, dtype=np.int64)
        w, Nr, dN = (np.ascontiguousarray(a, dtype=np.float64) for a in (tables.weights, tables.N, tables.dN))
        dNn = np.ascontiguousarray(dN_nodes, dtype=np.float64)
        if dNn.shape != (fc.shape[1], fc.shape[1], 3):
            raise ValueError(f"pressure: dN_nodes has shape {dNn.shape}, expected {(fc.shape[1], fc.shape[1], 3)}")
        pn = None if p_nodal is None else np.ascontiguousarray(p_nodal, dtype=np.float64)
        if pn is not None and pn.size != self.N:
            raise ValueError(f"pressure: p_nodal has {pn.size} values, the mesh {self.N} nodes")
        f = np.empty(3 * self.N, dtype=np.float64)
        check(self.ctx.lib.llfem_load_pressure(self.h, int(face_type), fc.shape[1], ptr(fc), fc.shape[0], 0,
                                               tables.q, ptr(w), ptr(Nr), ptr(dN), ptr(dNn), float(p), ptr(pn),
                                               ptr(f)))
        return f

    def solve(self, f, fixed, ufix=None, method=PCG_JACOBI_FP64, reltol=1e-10, maxiter=100000):
        f = np.ascontiguousarray(f, dtype=np.float64)
        fx = np.ascontiguousarray(fixed, dtype=np.int64)
        uf = None if ufix is None else np.ascontiguousarray(ufix, dtype=np.float64)
        u = np.empty_like(f)
        st = Stats()
        code = self.ctx.lib.llfem_solve(self.h, ptr(f), ptr(fx), fx.size, 0, ptr(uf), method,
                                        float(reltol), int(maxiter), ptr(u), C.byref(st))
        check(code)
        return u, st

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.llfem_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_csc(ctx: Context, colptr, rowval, nzval, f, fixed, ufix=None, method=PCG_JACOBI_FP64,
              reltol=1e-10, maxiter=100000):
    cp = np.ascontiguousarray(colptr, dtype=np.int64)
    rv = np.ascontiguousarray(rowval, dtype=np.int64)
    nz = np.ascontiguousarray(nzval, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    fx = np.ascontiguousarray(fixed, dtype=np.int64)
    uf = None if ufix is None else np.ascontiguousarray(ufix, dtype=np.float64)
    u = np.empty_like(f)
    st = Stats()
    check(ctx.lib.llfem_solve_csc(ctx.h, cp.size - 1, ptr(cp), ptr(rv), ptr(nz), 0, ptr(f), ptr(fx),
                                  fx.size, ptr(uf), method, float(reltol), int(maxiter), ptr(u),
                                  C.byref(st)))
    return u, st


def solve_csc_noalloc(ctx: Context, ndof, d_colptr, d_rowval, d_nzval, d_f, d_fixed, d_ufix, d_u,
                      reltol=1e-10, maxiter=100000):
    """llfem_solve_csc_noalloc on device addresses (ints): zero-based int64
    colptr, int32 rowval, FP64 nzval/f/u_fixed/u, uint8 fixed mask."""
    st = Stats()
    check(ctx.lib.llfem_solve_csc_noalloc(ctx.h, int(ndof), d_colptr, d_rowval, d_nzval, d_f, d_fixed,
                                          d_ufix, float(reltol), int(maxiter), d_u, C.byref(st)))
    return st


class Server:
    """Assemble-and-solve server (llfem_server): continuous batching of
    requests on one device, meshes and their patterns cached across requests."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, max_active: int = 64):
        self.lib = load_library()
        h = C.c_void_p()
        check(self.lib.llfem_server_create(device, hbm_budget, max_active, C.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, mesh, tables, face_tables, face, traction, fixed, E=2e5, nu=0.3,
               method=PCG_MIXED_RELIABLE, reltol=1e-10, maxiter=200000, mesh_key=0):
        """mesh: lowlevelfem.mesh.Mesh; face: a FaceSet of it; fixed: 0-based DoFs."""
        xyz = np.ascontiguousarray(mesh.coords, dtype=np.float64)
        cn = np.ascontiguousarray(mesh.conn, dtype=np.int64)
        fc = np.ascontiguousarray(face.conn, dtype=np.int64)
        fx = np.ascontiguousarray(fixed, dtype=np.int64)
        fw = np.ascontiguousarray(face_tables.weights, dtype=np.float64)
        fN = np.ascontiguousarray(face_tables.N, dtype=np.float64)
        fdN = np.ascontiguousarray(face_tables.dN, dtype=np.float64)
        desc, keep_desc = make_desc(OP_STIFFNESS, tables, E, nu)
        u = np.empty(3 * xyz.shape[0], dtype=np.float64)
        r = RequestSpec()
        r.xyz, r.N, r.elem_type, r.npe = ptr(xyz).value, xyz.shape[0], int(mesh.etype), cn.shape[1]
        r.conn, r.E, r.stiffness = ptr(cn).value, cn.shape[0], desc
        r.face_type, r.nf, r.face_conn, r.F = int(face.etype), fc.shape[1], ptr(fc).value, fc.shape[0]
        r.face_q, r.face_weights = face_tables.q, ptr(fw).value
        r.face_N_ref, r.face_dN_ref = ptr(fN).value, ptr(fdN).value
        r.traction = (C.c_double * 3)(*[float(t) for t in traction])
        r.fixed, r.nfixed = (ptr(fx).value if fx.size else None), fx.size
        r.method, r.reltol, r.maxiter, r.mesh_key = method, float(reltol), int(maxiter), int(mesh_key)
        r.u_out = ptr(u).value
        rid = C.c_int64()
        check(self.lib.llfem_server_submit(self.h, C.byref(r), C.byref(rid)))
        self._keep[rid.value] = (xyz, cn, fc, fx, fw, fN, fdN, keep_desc, u, r)
        return rid.value

    def wait(self, rid):
        st = RequestStats()
        code = self.lib.llfem_server_wait(self.h, int(rid), C.byref(st))
        keep = self._keep.pop(int(rid))
        check(code)
        return keep[8], st

    def close(self):
        if getattr(self, "h", None):
            self.lib.llfem_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
