"""Headline benchmark: FP64 stiffness assembly of the 40x40x160 quadratic-tet
beam (Kuhn split, 1,536,000 tet10, 2,106,081 nodes, 6,318,243 DoF,
535,898,889 structural non-zeros, 4-point rule, E = 2e5, nu = 0.3).

A step is one full assembly of K into the device-resident CSC values with
the mesh, pattern and element->nnz map already in HBM.  Prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lowlevelfem.jl_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CELLS = (40, 40, 160)
LZ = 4.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def cpu_baseline(cells):
    """Oracle direct-CSC assembly (restating assemble_csc_worker! with
    per-thread buffers) on a bounded slab of the same mesh family."""
    from lowlevelfem import mesh as M, refelem as R
    from oracle import oracle as O
    m = M.structured_box_mesh(lz=LZ * cells[2] / CELLS[2], cells=cells, elem="tet", order=2)
    tab = R.rule_tables(11, 2)
    cp, rv = O.csc_pattern(m)
    D = O.hooke_D(2e5, 0.3)
    T = O.num_threads()
    O.stiffness_csc(m, tab, D, cp, rv, nthreads=T)        # warm-up
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        O.stiffness_csc(m, tab, D, cp, rv, nthreads=T)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": m.E / t, "unit": "elem/s", "cores": T, "kind": "port",
            "sample": f"tet10 Kuhn slab {cells[0]}x{cells[1]}x{cells[2]} cells ({m.E} elements), "
                      f"oracle direct-CSC FP64 assembly, median of 3 after 1 warm-up, "
                      f"{T} OpenMP threads on {cpu_model()}"}


PROFILE = os.path.join(ROOT, "profiles", "panel_assembly_tet10_beam_40x40x160.md")


def measured_traffic():
    """HBM bytes per assembly from the committed rocprofv3 FETCH/WRITE summary
    of this workload (null when the summary is absent)."""
    try:
        for line in open(PROFILE):
            if "traffic_bytes_per_assembly:" in line:
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def calib_copy(nbytes):
    """One device-to-device hipMemcpy of nbytes (HIP's blit kernel): the
    reference dispatch whose FETCH_SIZE/WRITE_SIZE scale the counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0     # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cells", type=str, default=None, help="override nx,ny,nz (testing only)")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="after the steps, copy this many bytes device-to-device once "
                         "(reference dispatch for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is described in DESIGN.md, not run.", file=sys.stderr)
        sys.exit(2)

    from lowlevelfem import _lib, mesh as M, refelem as R
    cells = tuple(int(x) for x in args.cells.split(",")) if args.cells else CELLS
    t0 = time.perf_counter()
    m = M.structured_box_mesh(lz=LZ * cells[2] / CELLS[2], cells=cells, elem="tet", order=2)
    log(f"mesh: {m.N} nodes, {m.E} tet10 ({time.perf_counter() - t0:.1f} s)")
    tab = R.rule_tables(11, 2)
    ctx = _lib.default_context()
    t0 = time.perf_counter()
    dev = _lib.DeviceMesh(ctx, m.coords, m.etype, m.conn)
    ctx.synchronize()
    setup_s = time.perf_counter() - t0
    nnz = dev.pattern(with_arrays=False)
    log(f"device mesh + pattern + element->nnz map: {setup_s:.2f} s, nnz {nnz}")

    for _ in range(args.warmup):
        dev.assemble(_lib.OP_STIFFNESS, tab, copy_out=False)
    ctx.synchronize()
    kernel_ms, main_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, st = dev.assemble(_lib.OP_STIFFNESS, tab, copy_out=False)
        kernel_ms.append(st.kernel_ms)
        main_ms.append(st.main_kernel_ms)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    alg_bytes = 24.0 * m.N + 4.0 * 10 * m.E + 8.0 * nnz
    k_ms = float(np.mean(kernel_ms))
    p_ms = float(np.mean(main_ms))
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9

    # FP32-integrate / FP64-sum variant (reported beside the FP64 headline)
    _, st32 = dev.assemble(_lib.OP_STIFFNESS, tab, copy_out=False, precision=_lib.FP32_INT_FP64_SUM)
    for _ in range(3):
        _, st32 = dev.assemble(_lib.OP_STIFFNESS, tab, copy_out=False,
                               precision=_lib.FP32_INT_FP64_SUM)

    out = {
        "metric": "assembled elem/s (FP64 stiffness, CSC values in HBM)",
        "value": m.E / (ms_step * 1e-3),
        "unit": "elem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"tet10 Kuhn beam {cells[0]}x{cells[1]}x{cells[2]} cells, "
                               "1x1x4 box, 4-point rule, FP64 stiffness into CSC",
                   "elements": m.E, "nodes": m.N, "dofs": 3 * m.N, "nnz": nnz},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic() if cells == CELLS else None},
        "kernel_ms": {"assembly_events": k_ms, "panel_kernels_events": p_ms,
                      "geometry_events": k_ms - p_ms},
        "fp32_integrate_fp64_sum": {"kernel_ms": st32.kernel_ms,
                                    "elem_per_s": m.E / (st32.kernel_ms * 1e-3)},
        "setup_s": setup_s,
    }
    if args.calib_copy:
        calib_copy(args.calib_copy)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline((40, 40, 16) if not args.cells else cells)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
