"""Synthetic workloads of the benchmarks, generated from a recorded seed.

* ``batched_problems``: independent structured-box problems of mixed element
  type (hex8, tet10 Kuhn split, wedge6), edge counts uniform in [4, 48] per
  axis, resampled until the DoF count lies in [2k, 400k]; Hooke materials
  with E ~ U[1e5, 3e5], nu ~ U[0.2, 0.45].
* ``serving_requests``: a Poisson stream of assemble-and-solve requests on
  hex8 / tet10 boxes with log-uniform DoF counts, random materials and a
  random traction magnitude on one face (the opposite face clamped).

Only numpy's seeded generator is used, so a seed fixes the workload.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import mesh as M
from . import refelem as R

KINDS = {"hex8": ("hex", 1), "tet10": ("tet", 2), "wedge6": ("prism", 1)}


def box_dofs(kind: str, cells) -> int:
    """Nodal DoFs (3 per node) of a structured box of the given kind."""
    _, order = KINDS[kind]
    nx, ny, nz = cells
    return 3 * (order * nx + 1) * (order * ny + 1) * (order * nz + 1)


def rules(kind: str):
    """(stiffness rule, consistent-mass rule) per element kind: hex8 and
    wedge6 use degree 3 (2x2x2 Gauss; the 6-point wedge rule) for both;
    tet10 uses the 4-point degree-2 rule for stiffness (exact on straight
    sides) and a degree-4 rule for the mass (P2 x P2 is degree 4)."""
    elem, order = KINDS[kind]
    et = M.structured_box_mesh(n=1, elem=elem, order=order).etype
    if kind == "tet10":
        return R.rule_tables(et, 2), R.rule_tables(et, 4)
    t = R.rule_tables(et, 3)
    return t, t


@dataclass
class BatchProblem:
    kind: str
    cells: tuple
    E: float
    nu: float
    rho: float

    def mesh(self) -> M.Mesh:
        elem, order = KINDS[self.kind]
        nx, ny, nz = self.cells
        return M.structured_box_mesh(lx=1.0, ly=ny / nx, lz=nz / nx, cells=self.cells, elem=elem,
                                     order=order)


def batched_problems(seed: int = 2024, count: int = 256, dof_range=(2_000, 400_000),
                     edge_range=(4, 48)) -> list[BatchProblem]:
    rng = np.random.default_rng(seed)
    kinds = list(KINDS)
    out = []
    while len(out) < count:
        kind = kinds[int(rng.integers(0, len(kinds)))]
        cells = tuple(int(c) for c in rng.integers(edge_range[0], edge_range[1] + 1, size=3))
        if not (dof_range[0] <= box_dofs(kind, cells) <= dof_range[1]):
            continue
        out.append(BatchProblem(kind, cells, float(rng.uniform(1e5, 3e5)), float(rng.uniform(0.2, 0.45)),
                                7.85e-9))
    return out


@dataclass
class Request:
    rid: int
    arrival_s: float
    kind: str            # hex8 | tet10
    cells: tuple
    E: float
    nu: float
    traction: float      # fy on the loaded face
    load_face: str       # the clamped face is the opposite one

    @property
    def clamp_face(self) -> str:
        return {"right": "left", "left": "right", "top": "bottom", "bottom": "top",
                "front": "rear", "rear": "front"}[self.load_face]

    def mesh(self) -> M.Mesh:
        elem, order = KINDS[self.kind]
        nx, ny, nz = self.cells
        return M.structured_box_mesh(lx=1.0, ly=ny / nx, lz=nz / nx, cells=self.cells, elem=elem,
                                     order=order)


def _cells_for_dofs(kind: str, dofs: float, rng) -> tuple:
    """A box of aspect ratio 1:1:a (a ~ U[1, 4]) whose DoF count is near `dofs`."""
    _, order = KINDS[kind]
    a = float(rng.uniform(1.0, 4.0))
    nodes = dofs / 3.0
    n = max(1, int(round(((nodes / a) ** (1.0 / 3.0) - 1.0) / order)))
    nz = max(1, int(round(a * n)))
    return (n, n, nz)


def serving_requests(seed: int = 7, count: int = 64, rate_hz: float = 20.0, dof_range=(5_000, 2_000_000),
                     distinct_meshes: int | None = 12) -> list[Request]:
    """Poisson arrivals at `rate_hz`; DoF counts log-uniform in `dof_range`.
    With `distinct_meshes`, requests draw their box from a pool of that many
    shapes, so same-mesh requests recur (device state is reused for them)."""
    rng = np.random.default_rng(seed)
    pool = None
    if distinct_meshes:
        pool = []
        for _ in range(distinct_meshes):
            kind = ("hex8", "tet10")[int(rng.integers(0, 2))]
            dofs = math.exp(rng.uniform(math.log(dof_range[0]), math.log(dof_range[1])))
            pool.append((kind, _cells_for_dofs(kind, dofs, rng)))
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate_hz))
        if pool is not None:
            kind, cells = pool[int(rng.integers(0, len(pool)))]
        else:
            kind = ("hex8", "tet10")[int(rng.integers(0, 2))]
            dofs = math.exp(rng.uniform(math.log(dof_range[0]), math.log(dof_range[1])))
            cells = _cells_for_dofs(kind, dofs, rng)
        out.append(Request(i, t, kind, cells, float(rng.uniform(1e5, 3e5)), float(rng.uniform(0.2, 0.45)),
                           float(rng.uniform(-1.0, 1.0)), "right"))
    return out
