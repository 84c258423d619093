"""Synthetic workloads of the tumour-spheroid model: simulation descriptors
and seeded initial cell state for the five benchmark configurations.

Cell schema of the built-in tumour rule: pos.x, pos.y, pos.z (f32), id (u32),
age (f32), generation (u32) = 24 bytes per cell.

Initial positions and ages are inputs, drawn on the host with numpy's PCG64
generator from the seed; the run itself draws from Philox4x32-10 keyed by
the same seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .hip import _abi
from .hip._abi import FieldDesc, SimDesc

TUMOR_COL_TYPES = [_abi.COL_F32, _abi.COL_F32, _abi.COL_F32, _abi.COL_U32, _abi.COL_F32, _abi.COL_U32]
TUMOR_COL_NAMES = ["pos.x", "pos.y", "pos.z", "id", "age", "generation"]
CELL_BYTES = 24


@dataclass
class FieldSpec:
    rate: float
    decay: float = 0.0
    init: float = 0.0
    bc: int = _abi.BC_DIRICHLET
    bc_value: float = 0.0
    src_radius: float = 0.0
    uptake: float = 0.0
    lo: float = 0.0
    hi: float = 1.0

    def to_c(self) -> FieldDesc:
        return FieldDesc(self.rate, self.decay, self.init, self.bc, self.bc_value, self.src_radius,
                         self.uptake, self.lo, self.hi)


@dataclass
class Workload:
    name: str
    dims: tuple[int, int, int]
    dx: float
    key_format: int
    capacity: int
    seed: int
    fields: list[FieldSpec]
    cell_radius: float
    interaction_radius: float
    k_rep: float
    k_adh: float
    mobility: float
    max_disp: float
    params: list[float]            # div_age, div_o2, hypoxia, drug_kill, daughter_offset
    steps: int
    n_cells: int
    col_types: list[int] = field(default_factory=lambda: list(TUMOR_COL_TYPES))

    def desc(self) -> SimDesc:
        d = SimDesc()
        d.dims[:] = list(self.dims)
        d.dx = self.dx
        d.key_format = self.key_format
        d.capacity = self.capacity
        d.seed = self.seed
        d.cell_radius = self.cell_radius
        d.interaction_radius = self.interaction_radius
        d.k_rep = self.k_rep
        d.k_adh = self.k_adh
        d.mobility = self.mobility
        d.max_disp = self.max_disp
        d.n_fields = len(self.fields)
        for i, f in enumerate(self.fields):
            d.fields[i] = f.to_c()
        d.n_cols = len(self.col_types)
        for i, t in enumerate(self.col_types):
            d.col_types[i] = t
        d.n_params = len(self.params)
        for i, p in enumerate(self.params):
            d.params[i] = p
        return d

    @property
    def nvox(self) -> int:
        return self.dims[0] * self.dims[1] * self.dims[2]


def _sphere(rng: np.random.Generator, n: int, center, radius: float) -> np.ndarray:
    u = rng.random(n)
    r = radius * np.cbrt(u)
    z = rng.random(n) * 2.0 - 1.0
    phi = rng.random(n) * 2.0 * math.pi
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    p = np.stack([r * s * np.cos(phi), r * s * np.sin(phi), r * z], axis=1) + np.asarray(center, dtype=np.float64)
    return p


def tumor_columns(pos: np.ndarray, ages: np.ndarray, dims, dx: float) -> list[np.ndarray]:
    n = len(pos)
    hi = np.array(dims, dtype=np.float64) * dx
    p = np.clip(pos, 0.0, hi).astype(np.float32)
    return [np.ascontiguousarray(p[:, 0]), np.ascontiguousarray(p[:, 1]), np.ascontiguousarray(p[:, 2]),
            np.arange(n, dtype=np.uint32), ages.astype(np.float32), np.zeros(n, dtype=np.uint32)]


def initial_cells(w: Workload, n: Optional[int] = None) -> list[np.ndarray]:
    """Seeded initial state: positions per workload, integer ages in [0, div_age]."""
    n = w.n_cells if n is None else n
    rng = np.random.Generator(np.random.PCG64(w.seed))
    c = [d * w.dx * 0.5 for d in w.dims]
    if w.name.startswith("plain"):
        pos = _sphere(rng, n, c, 8.0)
    elif w.name.startswith("large"):
        pos = _sphere(rng, n, c, 80.0) + rng.normal(0.0, 0.3, size=(n, 3))
    elif w.name.startswith("stress"):
        k = 32
        centers = rng.uniform(0.25, 0.75, size=(k, 3)) * np.array(w.dims) * w.dx
        lab = rng.integers(0, k, size=n)
        pos = centers[lab] + rng.normal(0.0, 20.0, size=(n, 3))
    else:
        side = w.dims[0] * w.dx
        radius = min(0.45 * side, (3.0 * n / (4.0 * math.pi * 0.5)) ** (1.0 / 3.0))
        pos = _sphere(rng, n, c, radius)
    ages = rng.integers(0, int(w.params[0]) + 1, size=n).astype(np.float32)
    return tumor_columns(pos, ages, w.dims, w.dx)


def plain_c1(seed: int = 42) -> Workload:
    """4,096 cells in a radius-8 sphere, 64^3, one O2 field, 100 steps."""
    return Workload(
        name="plain_4k_64",
        dims=(64, 64, 64), dx=1.0, key_format=_abi.KEY_MORTON30, capacity=1 << 17, seed=seed,
        fields=[FieldSpec(rate=0.1, init=1.0, bc=_abi.BC_DIRICHLET, bc_value=1.0, uptake=0.01)],
        cell_radius=0.5, interaction_radius=1.0, k_rep=1.0, k_adh=0.0, mobility=0.25, max_disp=0.25,
        params=[24.0, 0.3, 0.05, 0.0, 0.5], steps=100, n_cells=4096)


def large_c2(seed: int = 42, n_cells: int = 1_000_000) -> Workload:
    """1M cells in a radius-80 sphere (+N(0, 0.3^2) jitter), 256^3, O2 + drug.

    The configuration does not give a cell radius; 0.4 is used (contact
    distance 0.8) so that the adhesion taper between contact and the 1.0
    interaction radius is not empty: with the 4,096-cell case's 0.5 the
    contact distance equals the interaction radius and k_adh would act on no
    pair."""
    return Workload(
        name="large_1m_256",
        dims=(256, 256, 256), dx=1.0, key_format=_abi.KEY_MORTON32, capacity=2 * n_cells, seed=seed,
        fields=[FieldSpec(rate=0.1, init=1.0, bc=_abi.BC_DIRICHLET, bc_value=1.0, uptake=0.01),
                FieldSpec(rate=0.05, decay=0.001, init=0.0, bc=_abi.BC_POINT_SOURCE, bc_value=1.0,
                          src_radius=16.0)],
        cell_radius=0.4, interaction_radius=1.0, k_rep=1.0, k_adh=0.2, mobility=0.25, max_disp=0.25,
        params=[24.0, 0.3, 0.05, 0.02, 0.5], steps=1000, n_cells=n_cells)


def stress_c5(seed: int = 42, n_cells: int = 16_000_000, side: int = 512) -> Workload:
    """16M cells in 32 Gaussian clusters (sigma 20 voxels), 512^3, three
    fields: O2, drug, glucose (the built-in rule reads field 0 as O2 and field
    1 as the drug; glucose diffuses and is consumed)."""
    return Workload(
        name="stress_16m_512",
        dims=(side, side, side), dx=1.0, key_format=_abi.KEY_MORTON32, capacity=2 * n_cells, seed=seed,
        fields=[FieldSpec(rate=0.1, init=1.0, bc=_abi.BC_DIRICHLET, bc_value=1.0, uptake=0.01),
                FieldSpec(rate=0.05, decay=0.001, init=0.0, bc=_abi.BC_POINT_SOURCE, bc_value=1.0,
                          src_radius=32.0),
                FieldSpec(rate=0.1, init=1.0, bc=_abi.BC_DIRICHLET, bc_value=1.0, uptake=0.005)],
        cell_radius=0.4, interaction_radius=1.0, k_rep=1.0, k_adh=0.2, mobility=0.25, max_disp=0.25,
        params=[24.0, 0.3, 0.05, 0.02, 0.5], steps=300, n_cells=n_cells)


def ensemble_c3(seed: int = 42, n_sims: int = 64, scale: float = 1.0) -> list[Workload]:
    """64 sims, cells log-uniform in [10k, 500k] (times scale), sides from
    {64, 128, 192}, per-sim uptake, division age and hypoxia threshold."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n_sims):
        n = int(round(math.exp(rng.uniform(math.log(10_000), math.log(500_000))) * scale))
        n = max(16, n)
        side = int(rng.choice([64, 128, 192]))
        uptake = float(rng.uniform(0.005, 0.02))
        div_age = float(rng.uniform(18.0, 30.0))
        hyp = float(rng.uniform(0.03, 0.08))
        out.append(Workload(
            name=f"ensemble_{i}", dims=(side, side, side), dx=1.0, key_format=_abi.KEY_MORTON30,
            capacity=2 * n, seed=seed * 1000 + i,
            fields=[FieldSpec(rate=0.1, init=1.0, bc=_abi.BC_DIRICHLET, bc_value=1.0, uptake=uptake)],
            cell_radius=0.5, interaction_radius=1.0, k_rep=1.0, k_adh=0.0, mobility=0.25, max_disp=0.25,
            params=[div_age, 0.3, hyp, 0.0, 0.5], steps=500, n_cells=n))
    return out


def served_job(rng: np.random.Generator, job_id: int, seed: int = 42, scale: float = 1.0,
               max_cells: int = 4_000_000) -> tuple[Workload, int]:
    """One served job: log-normal cells (median 100k, sigma 1.0, capped),
    side = roundup32(max(64, 2 cbrt(cells))), horizon U[100, 2000]."""
    n = int(min(max_cells, math.exp(rng.normal(math.log(100_000), 1.0))) * scale)
    n = max(16, n)
    side = max(64, int(2.0 * n ** (1.0 / 3.0)))
    side = (side + 31) // 32 * 32
    horizon = int(rng.integers(100, 2001))
    w = Workload(
        name=f"served_{job_id}", dims=(side, side, side), dx=1.0,
        key_format=_abi.KEY_MORTON30 if side <= 1024 else _abi.KEY_MORTON32,
        capacity=2 * n, seed=seed * 100_000 + job_id,
        fields=[FieldSpec(rate=0.1, init=1.0, bc=_abi.BC_DIRICHLET, bc_value=1.0, uptake=0.01)],
        cell_radius=0.5, interaction_radius=1.0, k_rep=1.0, k_adh=0.0, mobility=0.25, max_disp=0.25,
        params=[24.0, 0.3, 0.05, 0.0, 0.5], steps=horizon, n_cells=n)
    return w, horizon
