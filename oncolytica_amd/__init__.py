"""oncolytica_amd (ol): the MI355X (gfx950) implementation of the Oncolytica
per-timestep hot path behind the reference's rule API.

Same surface as the reference package (framework/src/oncolytica/__init__.py:
9-59): data-layout bases, vec3 / ivec3, scalar aliases, rule decorators,
Simulation, Engine, math intrinsics, RNG entry points, CompilationError.
Additions: ol.diffusing / ol.dirichlet / ol.point_source (declarative
substrate fields) and Engine.setup_mechanics (built-in force law).
"""
from __future__ import annotations

from ._types import (f32, i32, u32, bool, vec3, ivec3, BaseData, Tissue, Chemistry, Cell, Metrics, Params,  # noqa: A004
                     Simulation, random, random_dir, diffusing, dirichlet, point_source, Diffusing, Boundary)
from . import _math as math  # noqa: A004
from ._containers import Grid, AgentList
from ._decorators import tissue_rule, chemistry_rule, cell_rule, metric_rule
from ._engine import Engine
from ._errors import CompilationError

__all__ = [
    "f32", "i32", "u32", "bool", "vec3", "ivec3",
    "BaseData", "Tissue", "Chemistry", "Cell", "Metrics", "Params",
    "math", "Grid", "AgentList",
    "tissue_rule", "chemistry_rule", "cell_rule", "metric_rule",
    "Simulation", "Engine", "CompilationError",
    "random", "random_dir", "diffusing", "dirichlet", "point_source",
]

__version__ = "2.0.0+mi355x"
