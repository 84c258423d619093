"""Engine: owns the model's host state and drives a backend, with the
lifecycle of framework/src/oncolytica/core/runtime/_engine.py:19-212
(setup_geometry / setup_tissue / load_model / run_step / sync_to_host).

Differences, each a deterministic-spec decision:
  * backend "hip" (alias "gpu") is the MI355X backend; there is no CPU backend;
  * grid dimensions may be any size up to 1024 per axis (2048 in x and y with
    the 32-bit key format) instead of even sizes up to 240
    (_engine.py:84-96), since bins are dense and the scan is single-pass;
  * update_params reaches the device before the next step (the reference
    never re-uploads uniforms, _engine.py:64-68);
  * setup_mechanics enables the built-in neighbour force law (off by default,
    as the reference has no built-in forces).
"""
from __future__ import annotations

from typing import Any, Optional, Type

from ._containers import AgentList, Grid
from ._errors import CompilationError
from ._types import Chemistry, Simulation, Tissue


class Engine:
    def __init__(self, backend: str = "hip", max_agents: int = 1024 * 256, device: int = 0, seed: int = 0):
        if backend not in ("hip", "gpu"):
            raise ValueError(f"Unknown backend {backend!r}. Choose 'hip' (alias 'gpu').")
        self.backend = backend
        self.max_agents = int(max_agents)
        self.device = int(device)
        self.seed = int(seed)
        self.tissue_voxel_size = 1.4
        self.chemistry_voxel_size = 2.8
        self.cell_diameter = 1.0
        self.tissue_shape: tuple = (64, 64, 64)
        self.chemistry_shape: tuple = (32, 32, 32)
        self.key_format = 30
        self.mechanics = {"k_rep": 0.0, "k_adh": 0.0, "mobility": 0.25, "max_disp": 0.25,
                          "interaction_radius": None, "cell_radius": None}
        self._step = 0
        self.cells = AgentList(self)
        self.tissue: Optional[Grid] = None
        self.chemistry: Optional[Grid] = None
        self._metrics = None
        self._sim: Optional[Simulation] = None
        self._tissue_rules: list = []
        self._chemistry_rules: list = []
        self._cell_rules: list = []
        self._metric_rules: list = []
        self._backend_impl = None

    # ---- setup ----------------------------------------------------------------------
    def setup_params(self, params) -> None:
        self._sim._params = params
        if self._backend_impl is not None:
            self._backend_impl.update_params()

    def update_params(self, **kwargs: Any) -> None:
        for k, v in kwargs.items():
            self._sim._params._validate_field(k, v)
        for k, v in kwargs.items():
            object.__setattr__(self._sim._params, k, v)
        if self._backend_impl is not None:
            self._backend_impl.update_params()

    def setup_geometry(self, tissue_shape: tuple = (120, 120, 120), tissue_voxel_size: float = 1.4,
                       cell_diameter: float = 1.0, key_format: Optional[int] = None) -> None:
        if not isinstance(tissue_shape, (list, tuple)) or len(tissue_shape) != 3:
            raise TypeError(f"tissue_shape must be a tuple of 3 ints, got {tissue_shape!r}")
        for i, d in enumerate(tissue_shape):
            if not isinstance(d, int) or isinstance(d, bool):
                raise TypeError(f"tissue_shape component {i} must be int, got {type(d).__name__!r}")
            if d <= 0:
                raise ValueError(f"tissue_shape component {i} must be positive, got {d}")
        fmt = key_format or (30 if max(tissue_shape) <= 1024 else 32)
        lim = (2048, 2048, 1024) if fmt == 32 else (1024, 1024, 1024)
        for i, d in enumerate(tissue_shape):
            if d > lim[i]:
                raise ValueError(f"tissue_shape component {i} must be <= {lim[i]}, got {d}")
        if not tissue_voxel_size > 0:
            raise ValueError("tissue_voxel_size must be > 0")
        self.key_format = fmt
        self.tissue_shape = tuple(tissue_shape)
        self.chemistry_shape = tuple(max(1, d // 2) for d in tissue_shape)
        self.cell_diameter = float(cell_diameter)
        self.tissue_voxel_size = float(tissue_voxel_size)
        self.chemistry_voxel_size = 2.0 * float(tissue_voxel_size)

    def setup_tissue(self, tissue_class: Type) -> None:
        if not (isinstance(tissue_class, type) and issubclass(tissue_class, Tissue)):
            raise ValueError(f"tissue_class must be a subclass of ol.Tissue, got {tissue_class}")
        self.tissue = Grid(tissue_class, self.tissue_shape, self)

    def setup_chemistry(self, chemistry_class: Type) -> None:
        if not (isinstance(chemistry_class, type) and issubclass(chemistry_class, Chemistry)):
            raise ValueError(f"chemistry_class must be a subclass of ol.Chemistry, got {chemistry_class}")
        self.chemistry = Grid(chemistry_class, self.chemistry_shape, self)

    def setup_mechanics(self, k_rep: float = 1.0, k_adh: float = 0.0, mobility: float = 0.25,
                        max_disp: float = 0.25, interaction_radius: Optional[float] = None,
                        cell_radius: Optional[float] = None) -> None:
        """Built-in pairwise force law over the 27 neighbour voxels: linear
        repulsion k_rep (2 r - d) inside contact, linear adhesion taper k_adh
        up to interaction_radius (default: the voxel size), overdamped step
        of mobility * F clamped to max_disp."""
        self.mechanics = {"k_rep": float(k_rep), "k_adh": float(k_adh), "mobility": float(mobility),
                          "max_disp": float(max_disp), "interaction_radius": interaction_radius,
                          "cell_radius": cell_radius}
        if self._backend_impl is not None:
            raise RuntimeError("setup_mechanics must be called before load_model")

    @property
    def step(self) -> int:
        return self._step

    # ---- model ------------------------------------------------------------------------
    def load_model(self, sim_instance: Any) -> None:
        if not isinstance(sim_instance, Simulation):
            raise CompilationError(f"Expected an instance of 'ol.Simulation', got '{type(sim_instance).__name__}'.")
        self._sim = sim_instance
        self._metrics = sim_instance._spec.metrics_class(sim_instance)
        sim_instance._engine = self
        self._tissue_rules.clear()
        self._chemistry_rules.clear()
        self._cell_rules.clear()
        self._metric_rules.clear()
        for name in dir(type(sim_instance)):
            if name.startswith("_"):
                continue
            attr = getattr(type(sim_instance), name, None)
            if isinstance(attr, property) or not callable(attr):
                continue
            rt = getattr(attr, "_rule_type", None)
            if rt is None:
                continue
            bound = getattr(sim_instance, name)
            {"tissue": self._tissue_rules, "chemistry": self._chemistry_rules, "cell": self._cell_rules,
             "metric": self._metric_rules}[rt].append(bound)
        # declaration order (dir() is alphabetical)
        order = {n: i for i, n in enumerate(_declared_order(type(sim_instance)))}
        for lst in (self._tissue_rules, self._chemistry_rules, self._cell_rules, self._metric_rules):
            lst.sort(key=lambda m: order.get(m.__name__, 1 << 30))
        from .hip._backend import HipBackend
        # the reference passes its validator's context here (core/runtime/_engine.py:153-157);
        # this package has no separate validator: the translator checks the rules
        self._backend_impl = HipBackend(self, None)
        self._backend_impl.compile(sim_instance)

    # ---- execution ----------------------------------------------------------------------
    def run_step(self, collect_metrics: bool = False) -> None:
        if self._sim is None or self._backend_impl is None:
            raise RuntimeError("Engine.run_step() called before load_model().")
        self._step += 1
        self._backend_impl.run_step(collect_metrics=collect_metrics)

    def run_steps(self, n: int) -> None:
        """n steps without host synchronisation (one launch sequence per step)."""
        if self._sim is None or self._backend_impl is None:
            raise RuntimeError("Engine.run_steps() called before load_model().")
        self._step += n
        self._backend_impl.run_steps(n)

    def get_metrics(self):
        return self._metrics

    def sync_to_host(self) -> None:
        self._backend_impl.sync_to_host()

    def sync_to_device(self) -> None:
        self._backend_impl.sync_to_device()

    def __repr__(self) -> str:
        return f"Engine(backend={self.backend!r}, step={self._step}, cells={len(self.cells)})"


def _declared_order(cls) -> list[str]:
    out = []
    for klass in reversed(cls.__mro__):
        for name in klass.__dict__:
            if name not in out:
                out.append(name)
    return out
