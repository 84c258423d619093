"""HipBackend driven the way the reference Engine drives its GPU backend.

RefEngine below holds exactly the attributes the reference Engine sets in
__init__ (framework/src/oncolytica/core/runtime/_engine.py:26-55); __slots__
makes any other attribute a read error. RefAgentList is a list of cell
objects (core/utils/_containers.py:85-123) and RefGrid an [x, y, z] object
array of voxel records (_containers.py:17-82). load_model follows
_engine.py:122-161: rule discovery through dir(), then
HipBackend(engine, ctx) and compile(sim). The model is this package's
spheroid model (the reference package itself needs Python 3.12 and wgpu).
"""
import numpy as np
import pytest

import oncolytica_amd as ol
from oncolytica_amd import workloads as W
from oncolytica_amd.hip._backend import HipBackend, _grid_field, _set_grid_field
from oncolytica_amd.spheroid import Spheroid, SpheroidTissue, TumorCell


class RefAgentList:
    def __init__(self, engine):
        self._engine = engine
        self._data = []

    def __iter__(self):
        return iter(self._data)

    def __len__(self):
        return len(self._data)

    def add(self, agent):
        agent._alive = True
        agent._agent_list = self
        self._data.append(agent)

    def _remove_dead(self):
        before = len(self._data)
        self._data = [a for a in self._data if getattr(a, "_alive", True)]
        return before - len(self._data)


class RefGrid:
    def __init__(self, struct_class, shape, engine):
        self._struct_class = struct_class
        self._engine = engine
        self._shape = tuple(shape)
        self._data = np.empty(self._shape, dtype=object)
        for idx in np.ndindex(*self._shape):
            inst = struct_class()
            inst._grid = self
            inst._engine = engine
            inst._coord = ol.ivec3(*idx)
            self._data[idx] = inst

    @property
    def shape(self):
        return self._shape

    def __getitem__(self, idx):
        return self._data[idx]

    def __len__(self):
        return self._data.size


class RefEngine:
    __slots__ = ("backend", "max_agents", "tissue_voxel_size", "chemistry_voxel_size", "cell_diameter",
                 "tissue_shape", "chemistry_shape", "_step", "cells", "tissue", "chemistry", "_metrics", "_sim",
                 "_tissue_rules", "_chemistry_rules", "_cell_rules", "_metric_rules", "_backend_impl")

    def __init__(self, max_agents=1024 * 256):
        self.backend = "hip"
        self.max_agents = max_agents
        self.tissue_voxel_size = 1.4
        self.chemistry_voxel_size = 2.8
        self.cell_diameter = 1.0
        self.tissue_shape = (64, 64, 64)
        self.chemistry_shape = (32, 32, 32)
        self._step = 0
        self.cells = RefAgentList(self)
        self.tissue = None
        self.chemistry = None
        self._metrics = None
        self._sim = None
        self._tissue_rules = []
        self._chemistry_rules = []
        self._cell_rules = []
        self._metric_rules = []
        self._backend_impl = None

    def setup_geometry(self, tissue_shape, tissue_voxel_size, cell_diameter):
        self.tissue_shape = tissue_shape
        self.chemistry_shape = tuple(d // 2 for d in tissue_shape)
        self.cell_diameter = cell_diameter
        self.tissue_voxel_size = tissue_voxel_size
        self.chemistry_voxel_size = 2 * tissue_voxel_size

    def setup_tissue(self, cls):
        self.tissue = RefGrid(cls, self.tissue_shape, self)

    def load_model(self, sim, backend_ctx=None):
        self._sim = sim
        self._metrics = sim._spec.metrics_class(sim)
        sim._engine = self
        for name in dir(sim):
            if name.startswith("_"):
                continue
            try:
                method = getattr(sim, name)
            except Exception:
                continue
            if not callable(method):
                continue
            rt = getattr(method, "_rule_type", None)
            {"tissue": self._tissue_rules, "chemistry": self._chemistry_rules, "cell": self._cell_rules,
             "metric": self._metric_rules}.get(rt, []).append(method)
        self._backend_impl = HipBackend(self, backend_ctx)
        self._backend_impl.compile(sim)

    def run_step(self, collect_metrics=False):
        self._step += 1
        self._backend_impl.run_step(collect_metrics=collect_metrics)


def _seed_engine(cap=1 << 16):
    w = W.plain_c1()
    cols = W.initial_cells(w)
    e = RefEngine(max_agents=cap)
    e.setup_geometry((64, 64, 64), 1.0, 1.0)
    e.setup_tissue(SpheroidTissue)
    for i in range(len(cols[0])):
        e.cells.add(TumorCell(pos=ol.vec3(float(cols[0][i]), float(cols[1][i]), float(cols[2][i])),
                              age=float(cols[4][i]), generation=int(cols[5][i])))
    return e, cols


def test_desc_reads_only_reference_engine_attributes():
    e, _ = _seed_engine()
    sim = Spheroid()
    sim._engine = e
    be = HipBackend(e, object())
    from oncolytica_amd.hip._translator import CellSchema, TissueSchema
    be.schema = CellSchema.of(TumorCell)
    be.tissue = TissueSchema.of(SpheroidTissue)
    d = be._desc(sim)
    assert tuple(d.dims) == (64, 64, 64) and d.dx == 1.0 and d.capacity == 1 << 16
    assert d.key_format == 30 and d.seed == 0
    assert d.k_rep == 0.0 and d.k_adh == 0.0            # the reference has no built-in forces
    assert d.cell_radius == 0.5 and d.interaction_radius == 1.0
    assert d.n_fields == 2 and d.n_cols == 6


def test_object_grid_and_cells_round_trip():
    e, cols = _seed_engine()
    g = e.tissue
    g[3, 4, 5].oxygen = 0.25
    g[0, 0, 1].drug = 0.75
    ox = _grid_field(g, "oxygen")
    assert ox.shape == (64 ** 3,) and ox[3 + 64 * (4 + 64 * 5)] == 0.25 and np.count_nonzero(ox != 1.0) == 1
    dr = _grid_field(g, "drug")
    assert dr[64 * 64] == 0.75
    new = np.arange(64 ** 3, dtype=np.float32)
    _set_grid_field(g, "drug", new)
    assert g[1, 2, 3].drug == float(1 + 64 * (2 + 64 * 3)) and isinstance(g[1, 2, 3].drug, float)
    be = HipBackend(e)
    from oncolytica_amd.hip._translator import CellSchema
    be.schema = CellSchema.of(TumorCell)
    got = be._columns_from_objects(e.cells._data)
    for k in (0, 1, 2, 4, 5):
        np.testing.assert_array_equal(got[k], cols[k])
    np.testing.assert_array_equal(got[3], np.arange(len(cols[0]), dtype=np.uint32))


@pytest.mark.gpu
def test_reference_shaped_engine_runs_bit_exact_with_the_oracle():
    from oracle import oracle as O
    e, cols = _seed_engine()
    # the grid as the oracle initialises it (fixed voxels at their boundary value)
    ow = W.large_c2()
    ow.dims, ow.capacity, ow.key_format, ow.cell_radius, ow.k_adh, ow.k_rep, ow.seed = \
        (64, 64, 64), 1 << 16, 30, 0.5, 0.0, 0.0, 0
    d = ow.desc()
    _set_grid_field(e.tissue, "oxygen", O.field_init(d.fields[0], (64, 64, 64)))
    _set_grid_field(e.tissue, "drug", O.field_init(d.fields[1], (64, 64, 64)))
    e.load_model(Spheroid(), backend_ctx={"validated": True})
    be = e._backend_impl
    assert be.validation_ctx == {"validated": True}
    steps = 30
    for s in range(steps):
        e.run_step(collect_metrics=(s == steps - 1))
    be.sync_to_host()
    st = O.TumorState(d, cols)
    st.run(steps)
    oc = st.columns()
    objs = e.cells._data
    assert len(objs) == len(oc[0]) and all(o._alive and o._agent_list is e.cells for o in objs)
    for k, a in enumerate("xyz"):
        got = np.array([getattr(o.pos, a) for o in objs], dtype=np.float32)
        np.testing.assert_array_equal(got.view(np.uint32), oc[k].view(np.uint32))
    np.testing.assert_array_equal(np.array([o._id for o in objs], np.uint32), oc[3])
    np.testing.assert_array_equal(np.array([o.age for o in objs], np.float32).view(np.uint32),
                                  oc[4].view(np.uint32))
    np.testing.assert_array_equal(np.array([o.generation for o in objs], np.uint32), oc[5])
    np.testing.assert_array_equal(_grid_field(e.tissue, "oxygen").astype(np.float32).view(np.uint32),
                                  st.fields[0].view(np.uint32))
    assert e._metrics.alive == len(oc[0])
