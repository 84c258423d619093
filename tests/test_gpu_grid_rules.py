"""Translated rules of every kind against the rule executor (oracle/rule_exec.py).

Each model runs through Engine(backend="hip") and, from the same device
state, through the binary32 executor, which runs the model's own Python
rules; cells, tissue fields and chemistry columns must agree bit for bit
after every checked step (a model that calls transcendental ol.math
functions agrees to the stated tolerance instead). Models: a glioma-like
model with a tissue rule, a two-iteration chemistry rule over the coarse grid
(`.cells`, `.neighbors`, `.tissues`, helper methods, tuples) and a cell rule
reading chemistry_at / tissue_at; the broad cell-rule model of
tests/test_translator.py (helpers, tuples, while / break, range loops,
neighbour walks, integer floor division and modulo); and restatements of the
reference's grid-rule tests (framework/tests/test_oncolytica.py:266-296,
407-434).
"""
import numpy as np
import pytest

import oncolytica_amd as ol
from oracle.rule_exec import RuleState

pytestmark = pytest.mark.gpu

TYPE_INACTIVE: ol.i32 = 0
TYPE_INFECTED: ol.i32 = 1
TYPE_RECRUITED: ol.i32 = 2
PDGF_D: ol.f32 = 0.0417
PDGF_DECAY: ol.f32 = 0.005
PDGF_SECRETE: ol.f32 = 0.417
PDGF_CONSUME: ol.f32 = 0.042


class GTissue(ol.Tissue):
    oxygen: ol.f32 = ol.diffusing(rate=0.1, init=1.0, boundary=ol.dirichlet(1.0), uptake=0.002)
    crowd: ol.i32 = 0
    carrying_capacity: ol.i32 = 3

    def is_overcrowded(self, n: ol.i32) -> ol.bool:
        return n >= self.carrying_capacity


class GChem(ol.Chemistry):
    pdgf: ol.f32 = 0.0
    hits: ol.u32 = 0

    def compute_secretion_consumption(self) -> tuple[ol.f32, ol.f32]:
        secretion: ol.f32 = 0.0
        consumption: ol.f32 = 0.0
        for c in self.cells:
            if c.cell_type == TYPE_INFECTED:
                secretion += PDGF_SECRETE
            if c.cell_type != TYPE_INACTIVE:
                consumption += PDGF_CONSUME
        return secretion, consumption

    def compute_laplacian(self) -> ol.f32:
        lap: ol.f32 = 0.0
        for nb in self.neighbors:
            lap += nb.pdgf - self.pdgf
        return lap

    def step_pdgf(self):
        secretion, consumption = self.compute_secretion_consumption()
        lap: ol.f32 = self.compute_laplacian()
        new_val: ol.f32 = (self.pdgf + secretion - consumption + PDGF_D * 0.5 * lap) * (1.0 - PDGF_DECAY)
        self.pdgf = ol.math.clamp(new_val, 0.0, 1000000.0)


class GCell(ol.Cell):
    pos: ol.vec3
    cell_type: ol.i32 = 0
    energy: ol.f32 = 1.0


class GMetrics(ol.Metrics):
    n: ol.u32 = 0


class GParams(ol.Params):
    thr: ol.f32 = 0.05
    step_len: ol.f32 = 0.2


class Glioma(ol.Simulation[GTissue, GChem, GCell, GMetrics, GParams]):

    @ol.tissue_rule
    def count_cells(self, t: GTissue):
        n: ol.i32 = 0
        for c in t.cells:
            if c.energy > 0.25:
                n += 1
        t.crowd = n
        if ol.random() < 0.05:
            t.oxygen = t.oxygen * 0.99

    @ol.chemistry_rule(iterations=2)
    def update_pdgf(self, chem: GChem):
        chem.step_pdgf()
        for t in chem.tissues:
            if t.crowd > 0:
                chem.hits += 1

    @ol.cell_rule
    def behave(self, cell: GCell):
        chem: GChem = self.chemistry_at(cell.pos)
        tissue: GTissue = self.tissue_at(cell.pos)
        if cell.cell_type == TYPE_INACTIVE:
            if chem.pdgf > self.params.thr:
                cell.cell_type = TYPE_RECRUITED
            return
        if tissue.is_overcrowded(tissue.crowd):
            cell.energy -= 0.1
        if cell.energy < 0.0:
            cell.die()
            return
        cell.pos += ol.random_dir() * self.params.step_len
        if ol.random() < 0.02 * tissue.oxygen:
            cell.divide(GCell(pos=cell.pos, cell_type=cell.cell_type, energy=cell.energy * 0.5))

    @ol.metric_rule
    def count(self, cell: GCell, m: GMetrics):
        m.n += 1


def _engine(sim, tissue_cls, chem_cls, shape, cols: dict, seed=7, cap=1 << 14, k_rep=1.0):
    eng = ol.Engine(backend="hip", max_agents=cap, seed=seed)
    eng.setup_geometry(shape, 1.0, cell_diameter=1.0)
    eng.setup_mechanics(k_rep=k_rep, k_adh=0.0, mobility=0.25, max_disp=0.25, interaction_radius=1.0,
                        cell_radius=0.5)
    eng.setup_tissue(tissue_cls)
    if chem_cls is not None:
        eng.setup_chemistry(chem_cls)
    eng.cells.set_columns(cols)
    return eng


def _executor(eng, sim):
    be = eng._backend_impl
    s = be.sim
    tissue = {name: s.download_field(f) for f, (name, _, _) in enumerate(be.tissue.fields)}
    chem = None
    if be.chem is not None and be.chem.cls is not None:
        chem = {name: s.download_chem(k) for k, (name, _) in enumerate(be.chem.fields)}
    return RuleState(s.desc, sim, s.download_cells(), tissue, chem, cell_rules=eng._cell_rules,
                     tissue_rules=eng._tissue_rules, chem_rules=eng._chemistry_rules, next_id=s.counts()[1],
                     step=s.counts()[2])


def _compare(eng, ex, tag, atol_ulp=0):
    be = eng._backend_impl
    s = be.sim
    got = s.download_cells()
    assert len(got[0]) == len(ex.cols[0]), f"{tag}: {len(got[0])} cells vs executor {len(ex.cols[0])}"
    for k, (g, o) in enumerate(zip(got, ex.cols)):
        g32, o32 = g.view(np.uint32), np.ascontiguousarray(o).astype(g.dtype).view(np.uint32)
        if atol_ulp == 0 or g.dtype != np.float32:
            bad = np.nonzero(g32 != o32)[0]
        else:
            bad = np.nonzero(np.abs(g32.astype(np.int64) - o32.astype(np.int64)) > atol_ulp)[0]
        assert bad.size == 0, f"{tag}: column {be.schema.columns[k].name} differs at {bad[:5]}: {g[bad[:5]]} vs {o[bad[:5]]}"
    for f, (name, _, _) in enumerate(be.tissue.fields):
        g = s.download_field(f)
        bad = np.nonzero(g.view(np.uint32) != ex.tissue[name].view(np.uint32))[0]
        assert bad.size == 0, f"{tag}: tissue field {name} differs at {bad[:5]}"
    if be.chem is not None and be.chem.cls is not None:
        for k, (name, _) in enumerate(be.chem.fields):
            g = s.download_chem(k)
            o = np.asarray(ex.chem[name]).astype(g.dtype)
            bad = np.nonzero(g.view(np.uint32) != o.view(np.uint32))[0]
            assert bad.size == 0, f"{tag}: chemistry column {name} differs at {bad[:5]}: {g[bad[:5]]} vs {o[bad[:5]]}"


def _ball(rng, n, centre, radius):
    u = rng.random(n)
    r = radius * np.cbrt(u)
    z = rng.random(n) * 2 - 1
    phi = rng.random(n) * 2 * np.pi
    s = np.sqrt(1 - z * z)
    return (np.stack([r * s * np.cos(phi), r * s * np.sin(phi), r * z], 1) + centre).astype(np.float32)


def test_glioma_grid_rules_match_the_executor():
    rng = np.random.default_rng(5)
    n = 1200
    pos = _ball(rng, n, np.array([12.0, 12.0, 12.0]), 6.0)
    cols = {"pos": pos, "_id": np.arange(n, dtype=np.uint32),
            "cell_type": rng.integers(0, 3, n).astype(np.int32),
            "energy": rng.uniform(0.2, 1.0, n).astype(np.float32)}
    sim = Glioma()
    eng = _engine(sim, GTissue, GChem, (24, 24, 24), cols)
    eng.chemistry.flat("pdgf")[:] = rng.uniform(0.0, 0.2, eng.chemistry.flat("pdgf").size).astype(np.float32)
    eng.load_model(sim)
    ex = _executor(eng, sim)
    done = 0
    for k in (1, 2, 5):
        eng.run_steps(k - done)
        ex.run(k - done)
        done = k
        print(f"glioma: step {k}, {len(ex.cols[0])} cells", flush=True)
        _compare(eng, ex, f"step {k}")
    assert ex.next_id > n  # divisions happened
    assert int(ex.chem["hits"].max()) > 0 and float(ex.chem["pdgf"].max()) > 0.0


# the broad cell-rule model of the translator tests
from tests.test_translator import BCell, BTissue, Broad  # noqa: E402


def test_broad_cell_rules_match_the_executor():
    rng = np.random.default_rng(9)
    n = 900
    pos = _ball(rng, n, np.array([10.0, 10.0, 10.0]), 5.0)
    cols = {"pos": pos, "_id": np.arange(n, dtype=np.uint32), "vel": rng.normal(0, 1, (n, 3)).astype(np.float32),
            "kind": rng.integers(0, 3, n).astype(np.int32), "energy": rng.uniform(0.0, 6.0, n).astype(np.float32),
            "flag": np.zeros(n, bool), "tag": np.zeros(n, np.uint32)}
    sim = Broad()
    eng = _engine(sim, BTissue, None, (20, 20, 20), cols, seed=11)
    eng.load_model(sim)
    ex = _executor(eng, sim)
    for k in range(1, 4):
        eng.run_steps(1)
        ex.run(1)
        print(f"broad: step {k}, {len(ex.cols[0])} cells", flush=True)
        _compare(eng, ex, f"step {k}")


class ICell(ol.Cell):
    pos: ol.vec3
    health: ol.f32 = 100.0


class IChem(ol.Chemistry):
    drug_conc: ol.f32 = 0.0


class ITissue(ol.Tissue):
    oxygen: ol.f32 = 1.0


class IMetrics(ol.Metrics):
    n: ol.u32 = 0


class IParams(ol.Params):
    pass


class IterSim(ol.Simulation[ITissue, IChem, ICell, IMetrics, IParams]):
    @ol.chemistry_rule(iterations=3)
    def diffuse(self, chem: IChem):
        chem.drug_conc += 1.0

    @ol.tissue_rule
    def tissue_logic(self, t: ITissue):
        t.oxygen -= 0.01

    @ol.cell_rule
    def cell_logic(self, cell: ICell):
        if cell.health > 50.0:
            cell.pos.x += 0.5


def test_reference_grid_rule_cases():
    """iterations=3 adds 3.0 per step to every chemistry voxel
    (test_oncolytica.py:407-434); a tissue rule lowers every voxel's oxygen
    by 0.01 and cells move 0.5 in x (test_oncolytica.py:266-296), on a
    non-cubic grid."""
    cols = {"pos": np.array([[1.0 + 5 * i, 1.0, 1.0] for i in range(4)], np.float32),
            "_id": np.arange(4, dtype=np.uint32)}
    sim = IterSim()
    eng = _engine(sim, ITissue, IChem, (120, 64, 32), cols, k_rep=0.0)
    eng.load_model(sim)
    ex = _executor(eng, sim)
    eng.run_step()
    eng.run_step()
    ex.run(2)
    _compare(eng, ex, "2 steps")
    eng.sync_to_host()
    assert np.all(eng.chemistry.flat("drug_conc") == 6.0)
    assert np.allclose(eng.tissue.flat("oxygen"), np.float32(1.0) - np.float32(0.01) - np.float32(0.01))
    np.testing.assert_array_equal(eng.cells.columns()["pos"][:, 0], [2.0, 7.0, 12.0, 17.0])


class TCell(ol.Cell):
    pos: ol.vec3
    e: ol.f32 = 1.0


class TSim(ol.Simulation[ITissue, IChem, TCell, IMetrics, IParams]):
    @ol.cell_rule
    def decay(self, cell: TCell):
        cell.e = ol.math.exp(-cell.e) + ol.math.sin(cell.pos.x) * 0.5 + ol.math.pow(cell.e, 0.5)


def test_transcendental_rules_within_tolerance():
    """Device exp / sin / pow against numpy's binary32 routines: a few ulp
    (the device math library is not bit-matched)."""
    rng = np.random.default_rng(2)
    n = 500
    cols = {"pos": _ball(rng, n, np.array([8.0, 8.0, 8.0]), 4.0), "_id": np.arange(n, dtype=np.uint32),
            "e": rng.uniform(0.1, 3.0, n).astype(np.float32)}
    sim = TSim()
    eng = _engine(sim, ITissue, None, (16, 16, 16), cols, k_rep=0.0)
    eng.load_model(sim)
    ex = _executor(eng, sim)
    eng.run_steps(1)
    ex.run(1)
    _compare(eng, ex, "1 step", atol_ulp=4)


def test_glioma_bundle_round_trip(tmp_path):
    """A bundle of a model with tissue and chemistry rules holds the
    chemistry grid: the restored sim continues bit for bit (an odd step
    count, so the two chemistry buffers have swapped roles)."""
    from oncolytica_amd import bundle as B
    rng = np.random.default_rng(8)
    n = 600
    cols = {"pos": _ball(rng, n, np.array([12.0, 12.0, 12.0]), 5.0), "_id": np.arange(n, dtype=np.uint32),
            "cell_type": rng.integers(0, 3, n).astype(np.int32), "energy": rng.uniform(0.2, 1.0, n).astype(np.float32)}
    sim = Glioma()
    eng = _engine(sim, GTissue, GChem, (24, 24, 24), cols)
    eng.chemistry.flat("pdgf")[:] = rng.uniform(0.0, 0.2, eng.chemistry.flat("pdgf").size).astype(np.float32)
    eng.load_model(sim)
    eng.run_steps(3)
    be = eng._backend_impl
    r, mod, cfg = B.load_bundle(be.ctx, be.export_bundle(str(tmp_path / "g")))
    try:
        assert cfg["desc"]["chem_col_types"] and cfg["state"]["step"] == 3
        for k in range(2):
            np.testing.assert_array_equal(r.download_chem(k).view(np.uint32), be.sim.download_chem(k).view(np.uint32))
        r.step(3)
        be.sim.step(3)
        for a, b in zip(r.download_cells(), be.sim.download_cells()):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
        for k in range(2):
            np.testing.assert_array_equal(r.download_chem(k).view(np.uint32), be.sim.download_chem(k).view(np.uint32))
        for f in range(len(be.tissue.fields)):
            np.testing.assert_array_equal(r.download_field(f).view(np.uint32), be.sim.download_field(f).view(np.uint32))
    finally:
        r.close()
        mod.close()
