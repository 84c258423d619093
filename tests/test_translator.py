"""The Python-AST -> HIP translator on the CPU: emitted modules compile for
gfx950 with hipcc (the same front end hiprtc uses), and rules the backend
cannot run are rejected with CompilationError."""
import os
import subprocess
import tempfile

import pytest

import oncolytica_amd as ol
from oncolytica_amd.hip._translator import CellSchema, TissueSchema, Translator
from oncolytica_amd.spheroid import Spheroid, SpheroidTissue, TumorCell

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "oncolytica_amd", "csrc")

TYPE_A: ol.i32 = 1
GAIN: ol.f32 = 0.5


class BTissue(ol.Tissue):
    oxygen: ol.f32 = ol.diffusing(rate=0.1, init=1.0, boundary=ol.dirichlet(1.0))
    capacity: ol.i32 = 10

    def crowded(self, n: ol.i32) -> ol.bool:
        return n >= self.capacity


class BChem(ol.Chemistry):
    pad: ol.f32 = 0.0


class BCell(ol.Cell):
    pos: ol.vec3
    vel: ol.vec3
    kind: ol.i32 = 0
    energy: ol.f32 = 1.0
    flag: ol.bool = False
    tag: ol.u32 = 0

    def gamma(self, c: ol.f32) -> tuple[ol.f32, ol.f32]:
        return c / (c + 1.0), c * GAIN

    def spawn(self, d: ol.f32):
        daughter: BCell = BCell(pos=self.pos + ol.random_dir() * d, kind=self.kind, energy=self.energy * 0.5)
        self.divide(daughter)


class BMetrics(ol.Metrics):
    n: ol.u32 = 0


class BParams(ol.Params):
    speed: ol.f32 = 0.1
    steps: ol.i32 = 3
    on: ol.bool = True
    drift: ol.vec3 = ol.vec3(0.0, 0.0, 1.0)


class Broad(ol.Simulation[BTissue, BChem, BCell, BMetrics, BParams]):
    cutoff: ol.f32 = 0.8

    def score(self, a: ol.f32, b: ol.f32) -> ol.f32:
        return ol.math.max(a, b) - ol.math.min(a, b)

    @ol.cell_rule
    def move(self, cell: BCell):
        t: BTissue = self.tissue_at(cell.pos)
        n: ol.i32 = 0
        near: ol.f32 = 0.0
        for nb in cell.neighbors:
            if nb.kind == TYPE_A:
                n += 1
            d = ol.math.distance(cell.pos, nb.pos)
            if d < self.cutoff:
                near += 1.0 / (d + 1.0)
            if n > 20:
                break
        p, m = cell.gamma(t.oxygen)
        if not t.crowded(n) and self.params.on:
            cell.vel = ol.math.normalize(cell.vel + self.params.drift) * self.params.speed
            cell.pos += cell.vel * m
        k = 0
        while k < self.params.steps:
            cell.energy -= 0.01 * float(k)
            k += 1
        for i in range(2, 10, 3):
            cell.tag += ol.u32(i)
        cell.energy = ol.math.clamp(cell.energy - p * 0.1 + self.score(near, 1.0), 0.0, 10.0)
        cell.flag = cell.energy > 5.0 or n == 0
        cell.kind = 7 // 2 + (-7) % 3
        if cell.energy < 0.05:
            cell.die()
            return
        if ol.random() < 0.01 and cell.energy > 0.5:
            cell.spawn(0.5)


def _compile(src: str) -> None:
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "rule.hip")
        open(p, "w").write(src)
        r = subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O1", "-ffp-contract=off",
                            "-include", "hip/hip_runtime.h", "-I", CSRC, "-c", p, "-o", os.path.join(d, "r.o")],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:] + "\n" + src


def translate(sim, cell_cls, tissue_cls, rules):
    return Translator(sim, CellSchema.of(cell_cls), TissueSchema.of(tissue_cls), sim._params).translate(rules)


def test_spheroid_translates_and_compiles():
    sim = Spheroid()
    src = translate(sim, TumorCell, SpheroidTissue, [sim.lifecycle])
    assert "onc_rule_main" in src and "onc_v3_rand_dir" in src and "ctx.div = true" in src
    _compile(src)


def test_broad_model_translates_and_compiles():
    sim = Broad()
    src = translate(sim, BCell, BTissue, [sim.move])
    for frag in ("Tup_f32_f32", "h_cell_BCell_gamma", "h_tissue_BTissue_crowded", "h_sim_score",
                 "goto _nbend", "onc_floordiv_i", "onc_mod_i", "ctx.S->params["):
        assert frag in src, frag
    _compile(src)


def test_model_declared_inside_a_function():
    """test_oncolytica.py declares its models inside test functions; the rule
    then closes over its cell class."""
    class DivCell(ol.Cell):
        pos: ol.vec3
        generation: ol.i32 = 0

    class DivideSim(ol.Simulation[BTissue, BChem, DivCell, BMetrics, BParams]):
        @ol.cell_rule
        def rule(self, cell: DivCell):
            if cell.generation == 0:
                new_cell = DivCell()
                new_cell.pos = cell.pos + ol.vec3(1.0, 0.0, 0.0)
                new_cell.generation = 1
                cell.divide(new_cell)
                cell.generation = 1

    sim = DivideSim()
    _compile(translate(sim, DivCell, None, [sim.rule]))


def test_schema_columns():
    s = CellSchema.of(BCell)
    assert [c.name for c in s.columns] == ["pos.x", "pos.y", "pos.z", "_id", "vel.x", "vel.y", "vel.z", "kind",
                                           "energy", "flag", "tag"]
    assert [c.coltype for c in s.columns] == [0, 0, 0, 2, 0, 0, 0, 1, 0, 2, 2]


def _bad(body_rule):
    class S(ol.Simulation[BTissue, BChem, BCell, BMetrics, BParams]):
        rule = ol.cell_rule(body_rule)

    sim = S()
    return lambda: translate(sim, BCell, BTissue, [sim.rule])


def test_rejects_writes_to_tissue_and_neighbours():
    def w_tissue(self, cell: BCell):
        t: BTissue = self.tissue_at(cell.pos)
        t.capacity = 3

    def w_nb(self, cell: BCell):
        for nb in cell.neighbors:
            nb.energy = 0.0

    def w_unknown(self, cell: BCell):
        cell.energy = missing_name + 1.0  # noqa: F821

    def w_subscript(self, cell: BCell):
        x = [1, 2]
        cell.kind = x[0]

    for f in (w_tissue, w_nb, w_unknown, w_subscript):
        with pytest.raises(ol.CompilationError):
            _bad(f)()


def test_chemistry_at_without_a_chemistry_grid_rejected():
    def uses_chem(self, cell: BCell):
        c = self.chemistry_at(cell.pos)

    with pytest.raises(ol.CompilationError, match="setup_chemistry"):
        _bad(uses_chem)()


def test_grid_rules_translate_and_compile():
    """Tissue and chemistry rules (voxel loops, helper methods on voxels,
    tuples, random draws) and chemistry_at in a cell rule: the module holds
    the two pass kernels and their pass counts, and compiles for gfx950."""
    from oncolytica_amd.hip._translator import ChemSchema
    from tests.test_gpu_grid_rules import GCell, GChem, GMetrics, GTissue, Glioma
    sim = Glioma()
    tr = Translator(sim, CellSchema.of(GCell), TissueSchema.of(GTissue), sim._params, chem=ChemSchema.of(GChem))
    src = tr.translate([sim.behave], [sim.count], GMetrics, tissue_rules=[sim.count_cells],
                       chem_rules=[sim.update_pdgf])
    for frag in ("onc_tissue_pass", "onc_chem_pass", "onc_grid_tissue_passes = 1u", "onc_grid_chem_passes = 2u",
                 "Chem_GChem_at(ctx", "h_chem_GChem_step_pdgf", "ctx.rng.tag = 0x20000u + a.pass"):
        assert frag in src, frag
    _compile(src)


def test_grid_rule_misuse_rejected():
    from oncolytica_amd.hip._translator import ChemSchema
    from tests.test_gpu_grid_rules import GCell, GChem, GTissue, Glioma

    def writes_neighbour(self, t: GTissue):
        for nb in t.neighbors:
            nb.crowd = 1

    def cell_writes_tissue(self, cell: GCell):
        t: GTissue = self.tissue_at(cell.pos)
        t.crowd = 2

    sim = Glioma()
    for fn, kind in ((writes_neighbour, "tissue"), (cell_writes_tissue, "cell")):
        tr = Translator(sim, CellSchema.of(GCell), TissueSchema.of(GTissue), sim._params, chem=ChemSchema.of(GChem))
        with pytest.raises(ol.CompilationError, match="read-only|may not write"):
            if kind == "tissue":
                tr.translate([], tissue_rules=[fn])
            else:
                tr.translate([fn])


class MMetrics(ol.Metrics):
    alive: ol.u32 = 0
    energy_sum: ol.f32 = 0.0
    net: ol.i32 = 0


class MSim(ol.Simulation[BTissue, BChem, BCell, MMetrics, BParams]):
    @ol.cell_rule
    def move(self, cell: BCell):
        cell.energy -= 0.01

    @ol.metric_rule
    def count(self, cell: BCell, m: MMetrics):
        m.alive += 1
        m.energy_sum += cell.energy
        if cell.kind > 2:
            m.net -= 1
        else:
            m.net += cell.kind

    @ol.metric_rule
    def rich(self, cell: BCell, m: MMetrics):
        t: BTissue = self.tissue_at(cell.pos)
        if t.oxygen > self.params.speed:
            m.net += 2


def translate_metrics(sim, metric_rules):
    tr = Translator(sim, CellSchema.of(BCell), TissueSchema.of(BTissue), sim._params)
    src = tr.translate([sim.move], metric_rules, MMetrics)
    return tr, src


def test_metric_rules_translate_and_compile():
    """@ol.metric_rule methods become onc_metric_main (per-thread sums in cell
    order, a fixed tree per block) and onc_metric_final; the word order is the
    Metrics class's field order."""
    sim = MSim()
    tr, src = translate_metrics(sim, [sim.count, sim.rich])
    assert tr.metric_fields == [("alive", "u32"), ("energy_sum", "f32"), ("net", "i32")]
    for frag in ("onc_metric_main", "onc_metric_final", "metric_count(cell, m, ctx);", "metric_rich(cell, m, ctx);",
                 "(int)((unsigned)m_.net - (unsigned)(1))", "m_.energy_sum = m_.energy_sum + "):
        assert frag in src, frag
    _compile(src)


def test_metric_rule_restrictions():
    def reads(self, cell: BCell, m: MMetrics):
        if m.alive > 3:
            m.alive += 1

    def assigns(self, cell: BCell, m: MMetrics):
        m.alive = 1

    def multiplies(self, cell: BCell, m: MMetrics):
        m.energy_sum *= 2.0

    def draws(self, cell: BCell, m: MMetrics):
        m.energy_sum += ol.random()

    def walks(self, cell: BCell, m: MMetrics):
        for nb in cell.neighbors:
            m.alive += 1

    def writes_cell(self, cell: BCell, m: MMetrics):
        cell.energy = 0.0
        m.alive += 1

    def kills(self, cell: BCell, m: MMetrics):
        cell.die()

    for f in (reads, assigns, multiplies, draws, walks, writes_cell, kills):
        class S(MSim):
            extra = ol.metric_rule(f)

        sim = S()
        with pytest.raises(ol.CompilationError):
            translate_metrics(sim, [sim.extra])


def test_metric_fields_must_be_scalars():
    class VMetrics(ol.Metrics):
        centre: ol.vec3

    class VSim(ol.Simulation[BTissue, BChem, BCell, VMetrics, BParams]):
        @ol.metric_rule
        def c(self, cell: BCell, m: VMetrics):
            pass

    sim = VSim()
    tr = Translator(sim, CellSchema.of(BCell), TissueSchema.of(BTissue), sim._params)
    with pytest.raises(ol.CompilationError, match="metric field"):
        tr.translate([], [sim.c], VMetrics)


def test_vector_scalar_mix_rejected_with_the_types():
    """framework/tests/test_oncolytica.py:172-194: `cell.pos += 1.0` must not
    load; the message names the mismatched types."""
    def bad(self, cell: BCell):
        cell.pos += 1.0

    with pytest.raises(ol.CompilationError, match="types vec3 and f32"):
        _bad(bad)()
