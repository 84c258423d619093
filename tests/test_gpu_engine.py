"""End-to-end tests through the rule API and Engine(backend="hip"): the
reference's own runtime tests (framework/tests/test_gpu.py,
framework/tests/test_oncolytica.py) restated for this backend, plus
bit-exactness of a translated model against the built-in rule and the oracle."""
import math
import os

import numpy as np
import pytest

import oncolytica_amd as ol
from oncolytica_amd import workloads as W
from oncolytica_amd.spheroid import Spheroid, SpheroidTissue, TumorCell
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---- shared model classes (test_oncolytica.py:16-33) --------------------------------
class MyCell(ol.Cell):
    pos: ol.vec3 = ol.vec3(50.0, 50.0, 50.0)
    health: ol.f32
    mutated: ol.bool


class MyTissue(ol.Tissue):
    oxygen: ol.f32 = 1.0


class MyChem(ol.Chemistry):
    drug_conc: ol.f32 = 0.0


class MyMetrics(ol.Metrics):
    alive_cells: ol.u32 = 0


class MyParams(ol.Params):
    pass


class MockCell(ol.Cell):
    pos: ol.vec3
    id: int


class MockTissue(ol.Tissue):
    pad: int = 0


class MockChemistry(ol.Chemistry):
    pad: int = 0


class MockMetrics(ol.Metrics):
    pad: int = 0


class FakeSim(ol.Simulation[MockTissue, MockChemistry, MockCell, MockMetrics, MyParams]):
    pass


def test_sort_highlevel_reference_golden():
    """framework/tests/test_gpu.py:17-61: 10,000 cells, 50^3 grid, voxel 2.0,
    random.seed(42) positions; one step sorts cells into Morton buckets. Here
    the order inside a bucket is also checked (stable: ascending id)."""
    g = np.load(os.path.join(GOLDEN, "sort_golden.npz"))
    n = len(g["pos"])
    eng = ol.Engine(backend="hip", max_agents=n)
    eng.setup_geometry((50, 50, 50), 2.0)
    for i in range(n):
        c = MockCell()
        c.pos, c.id = ol.vec3(*g["pos"][i]), i
        eng.cells.add(c)
    eng.load_model(FakeSim())
    eng.run_step()
    eng.sync_to_host()
    ids = np.array([c.id for c in eng.cells._data])
    assert sorted(ids.tolist()) == list(range(n))
    order = np.argsort(g["keys"], kind="stable")
    np.testing.assert_array_equal(ids, order)
    keys = O.keys(*(np.array([getattr(c.pos, a) for c in eng.cells._data], dtype=np.float32) for a in "xyz"),
                  (50, 50, 50), 2.0)
    np.testing.assert_array_equal(keys, g["keys"][order])


def test_die_and_metrics():
    """test_oncolytica.py:300-364: a cell with type_id 1 dies; the metric
    counts one live cell on this step and the next."""
    class TestCell(ol.Cell):
        pos: ol.vec3
        type_id: ol.i32 = 0

    class TestMetrics(ol.Metrics):
        alive_count: ol.u32 = 0

    class DieSim(ol.Simulation[MyTissue, MyChem, TestCell, TestMetrics, MyParams]):
        @ol.cell_rule
        def rule(self, cell: TestCell):
            if cell.type_id == 1:
                cell.die()

        @ol.metric_rule
        def calc_metrics(self, cell: TestCell, m: TestMetrics):
            m.alive_count += 1

    eng = ol.Engine(backend="hip", max_agents=4)
    eng.setup_geometry((10, 10, 10), 2.0)
    eng.cells.add(TestCell(pos=ol.vec3(5.0, 5.0, 5.0), type_id=2))
    eng.cells.add(TestCell(pos=ol.vec3(1.0, 1.0, 1.0), type_id=1))
    eng.load_model(DieSim())
    eng.run_step(collect_metrics=True)
    assert eng.get_metrics().alive_count == 1
    eng.sync_to_host()  # metrics are reduced on the device; the cells stay there until synced
    assert [c.type_id for c in eng.cells._data] == [2]
    assert math.isclose(eng.cells._data[0].pos.x, 5.0)
    eng.run_step(collect_metrics=True)
    assert eng.get_metrics().alive_count == 1


def test_division():
    """test_oncolytica.py:367-404: one cell divides once; both carry generation 1."""
    class DivCell(ol.Cell):
        pos: ol.vec3
        generation: ol.i32 = 0

    class DivideSim(ol.Simulation[MyTissue, MyChem, DivCell, MyMetrics, MyParams]):
        @ol.cell_rule
        def rule(self, cell: DivCell):
            if cell.generation == 0:
                new_cell = DivCell()
                new_cell.pos = cell.pos + ol.vec3(1.0, 0.0, 0.0)
                new_cell.generation = 1
                cell.divide(new_cell)
                cell.generation = 1

    eng = ol.Engine(backend="hip", max_agents=10)
    eng.setup_geometry((10, 10, 10))
    eng.cells.add(DivCell(pos=ol.vec3(5.0, 5.0, 5.0)))
    eng.load_model(DivideSim())
    eng.run_step()
    eng.sync_to_host()
    assert len(eng.cells._data) == 2
    assert [c.generation for c in eng.cells._data] == [1, 1]
    assert [round(c.pos.x, 5) for c in eng.cells._data] == [5.0, 6.0]
    eng.run_step()
    eng.sync_to_host()
    assert len(eng.cells._data) == 2


def test_sequential_rules_and_lifecycle():
    """test_oncolytica.py:461-482 (100 - 5 - 15 = 80) and :58-93 (health drops
    by 10 per step, cells die at <= 0 and are compacted away)."""
    class MultiRuleSim(ol.Simulation[MyTissue, MyChem, MyCell, MyMetrics, MyParams]):
        @ol.cell_rule
        def rule_one(self, cell: MyCell):
            cell.health -= 5.0

        @ol.cell_rule
        def rule_two(self, cell: MyCell):
            cell.health -= 15.0

    eng = ol.Engine(backend="hip")
    eng.cells.add(MyCell(health=100.0))
    eng.load_model(MultiRuleSim())
    eng.run_step()
    eng.sync_to_host()
    assert math.isclose(eng.cells._data[0].health, 80.0)

    class BasicSim(ol.Simulation[MyTissue, MyChem, MyCell, MyMetrics, MyParams]):
        @ol.cell_rule
        def cell_logic(self, cell: MyCell):
            cell.health -= 10.0
            if cell.health <= 0.0:
                cell.die()

        @ol.metric_rule
        def track(self, cell: MyCell, metrics: MyMetrics):
            metrics.alive_cells += 1

    eng = ol.Engine(backend="hip")
    for i in range(5):
        c = MyCell(health=100.0)
        c.pos = ol.vec3(50.0, 50.0, 50.0) + 5.0 * ol.vec3(float(i))
        eng.cells.add(c)
    eng.load_model(BasicSim())
    for _ in range(9):
        eng.run_step()
    eng.run_step(collect_metrics=True)
    assert eng.get_metrics().alive_cells == 0
    eng.sync_to_host()
    assert len(eng.cells._data) == 0


def test_randomness_and_locals():
    """test_oncolytica.py:437-458 (random / random_dir compile and run) and
    :108-128 (a local first assigned inside a dead branch)."""
    class RandSim(ol.Simulation[MyTissue, MyChem, MyCell, MyMetrics, MyParams]):
        @ol.cell_rule
        def rule(self, cell: MyCell):
            if False:
                b = 5
            b = 4
            dist = ol.random() * 2.0
            direction = ol.random_dir()
            cell.pos += direction * dist
            cell.health = float(b)

    eng = ol.Engine(backend="hip", seed=7)
    eng.cells.add(MyCell(pos=ol.vec3(10.0, 10.0, 10.0)))
    eng.load_model(RandSim())
    eng.run_step()
    eng.sync_to_host()
    c = eng.cells._data[0]
    d = math.dist((c.pos.x, c.pos.y, c.pos.z), (10.0, 10.0, 10.0))
    assert 0.0 <= d <= 2.0 and c.health == 4.0


def test_neighbor_count_matches_voxel_membership():
    """for nb in cell.neighbors visits every other cell of the 27 Moore voxels."""
    class NCell(ol.Cell):
        pos: ol.vec3
        count: ol.i32 = 0

    class NSim(ol.Simulation[MyTissue, MyChem, NCell, MyMetrics, MyParams]):
        @ol.cell_rule
        def rule(self, cell: NCell):
            k: ol.i32 = 0
            for nb in cell.neighbors:
                k += 1
            cell.count = k

    rng = np.random.default_rng(3)
    n, dims = 3000, (12, 12, 12)
    p = rng.uniform(0.0, 12.0, (n, 3)).astype(np.float32)
    eng = ol.Engine(backend="hip", max_agents=n)
    eng.setup_geometry(dims, 1.0)
    eng.cells.set_columns({"pos": p, "count": np.zeros(n, np.int32)})
    eng.load_model(NSim())
    eng.run_step()
    eng.sync_to_host()
    out = eng.cells.columns()
    v = np.floor(out["pos"]).astype(int).clip(0, 11)
    diff = np.abs(v[:, None, :] - v[None, :, :]).max(-1)
    expect = (diff <= 1).sum(1) - 1
    np.testing.assert_array_equal(out["count"], expect)


def _c1_like_oracle_desc(cap):
    w = W.large_c2()
    w.dims = (64, 64, 64)
    w.capacity = cap
    w.key_format = 30
    w.cell_radius = 0.5
    w.k_adh = 0.0
    return w


def test_translated_spheroid_bit_exact_with_builtin_and_oracle():
    """The DSL form of the benchmark model (oncolytica_amd/spheroid.py),
    translated and built by hiprtc, reproduces the oracle bit for bit."""
    w = W.plain_c1()
    cols = W.initial_cells(w)
    cap = 1 << 16
    eng = ol.Engine(backend="hip", max_agents=cap, seed=w.seed)
    eng.setup_geometry((64, 64, 64), 1.0, cell_diameter=1.0)
    eng.setup_mechanics(k_rep=1.0, k_adh=0.0, mobility=0.25, max_disp=0.25, interaction_radius=1.0,
                        cell_radius=0.5)
    eng.setup_tissue(SpheroidTissue)
    eng.cells.set_columns({"pos": np.stack(cols[:3], 1), "_id": cols[3], "age": cols[4], "generation": cols[5]})
    eng.load_model(Spheroid())
    steps = 40
    eng.run_steps(steps)
    eng.sync_to_host()
    got = eng.cells.columns()

    ow = _c1_like_oracle_desc(cap)
    st = O.TumorState(ow.desc(), cols)
    st.run(steps)
    oc = st.columns()
    assert len(got["pos"]) == len(oc[0])
    for k, a in enumerate("xyz"):
        np.testing.assert_array_equal(got["pos"][:, k].view(np.uint32), oc[k].view(np.uint32))
    np.testing.assert_array_equal(got["_id"], oc[3])
    np.testing.assert_array_equal(got["age"].view(np.uint32), oc[4].view(np.uint32))
    np.testing.assert_array_equal(got["generation"], oc[5])
    np.testing.assert_array_equal(eng.tissue.flat("oxygen").view(np.uint32), st.fields[0].view(np.uint32))
    np.testing.assert_array_equal(eng.tissue.flat("drug").view(np.uint32), st.fields[1].view(np.uint32))


def test_update_params_reaches_device():
    class PCell(ol.Cell):
        pos: ol.vec3
        v: ol.f32 = 0.0

    class PParams(ol.Params):
        inc: ol.f32 = 1.0

    class PSim(ol.Simulation[MyTissue, MyChem, PCell, MyMetrics, PParams]):
        @ol.cell_rule
        def rule(self, cell: PCell):
            cell.v += self.params.inc

    eng = ol.Engine(backend="hip")
    eng.cells.add(PCell(pos=ol.vec3(1.0, 1.0, 1.0)))
    eng.load_model(PSim())
    eng.run_step()
    eng.update_params(inc=10.0)
    eng.run_step()
    eng.sync_to_host()
    assert eng.cells._data[0].v == 11.0


def test_metric_rules_reduce_on_device_like_the_oracle():
    """@ol.metric_rule on the device (onc_sim_metrics): u32 and i32 sums are
    exact (mod 2^32), the binary32 sum is bit-identical to the oracle's fixed
    summation order over the committed cells, and a metric that reads the
    swept tissue counts the same cells as numpy does."""
    class MCell(ol.Cell):
        pos: ol.vec3
        energy: ol.f32 = 1.0
        kind: ol.i32 = 0

    class MTissue(ol.Tissue):
        oxygen: ol.f32 = ol.diffusing(rate=0.1, init=1.0, boundary=ol.dirichlet(0.2))

    class MMetrics(ol.Metrics):
        alive: ol.u32 = 0
        energy_sum: ol.f32 = 0.0
        net: ol.i32 = 0
        rich: ol.u32 = 0

    class MSim(ol.Simulation[MTissue, MyChem, MCell, MMetrics, MyParams]):
        @ol.cell_rule
        def burn(self, cell: MCell):
            cell.energy = cell.energy * 0.75 + 0.001 * ol.f32(cell.kind)
            if cell.kind == 6 and cell.energy < 0.3:
                cell.die()

        @ol.metric_rule
        def count(self, cell: MCell, m: MMetrics):
            m.alive += 1
            m.energy_sum += cell.energy
            if cell.kind > 2:
                m.net -= 1
            else:
                m.net += cell.kind

        @ol.metric_rule
        def oxygenated(self, cell: MCell, m: MMetrics):
            t: MTissue = self.tissue_at(cell.pos)
            if t.oxygen > 0.6:
                m.rich += 1

    n = 20_000
    rng = np.random.default_rng(11)
    pos = rng.random((n, 3)) * 60.0 + 2.0
    energy = rng.random(n).astype(np.float32) * 3.0
    kind = rng.integers(0, 8, n)
    eng = ol.Engine(backend="hip", max_agents=n)
    eng.setup_geometry((64, 64, 64), 1.0)
    eng.setup_tissue(MTissue)
    for i in range(n):
        eng.cells.add(MCell(pos=ol.vec3(*pos[i]), energy=float(energy[i]), kind=int(kind[i])))
    eng.load_model(MSim())
    for step in range(4):
        eng.run_step(collect_metrics=True)
        m = eng.get_metrics()
        eng.sync_to_host()
        cells = eng.cells._data
        e = np.array([c.energy for c in cells], dtype=np.float32)
        k = np.array([c.kind for c in cells], dtype=np.int64)
        p = np.array([[c.pos.x, c.pos.y, c.pos.z] for c in cells], dtype=np.float32)
        ox = eng._backend_impl.sim.download_field(0)
        v = np.clip(np.floor(p).astype(np.int64), 0, 63)
        o = ox[v[:, 0] + 64 * (v[:, 1] + 64 * v[:, 2])]
        print(f"step {step}: {len(cells)} cells, metrics {m.alive} {m.energy_sum} {m.net} {m.rich}", flush=True)
        assert m.alive == len(cells)
        want = O.metric_reduce(e, "f32")
        got = int(np.array([m.energy_sum], dtype=np.float32).view(np.uint32)[0])
        assert got == want, (m.energy_sum, np.uint32(want).view(np.float32))
        assert m.net == int(np.where(k > 2, -1, k).sum())
        assert m.rich == int((o > np.float32(0.6)).sum())
    assert len(cells) < n  # some cells died


def test_spheroid_metrics_match_the_committed_cells():
    """The benchmark model's metric rule (population, summed age, summed O2
    at the cells, cells ready to divide) after 30 steps of the 4,096-cell
    workload: the integer metrics equal numpy's counts and the binary32 sums
    the oracle's fixed order, over the downloaded cells and field."""
    w = W.plain_c1()
    cols = W.initial_cells(w)
    eng = ol.Engine(backend="hip", max_agents=1 << 16, seed=w.seed)
    eng.setup_geometry((64, 64, 64), 1.0, cell_diameter=1.0)
    eng.setup_mechanics(k_rep=1.0, k_adh=0.0, mobility=0.25, max_disp=0.25, interaction_radius=1.0,
                        cell_radius=0.5)
    eng.setup_tissue(SpheroidTissue)
    eng.cells.set_columns({"pos": np.stack(cols[:3], 1), "_id": cols[3], "age": cols[4], "generation": cols[5]})
    eng.load_model(Spheroid())
    eng.run_steps(29)
    eng.run_step(collect_metrics=True)
    m = eng.get_metrics()
    eng.sync_to_host()
    got = eng.cells.columns()
    age = got["age"].astype(np.float32)
    p = got["pos"].astype(np.float32)
    v = np.clip(np.floor(p).astype(np.int64), 0, 63)
    o2 = eng.tissue.flat("oxygen")[v[:, 0] + 64 * (v[:, 1] + 64 * v[:, 2])]
    assert m.alive == len(age) > 4096
    assert np.float32(m.age_sum).view(np.uint32) == O.metric_reduce(age, "f32")
    assert np.float32(m.oxygen_sum).view(np.uint32) == O.metric_reduce(o2, "f32")
    assert m.dividing == int(((age + np.float32(1.0) > np.float32(24.0)) & (o2 > np.float32(0.3))).sum())


def test_backend_compile_errors_reach_load_model(monkeypatch):
    """A hiprtc failure (ONC_ERR_COMPILE) reaches Engine.load_model as an
    exception that carries the compiler log (framework/tests/
    test_oncolytica.py:172-219 expect the driver's shader errors there)."""
    from oncolytica_amd.hip import _abi
    from oncolytica_amd.hip._translator import Translator

    ctx = _abi.Context(0)
    try:
        with pytest.raises(_abi.OncError) as ei:
            from oncolytica_amd.hip._sim import Module
            Module(ctx, "extern \"C\" __global__ void k() { undeclared_thing = 1; }")
        assert ei.value.code == _abi.ERR_COMPILE and "undeclared_thing" in str(ei.value)
    finally:
        ctx.close()

    real = Translator.translate

    def broken(self, *a, **kw):
        return real(self, *a, **kw) + "\n__device__ int broken_decl = not_a_symbol;\n"

    monkeypatch.setattr(Translator, "translate", broken)

    class PCell(ol.Cell):
        pos: ol.vec3
        v: ol.f32 = 0.0

    class PSim(ol.Simulation[MyTissue, MyChem, PCell, MyMetrics, MyParams]):
        @ol.cell_rule
        def rule(self, cell: PCell):
            cell.v += 1.0

    eng = ol.Engine(backend="hip")
    eng.cells.add(PCell(pos=ol.vec3(1.0, 1.0, 1.0)))
    with pytest.raises(RuntimeError, match="not_a_symbol"):
        eng.load_model(PSim())
