"""Bundles (oncolytica_amd/bundle.py): a sim exported mid-run and restored
continues bit for bit like the original, with the built-in rule and with a
translated rule module, including states with daughters not yet merged."""
import numpy as np
import pytest

import oncolytica_amd as ol
from oncolytica_amd import bundle as B
from oncolytica_amd import workloads as W
from oncolytica_amd.hip._sim import Sim
from oncolytica_amd.spheroid import Spheroid, SpheroidTissue

pytestmark = pytest.mark.gpu


def _same(a: Sim, b: Sim, tag):
    ca, cb = a.download_cells(), b.download_cells()
    assert len(ca[0]) == len(cb[0]), tag
    for k, (x, y) in enumerate(zip(ca, cb)):
        np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32), err_msg=f"{tag} column {k}")
    for f in range(a.desc.n_fields):
        np.testing.assert_array_equal(a.download_field(f).view(np.uint32), b.download_field(f).view(np.uint32),
                                      err_msg=f"{tag} field {f}")
    assert a.counts() == b.counts(), tag


@pytest.mark.parametrize("at", [30, 60])
def test_builtin_rule_round_trip(ctx, tmp_path, at):
    w = W.plain_c1()
    d = w.desc()
    sim = Sim(ctx, d)
    sim.upload_cells(W.initial_cells(w))
    sim.step(at)
    sim.sync()
    path = B.export_bundle(sim, str(tmp_path / "b"), W.TUMOR_COL_NAMES)
    r, mod, cfg = B.load_bundle(ctx, path)
    try:
        assert mod is None and cfg["state"]["step"] == at and cfg["state"]["n"] == sim.counts()[0]
        _same(sim, r, f"restored at {at}")
        for k in range(3):
            sim.step(10)
            r.step(10)
            sim.sync()
            r.sync()
            _same(sim, r, f"{at} + {10 * (k + 1)} steps")
        assert sim.counts()[1] > 4096  # divisions happened
    finally:
        r.close()
        sim.close()


def test_translated_rule_round_trip(ctx, tmp_path):
    w = W.plain_c1()
    cols = W.initial_cells(w)
    eng = ol.Engine(backend="hip", max_agents=1 << 16, seed=w.seed)
    eng.setup_geometry((64, 64, 64), 1.0, cell_diameter=1.0)
    eng.setup_mechanics(k_rep=1.0, k_adh=0.0, mobility=0.25, max_disp=0.25, interaction_radius=1.0,
                        cell_radius=0.5)
    eng.setup_tissue(SpheroidTissue)
    eng.cells.set_columns({"pos": np.stack(cols[:3], 1), "_id": cols[3], "age": cols[4], "generation": cols[5]})
    eng.load_model(Spheroid())
    eng.run_steps(35)
    be = eng._backend_impl
    path = be.export_bundle(str(tmp_path / "t"))
    cfg = be.get_config()
    assert cfg["module"]["kernel"] == "onc_rule_main" and cfg["columns"][:3] == ["pos.x", "pos.y", "pos.z"]
    r, mod, _ = B.load_bundle(ctx, path)
    try:
        assert mod is not None
        _same(be.sim, r, "restored")
        be.sim.step(25)
        r.step(25)
        _same(be.sim, r, "35 + 25 steps")
        # the restored module's metric kernels answer like the original's
        nv = len(be.metric_fields)
        np.testing.assert_array_equal(be.sim.metrics(be.module, nv), r.metrics(mod, nv))
    finally:
        r.close()
        mod.close()


def test_empty_sim_metrics_and_bundle(ctx, tmp_path):
    """A sim with no cells: every metric is zero, and its bundle restores to
    an empty sim that keeps its step counter and fields."""
    from oncolytica_amd.hip._sim import Module
    from oncolytica_amd.hip._translator import CellSchema, TissueSchema, Translator
    from oncolytica_amd.spheroid import SpheroidMetrics, TumorCell
    w = W.plain_c1()
    sim = Sim(ctx, w.desc())
    model = Spheroid()
    tr = Translator(model, CellSchema.of(TumorCell), TissueSchema.of(SpheroidTissue), model._params)
    mod = Module(ctx, tr.translate([model.lifecycle], [model.count], SpheroidMetrics))
    try:
        sim.upload_cells([c[:0] for c in W.initial_cells(w)])
        sim.step(3)
        sim.sync()
        assert sim.counts()[0] == 0
        np.testing.assert_array_equal(sim.metrics(mod, len(tr.metric_fields)), np.zeros(4, np.uint32))
        r, _, cfg = B.load_bundle(ctx, B.export_bundle(sim, str(tmp_path / "e"), W.TUMOR_COL_NAMES))
        try:
            assert cfg["state"] == {"n": 0, "next_id": sim.counts()[1], "step": 3}
            _same(sim, r, "empty")
        finally:
            r.close()
    finally:
        mod.close()
        sim.close()


def test_bundle_after_update_params_continues_bit_for_bit(ctx, tmp_path):
    """Parameters changed with Engine.update_params reach the bundle: the
    restored sim divides with the new division age, like the original."""
    w = W.plain_c1()
    cols = W.initial_cells(w)
    eng = ol.Engine(backend="hip", max_agents=1 << 16, seed=w.seed)
    eng.setup_geometry((64, 64, 64), 1.0, cell_diameter=1.0)
    eng.setup_mechanics(k_rep=1.0, k_adh=0.0, mobility=0.25, max_disp=0.25, interaction_radius=1.0,
                        cell_radius=0.5)
    eng.setup_tissue(SpheroidTissue)
    eng.cells.set_columns({"pos": np.stack(cols[:3], 1), "_id": cols[3], "age": cols[4], "generation": cols[5]})
    eng.load_model(Spheroid())
    eng.run_steps(10)
    eng.update_params(div_age=12.0)
    eng.run_steps(10)
    be = eng._backend_impl
    path = be.export_bundle(str(tmp_path / "p"))
    r, mod, cfg = B.load_bundle(ctx, path)
    try:
        assert 12.0 in cfg["desc"]["params"] and 24.0 not in cfg["desc"]["params"]
        _same(be.sim, r, "restored")
        be.sim.step(20)
        r.step(20)
        _same(be.sim, r, "20 + 20 steps")
    finally:
        r.close()
        mod.close()


def test_metrics_reject_a_word_count_other_than_the_modules(ctx):
    """onc_sim_metrics takes exactly the module's metric count: fewer words
    would let its kernels write past the scratch it sizes, more would return
    words nobody wrote."""
    from oncolytica_amd.hip import _abi
    from oncolytica_amd.hip._sim import Module
    from oncolytica_amd.hip._translator import CellSchema, TissueSchema, Translator
    from oncolytica_amd.spheroid import SpheroidMetrics, TumorCell
    w = W.plain_c1()
    sim = Sim(ctx, w.desc())
    model = Spheroid()
    tr = Translator(model, CellSchema.of(TumorCell), TissueSchema.of(SpheroidTissue), model._params)
    mod = Module(ctx, tr.translate([model.lifecycle], [model.count], SpheroidMetrics))
    try:
        sim.upload_cells(W.initial_cells(w))
        nv = len(tr.metric_fields)
        assert nv == 4
        for bad in (nv - 1, nv + 1):
            with pytest.raises(_abi.OncError, match="metric count"):
                sim.metrics(mod, bad)
        got = sim.metrics(mod, nv)
        assert int(got[0]) == 4096
    finally:
        mod.close()
        sim.close()
