"""Whole-timestep parity of the HIP path against the oracle on the same seeded
inputs: every cell column and every field value bit-identical (binary32),
plus determinism, batch == alone, and graph replay == eager launches."""
import numpy as np
import pytest

from oracle import oracle as O
from oncolytica_amd import workloads as W
from oncolytica_amd.hip._sim import Batch, Sim

pytestmark = pytest.mark.gpu

NAMES = W.TUMOR_COL_NAMES


def compare(step, gcols, gfields, st):
    ocols = st.columns()
    assert len(gcols[0]) == len(ocols[0]), f"step {step}: live cells {len(gcols[0])} vs oracle {len(ocols[0])}"
    for name, g, o in zip(NAMES, gcols, ocols):
        gv, ov = g.view(np.uint32), o.view(np.uint32)
        bad = np.nonzero(gv != ov)[0]
        assert bad.size == 0, f"step {step}: column {name} differs at {bad[:5]} ({g[bad[:5]]} vs {o[bad[:5]]})"
    for f, (g, o) in enumerate(zip(gfields, st.fields)):
        bad = np.nonzero(g.view(np.uint32) != o.view(np.uint32))[0]
        assert bad.size == 0, f"step {step}: field {f} differs at {bad[:5]} ({g[bad[:5]]} vs {o[bad[:5]]})"


def run_pair(ctx, w, n_cells, steps, check_every=1, capacity=None):
    d = w.desc()
    if capacity:
        d.capacity = capacity
    cols = W.initial_cells(w, n_cells)
    sim = Sim(ctx, d)
    sim.upload_cells(cols)
    st = O.TumorState(d, cols)
    done = 0
    try:
        while done < steps:
            k = min(check_every, steps - done)
            sim.step(k)
            sim.sync()
            st.run(k)
            done += k
            print(f"{w.name}: step {done} n={st.n.value}", flush=True)
            compare(done, sim.download_cells(), [sim.download_field(f) for f in range(d.n_fields)], st)
    finally:
        sim.close()
    return st


def test_plain_4k_64_first_steps_bit_exact(ctx):
    run_pair(ctx, W.plain_c1(), None, 5, check_every=1)


def test_plain_4k_64_100_steps_bit_exact(ctx):
    st = run_pair(ctx, W.plain_c1(), None, 100, check_every=25)
    assert st.next_id.value > 4096  # divisions happened
    assert st.n.value > 0


def test_large_1m_256_bit_exact(ctx):
    w = W.large_c2()
    run_pair(ctx, w, None, 3, check_every=1)


def test_ensemble_member_bit_exact(ctx):
    w = W.ensemble_c3(n_sims=4, scale=0.2)[1]
    run_pair(ctx, w, None, 30, check_every=10)


def test_empty_and_single_cell(ctx):
    w = W.plain_c1()
    run_pair(ctx, w, 1, 40, check_every=20)
    d = w.desc()
    sim = Sim(ctx, d)
    try:
        sim.upload_cells(W.initial_cells(w, 0))
        sim.step(3)
        sim.sync()
        assert sim.counts() == (0, 0, 3)
        st = O.TumorState(d, W.initial_cells(w, 0))
        st.run(3)
        np.testing.assert_array_equal(sim.download_field(0).view(np.uint32), st.fields[0].view(np.uint32))
    finally:
        sim.close()


def test_capacity_clamp_matches_oracle(ctx):
    w = W.plain_c1()
    run_pair(ctx, w, 2000, 40, check_every=10, capacity=2600)


def test_deterministic_repeat_and_graph(ctx):
    """Eager launches, graph replays and a second graph run give the same
    bits over 110 steps. A batch replays its one-step graph 64 times before
    it captures a 16-step graph, so the replays of 70 + 40 steps cover both
    graphs, two replays of the 16-step one, and one-step replays after it."""
    w = W.plain_c1()
    outs = []
    for graph in (False, True, True):
        sim = Sim(ctx, w.desc())
        sim.upload_cells(W.initial_cells(w))
        b = Batch(ctx)
        b.add(sim)
        b.step(70, graph=graph)
        b.step(40, graph=graph)
        b.sync()
        outs.append((sim.download_cells(), sim.download_field(0)))
        b.close()
        sim.close()
    for cols, f in outs[1:]:
        for a, b_ in zip(outs[0][0], cols):
            np.testing.assert_array_equal(a.view(np.uint32), b_.view(np.uint32))
        np.testing.assert_array_equal(outs[0][1].view(np.uint32), f.view(np.uint32))


def test_tile_sizes_follow_the_population_bit_exact(ctx):
    """A batch switches to 16 / 8 cells per thread once its live cells reach
    big_from and back below small_below, from live counts read back every 16
    steps. The 4,096-cell workload grows past 9,000 cells by step 30, falls
    to ~8,800 by step 40 and grows again: with thresholds 9,000 / 8,900 it
    switches both ways, through eager steps and graph replays, and stays
    bit-identical to the oracle (tile sizes never change results)."""
    w = W.plain_c1()
    d = w.desc()
    cols = W.initial_cells(w)
    st = O.TumorState(d, cols)
    sim = Sim(ctx, d)
    sim.upload_cells(cols)
    b = Batch(ctx)
    b.add(sim)
    b.set_tile_policy(9000, 8900)
    try:
        assert b.tile_items() == (8, 4)
        seen = set()
        done = 0
        for k, graph in [(20, False), (16, True), (8, False), (16, True), (20, False), (20, True)]:
            b.step(k, graph=graph)
            b.sync()
            done += k
            seen.add(b.tile_items())
        st.run(done)
        print(f"tile policy: {done} steps, sizes seen {sorted(seen)}", flush=True)
        compare(done, sim.download_cells(), [sim.download_field(f) for f in range(d.n_fields)], st)
        assert (16, 8) in seen and (8, 4) in seen
        b.set_tile_policy(0, 0)
        assert b.tile_items() == (8, 4) or b.tile_items() == (16, 8)
    finally:
        b.close()
        sim.close()


def test_batch_equals_alone(ctx):
    """Five ensemble members stepped as one batch (graph replays of 5 and 7
    steps) equal each member stepped alone, and each equals the oracle run
    for the same 12 steps."""
    ws = W.ensemble_c3(n_sims=5, scale=0.05)
    alone = []
    for w in ws:
        sim = Sim(ctx, w.desc())
        sim.upload_cells(W.initial_cells(w))
        sim.step(12)
        sim.sync()
        alone.append((sim.download_cells(), sim.download_field(0)))
        sim.close()
    sims = []
    b = Batch(ctx)
    for w in ws:
        s = Sim(ctx, w.desc())
        s.upload_cells(W.initial_cells(w))
        b.add(s)
        sims.append(s)
    b.step(5, graph=True)
    b.step(7, graph=True)
    b.sync()
    for k, (w, s, (cols, f)) in enumerate(zip(ws, sims, alone)):
        got = s.download_cells()
        for a, g in zip(cols, got):
            np.testing.assert_array_equal(a.view(np.uint32), g.view(np.uint32))
        np.testing.assert_array_equal(f.view(np.uint32), s.download_field(0).view(np.uint32))
        st = O.TumorState(w.desc(), W.initial_cells(w))
        st.run(12)
        print(f"batch member {k}: {len(got[0])} cells vs oracle", flush=True)
        compare(12, got, [s.download_field(q) for q in range(w.desc().n_fields)], st)
    b.close()
    for s in sims:
        s.close()


def test_voxel_over_force_walk_limit_is_reported(ctx):
    """More than 65535 cells in one voxel (here 66,000 at one point) exceed
    the force walk's bin length: the step's sync reports a device error
    instead of returning forces computed from a clipped bin."""
    from oncolytica_amd.hip import _abi
    w = W.plain_c1()
    d = w.desc()
    d.capacity = 140_000
    n = 66_000
    cols = W.initial_cells(w, n)
    for c in range(3):
        cols[c][:] = np.float32(32.5)
    sim = Sim(ctx, d)
    try:
        sim.upload_cells(cols)
        sim.step(1)
        with pytest.raises(_abi.OncError, match="65535"):
            sim.sync()
    finally:
        sim.close()


def test_set_step_backwards_then_new_state_matches_oracle(ctx):
    """Run 30 steps, set the step counter back to 10 and upload the step-10
    state without its cells at x < 32: the next 20 steps equal the oracle's
    from that state. Bin entries written by the first run's steps 10..29
    carry the tags the next 20 steps use, and those of the removed half lie
    in the neighbourhoods of the cells left at x ~ 32: they must read as empty."""
    w = W.plain_c1()
    d = w.desc()
    cols = W.initial_cells(w)
    ref = O.TumorState(d, cols)
    ref.run(10)
    keep = ref.columns()[0] >= np.float32(32.0)
    half = O.TumorState(d, [c[keep] for c in ref.columns()], fields=ref.fields, next_id=ref.next_id.value)
    half.step.value = 10
    assert 0 < keep.sum() < len(keep)
    sim = Sim(ctx, d)
    try:
        sim.upload_cells(cols)
        sim.step(30)
        sim.sync()
        print("set_step: 30 steps run, rewinding to 10", flush=True)
        sim.set_step(10)
        sim.upload_cells(half.columns(), next_id=half.next_id.value)
        for f in range(d.n_fields):
            sim.upload_field(f, half.fields[f])
        assert sim.counts()[2] == 10
        sim.step(20)
        sim.sync()
        half.run(20)
        compare(30, sim.download_cells(), [sim.download_field(f) for f in range(d.n_fields)], half)
    finally:
        sim.close()


def test_steps_across_the_bin_tag_wrap_match_oracle(ctx):
    """Bin entries carry a 16-bit tag, step mod 65534 + 1. The step-10 state
    of the 4,096-cell workload, restored at step 65,520, runs 30 steps across
    the wrap (the step with 65,534 completed takes tag 1 again, after tag
    65,534) bit for bit like the oracle from the same state (the step count
    keys the Philox draws, so the oracle runs those same steps)."""
    w = W.plain_c1()
    d = w.desc()
    ref = O.TumorState(d, W.initial_cells(w))
    ref.run(10)
    st = O.TumorState(d, ref.columns(), fields=ref.fields, next_id=ref.next_id.value)
    st.step.value = 65_520
    sim = Sim(ctx, d)
    try:
        sim.upload_cells(st.columns(), next_id=st.next_id.value)
        for f in range(d.n_fields):
            sim.upload_field(f, st.fields[f])
        sim.set_step(65_520)
        for k in (10, 20):
            sim.step(k)
            sim.sync()
            st.run(k)
            print(f"tag wrap: step {st.step.value}", flush=True)
            compare(st.step.value, sim.download_cells(), [sim.download_field(f) for f in range(d.n_fields)], st)
        assert st.step.value == 65_550
    finally:
        sim.close()


@pytest.mark.parametrize("fused", [1, 0])
def test_quiet_sweep_tiles_match_the_oracle(ctx, fused):
    """The sweep skips tiles whose input neighbourhood holds one value that
    the update leaves unchanged. Four fields exercise it on a 64^3 grid: O2
    (Dirichlet 1.0, consumed by the cells), a uniform zero-flux field (never
    changes, so its tiles are kept), a field at 0.5 with Dirichlet faces at
    1.0 (the front moves inward tile by tile), and a decaying zero-flux field
    (no fixed point: always swept). At step 20 field 2 is replaced through the
    host (a uniform 0.75 with a block at 0.2), and the step counter is set to
    its own value (both must drop what the sweep knew). Every column and field
    is bit-identical to the oracle at steps 20 and 40, with graph replays and
    eager steps, in one launch with the forces and as its own launch."""
    from oncolytica_amd.hip._abi import BC_DIRICHLET, BC_NONE
    w = W.plain_c1()
    w.fields = [W.FieldSpec(rate=0.1, init=1.0, bc=BC_DIRICHLET, bc_value=1.0, uptake=0.01),
                W.FieldSpec(rate=0.05, init=0.25, bc=BC_NONE),
                W.FieldSpec(rate=0.1, init=0.5, bc=BC_DIRICHLET, bc_value=1.0),
                W.FieldSpec(rate=0.1, decay=0.01, init=0.5, bc=BC_NONE)]
    d = w.desc()
    cols = W.initial_cells(w)
    st = O.TumorState(d, cols)
    sim = Sim(ctx, d)
    b = Batch(ctx)
    try:
        b.set_fused(fused)
        sim.upload_cells(cols)
        b.add(sim)
        b.step(10, graph=True)
        b.step(10, graph=False)
        b.sync()
        st.run(20)
        print(f"quiet tiles fused={fused}: step 20", flush=True)
        compare(20, sim.download_cells(), [sim.download_field(f) for f in range(d.n_fields)], st)
        new = np.full((64, 64, 64), 0.75, dtype=np.float32)
        new[20:40, 8:30, 30:50] = np.float32(0.2)
        new = new.reshape(-1)
        sim.upload_field(2, new)
        st.fields[2][:] = new
        sim.set_step(20)
        b.step(10, graph=True)
        b.step(10, graph=False)
        b.sync()
        st.run(20)
        print(f"quiet tiles fused={fused}: step 40", flush=True)
        compare(40, sim.download_cells(), [sim.download_field(f) for f in range(d.n_fields)], st)
    finally:
        b.close()
        sim.close()
