"""Pins the rule executor (oracle/rule_exec.py), the oracle of translated
rules, on the CPU: the benchmark model written with the rule API, run by the
executor, reproduces the C oracle's built-in tumour step bit for bit; the
voxel loops visit what their definitions say (checked against numpy); the
reference's grid-rule expectations hold (framework/tests/test_oncolytica.py:
266-296, 407-434)."""
import numpy as np
import pytest

import oncolytica_amd as ol
from oracle import oracle as O
from oracle.rule_exec import RuleState
from oncolytica_amd import workloads as W
from oncolytica_amd.hip._abi import BC_NONE
from oncolytica_amd.spheroid import Spheroid


def _spheroid_desc():
    w = W.large_c2()
    w.dims, w.capacity, w.key_format, w.cell_radius, w.k_adh = (64, 64, 64), 1 << 16, 30, 0.5, 0.0
    return w.desc()


def test_spheroid_rules_reproduce_the_builtin_step():
    d = _spheroid_desc()
    cols = W.initial_cells(W.plain_c1())
    sim = Spheroid()
    tis = {"oxygen": O.field_init(d.fields[0], (64, 64, 64)), "drug": O.field_init(d.fields[1], (64, 64, 64))}
    ex = RuleState(d, sim, cols, tis, cell_rules=[sim.lifecycle])
    st = O.TumorState(d, cols)
    for k in range(10):
        ex.run(1)
        st.run(1)
        oc = st.columns()
        assert len(oc[0]) == len(ex.cols[0]), k
        for a, b in zip(oc, ex.cols):
            np.testing.assert_array_equal(a.view(np.uint32), b.astype(a.dtype).view(np.uint32))
        for f, name in enumerate(("oxygen", "drug")):
            np.testing.assert_array_equal(st.fields[f].view(np.uint32), ex.tissue[name].view(np.uint32))
    assert ex.next_id == st.next_id.value > 4096


class VTissue(ol.Tissue):
    count: ol.i32 = 0
    nbs: ol.i32 = 0
    xs: ol.f32 = 0.0


class VChem(ol.Chemistry):
    kids: ol.i32 = 0
    ncells: ol.i32 = 0
    ksum: ol.i32 = 0
    nsum: ol.f32 = 0.0
    val: ol.f32 = 0.0


class VCell(ol.Cell):
    pos: ol.vec3


class VMetrics(ol.Metrics):
    n: ol.u32 = 0


class VParams(ol.Params):
    pass


class VoxelSim(ol.Simulation[VTissue, VChem, VCell, VMetrics, VParams]):
    @ol.tissue_rule
    def census(self, t: VTissue):
        n: ol.i32 = 0
        for c in t.cells:
            n += 1
        t.count = n
        k: ol.i32 = 0
        for nb in t.neighbors:
            k += 1
        t.nbs = k
        t.xs = t.pos.x

    @ol.chemistry_rule(iterations=1)
    def gather(self, c: VChem):
        k: ol.i32 = 0
        s: ol.i32 = 0
        for t in c.tissues:
            k += 1
            s += t.coord.x + 10 * t.coord.y + 100 * t.coord.z
        c.kids = k
        c.ksum = s
        m: ol.i32 = 0
        for cell in c.cells:
            m += 1
        c.ncells = m
        acc: ol.f32 = 0.0
        for nb in c.neighbors:
            acc += nb.val
        c.nsum = acc


def _voxel_state(n=400, dims=(9, 8, 7), seed=3):
    rng = np.random.default_rng(seed)
    pos = (rng.random((n, 3)) * np.array(dims)).astype(np.float32)
    cols = [pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(), np.arange(n, dtype=np.uint32)]
    d = W.plain_c1().desc()
    d.dims[:] = list(dims)
    d.k_rep = 0.0
    d.capacity = 4096
    d.n_cols = 4
    d.n_fields = 3
    for f in range(3):
        d.fields[f].rate, d.fields[f].decay, d.fields[f].bc, d.fields[f].uptake = 0.0, 0.0, BC_NONE, 0.0
        d.fields[f].lo, d.fields[f].hi = -3e38, 3e38
    nv = int(np.prod(dims))
    cd = tuple(max(1, v // 2) for v in dims)
    nc = int(np.prod(cd))
    tissue = {k: np.zeros(nv, np.float32) for k in ("count", "nbs", "xs")}
    chem = {"kids": np.zeros(nc, np.int32), "ncells": np.zeros(nc, np.int32), "ksum": np.zeros(nc, np.int32),
            "nsum": np.zeros(nc, np.float32), "val": rng.random(nc).astype(np.float32)}
    return d, cols, tissue, chem, pos, cd


def test_voxel_loops_visit_what_they_define():
    dims = (9, 8, 7)
    d, cols, tissue, chem, pos, cd = _voxel_state(dims=dims)
    sim = VoxelSim()
    ex = RuleState(d, sim, cols, tissue, chem, tissue_rules=[sim.census], chem_rules=[sim.gather])
    ex.run(1)
    X, Y, Z = dims
    vox = np.floor(pos).astype(np.int64)
    lin = vox[:, 0] + X * (vox[:, 1] + Y * vox[:, 2])
    np.testing.assert_array_equal(ex.tissue["count"].astype(np.int64), np.bincount(lin, minlength=X * Y * Z))
    # 26 neighbours inside, fewer at faces: (clipped extent product) - 1
    g = np.indices((Z, Y, X))
    ext = [np.minimum(g[k] + 1, n - 1) - np.maximum(g[k] - 1, 0) + 1 for k, n in zip(range(3), (Z, Y, X))]
    np.testing.assert_array_equal(ex.tissue["nbs"], (ext[0] * ext[1] * ext[2] - 1).reshape(-1))
    np.testing.assert_array_equal(ex.tissue["xs"], (g[2] + np.float32(0.5)).reshape(-1).astype(np.float32))
    CX, CY, CZ = cd
    cg = np.indices((CZ, CY, CX))
    kids, ksum, ccount = [], [], []
    for z, y, x in zip(cg[0].ravel(), cg[1].ravel(), cg[2].ravel()):
        ch = [(2 * x + a, 2 * y + b, 2 * z + c) for c in (0, 1) for b in (0, 1) for a in (0, 1)
              if 2 * x + a < X and 2 * y + b < Y and 2 * z + c < Z]
        kids.append(len(ch))
        ksum.append(sum(p + 10 * q + 100 * r for p, q, r in ch))
        ccount.append(sum(int(np.sum(lin == p + X * (q + Y * r))) for p, q, r in ch))
    np.testing.assert_array_equal(ex.chem["kids"], kids)
    np.testing.assert_array_equal(ex.chem["ksum"], ksum)
    np.testing.assert_array_equal(ex.chem["ncells"], ccount)
    # the neighbour sum of the chemistry values (binary32, Moore order)
    v = chem["val"].reshape(CZ, CY, CX)
    ref = np.zeros(CX * CY * CZ, np.float32)
    for i, (z, y, x) in enumerate(zip(cg[0].ravel(), cg[1].ravel(), cg[2].ravel())):
        acc = np.float32(0.0)
        for oz in (-1, 0, 1):
            for oy in (-1, 0, 1):
                for ox in (-1, 0, 1):
                    if (ox, oy, oz) == (0, 0, 0):
                        continue
                    if 0 <= x + ox < CX and 0 <= y + oy < CY and 0 <= z + oz < CZ:
                        acc = np.float32(acc + v[z + oz, y + oy, x + ox])
        ref[i] = acc
    np.testing.assert_array_equal(ex.chem["nsum"].view(np.uint32), ref.view(np.uint32))


class IChem(ol.Chemistry):
    drug_conc: ol.f32 = 0.0


class ITissue(ol.Tissue):
    oxygen: ol.f32 = 1.0


class IterSim(ol.Simulation[ITissue, IChem, VCell, VMetrics, VParams]):
    @ol.chemistry_rule(iterations=3)
    def diffuse(self, chem: IChem):
        chem.drug_conc += 1.0

    @ol.tissue_rule
    def tissue_logic(self, t: ITissue):
        t.oxygen -= 0.01


def test_reference_iterations_and_tissue_rule():
    d, cols, _, _, _, cd = _voxel_state(n=4, dims=(2, 2, 2))
    d.n_fields = 1
    sim = IterSim()
    ex = RuleState(d, sim, cols, {"oxygen": np.ones(8, np.float32)}, {"drug_conc": np.zeros(1, np.float32)},
                   tissue_rules=[sim.tissue_logic], chem_rules=[sim.diffuse])
    ex.run(1)
    assert ex.chem["drug_conc"][0] == 3.0
    assert np.all(ex.tissue["oxygen"] == np.float32(1.0) - np.float32(0.01))
