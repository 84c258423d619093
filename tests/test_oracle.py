"""Pins the CPU oracle against independent checks: the reference's own test
inputs (re-created), published known-answer vectors, library routines and
closed forms. CPU only."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from oncolytica_amd import workloads as W
from oncolytica_amd.hip._abi import BC_DIRICHLET, BC_POINT_SOURCE, FieldDesc, SimDesc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def morton_bits(x, y, z, fmt=30):
    x = np.asarray(x, dtype=np.uint64)
    y = np.asarray(y, dtype=np.uint64)
    z = np.asarray(z, dtype=np.uint64)
    k = np.zeros(x.shape, dtype=np.uint64)
    for b in range(10):
        k |= ((x >> b) & 1) << (3 * b)
        k |= ((y >> b) & 1) << (3 * b + 1)
        k |= ((z >> b) & 1) << (3 * b + 2)
    if fmt == 32:
        k |= ((x >> 10) & 1) << 30
        k |= ((y >> 10) & 1) << 31
    return k.astype(np.uint32)


# ---- Morton ------------------------------------------------------------------

def test_morton30_equals_bit_interleave_on_all_64cubed():
    g = np.arange(64, dtype=np.uint32)
    x, y, z = (a.ravel() for a in np.meshgrid(g, g, g, indexing="ij"))
    np.testing.assert_array_equal(O.morton(x, y, z, 30), morton_bits(x, y, z))


def test_morton32_equals_bit_interleave_up_to_2047():
    rng = np.random.default_rng(1)
    x = rng.integers(0, 2048, 200_000).astype(np.uint32)
    y = rng.integers(0, 2048, 200_000).astype(np.uint32)
    z = rng.integers(0, 1024, 200_000).astype(np.uint32)
    np.testing.assert_array_equal(O.morton(x, y, z, 32), morton_bits(x, y, z, 32))
    # below 1024 both formats give the same key
    m = (x < 1024) & (y < 1024)
    np.testing.assert_array_equal(O.morton(x[m], y[m], z[m], 30), O.morton(x[m], y[m], z[m], 32))


@pytest.mark.parametrize("d,table", [(64, 1_835_009), (128, 14_680_065), (240, 16_773_121),
                                     (256, 117_440_513)])
def test_voxel_table_size_closed_form(d, table):
    # voxel_table_size(dims) = key(dims) + 1 (core/runtime/_geometry.py:45-47);
    # table sizes as tabulated from the reference closed form
    assert int(O.morton([d], [d], [d], 30)[0]) + 1 == table


def test_morton_monotone_componentwise():
    rng = np.random.default_rng(2)
    a = rng.integers(0, 500, (10_000, 3)).astype(np.uint32)
    b = a + rng.integers(0, 3, (10_000, 3)).astype(np.uint32)
    ka = O.morton(a[:, 0], a[:, 1], a[:, 2])
    kb = O.morton(b[:, 0], b[:, 1], b[:, 2])
    assert np.all(kb >= ka)


def test_reference_sort_golden_buckets():
    """framework/tests/test_gpu.py:17-61 inputs; the oracle's keys of the
    float32 positions equal the bucket of the reference's float64 test."""
    g = np.load(os.path.join(GOLDEN, "sort_golden.npz"))
    pos = g["pos"].astype(np.float32)
    k = O.keys(pos[:, 0], pos[:, 1], pos[:, 2], g["dims"], float(g["voxel"][0]), 30)
    np.testing.assert_array_equal(k, g["keys"])
    # the stable sort groups each bucket with ids in input order
    ks, ids = O.radix_sort_pairs(k, np.arange(len(k), dtype=np.uint32), 18)
    ref = {}
    for i, kk in enumerate(g["keys"]):
        ref.setdefault(int(kk), []).append(i)
    pos_ = 0
    for kk in sorted(ref):
        n = len(ref[kk])
        assert list(ids[pos_:pos_ + n]) == ref[kk]
        assert np.all(ks[pos_:pos_ + n] == kk)
        pos_ += n
    assert pos_ == len(k)


def test_voxel_floor_and_clamp():
    dims = (8, 8, 8)
    p = np.array([-1.0, 0.0, 0.999999, 1.0, 7.99, 8.0, 100.0], dtype=np.float32)
    k = O.keys(p, np.zeros_like(p), np.zeros_like(p), dims, 1.0)
    expect = morton_bits(np.array([0, 0, 0, 1, 7, 7, 7]), np.zeros(7, int), np.zeros(7, int))
    np.testing.assert_array_equal(k, expect)


# ---- sort / scan / bins --------------------------------------------------------

@pytest.mark.parametrize("n,bits", [(0, 18), (1, 18), (5, 8), (4097, 18), (100_000, 24), (70_000, 32)])
def test_radix_sort_matches_stable_argsort(n, bits):
    rng = np.random.default_rng(n)
    k = rng.integers(0, 2**bits, n, dtype=np.uint64).astype(np.uint32)
    if n > 10:
        k[: n // 3] = k[0]  # many duplicates
    v = rng.permutation(n).astype(np.uint32)
    ks, vs = O.radix_sort_pairs(k, v, bits)
    order = np.argsort(k, kind="stable")
    np.testing.assert_array_equal(ks, k[order])
    np.testing.assert_array_equal(vs, v[order])


@pytest.mark.parametrize("n", [0, 1, 2, 255, 256, 1000, 4097, 1 << 20])
def test_exscan_matches_cumsum(n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 100, n).astype(np.uint32)
    out, tot = O.exscan(a)
    ref = np.concatenate([[0], np.cumsum(a.astype(np.uint64))[:-1]]).astype(np.uint32) if n else a
    np.testing.assert_array_equal(out, ref)
    assert tot == int(a.astype(np.uint64).sum()) % 2**32
    if n:
        assert out[0] == 0
        np.testing.assert_array_equal(np.diff(out.astype(np.int64)), a[:-1].astype(np.int64))


def test_exscan_wraps_mod_2_32():
    a = np.full(10, 2**31, dtype=np.uint32)
    out, tot = O.exscan(a)
    assert tot == 0 and out[1] == 2**31 and out[2] == 0


def test_bins_match_searchsorted():
    rng = np.random.default_rng(3)
    k = np.sort(rng.integers(0, 5000, 20_000).astype(np.uint32))
    s, e = O.bins(k, 5000)
    occ = np.unique(k)
    np.testing.assert_array_equal(s[occ], np.searchsorted(k, occ, "left"))
    np.testing.assert_array_equal(e[occ], np.searchsorted(k, occ, "right"))
    empty = np.setdiff1d(np.arange(5000), occ)
    assert np.all(s[empty] == 0xFFFFFFFF) and np.all(e[empty] == 0xFFFFFFFF)


# ---- Philox ------------------------------------------------------------------

def test_philox_known_answers():
    d = json.load(open(os.path.join(GOLDEN, "philox_kat.json")))
    for src in ("random123", "rocrand"):
        for ctr, key, out in d[src]:
            got = O.philox(np.array(ctr, dtype=np.uint32), key[0] | (key[1] << 32))[0]
            assert [int(v) for v in got] == out, (src, ctr, key)


# ---- RNG-derived draws ---------------------------------------------------------

def test_sincos_polynomial_accuracy():
    a = np.linspace(0.0, math.pi / 2, 2001)
    for x in a:
        s, c = O.sincos_q(float(np.float32(x)), "f32")
        xf = float(np.float32(x))
        assert abs(s - math.sin(xf)) < 4e-7 and abs(c - math.cos(xf)) < 4e-7
        s64, c64 = O.sincos_q(x, "f64")
        assert abs(s64 - math.sin(x)) < 1e-9 and abs(c64 - math.cos(x)) < 1e-9


def test_random_dir_unit_and_isotropic():
    v = np.array([O.random_dir(i, 7, 42) for i in range(4000)], dtype=np.float64)
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=2e-6)
    assert np.all(np.abs(v.mean(axis=0)) < 0.05)
    # a change of seed, id or step changes the draw
    assert not np.allclose(O.random_dir(1, 7, 42), O.random_dir(1, 7, 43))
    assert not np.allclose(O.random_dir(1, 7, 42), O.random_dir(1, 8, 42))


# ---- stencil ------------------------------------------------------------------

def _fd(**kw):
    f = FieldDesc()
    f.rate, f.decay, f.init = kw.get("rate", 0.1), kw.get("decay", 0.0), kw.get("init", 0.0)
    f.bc, f.bc_value = kw.get("bc", BC_DIRICHLET), kw.get("bc_value", 0.0)
    f.src_radius, f.uptake = kw.get("src_radius", 0.0), kw.get("uptake", 0.0)
    f.lo, f.hi = kw.get("lo", -1e30), kw.get("hi", 1e30)
    return f


def test_stencil_constant_dirichlet_is_fixed_point():
    dims = (12, 10, 9)
    f = _fd(init=1.0, bc_value=1.0)
    u = O.field_init(f, dims)
    np.testing.assert_array_equal(O.stencil7(u, dims, f), u)


@pytest.mark.parametrize("k", [(1, 1, 1), (2, 1, 3)])
def test_stencil_sine_mode_decay_factor(k):
    """Discrete eigenmode of the Dirichlet Laplacian: one sweep multiplies it
    by 1 - 4 r sum_d sin^2(pi k_d / (2 (n_d + 1))) - decay."""
    n = (14, 12, 10)
    dims = tuple(m + 2 for m in n)
    r, dec = 0.1, 0.003
    f = _fd(rate=r, decay=dec, bc_value=0.0)
    idx = [np.arange(d) for d in dims]
    X, Y, Z = np.meshgrid(*idx, indexing="ij")
    mode = (np.sin(math.pi * k[0] * X / (n[0] + 1)) * np.sin(math.pi * k[1] * Y / (n[1] + 1))
            * np.sin(math.pi * k[2] * Z / (n[2] + 1)))
    u = np.transpose(mode, (2, 1, 0)).reshape(-1)  # x fastest
    lam = 1 - 4 * r * sum(math.sin(math.pi * k[d] / (2 * (n[d] + 1))) ** 2 for d in range(3)) - dec
    out = O.stencil7(u, dims, f, "f64")
    np.testing.assert_allclose(out, lam * u, atol=1e-13)


def test_stencil_zero_flux_conserves_mass():
    dims = (10, 12, 7)  # even x, y: no voxel centre sits on the (empty) source disc
    f = _fd(rate=0.15, bc=BC_POINT_SOURCE, src_radius=0.0)
    rng = np.random.default_rng(5)
    u = rng.random(int(np.prod(dims)))
    out = O.stencil7(u, dims, f, "f64")
    assert abs(out.sum() - u.sum()) < 1e-10 * u.sum()


def test_stencil_point_source_disc():
    dims = (32, 32, 8)
    f = _fd(bc=BC_POINT_SOURCE, bc_value=1.0, src_radius=4.0)
    u = O.field_init(f, dims).reshape(dims[2], dims[1], dims[0])
    fixed = u[0] == 1.0
    yy, xx = np.mgrid[0:32, 0:32]
    expect = (xx + 0.5 - 16) ** 2 + (yy + 0.5 - 16) ** 2 <= 16
    np.testing.assert_array_equal(fixed, expect)
    assert np.all(u[1:] == 0.0)


def test_stencil_f32_tracks_f64():
    dims = (20, 18, 16)
    f = _fd(rate=0.1, decay=0.001, init=0.5, bc_value=1.0, lo=0.0, hi=1.0)
    u32 = O.field_init(f, dims, "f32")
    u64 = u32.astype(np.float64)
    for _ in range(50):
        u32 = O.stencil7(u32, dims, f, "f32")
        u64 = O.stencil7(u64, dims, f, "f64")
    assert np.max(np.abs(u32 - u64)) < 1e-5


# ---- forces -------------------------------------------------------------------

def _desc(dims=(16, 16, 16), **kw):
    d = SimDesc()
    d.dims[:] = list(dims)
    d.dx = kw.get("dx", 1.0)
    d.key_format = 30
    d.capacity = kw.get("cap", 1 << 14)
    d.cell_radius = kw.get("cell_radius", 0.4)
    d.interaction_radius = kw.get("R", 1.0)
    d.k_rep = kw.get("k_rep", 1.0)
    d.k_adh = kw.get("k_adh", 0.2)
    d.mobility, d.max_disp = 0.25, 0.25
    return d


def _brute_forces(p, S, R, krep, kadh):
    d = p[:, None, :] - p[None, :, :]
    dist = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(dist, np.inf)
    f = np.where(dist < S, krep * (S - dist), np.where(dist < R, -kadh * (R - dist) / (R - S), 0.0))
    f = np.where(np.isfinite(dist) & (dist > 0) & (dist < R), f, 0.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        n = np.where(dist[..., None] > 0, d / dist[..., None], 0.0)
    n = np.nan_to_num(n)
    return (f[..., None] * n).sum(1)


def test_forces_match_all_pairs_brute_force():
    rng = np.random.default_rng(11)
    dims = (10, 10, 10)
    p = rng.uniform(2.0, 8.0, (600, 3))
    d = _desc(dims)
    k = O.keys(p[:, 0], p[:, 1], p[:, 2], dims, 1.0, real="f64")
    order = np.argsort(k, kind="stable")
    p = p[order]
    fx, fy, fz = O.forces_raw(d, p[:, 0], p[:, 1], p[:, 2], "f64")
    f32 = lambda v: float(np.float32(v))  # the descriptor holds binary32 parameters
    ref = _brute_forces(p, 2 * f32(0.4), 1.0, 1.0, f32(0.2))
    np.testing.assert_allclose(np.stack([fx, fy, fz], 1), ref, atol=1e-12)


def test_forces_newton_third_law():
    rng = np.random.default_rng(12)
    dims = (12, 12, 12)
    p = rng.uniform(3.0, 9.0, (3000, 3))
    d = _desc(dims)
    k = O.keys(p[:, 0], p[:, 1], p[:, 2], dims, 1.0, real="f64")
    p = p[np.argsort(k, kind="stable")]
    fx, fy, fz = O.forces_raw(d, p[:, 0], p[:, 1], p[:, 2], "f64")
    tot = np.array([fx.sum(), fy.sum(), fz.sum()])
    assert np.all(np.abs(tot) < 1e-9)
    assert np.abs(fx).max() > 0.1


def test_forces_pair_law_two_cells():
    d = _desc((8, 8, 8))
    for gap, expect in [(0.5, 1.0 * (0.8 - 0.5)), (0.9, -0.2 * (1.0 - 0.9) / 0.2), (1.1, 0.0)]:
        p = np.array([[4.0, 4.0, 4.0], [4.0 + gap, 4.0, 4.0]])
        k = O.keys(p[:, 0], p[:, 1], p[:, 2], (8, 8, 8), 1.0, real="f64")
        o = np.argsort(k, kind="stable")
        fx, _, _ = O.forces_raw(d, p[o, 0], p[o, 1], p[o, 2], "f64")
        f_first = fx[list(o).index(0)]
        assert f_first == pytest.approx(-expect, abs=1e-6)  # cell 0 sits at lower x; binary32 params


# ---- whole step ----------------------------------------------------------------

def test_tumor_step_die_divide_and_append_order():
    w = W.plain_c1()
    d = w.desc()
    d.capacity = 16
    # cell 0: hypoxic -> dies; cell 1: old and oxygenated -> divides; cell 2: young
    pos = np.array([[30.2, 30.2, 30.2], [20.5, 20.5, 20.5], [40.5, 40.5, 40.5]])
    cols = W.tumor_columns(pos, np.array([3.0, 30.0, 2.0]), w.dims, w.dx)
    st = O.TumorState(d, cols)
    f = st.fields[0].reshape(64, 64, 64)  # z, y, x
    f[29:32, 29:32, 29:32] = 0.0  # stays 0 at the centre after one sweep
    st.run(1)
    c = st.columns()
    assert st.n.value == 3 and st.next_id.value == 4
    # survivors in sorted (Morton) order, then the daughter with id 3
    assert list(c[3][:2]) == [1, 2] and c[3][2] == 3
    assert c[4][0] == 0.0 and c[5][0] == 1  # parent: age reset, generation + 1
    assert c[4][1] == 3.0                  # young cell aged by one step
    assert c[4][2] == 0.0 and c[5][2] == 1  # daughter
    off = np.array([c[0][2] - c[0][0], c[1][2] - c[1][0], c[2][2] - c[2][0]], dtype=np.float64)
    assert np.linalg.norm(off) == pytest.approx(0.5, abs=1e-5)


def test_tumor_step_capacity_drops_late_daughters():
    w = W.plain_c1()
    d = w.desc()
    d.capacity = 5
    pos = np.array([[10.5 + 4 * i, 20.5, 20.5] for i in range(4)])
    cols = W.tumor_columns(pos, np.full(4, 30.0), w.dims, w.dx)
    st = O.TumorState(d, cols)
    st.run(1)
    assert st.n.value == 5 and st.next_id.value == 5  # one of four daughters fits


def test_tumor_c1_f32_close_to_f64_over_10_steps():
    w = W.plain_c1()
    d = w.desc()
    cols = W.initial_cells(w)
    a = O.TumorState(d, cols, real="f32")
    b = O.TumorState(d, cols, real="f64")
    a.run(10)
    b.run(10)
    assert a.n.value == b.n.value
    ca, cb = a.columns(), b.columns()
    ia, ib = np.argsort(ca[3]), np.argsort(cb[3])
    for c in range(3):
        assert np.max(np.abs(ca[c][ia] - cb[c][ib])) < 1e-4
    # a cell whose binary32 and binary64 positions fall in different voxels moves
    # its uptake; everywhere else the fields agree to binary32 rounding
    dfield = np.abs(a.fields[0] - b.fields[0])
    assert np.mean(dfield > 1e-5) < 0.01 and dfield.max() < 0.02


def test_step_identical_for_any_thread_count():
    """The oracle's OpenMP loops write disjoint outputs: one thread and many
    give the same bits."""
    w = W.large_c2(n_cells=20_000)
    d = w.desc()
    cols = W.initial_cells(w)
    n_threads = O.threads()
    out = []
    try:
        for t in (1, max(2, n_threads)):
            O.set_threads(t)
            st = O.TumorState(d, cols)
            st.run(3)
            out.append((st.columns(), [f.copy() for f in st.fields]))
    finally:
        O.set_threads(n_threads)
    for a, b in zip(out[0][0] + out[0][1], out[1][0] + out[1][1]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ---- metric reductions ----------------------------------------------------------------
def _metric_reduce_scalar(contrib, cpt=8):
    """The same order as O.metric_reduce, one binary32 addition at a time."""
    f = np.float32
    n = len(contrib)
    per = 256 * cpt
    nb = max(1, -(-n // per))
    part = []
    for b in range(nb):
        th = []
        for t in range(256):
            a = f(0.0)
            for j in range(cpt):
                i = (b * cpt + j) * 256 + t
                if i < n:
                    a = f(a + f(contrib[i]))
            th.append(a)
        h = 128
        while h >= 1:
            for t in range(h):
                th[t] = f(th[t] + th[t + h])
            h //= 2
        part.append(th[0])
    v = [f(0.0)] * 256
    for q in range(nb):
        v[q % 256] = f(v[q % 256] + part[q])
    h = 128
    while h >= 1:
        for t in range(h):
            v[t] = f(v[t] + v[t + h])
        h //= 2
    return int(np.array([v[0]], dtype=np.float32).view(np.uint32)[0])


def test_metric_reduce_f32_matches_scalar_order():
    rng = np.random.default_rng(3)
    for n in (0, 1, 255, 2048, 2049, 5000):
        c = (rng.standard_normal(n) * 10).astype(np.float32)
        assert O.metric_reduce(c, "f32") == _metric_reduce_scalar(c), n


def test_metric_reduce_exact_cases():
    # sums that binary32 holds exactly give the exact value in any order
    for n in (0, 1, 3000, 70_000):
        assert np.uint32(O.metric_reduce(np.ones(n, np.float32), "f32")).view(np.float32) == n
    c = np.arange(10_000) % 7
    assert np.uint32(O.metric_reduce(c.astype(np.float32), "f32")).view(np.float32) == c.sum()
    # integers wrap mod 2^32 (atomicAdd / atomicSub semantics)
    rng = np.random.default_rng(4)
    c = rng.integers(-2**31, 2**31, 100_000)
    assert O.metric_reduce(c, "i32") == int(c.sum()) & 0xFFFFFFFF
    assert O.metric_reduce(np.full(600_000, 10_000), "u32") == (6_000_000_000 & 0xFFFFFFFF)
    # the order in closed form: 2^24 then 4095 ones. Thread 0 of block 0
    # holds cells 0, 256, ..., 1792 and absorbs its seven ones into 2^24 (a
    # plain sequential sum absorbs all of them); every other thread sums its
    # eight ones exactly, so the result is 2^24 + 4088.
    c = np.ones(4096, dtype=np.float32)
    c[0] = 2.0 ** 24
    assert np.uint32(O.metric_reduce(c, "f32")).view(np.float32) == 2.0 ** 24 + 4088
