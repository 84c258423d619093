"""Per-kernel parity of liboncohip.so against the CPU oracle: bit-exact for
every integer output and for the binary32 stencil."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from oncolytica_amd.hip import _abi
from oncolytica_amd.hip._abi import BC_DIRICHLET, BC_NONE, BC_POINT_SOURCE, FieldDesc, ptr

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _dims(d):
    return np.ascontiguousarray(np.asarray(d, dtype=np.int32))


def gpu_keys(ctx, p, dims, dx, fmt):
    n = len(p)
    bx, by, bz = (ctx.upload(np.ascontiguousarray(p[:, i], dtype=np.float32)) for i in range(3))
    out = ctx.alloc(4 * n)
    _abi.call("onc_morton_keys", ctx.h, bx.p, by.p, bz.p, n, _dims(dims).ctypes.data_as(_abi._i32p),
              dx, fmt, out.p)
    return out.download(np.uint32, n)


@pytest.mark.parametrize("dims,dx,fmt", [((64, 64, 64), 1.0, 30), ((50, 50, 50), 2.0, 30),
                                         ((256, 256, 256), 1.0, 32), ((2000, 1500, 700), 0.5, 32)])
def test_morton_keys(ctx, dims, dx, fmt):
    rng = np.random.default_rng(1)
    n = 300_001
    p = (rng.random((n, 3)) * np.array(dims) * dx * 1.02 - 0.01 * dx).astype(np.float32)
    k = gpu_keys(ctx, p, dims, dx, fmt)
    np.testing.assert_array_equal(k, O.keys(p[:, 0], p[:, 1], p[:, 2], dims, dx, fmt))


def test_reference_sort_golden_through_kernels(ctx):
    """framework/tests/test_gpu.py:17-61 inputs: GPU keys + stable sort give
    the reference's buckets with ids in input order."""
    g = np.load(os.path.join(GOLDEN, "sort_golden.npz"))
    pos = g["pos"].astype(np.float32)
    k = gpu_keys(ctx, pos, tuple(g["dims"]), float(g["voxel"][0]), 30)
    np.testing.assert_array_equal(k, g["keys"])
    ks, vs = gpu_sort(ctx, k, np.arange(len(k), dtype=np.uint32), 18)
    order = np.argsort(g["keys"], kind="stable")
    np.testing.assert_array_equal(vs, order)
    np.testing.assert_array_equal(ks, g["keys"][order])


def gpu_sort(ctx, k, v, bits):
    n = len(k)
    kin, vin = ctx.upload(k), ctx.upload(v)
    ko, vo = ctx.alloc(4 * n), ctx.alloc(4 * n)
    _abi.call("onc_sort_pairs", ctx.h, kin.p, vin.p, n, bits, ko.p, vo.p)
    return ko.download(np.uint32, n), vo.download(np.uint32, n)


@pytest.mark.parametrize("n,bits", [(0, 18), (1, 18), (7, 8), (4095, 18), (4096, 18), (4097, 18),
                                    (100_003, 24), (1_000_000, 24), (1_000_000, 32), (2_105_345, 32),
                                    (3_000_017, 27), (16_777_216, 27)])
def test_sort_pairs(ctx, n, bits):
    print(f"sort n={n} bits={bits}", flush=True)
    rng = np.random.default_rng(n + bits)
    k = rng.integers(0, 2**bits, n, dtype=np.uint64).astype(np.uint32)
    if n > 100:
        k[n // 4: n // 2] = k[n // 4]          # a long run of equal keys
        k[: n // 8] = np.sort(k[: n // 8])      # a presorted stretch
    v = np.arange(n, dtype=np.uint32)[::-1].copy()
    ks, vs = gpu_sort(ctx, k, v, bits)
    ok, ov = O.radix_sort_pairs(k, v, bits)
    np.testing.assert_array_equal(ks, ok)
    np.testing.assert_array_equal(vs, ov)


def test_sort_morton_like_keys_many_duplicates(ctx):
    rng = np.random.default_rng(5)
    k = np.sort(rng.integers(0, 2**18, 500_000)).astype(np.uint32)
    k = k[rng.permutation(len(k))]
    k[::3] = 17
    v = np.arange(len(k), dtype=np.uint32)
    ks, vs = gpu_sort(ctx, k, v, 18)
    order = np.argsort(k, kind="stable")
    np.testing.assert_array_equal(vs, order)


@pytest.mark.parametrize("n", [0, 1, 100, 4096, 4097, 16_383, 16_384, 16_385, 1_000_000, 16_777_215,
                               16_777_216, 16_809_985])
def test_exscan(ctx, n):
    print(f"exscan n={n}", flush=True)
    rng = np.random.default_rng(n)
    a = rng.integers(0, 1000, n).astype(np.uint32)
    if n > 10:
        a[-5:] = 0xFFFFFFF0  # wraps mod 2^32
    bi = ctx.upload(a) if n else ctx.alloc(4)
    bo = ctx.alloc(max(4, 4 * n))
    tot = ctx.alloc(4)
    _abi.call("onc_exscan_u32", ctx.h, bi.p, bo.p, n, tot.p)
    out = bo.download(np.uint32, n)
    ref, rt = O.exscan(a)
    np.testing.assert_array_equal(out, ref)
    assert int(tot.download(np.uint32, 1)[0]) == rt


@pytest.mark.parametrize("shift_in,shift_out", [(1, 0), (0, 3), (2, 1)])
def test_exscan_unaligned_pointers(ctx, shift_in, shift_out):
    """Input or output pointers that are not 16-byte aligned (the kernel then
    moves one element at a time) give the same scan."""
    n = 50_001
    rng = np.random.default_rng(shift_in * 7 + shift_out)
    a = rng.integers(0, 1 << 20, n).astype(np.uint32)
    bi = ctx.upload(np.concatenate([np.zeros(shift_in, np.uint32), a]))
    bo = ctx.alloc(4 * (n + shift_out))
    tot = ctx.alloc(4)
    _abi.call("onc_exscan_u32", ctx.h, bi.p.value + 4 * shift_in, bo.p.value + 4 * shift_out, n, tot.p)
    out = bo.download(np.uint32, n + shift_out)[shift_out:]
    ref, rt = O.exscan(a)
    np.testing.assert_array_equal(out, ref)
    assert int(tot.download(np.uint32, 1)[0]) == rt


def test_bins(ctx):
    rng = np.random.default_rng(3)
    table = 1 << 18
    k = np.sort(rng.integers(0, table, 300_000)).astype(np.uint32)
    fill = np.full(table, 0xFFFFFFFF, dtype=np.uint32)
    kb, sb, eb = ctx.upload(k), ctx.upload(fill), ctx.upload(fill)
    _abi.call("onc_bins", ctx.h, kb.p, len(k), sb.p, eb.p)
    s, e = O.bins(k, table)
    np.testing.assert_array_equal(sb.download(np.uint32, table), s)
    np.testing.assert_array_equal(eb.download(np.uint32, table), e)


def _fd(**kw):
    f = FieldDesc()
    f.rate, f.decay, f.init = kw.get("rate", 0.1), kw.get("decay", 0.0), kw.get("init", 0.0)
    f.bc, f.bc_value = kw.get("bc", BC_DIRICHLET), kw.get("bc_value", 0.0)
    f.src_radius, f.uptake = kw.get("src_radius", 0.0), kw.get("uptake", 0.0)
    f.lo, f.hi = kw.get("lo", 0.0), kw.get("hi", 1.0)
    return f


@pytest.mark.parametrize("dims,kw", [
    ((64, 64, 64), dict(rate=0.1, bc_value=1.0)),
    ((256, 256, 256), dict(rate=0.05, decay=0.001, bc=BC_POINT_SOURCE, bc_value=1.0, src_radius=16.0)),
    ((37, 91, 13), dict(rate=0.15, decay=0.01, bc=BC_POINT_SOURCE, bc_value=0.7, src_radius=5.0)),
    ((130, 66, 200), dict(rate=0.1, bc_value=0.25, lo=0.1, hi=0.9)),
    # float4 rows whose X/4 is not a power of two (row segments padded to the
    # next power of two, x halos shuffled between lanes): 48, 25, 129, 2, 1
    ((192, 40, 24), dict(rate=0.1, bc_value=1.0)),
    ((100, 30, 20), dict(rate=0.12, decay=0.002, bc=BC_POINT_SOURCE, bc_value=0.9, src_radius=9.0)),
    ((516, 9, 7), dict(rate=0.1, bc=BC_NONE, lo=0.0, hi=1.0)),
    ((8, 5, 6), dict(rate=0.1, bc_value=0.5)),
    ((4, 3, 3), dict(rate=0.1, bc=BC_NONE)),
])
def test_stencil_bit_exact(ctx, dims, kw):
    rng = np.random.default_rng(sum(dims))
    u = rng.random(int(np.prod(dims))).astype(np.float32)
    f = _fd(**kw)
    bi = ctx.upload(u)
    bo = ctx.alloc(u.nbytes)
    _abi.call("onc_stencil7", ctx.h, bi.p, bo.p, _dims(dims).ctypes.data_as(_abi._i32p), f)
    got = bo.download(np.float32, u.size)
    ref = O.stencil7(u, dims, f, "f32")
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_philox_known_answers_and_random(ctx):
    d = json.load(open(os.path.join(GOLDEN, "philox_kat.json")))
    for src in ("random123", "rocrand"):
        for ctr, key, out in d[src]:
            c = ctx.upload(np.array(ctr, dtype=np.uint32))
            o = ctx.alloc(16)
            _abi.call("onc_philox", ctx.h, c.p, 1, key[0] | (key[1] << 32), o.p)
            assert [int(v) for v in o.download(np.uint32, 4)] == out
    rng = np.random.default_rng(9)
    ctr = rng.integers(0, 2**32, (100_000, 4), dtype=np.uint64).astype(np.uint32)
    c = ctx.upload(ctr)
    o = ctx.alloc(ctr.nbytes)
    _abi.call("onc_philox", ctx.h, c.p, len(ctr), 0x123456789ABCDEF, o.p)
    np.testing.assert_array_equal(o.download(np.uint32, ctr.size).reshape(-1, 4),
                                  O.philox(ctr, 0x123456789ABCDEF))


def test_pure_kernel_timing(ctx):
    """onc_ctx_timing: a timed call reports a positive GPU span, results are
    unchanged, and reading the span with timing off is an error."""
    dims = (64, 64, 64)
    u = np.random.default_rng(2).random(int(np.prod(dims))).astype(np.float32)
    f = _fd(rate=0.1, bc_value=1.0)
    bi, bo = ctx.upload(u), ctx.alloc(u.nbytes)
    ctx.timing(True)
    try:
        _abi.call("onc_stencil7", ctx.h, bi.p, bo.p, _dims(dims).ctypes.data_as(_abi._i32p), f)
        assert ctx.last_kernel_us() > 0.0
        np.testing.assert_array_equal(bo.download(np.float32, u.size).view(np.uint32),
                                      O.stencil7(u, dims, f, "f32").view(np.uint32))
    finally:
        ctx.timing(False)
    with pytest.raises(_abi.OncError):
        ctx.last_kernel_us()


@pytest.mark.parametrize("n,shift", [(1, 0), (3, 0), (4, 0), (5, 0), (1025, 0), (1_000_003, 0), (1_000_003, 1),
                                     (4099, 2)])
def test_copy_u32(ctx, n, shift):
    """The plain copy (the byte counters' calibration kernel): 16-byte
    streaming copies with a tail when the pointers are aligned, one word at a
    time when one is shifted by `shift` words."""
    rng = np.random.default_rng(n + shift)
    a = rng.integers(0, 2**32, n + shift, dtype=np.uint64).astype(np.uint32)
    src = ctx.upload(a)
    dst = ctx.upload(np.full(n + 8, 0xDEADBEEF, dtype=np.uint32))
    _abi.call("onc_copy_u32", ctx.h, src.p.value + 4 * shift, dst.p, n)
    got = dst.download(np.uint32, n + 8)
    np.testing.assert_array_equal(got[:n], a[shift:])
    assert (got[n:] == 0xDEADBEEF).all()


def _forces_case(ctx, d, p):
    k = O.keys(p[:, 0], p[:, 1], p[:, 2], tuple(d.dims), d.dx, d.key_format)
    p = np.ascontiguousarray(p[np.argsort(k, kind="stable")])
    n = len(p)
    bufs = [ctx.upload(np.ascontiguousarray(p[:, a])) for a in range(3)]
    outs = [ctx.alloc(4 * max(1, n)) for _ in range(3)]
    _abi.call("onc_forces", ctx.h, d, *[b.p for b in bufs], n, *[o.p for o in outs])
    got = [o.download(np.float32, n) for o in outs]
    ref = O.forces_raw(d, p[:, 0], p[:, 1], p[:, 2], "f32")
    for a in range(3):
        np.testing.assert_array_equal(got[a].view(np.uint32), ref[a].view(np.uint32))
    return got


@pytest.mark.parametrize("case", ["c1_sphere", "dense_adhesion", "single"])
def test_forces_bit_exact_with_oracle(ctx, case):
    """onc_forces (the force walk and pair law alone, on key-sorted cells)
    gives the oracle's binary32 forces bit for bit."""
    from oncolytica_amd import workloads as Wl
    if case == "c1_sphere":
        w = Wl.plain_c1()
        d = w.desc()
        cols = Wl.initial_cells(w)
        p = np.stack(cols[:3], 1).astype(np.float32)
    elif case == "dense_adhesion":
        w = Wl.plain_c1()
        d = w.desc()
        d.k_adh, d.interaction_radius = 0.2, 1.0
        rng = np.random.default_rng(12)
        p = (rng.random((30_000, 3)) * 12.0 + 26.0).astype(np.float32)  # ~17 cells per voxel
    else:
        w = Wl.plain_c1()
        d = w.desc()
        p = np.array([[10.5, 20.25, 30.0]], dtype=np.float32)
    got = _forces_case(ctx, d, p)
    if case == "single":
        assert all(g[0] == 0.0 for g in got)
    else:
        assert max(np.abs(g).max() for g in got) > 0.0


def _pairs_at_reach(d, n, rng):
    """n pairs of cells straddling a face, edge or corner of a voxel, whose
    distance is R (1 + e) with e in [-3e-6, 3e-6]: the first cell's gaps to
    the neighbour voxel add up (in quadrature) to about R, the second sits on
    or just past the shared boundary."""
    dx, R = float(d.dx), float(d.interaction_radius)
    X = d.dims[0]
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1) if (a, b, c) != (0, 0, 0)]
    pts = []
    for _ in range(n):
        o = np.array(offs[rng.integers(len(offs))])
        v = rng.integers(2, X - 2, size=3)
        nz = np.nonzero(o)[0]
        w = rng.random(len(nz)) + 0.2
        g = np.zeros(3)
        g[nz] = w / np.linalg.norm(w) * R * (1.0 + rng.uniform(-3e-6, 3e-6))
        g = np.minimum(g, dx)
        a = (v + rng.random(3)) * dx
        b = a.copy()
        past = rng.choice([0.0, 1e-7, 1e-6]) * dx
        for ax in nz:
            if o[ax] > 0:
                a[ax] = (v[ax] + 1) * dx - g[ax]
                b[ax] = (v[ax] + 1) * dx + past
            else:
                a[ax] = v[ax] * dx + g[ax]
                b[ax] = v[ax] * dx - past
        pts += [a, b]
    return np.array(pts, dtype=np.float32)


@pytest.mark.parametrize("dx", [1.0, 2.5])
def test_forces_at_the_interaction_radius_bit_exact(ctx, dx):
    """Pairs at R (1 +- 3e-6) across voxel faces, edges and corners, with
    voxels as wide as R and wider: the force walk skips only bins out of reach
    of the cell, so it finds every pair the oracle finds, bit for bit."""
    from oncolytica_amd import workloads as Wl
    d = Wl.plain_c1().desc()
    d.dx = dx
    d.k_adh = 0.2
    rng = np.random.default_rng(int(dx * 10))
    p = _pairs_at_reach(d, 6000, rng)
    k = O.keys(p[:, 0], p[:, 1], p[:, 2], tuple(d.dims), d.dx, d.key_format)
    p = np.ascontiguousarray(p[np.argsort(k, kind="stable")])
    ref = O.forces_raw(d, p[:, 0], p[:, 1], p[:, 2], "f32")
    assert (np.abs(np.stack(ref)).max(0) > 0).mean() > 0.3  # many pairs fall inside R
    _forces_case(ctx, d, p)


def test_pinned_host_buffers_round_trip(ctx):
    """onc_host_alloc: page-locked host memory moves data to and from the
    device like ordinary host memory."""
    n = 3_000_001
    src = _abi.PinnedBuffer(ctx, n, np.uint32)
    dst = _abi.PinnedBuffer(ctx, n, np.uint32)
    src.array[:] = np.arange(n, dtype=np.uint32) * 2654435761
    dev = ctx.alloc(4 * n)
    _abi.call("onc_h2d", ctx.h, dev.p, src.p, 4 * n)
    _abi.call("onc_d2h", ctx.h, dst.p, dev.p, 4 * n)
    np.testing.assert_array_equal(dst.array, src.array)
    # a view taken from the buffer keeps its memory after the handle is freed
    view = dst.array[10:20]
    dst.free()
    del dst
    import gc
    gc.collect()
    np.testing.assert_array_equal(view, np.arange(10, 20, dtype=np.uint32) * np.uint32(2654435761))
    src.free()
