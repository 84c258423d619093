"""onc_exscan_u32_noalloc, the entry point that allocates nothing, gives the
same prefix sums and totals as onc_exscan_u32, through the C ABI, on two
calls in a row (the second smaller, so it runs over words the first left in
the shared scratch), including unaligned pointers and its size limit."""
import ctypes as C

import numpy as np
import pytest

from oncolytica_amd.hip import _abi

pytestmark = pytest.mark.gpu


def _scan(ctx, name, dev_in, dev_out, dev_tot, n, in_off=0, out_off=0):
    _abi.call(name, ctx.h, C.c_void_p(dev_in.p.value + 4 * in_off), C.c_void_p(dev_out.p.value + 4 * out_off), n,
              dev_tot.p)
    return dev_out.download(np.uint32, out_off + n)[out_off:], int(dev_tot.download(np.uint32, 1)[0])


@pytest.mark.parametrize("sizes,offset", [((5_000_001, 70_001), 0), ((20_000_000, 16_384), 0),
                                          ((300_003, 4_097), 1)])
def test_noalloc_scan_equals_scan(ctx, sizes, offset):
    rng = np.random.default_rng(sizes[0])
    n0 = sizes[0]
    data = rng.integers(0, 1 << 12, n0 + 1, dtype=np.uint32)
    dev_in = ctx.upload(data)
    out_a = ctx.alloc(4 * (n0 + 1))
    out_b = ctx.alloc(4 * (n0 + 1))
    tot_a, tot_b = ctx.alloc(4), ctx.alloc(4)
    for n in sizes:
        print(f"noalloc scan: n={n} offset={offset}", flush=True)
        ga, ta = _scan(ctx, "onc_exscan_u32", dev_in, out_a, tot_a, n, offset, offset)
        gb, tb = _scan(ctx, "onc_exscan_u32_noalloc", dev_in, out_b, tot_b, n, offset, offset)
        np.testing.assert_array_equal(gb, ga)
        assert tb == ta
        x = data[offset:offset + n].astype(np.uint64)
        assert ta == int(x.sum()) & 0xFFFFFFFF and ga[0] == 0 and ga[-1] == int(x[:-1].sum()) & 0xFFFFFFFF
    for b in (dev_in, out_a, out_b, tot_a, tot_b):
        b.free()


def test_noalloc_scan_limits(ctx):
    tot = ctx.alloc(4)
    tot.write(np.array([7], np.uint32))
    _abi.call("onc_exscan_u32_noalloc", ctx.h, None, None, 0, tot.p)
    assert int(tot.download(np.uint32, 1)[0]) == 0
    with pytest.raises(_abi.OncError, match="2\\^30"):
        _abi.call("onc_exscan_u32_noalloc", ctx.h, tot.p, tot.p, (1 << 30) + 1, None)
    tot.free()
