"""The C ABI library loads and exports every entry point include/onco.h
declares; descriptor layouts agree between C and ctypes. No GPU calls."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from oncolytica_amd.hip import _abi
from oncolytica_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "onco.h")


def header_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\*|int)\s+(onc_\w+)\s*\(", txt, re.M)))


def test_header_and_binding_agree():
    assert header_symbols() == _abi.declared_symbols()


def test_library_exports_every_declared_symbol():
    L = _abi.lib()
    for name in header_symbols():
        assert hasattr(L, name), name
    assert L.onc_abi_version() == 1


def test_struct_layouts_match_c():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "onco.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(onc_field_desc), sizeof(onc_sim_desc), sizeof(onc_device_info),
         offsetof(onc_sim_desc, fields), offsetof(onc_sim_desc, col_types), offsetof(onc_sim_desc, params));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "l.c")
        open(p, "w").write(src)
        exe = os.path.join(d, "l")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), p, "-o", exe], check=True)
        got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [C.sizeof(_abi.FieldDesc), C.sizeof(_abi.SimDesc), C.sizeof(_abi.DeviceInfo),
                   _abi.SimDesc.fields.offset, _abi.SimDesc.col_types.offset, _abi.SimDesc.params.offset]


def test_footprint_and_validation_without_device():
    d = W.large_c2().desc()
    b = _abi.footprint(d)
    # two state buffers + daughters (4 x 6 cols x 4 B), sort keys/vals/flags, packed
    # float4 positions, two
    # fields double-buffered, a 2^24-entry {start, len << 16 | tag} bin table,
    # one 4-byte uptake stamp per 8^3 cube (2^24 / 512 of them)
    cap = 2_000_000
    expect = (4 * 6 * cap * 4 + 5 * cap * 4 + 16 * cap + 2 * 2 * 256 ** 3 * 4 + 8 * 2 ** 24 + 4 * 2 ** 15
              + 256)
    assert b == expect
    bad = W.plain_c1().desc()
    bad.interaction_radius = 2.0
    with pytest.raises(_abi.OncError) as e:
        _abi.footprint(bad)
    assert e.value.code == _abi.ERR_INVALID and "interaction_radius" in str(e.value)
    bad = W.plain_c1().desc()
    bad.dims[0] = 2000
    with pytest.raises(_abi.OncError):
        _abi.footprint(bad)


def test_bundle_descriptor_round_trip_and_format_check(tmp_path):
    """A sim descriptor survives config.json exactly (binary32 constants
    through JSON doubles), and a directory with another format is refused
    before any device work."""
    import json
    from oncolytica_amd import bundle as B
    from oncolytica_amd import workloads as W
    d = W.large_c2(n_cells=1000).desc()
    d2 = B.desc_from_dict(json.loads(json.dumps(B.desc_to_dict(d))))
    assert bytes(d) == bytes(d2)
    (tmp_path / "config.json").write_text(json.dumps({"format": "something-else", "version": 1}))
    with pytest.raises(ValueError, match="bundle"):
        B.load_bundle(None, str(tmp_path))


def test_new_entry_points_reject_null_handles():
    """The metric, bundle, forces, timing-sample and pinned-memory entry
    points refuse NULL handles with ONC_ERR_INVALID before any device work."""
    L = _abi.lib()
    out = (C.c_uint32 * 4)()
    p = C.c_void_p()
    assert L.onc_sim_metrics(None, None, 4, out) == _abi.ERR_INVALID
    assert L.onc_sim_set_step(None, 3) == _abi.ERR_INVALID
    assert L.onc_forces(None, None, None, None, None, 0, None, None, None) == _abi.ERR_INVALID
    assert L.onc_batch_profile_every(None, 1, 2) == _abi.ERR_INVALID
    assert L.onc_host_alloc(None, 64, C.byref(p)) == _abi.ERR_INVALID
    assert L.onc_host_free(None, None) == _abi.ERR_INVALID
    assert b"NULL" in L.onc_last_error()


def test_bundle_layout_has_the_reference_launch_config_keys():
    """The bundle's `layout` carries every key of the reference's launch
    config layout (framework/tests/launch_config.json, written for a
    50^3 grid and 10,000 agents), with the same shapes; strides are those of
    this build's 32-bit columns."""
    import json
    from oncolytica_amd import bundle as B
    with open(os.path.join(ROOT, "tests", "golden", "launch_layout.json")) as f:
        ref_layout = json.load(f)
    d = W.plain_c1().desc()
    d.dims[:] = [50, 50, 50]
    d.capacity = 10000
    d.n_cols = 5            # pos.x, pos.y, pos.z, _id and the model's `id`
    d.n_fields = 1
    d.n_chem_cols = 1
    lay = B.layout(d, n_metrics=1)
    assert set(ref_layout) <= set(lay)
    for k in ("tissue_shape", "chemistry_shape", "max_agents"):
        assert lay[k] == ref_layout[k], k
    assert (lay["cell_stride"], lay["tissue_stride"], lay["chem_stride"], lay["metric_stride"]) == (20, 4, 4, 4)
