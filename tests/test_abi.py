"""The C-ABI library: loads, exports every symbol the header declares, and
the Params layout matches struct Params (src/rainbow_fmm.wgsl:18-50).  No
device calls."""

import ctypes

import rainbow_fmm as rf


def test_library_exports_every_declared_symbol():
    lib = rf.lib()
    names = rf.header_functions()
    assert len(names) >= 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_binding_covers_every_declared_symbol():
    assert sorted(rf.header_functions()) == sorted(rf._SIGNATURES)


def test_params_layout():
    assert ctypes.sizeof(rf.Params) == 128
    assert rf.Params.dt.offset == 32
    assert rf.Params.speed.offset == 76
    assert rf.Params.control_point.offset == 80
    assert rf.Params.control_axis_mode.offset == 96
    assert rf.Params.control_settings.offset == 112


def test_write_params_matches_harness_constants():
    # native/main.rs:57-82
    p = rf.write_params(1 << 20, 6, 3, 73, 8, 9, 4, 299593, 1024.0)
    raw = bytes(p)
    import struct

    u = struct.unpack("<8I", raw[:32])
    f = struct.unpack("<24f", raw[32:])
    assert u == (1 << 20, 6, 3, 73, 8, 9, 4, 299593)
    want = [0.0009, 0.34, 0.012, 5.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.84, 1.0,
            0, 0, 0, 0, 0, 0, 0, 0, 0.7, 0.22, 1024.0, 0.0]
    assert list(f) == [struct.unpack("<f", struct.pack("<f", w))[0] for w in want]


def test_plan_table_matches_oracle_restatement():
    import numpy as np
    import oracle as orc

    for n in (1, 1000, 16384, 65536, 262145, 1 << 20, 5242880):
        a = np.frombuffer(rf.build_plans(n), dtype=np.uint8)
        b = orc.build_plans(n)
        assert a.tobytes() == b.tobytes(), n


def test_shader_payload_is_present():
    # src/lib.rs:212-217 checks the WGSL text; the payload here is the
    # compiled kernel source.
    text = rf.shader_payload().decode()
    assert "k_tree" in text and "k_integrate" in text
    assert len(text) > 10000
    assert rf.algorithm() == "Sparse p=1 Cartesian FMM / BLDTT-inspired adaptive hierarchy"
    assert rf.api_version() == 1
