"""TEST INFRASTRUCTURE ONLY: ctypes binding of the oracle (liboracle.so).

The oracle is the CPU restatement of fmm-rainbow's p=1 FMM step in
rainbow_oracle.cc.  Only tests/, the smoke check and bench.py's CPU-baseline
leg use it, as the checker; the product never imports it.

Arrays use the reference layouts: particles float32[N, 8], moments
int32[C, 4], fields float32[C, 12] (float64[C, 12] for the f64 twin).
Params slots are the 16 x 256-byte table of native/main.rs:89-128.
"""

from __future__ import annotations

import contextlib
import ctypes
import struct
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "liboracle.so"

_u32, _u64, _i32, _f32, _f64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float, ctypes.c_double
_vp, _sz = ctypes.c_void_p, ctypes.c_size_t
_P = ctypes.POINTER

_SIGNATURES = {
    "oracle_set_threads": (None, [_i32]),
    "oracle_threads": (_i32, []),
    "oracle_plan": (None, [_u32, _P(_u32), _P(_u32), _P(_f32)]),
    "oracle_build_plans": (_u32, [_u32, _vp]),
    "oracle_fnv1a": (_u32, [_vp, _sz]),
    "oracle_position_cell": (_u32, [_f32, _u32]),
    "oracle_clear_cells": (None, [_vp, _vp]),
    "oracle_deposit": (None, [_vp, _vp, _vp]),
    "oracle_upward": (None, [_vp, _vp]),
    "oracle_m2l_all": (None, [_vp, _vp, _vp]),
    "oracle_downward": (None, [_vp, _vp, _vp]),
    "oracle_near_all": (None, [_vp, _vp, _vp]),
    "oracle_integrate": (None, [_vp, _vp, _vp, _vp]),
    "oracle_solve_fields": (None, [_vp, _vp, _vp, _vp, _i32]),
    "oracle_step": (None, [_vp, _vp, _vp, _vp, _u32, _i32]),
    "oracle_step_timed": (None, [_vp, _vp, _vp, _vp, _u32, _vp]),
    "oracle_solve_fields_f64": (None, [_vp, _vp, _vp, _vp]),
    "oracle_accelerations": (None, [_vp, _vp, _vp, _vp, _vp, _i32]),
    "oracle_accelerations_f64": (None, [_vp, _vp, _vp, _vp, _vp]),
    "oracle_leaf_cells": (None, [_vp, _vp, _vp, _vp]),
    "oracle_m2l_sources": (_u32, [_u32, _vp]),
    "oracle_near_sources": (_u32, [_u32, _vp]),
    "oracle_pair_field": (None, [_vp, _vp, _vp, _f32, _vp]),
    "oracle_direct_sum_f64": (None, [_u32, _vp, _f64, _f64, _u32, _vp, _vp]),
    "oracle_conserved_f64": (None, [_u32, _vp, _vp]),
    "oracle_seed_galaxy": (None, [_u32, _f32, _vp]),
    "oracle_hash32": (_u32, [_u32]),
    "oracle_enclosed_mass": (_f32, [_f32]),
    "oracle_comp_sum_f32": (_f32, [_vp, _u32, _f32]),
}

_lib = None


def _load(path: Path) -> ctypes.CDLL:
    h = ctypes.CDLL(str(path))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(h, name)
        fn.restype = res
        fn.argtypes = args
    return h


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {ORACLE_DIR}`")
        _lib = _load(LIB_PATH)
    return _lib


@contextlib.contextmanager
def use_library(path):
    """Temporarily route every call through another build of the oracle
    (the mutation tests load builds with a planted mistake)."""
    global _lib
    saved = lib()
    _lib = _load(Path(path))
    try:
        yield
    finally:
        _lib = saved


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def set_threads(n: int):
    lib().oracle_set_threads(n)


def threads() -> int:
    return lib().oracle_threads()


def plan(n: int):
    d, c, s = _u32(), _u32(), _f32()
    lib().oracle_plan(n, ctypes.byref(d), ctypes.byref(c), ctypes.byref(s))
    return d.value, c.value, s.value


# Byte offsets in a Params block (src/rainbow_fmm.wgsl:18-50).
_OFF = {"dt": 32, "gravity": 36, "softening": 40, "speed": 76, "control_point": 80,
        "control_axis_mode": 96, "control_settings": 112}


def build_plans(n: int, dt=None, gravity=None, softening=None, speed=None, control=None,
                scale=None) -> np.ndarray:
    """Slot table of native/main.rs:89-128 with the physics (and optionally
    the fixed-point scale, control_settings.z) overridden."""
    slots = np.zeros(4096, dtype=np.uint8)
    depth = lib().oracle_build_plans(n, _p(slots))
    raw = bytearray(slots.tobytes())
    used = list(range(depth + 1)) + [8 + l for l in range(depth)]
    for s in used:
        base = 256 * s
        for name, v in (("dt", dt), ("gravity", gravity), ("softening", softening), ("speed", speed)):
            if v is not None:
                raw[base + _OFF[name]:base + _OFF[name] + 4] = struct.pack("<f", v)
        if control is not None:
            point, axis_mode, strength, radius = control
            raw[base + 80:base + 96] = struct.pack("<4f", *point)
            raw[base + 96:base + 112] = struct.pack("<4f", *axis_mode)
            raw[base + 112:base + 120] = struct.pack("<2f", strength, radius)
        if scale is not None:
            raw[base + 120:base + 124] = struct.pack("<f", scale)
    return np.frombuffer(bytes(raw), dtype=np.uint8).copy()


def slot(slots: np.ndarray, index: int) -> np.ndarray:
    return np.ascontiguousarray(slots[256 * index:256 * index + 128])


def depth_of(slots: np.ndarray) -> int:
    return int(np.frombuffer(slots[4:8].tobytes(), dtype=np.uint32)[0])


def cells_of(slots: np.ndarray) -> int:
    return int(np.frombuffer(slots[28:32].tobytes(), dtype=np.uint32)[0])


def fnv1a(data) -> int:
    b = bytes(data) if not isinstance(data, np.ndarray) else np.ascontiguousarray(data).tobytes()
    return lib().oracle_fnv1a(b, len(b))


def position_cell(p: float, dim: int) -> int:
    return lib().oracle_position_cell(p, dim)


def solve_fields(slots: np.ndarray, particles: np.ndarray, compensated: bool = False,
                 fields: np.ndarray | None = None):
    """One field solve; returns (moments, fields).  `fields` is the prior
    field buffer (zero if omitted), as the reference never clears it."""
    c = cells_of(slots)
    m = np.zeros((c, 4), dtype=np.int32)
    f = np.zeros((c, 12), dtype=np.float32) if fields is None else np.ascontiguousarray(fields).copy()
    parts = np.ascontiguousarray(particles, dtype=np.float32)
    lib().oracle_solve_fields(_p(slots), _p(parts), _p(m), _p(f), 1 if compensated else 0)
    return m, f


def step(slots: np.ndarray, particles: np.ndarray, steps: int, compensated: bool = False):
    """`steps` full steps from zeroed moments/fields; returns (particles, moments, fields)."""
    c = cells_of(slots)
    m = np.zeros((c, 4), dtype=np.int32)
    f = np.zeros((c, 12), dtype=np.float32)
    parts = np.ascontiguousarray(particles, dtype=np.float32).copy()
    lib().oracle_step(_p(slots), _p(parts), _p(m), _p(f), steps, 1 if compensated else 0)
    return parts, m, f


PHASES = ("clear", "deposit", "upward", "m2l", "downward", "near", "integrate")


def step_timed(slots: np.ndarray, particles: np.ndarray, steps: int):
    """`step` (plain sums) with the wall seconds of each phase, summed over
    the steps: returns (particles, {phase: seconds})."""
    c = cells_of(slots)
    m = np.zeros((c, 4), dtype=np.int32)
    f = np.zeros((c, 12), dtype=np.float32)
    parts = np.ascontiguousarray(particles, dtype=np.float32).copy()
    sec = np.zeros(len(PHASES), dtype=np.float64)
    lib().oracle_step_timed(_p(slots), _p(parts), _p(m), _p(f), steps, _p(sec))
    return parts, dict(zip(PHASES, sec.tolist()))


def solve_fields_f64(slots: np.ndarray, particles: np.ndarray):
    c = cells_of(slots)
    m = np.zeros((c, 4), dtype=np.int32)
    f = np.zeros((c, 12), dtype=np.float64)
    parts = np.ascontiguousarray(particles, dtype=np.float32)
    lib().oracle_solve_fields_f64(_p(slots), _p(parts), _p(m), _p(f))
    return m, f


def accelerations(slots: np.ndarray, particles, moments, fields, compensated=False) -> np.ndarray:
    d = depth_of(slots)
    parts = np.ascontiguousarray(particles, dtype=np.float32)
    out = np.zeros((parts.shape[0], 4), dtype=np.float32)
    lib().oracle_accelerations(_p(slot(slots, d)), _p(parts), _p(np.ascontiguousarray(moments)),
                               _p(np.ascontiguousarray(fields)), _p(out), 1 if compensated else 0)
    return out


def accelerations_f64(slots: np.ndarray, particles, moments, fields64) -> np.ndarray:
    d = depth_of(slots)
    parts = np.ascontiguousarray(particles, dtype=np.float32)
    out = np.zeros((parts.shape[0], 4), dtype=np.float64)
    lib().oracle_accelerations_f64(_p(slot(slots, d)), _p(parts), _p(np.ascontiguousarray(moments)),
                                   _p(np.ascontiguousarray(fields64)), _p(out))
    return out


def leaf_cells(slots: np.ndarray, particles, moments) -> np.ndarray:
    d = depth_of(slots)
    parts = np.ascontiguousarray(particles, dtype=np.float32)
    out = np.zeros(parts.shape[0], dtype=np.uint32)
    lib().oracle_leaf_cells(_p(slot(slots, d)), _p(parts), _p(np.ascontiguousarray(moments)), _p(out))
    return out


def m2l_sources(g: int) -> np.ndarray:
    buf = np.zeros(256, dtype=np.uint32)
    n = lib().oracle_m2l_sources(g, _p(buf))
    return buf[:n].copy()


def near_sources(g: int) -> np.ndarray:
    buf = np.zeros(32, dtype=np.uint32)
    n = lib().oracle_near_sources(g, _p(buf))
    return buf[:n].copy()


def pair_field(slot_bytes: np.ndarray, moment, target, epsilon: float) -> np.ndarray:
    m = np.ascontiguousarray(moment, dtype=np.int32).reshape(4)
    t = np.ascontiguousarray(target, dtype=np.float32).reshape(3)
    out = np.zeros(12, dtype=np.float32)
    lib().oracle_pair_field(_p(slot_bytes), _p(m), _p(t), epsilon, _p(out))
    return out


def direct_sum_f64(particles, gravity: float, softening: float, targets) -> np.ndarray:
    parts = np.ascontiguousarray(particles, dtype=np.float32)
    t = np.ascontiguousarray(targets, dtype=np.uint32)
    out = np.zeros((t.shape[0], 4), dtype=np.float64)
    lib().oracle_direct_sum_f64(parts.shape[0], _p(parts), gravity, softening, t.shape[0], _p(t), _p(out))
    return out


def conserved_f64(particles) -> np.ndarray:
    parts = np.ascontiguousarray(particles, dtype=np.float32)
    out = np.zeros(8, dtype=np.float64)
    lib().oracle_conserved_f64(parts.shape[0], _p(parts), _p(out))
    return out


def seed_galaxy(n: int, gravity: float) -> np.ndarray:
    """initialize (src/rainbow_fmm.wgsl:232-374): particles float32[n, 8]."""
    out = np.zeros((n, 8), dtype=np.float32)
    lib().oracle_seed_galaxy(n, gravity, _p(out))
    return out


def hash32(x: int) -> int:
    return int(lib().oracle_hash32(x))


def enclosed_mass(r: float) -> float:
    return float(lib().oracle_enclosed_mass(r))


def comp_sum_f32(terms, init: float = 0.0) -> np.float32:
    """The Neumaier accumulator of the compensated solve over f32 terms."""
    t = np.ascontiguousarray(terms, dtype=np.float32)
    return np.float32(lib().oracle_comp_sum_f32(_p(t), t.shape[0], init))
