"""ctypes binding of librainbow_fmm.so, the gfx950 p=1 FMM step.

The planner functions mirror the reference's wasm exports (src/lib.rs:88-157)
and the harness helpers (native/main.rs:38-163); ``Solver`` drives one
system the way native/main.rs drives the WGSL passes; ``Batch`` and
``Server`` are the batched and continuous-batching forms.

Buffers are numpy arrays in the reference layout:
particles ``float32[N, 8]`` ({pos.xyz, hue}, {vel.xyz, excitation}),
moments ``int32[C, 4]`` ({count, sum_x, sum_y, sum_z}),
fields ``float32[C, 12]`` ({a.xyz, phi}, {Jxx, Jxy, Jxz, Jyy}, {Jyz, Jzz, 0, 0}).

The library must have been built (``make -C fmm-rainbow_amd``); there is no
fallback: every call goes to the native library, and a missing library
raises at import of the first call.
"""

from __future__ import annotations

import ctypes
import os
import re
from pathlib import Path

import numpy as np

PACKAGE_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = PACKAGE_DIR / "lib" / "librainbow_fmm.so"
HEADER_PATH = PACKAGE_DIR.parent / "include" / "rainbow_fmm.h"

PARTICLES, MOMENTS, FIELDS = 0, 1, 2
OPT_COMPENSATED = 1
REQUEST_BORROW = 1
OK, E_INVALID, E_HIP, E_NOMEM, E_BUDGET, E_STATE, E_CANCELLED, E_TIMEOUT = 0, -1, -2, -3, -4, -5, -6, -7


class RainbowError(RuntimeError):
    """A negative status from the library, with rainbow_last_error()."""

    def __init__(self, code: int, message: str):
        super().__init__(f"rainbow error {code}: {message}")
        self.code = code


class Params(ctypes.Structure):
    """struct Params of src/rainbow_fmm.wgsl:18-50 (128 bytes)."""

    _fields_ = [
        ("particle_count", ctypes.c_uint32),
        ("depth", ctypes.c_uint32),
        ("level", ctypes.c_uint32),
        ("level_offset", ctypes.c_uint32),
        ("level_dim", ctypes.c_uint32),
        ("parent_offset", ctypes.c_uint32),
        ("parent_dim", ctypes.c_uint32),
        ("total_cells", ctypes.c_uint32),
        ("dt", ctypes.c_float),
        ("gravity", ctypes.c_float),
        ("softening", ctypes.c_float),
        ("point_size", ctypes.c_float),
        ("time", ctypes.c_float),
        ("aspect", ctypes.c_float),
        ("camera_yaw", ctypes.c_float),
        ("camera_pitch", ctypes.c_float),
        ("viewport_width", ctypes.c_float),
        ("viewport_height", ctypes.c_float),
        ("exposure", ctypes.c_float),
        ("speed", ctypes.c_float),
        ("control_point", ctypes.c_float * 4),
        ("control_axis_mode", ctypes.c_float * 4),
        ("control_settings", ctypes.c_float * 4),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("particles", ctypes.c_uint32),
        ("steps", ctypes.c_uint32),
        ("data", ctypes.c_void_p),
        ("timeout_s", ctypes.c_double),
        ("flags", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class Completion(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("particles", ctypes.c_uint32),
        ("steps", ctypes.c_uint32),
        ("submit_s", ctypes.c_double),
        ("admit_s", ctypes.c_double),
        ("finish_s", ctypes.c_double),
        ("status", ctypes.c_int32),
        ("steps_run", ctypes.c_uint32),
    ]


_u32, _u64, _i32, _f32, _f64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float, ctypes.c_double
_vp, _sz = ctypes.c_void_p, ctypes.c_size_t
_P = ctypes.POINTER

_SIGNATURES = {
    "rainbow_api_version": (_u32, []),
    "rainbow_shader_ptr": (_vp, []),
    "rainbow_shader_len": (_sz, []),
    "rainbow_algorithm_ptr": (_vp, []),
    "rainbow_algorithm_len": (_sz, []),
    "rainbow_depth_for_particles": (_u32, [_u32]),
    "rainbow_level_cells": (_u32, [_u32]),
    "rainbow_level_offset": (_u32, [_u32]),
    "rainbow_total_cells": (_u32, [_u32]),
    "rainbow_clamp_particle_count": (_u32, [_u32]),
    "rainbow_fixed_point_scale": (_f32, [_u32]),
    "rainbow_linear_work_lo": (_u32, [_u32, _u32]),
    "rainbow_plan_depth": (_u32, [_u32]),
    "rainbow_plan_fixed_point_scale": (_f32, [_u32]),
    "rainbow_plan_linear_work": (_u64, [_u32, _u32]),
    "rainbow_write_params": (None, [_u32] * 8 + [_f32, _P(Params)]),
    "rainbow_build_plans": (_u32, [_u32, _vp]),
    "rainbow_fnv1a": (_u32, [_vp, _sz]),
    "rainbow_ic_uniform_ball": (_i32, [_u32, _u64, _f64, _vp]),
    "rainbow_ic_plummer": (_i32, [_u32, _u64, _f64, _f64, _f64, _vp]),
    "rainbow_ic_clumps": (_i32, [_u32, _u64, _u32, _f64, _f64, _f64, _vp]),
    "rainbow_hip_create": (_i32, [_u32, _P(Params), _P(_vp)]),
    "rainbow_hip_create_ex": (_i32, [_u32, _P(Params), _u32, _P(_vp)]),
    "rainbow_hip_destroy": (None, [_vp]),
    "rainbow_hip_set_params": (_i32, [_vp, _P(Params)]),
    "rainbow_hip_plan": (_i32, [_vp, _P(_u32), _P(_u32), _P(_f32)]),
    "rainbow_hip_upload_particles": (_i32, [_vp, _vp, _sz]),
    "rainbow_hip_seed_galaxy": (_i32, [_vp]),
    "rainbow_hip_step_frame": (_i32, [_vp, ctypes.c_double]),
    "rainbow_frame_dt": (_f32, [ctypes.c_double]),
    "rainbow_hip_solve_fields": (_i32, [_vp]),
    "rainbow_hip_step": (_i32, [_vp, _u32]),
    "rainbow_hip_synchronize": (_i32, [_vp]),
    "rainbow_hip_download": (_i32, [_vp, _i32, _vp, _sz]),
    "rainbow_hip_digest": (_i32, [_vp, _i32, _P(_u32)]),
    "rainbow_hip_time_steps": (_i32, [_vp, _u32, _P(_f32)]),
    "rainbow_hip_phase_ms": (_i32, [_vp, _u32, _P(_f32), _u32]),
    "rainbow_hip_set_tree_fusion": (_i32, [_vp, _u32]),
    "rainbow_hip_set_tree_split": (_i32, [_vp, _u32]),
    "rainbow_batch_set_tree_split": (_i32, [_vp, _u32]),
    "rainbow_hip_direct_sum_f64": (_i32, [_vp, _P(_u32), _u32, _P(_f64)]),
    "rainbow_hip_direct_sum_f64_noalloc": (_i32, [_vp, _P(_u32), _u32, _P(_f64)]),
    "rainbow_hip_tree_launches": (_i32, [_vp, _P(_u32), _u32]),
    "rainbow_batch_set_tree_fusion": (_i32, [_vp, _u32]),
    "rainbow_hip_tree_stats": (_i32, [_vp, _P(_u32)]),
    "rainbow_hip_pair_counts": (_i32, [_vp, _P(_u64)]),
    "rainbow_hip_set_sort_interval": (_i32, [_vp, _u32]),
    "rainbow_hip_autotune": (_i32, [_vp, _u32, _P(_u32)]),
    "rainbow_hip_set_shape": (_i32, [_vp, _u32, _u32]),
    "rainbow_hip_get_shape": (_i32, [_vp, _P(_u32), _P(_u32)]),
    "rainbow_batch_set_shape": (_i32, [_vp, _u32, _u32]),
    "rainbow_batch_get_shape": (_i32, [_vp, _P(_u32), _P(_u32)]),
    "rainbow_batch_set_sort_interval": (_i32, [_vp, _u32]),
    "rainbow_batch_autotune": (_i32, [_vp, _u32, _P(_u32)]),
    "rainbow_hip_diagnostics": (_i32, [_vp, _P(_f64)]),
    "rainbow_hip_accelerations": (_i32, [_vp, _vp, _sz]),
    "rainbow_last_error": (ctypes.c_char_p, []),
    "rainbow_batch_create": (_i32, [_u32, _P(_u32), _P(Params), _u32, _P(_vp)]),
    "rainbow_batch_destroy": (None, [_vp]),
    "rainbow_batch_upload": (_i32, [_vp, _u32, _vp, _sz]),
    "rainbow_batch_seed_galaxy": (_i32, [_vp, _u32]),
    "rainbow_batch_download": (_i32, [_vp, _u32, _i32, _vp, _sz]),
    "rainbow_batch_solve_fields": (_i32, [_vp]),
    "rainbow_batch_step": (_i32, [_vp, _u32]),
    "rainbow_batch_time_steps": (_i32, [_vp, _u32, _P(_f32)]),
    "rainbow_batch_synchronize": (_i32, [_vp]),
    "rainbow_batch_particles": (_u64, [_vp]),
    "rainbow_server_create": (_i32, [_u64, _u32, _P(Params), _u32, _P(_vp)]),
    "rainbow_server_destroy": (None, [_vp]),
    "rainbow_server_request_bytes": (_u64, [_u32]),
    "rainbow_server_submit": (_i32, [_vp, _P(Request)]),
    "rainbow_server_poll": (_i32, [_vp, _P(_u32)]),
    "rainbow_server_cancel": (_i32, [_vp, _u64]),
    "rainbow_server_result_view": (_i32, [_vp, _u64, _P(_vp), _P(_sz)]),
    "rainbow_server_release": (_i32, [_vp, _u64]),
    "rainbow_host_alloc": (_vp, [_sz]),
    "rainbow_host_free": (None, [_vp]),
    "rainbow_server_drain": (_i32, [_vp]),
    "rainbow_server_stats": (_i32, [_vp, _P(_f64)]),
    "rainbow_server_completions": (_i32, [_vp, _P(Completion), _u32, _P(_u32)]),
    "rainbow_server_result": (_i32, [_vp, _u64, _vp, _sz]),
    "rainbow_server_state": (_i32, [_vp, _P(_u32), _P(_u32), _P(_u64)]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C {PACKAGE_DIR}` "
                "(there is no fallback implementation)"
            )
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def header_functions() -> list[str]:
    """Names of the functions include/rainbow_fmm.h declares."""
    text = HEADER_PATH.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(rainbow_[a-z0-9_]+)\s*\(", text)
    seen = []
    for n in names:
        if n not in seen:
            seen.append(n)
    return seen


def _check(code: int) -> int:
    if code < 0:
        msg = lib().rainbow_last_error()
        raise RainbowError(code, msg.decode() if msg else "")
    return code


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


# ---- planner (src/lib.rs:88-157) ------------------------------------------


def api_version() -> int:
    return lib().rainbow_api_version()


def shader_payload() -> bytes:
    return ctypes.string_at(lib().rainbow_shader_ptr(), lib().rainbow_shader_len())


def algorithm() -> str:
    return ctypes.string_at(lib().rainbow_algorithm_ptr(), lib().rainbow_algorithm_len()).decode()


def depth_for_particles(n: int) -> int:
    return lib().rainbow_depth_for_particles(n)


def level_cells(level: int) -> int:
    return lib().rainbow_level_cells(level)


def level_offset(level: int) -> int:
    return lib().rainbow_level_offset(level)


def total_cells(depth: int) -> int:
    return lib().rainbow_total_cells(depth)


def clamp_particle_count(n: int) -> int:
    return lib().rainbow_clamp_particle_count(n)


def fixed_point_scale(n: int) -> float:
    return lib().rainbow_fixed_point_scale(n)


def linear_work_lo(n: int, depth: int) -> int:
    return lib().rainbow_linear_work_lo(n, depth)


def plan_depth(n: int) -> int:
    return lib().rainbow_plan_depth(n)


def plan_scale(n: int) -> float:
    return lib().rainbow_plan_fixed_point_scale(n)


def plan_linear_work(n: int, depth: int) -> int:
    return lib().rainbow_plan_linear_work(n, depth)


def plan_cells(n: int) -> int:
    d = plan_depth(n)
    return (8 ** (d + 1) - 1) // 7


# ---- harness helpers (native/main.rs:38-163) --------------------------------


def write_params(count, depth, level, level_offset_, level_dim, parent_offset, parent_dim, cells, scale) -> Params:
    p = Params()
    lib().rainbow_write_params(count, depth, level, level_offset_, level_dim, parent_offset, parent_dim, cells,
                               scale, ctypes.byref(p))
    return p


def build_plans(count: int) -> bytes:
    buf = (ctypes.c_uint8 * 4096)()
    lib().rainbow_build_plans(count, buf)
    return bytes(buf)


def frame_dt(frame_interval_ms: float) -> float:
    """The browser's per-frame dt (app.js:42-43, 716-719)."""
    return float(lib().rainbow_frame_dt(frame_interval_ms))


def fnv1a(data) -> int:
    b = bytes(data) if not isinstance(data, np.ndarray) else np.ascontiguousarray(data).tobytes()
    return lib().rainbow_fnv1a(b, len(b))


def physics(dt=0.0009, gravity=0.34, softening=0.012, speed=1.0, scale=0.0, control=None) -> Params:
    """A Params block carrying the physics of a run (harness defaults)."""
    p = write_params(0, 0, 0, 0, 1, 0, 1, 0, scale)
    p.dt, p.gravity, p.softening, p.speed = dt, gravity, softening, speed
    if control is not None:
        point, axis_mode, strength, radius = control
        for k in range(4):
            p.control_point[k] = point[k]
            p.control_axis_mode[k] = axis_mode[k]
        p.control_settings[0] = strength
        p.control_settings[1] = radius
    return p


# ---- initial conditions -----------------------------------------------------


def ic_uniform_ball(n: int, seed: int, radius: float = 1.0) -> np.ndarray:
    out = np.zeros((n, 8), dtype=np.float32)
    _check(lib().rainbow_ic_uniform_ball(n, seed, radius, _ptr(out)))
    return out


def ic_plummer(n: int, seed: int, a: float = 0.1, rmax: float = 1.5, gravity: float = 1.0) -> np.ndarray:
    out = np.zeros((n, 8), dtype=np.float32)
    _check(lib().rainbow_ic_plummer(n, seed, a, rmax, gravity, _ptr(out)))
    return out


def ic_clumps(n: int, seed: int, clumps: int = 64, a: float = 0.02, rcut_over_a: float = 10.0,
              gravity: float = 1.0) -> np.ndarray:
    out = np.zeros((n, 8), dtype=np.float32)
    _check(lib().rainbow_ic_clumps(n, seed, clumps, a, rcut_over_a, gravity, _ptr(out)))
    return out


# ---- solver -------------------------------------------------------------------


class Solver:
    """One system on the GPU (rainbow_hip_*)."""

    def __init__(self, n: int, params: Params | None = None, compensated: bool = False):
        self.n = n
        h = ctypes.c_void_p()
        opts = OPT_COMPENSATED if compensated else 0
        _check(lib().rainbow_hip_create_ex(n, ctypes.byref(params) if params is not None else None, opts,
                                           ctypes.byref(h)))
        self._h = h
        d, c, s = _u32(), _u32(), _f32()
        _check(lib().rainbow_hip_plan(self._h, ctypes.byref(d), ctypes.byref(c), ctypes.byref(s)))
        self.depth, self.cells, self.scale = d.value, c.value, s.value

    def close(self):
        if self._h:
            lib().rainbow_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_params(self, params: Params):
        _check(lib().rainbow_hip_set_params(self._h, ctypes.byref(params)))

    def upload(self, particles: np.ndarray):
        a = np.ascontiguousarray(particles, dtype=np.float32)
        assert a.shape == (self.n, 8)
        _check(lib().rainbow_hip_upload_particles(self._h, _ptr(a), a.nbytes))

    def step_frame(self, frame_interval_ms: float):
        """One step of a browser frame: dt from the frame interval."""
        _check(lib().rainbow_hip_step_frame(self._h, frame_interval_ms))

    def seed_galaxy(self):
        """The reference's default scene (its initialize pass) on the device,
        in place of an upload."""
        _check(lib().rainbow_hip_seed_galaxy(self._h))

    def solve_fields(self):
        _check(lib().rainbow_hip_solve_fields(self._h))

    def step(self, steps: int = 1):
        _check(lib().rainbow_hip_step(self._h, steps))

    def sync(self):
        _check(lib().rainbow_hip_synchronize(self._h))

    def download(self, which: int, out: np.ndarray | None = None) -> np.ndarray:
        """Copies a buffer out, into `out` when given (e.g. a pinned
        host_array, for a DMA at full bus speed)."""
        shape, dtype = {PARTICLES: ((self.n, 8), np.float32), MOMENTS: ((self.cells, 4), np.int32)}.get(
            which, ((self.cells, 12), np.float32))
        if out is None:
            out = np.empty(shape, dtype=dtype)
        elif out.shape != shape or out.dtype != dtype or not out.flags["C_CONTIGUOUS"]:
            raise ValueError(f"out must be a contiguous {np.dtype(dtype).name}{list(shape)} array")
        _check(lib().rainbow_hip_download(self._h, which, _ptr(out), out.nbytes))
        return out

    def particles(self, out: np.ndarray | None = None):
        return self.download(PARTICLES, out)

    def moments(self):
        return self.download(MOMENTS)

    def fields(self):
        return self.download(FIELDS)

    def digest(self, which: int) -> int:
        v = _u32()
        _check(lib().rainbow_hip_digest(self._h, which, ctypes.byref(v)))
        return v.value

    def time_steps(self, steps: int) -> float:
        ms = _f32()
        _check(lib().rainbow_hip_time_steps(self._h, steps, ctypes.byref(ms)))
        return ms.value

    def phase_ms(self, steps: int) -> list[float]:
        buf = (_f32 * 16)()
        n = _check(lib().rainbow_hip_phase_ms(self._h, steps, buf, 16))
        return [buf[i] for i in range(n)]

    def set_sort_interval(self, interval: int):
        _check(lib().rainbow_hip_set_sort_interval(self._h, interval))

    def direct_sum_f64(self, targets) -> np.ndarray:
        """FP64 direct-sum {ax, ay, az, phi} at the given bodies (caller indices)."""
        t = np.ascontiguousarray(targets, dtype=np.uint32)
        out = np.zeros((t.shape[0], 4), dtype=np.float64)
        _check(lib().rainbow_hip_direct_sum_f64(self._h, t.ctypes.data_as(_P(_u32)), t.shape[0],
                                                out.ctypes.data_as(_P(_f64))))
        return out

    def set_tree_fusion(self, hi: int):
        """Tree levels 2..hi in one launch; each deeper level alone."""
        _check(lib().rainbow_hip_set_tree_fusion(self._h, hi))

    def set_tree_split(self, starts: int):
        """Bit l of `starts` begins a tree launch at level l (bit 2 implied);
        starts=1 selects the two-launch form (M2L sums, then fields)."""
        _check(lib().rainbow_hip_set_tree_split(self._h, starts))

    def tree_launches(self) -> list[tuple[int, int]]:
        """(lo, hi) levels of each tree launch of a step, in order."""
        return [(lo, hi) for _, lo, hi in self.tree_launch_forms()]

    def tree_launch_forms(self) -> list[tuple[str, int, int]]:
        """(form, lo, hi) of each tree launch: form "run" (M2L, L2L and near
        of levels lo..hi), "m2l" (the M2L sums below the finest level),
        "m2l_top" (the same for one system, also running the top levels and
        the levels 2..depth-3; lo..hi are the levels upward lists) or
        "finish" (every level's fields from those sums)."""
        buf = (_u32 * 16)()
        n = _check(lib().rainbow_hip_tree_launches(self._h, buf, 16))
        forms = {0: "run", 1: "m2l", 2: "finish", 3: "m2l_top"}
        return [(forms[buf[2 * i] >> 8], buf[2 * i] & 0xFF, buf[2 * i + 1]) for i in range(n)]

    def phase_names(self) -> list[str]:
        """Names of the entries phase_ms returns."""
        tree = []
        for form, lo, hi in self.tree_launch_forms():
            levels = f"l{lo}" if lo == hi else f"l{lo}-{hi}"
            tree.append(f"tree_{levels}" if form == "run" else f"tree_{form}_{levels}")
        return ["deposit", "upward", "top"] + tree + ["integrate", "sort"]

    def set_shape(self, integrate_threads: int = 1024, tree_chunks: int = 8):
        """Block shapes: integrate+deposit block width (1024 or 512 threads
        over a 1024-body range) and the tree's pair tile (8 or 4 chunks of
        8 x 8 pairs per round; 1: an item per wave; 0: chosen per launch
        from the measured items)."""
        _check(lib().rainbow_hip_set_shape(self._h, integrate_threads, tree_chunks))

    @property
    def shape(self) -> tuple[int, int]:
        t, c = _u32(), _u32()
        _check(lib().rainbow_hip_get_shape(self._h, ctypes.byref(t), ctypes.byref(c)))
        return t.value, c.value

    def autotune(self, steps: int = 32) -> int:
        """Times block shapes, tree splits and sort intervals; keeps the
        fastest of each and returns the sort interval (see `shape`)."""
        v = _u32()
        _check(lib().rainbow_hip_autotune(self._h, steps, ctypes.byref(v)))
        return v.value

    def pair_counts(self) -> tuple[list[int], list[int]]:
        """(M2L pairs per level, near pairs per level) of the last solve."""
        buf = (_u64 * 16)()
        _check(lib().rainbow_hip_pair_counts(self._h, buf))
        return [buf[i] for i in range(8)], [buf[8 + i] for i in range(8)]

    def tree_stats(self) -> list[int]:
        buf = (_u32 * 8)()
        _check(lib().rainbow_hip_tree_stats(self._h, buf))
        return list(buf)

    def diagnostics(self) -> np.ndarray:
        buf = (_f64 * 8)()
        _check(lib().rainbow_hip_diagnostics(self._h, buf))
        return np.array(list(buf))

    def accelerations(self) -> np.ndarray:
        out = np.empty((self.n, 4), dtype=np.float32)
        _check(lib().rainbow_hip_accelerations(self._h, _ptr(out), out.nbytes))
        return out


class Batch:
    """Independent systems advanced together (rainbow_batch_*)."""

    def __init__(self, counts, params: Params | None = None, compensated: bool = False):
        self.counts = [int(c) for c in counts]
        arr = (_u32 * len(self.counts))(*self.counts)
        h = ctypes.c_void_p()
        _check(lib().rainbow_batch_create(len(self.counts), arr, ctypes.byref(params) if params is not None else None,
                                          OPT_COMPENSATED if compensated else 0, ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().rainbow_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, system: int, particles: np.ndarray):
        a = np.ascontiguousarray(particles, dtype=np.float32)
        _check(lib().rainbow_batch_upload(self._h, system, _ptr(a), a.nbytes))

    def seed_galaxy(self, system: int):
        _check(lib().rainbow_batch_seed_galaxy(self._h, system))

    def download(self, system: int, which: int) -> np.ndarray:
        n = self.counts[system]
        cells = plan_cells(n)
        if which == PARTICLES:
            out = np.empty((n, 8), dtype=np.float32)
        elif which == MOMENTS:
            out = np.empty((cells, 4), dtype=np.int32)
        else:
            out = np.empty((cells, 12), dtype=np.float32)
        _check(lib().rainbow_batch_download(self._h, system, which, _ptr(out), out.nbytes))
        return out

    def solve_fields(self):
        _check(lib().rainbow_batch_solve_fields(self._h))

    def step(self, steps: int = 1):
        _check(lib().rainbow_batch_step(self._h, steps))

    def sync(self):
        _check(lib().rainbow_batch_synchronize(self._h))

    def time_steps(self, steps: int) -> float:
        ms = _f32()
        _check(lib().rainbow_batch_time_steps(self._h, steps, ctypes.byref(ms)))
        return ms.value

    def set_sort_interval(self, interval: int):
        _check(lib().rainbow_batch_set_sort_interval(self._h, interval))

    def set_tree_fusion(self, hi: int):
        _check(lib().rainbow_batch_set_tree_fusion(self._h, hi))

    def set_tree_split(self, starts: int):
        _check(lib().rainbow_batch_set_tree_split(self._h, starts))

    def set_shape(self, integrate_threads: int = 1024, tree_chunks: int = 8):
        _check(lib().rainbow_batch_set_shape(self._h, integrate_threads, tree_chunks))

    @property
    def shape(self) -> tuple[int, int]:
        t, c = _u32(), _u32()
        _check(lib().rainbow_batch_get_shape(self._h, ctypes.byref(t), ctypes.byref(c)))
        return t.value, c.value

    def autotune(self, steps: int = 32) -> int:
        v = _u32()
        _check(lib().rainbow_batch_autotune(self._h, steps, ctypes.byref(v)))
        return v.value

    @property
    def particles_total(self) -> int:
        return lib().rainbow_batch_particles(self._h)


class Server:
    """Continuous-batching request loop (rainbow_server_*)."""

    def __init__(self, budget_bytes: int, max_slots: int, params: Params | None = None, compensated: bool = False):
        h = ctypes.c_void_p()
        _check(lib().rainbow_server_create(budget_bytes, max_slots, ctypes.byref(params) if params is not None else None,
                                           OPT_COMPENSATED if compensated else 0, ctypes.byref(h)))
        self._h = h
        self._keep = {}

    def close(self):
        if self._h:
            lib().rainbow_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def request_bytes(n: int) -> int:
        return lib().rainbow_server_request_bytes(n)

    def submit(self, rid: int, particles: np.ndarray, steps: int, timeout_s: float = 0.0, borrow: bool = False):
        """Queues a request.  borrow=True: the server reads `particles` in
        place (no copy); keep the array alive and unchanged until the
        request's completion has been popped (pinned arrays from
        host_array() upload by DMA)."""
        if borrow:
            if particles.dtype != np.float32 or not particles.flags["C_CONTIGUOUS"]:
                raise ValueError("a borrowed request needs a C-contiguous float32 array")
            a = particles
        else:
            a = np.ascontiguousarray(particles, dtype=np.float32)
        r = Request(rid, a.shape[0], steps, _ptr(a), timeout_s, REQUEST_BORROW if borrow else 0, 0)
        _check(lib().rainbow_server_submit(self._h, ctypes.byref(r)))
        if borrow:
            self._keep[rid] = a

    def cancel(self, rid: int):
        _check(lib().rainbow_server_cancel(self._h, rid))

    def drain(self):
        """Polls until every submitted request has completed."""
        _check(lib().rainbow_server_drain(self._h))

    def poll(self) -> int:
        n = _u32()
        _check(lib().rainbow_server_poll(self._h, ctypes.byref(n)))
        return n.value

    def completions(self, cap: int = 1024) -> list[Completion]:
        buf = (Completion * cap)()
        n = _u32()
        _check(lib().rainbow_server_completions(self._h, buf, cap, ctypes.byref(n)))
        out = [buf[i] for i in range(n.value)]
        for c in out:
            self._keep.pop(c.id, None)  # a borrowed request's array is free again
        return out

    def result(self, rid: int, n: int) -> np.ndarray:
        out = np.empty((n, 8), dtype=np.float32)
        _check(lib().rainbow_server_result(self._h, rid, _ptr(out), out.nbytes))
        return out

    def result_view(self, rid: int) -> np.ndarray:
        """The result in the server's pinned buffer, without a copy; valid
        until release(rid)."""
        p, b = _vp(), _sz()
        _check(lib().rainbow_server_result_view(self._h, rid, ctypes.byref(p), ctypes.byref(b)))
        buf = (ctypes.c_float * (b.value // 4)).from_address(p.value)
        return np.frombuffer(buf, dtype=np.float32).reshape(-1, 8)

    def release(self, rid: int):
        _check(lib().rainbow_server_release(self._h, rid))

    def stats(self) -> dict:
        """Cumulative host seconds per scheduling activity, steps and polls."""
        buf = (_f64 * 8)()
        _check(lib().rainbow_server_stats(self._h, buf))
        return {"admit_s": buf[0], "configure_s": buf[1], "step_s": buf[2], "retire_s": buf[3],
                "steps": int(buf[4]), "polls": int(buf[5]), "gpu_step_s": buf[6], "wait_s": buf[7]}

    def state(self):
        q, r, b = _u32(), _u32(), _u64()
        _check(lib().rainbow_server_state(self._h, ctypes.byref(q), ctypes.byref(r), ctypes.byref(b)))
        return q.value, r.value, b.value


class HostArray(np.ndarray):
    """A float32[N, 8] particle array in page-locked host memory from
    rainbow_host_alloc, freed with the array."""

    def __array_finalize__(self, obj):
        self._owner = getattr(obj, "_owner", None)


class _Pinned:
    def __init__(self, nbytes):
        self.ptr = lib().rainbow_host_alloc(nbytes)
        if not self.ptr:
            raise RainbowError(E_NOMEM, f"no pinned host memory for {nbytes} bytes")
        self.nbytes = nbytes

    def __del__(self):
        if self.ptr:
            lib().rainbow_host_free(self.ptr)
            self.ptr = None


def host_array(n: int) -> np.ndarray:
    """An uninitialised float32[n, 8] particle array in pinned host memory
    (for zero-copy server requests)."""
    owner = _Pinned(32 * n)
    buf = (ctypes.c_float * (8 * n)).from_address(owner.ptr)
    a = np.frombuffer(buf, dtype=np.float32).reshape(n, 8).view(HostArray)
    a._owner = owner
    return a


def gpu_available() -> bool:
    """True when a HIP device is visible (checked without initialising HIP)."""
    if os.environ.get("HIP_VISIBLE_DEVICES", None) == "":
        return False
    return Path("/dev/kfd").exists() and any(Path("/dev/dri").glob("renderD*"))
