"""ctypes binding of ``libglisser_amd.so`` (the C ABI in include/glisser_amd.h).

The library is loaded from the package folder only.  If it is missing or
cannot be loaded the import fails loudly: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libglisser_amd.so"

GLX_OK, GLX_EINVAL, GLX_EHIP, GLX_ENOMEM, GLX_ESTATE = 0, 1, 2, 3, 4
GLX_MAX_BODIES = 16
FORM_FP64, FORM_FAST, FORM_MIXED = 0, 1, 2
FORMS = {"fp64": FORM_FP64, "fast": FORM_FAST, "mixed": FORM_MIXED}
PART_UNFLAGGED, PART_THREE_WAY = 0, 1
MAX_BODIES = 16

_dp = C.POINTER(C.c_double)
_u16p = C.POINTER(C.c_uint16)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


class GlxError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"glx error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("n_particles", C.c_uint64),
        ("n_bodies_max", C.c_uint32),
        ("tbsize_max", C.c_uint32),
        ("device", C.c_int32),
        ("keep_rollback", C.c_int32),
        ("n_planet_buffers", C.c_uint32),
    ]


class SystemDesc(C.Structure):
    """glx_system: one planetary system of a batch."""
    _fields_ = [
        ("n_bodies", C.c_uint32),
        ("n_particles", C.c_uint32),
        ("t0", C.c_double),
        ("dt", C.c_double),
        ("outer_radius", C.c_double),
        ("m", _dp),
        ("id", _u32p),
        ("rh", _dp),
        ("n_samples", C.c_uint64),
        ("t", _dp),
        ("r", _dp),
        ("v", _dp),
    ]


class BatchConfig(C.Structure):
    """glx_batch_config: launch shape of a batch."""
    _fields_ = [
        ("steps_per_launch", C.c_uint32),
        ("lanes", C.c_uint32),
        ("persistent", C.c_uint32),
        ("waves_per_simd", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


BATCH_MAX_BODIES = 9
STATE_ASCII, STATE_BINARY, STATE_SPLIT = 0, 1, 2


class StateC(C.Structure):
    """glx_state: caller-owned arrays of a state file."""
    _fields_ = [("n_planets", C.c_uint32), ("n_particles", C.c_uint64), ("pm", _dp), ("pr", _dp), ("pv", _dp),
                ("pid", _u32p), ("r", _dp), ("v", _dp), ("id", _u32p), ("flags", _u16p), ("deathtime", _dp)]

# numpy view of glx_encounter records (layout of EncounterRec below)
_ENC_DTYPE = np.dtype({"names": ["id", "flags", "step", "t_block", "r0", "v0", "r", "v"],
                       "formats": [np.uint32, np.uint16, np.uint32, np.float64, (np.float64, 3), (np.float64, 3),
                                   (np.float64, 3), (np.float64, 3)],
                       "offsets": [0, 4, 8, 16, 24, 48, 72, 96], "itemsize": 120})


class DeviceParticles(C.Structure):
    """glx_device_particles: caller-owned device arrays of one particle set."""
    _fields_ = [("r", C.c_void_p * 3), ("v", C.c_void_p * 3), ("a", C.c_void_p * 3), ("flags", C.c_void_p),
                ("deathtime_index", C.c_void_p), ("id", C.c_void_p)]


PARTITION_STATIC_MAX = 1 << 25


class ExecParams(C.Structure):
    """glx_exec_params: configuration of the host time-block loop."""
    _fields_ = [
        ("t0", C.c_double),
        ("dt", C.c_double),
        ("tbsize", C.c_uint32),
        ("outer_radius", C.c_double),
        ("cull_radius", C.c_double),
        ("rh_factor", C.c_double),
        ("resolve_encounters", C.c_int32),
        ("form", C.c_int32),
        ("device", C.c_int32),
        ("planet_history", C.c_char_p),
        ("history_table_size", C.c_uint32),
        ("resync_every", C.c_uint32),
        ("reference_deathtime", C.c_int32),
        ("encounter_log", C.c_char_p),
    ]


class EncounterRec(C.Structure):
    """glx_encounter: one record of the host close-encounter queue."""
    _fields_ = [
        ("id", C.c_uint32),
        ("flags", C.c_uint16),
        ("step", C.c_uint32),
        ("t_block", C.c_double),
        ("r0", C.c_double * 3),
        ("v0", C.c_double * 3),
        ("r", C.c_double * 3),
        ("v", C.c_double * 3),
    ]

# name -> (restype, argtypes); every symbol declared in include/glisser_amd.h
SIGNATURES = {
    "glx_create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "glx_destroy": (C.c_int, [C.c_void_p]),
    "glx_last_error": (C.c_char_p, []),
    "glx_upload_particles": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, _dp, _dp, _dp, _u16p, _u32p, _u32p]),
    "glx_download_particles": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, _dp, _dp, _dp, _u16p, _u32p, _u32p]),
    "glx_set_counts": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64]),
    "glx_get_counts": (C.c_int, [C.c_void_p, _u64p, _u64p]),
    "glx_upload_planet_block": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32, _dp, _u32p, _dp, _dp, _dp, C.c_uint32]),
    "glx_upload_history": (C.c_int, [C.c_void_p, C.c_uint32, _dp, _u32p, C.c_uint64, _dp, _dp, _dp]),
    "glx_interp_block": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_uint32, _dp]),
    "glx_interp_blocks": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32, _dp, C.c_double, C.c_uint32, _dp]),
    "glx_interp_at": (C.c_int, [C.c_void_p, C.c_uint32, _dp, _dp, _dp]),
    "glx_download_planet_block": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, _dp, _dp]),
    "glx_run_block": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int]),
    "glx_run_blocks": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_double, C.c_double, C.c_int, C.c_uint32,
                                 C.c_uint32, C.c_uint32]),
    "glx_run_blocks_queued": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_double, C.c_double, C.c_int,
                                        C.c_uint32]),
    "glx_init_accel": (C.c_int, [C.c_void_p, C.c_uint32, _dp, _u32p, _dp, _dp, _dp, C.c_double]),
    "glx_kill_encounters": (C.c_int, [C.c_void_p]),
    "glx_partition": (C.c_int, [C.c_void_p, C.c_int, _u64p, _u64p]),
    "glx_snapshot_rollback": (C.c_int, [C.c_void_p]),
    "glx_download_rollback": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, _dp, _dp]),
    "glx_sync": (C.c_int, [C.c_void_p]),
    "glx_last_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "glx_sum_steps": (C.c_int, [C.c_void_p, C.c_uint64, _u64p]),
    "glx_sincos_device": (C.c_int, [C.c_int, C.c_uint64, _dp, _dp, _dp]),
    "glx_partition_static": (C.c_int, [C.POINTER(DeviceParticles), C.POINTER(DeviceParticles), C.c_uint64, C.c_int,
                                       C.c_void_p, C.c_void_p]),
    "glx_sincos_host": (C.c_int, [C.c_uint64, _dp, _dp, _dp]),
    "glx_track_elements_device": (C.c_int, [C.c_int, C.c_uint64, C.c_double, _dp, _dp]),
    "glx_track_elements_host": (C.c_int, [C.c_uint64, C.c_double, _dp, _dp]),
    "glx_cr_math_device": (C.c_int, [C.c_int, C.c_int, C.c_uint64, _dp, _dp, _dp]),
    "glx_cr_math_host": (C.c_int, [C.c_int, C.c_uint64, _dp, _dp, _dp]),
    "glx_batch_create": (C.c_int, [C.c_int, C.POINTER(SystemDesc), C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "glx_batch_destroy": (C.c_int, [C.c_void_p]),
    "glx_batch_upload": (C.c_int, [C.c_void_p, C.c_uint32, _dp, _dp, _u16p, _u32p]),
    "glx_batch_download": (C.c_int, [C.c_void_p, C.c_uint32, _dp, _dp, _u16p, _u32p, _u32p]),
    "glx_batch_init_accel": (C.c_int, [C.c_void_p]),
    "glx_batch_set_config": (C.c_int, [C.c_void_p, C.POINTER(BatchConfig)]),
    "glx_batch_get_config": (C.c_int, [C.c_void_p, C.POINTER(BatchConfig)]),
    "glx_batch_run": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.POINTER(C.c_float)]),
    "glx_batch_tune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(BatchConfig), C.c_uint32, C.POINTER(BatchConfig),
                                 C.POINTER(C.c_double)]),
    "glx_batch_sum_steps": (C.c_int, [C.c_void_p, _u64p]),
    "glx_pack_encounters": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_double, C.POINTER(EncounterRec)]),
    "glx_accel_range": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint32, _dp, _u32p, _dp, _dp, _dp,
                                  C.c_double]),
    "glx_exec_create": (C.c_int, [C.POINTER(ExecParams), C.c_uint32, _dp, _dp, _dp, _u32p, C.c_uint64, _dp, _dp,
                                  _u32p, C.POINTER(C.c_void_p)]),
    "glx_exec_destroy": (C.c_int, [C.c_void_p]),
    "glx_exec_step": (C.c_int, [C.c_void_p, C.c_uint32]),
    "glx_exec_status": (C.c_int, [C.c_void_p, _dp, _u64p, _u64p, _u64p]),
    "glx_exec_take_encounters": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(EncounterRec), _u64p]),
    "glx_exec_return": (C.c_int, [C.c_void_p, C.c_uint64, _u32p, _dp, _dp]),
    "glx_exec_download": (C.c_int, [C.c_void_p, _dp, _dp, _u16p, _u32p, _dp]),
    "glx_exec_planets": (C.c_int, [C.c_void_p, _dp]),
    "glx_exec_bodies": (C.c_int, [C.c_void_p, _u32p, _dp, _u32p]),
    "glx_exec_times": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "glx_exec_kernel_total": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_uint64)]),
    "glx_config_normalize": (C.c_int, [C.c_char_p, C.c_char_p, C.c_uint64, _u64p]),
    "glx_state_save": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.POINTER(StateC)]),
    "glx_state_sizes": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_uint64, _u32p, _u64p]),
    "glx_state_load": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_uint64, C.c_int, C.c_double,
                                 C.POINTER(StateC)]),
    "glx_track_write": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.POINTER(StateC), C.c_int, C.c_int]),
    "glx_track_pack": (C.c_int, [C.c_void_p, C.c_uint32, _dp, _dp, _dp, C.c_int, C.c_int, C.c_void_p, C.c_uint64,
                                 _u64p]),
    "glx_track_write_records": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.POINTER(StateC), C.c_int, C.c_int,
                                          C.c_void_p, C.c_uint64]),
    "glx_plhist_write": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.c_uint32, _dp, _dp, _dp, _u32p]),
    "glx_plhist_open": (C.c_int, [C.c_char_p, C.c_double, C.c_double, C.c_uint32, C.POINTER(C.c_void_p)]),
    "glx_plhist_close": (C.c_int, [C.c_void_p]),
    "glx_plhist_state": (C.c_int, [C.c_void_p, _u32p, _dp, _u32p, _dp, _dp]),
    "glx_plhist_block": (C.c_int, [C.c_void_p, C.c_uint32, _dp, _dp, _u32p, _u32p, _dp, _u32p, _dp, _dp]),
    "glx_planets_create": (C.c_int, [C.c_uint32, _dp, _dp, _dp, _u32p, C.c_double, C.c_double, C.c_double,
                                     C.POINTER(C.c_void_p)]),
    "glx_planets_destroy": (C.c_int, [C.c_void_p]),
    "glx_planets_h0_initial": (C.c_int, [C.c_void_p, _dp]),
    "glx_planets_block": (C.c_int, [C.c_void_p, C.c_uint32, C.c_double, _dp, _dp, _dp]),
    "glx_planets_state": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "glx_planets_recalculate_rh": (C.c_int, [C.c_void_p]),
    "glx_planets_metrics": (C.c_int, [C.c_void_p, _dp, _dp]),
    "glx_load_h0": (C.c_int, [C.c_uint32, _dp, _dp, C.c_uint32, _dp]),
    "glx_to_elements": (C.c_int, [C.c_uint64, _dp, _dp, _dp, _dp]),
    "glx_from_elements": (C.c_int, [C.c_uint64, _dp, _dp, C.c_int, _dp, _dp]),
}

_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m glisser_amd.build` "
                "(the particle integrator has no CPU fallback)")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _ptr(a, typ):
    if a is None:
        return None
    return a.ctypes.data_as(typ)


def _vec3(a, n):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != (n, 3):
        raise ValueError(f"expected shape ({n}, 3), got {a.shape}")
    return a


class Context:
    """One device context: particle arrays, planet-block buffers, streams."""

    def __init__(self, n_particles: int, n_bodies_max: int, tbsize_max: int, device: int = 0,
                 keep_rollback: bool = False, n_planet_buffers: int = 2):
        self.lib = load()
        p = Params(n_particles, n_bodies_max, tbsize_max, device, int(keep_rollback), n_planet_buffers)
        h = C.c_void_p()
        self._check(self.lib.glx_create(C.byref(p), C.byref(h)))
        self.h = h
        self.capacity = n_particles

    def _check(self, rc: int) -> None:
        if rc != GLX_OK:
            raise GlxError(rc, self.lib.glx_last_error().decode())

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.glx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- particles -----------------------------------------------------
    def upload(self, r, v, flags, ids, a=None, dti=None, begin: int = 0) -> None:
        n = len(flags)
        r, v = _vec3(r, n), _vec3(v, n)
        a = None if a is None else _vec3(a, n)
        flags = np.ascontiguousarray(flags, dtype=np.uint16)
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        dti = None if dti is None else np.ascontiguousarray(dti, dtype=np.uint32)
        self._check(self.lib.glx_upload_particles(self.h, begin, n, _ptr(r, _dp), _ptr(v, _dp), _ptr(a, _dp),
                                                  _ptr(flags, _u16p), _ptr(ids, _u32p), _ptr(dti, _u32p)))

    def download(self, n: int, begin: int = 0) -> dict:
        out = {
            "r": np.empty((n, 3)), "v": np.empty((n, 3)), "a": np.empty((n, 3)),
            "flags": np.empty(n, np.uint16), "id": np.empty(n, np.uint32), "dti": np.empty(n, np.uint32),
        }
        self._check(self.lib.glx_download_particles(
            self.h, begin, n, _ptr(out["r"], _dp), _ptr(out["v"], _dp), _ptr(out["a"], _dp),
            _ptr(out["flags"], _u16p), _ptr(out["id"], _u32p), _ptr(out["dti"], _u32p)))
        return out

    def set_counts(self, n_alive: int, n_total: int) -> None:
        self._check(self.lib.glx_set_counts(self.h, n_alive, n_total))

    def counts(self) -> tuple[int, int]:
        a, t = C.c_uint64(), C.c_uint64()
        self._check(self.lib.glx_get_counts(self.h, C.byref(a), C.byref(t)))
        return a.value, t.value

    # --- planets / kernels -----------------------------------------------
    def upload_planet_block(self, buf: int, m, rh, r_log, h0_log, ids=None) -> None:
        m = np.ascontiguousarray(m, dtype=np.float64)
        nb = len(m)
        rh = np.ascontiguousarray(rh, dtype=np.float64)
        h0 = np.ascontiguousarray(h0_log, dtype=np.float64).reshape(-1, 3)
        steps = h0.shape[0]
        rl = np.ascontiguousarray(r_log, dtype=np.float64).reshape(steps, nb - 1, 3)
        ids = np.arange(nb, dtype=np.uint32) if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
        self._check(self.lib.glx_upload_planet_block(self.h, buf, nb, _ptr(m, _dp), _ptr(ids, _u32p),
                                                     _ptr(rh, _dp), _ptr(rl, _dp), _ptr(h0, _dp), steps))

    def upload_history(self, m, t, r, v, ids=None) -> None:
        """Sampled planet history: times t (k,), heliocentric r, v of the
        planets (k, n_bodies-1, 3)."""
        m = np.ascontiguousarray(m, dtype=np.float64)
        nb = len(m)
        t = np.ascontiguousarray(t, dtype=np.float64)
        k = len(t)
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(k, nb - 1, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(k, nb - 1, 3)
        ids = np.arange(nb, dtype=np.uint32) if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
        self._hist_bodies = nb
        self._check(self.lib.glx_upload_history(self.h, nb, _ptr(m, _dp), _ptr(ids, _u32p), k, _ptr(t, _dp),
                                                _ptr(r, _dp), _ptr(v, _dp)))

    def interp_block(self, buf: int, t_start: float, dt: float, n_steps: int, rh) -> None:
        rh = np.ascontiguousarray(rh, dtype=np.float64)
        self._check(self.lib.glx_interp_block(self.h, buf, t_start, dt, n_steps, _ptr(rh, _dp)))

    def interp_blocks(self, first_buf: int, t_start, dt: float, n_steps: int, rh) -> None:
        """interp_block for len(t_start) consecutive buffers in one launch."""
        rh = np.ascontiguousarray(rh, dtype=np.float64)
        t_start = np.ascontiguousarray(np.atleast_1d(t_start), dtype=np.float64)
        self._check(self.lib.glx_interp_blocks(self.h, first_buf, len(t_start), _ptr(t_start, _dp), dt, n_steps,
                                               _ptr(rh, _dp)))

    def interp_at(self, times):
        """(r (n, n_bodies-1, 3), h0 (n, 3)) interpolated at the given times."""
        times = np.ascontiguousarray(np.atleast_1d(times), dtype=np.float64)
        nb = self._hist_bodies
        r = np.empty((len(times), nb - 1, 3))
        h0 = np.empty((len(times), 3))
        self._check(self.lib.glx_interp_at(self.h, len(times), _ptr(times, _dp), _ptr(r, _dp), _ptr(h0, _dp)))
        return r, h0

    def download_planet_block(self, buf: int, n_bodies: int, n_steps: int):
        r = np.empty((n_steps, n_bodies - 1, 3))
        h0 = np.empty((n_steps, 3))
        self._check(self.lib.glx_download_planet_block(self.h, buf, n_bodies, n_steps, _ptr(r, _dp), _ptr(h0, _dp)))
        return r, h0

    def run_block(self, buf: int, dt: float, outer_radius: float, form: str | int = "fp64") -> None:
        f = FORMS[form] if isinstance(form, str) else int(form)
        self._check(self.lib.glx_run_block(self.h, buf, dt, outer_radius, f))

    def run_blocks(self, n_blocks: int, first_buf: int, dt: float, outer_radius: float, form: str | int = "fp64",
                   n_shards: int = 4, waves_per_simd: int = 3, launches_per_block: int = 1) -> None:
        f = FORMS[form] if isinstance(form, str) else int(form)
        self._check(self.lib.glx_run_blocks(self.h, n_blocks, first_buf, dt, outer_radius, f, n_shards,
                                            waves_per_simd, launches_per_block))

    def run_blocks_queued(self, n_blocks: int, first_buf: int, dt: float, outer_radius: float,
                          form: str | int = "fp64", segment_steps: int = 0) -> None:
        """run_blocks as one persistent launch taking (tile, segment) tasks
        from a device queue (segment_steps: a multiple of 64, 0 = 256)."""
        f = FORMS[form] if isinstance(form, str) else int(form)
        self._check(self.lib.glx_run_blocks_queued(self.h, n_blocks, first_buf, dt, outer_radius, f, segment_steps))

    def init_accel(self, m, rh, planet_r, h0, outer_radius: float, ids=None) -> None:
        m = np.ascontiguousarray(m, dtype=np.float64)
        nb = len(m)
        rh = np.ascontiguousarray(rh, dtype=np.float64)
        pr = np.ascontiguousarray(planet_r, dtype=np.float64).reshape(nb - 1, 3)
        h0 = np.ascontiguousarray(h0, dtype=np.float64).reshape(3)
        ids = np.arange(nb, dtype=np.uint32) if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
        self._check(self.lib.glx_init_accel(self.h, nb, _ptr(m, _dp), _ptr(ids, _u32p), _ptr(rh, _dp),
                                            _ptr(pr, _dp), _ptr(h0, _dp), outer_radius))

    def kill_encounters(self) -> None:
        self._check(self.lib.glx_kill_encounters(self.h))

    def partition(self, three_way: bool = False) -> tuple[int, int]:
        a, e = C.c_uint64(), C.c_uint64()
        mode = PART_THREE_WAY if three_way else PART_UNFLAGGED
        self._check(self.lib.glx_partition(self.h, mode, C.byref(a), C.byref(e)))
        return a.value, e.value

    def snapshot_rollback(self) -> None:
        self._check(self.lib.glx_snapshot_rollback(self.h))

    def download_rollback(self, n: int, begin: int = 0):
        r, v = np.empty((n, 3)), np.empty((n, 3))
        self._check(self.lib.glx_download_rollback(self.h, begin, n, _ptr(r, _dp), _ptr(v, _dp)))
        return r, v

    def track_records(self, m, r, v, to_elements: bool = True, barycentric: bool = True) -> np.ndarray:
        """Track records (structured: id, x[6] float32) of the alive particles,
        ordered by id, formed on the device from planets m, r, v (Sun first)."""
        m = np.ascontiguousarray(m, dtype=np.float64)
        r = _vec3(r, len(m))
        v = _vec3(v, len(m))
        n_alive, _ = self.counts()
        out = np.empty(max(n_alive, 1), dtype=TRACK_DTYPE)
        got = C.c_uint64(0)
        self._check(self.lib.glx_track_pack(self.h, len(m), _ptr(m, _dp), _ptr(r, _dp), _ptr(v, _dp),
                                            int(to_elements), int(barycentric), out.ctypes.data, len(out),
                                            C.byref(got)))
        return out[:got.value]

    def sync(self) -> None:
        self._check(self.lib.glx_sync(self.h))

    def last_times(self) -> tuple[float, float]:
        k, p = C.c_float(), C.c_float()
        self._check(self.lib.glx_last_times(self.h, C.byref(k), C.byref(p)))
        return k.value, p.value

    def sum_steps(self, n: int) -> int:
        t = C.c_uint64()
        self._check(self.lib.glx_sum_steps(self.h, n, C.byref(t)))
        return t.value


def _check_host(rc: int) -> None:
    if rc != GLX_OK:
        raise GlxError(rc, load().glx_last_error().decode())


class Planets:
    """Host planet producer (Jacobi WH): one planet log + h0 log per block."""

    def __init__(self, m, r, v, ids=None, cull_radius: float = 0.5, rh_factor: float = 0.0,
                 outer_radius: float = 300.0):
        self.lib = load()
        self.m = np.ascontiguousarray(m, dtype=np.float64)
        self.n = len(self.m)
        r = _vec3(r, self.n)
        v = _vec3(v, self.n)
        self.ids = (np.arange(self.n, dtype=np.uint32) if ids is None
                    else np.ascontiguousarray(ids, dtype=np.uint32))
        self.outer_radius = outer_radius
        h = C.c_void_p()
        _check_host(self.lib.glx_planets_create(self.n, _ptr(self.m, _dp), _ptr(r, _dp), _ptr(v, _dp),
                                                _ptr(self.ids, _u32p), cull_radius, rh_factor, outer_radius,
                                                C.byref(h)))
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.glx_planets_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def h0_initial(self) -> np.ndarray:
        out = np.empty(3)
        _check_host(self.lib.glx_planets_h0_initial(self.h, _ptr(out, _dp)))
        return out

    def block(self, n_steps: int, dt: float):
        """Advance n_steps; returns (r_log, v_log, h0_log)."""
        rl = np.empty((n_steps, self.n - 1, 3))
        vl = np.empty((n_steps, self.n - 1, 3))
        h0 = np.empty((n_steps, 3))
        _check_host(self.lib.glx_planets_block(self.h, n_steps, dt, _ptr(rl, _dp), _ptr(vl, _dp), _ptr(h0, _dp)))
        return rl, vl, h0

    def state(self):
        r, v, rh = np.empty((self.n, 3)), np.empty((self.n, 3)), np.empty(self.n)
        _check_host(self.lib.glx_planets_state(self.h, _ptr(r, _dp), _ptr(v, _dp), _ptr(rh, _dp)))
        return r, v, rh

    def recalculate_rh(self) -> None:
        _check_host(self.lib.glx_planets_recalculate_rh(self.h))

    def metrics(self):
        e = np.empty(1)
        lv = np.empty(3)
        _check_host(self.lib.glx_planets_metrics(self.h, _ptr(e, _dp), _ptr(lv, _dp)))
        return float(e[0]), lv


class Batch:
    """Many independent systems advanced together (glx_batch_*).

    systems: sequence of dicts with keys m, rh, n_particles, dt, history
    (an object with t, r, v), optional t0 (default: first history time),
    outer_radius (default 300) and ids (planet ids)."""

    def __init__(self, systems, steps_max: int = 2048, device: int = 0):
        self.lib = load()
        self._keep = []
        descs = (SystemDesc * len(systems))()
        self.sizes = []
        for k, y in enumerate(systems):
            m = np.ascontiguousarray(y["m"], dtype=np.float64)
            nb = len(m)
            rh = np.ascontiguousarray(y["rh"], dtype=np.float64)
            h = y["history"]
            t = np.ascontiguousarray(h.t, dtype=np.float64)
            r = np.ascontiguousarray(h.r, dtype=np.float64).reshape(len(t), nb - 1, 3)
            v = np.ascontiguousarray(h.v, dtype=np.float64).reshape(len(t), nb - 1, 3)
            ids = y.get("ids")
            ids = None if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
            self._keep += [m, rh, t, r, v, ids]
            d = descs[k]
            d.n_bodies = nb
            d.n_particles = int(y["n_particles"])
            d.t0 = float(y.get("t0", t[0]))
            d.dt = float(y["dt"])
            d.outer_radius = float(y.get("outer_radius", 300.0))
            d.m, d.rh, d.t, d.r, d.v = (_ptr(m, _dp), _ptr(rh, _dp), _ptr(t, _dp), _ptr(r, _dp), _ptr(v, _dp))
            d.id = _ptr(ids, _u32p)
            d.n_samples = len(t)
            self.sizes.append(int(y["n_particles"]))
        h = C.c_void_p()
        rc = self.lib.glx_batch_create(device, descs, len(systems), steps_max, C.byref(h))
        self._keep = None
        if rc != 0:
            raise GlxError(rc, self.lib.glx_last_error().decode())
        self.h = h

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise GlxError(rc, self.lib.glx_last_error().decode())

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.glx_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, sys: int, r, v, flags=None, ids=None) -> None:
        n = self.sizes[sys]
        r = _vec3(r, n)
        v = _vec3(v, n)
        flags = np.zeros(n, np.uint16) if flags is None else np.ascontiguousarray(flags, dtype=np.uint16)
        ids = np.arange(n, dtype=np.uint32) if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
        self._check(self.lib.glx_batch_upload(self.h, sys, _ptr(r, _dp), _ptr(v, _dp), _ptr(flags, _u16p),
                                              _ptr(ids, _u32p)))

    def download(self, sys: int) -> dict:
        n = self.sizes[sys]
        out = {"r": np.empty((n, 3)), "v": np.empty((n, 3)), "flags": np.empty(n, np.uint16),
               "id": np.empty(n, np.uint32), "dti": np.empty(n, np.uint32)}
        self._check(self.lib.glx_batch_download(self.h, sys, _ptr(out["r"], _dp), _ptr(out["v"], _dp),
                                                _ptr(out["flags"], _u16p), _ptr(out["id"], _u32p),
                                                _ptr(out["dti"], _u32p)))
        return out

    def init_accel(self) -> None:
        self._check(self.lib.glx_batch_init_accel(self.h))

    def set_config(self, steps_per_launch: int, lanes: int = 256, persistent: int = 0,
                   waves_per_simd: int = 0) -> None:
        c = BatchConfig(steps_per_launch, lanes, persistent, waves_per_simd)
        self._check(self.lib.glx_batch_set_config(self.h, C.byref(c)))

    def config(self) -> dict:
        c = BatchConfig()
        self._check(self.lib.glx_batch_get_config(self.h, C.byref(c)))
        return c.as_dict()

    def run(self, n_launches: int, form: str | int = "fp64") -> float:
        f = FORMS[form] if isinstance(form, str) else int(form)
        ms = C.c_float(0.0)
        self._check(self.lib.glx_batch_run(self.h, n_launches, f, C.byref(ms)))
        return float(ms.value)

    def tune(self, form: str | int = "fast", candidates=None):
        """Returns (best config dict, particle-steps/s of the best)."""
        f = FORMS[form] if isinstance(form, str) else int(form)
        cand = None
        n = 0
        if candidates:
            cand = (BatchConfig * len(candidates))(*[BatchConfig(*c) for c in candidates])
            n = len(candidates)
        best = BatchConfig()
        rate = C.c_double(0.0)
        self._check(self.lib.glx_batch_tune(self.h, f, cand, n, C.byref(best), C.byref(rate)))
        return best.as_dict(), float(rate.value)

    def sum_steps(self) -> int:
        out = C.c_uint64(0)
        self._check(self.lib.glx_batch_sum_steps(self.h, C.byref(out)))
        return int(out.value)


class Executor:
    """Host time-block loop (glx_exec_*): planets from the host Jacobi WH
    producer, particles on the device, optional encounter hand-off."""

    def __init__(self, m, r, v, pr, pv, ids=None, planet_ids=None, t0: float = 0.0, dt: float = 180.0,
                 tbsize: int = 2048, outer_radius: float = 300.0, cull_radius: float = 0.5, rh_factor: float = 0.0,
                 resolve_encounters: bool = False, form: str | int = "fp64", device: int = 0,
                 planet_history=None, history_table_size: int = 16, resync_every: int = 1,
                 reference_deathtime: bool = False, encounter_log=None):
        """m, r, v: planets (Sun first); with planet_history (a history file)
        they are not used and may be None.  resync_every: Resync-Interval;
        reference_deathtime: death times as the reference records them (end
        of the resync block, cast to float); encounter_log: a file for the
        reference's encounter.out lines."""
        self.lib = load()
        if planet_history is None:
            m = np.ascontiguousarray(m, dtype=np.float64)
            nb = len(m)
            r = _vec3(r, nb)
            v = _vec3(v, nb)
        else:
            m = r = v = None
            nb = 0
        n = len(pr)
        pr = _vec3(pr, n)
        pv = _vec3(pv, n)
        ids = np.arange(n, dtype=np.uint32) if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
        pid = None if planet_ids is None else np.ascontiguousarray(planet_ids, dtype=np.uint32)
        f = FORMS[form] if isinstance(form, str) else int(form)
        hist = None if planet_history is None else str(planet_history).encode()
        log = None if encounter_log is None else str(encounter_log).encode()
        p = ExecParams(t0, dt, tbsize, outer_radius, cull_radius, rh_factor, int(resolve_encounters), f, device, hist,
                       history_table_size, resync_every, int(reference_deathtime), log)
        h = C.c_void_p()
        rc = self.lib.glx_exec_create(C.byref(p), nb, _ptr(m, _dp), _ptr(r, _dp), _ptr(v, _dp),
                                      _ptr(pid, _u32p), n, _ptr(pr, _dp), _ptr(pv, _dp), _ptr(ids, _u32p),
                                      C.byref(h))
        if rc != 0:
            raise GlxError(rc, self.lib.glx_last_error().decode())
        self.h = h
        self.n = n

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise GlxError(rc, self.lib.glx_last_error().decode())

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.glx_exec_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n_blocks: int = 1) -> None:
        self._check(self.lib.glx_exec_step(self.h, n_blocks))

    def status(self) -> dict:
        t = C.c_double()
        a, tot, q = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self._check(self.lib.glx_exec_status(self.h, C.byref(t), C.byref(a), C.byref(tot), C.byref(q)))
        return {"t": t.value, "n_alive": a.value, "n_total": tot.value, "n_queued": q.value}

    def take_encounters(self) -> dict:
        """Queued records as arrays: id, flags, step, t_block, r0, v0, r, v."""
        n = self.status()["n_queued"]
        a = np.empty(max(n, 1), dtype=_ENC_DTYPE)
        got = C.c_uint64()
        self._check(self.lib.glx_exec_take_encounters(self.h, n, a.ctypes.data_as(C.POINTER(EncounterRec)),
                                                      C.byref(got)))
        a = a[:got.value]
        out = {k: a[k].copy() for k in ("id", "flags", "step", "t_block")}
        for k in ("r0", "v0", "r", "v"):
            out[k] = a[k].reshape(-1, 3).copy()
        return out

    def return_particles(self, ids, r, v) -> None:
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        r = _vec3(r, len(ids))
        v = _vec3(v, len(ids))
        self._check(self.lib.glx_exec_return(self.h, len(ids), _ptr(ids, _u32p), _ptr(r, _dp), _ptr(v, _dp)))

    def download(self) -> dict:
        n = self.n
        out = {"r": np.empty((n, 3)), "v": np.empty((n, 3)), "flags": np.empty(n, np.uint16),
               "id": np.empty(n, np.uint32), "deathtime": np.empty(n)}
        self._check(self.lib.glx_exec_download(self.h, _ptr(out["r"], _dp), _ptr(out["v"], _dp),
                                               _ptr(out["flags"], _u16p), _ptr(out["id"], _u32p),
                                               _ptr(out["deathtime"], _dp)))
        out.update(self.status())
        return out

    def bodies(self) -> dict:
        """The bodies at the current time: n_bodies, GM values and ids."""
        nb = C.c_uint32()
        m, ids = np.empty(GLX_MAX_BODIES), np.empty(GLX_MAX_BODIES, np.uint32)
        self._check(self.lib.glx_exec_bodies(self.h, C.byref(nb), _ptr(m, _dp), _ptr(ids, _u32p)))
        return {"n_bodies": nb.value, "m": m[:nb.value].copy(), "id": ids[:nb.value].copy()}

    @property
    def nb(self) -> int:
        return self.bodies()["n_bodies"]

    def planet_positions(self) -> np.ndarray:
        out = np.empty((GLX_MAX_BODIES - 1) * 3)
        self._check(self.lib.glx_exec_planets(self.h, _ptr(out, _dp)))
        return out[: (self.nb - 1) * 3].reshape(-1, 3).copy()

    def times(self) -> tuple[float, float, float]:
        k, q, p = C.c_double(), C.c_double(), C.c_double()
        self._check(self.lib.glx_exec_times(self.h, C.byref(k), C.byref(q), C.byref(p)))
        return k.value, q.value, p.value

    def kernel_total(self) -> tuple[float, int]:
        """Device time of every block kernel so far (ms) and their count."""
        k, n = C.c_double(), C.c_uint64()
        self._check(self.lib.glx_exec_kernel_total(self.h, C.byref(k), C.byref(n)))
        return k.value, n.value


def _state_c(st: dict):
    """glx_state over numpy arrays (returned with the arrays it points to)."""
    n = len(st["id"])
    a = {"pm": np.ascontiguousarray(st["pm"], np.float64), "pr": np.ascontiguousarray(st["pr"], np.float64),
         "pv": np.ascontiguousarray(st["pv"], np.float64),
         "pid": np.ascontiguousarray(st.get("pid", np.arange(len(st["pm"]))), np.uint32),
         "r": np.ascontiguousarray(st["r"], np.float64), "v": np.ascontiguousarray(st["v"], np.float64),
         "id": np.ascontiguousarray(st["id"], np.uint32),
         "flags": np.ascontiguousarray(st.get("flags", np.zeros(n)), np.uint16),
         "deathtime": np.ascontiguousarray(st.get("deathtime", np.zeros(n)), np.float64)}
    c = StateC(len(a["pm"]), n, _ptr(a["pm"], _dp), _ptr(a["pr"], _dp), _ptr(a["pv"], _dp), _ptr(a["pid"], _u32p),
               _ptr(a["r"], _dp), _ptr(a["v"], _dp), _ptr(a["id"], _u32p), _ptr(a["flags"], _u16p),
               _ptr(a["deathtime"], _dp))
    return c, a


def config_normalize(text: str) -> str:
    """Parse a configuration text and return its canonical form."""
    lib = load()
    buf = C.create_string_buffer(1 << 16)
    n = C.c_uint64(0)
    _check_host(lib.glx_config_normalize(text.encode(), buf, len(buf), C.byref(n)))
    return buf.value.decode()


def save_state(path, st: dict, fmt: int = STATE_ASCII, write_momenta: bool = False, particle_path=None) -> None:
    c, _keep = _state_c(st)
    _check_host(load().glx_state_save(str(path).encode(),
                                      None if particle_path is None else str(particle_path).encode(), fmt,
                                      int(write_momenta), C.byref(c)))


def load_state(path, fmt: int = STATE_ASCII, particle_path=None, max_particles: int = 0xFFFFFFFF,
               read_momenta: bool = False, big_g: float = 1.0) -> dict:
    lib = load()
    pp = None if particle_path is None else str(particle_path).encode()
    npl, n = C.c_uint32(0), C.c_uint64(0)
    _check_host(lib.glx_state_sizes(str(path).encode(), pp, fmt, max_particles, C.byref(npl), C.byref(n)))
    npl, n = npl.value, n.value
    st = {"pm": np.empty(npl), "pr": np.empty((npl, 3)), "pv": np.empty((npl, 3)), "pid": np.empty(npl, np.uint32),
          "r": np.empty((n, 3)), "v": np.empty((n, 3)), "id": np.empty(n, np.uint32), "flags": np.empty(n, np.uint16),
          "deathtime": np.empty(n)}
    c, _keep = _state_c(st)
    _check_host(lib.glx_state_load(str(path).encode(), pp, fmt, max_particles, int(read_momenta), big_g,
                                   C.byref(c)))
    return st


TRACK_DTYPE = np.dtype([("id", np.uint32), ("x", np.float32, 6)])  # 28-byte track record


def write_track_records(path, time: float, planets: dict, records: np.ndarray, to_elements: bool = True,
                        barycentric: bool = True, append: bool = False) -> None:
    """Track snapshot from the planets of `planets` and device-formed records."""
    st = dict(planets, r=np.zeros((0, 3)), v=np.zeros((0, 3)), id=np.zeros(0, np.uint32))
    c, _keep = _state_c(st)
    rec = np.ascontiguousarray(records, dtype=TRACK_DTYPE)
    _check_host(load().glx_track_write_records(str(path).encode(), int(append), time, C.byref(c), int(to_elements),
                                               int(barycentric), rec.ctypes.data, len(rec)))


def write_track(path, time: float, st: dict, to_elements: bool = True, barycentric: bool = True,
                append: bool = False) -> None:
    c, _keep = _state_c(st)
    _check_host(load().glx_track_write(str(path).encode(), int(append), time, C.byref(c), int(to_elements),
                                       int(barycentric)))


def write_plhist(path, time: float, m, r, v, ids=None, append: bool = False) -> None:
    """Start (or, with append, extend) a planet history file with one record
    of heliocentric bodies (entry 0 the Sun)."""
    m = np.ascontiguousarray(m, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    ids = None if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
    _check_host(load().glx_plhist_write(str(path).encode(), int(append), time, len(m), _ptr(m, _dp), _ptr(r, _dp),
                                        _ptr(v, _dp), _ptr(ids, _u32p)))


class PlanetHistory:
    """A planet history file read block by block (element-linear
    interpolation, blocks aligned to its records)."""

    def __init__(self, path, t0: float, dt: float, table_size: int = 16):
        self.lib = load()
        self.table_size = table_size
        h = C.c_void_p()
        _check_host(self.lib.glx_plhist_open(str(path).encode(), t0, dt, table_size, C.byref(h)))
        self.h = h

    def close(self) -> None:
        if self.h:
            self.lib.glx_plhist_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def state(self) -> dict:
        nb = np.zeros(1, np.uint32)
        m, ids = np.empty(self.table_size), np.empty(self.table_size, np.uint32)
        r, v = np.empty((self.table_size, 3)), np.empty((self.table_size, 3))
        _check_host(self.lib.glx_plhist_state(self.h, _ptr(nb, _u32p), _ptr(m, _dp), _ptr(ids, _u32p), _ptr(r, _dp),
                                              _ptr(v, _dp)))
        return {"n_bodies": int(nb[0]), "m": m, "id": ids, "r": r, "v": v}

    def block(self, tbsize: int) -> dict:
        t, dt = np.zeros(1), np.zeros(1)
        ns, nb = np.zeros(1, np.uint32), np.zeros(1, np.uint32)
        m, ids = np.empty(self.table_size), np.empty(self.table_size, np.uint32)
        rl = np.empty(tbsize * max(self.table_size - 1, 1) * 3)
        h0 = np.empty((tbsize, 3))
        _check_host(self.lib.glx_plhist_block(self.h, tbsize, _ptr(t, _dp), _ptr(dt, _dp), _ptr(ns, _u32p),
                                              _ptr(nb, _u32p), _ptr(m, _dp), _ptr(ids, _u32p), _ptr(rl, _dp),
                                              _ptr(h0, _dp)))
        n, b = int(ns[0]), int(nb[0])
        return {"t": float(t[0]), "dt": float(dt[0]), "n_steps": n, "n_bodies": b, "m": m[:b].copy(),
                "id": ids[:b].copy(), "r_log": rl[: n * (b - 1) * 3].reshape(n, b - 1, 3).copy(),
                "h0_log": h0[:n].copy()}


def load_h0(m, r_log) -> np.ndarray:
    m = np.ascontiguousarray(m, dtype=np.float64)
    rl = np.ascontiguousarray(r_log, dtype=np.float64)
    steps = rl.shape[0]
    h0 = np.empty((steps, 3))
    _check_host(load().glx_load_h0(len(m), _ptr(m, _dp), _ptr(rl, _dp), steps, _ptr(h0, _dp)))
    return h0


def to_elements(mu, r, v) -> np.ndarray:
    r = np.ascontiguousarray(np.atleast_2d(r), dtype=np.float64)
    v = np.ascontiguousarray(np.atleast_2d(v), dtype=np.float64)
    n = r.shape[0]
    mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (n,)))
    el = np.empty((n, 6))
    _check_host(load().glx_to_elements(n, _ptr(mu, _dp), _ptr(r, _dp), _ptr(v, _dp), _ptr(el, _dp)))
    return el


def from_elements(mu, el, mean_anomaly: bool = True):
    el = np.ascontiguousarray(np.atleast_2d(el), dtype=np.float64)
    n = el.shape[0]
    mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (n,)))
    r, v = np.empty((n, 3)), np.empty((n, 3))
    _check_host(load().glx_from_elements(n, _ptr(mu, _dp), _ptr(el, _dp), int(mean_anomaly), _ptr(r, _dp),
                                         _ptr(v, _dp)))
    return r, v


def track_elements(mu: float, rv, device: int | None = None):
    """Elements (a, e, i, Omega, omega, f) of Cartesian rows rv (n x 6) about
    central GM mu, as the device track output forms them: on GPU `device`,
    or with the host twin of the same code when device is None."""
    rv = np.ascontiguousarray(np.atleast_2d(rv), dtype=np.float64)
    if rv.ndim != 2 or rv.shape[1] != 6:
        raise ValueError("rv must have shape (n, 6)")
    n = rv.shape[0]
    el = np.empty((n, 6))
    lib = load()
    if device is None:
        _check_host(lib.glx_track_elements_host(n, float(mu), _ptr(rv, _dp), _ptr(el, _dp)))
    else:
        _check_host(lib.glx_track_elements_device(int(device), n, float(mu), _ptr(rv, _dp), _ptr(el, _dp)))
    return el


CR_FN = {"atan2": 0, "acos": 1, "asin": 2}


def cr_math(fn: str, a, b=None, device: int | None = None):
    """The track conversion's correctly rounded atan2(a, b), acos(a) or
    asin(a) over arrays: on GPU `device`, or on the host when device is None."""
    a = np.ascontiguousarray(np.atleast_1d(a), dtype=np.float64)
    n = a.shape[0]
    if fn == "atan2":
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(b, dtype=np.float64), (n,)))
    out = np.empty(n)
    bp = _ptr(b, _dp) if b is not None else None
    lib = load()
    if device is None:
        _check_host(lib.glx_cr_math_host(CR_FN[fn], n, _ptr(a, _dp), bp, _ptr(out, _dp)))
    else:
        _check_host(lib.glx_cr_math_device(int(device), CR_FN[fn], n, _ptr(a, _dp), bp, _ptr(out, _dp)))
    return out


def sincos(x, device: int | None = None):
    """sin and cos of x as the fp64 form's Kepler drift evaluates them
    (correctly rounded for |x| < 2^20): on GPU `device`, or with the host
    twin of the same routine when device is None."""
    x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
    n = x.shape[0]
    s, c = np.empty(n), np.empty(n)
    lib = load()
    if device is None:
        _check_host(lib.glx_sincos_host(n, _ptr(x, _dp), _ptr(s, _dp), _ptr(c, _dp)))
    else:
        _check_host(lib.glx_sincos_device(int(device), n, _ptr(x, _dp), _ptr(s, _dp), _ptr(c, _dp)))
    return s, c
