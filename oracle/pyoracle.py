"""ctypes wrapper of the CPU oracle libraries (TEST INFRASTRUCTURE ONLY).

``Oracle()`` loads ``oracle/liboracle.so`` (the repo's restatement of the
reference CPU WH path); ``Oracle(reference=True)`` loads
``oracle/_ref/libglisser_ref.so`` (the reference's own CPU sources, built in
the container that holds them).  Both expose the same entry points.
Only tests/, the smoke entry point and bench.py's CPU-baseline leg use it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libglisser_ref.so"

_dp = C.POINTER(C.c_double)
_u16p = C.POINTER(C.c_uint16)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)

_SIGS = {
    "last_error": (C.c_char_p, []),
    "particle_accels": (None, [C.c_uint32, _dp, _dp, C.c_double, _dp, _dp, C.c_double, C.c_uint32, _dp, _dp,
                               _u16p, _u32p, _dp]),
    "particle_block": (C.c_int, [C.c_uint32, _dp, _dp, C.c_double, _dp, _dp, C.c_uint32, C.c_double, C.c_double,
                                 C.c_uint32, _dp, _dp, _dp, _u16p, _u32p, _u32p, _dp, _u64p]),
    "planets_new": (C.c_void_p, [C.c_uint32, _dp, _dp, _dp, _u32p, C.c_double, C.c_double, C.c_double, _dp]),
    "planets_free": (None, [C.c_void_p]),
    "planets_block": (C.c_int, [C.c_void_p, C.c_uint32, C.c_double, _dp, _dp, _dp]),
    "planets_get": (None, [C.c_void_p, _dp, _dp, _dp]),
    "planets_recalculate_rh": (None, [C.c_void_p]),
    "load_h0": (None, [C.c_uint32, _dp, _dp, C.c_uint32, _dp]),
    "planet_metrics": (None, [C.c_uint32, _dp, _dp, _dp, _dp, _dp]),
    "stable_partition": (C.c_uint64, [_u16p, C.c_uint64, C.c_uint64, C.c_int, _u64p]),
    "to_elements": (None, [C.c_double, _dp, _dp, _dp]),
    "from_elements": (None, [C.c_double, _dp, _dp, _dp]),
    "from_elements_mean": (None, [C.c_double, _dp, _dp, _dp]),
    "plhist_write": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.c_uint32, _dp, _dp, _dp, _u32p]),
    "plhist_open": (C.c_void_p, [C.c_char_p, C.c_double, C.c_double, C.c_uint32]),
    "plhist_free": (None, [C.c_void_p]),
    "plhist_state": (None, [C.c_void_p, _u32p, _dp, _u32p, _dp, _dp]),
    "plhist_block": (C.c_int, [C.c_void_p, C.c_uint32, _dp, _dp, _u32p, _u32p, _dp, _u32p, _dp, _dp]),
}
_REF_ONLY = {
    "config_normalize": (C.c_int, [C.c_char_p, C.c_char_p, C.c_uint64, _u64p]),
    "state_save": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_uint32, _dp, _dp, _dp, _u32p, C.c_uint64,
                             _dp, _dp, _u32p, _u16p, _dp]),
    "state_load": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_uint32, C.c_uint64, _u32p, _dp, _dp, _dp, _u32p,
                             _u64p, _dp, _dp, _u32p, _u16p, _dp]),
    "track_write": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.c_uint32, _dp, _dp, _dp, _u32p, C.c_uint64, _dp,
                              _dp, _u32p, C.c_int, C.c_int]),
}
_ORACLE_ONLY = {
    "set_particle_libm": (None, [C.c_int]),
    "particle_libm_counts": (None, [_u64p, _u64p]),
    "sincos_cr": (None, [C.c_uint64, _dp, _dp, _dp]),
    "drift": (C.c_int, [C.c_double, C.c_uint32, _dp, _dp, _dp, C.c_int, _u64p]),
    "hermite_log": (C.c_int, [C.c_uint32, _dp, C.c_uint64, _dp, _dp, _dp, C.c_double, C.c_double, C.c_uint32,
                              _dp, _dp]),
}


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def reference_available() -> bool:
    return REF_LIB.exists()


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"oracle status {code}: {msg}")
        self.code = code


class Oracle:
    def __init__(self, reference: bool = False):
        path = REF_LIB if reference else ORACLE_LIB
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (make -C oracle {'ref' if reference else 'oracle'})")
        self.lib = C.CDLL(str(path))
        self.prefix = "ref_" if reference else "orc_"
        self.reference = reference
        sigs = dict(_SIGS)
        sigs.update(_REF_ONLY if reference else _ORACLE_ONLY)
        for name, (res, args) in sigs.items():
            fn = getattr(self.lib, self.prefix + name)
            fn.restype = res
            fn.argtypes = args

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    def err(self) -> str:
        return self._f("last_error")().decode()

    def set_particle_libm(self, correctly_rounded: bool) -> None:
        """sin/cos of the particle drift on the calling thread: glibc (the
        reference's, default) or correctly rounded (libquadmath's sinq /
        cosq rounded to double); resets the thread's call counters."""
        if self.reference:
            raise ValueError("the reference build always calls glibc")
        self._f("set_particle_libm")(1 if correctly_rounded else 0)

    def sincos_cr(self, x):
        """The correctly rounded sin and cos of the alternative particle libm."""
        x = np.ascontiguousarray(x, np.float64)
        s, c = np.empty_like(x), np.empty_like(x)
        self._f("sincos_cr")(x.size, _p(x, _dp), _p(s, _dp), _p(c, _dp))
        return s, c

    def particle_libm_counts(self) -> tuple[int, int]:
        """(sin/cos calls in correctly rounded mode on this thread, calls in
        which glibc returns another double)."""
        calls, differ = C.c_uint64(0), C.c_uint64(0)
        self._f("particle_libm_counts")(C.byref(calls), C.byref(differ))
        return calls.value, differ.value

    # -- particles ------------------------------------------------------
    def particle_accels(self, m, rh, outer, planet_r, h0, r, flags, ids, time=0.0, deathtime=None):
        m = np.ascontiguousarray(m, np.float64)
        rh = np.ascontiguousarray(rh, np.float64)
        pr = np.ascontiguousarray(planet_r, np.float64)
        h0 = np.ascontiguousarray(h0, np.float64)
        r = np.ascontiguousarray(r, np.float64)
        n = r.shape[0]
        a = np.empty((n, 3))
        flags = np.array(flags, dtype=np.uint16, copy=True)
        ids = np.ascontiguousarray(ids, np.uint32)
        self._f("particle_accels")(len(m), _p(m, _dp), _p(rh, _dp), outer, _p(pr, _dp), _p(h0, _dp), time, n,
                                   _p(r, _dp), _p(a, _dp), _p(flags, _u16p), _p(ids, _u32p), _p(deathtime, _dp))
        return a, flags

    def particle_block(self, m, rh, outer, r_log, h0_log, t, dt, r, v, a, flags, ids, dti=None, deathtime=None):
        """One CPU time-block; returns dict of updated arrays (inputs are
        copied).  dti / deathtime are indexed by particle id."""
        m = np.ascontiguousarray(m, np.float64)
        rh = np.ascontiguousarray(rh, np.float64)
        rl = np.ascontiguousarray(r_log, np.float64)
        h0 = np.ascontiguousarray(h0_log, np.float64).reshape(-1, 3)
        steps = h0.shape[0]
        r = np.array(r, np.float64, copy=True)
        v = np.array(v, np.float64, copy=True)
        a = np.array(a, np.float64, copy=True)
        flags = np.array(flags, np.uint16, copy=True)
        ids = np.ascontiguousarray(ids, np.uint32)
        n = r.shape[0]
        size = max(n, int(ids.max()) + 1 if n else 0)
        dti = np.zeros(size, np.uint32) if dti is None else np.array(dti, np.uint32, copy=True)
        if deathtime is None:
            deathtime = np.full(size, np.nan)
        bad = C.c_uint64(0)
        rc = self._f("particle_block")(len(m), _p(m, _dp), _p(rh, _dp), outer, _p(rl, _dp), _p(h0, _dp), steps,
                                       t, dt, n, _p(r, _dp), _p(v, _dp), _p(a, _dp), _p(flags, _u16p),
                                       _p(ids, _u32p), _p(dti, _u32p), _p(deathtime, _dp), C.byref(bad))
        return {"rc": rc, "r": r, "v": v, "a": a, "flags": flags, "dti": dti, "deathtime": deathtime,
                "bad": bad.value}

    def drift(self, t, r, v, mu, fixed=True):
        r = np.array(r, np.float64, copy=True)
        v = np.array(v, np.float64, copy=True)
        n = r.shape[0]
        mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, np.float64), (n,)))
        bad = C.c_uint64(0)
        rc = self._f("drift")(t, n, _p(r, _dp), _p(v, _dp), _p(mu, _dp), int(fixed), C.byref(bad))
        return rc, r, v

    def hermite_log(self, m, t_samples, r_samples, v_samples, t_start, dt, n_steps):
        """Cubic Hermite planet log from a sampled history: (r_log, h0_log)
        for times t_start + (s+1) dt, s < n_steps."""
        m = np.ascontiguousarray(m, np.float64)
        ts = np.ascontiguousarray(t_samples, np.float64)
        rs = np.ascontiguousarray(r_samples, np.float64)
        vs = np.ascontiguousarray(v_samples, np.float64)
        nb = len(m)
        rl = np.empty((n_steps, nb - 1, 3))
        h0 = np.empty((n_steps, 3))
        rc = self._f("hermite_log")(nb, _p(m, _dp), len(ts), _p(ts, _dp), _p(rs, _dp), _p(vs, _dp), t_start, dt,
                                    n_steps, _p(rl, _dp), _p(h0, _dp))
        if rc != 0:
            raise OracleError(rc, self.err())
        return rl, h0

    # -- file formats (reference build only) ---------------------------------
    def config_normalize(self, text: str) -> str:
        """The reference's read_configuration + write_configuration."""
        buf = C.create_string_buffer(1 << 16)
        n = C.c_uint64(0)
        rc = self._f("config_normalize")(text.encode(), buf, len(buf), C.byref(n))
        if rc != 0:
            raise OracleError(rc, self.err())
        return buf.value.decode()

    def state_save(self, path, fmt: int, st: dict, write_momenta=False, particle_path=None) -> None:
        a = _state_arrays(st)
        self._f("state_save")(str(path).encode(), None if particle_path is None else str(particle_path).encode(),
                              fmt, int(write_momenta), len(a["pm"]), _p(a["pm"], _dp), _p(a["pr"], _dp),
                              _p(a["pv"], _dp), _p(a["pid"], _u32p), len(a["id"]), _p(a["r"], _dp),
                              _p(a["v"], _dp), _p(a["id"], _u32p), _p(a["flags"], _u16p), _p(a["deathtime"], _dp))

    def state_load(self, path, fmt: int, particle_path=None, cap_pl=64, cap=1 << 16) -> dict:
        out = {"pm": np.empty(cap_pl), "pr": np.empty((cap_pl, 3)), "pv": np.empty((cap_pl, 3)),
               "pid": np.empty(cap_pl, np.uint32), "r": np.empty((cap, 3)), "v": np.empty((cap, 3)),
               "id": np.empty(cap, np.uint32), "flags": np.empty(cap, np.uint16), "deathtime": np.empty(cap)}
        npl, n = C.c_uint32(0), C.c_uint64(0)
        rc = self._f("state_load")(str(path).encode(), None if particle_path is None else str(particle_path).encode(),
                                   fmt, cap_pl, cap, C.byref(npl), _p(out["pm"], _dp), _p(out["pr"], _dp),
                                   _p(out["pv"], _dp), _p(out["pid"], _u32p), C.byref(n), _p(out["r"], _dp),
                                   _p(out["v"], _dp), _p(out["id"], _u32p), _p(out["flags"], _u16p),
                                   _p(out["deathtime"], _dp))
        if rc != 0:
            raise OracleError(rc, "state_load")
        for k in ("pm", "pr", "pv", "pid"):
            out[k] = out[k][:npl.value]
        for k in ("r", "v", "id", "flags", "deathtime"):
            out[k] = out[k][:n.value]
        return out

    def track_write(self, path, time: float, st: dict, to_elements=True, barycentric=True, append=False) -> None:
        a = _state_arrays(st)
        self._f("track_write")(str(path).encode(), int(append), time, len(a["pm"]), _p(a["pm"], _dp),
                               _p(a["pr"], _dp), _p(a["pv"], _dp), _p(a["pid"], _u32p), len(a["id"]),
                               _p(a["r"], _dp), _p(a["v"], _dp), _p(a["id"], _u32p), int(to_elements),
                               int(barycentric))

    # -- planets ----------------------------------------------------------
    def planets(self, m, r, v, ids=None, cull_radius=0.5, rh_factor=0.0, outer=300.0):
        return OraclePlanets(self, m, r, v, ids, cull_radius, rh_factor, outer)

    def load_h0(self, m, r_log):
        m = np.ascontiguousarray(m, np.float64)
        rl = np.ascontiguousarray(r_log, np.float64)
        steps = rl.shape[0]
        out = np.empty((steps, 3))
        self._f("load_h0")(len(m), _p(m, _dp), _p(rl, _dp), steps, _p(out, _dp))
        return out

    def planet_metrics(self, m, r, v):
        m = np.ascontiguousarray(m, np.float64)
        r = np.ascontiguousarray(r, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        e = np.empty(1)
        lv = np.empty(3)
        self._f("planet_metrics")(len(m), _p(m, _dp), _p(r, _dp), _p(v, _dp), _p(e, _dp), _p(lv, _dp))
        return float(e[0]), lv

    def stable_partition(self, flags, begin, length, mode):
        flags = np.ascontiguousarray(flags, np.uint16)
        perm = np.empty(length, np.uint64)
        split = self._f("stable_partition")(_p(flags, _u16p), begin, length, mode, _p(perm, _u64p))
        return int(split), perm

    def to_elements(self, mu, r, v):
        r = np.ascontiguousarray(r, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        out = np.empty(6)
        self._f("to_elements")(mu, _p(r, _dp), _p(v, _dp), _p(out, _dp))
        return out

    def from_elements(self, mu, el):
        el = np.ascontiguousarray(el, np.float64)
        r, v = np.empty(3), np.empty(3)
        self._f("from_elements")(mu, _p(el, _dp), _p(r, _dp), _p(v, _dp))
        return r, v

    def from_elements_mean(self, mu, el):
        el = np.ascontiguousarray(el, np.float64)
        r, v = np.empty(3), np.empty(3)
        self._f("from_elements_mean")(mu, _p(el, _dp), _p(r, _dp), _p(v, _dp))
        return r, v

    # -- planet history file ---------------------------------------------
    def plhist_write(self, path, time: float, m, r, v, ids=None, append=False) -> None:
        m = np.ascontiguousarray(m, np.float64)
        r = np.ascontiguousarray(r, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        ids = np.arange(len(m), dtype=np.uint32) if ids is None else np.ascontiguousarray(ids, np.uint32)
        rc = self._f("plhist_write")(str(path).encode(), int(append), time, len(m), _p(m, _dp), _p(r, _dp),
                                     _p(v, _dp), _p(ids, _u32p))
        if rc != 0:
            raise OracleError(rc, self.err())

    def plhist(self, path, t0: float, dt: float, maxpl: int = 16):
        return OracleHistory(self, path, t0, dt, maxpl)


def _state_arrays(st: dict) -> dict:
    n = len(st["id"])
    return {"pm": np.ascontiguousarray(st["pm"], np.float64), "pr": np.ascontiguousarray(st["pr"], np.float64),
            "pv": np.ascontiguousarray(st["pv"], np.float64), "pid": np.ascontiguousarray(st["pid"], np.uint32),
            "r": np.ascontiguousarray(st["r"], np.float64), "v": np.ascontiguousarray(st["v"], np.float64),
            "id": np.ascontiguousarray(st["id"], np.uint32),
            "flags": np.ascontiguousarray(st.get("flags", np.zeros(n)), np.uint16),
            "deathtime": np.ascontiguousarray(st.get("deathtime", np.zeros(n)), np.float64)}


class OraclePlanets:
    def __init__(self, o: Oracle, m, r, v, ids, cull_radius, rh_factor, outer):
        self.o = o
        self.m = np.ascontiguousarray(m, np.float64)
        self.n = len(self.m)
        r = np.ascontiguousarray(r, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        self.ids = np.arange(self.n, dtype=np.uint32) if ids is None else np.ascontiguousarray(ids, np.uint32)
        self.h0_0 = np.empty(3)
        self.h = o._f("planets_new")(self.n, _p(self.m, _dp), _p(r, _dp), _p(v, _dp), _p(self.ids, _u32p),
                                     cull_radius, rh_factor, outer, _p(self.h0_0, _dp))

    def __del__(self):
        try:
            if self.h:
                self.o._f("planets_free")(self.h)
                self.h = None
        except Exception:
            pass

    def block(self, n_steps, dt):
        rl = np.empty((n_steps, self.n - 1, 3))
        vl = np.empty((n_steps, self.n - 1, 3))
        h0 = np.empty((n_steps, 3))
        rc = self.o._f("planets_block")(self.h, n_steps, dt, _p(rl, _dp), _p(vl, _dp), _p(h0, _dp))
        if rc != 0:
            raise OracleError(rc, self.o.err())
        return rl, vl, h0

    def state(self):
        r, v, rh = np.empty((self.n, 3)), np.empty((self.n, 3)), np.empty(self.n)
        self.o._f("planets_get")(self.h, _p(r, _dp), _p(v, _dp), _p(rh, _dp))
        return r, v, rh

    def recalculate_rh(self):
        self.o._f("planets_recalculate_rh")(self.h)


class OracleHistory:
    """Planet history file driven as the reference's executor drives its
    Interpolator: state() is the planet table, block(tbsize) the next block."""

    def __init__(self, o: Oracle, path, t0, dt, maxpl):
        self.o = o
        self.maxpl = maxpl
        self.h = o._f("plhist_open")(str(path).encode(), t0, dt, maxpl)
        if not self.h:
            raise OracleError(3, o.err())

    def __del__(self):
        try:
            if self.h:
                self.o._f("plhist_free")(self.h)
                self.h = None
        except Exception:
            pass

    def state(self):
        nb = np.zeros(1, np.uint32)
        m, ids = np.empty(self.maxpl), np.empty(self.maxpl, np.uint32)
        r, v = np.empty((self.maxpl, 3)), np.empty((self.maxpl, 3))
        self.o._f("plhist_state")(self.h, _p(nb, _u32p), _p(m, _dp), _p(ids, _u32p), _p(r, _dp), _p(v, _dp))
        return {"n_bodies": int(nb[0]), "m": m, "id": ids, "r": r, "v": v}

    def block(self, tbsize: int):
        t = np.zeros(1)
        dt = np.zeros(1)
        ns = np.zeros(1, np.uint32)
        nb = np.zeros(1, np.uint32)
        m, ids = np.empty(self.maxpl), np.empty(self.maxpl, np.uint32)
        rl = np.empty((tbsize, self.maxpl - 1, 3))
        h0 = np.empty((tbsize, 3))
        rc = self.o._f("plhist_block")(self.h, tbsize, _p(t, _dp), _p(dt, _dp), _p(ns, _u32p), _p(nb, _u32p),
                                       _p(m, _dp), _p(ids, _u32p), _p(rl, _dp), _p(h0, _dp))
        if rc != 0:
            raise OracleError(rc, self.o.err())
        n, b = int(ns[0]), int(nb[0])
        r_log = rl.reshape(-1)[: n * (b - 1) * 3].reshape(n, b - 1, 3)
        return {"t": float(t[0]), "dt": float(dt[0]), "n_steps": n, "n_bodies": b, "m": m[:b].copy(),
                "id": ids[:b].copy(), "r_log": r_log.copy(), "h0_log": h0[:n].copy()}
