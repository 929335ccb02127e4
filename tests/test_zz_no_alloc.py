"""glx_partition_static (the partition entry point that allocates nothing)
against glx_partition on the same particles, through the C ABI: identical
arrays and class sizes, both partition modes, two calls in a row (the second
smaller, so the shared scratch holds the first call's counts), and the size
limit refused.  Device memory for the caller-owned arrays comes from the HIP
runtime the library itself uses."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_FIELDS = [("r", 3), ("v", 3), ("a", 3)]


class Hip:
    """The HIP runtime the library links (same soname: the same instance)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        self.ptrs = []

    def check(self, rc, what):
        assert rc == 0, f"{what} failed: {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.lib.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 8))), "hipMalloc")
        self.ptrs.append(p.value)
        return p.value

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        self.check(self.lib.hipMemcpy(C.c_void_p(p), arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1),
                   "hipMemcpy H2D")
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype)
        if n:
            self.check(self.lib.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(p), C.c_size_t(out.nbytes), 2),
                       "hipMemcpy D2H")
        return out

    def sync(self):
        self.check(self.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free_all(self):
        for p in self.ptrs:
            self.lib.hipFree(C.c_void_p(p))
        self.ptrs.clear()


def _particles(n, seed):
    rng = np.random.default_rng(seed)
    flags = np.zeros(n, np.uint16)
    pick = rng.random(n)
    flags[pick < 0.2] = 0x0201
    flags[(pick >= 0.2) & (pick < 0.3)] = 0x0080
    flags[(pick >= 0.3) & (pick < 0.35)] = 0x0002
    flags[(pick >= 0.35) & (pick < 0.4)] = 0x0001
    return {"r": rng.normal(size=(n, 3)), "v": rng.normal(size=(n, 3)), "a": rng.normal(size=(n, 3)),
            "flags": flags, "dti": rng.integers(0, 2048, n).astype(np.uint32),
            "id": rng.permutation(n).astype(np.uint32)}


def _device_set(hip, p, n):
    from glisser_amd import native
    d = native.DeviceParticles()
    for name, _ in _FIELDS:
        for k in range(3):
            arr = np.ascontiguousarray(p[name][:, k]) if p is not None else np.zeros(n)
            getattr(d, name)[k] = hip.put(arr)
    d.flags = hip.put(p["flags"] if p is not None else np.zeros(n, np.uint16))
    d.deathtime_index = hip.put(p["dti"] if p is not None else np.zeros(n, np.uint32))
    d.id = hip.put(p["id"] if p is not None else np.zeros(n, np.uint32))
    return d


def _read_set(hip, d, n):
    out = {}
    for name, _ in _FIELDS:
        out[name] = np.stack([hip.get(getattr(d, name)[k], n, np.float64) for k in range(3)], axis=1)
    out["flags"] = hip.get(d.flags, n, np.uint16)
    out["dti"] = hip.get(d.deathtime_index, n, np.uint32)
    out["id"] = hip.get(d.id, n, np.uint32)
    return out


def _context_partition(native, p, n, three_way):
    with native.Context(n, 2, 8) as ctx:
        ctx.upload(p["r"], p["v"], p["flags"], p["id"], a=p["a"], dti=p["dti"])
        ctx.set_counts(n, n)
        n0, n1 = ctx.partition(three_way=three_way)
        d = ctx.download(n)
    return d, n0, n1


@pytest.mark.parametrize("three_way", [False, True])
def test_partition_static_equals_partition(three_way):
    from glisser_amd import native
    lib = native.load()
    hip = Hip()
    try:
        for n, seed in ((100_000, 1 + three_way), (777, 5 + three_way)):
            p = _particles(n, seed)
            ref, n0, n1 = _context_partition(native, p, n, three_way)
            din, dout = _device_set(hip, p, n), _device_set(hip, None, n)
            sizes = hip.put(np.full(2, 0xFFFFFFFF, np.uint32))
            rc = lib.glx_partition_static(C.byref(din), C.byref(dout), n, int(three_way), C.c_void_p(sizes), None)
            assert rc == 0, lib.glx_last_error().decode()
            hip.sync()
            got = _read_set(hip, dout, n)
            cls = hip.get(sizes, 2, np.uint32)
            print(f"n={n} three_way={three_way}: classes {cls.tolist()} (glx_partition: {n0}, {n1})")
            assert int(cls[0]) == n0 and int(cls[1]) == n1
            for key in ("r", "v", "a", "flags", "dti", "id"):
                np.testing.assert_array_equal(got[key], ref[key], err_msg=key)
        # beyond the limit: refused before anything runs
        rc = lib.glx_partition_static(C.byref(din), C.byref(dout), native.PARTITION_STATIC_MAX + 1, 0,
                                      C.c_void_p(sizes), None)
        assert rc == native.GLX_EINVAL
        assert "exceed" in lib.glx_last_error().decode()
    finally:
        hip.free_all()
