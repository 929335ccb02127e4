"""Track records formed on the device (glx_track_pack) against the host
track writer, which writes the reference's bytes (tests/test_formats.py):
ids in id order, Cartesian and element records byte for byte (the device's
element conversion uses correctly rounded atan2, acos, asin, sin and cos,
elem_cr.h and sincos_cr.h, the values glibc returns in all but rare cases);
and the device's element conversion and those functions equal to their host
twins (the same code compiled for the CPU, pinned to the reference's writer
and to 200-bit arithmetic in tests/test_track_elements.py) bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from glisser_amd import native as nat
    nat.load()
    return nat


def _swarm(n, seed):
    from glisser_amd import systems
    return systems.outer_solar_system(n, seed=seed, a_range=(8.0, 90.0), e_hi=0.6, i_hi_deg=60.0)


def _ctx_with(native, s, ids, flags):
    n = s.n
    ctx = native.Context(n, s.n_bodies, 16)
    ctx.upload(s.pr, s.pv, flags, ids, a=np.zeros((n, 3)), dti=np.zeros(n, np.uint32))
    ctx.set_counts(n, n)
    ctx.partition()  # alive = unflagged, in slot order
    return ctx


def _host_records(native, tmp_path, s, ids, flags, to_elements, bary):
    alive = flags == 0
    order = np.argsort(ids[alive])
    st = {"pm": s.m, "pr": s.r, "pv": s.v, "pid": np.arange(s.n_bodies), "r": s.pr[alive][order],
          "v": s.pv[alive][order], "id": ids[alive][order]}
    path = tmp_path / "host.trk"
    native.write_track(path, 10.0, st, to_elements, bary)
    data = path.read_bytes()
    head = 8 + 8 + (s.n_bodies - 1) * 28 + 8
    return np.frombuffer(data[head:], dtype=native.TRACK_DTYPE), st


def _ulps(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    return np.abs(ai - bi)


@pytest.mark.parametrize("to_elements,bary", [(True, True), (True, False), (False, False)])
def test_device_records_match_host_writer(native, tmp_path, to_elements, bary):
    s = _swarm(5000, 91)
    rng = np.random.default_rng(91)
    ids = (rng.permutation(s.n) + 3).astype(np.uint32)
    flags = np.where(rng.random(s.n) < 0.2, 0x80, 0).astype(np.uint16)
    ctx = _ctx_with(native, s, ids, flags)
    try:
        dev = ctx.track_records(s.m, s.r, s.v, to_elements, bary)
    finally:
        ctx.close()
    host, st = _host_records(native, tmp_path, s, ids, flags, to_elements, bary)
    np.testing.assert_array_equal(dev["id"], host["id"])
    assert np.all(np.diff(dev["id"].astype(np.int64)) > 0)
    u = _ulps(dev["x"], host["x"])
    assert u.max() == 0, (int(np.count_nonzero(u)), int(u.max()))
    # a whole snapshot written from the device records has the host file's layout
    native.write_track_records(tmp_path / "dev.trk", 10.0, st, dev, to_elements, bary)
    a, b = (tmp_path / "dev.trk").read_bytes(), (tmp_path / "host.trk").read_bytes()
    assert len(a) == len(b) and a[:8 + 8 + (s.n_bodies - 1) * 28 + 8] == b[:8 + 8 + (s.n_bodies - 1) * 28 + 8]


def test_sparse_ids_and_errors(native):
    s = _swarm(600, 92)
    rng = np.random.default_rng(92)
    ids = rng.choice(10**9, s.n, replace=False).astype(np.uint32)  # far above 4 x capacity: host ordering
    flags = np.zeros(s.n, np.uint16)
    ctx = _ctx_with(native, s, ids, flags)
    try:
        rec = ctx.track_records(s.m, s.r, s.v, False, False)
        np.testing.assert_array_equal(rec["id"], np.sort(ids))
        order = np.argsort(ids)
        np.testing.assert_array_equal(rec["x"][:, :3], s.pr[order].astype(np.float32))
    finally:
        ctx.close()
    dup = np.arange(s.n, dtype=np.uint32)
    dup[5] = dup[6]
    ctx = _ctx_with(native, s, dup, flags)
    try:
        with pytest.raises(native.GlxError, match="not unique"):
            ctx.track_records(s.m, s.r, s.v)
    finally:
        ctx.close()


def test_rank_scan_many_tiles(native):
    # 400,000 particles with ids drawn from [0, 1.5M): 1,465 tiles of the id
    # rank scan, more than its 1,024 lanes, so lanes carry two tiles each;
    # 1 in 5 particles flagged (not in the records)
    rng = np.random.default_rng(93)
    n = 400_000
    pr = rng.normal(0.0, 30.0, (n, 3))
    pv = rng.normal(0.0, 3e-3, (n, 3))
    ids = rng.choice(1_500_000, n, replace=False).astype(np.uint32)
    flags = np.where(rng.random(n) < 0.2, 0x80, 0).astype(np.uint16)
    m = np.array([2.959122e-4, 2.8e-7])
    r_pl = np.array([[0.0, 0.0, 0.0], [5.2, 0.0, 0.0]])
    v_pl = np.array([[0.0, 0.0, 0.0], [0.0, 7.5e-3, 0.0]])
    ctx = native.Context(n, 2, 16)
    try:
        ctx.upload(pr, pv, flags, ids, a=np.zeros((n, 3)), dti=np.zeros(n, np.uint32))
        ctx.set_counts(n, n)
        ctx.partition()
        rec = ctx.track_records(m, r_pl, v_pl, False, False)
    finally:
        ctx.close()
    alive = flags == 0
    order = np.argsort(ids[alive])
    np.testing.assert_array_equal(rec["id"], ids[alive][order])
    np.testing.assert_array_equal(rec["x"][:, :3], pr[alive][order].astype(np.float32))
    np.testing.assert_array_equal(rec["x"][:, 3:], pv[alive][order].astype(np.float32))


def _same(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.int64), b[~nb].view(np.int64))


def test_device_elements_equal_host_twin(native):
    """60,000 orbits (a 8-90 au, e < 0.6, i < 60 deg) with unbound,
    near-planar, retrograde and exactly planar ones mixed in, about the Sun:
    the device's doubles equal the host twin's bit for bit."""
    s = _swarm(60_000, 93)
    rng = np.random.default_rng(93)
    k = rng.choice(s.n, 4000, replace=False)
    s.pv[k[:1000]] *= 1.6
    s.pr[k[1000:2000], 2] *= 1e-15
    s.pv[k[1000:2000], 2] *= 1e-15
    s.pv[k[2000:3000]] *= -1.0
    s.pr[k[3000:], 2] = 0.0
    s.pv[k[3000:], 2] = 0.0
    rv = np.hstack([s.pr, s.pv])
    dev = native.track_elements(s.m[0], rv, device=0)
    host = native.track_elements(s.m[0], rv)
    assert _same(dev, host), int(np.count_nonzero(dev.view(np.int64) != host.view(np.int64)))


@pytest.mark.parametrize("fn", ["atan2", "acos", "asin"])
def test_device_cr_math_equal_host_twin(native, fn):
    rng = np.random.default_rng(94)
    n = 200_000
    if fn == "atan2":
        a = rng.normal(size=n) * np.exp(rng.uniform(-30, 30, n))
        b = rng.normal(size=n) * np.exp(rng.uniform(-30, 30, n))
        a[:20] = [0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 5e-324, 1e308, 3.0] * 2
        b[:20] = [1.0, -1.0, 2.0, np.inf, 1.0, 0.0, -0.0, 1.0, 1e-300, -3.0] * 2
    else:
        a = np.concatenate([rng.uniform(-1, 1, n - 40_000), 1.0 - np.exp(rng.uniform(-36, -0.7, 20_000)),
                            -1.0 + np.exp(rng.uniform(-36, -0.7, 20_000))])
        a[:8] = [1.0, -1.0, 0.0, -0.0, 1.5, np.inf, np.nan, 5e-324]
        b = None
    dev = native.cr_math(fn, a, b, device=0)
    host = native.cr_math(fn, a, b)
    assert _same(dev, host), int(np.count_nonzero(dev.view(np.int64) != host.view(np.int64)))
