"""The track output's element conversion (glisser_amd/csrc/track.hip
elements_of, elem_cr.h), through its host twin: the same code the device
runs, compiled for the CPU.

* atan2, acos and asin are correctly rounded: against the same functions in
  200-bit arithmetic (mpmath) rounded to the nearest double, over wide
  ranges, the table breakpoints of the reduction (t = k/32 and k/32 +- 1/64),
  arguments near +-1, tiny and huge quotients and the special values;
* the elements written as float32 equal the reference's own track writer
  (oracle/_ref: src/convert.cpp:310-463 to_elements inside
  src/data.cpp:770-812 save_binary_track, built from the reference's
  sources) byte for byte, on the benchmark's 200,000-particle swarm (about
  the barycentre and about the Sun) and on a wide swarm (a 8-90 au, e up to
  0.6, i up to 60 degrees) with hyperbolic and near-circular, near-planar
  orbits added.
The device equals this host twin bit for bit (tests/test_gpu_track.py).
"""
import math
import tempfile
from pathlib import Path

import numpy as np
import pytest

mp = pytest.importorskip("mpmath")


@pytest.fixture(scope="module")
def native():
    from glisser_amd import native as nat
    nat.load()
    return nat


def _cr(f, *args):
    mp.mp.prec = 200
    return np.array([float(f(*[mp.mpf(float(v)) for v in vals])) for vals in zip(*args)])


def _same(a, b):
    """Equal as doubles, +0 distinct from -0, any NaN equal to any NaN."""
    a, b = np.atleast_1d(np.asarray(a, dtype=np.float64)), np.atleast_1d(np.asarray(b, dtype=np.float64))
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.int64), b[~nb].view(np.int64))


def _atan2_args(seed=0):
    rng = np.random.default_rng(seed)
    n = 6000
    y = rng.normal(size=n) * np.exp(rng.uniform(-30, 30, n))
    x = rng.normal(size=n) * np.exp(rng.uniform(-30, 30, n))
    # quotients at and around the reduction's breakpoints k/32 and k/32 +- 1/64
    k = rng.integers(0, 33, 3000)
    t = k / 32.0 + rng.choice([0.0, 1 / 64, -1 / 64], 3000) + rng.normal(size=3000) * 2.0**-40
    t = np.clip(np.abs(t), 2.0**-30, 1.0)
    big = np.exp(rng.uniform(-5, 5, 3000)) * rng.choice([-1.0, 1.0], 3000)
    yb, xb = t * big, big * rng.choice([-1.0, 1.0], 3000)
    yb[::2], xb[::2] = xb[::2].copy(), yb[::2].copy()  # both octants
    # |y| near |x|, quotients near the 2^-900 switch, near-equal tiny and huge operands
    m = 500
    e = rng.uniform(0.5, 2.0, m)
    ys = np.concatenate([e, e * 2.0**-899, e * 2.0**-901, e * 1e-300, e * 1e300])
    xs = np.concatenate([np.nextafter(e, np.inf), e, e, np.nextafter(e, 0) * 1e-300, e * -1e300])
    return np.concatenate([y, yb, ys]), np.concatenate([x, xb, xs])


def test_atan2_correctly_rounded(native):
    y, x = _atan2_args()
    got = native.cr_math("atan2", y, x)
    want = _cr(mp.atan2, y, x)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, [(y[i].hex(), x[i].hex(), got[i], want[i]) for i in bad[:5]]


def test_atan2_special_values(native):
    inf, nan = math.inf, math.nan
    pairs = [(0.0, 1.0), (-0.0, 1.0), (0.0, -1.0), (-0.0, -1.0), (0.0, 0.0), (-0.0, -0.0), (1.0, 0.0),
             (-1.0, -0.0), (inf, 1.0), (-inf, -3.0), (inf, inf), (-inf, -inf), (2.0, inf), (-2.0, -inf),
             (nan, 1.0), (1.0, nan), (5e-324, 1.0), (1.0, 5e-324), (1.7976931348623157e308, 1.0),
             (1.7976931348623157e308, 1.7976931348623157e308), (3.0, 3.0), (-3.0, 3.0), (3.0, -3.0)]
    y = np.array([p[0] for p in pairs])
    x = np.array([p[1] for p in pairs])
    got = native.cr_math("atan2", y, x)
    want = np.array([math.atan2(a, b) for a, b in pairs])
    assert _same(got, want), [(p, g, w) for p, g, w in zip(pairs, got, want) if not _same(g, w)]


@pytest.mark.parametrize("fn", ["acos", "asin"])
def test_acos_asin_correctly_rounded(native, fn):
    rng = np.random.default_rng(1)
    u = np.concatenate([
        rng.uniform(-1, 1, 6000),
        1.0 - np.exp(rng.uniform(np.log(2.0**-53), np.log(0.5), 3000)),  # near +1
        -1.0 + np.exp(rng.uniform(np.log(2.0**-53), np.log(0.5), 3000)),  # near -1
        np.exp(rng.uniform(np.log(1e-300), np.log(1e-3), 1000)) * rng.choice([-1.0, 1.0], 1000),
        np.array([0.5, -0.5, np.nextafter(1.0, 0), np.nextafter(-1.0, 0), 2.0**-1074]),
    ])
    got = native.cr_math(fn, u)
    want = _cr(getattr(mp, fn), u)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, [(u[i].hex(), got[i], want[i]) for i in bad[:5]]
    special = np.array([1.0, -1.0, 0.0, -0.0, 1.5, -2.0, math.inf, math.nan])
    lib = getattr(math, fn)
    want_s = np.array([lib(v) if abs(v) <= 1 else math.nan for v in special])
    assert _same(native.cr_math(fn, special), want_s)


def test_cr_math_argument_checks(native):
    lib = native.load()
    with pytest.raises(native.GlxError, match="GLX_CR_ATAN2"):
        native._check_host(lib.glx_cr_math_host(7, 0, None, None, None))
    with pytest.raises(native.GlxError, match="null argument"):
        native._check_host(lib.glx_cr_math_host(0, 3, None, None, None))


def _reference_records(s, bary):
    from oracle.pyoracle import Oracle
    order = np.argsort(s.ids)
    st = {"pm": s.m, "pr": s.r, "pv": s.v, "pid": np.arange(s.n_bodies, dtype=np.uint32), "r": s.pr[order],
          "v": s.pv[order], "id": s.ids[order]}
    with tempfile.TemporaryDirectory() as tmp:
        path = Path(tmp) / "t.trk"
        Oracle(reference=True).track_write(path, 0.0, st, True, bary)
        data = path.read_bytes()
    head = 8 + 8 + (s.n_bodies - 1) * 28 + 8
    return np.frombuffer(data[head:], dtype=np.dtype([("id", np.uint32), ("x", np.float32, 6)])), st


def _centre(s, bary):
    """The writer's centre (sum of r m and v m over the bodies, divided by
    the summed mass, component by component), or the Sun."""
    if not bary:
        return s.m[0], s.r[0], s.v[0]
    rs, vs, ms = [0.0] * 3, [0.0] * 3, 0.0
    for j in range(s.n_bodies):
        for c in range(3):
            rs[c] = rs[c] + float(s.r[j][c]) * float(s.m[j])
            vs[c] = vs[c] + float(s.v[j][c]) * float(s.m[j])
        ms = ms + float(s.m[j])
    return ms, np.array([v / ms for v in rs]), np.array([v / ms for v in vs])


def _wide_swarm():
    """a 8-90 au, e < 0.6, i < 60 deg, plus hyperbolic, near-circular and
    near-planar (i ~ 1e-15, i ~ pi) orbits."""
    from glisser_amd import systems
    s = systems.outer_solar_system(60_000, seed=91, a_range=(8.0, 90.0), e_hi=0.6, i_hi_deg=60.0)
    rng = np.random.default_rng(5)
    k = rng.choice(s.n, 3000, replace=False)
    s.pv[k[:1000]] *= 1.6                      # unbound
    s.pr[k[1000:2000], 2] *= 1e-15             # nearly in the plane
    s.pv[k[1000:2000], 2] *= 1e-15
    s.pv[k[2000:2500]] *= -1.0                 # retrograde
    s.pr[k[2500:], 2] = 0.0                    # exactly in the plane
    s.pv[k[2500:], 2] = 0.0
    return s


@pytest.mark.parametrize("case", ["swarm200k_bary", "swarm200k_sun", "wide_bary"])
def test_host_twin_equals_reference_writer(native, case):
    from oracle.pyoracle import reference_available
    if not reference_available():
        pytest.skip("oracle/_ref (the reference's own sources, built by oracle/) is not built")
    from glisser_amd import systems
    bary = case.endswith("bary")
    s = systems.outer_solar_system(200_000, seed=1) if case.startswith("swarm") else _wide_swarm()
    ref, st = _reference_records(s, bary)
    mu, cr, cv = _centre(s, bary)
    el = native.track_elements(mu, np.hstack([st["r"] - cr, st["v"] - cv])).astype(np.float32)
    bad = np.flatnonzero((el.view(np.int32) != ref["x"].view(np.int32)).any(axis=1))
    assert bad.size == 0, (bad.size, [(int(ref["id"][i]), el[i], ref["x"][i]) for i in bad[:3]])
