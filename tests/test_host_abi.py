"""CPU-side checks of the product library: it loads, exports every entry
point declared in include/glisser_amd.h, and its host-side producer (planet
Jacobi WH integrator, h0 rows, element conversions) is bit-identical to the
oracle -- no GPU calls."""
import re

import numpy as np
import pytest

from tests.conftest import GOLDEN, ROOT


def _declared():
    text = (ROOT / "include" / "glisser_amd.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(glx_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "glx_run_block" in names and "glx_partition" in names and "glx_planets_block" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol(glx_lib):
    from glisser_amd import native
    for name in _declared():
        assert hasattr(glx_lib, name), name
        assert name in native.SIGNATURES, f"{name} has no ctypes signature"


def test_library_is_gfx950_code_object():
    data = (ROOT / "glisser_amd" / "libglisser_amd.so").read_bytes()
    assert b"gfx950" in data


def test_product_planets_bit_exact_vs_oracle(oracle):
    from glisser_amd import native, systems
    s = systems.outer_solar_system(1, seed=51)
    for rh_factor in (0.0, 3.5):
        pp = native.Planets(s.m, s.r, s.v, rh_factor=rh_factor)
        po = oracle.planets(s.m, s.r, s.v, rh_factor=rh_factor)
        np.testing.assert_array_equal(pp.h0_initial(), po.h0_0)
        for _ in range(2):
            for x, y in zip(pp.block(1000, 180.0), po.block(1000, 180.0)):
                np.testing.assert_array_equal(x, y)
            pp.recalculate_rh()
            po.recalculate_rh()
            for x, y in zip(pp.state(), po.state()):
                np.testing.assert_array_equal(x, y)
        e, lv = pp.metrics()
        r, v, _ = pp.state()
        e2, l2 = oracle.planet_metrics(s.m, r, v)
        assert e == e2
        np.testing.assert_array_equal(lv, l2)


def test_product_planets_match_golden():
    from glisser_amd import native
    g = np.load(GOLDEN / "planets.npz")
    pp = native.Planets(g["m"], g["r"], g["v"])
    np.testing.assert_array_equal(pp.h0_initial(), g["h0_0"])
    for b in range(2):
        rl, vl, h0 = pp.block(256, float(g["dt"]))
        np.testing.assert_array_equal(rl, g["r_log"][b])
        np.testing.assert_array_equal(vl, g["v_log"][b])
        np.testing.assert_array_equal(h0, g["h0_log"][b])


def test_product_load_h0_and_elements(oracle):
    from glisser_amd import native
    g = np.load(GOLDEN / "planets.npz")
    rl = g["r_log"][0]
    np.testing.assert_array_equal(native.load_h0(g["m"], rl), g["h0_log"][0])
    np.testing.assert_array_equal(native.load_h0(g["m"], rl), oracle.load_h0(g["m"], rl))
    e = np.load(GOLDEN / "elements.npz")
    r, v = native.from_elements(float(e["mu"]), e["el"], mean_anomaly=False)
    np.testing.assert_array_equal(r, e["r"])
    np.testing.assert_array_equal(v, e["v"])
    np.testing.assert_array_equal(native.to_elements(float(e["mu"]), r, v), e["el_back"])


def test_planet_errors_are_status_codes():
    from glisser_amd import native
    with pytest.raises(native.GlxError):
        native.Planets(np.array([0.0, 1e-7]), np.zeros((2, 3)), np.zeros((2, 3)))
    m = np.array([2.959122e-4, 1e-7])
    r = np.array([[0.0, 0, 0], [10.0, 0, 0]])
    v = np.array([[0.0, 0, 0], [0.0, 0.1, 0]])  # far above escape speed
    p = native.Planets(m, r, v)
    with pytest.raises(native.GlxError, match="unbound"):
        p.block(4, 10.0)


def test_generator_is_seeded_and_in_range():
    from glisser_amd import native, systems
    a = systems.outer_solar_system(500, seed=9)
    b = systems.outer_solar_system(500, seed=9)
    np.testing.assert_array_equal(a.pr, b.pr)
    el = native.to_elements(a.m[0], a.pr, a.pv)
    assert el[:, 0].min() > 59.999 and el[:, 0].max() < 65.001
    assert el[:, 1].max() < 0.1 + 1e-9 and el[:, 2].max() < 5 * np.pi / 180 + 1e-9
    pel = native.to_elements(a.m[0] + a.m[1:], a.r[1:], a.v[1:])
    np.testing.assert_allclose(pel[:, 0], [5.20, 9.55, 19.2, 30.1], rtol=1e-12)


def test_batch_generator_properties():
    """Batched workload: 2-8 planets at >= 8 mutual Hill radii, dt = P_inner / 20,
    particles outside the outermost planet, histories sampled every dt."""
    from glisser_amd import native, systems
    members = systems.batch_systems(6, seed=5, n_range=(100, 300), n_steps=16)
    for y in members:
        s = y["system"]
        mus = s.m[1:] / s.m[0]
        pel = native.to_elements(s.m[0] + s.m[1:], s.r[1:], s.v[1:])
        a = pel[:, 0]
        assert 2 <= len(mus) <= 8 and np.all(np.diff(a) > 0)
        rh_mut = ((mus[:-1] + mus[1:]) / 3) ** (1 / 3) * (a[:-1] + a[1:]) / 2
        assert np.all(np.diff(a) >= 8 * rh_mut * 0.99)  # osculating a differs slightly from the drawn one
        p_in = 2 * np.pi * np.sqrt(s.meta["axes"][0] ** 3 / (s.m[0] + s.m[1]))
        assert abs(y["dt"] - p_in / 20) < 1e-9 * y["dt"]
        h = y["history"]
        np.testing.assert_allclose(np.diff(h.t), y["dt"], rtol=1e-12)
        assert h.t[-1] >= 17 * y["dt"]
        pa = native.to_elements(np.full(s.n, s.m[0]), s.pr, s.pv)[:, 0]
        assert pa.min() > 1.29 * s.meta["axes"][-1]


def test_batch_rejects_bad_systems():
    """Descriptor validation happens on the host, before any device work."""
    from glisser_amd import native, systems
    s = systems.three_body(8, seed=1)
    h = systems.planet_history(s, 400.0, spacing=100.0)
    good = {"m": s.m, "rh": np.array([0.5, 1.0]), "n_particles": 8, "dt": 50.0, "history": h}
    bad_cases = [
        dict(good, dt=0.0),
        dict(good, m=np.array([0.0, 1e-7])),
        dict(good, history=systems.History(h.m, h.t[::-1].copy(), h.r, h.v)),
        dict(good, t0=-1000.0),
    ]
    for case in bad_cases:
        with pytest.raises(native.GlxError):
            native.Batch([case])
    with pytest.raises(native.GlxError):
        native.Batch([])
