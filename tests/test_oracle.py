"""Pins the CPU oracle (oracle/wh_oracle.cpp) before it is trusted as the
checker of the HIP path:

* bit-for-bit against the reference's own CPU sources compiled here
  (oracle/_ref) on fresh seeded inputs, when that library exists;
* bit-for-bit against the golden vectors in tests/golden (written by the
  reference build), which travel with the repo;
* against mathematics that does not depend on the reference: closed-form
  two-body propagation, drift composition and reversibility, energy and
  angular-momentum invariance, a brute-force direct sum, the zero-mass
  limit (pure Kepler), std::stable_partition properties, element round
  trips.
"""
import numpy as np
import pytest

from tests.conftest import GOLDEN, rel_err
from tests.helpers import oracle_run, planet_blocks

MU = 2.959122e-4


# --------------------------------------------------------------- golden ----

def test_golden_planets(oracle):
    g = np.load(GOLDEN / "planets.npz")
    pl = oracle.planets(g["m"], g["r"], g["v"])
    np.testing.assert_array_equal(pl.h0_0, g["h0_0"])
    for b in range(2):
        rl, vl, h0 = pl.block(256, float(g["dt"]))
        np.testing.assert_array_equal(rl, g["r_log"][b])
        np.testing.assert_array_equal(vl, g["v_log"][b])
        np.testing.assert_array_equal(h0, g["h0_log"][b])
    r, v, rh = pl.state()
    np.testing.assert_array_equal(r, g["r_end"])
    np.testing.assert_array_equal(v, g["v_end"])
    np.testing.assert_array_equal(rh, g["rh"])
    e, lv = oracle.planet_metrics(g["m"], r, v)
    assert e == float(g["energy_end"])
    np.testing.assert_array_equal(lv, g["l_end"])


def test_golden_particle_block(oracle):
    g = np.load(GOLDEN / "particle_block.npz")
    a0, f0 = oracle.particle_accels(g["m"], g["rh"], float(g["outer"]), g["planet_r0"], g["h0_0"], g["r"],
                                    np.zeros(len(g["ids"])), g["ids"])
    np.testing.assert_array_equal(a0, g["a"])
    np.testing.assert_array_equal(f0, g["flags"])
    out = oracle.particle_block(g["m"], g["rh"], float(g["outer"]), g["r_log"], g["h0_log"], 0.0, float(g["dt"]),
                                g["r"], g["v"], g["a"], g["flags"], g["ids"])
    assert out["rc"] == 0
    for k in ("r", "v", "a", "flags", "dti"):
        np.testing.assert_array_equal(out[k], g[k + "_out"], err_msg=k)


def test_golden_long_run(oracle):
    """64 particles x 100,352 steps with partitions between blocks."""
    from glisser_amd.systems import System
    g = np.load(GOLDEN / "long_run.npz")
    n = int(g["n"])
    s = System(g["m"], g["pl_r"], g["pl_v"], g["r0"], g["v0"], np.arange(n, dtype=np.uint32), int(g["seed"]))
    h0_0, rh0, blocks = planet_blocks(oracle, s, int(g["n_blocks"]), int(g["tbsize"]), float(g["dt"]))
    out = oracle_run(oracle, s, h0_0, rh0, blocks, float(g["dt"]), float(g["outer"]))
    for k in ("r", "v", "a", "flags"):
        np.testing.assert_array_equal(out[k], g[k], err_msg=k)
    np.testing.assert_array_equal(out["id"], g["ids"])
    np.testing.assert_array_equal(out["dti_by_id"], g["dti_by_id"])
    assert out["n_alive"] == int(g["n_alive"])


def test_golden_encounters(oracle):
    from glisser_amd.systems import System
    g = np.load(GOLDEN / "encounters.npz")
    n = int(g["n"])
    s = System(g["m"], g["pl_r"], g["pl_v"], g["r0"], g["v0"], np.arange(n, dtype=np.uint32), int(g["seed"]))
    h0_0, rh0, blocks = planet_blocks(oracle, s, 3, 512, 30.0, cull_radius=3.0, rh_factor=3.5)
    out = oracle_run(oracle, s, h0_0, rh0, blocks, 30.0, 300.0)
    np.testing.assert_array_equal(out["flags"], g["flags"])
    np.testing.assert_array_equal(out["id"], g["ids"])
    np.testing.assert_array_equal(out["dti_by_id"], g["dti_by_id"])
    np.testing.assert_array_equal(out["r"], g["r"])
    assert out["n_alive"] == int(g["n_alive"])


def test_golden_partition(oracle):
    g = np.load(GOLDEN / "partition.npz")
    flags = g["flags"]
    s0, p0 = oracle.stable_partition(flags, 0, len(flags), 0)
    assert s0 == int(g["split_unflagged"])
    np.testing.assert_array_equal(p0, g["perm_unflagged"])
    s1, p1 = oracle.stable_partition(flags, 100, len(flags) - 100, 1)
    assert s1 == int(g["split_alive_from100"])
    np.testing.assert_array_equal(p1, g["perm_alive_from100"])


def test_golden_elements(oracle):
    g = np.load(GOLDEN / "elements.npz")
    for k in range(len(g["el"])):
        r, v = oracle.from_elements(float(g["mu"]), g["el"][k])
        np.testing.assert_array_equal(r, g["r"][k])
        np.testing.assert_array_equal(v, g["v"][k])
        np.testing.assert_array_equal(oracle.to_elements(float(g["mu"]), r, v), g["el_back"][k])


# ------------------------------------------------ reference build (live) ----

@pytest.mark.parametrize("seed", [21, 22])
def test_reference_particle_block_bit_exact(oracle, reference, seed):
    from glisser_amd import systems
    s = systems.three_body(300, seed=seed, a_range=(20.0, 40.0), e_hi=0.5)
    out = {}
    for lib in (oracle, reference):
        h0_0, rh0, blocks = planet_blocks(lib, s, 2, 300, 30.0, cull_radius=3.0, rh_factor=3.5)
        out[lib.prefix] = oracle_run(lib, s, h0_0, rh0, blocks, 30.0, 300.0, kill=False)
    a, b = out["orc_"], out["ref_"]
    for k in ("r", "v", "a", "flags", "id", "dti_by_id"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_reference_planets_bit_exact(oracle, reference):
    from glisser_amd import systems
    s = systems.outer_solar_system(1, seed=31)
    po = oracle.planets(s.m, s.r, s.v, rh_factor=3.5)
    pr = reference.planets(s.m, s.r, s.v, rh_factor=3.5)
    np.testing.assert_array_equal(po.h0_0, pr.h0_0)
    for _ in range(3):
        for x, y in zip(po.block(700, 180.0), pr.block(700, 180.0)):
            np.testing.assert_array_equal(x, y)
        po.recalculate_rh()
        pr.recalculate_rh()
        for x, y in zip(po.state(), pr.state()):
            np.testing.assert_array_equal(x, y)


def test_reference_load_h0_and_metrics(oracle, reference):
    rng = np.random.default_rng(3)
    m = np.array([MU, 1e-7, 3e-8, 4e-9])
    rl = rng.normal(scale=20.0, size=(50, 3, 3))
    np.testing.assert_array_equal(oracle.load_h0(m, rl), reference.load_h0(m, rl))
    r, v = rng.normal(size=(4, 3)), rng.normal(size=(4, 3)) * 1e-3
    e1, l1 = oracle.planet_metrics(m, r, v)
    e2, l2 = reference.planet_metrics(m, r, v)
    assert e1 == e2
    np.testing.assert_array_equal(l1, l2)


def test_reference_partition(oracle, reference):
    rng = np.random.default_rng(4)
    flags = rng.choice(np.array([0, 1, 0x80, 0x0201, 2, 0x0100], np.uint16), 5000)
    for mode in (0, 1):
        for begin in (0, 17):
            a = oracle.stable_partition(flags, begin, len(flags) - begin, mode)
            b = reference.stable_partition(flags, begin, len(flags) - begin, mode)
            assert a[0] == b[0]
            np.testing.assert_array_equal(a[1], b[1])


# ------------------------------------------------------------ mathematics ----

def _kepler_closed_form(mu, r, v, t):
    """Two-body propagation by elements: advance the mean anomaly by n t and
    convert back (independent Newton solve in numpy)."""
    r = np.asarray(r, float)
    v = np.asarray(v, float)
    h = np.cross(r, v)
    rn = np.linalg.norm(r)
    evec = np.cross(v, h) / mu - r / rn
    e = np.linalg.norm(evec)
    a = 1.0 / (2.0 / rn - v @ v / mu)
    n = np.sqrt(mu / a**3)
    # eccentric anomaly from r and r.v
    E0 = np.arctan2((r @ v) / np.sqrt(mu * a), 1.0 - rn / a)
    M = E0 - e * np.sin(E0) + n * t
    E = M
    for _ in range(50):
        E = E - (E - e * np.sin(E) - M) / (1.0 - e * np.cos(E))
    dE = E - E0
    f = 1.0 - a / rn * (1.0 - np.cos(dE))
    g = t - (dE - np.sin(dE)) / n
    r1 = f * r + g * v
    r1n = np.linalg.norm(r1)
    fd = -np.sqrt(mu * a) / (r1n * rn) * np.sin(dE)
    gd = 1.0 - a / r1n * (1.0 - np.cos(dE))
    return r1, fd * r + gd * v


def _swarm(n, seed, a=(5.0, 60.0), e_hi=0.5):
    from glisser_amd import systems
    return systems.outer_solar_system(n, seed=seed, a_range=a, e_hi=e_hi, i_hi_deg=60.0)


def test_drift_matches_closed_form(oracle):
    s = _swarm(50, 41)
    for t in (30.0, 180.0, 1000.0):
        rc, r1, v1 = oracle.drift(t, s.pr, s.pv, MU, fixed=True)
        assert rc == 0
        for k in range(s.n):
            rr, vv = _kepler_closed_form(MU, s.pr[k], s.pv[k], t)
            assert np.linalg.norm(r1[k] - rr) / np.linalg.norm(rr) < 1e-12
            assert np.linalg.norm(v1[k] - vv) / np.linalg.norm(vv) < 1e-11


def test_drift_composition_and_reversibility(oracle):
    s = _swarm(64, 42)
    _, ra, va = oracle.drift(180.0, s.pr, s.pv, MU)
    _, rb, vb = oracle.drift(180.0, ra, va, MU)
    _, rc_, vc = oracle.drift(360.0, s.pr, s.pv, MU)
    assert rel_err(rb, rc_).max() < 1e-12 and rel_err(vb, vc).max() < 1e-12
    _, rback, vback = oracle.drift(-180.0, ra, va, MU)
    assert rel_err(rback, s.pr).max() < 1e-13 and rel_err(vback, s.pv).max() < 1e-13


def test_drift_conserves_energy_and_angular_momentum(oracle):
    s = _swarm(64, 43)
    _, r1, v1 = oracle.drift(500.0, s.pr, s.pv, MU)

    def energy(r, v):
        return 0.5 * np.sum(v * v, axis=1) - MU / np.linalg.norm(r, axis=1)

    e0, e1 = energy(s.pr, s.pv), energy(r1, v1)
    assert np.max(np.abs((e1 - e0) / e0)) < 1e-13
    assert rel_err(np.cross(r1, v1), np.cross(s.pr, s.pv)).max() < 1e-13


def test_drift_flags_unbound(oracle):
    r = np.array([[50.0, 0, 0]])
    v = np.array([[0, np.sqrt(2.5 * MU / 50.0), 0]])
    rc, _, _ = oracle.drift(180.0, r, v, MU)
    assert rc == 1


def test_accel_brute_force(oracle):
    """Direct sum a = h0 - sum_j m_j (r - r_j) / |r - r_j|^3 and the
    encounter / cull / outer flags, against numpy brute force."""
    rng = np.random.default_rng(44)
    m = np.array([MU, 3e-7, 1e-7, 5e-8])
    pr = rng.normal(scale=20.0, size=(3, 3))
    h0 = rng.normal(size=3) * 1e-9
    rh = np.array([1.0, 4.0, 3.0, 2.0])
    r = np.vstack([rng.normal(scale=30.0, size=(200, 3)), pr[1] + 0.5, pr[0] - 2.0, [[0.3, 0.2, 0.1]],
                   [[400.0, 0.0, 0.0]]])
    ids = np.arange(len(r), dtype=np.uint32)
    a, flags = oracle.particle_accels(m, rh, 300.0, pr, h0, r, np.zeros(len(r)), ids)
    for k in range(len(r)):
        d = r[k] - pr
        rad = np.sum(d * d, axis=1)
        expect = h0 - np.sum(d * (m[1:] / (rad * np.sqrt(rad)))[:, None], axis=0)
        np.testing.assert_allclose(a[k], expect, rtol=1e-12, atol=1e-25)
        f = 0
        for j in range(3):  # last planet within rh wins, high byte = index
            if rad[j] < rh[j + 1] ** 2:
                f = ((j + 1) << 8) | 1
        if r[k] @ r[k] < rh[0] ** 2:
            f = 1
        if r[k] @ r[k] > 300.0 ** 2:
            f = 2
        assert flags[k] == f, (k, flags[k], f)
    assert flags[-4] == (2 << 8 | 1) and flags[-3] == (1 << 8 | 1) and flags[-2] == 1 and flags[-1] == 2


def test_zero_mass_planets_reduce_to_kepler(oracle):
    """With massless planets and no indirect term the WH map is the exact
    Kepler flow: 4096 kick-drift-kick steps == closed-form propagation."""
    s = _swarm(32, 45, a=(20.0, 60.0), e_hi=0.3)
    m = np.array([MU, 0.0, 0.0])
    steps = 4096
    rl = np.full((steps, 2, 3), 1.0e4)
    h0 = np.zeros((steps, 3))
    rh = np.full(3, 0.5)
    a0 = np.zeros((s.n, 3))
    out = oracle.particle_block(m, rh, 1e5, rl, h0, 0.0, 180.0, s.pr, s.pv, a0, np.zeros(s.n), s.ids)
    assert out["rc"] == 0
    for k in range(s.n):
        rr, vv = _kepler_closed_form(MU, s.pr[k], s.pv[k], steps * 180.0)
        assert np.linalg.norm(out["r"][k] - rr) / np.linalg.norm(rr) < 1e-10


def test_stable_partition_properties(oracle):
    rng = np.random.default_rng(46)
    for n in (0, 1, 7, 1000):
        flags = rng.choice(np.array([0, 1, 0x80, 0x0301, 4], np.uint16), n)
        for mode, keep in ((0, lambda f: (f & 0xFF) == 0), (1, lambda f: (f & 0xFE) == 0)):
            split, perm = oracle.stable_partition(flags, 0, n, mode)
            expect = [i for i in range(n) if keep(flags[i])] + [i for i in range(n) if not keep(flags[i])]
            np.testing.assert_array_equal(perm, np.array(expect, dtype=np.uint64))
            assert split == sum(keep(f) for f in flags)


def test_elements_round_trip(oracle):
    rng = np.random.default_rng(47)
    for _ in range(100):
        el = np.array([rng.uniform(1, 80), rng.uniform(0.01, 0.9), rng.uniform(0.05, 3.0),
                       rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi), rng.uniform(-3.0, 3.0)])
        r, v = oracle.from_elements(MU, el)
        back = oracle.to_elements(MU, r, v)
        assert abs(back[0] - el[0]) / el[0] < 1e-12
        assert abs(back[1] - el[1]) < 1e-11
        assert abs(back[2] - el[2]) < 1e-11
        ang = lambda x: np.angle(np.exp(1j * x))  # noqa: E731
        assert abs(ang(back[3] - el[3])) < 1e-9
        assert abs(ang(back[4] + back[5] - el[4] - el[5])) < 1e-9
    # circular, equatorial special cases
    r, v = oracle.from_elements(MU, np.array([10.0, 0.0, 0.0, 0.0, 0.0, 0.7]))
    back = oracle.to_elements(MU, r, v)
    assert back[1] < 1e-13 and back[2] == 0.0 and abs(back[5] - 0.7) < 1e-12


# ----------------------------------------------------- planet history ----

def _cubic_history(rng, t, n_planets):
    """Planet tracks that are exact cubics in t: r(t) = c0 + c1 t + c2 t^2 +
    c3 t^3 per component, v = r'(t)."""
    c = rng.normal(size=(4, n_planets, 3)) * np.array([10.0, 1e-2, 1e-5, 1e-8])[:, None, None]

    def r_of(tt):
        tt = np.asarray(tt)[..., None, None]
        return c[0] + c[1] * tt + c[2] * tt**2 + c[3] * tt**3

    def v_of(tt):
        tt = np.asarray(tt)[..., None, None]
        return c[1] + 2 * c[2] * tt + 3 * c[3] * tt**2

    return r_of, v_of


def test_hermite_exact_for_cubics(oracle):
    """Cubic Hermite reproduces any cubic trajectory (to rounding): a
    dropped basis term, a missing h factor on the velocities or swapped
    end points all fail this."""
    rng = np.random.default_rng(51)
    t = np.cumsum(rng.uniform(200.0, 500.0, 40))  # irregular sample times
    m = np.array([MU, 1e-6, 2e-6, 3e-7])
    r_of, v_of = _cubic_history(rng, t, 3)
    dt = 37.0
    t_start = t[0] - dt + 0.5
    n_steps = int((t[-1] - t[0] - 1.0) / dt)
    rl, h0 = oracle.hermite_log(m, t, r_of(t), v_of(t), t_start, dt, n_steps)
    times = t_start + (np.arange(n_steps) + 1.0) * dt
    exact = r_of(times)
    scale = np.abs(exact).max()
    assert np.abs(rl - exact).max() / scale < 1e-13
    # h0 is load_h0 of the interpolated rows, bit for bit
    np.testing.assert_array_equal(h0, oracle.load_h0(m, rl))


def test_hermite_reproduces_nodes(oracle):
    """Rows that fall on a sample time return the sample exactly (u = 0)."""
    from glisser_amd import systems
    s = systems.outer_solar_system(1, seed=3)
    h = systems.planet_history(s, 20 * 365.0, spacing=365.0)
    dt = 365.0 / 4  # 91.25: exact in binary, so t_start + (s+1) dt hits the nodes
    rl, _ = oracle.hermite_log(h.m, h.t, h.r, h.v, -dt, dt, 80)
    for k in range(20):  # row s is at t = s dt
        np.testing.assert_array_equal(rl[4 * k], h.r[k])


def test_hermite_fourth_order_on_kepler_orbits(oracle):
    """On real orbits (sampled with the pinned Kepler drift) the error at
    interval midpoints falls ~16x when the spacing halves."""
    from glisser_amd import systems
    s = systems.outer_solar_system(1, seed=5)
    r0, v0 = s.r[1:], s.v[1:]
    mu = s.m[0] + s.m[1:]
    errs = []
    for spacing in (200.0, 100.0, 50.0):
        t = np.arange(0.0, 4001.0, spacing)
        rs = np.empty((len(t), 4, 3))
        vs = np.empty((len(t), 4, 3))
        for k, tk in enumerate(t):
            _, rs[k], vs[k] = oracle.drift(tk, r0, v0, mu, fixed=False)
        dt = spacing
        rl, _ = oracle.hermite_log(s.m, t, rs, vs, -dt / 2, dt, len(t) - 1)  # midpoints
        mid = t[:-1] + dt / 2
        exact = np.array([oracle.drift(tm, r0, v0, mu, fixed=False)[1] for tm in mid])
        errs.append(np.abs(rl - exact).max())
    for coarse, fine in zip(errs, errs[1:]):
        assert 12.0 < coarse / fine < 20.0, errs


def test_hermite_rejects_short_history(oracle):
    from oracle.pyoracle import OracleError
    with pytest.raises(OracleError):
        oracle.hermite_log(np.array([MU, 1e-6]), np.array([0.0]), np.zeros((1, 1, 3)), np.zeros((1, 1, 3)),
                           0.0, 1.0, 1)
