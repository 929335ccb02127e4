"""Planet history files (the reference's "planetary history", README.md:154-171):
the writer and the element-linear interpolator of the product host code
against the oracle restatement, the oracle against the reference's own
Interpolator and save_swift_plhist (oracle/_ref, src/interp.cpp,
src/data.cpp:817-873) bit for bit, against golden blocks the reference
wrote (tests/golden/plhist), and against what the interpolation fixes
mathematically: the records themselves at interval ends, linear elements
in between, Kepler motion for a planet whose records are Keplerian."""
import numpy as np
import pytest

from tests.conftest import GOLDEN
from tests.helpers import GM_SUN_YR, sample_history, write_plhist_bytes

KEYS = ("m", "id", "r_log", "h0_log")


def _cases():
    # (seed, history options, t0, dt, tbsize)
    return [
        (1, {}, 0.0, 0.05, 2048),
        (2, {"dropout": True}, 2.5, 0.03, 7),
        (3, {"newcomer": True}, 3.0, 0.25, 1),
        (4, {"spacing": 10.0}, 0.0, 0.5, 8),
        (5, {"e_mode": 2}, 1.25, 0.01, 64),
        (6, {"dropout": True, "newcomer": True, "spacing": 3.7}, 7.4, 0.2, 5),
    ]


def _blocks(h, tbsize, limit=400):
    out = []
    for _ in range(limit):
        try:
            out.append(h.block(tbsize))
        except Exception as e:  # end of the file
            return out, str(e)
    return out, None


def _same_block(a, b):
    for k in KEYS:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    for k in ("t", "dt", "n_steps", "n_bodies"):
        assert a[k] == b[k], k


def test_writer_matches_reference(reference, oracle, tmp_path):
    from glisser_amd import native, systems
    s = systems.outer_solar_system(10, seed=3)
    pl = reference.planets(s.m, s.r, s.v)
    states = [(0.0, s.r, s.v)]
    for k in range(1, 12):
        pl.block(10, 36.5)
        r, v, _ = pl.state()
        states.append((365.0 * k, r, v))
    files = {}
    for name, write in (("ref", reference.plhist_write), ("orc", oracle.plhist_write),
                        ("glx", lambda p, t, m, r, v, ids, append: native.write_plhist(p, t, m, r, v, ids, append))):
        path = tmp_path / f"{name}.bin"
        for k, (t, r, v) in enumerate(states):
            write(path, t, s.m, r, v, np.array([0, 5, 6, 7, 8], np.uint32), append=k > 0)
        files[name] = path.read_bytes()
    assert files["orc"] == files["ref"]
    assert files["glx"] == files["ref"]
    assert len(files["ref"]) == 32 + 12 * (32 + 4 * 32)


@pytest.mark.parametrize("case", range(len(_cases())))
def test_oracle_matches_reference(reference, oracle, tmp_path, case):
    seed, opts, t0, dt, tb = _cases()[case]
    path = tmp_path / "h.bin"
    write_plhist_bytes(path, GM_SUN_YR, sample_history(seed, **opts))
    hr, ho = reference.plhist(path, t0, dt), oracle.plhist(path, t0, dt)
    sr, so = hr.state(), ho.state()
    assert sr["n_bodies"] == so["n_bodies"]
    for k in ("m", "id", "r", "v"):
        np.testing.assert_array_equal(so[k], sr[k], err_msg=k)
    br, er = _blocks(hr, tb)
    bo, eo = _blocks(ho, tb)
    assert len(br) == len(bo) > 0 and er is not None and eo is not None
    assert "EOS reached" in er and "EOS reached" in eo
    for a, b in zip(bo, br):
        _same_block(a, b)


@pytest.mark.parametrize("case", range(len(_cases())))
def test_product_matches_oracle(oracle, glx_lib, tmp_path, case):
    from glisser_amd import native
    seed, opts, t0, dt, tb = _cases()[case]
    path = tmp_path / "h.bin"
    write_plhist_bytes(path, GM_SUN_YR, sample_history(seed, **opts))
    hg, ho = native.PlanetHistory(path, t0, dt), oracle.plhist(path, t0, dt)
    sg, so = hg.state(), ho.state()
    assert sg["n_bodies"] == so["n_bodies"]
    for k in ("m", "id", "r", "v"):
        np.testing.assert_array_equal(sg[k], so[k], err_msg=k)
    bg, eg = _blocks(hg, tb)
    bo, _ = _blocks(ho, tb)
    assert len(bg) == len(bo) > 0 and "EOS reached" in eg
    for a, b in zip(bg, bo):
        _same_block(a, b)


def test_golden_blocks(oracle, glx_lib, tmp_path):
    """Blocks the reference's own Interpolator produced (no reference build needed)."""
    from glisser_amd import native
    d = GOLDEN / "plhist"
    g = np.load(d / "blocks.npz")
    for name, h in (("oracle", oracle.plhist(d / "hist.bin", float(g["t0"]), float(g["dt"]))),
                    ("product", native.PlanetHistory(d / "hist.bin", float(g["t0"]), float(g["dt"])))):
        for k in range(int(g["n_blocks"])):
            b = h.block(int(g["tbsize"]))
            assert b["t"] == g["t"][k] and b["dt"] == g["dt_block"][k] and b["n_steps"] == g["n_steps"][k], name
            np.testing.assert_array_equal(b["r_log"], g[f"r_log_{k}"], err_msg=f"{name} block {k}")
            np.testing.assert_array_equal(b["h0_log"], g[f"h0_log_{k}"], err_msg=f"{name} block {k}")
            np.testing.assert_array_equal(b["id"], g[f"id_{k}"], err_msg=f"{name} block {k}")


def test_blocks_end_on_records_and_reproduce_them(glx_lib, tmp_path):
    """An interval of length T holds round(T / dt) equal steps, blocks never
    cross a record, and the last row of an interval is the record's own
    elements (from_elements_M of its float32 values), up to rounding."""
    from glisser_amd import native
    recs = sample_history(11, n_records=6, spacing=1.0, n_planets=3)
    path = tmp_path / "h.bin"
    write_plhist_bytes(path, GM_SUN_YR, recs)
    h = native.PlanetHistory(path, 0.0, 0.07, 16)
    n_ts = round(1.0 / 0.07)
    t_expect = 0.0
    for rec in range(1, 5):
        steps = 0
        while steps < n_ts:
            b = h.block(5)
            assert b["t"] == pytest.approx(t_expect, abs=1e-12)
            assert b["dt"] == pytest.approx(1.0 / n_ts, rel=1e-15)
            steps += b["n_steps"]
            t_expect += b["n_steps"] * b["dt"]
        assert steps == n_ts
        # last row = record `rec`
        end = recs[rec][1]
        for j, (pid, gm, a, e, inc, node, peri, mean) in enumerate(end):
            el = np.array([a, e, inc, node, peri, mean], dtype=np.float32).astype(np.float64)
            mu = np.array([GM_SUN_YR + np.float32(recs[0][1][j][1])])
            r, _ = native.from_elements(mu, el[None], mean_anomaly=True)
            np.testing.assert_allclose(b["r_log"][-1, j], r[0], rtol=2e-12, atol=1e-12)


def test_keplerian_records_give_kepler_motion(glx_lib, tmp_path):
    """Records of fixed Keplerian orbits (constant a, e, i, Omega, omega; mean
    anomaly at the two-body rate): every interpolated row is the Kepler
    orbit at that time, to the float32 precision of the records."""
    from glisser_amd import native
    recs = sample_history(12, n_records=5, spacing=1.0, n_planets=3, drift=False, e_mode=0)
    path = tmp_path / "h.bin"
    write_plhist_bytes(path, GM_SUN_YR, recs)
    h = native.PlanetHistory(path, 0.0, 0.1, 16)
    b = h.block(10)
    for s in range(b["n_steps"]):
        t = b["t"] + (s + 1) * b["dt"]
        for j, (pid, gm, a, e, inc, node, peri, mean0) in enumerate(recs[0][1]):
            a32, e32 = np.float32(a), np.float32(e)
            mu = GM_SUN_YR + np.float32(gm)
            n = np.sqrt((GM_SUN_YR + gm) / a ** 3)
            el = np.array([a32, e32, np.float32(inc), np.float32(node), np.float32(peri), np.float32(mean0) + n * t])
            r, _ = native.from_elements(np.array([mu]), el[None], mean_anomaly=True)
            np.testing.assert_allclose(b["r_log"][s, j], r[0], rtol=1e-5, atol=1e-6 * a)


def test_errors(glx_lib, tmp_path):
    from glisser_amd import native
    path = tmp_path / "h.bin"
    write_plhist_bytes(path, GM_SUN_YR, sample_history(13, n_records=3))
    with pytest.raises(native.GlxError, match="starts after t0"):
        native.PlanetHistory(path, -1.0, 0.1)
    with pytest.raises(native.GlxError, match="EOS reached"):
        native.PlanetHistory(path, 5.0, 0.1)  # past the last record
    with pytest.raises(native.GlxError, match="Max-Planet-Count"):
        native.PlanetHistory(path, 0.0, 0.1, table_size=3)
    h = native.PlanetHistory(path, 0.0, 0.5)
    for _ in range(2):  # records at 0, 1, 2: two intervals of two steps
        assert h.block(2048)["n_steps"] == 2
    with pytest.raises(native.GlxError, match="EOS reached"):
        h.block(2048)


def test_truncated_record_is_an_error(glx_lib, tmp_path):
    from glisser_amd import native
    path = tmp_path / "h.bin"
    write_plhist_bytes(path, GM_SUN_YR, sample_history(14, n_records=6))
    data = path.read_bytes()
    cut = tmp_path / "cut.bin"
    cut.write_bytes(data[: 32 + 3 * (32 + 4 * 32) + 32 + 50])  # inside the fourth record's planets
    h = native.PlanetHistory(cut, 0.0, 0.5)
    h.block(2048)
    h.block(2048)
    with pytest.raises(native.GlxError, match="truncated"):
        h.block(2048)
    first = tmp_path / "cut0.bin"
    first.write_bytes(data[: 32 + 32 + 40])  # inside the first record
    with pytest.raises(native.GlxError, match="truncated"):
        native.PlanetHistory(first, 0.0, 0.5)
