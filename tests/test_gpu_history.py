"""The executor reading a planet history file (Read-Planet-History): blocks
that follow the file's records (variable block length and step, a planet
dropping out and returning) against the CPU oracle's loop over the same
blocks, through the C ABI."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import GM_SUN_YR, kill_encounters, sample_history, write_plhist_bytes

pytestmark = pytest.mark.gpu

CULL = 0.5
OUTER = 100.0


@pytest.fixture(scope="module")
def native():
    from glisser_amd import native as nat
    nat.load()
    return nat


def _particles(native, n, seed):
    rng = np.random.default_rng(seed)
    el = np.column_stack([rng.uniform(20, 45, n), rng.uniform(0, 0.6, n), rng.uniform(0, 0.3, n),
                          rng.uniform(-np.pi, np.pi, n), rng.uniform(-np.pi, np.pi, n), rng.uniform(-np.pi, np.pi, n)])
    r, v = native.from_elements(np.full(n, GM_SUN_YR), el, mean_anomaly=True)
    return r, v, (np.arange(n) * 3 + 11).astype(np.uint32)


def _oracle_loop(oracle, path, t0, dt, tb, r, v, ids, n_blocks):
    """The reference CPU path with interpolated planets: Executor::init's
    initial accelerations from the planets at t0, then per block the
    interpolated planet log, the particle time-block, the encounter kill and
    the stable partition (src/executor.cu:93-133,226-299,598-877)."""
    h = oracle.plhist(path, t0, dt)
    st = h.state()
    nb = st["n_bodies"]
    m0 = st["m"][:nb]
    pr0 = st["r"][1:nb]
    h0_0 = oracle.load_h0(m0, pr0[None])[0]
    n = len(ids)
    r, v, ids = r.copy(), v.copy(), ids.copy()
    a, flags = oracle.particle_accels(m0, np.full(nb, CULL), OUTER, pr0, h0_0, r, np.zeros(n, np.uint16), ids)
    flags = kill_encounters(flags)
    split, perm = oracle.stable_partition(flags, 0, n, 0)
    r, v, a, flags, ids = r[perm], v[perm], a[perm], flags[perm], ids[perm]
    n_alive = split
    dti = np.zeros(int(ids.max()) + 1, np.uint32)
    t_end = t0
    for _ in range(n_blocks):
        b = h.block(tb)
        out = oracle.particle_block(b["m"], np.full(b["n_bodies"], CULL), OUTER, b["r_log"], b["h0_log"], b["t"],
                                    b["dt"], r[:n_alive], v[:n_alive], a[:n_alive], flags[:n_alive], ids[:n_alive],
                                    dti)
        assert out["rc"] == 0, oracle.err()
        sl = slice(0, n_alive)
        r[sl], v[sl], a[sl], flags[sl] = out["r"], out["v"], out["a"], kill_encounters(out["flags"])
        dti = out["dti"]
        split, perm = oracle.stable_partition(flags, 0, n_alive, 0)
        r[sl], v[sl], a[sl], flags[sl], ids[sl] = r[sl][perm], v[sl][perm], a[sl][perm], flags[sl][perm], ids[sl][perm]
        n_alive = split
        t_end = b["t"] + b["n_steps"] * b["dt"]
    return {"r": r, "v": v, "flags": flags, "id": ids, "n_alive": n_alive, "t": t_end}


@pytest.mark.parametrize("t0,dt,tb", [(0.0, 0.5, 16), (30.0, 0.35, 2048)])
def test_executor_with_planet_history_matches_oracle(native, oracle, tmp_path, t0, dt, tb):
    path = tmp_path / "plhist.bin"
    write_plhist_bytes(path, GM_SUN_YR, sample_history(31, n_records=16, spacing=20.0, n_planets=4, dropout=True,
                                                       e_mode=0))
    r, v, ids = _particles(native, 1500, 32)
    n_blocks = 12
    cpu = _oracle_loop(oracle, path, t0, dt, tb, r, v, ids, n_blocks)
    with native.Executor(None, None, None, r, v, ids, t0=t0, dt=dt, tbsize=tb, outer_radius=OUTER,
                         cull_radius=CULL, planet_history=path) as ex:
        ex.step(n_blocks)
        gpu = ex.download()
    assert gpu["t"] == cpu["t"]
    assert gpu["n_alive"] == cpu["n_alive"]
    np.testing.assert_array_equal(gpu["id"], cpu["id"])
    # the GPU kernel writes the planet id into the high byte of an encounter
    # flag, the CPU path the planet's index in the block (src/wh.cu:151-170
    # vs src/wh.cpp:640); the history's ids are not indices
    np.testing.assert_array_equal(gpu["flags"] & 0xFF, cpu["flags"] & 0xFF)
    enc = (cpu["flags"] >> 8) != 0
    assert np.all(np.isin(gpu["flags"][enc] >> 8, [100, 107, 114, 121]))
    np.testing.assert_array_equal(gpu["flags"][~enc], cpu["flags"][~enc])
    assert rel_err(gpu["r"], cpu["r"]).max() < 1e-12
    assert rel_err(gpu["v"], cpu["v"]).max() < 1e-12
    assert (gpu["flags"] != 0).any()  # some particles were culled on the way


def test_history_bodies_follow_the_file(native, tmp_path):
    """A planet absent from some records leaves the planet set for the
    intervals that miss it and returns afterwards."""
    path = tmp_path / "plhist.bin"
    write_plhist_bytes(path, GM_SUN_YR, sample_history(33, n_records=16, spacing=20.0, n_planets=3, dropout=True,
                                                       e_mode=0))
    r, v, ids = _particles(native, 64, 34)
    counts = []
    with native.Executor(None, None, None, r, v, ids, dt=5.0, tbsize=4, outer_radius=OUTER, cull_radius=CULL,
                         planet_history=path) as ex:
        for _ in range(15):  # 16 records: 15 intervals of one block each
            ex.step(1)
            counts.append(ex.bodies()["n_bodies"])
            assert ex.planet_positions().shape == (counts[-1] - 1, 3)
        assert ex.status()["t"] == 300.0
        with pytest.raises(native.GlxError, match="planet history ended"):
            ex.step(1)
    # records 10..13 miss the first planet: the intervals touching them have 3 bodies, the rest 4
    assert counts[0] == 4 and min(counts) == 3 and counts[-1] == 4
    assert counts == [4] * 9 + [3] * 5 + [4]
