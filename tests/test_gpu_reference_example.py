"""The reference's three-body example end to end: the planets of
example/threebody-planets integrated into a planet history (as in
tests/test_reference_inputs.py), then the 1000 particles of
example/threebody-particles/state.in.full run through the executor on that
history with the example's settings (Time-Step 0.2, Time-Block-Size 100,
Encounter-RH-Factor 3.5, Cull-Radius 0.5), against the CPU oracle's loop
over the same blocks.  The example resolves encounters with SWIFT, which is
outside this build: the comparison runs with encounters killed, and the
resolve-mode queue is checked to hold exactly the particles the kill mode
removes for encounters."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import kill_encounters
from tests.test_reference_inputs import D, _native_writer, example_planet_history

pytestmark = pytest.mark.gpu

DT, TB, RH_FACTOR, CULL, OUTER = 0.2, 100, 3.5, 0.5, 300.0


@pytest.fixture(scope="module")
def native():
    from glisser_amd import native as nat
    nat.load()
    return nat


@pytest.fixture(scope="module")
def example(native, tmp_path_factory):
    path = tmp_path_factory.mktemp("threebody") / "plhist.out"
    example_planet_history(path, _native_writer, lambda m, r, v, ids: native.Planets(m, r, v, ids))
    st = native.load_state(D / "threebody-particles.state.in.full", 0)
    return path, st


def _radii(oracle, st):
    """Encounter radii of the planet table (src/wh.cpp:472-495)."""
    nb, m, r, v = st["n_bodies"], st["m"], st["r"], st["v"]
    rh = np.empty(nb)
    rh[0] = CULL
    for i in range(1, nb):
        el = oracle.to_elements(m[0] + m[i], r[i], v[i])
        rh[i] = RH_FACTOR * el[0] * (1 - el[1]) * (m[i] / (3 * m[0])) ** (1.0 / 3)
    return rh


def _oracle_loop(oracle, path, st, n_blocks):
    h = oracle.plhist(path, 0.0, DT)
    s0 = h.state()
    nb = s0["n_bodies"]
    pr0 = s0["r"][1:nb]
    h0_0 = oracle.load_h0(s0["m"][:nb], pr0[None])[0]
    r, v, ids = st["r"].copy(), st["v"].copy(), st["id"].copy()
    n = len(ids)
    a, flags = oracle.particle_accels(s0["m"][:nb], _radii(oracle, s0), OUTER, pr0, h0_0, r, np.zeros(n, np.uint16),
                                      ids)
    flags = kill_encounters(flags)
    split, perm = oracle.stable_partition(flags, 0, n, 0)
    r, v, a, flags, ids = r[perm], v[perm], a[perm], flags[perm], ids[perm]
    n_alive = split
    dti = np.zeros(int(ids.max()) + 1, np.uint32)
    for _ in range(n_blocks):
        rh = _radii(oracle, h.state())
        b = h.block(TB)
        out = oracle.particle_block(b["m"], rh[: b["n_bodies"]], OUTER, b["r_log"], b["h0_log"], b["t"], b["dt"],
                                    r[:n_alive], v[:n_alive], a[:n_alive], flags[:n_alive], ids[:n_alive], dti)
        assert out["rc"] == 0, oracle.err()
        sl = slice(0, n_alive)
        r[sl], v[sl], a[sl], flags[sl] = out["r"], out["v"], out["a"], kill_encounters(out["flags"])
        dti = out["dti"]
        split, perm = oracle.stable_partition(flags, 0, n_alive, 0)
        r[sl], v[sl], a[sl], flags[sl], ids[sl] = r[sl][perm], v[sl][perm], a[sl][perm], flags[sl][perm], ids[sl][perm]
        n_alive = split
    return {"r": r, "v": v, "flags": flags, "id": ids, "n_alive": n_alive}


def _executor(native, path, st, resolve):
    return native.Executor(None, None, None, st["r"], st["v"], st["id"], dt=DT, tbsize=TB, outer_radius=OUTER,
                           cull_radius=CULL, rh_factor=RH_FACTOR, resolve_encounters=resolve, planet_history=path)


def test_threebody_example_matches_oracle(native, oracle, example):
    path, st = example
    n_blocks = 20
    cpu = _oracle_loop(oracle, path, st, n_blocks)
    with _executor(native, path, st, False) as ex:
        ex.step(n_blocks)
        gpu = ex.download()
    assert gpu["t"] == pytest.approx(n_blocks * TB * DT, rel=1e-12)
    assert gpu["n_alive"] == cpu["n_alive"] < len(st["id"])  # the planet removes particles
    np.testing.assert_array_equal(gpu["id"], cpu["id"])
    np.testing.assert_array_equal(gpu["flags"], cpu["flags"])
    err = rel_err(gpu["r"][: gpu["n_alive"]], cpu["r"][: cpu["n_alive"]])
    assert np.median(err) < 1e-13 and err.max() < 1e-10


def test_threebody_example_queue_holds_the_encounters(native, example):
    """Resolve mode queues (and keeps out of the alive set) exactly the
    particles kill mode removes for planet encounters; the others evolve
    alike (test particles do not interact)."""
    path, st = example
    n_blocks = 10
    with _executor(native, path, st, False) as ex:
        ex.step(n_blocks)
        killed = ex.download()
    enc = ((killed["flags"] & 0xFF) == 0x80) & ((killed["flags"] >> 8) != 0)
    enc_ids = set(killed["id"][enc].tolist())
    with _executor(native, path, st, True) as ex:
        queued = set(ex.take_encounters()["id"].tolist())
        for _ in range(n_blocks):
            ex.step(1)
            queued |= set(ex.take_encounters()["id"].tolist())
        kept = ex.download()
    assert len(enc_ids) > 0 and queued == enc_ids
    assert kept["n_alive"] == killed["n_alive"]
    np.testing.assert_array_equal(kept["id"][: kept["n_alive"]], killed["id"][: killed["n_alive"]])
    np.testing.assert_array_equal(kept["r"][: kept["n_alive"]], killed["r"][: killed["n_alive"]])
