"""Host time-block loop (glx::Executor through glx_exec_*) against the CPU
oracle's loop, and the encounter hand-off: rollback states in the queue,
flag-set equality with the CPU path, Jacobi-constant conservation of the
particles that stay on the GPU, and particles handed back after a
(stand-in) host resolution continuing exactly like the CPU path."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import oracle_run, planet_blocks

pytestmark = pytest.mark.gpu

DT = 180.0
TB = 512


@pytest.fixture(scope="module")
def native():
    from glisser_amd import native as nat
    nat.load()
    return nat


def jacobi_constant(m, r, v, rp, vp, a_p):
    """C_J of test particles in the circular restricted three-body problem
    (Sun m[0], planet m[1] on a circular orbit of radius a_p in the xy plane),
    from heliocentric states: 2 (m0/d0 + m1/d1) + 2 n (x u_y - y u_x) - |u|^2
    with barycentric x, u and n^2 = (m0 + m1) / a_p^3."""
    mu_s, mu_p = m
    M = mu_s + mu_p
    n = np.sqrt(M / a_p**3)
    x = r - mu_p * rp / M
    u = v - mu_p * vp / M
    d0 = np.linalg.norm(r, axis=1)
    d1 = np.linalg.norm(r - rp, axis=1)
    return 2 * (mu_s / d0 + mu_p / d1) + 2 * n * (x[:, 0] * u[:, 1] - x[:, 1] * u[:, 0]) - np.sum(u * u, axis=1)


def test_executor_matches_oracle_loop(native, oracle):
    """Kill mode (Resolve-Encounters 0) on an outer-solar-system swarm whose
    aphelia reach past an 80 au outer limit."""
    from glisser_amd import systems
    s = systems.outer_solar_system(700, seed=81, a_range=(20.0, 65.0), e_hi=0.4)
    h0_0, rh0, blocks = planet_blocks(oracle, s, 3, TB, DT, outer=80.0)
    cpu = oracle_run(oracle, s, h0_0, rh0, blocks, DT, 80.0, kill=True)
    with native.Executor(s.m, s.r, s.v, s.pr, s.pv, s.ids, dt=DT, tbsize=TB, outer_radius=80.0) as ex:
        ex.step(3)
        gpu = ex.download()
    assert gpu["n_alive"] == cpu["n_alive"] and gpu["t"] == 3 * TB * DT
    np.testing.assert_array_equal(gpu["id"], cpu["id"])
    np.testing.assert_array_equal(gpu["flags"], cpu["flags"])
    assert rel_err(gpu["r"], cpu["r"]).max() < 1e-12
    assert rel_err(gpu["v"], cpu["v"]).max() < 1e-12
    dead = gpu["flags"] != 0
    assert dead.any() and np.all(np.isfinite(gpu["deathtime"][dead])) and np.all(np.isnan(gpu["deathtime"][~dead]))


@pytest.fixture(scope="module")
def three_body_run(native, oracle):
    """Resolve mode on the three-body setup (planet of 1e-3 Msun on a circular
    orbit at 30 au, encounter radius 3.5 Hill radii)."""
    from glisser_amd import systems
    s = systems.three_body(3000, seed=82)
    rh_factor = 3.5  # rh = 3.5 a (1 - e) (m / 3 m0)^(1/3): 3.5 Hill radii on the circular orbit
    ex = native.Executor(s.m, s.r, s.v, s.pr, s.pv, s.ids, dt=DT, tbsize=TB, rh_factor=rh_factor,
                         resolve_encounters=True)
    records = []
    for _ in range(3):
        ex.step(1)
        records.append(ex.take_encounters())
    out = ex.download()
    return s, ex, records, out, rh_factor


def test_encounter_flags_match_cpu_path(oracle, three_body_run):
    s, ex, records, gpu, rh_factor = three_body_run
    h0_0, rh0, blocks = planet_blocks(oracle, s, 3, TB, DT, rh_factor=rh_factor)
    cpu = oracle_run(oracle, s, h0_0, rh0, blocks, DT, 300.0, kill=False)
    og, oc = np.argsort(gpu["id"]), np.argsort(cpu["id"])
    np.testing.assert_array_equal(gpu["flags"][og], cpu["flags"][oc])
    assert rel_err(gpu["r"][og], cpu["r"][oc]).max() < 1e-12
    enc = (gpu["flags"] & 0xFE) == 0
    enc &= gpu["flags"] != 0
    queued = np.concatenate([r["id"] for r in records])
    # every encounter-class particle was queued exactly once
    np.testing.assert_array_equal(np.sort(queued), np.sort(gpu["id"][enc]))
    assert len(queued) > 50
    print(f"{len(queued)} encounters queued, {int((gpu['flags'] & 0x80 != 0).sum())} dead")


def test_queue_records_are_rollback_states(native, oracle, three_body_run):
    """r0, v0 of a record are the particle's state at the start of its block:
    re-running the CPU path from them for `step` steps reproduces r, v at the
    flagging step."""
    s, ex, records, gpu, rh_factor = three_body_run
    pl = oracle.planets(s.m, s.r, s.v, cull_radius=0.5, rh_factor=rh_factor)
    logs = []
    for _ in range(3):
        pl.recalculate_rh()
        rh = pl.state()[2]
        rl, _vl, h0 = pl.block(TB, DT)
        logs.append((rl, h0, rh))
    checked = 0
    for k, rec in enumerate(records):
        if len(rec["id"]) == 0:
            continue
        assert np.all(rec["t_block"] == k * TB * DT)
        rl, h0, rh = logs[k]
        row_prev = (logs[k - 1][0][-1], logs[k - 1][1][-1]) if k else (s.r[1:], pl.h0_0)
        n = len(rec["id"])
        a0, f0 = oracle.particle_accels(s.m, rh, 300.0, row_prev[0], row_prev[1], rec["r0"], np.zeros(n), rec["id"])
        # flagged by the initial accelerations (t0, step 0): the state did not move
        at_start = rec["step"] == 0
        if at_start.any():
            assert k == 0
            np.testing.assert_array_equal(rec["r"][at_start], rec["r0"][at_start])
            np.testing.assert_array_equal(f0[at_start], rec["flags"][at_start])
        live = ~at_start
        out = oracle.particle_block(s.m, rh, 300.0, rl, h0, k * TB * DT, DT, rec["r0"][live], rec["v0"][live],
                                    a0[live], np.zeros(int(live.sum()), np.uint16), rec["id"][live])
        np.testing.assert_array_equal(out["flags"], rec["flags"][live])
        np.testing.assert_array_equal(out["dti"][rec["id"][live]], rec["step"][live])
        assert rel_err(out["r"], rec["r"][live]).max() < 1e-12
        checked += n
    assert checked > 50


def test_jacobi_constant_conserved(three_body_run):
    """Particles never flagged keep their Jacobi constant to the WH map's
    own accuracy (the CPU path gives at most ~4e-7 at this dt over 8192
    steps); a wrong sign or a missing indirect term shows as >= 1e-4."""
    s, ex, records, gpu, _ = three_body_run
    rp = ex.planet_positions()[0]
    a_p = 30.0
    vp = np.cross([0, 0, 1.0], rp) * np.sqrt((s.m[0] + s.m[1]) / a_p**3)  # circular, prograde in xy
    order = np.argsort(gpu["id"])
    r, v, f = gpu["r"][order], gpu["v"][order], gpu["flags"][order]
    c0 = jacobi_constant(s.m, s.pr, s.pv, s.r[1], s.v[1], a_p)
    c1 = jacobi_constant(s.m, r, v, rp, vp, a_p)
    ok = f == 0
    drift = np.abs(c1 - c0)[ok] / np.abs(c0[ok])
    print(f"C_J drift over {3 * TB} steps: max {drift.max():.2e}, median {np.median(drift):.2e}")
    assert ok.sum() > 1000 and drift.max() < 1e-6


def test_returned_particles_continue_like_cpu(native, oracle):
    """Queued particles handed back (here after a pure-Kepler stand-in
    resolver) rejoin the alive set with fresh accelerations and continue
    exactly like the CPU path started from the same state."""
    from glisser_amd import systems
    s = systems.three_body(1500, seed=83)
    with native.Executor(s.m, s.r, s.v, s.pr, s.pv, s.ids, dt=DT, tbsize=TB, rh_factor=3.5,
                         resolve_encounters=True) as ex:
        ex.step(1)
        rec = ex.take_encounters()
        assert len(rec["id"]) > 10
        st = ex.status()
        t_now = st["t"]
        _, r_now, v_now = oracle.drift(t_now - rec["t_block"][0], rec["r0"], rec["v0"], s.m[0])
        ex.return_particles(rec["id"], r_now, v_now)
        st2 = ex.status()
        back = ex.download()
        rp_now = ex.planet_positions()
        ex.step(1)
        after = ex.download()
    # the CPU path for the returned particles over the next block
    pl = oracle.planets(s.m, s.r, s.v, cull_radius=0.5, rh_factor=3.5)
    pl.recalculate_rh()
    rl0, _, h00 = pl.block(TB, DT)
    np.testing.assert_array_equal(rl0[-1], rp_now)
    pl.recalculate_rh()
    rh1 = pl.state()[2]
    rl1, _, h01 = pl.block(TB, DT)
    n = len(rec["id"])
    a, f = oracle.particle_accels(s.m, rh1, 300.0, rl0[-1], h00[-1], r_now, np.zeros(n), rec["id"])
    f = np.where((f & 0xFF) == 0x01, 0, f).astype(np.uint16)
    # slots of the returned particles right after the hand-back
    pos = {int(i): k for k, i in enumerate(back["id"])}
    idx = np.array([pos[int(i)] for i in rec["id"]])
    np.testing.assert_array_equal(back["flags"][idx], f)
    assert st2["n_alive"] == st["n_alive"] + int(np.sum(f == 0))
    live = f == 0
    out = oracle.particle_block(s.m, rh1, 300.0, rl1, h01, t_now, DT, r_now[live], v_now[live], a[live], f[live],
                                rec["id"][live])
    pos = {int(i): k for k, i in enumerate(after["id"])}
    idx = np.array([pos[int(i)] for i in rec["id"][live]])
    np.testing.assert_array_equal(after["flags"][idx], out["flags"])
    assert rel_err(after["r"][idx], out["r"]).max() < 1e-12
