"""Batched form (glx_batch_*): several systems with different bodies, particle
counts and timesteps in one launch, checked against the CPU oracle run
system by system on the same planet logs, and for independence from the
launch shape (steps per launch, lanes, persistent / dispatch)."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import oracle_run

pytestmark = pytest.mark.gpu

T = 256       # steps per launch in these tests
LAUNCHES = 3


@pytest.fixture(scope="module")
def native():
    from glisser_amd import native as nat
    nat.load()
    return nat


@pytest.fixture(scope="module")
def members():
    """Three systems: Sun + 4 giants (dt 180 d), Sun + one Jupiter-mass planet
    at 30 au with a swarm across its orbit (encounters, dt 100 d), and a
    random 8-planet system (dt = P_inner / 20)."""
    from glisser_amd import systems
    steps = 6 * T
    out = []
    s = systems.outer_solar_system(300, seed=71)
    h = systems.planet_history(s, (steps + 2) * 180.0, spacing=365.0)
    out.append({"system": s, "m": s.m, "rh": np.full(5, 0.5), "n_particles": s.n, "dt": 180.0, "t0": 0.0,
                "outer_radius": 300.0, "history": h})
    s = systems.three_body(200, seed=72)
    rh_pl = 3.5 * 30.0 * (1e-3 / 3) ** (1 / 3)
    h = systems.planet_history(s, (steps + 2) * 100.0, spacing=200.0)
    out.append({"system": s, "m": s.m, "rh": np.array([0.5, rh_pl]), "n_particles": s.n, "dt": 100.0, "t0": 0.0,
                "outer_radius": 300.0, "history": h})
    rng = np.random.Generator(np.random.PCG64(73))
    s, dt, rh, outer = systems.random_system(rng, 150, 8)
    h = systems.planet_history(s, (steps + 2) * dt, spacing=dt, substeps=2, cull_radius=rh[0])
    out.append({"system": s, "m": s.m, "rh": rh, "n_particles": s.n, "dt": dt, "t0": 0.0, "outer_radius": outer,
                "history": h})
    return out


def _batch(native, members, steps_max=2048):
    b = native.Batch(members, steps_max=steps_max)
    for k, y in enumerate(members):
        s = y["system"]
        b.upload(k, s.pr, s.pv, np.zeros(s.n, np.uint16), s.ids)
    b.init_accel()
    return b


def _oracle(oracle, y, n_steps):
    """CPU path for one member over n_steps steps in blocks of T, on the
    same rows the device interpolates (step s at t0 + (s+1) dt)."""
    s, h = y["system"], y["history"]
    rl, h0 = oracle.hermite_log(y["m"], h.t, h.r, h.v, y["t0"], y["dt"], n_steps)
    r0, h00 = oracle.hermite_log(y["m"], h.t, h.r, h.v, y["t0"] - y["dt"], y["dt"], 1)
    np.testing.assert_array_equal(r0[0], s.r[1:])
    blocks = [(rl[k:k + T], h0[k:k + T], y["rh"]) for k in range(0, n_steps, T)]
    return oracle_run(oracle, s, h00[0], y["rh"], blocks, y["dt"], y["outer_radius"], kill=False)


def _by_id(d):
    o = np.argsort(d["id"])
    return {k: d[k][o] for k in ("r", "v", "flags", "id")}


@pytest.mark.parametrize("form,tol", [("fp64", 1e-12), ("fast", 1e-10), ("mixed", 1e-10)])
def test_batch_matches_oracle(native, oracle, members, form, tol):
    with _batch(native, members) as b:
        b.set_config(T, 256, 0, 0)
        b.run(LAUNCHES, form)
        gpu = [b.download(k) for k in range(len(members))]
        total = b.sum_steps()
    expect_steps = 0
    for k, y in enumerate(members):
        cpu = _by_id(_oracle(oracle, y, LAUNCHES * T))
        g = _by_id(gpu[k])
        np.testing.assert_array_equal(g["flags"], cpu["flags"], err_msg=f"system {k}")
        assert rel_err(g["r"], cpu["r"]).max() < tol, (k, rel_err(g["r"], cpu["r"]).max())
        assert rel_err(g["v"], cpu["v"]).max() < tol, (k, rel_err(g["v"], cpu["v"]).max())
        alive = gpu[k]["flags"] == 0
        assert np.all(gpu[k]["dti"][alive] == LAUNCHES * T)
        expect_steps += int(gpu[k]["dti"].astype(np.int64).sum())
        print(f"system {k} ({form}): {int((~alive).sum())} flagged, max rel r err "
              f"{rel_err(g['r'], cpu['r']).max():.2e}")
    assert total == expect_steps
    # the encounter system does flag particles
    assert np.count_nonzero(gpu[1]["flags"]) > 0


@pytest.mark.parametrize("form", ["fp64", "fast"])
def test_batch_shape_independent(native, members, form):
    """Steps per launch, lanes per workgroup and persistent/dispatch change
    nothing in the results (bit for bit)."""
    shapes = [((T, 256, 0, 0), LAUNCHES), ((T // 2, 64, 1, 4), 2 * LAUNCHES), ((3 * T, 128, 0, 3), 1),
              ((T, 128, 1, 2), LAUNCHES)]
    results = []
    for cfg, n in shapes:
        with _batch(native, members) as b:
            b.set_config(*cfg)
            assert b.config() == dict(zip(("steps_per_launch", "lanes", "persistent", "waves_per_simd"), cfg))
            b.run(n, form)
            results.append([b.download(k) for k in range(len(members))])
    for other in results[1:]:
        for k in range(len(members)):
            for key in ("r", "v", "flags", "dti"):
                np.testing.assert_array_equal(other[k][key], results[0][k][key], err_msg=f"system {k} {key}")


def test_batch_matches_single_system_kernel(native, oracle, members):
    """The batched kernel and the single-system kernel give the same bits on
    the same planet rows (fp64 form: identical arithmetic)."""
    y = members[0]
    s, h = y["system"], y["history"]
    rl, h0 = oracle.hermite_log(y["m"], h.t, h.r, h.v, 0.0, y["dt"], LAUNCHES * T)
    r0, h00 = oracle.hermite_log(y["m"], h.t, h.r, h.v, -y["dt"], y["dt"], 1)
    with native.Context(s.n, s.n_bodies, T) as ctx:
        ctx.upload(s.pr, s.pv, np.zeros(s.n, np.uint16), s.ids, a=np.zeros((s.n, 3)), dti=np.zeros(s.n, np.uint32))
        ctx.set_counts(s.n, s.n)
        ctx.init_accel(s.m, y["rh"], r0[0], h00[0], y["outer_radius"])
        for k in range(LAUNCHES):
            ctx.upload_planet_block(k % 2, s.m, y["rh"], rl[k * T:(k + 1) * T], h0[k * T:(k + 1) * T])
            ctx.run_block(k % 2, y["dt"], y["outer_radius"], "fp64")
        single = ctx.download(s.n)
    with _batch(native, members) as b:
        b.set_config(T, 256, 0, 0)
        b.run(LAUNCHES, "fp64")
        batch = b.download(0)
    np.testing.assert_array_equal(batch["r"], single["r"])
    np.testing.assert_array_equal(batch["v"], single["v"])
    np.testing.assert_array_equal(batch["flags"], single["flags"])


def test_batch_tune_restores_state(native, members):
    cands = [(T, 256, 0, 0), (T, 64, 1, 4), (2 * T, 128, 0, 0)]
    with _batch(native, members) as b:
        before = [b.download(k) for k in range(len(members))]
        best, rate = b.tune("fast", candidates=cands)
        assert tuple(best.values()) in cands and rate > 0
        assert b.config() == best
        after = [b.download(k) for k in range(len(members))]
        for x, y in zip(before, after):
            for key in x:
                np.testing.assert_array_equal(x[key], y[key])
        b.run(1, "fast")
        tuned = [b.download(k) for k in range(len(members))]
    with _batch(native, members) as b:
        b.set_config(*best.values())
        b.run(1, "fast")
        direct = [b.download(k) for k in range(len(members))]
    for x, y in zip(tuned, direct):
        for key in x:
            np.testing.assert_array_equal(x[key], y[key])


def test_batch_errors(native, members):
    with _batch(native, members, steps_max=2 * T) as b:
        with pytest.raises(native.GlxError):
            b.set_config(4 * T, 256, 0, 0)  # above steps_max
        with pytest.raises(native.GlxError):
            b.set_config(T, 96, 0, 0)  # lanes
        with pytest.raises(native.GlxError):
            b.set_config(T, 256, 1, 0)  # persistent needs a wave cap
        b.set_config(2 * T, 256, 0, 0)
        with pytest.raises(native.GlxError):
            b.run(100, "fast")  # runs past the histories
        b.run(1, "fast")  # still usable
