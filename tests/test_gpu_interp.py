"""Device planet-history interpolation (cubic Hermite + h0) vs the CPU
restatement of the same formula, and the particle path fed by it.  The
rows are computed without FMA contraction on both sides, so the bar is
bit-for-bit equality."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import gpu_run, oracle_run

pytestmark = pytest.mark.gpu

DT = 180.0
OUTER = 300.0


@pytest.fixture(scope="module")
def native():
    from glisser_amd import native as nat
    nat.load()
    return nat


@pytest.fixture(scope="module")
def hist_system():
    from glisser_amd import systems
    s = systems.outer_solar_system(512, seed=61)
    h = systems.planet_history(s, 3 * 2048 * DT + 365.0)
    return s, h


@pytest.mark.parametrize("t_start,n_steps", [(0.0, 2048), (2048 * DT, 2048), (12345.5, 37), (0.0, 1)])
def test_interp_block_bit_exact(native, oracle, hist_system, t_start, n_steps):
    s, h = hist_system
    rh = np.full(s.n_bodies, 0.5)
    rl_o, h0_o = oracle.hermite_log(h.m, h.t, h.r, h.v, t_start, DT, n_steps)
    with native.Context(16, s.n_bodies, 2048) as ctx:
        ctx.upload_history(h.m, h.t, h.r, h.v)
        ctx.interp_block(1, t_start, DT, n_steps, rh)
        rl_g, h0_g = ctx.download_planet_block(1, s.n_bodies, n_steps)
        times = t_start + (np.arange(n_steps) + 1.0) * DT
        rl_q, h0_q = ctx.interp_at(times)
    np.testing.assert_array_equal(rl_g, rl_o)
    np.testing.assert_array_equal(h0_g, h0_o)
    np.testing.assert_array_equal(rl_q, rl_o)
    np.testing.assert_array_equal(h0_q, h0_o)


def test_interp_nodes_and_range(native, hist_system):
    s, h = hist_system
    rh = np.full(s.n_bodies, 0.5)
    with native.Context(16, s.n_bodies, 2048) as ctx:
        ctx.upload_history(h.m, h.t, h.r, h.v)
        r, _ = ctx.interp_at(h.t[:50])
        np.testing.assert_array_equal(r, h.r[:50])
        with pytest.raises(native.GlxError):
            ctx.interp_block(0, h.t[-1] - 10 * DT, DT, 2048, rh)  # runs past the last sample
        with pytest.raises(native.GlxError):
            ctx.interp_at(np.array([h.t[0] - 1.0]))
        with pytest.raises(native.GlxError):
            ctx.upload_history(h.m, h.t[::-1], h.r, h.v)  # times must increase


def test_particles_from_device_history(native, oracle, hist_system):
    """Three blocks whose planet logs are interpolated on the device give the
    same particles, bit for bit, as the same blocks with the CPU-interpolated
    logs run through the oracle's particle path -- i.e. the full hot path
    with the device producer matches the CPU path."""
    s, h = hist_system
    rh = np.full(s.n_bodies, 0.5)
    n_blocks, tb = 3, 2048
    r0, h00 = oracle.hermite_log(h.m, h.t, h.r, h.v, -DT, DT, 1)  # state at t = 0
    blocks = []
    for k in range(n_blocks):
        rl, h0 = oracle.hermite_log(h.m, h.t, h.r, h.v, k * tb * DT, DT, tb)
        blocks.append((rl, h0, rh))
    cpu = oracle_run(oracle, s, h00[0], rh, blocks, DT, OUTER)

    with native.Context(s.n, s.n_bodies, tb) as ctx:
        ctx.upload_history(h.m, h.t, h.r, h.v)
        n = s.n
        ctx.upload(s.pr, s.pv, np.zeros(n, np.uint16), s.ids, a=np.zeros((n, 3)), dti=np.zeros(n, np.uint32))
        ctx.set_counts(n, n)
        pr0, h0_0 = ctx.interp_at(np.array([0.0]))
        np.testing.assert_array_equal(pr0[0], h.r[0])
        ctx.init_accel(s.m, rh, pr0[0], h0_0[0], OUTER)
        ctx.kill_encounters()
        ctx.partition()
        for k in range(n_blocks):
            ctx.interp_block(k % 2, k * tb * DT, DT, tb, rh)
            ctx.run_block(k % 2, DT, OUTER, "fp64")
            ctx.kill_encounters()
            ctx.partition()
        n_alive, _ = ctx.counts()
        gpu = ctx.download(n)
    assert n_alive == cpu["n_alive"]
    np.testing.assert_array_equal(gpu["id"], cpu["id"])
    np.testing.assert_array_equal(gpu["flags"], cpu["flags"])
    assert rel_err(gpu["r"], cpu["r"]).max() < 1e-12
    assert rel_err(gpu["v"], cpu["v"]).max() < 1e-12


def test_interp_blocks_equals_per_block_calls():
    """glx_interp_blocks (one launch for several blocks, 40 here: two launches
    of at most 32) fills every buffer with the rows n glx_interp_block calls
    write: bit for bit, and the particle kernel runs on them alike."""
    from glisser_amd import native, systems
    s = systems.outer_solar_system(8, seed=5)
    nblk, tb, dt = 40, 96, 180.0
    h = systems.planet_history(s, (nblk + 1) * tb * dt + 365.0, spacing=365.0, substeps=10)
    rh = np.full(s.n_bodies, 0.5)
    starts = [k * tb * dt for k in range(nblk)]
    rows = []
    for multi in (False, True):
        with native.Context(8, s.n_bodies, tb, n_planet_buffers=nblk) as ctx:
            ctx.upload_history(h.m, h.t, h.r, h.v)
            if multi:
                ctx.interp_blocks(0, starts, dt, tb, rh)
            else:
                for k, t0 in enumerate(starts):
                    ctx.interp_block(k, t0, dt, tb, rh)
            ctx.sync()
            rows.append([ctx.download_planet_block(k, s.n_bodies, tb) for k in range(nblk)])
    for (r0, h0), (r1, h1) in zip(*rows):
        np.testing.assert_array_equal(r0, r1)
        np.testing.assert_array_equal(h0, h1)
