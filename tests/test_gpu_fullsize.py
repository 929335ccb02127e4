"""Config 1 at its full size: 200,000 particles of the outer-solar-system
swarm (Sun + 4 giants, a ~ U[60, 65] au), planet rows interpolated on the
device from a 365-d history, two 2048-step blocks through the multi-block
launch the benchmark uses (4 particle slices on 4 streams, 4 launches per
block, no resident-wave cap).  Checked through properties that do not need
a CPU run of all 200,000 particles:
  * every particle's result is bit-identical to a 1,000-particle run of the
    same particles with another launch shape (1 slice, whole blocks, capped
    at 3 waves per SIMD: other chunking, slicing and wave grouping);
  * the fast and mixed forms agree with the fp64 form (osculating a within
    1e-12 relative, positions within 1e-10, flags identical);
  * the device step count equals 200,000 x 4096 (no particle of this swarm is
    flagged in 2 blocks);
  * a 256-particle sample equals the oracle's particle path on the same
    (CPU-interpolated) planet rows to 1e-12;
  * the queued persistent launch gives the same bits as the sliced one."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import oracle_run

pytestmark = pytest.mark.gpu

N = 200_000
TB = 2048
BLOCKS = 2
DT = 180.0
OUTER = 300.0
CULL = 0.5


@pytest.fixture(scope="module")
def swarm():
    from glisser_amd import systems
    s = systems.outer_solar_system(N, seed=1)
    h = systems.planet_history(s, BLOCKS * TB * DT + 365.0, spacing=365.0, substeps=10, cull_radius=CULL)
    return s, h, np.full(s.n_bodies, CULL)


def run(s, h, rh, form, n, shards, cap, parts, queued=False):
    from glisser_amd import native
    with native.Context(n, s.n_bodies, TB, n_planet_buffers=BLOCKS) as ctx:
        ctx.upload(s.pr[:n], s.pv[:n], np.zeros(n, np.uint16), s.ids[:n], a=np.zeros((n, 3)),
                   dti=np.zeros(n, np.uint32))
        ctx.set_counts(n, n)
        ctx.upload_history(h.m, h.t, h.r, h.v)
        pr0, h00 = ctx.interp_at(np.array([0.0]))
        ctx.init_accel(s.m, rh, pr0[0], h00[0], OUTER)
        ctx.kill_encounters()
        ctx.partition()
        n_alive, _ = ctx.counts()
        for k in range(BLOCKS):
            ctx.interp_block(k, k * TB * DT, DT, TB, rh)
        if queued:
            ctx.run_blocks_queued(BLOCKS, 0, DT, OUTER, form, 0)
        else:
            ctx.run_blocks(BLOCKS, 0, DT, OUTER, form, shards, cap, parts)
        ctx.kill_encounters()
        steps = ctx.sum_steps(n_alive)
        ctx.partition()
        ctx.sync()
        d = ctx.download(n)
    o = np.argsort(d["id"])
    return {k: d[k][o] for k in ("r", "v", "a", "flags", "id")}, steps


@pytest.fixture(scope="module")
def full(swarm):
    s, h, rh = swarm
    return {form: run(s, h, rh, form, N, 4, 0, 4) for form in ("fp64", "fast", "mixed")}


@pytest.mark.parametrize("form", ["fp64", "fast", "mixed"])
def test_fullsize_equals_small_run(swarm, full, form):
    s, h, rh = swarm
    small, _ = run(s, h, rh, form, 1000, 1, 3, 1)
    big, _ = full[form]
    for key in ("r", "v", "a", "flags", "id"):
        np.testing.assert_array_equal(big[key][:1000], small[key], err_msg=key)


@pytest.mark.parametrize("form", ["fp64", "fast", "mixed"])
def test_fullsize_queued_equals_sliced(swarm, full, form):
    """The queued persistent launch (782 tiles x 16 segments of 256 steps)
    gives every particle the bits of the sliced launch."""
    s, h, rh = swarm
    got, steps = run(s, h, rh, form, N, 0, 0, 0, queued=True)
    ref, ref_steps = full[form]
    assert steps == ref_steps
    for key in ("r", "v", "a", "flags", "id"):
        np.testing.assert_array_equal(got[key], ref[key], err_msg=key)


def test_fullsize_forms_agree(swarm, full):
    from glisser_amd import native
    s = swarm[0]
    ref = full["fp64"][0]
    ea = native.to_elements(s.m[0], ref["r"], ref["v"])
    for form in ("fast", "mixed"):
        got = full[form][0]
        np.testing.assert_array_equal(got["flags"], ref["flags"])
        eg = native.to_elements(s.m[0], got["r"], got["v"])
        da = np.abs(eg[:, 0] - ea[:, 0]) / ea[:, 0]
        er = rel_err(got["r"], ref["r"])
        print(f"{form}: max |da/a| {da.max():.2e}, max r-err {er.max():.2e}")
        assert da.max() < 1e-12 and er.max() < 1e-10


def test_fullsize_step_count(full):
    for form, (d, steps) in full.items():
        assert np.all(d["flags"] == 0), form
        assert steps == N * BLOCKS * TB, form


def test_fullsize_oracle_sample(swarm, full, oracle):
    from glisser_amd import systems
    s, h, rh = swarm
    pick = np.arange(0, N, N // 256)[:256]
    sub = systems.System(s.m, s.r, s.v, s.pr[pick], s.pv[pick], np.arange(len(pick), dtype=np.uint32), s.seed)
    r0, h00 = oracle.hermite_log(h.m, h.t, h.r, h.v, -DT, DT, 1)
    blocks = [(*oracle.hermite_log(h.m, h.t, h.r, h.v, k * TB * DT, DT, TB), rh) for k in range(BLOCKS)]
    cpu = oracle_run(oracle, sub, h00[0], rh, blocks, DT, OUTER)
    gpu = full["fp64"][0]
    oc = np.argsort(cpu["id"])
    np.testing.assert_array_equal(gpu["flags"][pick], cpu["flags"][oc])
    assert rel_err(gpu["r"][pick], cpu["r"][oc]).max() < 1e-12
    assert rel_err(gpu["v"][pick], cpu["v"][oc]).max() < 1e-12
