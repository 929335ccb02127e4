"""The queued multi-block launch (glx_run_blocks_queued: one persistent grid
taking (256-particle tile, segment) tasks from a device FIFO of runnable
tasks, where finishing a tile's segment appends its next one) against the
multi-block launch it replaces (glx_run_blocks), through the C ABI.

Both run the same particles through the same device-interpolated planet
blocks; the queued launch must give every particle the same bits (r, v, a,
flags, deathtime index) in every form, for segment lengths from one radius
period (64 steps) to a whole block, blocks whose length is not a multiple of
the segment, per-block encounter radii, and a swarm in which a planet flags
particles throughout the run (encounters at 3.5 Hill radii)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 180.0
OUTER = 300.0
CULL = 0.5
SPACING = 365.0


def _setup(kind: str, n: int):
    from glisser_amd import systems
    if kind == "three_body":
        s = systems.three_body(n, seed=4)
        rh_planet = 3.5 * 30.0 * (1e-3 / 3.0) ** (1.0 / 3.0)
        rh = np.array([CULL, rh_planet])
    else:
        s = systems.outer_solar_system(n, seed=2)
        rh = np.full(s.n_bodies, CULL)
    return s, rh


def _run(s, rh, lengths, form, launch):
    """Particles through consecutive blocks of the given lengths (block k's
    encounter radii scaled by 1 + k/100), raw arrays by id before any
    compaction."""
    from glisser_amd import native, systems
    n = s.n
    h = systems.planet_history(s, sum(lengths) * DT + SPACING, spacing=SPACING, substeps=10, cull_radius=CULL)
    with native.Context(n, s.n_bodies, max(lengths), n_planet_buffers=len(lengths)) as ctx:
        ctx.upload(s.pr, s.pv, np.zeros(n, np.uint16), s.ids, a=np.zeros((n, 3)), dti=np.zeros(n, np.uint32))
        ctx.set_counts(n, n)
        ctx.upload_history(h.m, h.t, h.r, h.v)
        pr0, h00 = ctx.interp_at(np.array([0.0]))
        ctx.init_accel(s.m, rh, pr0[0], h00[0], OUTER)
        ctx.kill_encounters()
        ctx.partition()
        n_alive, _ = ctx.counts()
        t = 0.0
        for k, length in enumerate(lengths):
            ctx.interp_block(k, t, DT, length, rh * (1.0 + k / 100.0))
            t += length * DT
        launch(ctx, len(lengths))
        ctx.sync()
        d = ctx.download(n_alive)
    o = np.argsort(d["id"])
    return {k: d[k][o] for k in ("r", "v", "a", "flags", "dti", "id")}


def _sliced(form):
    return lambda ctx, nb: ctx.run_blocks(nb, 0, DT, OUTER, form, 1, 0, 1)


def _queued(form, seg):
    return lambda ctx, nb: ctx.run_blocks_queued(nb, 0, DT, OUTER, form, seg)


def _assert_same(got, ref):
    for key in ("id", "flags", "dti", "r", "v", "a"):
        np.testing.assert_array_equal(got[key], ref[key], err_msg=key)


@pytest.mark.parametrize("form", ["fp64", "fast", "mixed"])
def test_queued_equals_sliced_encounters(form):
    s, rh = _setup("three_body", 20_000)
    lengths = [2048, 1000, 576]
    ref = _run(s, rh, lengths, form, _sliced(form))
    flagged = int(np.count_nonzero(ref["flags"]))
    print(f"{form}: {flagged} of {len(ref['id'])} particles flagged during the run")
    assert flagged > 1000  # the case exercises flags and death indices
    for seg in (64, 256, 2048):
        got = _run(s, rh, lengths, form, _queued(form, seg))
        _assert_same(got, ref)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_queued_equals_sliced_swarm(form):
    s, rh = _setup("swarm", 50_000)
    lengths = [2048, 2048]
    ref = _run(s, rh, lengths, form, lambda ctx, nb: ctx.run_blocks(nb, 0, DT, OUTER, form, 4, 0, 4))
    got = _run(s, rh, lengths, form, _queued(form, 0))
    _assert_same(got, ref)


def test_queued_argument_checks():
    from glisser_amd import native
    s, rh = _setup("swarm", 300)
    with native.Context(s.n, s.n_bodies, 128, n_planet_buffers=2) as ctx:
        with pytest.raises(native.GlxError, match="multiple of 64 up to 4096"):
            ctx.run_blocks_queued(1, 0, DT, OUTER, "fast", 100)
        with pytest.raises(native.GlxError, match="holds no block"):
            ctx.run_blocks_queued(1, 0, DT, OUTER, "fast", 64)
        with pytest.raises(native.GlxError, match="n_blocks"):
            ctx.run_blocks_queued(3, 0, DT, OUTER, "fast", 64)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_run_block_above_the_cu_count_is_queued_and_equal(form):
    """glx_run_block with more 256-particle tiles than compute units runs as a
    queued launch of its one block: same bits as the plain single launch
    (glx_run_blocks with one slice and one launch per block)."""
    s, rh = _setup("three_body", 100_000)  # 391 tiles
    ref = _run(s, rh, [2048], form, _sliced(form))
    got = _run(s, rh, [2048], form, lambda ctx, nb: ctx.run_block(0, DT, OUTER, form))
    assert int(np.count_nonzero(ref["flags"])) > 1000
    _assert_same(got, ref)
