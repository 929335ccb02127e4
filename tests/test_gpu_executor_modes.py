"""glx::Executor options that reproduce the reference executor's host-side
records (src/executor.cu:729-735,816-838):

* reference_deathtime: a particle's death time is the end time of the block
  whose resync removed it from the active range, cast to float -- checked
  against the CPU path's flagging steps (a particle flagged in block b dies
  at float((b + 1) T dt); one flagged for an encounter at t0 dies with the
  first block, the others flagged at t0 never get a death time);
* encounter_log: the reference's encounter.out lines ("id death t" when the
  0x80 bit is set, else "id encounter t", t with 17 significant digits) for
  exactly the particles each resync removes;
* resync_every (Resync-Interval): the active range shrinks, death times are
  set and log lines written only every resync_every blocks, while every
  particle's trajectory stays bit for bit the same as with a resync per
  block; refused together with resolve_encounters, as the reference does."""
import numpy as np
import pytest

from tests.helpers import oracle_run, planet_blocks

pytestmark = pytest.mark.gpu

DT = 180.0
TB = 512
OUTER = 80.0


@pytest.fixture(scope="module")
def case(oracle):
    """Swarm whose aphelia cross an 80 au outer limit and whose perihelia
    reach the giants (encounter radius 3.5 Hill radii): culls and killed
    encounters in every block."""
    from glisser_amd import systems
    s = systems.outer_solar_system(1500, seed=91, a_range=(8.0, 65.0), e_hi=0.5)
    h0_0, rh0, blocks = planet_blocks(oracle, s, 6, TB, DT, rh_factor=3.5, outer=OUTER)
    cpu = oracle_run(oracle, s, h0_0, rh0, blocks, DT, OUTER, kill=True)
    return s, cpu


def _run(s, tmp_path, resync_every, reference_deathtime, blocks=6, name="enc.out"):
    from glisser_amd import native
    log = tmp_path / name
    counts = []
    with native.Executor(s.m, s.r, s.v, s.pr, s.pv, s.ids, dt=DT, tbsize=TB, outer_radius=OUTER, rh_factor=3.5,
                         resync_every=resync_every, reference_deathtime=reference_deathtime,
                         encounter_log=log) as ex:
        counts.append(ex.status()["n_alive"])
        for _ in range(blocks):
            ex.step(1)
            counts.append(ex.status()["n_alive"])
        d = ex.download()
    o = np.argsort(d["id"])
    lines = [ln.split() for ln in log.read_text().splitlines()]
    return {k: d[k][o] for k in ("r", "v", "flags", "id", "deathtime")}, counts, lines


def _expected(cpu, resync_every, n_blocks=6):
    """Reference death times and log lines from the CPU path's flagging
    steps: a particle flagged in block b leaves at the resync ending block
    ceil((b + 1) / k) k."""
    death = cpu["death_by_id"]
    flags = np.zeros(len(death), np.uint16)
    flags[cpu["id"]] = cpu["flags"]
    dt_ref = np.full(len(death), np.nan)
    lines = set()
    for pid in np.flatnonzero(death >= 0):
        if death[pid] == 0:
            # flagged at t0: an encounter stays active until the first
            # block's kill; anything else is never given a death time
            if (flags[pid] & 0x80) == 0:
                continue
            block = 0
        else:
            block = (int(death[pid]) - 1) // TB
        end_block = ((block // resync_every) + 1) * resync_every
        if end_block > n_blocks:
            continue  # removed at a later resync
        t_end = end_block * TB * DT
        dt_ref[pid] = float(np.float32(t_end))
        kind = "death" if flags[pid] & 0x80 else "encounter"
        lines.add((str(pid), kind, format(t_end, ".17g")))
    return dt_ref, lines


@pytest.mark.parametrize("resync_every", [1, 3])
def test_reference_deathtime_and_log(case, tmp_path, resync_every):
    s, cpu = case
    gpu, counts, lines = _run(s, tmp_path, resync_every, True)
    oc = np.argsort(cpu["id"])
    np.testing.assert_array_equal(gpu["flags"], cpu["flags"][oc])
    dt_ref, lines_ref = _expected(cpu, resync_every)
    n_dead = int(np.isfinite(dt_ref).sum())
    kinds = {k: sum(1 for x in lines if x[1] == k) for k in ("death", "encounter")}
    print(f"resync every {resync_every}: {n_dead} death times, log lines {kinds}, n_alive by block {counts}")
    assert n_dead > 50 and kinds["death"] > 10 and kinds["encounter"] > 10
    np.testing.assert_array_equal(np.isnan(gpu["deathtime"]), np.isnan(dt_ref))
    np.testing.assert_array_equal(gpu["deathtime"][np.isfinite(dt_ref)], dt_ref[np.isfinite(dt_ref)])
    assert len(lines) == len(lines_ref) and set(map(tuple, lines)) == lines_ref
    # the active range shrinks only at resyncs
    for b in range(1, len(counts)):
        if b % resync_every:
            assert counts[b] == counts[b - 1], (b, counts)


def test_resync_interval_keeps_trajectories(case, tmp_path):
    """Every particle's state and flags are the same bits with a resync every
    block and every third block; in the default mode (time of the flagging
    step) the death times are the same too."""
    s, _ = case
    one, c1, _ = _run(s, tmp_path, 1, False, name="a.out")
    three, c3, _ = _run(s, tmp_path, 3, False, name="b.out")
    for key in ("r", "v", "flags", "deathtime"):
        np.testing.assert_array_equal(one[key], three[key], err_msg=key)
    assert c3[3] == c1[3] and c3[6] == c1[6] and c3[1] == c3[2] == c3[0]


def test_resync_interval_refused_with_encounters(tmp_path):
    from glisser_amd import native, systems
    s = systems.three_body(10, seed=3)
    with pytest.raises(native.GlxError, match="Resync-Interval must be 1"):
        native.Executor(s.m, s.r, s.v, s.pr, s.pv, s.ids, resolve_encounters=True, resync_every=2)
