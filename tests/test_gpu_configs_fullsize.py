"""The benchmark configurations beyond the headline at their full sizes.

* Mixed-precision form over more than 1e5 years: the 200,000-particle
  outer-solar-system swarm, 100 blocks of 2048 steps (204,800 steps of
  180 d), planet rows interpolated on the device.  The mixed and fast forms
  against the fp64 form over all 200,000 particles (a, e, i within 1e-9,
  flags identical); a 256-particle sample against the CPU oracle: the fp64
  form bit for bit equal to the CPU path with correctly rounded sin/cos,
  the mixed and fast forms within 1e-9 in a, e and i of the CPU path.
* Batched uneven systems: the 64-system batch (2-8 planets, 1e3-1e6
  particles each, about 9.4 million in all, dt = P_inner / 20) in one
  launch of 2048 steps per form; the first 256 particles of every system
  through the CPU oracle on the same interpolated rows: fp64 bit for bit
  (correctly rounded sin/cos), fast and mixed within 1e-10 in position,
  flags identical except within 1e-9 au of a threshold.
* Encounter hand-off at 1,000,000 particles: Sun + a 1e-3 Msun planet on a
  circular orbit at 30 au, particles a ~ U[25, 40] au, e < 0.3, i < 10 deg,
  encounter radius 3.5 Hill radii, Resolve-Encounters on, three blocks
  through glx::Executor; a sample of 2,000 particles against the CPU path
  (fp64 bit for bit with correctly rounded sin/cos, flags identical; mixed
  within 1e-10, flag differences only within 1e-9 au of a threshold), every
  encounter-class particle queued once, and the Jacobi constant of the
  particles never flagged conserved to the WH map's accuracy."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import oracle_run, oracle_run_parallel, planet_blocks, unexplained_flag_mismatches

pytestmark = pytest.mark.gpu

DT = 180.0
TB = 2048
OUTER = 300.0
CULL = 0.5


def _elements_diff(mu, r0, v0, r1, v1):
    from glisser_amd import native
    e0 = native.to_elements(mu, r0, v0)
    e1 = native.to_elements(mu, r1, v1)
    return (np.abs(e1[:, 0] - e0[:, 0]) / e0[:, 0], np.abs(e1[:, 1] - e0[:, 1]), np.abs(e1[:, 2] - e0[:, 2]))


# ------------------------------------------------------- mixed, 1e5 years --

BLOCKS_1E5 = 100  # 204,800 steps of 180 d: 1.009e5 years


def test_mixed_form_1e5_years_200k(oracle):
    from glisser_amd import native, systems
    n = 200_000
    s = systems.outer_solar_system(n, seed=1)
    hist = systems.planet_history(s, BLOCKS_1E5 * TB * DT + 365.0, spacing=365.0, substeps=10, cull_radius=CULL)
    rh = np.full(s.n_bodies, CULL)
    print(f"{n} particles, {BLOCKS_1E5 * TB} steps = {BLOCKS_1E5 * TB * DT / 365.25:.0f} years", flush=True)
    final = {}
    for form in ("fp64", "fast", "mixed"):
        with native.Context(n, s.n_bodies, TB, n_planet_buffers=BLOCKS_1E5) as ctx:
            ctx.upload(s.pr, s.pv, np.zeros(n, np.uint16), s.ids, a=np.zeros((n, 3)), dti=np.zeros(n, np.uint32))
            ctx.set_counts(n, n)
            ctx.upload_history(hist.m, hist.t, hist.r, hist.v)
            pr0, h00 = ctx.interp_at(np.array([0.0]))
            ctx.init_accel(s.m, rh, pr0[0], h00[0], OUTER)
            ctx.kill_encounters()
            ctx.partition()
            for k in range(BLOCKS_1E5):
                ctx.interp_block(k, k * TB * DT, DT, TB, rh)
            ctx.run_blocks(BLOCKS_1E5, 0, DT, OUTER, form, 4, 3 if form == "fp64" else 0, 4)
            ctx.kill_encounters()
            ctx.partition()
            ctx.sync()
            d = ctx.download(n)
        o = np.argsort(d["id"])
        final[form] = {k: d[k][o] for k in ("r", "v", "flags", "id")}
        print(f"{form} done", flush=True)
    ref = final["fp64"]
    for form in ("fast", "mixed"):
        got = final[form]
        np.testing.assert_array_equal(got["flags"], ref["flags"])
        da, de, di = _elements_diff(s.m[0], ref["r"], ref["v"], got["r"], got["v"])
        print(f"{form} vs fp64 over {n}: |da/a| max {da.max():.2e} median {np.median(da):.2e}, |de| max "
              f"{de.max():.2e}, |di| max {di.max():.2e}")
        assert da.max() < 1e-9 and de.max() < 1e-9 and di.max() < 1e-9
    # the CPU path over a sample, on the rows the device interpolates
    pick = np.arange(0, n, n // 256)[:256]
    sub = systems.System(s.m, s.r, s.v, s.pr[pick], s.pv[pick], np.arange(len(pick), dtype=np.uint32), s.seed)
    _, h00c = oracle.hermite_log(hist.m, hist.t, hist.r, hist.v, -DT, DT, 1)
    blocks = [(*oracle.hermite_log(hist.m, hist.t, hist.r, hist.v, k * TB * DT, DT, TB), rh)
              for k in range(BLOCKS_1E5)]
    cr = oracle_run_parallel(oracle, sub, h00c[0], rh, blocks, DT, OUTER, correctly_rounded=True, chunks=16)
    cpu = oracle_run_parallel(oracle, sub, h00c[0], rh, blocks, DT, OUTER, chunks=16)
    g = final["fp64"]
    np.testing.assert_array_equal(g["r"][pick], cr["r"])
    np.testing.assert_array_equal(g["v"][pick], cr["v"])
    np.testing.assert_array_equal(g["flags"][pick], cr["flags"])
    er = rel_err(g["r"][pick], cpu["r"])
    print(f"fp64 vs the CPU path (glibc sin/cos misround in {cr['libm_differ']} of {cr['libm_calls']} calls): "
          f"identical {(er == 0).sum()}/256, r-err median {np.median(er):.2e} max {er.max():.2e}")
    for form in ("fp64", "fast", "mixed"):
        got = final[form]
        np.testing.assert_array_equal(got["flags"][pick], cpu["flags"])
        da, de, di = _elements_diff(s.m[0], cpu["r"], cpu["v"], got["r"][pick], got["v"][pick])
        print(f"{form} vs the CPU path (256): |da/a| max {da.max():.2e} median {np.median(da):.2e}, |de| max "
              f"{de.max():.2e}, |di| max {di.max():.2e}")
        assert da.max() < 1e-9 and de.max() < 1e-9 and di.max() < 1e-9


# ------------------------------------------------------ 64-system batch --

def test_batch_64_systems(oracle):
    from concurrent.futures import ThreadPoolExecutor
    from glisser_amd import native, systems
    members = systems.batch_systems(64, seed=3, n_steps=2 * TB)
    total = sum(y["n_particles"] for y in members)
    print(f"64 systems, {total} particles", flush=True)
    assert total > 5_000_000
    out = {}
    for form in ("fp64", "fast", "mixed"):
        with native.Batch(members, steps_max=TB) as b:
            for k, y in enumerate(members):
                b.upload(k, y["system"].pr, y["system"].pv)
            b.init_accel()
            b.set_config(TB, 256, 0, 0)
            b.run(1, form)
            out[form] = [b.download(k) for k in range(len(members))]
            steps = b.sum_steps()
        assert steps == sum(int(d["dti"].astype(np.int64).sum()) for d in out[form])
        print(f"{form} done", flush=True)

    def first_tile(k):
        y = members[k]
        s, h = y["system"], y["history"]
        m = min(256, s.n)
        sub = systems.System(s.m, s.r, s.v, s.pr[:m], s.pv[:m], np.arange(m, dtype=np.uint32), s.seed)
        rl, h0 = oracle.hermite_log(y["m"], h.t, h.r, h.v, y["t0"], y["dt"], TB)
        _, h00 = oracle.hermite_log(y["m"], h.t, h.r, h.v, y["t0"] - y["dt"], y["dt"], 1)
        blocks = [(rl, h0, y["rh"])]
        cpu = oracle_run(oracle, sub, h00[0], y["rh"], blocks, y["dt"], y["outer_radius"], kill=False)
        cr = oracle_run(oracle, sub, h00[0], y["rh"], blocks, y["dt"], y["outer_radius"], kill=False,
                        correctly_rounded=True)
        return sub, h00[0], blocks, cpu, cr

    with ThreadPoolExecutor(16) as pool:
        cpu_runs = list(pool.map(first_tile, range(len(members))))
    worst = {"fast": 0.0, "mixed": 0.0}
    flagged = 0
    for k, (sub, h00, blocks, cpu, cr) in enumerate(cpu_runs):
        m = sub.n
        y = members[k]
        oc, ocr = np.argsort(cpu["id"]), np.argsort(cr["id"])
        flagged += int(np.count_nonzero(cpu["flags"]))
        for form in ("fp64", "fast", "mixed"):
            d = out[form][k]
            g = {"r": d["r"][:m], "v": d["v"][:m], "flags": d["flags"][:m], "id": np.arange(m, dtype=np.uint32)}
            if form == "fp64":
                np.testing.assert_array_equal(g["r"], cr["r"][ocr], err_msg=f"system {k}")
                np.testing.assert_array_equal(g["v"], cr["v"][ocr], err_msg=f"system {k}")
                np.testing.assert_array_equal(g["flags"], cr["flags"][ocr], err_msg=f"system {k}")
                continue
            ok = (g["flags"] == 0) & (cpu["flags"][oc] == 0)
            if ok.any():
                worst[form] = max(worst[form], float(rel_err(g["r"][ok], cpu["r"][oc][ok]).max()))
            if not np.array_equal(g["flags"], cpu["flags"][oc]):
                g["death_by_id"] = np.where(g["flags"] != 0, d["dti"][:m].astype(np.int64), -1)
                bad, dist, unexplained = unexplained_flag_mismatches(oracle, sub, h00, y["rh"], blocks, y["dt"],
                                                                     y["outer_radius"], cpu, g)
                print(f"system {k} {form}: {len(bad)} flag differences at threshold distances {dist}")
                assert unexplained == [], (k, form, unexplained)
    print(f"first tiles: {flagged} flagged by the CPU path; max rel r err fast {worst['fast']:.2e}, "
          f"mixed {worst['mixed']:.2e}")
    assert worst["fast"] < 1e-10 and worst["mixed"] < 1e-10


# ------------------------------------------- 1M three-body hand-off --

def _jacobi(m, r, v, rp, vp, a_p):
    mu_s, mu_p = m
    M = mu_s + mu_p
    n = np.sqrt(M / a_p**3)
    x = r - mu_p * rp / M
    u = v - mu_p * vp / M
    return (2 * (mu_s / np.linalg.norm(r, axis=1) + mu_p / np.linalg.norm(r - rp, axis=1))
            + 2 * n * (x[:, 0] * u[:, 1] - x[:, 1] * u[:, 0]) - np.sum(u * u, axis=1))


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_three_body_handoff_1m(oracle, form):
    from glisser_amd import native, systems
    n, blocks_n, a_p = 1_000_000, 3, 30.0
    s = systems.three_body(n, seed=4, a_range=(25.0, 40.0), e_hi=0.3, i_hi_deg=10.0)
    queued = []
    with native.Executor(s.m, s.r, s.v, s.pr, s.pv, s.ids, dt=DT, tbsize=TB, rh_factor=3.5,
                         resolve_encounters=True, form=form) as ex:
        queued.append(ex.take_encounters()["id"])
        for _ in range(blocks_n):
            ex.step(1)
            queued.append(ex.take_encounters()["id"])
        gpu = ex.download()
        rp = ex.planet_positions()[0]
    og = np.argsort(gpu["id"])
    r, v, f = gpu["r"][og], gpu["v"][og], gpu["flags"][og]
    dtime = gpu["deathtime"][og]
    q = np.concatenate(queued)
    enc = (f != 0) & ((f & 0xFE) == 0)
    np.testing.assert_array_equal(np.sort(q), np.flatnonzero(enc))
    print(f"{form}: {len(q)} particles handed to the host queue, {int(np.count_nonzero(f & 0x80))} culled, "
          f"{int((f == 0).sum())} on the GPU", flush=True)
    # Jacobi constant of the particles never flagged
    vp = np.cross([0.0, 0.0, 1.0], rp) * np.sqrt((s.m[0] + s.m[1]) / a_p**3)
    c0 = _jacobi(s.m, s.pr, s.pv, s.r[1], s.v[1], a_p)
    c1 = _jacobi(s.m, r, v, rp, vp, a_p)
    drift = (np.abs(c1 - c0) / np.abs(c0))[f == 0]
    print(f"{form}: C_J drift max {drift.max():.2e}, median {np.median(drift):.2e}")
    assert drift.max() < 1e-6 and np.median(drift) < 1e-8
    # a sample through the CPU path (resolve mode: flagged particles stop)
    pick = np.sort(np.random.default_rng(11).choice(n, 2000, replace=False))
    sub = systems.System(s.m, s.r, s.v, s.pr[pick], s.pv[pick], np.arange(len(pick), dtype=np.uint32), s.seed)
    h0_0, rh0, blocks = planet_blocks(oracle, s, blocks_n, TB, DT, rh_factor=3.5)
    cpu = oracle_run(oracle, sub, h0_0, rh0, blocks, DT, OUTER, kill=False)
    oc = np.argsort(cpu["id"])
    g = {"r": r[pick], "v": v[pick], "flags": f[pick], "id": np.arange(len(pick), dtype=np.uint32),
         "death_by_id": np.where(np.isnan(dtime[pick]), -1, np.rint(np.nan_to_num(dtime[pick]) / DT)).astype(np.int64)}
    assert np.count_nonzero(cpu["flags"]) > 100
    if form == "fp64":
        cr = oracle_run_parallel(oracle, sub, h0_0, rh0, blocks, DT, OUTER, correctly_rounded=True, chunks=16,
                                 kill=False)
        np.testing.assert_array_equal(g["r"], cr["r"])
        np.testing.assert_array_equal(g["v"], cr["v"])
        np.testing.assert_array_equal(g["flags"], cr["flags"])
        np.testing.assert_array_equal(g["flags"], cpu["flags"][oc])
    else:
        ok = (g["flags"] == 0) & (cpu["flags"][oc] == 0)
        er = rel_err(g["r"][ok], cpu["r"][oc][ok])
        print(f"mixed vs the CPU path (2000): r-err max {er.max():.2e} median {np.median(er):.2e}")
        assert er.max() < 1e-10
        bad, dist, unexplained = unexplained_flag_mismatches(oracle, sub, h0_0, rh0, blocks, DT, OUTER, cpu, g)
        print(f"mixed: {len(bad)} flag differences, threshold distances {dist}")
        assert unexplained == []
