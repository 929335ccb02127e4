"""Throughput of the time-block particle integrator on one MI355X.

Workload: 200,000 test particles around the Sun + 4 giant planets (masses
9.55e-4, 2.86e-4, 4.37e-5, 5.15e-5 Msun at 5.20/9.55/19.2/30.1 au;
particles a ~ U[60,65] au, e ~ U[0,0.1], i ~ U[0,5 deg]), dt = 180 d,
Time-Block-Size 2048.  The planets come from a history sampled every
365 d (host Jacobi WH at 36.5 d, every tenth state kept), uploaded once;
each block's planet log is interpolated on the device (cubic Hermite + h0).
One bench "step" is one time-block.  The timed region is the interpolation
of `--steps` planet logs and one launch of `--steps` consecutive blocks,
then the encounter-kill transform, the step count and the stable partition.
`--warmup` blocks of the same integration run first, untimed, so the timed
blocks see the GPU at its sustained clock.

Every arithmetic form is measured on the same inputs: fp64 (the
reference's arithmetic, bit-for-bit with the CPU path in every basic
operation), fast (FP64 throughout, restructured) and mixed (fast with each
planet's r^-3 started from the FP32 reciprocal square root and corrected
to first order in FP64).  `value` is the mixed form's rate (the form the
metric names; fast is within 1-2 % of it); the others are reported beside
it, with their orbital-element differences from the fp64 form at the end
of the run.

Beside the headline: a batched launch of 64 systems (9.45M particles), a
1M-particle three-body run that hands encounter records to the host, the
reference's own CPU-runnable case end to end (1,000 particles, planets
integrated on the host, 50 blocks), and the track output of the 200k swarm (orbital elements formed on the device
as the reference's 28-byte records, then written to a track file).  The
CPU leg (skipped by --no-cpu) times the reference's CPU path on the host,
and runs the oracle over a particle sample as the checker.

Prints one JSON line.  --gpus N with N > 1 exits: this build uses one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

N_PARTICLES = 200_000
TBSIZE = 2048
DT = 180.0
OUTER = 300.0
CULL = 0.5
FP64_PEAK_TFLOPS = 78.6
HEADLINE_FORM = "mixed"

# FP64 FLOPs per particle-step counted by rocprofv3 on this build
# (64 x (ADD_F64 + MUL_F64 + TRANS_F64 + 2 FMA_F64) per wave instruction,
# summed over the block kernel and divided by its particle-steps); see
# profiles/wh_queue_200k_counters.md.  Used with the live rate to report
# the counter-based share of FP64 vector peak.
COUNTER_FLOPS_PER_PSTEP = {"fp64": 880.36, "fast": 210.39, "mixed": 206.39}

# VALU-busy SIMD cycles per particle-step (rocprofv3 SQ_ACTIVE_INST_VALU, in
# quad-cycles, x 4 / particle-steps, the benchmark's queued launch) and the
# shader clock during the benchmark's launches (the highest seen on any box,
# by GRBM_GUI_ACTIVE in the counter runs or by s_memtime / s_memrealtime
# read at each workgroup's start and end in an instrumented build, so the
# figures derived from it err low); see profiles/wh_queue_200k_counters.md.
# Used with the live rate to report the share of the SIMDs' issue cycles the
# kernel keeps busy and the roofline fraction at the measured clock.
VALU_CYCLES_PER_PSTEP = {"fp64": 51.87, "fast": 10.588, "mixed": 10.588}
CLOCK_GHZ_MEASURED = {"fp64": 2.370, "fast": 2.280, "mixed": 2.289}
N_SIMDS = 1024  # 256 CUs x 4 SIMDs
PEAK_CLOCK_GHZ = 2.4

# HBM bytes per time-block of the mixed form at this workload, by launch
# shape (queued: every 256-step task reads and writes its tile's state, 8
# times per block; sliced: 4 slices x 4 launches per block, 4 times), from
# rocprofv3 FETCH_SIZE (x 2.0, the factor a copy kernel of known size needs)
# and WRITE_SIZE (x 1.0); see profiles/wh_queue_200k_counters.md.
TRAFFIC_BYTES_PER_BLOCK = {("mixed", "queued"): 247.2e6, ("mixed", "sliced"): 133.4e6}

TRACK_RECORD_BYTES = 28  # u32 id + 6 float32 per particle in a track snapshot


_T_START = time.perf_counter()


def note(msg: str) -> None:
    """One progress line on stderr (the JSON line goes to stdout)."""
    print(f"bench [{time.perf_counter() - _T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)


def copy_probe(n: int = 1_000_000, repeats: int = 3) -> None:
    """Calibration kernel for the HBM byte counters: the rollback snapshot
    copies r, v, flags, deathtime index and id (58 B per particle) of n
    particles, `repeats` times."""
    from glisser_amd import native

    with native.Context(n, 2, 8, keep_rollback=True) as ctx:
        z = np.zeros((n, 3))
        ctx.upload(z, z, np.zeros(n, np.uint16), np.arange(n, dtype=np.uint32), a=z, dti=np.zeros(n, np.uint32))
        ctx.set_counts(n, n)
        for _ in range(repeats):
            ctx.snapshot_rollback()
        ctx.sync()
    print(json.dumps({"copy_probe_bytes_read": 58 * n, "copy_probe_bytes_written": 58 * n, "repeats": repeats}))


def waves_per_simd(args, form: str) -> int:
    """Resident-wave cap of a form's launch: --waves-per-simd if given, else
    4 for fp64 and none for the fast forms (see the option's note)."""
    if args.waves_per_simd is not None:
        return args.waves_per_simd
    return 4 if form == "fp64" else 0


def flops_per_particle_step(n_planets: int) -> int:
    """Algorithmic FP64 operations of one particle-step (reference
    src/wh.cu:74-177): 135 + 16 n_p basic ops, 16 + n_p divides, 2 + n_p
    square roots, 14 sin/cos (5 Newton iterations) -- each special op
    counted once."""
    return (135 + 16 * n_planets) + (16 + n_planets) + (2 + n_planets) + 14


def hbm_bytes_per_block(n_particles: int, n_bodies: int, tbsize: int) -> int:
    """Minimum HBM bytes of one time-block: the particle state read and
    written once (r, v, a: 72 B; flags 2 B; deathtime index 4 B) and the
    planet block read once (n_bodies x 24 B per step; the workgroups'
    further reads of it come from cache)."""
    return n_particles * (72 + 2 + 4) * 2 + tbsize * n_bodies * 24


# Source files of the time-block kernel; their SHA-256 when the counters
# above were taken.  The counter-based fields are reported only while the
# kernel is the profiled one.
KERNEL_SOURCES = ["glisser_amd/csrc/wh_kernels.hip", "glisser_amd/csrc/wh_fast.h", "glisser_amd/csrc/wh_device.h",
                  "glisser_amd/csrc/sincos_cr.h", "glisser_amd/csrc/launch.h"]
PROFILED_KERNEL_SHA256 = "3f003a1ac264c57c115d5513e46aa8ca5f35ab7f569b8eacb3a7118f07685d5f"


def kernel_fingerprint() -> str:
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        h.update((ROOT / rel).read_bytes())
    return h.hexdigest()


HIST_SPACING = 365.0
HIST_SUBSTEPS = 10


def make_inputs(n_particles: int, n_blocks: int, seed: int):
    """Swarm + planet history covering n_blocks blocks (plus one spacing)."""
    from glisser_amd import systems

    s = systems.outer_solar_system(n_particles, seed=seed)
    hist = systems.planet_history(s, n_blocks * TBSIZE * DT + HIST_SPACING, spacing=HIST_SPACING,
                                  substeps=HIST_SUBSTEPS, cull_radius=CULL)
    rh = np.full(s.n_bodies, CULL)  # Encounter-RH-Factor 0: every body's radius is the cull radius
    return s, hist, rh


def prepare_gpu(inputs, steps: int, warmup: int):
    """A context holding the swarm, the planet history, the initial
    accelerations and flags, compacted: everything before the first block.
    bench.py prepares every form's context before timing any of them, so the
    forms then run back to back on a GPU at its sustained clock."""
    from glisser_amd import native

    s, hist, rh = inputs
    n = s.n
    ctx = native.Context(n, s.n_bodies, TBSIZE, n_planet_buffers=warmup + steps)
    try:
        ctx.upload(s.pr, s.pv, np.zeros(n, np.uint16), s.ids, a=np.zeros((n, 3)), dti=np.zeros(n, np.uint32))
        ctx.set_counts(n, n)
        ctx.upload_history(hist.m, hist.t, hist.r, hist.v)
        pr0, h00 = ctx.interp_at(np.array([0.0]))
        ctx.init_accel(s.m, rh, pr0[0], h00[0], OUTER)
        ctx.kill_encounters()
        ctx.partition()
        ctx.sync()
    except BaseException:
        ctx.close()
        raise
    return ctx


def run_gpu(inputs, form: str, steps: int, warmup: int, shards: int, waves_per_simd: int, parts: int = 1,
            ctx=None, launch: str = "sliced", segment_steps: int = 0):
    """`warmup` untimed blocks, then the timed region: the planet rows of
    `steps` blocks interpolated on the device, one launch of the `steps`
    blocks, the encounter-kill transform, the step count and the stable
    partition.  Closes the context."""
    s, hist, rh = inputs
    n = s.n
    if ctx is None:
        ctx = prepare_gpu(inputs, steps, warmup)
    try:
        if warmup:
            ctx.interp_blocks(0, [k * TBSIZE * DT for k in range(warmup)], DT, TBSIZE, rh)
            if launch == "queued":
                ctx.run_blocks_queued(warmup, 0, DT, OUTER, form, segment_steps)
            else:
                ctx.run_blocks(warmup, 0, DT, OUTER, form, shards, waves_per_simd, parts)
            ctx.kill_encounters()
            ctx.partition()
        ctx.sync()
        n_alive, _ = ctx.counts()
        t0 = time.perf_counter()
        ctx.interp_blocks(warmup, [k * TBSIZE * DT for k in range(warmup, warmup + steps)], DT, TBSIZE, rh)
        if launch == "queued":
            ctx.run_blocks_queued(steps, warmup, DT, OUTER, form, segment_steps)
        else:
            ctx.run_blocks(steps, warmup, DT, OUTER, form, shards, waves_per_simd, parts)
        ctx.kill_encounters()
        psteps = ctx.sum_steps(n_alive)
        ctx.partition()
        ctx.sync()
        elapsed = time.perf_counter() - t0
        kernel_ms, part_ms = ctx.last_times()
        n_end, _ = ctx.counts()
        final = ctx.download(n)
    finally:
        ctx.close()
    return {"elapsed_s": elapsed, "particle_steps": psteps, "kernel_ms": kernel_ms, "partition_ms": part_ms,
            "n_alive_start": n_alive, "n_alive_end": n_end, "final": final}


def element_differences(mu: float, ref: dict, other: dict) -> dict:
    """Orbital-element and energy differences of `other` from `ref` over the
    particles unflagged in both (matched by id)."""
    from glisser_amd import native

    def by_id(d):
        order = np.argsort(d["id"])
        return d["r"][order], d["v"][order], d["flags"][order]

    r0, v0, f0 = by_id(ref)
    r1, v1, f1 = by_id(other)
    keep = (f0 == 0) & (f1 == 0)
    if not keep.any():
        return {"n": 0}
    mus = np.full(int(keep.sum()), mu)
    e0 = native.to_elements(mus, r0[keep], v0[keep])
    e1 = native.to_elements(mus, r1[keep], v1[keep])

    def energy(r, v):
        return 0.5 * np.sum(v * v, axis=1) - mu / np.linalg.norm(r, axis=1)

    da = np.abs(e1[:, 0] - e0[:, 0]) / e0[:, 0]
    de = np.abs(e1[:, 1] - e0[:, 1])
    di = np.abs(e1[:, 2] - e0[:, 2])
    den = np.abs(energy(r1[keep], v1[keep]) - energy(r0[keep], v0[keep])) / np.abs(energy(r0[keep], v0[keep]))
    dr = np.linalg.norm(r1[keep] - r0[keep], axis=1) / np.linalg.norm(r0[keep], axis=1)
    dv = np.linalg.norm(v1[keep] - v0[keep], axis=1) / np.linalg.norm(v0[keep], axis=1)
    stat = lambda x: {"max": float(x.max()), "median": float(np.median(x))}  # noqa: E731
    return {"n": int(keep.sum()), "dr_rel": stat(dr), "dv_rel": stat(dv), "da_over_a": stat(da), "de": stat(de),
            "di_rad": stat(di),
            "denergy_rel": stat(den), "flags_differ": int(np.sum(f0 != f1))}


def oracle_sample_run(inputs, n_sample: int, n_blocks: int) -> dict:
    """The CPU oracle (restatement of the reference's src/wh.cpp) over the
    first n_sample particles for n_blocks blocks on the same planet logs;
    flagged particles stop, as on the GPU."""
    from oracle.pyoracle import Oracle

    s, hist, rh = inputs
    o = Oracle()
    n = n_sample
    pr0, h00 = o.hermite_log(hist.m, hist.t, hist.r, hist.v, -DT, DT, 1)
    r, v = s.pr[:n].copy(), s.pv[:n].copy()
    a, flags = o.particle_accels(s.m, rh, OUTER, pr0[0], h00[0], r, np.zeros(n), s.ids[:n])
    dti = None
    for k in range(n_blocks):
        rl, h0 = o.hermite_log(hist.m, hist.t, hist.r, hist.v, k * TBSIZE * DT, DT, TBSIZE)
        out = o.particle_block(s.m, rh, OUTER, rl, h0, k * TBSIZE * DT, DT, r, v, a, flags, s.ids[:n], dti)
        if out["rc"] != 0:
            raise RuntimeError(f"oracle failed: {o.err()}")
        r, v, a, flags, dti = out["r"], out["v"], out["a"], out["flags"], out["dti"]
    return {"id": s.ids[:n].copy(), "r": r, "v": v, "flags": flags}


def subset_by_id(d: dict, ids: np.ndarray) -> dict:
    order = np.argsort(d["id"])
    pos = order[np.searchsorted(d["id"][order], ids)]
    return {k: d[k][pos] for k in ("id", "r", "v", "flags")}


BATCH_SYSTEMS = 64
BATCH_LAUNCHES = 4


def run_batch(forms, n_systems: int, launches: int, seed: int) -> dict:
    """Batched workload: n_systems independent systems (2-8 planets each,
    particle counts log-uniform 1e3-1e6, dt = P_inner / 20), all in one launch
    per time-block, captured in a HIP graph.  Per form: the autotuner picks
    steps per launch, lanes per workgroup and dispatch / persistent mode
    (outside the timed region, state restored), then one warm-up launch and
    `launches` timed launches."""
    from glisser_amd import native, systems

    members = systems.batch_systems(n_systems, seed=seed, n_steps=(launches + 1) * 2048)
    n_total = sum(y["n_particles"] for y in members)
    out = {"systems": n_systems, "particles": n_total,
           "bodies_histogram": np.bincount([len(y["m"]) for y in members], minlength=10)[2:].tolist(),
           "launches": launches, "forms": {}}
    for form in forms:
        with native.Batch(members, steps_max=2048) as b:
            for k, y in enumerate(members):
                s = y["system"]
                b.upload(k, s.pr, s.pv)
            b.init_accel()
            cfg, tune_rate = b.tune(form)
            b.run(1, form)
            s0 = b.sum_steps()
            ms = b.run(launches, form)
            s1 = b.sum_steps()
        out["forms"][form] = {"value": (s1 - s0) / (ms * 1e-3), "ms_per_launch": ms / launches,
                              "steps_per_launch": cfg["steps_per_launch"], "config": cfg,
                              "tune_rate": tune_rate, "particle_steps": s1 - s0}
    if "fp64" in out["forms"] and "mixed" in out["forms"]:
        out["mixed_vs_fp64"] = out["forms"]["mixed"]["value"] / out["forms"]["fp64"]["value"]
    return out


ENC_PARTICLES = 1_000_000
ENC_BLOCKS = 3


def run_encounter(forms, n_particles: int, n_blocks: int, seed: int) -> dict:
    """Encounter hand-off on the three-body setup: Sun + one planet of 1e-3
    Msun on a circular orbit at 30 au, particles a ~ U[25, 40] au, e < 0.3,
    i < 10 deg; encounter radius 3.5 Hill radii; Resolve-Encounters on, so
    every block is rollback snapshot -> kernel -> stable three-way compaction
    -> encounter records (block-start and flagging states) to the host queue,
    with the next planet block produced and the previous block's records
    queued on the host meanwhile (glx::Executor).  Timed: one call running
    n_blocks blocks after one warm-up block; particle-steps counted from each
    particle's death time.  Also reports the
    Jacobi-constant drift of the particles never flagged."""
    from glisser_amd import native, systems

    s = systems.three_body(n_particles, seed=seed, a_range=(25.0, 40.0), e_hi=0.3, i_hi_deg=10.0)
    a_p = 30.0
    out = {"particles": n_particles, "blocks": n_blocks, "tbsize": TBSIZE, "dt_days": DT,
           "encounter_radius_hill": 3.5, "forms": {}}

    def jacobi(r, v, rp, vp):
        mu_s, mu_p = s.m
        M = mu_s + mu_p
        n = np.sqrt(M / a_p**3)
        x = r - mu_p * rp / M
        u = v - mu_p * vp / M
        return (2 * (mu_s / np.linalg.norm(r, axis=1) + mu_p / np.linalg.norm(r - rp, axis=1))
                + 2 * n * (x[:, 0] * u[:, 1] - x[:, 1] * u[:, 0]) - np.sum(u * u, axis=1))

    c0 = jacobi(s.pr, s.pv, s.r[1], s.v[1])
    for form in forms:
        with native.Executor(s.m, s.r, s.v, s.pr, s.pv, s.ids, dt=DT, tbsize=TBSIZE, rh_factor=3.5,
                             resolve_encounters=True, form=form) as ex:
            ex.step(1)
            queued0 = len(ex.take_encounters()["id"])
            st = ex.status()
            n0, t_s = st["n_alive"], st["t"]
            # one call for all timed blocks: each block's records are taken
            # on the host while the next block's kernel runs
            k0, _ = ex.kernel_total()
            t0 = time.perf_counter()
            ex.step(n_blocks)  # the records are in the host queue when this returns
            el = time.perf_counter() - t0
            k1, nk = ex.kernel_total()
            rec = ex.take_encounters()
            handed = len(rec["id"])
            st = ex.status()
            t_e = st["t"]
            # particle-steps, exactly: every particle alive at t_s ran until its
            # death time (the end of its flagging step) or to t_e
            dd = ex.download()
            death = np.where(np.isnan(dd["deathtime"]), t_e, np.minimum(dd["deathtime"], t_e))
            live = np.isnan(dd["deathtime"]) | (dd["deathtime"] > t_s)
            psteps = int(np.rint((death[live] - t_s) / DT).sum())
            other_flagged = int(n0 - st["n_alive"] - handed)
            kernel_ms, part_ms, planet_ms = ex.times()
            d = dd
            rp = ex.planet_positions()[0]
        vp = np.cross([0.0, 0.0, 1.0], rp) * np.sqrt((s.m[0] + s.m[1]) / a_p**3)
        order = np.argsort(d["id"])
        ok = d["flags"][order] == 0
        c1 = jacobi(d["r"][order], d["v"][order], rp, vp)
        drift = np.abs(c1 - c0)[ok] / np.abs(c0[ok])
        out["forms"][form] = {
            "value": psteps / el, "unit": "particle-steps/s", "ms_per_block": el / n_blocks * 1e3,
            "kernel_ms_per_block": (k1 - k0) / n_blocks, "host_overhead": el * 1e3 / (k1 - k0) - 1.0,
            "kernel_ms_last_block": kernel_ms, "partition_ms_last_block": part_ms, "other_flagged": other_flagged,
            "planet_block_host_ms": planet_ms, "alive_start": int(n0), "alive_end": int(d["n_alive"]),
            "handed_to_host": handed, "handed_warmup_block": queued0,
            "jacobi_drift_max": float(drift.max()), "jacobi_drift_median": float(np.median(drift))}
    return out


CONFIG0_PARTICLES = 1000
CONFIG0_BLOCKS = 50


def run_config0(forms, seed: int = 1) -> dict:
    """The reference's own CPU-runnable case end to end: Sun + 4 giants
    integrated on the host (the Jacobi WH planet producer, overlapped with
    the kernel), 1,000 particles a ~ U[60, 65] au, dt = 180 d, 50 blocks of
    2048 steps (102,400 steps), cull radius 0.5 au, kill mode, through
    glx::Executor: one call of 50 blocks per form, host clock; particle-steps
    from the death times.  Returns the fp64 form's final state under
    "final"."""
    from glisser_amd import native, systems

    s = systems.outer_solar_system(CONFIG0_PARTICLES, seed=seed)
    t_e = CONFIG0_BLOCKS * TBSIZE * DT
    out = {"particles": s.n, "blocks": CONFIG0_BLOCKS, "tbsize": TBSIZE, "dt_days": DT, "forms": {}}
    for form in forms:
        with native.Executor(s.m, s.r, s.v, s.pr, s.pv, s.ids, dt=DT, tbsize=TBSIZE, outer_radius=OUTER,
                             form=form) as ex:
            t0 = time.perf_counter()
            ex.step(CONFIG0_BLOCKS)
            el = time.perf_counter() - t0
            d = ex.download()
        death = np.where(np.isnan(d["deathtime"]), t_e, np.minimum(d["deathtime"], t_e))
        psteps = int(np.rint(death / DT).sum())
        out["forms"][form] = {"value": psteps / el, "unit": "particle-steps/s", "ms": el * 1e3,
                              "particle_steps": psteps, "alive_end": int(d["n_alive"])}
        if form == "fp64":
            out["final"] = d
    return out


def cpu_config0(gpu_final=None, seed: int = 1) -> dict:
    """The reference's CPU path on the same case end to end, one thread, host
    clock (oracle/_ref: its planet producer, particle time-blocks and stable
    partition, encounters killed, as targets/main.cpp drives them); as the
    checker, the GPU's fp64 final state against it (survivor count,
    relative position error by id)."""
    from glisser_amd import systems
    from oracle.pyoracle import Oracle, reference_available

    s = systems.outer_solar_system(CONFIG0_PARTICLES, seed=seed)
    kind = "reference" if reference_available() else "port"
    o = Oracle(reference=(kind == "reference"))

    def kill(f):
        f = f.copy()
        enc = (f & 0x01) == 0x01
        f[enc] = (f[enc] & 0xFF00) | 0x80
        return f

    t0 = time.perf_counter()
    pl = o.planets(s.m, s.r, s.v, cull_radius=CULL, rh_factor=0.0, outer=OUTER)
    n = s.n
    r, v, ids = s.pr.copy(), s.pv.copy(), s.ids.copy()
    a, flags = o.particle_accels(s.m, pl.state()[2], OUTER, s.r[1:], pl.h0_0, r, np.zeros(n, np.uint16), ids)
    flags = kill(flags)
    n_alive, perm = o.stable_partition(flags, 0, n, 0)
    r, v, a, flags, ids = r[perm], v[perm], a[perm], flags[perm], ids[perm]
    dti = np.zeros(n, np.uint32)  # indexed by id, as in the reference
    t, psteps = 0.0, 0
    for _ in range(CONFIG0_BLOCKS):
        pl.recalculate_rh()
        rh = pl.state()[2]
        rl, _vl, h0 = pl.block(TBSIZE, DT)
        res = o.particle_block(s.m, rh, OUTER, rl, h0, t, DT, r[:n_alive], v[:n_alive], a[:n_alive],
                               flags[:n_alive], ids[:n_alive], dti)
        if res["rc"] != 0:
            raise RuntimeError(f"reference CPU path failed: {o.err()}")
        dti = res["dti"]
        psteps += int(dti[ids[:n_alive]].astype(np.int64).sum())
        sl = slice(0, n_alive)
        r[sl], v[sl], a[sl], flags[sl] = res["r"], res["v"], res["a"], kill(res["flags"])
        split, perm = o.stable_partition(flags, 0, n_alive, 0)
        r[sl], v[sl], a[sl], flags[sl], ids[sl] = r[sl][perm], v[sl][perm], a[sl][perm], flags[sl][perm], ids[sl][perm]
        n_alive = split
        t += DT * TBSIZE
    el = time.perf_counter() - t0
    out = {"value": psteps / el, "unit": "particle-steps/s", "cores": 1, "kind": kind,
           "sample": f"{n} particles + Sun and 4 giants integrated on the host, {CONFIG0_BLOCKS} blocks of "
                     f"{TBSIZE} steps ({psteps} particle-steps, {el:.1f} s, single thread)",
           "particle_steps": psteps, "alive_end": int(n_alive)}
    if gpu_final is not None:
        og, oc = np.argsort(gpu_final["id"]), np.argsort(ids)
        same = bool(np.array_equal(gpu_final["id"][og], ids[oc]) and gpu_final["n_alive"] == n_alive)
        out["fp64_vs_cpu"] = {"same_survivors": same}
        if same:
            err = (np.linalg.norm(gpu_final["r"][og] - r[oc], axis=1) / np.linalg.norm(r[oc], axis=1))
            out["fp64_vs_cpu"].update({"r_rel_max": float(err.max()), "r_rel_median": float(np.median(err))})
    return out


def run_track(inputs, repeats: int = 5, seed: int = 5) -> dict:
    """Track output of the whole swarm, the reference's particle-history pass
    (targets/main.cpp:148-156 sort_by_id, src/data.cpp:770-812
    save_binary_track): each particle's orbital elements about the planets'
    barycentre, formed on the device as 28-byte records in id order
    (glx_track_pack: elements, id ranking, record pack, copy to host memory),
    then one snapshot appended to a track file (glx_track_write_records).
    The particles sit in the context in a shuffled slot order, as after
    compactions.  Both calls are synchronous: host clock, median of `repeats`
    after one untimed call.  Returns the records under "records"."""
    import tempfile
    from glisser_amd import native

    s = inputs[0]
    n = s.n
    order = np.random.default_rng(seed).permutation(n)
    planets = {"pm": s.m, "pr": s.r, "pv": s.v, "pid": np.arange(s.n_bodies, dtype=np.uint32)}
    with native.Context(n, s.n_bodies, 16) as ctx, tempfile.TemporaryDirectory() as tmp:
        ctx.upload(s.pr[order], s.pv[order], np.zeros(n, np.uint16), s.ids[order], a=np.zeros((n, 3)),
                   dti=np.zeros(n, np.uint32))
        ctx.set_counts(n, n)
        path = Path(tmp) / "tracks.bin"
        rec = ctx.track_records(s.m, s.r, s.v)
        native.write_track_records(path, 0.0, planets, rec)
        pack, write = [], []
        for _ in range(repeats):
            t0 = time.perf_counter()
            rec = ctx.track_records(s.m, s.r, s.v)
            t1 = time.perf_counter()
            native.write_track_records(path, 0.0, planets, rec)
            write.append(time.perf_counter() - t1)
            pack.append(t1 - t0)
        size = path.stat().st_size
    pack_ms, write_ms = 1e3 * float(np.median(pack)), 1e3 * float(np.median(write))
    return {"particles": n, "record_bytes": TRACK_RECORD_BYTES, "file_bytes": size,
            "pack_ms": pack_ms, "write_ms": write_ms, "value": n / ((pack_ms + write_ms) / 1e3),
            "unit": "particles/s", "records": rec}


def cpu_track(inputs, records) -> dict:
    """The reference's own save_binary_track (oracle/_ref) over the same
    particles, already in id order (its sort_by_id is not timed), to a track
    file: host clock, median of 3; and, as the checker, the device records
    against the records of that file (same ids; the share of element floats
    equal bit for bit, and the largest float32 ulp distance)."""
    import tempfile
    from oracle.pyoracle import Oracle

    s = inputs[0]
    order = np.argsort(s.ids)
    st = {"pm": s.m, "pr": s.r, "pv": s.v, "pid": np.arange(s.n_bodies, dtype=np.uint32), "r": s.pr[order],
          "v": s.pv[order], "id": s.ids[order]}
    o = Oracle(reference=True)
    times = []
    with tempfile.TemporaryDirectory() as tmp:
        path = Path(tmp) / "tracks.bin"
        for _ in range(3):
            t = time.perf_counter()
            o.track_write(path, 0.0, st)
            times.append(time.perf_counter() - t)
        data = path.read_bytes()
    head = 8 + 8 + (s.n_bodies - 1) * TRACK_RECORD_BYTES + 8
    ref = np.frombuffer(data[head:], dtype=np.dtype([("id", np.uint32), ("x", np.float32, 6)]))
    same_ids = bool(len(ref) == len(records) and np.array_equal(ref["id"], records["id"]))
    out = {"value": len(ref) / float(np.median(times)), "unit": "particles/s", "cores": 1, "kind": "reference",
           "ms": 1e3 * float(np.median(times)), "ids_equal": same_ids}
    if same_ids:
        ulps = np.abs(ref["x"].view(np.int32).astype(np.int64) - records["x"].view(np.int32).astype(np.int64))
        out["floats_equal"] = float(np.mean(ulps == 0))
        out["max_ulps"] = int(ulps.max())
    return out


def cpu_baseline(inputs, sample_particles: int, runs: int = 3, threads: int = 0):
    """Reference CPU path (oracle/_ref: the reference's own src/wh.cpp built
    from its sources) or, where that library is absent, the oracle
    restatement: one 2048-step block over the first `sample_particles`
    particles of the same swarm and planet log, single thread, median of
    `runs` runs.  `all_cores` is a labelled extension, not the reference's
    behaviour (its CPU path is single-threaded): the same block over the whole
    swarm split into `threads` disjoint particle slices, one host thread each
    (the library call releases the GIL; each call owns all of its state)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle.pyoracle import Oracle, reference_available

    s, hist, rh = inputs
    kind = "reference" if reference_available() else "port"
    checker = Oracle()
    rl, h0 = checker.hermite_log(hist.m, hist.t, hist.r, hist.v, 0.0, DT, TBSIZE)
    pr0, h00 = checker.hermite_log(hist.m, hist.t, hist.r, hist.v, -DT, DT, 1)
    o = Oracle(reference=(kind == "reference"))
    n_all = len(s.ids)
    a0, f0 = o.particle_accels(s.m, rh, OUTER, pr0[0], h00[0], s.pr, np.zeros(n_all), s.ids)

    def block(lo, hi):
        out = o.particle_block(s.m, rh, OUTER, rl, h0, 0.0, DT, s.pr[lo:hi], s.pv[lo:hi], a0[lo:hi],
                               f0[lo:hi], s.ids[lo:hi])
        return int(out["dti"].astype(np.int64).sum())

    n = min(sample_particles, n_all)
    times = []
    for _ in range(runs):
        t = time.perf_counter()
        steps = block(0, n)
        times.append(time.perf_counter() - t)
    el = float(np.median(times))
    res = {"value": steps / el, "unit": "particle-steps/s", "cores": 1, "kind": kind,
           "sample": f"{n} particles x 1 block of {TBSIZE} steps of the same swarm and planet log "
                     f"({steps} particle-steps, median of {runs} runs: "
                     + ", ".join(f"{x:.2f}" for x in times) + " s; single thread)",
           "host": host_description()}
    if threads <= 0:
        threads = min(16, len(os.sched_getaffinity(0)))
    if threads > 1:
        cuts = np.linspace(0, n_all, threads + 1).astype(int)
        t = time.perf_counter()
        with ThreadPoolExecutor(threads) as pool:
            steps = sum(pool.map(block, cuts[:-1], cuts[1:]))
        el = time.perf_counter() - t
        res["all_cores"] = {"value": steps / el, "unit": "particle-steps/s", "cores": threads,
                            "sample": f"{n_all} particles x 1 block of {TBSIZE} steps in {threads} disjoint "
                                      f"slices, one host thread each ({steps} particle-steps, {el:.2f} s)",
                            "note": "extension: the reference's CPU path is single-threaded"}
    return res


def host_description() -> dict:
    """The CPU the baseline ran on: model, logical CPUs, C library, and the
    compiler and flags of the CPU build (oracle/Makefile)."""
    import platform
    import subprocess
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cxx = subprocess.run(["g++", "--version"], capture_output=True, text=True, timeout=10).stdout.splitlines()[0]
    except (OSError, IndexError, subprocess.SubprocessError):
        cxx = "unknown"
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "libc": " ".join(platform.libc_ver()),
            "compiler": f"{cxx}; -O3 -ffp-contract=off"}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed time-blocks (one launch)")
    # 20 untimed blocks (~50 ms of mixed-form load): the MI355X reaches its
    # full clock only after tens of ms of sustained load; with 1 warm-up
    # block the same 10 timed blocks take 2.66 ms each instead of 2.37
    ap.add_argument("--warmup", type=int, default=20, help="untimed time-blocks (run before the timed launch)")
    ap.add_argument("--form", default="all", choices=["all", "fp64", "fast", "mixed"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--particles", type=int, default=N_PARTICLES,
                    help="swarm size (the benchmark workload is 200,000)")
    ap.add_argument("--shards", type=int, default=4, help="particle slices, one stream each")
    # Resident-wave cap through the LDS request.  The fast forms use >= 147
    # VGPRs, which already limits them to 3 waves per SIMD; without the cap
    # they stage 273-step chunks in 32 KB (3 % faster than the capped 455-step
    # chunks on the same box).  The fp64 form (114 VGPRs, up to 4 waves) is
    # 2.5 % faster capped at 4 than at 3 and 0.8 % faster than uncapped.
    ap.add_argument("--waves-per-simd", type=int, default=None,
                    help="resident-wave cap (0: none; default: 4 for fp64, none for fast and mixed)")
    ap.add_argument("--launches-per-block", type=int, default=4,
                    help="launches each 2048-step block is split into (same results, shorter chain end)")
    ap.add_argument("--launch", default="queued", choices=["queued", "sliced"],
                    help="queued: one persistent launch taking (tile, segment) tasks from a device queue; "
                         "sliced: --shards streams of --launches-per-block launches per block")
    ap.add_argument("--segment-steps", type=int, default=0,
                    help="steps per task of the queued launch (multiple of 64; 0: 256)")
    ap.add_argument("--cpu-sample", type=int, default=16000,
                    help="particles in the single-thread CPU-baseline sample (~4 s per run, 3 runs)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the all-cores CPU figure (0: the CPUs this process may use, at most 16)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--oracle-sample", type=int, default=256,
                    help="particles also run through the CPU oracle for the precision report (0: none)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-system workload")
    ap.add_argument("--batch-systems", type=int, default=BATCH_SYSTEMS)
    ap.add_argument("--batch-launches", type=int, default=BATCH_LAUNCHES)
    ap.add_argument("--no-encounter", action="store_true", help="skip the encounter hand-off workload")
    ap.add_argument("--encounter-particles", type=int, default=ENC_PARTICLES)
    ap.add_argument("--no-track", action="store_true", help="skip the track-output workload")
    ap.add_argument("--no-config0", action="store_true",
                    help="skip the end-to-end 1,000-particle case with host-integrated planets")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the known-size copy kernel (HBM counter calibration)")
    args = ap.parse_args(argv)
    if args.copy_probe:
        copy_probe()
        return 0

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(particles shard by slices across GPUs; see the design notes).", file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))

    inputs = make_inputs(args.particles, args.warmup + args.steps, args.seed)
    note(f"inputs ready: {args.particles} particles, {args.warmup} + {args.steps} blocks")
    forms = ["fp64", "fast", "mixed"] if args.form == "all" else [args.form]
    head = HEADLINE_FORM if HEADLINE_FORM in forms else forms[0]
    n_planets = inputs[0].n_bodies - 1
    finals = {}
    f_alg = flops_per_particle_step(n_planets)
    per_form = {}
    counters_current = PROFILED_KERNEL_SHA256 == kernel_fingerprint()
    ctxs = {form: prepare_gpu(inputs, args.steps, args.warmup) for form in forms}
    for form in forms:
        r = run_gpu(inputs, form, args.steps, args.warmup, args.shards, waves_per_simd(args, form),
                    args.launches_per_block, ctx=ctxs.pop(form), launch=args.launch,
                    segment_steps=args.segment_steps)
        finals[form] = r.pop("final")
        block_s = r["kernel_ms"] / 1e3 / args.steps
        per_block = r["particle_steps"] / args.steps
        entry = {
            "value": r["particle_steps"] / r["elapsed_s"],
            "kernel_ms_per_block": block_s * 1e3,
            "particle_steps_per_s_kernel": per_block / block_s,
            "alg_tflops": per_block * f_alg / block_s / 1e12,
            "partition_ms": r["partition_ms"],
            "elapsed_s": r["elapsed_s"],
            "particle_steps": r["particle_steps"],
        }
        note(f"{form}: {entry['value']:.4g} particle-steps/s, {entry['kernel_ms_per_block']:.3f} ms per block "
             "in the kernel")
        cf = COUNTER_FLOPS_PER_PSTEP.get(form)
        if cf and counters_current:
            entry["counter_fp64_frac"] = per_block * cf / block_s / 1e12 / FP64_PEAK_TFLOPS
            clk = CLOCK_GHZ_MEASURED[form]
            entry["valu_busy"] = per_block * VALU_CYCLES_PER_PSTEP[form] / block_s / (N_SIMDS * clk * 1e9)
            entry["frac_at_measured_clock"] = entry["alg_tflops"] / (FP64_PEAK_TFLOPS * clk / PEAK_CLOCK_GHZ)
        per_form[form] = entry
    h = per_form[head]
    line = {
        "metric": "particle-steps/s at 200k particles: FP64 vs mixed; % FP64 vector peak",
        "value": h["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["elapsed_s"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if head == "mixed" else "f64",
        "dtype_note": ("mixed form: FP64 state, Kepler drift, sums and tests; each planet's r^-3 seeded by an FP32 "
                       "reciprocal square root and corrected in FP64" if head == "mixed" else "FP64 throughout"),
        "data": "synthetic (seeded Sun + 4 giants, 200k-particle swarm at 60-65 au)",
        "config": {
            "workload": "wh-timeblock-200k",
            "n_particles": args.particles,
            "n_bodies": inputs[0].n_bodies,
            "tbsize": TBSIZE,
            "dt_days": DT,
            "form": head,
            "blocks_per_launch": args.steps,
            "launch": args.launch,
            **({"segment_steps": args.segment_steps or 256} if args.launch == "queued" else
               {"shards": args.shards, "waves_per_simd": waves_per_simd(args, head),
                "launches_per_block": args.launches_per_block}),
            "planet_log": f"device cubic-Hermite interpolation of a {HIST_SPACING:g}-d history",
            "seed": args.seed,
            "parallelism": "single-gpu",
        },
        "forms": per_form,
        "roofline": {
            "bound": "fp64-valu",
            "bound_note": ("FP64 vector (VALU) issue: no matrix-core work on this path; MI355X's FP64 vector "
                           "peak equals its dense FP64 MFMA peak, 78.6 TFLOP/s"),
            "achieved": h["alg_tflops"],
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": h["alg_tflops"] / FP64_PEAK_TFLOPS,
            "traffic": (TRAFFIC_BYTES_PER_BLOCK[head, args.launch] * args.steps
                        if (head, args.launch) in TRAFFIC_BYTES_PER_BLOCK and args.particles == N_PARTICLES
                        and counters_current and args.segment_steps in (0, 256)
                        else None),
            "traffic_note": "HBM bytes per launch of `steps` blocks (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE)",
            "counter_fp64_frac": h.get("counter_fp64_frac"),
            "clock_ghz_measured": CLOCK_GHZ_MEASURED[head] if counters_current else None,
            "frac_at_measured_clock": h.get("frac_at_measured_clock"),
            "valu_busy": h.get("valu_busy"),
            "counters_note": ("from the counters of profiles/wh_queue_200k_counters.md (same kernel sources)"
                              if counters_current else
                              "kernel sources differ from the profiled build: counter-based fields omitted"),
            "flops_per_particle_step": f_alg,
            "hbm_gbs_algorithmic": hbm_bytes_per_block(args.particles, inputs[0].n_bodies, TBSIZE)
            / (h["kernel_ms_per_block"] / 1e3) / 1e9,
        },
    }
    if "fp64" in per_form and "mixed" in per_form:
        line["mixed_vs_fp64"] = per_form["mixed"]["value"] / per_form["fp64"]["value"]
    if "fp64" in per_form and "fast" in per_form:
        line["fast_vs_fp64"] = per_form["fast"]["value"] / per_form["fp64"]["value"]
    if "fast" in per_form and "mixed" in per_form:
        line["mixed_vs_fast"] = per_form["mixed"]["value"] / per_form["fast"]["value"]
    if "fp64" in finals:
        line["precision_vs_fp64"] = {
            "after_steps": (args.warmup + args.steps) * TBSIZE,
            **{f: element_differences(inputs[0].m[0], finals["fp64"], finals[f]) for f in finals if f != "fp64"}}
    if not args.no_batch:
        line["batch"] = run_batch(forms, args.batch_systems, args.batch_launches, args.seed + 2)
        note("batch done")
    if not args.no_encounter:
        line["encounter"] = run_encounter(forms, args.encounter_particles, ENC_BLOCKS, args.seed + 3)
        note("encounter hand-off done")
    c0_final = None
    if not args.no_config0:
        line["config0"] = run_config0(forms)
        note("config-0 case done")
        c0_final = line["config0"].pop("final", None)
    records = None
    if not args.no_track:
        line["track"] = run_track(inputs)
        note("track done")
        records = line["track"].pop("records")
    if rank == 0 and not args.no_cpu:
        # the CPU leg: the reference's CPU path timed on the host and, as the
        # checker, the oracle over a particle sample of the timed run
        cpu = cpu_baseline(inputs, args.cpu_sample, threads=args.cpu_threads)
        note("CPU baseline done")
        if args.oracle_sample > 0:
            n_s = min(args.oracle_sample, args.particles)
            ref = oracle_sample_run(inputs, n_s, args.warmup + args.steps)
            note("oracle sample done")
            cpu["precision_vs_oracle"] = {
                "after_steps": (args.warmup + args.steps) * TBSIZE, "sample": n_s,
                **{f: element_differences(inputs[0].m[0], ref, subset_by_id(finals[f], ref["id"])) for f in finals}}
        if records is not None and cpu["kind"] == "reference":
            cpu["track"] = cpu_track(inputs, records)
        if not args.no_config0:
            cpu["config0"] = cpu_config0(c0_final)
        line["cpu_baseline"] = cpu
    if rank == 0:
        print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
