"""Throughput of the time-block particle integrator on one MI355X.

Workload (one "step" = one time-block): 200,000 test particles around the
Sun + 4 giant planets (masses 9.55e-4, 2.86e-4, 4.37e-5, 5.15e-5 Msun at
5.20/9.55/19.2/30.1 au; particles a ~ U[60,65] au, e ~ U[0,0.1],
i ~ U[0,5 deg]), dt = 180 d, Time-Block-Size 2048.  Each step runs the
time-block kernel over the alive particles, the encounter-kill transform
and the stable partition; the next block's planet log is uploaded on the
second stream while the kernel runs.  Planet logs for every block are
produced on the host before the timed region (inputs resident).

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

GM_SUN = 2.959122e-4
N_PARTICLES = 200_000
TBSIZE = 2048
DT = 180.0
OUTER = 300.0
CULL = 0.5
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def flops_per_particle_step(n_planets: int) -> int:
    """Algorithmic FP64 operations of one particle-step (reference
    src/wh.cu:74-177): 135 + 16 n_p basic ops, 16 + n_p divides, 2 + n_p
    square roots, 14 sin/cos (5 Newton iterations) -- each special op
    counted once."""
    return (135 + 16 * n_planets) + (16 + n_planets) + (2 + n_planets) + 14


def hbm_bytes_per_block(n_particles: int, n_bodies: int, tbsize: int, n_workgroups: int) -> int:
    """Algorithmic HBM bytes of one time-block launch: particle state in and
    out (r, v, a: 72 B; flags 2 B in and out; deathtime index 4 B out) plus
    the planet block read once per workgroup (n_bodies x 24 B per step)."""
    state = n_particles * (72 + 2) * 2 + n_particles * 4
    planets = n_workgroups * tbsize * n_bodies * 24
    return state + planets


def run_gpu(form: str, steps: int, warmup: int, seed: int, n_particles: int = N_PARTICLES):
    from glisser_amd import native, systems

    s = systems.outer_solar_system(n_particles, seed=seed)
    nb = s.n_bodies
    planets = native.Planets(s.m, s.r, s.v, cull_radius=CULL, rh_factor=0.0, outer_radius=OUTER)
    h0_0 = planets.h0_initial()
    rh0 = planets.state()[2]
    blocks = []
    for _ in range(warmup + steps):
        planets.recalculate_rh()
        rh = planets.state()[2]
        rl, _vl, h0 = planets.block(TBSIZE, DT)
        blocks.append((rl, h0, rh))

    ctx = native.Context(n_particles, nb, TBSIZE)
    ctx.upload(s.pr, s.pv, np.zeros(s.n, np.uint16), s.ids, a=np.zeros((s.n, 3)), dti=np.zeros(s.n, np.uint32))
    ctx.set_counts(s.n, s.n)
    ctx.init_accel(s.m, rh0, s.r[1:], h0_0, OUTER)
    ctx.kill_encounters()
    ctx.partition()

    kernel_ms = []
    part_ms = []
    psteps = 0
    t0 = None
    ctx.upload_planet_block(0, s.m, blocks[0][2], blocks[0][0], blocks[0][1])
    for k in range(warmup + steps):
        if k == warmup:
            ctx.sync()
            t0 = time.perf_counter()
        buf = k % 2
        n_alive, _ = ctx.counts()
        ctx.run_block(buf, DT, OUTER, form)
        if k + 1 < warmup + steps:
            nxt = blocks[k + 1]
            ctx.upload_planet_block(1 - buf, s.m, nxt[2], nxt[0], nxt[1])
        ctx.kill_encounters()
        done = ctx.sum_steps(n_alive)
        ctx.partition()
        km, pm = ctx.last_times()
        if k >= warmup:
            kernel_ms.append(km)
            part_ms.append(pm)
            psteps += done
    ctx.sync()
    elapsed = time.perf_counter() - t0
    n_alive, _ = ctx.counts()
    ctx.close()
    return {
        "elapsed_s": elapsed,
        "particle_steps": psteps,
        "kernel_ms": kernel_ms,
        "partition_ms": part_ms,
        "n_alive_end": n_alive,
        "n_bodies": nb,
    }


def cpu_baseline(seed: int, sample_particles: int):
    """Reference CPU path (oracle/_ref, the reference's own src/wh.cpp built
    here) or, where that library is absent, the oracle restatement: one
    2048-step block over the first `sample_particles` particles of the same
    swarm on one core."""
    from glisser_amd import native, systems
    from oracle.pyoracle import Oracle, reference_available

    kind = "reference" if reference_available() else "port"
    o = Oracle(reference=(kind == "reference"))
    s = systems.outer_solar_system(N_PARTICLES, seed=seed)
    pl = native.Planets(s.m, s.r, s.v, cull_radius=CULL, rh_factor=0.0, outer_radius=OUTER)
    rh = pl.state()[2]
    n = sample_particles
    a0, f0 = o.particle_accels(s.m, rh, OUTER, s.r[1:], pl.h0_initial(), s.pr[:n], np.zeros(n), s.ids[:n])
    rl, _vl, h0 = pl.block(TBSIZE, DT)
    t = time.perf_counter()
    out = o.particle_block(s.m, rh, OUTER, rl, h0, 0.0, DT, s.pr[:n], s.pv[:n], a0, f0, s.ids[:n])
    el = time.perf_counter() - t
    steps = int(out["dti"].astype(np.int64).sum())
    return {"value": steps / el, "unit": "particle-steps/s", "cores": 1, "kind": kind,
            "sample": f"{n} particles x 1 block of {TBSIZE} steps of the same swarm ({steps} particle-steps, "
                      f"{el:.1f} s, single thread)"}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed time-blocks")
    ap.add_argument("--warmup", type=int, default=1, help="untimed time-blocks")
    ap.add_argument("--form", default="fp64", choices=["fp64", "fast", "mixed"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--particles", type=int, default=N_PARTICLES,
                    help="swarm size (the benchmark workload is 200,000)")
    ap.add_argument("--cpu-sample", type=int, default=20000, help="particles in the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(particles shard by slices across GPUs; see the design notes).", file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))

    res = run_gpu(args.form, args.steps, args.warmup, args.seed, args.particles)
    value = res["particle_steps"] / res["elapsed_s"]
    n_planets = res["n_bodies"] - 1
    f_alg = flops_per_particle_step(n_planets)
    k_avg_s = float(np.mean(res["kernel_ms"])) / 1e3
    per_launch = res["particle_steps"] / max(len(res["kernel_ms"]), 1)
    achieved_tflops = per_launch * f_alg / k_avg_s / 1e12
    workgroups = (args.particles + 255) // 256
    line = {
        "metric": "particle-steps/s at 200k particles: FP64 vs mixed; % FP64 vector peak",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["elapsed_s"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Sun + 4 giants, 200k-particle swarm at 60-65 au)",
        "config": {
            "workload": "wh-timeblock-200k",
            "n_particles": args.particles,
            "n_bodies": res["n_bodies"],
            "tbsize": TBSIZE,
            "dt_days": DT,
            "form": args.form,
            "planet_log": "host Jacobi-WH producer, precomputed",
            "seed": args.seed,
            "parallelism": "single-gpu",
        },
        "kernel_ms_avg": k_avg_s * 1e3,
        "partition_ms_avg": float(np.mean(res["partition_ms"])),
        "particle_steps_per_s_kernel": per_launch / k_avg_s,
        "roofline": {
            "bound": "fp64-valu",
            "achieved": achieved_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_PEAK_TFLOPS,
            "traffic": None,
            "flops_per_particle_step": f_alg,
            "hbm_gbs_algorithmic": hbm_bytes_per_block(args.particles, res["n_bodies"], TBSIZE, workgroups)
            / k_avg_s / 1e9,
        },
    }
    if rank == 0 and not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.seed, args.cpu_sample)
    if rank == 0:
        print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
