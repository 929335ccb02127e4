"""Shared drivers for the parity tests: the same time-block loop run through
the CPU oracle and through the HIP path (C ABI), on the same inputs."""
from __future__ import annotations

import numpy as np


def planet_blocks(oracle, system, n_blocks, tbsize, dt, cull_radius=0.5, rh_factor=0.0, outer=300.0):
    """Planet logs of n_blocks blocks from the oracle's planet producer.
    Returns (h0_initial, rh_initial, [(r_log, h0_log, rh), ...]); rh is
    recomputed at each block start as the executor does
    (src/executor.cu:233)."""
    pl = oracle.planets(system.m, system.r, system.v, cull_radius=cull_radius, rh_factor=rh_factor, outer=outer)
    rh0 = pl.state()[2]
    blocks = []
    for _ in range(n_blocks):
        pl.recalculate_rh()
        rh = pl.state()[2]
        rl, _vl, h0 = pl.block(tbsize, dt)
        blocks.append((rl, h0, rh))
    return pl.h0_0.copy(), rh0, blocks


def kill_encounters(flags):
    f = np.asarray(flags, dtype=np.uint16).copy()
    enc = (f & 0x01) == 0x01
    f[enc] = (f[enc] & 0xFF00) | 0x80
    return f


def oracle_run(oracle, system, h0_0, rh0, blocks, dt, outer=300.0, kill=True):
    """CPU path of the reference (call stack: WHIntegrator ctor, then per
    block integrate_particles_timeblock over the alive range and a stable
    partition of the unflagged particles).  Returns final per-slot arrays
    and n_alive."""
    n = system.n
    r, v = system.pr.copy(), system.pv.copy()
    ids = system.ids.copy()
    flags = np.zeros(n, np.uint16)
    a, flags = oracle.particle_accels(system.m, rh0, outer, system.r[1:], h0_0, r, flags, ids)
    if kill:
        flags = kill_encounters(flags)
    split, perm = oracle.stable_partition(flags, 0, n, 0)
    r, v, a, flags, ids = r[perm], v[perm], a[perm], flags[perm], ids[perm]
    n_alive = split
    dti = np.zeros(n, np.uint32)  # indexed by id, as in the reference
    t = 0.0
    for rl, h0, rh in blocks:
        out = oracle.particle_block(system.m, rh, outer, rl, h0, t, dt, r[:n_alive], v[:n_alive], a[:n_alive],
                                    flags[:n_alive], ids[:n_alive], dti)
        if out["rc"] != 0:
            raise RuntimeError(f"oracle failed: {oracle.err()}")
        r[:n_alive], v[:n_alive], a[:n_alive], flags[:n_alive] = out["r"], out["v"], out["a"], out["flags"]
        dti = out["dti"]
        if kill:
            flags[:n_alive] = kill_encounters(flags[:n_alive])
        split, perm = oracle.stable_partition(flags, 0, n_alive, 0)
        sl = slice(0, n_alive)
        r[sl], v[sl], a[sl], flags[sl], ids[sl] = r[sl][perm], v[sl][perm], a[sl][perm], flags[sl][perm], ids[sl][perm]
        n_alive = split
        t += dt * len(h0)
    return {"r": r, "v": v, "a": a, "flags": flags, "id": ids, "dti_by_id": dti, "n_alive": n_alive}


def gpu_run(ctx, system, h0_0, rh0, blocks, dt, outer=300.0, form="fp64", kill=True, per_block=None):
    """Same loop through the C ABI: upload, initial accelerations, then per
    block upload planet log -> time-block kernel -> kill -> stable partition."""
    n = system.n
    flags = np.zeros(n, np.uint16)
    ctx.upload(system.pr, system.pv, flags, system.ids, a=np.zeros((n, 3)), dti=np.zeros(n, np.uint32))
    ctx.set_counts(n, n)
    ctx.init_accel(system.m, rh0, system.r[1:], h0_0, outer)
    if kill:
        ctx.kill_encounters()
    ctx.partition()
    dti_by_id = np.zeros(n, np.uint32)
    for k, (rl, h0, rh) in enumerate(blocks):
        n_alive, _ = ctx.counts()
        ctx.upload_planet_block(k % 2, system.m, rh, rl, h0)
        ctx.run_block(k % 2, dt, outer, form)
        if kill:
            ctx.kill_encounters()
        ctx.sync()
        d = ctx.download(n_alive)
        dti_by_id[d["id"]] = d["dti"]
        ctx.partition()
        if per_block:
            per_block(k, ctx)
    n_alive, _ = ctx.counts()
    d = ctx.download(n)
    d["n_alive"] = n_alive
    d["dti_by_id"] = dti_by_id
    return d


def compare_flags(cpu_flags, gpu_flags, cpu_ids, gpu_ids):
    """Flag-set comparison keyed by particle id: returns (ids whose flagged
    state or low byte differ, ids whose high byte differs)."""
    order_c = np.argsort(cpu_ids)
    order_g = np.argsort(gpu_ids)
    fc = np.asarray(cpu_flags)[order_c].astype(np.uint32)
    fg = np.asarray(gpu_flags)[order_g].astype(np.uint32)
    ids = np.asarray(cpu_ids)[order_c]
    low = ids[(fc & 0xFF) != (fg & 0xFF)]
    high = ids[(fc >> 8) != (fg >> 8)]
    return low, high


# au, years: GM of the Sun = 4 pi^2, the units the reference's mean-motion
# estimate of a history interval assumes (src/interp.cpp:271-272)
GM_SUN_YR = 4 * np.pi ** 2


def write_plhist_bytes(path, sun_gm, records):
    """A planet history file in the reference's layout (README.md:154-171):
    records = [(time, [(id, gm, a, e, i, Omega, omega, M), ...]), ...]."""
    import struct
    with open(path, "wb") as f:
        f.write(struct.pack("<d", sun_gm) + b"\0" * 24)
        for t, pls in records:
            f.write(struct.pack("<dI", t, len(pls)) + b"\0" * 20)
            for p in pls:
                f.write(struct.pack("<I7f", *p))


def sample_history(seed, n_records=30, spacing=1.0, n_planets=4, dropout=False, newcomer=False, e_mode=None,
                   sun_gm=GM_SUN_YR, drift=True):
    """Synthetic history records (au, years): planets whose mean anomaly
    advances at the two-body rate (plus a small drift unless drift=False),
    every eccentricity regime of the reference's Kepler solver (e < 0.18,
    < 0.8, >= 0.8), optionally a planet absent for a few records (dropout)
    and one that appears part way (newcomer)."""
    rng = np.random.default_rng(seed)
    base = []
    for j in range(n_planets):
        regime = e_mode if e_mode is not None else j % 3
        e = (rng.uniform(0, 0.17), rng.uniform(0.18, 0.79), rng.uniform(0.8, 0.95))[regime]
        base.append((100 + 7 * j, sun_gm * 10 ** rng.uniform(-5, -3), rng.uniform(1, 40), e, rng.uniform(0, 0.5),
                     rng.uniform(-np.pi, np.pi), rng.uniform(-np.pi, np.pi), rng.uniform(-np.pi, np.pi)))
    wrap = lambda x: (x + np.pi) % (2 * np.pi) - np.pi  # noqa: E731
    records = []
    for k in range(n_records):
        pls = []
        for j, (pid, gm, a, e, inc, node, peri, m0) in enumerate(base):
            if dropout and j == 0 and 10 <= k < 14:
                continue
            n = np.sqrt((sun_gm + gm) / a ** 3)
            if drift:
                pls.append((pid, gm, a * (1 + 1e-4 * np.sin(0.3 * k)), e * (1 + 1e-3 * np.cos(0.2 * k)), inc,
                            wrap(node + 1e-3 * k), wrap(peri - 2e-3 * k), wrap(m0 + (n + 1e-6 * j) * spacing * k)))
            else:
                pls.append((pid, gm, a, e, inc, node, peri, wrap(m0 + n * spacing * k)))
        if newcomer and k >= 8:
            pls.append((999, sun_gm * 1e-4, 12.0, 0.3, 0.1, 1.0, 2.0, wrap(0.5 + 0.2 * k)))
        records.append((k * spacing, pls))
    return records
