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


def oracle_run(oracle, system, h0_0, rh0, blocks, dt, outer=300.0, kill=True, correctly_rounded=False):
    """CPU path of the reference (call stack: WHIntegrator ctor, then per
    block integrate_particles_timeblock over the alive range and a stable
    partition of the unflagged particles).  Returns final per-slot arrays
    and n_alive.  correctly_rounded: the particle drift's sin/cos are the
    correctly rounded ones instead of glibc's (the result then also holds
    "libm_calls" and "libm_differ", the calls where glibc differs)."""
    if correctly_rounded:
        oracle.set_particle_libm(True)
        try:
            out = oracle_run(oracle, system, h0_0, rh0, blocks, dt, outer, kill)
            out["libm_calls"], out["libm_differ"] = oracle.particle_libm_counts()
        finally:
            oracle.set_particle_libm(False)
        return out
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
    death = np.full(n, -1, np.int64)  # by id: steps done when flagged (-1: never)
    death[ids[flags != 0]] = 0
    t = 0.0
    done = 0
    for rl, h0, rh in blocks:
        out = oracle.particle_block(system.m, rh, outer, rl, h0, t, dt, r[:n_alive], v[:n_alive], a[:n_alive],
                                    flags[:n_alive], ids[:n_alive], dti)
        if out["rc"] != 0:
            raise RuntimeError(f"oracle failed: {oracle.err()}")
        r[:n_alive], v[:n_alive], a[:n_alive], flags[:n_alive] = out["r"], out["v"], out["a"], out["flags"]
        dti = out["dti"]
        newly = ids[:n_alive][flags[:n_alive] != 0]
        death[newly] = done + dti[newly].astype(np.int64)
        done += len(h0)
        if kill:
            flags[:n_alive] = kill_encounters(flags[:n_alive])
        split, perm = oracle.stable_partition(flags, 0, n_alive, 0)
        sl = slice(0, n_alive)
        r[sl], v[sl], a[sl], flags[sl], ids[sl] = r[sl][perm], v[sl][perm], a[sl][perm], flags[sl][perm], ids[sl][perm]
        n_alive = split
        t += dt * len(h0)
    return {"r": r, "v": v, "a": a, "flags": flags, "id": ids, "dti_by_id": dti, "death_by_id": death,
            "n_alive": n_alive}


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
    death = np.full(n, -1, np.int64)  # by id: steps done when flagged (-1: never)
    d0 = ctx.download(n)
    death[d0["id"][d0["flags"] != 0]] = 0
    done = 0
    for k, (rl, h0, rh) in enumerate(blocks):
        n_alive, _ = ctx.counts()
        ctx.upload_planet_block(k % 2, system.m, rh, rl, h0)
        ctx.run_block(k % 2, dt, outer, form)
        if kill:
            ctx.kill_encounters()
        ctx.sync()
        d = ctx.download(n_alive)
        dti_by_id[d["id"]] = d["dti"]
        fl = d["flags"] != 0
        death[d["id"][fl]] = done + d["dti"][fl].astype(np.int64)
        done += len(h0)
        ctx.partition()
        if per_block:
            per_block(k, ctx)
    n_alive, _ = ctx.counts()
    d = ctx.download(n)
    d["n_alive"] = n_alive
    d["dti_by_id"] = dti_by_id
    d["death_by_id"] = death
    return d


def oracle_run_parallel(oracle, system, h0_0, rh0, blocks, dt, outer=300.0, correctly_rounded=False, chunks=8,
                        kill=True):
    """oracle_run over disjoint particle chunks on `chunks` host threads
    (particles are independent given the planet blocks; the oracle's calls
    release the interpreter lock).  Returns r, v, flags, death_by_id and the
    libm counters in id order (key "id": the sorted ids)."""
    import copy
    from concurrent.futures import ThreadPoolExecutor
    order = np.argsort(system.ids)
    parts = [p for p in np.array_split(order, chunks) if len(p)]

    def one(idx):
        sub = copy.copy(system)
        sub.pr, sub.pv = system.pr[idx], system.pv[idx]
        sub.ids = np.arange(len(idx), dtype=np.uint32)
        return oracle_run(oracle, sub, h0_0, rh0, blocks, dt, outer, kill=kill, correctly_rounded=correctly_rounded)

    with ThreadPoolExecutor(len(parts)) as pool:
        outs = list(pool.map(one, parts))
    res = {"id": system.ids[order].copy(), "r": np.empty((system.n, 3)), "v": np.empty((system.n, 3)),
           "flags": np.empty(system.n, np.uint16), "death_by_id": np.full(int(system.ids.max()) + 1, -1, np.int64),
           "libm_calls": 0, "libm_differ": 0}
    pos = 0
    for idx, out in zip(parts, outs):
        o = np.argsort(out["id"])  # local ids 0..len-1 in chunk order
        k = len(idx)
        res["r"][pos:pos + k], res["v"][pos:pos + k] = out["r"][o], out["v"][o]
        res["flags"][pos:pos + k] = out["flags"][o]
        res["death_by_id"][system.ids[idx]] = out["death_by_id"][:k]
        res["libm_calls"] += out.get("libm_calls", 0)
        res["libm_differ"] += out.get("libm_differ", 0)
        pos += k
    return res


def nudge_response(oracle, system, h0_0, rh0, blocks, dt, outer, cpu):
    """The reference arithmetic's own sensitivity: per particle (in id
    order), the relative position change after the same blocks when every
    initial position is moved by one ulp.  Particles whose response stays
    below 1e-10 are the 'regular' (non-chaotic) ones the 1e-10 bar is for."""
    import copy
    s2 = copy.deepcopy(system)
    s2.pr = np.nextafter(system.pr, np.inf)
    cpu2 = oracle_run(oracle, s2, h0_0, rh0, blocks, dt, outer)
    oc, o2 = np.argsort(cpu["id"]), np.argsort(cpu2["id"])
    r1, r2 = cpu["r"][oc], cpu2["r"][o2]
    return np.linalg.norm(r2 - r1, axis=1) / np.linalg.norm(r1, axis=1)


def threshold_distance(oracle, system, h0_0, rh0, blocks, dt, outer, pid, steps):
    """Distance (au) of particle `pid` on the oracle's trajectory, after
    `steps` steps, from the nearest flag threshold the reference tests
    (src/wh.cpp:637-656, src/wh.cu:151-170): each planet's encounter radius,
    the Sun's cull radius and the outer limit, at that step's planet
    positions.  The oracle runs the particle alone over the first `steps`
    steps of the same planet blocks (a flagged particle stops, as in the
    block runs)."""
    k = int(np.flatnonzero(system.ids == pid)[0])
    r, v = system.pr[k:k + 1].copy(), system.pv[k:k + 1].copy()
    ids = system.ids[k:k + 1].copy()
    a, flags = oracle.particle_accels(system.m, rh0, outer, system.r[1:], h0_0, r, np.zeros(1, np.uint16), ids)
    rows, radii = system.r[1:], rh0  # step 0: the initial positions
    left, t = int(steps), 0.0
    dti = np.zeros(int(system.ids.max()) + 1, np.uint32)
    for rl, h0, rh in blocks:
        if left == 0:
            break
        m = min(left, len(h0))
        out = oracle.particle_block(system.m, rh, outer, rl[:m], h0[:m], t, dt, r, v, a, flags, ids, dti)
        r, v, a, flags, dti = out["r"], out["v"], out["a"], out["flags"], out["dti"]
        left -= m
        t += dt * m
        if left == 0:
            rows, radii = rl[m - 1], rh
            break
    d_pl = np.linalg.norm(r[0] - rows, axis=1)
    d_sun = np.linalg.norm(r[0])
    gaps = np.concatenate([np.abs(d_pl - radii[1:]), [abs(d_sun - radii[0]), abs(d_sun - outer)]])
    return float(gaps.min())


def unexplained_flag_mismatches(oracle, system, h0_0, rh0, blocks, dt, outer, cpu, gpu, tol=1e-9):
    """Flag-set comparison with the one exception the bar allows: a particle
    whose flags differ between the CPU path and the GPU path must sit within
    `tol` au of a threshold on the CPU trajectory at the earlier of the two
    flagging steps.  Returns (mismatched ids, their threshold distances, ids
    not explained)."""
    low, high = compare_flags(cpu["flags"], gpu["flags"], cpu["id"], gpu["id"])
    dc, dg = cpu["death_by_id"], gpu["death_by_id"]
    bad = np.union1d(low, high)
    dist, unexplained = [], []
    for pid in bad:
        steps = [x for x in (dc[pid], dg[pid]) if x >= 0]
        g = threshold_distance(oracle, system, h0_0, rh0, blocks, dt, outer, pid, min(steps))
        dist.append(g)
        if not g < tol:
            unexplained.append(int(pid))
    return bad, np.array(dist), unexplained


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
