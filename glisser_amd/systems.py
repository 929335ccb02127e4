"""Seeded synthetic planetary systems and test-particle swarms.

Units: au, day; masses in GM units (Big-G = 1), GM_sun = 2.959122e-4.
Elements become Cartesian states with the reference's ``from_elements_M``
convention (src/convert.cpp:230-308): planets about GM_sun (1 + mu_i),
particles about GM_sun.  The generator is numpy's PCG64 seeded per call;
the seed is recorded in every returned system.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import native

GM_SUN = 2.959122e-4
DEG = np.pi / 180.0

# Sun + 4 giants: mass ratios and semi-major axes (au)
GIANTS_MU = np.array([9.55e-4, 2.86e-4, 4.37e-5, 5.15e-5])
GIANTS_A = np.array([5.20, 9.55, 19.2, 30.1])


@dataclass
class System:
    m: np.ndarray          # (nb,) GM of Sun + planets
    r: np.ndarray          # (nb, 3) heliocentric planets (Sun at 0)
    v: np.ndarray
    pr: np.ndarray         # (n, 3) particles
    pv: np.ndarray
    ids: np.ndarray        # (n,) particle ids
    seed: int
    meta: dict = field(default_factory=dict)

    @property
    def n_bodies(self) -> int:
        return len(self.m)

    @property
    def n(self) -> int:
        return len(self.ids)


def _elements(rng, n, a_lo, a_hi, e_hi, i_hi_deg):
    el = np.empty((n, 6))
    el[:, 0] = rng.uniform(a_lo, a_hi, n)
    el[:, 1] = rng.uniform(0.0, e_hi, n)
    el[:, 2] = rng.uniform(0.0, i_hi_deg * DEG, n)
    el[:, 3:6] = rng.uniform(0.0, 2 * np.pi, (n, 3))  # Omega, omega, M
    return el


def planets(rng, mus=GIANTS_MU, axes=GIANTS_A, e_hi=0.05, i_hi_deg=2.0):
    """Sun + planets on seeded low-e, low-i orbits; returns (m, r, v)."""
    npl = len(mus)
    el = np.empty((npl, 6))
    el[:, 0] = axes
    el[:, 1] = rng.uniform(0.0, e_hi, npl)
    el[:, 2] = rng.uniform(0.0, i_hi_deg * DEG, npl)
    el[:, 3:6] = rng.uniform(0.0, 2 * np.pi, (npl, 3))
    m = np.concatenate([[GM_SUN], GM_SUN * np.asarray(mus, dtype=np.float64)])
    pr, pv = native.from_elements(m[0] + m[1:], el, mean_anomaly=True)
    r = np.vstack([np.zeros(3), pr])
    v = np.vstack([np.zeros(3), pv])
    return m, r, v


def outer_solar_system(n_particles: int, seed: int = 1, a_range=(60.0, 65.0), e_hi=0.1,
                       i_hi_deg=5.0) -> System:
    """Sun + 4 giants (a = 5.20/9.55/19.2/30.1 au, e < 0.05, i < 2 deg) and
    a particle swarm with a ~ U[a_range], e ~ U[0, e_hi], i ~ U[0, i_hi]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m, r, v = planets(rng)
    el = _elements(rng, n_particles, a_range[0], a_range[1], e_hi, i_hi_deg)
    pr, pv = native.from_elements(np.full(n_particles, m[0]), el, mean_anomaly=True)
    return System(m, r, v, pr, pv, np.arange(n_particles, dtype=np.uint32), seed,
                  {"kind": "outer_solar_system", "a_range": a_range, "e_hi": e_hi, "i_hi_deg": i_hi_deg})


def three_body(n_particles: int, seed: int = 4, a_pl=30.0, mu_pl=1e-3, a_range=(25.0, 40.0), e_hi=0.3,
               i_hi_deg=10.0) -> System:
    """Sun + one planet on a circular orbit and a swarm around it."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = np.array([GM_SUN, GM_SUN * mu_pl])
    el = np.array([[a_pl, 0.0, 0.0, 0.0, 0.0, 0.0]])
    pr, pv = native.from_elements(m[0] + m[1:], el, mean_anomaly=True)
    r = np.vstack([np.zeros(3), pr])
    v = np.vstack([np.zeros(3), pv])
    el = _elements(rng, n_particles, a_range[0], a_range[1], e_hi, i_hi_deg)
    ppr, ppv = native.from_elements(np.full(n_particles, m[0]), el, mean_anomaly=True)
    return System(m, r, v, ppr, ppv, np.arange(n_particles, dtype=np.uint32), seed,
                  {"kind": "three_body", "a_pl": a_pl, "mu_pl": mu_pl})


@dataclass
class History:
    """Sampled planet history: times t (k,), heliocentric planet r, v
    (k, n_bodies - 1, 3), body masses m (n_bodies,)."""
    m: np.ndarray
    t: np.ndarray
    r: np.ndarray
    v: np.ndarray

    @property
    def n_bodies(self) -> int:
        return len(self.m)


def planet_history(system: System, t_end: float, spacing: float = 365.0, substeps: int = 10,
                   cull_radius: float = 0.5) -> History:
    """Integrate the planets from t = 0 with the host Jacobi WH producer at
    dt = spacing / substeps and keep every substeps-th state: samples at
    t = 0, spacing, 2 spacing, ... up to the first sample >= t_end."""
    n_samples = int(np.ceil(t_end / spacing)) + 1
    pl = native.Planets(system.m, system.r, system.v, cull_radius=cull_radius)
    try:
        npl = system.n_bodies - 1
        r = np.empty((n_samples, npl, 3))
        v = np.empty((n_samples, npl, 3))
        r[0] = system.r[1:]
        v[0] = system.v[1:]
        dt = spacing / substeps
        chunk = max(1, 4096 // substeps) * substeps
        k = 1
        while k < n_samples:
            steps = min(chunk, (n_samples - k) * substeps)
            rl, vl, _ = pl.block(steps, dt)
            got = steps // substeps
            r[k:k + got] = rl[substeps - 1::substeps]
            v[k:k + got] = vl[substeps - 1::substeps]
            k += got
    finally:
        pl.close()
    t = np.arange(n_samples, dtype=np.float64) * spacing
    return History(system.m.copy(), t, r, v)


def _hill_spaced_axes(rng, mus, a_lo, a_hi, min_hill, tries=2000):
    """Semi-major axes log-uniform in [a_lo, a_hi], sorted, with every
    neighbouring pair at least `min_hill` mutual Hill radii apart
    (R_H = ((mu_k + mu_k+1) / 3)^(1/3) (a_k + a_k+1) / 2); None if rejection
    fails `tries` times."""
    n = len(mus)
    for _ in range(tries):
        a = np.sort(np.exp(rng.uniform(np.log(a_lo), np.log(a_hi), n)))
        mu = mus  # masses keep their draw order; pair them with sorted axes
        rh = ((mu[:-1] + mu[1:]) / 3.0) ** (1.0 / 3.0) * (a[:-1] + a[1:]) / 2.0
        if np.all(np.diff(a) >= min_hill * rh):
            return a
    return None


def random_system(rng, n_particles: int, n_planets: int, mu_range=(1e-5, 1e-3), a_range=(1.0, 40.0),
                  min_hill=8.0, e_hi=0.05, i_hi_deg=2.0):
    """One system of the batched workload: Sun + n_planets planets with
    log-uniform mass ratios and semi-major axes at least `min_hill` mutual
    Hill radii apart, e < e_hi; a particle belt outside the outermost planet
    (a in [1.3, 2.0] a_out, e < 0.1, i < 5 deg).  dt = P_inner / 20.  Returns
    (System, dt, rh, outer_radius)."""
    while True:
        mus = np.exp(rng.uniform(np.log(mu_range[0]), np.log(mu_range[1]), n_planets))
        axes = None
        for _ in range(50):
            axes = _hill_spaced_axes(rng, mus, a_range[0], a_range[1], min_hill)
            if axes is not None:
                break
            mus = np.exp(rng.uniform(np.log(mu_range[0]), np.log(mu_range[1]), n_planets))
        if axes is not None:
            break
        n_planets -= 1  # too many planets for the range at this spacing
    m, r, v = planets(rng, mus=mus, axes=axes, e_hi=e_hi, i_hi_deg=i_hi_deg)
    a_out = axes[-1]
    el = _elements(rng, n_particles, 1.3 * a_out, 2.0 * a_out, 0.1, 5.0)
    pr, pv = native.from_elements(np.full(n_particles, m[0]), el, mean_anomaly=True)
    p_inner = 2 * np.pi * np.sqrt(axes[0] ** 3 / (m[0] + m[1]))
    dt = p_inner / 20.0
    rh = np.concatenate([[0.25 * axes[0]], 3.5 * axes * (mus / 3.0) ** (1.0 / 3.0)])
    s = System(m, r, v, pr, pv, np.arange(n_particles, dtype=np.uint32), 0,
               {"kind": "batch_member", "axes": axes, "mus": mus})
    return s, dt, rh, 10.0 * a_out


def batch_systems(n_systems: int = 64, seed: int = 3, n_range=(1e3, 1e6), planets_range=(2, 8),
                  n_steps: int = 4096, substeps: int = 2):
    """The batched workload: n_systems independent systems with 2-8 planets
    each (random_system), particle counts log-uniform in n_range, and a
    planet history per system sampled every dt (host Jacobi WH at
    dt / substeps) covering n_steps steps.  Returns a list of dicts with the
    keys Batch() takes plus "system"."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n_systems):
        n = int(round(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1])))))
        npl = int(rng.integers(planets_range[0], planets_range[1] + 1))
        s, dt, rh, outer = random_system(rng, n, npl)
        hist = planet_history(s, (n_steps + 1) * dt, spacing=dt, substeps=substeps, cull_radius=rh[0])
        out.append({"system": s, "m": s.m, "rh": rh, "n_particles": n, "dt": dt, "t0": 0.0,
                    "outer_radius": outer, "history": hist})
    return out
