"""The fp64 form's guarded divide and square-root cores (core_div_pos,
core_sqrt in glisser_amd/csrc/wh_device.h): inside the guards they return the
IEEE operators' bits, and a lane outside them is recomputed with the IEEE
operators, so every lane's result is the reference arithmetic's.

* One wave mixing swarm particles (inside every guard) with particles
  outside them: one at 1e31 au (|r|^2 above 2^200, and every planet distance
  too) and one on a near-parabolic orbit at 1e16 au (semi-major axis above
  2^100), around a Sun whose four planets weigh GM = 1e-30 (the indirect
  term of real giants, ~1e-11 au/d^2, would make any particle that far
  unbound); three blocks of 256 steps with an outer limit of 1e40 au: every
  particle bit for bit equal to the CPU path with correctly rounded sin/cos.
  (A particle beyond the |v|^2 guard would be unbound, which the CPU path
  refuses, as the reference's does.)
* A launch whose dt (1e-31 d) lies outside [2^-100, 2^100]: every lane runs
  the IEEE operators; the same equality.
The guard conditions the special particles violate are checked on the CPU
first, so the fallback is known to run.
"""
import copy

import numpy as np
import pytest

from tests.helpers import gpu_run, oracle_run, planet_blocks

pytestmark = pytest.mark.gpu

OUTER_FAR = 1e40


@pytest.fixture(scope="module")
def native():
    from glisser_amd import native as nat
    nat.load()
    return nat


def _near_parabolic(mu, dist):
    """A speed just below escape at `dist` whose energy (computed as the
    drift computes it) is negative with a semi-major axis in (2^100,
    2^102): far enough below escape that the planets' kicks over the run
    cannot unbind it."""
    v_esc = np.sqrt(2.0 * mu / dist)
    for k in range(1, 64):
        v = v_esc * (1.0 - k * 2.0**-53)
        energy = (v * v) * 0.5 - mu / dist
        if energy < 0.0 and 2.0**100 < -0.5 * mu / energy < 2.0**102:
            return v
    raise AssertionError("no near-parabolic speed found")


def _mixed_wave(seed=21):
    from glisser_amd import systems
    s = copy.deepcopy(systems.outer_solar_system(64, seed=seed))
    s.m[1:] = 1e-30
    mu = s.m[0]
    # lane 5: far away, at rest (|r|^2 > 2^200, planet distances too)
    s.pr[5] = [1e31, 2e30, -3e29]
    s.pv[5] = [0.0, 0.0, 0.0]
    # lane 40: near-parabolic at 1e16 au (a > 2^100)
    d = 1e16
    s.pr[40] = [d, 0.0, 0.0]
    s.pv[40] = [0.0, _near_parabolic(mu, d), 0.0]
    r2 = np.einsum("ij,ij->i", s.pr, s.pr)
    v2 = np.einsum("ij,ij->i", s.pv, s.pv)
    assert r2[5] > 2.0**200
    inside = np.ones(64, bool)
    inside[[5, 40]] = False
    assert np.all((r2[inside] >= 2.0**-180) & (r2[inside] <= 2.0**200) & (v2[inside] <= 2.0**200))
    return s


def _assert_same(gpu, cpu):
    og, oc = np.argsort(gpu["id"]), np.argsort(cpu["id"])
    np.testing.assert_array_equal(gpu["flags"][og], cpu["flags"][oc])
    np.testing.assert_array_equal(gpu["r"][og], cpu["r"][oc])
    np.testing.assert_array_equal(gpu["v"][og], cpu["v"][oc])
    np.testing.assert_array_equal(gpu["death_by_id"], cpu["death_by_id"])


def test_guard_fallback_lanes_in_a_wave(native, oracle):
    s = _mixed_wave()
    h0_0, rh0, blocks = planet_blocks(oracle, s, 3, 256, 180.0, outer=OUTER_FAR)
    cpu = oracle_run(oracle, s, h0_0, rh0, blocks, 180.0, OUTER_FAR, correctly_rounded=True)
    with native.Context(s.n, s.n_bodies, 256) as ctx:
        gpu = gpu_run(ctx, s, h0_0, rh0, blocks, 180.0, OUTER_FAR)
    _assert_same(gpu, cpu)
    fl = dict(zip(cpu["id"].tolist(), cpu["flags"].tolist()))
    print(f"flags of the special lanes: far {fl[5]:#x}, near-parabolic {fl[40]:#x}; "
          f"{int(np.count_nonzero(cpu['flags']))} flagged in all")
    assert fl[5] == 0 and fl[40] == 0


def test_ieee_operators_when_dt_outside_the_core_range(native, oracle):
    from glisser_amd import systems
    dt = 1e-31
    assert dt < 2.0**-100
    s = systems.outer_solar_system(130, seed=22)
    h0_0, rh0, blocks = planet_blocks(oracle, s, 2, 128, dt)
    cpu = oracle_run(oracle, s, h0_0, rh0, blocks, dt, 300.0, correctly_rounded=True)
    with native.Context(s.n, s.n_bodies, 128) as ctx:
        gpu = gpu_run(ctx, s, h0_0, rh0, blocks, dt, 300.0)
    _assert_same(gpu, cpu)
