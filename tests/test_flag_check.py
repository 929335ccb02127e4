"""The flag-set check of the parity tests (tests/helpers.py
unexplained_flag_mismatches), itself checked on the CPU: a run compared
with itself has no mismatch, and a run whose start is moved by 1e-4
(a planted mistake, far beyond rounding) is reported with mismatches that
no threshold explains."""
import copy

import numpy as np

from tests.helpers import oracle_run, planet_blocks, threshold_distance, unexplained_flag_mismatches


def _case(oracle):
    from glisser_amd import systems
    s = systems.three_body(2000, seed=5, a_range=(20.0, 40.0), e_hi=0.6, i_hi_deg=5.0)
    h0_0, rh0, blocks = planet_blocks(oracle, s, 3, 512, 30.0, cull_radius=3.0, rh_factor=3.5)
    return s, h0_0, rh0, blocks


def test_flag_check_accepts_identical_runs(oracle):
    s, h0_0, rh0, blocks = _case(oracle)
    cpu = oracle_run(oracle, s, h0_0, rh0, blocks, 30.0, 300.0)
    bad, dist, unexplained = unexplained_flag_mismatches(oracle, s, h0_0, rh0, blocks, 30.0, 300.0, cpu, cpu)
    assert (cpu["flags"] != 0).sum() > 100
    assert len(bad) == 0 and unexplained == []


def test_flag_check_rejects_planted_deviation(oracle):
    s, h0_0, rh0, blocks = _case(oracle)
    cpu = oracle_run(oracle, s, h0_0, rh0, blocks, 30.0, 300.0)
    s2 = copy.deepcopy(s)
    s2.pr = s.pr * (1 + 1e-4)
    off = oracle_run(oracle, s2, h0_0, rh0, blocks, 30.0, 300.0)
    bad, dist, unexplained = unexplained_flag_mismatches(oracle, s, h0_0, rh0, blocks, 30.0, 300.0, cpu, off)
    assert len(bad) > 0 and len(unexplained) == len(bad)
    assert dist.min() > 1e-6


def test_threshold_distance_at_flagging_step(oracle):
    """At the step where a particle is flagged it lies inside a radius, by
    at most one step's travel; one step earlier it was outside."""
    s, h0_0, rh0, blocks = _case(oracle)
    cpu = oracle_run(oracle, s, h0_0, rh0, blocks, 30.0, 300.0)
    pid = int(np.flatnonzero(cpu["death_by_id"] > 1)[0])
    k = int(cpu["death_by_id"][pid])
    d_at = threshold_distance(oracle, s, h0_0, rh0, blocks, 30.0, 300.0, pid, k)
    d_before = threshold_distance(oracle, s, h0_0, rh0, blocks, 30.0, 300.0, pid, k - 1)
    assert 0 < d_at < 1.0 and 0 < d_before < 1.0
