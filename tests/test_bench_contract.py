"""bench.py's contract pieces that run without a GPU: the algorithmic FLOP
count it prices the roofline with, the refusal of more than one GPU, the
per-form launch defaults, the CPU baseline's single-thread and all-cores
legs, its track-record check and its end-to-end run of the reference's own
case."""
import argparse
import subprocess
import sys

import numpy as np
import pytest

from tests.conftest import ROOT


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, str(ROOT))
    import bench as b
    return b


def test_flops_per_particle_step(bench):
    # (135 + 16 n_p) basic + (16 + n_p) div + (2 + n_p) sqrt + 14 sin/cos
    assert bench.flops_per_particle_step(4) == 199 + 20 + 6 + 14 == 239
    assert bench.flops_per_particle_step(1) == 151 + 17 + 3 + 14


def test_profiled_counters_match_the_kernel_sources(bench):
    """The counter-based fields of the bench line (counter FP64 share, VALU
    busy, HBM traffic, roofline fraction at the measured clock) come from a
    profile of the kernel sources whose SHA-256 bench.py records; the
    sources must still be those, or the profile is stale."""
    assert bench.PROFILED_KERNEL_SHA256 == bench.kernel_fingerprint(), (
        "kernel sources changed since profiles/wh_queue_200k_counters.md: re-profile and update bench.py")
    for table in (bench.COUNTER_FLOPS_PER_PSTEP, bench.VALU_CYCLES_PER_PSTEP, bench.CLOCK_GHZ_MEASURED):
        assert set(table) == {"fp64", "fast", "mixed"}
    assert all(2.0 < c <= bench.PEAK_CLOCK_GHZ for c in bench.CLOCK_GHZ_MEASURED.values())


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr + r.stdout and not r.stdout.startswith("{")


def test_wave_cap_defaults(bench):
    auto = argparse.Namespace(waves_per_simd=None)
    assert bench.waves_per_simd(auto, "fp64") == 4
    assert bench.waves_per_simd(auto, "fast") == 0 and bench.waves_per_simd(auto, "mixed") == 0
    assert bench.waves_per_simd(argparse.Namespace(waves_per_simd=4), "mixed") == 4


def test_cpu_baseline_slices_cover_the_swarm(bench):
    # the all-cores leg splits the swarm into disjoint slices: every particle
    # of this all-alive swarm is counted once, TBSIZE steps each
    inputs = bench.make_inputs(96, 1, 3)
    r = bench.cpu_baseline(inputs, 40, runs=2, threads=3)
    assert r["cores"] == 1 and r["kind"] in ("reference", "port") and r["value"] > 0
    assert f"({40 * bench.TBSIZE} particle-steps, median of 2 runs" in r["sample"]
    ac = r["all_cores"]
    assert ac["cores"] == 3 and ac["value"] > 0
    assert f"({96 * bench.TBSIZE} particle-steps" in ac["sample"]
    assert set(r["host"]) == {"cpu_model", "logical_cpus", "libc", "compiler"}


def test_cpu_track_check(bench, tmp_path):
    # records from the host track writer (byte-equal to the reference's) pass
    # the check exactly; one float moved by one ulp shows up as such
    from glisser_amd import native
    from oracle.pyoracle import reference_available
    if not reference_available():
        pytest.skip("reference build absent")
    inputs = bench.make_inputs(300, 1, 4)
    s = inputs[0]
    o = np.argsort(s.ids)
    st = {"pm": s.m, "pr": s.r, "pv": s.v, "pid": np.arange(s.n_bodies, dtype=np.uint32), "r": s.pr[o],
          "v": s.pv[o], "id": s.ids[o]}
    native.write_track(tmp_path / "h.trk", 0.0, st)
    head = 8 + 8 + (s.n_bodies - 1) * bench.TRACK_RECORD_BYTES + 8
    rec = np.frombuffer((tmp_path / "h.trk").read_bytes()[head:], dtype=native.TRACK_DTYPE).copy()
    r = bench.cpu_track(inputs, rec)
    assert r["ids_equal"] and r["floats_equal"] == 1.0 and r["max_ulps"] == 0 and r["value"] > 0
    rec["x"][7, 2] = np.nextafter(rec["x"][7, 2], np.float32(np.inf))
    r = bench.cpu_track(inputs, rec)
    assert r["max_ulps"] == 1 and r["floats_equal"] == 1.0 - 1.0 / rec["x"].size


def test_cpu_config0_loop(bench, oracle, monkeypatch):
    # the end-to-end CPU leg (2 of its 50 blocks here) is the same loop as
    # the parity tests' CPU path: same survivors, same positions, exact
    # particle-step count, and its checker reads that as zero error
    from glisser_amd import systems
    from tests.helpers import oracle_run, planet_blocks
    monkeypatch.setattr(bench, "CONFIG0_BLOCKS", 2)
    s = systems.outer_solar_system(bench.CONFIG0_PARTICLES, seed=1)
    h00, rh0, blocks = planet_blocks(oracle, s, 2, bench.TBSIZE, bench.DT)
    ref = oracle_run(oracle, s, h00, rh0, blocks, bench.DT, bench.OUTER)
    r = bench.cpu_config0({"id": ref["id"], "r": ref["r"], "n_alive": ref["n_alive"]})
    assert r["particle_steps"] == bench.CONFIG0_PARTICLES * 2 * bench.TBSIZE
    assert r["alive_end"] == ref["n_alive"] == bench.CONFIG0_PARTICLES
    assert r["fp64_vs_cpu"] == {"same_survivors": True, "r_rel_max": 0.0, "r_rel_median": 0.0}
