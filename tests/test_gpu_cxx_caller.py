"""The library from a compiled C++ caller (tests/cxx/executor_drive, built
with g++ against include/glisser_amd.h and include/glisser_amd_executor.hpp
and linked to libglisser_amd.so), as GLISSER's C++ host would use it:

* "abi": the reference executor's device calls replaced one by one --
  context, particle upload, initial accelerations, then per block the
  planet-log upload, the time-block kernel, the encounter kill and the
  partition, and the download;
* "exec": the library's glx::Executor loop (section 2b).

Both on an outer-solar-system swarm whose aphelia cross an 80 au outer
limit, three 512-step blocks, fp64 form, against the oracle's CPU path:
bit for bit with correctly rounded sin/cos, flags and survivors equal to
the CPU path with glibc, death times (exec) at the flagging steps."""
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT
from tests.helpers import oracle_run, oracle_run_parallel, planet_blocks

pytestmark = pytest.mark.gpu

DRIVER = ROOT / "tests" / "cxx" / "executor_drive"
DT, TB, BLOCKS, OUTER, CULL = 180.0, 512, 3, 80.0, 0.5


def _write_input(path, s, h0_0, rh0, blocks, form=0):
    n, nb = s.n, s.n_bodies
    with open(path, "wb") as f:
        f.write(np.array([n], np.uint64).tobytes())
        f.write(np.array([nb, TB, len(blocks), form], np.uint32).tobytes())
        f.write(np.array([DT, OUTER, CULL, 0.0], np.float64).tobytes())
        for a in (s.m, s.r, s.v, s.pr, s.pv):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        f.write(np.ascontiguousarray(s.ids, np.uint32).tobytes())
        f.write(np.ascontiguousarray(h0_0, np.float64).tobytes())
        f.write(np.ascontiguousarray(rh0, np.float64).tobytes())
        for rl, h0, rh in blocks:
            for a in (rh, rl, h0):
                f.write(np.ascontiguousarray(a, np.float64).tobytes())


def _read_output(path, n):
    raw = path.read_bytes()
    off = 8
    out = {"n_alive": int(np.frombuffer(raw[:8], np.uint64)[0])}
    for key, dt, count in (("r", np.float64, 3 * n), ("v", np.float64, 3 * n), ("flags", np.uint16, n),
                           ("id", np.uint32, n), ("deathtime", np.float64, n)):
        a = np.frombuffer(raw[off:off + count * np.dtype(dt).itemsize], dt)
        off += count * np.dtype(dt).itemsize
        out[key] = a.reshape(n, 3) if count == 3 * n else a
    assert off == len(raw)
    return out


@pytest.fixture(scope="module")
def case(oracle):
    from glisser_amd import systems
    assert DRIVER.exists(), "tests/cxx/executor_drive is built by __graft_entry__.build()"
    s = systems.outer_solar_system(600, seed=95, a_range=(20.0, 65.0), e_hi=0.4)
    h0_0, rh0, blocks = planet_blocks(oracle, s, BLOCKS, TB, DT, cull_radius=CULL, outer=OUTER)
    cpu = oracle_run(oracle, s, h0_0, rh0, blocks, DT, OUTER)
    cr = oracle_run_parallel(oracle, s, h0_0, rh0, blocks, DT, OUTER, correctly_rounded=True)
    return s, h0_0, rh0, blocks, cpu, cr


@pytest.mark.parametrize("mode", ["abi", "exec"])
def test_cxx_caller_against_oracle(case, tmp_path, mode):
    s, h0_0, rh0, blocks, cpu, cr = case
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(inp, s, h0_0, rh0, blocks)
    r = subprocess.run([str(DRIVER), mode, str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = _read_output(outp, s.n)
    assert got["n_alive"] == cpu["n_alive"]
    # survivors in the same slots and order as the CPU path's stable partition
    np.testing.assert_array_equal(got["id"][:got["n_alive"]], cpu["id"][:cpu["n_alive"]])
    o = np.argsort(got["id"])
    oc = np.argsort(cpu["id"])
    np.testing.assert_array_equal(got["flags"][o], cpu["flags"][oc])
    np.testing.assert_array_equal(got["r"][o], cr["r"])
    np.testing.assert_array_equal(got["v"][o], cr["v"])
    dead = cpu["flags"][oc] != 0
    print(f"{mode}: {got['n_alive']} alive, {int(dead.sum())} flagged")
    assert dead.sum() > 20
    if mode == "exec":
        death = cpu["death_by_id"]
        np.testing.assert_array_equal(got["deathtime"][o][dead], death[dead] * DT)
        assert np.all(np.isnan(got["deathtime"][o][~dead]))
