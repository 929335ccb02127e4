"""The reference's own CPU-runnable case end to end: Sun + 4 giants
integrated directly on the host (Jacobi WH producer), 1,000 particles with
a ~ U[60, 65] au, e < 0.1, i < 5 deg, dt = 180 d, time-block 2048, 50 blocks
(102,400 steps), cull radius 0.5 au, kill mode -- through glx::Executor on
the GPU and through the oracle's loop (the reference CPU path) on the CPU.

The fp64 form rounds every basic operation as the CPU path does, and its
sin/cos are correctly rounded.  So it must equal, bit for bit, the CPU path
run with a correctly rounded sin/cos (the oracle's alternative libm: glibc's
long-double values, or 113-bit ones where those cannot decide): that is
asserted for every particle.  Against the CPU path with glibc itself the
only differences are then the calls in which glibc's sin/cos misround
(counted in the same run: a few per million), each a one-ulp perturbation
of one trajectory (its effect need not follow the particle's response to a
one-ulp nudge of its start: a step late in a run can move a particle more);
flags are identical, at least 99 % of the particles are within the 1e-10
relative bar, all within 1e-9, and the median is below 1e-11.
(No particle reaches the cull radius or the outer limit in 5e4 years.)"""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import compare_flags, oracle_run, oracle_run_parallel, planet_blocks

pytestmark = pytest.mark.gpu

DT = 180.0
TB = 2048
BLOCKS = 50
OUTER = 300.0


def test_config0_executor_vs_oracle(oracle):
    import copy
    from concurrent.futures import ThreadPoolExecutor

    from glisser_amd import native, systems
    s = systems.outer_solar_system(1000, seed=1)
    h0_0, rh0, blocks = planet_blocks(oracle, s, BLOCKS, TB, DT)
    s2 = copy.deepcopy(s)
    s2.pr = np.nextafter(s.pr, np.inf)
    # the oracle's calls release the interpreter lock: both CPU runs at once
    with ThreadPoolExecutor(2) as pool:
        runs = [pool.submit(oracle_run, oracle, x, h0_0, rh0, blocks, DT, OUTER) for x in (s, s2)]
        cpu, cpu2 = (r.result() for r in runs)
    cr = oracle_run_parallel(oracle, s, h0_0, rh0, blocks, DT, OUTER, correctly_rounded=True)
    with native.Executor(s.m, s.r, s.v, s.pr, s.pv, s.ids, dt=DT, tbsize=TB, outer_radius=OUTER) as ex:
        ex.step(BLOCKS)
        gpu = ex.download()
    assert gpu["t"] == BLOCKS * TB * DT
    assert gpu["n_alive"] == cpu["n_alive"]
    low, high = compare_flags(cpu["flags"], gpu["flags"], cpu["id"], gpu["id"])
    assert len(low) == 0 and len(high) == 0
    oc, og, o2 = np.argsort(cpu["id"]), np.argsort(gpu["id"]), np.argsort(cpu2["id"])
    # the reference arithmetic with correctly rounded sin/cos: bit for bit
    np.testing.assert_array_equal(gpu["id"][og], cr["id"])
    np.testing.assert_array_equal(gpu["r"][og], cr["r"])
    np.testing.assert_array_equal(gpu["v"][og], cr["v"])
    np.testing.assert_array_equal(gpu["flags"][og], cr["flags"])
    # the reference arithmetic with glibc
    nudge = rel_err(cpu2["r"][o2], cpu["r"][oc])
    er = rel_err(gpu["r"][og], cpu["r"][oc])
    ev = rel_err(gpu["v"][og], cpu["v"][oc])
    out = np.flatnonzero(er >= 1e-10)
    print(f"glibc sin/cos misround in {cr['libm_differ']} of {cr['libm_calls']} particle-drift calls; "
          f"vs the CPU path: bit-identical {(er == 0).sum()}/1000, within 1e-10 {(er < 1e-10).sum()}, "
          f"r-err median {np.median(er):.2e}, max {er.max():.2e}; one-ulp nudge median {np.median(nudge):.2e}, "
          f"max {nudge.max():.2e}; above 1e-10 (r-err, nudge): "
          + ", ".join(f"({er[i]:.2e}, {nudge[i]:.2e})" for i in out))
    assert (er < 1e-10).mean() >= 0.99 and (ev < 1e-10).mean() >= 0.99
    assert er.max() < 1e-9 and ev.max() < 1e-9
    assert np.median(er) < 1e-11
