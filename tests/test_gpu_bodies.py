"""Every planet count the kernel is built for: the specialised instances
(2..9 bodies) and the generic one (up to 16 bodies), in every arithmetic
form, against the CPU oracle on systems whose particles cross the planets'
orbits (encounters, Sun and outer-limit culls)."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import compare_flags, gpu_run, oracle_run, planet_blocks

pytestmark = pytest.mark.gpu

T = 256


@pytest.fixture(scope="module")
def native():
    from glisser_amd import native as nat
    nat.load()
    return nat


def _system(native, n_planets, n=3000):
    from glisser_amd import systems
    rng = np.random.default_rng(300 + n_planets)
    s, dt, _rh, _outer = systems.random_system(rng, 0, n_planets, mu_range=(1e-7, 1e-5), a_range=(1.0, 300.0),
                                               min_hill=4.0)
    assert s.n_bodies == n_planets + 1
    axes = s.meta["axes"]
    el = np.column_stack([rng.uniform(0.7 * axes[0], 1.6 * axes[-1], n), rng.uniform(0, 0.5, n),
                          rng.uniform(0, 0.2, n), rng.uniform(0, 2 * np.pi, (n, 3))])
    pr, pv = native.from_elements(np.full(n, s.m[0]), el, mean_anomaly=True)
    s = systems.System(s.m, s.r, s.v, pr, pv, np.arange(n, dtype=np.uint32), 0, s.meta)
    return s, dt, 2.2 * axes[-1]


@pytest.mark.parametrize("form", ["fp64", "fast", "mixed"])
@pytest.mark.parametrize("n_planets", [1, 2, 5, 8, 9, 12, 15])
def test_body_counts_match_oracle(native, oracle, n_planets, form):
    s, dt, outer = _system(native, n_planets)
    h0_0, rh0, blocks = planet_blocks(oracle, s, 2, T, dt, cull_radius=0.3 * s.meta["axes"][0], rh_factor=3.0,
                                      outer=outer)
    cpu = oracle_run(oracle, s, h0_0, rh0, blocks, dt, outer)
    with native.Context(s.n, s.n_bodies, T) as ctx:
        gpu = gpu_run(ctx, s, h0_0, rh0, blocks, dt, outer, form=form)
    flagged = int((cpu["flags"] != 0).sum())
    assert 0 < flagged < s.n  # some culls and encounters, most particles alive
    low, high = compare_flags(cpu["flags"], gpu["flags"], cpu["id"], gpu["id"])
    assert len(low) == 0 and len(high) == 0
    np.testing.assert_array_equal(np.sort(gpu["id"][: gpu["n_alive"]]), np.sort(cpu["id"][: cpu["n_alive"]]))
    alive = np.sort(cpu["id"][: cpu["n_alive"]])
    oc = np.argsort(cpu["id"])[np.searchsorted(np.sort(cpu["id"]), alive)]
    og = np.argsort(gpu["id"])[np.searchsorted(np.sort(gpu["id"]), alive)]
    err = rel_err(gpu["r"][og], cpu["r"][oc])
    # these particles move among the planets, where close approaches amplify
    # the last-bit differences of the device's sin/cos (and, in the fast
    # forms, of the reordered arithmetic) within 512 steps: measured medians
    # <= 6e-13, 99th percentiles <= 6.3e-12, maxima <= 2.3e-11
    assert np.median(err) < 1e-12 and np.quantile(err, 0.99) < 1e-11 and err.max() < 1e-10
    if form == "fp64":
        np.testing.assert_array_equal(gpu["id"], cpu["id"])


@pytest.mark.parametrize("n_planets,tb,n_blocks", [(15, 3000, 1), (8, 1, 5), (3, 4097, 1)])
def test_block_lengths(native, oracle, n_planets, tb, n_blocks):
    """Blocks longer than the planet rows that fit in LDS at once (staged in
    chunks) and one-step blocks, fp64 against the oracle."""
    s, dt, outer = _system(native, n_planets, n=700)
    h0_0, rh0, blocks = planet_blocks(oracle, s, n_blocks, tb, dt, cull_radius=0.3 * s.meta["axes"][0],
                                      rh_factor=3.0, outer=outer)
    cpu = oracle_run(oracle, s, h0_0, rh0, blocks, dt, outer)
    with native.Context(s.n, s.n_bodies, tb) as ctx:
        gpu = gpu_run(ctx, s, h0_0, rh0, blocks, dt, outer)
    np.testing.assert_array_equal(gpu["id"], cpu["id"])
    np.testing.assert_array_equal(gpu["flags"], cpu["flags"])
    np.testing.assert_array_equal(gpu["dti_by_id"], cpu["dti_by_id"])
    err = rel_err(gpu["r"], cpu["r"])
    # close approaches among the planets amplify last-bit differences, as
    # above, over up to 4097 steps (measured: 99th percentile 1.2e-11)
    assert np.median(err) < 1e-12 and np.quantile(err, 0.99) < 5e-11 and err.max() < 1e-9
