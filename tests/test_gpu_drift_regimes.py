"""Every branch of the FAST / MIXED Kepler solve against the CPU oracle.

Near-Keplerian particles (Sun plus one planet of 1e-12 Msun at 1000 au, so
the planet term is negligible and no particle is flagged) whose a and e put
every lane of a wave on one branch of the restructured solve:
  * a = 60 au, e < 0.1: |dM| ~ 0.007, the short path (3-term Taylor start,
    one short correction);
  * a = 20 au: |dM| ~ 0.035, start angle above 0.02, the retry with the
    5-term polynomial and full corrections;
  * a = 8 au: |dM| ~ 0.14, retry with a start error ~1e-5;
  * a = 60 au, e = 0.9-0.95: eccentric, second corrections near perihelion;
  * a = 3 and 6 au: |dM| > 0.2, the reference-arithmetic f and g functions.
The reference CPU path (the oracle) runs the same 256 steps; positions and
velocities agree to 1e-11 relative, flags exactly.  Unbound particles are
flagged 0x08 in the step they are met and keep their position, as the fp64
form does."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import gpu_run, oracle_run, planet_blocks

pytestmark = pytest.mark.gpu

DT = 180.0
OUTER = 300.0
GM_SUN = 2.959122e-4

# (a, e_lo, e_hi): one regime of the solve each, 64 particles (one wave)
REGIMES = [(60.0, 0.0, 0.1), (20.0, 0.0, 0.1), (8.0, 0.0, 0.1), (60.0, 0.9, 0.95), (6.0, 0.0, 0.1),
           (3.0, 0.0, 0.2)]


def regime_system(seed=5, unbound=0):
    """Sun + a 1e-12 Msun planet at 1000 au and 64 particles per regime
    (mean anomalies spread over the whole orbit); `unbound` more particles
    on hyperbolic orbits at 60 au."""
    from glisser_amd import native, systems
    rng = np.random.default_rng(seed)
    m = np.array([GM_SUN, GM_SUN * 1e-12])
    pl_r, pl_v = native.from_elements(GM_SUN * (1 + 1e-12), np.array([[1000.0, 0.0, 0.0, 0.0, 0.0, 0.0]]))
    els = []
    for a, e_lo, e_hi in REGIMES:
        el = np.empty((64, 6))
        el[:, 0] = a
        el[:, 1] = rng.uniform(e_lo, e_hi, 64)
        el[:, 2] = rng.uniform(0.0, 0.1, 64)
        el[:, 3:5] = rng.uniform(0.0, 2 * np.pi, (64, 2))
        el[:, 5] = np.linspace(0.0, 2 * np.pi, 64, endpoint=False)
        els.append(el)
    el = np.concatenate(els)
    pr, pv = native.from_elements(GM_SUN, el)
    if unbound:
        ur, uv = native.from_elements(GM_SUN, np.tile([[60.0, 0.05, 0.05, 1.0, 2.0, 0.0]], (unbound, 1)))
        uv *= np.linspace(1.6, 2.5, unbound)[:, None]  # above escape speed (sqrt 2 x circular)
        pr, pv = np.concatenate([pr, ur]), np.concatenate([pv, uv])
    n = len(pr)
    return systems.System(m, np.concatenate([[np.zeros(3)], pl_r]), np.concatenate([[np.zeros(3)], pl_v]), pr, pv,
                          np.arange(n, dtype=np.uint32), seed, {"kind": "drift_regimes"})


def test_regimes_cover_the_branches():
    """The regimes are what the docstring says (host-side check of dM and of
    the start angle the solve evaluates first)."""
    n = np.sqrt(GM_SUN / np.array([a for a, _, _ in REGIMES]) ** 3)
    dM = n * DT
    assert dM[0] < 0.01 and 0.02 < dM[1] < 0.2 and 0.02 < dM[2] < 0.2 and dM[4] > 0.2 and dM[5] > 0.2
    # perihelion start angle dM / (1 - e) of the eccentric regime exceeds 0.02
    assert dM[3] / (1 - 0.9) > 0.02


@pytest.mark.parametrize("form", ["fp64", "fast", "mixed"])
def test_drift_regimes_match_oracle(oracle, form):
    from glisser_amd import native
    s = regime_system()
    h0_0, rh0, blocks = planet_blocks(oracle, s, 1, 256, DT)
    cpu = oracle_run(oracle, s, h0_0, rh0, blocks, DT, OUTER)
    with native.Context(s.n, 2, 256) as ctx:
        gpu = gpu_run(ctx, s, h0_0, rh0, blocks, DT, OUTER, form=form)
    assert cpu["n_alive"] == s.n and gpu["n_alive"] == s.n
    oc, og = np.argsort(cpu["id"]), np.argsort(gpu["id"])
    np.testing.assert_array_equal(gpu["flags"][og], cpu["flags"][oc])
    er = rel_err(gpu["r"][og], cpu["r"][oc])
    ev = rel_err(gpu["v"][og], cpu["v"][oc])
    for k, (a, e_lo, e_hi) in enumerate(REGIMES):
        sl = slice(64 * k, 64 * (k + 1))
        print(f"{form} a={a} e=[{e_lo},{e_hi}]: max r-err {er[sl].max():.2e}, v-err {ev[sl].max():.2e}")
    assert er.max() < 1e-11 and ev.max() < 1e-11


@pytest.mark.parametrize("form", ["fast", "mixed"])
def test_unbound_lanes_in_fast_forms(oracle, form):
    """Hyperbolic particles among bound ones: flagged 0x08 at the first step
    (death index 1), state equal to the fp64 form's (both keep r through the
    drift and complete the step's kicks); the bound ones are unaffected."""
    from glisser_amd import native
    s = regime_system(unbound=8)
    h0_0, rh0, blocks = planet_blocks(oracle, s, 1, 64, DT)
    out = {}
    for f in ("fp64", form):
        with native.Context(s.n, 2, 64) as ctx:
            out[f] = gpu_run(ctx, s, h0_0, rh0, blocks, DT, OUTER, form=f, kill=False)
    ref, got = out["fp64"], out[form]
    o1, o2 = np.argsort(ref["id"]), np.argsort(got["id"])
    nb = len(REGIMES) * 64
    np.testing.assert_array_equal(got["flags"][o2], ref["flags"][o1])
    assert np.all(got["flags"][o2][nb:] == 0x08) and np.all(got["flags"][o2][:nb] == 0)
    assert np.all(got["dti_by_id"][nb:] == 1)
    np.testing.assert_allclose(got["r"][o2][nb:], s.pr[nb:], rtol=0, atol=0)
    assert rel_err(got["v"][o2][nb:], ref["v"][o1][nb:]).max() < 1e-12


@pytest.mark.parametrize("form", ["fp64", "fast", "mixed"])
def test_results_independent_of_wave_grouping(oracle, form):
    """A particle's result depends on its own state only, not on which other
    particles share its wave: the regime mix above in two different orders
    (so every wave holds a different mix of branches) gives bit-identical
    r, v, a and flags per particle id."""
    from glisser_amd import native
    s = regime_system(unbound=8)
    h0_0, rh0, blocks = planet_blocks(oracle, s, 1, 256, DT)
    perm = np.random.default_rng(11).permutation(s.n)
    out = []
    for order in (np.arange(s.n), perm):
        t = type(s)(s.m, s.r, s.v, s.pr[order], s.pv[order], s.ids[order], s.seed, s.meta)
        with native.Context(s.n, 2, 256) as ctx:
            d = gpu_run(ctx, t, h0_0, rh0, blocks, DT, OUTER, form=form, kill=False)
        o = np.argsort(d["id"])
        out.append({k: d[k][o] for k in ("r", "v", "a", "flags")})
    for k in ("r", "v", "a", "flags"):
        np.testing.assert_array_equal(out[1][k], out[0][k], err_msg=k)
