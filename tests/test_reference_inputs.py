"""The reference's own input files (data/*.in, example/threebody-*: copied
unchanged under tests/golden/reference_inputs) through the product's
readers, against what the reference's readers make of them, and the first
half of the reference's three-body example (planets integrated and written
to a planet history every 100 blocks, as its threebody-planets
configuration asks) against the oracle and the reference build."""
import numpy as np
import pytest

from tests.conftest import GOLDEN

D = GOLDEN / "reference_inputs"
# (file, state format, particle file) as the reference reads them
FILES = [("threebody-planets.state.in", 0, None), ("threebody-particles.state.in.full", 0, None),
         ("outer-planets.in", 2, "empty_particles.in"), ("three-body.in", 2, "empty_particles.in")]


@pytest.mark.parametrize("name,fmt,ppath", FILES)
def test_state_files_load_like_reference(glx_lib, name, fmt, ppath):
    from glisser_amd import native
    g = np.load(D / "loaded.npz")
    st = native.load_state(D / name, fmt, particle_path=None if ppath is None else D / ppath)
    for k in ("pm", "pr", "pv", "pid", "r", "v", "id", "flags", "deathtime"):
        np.testing.assert_array_equal(st[k], g[f"{name}:{k}"], err_msg=f"{name} {k}")


@pytest.mark.parametrize("name", ["threebody-planets.config.in", "threebody-particles.config.in"])
def test_example_configs_parse_like_reference(glx_lib, name):
    from glisser_amd import native
    ours = native.config_normalize((D / name).read_text()).splitlines()
    ref = (D / (name + ".normalized")).read_text().splitlines()
    # two keys are written differently on purpose (see tests/test_formats.py)
    changed = ("Encounter-RH-Factor", "Big-G")
    assert [x for x in ours if x.split(" ")[0] not in changed] == [x for x in ref if x.split(" ")[0] not in changed]


def example_planet_history(path, writer, planets_factory, n_records=11):
    """threebody-planets: the planets of its state file integrated with
    Time-Step 0.2 and Time-Block-Size 5, one history record every 100
    blocks (Swift-History-Interval 100) from t = 0."""
    from glisser_amd import native
    st = native.load_state(D / "threebody-planets.state.in", 0)
    m, r, v, ids = st["pm"], st["pr"], st["pv"], st["pid"]
    pl = planets_factory(m, r, v, ids)
    writer(path, 0.0, m, r, v, ids, False)
    for k in range(1, n_records):
        for _ in range(100):
            pl.block(5, 0.2)
        r, v = pl.state()[:2]
        writer(path, k * 100.0, m, r, v, ids, True)
    return st


def _native_writer(path, t, m, r, v, ids, append):
    from glisser_amd import native
    native.write_plhist(path, t, m, r, v, ids, append)


def test_example_planet_history_matches_oracle_and_reference(glx_lib, oracle, tmp_path):
    from glisser_amd import native
    example_planet_history(tmp_path / "glx.plhist", _native_writer,
                           lambda m, r, v, ids: native.Planets(m, r, v, ids))
    example_planet_history(tmp_path / "orc.plhist",
                           lambda p, t, m, r, v, ids, a: oracle.plhist_write(p, t, m, r, v, ids, a),
                           lambda m, r, v, ids: oracle.planets(m, r, v, ids))
    glx = (tmp_path / "glx.plhist").read_bytes()
    assert glx == (tmp_path / "orc.plhist").read_bytes()
    assert len(glx) == 32 + 11 * (32 + 2 * 32)
    from oracle.pyoracle import Oracle, reference_available
    if reference_available():
        ref = Oracle(reference=True)
        example_planet_history(tmp_path / "ref.plhist",
                               lambda p, t, m, r, v, ids, a: ref.plhist_write(p, t, m, r, v, ids, a),
                               lambda m, r, v, ids: ref.planets(m, r, v, ids))
        assert glx == (tmp_path / "ref.plhist").read_bytes()
    # and it reads back into record-aligned blocks of the example's particle run
    h = native.PlanetHistory(tmp_path / "glx.plhist", 0.0, 0.2)
    b = h.block(100)
    assert b["n_steps"] == 100 and b["n_bodies"] == 3 and b["dt"] == pytest.approx(0.2, rel=1e-15)
