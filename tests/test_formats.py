"""File formats (configuration, state files, binary track) of the product
host code against the reference's own writers and readers (oracle/_ref,
src/data.cpp) byte for byte, and against golden files the reference wrote
(tests/golden/formats, made by tests/golden/make_golden.py)."""
import numpy as np
import pytest

from tests.conftest import GOLDEN

FMT = GOLDEN / "formats"


def sample_state(seed=5, n=300):
    from glisser_amd import systems
    s = systems.outer_solar_system(n, seed=seed, a_range=(10.0, 90.0), e_hi=0.5, i_hi_deg=40.0)
    rng = np.random.default_rng(seed)
    ids = rng.permutation(n).astype(np.uint32) + 7
    flags = rng.choice(np.array([0, 0, 0, 0x80, 0x0101, 0x02, 0x0301], np.uint16), n)
    deathtime = np.where(flags != 0, rng.uniform(0, 1e6, n), 0.0)
    return {"pm": s.m, "pr": s.r, "pv": s.v, "pid": np.array([0, 3, 1, 4, 2], np.uint32), "r": s.pr, "v": s.pv,
            "id": ids, "flags": flags, "deathtime": deathtime}


CONFIG = """# comment
Initial-Time 12.5
Time-Step 180
Final-Time 3650000
Time-Block-Size 2048
Encounter-RH-Factor 3.5
Cull-Radius 0.25
Outer-Limit 400
Resolve-Encounters 1
Read-Planet-History 1
Planet-History-File hist.bin
Input-File state.in
Output-File state.out
Output-Folder out/
Track-Interval 10
Write-Barycentric-Track 0
Big-G 1.5
"""

# lines the product writes differently on purpose: the reference writes
# Cull-Radius under Encounter-RH-Factor and leaves Big-G out
CHANGED = ("Encounter-RH-Factor", "Big-G")


def _strip(text):
    return [ln for ln in text.splitlines() if ln.split(" ")[0] not in CHANGED]


def test_config_normalize_matches_reference(reference):
    from glisser_amd import native
    for text in (CONFIG, "Input-File a\nOutput-File b\n"):
        ours, ref = native.config_normalize(text), reference.config_normalize(text)
        assert _strip(ours) == _strip(ref)
    ours = native.config_normalize(CONFIG)
    assert "Encounter-RH-Factor 3.5" in ours and "Big-G 1.5" in ours
    # the canonical form parses back to itself
    assert native.config_normalize(ours) == ours


@pytest.mark.parametrize("bad", ["Time-Step\n", "Time-Step abc\nInput-File a\n", "No-Such-Key 1\nInput-File a\n",
                                 "Input-File a\nResolve-Encounters 1\n", "Output-File \n"])
def test_config_errors_match_reference(reference, bad):
    from glisser_amd import native
    from oracle.pyoracle import OracleError
    with pytest.raises(OracleError) as ref_err:
        reference.config_normalize(bad)
    with pytest.raises(native.GlxError) as our_err:
        native.config_normalize(bad)
    assert ref_err.value.args[0].split(": ", 1)[1] in str(our_err.value)


def test_config_golden():
    from glisser_amd import native
    assert _strip(native.config_normalize(CONFIG)) == _strip((FMT / "config.out").read_text())


@pytest.mark.parametrize("fmt,name", [(0, "state.txt"), (1, "state.bin")])
@pytest.mark.parametrize("momenta", [False, True])
def test_state_writers_match_reference(reference, tmp_path, fmt, name, momenta):
    from glisser_amd import native
    st = sample_state()
    ours, ref = tmp_path / ("ours_" + name), tmp_path / ("ref_" + name)
    native.save_state(ours, st, fmt, write_momenta=momenta)
    reference.state_save(ref, fmt, st, write_momenta=momenta)
    assert ours.read_bytes() == ref.read_bytes()


def test_split_writer_matches_reference(reference, tmp_path):
    from glisser_amd import native
    st = sample_state()
    native.save_state(tmp_path / "pl", st, native.STATE_SPLIT, particle_path=tmp_path / "ics")
    reference.state_save(tmp_path / "rpl", 2, st, particle_path=tmp_path / "rics")
    assert (tmp_path / "pl").read_bytes() == (tmp_path / "rpl").read_bytes()
    assert (tmp_path / "ics").read_bytes() == (tmp_path / "rics").read_bytes()


def test_state_golden_files(tmp_path):
    """The reference wrote these files from sample_state(); the product
    writes the same bytes and reads them back exactly."""
    from glisser_amd import native
    st = sample_state()
    for fmt, name in ((0, "state.txt"), (1, "state.bin")):
        native.save_state(tmp_path / name, st, fmt)
        assert (tmp_path / name).read_bytes() == (FMT / name).read_bytes(), name
        back = native.load_state(FMT / name, fmt)
        for k in st:
            np.testing.assert_array_equal(back[k], st[k], err_msg=f"{name} {k}")
    back = native.load_state(FMT / "pl.txt", native.STATE_SPLIT, particle_path=FMT / "ics.txt")
    for k in ("pm", "pr", "pv", "r", "v", "id", "flags"):
        np.testing.assert_array_equal(back[k], st[k] if k != "pid" else np.arange(5), err_msg=k)
    # the split particle file keeps deathtime as float32 text, as the reference reads it
    np.testing.assert_array_equal(back["deathtime"], st["deathtime"].astype(np.float32).astype(np.float64))


def test_state_readers_match_reference(reference, tmp_path):
    from glisser_amd import native
    st = sample_state()
    reference.state_save(tmp_path / "s.txt", 0, st)
    a, b = native.load_state(tmp_path / "s.txt"), reference.state_load(tmp_path / "s.txt", 0)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    reference.state_save(tmp_path / "pl", 2, st, particle_path=tmp_path / "ics")
    a = native.load_state(tmp_path / "pl", native.STATE_SPLIT, particle_path=tmp_path / "ics")
    b = reference.state_load(tmp_path / "pl", 2, particle_path=tmp_path / "ics")
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_state_units_and_limits(tmp_path):
    from glisser_amd import native
    st = sample_state()
    native.save_state(tmp_path / "m.bin", st, native.STATE_BINARY, write_momenta=True)
    back = native.load_state(tmp_path / "m.bin", native.STATE_BINARY, read_momenta=True, big_g=2.0,
                             max_particles=100)
    assert len(back["id"]) == 100
    np.testing.assert_allclose(back["pv"], st["pv"], rtol=1e-15)
    np.testing.assert_array_equal(back["pm"], st["pm"] * 2.0)
    with pytest.raises(native.GlxError):
        native.load_state(tmp_path / "missing.txt")


@pytest.mark.parametrize("to_elements", [True, False])
@pytest.mark.parametrize("bary", [True, False])
def test_track_matches_reference(reference, tmp_path, to_elements, bary):
    from glisser_amd import native
    st = sample_state()
    native.write_track(tmp_path / "t", 123.25, st, to_elements, bary)
    native.write_track(tmp_path / "t", 456.5, st, to_elements, bary, append=True)
    reference.track_write(tmp_path / "r", 123.25, st, to_elements, bary)
    reference.track_write(tmp_path / "r", 456.5, st, to_elements, bary, append=True)
    assert (tmp_path / "t").read_bytes() == (tmp_path / "r").read_bytes()


def test_track_golden(tmp_path):
    from glisser_amd import native
    st = sample_state()
    native.write_track(tmp_path / "t", 123.25, st, True, True)
    data = (tmp_path / "t").read_bytes()
    assert data == (FMT / "track.bin").read_bytes()
    # layout: f64 time, u64 planets-1, 4 x 28 B, u64 n, n x 28 B
    assert len(data) == 8 + 8 + 4 * 28 + 8 + 300 * 28
