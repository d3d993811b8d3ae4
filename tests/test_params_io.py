"""Reference file formats: astronomy parameter files (parameters.c:107-305)
and star files (star_points.c:69-93), read from the reference's own fixtures."""
import numpy as np
import pytest

from helpers import GOLDEN
from milkyway_client_amd import params_io


@pytest.mark.parametrize("name", sorted(p.name for p in GOLDEN.glob("astronomy_parameters-*.txt")))
def test_fixture_files_parse_and_round_trip(name):
    p = params_io.read_astronomy_parameters(GOLDEN / name)
    assert p.convolve == 120 and p.number_streams >= 1
    text = params_io.format_astronomy_parameters(p)
    q = params_io.parse_astronomy_parameters(text)
    assert q.background_parameters == pytest.approx(p.background_parameters, abs=1e-12)
    assert q.regions == p.regions
    assert np.allclose([s.parameters for s in q.streams], [s.parameters for s in p.streams], rtol=1e-15)
    assert q.optimized_parameter_count() == p.optimized_parameter_count()


def test_fixture_86_contents():
    p = params_io.read_astronomy_parameters(GOLDEN / "astronomy_parameters-86-small.txt")
    assert p.wedge == 86 and p.sgr_coordinates == 0 and p.aux_bg_profile == 0
    assert p.background_parameters == pytest.approx([1.0, 0.6676872870, 26.0005055194, 1.0])
    assert p.background_optimize == [0, 1, 1, 0]
    assert len(p.regions) == 2
    main, cut = p.regions
    assert (main.r_min, main.r_max, main.r_steps) == (16.0, 22.5, 70)
    assert (main.mu_min, main.mu_max, main.mu_steps) == (310.0, 420.0, 160)
    assert (cut.mu_min, cut.mu_max, cut.mu_steps) == (21.6, 22.3, 40)
    assert (cut.nu_steps, cut.r_steps) == (40, 70)
    assert p.streams[0].weight == pytest.approx(-17.287176)


def test_version_0_01_has_no_aux_line():
    p = params_io.read_astronomy_parameters(GOLDEN / "astronomy_parameters-79-small.txt")
    text = params_io.format_astronomy_parameters(p)
    text = text.replace("parameters_version: 0.040000\n", "").replace("aux_bg_profile: 0\n", "")
    q = params_io.parse_astronomy_parameters(text)
    assert q.parameters_version == 0.01 and q.aux_bg_profile == 0


def test_star_files():
    s = params_io.read_star_points(GOLDEN / "stars-86-first2000.txt")
    assert s.shape == (2000, 3)
    assert s[0] == pytest.approx([195.338577, -44.234104, 7.890563])
    t = params_io.parse_star_points(params_io.format_star_points(s[:5]))
    assert np.allclose(t, s[:5], atol=1e-6)


def test_malformed_inputs_raise():
    with pytest.raises(params_io.ParameterFileError):
        params_io.parse_astronomy_parameters("parameters_version: 0.04\nnumber_parameters: x\n")
    with pytest.raises(params_io.ParameterFileError):
        params_io.parse_star_points("3\n1 2 3\n4 5 6\n")
