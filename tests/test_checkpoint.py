"""Checkpoint files (evaluation_state.c:214-290) and progress
(evaluation_optimized.c:301-333), on the CPU.

The writer must produce the reference's text, the reader must restore every
field it holds (exactly where twenty decimals represent the double exactly,
as for the integrals' magnitudes here), a missing file must leave the state
alone, and a damaged one must be refused without touching it.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import holders, fixture
from milkyway_client_amd import separation as sep


@pytest.fixture(scope="module")
def lib():
    return sep.load_library()


def state_for(fid="86", n_stars=300):
    p, stars = fixture(fid, n_stars=n_stars)
    return p, holders(p, stars)


def fill(es, seed=1):
    rng = np.random.default_rng(seed)
    e = es.es
    e.background_integral = 0.123456789
    for s in range(e.number_streams):
        e.stream_integrals[s] = 0.5 + s
    e.prob_sum = -1234.5
    e.num_zero = 3
    e.bad_jacobians = 0
    e.current_star_point = 170
    e.current_integral = 1
    for g in range(e.number_integrals):
        a = es.area(g)
        a.current_calculation = a.min_calculation + 7 * (g + 1)
        a.background_integral = float(rng.uniform(0.25, 0.75))
        for s in range(a.number_streams):
            a.stream_integrals[s] = float(rng.uniform(0.25, 0.75))


def snapshot(es):
    e = es.es
    out = [e.background_integral, e.prob_sum, e.num_zero, e.bad_jacobians, e.current_star_point,
           e.current_integral] + [e.stream_integrals[s] for s in range(e.number_streams)]
    for g in range(e.number_integrals):
        a = es.area(g)
        out += [a.mu_min, a.mu_max, a.mu_steps, a.nu_min, a.nu_max, a.nu_steps, a.r_min, a.r_max,
                a.r_steps, a.mu_step_size, a.nu_step_size, a.r_step_size, a.min_calculation,
                a.max_calculation, a.current_calculation, a.background_integral]
        out += [a.stream_integrals[s] for s in range(a.number_streams)]
    return out


def test_checkpoint_text_is_the_reference_format(lib, tmp_path):
    p, (ph, sh, es) = state_for()
    fill(es)
    f = tmp_path / "astronomy_checkpoint"
    assert lib.mw_write_checkpoint(str(f).encode(), es.ptr) == 0
    lines = f.read_text().splitlines()
    e = es.es
    assert lines[0] == "background_integral: %.20f" % e.background_integral
    assert lines[1] == "stream_integrals[1]: %.20f" % e.stream_integrals[0]
    assert lines[2] == "prob_sum: %.20f, num_zero: 3, bad_jacobians: 0" % e.prob_sum
    assert lines[3:6] == ["current_star_point: 170", "current_integral: 1",
                          "number_integrals: %d" % e.number_integrals]
    a = es.area(0)
    assert lines[6] == "mu[min,max,steps]: %.3f, %.3f, %d" % (a.mu_min, a.mu_max, a.mu_steps)
    assert lines[7] == "nu[min,max,steps]: %.3f, %.3f, %d" % (a.nu_min, a.nu_max, a.nu_steps)
    assert lines[8] == " r[min,max,steps]: %.3f, %.3f, %d" % (a.r_min, a.r_max, a.r_steps)
    assert lines[9] == ("min_calculation: %d, max_calculation: %d, current_calculation: %d"
                        % (a.min_calculation, a.max_calculation, a.current_calculation))
    assert lines[10] == "background_integral: %.20f" % a.background_integral
    assert lines[11] == "stream_integrals[1]: %.20f" % a.stream_integrals[0]
    assert len(lines) == 6 + 6 * e.number_integrals


@pytest.mark.parametrize("fid", ["86", "11"])
def test_checkpoint_round_trip(lib, tmp_path, fid):
    p, (ph, sh, es) = state_for(fid)
    fill(es, seed=int(fid))
    want = snapshot(es)
    f = tmp_path / "ckpt"
    assert lib.mw_write_checkpoint(str(f).encode(), es.ptr) == 0
    p2, (ph2, sh2, es2) = state_for(fid)
    assert lib.mw_read_checkpoint(str(f).encode(), es2.ptr) == 0
    assert snapshot(es2) == want


def test_missing_checkpoint_leaves_state(lib, tmp_path):
    p, (ph, sh, es) = state_for()
    before = snapshot(es)
    assert lib.mw_read_checkpoint(str(tmp_path / "absent").encode(), es.ptr) == 0
    assert snapshot(es) == before


@pytest.mark.parametrize("cut", [1, 3, 9, 14])
def test_damaged_checkpoint_is_refused(lib, tmp_path, cut):
    p, (ph, sh, es) = state_for()
    fill(es)
    f = tmp_path / "ckpt"
    assert lib.mw_write_checkpoint(str(f).encode(), es.ptr) == 0
    f.write_text("\n".join(f.read_text().splitlines()[:cut]) + "\n")
    p2, (ph2, sh2, es2) = state_for()
    before = snapshot(es2)
    assert lib.mw_read_checkpoint(str(f).encode(), es2.ptr) == 1
    assert snapshot(es2) == before


def test_checkpoint_of_another_model_is_refused(lib, tmp_path):
    p, (ph, sh, es) = state_for("11")          # 3 streams, 1 region
    f = tmp_path / "ckpt"
    assert lib.mw_write_checkpoint(str(f).encode(), es.ptr) == 0
    p2, (ph2, sh2, es2) = state_for("86")      # 1 stream, 2 regions
    assert lib.mw_read_checkpoint(str(f).encode(), es2.ptr) == 1


def test_progress(lib):
    """Fraction of cells and stars done (evaluation_optimized.c:301-333)."""
    p, (ph, sh, es) = state_for("86", n_stars=500)
    cells = [g.r_steps * g.mu_steps * g.nu_steps for g in p.regions]
    total = sum(cells) + 500
    assert lib.mw_calculate_progress(es.ptr) == 0.0
    es.area(0).current_calculation = 12345
    assert lib.mw_calculate_progress(es.ptr) == pytest.approx(12345 / total, rel=1e-15)
    es.es.current_integral = 1
    es.area(1).current_calculation = 77
    assert lib.mw_calculate_progress(es.ptr) == pytest.approx((cells[0] + 77) / total, rel=1e-15)
    es.es.current_integral = 2
    es.es.current_star_point = 250
    assert lib.mw_calculate_progress(es.ptr) == pytest.approx((sum(cells) + 250) / total, rel=1e-15)
    es.es.current_star_point = 500
    assert lib.mw_calculate_progress(es.ptr) == 1.0
