"""Stream separation on the GPU (mw_ctx_separation, mw_hip_separation) against
the oracle's restatement of separation.c:37-254.

Bars (FP64): membership probabilities within 1e-12 absolute, output
positions within 1e-11 kpc, the summed probabilities within 1e-11 relative,
and the assignment equal star for star: both sides draw glibc's drand48
sequence for the same seed (the GPU by jump-ahead), so only a star whose
random number lies within 1e-9 of a cumulative threshold may differ, and the
test counts those.  Mixed precision: probabilities within 1e-7 absolute.
"""
import numpy as np
import pytest

from helpers import fixture
from milkyway_client_amd import separation as sep
from milkyway_client_amd import synth
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder
from oracle import oracle as O

pytestmark = pytest.mark.gpu

P_TOL = 1e-12
XYZ_TOL = 1e-11
MIXED_P_TOL = 1e-7


@pytest.fixture(scope="module", autouse=True)
def _device():
    sep.load_library()
    sep.require_device()


def near_threshold(prob, r, tol):
    cum = np.cumsum(prob, axis=1)
    return np.min(np.abs(cum - r[:, None]), axis=1) < tol


def drand48_sequence(seed, n):
    import ctypes as C
    libc = C.CDLL("libc.so.6")
    libc.srand48.argtypes = [C.c_long]
    libc.drand48.restype = C.c_double
    libc.srand48(seed)
    return np.array([libc.drand48() for _ in range(n)])


def compare(p, stars, seed, precision=sep.PREC_FP64, p_tol=P_TOL):
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    ref_eval = O.evaluate_full(ph, sh)
    I = np.concatenate([[ref_eval.background_integral], ref_eval.stream_integrals])
    ref = O.separation(ph, sh, I[0], I[1:], seed)
    with sep.Context(ph, sh) as ctx:
        a, xyz, prob, summ = ctx.separation(p.parameter_vector(), precision, integrals=I, seed=seed)
    S = len(p.streams)
    print(f"S={S} N={len(stars)} max|dp|={np.max(np.abs(prob - ref.probability)):.2e} "
          f"max|dxyz|={np.max(np.abs(xyz - ref.xyz)):.2e}")
    assert np.max(np.abs(prob - ref.probability)) <= p_tol
    assert np.max(np.abs(xyz - ref.xyz)) <= XYZ_TOL
    r = drand48_sequence(seed, len(stars))
    edge = near_threshold(ref.probability, r, max(1e-9, 10 * p_tol))
    diff = a != ref.assignment
    assert not np.any(diff & ~edge), np.flatnonzero(diff & ~edge)[:10]
    assert summ.number_streams == S and summ.number_stars == len(stars)
    e_tol = 1e-11 if precision == sep.PREC_FP64 else p_tol * len(stars)
    for s in range(S):
        assert abs(summ.expected[s] - ref.expected[s]) <= e_tol * max(1.0, abs(ref.expected[s]))
        assert summ.separated[s] == int(np.sum(a == s + 1))
    return a, xyz, prob, summ, ref


@pytest.mark.parametrize("fid", ["86", "79", "11"])
def test_fixture_separation_fp64(fid):
    p, stars = fixture(fid, n_stars=2000)
    a, _, _, _, ref = compare(p, stars, seed=1234567)
    assert np.array_equal(a, ref.assignment)


def test_aux_and_general_background():
    p, stars = fixture("86", n_stars=1500, aux=True, convolve=40)
    compare(p, stars, seed=99)
    p, stars = synth.workload(3, synth.WedgeShape(20, 24, 16, 15, 800), 2)
    p.background_parameters[0] = 1.3
    p.background_parameters[3] = 0.8
    compare(p, stars, seed=5)


def test_sgr_four_streams():
    p, stars = synth.workload(7, synth.WedgeShape(16, 24, 12, 30, 3000), 4)
    p.sgr_coordinates = 1
    compare(p, stars, seed=2**31 + 17)


def test_seed_reproducibility():
    p, stars = fixture("79", n_stars=2000, convolve=30)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    with sep.Context(ph, sh) as ctx:
        a1, *_ = ctx.separation(seed=42)
        a2, *_ = ctx.separation(seed=42)
        a3, *_ = ctx.separation(seed=43)
    assert np.array_equal(a1, a2)
    assert not np.array_equal(a1, a3)


def test_integrals_computed_when_not_given():
    p, stars = fixture("86", n_stars=1000, convolve=60)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    ref = O.evaluate_full(ph, sh)
    with sep.Context(ph, sh) as ctx:
        _, _, prob, summ = ctx.separation(p.parameter_vector(), seed=3)
    assert abs(summ.background_integral - ref.background_integral) <= 1e-12 * abs(ref.background_integral)
    assert abs(summ.stream_integrals[0] - ref.stream_integrals[0]) <= 1e-12 * abs(ref.stream_integrals[0])
    want = O.separation(ph, sh, ref.background_integral, ref.stream_integrals, 3)
    assert np.max(np.abs(prob - want.probability)) <= 1e-12


def test_mixed_precision_separation():
    p, stars = fixture("86", n_stars=2000)
    compare(p, stars, seed=11, precision=sep.PREC_MIXED, p_tol=MIXED_P_TOL)


def test_large_star_list_properties():
    """200k stars of the full-size workload: a 4000-star slice matches the
    oracle, and the assignment frequencies follow the probabilities."""
    p, stars = synth.workload(0, synth.FULL_SIZE, synth.FULL_SIZE_STREAMS)
    print("stars:", len(stars), flush=True)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    I = np.array([3.1e3, 21.0, 17.5, 9.0])[: 1 + len(p.streams)]
    with sep.Context(ph, sh) as ctx:
        a, xyz, prob, summ = ctx.separation(p.parameter_vector(), integrals=I, seed=8)
    k = 4000
    sl = StarsHolder(stars[:k])
    ref = O.separation(ph, sl, I[0], I[1:], 8)
    assert np.max(np.abs(prob[:k] - ref.probability)) <= P_TOL
    assert np.max(np.abs(xyz[:k] - ref.xyz)) <= XYZ_TOL
    assert np.array_equal(a[:k], ref.assignment)
    assert np.all((a >= 0) & (a <= len(p.streams)))
    for s in range(len(p.streams)):
        e = prob[:, s].sum()
        sd = np.sqrt(np.sum(prob[:, s] * (1 - prob[:, s])))
        assert abs(summ.expected[s] - e) <= 1e-9 * max(1.0, e)
        assert abs(summ.separated[s] - e) <= 6 * sd + 1


def test_hip_separation_writes_reference_file(tmp_path):
    p, stars = fixture("86", n_stars=500, convolve=30)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    ref_eval = O.evaluate_full(ph, sh)
    out = tmp_path / "separated.txt"
    sep.set_precision(sep.PREC_FP64)
    sep.hip_separation(out, ref_eval.background_integral, ref_eval.stream_integrals, ph, sh, 2024)
    ref = O.separation(ph, sh, ref_eval.background_integral, ref_eval.stream_integrals, 2024)
    rows = np.loadtxt(out)
    assert rows.shape == (500, 4)
    assert np.array_equal(rows[:, 0].astype(int), ref.assignment)
    assert np.max(np.abs(rows[:, 1:] - ref.xyz)) <= 5e-7   # "%lf": six decimals
    # the text is the reference's format
    first = out.read_text().splitlines()[0].split()
    assert len(first) == 4 and all(len(v.split(".")[1]) == 6 for v in first[1:])


def test_no_stream_is_rejected():
    p, stars = synth.workload(1, synth.WedgeShape(8, 8, 8, 5, 50), 0)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    with sep.Context(ph, sh) as ctx:
        with pytest.raises(sep.MwError):
            ctx.separation(seed=1)


def test_separation_program(tmp_path):
    """The separation program evaluates the parameter file's own values and
    separates with that evaluation's integrals (separation.c:256-322)."""
    import subprocess
    from helpers import GOLDEN
    from milkyway_client_amd import params_io
    prog = sep.CLIENT_PATH.parent / "separation_hip"
    assert prog.exists(), "separation program not built"
    a = GOLDEN / "astronomy_parameters-79-small.txt"
    s = GOLDEN / "stars-79-first2000.txt"
    out = tmp_path / "sep79.txt"
    print("separation program on fixture 79", flush=True)
    r = subprocess.run([str(prog), "-parameters", str(a), "-stars", str(s), "-outfile", str(out),
                        "-seed", "31337"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "Output written to" in r.stdout
    p = params_io.read_astronomy_parameters(a)
    stars = params_io.read_star_points(s)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    ev = O.evaluate_full(ph, sh)
    ref = O.separation(ph, sh, ev.background_integral, ev.stream_integrals, 31337)
    rows = np.loadtxt(out)
    assert rows.shape == (len(stars), 4)
    assert np.array_equal(rows[:, 0].astype(int), ref.assignment)
    assert np.max(np.abs(rows[:, 1:] - ref.xyz)) <= 5e-7
    assert f"{ref.separated[0]} stars separated into stream" in r.stdout
