"""Pins of the shared float32 elementary functions and Philox against
published known answers and float64 numpy (the oracle and the kernels share
them, so they are checked on their own here)."""
import numpy as np
import pytest

import oracle as O

# Random123 kat_vectors, Philox4x32-10 (Salmon et al., SC'11).
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    assert O.philox(ctr, key) == expect


def _max_ulps(name, ref, xs):
    f = O.elementary(name)
    got = np.array([f(float(x)) for x in xs], np.float32).astype(np.float64)
    r = ref(xs.astype(np.float64))
    ulp = np.spacing(np.abs(r).astype(np.float32)).astype(np.float64)
    return float(np.max(np.abs(got - r) / ulp))


RNG = np.random.default_rng(7)


@pytest.mark.parametrize("name,ref,xs,bound", [
    ("sin", np.sin, np.concatenate([RNG.uniform(-8, 8, 4000), RNG.uniform(-1e-3, 1e-3, 500)]), 2.0),
    ("cos", np.cos, np.concatenate([RNG.uniform(-8, 8, 4000), RNG.uniform(-1e-3, 1e-3, 500)]), 2.0),
    ("exp", np.exp, np.concatenate([RNG.uniform(-100, 88, 4000), RNG.uniform(-1, 1, 500)]), 1.5),
    ("log", np.log, np.concatenate([RNG.uniform(1e-6, 1, 3000), 10 ** RNG.uniform(-37, 37, 1000),
                                    np.float32(2.0) ** -np.arange(127, 149)]), 2.0),
    ("tanh", np.tanh, np.concatenate([RNG.uniform(-10, 10, 4000), RNG.uniform(-0.6, 0.6, 2000)]), 2.0),
])
def test_elementary_accuracy(name, ref, xs, bound):
    xs = xs.astype(np.float32)
    assert _max_ulps(name, ref, xs) <= bound


def test_surrogate_tanh_accuracy():
    """The bf16 surrogate's branch-free tanh (host form: libm exp2f, a true
    division; the GPU form swaps in its approximate exp2 and reciprocal, about
    an ulp each): within a few ulp of tanh, the right limits and NaN."""
    xs = np.concatenate([RNG.uniform(-10, 10, 20000), RNG.uniform(-0.6, 0.6, 20000),
                         RNG.uniform(0.5, 0.6, 5000), 10.0 ** RNG.uniform(-30, -1, 2000)]).astype(np.float32)
    assert _max_ulps("tanh_mlp", np.tanh, xs) <= 4.0
    f = O.elementary("tanh_mlp")
    assert f(60.0) == 1.0 and f(-60.0) == -1.0 and f(float("inf")) == 1.0 and f(0.0) == 0.0
    assert np.isnan(f(float("nan")))


def test_special_values():
    assert O.elementary("sin")(0.0) == 0.0 and O.elementary("cos")(0.0) == 1.0
    assert O.elementary("exp")(0.0) == 1.0 and O.elementary("log")(1.0) == 0.0
    assert O.elementary("tanh")(20.0) == 1.0 and O.elementary("tanh")(-20.0) == -1.0
    assert np.isinf(O.elementary("exp")(100.0)) and O.elementary("exp")(-200.0) == 0.0
    assert np.isneginf(O.elementary("log")(0.0)) and np.isnan(O.elementary("log")(-1.0))
    assert np.isnan(O.elementary("tanh")(float("nan")))


def test_uniform_and_normal_statistics():
    u = np.array([O.uniform4(42, i, 3) for i in range(5000)]).ravel()
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.003
    z = np.array([O.normal4(43, i, 0) for i in range(10000)]).ravel()
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    assert abs(np.mean(z ** 3)) < 0.06 and abs(np.mean(z ** 4) - 3) < 0.15
    # distinct (seed, subsequence, offset) give distinct streams
    assert not np.array_equal(O.normal4(1, 0, 0), O.normal4(1, 1, 0))
    assert not np.array_equal(O.normal4(1, 0, 0), O.normal4(1, 0, 1))


@pytest.mark.parametrize("cell", [0.025, 0.05, 0.1, 0.0625])
def test_nearest_index_shortcut_exhaustive(cell):
    """racer_nearest_index equals floorf(a / cell + 0.5f) for every float a in
    [-40, 40], and for every float a with 40 <= |a| <= 4096 (grids up to 160 k
    cells wide)."""
    O.set_threads(8)
    assert O.lib().o_check_nearest_index(cell, -40.0, 40.0) == 0
    assert O.lib().o_check_nearest_index(cell, 40.0, 4096.0) == 0
    assert O.lib().o_check_nearest_index(cell, -4096.0, -40.0) == 0


@pytest.mark.parametrize("cell", [0.025, 0.05])
def test_nearest_index_with_run_limit_exhaustive(cell):
    """The lidar's form, one limit for a run of calls whose |a / cell| is
    bounded: equal to floorf(a / cell + 0.5f) for every float of each run."""
    O.set_threads(8)
    for lo, hi in ((-40.0, 40.0), (-12.0, 3.0), (100.0, 130.0), (-2000.0, -1900.0)):
        assert O.lib().o_check_nearest_index_lim(cell, lo, hi) == 0, (lo, hi)


@pytest.mark.parametrize("cell,stride", [(0.025, 1), (0.05, 13), (1.0 / 3.0, 13), (7.3, 13)])
def test_nearest_cell_by_reciprocal(cell, stride):
    """The lidar's nearest cell from the reciprocal quotient: the same index as
    floorf(a / cell + 0.5f) for every float a on the grid, off the grid where
    that is (all 2^32 patterns for the built-in tracks' cell, every 13th
    pattern for the others)."""
    O.set_threads(8)
    assert O.lib().o_check_nearest_by_reciprocal(cell, stride) == 0


@pytest.mark.parametrize("cell,stride", [(0.025, 1), (0.05, 13), (0.1, 13), (1.0 / 3.0, 13), (7.3, 13)])
def test_nearest_cell_by_one_step_reciprocal(cell, stride):
    """The padded lidar march's nearest cell (one correction step from RN(1 /
    cell)): the index floorf(a / cell + 0.5f) for every float a whose index is
    a number below 2^22 in magnitude, and off every grid where it is not (all
    2^32 patterns for the built-in tracks' cell, every 13th for the others)."""
    O.set_threads(8)
    assert O.lib().o_check_nearest_by_reciprocal1(cell, stride) == 0


@pytest.mark.parametrize("cell", [0.025, 0.05, 0.1, 0.0625, 0.03, 1.0 / 3.0])
def test_per_map_nearest_form_check(cell):
    """The check the library runs at sim creation, which decides whether the
    lidar may take the one-step nearest cell: it passes the one-step form over
    a padded grid's index range, and it finds the boundary cells where the
    multiply-only form RN(a RN(1 / cell)) takes the wrong index (for 1/16 m,
    whose reciprocal is exact, that form has none)."""
    lib = O.lib()
    assert lib.o_nearest_form_bad(cell, -460, 1900, 1) == 0
    bad2 = lib.o_nearest_form_bad(cell, -460, 1900, 2)
    if cell == 0.0625:
        assert bad2 == 0
    else:
        assert bad2 > 0


def test_div_by_reciprocal_is_correctly_rounded():
    """The lidar march step's quotient from the beam's RN(1 / descent): bit
    for bit the IEEE division over its stated range (random operands, quotients
    next to rounding midpoints, all-ones and power-of-two divisors, +0)."""
    O.set_threads(8)
    assert O.lib().o_check_div_by_reciprocal(16_000_000, 11) == 0


def test_branch_free_tanh_is_racer_tanh():
    """racer_tanh_select (both paths evaluated, one selected; the float32 fused
    forward's epilogue) equals racer_tanh bit for bit: every 61st float bit
    pattern here (all 2^32 agree as well, checked with stride 1)."""
    O.set_threads(8)
    assert O.lib().o_check_tanh_select(61) == 0
